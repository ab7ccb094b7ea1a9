// C++ drop-in checks: the reference's own test flows (tests/test_pcg.cpp, test_patch.cpp,
// test_bc.cpp, test_csr.cpp of the reference) run against the GPU path through
// include/cstfem_hip.hpp.  Exit code 0 when every check passes.  Runs on a GPU.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>
#include <string>

#include "cstfem_hip.hpp"

using namespace cstfem_hip;

static int g_fail = 0, g_pass = 0;
#define CHECK(c)                                                                  \
  do {                                                                            \
    if (c) {                                                                      \
      ++g_pass;                                                                   \
    } else {                                                                      \
      ++g_fail;                                                                   \
      std::printf("  CHECK failed: %s (%s:%d)\n", #c, __FILE__, __LINE__);        \
    }                                                                             \
  } while (0)

template <class F>
static void check_throws(F&& f, const std::string& what) {
  try {
    f();
    ++g_fail;
    std::printf("  expected an exception containing '%s'\n", what.c_str());
  } catch (const std::exception& e) {
    if (std::string(e.what()).find(what) == std::string::npos) {
      ++g_fail;
      std::printf("  exception '%s' lacks '%s'\n", e.what(), what.c_str());
    } else {
      ++g_pass;
    }
  }
}

using Dense = std::vector<double>;

static CsrMatrix dense_to_csr(const Dense& A, int n) {
  CsrMatrix K;
  K.num_rows = K.num_cols = n;
  K.row_ptr.push_back(0);
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j)
      if (A[i * n + j] != 0.0) {
        K.cols.push_back(j);
        K.vals.push_back(A[i * n + j]);
      }
    K.row_ptr.push_back(static_cast<int>(K.cols.size()));
  }
  return K;
}

static std::vector<double> cholesky_solve(Dense A, int n, std::vector<double> b) {
  for (int j = 0; j < n; ++j) {
    double d = A[j * n + j];
    for (int k = 0; k < j; ++k) d -= A[j * n + k] * A[j * n + k];
    A[j * n + j] = std::sqrt(d);
    for (int i = j + 1; i < n; ++i) {
      double s = A[i * n + j];
      for (int k = 0; k < j; ++k) s -= A[i * n + k] * A[j * n + k];
      A[i * n + j] = s / A[j * n + j];
    }
  }
  for (int i = 0; i < n; ++i) {
    for (int k = 0; k < i; ++k) b[i] -= A[i * n + k] * b[k];
    b[i] /= A[i * n + i];
  }
  for (int i = n - 1; i >= 0; --i) {
    for (int k = i + 1; k < n; ++k) b[i] -= A[k * n + i] * b[k];
    b[i] /= A[i * n + i];
  }
  return b;
}

static Dense random_spd(int n, std::mt19937& gen) {
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  Dense M(n * n), A(n * n, 0.0);
  for (double& v : M) v = u(gen);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      for (int k = 0; k < n; ++k) A[i * n + j] += M[k * n + i] * M[k * n + j];
      if (i == j) A[i * n + j] += n;
    }
  std::vector<double> d(n);
  for (int i = 0; i < n; ++i) d[i] = std::pow(10.0, 3.0 * (u(gen) + 1.0) / 2.0);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) A[i * n + j] *= d[i] * d[j];
  return A;
}

static void test_random_spd() {
  std::mt19937 gen(2026);
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  for (int n : {5, 40, 120}) {
    const Dense A = random_spd(n, gen);
    std::vector<double> b(n);
    for (double& v : b) v = u(gen);
    const auto x_ref = cholesky_solve(A, n, b);
    HipCgJacobiSolver solver({1e-12, 1000, true});
    solver.setup(dense_to_csr(A, n));
    std::vector<double> x;
    const SolverStats st = solver.solve(b, x);
    double err = 0.0, ref = 0.0;
    for (int i = 0; i < n; ++i) {
      err = std::max(err, std::abs(x[i] - x_ref[i]));
      ref = std::max(ref, std::abs(x_ref[i]));
    }
    std::printf("  n = %3d: %d iterations, rel. residual %.2e, max rel. error %.2e\n", n,
                st.iterations, st.rel_residual, err / ref);
    CHECK(st.converged);
    CHECK(st.rel_residual < 1e-12);
    CHECK(err / ref < 1e-9);
    CHECK(st.residual_history.size() == static_cast<std::size_t>(st.iterations) + 1);
    CHECK(st.residual_history.front() == 1.0 && st.residual_history.back() < 1e-12);
  }
}

static void test_edge_cases() {
  std::mt19937 gen(1);
  {
    HipCgJacobiSolver solver;
    solver.setup(dense_to_csr(random_spd(6, gen), 6));
    std::vector<double> x(6, 123.0);
    const SolverStats st = solver.solve(std::vector<double>(6, 0.0), x);
    CHECK(st.converged && st.iterations == 0);
    CHECK(std::all_of(x.begin(), x.end(), [](double v) { return v == 0.0; }));
  }
  {
    std::mt19937 g3(3);
    HipCgJacobiSolver solver({1e-14, 2, false});
    solver.setup(dense_to_csr(random_spd(30, g3), 30));
    std::vector<double> x;
    const SolverStats st = solver.solve(std::vector<double>(30, 1.0), x);
    std::printf("  max_iter = 2: %s\n", st.message.c_str());
    CHECK(!st.converged && st.iterations == 2);
    CHECK(st.message.find("NOT converged") != std::string::npos);
    CHECK(!solver.stats().converged);
  }
  {
    HipCgJacobiSolver solver;
    check_throws([&] { solver.setup(dense_to_csr({2.0, 1.0, 1.0, -1.0}, 2)); },
                 "non-positive diagonal");
    HipCgJacobiSolver fresh;
    std::vector<double> x;
    check_throws([&] { (void)fresh.solve({1.0}, x); }, "before setup");
  }
}

static void test_breakdown_accounting() {
  std::mt19937 gen(21);
  const int n = 400;
  const CsrMatrix K = dense_to_csr(random_spd(n, gen), n);
  std::vector<double> b(n, 1.0), x_off, x_on;
  HipCgJacobiSolver off({1e-12, 1000, false}), on({1e-12, 1000, false});
  SolveBreakdown bd;
  on.set_breakdown(&bd);
  off.setup(K);
  on.setup(K);
  const SolverStats so = off.solve(b, x_off), sn = on.solve(b, x_on);
  CHECK(so.iterations == sn.iterations);
  CHECK(x_off == x_on);
  CHECK(bd.spmv_calls == sn.iterations + 1);
  CHECK(bd.total > 0.0 && bd.accounted() <= bd.total * 1.05);
  CHECK(bd.halo_bytes_sent == 0 && bd.allreduce == 0.0 && bd.halo_wait == 0.0);
}

static void test_symmetric_elimination() {
  // test_bc.cpp:63-78 through the device elimination (generic CSR upload)
  Device dev;
  CsrMatrix K;
  K.num_rows = K.num_cols = 3;
  K.row_ptr = {0, 2, 5, 7};
  K.cols = {0, 1, 0, 1, 2, 1, 2};
  K.vals = {4, 1, 1, 3, 1, 1, 2};
  check(cstfem_upload_csr(dev.get(), 3, K.row_ptr.data(), K.cols.data(), K.vals.data()));
  std::vector<double> b = {1, 1, 1};
  DirichletDofs d;
  d.add(1, 2.0);
  apply_dirichlet(dev, d, K, b);
  CHECK(b[0] == -1.0 && b[1] == 6.0 && b[2] == -1.0);
  const std::vector<double> expected = {4, 0, 0, 3, 0, 0, 2};
  CHECK(K.vals == expected);
  DirichletDofs c;
  c.add(4, 1.0);
  c.add(4, 1.0);
  CHECK(c.size() == 1);
  check_throws([&] { c.add(4, 2.0); },
               "conflicting Dirichlet values for node 2 (0-based) component ux");
}

// Structured unit square, nx x ny cells, diagonal split (a,b,c), (a,c,d).
static Mesh square_mesh(int nx, int ny) {
  Mesh m;
  for (int j = 0; j <= ny; ++j)
    for (int i = 0; i <= nx; ++i) {
      m.xy.push_back(static_cast<double>(i) / nx);
      m.xy.push_back(static_cast<double>(j) / ny);
    }
  const int W = nx + 1;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const int a = j * W + i, b = a + 1, d = a + W, c = d + 1;
      for (int v : {a, b, c, a, c, d}) m.triangles.push_back(v);
    }
  for (int i = 0; i < nx; ++i) {
    m.boundary_edges["bottom"].push_back({i, i + 1});
    m.boundary_edges["top"].push_back({ny * W + i, ny * W + i + 1});
  }
  for (int j = 0; j < ny; ++j) {
    m.boundary_edges["left"].push_back({j * W, (j + 1) * W});
    m.boundary_edges["right"].push_back({j * W + nx, (j + 1) * W + nx});
  }
  return m;
}

static void test_pattern_and_rigid_modes() {
  // test_csr.cpp:36-98 on a structured mesh
  Device dev;
  const Mesh m = square_mesh(12, 9);
  CsrMatrix K = build_stiffness_pattern(dev, m);
  const std::size_t nodes = m.num_nodes(), edges = 12 * 10 + 9 * 13 + 12 * 9;
  CHECK(K.nnz() == 4 * (nodes + 2 * edges));
  bool sorted = true, diag = true;
  for (int i = 0; i < K.num_rows; ++i) {
    for (int k = K.row_ptr[i] + 1; k < K.row_ptr[i + 1]; ++k)
      if (K.cols[k - 1] >= K.cols[k]) sorted = false;
    if (K.find(i, i) < 0) diag = false;
  }
  CHECK(sorted && diag);
  assemble_stiffness(dev, m, constitutive_matrix(210e9, 0.3, PlaneCondition::Stress), 0.01, K);
  double kmax = 0.0, knorm = 0.0;
  for (double v : K.vals) kmax = std::max(kmax, std::abs(v));
  int asym = 0;
  for (int i = 0; i < K.num_rows; ++i) {
    double rs = 0.0;
    for (int k = K.row_ptr[i]; k < K.row_ptr[i + 1]; ++k) {
      rs += std::abs(K.vals[k]);
      const int t = K.find(K.cols[k], i);
      if (t < 0 || std::abs(K.vals[k] - K.vals[t]) > 1e-12 * kmax) ++asym;
    }
    knorm = std::max(knorm, rs);
  }
  CHECK(asym == 0);
  for (int mode = 0; mode < 3; ++mode) {
    std::vector<double> r(K.num_rows), y(K.num_rows, 0.0);
    double rmax = 0.0, ymax = 0.0;
    for (std::size_t n = 0; n < nodes; ++n) {
      r[2 * n] = mode == 0 ? 1.0 : (mode == 2 ? -m.xy[2 * n + 1] : 0.0);
      r[2 * n + 1] = mode == 1 ? 1.0 : (mode == 2 ? m.xy[2 * n] : 0.0);
    }
    for (int i = 0; i < K.num_rows; ++i) {
      for (int k = K.row_ptr[i]; k < K.row_ptr[i + 1]; ++k) y[i] += K.vals[k] * r[K.cols[k]];
      rmax = std::max(rmax, std::abs(r[i]));
      ymax = std::max(ymax, std::abs(y[i]));
    }
    CHECK(ymax / (knorm * rmax) <= 1e-12);
  }
}

static void test_dirichlet_patch() {
  // test_patch.cpp:52-73: a linear field prescribed on the whole boundary is reproduced
  const double a[3] = {1e-4, 2e-3, -1e-3}, bb[3] = {-5e-4, 1e-3, 3e-3};
  for (PlaneCondition plane : {PlaneCondition::Stress, PlaneCondition::Strain}) {
    Device dev;
    Mesh m = square_mesh(16, 11);
    for (double& v : m.xy) v = v * 1.3 + 0.1;  // move off the axes
    CsrMatrix K = build_stiffness_pattern(dev, m);
    assemble_stiffness(dev, m, constitutive_matrix(210e9, 0.3, plane), 0.01, K);
    DirichletDofs dofs;
    for (const auto& [g, es] : m.boundary_edges)
      for (const auto& e : es)
        for (int n : e) {
          const double x = m.xy[2 * n], y = m.xy[2 * n + 1];
          dofs.add(2 * n, a[0] + a[1] * x + a[2] * y);
          dofs.add(2 * n + 1, bb[0] + bb[1] * x + bb[2] * y);
        }
    std::vector<double> f(K.num_rows, 0.0);
    apply_dirichlet(dev, dofs, K, f);
    HipCgJacobiSolver solver({1e-12, 20000, false});
    solver.setup(K);
    std::vector<double> u;
    const SolverStats st = solver.solve(f, u);
    double err = 0.0, ref = 0.0;
    for (std::size_t n = 0; n < m.num_nodes(); ++n) {
      const double x = m.xy[2 * n], y = m.xy[2 * n + 1];
      const double ex[2] = {a[0] + a[1] * x + a[2] * y, bb[0] + bb[1] * x + bb[2] * y};
      for (int c = 0; c < 2; ++c) {
        err = std::max(err, std::abs(u[2 * n + c] - ex[c]));
        ref = std::max(ref, std::abs(ex[c]));
      }
    }
    std::printf("  Dirichlet patch: %d iterations, max rel. error %.2e\n", st.iterations, err / ref);
    CHECK(st.converged);
    CHECK(err / ref < 1e-9);
  }
}

int main() {
  struct T {
    const char* name;
    void (*f)();
  } tests[] = {{"test_random_spd", test_random_spd},
               {"test_edge_cases", test_edge_cases},
               {"test_breakdown_accounting", test_breakdown_accounting},
               {"test_symmetric_elimination", test_symmetric_elimination},
               {"test_pattern_and_rigid_modes", test_pattern_and_rigid_modes},
               {"test_dirichlet_patch", test_dirichlet_patch}};
  for (const auto& t : tests) {
    std::printf("%s\n", t.name);
    std::fflush(stdout);
    try {
      t.f();
    } catch (const std::exception& e) {
      ++g_fail;
      std::printf("  unexpected exception: %s\n", e.what());
    }
  }
  std::printf("%d checks passed, %d failed\n", g_pass, g_fail);
  return g_fail == 0 ? 0 : 1;
}
