// The MI355X path behind the reference's own C++ types (include/cstfem_hip_reference.hpp).
// The reference's serial core (oracle/_ref/libcstfem_ref.so, compiled from the reference's
// sources) reads the reference's Kirsch L1 mesh, assembles K, adds a traction and eliminates
// the symmetry constraints; then the same cstfem::CsrMatrix and vectors go, without copies
// into another type, to
//   1. cstfem::CgJacobiSolver (the reference, on the host),
//   2. cstfem_hip::dropin::HipCgJacobiSolver through a cstfem::LinearSolver reference,
//   3. the reference's own pcg_jacobi loop over DeviceOperator / DeviceReduction,
//   4. an unqualified pcg_jacobi call on DeviceOperator (the device solve, by ADL),
// and cst_stiffness is compared on random triangles.  Exit code 0 when every check passes.
#include <cmath>
#include <cstdio>
#include <memory>
#include <random>
#include <string>

#include "cstfem/assembly.hpp"
#include "cstfem/bc.hpp"
#include "cstfem/element.hpp"
#include "cstfem/mesh.hpp"
#include "cstfem/solver.hpp"
#include "cstfem_hip_reference.hpp"

static int g_fail = 0, g_pass = 0;
#define CHECK(c)                                                           \
  do {                                                                     \
    if (c) {                                                               \
      ++g_pass;                                                            \
    } else {                                                               \
      ++g_fail;                                                            \
      std::printf("  CHECK failed: %s (%s:%d)\n", #c, __FILE__, __LINE__); \
    }                                                                      \
  } while (0)

static double rel_diff(const std::vector<double>& a, const std::vector<double>& b) {
  double d = 0.0, m = 0.0;
  for (size_t i = 0; i < a.size(); ++i) {
    d = std::max(d, std::fabs(a[i] - b[i]));
    m = std::max(m, std::fabs(b[i]));
  }
  return d / m;
}

static bool starts_with(const std::string& s, const char* p) { return s.rfind(p, 0) == 0; }

int main(int argc, char** argv) {
  const std::string path = argc > 1 ? argv[1] : "tests/golden/kirsch_L1.msh";
  const cstfem::Mesh mesh = cstfem::read_gmsh(path);
  const cstfem::Mat3 D = cstfem::constitutive_matrix(210e9, 0.3, cstfem::PlaneCondition::Stress);
  const double t = 0.01;
  cstfem::CsrMatrix K = cstfem::build_stiffness_pattern(mesh);
  cstfem::assemble_stiffness(mesh, D, t, K);
  std::vector<double> f(K.num_rows, 0.0);
  cstfem::add_traction(mesh, "right", t, 1e6, 0.0, f);
  cstfem::DirichletDofs dofs;
  cstfem::add_dirichlet_group(mesh, "left_sym", 0, 0.0, dofs);
  cstfem::add_dirichlet_group(mesh, "bottom_sym", 1, 0.0, dofs);
  cstfem::apply_dirichlet(dofs, K, f);
  cstfem::SolverOptions opt;
  opt.rtol = 1e-12;
  opt.record_history = true;

  // 1. the reference
  cstfem::CgJacobiSolver ref(opt);
  ref.setup(K);
  std::vector<double> x_ref;
  const cstfem::SolverStats s_ref = ref.solve(f, x_ref);
  std::printf("reference CgJacobiSolver: %s\n", s_ref.message.c_str());
  CHECK(s_ref.converged);

  // 2. the device solver through the reference's abstract interface
  {
    std::unique_ptr<cstfem::LinearSolver> s =
        std::make_unique<cstfem_hip::dropin::HipCgJacobiSolver>(opt);
    s->setup(K);
    std::vector<double> x;
    const cstfem::SolverStats st = s->solve(f, x);
    std::printf("dropin HipCgJacobiSolver: %s\n", st.message.c_str());
    CHECK(st.converged && starts_with(st.message, "converged in "));
    CHECK(std::abs(st.iterations - s_ref.iterations) <= 2);
    CHECK(rel_diff(x, x_ref) < 1e-9);
    CHECK(st.residual_history.size() == static_cast<size_t>(st.iterations) + 1);
    CHECK(st.true_rel_residual < 1e-11 && st.warning.empty());
    CHECK(&s->stats() != nullptr && s->stats().iterations == st.iterations);
  }
  // solve() before setup(): the reference's logic_error
  {
    cstfem_hip::dropin::HipCgJacobiSolver s(opt);
    std::vector<double> x;
    bool thrown = false;
    try {
      (void)s.solve(f, x);
    } catch (const std::logic_error& e) {
      thrown = std::string(e.what()).find("before setup") != std::string::npos;
    }
    CHECK(thrown);
  }

  // 3. the reference's pcg_jacobi loop over the device operator and reduction
  const cstfem_hip::dropin::DeviceOperator A(K);
  const cstfem_hip::dropin::DeviceReduction red(A);
  std::vector<double> inv_diag(K.num_rows);
  for (int i = 0; i < K.num_rows; ++i) inv_diag[i] = 1.0 / K.diagonal(i);
  {
    std::vector<double> x;
    const cstfem::SolverStats st =
        cstfem::pcg_jacobi(static_cast<const cstfem::LinearOperator&>(A), inv_diag,
                           static_cast<const cstfem::Reduction&>(red), f, x, opt);
    std::printf("reference pcg_jacobi loop, device SpMV and dots: %s\n", st.message.c_str());
    CHECK(st.converged && std::abs(st.iterations - s_ref.iterations) <= 2);
    CHECK(rel_diff(x, x_ref) < 1e-9);
  }

  // 4. unqualified call on the device operator: the device solve (found by ADL)
  {
    std::vector<double> x, x_timed;
    using cstfem::pcg_jacobi;
    const cstfem::SolverStats st = pcg_jacobi(A, inv_diag, red, f, x, opt);
    std::printf("device pcg_jacobi: %s\n", st.message.c_str());
    CHECK(st.converged && std::abs(st.iterations - s_ref.iterations) <= 2);
    CHECK(rel_diff(x, x_ref) < 1e-9);
    cstfem::SolveBreakdown bd;
    bd.clock = cstfem::steady_seconds;
    const cstfem::SolverStats st2 = pcg_jacobi(A, inv_diag, red, f, x_timed, opt, &bd);
    CHECK(st2.iterations == st.iterations && x_timed == x);  // timing on = timing off
    CHECK(bd.spmv_calls == st.iterations + 1 && bd.total > 0.0);
  }

  // cst_stiffness on random triangles, bit for bit
  {
    cstfem_hip::Device dev;
    std::mt19937 gen(7);
    std::uniform_real_distribution<double> u(-3.0, 3.0);
    int same = 0;
    for (int k = 0; k < 200; ++k) {
      cstfem::TriCoords xy{u(gen), u(gen), u(gen), u(gen), u(gen), u(gen)};
      if (cstfem::cst_signed_area(xy) < 0.0) {
        std::swap(xy[2], xy[4]);
        std::swap(xy[3], xy[5]);
      }
      if (std::fabs(cstfem::cst_signed_area(xy)) < 1e-3) continue;
      const cstfem::Mat6 a = cstfem::cst_stiffness(xy, D, t);
      const cstfem::Mat6 b = cstfem_hip::dropin::cst_stiffness(dev, xy, D, t);
      same += a == b;
      CHECK(a == b);
    }
    std::printf("cst_stiffness: %d triangles bit-identical\n", same);
  }

  std::printf("%d checks passed, %d failed\n", g_pass, g_fail);
  return g_fail ? 1 : 0;
}
