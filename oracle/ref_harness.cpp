// ref_harness.cpp -- C-ABI shim over the reference's own serial core, compiled from the
// reference sources where they lie (see oracle/Makefile).  TEST INFRASTRUCTURE ONLY:
// tests/ and bench.py's cpu_baseline leg load oracle/_ref/libcstfem_ref.so to check the
// oracle restatement and to time the reference's CPU path.  Nothing in the product links it.
//
// Each ref_* function has the same signature as the oc_* function of cst_oracle.c and
// forwards to the reference function named beside it.
#include <cstdint>
#include <cstring>
#include <exception>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "cstfem/assembly.hpp"
#include "cstfem/bc.hpp"
#include "cstfem/csr.hpp"
#include "cstfem/element.hpp"
#include "cstfem/kirsch.hpp"
#include "cstfem/mesh.hpp"
#include "cstfem/postprocess.hpp"
#include "cstfem/solver.hpp"

using namespace cstfem;

namespace {

thread_local std::string g_error;

Mesh make_mesh(int nn, int64_t ne, const int* tri, const double* xy) {
  Mesh m;
  m.xy.assign(xy, xy + 2 * static_cast<size_t>(nn));
  m.node_tags.resize(nn);
  for (int i = 0; i < nn; ++i) m.node_tags[i] = static_cast<size_t>(i) + 1;
  m.triangles.assign(tri, tri + 3 * ne);
  return m;
}

CsrMatrix make_csr(int n, const int* row_ptr, const int* cols, const double* vals) {
  CsrMatrix K;
  K.num_rows = K.num_cols = n;
  K.row_ptr.assign(row_ptr, row_ptr + n + 1);
  K.cols.assign(cols, cols + row_ptr[n]);
  K.vals.assign(vals, vals + row_ptr[n]);
  return K;
}

PlaneCondition plane_of(int plane_strain) {
  return plane_strain ? PlaneCondition::Strain : PlaneCondition::Stress;
}

struct MeshHandle {
  Mesh mesh;
  std::vector<std::string> names;
};

}  // namespace

extern "C" {

const char* ref_last_error() { return g_error.c_str(); }

// element.cpp:7-15
void ref_constitutive(double E, double nu, int plane_strain, double D[9]) {
  const Mat3 d = constitutive_matrix(E, nu, plane_of(plane_strain));
  std::memcpy(D, d.data(), sizeof(double) * 9);
}

// element.cpp:37-55
void ref_stiffness(const double xy[6], const double D[9], double t, double K[36]) {
  TriCoords c;
  Mat3 d;
  std::memcpy(c.data(), xy, sizeof(double) * 6);
  std::memcpy(d.data(), D, sizeof(double) * 9);
  const Mat6 k = cst_stiffness(c, d, t);
  std::memcpy(K, k.data(), sizeof(double) * 36);
}

// assembly.cpp:9-42
int64_t ref_build_pattern(int nn, int64_t ne, const int* tri, int* row_ptr, int* cols) {
  const Mesh m = make_mesh(nn, ne, tri, std::vector<double>(2 * nn, 0.0).data());
  const CsrMatrix K = build_stiffness_pattern(m);
  std::memcpy(row_ptr, K.row_ptr.data(), sizeof(int) * K.row_ptr.size());
  if (cols) std::memcpy(cols, K.cols.data(), sizeof(int) * K.cols.size());
  return static_cast<int64_t>(K.nnz());
}

// assembly.cpp:54-74 (uniform E only: E_elem must be NULL)
int ref_assemble(int nn, int64_t ne, const int* tri, const double* xy, const double* E_elem,
                 double E, double nu, int plane_strain, double t, const int* row_ptr,
                 const int* cols, double* vals, int* bad_row, int* bad_col) {
  (void)bad_row;
  (void)bad_col;
  if (E_elem) {
    g_error = "the reference has no per-element modulus";
    return -1;
  }
  try {
    const Mesh m = make_mesh(nn, ne, tri, xy);
    CsrMatrix K = make_csr(2 * nn, row_ptr, cols, vals);
    assemble_stiffness(m, constitutive_matrix(E, nu, plane_of(plane_strain)), t, K);
    std::memcpy(vals, K.vals.data(), sizeof(double) * K.vals.size());
    return 0;
  } catch (const std::exception& e) {
    g_error = e.what();
    return -4;
  }
}

// csr.cpp:19-26
void ref_spmv(int n, const int* row_ptr, const int* cols, const double* vals, const double* x,
              double* y) {
  const CsrMatrix K = make_csr(n, row_ptr, cols, vals);
  std::vector<double> xv(x, x + n), yv;
  K.multiply(xv, yv);
  std::memcpy(y, yv.data(), sizeof(double) * n);
}

// bc.cpp:107-166 (DirichletDofs::add + apply_dirichlet)
int ref_apply_dirichlet(int n, const int* row_ptr, const int* cols, double* vals, double* b,
                        int m, const int* dofs, const double* values, int* bad) {
  (void)bad;
  try {
    CsrMatrix K = make_csr(n, row_ptr, cols, vals);
    DirichletDofs d;
    for (int k = 0; k < m; ++k) d.add(dofs[k], values[k]);
    std::vector<double> bv(b, b + n);
    apply_dirichlet(d, K, bv);
    std::memcpy(vals, K.vals.data(), sizeof(double) * K.vals.size());
    std::memcpy(b, bv.data(), sizeof(double) * n);
    return 0;
  } catch (const std::exception& e) {
    g_error = e.what();
    return -3;
  }
}

// bc.cpp:190-253 (mode 0: constant traction) and kirsch.cpp:40-45 (mode 1: Kirsch)
int ref_add_traction(int nn, int64_t ne, const int* tri, const double* xy, int nedge,
                     const int* edges, double t, int mode, double p0, double p1, double* f,
                     int* bad_edge) {
  (void)bad_edge;
  try {
    Mesh m = make_mesh(nn, ne, tri, xy);
    auto& g = m.boundary_edges["g"];
    for (int k = 0; k < nedge; ++k) g.push_back({edges[2 * k], edges[2 * k + 1]});
    std::vector<double> fv(f, f + 2 * nn);
    if (mode == 0)
      add_traction(m, "g", t, p0, p1, fv);
    else
      add_traction(m, "g", t, kirsch_traction(KirschParams{p0, p1}), fv);
    std::memcpy(f, fv.data(), sizeof(double) * 2 * nn);
    return 0;
  } catch (const std::exception& e) {
    g_error = e.what();
    return -7;
  }
}

// Stats layout shared with cst_oracle.c.
struct ref_stats {
  int iterations;
  int converged;
  int breakdown;
  int zero_rhs;
  double rel_residual;
  double true_rel_residual;
  double breakdown_pap;
};

// solver.cpp:109-157 (CgJacobiSolver::setup + solve -> pcg_jacobi)
int ref_pcg_jacobi(int n, const int* row_ptr, const int* cols, const double* vals,
                   const double* b, double* x, double rtol, int max_iter, double* history,
                   ref_stats* st) {
  std::memset(st, 0, sizeof *st);
  const CsrMatrix K = make_csr(n, row_ptr, cols, vals);
  CgJacobiSolver solver(SolverOptions{rtol, max_iter, history != nullptr});
  try {
    solver.setup(K);
  } catch (const std::exception& e) {
    g_error = e.what();
    return -3;
  }
  std::vector<double> bv(b, b + n), xv;
  const SolverStats s = solver.solve(bv, xv);
  std::memcpy(x, xv.data(), sizeof(double) * n);
  st->iterations = s.iterations;
  st->converged = s.converged ? 1 : 0;
  st->breakdown = s.message.find("breakdown") != std::string::npos ? 1 : 0;
  st->zero_rhs = s.message.find("zero right-hand side") != std::string::npos ? 1 : 0;
  st->rel_residual = s.rel_residual;
  st->true_rel_residual = s.true_rel_residual;
  if (history)
    std::memcpy(history, s.residual_history.data(), sizeof(double) * s.residual_history.size());
  g_error = s.message + (s.warning.empty() ? "" : "\n" + s.warning);
  return s.converged ? 0 : 2;
}

// postprocess.cpp:35-60
void ref_element_stresses(int64_t ne, const int* tri, const double* xy, const double* u,
                          const double* E_elem, double E, double nu, int plane_strain,
                          double* sig, double* cxy, double* area) {
  (void)E_elem;
  int nn = 0;
  for (int64_t k = 0; k < 3 * ne; ++k) nn = std::max(nn, tri[k] + 1);
  const Mesh m = make_mesh(nn, ne, tri, xy);
  const std::vector<double> uv(u, u + 2 * nn);
  const ElementStresses s =
      compute_element_stresses(m, uv, constitutive_matrix(E, nu, plane_of(plane_strain)), nu,
                               plane_of(plane_strain));
  for (int64_t e = 0; e < ne; ++e) {
    sig[3 * e] = s.sxx[e];
    sig[3 * e + 1] = s.syy[e];
    sig[3 * e + 2] = s.txy[e];
    cxy[2 * e] = s.cx[e];
    cxy[2 * e + 1] = s.cy[e];
    area[e] = s.area[e];
  }
}

// postprocess.cpp:142-150 (kirsch_metrics); same out layout as oc_kirsch_metrics
void ref_kirsch_metrics(int nn, int64_t ne, const int* tri, const double* xy, const double* u,
                        double sigma_inf, double a, double E, double nu, int plane_strain,
                        double out[8]) {
  const Mesh m = make_mesh(nn, ne, tri, xy);
  const std::vector<double> uv(u, u + 2 * nn);
  const PlaneCondition pl = plane_of(plane_strain);
  const ElementStresses s = compute_element_stresses(m, uv, constitutive_matrix(E, nu, pl), nu, pl);
  const KirschMetrics k = kirsch_metrics(m, uv, s, KirschParams{sigma_inf, a}, {E, nu, pl});
  out[0] = k.l2_rel_stress_error;
  out[1] = k.l2_rel_displacement_error;
  out[2] = k.energy_rel_error;
  out[3] = k.scf_element;
  out[4] = k.scf_x;
  out[5] = k.scf_y;
  out[6] = k.scf_numerical;
  out[7] = k.scf_analytical_centroid;
}

// mesh_io.cpp:352-356 (read_gmsh): opens a .msh file; the handle exposes nodes, triangles
// (counter-clockwise after the reader's flips) and named boundary groups.
void* ref_mesh_open(const char* path) {
  try {
    auto h = std::make_unique<MeshHandle>();
    h->mesh = read_gmsh(path);
    for (const auto& kv : h->mesh.boundary_edges) h->names.push_back(kv.first);
    return h.release();
  } catch (const std::exception& e) {
    g_error = e.what();
    return nullptr;
  }
}

void ref_mesh_sizes(void* hp, int* nn, int64_t* ne, int* ngroups) {
  const auto* h = static_cast<MeshHandle*>(hp);
  *nn = static_cast<int>(h->mesh.num_nodes());
  *ne = static_cast<int64_t>(h->mesh.num_triangles());
  *ngroups = static_cast<int>(h->names.size());
}

void ref_mesh_arrays(void* hp, double* xy, int* tri) {
  const auto* h = static_cast<MeshHandle*>(hp);
  std::memcpy(xy, h->mesh.xy.data(), sizeof(double) * h->mesh.xy.size());
  std::memcpy(tri, h->mesh.triangles.data(), sizeof(int) * h->mesh.triangles.size());
}

const char* ref_mesh_group_name(void* hp, int g) {
  return static_cast<MeshHandle*>(hp)->names[g].c_str();
}

int ref_mesh_group_edges(void* hp, int g, int* edges) {
  const auto* h = static_cast<MeshHandle*>(hp);
  const auto& es = h->mesh.boundary_edges.at(h->names[g]);
  if (edges)
    for (size_t k = 0; k < es.size(); ++k) {
      edges[2 * k] = es[k][0];
      edges[2 * k + 1] = es[k][1];
    }
  return static_cast<int>(es.size());
}

void ref_mesh_close(void* hp) { delete static_cast<MeshHandle*>(hp); }

}  // extern "C"
