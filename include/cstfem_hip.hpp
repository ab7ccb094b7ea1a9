// cstfem_hip.hpp -- C++ host layer over the C ABI (cstfem_hip.h), mirroring the reference's
// interface for the assemble-and-solve path (namespace cstfem, include/cstfem/*.hpp):
//
//   reference (file:line)                         here
//   constitutive_matrix      element.hpp:18       cstfem_hip::constitutive_matrix
//   cst_stiffness            element.hpp:27       cstfem_hip::cst_stiffness(Device&, xy, D, t) (and
//                                                 a batch form); assembly uses it through the slot map
//   CsrMatrix                csr.hpp:14-32        cstfem_hip::CsrMatrix (same fields)
//   build_stiffness_pattern  assembly.hpp:15      cstfem_hip::build_stiffness_pattern(Device&, Mesh)
//   assemble_stiffness       assembly.hpp:22      cstfem_hip::assemble_stiffness(Device&, Mesh, D, t, K)
//   DirichletDofs            bc.hpp:28-41         cstfem_hip::DirichletDofs (same add() rule)
//   apply_dirichlet          bc.hpp:53            cstfem_hip::apply_dirichlet(Device&, dofs, K, b)
//   add_traction             bc.hpp:72-76         cstfem_hip::add_traction(Device&, edges, t, tx, ty)
//   SolverOptions/SolverStats/SolveBreakdown      same names and fields (solver.hpp:17-107)
//   LinearSolver / CgJacobiSolver solver.hpp:120-146  cstfem_hip::LinearSolver / HipCgJacobiSolver
//
// Errors are thrown as the reference throws them: std::runtime_error for bad input and
// non-positive diagonals, std::logic_error for a missing pattern entry or solve() before
// setup(), with the reference's message prefixes.  Non-convergence is reported in
// SolverStats (converged = false, message "NOT converged: ..."), never thrown.
#pragma once

#include <array>
#include <cmath>
#include <cstdio>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "cstfem_hip.h"

namespace cstfem_hip {

enum class PlaneCondition { Stress, Strain };
using Mat3 = std::array<double, 9>;

// element.cpp:7-15
inline Mat3 constitutive_matrix(double E, double nu, PlaneCondition plane) {
  if (plane == PlaneCondition::Stress) {
    const double c = E / (1.0 - nu * nu);
    return {c, c * nu, 0.0, c * nu, c, 0.0, 0.0, 0.0, c * (1.0 - nu) / 2.0};
  }
  const double c = E / ((1.0 + nu) * (1.0 - 2.0 * nu));
  return {c * (1.0 - nu), c * nu, 0.0, c * nu, c * (1.0 - nu), 0.0, 0.0, 0.0,
          c * (1.0 - 2.0 * nu) / 2.0};
}

struct PhysicalGroup {  // mesh.hpp:13-17
  int dim = 0;
  int tag = 0;
  std::string name;
};

// The reference's Mesh (mesh.hpp:20-31).  Node and element indices are 0-based.
struct Mesh {
  std::vector<double> xy;                                          // 2 per node
  std::vector<std::size_t> node_tags;                              // Gmsh tag per node (read)
  std::vector<int> triangles;                                      // 3 per element, CCW
  std::map<std::string, std::vector<std::array<int, 2>>> boundary_edges;
  std::vector<PhysicalGroup> physical_groups;                      // from $PhysicalNames
  int flipped_triangles = 0;  // triangles reoriented to counter-clockwise on read
  std::size_t num_nodes() const { return xy.size() / 2; }
  std::size_t num_triangles() const { return triangles.size() / 3; }
  bool has_physical_group(const std::string& name) const {
    for (const auto& g : physical_groups)
      if (g.name == name) return true;
    return false;
  }
};

// csr.hpp:14-32 (host copy; the device keeps its own)
struct CsrMatrix {
  int num_rows = 0;
  int num_cols = 0;
  int first_row = 0;
  std::vector<int> row_ptr;
  std::vector<int> cols;
  std::vector<double> vals;
  std::size_t nnz() const { return cols.size(); }
  int find(int row, int col) const {  // csr.cpp:7-12
    int lo = row_ptr[row], hi = row_ptr[row + 1];
    while (lo < hi) {
      const int mid = lo + (hi - lo) / 2;
      if (cols[mid] < col) lo = mid + 1; else hi = mid;
    }
    return (lo < row_ptr[row + 1] && cols[lo] == col) ? lo : -1;
  }
  double diagonal(int row) const {
    const int k = find(row, first_row + row);
    return k >= 0 ? vals[k] : 0.0;
  }
};

struct SolveBreakdown {  // solver.hpp:17-35 (device kernel-class times, seconds)
  double setup = 0.0, spmv_local = 0.0, halo_post = 0.0, halo_wait = 0.0, spmv_ghost = 0.0;
  double dot_local = 0.0, allreduce = 0.0, vector_ops = 0.0, total = 0.0;
  long long spmv_calls = 0, reductions = 0, halo_bytes_sent = 0, halo_messages_sent = 0;
  double accounted() const {
    return setup + spmv_local + halo_post + halo_wait + spmv_ghost + dot_local + allreduce +
           vector_ops;
  }
};

struct SolverOptions {  // solver.hpp:93-97
  double rtol = 1e-10;
  int max_iter = 20000;
  bool record_history = false;
};

struct SolverStats {  // solver.hpp:99-107
  int iterations = 0;
  double rel_residual = 0.0;
  double true_rel_residual = 0.0;
  bool converged = false;
  std::string message;
  std::string warning;
  std::vector<double> residual_history;
  int outer_iterations = 0;  // refinements (refined methods only)
};

[[noreturn]] inline void throw_status(int rc) {
  const std::string msg = cstfem_last_error();
  if (rc == CSTFEM_ERR_MISSING_ENTRY || rc == CSTFEM_ERR_ORDER) throw std::logic_error(msg);
  throw std::runtime_error(msg);
}

inline int check(int rc) {
  if (rc != CSTFEM_OK && rc != CSTFEM_NOT_CONVERGED) throw_status(rc);
  return rc;
}

// One GPU context (device memory, stream); movable, not copyable.
class Device {
 public:
  explicit Device(int device = 0) { check(cstfem_ctx_create(device, &ctx_)); }
  ~Device() { cstfem_ctx_destroy(ctx_); }
  Device(const Device&) = delete;
  Device& operator=(const Device&) = delete;
  cstfem_ctx* get() const { return ctx_; }

 private:
  cstfem_ctx* ctx_ = nullptr;
};

using TriCoords = std::array<double, 6>;  // element.hpp:15: (x1, y1, x2, y2, x3, y3)
using Mat6 = std::array<double, 36>;      // element.hpp:13, row-major

// cst_stiffness (element.cpp:37-55) of a batch of triangles on the device, bit for bit the
// reference's Ke; the single-triangle form below keeps the reference's signature plus the
// device.
inline std::vector<Mat6> cst_stiffness(Device& dev, const std::vector<TriCoords>& xy,
                                       const Mat3& D, double thickness) {
  std::vector<Mat6> K(xy.size());
  check(cstfem_element_stiffness(dev.get(), static_cast<int64_t>(xy.size()),
                                 xy.empty() ? nullptr : xy[0].data(), D.data(), thickness,
                                 K.empty() ? nullptr : K[0].data()));
  return K;
}

inline Mat6 cst_stiffness(Device& dev, const TriCoords& xy, const Mat3& D, double thickness) {
  return cst_stiffness(dev, std::vector<TriCoords>{xy}, D, thickness)[0];
}

// assembly.cpp:9-42 on the device; returns the host CSR with zeroed values.  The mesh,
// its pattern and the element-to-slot map stay resident on `dev`.
inline CsrMatrix build_stiffness_pattern(Device& dev, const Mesh& mesh) {
  const int nn = static_cast<int>(mesh.num_nodes());
  check(cstfem_upload_mesh(dev.get(), nn, mesh.xy.data(),
                           static_cast<int64_t>(mesh.num_triangles()), mesh.triangles.data()));
  int64_t nnz = 0;
  check(cstfem_build_pattern(dev.get(), &nnz));
  CsrMatrix K;
  K.num_rows = K.num_cols = 2 * nn;
  K.row_ptr.resize(K.num_rows + 1);
  K.cols.resize(nnz);
  check(cstfem_download_pattern(dev.get(), K.row_ptr.data(), K.cols.data()));
  K.vals.assign(nnz, 0.0);
  return K;
}

// assembly.cpp:54-74 on the device with a uniform D (as the reference), or per-element moduli
// E_elem with D_e = constitutive_matrix(E_e, nu, plane).  K.vals receives the values.
inline void assemble_stiffness(Device& dev, const Mesh& mesh, const Mat3& D, double thickness,
                               CsrMatrix& K) {
  (void)mesh;
  check(cstfem_set_constitutive(dev.get(), D.data(), thickness));
  check(cstfem_assemble(dev.get()));
  K.vals.resize(K.cols.size());
  check(cstfem_download_values(dev.get(), K.vals.data()));
}

inline void assemble_stiffness(Device& dev, const std::vector<double>& E_elem, double nu,
                               PlaneCondition plane, double thickness, CsrMatrix& K) {
  check(cstfem_set_material(dev.get(), 1.0, nu,
                            plane == PlaneCondition::Strain ? CSTFEM_PLANE_STRAIN
                                                            : CSTFEM_PLANE_STRESS,
                            thickness, E_elem.data()));
  check(cstfem_assemble(dev.get()));
  K.vals.resize(K.cols.size());
  check(cstfem_download_values(dev.get(), K.vals.data()));
}

// bc.hpp:28-41: DOF -> prescribed value; a repeated DOF must repeat its value.
class DirichletDofs {
 public:
  void add(int dof, double value) {  // bc.cpp:107-116
    const auto [it, inserted] = values_.emplace(dof, value);
    if (!inserted && it->second != value) {
      std::ostringstream msg;
      msg.precision(17);
      msg << "conflicting Dirichlet values for node " << dof / 2 << " (0-based) component "
          << (dof % 2 == 0 ? "ux" : "uy") << ": " << it->second << " vs " << value;
      throw std::runtime_error(msg.str());
    }
  }
  const std::map<int, double>& values() const { return values_; }
  std::size_t size() const { return values_.size(); }

 private:
  std::map<int, double> values_;
};

// bc.cpp:118-127: every node of every edge of a group
inline void add_dirichlet_group(const Mesh& mesh, const std::string& group, int component,
                                double value, DirichletDofs& dofs) {
  const auto it = mesh.boundary_edges.find(group);
  if (it == mesh.boundary_edges.end())
    throw std::runtime_error("unknown boundary group '" + group + "'");
  for (const auto& e : it->second)
    for (int n : e) dofs.add(2 * n + component, value);
}

// bc.cpp:190-253 for a constant traction on one group (device RHS `b` accumulates; the host
// vector f, if given, receives the device RHS afterwards)
inline void add_traction(Device& dev, const Mesh& mesh, const std::string& group, double tx,
                         double ty) {
  const auto it = mesh.boundary_edges.find(group);
  if (it == mesh.boundary_edges.end())
    throw std::runtime_error("unknown boundary group '" + group + "'");
  std::vector<int> e;
  for (const auto& ed : it->second) {
    e.push_back(ed[0]);
    e.push_back(ed[1]);
  }
  check(cstfem_add_traction(dev.get(), static_cast<int>(it->second.size()), e.data(), 0, tx, ty));
}

// bc.cpp:137-166 on the device matrix and RHS; K.vals and b are refreshed from the device.
inline void apply_dirichlet(Device& dev, const DirichletDofs& dofs, CsrMatrix& K,
                            std::vector<double>& b) {
  std::vector<int> d;
  std::vector<double> v;
  for (const auto& [dof, val] : dofs.values()) {
    d.push_back(dof);
    v.push_back(val);
  }
  check(cstfem_set_rhs(dev.get(), b.data(), static_cast<int>(b.size())));
  check(cstfem_apply_dirichlet(dev.get(), static_cast<int>(d.size()), d.data(), v.data()));
  K.vals.resize(K.cols.size());
  check(cstfem_download_values(dev.get(), K.vals.data()));
  check(cstfem_download_rhs(dev.get(), b.data(), static_cast<int>(b.size())));
}

// SolverStats with the reference's messages (solver.cpp:46-104, 123-129) from the C struct
inline SolverStats make_stats(const cstfem_solver_stats& s, double rtol, int max_iter) {
  SolverStats st;
  st.iterations = s.iterations;
  st.rel_residual = s.rel_residual;
  st.true_rel_residual = s.true_rel_residual;
  st.converged = s.converged != 0;
  st.outer_iterations = s.outer_iterations;
  std::ostringstream msg;
  if (s.zero_rhs)
    msg << "zero right-hand side, x = 0";
  else if (s.breakdown)
    msg << "CG breakdown at iteration " << s.iterations + 1 << ": p'Ap = " << s.breakdown_pap
        << " (matrix not SPD?)";
  else if (s.converged)
    msg << "converged in " << s.iterations << " iterations, relative residual " << s.rel_residual;
  else
    msg << "NOT converged: reached max_iter = " << max_iter << ", relative residual "
        << s.rel_residual << " (rtol " << rtol << ")";
  st.message = msg.str();
  if (s.warning) {
    std::ostringstream w;
    w << "WARNING: CG recursive residual converged (" << s.rel_residual
      << ") but the true residual ||b - Ax|| / ||b|| = " << s.true_rel_residual
      << " exceeds 10 * rtol = " << 10.0 * rtol;
    st.warning = w.str();
  }
  return st;
}

// solver.hpp:120-128
class LinearSolver {
 public:
  virtual ~LinearSolver() = default;
  virtual void setup(const CsrMatrix& A) = 0;
  [[nodiscard]] virtual SolverStats solve(const std::vector<double>& b,
                                          std::vector<double>& x) = 0;
  virtual const SolverStats& stats() const = 0;
};

// CgJacobiSolver (solver.hpp:131-146) on the GPU.  setup() copies A to the device (values
// and pattern; A need not outlive the solver) and forms D^-1; solve() runs the fused
// Jacobi-PCG and recomputes the true residual.  `method` selects the reference recurrence
// (default) or a refined variant.
class HipCgJacobiSolver : public LinearSolver {
 public:
  explicit HipCgJacobiSolver(SolverOptions options = {}, int device = 0,
                             int method = CSTFEM_METHOD_CG_FP64)
      : options_(options), method_(method), dev_(std::make_shared<Device>(device)) {}
  void set_breakdown(SolveBreakdown* bd) { bd_ = bd; }

  void setup(const CsrMatrix& A) override {
    check(cstfem_upload_csr(dev_->get(), A.num_rows, A.row_ptr.data(), A.cols.data(),
                            A.vals.data()));
    check(cstfem_setup(dev_->get()));
    n_ = A.num_rows;
  }

  [[nodiscard]] SolverStats solve(const std::vector<double>& b, std::vector<double>& x) override {
    if (n_ < 0) throw std::logic_error("cg_jacobi: solve() called before setup()");
    check(cstfem_set_rhs(dev_->get(), b.data(), static_cast<int>(b.size())));
    cstfem_solver_options o{};
    o.rtol = options_.rtol;
    o.max_iter = options_.max_iter;
    o.record_history = options_.record_history ? 1 : 0;
    o.instrument = bd_ ? 1 : 0;
    o.method = method_;
    cstfem_solver_stats s{};
    x.assign(b.size(), 0.0);
    std::vector<double> hist(options_.record_history ? options_.max_iter + 1 : 0);
    check(cstfem_solve(dev_->get(), &o, &s, x.data(), hist.empty() ? nullptr : hist.data()));
    stats_ = make_stats(s, options_.rtol, options_.max_iter);
    if (options_.record_history)
      stats_.residual_history.assign(hist.begin(),
                                     hist.begin() + (s.zero_rhs ? 1 : s.iterations + 1));
    if (bd_) {
      bd_->spmv_local += s.spmv_local;
      bd_->dot_local += s.dot_local;
      bd_->vector_ops += s.vector_ops;
      bd_->total += s.total;
      bd_->spmv_calls += s.spmv_calls;
      bd_->reductions += s.reductions;
    }
    return stats_;
  }

  const SolverStats& stats() const override { return stats_; }
  Device& device() { return *dev_; }

 private:
  SolverOptions options_;
  int method_;
  std::shared_ptr<Device> dev_;
  SolveBreakdown* bd_ = nullptr;
  SolverStats stats_;
  int n_ = -1;
};

}  // namespace cstfem_hip
