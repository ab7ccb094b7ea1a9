// cstfem_hip_reference.hpp -- the MI355X path behind the reference's own C++ types.
//
// For a build of the reference (namespace cstfem, include/cstfem/*.hpp) that links
// libcstfem_hip.so: the classes below derive from the reference's abstract interfaces and take
// the reference's CsrMatrix, SolverOptions, SolverStats and SolveBreakdown as they are, so no
// second matrix type or copy sits between the reference's code and the device:
//
//   reference (file:line)                              here (namespace cstfem_hip::dropin)
//   LinearSolver / CgJacobiSolver  solver.hpp:120-146  HipCgJacobiSolver : cstfem::LinearSolver
//   LinearOperator / CsrOperator   solver.hpp:60-79    DeviceOperator : cstfem::LinearOperator
//   Reduction / SerialReduction    solver.hpp:64-91    DeviceReduction : cstfem::Reduction
//   pcg_jacobi                     solver.hpp:115-118  pcg_jacobi(const DeviceOperator&, ...)
//   cst_stiffness                  element.hpp:27      cst_stiffness(const TriCoords&, D, t)
//
// HipCgJacobiSolver::setup passes A.row_ptr / A.cols / A.vals straight to cstfem_upload_csr;
// the device holds its own copy, so unlike CgJacobiSolver (which keeps a pointer to A) the
// matrix need not outlive the solver.  DeviceOperator / DeviceReduction run one SpMV / dot on
// the device per call, for code that drives the reference's own pcg_jacobi loop; the
// pcg_jacobi overload for a DeviceOperator runs the whole solve on the device instead and is
// found by argument-dependent lookup from an unqualified call in the reference's namespace.
// Errors are the reference's exceptions and message prefixes (cstfem_hip.hpp: throw_status).
#pragma once

#include <chrono>
#include <memory>
#include <vector>

#include "cstfem/csr.hpp"
#include "cstfem/element.hpp"
#include "cstfem/solver.hpp"
#include "cstfem_hip.h"
#include "cstfem_hip.hpp"

namespace cstfem_hip {
namespace dropin {

// the reference's SolverStats from the C struct, with its messages (solver.cpp:46-130)
inline cstfem::SolverStats to_reference(const cstfem_solver_stats& s, double rtol, int max_iter,
                                         const std::vector<double>& hist) {
  const SolverStats h = make_stats(s, rtol, max_iter);
  cstfem::SolverStats st;
  st.iterations = h.iterations;
  st.rel_residual = h.rel_residual;
  st.true_rel_residual = h.true_rel_residual;
  st.converged = h.converged;
  st.message = h.message;
  st.warning = h.warning;
  if (!hist.empty())
    st.residual_history.assign(hist.begin(), hist.begin() + (s.zero_rhs ? 1 : s.iterations + 1));
  return st;
}

// timing is on when the breakdown has a clock, as in the reference (BreakdownTimer)
inline bool timed(const cstfem::SolveBreakdown* bd) { return bd && bd->clock; }

inline void add_breakdown(cstfem::SolveBreakdown* bd, const cstfem_solver_stats& s) {
  if (!timed(bd)) return;
  bd->total += s.total;
  bd->spmv_local += s.spmv_local;
  bd->dot_local += s.dot_local;
  bd->vector_ops += s.vector_ops;
  bd->spmv_calls += s.spmv_calls;
  bd->reductions += s.reductions;
}

// CgJacobiSolver (solver.hpp:131-146, solver.cpp:133-157) on the GPU.
class HipCgJacobiSolver final : public cstfem::LinearSolver {
 public:
  explicit HipCgJacobiSolver(cstfem::SolverOptions options = {}, int device = 0,
                             int method = CSTFEM_METHOD_CG_FP64)
      : options_(options), method_(method), dev_(std::make_shared<Device>(device)) {}
  void set_breakdown(cstfem::SolveBreakdown* bd) { bd_ = bd; }

  void setup(const cstfem::CsrMatrix& A) override {
    const auto t0 = std::chrono::steady_clock::now();
    check(cstfem_upload_csr(dev_->get(), A.num_rows, A.row_ptr.data(), A.cols.data(),
                            A.vals.data()));
    check(cstfem_setup(dev_->get()));
    n_ = A.num_rows;
    if (timed(bd_))
      bd_->setup += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }

  [[nodiscard]] cstfem::SolverStats solve(const std::vector<double>& b,
                                          std::vector<double>& x) override {
    if (n_ < 0) throw std::logic_error("cg_jacobi: solve() called before setup()");
    check(cstfem_set_rhs(dev_->get(), b.data(), static_cast<int>(b.size())));
    cstfem_solver_options o{};
    o.rtol = options_.rtol;
    o.max_iter = options_.max_iter;
    o.record_history = options_.record_history ? 1 : 0;
    o.instrument = timed(bd_) ? 1 : 0;
    o.method = method_;
    cstfem_solver_stats s{};
    x.assign(b.size(), 0.0);
    std::vector<double> hist(options_.record_history ? options_.max_iter + 1 : 0);
    check(cstfem_solve(dev_->get(), &o, &s, x.data(), hist.empty() ? nullptr : hist.data()));
    stats_ = to_reference(s, options_.rtol, options_.max_iter, hist);
    add_breakdown(bd_, s);
    return stats_;
  }

  const cstfem::SolverStats& stats() const override { return stats_; }

 private:
  cstfem::SolverOptions options_;
  int method_;
  std::shared_ptr<Device> dev_;
  cstfem::SolveBreakdown* bd_ = nullptr;
  cstfem::SolverStats stats_;
  int n_ = -1;
};

// CsrOperator (solver.hpp:74-82): y = A x with A held on the device.
class DeviceOperator final : public cstfem::LinearOperator {
 public:
  explicit DeviceOperator(const cstfem::CsrMatrix& A, int device = 0)
      : dev_(std::make_shared<Device>(device)), n_(A.num_rows) {
    check(cstfem_upload_csr(dev_->get(), A.num_rows, A.row_ptr.data(), A.cols.data(),
                            A.vals.data()));
  }
  void apply(const std::vector<double>& x, std::vector<double>& y) const override {
    y.resize(n_);
    check(cstfem_spmv(dev_->get(), x.data(), y.data(), n_));
  }
  cstfem_ctx* context() const { return dev_->get(); }
  int size() const { return n_; }

 private:
  std::shared_ptr<Device> dev_;
  int n_;
};

// SerialReduction (solver.hpp:84-91): dot products summed on the device of an operator.
class DeviceReduction final : public cstfem::Reduction {
 public:
  explicit DeviceReduction(const DeviceOperator& A) : A_(A) {}
  double dot(const std::vector<double>& a, const std::vector<double>& b) const override {
    double out = 0.0;
    check(cstfem_dot(A_.context(), a.data(), b.data(), A_.size(), &out));
    return out;
  }

 private:
  const DeviceOperator& A_;
};

// pcg_jacobi (solver.hpp:115-118, solver.cpp:35-131) for an operator on the device: the whole
// recurrence, the true residual and the warning run there.  inv_diag must be 1 / diag(A), as
// CgJacobiSolver::setup forms it; the device forms the same values from its copy of A.
inline cstfem::SolverStats pcg_jacobi(const DeviceOperator& A, const std::vector<double>& inv_diag,
                                      const DeviceReduction& reduction,
                                      const std::vector<double>& b, std::vector<double>& x,
                                      const cstfem::SolverOptions& options,
                                      cstfem::SolveBreakdown* timing = nullptr) {
  (void)reduction;
  if (static_cast<int>(inv_diag.size()) != A.size() || static_cast<int>(b.size()) != A.size())
    throw std::runtime_error("pcg_jacobi: size mismatch");
  check(cstfem_setup(A.context()));
  check(cstfem_set_rhs(A.context(), b.data(), static_cast<int>(b.size())));
  cstfem_solver_options o{};
  o.rtol = options.rtol;
  o.max_iter = options.max_iter;
  o.record_history = options.record_history ? 1 : 0;
  o.instrument = timed(timing) ? 1 : 0;
  o.method = CSTFEM_METHOD_CG_FP64;
  cstfem_solver_stats s{};
  x.assign(b.size(), 0.0);
  std::vector<double> hist(options.record_history ? options.max_iter + 1 : 0);
  check(cstfem_solve(A.context(), &o, &s, x.data(), hist.empty() ? nullptr : hist.data()));
  add_breakdown(timing, s);
  return to_reference(s, options.rtol, options.max_iter, hist);
}

// cst_stiffness (element.cpp:37-55) of one triangle on the device of `dev`.
inline cstfem::Mat6 cst_stiffness(Device& dev, const cstfem::TriCoords& xy, const cstfem::Mat3& D,
                                  double thickness) {
  cstfem::Mat6 K{};
  check(cstfem_element_stiffness(dev.get(), 1, xy.data(), D.data(), thickness, K.data()));
  return K;
}

}  // namespace dropin
}  // namespace cstfem_hip
