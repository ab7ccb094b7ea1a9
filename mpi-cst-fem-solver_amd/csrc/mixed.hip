// mixed.hip -- FP32 inner Jacobi-PCG under FP64 iterative refinement (not yet built).
#include "internal.hpp"

namespace cstfem_hip {

int solve_ir_fp32(Context& c, const cstfem_solver_options& o, cstfem_solver_stats& st) {
  (void)c;
  (void)o;
  (void)st;
  throw Error(CSTFEM_ERR_ARGUMENT, "cstfem_solve: the FP32 refinement method is not built yet");
}

}  // namespace cstfem_hip
