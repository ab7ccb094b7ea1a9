// dirichlet.hip -- Dirichlet elimination and body-force load on the device.
//
// Reference: src/bc.cpp:137-166 (apply_dirichlet).  One thread owns a block row (the two
// scalar rows 2n, 2n+1) and visits its blocks in ascending column order, so the
// subtractions b_i -= K_ik u_k happen in the reference's column order and every value of
// K and b equals the reference's.  The pattern is kept (eliminated entries stay stored as
// zeros).
#include <climits>

#include "internal.hpp"

namespace cstfem_hip {

namespace {

__global__ void k_mark_fixed(int m, const int* __restrict__ dofs, const double* __restrict__ val,
                             unsigned char* __restrict__ fixed, double* __restrict__ ubar,
                             unsigned char* __restrict__ all_fixed) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  fixed[dofs[k]] = 1;
  all_fixed[dofs[k]] = 1;
  ubar[dofs[k]] = val[k];
}

__global__ void k_dirichlet(int nb, int n, const int* __restrict__ brow_ptr,
                            const int* __restrict__ bcol, const int* __restrict__ bdiag,
                            const int64_t* __restrict__ spos, double* __restrict__ val,
                            const unsigned char* __restrict__ fixed,
                            const double* __restrict__ ubar, double* __restrict__ b,
                            int* __restrict__ err) {
  const int nd = blockIdx.x * blockDim.x + threadIdx.x;
  if (nd >= nb) return;
  const int q0 = brow_ptr[nd], q1 = brow_ptr[nd + 1];
  for (int c = 0; c < 2; ++c) {
    const int i = 2 * nd + c;
    if (i >= n) break;
    if (fixed[i]) {
      const int qd = bdiag[nd];
      const double diag = qd >= 0 ? val[sell_vidx(spos[qd], 3 * c)] : 0.0;
      if (!(diag > 0.0)) {
        atomicMin(err, i);
        continue;
      }
      for (int q = q0; q < q1; ++q) {
        double* v = val + sell_vidx(spos[q], 2 * c);  // row c: v[0], v[1]
        const int m = bcol[q];
        if (2 * m != i) v[0] = 0.0;
        if (2 * m + 1 != i) v[1] = 0.0;
      }
      b[i] = diag * ubar[i];
    } else {
      double bi = b[i];
      for (int q = q0; q < q1; ++q) {
        const int m = bcol[q];
        double* v = val + sell_vidx(spos[q], 2 * c);  // row c: v[0], v[1]
#pragma unroll
        for (int d = 0; d < 2; ++d) {
          const int j = 2 * m + d;
          if (j < n && fixed[j]) {
            bi -= v[d] * ubar[j];
            v[d] = 0.0;
          }
        }
      }
      b[i] = bi;
    }
  }
}

// t*|A|/3*(bx, by) to each node of every element, per node in ascending element order
__global__ void k_body_force(int nn, const int* __restrict__ ptr, const int* __restrict__ list,
                             const int* __restrict__ tri, const double* __restrict__ xy,
                             double t, double bx, double by, double* __restrict__ f) {
  const int nd = blockIdx.x * blockDim.x + threadIdx.x;
  if (nd >= nn) return;
  double fx = f[2 * nd], fy = f[2 * nd + 1];
  for (int k = ptr[nd]; k < ptr[nd + 1]; ++k) {
    const int e = list[k];
    const int n0 = tri[3 * e], n1 = tri[3 * e + 1], n2 = tri[3 * e + 2];
    const double x1 = xy[2 * n0], y1 = xy[2 * n0 + 1], x2 = xy[2 * n1], y2 = xy[2 * n1 + 1];
    const double x3 = xy[2 * n2], y3 = xy[2 * n2 + 1];
    const double area = 0.5 * ((x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1));
    const double w = t * fabs(area) / 3.0;
    fx += w * bx;
    fy += w * by;
  }
  f[2 * nd] = fx;
  f[2 * nd + 1] = fy;
}

}  // namespace

void apply_dirichlet_device(Context& c, int m, const int* dofs, const double* values) {
  if (!c.have_pattern) throw Error(CSTFEM_ERR_ORDER, "apply_dirichlet: no matrix on the device");
  if (c.b.n == 0) throw Error(CSTFEM_ERR_ORDER, "apply_dirichlet: set the right-hand side first");
  const int n = c.n;
  // DirichletDofs::add (bc.cpp:107-116): a repeated DOF must repeat its value
  std::vector<std::pair<int, double>> seen;
  seen.reserve(m);
  for (int k = 0; k < m; ++k) {
    if (dofs[k] < 0 || dofs[k] >= n)
      throw Error(CSTFEM_ERR_OUT_OF_RANGE,
                  "apply_dirichlet: DOF " + std::to_string(dofs[k]) + " out of range");
    seen.emplace_back(dofs[k], values[k]);
  }
  std::stable_sort(seen.begin(), seen.end(),
                   [](const auto& a, const auto& b) { return a.first < b.first; });
  for (size_t k = 1; k < seen.size(); ++k)
    if (seen[k].first == seen[k - 1].first && seen[k].second != seen[k - 1].second) {
      char buf[256];
      const int dof = seen[k].first;
      snprintf(buf, sizeof buf,
               "conflicting Dirichlet values for node %d (0-based) component %s: %.17g vs %.17g",
               dof / 2, dof % 2 == 0 ? "ux" : "uy", seen[k - 1].second, seen[k].second);
      throw Error(CSTFEM_ERR_CONFLICT, buf);
    }
  DBuf<int> d_dofs;
  DBuf<double> d_vals, ubar;
  DBuf<unsigned char> fixed;
  fixed.alloc(n);
  ubar.alloc(n);
  HIP_CHECK(hipMemsetAsync(fixed.get(), 0, n, c.stream));
  HIP_CHECK(hipMemsetAsync(ubar.get(), 0, sizeof(double) * n, c.stream));
  if (c.fixed.n != (size_t)n) {  // union of all constrained DOFs since the last assembly
    c.fixed.alloc(n);
    HIP_CHECK(hipMemsetAsync(c.fixed.get(), 0, n, c.stream));
  }
  if (m > 0) {
    d_dofs.alloc(m);
    d_vals.alloc(m);
    HIP_CHECK(hipMemcpyAsync(d_dofs.get(), dofs, sizeof(int) * m, hipMemcpyHostToDevice, c.stream));
    HIP_CHECK(hipMemcpyAsync(d_vals.get(), values, sizeof(double) * m, hipMemcpyHostToDevice, c.stream));
    hipLaunchKernelGGL(k_mark_fixed, dim3(grid_for(m, 256, 0)), dim3(256), 0, c.stream, m,
                       d_dofs.get(), d_vals.get(), fixed.get(), ubar.get(), c.fixed.get());
    HIP_CHECK(hipGetLastError());
  }
  const int big = INT_MAX;
  HIP_CHECK(hipMemcpyAsync(c.err_flag.get(), &big, sizeof(int), hipMemcpyHostToDevice, c.stream));
  hipLaunchKernelGGL(k_dirichlet, dim3(grid_for(c.nb, 128, 0)), dim3(128), 0, c.stream, c.nb, n,
                     c.brow_ptr.get(), c.bcol.get(), c.bdiag.get(), c.sell_pos.get(),
                     c.sell_val.get(), fixed.get(), ubar.get(), c.b.get(), c.err_flag.get());
  HIP_CHECK(hipGetLastError());
  int bad = INT_MAX;
  HIP_CHECK(hipMemcpyAsync(&bad, c.err_flag.get(), sizeof(int), hipMemcpyDeviceToHost, c.stream));
  HIP_CHECK(hipStreamSynchronize(c.stream));
  c.setup_done = false;
  c.vals32_valid = false;
  if (bad != INT_MAX)
    throw Error(CSTFEM_ERR_BAD_DIAGONAL,
                "apply_dirichlet: non-positive diagonal at DOF " + std::to_string(bad));
}

void add_body_force_device(Context& c, double bx, double by) {
  if (c.ne_ptr.n == 0) build_node_elements(c);
  if (c.b.n == 0) throw Error(CSTFEM_ERR_ORDER, "add_body_force: set the right-hand side first");
  hipLaunchKernelGGL(k_body_force, dim3(grid_for(c.nn, 256, 0)), dim3(256), 0, c.stream, c.nn,
                     c.ne_ptr.get(), c.ne_list.get(), c.tri.get(), c.xy.get(), c.thick, bx, by,
                     c.b.get());
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(c.stream));
}

}  // namespace cstfem_hip
