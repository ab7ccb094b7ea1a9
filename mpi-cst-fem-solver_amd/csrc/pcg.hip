// pcg.hip -- Jacobi-preconditioned CG on the device: the reference's FP64 recurrence and an
// FP32 variant under FP64 refinement.
//
// Reference: src/solver.cpp:35-157 (pcg_iterate, pcg_jacobi, CgJacobiSolver::setup/solve).
// The recurrence, the stopping test (rel < rtol), the breakdown test !(p'Ap > 0), the
// zero-RHS exit and the true residual with its 10*rtol warning are the reference's.
//
// One iteration = three kernels, each one pass over HBM:
//   A  Ap = A p (SELL-64 2x2 blocks), per-workgroup partials of p'Ap
//   B  every workgroup reduces A's partials (fixed order) -> alpha = rz/p'Ap, breakdown test;
//      x += alpha p, r -= alpha Ap; per-workgroup partials of r'r and r'(D^-1 r)
//   C  every workgroup reduces B's partials -> convergence test, beta; p = D^-1 r + beta p
// No kernel needs a cross-workgroup fence (an agent-scope release writes back the whole L2
// of an XCD on gfx950 and costs more than the HBM traffic): scalars are recomputed
// redundantly, bit-identically, by each workgroup of the consumer kernel, and the kernel
// boundary provides visibility.  B's partials live in two slots selected by iteration
// parity, so C can read rz(k-1) and rz(k) without a race; graphs hold an even number of
// iterations.  Element-wise arithmetic and each SpMV row sum follow the reference's operation
// order (no FMA contraction); only dot products are summed in a different (fixed) order, so
// results are run-to-run reproducible.
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <optional>
#include <type_traits>

#include "internal.hpp"

namespace cstfem_hip {

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kMaxGrid = 1024;  // partial slots per kernel

typedef double v2d __attribute__((ext_vector_type(2)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Workgroup sum, broadcast to every thread.
__device__ inline double block_sum_all(double v, double* lds) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) lds[w] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < kWaves; ++i) s += lds[i];
  return s;
}

// Fixed-order sum of `count` partials (stride `stride`); identical in every workgroup.
__device__ inline double reduce_partials(const double* __restrict__ part, int count, int stride,
                                         double* lds) {
  double s = 0.0;
  for (int i = threadIdx.x; i < count; i += kThreads) s += part[(size_t)i * stride];
  return block_sum_all(s, lds);
}

// Maps workgroup b to its share of `nslices` so that the 8 XCDs (which receive workgroups
// round-robin) each sweep one contiguous range of slices: neighbouring x entries then stay
// in one XCD's L2.
struct SliceRange {
  int first, end, step;
};
__device__ inline SliceRange slice_range(int nslices) {
  const int G = gridDim.x;
  const int wave = threadIdx.x >> 6;
  if (G >= 64 && (G & 7) == 0) {
    const int xcd = blockIdx.x & 7, idx = blockIdx.x >> 3, per = G >> 3;
    const int lo = (int)((int64_t)nslices * xcd / 8), hi = (int)((int64_t)nslices * (xcd + 1) / 8);
    return {lo + idx * kWaves + wave, hi, per * kWaves};
  }
  return {(int)blockIdx.x * kWaves + wave, nslices, G * kWaves};
}

// ---- precision traits ------------------------------------------------------------------

struct F64 {
  using V = double2;   // 2 DOFs of one node
  using S = double;
};
struct F32 {
  using V = float2;
  using S = float;
};

// Rows 2r, 2r+1 of a SELL slice times x: sequential per-row sums in column order.
__device__ inline void sell_row(const int64_t* __restrict__ off, const int* __restrict__ col,
                                const double* __restrict__ val, const double2* __restrict__ x,
                                int s, int lane, double& y0, double& y1) {
  const int64_t o = off[s];
  const int w = (int)((off[s + 1] - o) >> 6);
  const int* c = col + o + lane;
  const v2d* v = reinterpret_cast<const v2d*>(val + 4 * o) + lane;
  double a0 = 0.0, a1 = 0.0;
#pragma unroll 4
  for (int k = 0; k < w; ++k) {
    const int m = c[(size_t)k * 64];
    const v2d va = __builtin_nontemporal_load(v + (size_t)k * 128);
    const v2d vb = __builtin_nontemporal_load(v + (size_t)k * 128 + 64);
    const double2 xm = x[m];
    a0 = a0 + va.x * xm.x;
    a0 = a0 + va.y * xm.y;
    a1 = a1 + vb.x * xm.x;
    a1 = a1 + vb.y * xm.y;
  }
  y0 = a0;
  y1 = a1;
}

__device__ inline void sell_row(const int64_t* __restrict__ off, const int* __restrict__ col,
                                const float* __restrict__ val, const float2* __restrict__ x,
                                int s, int lane, float& y0, float& y1) {
  const int64_t o = off[s];
  const int w = (int)((off[s + 1] - o) >> 6);
  const int* c = col + o + lane;
  const v4f* v = reinterpret_cast<const v4f*>(val) + (o + lane);
  float a0 = 0.f, a1 = 0.f;
#pragma unroll 4
  for (int k = 0; k < w; ++k) {
    const int m = c[(size_t)k * 64];
    const v4f vv = __builtin_nontemporal_load(v + (size_t)k * 64);
    const float2 xm = x[m];
    a0 = a0 + vv.x * xm.x;
    a0 = a0 + vv.y * xm.y;
    a1 = a1 + vv.z * xm.x;
    a1 = a1 + vv.w * xm.y;
  }
  y0 = a0;
  y1 = a1;
}

template <class P>
struct Mat {
  const int64_t* off;
  const int* col;
  const typename P::S* val;
  const int* row0;  // first block row of each slice
  const int* nrow;  // block rows in each slice (<= 64)
  int nslices, nb;
};

template <class P>
__global__ __launch_bounds__(kThreads) void k_spmv(Mat<P> A, const typename P::V* __restrict__ x,
                                                   typename P::V* __restrict__ y) {
  using S = typename P::S;
  const int lane = threadIdx.x & 63;
  const SliceRange sr = slice_range(A.nslices);
  for (int s = sr.first; s < sr.end; s += sr.step) {
    const int r = A.row0[s] + lane;
    S y0, y1;
    sell_row(A.off, A.col, A.val, x, s, lane, y0, y1);
    if (lane < A.nrow[s]) y[r] = {y0, y1};
  }
}

// A: Ap = A p; partials of p'Ap (accumulated in FP64)
template <class P>
__global__ __launch_bounds__(kThreads) void k_cg_spmv(Mat<P> A, const typename P::V* __restrict__ p,
                                                      typename P::V* __restrict__ Ap,
                                                      double* __restrict__ partA,
                                                      const PcgScalars* __restrict__ sc) {
  using S = typename P::S;
  __shared__ double lds[kWaves];
  if (sc->status != kRunning) return;
  const int lane = threadIdx.x & 63;
  const SliceRange sr = slice_range(A.nslices);
  double local = 0.0;
  for (int s = sr.first; s < sr.end; s += sr.step) {
    const int r = A.row0[s] + lane;
    S y0, y1;
    sell_row(A.off, A.col, A.val, p, s, lane, y0, y1);
    if (lane < A.nrow[s]) {
      Ap[r] = {y0, y1};
      const typename P::V pr = p[r];
      local += (double)pr.x * (double)y0 + (double)pr.y * (double)y1;
    }
  }
  const double bs = block_sum_all(local, lds);
  if (threadIdx.x == 0) partA[blockIdx.x] = bs;
}

// B (FP64): alpha, breakdown; x += alpha p, r -= alpha Ap; partials (r'r, r'z)
__global__ __launch_bounds__(kThreads) void k_cg_update(
    int nvec2, double2* __restrict__ x, double2* __restrict__ r, const double2* __restrict__ p,
    const double2* __restrict__ Ap, const double2* __restrict__ invd,
    const double* __restrict__ partA, int nA, const double* __restrict__ partPrev,
    double* __restrict__ partCur, int nB, PcgScalars* __restrict__ sc) {
  __shared__ double lds[kWaves];
  if (sc->status != kRunning) return;
  const double pAp = reduce_partials(partA, nA, 1, lds);
  const double rz = reduce_partials(partPrev + 1, nB, 2, lds);
  if (!(pAp > 0.0)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc->pAp = pAp;
      sc->status = kBreakdown;
    }
    return;
  }
  const double alpha = rz / pAp;
  double rr = 0.0, rzn = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < nvec2; i += gridDim.x * kThreads) {
    const double2 pi = p[i], api = Ap[i], di = invd[i];
    double2 xi = x[i], ri = r[i];
    xi.x = xi.x + alpha * pi.x;
    xi.y = xi.y + alpha * pi.y;
    ri.x = ri.x - alpha * api.x;
    ri.y = ri.y - alpha * api.y;
    x[i] = xi;
    r[i] = ri;
    rr += ri.x * ri.x + ri.y * ri.y;
    rzn += ri.x * (di.x * ri.x) + ri.y * (di.y * ri.y);
  }
  const double b_rr = block_sum_all(rr, lds);
  const double b_rz = block_sum_all(rzn, lds);
  if (threadIdx.x == 0) {
    partCur[2 * blockIdx.x] = b_rr;
    partCur[2 * blockIdx.x + 1] = b_rz;
  }
}

// B (FP32 inner, FP64 solution): x64 += alpha p32; r32 -= alpha Ap32 (rounded once)
__global__ __launch_bounds__(kThreads) void k_cg_update32(
    int nvec2, double2* __restrict__ x, float2* __restrict__ r, const float2* __restrict__ p,
    const float2* __restrict__ Ap, const float2* __restrict__ invd,
    const double* __restrict__ partA, int nA, const double* __restrict__ partPrev,
    double* __restrict__ partCur, int nB, PcgScalars* __restrict__ sc) {
  __shared__ double lds[kWaves];
  if (sc->status != kRunning) return;
  const double pAp = reduce_partials(partA, nA, 1, lds);
  const double rz = reduce_partials(partPrev + 1, nB, 2, lds);
  if (!(pAp > 0.0)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc->pAp = pAp;
      sc->status = kBreakdown;
    }
    return;
  }
  const double alpha = rz / pAp;
  double rr = 0.0, rzn = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < nvec2; i += gridDim.x * kThreads) {
    const float2 pi = p[i], api = Ap[i], di = invd[i];
    double2 xi = x[i];
    float2 ri = r[i];
    xi.x = xi.x + alpha * (double)pi.x;
    xi.y = xi.y + alpha * (double)pi.y;
    ri.x = (float)((double)ri.x - alpha * (double)api.x);
    ri.y = (float)((double)ri.y - alpha * (double)api.y);
    x[i] = xi;
    r[i] = ri;
    const double r0 = ri.x, r1 = ri.y;
    rr += r0 * r0 + r1 * r1;
    rzn += r0 * ((double)di.x * r0) + r1 * ((double)di.y * r1);
  }
  const double b_rr = block_sum_all(rr, lds);
  const double b_rz = block_sum_all(rzn, lds);
  if (threadIdx.x == 0) {
    partCur[2 * blockIdx.x] = b_rr;
    partCur[2 * blockIdx.x + 1] = b_rz;
  }
}

// C: convergence test, iteration count, beta; p = D^-1 r + beta p.  `cur` = iteration parity.
// FP32 mode: reaching rtol (or dropping by `sc->delta` since the last refinement) stops the
// chunk with kRefine so that the host refreshes r from the FP64 true residual.
template <class P>
__global__ __launch_bounds__(kThreads) void k_cg_direction(
    int nvec2, const typename P::V* __restrict__ r, const typename P::V* __restrict__ invd,
    typename P::V* __restrict__ p, const double* __restrict__ partPrev,
    const double* __restrict__ partCur, int nB, int cur, PcgScalars* __restrict__ sc,
    double* __restrict__ history, int test) {
  using S = typename P::S;
  __shared__ double lds[kWaves];
  if (sc->status != kRunning) return;
  const double rr = reduce_partials(partCur, nB, 2, lds);
  const double rz_new = reduce_partials(partCur + 1, nB, 2, lds);
  const double rz_old = reduce_partials(partPrev + 1, nB, 2, lds);
  const int k = sc->k_in[cur ^ 1] + 1;
  const double rel = sqrt(rr) / sc->bnorm;
  int status = kRunning;
  if (test != 1)
    status = kRunning;  // direction update after a refinement: the test already ran
  else if (rel < sc->rtol)
    status = sc->refine_mode ? kRefine : kConverged;
  else if (sc->refine_mode && rel < sc->delta * sc->rel_ref)
    status = kRefine;
  else if (k >= sc->max_iter)
    status = kMaxIter;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->k_in[cur] = k;
    sc->rel = rel;
    if (sc->record) history[k] = rel;
    if (status != kRunning) sc->status = status;
  }
  if (status != kRunning) return;
  const double beta = test == 2 ? 0.0 : rz_new / rz_old;  // 2: restart after a refinement
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < nvec2; i += gridDim.x * kThreads) {
    const typename P::V ri = r[i], di = invd[i];
    typename P::V pi = p[i];
    const S z0 = di.x * ri.x, z1 = di.y * ri.y;
    pi.x = z0 + (S)beta * pi.x;
    pi.y = z1 + (S)beta * pi.y;
    p[i] = pi;
  }
}

// x = 0, r = b, p = D^-1 b; partials (b'b, r'z) into slot 1 (the "previous" slot of k = 0)
__global__ __launch_bounds__(kThreads) void k_cg_init(int nvec2, const double2* __restrict__ b,
                                                      double2* __restrict__ x,
                                                      double2* __restrict__ r,
                                                      double2* __restrict__ p,
                                                      const double2* __restrict__ invd,
                                                      double* __restrict__ part) {
  __shared__ double lds[kWaves];
  double bb = 0.0, rz = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < nvec2; i += gridDim.x * kThreads) {
    const double2 bi = b[i], di = invd[i];
    x[i] = make_double2(0.0, 0.0);
    r[i] = bi;
    const double z0 = di.x * bi.x, z1 = di.y * bi.y;
    p[i] = make_double2(z0, z1);
    bb += bi.x * bi.x + bi.y * bi.y;
    rz += bi.x * z0 + bi.y * z1;
  }
  const double s0 = block_sum_all(bb, lds);
  const double s1 = block_sum_all(rz, lds);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s0;
    part[2 * blockIdx.x + 1] = s1;
  }
}

// one workgroup: bnorm, zero-RHS exit, counters
__global__ __launch_bounds__(kThreads) void k_cg_init_finish(const double* __restrict__ part,
                                                             int nB, PcgScalars* __restrict__ sc,
                                                             double* __restrict__ history) {
  __shared__ double lds[kWaves];
  const double bb = reduce_partials(part, nB, 2, lds);
  if (threadIdx.x == 0) {
    sc->bnorm = sqrt(bb);
    sc->k_in[0] = sc->k_in[1] = 0;
    sc->rel = 1.0;
    if (sc->bnorm == 0.0) {
      sc->status = kConverged;
      if (sc->record) history[0] = 0.0;
    } else {
      sc->status = kRunning;
      if (sc->record) history[0] = 1.0;
    }
  }
}

// r = b - y and partials of r'r (true residual of pcg_jacobi, solver.cpp:117-122)
__global__ __launch_bounds__(kThreads) void k_residual(int nvec2, const double2* __restrict__ b,
                                                       const double2* __restrict__ y,
                                                       double2* __restrict__ r,
                                                       double* __restrict__ part) {
  __shared__ double lds[kWaves];
  double rr = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < nvec2; i += gridDim.x * kThreads) {
    const double2 bi = b[i], yi = y[i];
    const double2 ri = make_double2(bi.x - yi.x, bi.y - yi.y);
    r[i] = ri;
    rr += ri.x * ri.x + ri.y * ri.y;
  }
  const double s = block_sum_all(rr, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(kThreads) void k_sum_into(const double* __restrict__ part, int n,
                                                       double* __restrict__ out) {
  __shared__ double lds[kWaves];
  const double s = reduce_partials(part, n, 1, lds);
  if (threadIdx.x == 0) *out = s;
}

// Partials of r64'p (the corrective step of a refinement needs it).
template <class P>
__global__ __launch_bounds__(kThreads) void k_dot_rp(int nvec2, const double2* __restrict__ r,
                                                     const typename P::V* __restrict__ p,
                                                     double* __restrict__ part) {
  __shared__ double lds[kWaves];
  double s = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < nvec2; i += gridDim.x * kThreads) {
    const double2 ri = r[i];
    const typename P::V pi = p[i];
    s += ri.x * (double)pi.x + ri.y * (double)pi.y;
  }
  const double t = block_sum_all(s, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// Refinement with r64 = b - A x (FP64).  correct = 1: corrective line search along the kept
// direction p, gamma = (r'p)/(p'Ap), x += gamma p, r -= gamma Ap, which restores r'p = 0
// (broken by the replacement).  Then the working residual (FP32 copy in FP32 mode) and the
// partials (r'r, r'z) of the refreshed residual into `part`.
template <class P>
__global__ __launch_bounds__(kThreads) void k_refresh(
    int nvec2, double2* __restrict__ x, double2* __restrict__ r64,
    const typename P::V* __restrict__ p, const typename P::V* __restrict__ Ap,
    const typename P::V* __restrict__ invd, float2* __restrict__ r32,
    const double* __restrict__ partA, int nA, const double* __restrict__ partRP, int nRP,
    int correct, double* __restrict__ part) {
  __shared__ double lds[kWaves];
  double gamma = 0.0;
  if (correct) {
    const double pAp = reduce_partials(partA, nA, 1, lds);
    const double rp = reduce_partials(partRP, nRP, 1, lds);
    gamma = pAp > 0.0 ? rp / pAp : 0.0;
  }
  double rr = 0.0, rz = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < nvec2; i += gridDim.x * kThreads) {
    double2 ri = r64[i];
    const typename P::V di = invd[i];
    if (correct) {
      const typename P::V pi = p[i], api = Ap[i];
      double2 xi = x[i];
      xi.x = xi.x + gamma * (double)pi.x;
      xi.y = xi.y + gamma * (double)pi.y;
      x[i] = xi;
      ri.x = ri.x - gamma * (double)api.x;
      ri.y = ri.y - gamma * (double)api.y;
      r64[i] = ri;
    }
    double r0 = ri.x, r1 = ri.y;
    if (r32) {
      const float2 f = make_float2((float)ri.x, (float)ri.y);
      r32[i] = f;
      r0 = f.x;
      r1 = f.y;
    }
    rr += r0 * r0 + r1 * r1;
    rz += r0 * ((double)di.x * r0) + r1 * ((double)di.y * r1);
  }
  const double s0 = block_sum_all(rr, lds);
  const double s1 = block_sum_all(rz, lds);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s0;
    part[2 * blockIdx.x + 1] = s1;
  }
}

// FP32 start: r32 = b, p32 = D^-1 r32, x64 = 0; partials (b'b as FP64, r32'z32) into `part`.
// Warm start (r64 != NULL): r32 = r64 = b - A x for the given x, which is kept.
__global__ __launch_bounds__(kThreads) void k_init32(int nvec2, const double2* __restrict__ b,
                                                     double2* __restrict__ x,
                                                     float2* __restrict__ r32,
                                                     float2* __restrict__ p32,
                                                     const float2* __restrict__ invd,
                                                     double* __restrict__ part,
                                                     const double2* __restrict__ r64 = nullptr) {
  __shared__ double lds[kWaves];
  double bb = 0.0, rz = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < nvec2; i += gridDim.x * kThreads) {
    const double2 bi = b[i];
    const double2 rs = r64 ? r64[i] : bi;
    const float2 ri = make_float2((float)rs.x, (float)rs.y), di = invd[i];
    if (!r64) x[i] = make_double2(0.0, 0.0);
    r32[i] = ri;
    p32[i] = make_float2(di.x * ri.x, di.y * ri.y);
    bb += bi.x * bi.x + bi.y * bi.y;
    rz += (double)ri.x * ((double)di.x * ri.x) + (double)ri.y * ((double)di.y * ri.y);
  }
  const double s0 = block_sum_all(bb, lds);
  const double s1 = block_sum_all(rz, lds);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s0;
    part[2 * blockIdx.x + 1] = s1;
  }
}

// inv_diag = 1 / K_ii (CgJacobiSolver::setup, solver.cpp:133-148), plus the FP32 copies
__global__ void k_inv_diag(int nb, int n, const int* __restrict__ bdiag,
                           const int64_t* __restrict__ spos, const double* __restrict__ val,
                           double* __restrict__ invd, float* __restrict__ invd32,
                           int* __restrict__ err) {
  const int nd = blockIdx.x * blockDim.x + threadIdx.x;
  if (nd >= nb) return;
  const int q = bdiag[nd];
  for (int c = 0; c < 2; ++c) {
    const int i = 2 * nd + c;
    double v = 0.0;
    if (i < n) {
      const double d = q >= 0 ? val[sell_vidx(spos[q], 3 * c)] : 0.0;
      if (!(d > 0.0))
        atomicMin(err, i);
      else
        v = 1.0 / d;
    }
    invd[i] = v;  // 0 on the padding row of an odd-sized generic matrix
    invd32[i] = (float)v;
  }
}

// FP32 copy of the FP64 values: planar FP64 (sell_vidx) -> 4 consecutive floats per block
__global__ void k_round_values(int64_t nblocks, const double* __restrict__ a, float* __restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nblocks;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double2 r0 = *reinterpret_cast<const double2*>(a + sell_vidx(i, 0));
    const double2 r1 = *reinterpret_cast<const double2*>(a + sell_vidx(i, 2));
    reinterpret_cast<float4*>(b)[i] = make_float4((float)r0.x, (float)r0.y, (float)r1.x, (float)r1.y);
  }
}

__global__ __launch_bounds__(kThreads) void k_copy(int64_t n2, const double2* __restrict__ a,
                                                   double2* __restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n2;
       i += (int64_t)gridDim.x * kThreads)
    reinterpret_cast<v2d*>(b)[i] = __builtin_nontemporal_load(reinterpret_cast<const v2d*>(a) + i);
}

constexpr int kReadGrid = 2048;  // workgroups of k_read

// Reads n2 double2 values once (non-temporal loads) and writes one sum per wavefront: the
// read-only streaming rate, the yardstick of kernels that read far more than they write
__global__ __launch_bounds__(kThreads) void k_read(int64_t n2, const double2* __restrict__ a,
                                                   double* __restrict__ part) {
  double s = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n2;
       i += (int64_t)gridDim.x * kThreads) {
    const v2d v = __builtin_nontemporal_load(reinterpret_cast<const v2d*>(a) + i);
    s += v[0] + v[1];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) part[blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6)] = s;
}

Mat<F64> mat64(Context& c) {
  return Mat<F64>{c.sell_off.get(), c.sell_col.get(), c.sell_val.get(), c.sell_row0.get(),
                  c.sell_nrow.get(), c.nslices, c.nb};
}
Mat<F32> mat32(Context& c) {
  return Mat<F32>{c.sell_off.get(), c.sell_col.get(), c.sell_val32.get(), c.sell_row0.get(),
                  c.sell_nrow.get(), c.nslices, c.nb};
}

int spmv_grid(Context& c) {
  const int want = (c.nslices + kWaves - 1) / kWaves;
  int g = want < kMaxGrid ? want : kMaxGrid;
  if (g >= 64) g &= ~7;
  return g < 1 ? 1 : g;
}

int vec_grid(Context& c) { return grid_for(c.nb, kThreads, kMaxGrid); }

double elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  HIP_CHECK(hipEventElapsedTime(&ms, a, b));
  return ms * 1e-3;
}

struct Graph {
  hipGraph_t g = nullptr;
  hipGraphExec_t e = nullptr;
  ~Graph() {
    if (e) (void)hipGraphExecDestroy(e);
    if (g) (void)hipGraphDestroy(g);
  }
};

// Partial slots: A at [0, kMaxGrid), B slot s at kMaxGrid + 2*kMaxGrid*s, scratch after.
double* partA(Context& c) { return c.partials.get(); }
double* partB(Context& c, int s) { return c.partials.get() + kMaxGrid + 2 * kMaxGrid * s; }
double* partX(Context& c) { return c.partials.get() + 5 * kMaxGrid; }

}  // namespace

void ensure_vectors(Context& c) {
  const size_t nv = 2 * (size_t)c.nb;
  c.x.alloc(nv);
  c.work.alloc(nv);
  c.r.alloc(nv);
  c.p.alloc(nv);
  c.Ap.alloc(nv);
  c.inv_diag.alloc(nv);
  c.invd32.alloc(nv);
  if (c.b.n != nv) {
    c.b.alloc(nv);
    HIP_CHECK(hipMemsetAsync(c.b.get(), 0, sizeof(double) * nv, c.stream));
  }
  c.partials.alloc(6 * kMaxGrid + 64);
  if (c.fixed.n != nv) {
    c.fixed.alloc(nv);
    c.fixed_valid = false;
  }
}

void setup_solver(Context& c) {
  if (!c.have_pattern) throw Error(CSTFEM_ERR_ORDER, "cg_jacobi: no matrix on the device");
  ensure_vectors(c);
  const int big = INT_MAX;
  HIP_CHECK(hipMemcpyAsync(c.err_flag.get(), &big, sizeof(int), hipMemcpyHostToDevice, c.stream));
  hipLaunchKernelGGL(k_inv_diag, dim3(grid_for(c.nb, 256, 0)), dim3(256), 0, c.stream, c.nb, c.n,
                     c.bdiag.get(), c.sell_pos.get(), c.sell_val.get(), c.inv_diag.get(),
                     c.invd32.get(), c.err_flag.get());
  HIP_CHECK(hipGetLastError());
  int bad = INT_MAX;
  HIP_CHECK(hipMemcpyAsync(&bad, c.err_flag.get(), sizeof(int), hipMemcpyDeviceToHost, c.stream));
  HIP_CHECK(hipStreamSynchronize(c.stream));
  if (bad != INT_MAX)
    throw Error(CSTFEM_ERR_BAD_DIAGONAL, "cg_jacobi: non-positive diagonal at row " + std::to_string(bad));
  c.setup_done = true;
  c.vals32_valid = false;
  c.mpnn.built = false;  // M depends on the values and the Jacobi scaling
  c.mg.clear();          // so does the multilevel hierarchy
}

// FP32 copy of the matrix values for the mixed-precision paths (rounded once per matrix)
void ensure_vals32(Context& c) {
  if (c.vals32_valid) return;
  c.sell_val32.alloc(4 * c.sell_blocks);
  hipLaunchKernelGGL(k_round_values, dim3(2048), dim3(256), 0, c.stream,
                     (int64_t)c.sell_blocks, (const double*)c.sell_val.get(),
                     c.sell_val32.get());
  HIP_CHECK(hipGetLastError());
  c.vals32_valid = true;
}

double dot_fp64(Context& c, const double* a, const double* b) {
  const int gv = vec_grid(c);
  hipLaunchKernelGGL(k_dot_rp<F64>, dim3(gv), dim3(kThreads), 0, c.stream, c.nb,
                     (const double2*)a, (const double2*)b, partX(c));
  hipLaunchKernelGGL(k_sum_into, dim3(1), dim3(kThreads), 0, c.stream, partX(c), gv,
                     &c.scal.get()->rr);
  HIP_CHECK(hipGetLastError());
  double out = 0.0;
  HIP_CHECK(hipMemcpyAsync(&out, &c.scal.get()->rr, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  HIP_CHECK(hipStreamSynchronize(c.stream));
  return out;
}

void spmv_fp64(Context& c, const double* x, double* y) {
  hipLaunchKernelGGL(k_spmv<F64>, dim3(spmv_grid(c)), dim3(kThreads), 0, c.stream, mat64(c),
                     (const double2*)x, (double2*)y);
  HIP_CHECK(hipGetLastError());
}

namespace {

// ||b - A x||^2 into sc->rr (leaves r = b - A x in FP64 in c.r)
void true_residual(Context& c, double* Ax_scratch) {
  const int gv = vec_grid(c);
  spmv_fp64(c, c.x.get(), Ax_scratch);
  hipLaunchKernelGGL(k_residual, dim3(gv), dim3(kThreads), 0, c.stream, c.nb,
                     (const double2*)c.b.get(), (const double2*)Ax_scratch, (double2*)c.r.get(),
                     partX(c));
  hipLaunchKernelGGL(k_sum_into, dim3(1), dim3(kThreads), 0, c.stream, partX(c), gv,
                     &c.scal.get()->rr);
  HIP_CHECK(hipGetLastError());
}

void enqueue_iteration64(Context& c, int cur, double* hist) {
  const int gs = spmv_grid(c), gv = vec_grid(c);
  hipLaunchKernelGGL(k_cg_spmv<F64>, dim3(gs), dim3(kThreads), 0, c.stream, mat64(c),
                     (const double2*)c.p.get(), (double2*)c.Ap.get(), partA(c),
                     (const PcgScalars*)c.scal.get());
  hipLaunchKernelGGL(k_cg_update, dim3(gv), dim3(kThreads), 0, c.stream, c.nb,
                     (double2*)c.x.get(), (double2*)c.r.get(), (const double2*)c.p.get(),
                     (const double2*)c.Ap.get(), (const double2*)c.inv_diag.get(),
                     (const double*)partA(c), gs, (const double*)partB(c, cur ^ 1), partB(c, cur),
                     gv, c.scal.get());
  hipLaunchKernelGGL(k_cg_direction<F64>, dim3(gv), dim3(kThreads), 0, c.stream, c.nb,
                     (const double2*)c.r.get(), (const double2*)c.inv_diag.get(),
                     (double2*)c.p.get(), (const double*)partB(c, cur ^ 1),
                     (const double*)partB(c, cur), gv, cur, c.scal.get(), hist, 1);
}

// FP32 direction update only (after a refinement refreshed the residual of iteration k-1)
void enqueue_direction32(Context& c, int cur, int test) {
  const int gv = vec_grid(c);
  hipLaunchKernelGGL(k_cg_direction<F32>, dim3(gv), dim3(kThreads), 0, c.stream, c.nb,
                     (const float2*)c.r32.get(), (const float2*)c.invd32.get(),
                     (float2*)c.p32.get(), (const double*)partB(c, cur ^ 1),
                     (const double*)partB(c, cur), gv, cur, c.scal.get(), (double*)nullptr, test);
}

void enqueue_iteration32(Context& c, int cur) {
  const int gs = spmv_grid(c), gv = vec_grid(c);
  hipLaunchKernelGGL(k_cg_spmv<F32>, dim3(gs), dim3(kThreads), 0, c.stream, mat32(c),
                     (const float2*)c.p32.get(), (float2*)c.Ap32.get(), partA(c),
                     (const PcgScalars*)c.scal.get());
  hipLaunchKernelGGL(k_cg_update32, dim3(gv), dim3(kThreads), 0, c.stream, c.nb,
                     (double2*)c.x.get(), (float2*)c.r32.get(), (const float2*)c.p32.get(),
                     (const float2*)c.Ap32.get(), (const float2*)c.invd32.get(),
                     (const double*)partA(c), gs, (const double*)partB(c, cur ^ 1), partB(c, cur),
                     gv, c.scal.get());
  enqueue_direction32(c, cur, 1);
}

// Replays `graph` (an even number of iterations) until the device status leaves kRunning;
// one chunk stays in flight while the previous chunk's status is read.
// `chunk` iterations as one unit of work: a captured hipGraph, or -- when the environment
// sets CSTFEM_GRAPH=0, for tools that cannot follow graph launches -- the same kernels
// launched directly.
// A one-iteration chunk holds two graphs, one per parity, launched alternately from parity
// 0 (restart()): a stop then leaves one iteration's launches in flight instead of two.
struct Chunk {
  Graph gr, gr1;
  std::function<void(int)> iteration;
  int n = 0;
  int next = 0;  // parity of the next one-iteration chunk
  bool graph = true;
  Chunk(Context& c, int count, std::function<void(int)> it) : iteration(std::move(it)), n(count) {
    const char* env = std::getenv("CSTFEM_GRAPH");
    graph = !(env && env[0] == '0');
    if (!graph) return;
    capture(c, gr, 0);
    if (n == 1) capture(c, gr1, 1);
  }
  void capture(Context& c, Graph& g, int parity) {
    HIP_CHECK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < n; ++i) iteration((i + parity) & 1);
    HIP_CHECK(hipStreamEndCapture(c.stream, &g.g));
    HIP_CHECK(hipGraphInstantiate(&g.e, g.g, nullptr, nullptr, 0));
  }
  void restart() { next = 0; }
  void launch(hipStream_t s) {
    const int p = n == 1 ? next : 0;
    next ^= 1;
    if (graph) {
      HIP_CHECK(hipGraphLaunch(p ? gr1.e : gr.e, s));
    } else {
      for (int i = 0; i < n; ++i) iteration((i + p) & 1);
      HIP_CHECK(hipGetLastError());
    }
  }
};

void run_until_stopped(Context& c, Chunk& graph, int chunk, int64_t max_iter) {
  PcgScalars* slot[2] = {c.h_scal, c.h_scal + 1};
  hipEvent_t done[2];
  HIP_CHECK(hipEventCreateWithFlags(&done[0], hipEventDisableTiming));
  HIP_CHECK(hipEventCreateWithFlags(&done[1], hipEventDisableTiming));
  const int64_t max_chunks = (max_iter + chunk - 1) / chunk + 2;
  int cur = 0;
  graph.restart();  // every run starts at parity 0
  graph.launch(c.stream);
  HIP_CHECK(hipMemcpyAsync(slot[cur], c.scal.get(), sizeof(PcgScalars), hipMemcpyDeviceToHost, c.stream));
  HIP_CHECK(hipEventRecord(done[cur], c.stream));
  for (int64_t ch = 1; ch <= max_chunks; ++ch) {
    const int nxt = cur ^ 1;
    graph.launch(c.stream);
    HIP_CHECK(hipMemcpyAsync(slot[nxt], c.scal.get(), sizeof(PcgScalars), hipMemcpyDeviceToHost, c.stream));
    HIP_CHECK(hipEventRecord(done[nxt], c.stream));
    HIP_CHECK(hipEventSynchronize(done[cur]));
    if (slot[cur]->status != kRunning) break;
    cur = nxt;
  }
  HIP_CHECK(hipStreamSynchronize(c.stream));
  (void)hipEventDestroy(done[0]);
  (void)hipEventDestroy(done[1]);
}

PcgScalars read_scalars(Context& c) {
  HIP_CHECK(hipMemcpyAsync(c.h_scal, c.scal.get(), sizeof(PcgScalars), hipMemcpyDeviceToHost, c.stream));
  HIP_CHECK(hipStreamSynchronize(c.stream));
  return *c.h_scal;
}

}  // namespace

int solve_cg_fp64(Context& c, const cstfem_solver_options& o, cstfem_solver_stats& st,
                  double* history) {
  if (!c.setup_done) throw Error(CSTFEM_ERR_ORDER, "cg_jacobi: solve() called before setup()");
  const int gv = vec_grid(c);
  DBuf<double> d_hist;
  if (o.record_history) d_hist.alloc((size_t)o.max_iter + 1);
  double* hist = o.record_history ? d_hist.get() : nullptr;

  PcgScalars init{};
  init.rtol = o.rtol;
  init.max_iter = o.max_iter;
  init.record = o.record_history ? 1 : 0;
  HIP_CHECK(hipMemcpyAsync(c.scal.get(), &init, sizeof init, hipMemcpyHostToDevice, c.stream));

  hipEvent_t ev0, ev1;
  HIP_CHECK(hipEventCreate(&ev0));
  HIP_CHECK(hipEventCreate(&ev1));
  HIP_CHECK(hipEventRecord(ev0, c.stream));
  hipLaunchKernelGGL(k_cg_init, dim3(gv), dim3(kThreads), 0, c.stream, c.nb,
                     (const double2*)c.b.get(), (double2*)c.x.get(), (double2*)c.r.get(),
                     (double2*)c.p.get(), (const double2*)c.inv_diag.get(), partB(c, 1));
  hipLaunchKernelGGL(k_cg_init_finish, dim3(1), dim3(kThreads), 0, c.stream,
                     (const double*)partB(c, 1), gv, c.scal.get(), hist);
  HIP_CHECK(hipGetLastError());
  const bool zero_rhs = read_scalars(c).bnorm == 0.0;

  double t_spmv = 0.0, t_vec = 0.0;
  if (!zero_rhs && o.max_iter > 0) {
    if (o.instrument) {
      // eager launches with events around each kernel class (same kernels, same results)
      hipEvent_t e0, e1, e2;
      HIP_CHECK(hipEventCreate(&e0));
      HIP_CHECK(hipEventCreate(&e1));
      HIP_CHECK(hipEventCreate(&e2));
      const int gs = spmv_grid(c);
      for (int it = 0;; ++it) {
        const int cur = it & 1;
        HIP_CHECK(hipEventRecord(e0, c.stream));
        hipLaunchKernelGGL(k_cg_spmv<F64>, dim3(gs), dim3(kThreads), 0, c.stream, mat64(c),
                           (const double2*)c.p.get(), (double2*)c.Ap.get(), partA(c),
                           (const PcgScalars*)c.scal.get());
        HIP_CHECK(hipEventRecord(e1, c.stream));
        hipLaunchKernelGGL(k_cg_update, dim3(gv), dim3(kThreads), 0, c.stream, c.nb,
                           (double2*)c.x.get(), (double2*)c.r.get(), (const double2*)c.p.get(),
                           (const double2*)c.Ap.get(), (const double2*)c.inv_diag.get(),
                           (const double*)partA(c), gs, (const double*)partB(c, cur ^ 1),
                           partB(c, cur), gv, c.scal.get());
        hipLaunchKernelGGL(k_cg_direction<F64>, dim3(gv), dim3(kThreads), 0, c.stream, c.nb,
                           (const double2*)c.r.get(), (const double2*)c.inv_diag.get(),
                           (double2*)c.p.get(), (const double*)partB(c, cur ^ 1),
                           (const double*)partB(c, cur), gv, cur, c.scal.get(), hist, 1);
        HIP_CHECK(hipEventRecord(e2, c.stream));
        const PcgScalars s = read_scalars(c);
        t_spmv += elapsed(e0, e1);
        t_vec += elapsed(e1, e2);
        if (s.status != kRunning) break;
      }
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      (void)hipEventDestroy(e2);
    } else {
      const int chunk = c.n > 200000 ? 16 : 64;  // even
      Chunk gr(c, chunk, [&](int cur) { enqueue_iteration64(c, cur, hist); });
      run_until_stopped(c, gr, chunk, o.max_iter);
    }
  }
  PcgScalars s = read_scalars(c);
  const int iters = s.k_in[0] > s.k_in[1] ? s.k_in[0] : s.k_in[1];

  // true residual b - A x (pcg_jacobi, solver.cpp:113-129)
  double true_rel = 0.0;
  if (!zero_rhs) {
    hipEvent_t t0, t1;
    HIP_CHECK(hipEventCreate(&t0));
    HIP_CHECK(hipEventCreate(&t1));
    HIP_CHECK(hipEventRecord(t0, c.stream));
    true_residual(c, c.work.get());
    HIP_CHECK(hipEventRecord(t1, c.stream));
    const PcgScalars s2 = read_scalars(c);
    true_rel = std::sqrt(s2.rr) / s.bnorm;
    t_spmv += elapsed(t0, t1);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
  }
  HIP_CHECK(hipEventRecord(ev1, c.stream));
  HIP_CHECK(hipEventSynchronize(ev1));
  const double total = elapsed(ev0, ev1);
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);

  if (history && o.record_history) {
    const int cnt = zero_rhs ? 1 : iters + 1;
    HIP_CHECK(hipMemcpy(history, hist, sizeof(double) * cnt, hipMemcpyDeviceToHost));
  }
  st.iterations = zero_rhs ? 0 : iters;
  st.converged = (zero_rhs || s.status == kConverged) ? 1 : 0;
  st.breakdown = s.status == kBreakdown ? 1 : 0;
  st.zero_rhs = zero_rhs ? 1 : 0;
  st.rel_residual = zero_rhs ? 0.0 : s.rel;
  st.true_rel_residual = true_rel;
  st.breakdown_pap = s.status == kBreakdown ? s.pAp : 0.0;
  st.warning = (st.converged && !zero_rhs && true_rel > 10.0 * o.rtol) ? 1 : 0;
  st.spmv_calls = zero_rhs ? 0 : (long long)st.iterations + (st.breakdown ? 1 : 0) + 1;
  // the reference's reduction count: norm(b), r'z, then per iteration p'Ap, ||r|| (+ r'z
  // unless it stopped), then pcg_jacobi's norm(b) and ||b - Ax||
  long long red = 2;
  if (!zero_rhs) {
    const long long k = st.iterations;
    red += 2 * k + (st.converged ? k - 1 : k) + (st.breakdown ? 1 : 0) + 2;
  }
  st.reductions = red;
  st.spmv_local = o.instrument ? t_spmv : 0.0;
  st.vector_ops = o.instrument ? t_vec : 0.0;
  st.dot_local = 0.0;  // fused into the SpMV, update and direction kernels
  st.total = total;
  if (zero_rhs || s.status == kConverged) return CSTFEM_OK;
  return CSTFEM_NOT_CONVERGED;
}

// Jacobi-PCG refined by FP64 true residuals.  FP32 mode: FP32 matrix values and vectors,
// FP64-accumulated dots and an FP64 solution; FP64 mode: the reference recurrence.  Whenever
// the recursive residual reaches rtol (or, FP32 mode, drops by `delta` since the last
// refinement), r is replaced by b - A x computed in FP64 -- the search direction is kept, so
// the Krylov information survives -- and the solve stops only when the FP64 true residual is
// below rtol.
int solve_refined(Context& c, const cstfem_solver_options& o, cstfem_solver_stats& st, bool fp32,
                  bool warm) {
  if (!c.setup_done) throw Error(CSTFEM_ERR_ORDER, "cg_jacobi: solve() called before setup()");
  if (warm && !fp32) throw Error(CSTFEM_ERR_ARGUMENT, "solve_refined: warm start needs FP32 mode");
  const size_t nv = 2 * (size_t)c.nb;
  if (fp32) {
    c.r32.alloc(nv);
    c.p32.alloc(nv);
    c.Ap32.alloc(nv);
    ensure_vals32(c);
  }
  const int gv = vec_grid(c);
  // FP64: refresh only where the reference stops (rtol); FP32 drifts faster
  const double delta = o.inner_rtol > 0 ? o.inner_rtol : (fp32 ? 1e-3 : 0.0);
  const int max_outer = o.max_outer > 0 ? o.max_outer : 200;
  auto iteration = [&](int cur) {
    if (fp32)
      enqueue_iteration32(c, cur);
    else
      enqueue_iteration64(c, cur, nullptr);
  };
  const int refresh = o.refresh;                // 0 restart, 1 corrective step, 2 keep p
  const int dir_mode = refresh == 0 ? 2 : 0;
  auto direction = [&](int cur) {
    if (fp32) {
      enqueue_direction32(c, cur, dir_mode);
    } else {
      hipLaunchKernelGGL(k_cg_direction<F64>, dim3(gv), dim3(kThreads), 0, c.stream, c.nb,
                         (const double2*)c.r.get(), (const double2*)c.inv_diag.get(),
                         (double2*)c.p.get(), (const double*)partB(c, cur ^ 1),
                         (const double*)partB(c, cur), gv, cur, c.scal.get(), (double*)nullptr,
                         dir_mode);
    }
  };

  PcgScalars init{};
  init.rtol = o.rtol;
  init.max_iter = o.max_iter;
  init.refine_mode = 1;
  init.delta = delta;
  init.rel_ref = 1.0;
  HIP_CHECK(hipMemcpyAsync(c.scal.get(), &init, sizeof init, hipMemcpyHostToDevice, c.stream));
  hipEvent_t ev0, ev1;
  HIP_CHECK(hipEventCreate(&ev0));
  HIP_CHECK(hipEventCreate(&ev1));
  HIP_CHECK(hipEventRecord(ev0, c.stream));
  if (warm) true_residual(c, c.work.get());  // r = b - A x for the given x
  if (fp32)
    hipLaunchKernelGGL(k_init32, dim3(gv), dim3(kThreads), 0, c.stream, c.nb,
                       (const double2*)c.b.get(), (double2*)c.x.get(), (float2*)c.r32.get(),
                       (float2*)c.p32.get(), (const float2*)c.invd32.get(), partB(c, 1),
                       warm ? (const double2*)c.r.get() : (const double2*)nullptr);
  else
    hipLaunchKernelGGL(k_cg_init, dim3(gv), dim3(kThreads), 0, c.stream, c.nb,
                       (const double2*)c.b.get(), (double2*)c.x.get(), (double2*)c.r.get(),
                       (double2*)c.p.get(), (const double2*)c.inv_diag.get(), partB(c, 1));
  hipLaunchKernelGGL(k_cg_init_finish, dim3(1), dim3(kThreads), 0, c.stream,
                     (const double*)partB(c, 1), gv, c.scal.get(), (double*)nullptr);
  HIP_CHECK(hipGetLastError());
  PcgScalars s = read_scalars(c);
  const double bnorm = s.bnorm;
  const bool zero_rhs = bnorm == 0.0;
  int outer = 0;
  double true_rel = zero_rhs ? 0.0 : 1.0;
  if (warm && !zero_rhs) {  // refinement thresholds relative to the starting residual
    true_rel = std::sqrt(s.rr) / bnorm;
    s.rel_ref = true_rel;
    HIP_CHECK(hipMemcpyAsync(c.scal.get(), &s, sizeof s, hipMemcpyHostToDevice, c.stream));
    HIP_CHECK(hipStreamSynchronize(c.stream));
  }
  bool converged = zero_rhs || (warm && true_rel < o.rtol);
  int last_status = kRunning;
  if (!converged && o.max_iter > 0) {
    const int chunk = c.n > 200000 ? 16 : 64;  // even
    Chunk gr(c, chunk, iteration);
    double prev_true = true_rel;
    int stagnant = 0;
    for (;;) {
      run_until_stopped(c, gr, chunk, o.max_iter);
      s = read_scalars(c);
      last_status = s.status;
      // iteration k-1 (parity (k-1)&1) stopped before its direction update
      const int k = s.k_in[0] > s.k_in[1] ? s.k_in[0] : s.k_in[1];
      if (s.status != kRefine) break;  // breakdown or max_iter
      // refinement: FP64 true residual r = b - A x; the residual of iteration k-1 and its
      // partials (r'r, r'z) are replaced by the refreshed ones
      const int cur = (k - 1) & 1;
      true_residual(c, c.work.get());
      const int correct = refresh == 1 ? 1 : 0;
      if (fp32) {
        if (correct)
          hipLaunchKernelGGL(k_dot_rp<F32>, dim3(gv), dim3(kThreads), 0, c.stream, c.nb,
                             (const double2*)c.r.get(), (const float2*)c.p32.get(), partX(c));
        hipLaunchKernelGGL(k_refresh<F32>, dim3(gv), dim3(kThreads), 0, c.stream, c.nb,
                           (double2*)c.x.get(), (double2*)c.r.get(), (const float2*)c.p32.get(),
                           (const float2*)c.Ap32.get(), (const float2*)c.invd32.get(),
                           (float2*)c.r32.get(), (const double*)partA(c), spmv_grid(c),
                           (const double*)partX(c), gv, correct, partB(c, cur));
      } else {
        if (correct)
          hipLaunchKernelGGL(k_dot_rp<F64>, dim3(gv), dim3(kThreads), 0, c.stream, c.nb,
                             (const double2*)c.r.get(), (const double2*)c.p.get(), partX(c));
        hipLaunchKernelGGL(k_refresh<F64>, dim3(gv), dim3(kThreads), 0, c.stream, c.nb,
                           (double2*)c.x.get(), (double2*)c.r.get(), (const double2*)c.p.get(),
                           (const double2*)c.Ap.get(), (const double2*)c.inv_diag.get(),
                           (float2*)nullptr, (const double*)partA(c), spmv_grid(c),
                           (const double*)partX(c), gv, correct, partB(c, cur));
      }
      HIP_CHECK(hipGetLastError());
      const PcgScalars s2 = read_scalars(c);
      true_rel = std::sqrt(s2.rr) / bnorm;
      ++outer;
      if (true_rel < o.rtol) {
        converged = true;
        break;
      }
      stagnant = true_rel > 0.5 * prev_true ? stagnant + 1 : 0;
      prev_true = true_rel;
      if (outer >= max_outer || k >= o.max_iter || stagnant >= 4) break;
      PcgScalars upd = s2;
      upd.status = kRunning;
      upd.rel_ref = true_rel;
      HIP_CHECK(hipMemcpyAsync(c.scal.get(), &upd, sizeof upd, hipMemcpyHostToDevice, c.stream));
      // the skipped direction update of iteration k-1, with the refreshed residual
      direction(cur);
      // the graph starts at parity 0: an odd next iteration index k runs once eagerly
      if (k & 1) iteration(1);
      HIP_CHECK(hipGetLastError());
      last_status = kRunning;
    }
  }
  s = read_scalars(c);
  HIP_CHECK(hipEventRecord(ev1, c.stream));
  HIP_CHECK(hipEventSynchronize(ev1));
  const int iters = s.k_in[0] > s.k_in[1] ? s.k_in[0] : s.k_in[1];
  st.iterations = zero_rhs ? 0 : iters;
  st.outer_iterations = outer;
  st.converged = converged ? 1 : 0;
  st.breakdown = last_status == kBreakdown ? 1 : 0;
  st.zero_rhs = zero_rhs ? 1 : 0;
  st.rel_residual = zero_rhs ? 0.0 : s.rel;
  st.true_rel_residual = true_rel;
  st.breakdown_pap = last_status == kBreakdown ? s.pAp : 0.0;
  st.warning = 0;
  st.spmv_calls = zero_rhs ? 0 : (long long)iters + outer;
  st.reductions = 0;
  st.total = elapsed(ev0, ev1);
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  return converged ? CSTFEM_OK : CSTFEM_NOT_CONVERGED;
}

// ---------------------------------------------------------------------------------------
// Flexible CG (Notay's FCG with one kept direction) with the learned preconditioner M (FP32
// values in the matrix's SELL layout).  P = F64: the reference's FP64 matrix and vectors, M
// applied with FP64 accumulation; P = F32: FP32 matrix and vectors, FP64 dots and solution.
// Both stop on the FP64 true residual.
//   A  q = A p; partials p'q, p'r
//   B  alpha = p'r / p'q (breakdown if !(p'q > 0)); x += alpha p; r -= alpha q; partial r'r
//   C  convergence / refinement / stall test; z = M r; partial z'q
//   D  beta = -z'q / p'q (0 on a restart); p = z + beta p

namespace {

// rows of a SELL slice of M (FP32 values) times an FP64 vector, FP64 accumulation
__device__ inline void sell_row(const int64_t* __restrict__ off, const int* __restrict__ col,
                                const float* __restrict__ val, const double2* __restrict__ x,
                                int s, int lane, double& y0, double& y1) {
  const int64_t o = off[s];
  const int w = (int)((off[s + 1] - o) >> 6);
  const int* c = col + o + lane;
  const v4f* v = reinterpret_cast<const v4f*>(val) + (o + lane);
  double a0 = 0.0, a1 = 0.0;
#pragma unroll 4
  for (int k = 0; k < w; ++k) {
    const int m = c[(size_t)k * 64];
    const v4f vv = __builtin_nontemporal_load(v + (size_t)k * 64);
    const double2 xm = x[m];
    a0 = a0 + (double)vv.x * xm.x;
    a0 = a0 + (double)vv.y * xm.y;
    a1 = a1 + (double)vv.z * xm.x;
    a1 = a1 + (double)vv.w * xm.y;
  }
  y0 = a0;
  y1 = a1;
}

template <class P>
__device__ inline typename P::V make_v(double a, double b) {
  if constexpr (std::is_same<typename P::V, double2>::value)
    return make_double2(a, b);
  else
    return make_float2((float)a, (float)b);
}

template <class P>
__global__ __launch_bounds__(kThreads) void k_fcg_spmv(Mat<P> A, const typename P::V* __restrict__ p,
                                                       const typename P::V* __restrict__ r,
                                                       typename P::V* __restrict__ q,
                                                       double* __restrict__ part,
                                                       const PcgScalars* __restrict__ sc) {
  __shared__ double lds[kWaves];
  if (sc->status != kRunning) return;
  const int lane = threadIdx.x & 63;
  const SliceRange sr = slice_range(A.nslices);
  double pq = 0.0, pr = 0.0;
  for (int s = sr.first; s < sr.end; s += sr.step) {
    const int row = A.row0[s] + lane;
    typename P::S y0, y1;
    sell_row(A.off, A.col, A.val, p, s, lane, y0, y1);
    if (lane < A.nrow[s]) {
      q[row] = make_v<P>(y0, y1);
      const typename P::V pi = p[row], ri = r[row];
      pq += (double)pi.x * (double)y0 + (double)pi.y * (double)y1;
      pr += (double)pi.x * (double)ri.x + (double)pi.y * (double)ri.y;
    }
  }
  const double a = block_sum_all(pq, lds);
  const double b = block_sum_all(pr, lds);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[kMaxGrid + blockIdx.x] = b;
  }
}

template <class P>
__global__ __launch_bounds__(kThreads) void k_fcg_update(
    int nvec2, double2* __restrict__ x, typename P::V* __restrict__ r,
    const typename P::V* __restrict__ p, const typename P::V* __restrict__ q,
    const double* __restrict__ partA, int nA, double* __restrict__ partB,
    PcgScalars* __restrict__ sc) {
  __shared__ double lds[kWaves];
  if (sc->status != kRunning) return;
  const double pq = reduce_partials(partA, nA, 1, lds);
  const double pr = reduce_partials(partA + kMaxGrid, nA, 1, lds);
  if (!(pq > 0.0)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc->pAp = pq;
      sc->status = kBreakdown;
    }
    return;
  }
  const double alpha = pr / pq;
  double rr = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < nvec2; i += gridDim.x * kThreads) {
    const typename P::V pi = p[i], qi = q[i];
    double2 xi = x[i];
    const typename P::V r0 = r[i];
    xi.x = xi.x + alpha * (double)pi.x;
    xi.y = xi.y + alpha * (double)pi.y;
    const typename P::V ri = make_v<P>((double)r0.x - alpha * (double)qi.x,
                                       (double)r0.y - alpha * (double)qi.y);
    x[i] = xi;
    r[i] = ri;
    rr += (double)ri.x * (double)ri.x + (double)ri.y * (double)ri.y;
  }
  const double t = block_sum_all(rr, lds);
  if (threadIdx.x == 0) partB[blockIdx.x] = t;
}

// The tests of flexible CG iteration k (parity cur) on its recursive ||r||/||b|| rel:
// refinement (rel below rtol, or below delta times the last refresh's), max_iter, and stall
// (NaN; no new best for stall_window iterations; or, with slow_window > 0 and far from rtol,
// the best not down to slow_factor times its value slow_window iterations before).  Reads
// the previous iteration's records (other parity slot: stable during the kernel); `writer`
// carries them forward into slot cur and posts a stop.  Returns the status.
__device__ inline int fcg_test_status(PcgScalars* __restrict__ sc, int cur, double rel, int k,
                                      bool writer) {
  const bool better = rel < sc->best_rel[cur ^ 1];
  const double best = better ? rel : sc->best_rel[cur ^ 1];
  const int best_k = better ? k : sc->best_k[cur ^ 1];
  double slow_rel = sc->slow_rel[cur ^ 1];
  int slow_k = sc->slow_k[cur ^ 1];
  bool slow = false;
  if (sc->slow_window > 0 && k - slow_k >= sc->slow_window) {
    slow = !(best <= sc->slow_factor * slow_rel) && best > 1e3 * sc->rtol;
    slow_rel = best;
    slow_k = k;
  }
  int status = kRunning;
  if (rel < sc->rtol || rel < sc->delta * sc->rel_ref)
    status = kRefine;
  else if (k >= sc->max_iter)
    status = kMaxIter;
  else if (!(rel == rel) || k - best_k > sc->stall_window || slow)
    status = kStalled;
  if (writer) {
    sc->k_in[cur] = k;
    sc->rel = rel;
    sc->best_rel[cur] = best;
    sc->best_k[cur] = best_k;
    sc->slow_rel[cur] = slow_rel;
    sc->slow_k[cur] = slow_k;
    if (status != kRunning) sc->status = status;
  }
  return status;
}

// test: 1 = convergence / refinement / stall test first; 0 = preconditioner application only
template <class P>
__global__ __launch_bounds__(kThreads) void k_fcg_precond(
    Mat<F32> G, const typename P::V* __restrict__ r, const typename P::V* __restrict__ q,
    typename P::V* __restrict__ z, const double* __restrict__ partB, int nB,
    double* __restrict__ partC, int cur, PcgScalars* __restrict__ sc, int test) {
  __shared__ double lds[kWaves];
  if (sc->status != kRunning) return;
  if (test) {
    const double rr = reduce_partials(partB, nB, 1, lds);
    const int k = sc->k_in[cur ^ 1] + 1;
    const double rel = sqrt(rr) / sc->bnorm;
    if (fcg_test_status(sc, cur, rel, k, blockIdx.x == 0 && threadIdx.x == 0) != kRunning) return;
  }
  const int lane = threadIdx.x & 63;
  const SliceRange sr = slice_range(G.nslices);
  double zq = 0.0;
  for (int s = sr.first; s < sr.end; s += sr.step) {
    const int row = G.row0[s] + lane;
    typename P::S y0, y1;
    sell_row(G.off, G.col, G.val, r, s, lane, y0, y1);
    if (lane < G.nrow[s]) {
      z[row] = make_v<P>(y0, y1);
      const typename P::V qi = q[row];
      zq += (double)y0 * (double)qi.x + (double)y1 * (double)qi.y;
    }
  }
  const double t = block_sum_all(zq, lds);
  if (threadIdx.x == 0) partC[blockIdx.x] = t;
}

// the convergence / refinement / stall test of k_fcg_precond alone (one workgroup), for
// preconditioners applied by a kernel sequence
__global__ __launch_bounds__(kThreads) void k_fcg_test(const double* __restrict__ partB, int nB,
                                                       int cur, PcgScalars* __restrict__ sc) {
  __shared__ double lds[kWaves];
  if (sc->status != kRunning) return;
  const double rr = reduce_partials(partB, nB, 1, lds);
  const int k = sc->k_in[cur ^ 1] + 1;
  const double rel = sqrt(rr) / sc->bnorm;
  fcg_test_status(sc, cur, rel, k, threadIdx.x == 0);
}

// partials of z'q
template <class P>
__global__ __launch_bounds__(kThreads) void k_fcg_zq(int nvec2, const typename P::V* __restrict__ z,
                                                     const typename P::V* __restrict__ q,
                                                     double* __restrict__ partC,
                                                     const PcgScalars* __restrict__ sc) {
  __shared__ double lds[kWaves];
  if (sc->status != kRunning) return;
  double s = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < nvec2; i += gridDim.x * kThreads) {
    const typename P::V zi = z[i], qi = q[i];
    s += (double)zi.x * (double)qi.x + (double)zi.y * (double)qi.y;
  }
  const double t = block_sum_all(s, lds);
  if (threadIdx.x == 0) partC[blockIdx.x] = t;
}

template <class P>
__global__ __launch_bounds__(kThreads) void k_fcg_direction(
    int nvec2, const typename P::V* __restrict__ z, typename P::V* __restrict__ p,
    const double* __restrict__ partA, int nA, const double* __restrict__ partC, int nC,
    const PcgScalars* __restrict__ sc, int restart) {
  __shared__ double lds[kWaves];
  if (sc->status != kRunning) return;
  double beta = 0.0;
  if (!restart) {
    const double pq = reduce_partials(partA, nA, 1, lds);
    const double zq = reduce_partials(partC, nC, 1, lds);
    beta = -zq / pq;
  }
  const typename P::S b = (typename P::S)beta;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < nvec2; i += gridDim.x * kThreads) {
    const typename P::V zi = z[i];
    if (restart) {  // p = z: the old direction is not read (it may hold a failed solve's NaN)
      p[i] = zi;
      continue;
    }
    typename P::V pi = p[i];
    pi.x = zi.x + b * pi.x;
    pi.y = zi.y + b * pi.y;
    p[i] = pi;
  }
}

// r = b (x = 0) or r = r64 (warm); bnorm partials as k_init32 does
template <class P>
__global__ __launch_bounds__(kThreads) void k_fcg_init(int nvec2, const double2* __restrict__ b,
                                                       double2* __restrict__ x,
                                                       typename P::V* __restrict__ r,
                                                       typename P::V* __restrict__ q,
                                                       double* __restrict__ part) {
  __shared__ double lds[kWaves];
  double bb = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < nvec2; i += gridDim.x * kThreads) {
    const double2 bi = b[i];
    x[i] = make_double2(0.0, 0.0);
    r[i] = make_v<P>(bi.x, bi.y);
    q[i] = make_v<P>(0.0, 0.0);
    bb += bi.x * bi.x + bi.y * bi.y;
  }
  const double s0 = block_sum_all(bb, lds);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s0;
    part[2 * blockIdx.x + 1] = 0.0;
  }
}

// warm start from the current x: r = b - A x (Ax in FP64), q = 0; partials (b'b, 0) as
// k_fcg_init
template <class P>
__global__ __launch_bounds__(kThreads) void k_fcg_init_warm(int nvec2, const double2* __restrict__ b,
                                                            const double2* __restrict__ Ax,
                                                            typename P::V* __restrict__ r,
                                                            typename P::V* __restrict__ q,
                                                            double* __restrict__ part) {
  __shared__ double lds[kWaves];
  double bb = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < nvec2; i += gridDim.x * kThreads) {
    const double2 bi = b[i], ai = Ax[i];
    r[i] = make_v<P>(bi.x - ai.x, bi.y - ai.y);
    q[i] = make_v<P>(0.0, 0.0);
    bb += bi.x * bi.x + bi.y * bi.y;
  }
  const double s0 = block_sum_all(bb, lds);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s0;
    part[2 * blockIdx.x + 1] = 0.0;
  }
}

// after a refinement: the working residual from r64 = b - A x (FP64)
template <class P>
__global__ __launch_bounds__(kThreads) void k_fcg_set_r(int nvec2, const double2* __restrict__ r64,
                                                        typename P::V* __restrict__ r) {
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < nvec2; i += gridDim.x * kThreads)
    r[i] = make_v<P>(r64[i].x, r64[i].y);
}

// warm: start from the current x (r = b - A x) instead of x = 0
template <class P>
int fcg_run(Context& c, const cstfem_solver_options& o, cstfem_solver_stats& st, bool fp32,
            bool warm = false) {
  using V = typename P::V;
  const size_t nv = 2 * (size_t)c.nb;
  const int gs = spmv_grid(c), gv = vec_grid(c);
  double* pA = c.partials.get();
  double* pB = pA + 2 * kMaxGrid;
  double* pC = pA + 3 * kMaxGrid;
  Mat<P> A;
  V *r, *p, *q, *z;
  DBuf<double>& z64 = c.z64;
  DBuf<double>& ax64 = c.ax64;
  if (fp32) {
    c.r32.alloc(nv);
    c.p32.alloc(nv);
    c.Ap32.alloc(nv);
    c.x32.alloc(nv);
    ensure_vals32(c);
    r = (V*)c.r32.get();
    p = (V*)c.p32.get();
    q = (V*)c.Ap32.get();
    z = (V*)c.x32.get();
  } else {
    z64.alloc(nv);
    ax64.alloc(nv);
    r = (V*)c.r.get();
    p = (V*)c.p.get();
    q = (V*)c.Ap.get();
    z = (V*)z64.get();
  }
  {
    void* vals = fp32 ? (void*)c.sell_val32.get() : (void*)c.sell_val.get();
    A = Mat<P>{c.sell_off.get(), c.sell_col.get(), (const typename P::S*)vals, c.sell_row0.get(),
               c.sell_nrow.get(), c.nslices, c.nb};
  }
  const Mat<F32> G{c.sell_off.get(), c.sell_col.get(), c.mpnn.gsell.get(), c.sell_row0.get(),
                   c.sell_nrow.get(), c.nslices, c.nb};
  // scratch of a refinement: r64 = b - A x and A x (FP64), apart from r, p, q, z
  double* r64 = fp32 ? c.r.get() : c.work.get();
  double* Ax = fp32 ? c.work.get() : ax64.get();
  // the multilevel preconditioner (FP64 flexible CG only) or the learned M alone
  const bool multilevel = !fp32 && c.mg.ready;
  int nC = multilevel ? gv : gs;  // partials of z'q
  auto precond = [&](int cur, int test) {
    if (multilevel) {
      if (test)
        hipLaunchKernelGGL(k_fcg_test, dim3(1), dim3(kThreads), 0, c.stream, (const double*)pB, gv,
                           cur, c.scal.get());
      // the cycle's last kernel forms the z'q partials where it can (learned smoother)
      const int nzq = mg_enqueue_apply(c, (const double*)r, (double*)z, c.scal.get(),
                                       (const double*)q, pC);
      if (nzq > 0) {
        nC = nzq;
      } else {
        hipLaunchKernelGGL(k_fcg_zq<P>, dim3(gv), dim3(kThreads), 0, c.stream, c.nb, (const V*)z,
                           (const V*)q, pC, (const PcgScalars*)c.scal.get());
        nC = gv;
      }
    } else {
      hipLaunchKernelGGL(k_fcg_precond<P>, dim3(gs), dim3(kThreads), 0, c.stream, G, (const V*)r,
                         (const V*)q, z, (const double*)pB, gv, pC, cur, c.scal.get(), test);
    }
  };
  auto direction = [&](int restart) {
    hipLaunchKernelGGL(k_fcg_direction<P>, dim3(gv), dim3(kThreads), 0, c.stream, c.nb,
                       (const V*)z, p, (const double*)pA, gs, (const double*)pC, nC,
                       (const PcgScalars*)c.scal.get(), restart);
  };
  auto iteration = [&](int cur) {
    hipLaunchKernelGGL(k_fcg_spmv<P>, dim3(gs), dim3(kThreads), 0, c.stream, A, (const V*)p,
                       (const V*)r, q, pA, (const PcgScalars*)c.scal.get());
    hipLaunchKernelGGL(k_fcg_update<P>, dim3(gv), dim3(kThreads), 0, c.stream, c.nb,
                       (double2*)c.x.get(), r, (const V*)p, (const V*)q, (const double*)pA, gs,
                       pB, c.scal.get());
    precond(cur, 1);
    direction(0);
  };
  // FP64: refresh only where the plain recurrence would stop; FP32: frequent refreshes that
  // keep the search direction (fewer iterations than restarts)
  const double delta = o.inner_rtol > 0 ? o.inner_rtol : (fp32 ? 1e-1 : 0.0);
  const int max_outer = o.max_outer > 0 ? o.max_outer : 200;
  PcgScalars init{};
  init.rtol = o.rtol;
  init.max_iter = o.max_iter;
  init.refine_mode = 1;
  init.delta = delta;
  init.rel_ref = 1.0;
  init.best_rel[0] = init.best_rel[1] = 1.0;
  // the multilevel preconditioner gains an order of magnitude every 10-20 iterations: 200
  // iterations without a new best residual are a stall
  init.stall_window = multilevel ? 200 : std::max(1000, o.max_iter / 20);
  // and must at least halve its best residual every 50 iterations: a learned smoother that
  // passed the set-up check but leaves some modes alone shows as a plateau whose small
  // fluctuations keep setting new bests
  init.slow_window = multilevel ? 50 : 0;
  init.slow_factor = 0.5;
  init.slow_rel[0] = init.slow_rel[1] = 1.0;
  HIP_CHECK(hipMemcpyAsync(c.scal.get(), &init, sizeof init, hipMemcpyHostToDevice, c.stream));
  hipEvent_t ev0, ev1;
  HIP_CHECK(hipEventCreate(&ev0));
  HIP_CHECK(hipEventCreate(&ev1));
  HIP_CHECK(hipEventRecord(ev0, c.stream));
  // x = 0 (or the current x), r = b - A x, q = 0; bnorm; z = M r; p = z
  if (warm) {
    spmv_fp64(c, c.x.get(), Ax);
    hipLaunchKernelGGL(k_fcg_init_warm<P>, dim3(gv), dim3(kThreads), 0, c.stream, c.nb,
                       (const double2*)c.b.get(), (const double2*)Ax, r, q, partB(c, 1));
  } else {
    hipLaunchKernelGGL(k_fcg_init<P>, dim3(gv), dim3(kThreads), 0, c.stream, c.nb,
                       (const double2*)c.b.get(), (double2*)c.x.get(), r, q, partB(c, 1));
  }
  hipLaunchKernelGGL(k_cg_init_finish, dim3(1), dim3(kThreads), 0, c.stream,
                     (const double*)partB(c, 1), gv, c.scal.get(), (double*)nullptr);
  precond(0, 0);
  direction(1);
  HIP_CHECK(hipGetLastError());
  // a multilevel iteration is ~1.3 ms of ~100 launches: one of them hides the host's status
  // read (a graph per parity), and a stop leaves one iteration's launches in flight (0.7%
  // less time per iteration than chunks of two on the 4M-DOF plate; CSTFEM_FCG_CHUNK sets 1
  // or an even count)
  const char* env_chunk = std::getenv("CSTFEM_FCG_CHUNK");
  const int want = env_chunk ? std::atoi(env_chunk) : 1;
  const int chunk = multilevel ? (want == 1 || (want >= 2 && want % 2 == 0) ? want : 2)
                               : (c.n > 200000 ? 16 : 64);
  // the iteration graph is captured while the device still runs the initialisation and the
  // first preconditioner application queued above: the capture records launches without
  // running them, and the work queued before it does not wait (a multilevel capture and its
  // instantiation take about as long as one cycle)
  std::optional<Chunk> gr;
  if (o.max_iter > 0) gr.emplace(c, chunk, iteration);
  PcgScalars s = read_scalars(c);
  const double bnorm = s.bnorm;
  const bool zero_rhs = bnorm == 0.0;
  int outer = 0, last_status = kRunning, iters = 0;
  double true_rel = zero_rhs ? 0.0 : 1.0;
  bool converged = zero_rhs, fallback = false;
  if (!zero_rhs && o.max_iter > 0) {
    double prev_true = 1.0;
    int stagnant = 0;
    for (;;) {
      run_until_stopped(c, *gr, chunk, o.max_iter);
      s = read_scalars(c);
      last_status = s.status;
      const int k = s.k_in[0] > s.k_in[1] ? s.k_in[0] : s.k_in[1];
      iters = k;
      if (s.status != kRefine) break;  // breakdown, stall or max_iter
      // r64 = b - A x in FP64 (sum into sc->rr); the working residual follows
      spmv_fp64(c, c.x.get(), Ax);
      hipLaunchKernelGGL(k_residual, dim3(gv), dim3(kThreads), 0, c.stream, c.nb,
                         (const double2*)c.b.get(), (const double2*)Ax, (double2*)r64, partX(c));
      hipLaunchKernelGGL(k_sum_into, dim3(1), dim3(kThreads), 0, c.stream, partX(c), gv,
                         &c.scal.get()->rr);
      hipLaunchKernelGGL(k_fcg_set_r<P>, dim3(gv), dim3(kThreads), 0, c.stream, c.nb,
                         (const double2*)r64, r);
      HIP_CHECK(hipGetLastError());
      const PcgScalars s2 = read_scalars(c);
      true_rel = std::sqrt(s2.rr) / bnorm;
      ++outer;
      if (true_rel < o.rtol) {
        converged = true;
        break;
      }
      // near rtol the FP64 residual itself is at its rounding floor: refreshes there are
      // not stagnation (max_outer bounds them)
      stagnant = (true_rel > 0.5 * prev_true && true_rel > 4.0 * o.rtol) ? stagnant + 1 : 0;
      prev_true = true_rel;
      if (outer >= max_outer || k >= o.max_iter || stagnant >= 4) break;
      PcgScalars upd = s2;
      upd.status = kRunning;
      upd.rel_ref = true_rel;
      upd.best_rel[0] = upd.best_rel[1] = true_rel;  // progress is measured from the refresh
      upd.best_k[0] = upd.best_k[1] = k;
      upd.slow_rel[0] = upd.slow_rel[1] = true_rel;
      upd.slow_k[0] = upd.slow_k[1] = k;
      HIP_CHECK(hipMemcpyAsync(c.scal.get(), &upd, sizeof upd, hipMemcpyHostToDevice, c.stream));
      precond((k - 1) & 1, 0);  // z = M r with the refreshed residual, z'q against the last q
      direction(0);             // p = z + beta p: the direction is kept
      if (k & 1) iteration(1);  // the graph starts at parity 0
      HIP_CHECK(hipGetLastError());
      last_status = kRunning;
    }
  }
  int fb_iters = 0, fb_outer = 0, fb_kind = 0;
  // a fallback continues from the current x only if its FP64 true residual is below that of
  // x = 0 (a diverged iterate would leave the FP64 residual floor above rtol); else from 0
  auto fallback_start = [&]() {
    spmv_fp64(c, c.x.get(), Ax);
    hipLaunchKernelGGL(k_residual, dim3(gv), dim3(kThreads), 0, c.stream, c.nb,
                       (const double2*)c.b.get(), (const double2*)Ax, (double2*)r64, partX(c));
    hipLaunchKernelGGL(k_sum_into, dim3(1), dim3(kThreads), 0, c.stream, partX(c), gv,
                       &c.scal.get()->rr);
    HIP_CHECK(hipGetLastError());
    const double cur = std::sqrt(read_scalars(c).rr) / bnorm;
    if (cur < 1.0) return true;
    HIP_CHECK(hipMemsetAsync(c.x.get(), 0, sizeof(double) * nv, c.stream));
    return false;
  };
  // a stop at max_iter is not a failure of the preconditioner: no fallback then (the
  // fallbacks are for stalls and breakdowns, with iterations left)
  const bool spent = last_status == kMaxIter || iters >= o.max_iter;
  if (!converged && !zero_rhs && !spent && multilevel && c.mg.fine_smoother == 0) {
    // the learned smoother failed on this matrix: the same hierarchy with the Chebyshev
    // smoother, from the current x (its own stall falls back to Jacobi-CG)
    cstfem_solver_options fo = o;
    fo.max_iter = std::max(1, o.max_iter - iters);
    cstfem_solver_stats fs{};
    const bool warm_x = fallback_start();
    mg_set_fine_smoother(c, 1);
    try {
      fcg_run<P>(c, fo, fs, false, warm_x);
    } catch (...) {
      mg_set_fine_smoother(c, 0);
      throw;
    }
    mg_set_fine_smoother(c, 0);
    fb_iters = fs.iterations;
    fb_outer = fs.outer_iterations;
    converged = fs.converged != 0;
    true_rel = fs.true_rel_residual;
    last_status = converged ? kConverged : kMaxIter;
    fallback = true;
    fb_kind = fs.fallback ? fs.fallback : 2;
  } else if (!converged && !spent) {
    // breakdown or stall of the learned preconditioner: Jacobi-CG (FP32 + FP64 refinement)
    // from the current x; the FP64 true residual decides
    fallback = true;
    fb_kind = 1;
    cstfem_solver_options fo = o;
    fo.max_iter = std::max(1, o.max_iter - iters);
    fo.inner_rtol = 0.0;
    fo.refresh = 0;
    cstfem_solver_stats fs{};
    const bool warm_x = fallback_start();
    solve_refined(c, fo, fs, true, warm_x);
    fb_iters = fs.iterations;
    fb_outer = fs.outer_iterations;
    converged = fs.converged != 0;
    true_rel = fs.true_rel_residual;
    last_status = converged ? kConverged : kMaxIter;
  }
  if (!converged && !zero_rhs && !fallback && o.max_iter > 0) {
    // stopped without converging: the FP64 true residual of the x it stopped at (the last
    // refresh's, if any, belongs to an earlier x)
    spmv_fp64(c, c.x.get(), Ax);
    hipLaunchKernelGGL(k_residual, dim3(gv), dim3(kThreads), 0, c.stream, c.nb,
                       (const double2*)c.b.get(), (const double2*)Ax, (double2*)r64, partX(c));
    hipLaunchKernelGGL(k_sum_into, dim3(1), dim3(kThreads), 0, c.stream, partX(c), gv,
                       &c.scal.get()->rr);
    HIP_CHECK(hipGetLastError());
    true_rel = std::sqrt(read_scalars(c).rr) / bnorm;
  }
  s = read_scalars(c);
  HIP_CHECK(hipEventRecord(ev1, c.stream));
  HIP_CHECK(hipEventSynchronize(ev1));
  st.iterations = zero_rhs ? 0 : iters + fb_iters;
  st.outer_iterations = outer + fb_outer;
  st.converged = converged ? 1 : 0;
  st.breakdown = last_status == kBreakdown ? 1 : 0;
  st.fallback = fallback ? fb_kind : 0;
  st.zero_rhs = zero_rhs ? 1 : 0;
  st.rel_residual = zero_rhs ? 0.0 : s.rel;
  st.true_rel_residual = true_rel;
  st.spmv_calls = zero_rhs ? 0 : (long long)st.iterations + st.outer_iterations;
  st.total = elapsed(ev0, ev1);
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  return converged ? CSTFEM_OK : CSTFEM_NOT_CONVERGED;
}

}  // namespace

int solve_fcg_learned(Context& c, const cstfem_solver_options& o, cstfem_solver_stats& st,
                      bool fp32) {
  if (!c.setup_done) throw Error(CSTFEM_ERR_ORDER, "cg_jacobi: solve() called before setup()");
  const bool multilevel = !fp32 && c.mg.ready;
  const bool need_m = !multilevel || c.mg.fine_smoother == 0;
  if (need_m && !c.mpnn.built)
    throw Error(CSTFEM_ERR_ORDER, "fcg: build the learned preconditioner first (cstfem_mpnn_build)");
  return fp32 ? fcg_run<F32>(c, o, st, true) : fcg_run<F64>(c, o, st, false);
}

// Average duration (seconds) of `reps` back-to-back launches of one hot kernel on the
// library stream, between two HIP events.  kind: 0 SpMV + p'Ap partials (k_cg_spmv, FP64),
// 1 plain FP64 SpMV, 2 update (k_cg_update), 3 direction (k_cg_direction), 4 copy of
// `bytes` bytes (achievable HBM bandwidth), 5 FP32 SpMV + partials; flexible CG (FP64):
// 6 SpMV + p'q, p'r (k_fcg_spmv), 7 z = M r + z'q (k_fcg_precond, needs a built M),
// 8 update (k_fcg_update), 9 direction (k_fcg_direction).  Solver state is not modified
// (private scalars and vectors).
double profile_kernel(Context& c, int kind, int reps, int64_t bytes) {
  const bool stream = kind == 4 || kind == 10;  // copy or read of `bytes`, no matrix needed
  if (!stream && !c.have_pattern) throw Error(CSTFEM_ERR_ORDER, "profile: no matrix on the device");
  if (kind == 7 && !c.mpnn.built)
    throw Error(CSTFEM_ERR_ORDER, "profile: build the learned preconditioner first");
  if (!stream) ensure_vectors(c);
  DBuf<PcgScalars> sc;
  sc.alloc(1);
  PcgScalars init{};
  init.bnorm = 1.0;
  init.rtol = 0.0;
  init.max_iter = INT_MAX;
  HIP_CHECK(hipMemcpyAsync(sc.get(), &init, sizeof init, hipMemcpyHostToDevice, c.stream));
  DBuf<double> a, b, w, part;
  DBuf<float> a32, v32;
  const size_t nv = 2 * (size_t)c.nb;
  part.alloc(std::max<size_t>(6 * kMaxGrid, 4 * kReadGrid));  // k_read: 4 sums per workgroup
  if (stream) {
    const int64_t n2 = bytes / 16;  // double2 elements in `bytes`
    a.alloc(2 * n2);
    b.alloc(kind == 4 ? 2 * n2 : 1);
    HIP_CHECK(hipMemsetAsync(a.get(), 0, sizeof(double) * 2 * n2, c.stream));
  } else {
    a.alloc(nv);
    b.alloc(nv);
    w.alloc(nv);
    HIP_CHECK(hipMemsetAsync(a.get(), 0, sizeof(double) * nv, c.stream));
    HIP_CHECK(hipMemsetAsync(w.get(), 0, sizeof(double) * nv, c.stream));
    // positive partials: p'Ap > 0 and rz > 0, so the update kernel takes its full path
    const std::vector<double> ones(6 * kMaxGrid, 1.0);
    HIP_CHECK(hipMemcpyAsync(part.get(), ones.data(), sizeof(double) * ones.size(),
                             hipMemcpyHostToDevice, c.stream));
    HIP_CHECK(hipStreamSynchronize(c.stream));
    HIP_CHECK(hipMemcpyAsync(b.get(), c.inv_diag.get(), sizeof(double) * nv, hipMemcpyDeviceToDevice, c.stream));
    if (kind == 5) {
      a32.alloc(nv);
      v32.alloc(4 * c.sell_blocks);
      hipLaunchKernelGGL(k_round_values, dim3(2048), dim3(256), 0, c.stream,
                         (int64_t)c.sell_blocks, (const double*)c.sell_val.get(), v32.get());
      HIP_CHECK(hipMemsetAsync(a32.get(), 0, sizeof(float) * nv, c.stream));
    }
  }
  const int gs = spmv_grid(c), gv = vec_grid(c);
  auto launch = [&]() {
    switch (kind) {
      case 0:
        hipLaunchKernelGGL(k_cg_spmv<F64>, dim3(gs), dim3(kThreads), 0, c.stream, mat64(c),
                           (const double2*)c.inv_diag.get(), (double2*)a.get(), part.get(),
                           (const PcgScalars*)sc.get());
        break;
      case 1:
        hipLaunchKernelGGL(k_spmv<F64>, dim3(gs), dim3(kThreads), 0, c.stream, mat64(c),
                           (const double2*)c.inv_diag.get(), (double2*)a.get());
        break;
      case 2:
        hipLaunchKernelGGL(k_cg_update, dim3(gv), dim3(kThreads), 0, c.stream, c.nb,
                           (double2*)a.get(), (double2*)w.get(), (const double2*)b.get(),
                           (const double2*)b.get(), (const double2*)c.inv_diag.get(),
                           (const double*)part.get(), gs, (const double*)(part.get() + kMaxGrid),
                           part.get() + 3 * kMaxGrid, gv, sc.get());
        break;
      case 3:
        hipLaunchKernelGGL(k_cg_direction<F64>, dim3(gv), dim3(kThreads), 0, c.stream, c.nb,
                           (const double2*)b.get(), (const double2*)c.inv_diag.get(),
                           (double2*)a.get(), (const double*)(part.get() + kMaxGrid),
                           (const double*)(part.get() + 3 * kMaxGrid), gv, 0, sc.get(),
                           (double*)nullptr, 1);
        break;
      case 5: {
        Mat<F32> m32{c.sell_off.get(), c.sell_col.get(), v32.get(), c.sell_row0.get(),
                     c.sell_nrow.get(), c.nslices, c.nb};
        hipLaunchKernelGGL(k_cg_spmv<F32>, dim3(gs), dim3(kThreads), 0, c.stream, m32,
                           (const float2*)c.invd32.get(), (float2*)a32.get(), part.get(),
                           (const PcgScalars*)sc.get());
        break;
      }
      case 6:
        hipLaunchKernelGGL(k_fcg_spmv<F64>, dim3(gs), dim3(kThreads), 0, c.stream, mat64(c),
                           (const double2*)b.get(), (const double2*)w.get(), (double2*)a.get(),
                           part.get(), (const PcgScalars*)sc.get());
        break;
      case 7: {
        const Mat<F32> G{c.sell_off.get(), c.sell_col.get(), c.mpnn.gsell.get(),
                         c.sell_row0.get(), c.sell_nrow.get(), c.nslices, c.nb};
        hipLaunchKernelGGL(k_fcg_precond<F64>, dim3(gs), dim3(kThreads), 0, c.stream, G,
                           (const double2*)b.get(), (const double2*)w.get(), (double2*)a.get(),
                           (const double*)part.get(), gv, part.get() + 3 * kMaxGrid, 0, sc.get(),
                           0);
        break;
      }
      case 8:
        hipLaunchKernelGGL(k_fcg_update<F64>, dim3(gv), dim3(kThreads), 0, c.stream, c.nb,
                           (double2*)a.get(), (double2*)w.get(), (const double2*)b.get(),
                           (const double2*)b.get(), (const double*)part.get(), gs,
                           part.get() + 3 * kMaxGrid, sc.get());
        break;
      case 9:
        hipLaunchKernelGGL(k_fcg_direction<F64>, dim3(gv), dim3(kThreads), 0, c.stream, c.nb,
                           (const double2*)b.get(), (double2*)a.get(), (const double*)part.get(),
                           gs, (const double*)(part.get() + 3 * kMaxGrid), gs,
                           (const PcgScalars*)sc.get(), 0);
        break;
      case 10:
        // the fastest shape of a sweep over grids (1024-65536) and loads per lane (1, 2, 4),
        // non-temporal and plain: 2048 workgroups, one 16-byte non-temporal load per lane
        hipLaunchKernelGGL(k_read, dim3(kReadGrid), dim3(kThreads), 0, c.stream, bytes / 16,
                           (const double2*)a.get(), part.get());
        break;
      default: {
        const int64_t n2 = bytes / 16;
        hipLaunchKernelGGL(k_copy, dim3(kMaxGrid), dim3(kThreads), 0, c.stream, n2,
                           (const double2*)a.get(), (double2*)b.get());
      }
    }
  };
  launch();  // warm-up
  HIP_CHECK(hipGetLastError());
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  HIP_CHECK(hipEventRecord(e0, c.stream));
  for (int i = 0; i < reps; ++i) launch();
  HIP_CHECK(hipEventRecord(e1, c.stream));
  HIP_CHECK(hipEventSynchronize(e1));
  const double t = elapsed(e0, e1) / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return t;
}

}  // namespace cstfem_hip
