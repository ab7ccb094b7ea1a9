// mg.hip -- aggregation multilevel preconditioner for the flexible CG.
//
// Hierarchy (set up once per matrix):
// Set-up: the aggregation runs on the host (it is sequential by definition) over a
// page-locked copy of the node graph; everything else (Gram-Schmidt, coarse pattern, SELL
// layout, Galerkin product, spectral bounds) runs on the device.
//   aggregation   greedy over the node graph of each level, in node order: a node whose
//                 neighbours are all free becomes a root and takes its whole neighbourhood;
//                 leftovers join the aggregate of their first aggregated neighbour (column
//                 order); nodes with none start their own.
//   prolongator   unsmoothed: on each aggregate, the near-null space B restricted to its DOFs
//                 is orthonormalised by modified Gram-Schmidt, P = Q (3 columns per
//                 aggregate), and the coarse near-null space is R.  Level 0's B holds the three
//                 global rigid-body modes (ux, uy, rotation about the centre of the mesh's
//                 bounding box) with the rows of constrained DOFs zeroed; a column that vanishes (rank < 3) is dropped
//                 as a dead coarse DOF (zero row and column of the coarse matrix).
//   Galerkin      A_{l+1} = P^T A_l P on the device: one thread per coarse 3x3 block sums the
//                 contributions of its aggregate's fine blocks in a fixed order (no atomics,
//                 reproducible to the bit).
//   coarsest      dense inverse (Cholesky on the host) applied as a matrix-vector product.
// Cycle applied to r (all kernels on the library stream, no host synchronisation, so the
// flexible CG captures it in its iteration graph):
//   level 0:   z = S(r) [learned M, or 2 Chebyshev steps on D^-1 A]; z += P0 K_1(P0^T (r - A z));
//              z = z + S(r - A z)
//   level l:   2 Chebyshev steps, coarse correction, 2 steps.  The coarse correction of level
//              l is the dense solve on the coarsest level, the K-cycle of level l+1 (two
//              flexible CG steps preconditioned by the cycle of level l+1) while l+1 <= kdepth,
//              and one cycle of level l+1 (V-cycle) below that.
#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <numeric>
#include <type_traits>

#include "aggregate.hpp"
#include "internal.hpp"

namespace cstfem_hip {

namespace {

constexpr int kT = 256;
constexpr int kW = kT / 64;
constexpr int kPart = 2048;  // partial slots per reduction on a coarse level

typedef double v2d __attribute__((ext_vector_type(2)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ inline bool stopped(const PcgScalars* sc) { return sc && sc->status != kRunning; }

// The coarse levels' kernels are latency-bound: each dependent round trip to memory costs about
// a microsecond of a few-microsecond kernel.  They therefore read the status word together
// with the row's own operands (status_of, then the loads, then hold, then halted), so
// that both share one round trip, instead of testing the status before issuing anything.
__device__ int kNoStatus = kRunning;  // read when a launch has no status word (never written)
__device__ inline int status_of(const PcgScalars* sc) {
  // a per-lane zero offset makes this a vector load, whose wait counts in order with the row's
  // loads (a scalar load's wait would also hold back every later kernel-argument load); no
  // branch, so that nothing waits for it before the row's loads are issued
  const int* p = sc ? &sc->status : &kNoStatus;
  int z = 0;
  asm volatile("" : "+v"(z));
  return p[z];
}
__device__ inline bool halted(int status) { return status != kRunning; }
// The values are needed here: their loads, issued above together, complete here, and none of
// them may be moved below this point (past the status test that follows)
template <class X>
__device__ inline void hold1(X& v) {
  asm volatile("" : "+v"(v));
}
template <class... V>
__device__ inline void hold(V&... v) {
  (hold1(v), ...);
}

__device__ inline double block_sum(double v, double* lds) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < kW; ++i) s += lds[i];
  return s;
}

__device__ inline double sum_parts(const double* __restrict__ p, int n, double* lds) {
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += kT) s += p[i];
  return block_sum(s, lds);
}

// ---- level 0 (2x2-block SELL, rows through row0/nrow) ----------------------------------

struct Mat0 {
  const int64_t* off;
  const int* col;
  const double* val;
  const int* row0;
  const int* nrow;
  int nslices;
};

struct MatM {  // learned M: FP32 values in the same layout
  const int64_t* off;
  const int* col;
  const float* val;
  const int* row0;
  const int* nrow;
  int nslices;
};

__device__ inline void row0_spmv(const Mat0& A, const double2* __restrict__ x, int s, int lane,
                                 double& y0, double& y1) {
  const int64_t o = A.off[s];
  const int w = (int)((A.off[s + 1] - o) >> 6);
  const int* c = A.col + o + lane;
  const v2d* v = reinterpret_cast<const v2d*>(A.val + 4 * o) + lane;
  double a0 = 0.0, a1 = 0.0;
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

__device__ inline void rowM_spmv(const MatM& A, const double2* __restrict__ x, int s, int lane,
                                 double& y0, double& y1) {
  const int64_t o = A.off[s];
  const int w = (int)((A.off[s + 1] - o) >> 6);
  const int* c = A.col + o + lane;
  // the values stream once per application: non-temporal loads, 2 entries in flight (1.7%
  // faster than 4 for these 16-byte blocks; 1 or 8 are 15% slower)
  const v4f* v = reinterpret_cast<const v4f*>(A.val) + (o + lane);
  double a0 = 0.0, a1 = 0.0;
#pragma unroll 2
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

// Slices of workgroup b: the 8 XCDs receive workgroups round-robin, so with a grid of 8k
// workgroups (k >= 8) each XCD sweeps one contiguous eighth of the slices and the gathered x
// entries of neighbouring rows stay in its L2; other grids interleave
struct Slices {
  int first, end, step;
};
__device__ inline Slices xcd_slices(int nslices) {
  const int G = gridDim.x, wave = threadIdx.x >> 6;
  if (G >= 64 && (G & 7) == 0) {
    const int xcd = blockIdx.x & 7, idx = blockIdx.x >> 3, per = G >> 3;
    const int lo = (int)((int64_t)nslices * xcd / 8), hi = (int)((int64_t)nslices * (xcd + 1) / 8);
    return {lo + idx * kW + wave, hi, per * kW};
  }
  return {(int)blockIdx.x * kW + wave, nslices, G * kW};
}

// res = r - A x
__global__ __launch_bounds__(kT) void k0_resid(Mat0 A, const double2* __restrict__ x,
                                               const double2* __restrict__ r,
                                               double2* __restrict__ res,
                                               const PcgScalars* __restrict__ sc) {
  if (stopped(sc)) return;
  const int lane = threadIdx.x & 63;
  const Slices sl = xcd_slices(A.nslices);
  for (int s = sl.first; s < sl.end; s += sl.step) {
    double y0, y1;
    row0_spmv(A, x, s, lane, y0, y1);
    if (lane < A.nrow[s]) {
      const int i = A.row0[s] + lane;
      const double2 ri = r[i];
      res[i] = make_double2(ri.x - y0, ri.y - y1);
    }
  }
}

// res = r - A x with the FP32 copy of the matrix values (4 floats per block, the layout M
// uses), FP64 sums
__global__ __launch_bounds__(kT) void k0_resid32(MatM A, const double2* __restrict__ x,
                                                 const double2* __restrict__ r,
                                                 double2* __restrict__ res,
                                                 const PcgScalars* __restrict__ sc) {
  if (stopped(sc)) return;
  const int lane = threadIdx.x & 63;
  const Slices sl = xcd_slices(A.nslices);
  for (int s = sl.first; s < sl.end; s += sl.step) {
    double y0, y1;
    rowM_spmv(A, x, s, lane, y0, y1);
    if (lane < A.nrow[s]) {
      const int i = A.row0[s] + lane;
      const double2 ri = r[i];
      res[i] = make_double2(ri.x - y0, ri.y - y1);
    }
  }
}

// out = base + M in   (base == nullptr: out = M in); with dpart, also the block partials of
// out'dq (the flexible CG's z'q, formed where z is written)
__global__ __launch_bounds__(kT) void k0_msmooth(MatM M, const double2* __restrict__ in,
                                                 const double2* base, double2* out,
                                                 const double2* __restrict__ dq,
                                                 double* __restrict__ dpart,
                                                 const PcgScalars* __restrict__ sc) {
  __shared__ double lds[kW];
  if (stopped(sc)) return;
  const int lane = threadIdx.x & 63;
  double dot = 0.0;
  const Slices sl = xcd_slices(M.nslices);
  for (int s = sl.first; s < sl.end; s += sl.step) {
    double y0, y1;
    rowM_spmv(M, in, s, lane, y0, y1);
    if (lane < M.nrow[s]) {
      const int i = M.row0[s] + lane;
      double2 o = make_double2(y0, y1);
      if (base) {
        const double2 b = base[i];
        o.x = b.x + o.x;
        o.y = b.y + o.y;
      }
      out[i] = o;
      if (dpart) {
        const double2 qi = dq[i];
        dot += o.x * qi.x + o.y * qi.y;
      }
    }
  }
  if (dpart) {
    const double t = block_sum(dot, lds);
    if (threadIdx.x == 0) dpart[blockIdx.x] = t;
  }
}

// one Chebyshev step on D^-1 A: res = r - A xin (xin == nullptr: 0), d = cd d + cr D^-1 res,
// xout = xin + d
__global__ __launch_bounds__(kT) void k0_cheb(Mat0 A, const double2* __restrict__ xin,
                                              double2* __restrict__ xout, double2* __restrict__ d,
                                              const double2* __restrict__ r,
                                              const double2* __restrict__ dinv, double cd,
                                              double cr, const PcgScalars* __restrict__ sc) {
  if (stopped(sc)) return;
  const int lane = threadIdx.x & 63;
  const Slices sl = xcd_slices(A.nslices);
  for (int s = sl.first; s < sl.end; s += sl.step) {
    double y0 = 0.0, y1 = 0.0;
    if (xin) row0_spmv(A, xin, s, lane, y0, y1);
    if (lane < A.nrow[s]) {
      const int i = A.row0[s] + lane;
      const double2 ri = r[i], di = dinv[i];
      double2 dd = cd != 0.0 ? d[i] : make_double2(0.0, 0.0);
      dd.x = cd * dd.x + cr * (di.x * (ri.x - y0));
      dd.y = cd * dd.y + cr * (di.y * (ri.y - y1));
      d[i] = dd;
      const double2 xo = xin ? xin[i] : make_double2(0.0, 0.0);
      xout[i] = make_double2(xo.x + dd.x, xo.y + dd.y);
    }
  }
}

// ---- levels >= 1: scalar SELL-64 ---------------------------------------------------------

// Scalar SELL-64 matrix of a level >= 1 with values of type T (the FP64 values, or an FP32
// copy for the cycle's row sums).  Entries come in triples of one 3x3 block; bcol holds one
// block column per triple (value index q = off[s] + k * 64 + lane -> block column index
// off[s] / 3 + (k / 3) * 64 + lane).
template <class T>
struct MatST {
  const int64_t* off;
  const int* bcol;
  const T* val;
  int n;
  // column of entry k of the row at `lane` of the slice starting at value index o
  __device__ int col(int64_t o, int k, int lane) const {
    return 3 * bcol[o / 3 + (int64_t)(k / 3) * 64 + lane] + k % 3;
  }
};
using MatS = MatST<double>;
using MatS32 = MatST<float>;

// x entries read straight from a vector
struct XVec {
  const double* __restrict__ x;
  __device__ double operator()(int j) const { return x[j]; }
};

// x = the first Chebyshev step from zero, c1r D^-1 rhs, evaluated where it is needed (the
// same expression as the separate step, so the same bits); RF gives the right-hand side
template <class RF>
struct XCheb1 {
  const double* __restrict__ dinv;
  RF rhs;
  double c1r;
  __device__ double operator()(int j) const { return 0.0 + c1r * (dinv[j] * (rhs(j) - 0.0)); }
};

// rhs = res - alpha q: the K-cycle residual after its first step (ks_update's expression),
// evaluated where it is needed
struct RhsFold {
  const double* __restrict__ res;
  const double* __restrict__ q;
  double alpha;
  __device__ double operator()(int j) const { return res[j] - alpha * q[j]; }
};

// x2 + P e after the coarse correction (k_prolong's expression), evaluated where it is
// needed; e = the coarse level's kx, plus alpha kp when that level's last K-cycle update
// (ks_update's expression) is left to this reader
struct XProlong {
  const double* __restrict__ x2;
  const int* __restrict__ agg;
  const double* __restrict__ P;
  const double* __restrict__ ec;
  const double* __restrict__ kp;  // null: no update left
  double alpha;
  int bs;
  __device__ double e(int k) const { return kp ? ec[k] + alpha * kp[k] : ec[k]; }
  // from x2[f], the aggregate a of f's node and P's row f, already loaded
  __device__ double at(double x2f, int a, double p0, double p1, double p2) const {
    return x2f + (p0 * e(3 * a) + p1 * e(3 * a + 1) + p2 * e(3 * a + 2));
  }
  __device__ double operator()(int f) const {
    const int a = agg[f / bs];
    const double* p = P + (size_t)f * 3;
    return at(x2[f], a, p[0], p[1], p[2]);
  }
};

// A reduction's partials, this thread's first kSlots slots tid, tid + kT, ... loaded ahead of
// their sum, so that the loads share a round trip with the kernel's others.  sum() adds them
// in slot order and then any slots beyond kSlots kT, which it reads itself: sum_parts's value
// bit for bit, for any count.  kSlots covers the counts that occur (kPart for the K-cycle
// steps, 2 kPart for the z'q partials of a smoothing grid).
template <int kSlots>
struct PartLoad {
  double v[kSlots];
  const double* p = nullptr;
  __device__ void load(const double* __restrict__ part, int n) {
    p = part;
#pragma unroll
    for (int u = 0; u < kSlots; ++u) {
      const int i = threadIdx.x + u * kT;
      v[u] = i < n ? part[i] : 0.0;
    }
  }
  __device__ void hold() {
#pragma unroll
    for (int u = 0; u < kSlots; ++u) hold1(v[u]);
  }
  __device__ double sum(int n, double* lds) const {
    double s = 0.0;
#pragma unroll
    for (int u = 0; u < kSlots; ++u)
      if (threadIdx.x + u * kT < n) s += v[u];
    for (int i = threadIdx.x + kSlots * kT; i < n; i += kT) s += p[i];
    return block_sum(s, lds);
  }
};
using StepParts = PartLoad<kPart / kT>;     // a K-cycle step's p'q or p'r (<= kPart)
using DotParts = PartLoad<2 * kPart / kT>;  // z'q of a smoothing step (<= 2 kPart)

// alpha = p'res / p'q of a K-cycle step (0 unless p'q > 0)
__device__ inline double alpha_of(double pq, double pr) { return pq > 0.0 ? pr / pq : 0.0; }

// Entries k = k0, k0 + step, ... (k < w) of one SELL row in ascending order, accumulated
// into a in that order (FP64).  Loads are issued in groups of G (all column indices and
// values, then all x entries) so that a row costs two memory round trips per group instead
// of two per entry.
template <int G, class XF, class T>
__device__ inline double row_entries(const MatST<T>& A, const XF& xf, int64_t o, int lane, int w,
                                     int k0, int step, double a) {
  for (int k = k0; k < w; k += G * step) {
    int cidx[G];
    double v[G], xv[G];
#pragma unroll
    for (int u = 0; u < G; ++u) {
      const int kk = k + u * step < w ? k + u * step : k0;  // in-row index for the tail
      cidx[u] = A.col(o, kk, lane);
      v[u] = (double)A.val[o + (int64_t)kk * 64 + lane];
    }
#pragma unroll
    for (int u = 0; u < G; ++u) xv[u] = xf(cidx[u]);
#pragma unroll
    for (int u = 0; u < G; ++u)
      if (k + u * step < w) a = a + v[u] * xv[u];
  }
  return a;
}

// One whole SELL row, its entries taken block by block (a row holds 3 entries per 3x3 block,
// one block column each): the block columns and values of B blocks are loaded first, then
// their 3 B x entries, so that a row of up to B blocks costs two memory round trips; the
// products are added in ascending entry order, as row_entries adds them (the same sum, bit
// for bit).
template <int B, class XF, class T>
__device__ inline double row_blocks(const MatST<T>& A, const XF& xf, int64_t o, int lane, int w,
                                    double a) {
  for (int k = 0; k < w; k += 3 * B) {
    int bc[B];
    T v[3 * B];
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const int kb = k + 3 * u < w ? k + 3 * u : 0;  // past the row's end: block 0 again
      bc[u] = A.bcol[o / 3 + (int64_t)(kb / 3) * 64 + lane];
#pragma unroll
      for (int c = 0; c < 3; ++c) v[3 * u + c] = A.val[o + (int64_t)(kb + c) * 64 + lane];
    }
    double xv[3 * B];
#pragma unroll
    for (int u = 0; u < B; ++u)
#pragma unroll
      for (int c = 0; c < 3; ++c) xv[3 * u + c] = xf(3 * bc[u] + c);
#pragma unroll
    for (int u = 0; u < B; ++u)
      if (k + 3 * u < w)
#pragma unroll
        for (int c = 0; c < 3; ++c) a = a + (double)v[3 * u + c] * xv[3 * u + c];
  }
  return a;
}

// Where row i of a scalar level sits in its slice: value index o of the slice's first entry,
// entries per row w, lane
struct RowAt {
  int64_t o, o1;  // value index of the slice's first entry, and of the next slice's
  int lane;
  __device__ int w() const { return (int)((o1 - o) >> 6); }
};

template <class T>
__device__ inline RowAt row_at(const MatST<T>& A, int i) {
  const int s = i >> 6;
  return RowAt{A.off[s], A.off[s + 1], i & 63};
}

// one lane per row: 8 blocks (24 entries) per round, so that most rows (7 blocks in the
// interior of a plate's level 1) take one round of two memory round trips.  ks_cheb on level
// 1 of the 4M-DOF plate: 18.6 µs, against 19.6 µs with 4 blocks per round and 20.4 µs with
// row_entries' groups of 8 entries
template <class XF, class T>
__device__ inline double rowS_f(const MatST<T>& A, const XF& xf, const RowAt& ra) {
  return row_blocks<8>(A, xf, ra.o, ra.lane, ra.w(), 0.0);
}

template <class XF, class T>
__device__ inline double rowS_f(const MatST<T>& A, const XF& xf, int i) {
  return rowS_f(A, xf, row_at(A, i));
}

__device__ inline double rowS(const MatS& A, const double* __restrict__ x, int i) {
  return rowS_f(A, XVec{x}, i);
}

// Row i of A x computed by LPR consecutive lanes (lane j takes the entries k = j, j + LPR, ...;
// the partial sums are combined by a butterfly, so every lane of the group returns the sum).
// The LPR lanes of a group are always active together.
template <int LPR, class XF, class T>
__device__ inline double rowS_lanes(const MatST<T>& A, const XF& xf, const RowAt& ra, int j) {
  if (LPR == 1) return rowS_f(A, xf, ra);
  // rows hold about 21 entries: with 4 or more lanes per row, groups of 4 per lane cover a row
  // in one or two rounds
  const int w = ra.w();
  double a = j < w ? row_entries<(LPR >= 4 ? 4 : 8)>(A, xf, ra.o, ra.lane, w, j, LPR, 0.0) : 0.0;
#pragma unroll
  for (int m = 1; m < LPR; m <<= 1) a += __shfl_xor(a, m, 64);
  return a;
}

template <int LPR, class XF, class T>
__device__ inline double rowS_lanes(const MatST<T>& A, const XF& xf, int i, int j) {
  return rowS_lanes<LPR>(A, xf, row_at(A, i), j);
}

// Lanes per row on a scalar level of n rows (about 21 entries per row): one lane per row while
// there are enough rows to fill the chip, more lanes on small levels, where one lane's chain
// of loads would set the kernel's duration.
inline int lanes_per_row(int n) { return n >= 200000 ? 1 : n >= 50000 ? 2 : n >= 10000 ? 4 : 8; }

// With dpart, also the block partials of xout'dq (the K-cycle's z'q of the cycle's result).
template <int LPR, class T>
__global__ __launch_bounds__(kT) void ks_cheb(MatST<T> A, const double* __restrict__ xin,
                                              double* __restrict__ xout, double* __restrict__ d,
                                              const double* __restrict__ r,
                                              const double* __restrict__ dinv, double cd, double cr,
                                              const double* __restrict__ dq,
                                              double* __restrict__ dpart,
                                              const PcgScalars* __restrict__ sc) {
  __shared__ double lds[kW];
  const int t = blockIdx.x * kT + threadIdx.x;
  const int i = t / LPR, j = t % LPR;
  const bool live = i < A.n;
  int st = status_of(sc);
  RowAt ra{0, 0, 0};
  double di = 0.0, ri = 0.0, dold = 0.0, xi = 0.0, qi = 0.0;
  if (live) {  // every pointer read here is valid (r stands in for the absent ones)
    ra = row_at(A, i);
    di = dinv[i];
    ri = r[i];
    dold = d[i];
    xi = (xin ? xin : r)[i];
    qi = (dpart ? dq : r)[i];
  }
  hold(st, ra.o, ra.o1, di, ri, dold, xi, qi);
  if (halted(st)) return;
  double dot = 0.0;
  if (live) {
    const double y = xin ? rowS_lanes<LPR>(A, XVec{xin}, ra, j) : 0.0;
    if (j == 0) {
      const double dd = (cd != 0.0 ? cd * dold : 0.0) + cr * (di * (ri - y));
      d[i] = dd;
      const double xo = (xin ? xi : 0.0) + dd;
      xout[i] = xo;
      if (dpart) dot = xo * qi;
    }
  }
  if (dpart) {
    const double s = block_sum(dot, lds);
    if (threadIdx.x == 0) dpart[blockIdx.x] = s;
  }
}

// Both pre-smoothing steps from x = 0 in one pass: the first step's x (= c1r D^-1 rhs) is
// evaluated at every column the second step's row sum needs.  Writes d and x2 of the second
// step; identical in value to the two separate steps.
template <int LPR, class T>
__global__ __launch_bounds__(kT) void ks_cheb_pre(MatST<T> A, double* __restrict__ x2,
                                                  double* __restrict__ d,
                                                  const double* __restrict__ r,
                                                  const double* __restrict__ dinv, double c1r,
                                                  double c2d, double c2r,
                                                  const PcgScalars* __restrict__ sc) {
  const int t = blockIdx.x * kT + threadIdx.x;
  const int i = t / LPR, j = t % LPR;
  if (i >= A.n) return;
  int st = status_of(sc);
  RowAt ra = row_at(A, i);
  double di = dinv[i], ri = r[i];
  hold(st, ra.o, ra.o1, di, ri);
  if (halted(st)) return;
  const XCheb1<XVec> x1{dinv, XVec{r}, c1r};
  const double y = rowS_lanes<LPR>(A, x1, ra, j);
  if (j != 0) return;
  const double d1 = 0.0 + c1r * (di * (ri - 0.0));
  const double dd = (c2d != 0.0 ? c2d * d1 : 0.0) + c2r * (di * (ri - y));
  d[i] = dd;
  x2[i] = (0.0 + d1) + dd;
}

// ks_cheb_pre of the K-cycle's second cycle with the first step's update folded in: the
// right-hand side res - alpha q (RhsFold, alpha from the first step's partials) is formed
// where it is needed and also stored into res2, and kx = alpha p is written: what ks_update
// of the first step would have left, bit for bit.
template <int LPR, class T>
__global__ __launch_bounds__(kT) void ks_cheb_pre_fold(MatST<T> A, double* __restrict__ x2,
                                                       double* __restrict__ d,
                                                       const double* __restrict__ res,
                                                       const double* __restrict__ q,
                                                       const double* __restrict__ p,
                                                       const double* __restrict__ part, int np,
                                                       double* __restrict__ res2,
                                                       double* __restrict__ kx,
                                                       const double* __restrict__ dinv,
                                                       double c1r, double c2d, double c2r,
                                                       const PcgScalars* __restrict__ sc) {
  __shared__ double lds[kW];
  const int t = blockIdx.x * kT + threadIdx.x;
  const int i = t / LPR, j = t % LPR;
  const bool live = i < A.n;
  int st = status_of(sc);
  RowAt ra{0, 0, 0};
  double di = 0.0, resi = 0.0, qi = 0.0, pi = 0.0;
  if (live) {
    ra = row_at(A, i);
    di = dinv[i];
    resi = res[i];
    qi = q[i];
    pi = p[i];
  }
  StepParts pq, pr;
  pq.load(part, np);
  pr.load(part + kPart, np);
  hold(st, ra.o, ra.o1, di, resi, qi, pi);
  pq.hold();
  pr.hold();
  const RhsFold rf{res, q, alpha_of(pq.sum(np, lds), pr.sum(np, lds))};
  if (halted(st) || !live) return;
  const XCheb1<RhsFold> x1{dinv, rf, c1r};
  const double y = rowS_lanes<LPR>(A, x1, ra, j);
  if (j != 0) return;
  const double ri = resi - rf.alpha * qi;
  res2[i] = ri;
  kx[i] = 0.0 + rf.alpha * pi;
  const double d1 = 0.0 + c1r * (di * (ri - 0.0));
  const double dd = (c2d != 0.0 ? c2d * d1 : 0.0) + c2r * (di * (ri - y));
  d[i] = dd;
  x2[i] = (0.0 + d1) + dd;
}

// The first post-smoothing step (ks_cheb with cd = 0) with the prolongation folded in: its
// input x2 + P e (XProlong) is formed where the row sums need it, so k_prolong and its pass
// over x2 go away; with xp.kp, alpha of the coarse K-cycle's last step comes from its partials.
template <int LPR, class T>
__global__ __launch_bounds__(kT) void ks_cheb_prolong(MatST<T> A, XProlong xp,
                                                      const double* __restrict__ part, int np,
                                                      double* __restrict__ xout,
                                                      double* __restrict__ d,
                                                      const double* __restrict__ r,
                                                      const double* __restrict__ dinv, double cr,
                                                      const PcgScalars* __restrict__ sc) {
  __shared__ double lds[kW];
  const int t = blockIdx.x * kT + threadIdx.x;
  const int i = t / LPR, j = t % LPR;
  const bool live = i < A.n;
  int st = status_of(sc);
  RowAt ra{0, 0, 0};
  double di = 0.0, ri = 0.0, x2i = 0.0, p0 = 0.0, p1 = 0.0, p2 = 0.0;
  int ai = 0;
  if (live) {
    ra = row_at(A, i);
    di = dinv[i];
    ri = r[i];
    x2i = xp.x2[i];
    ai = xp.agg[i / xp.bs];
    p0 = xp.P[(size_t)i * 3];
    p1 = xp.P[(size_t)i * 3 + 1];
    p2 = xp.P[(size_t)i * 3 + 2];
  }
  StepParts pq, pr;
  if (xp.kp) {
    pq.load(part, np);
    pr.load(part + kPart, np);
  }
  hold(st, ra.o, ra.o1, di, ri, x2i, ai, p0, p1, p2);
  if (xp.kp) {
    pq.hold();
    pr.hold();
    xp.alpha = alpha_of(pq.sum(np, lds), pr.sum(np, lds));
  }
  if (halted(st) || !live) return;
  const double xi = xp.at(x2i, ai, p0, p1, p2);
  const double y = rowS_lanes<LPR>(A, xp, ra, j);
  if (j == 0) {
    const double dd = 0.0 + cr * (di * (ri - y));
    d[i] = dd;
    xout[i] = xi + dd;
  }
}

template <int LPR, class T>
__global__ __launch_bounds__(kT) void ks_resid(MatST<T> A, const double* __restrict__ x,
                                               const double* __restrict__ r,
                                               double* __restrict__ res,
                                               const PcgScalars* __restrict__ sc) {
  const int t = blockIdx.x * kT + threadIdx.x;
  const int i = t / LPR, j = t % LPR;
  if (i >= A.n) return;
  int st = status_of(sc);
  RowAt ra = row_at(A, i);
  double ri = r[i];
  hold(st, ra.o, ra.o1, ri);
  if (halted(st)) return;
  const double y = rowS_lanes<LPR>(A, XVec{x}, ra, j);
  if (j == 0) res[i] = ri - y;
}

// K-cycle step: beta = first ? 0 : -z'q_old / p'q_old; q = A z + beta q; p = z + beta p;
// partials p'q -> part[b], p'res -> part[kPart + b]
template <int LPR, class T>
__global__ __launch_bounds__(kT) void ks_spmv_dir(MatST<T> A, const double* __restrict__ z,
                                                  double* __restrict__ p, double* __restrict__ q,
                                                  const double* __restrict__ res,
                                                  double* __restrict__ part,
                                                  const double* __restrict__ part_zq, int nzq,
                                                  const double* __restrict__ part_pq, int np,
                                                  int first, const PcgScalars* __restrict__ sc) {
  __shared__ double lds[kW];
  const int j = threadIdx.x % LPR;
  const int stride = gridDim.x * (kT / LPR);
  int i = (blockIdx.x * kT + threadIdx.x) / LPR;
  int st = status_of(sc);
  // the first row's operands are loaded with the status word and the partials
  RowAt ra{0, 0, 0};
  double zi = 0.0, qo = 0.0, po = 0.0, ri = 0.0;
  auto load_row = [&]() {
    ra = row_at(A, i);
    zi = z[i];
    qo = q[i];
    po = p[i];
    ri = res[i];
  };
  if (i < A.n) load_row();
  DotParts zq;
  StepParts pq;
  if (!first) {
    zq.load(part_zq, nzq);
    pq.load(part_pq, np);
  }
  hold(st, ra.o, ra.o1, zi, qo, po, ri);
  double beta = 0.0;
  if (!first) {
    zq.hold();
    pq.hold();
    const double szq = zq.sum(nzq, lds);
    const double spq = pq.sum(np, lds);
    beta = spq > 0.0 ? -szq / spq : 0.0;
  }
  if (halted(st)) return;
  double a = 0.0, b = 0.0;
  while (i < A.n) {
    const double y = rowS_lanes<LPR>(A, XVec{z}, ra, j);
    if (j == 0) {
      const double qi = first ? y : y + beta * qo;
      const double pi = first ? zi : zi + beta * po;
      q[i] = qi;
      p[i] = pi;
      a += pi * qi;
      b += pi * ri;
    }
    i += stride;
    if (i < A.n) load_row();
  }
  const double sa = block_sum(a, lds);
  const double sb = block_sum(b, lds);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = sa;
    part[kPart + blockIdx.x] = sb;
  }
}

// alpha = p'res / p'q (0 unless p'q > 0); x += alpha p; res -= alpha q (if upd_res)
__global__ __launch_bounds__(kT) void ks_update(int n, double* __restrict__ x,
                                                double* __restrict__ res,
                                                const double* __restrict__ p,
                                                const double* __restrict__ q,
                                                const double* __restrict__ part, int np,
                                                int first, int upd_res,
                                                const PcgScalars* __restrict__ sc) {
  __shared__ double lds[kW];
  int st = status_of(sc);
  int i = blockIdx.x * kT + threadIdx.x;
  // the first element's operands are loaded with the status word and the partials
  double xo = 0.0, pi = 0.0, ro = 0.0, qi = 0.0;
  auto load = [&]() {
    xo = x[i];
    pi = p[i];
    ro = res[i];
    qi = q[i];
  };
  if (i < n) load();
  StepParts pq, pr;
  pq.load(part, np);
  pr.load(part + kPart, np);
  hold(st, xo, pi, ro, qi);
  pq.hold();
  pr.hold();
  const double alpha = alpha_of(pq.sum(np, lds), pr.sum(np, lds));
  if (halted(st)) return;
  while (i < n) {
    x[i] = (first ? 0.0 : xo) + alpha * pi;
    if (upd_res) res[i] = ro - alpha * qi;
    i += gridDim.x * kT;
    if (i < n) load();
  }
}

__global__ __launch_bounds__(kT) void ks_dot(int n, const double* __restrict__ a,
                                             const double* __restrict__ b,
                                             double* __restrict__ part,
                                             const PcgScalars* __restrict__ sc) {
  __shared__ double lds[kW];
  if (stopped(sc)) return;
  double v = 0.0;
  for (int i = blockIdx.x * kT + threadIdx.x; i < n; i += gridDim.x * kT) v += a[i] * b[i];
  const double s = block_sum(v, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// sum over the member rows k = k0, k0 + step, ... (k < rows) of the aggregate whose member
// list starts at mem[m0] (rows in member order, then DOF order) of P[f * 3 + c] res[f], in
// that order; loads in groups of 8
template <class PT>
__device__ inline double restrict_sum(int m0, int rows, int c, int bs,
                                      const int* __restrict__ mem, const PT* __restrict__ P,
                                      const double* __restrict__ res, int k0, int step) {
  double s = 0.0;
  for (int k = k0; k < rows; k += 8 * step) {
    int f[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int kk = k + u * step < rows ? k + u * step : k;
      f[u] = bs * mem[m0 + kk / bs] + kk % bs;
    }
    double p[8], r[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      p[u] = (double)P[(size_t)f[u] * 3 + c];
      r[u] = res[f[u]];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (k + u * step < rows) s = s + p[u] * r[u];
  }
  return s;
}

// rc[3a + c] = sum over member nodes i of a, k < bs:  P[(bs i + k) * 3 + c] * res[bs i + k];
// LPR lanes per coarse DOF split the member rows (butterfly sum, as in rowS_lanes)
template <int LPR, class PT>
__global__ __launch_bounds__(kT) void k_restrict(int nagg, int bs, const int* __restrict__ mptr,
                                                 const int* __restrict__ mem,
                                                 const PT* __restrict__ P,
                                                 const double* __restrict__ res,
                                                 double* __restrict__ rc,
                                                 const PcgScalars* __restrict__ sc) {
  const int tt = blockIdx.x * kT + threadIdx.x;
  const int t = tt / LPR, j = tt % LPR;
  if (t >= 3 * nagg) return;
  const int a = t / 3, c = t % 3;
  int st = status_of(sc);
  int m0 = mptr[a], m1 = mptr[a + 1];
  hold(st, m0, m1);
  if (halted(st)) return;
  double s = restrict_sum(m0, bs * (m1 - m0), c, bs, mem, P, res, j, LPR);
#pragma unroll
  for (int m = 1; m < LPR; m <<= 1) s += __shfl_xor(s, m, 64);
  if (j == 0) rc[t] = s;
}

// Residual rhs - A x of the member rows of aggregates [ag b, ag b + ag) (LPR lanes per row),
// kept in LDS with the rows' indices, then their restriction into rc with lr lanes per coarse
// DOF: each lane sums the rows k = j, j + lr, ... of its aggregate in member order and the
// lanes combine by butterfly, which is k_restrict<lr>'s arithmetic on ks_resid's values, so
// rc is bit-identical to the two separate kernels.  Dynamic LDS: cap doubles, then cap ints.
template <int LPR, class T>
__global__ __launch_bounds__(kT) void ks_resid_restrict(MatST<T> A, const double* __restrict__ x,
                                                        const double* __restrict__ r, int nagg,
                                                        int ag, int cap, int bs, int lr,
                                                        const int* __restrict__ mptr,
                                                        const int* __restrict__ mem,
                                                        const double* __restrict__ P,
                                                        double* __restrict__ rc,
                                                        const PcgScalars* __restrict__ sc) {
  extern __shared__ double rr_lds[];
  double* res = rr_lds;
  int* row = (int*)(rr_lds + cap);
  const int a0 = blockIdx.x * ag;
  const int a1 = min(a0 + ag, nagg);
  int st = status_of(sc);
  int m0 = mptr[a0], m1 = mptr[a1];
  hold(st, m0, m1);
  if (halted(st)) return;
  const int rows = bs * (m1 - m0);  // <= cap (planned on the host)
  const int j = threadIdx.x % LPR;
  for (int t = threadIdx.x / LPR; t < rows; t += kT / LPR) {
    const int f = bs * mem[m0 + t / bs] + t % bs;
    const double y = rowS_lanes<LPR>(A, XVec{x}, f, j);
    if (j == 0) {
      res[t] = r[f] - y;
      row[t] = f;
    }
  }
  __syncthreads();
  for (int tt = threadIdx.x; tt < 3 * (a1 - a0) * lr; tt += kT) {
    const int u = tt / lr, jj = tt % lr;
    const int a = a0 + u / 3, c = u % 3;
    const int k0 = bs * (mptr[a] - m0), k1 = bs * (mptr[a + 1] - m0);
    double s = 0.0;
    for (int k = k0 + jj; k < k1; k += 8 * lr) {
      double pv[8], rv[8];
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        const int kk = k + v * lr < k1 ? k + v * lr : k;
        pv[v] = P[(size_t)row[kk] * 3 + c];
        rv[v] = res[kk];
      }
#pragma unroll
      for (int v = 0; v < 8; ++v)
        if (k + v * lr < k1) s = s + pv[v] * rv[v];
    }
    for (int m = 1; m < lr; m <<= 1) s += __shfl_xor(s, m, 64);
    if (jj == 0) rc[3 * a + c] = s;
  }
}

// x[f] += sum_c P[f * 3 + c] * e[3 agg(node of f) + c] (XProlong with xp.x2 = x); with xp.kp,
// e = kx + alpha kp, alpha from the coarse K-cycle's last partials
__global__ __launch_bounds__(kT) void k_prolong(int n, XProlong xp, const float* __restrict__ P32,
                                                const double* __restrict__ part, int np,
                                                double* __restrict__ x,
                                                const PcgScalars* __restrict__ sc) {
  __shared__ double lds[kW];
  const int f = blockIdx.x * kT + threadIdx.x;
  int st = status_of(sc);
  double x2f = 0.0, p0 = 0.0, p1 = 0.0, p2 = 0.0;
  int a = 0;
  if (f < n) {
    x2f = xp.x2[f];
    a = xp.agg[f / xp.bs];
    if (P32) {  // P32 non-null: P's rows rounded to FP32 instead of xp.P
      p0 = (double)P32[(size_t)f * 3];
      p1 = (double)P32[(size_t)f * 3 + 1];
      p2 = (double)P32[(size_t)f * 3 + 2];
    } else {
      p0 = xp.P[(size_t)f * 3];
      p1 = xp.P[(size_t)f * 3 + 1];
      p2 = xp.P[(size_t)f * 3 + 2];
    }
  }
  StepParts pq, pr;
  if (xp.kp) {
    pq.load(part, np);
    pr.load(part + kPart, np);
  }
  hold(st, x2f, a, p0, p1, p2);
  if (xp.kp) {
    pq.hold();
    pr.hold();
    xp.alpha = alpha_of(pq.sum(np, lds), pr.sum(np, lds));
  }
  if (halted(st)) return;
  if (f < n) x[f] = xp.at(x2f, a, p0, p1, p2);
}

// sum_j row[j] r[j] in ascending j, loads issued in groups of 8
__device__ inline double dense_row(const double* __restrict__ row, const double* __restrict__ r,
                                   int n) {
  double s = 0.0;
  for (int j = 0; j < n; j += 8) {
    double a[8], b[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int jj = j + u < n ? j + u : j;
      a[u] = row[jj];
      b[u] = r[jj];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (j + u < n) s = s + a[u] * b[u];
  }
  return s;
}

// x = Cinv r (dense, row-major n x n)
__global__ __launch_bounds__(kT) void k_dense(int n, const double* __restrict__ Cinv,
                                              const double* __restrict__ r, double* __restrict__ x,
                                              const PcgScalars* __restrict__ sc) {
  if (stopped(sc)) return;
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  x[i] = dense_row(Cinv + (size_t)i * n, r, n);
}

// ---- fused tail: the V-cycle of the smallest levels in one workgroup ------------------------

constexpr int kSmootherSteps = 12;       // power steps for lambda_max(M A) of the learned smoother
constexpr double kSmootherAccept = 1.5;  // largest lambda_max(M A) accepted for the smoother

constexpr int kTailLevels = 4;       // cycle levels in the fused tail (plus the dense coarsest)
constexpr int kTailThreads = 1024;
constexpr int kTailLds = 64 * 1024;  // bytes of LDS for the tail's vectors
constexpr int kTailLanes = 8;        // lanes per coarse DOF (restriction) and per dense row
constexpr int kTailRegW = 21;        // row entries a tail thread keeps in registers (7 blocks)

template <class T>
struct TailLevel {
  MatST<T> A;
  const double* dinv;
  const int* mptr;
  const int* mem;
  const int* agg;
  const double* P;
  int nagg, bs;
  int lds;  // offset (doubles) of this level's vectors rhs, x, x2, d, res in LDS
  double c1r, c2d, c2r;
};

template <class T>
struct Tail {
  int nlev;
  TailLevel<T> L[kTailLevels];
  const double* cinv;
  int nc;
  int lds_c;          // offset of the coarsest rhs and solution in LDS
  const double* rhs;  // right-hand side of the top tail level (global)
  double* out;        // its result (global)
};

// The same steps as cycle_s (two Chebyshev steps, residual, restriction, coarse correction,
// prolongation, two steps) on each tail level in turn, with the dense solve at the bottom, in
// one workgroup: the vectors live in LDS, the matrices are read from L2, and barriers separate
// the phases.  The restriction and the dense rows use kTailLanes lanes each (butterfly sums).
// Row i of the tail's top level, loaded once into registers (kReg: that level has at most
// kTailThreads rows of at most kTailRegW entries): its four row sums then read only LDS.  The
// sums run over the entries in the order of row_entries, so they are bit-identical to rowS_f.
// The values are held as FP64 whatever the stored type (the conversion is exact).
template <class T>
struct RegRow {
  double v[kTailRegW];
  int b[kTailRegW / 3];
  int w = 0;
  __device__ void load(const MatST<T>& A, int i) {
    const int s = i >> 6, lane = i & 63;
    const int64_t o = A.off[s];
    w = (int)((A.off[s + 1] - o) >> 6);
#pragma unroll
    for (int k = 0; k < kTailRegW; ++k)
      v[k] = k < w ? (double)A.val[o + (int64_t)k * 64 + lane] : 0.0;
#pragma unroll
    for (int k = 0; k < kTailRegW / 3; ++k)
      b[k] = 3 * k < w ? A.bcol[o / 3 + (int64_t)k * 64 + lane] : 0;
  }
  __device__ double sum(const double* __restrict__ x) const {
    double a = 0.0;
#pragma unroll
    for (int k = 0; k < kTailRegW; ++k)
      if (k < w) a = a + v[k] * x[3 * b[k / 3] + k % 3];
    return a;
  }
};

// What the top level (kReg) keeps in registers besides its row: its diagonal scaling and the
// aggregate of its DOF, loaded at the kernel's start with the row (one round trip; a deeper
// chain of loads there, such as the restriction's member rows and P entries, costs more at
// the start than it saves later: the tail's data comes from HBM or the MALL at every call)
struct TopRegs {
  double dinv = 0.0;
  int agg = 0;
};

template <bool kReg, class T>
__global__ __launch_bounds__(kTailThreads) void k_tail(Tail<T> T_, const PcgScalars* __restrict__ sc) {
  extern __shared__ double lds[];
  const Tail<T>& TT = T_;
  const int tid = threadIdx.x;
  const TailLevel<T>& L0 = TT.L[0];
  const int n0 = L0.A.n;
  int st = status_of(sc);
  // the right-hand side (one entry per thread while the top level fits), the top level's rows
  // and operands and the dense inverse's staging are loaded with the status word
  double rhs_first = tid < n0 ? TT.rhs[tid] : 0.0;
  RegRow<T> rr;
  TopRegs top;
  if (kReg) {
    if (tid < n0) {
      rr.load(L0.A, tid);
      top.dinv = L0.dinv[tid];
      top.agg = L0.agg[tid / L0.bs];
    }
  }
  hold(st, rhs_first);
  if (halted(st)) return;
  // row sum of level l's row i: from registers on the top level (kReg), else from L2 (with
  // the top level's registers taken, in groups of 4 loads)
  auto row = [&](int l, const TailLevel<T>& L, const double* x, int i) {
    if (kReg && l == 0) return rr.sum(x);
    const RowAt ra = row_at(L.A, i);
    return row_entries<kReg ? 4 : 8>(L.A, XVec{x}, ra.o, ra.lane, ra.w(), 0, 1, 0.0);
  };
  // diagonal scaling of level l's row i
  auto dinv_of = [&](int l, const TailLevel<T>& L, int i) {
    return (kReg && l == 0) ? top.dinv : L.dinv[i];
  };
  {
    double* rhs0 = lds + L0.lds;
    if (tid < n0) rhs0[tid] = rhs_first;
    for (int i = tid + kTailThreads; i < n0; i += kTailThreads) rhs0[i] = TT.rhs[i];
  }
  __syncthreads();
  for (int l = 0; l < TT.nlev; ++l) {
    const TailLevel<T>& L = TT.L[l];
    const int n = L.A.n;
    double* rhs = lds + L.lds;
    double *x = rhs + n, *x2 = x + n, *d = x2 + n, *res = d + n;
    for (int i = tid; i < n; i += kTailThreads) {
      const double dd = 0.0 + L.c1r * (dinv_of(l, L, i) * (rhs[i] - 0.0));
      d[i] = dd;
      x[i] = 0.0 + dd;
    }
    __syncthreads();
    for (int i = tid; i < n; i += kTailThreads) {
      const double y = row(l, L, x, i);
      const double dd =
          (L.c2d != 0.0 ? L.c2d * d[i] : 0.0) + L.c2r * (dinv_of(l, L, i) * (rhs[i] - y));
      d[i] = dd;
      x2[i] = x[i] + dd;
    }
    __syncthreads();
    for (int i = tid; i < n; i += kTailThreads) res[i] = rhs[i] - row(l, L, x2, i);
    __syncthreads();
    double* rc = lds + (l + 1 < TT.nlev ? TT.L[l + 1].lds : TT.lds_c);
    for (int tt = tid; tt < 3 * L.nagg * kTailLanes; tt += kTailThreads) {
      const int t = tt / kTailLanes, j = tt % kTailLanes;
      const int m0 = L.mptr[t / 3];
      double sum = restrict_sum(m0, L.bs * (L.mptr[t / 3 + 1] - m0), t % 3, L.bs, L.mem, L.P, res,
                                j, kTailLanes);
#pragma unroll
      for (int m = 1; m < kTailLanes; m <<= 1) sum += __shfl_xor(sum, m, 64);
      if (j == 0) rc[t] = sum;
    }
    __syncthreads();
  }
  {
    const double* crhs = lds + TT.lds_c;
    double* ckx = lds + TT.lds_c + TT.nc;
    for (int tt = tid; tt < TT.nc * kTailLanes; tt += kTailThreads) {
      const int i = tt / kTailLanes, j = tt % kTailLanes;
      const double* row = TT.cinv + (size_t)i * TT.nc;
      double sum = 0.0;
      // entries k = j, j + kTailLanes, ... in order, their loads issued 8 at a time
      for (int k = j; k < TT.nc; k += 8 * kTailLanes) {
        double a[8], b[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int kk = k + u * kTailLanes < TT.nc ? k + u * kTailLanes : k;
          a[u] = row[kk];
          b[u] = crhs[kk];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (k + u * kTailLanes < TT.nc) sum = sum + a[u] * b[u];
      }
#pragma unroll
      for (int m = 1; m < kTailLanes; m <<= 1) sum += __shfl_xor(sum, m, 64);
      if (j == 0) ckx[i] = sum;
    }
  }
  __syncthreads();
  for (int l = TT.nlev - 1; l >= 0; --l) {
    const TailLevel<T>& L = TT.L[l];
    const int n = L.A.n;
    double* rhs = lds + L.lds;
    double *x = rhs + n, *x2 = x + n, *d = x2 + n, *res = d + n;
    // the level below left its result in its res vector (the coarsest: after its rhs)
    const double* e = l + 1 < TT.nlev ? lds + TT.L[l + 1].lds + 4 * TT.L[l + 1].A.n
                                      : lds + TT.lds_c + TT.nc;
    for (int f = tid; f < n; f += kTailThreads) {
      const int a = (kReg && l == 0) ? top.agg : L.agg[f / L.bs];
      const double* p = L.P + (size_t)f * 3;
      const double* ea = e + 3 * a;
      x2[f] = x2[f] + (p[0] * ea[0] + p[1] * ea[1] + p[2] * ea[2]);
    }
    __syncthreads();
    for (int i = tid; i < n; i += kTailThreads) {
      const double y = row(l, L, x2, i);
      const double dd = 0.0 + L.c1r * (dinv_of(l, L, i) * (rhs[i] - y));
      d[i] = dd;
      x[i] = x2[i] + dd;
    }
    __syncthreads();
    double* out = l == 0 ? TT.out : res;
    for (int i = tid; i < n; i += kTailThreads) {
      const double y = row(l, L, x, i);
      const double dd =
          (L.c2d != 0.0 ? L.c2d * d[i] : 0.0) + L.c2r * (dinv_of(l, L, i) * (rhs[i] - y));
      d[i] = dd;
      out[i] = x[i] + dd;
    }
    __syncthreads();
  }
}

// ---- packed tail: one cycle level above the dense coarsest ----------------------------------
//
// k_tail spends most of its time in chains of dependent loads: a slice offset before the row,
// a member list before the restriction's rows and P entries, the dense inverse in two rounds.
// When the tail holds a single cycle level (rows in registers) over a small enough coarsest
// level, every operand it reads -- the rows, their scaling, each fine DOF's aggregate and P
// row, each restriction lane's (row, P entry) pairs and each dense lane's inverse entries --
// is gathered once per set-up into one thread-major array (k_tail_pack), and k_tail1 issues
// all its loads at its start, in one round trip.  k_tail1 keeps k_tail's arithmetic and order,
// so its result is k_tail's to the bit.
constexpr int kPackR = 6;   // restriction entries per lane (aggregates of <= 48 rows)
constexpr int kPackD = 14;  // dense entries per lane (a coarsest level of <= 112 DOFs)
// slot e of thread t is element e * kTailThreads + t of the double array d or the int array i
constexpr int kPdRow = 0;                   // the row's values (kTailRegW, zero beyond w)
constexpr int kPdP = kPdRow + kTailRegW;    // P row of fine DOF t (3)
constexpr int kPdDense = kPdP + 3;          // dense lane t's inverse entries (kPackD)
constexpr int kPdDinv = kPdDense + kPackD;  // diagonal scaling of row t
constexpr int kPdSlots = kPdDinv + 1;
constexpr int kPiCol = 0;                     // first column (3 b) of each 3x3 block of the row
constexpr int kPiW = kPiCol + kTailRegW / 3;  // entries of row t
constexpr int kPiAgg = kPiW + 1;              // aggregate of fine DOF t
constexpr int kPiResN = kPiAgg + 1;           // entries of restriction lane t
constexpr int kPiRes = kPiResN + 1;           // their products' LDS slots 3 f + c (kPackR)
constexpr int kPiSlots = kPiRes + kPackR;

// One workgroup of kTailThreads: fills the slots of every thread from the level's arrays and
// the dense inverse (row-major nc x nc), in the orders k_tail reads them.
__global__ __launch_bounds__(kTailThreads) void k_tail_pack(TailLevel<double> L,
                                                            const double* __restrict__ cinv,
                                                            int nc, double* __restrict__ pd,
                                                            int* __restrict__ pi) {
  const int t = threadIdx.x;
  auto D = [&](int e) -> double& { return pd[(size_t)e * kTailThreads + t]; };
  auto I = [&](int e) -> int& { return pi[(size_t)e * kTailThreads + t]; };
  int w = 0, a = 0;
  double dinv = 0.0, p[3] = {0.0, 0.0, 0.0};
  if (t < L.A.n) {
    const int s = t >> 6, lane = t & 63;
    const int64_t o = L.A.off[s];
    w = (int)((L.A.off[s + 1] - o) >> 6);
    for (int k = 0; k < kTailRegW; ++k) D(kPdRow + k) = k < w ? L.A.val[o + (int64_t)k * 64 + lane] : 0.0;
    for (int k = 0; k < kTailRegW / 3; ++k)
      I(kPiCol + k) = 3 * k < w ? 3 * L.A.bcol[o / 3 + (int64_t)k * 64 + lane] : 0;
    dinv = L.dinv[t];
    a = L.agg[t / L.bs];
    for (int c = 0; c < 3; ++c) p[c] = L.P[(size_t)t * 3 + c];
  } else {
    for (int k = 0; k < kTailRegW; ++k) D(kPdRow + k) = 0.0;
    for (int k = 0; k < kTailRegW / 3; ++k) I(kPiCol + k) = 0;
  }
  I(kPiW) = w;
  I(kPiAgg) = a;
  D(kPdDinv) = dinv;
  for (int c = 0; c < 3; ++c) D(kPdP + c) = p[c];
  // restriction lane t: coarse DOF q = t / kTailLanes, its member rows f for k = j,
  // j + kTailLanes, ... (restrict_sum's order), each as the slot 3 f + q % 3 of the product
  // P[f][q % 3] res[f]; the host checked that they fit kPackR
  int rn = 0;
  if (t < 3 * L.nagg * kTailLanes) {
    const int q = t / kTailLanes, j = t % kTailLanes;
    const int m0 = L.mptr[q / 3], rows = L.bs * (L.mptr[q / 3 + 1] - m0);
    for (int k = j; k < rows && rn < kPackR; k += kTailLanes, ++rn)
      I(kPiRes + rn) = 3 * (L.bs * L.mem[m0 + k / L.bs] + k % L.bs) + q % 3;
  }
  I(kPiResN) = rn;
  for (int e = rn; e < kPackR; ++e) I(kPiRes + e) = 0;
  // dense lane t: row t / kTailLanes of the inverse, entries k = j, j + kTailLanes, ...
  for (int e = 0; e < kPackD; ++e) {
    const int i = t / kTailLanes, k = t % kTailLanes + e * kTailLanes;
    D(kPdDense + e) = (i < nc && k < nc) ? cinv[(size_t)i * nc + k] : 0.0;
  }
}

struct TailPack {
  const double* d;
  const int* i;
  int n;   // DOFs of the cycle level
  int nr;  // restriction lanes (3 nagg kTailLanes)
  int nc;  // DOFs of the coarsest level
  double c1r, c2d, c2r;
  const double* rhs;
  double* out;
};

// k_tail's cycle for one level in registers (nlev = 1, kReg), operands from the pack.  The
// restriction sums products P[f][c] res[f] that thread f forms (the same rounded products
// k_tail adds).  The dense inverse's entries are loaded after the first barrier: they arrive
// during the smoothing steps, and the registers they take are not live beside the prologue's.
// LDS: x, x2 (n each), the products (3 n), the coarsest level's right-hand side and solution
// (nc each).
__global__ __launch_bounds__(kTailThreads) void k_tail1(TailPack T, const PcgScalars* __restrict__ sc) {
  extern __shared__ double lds[];
  const int t = threadIdx.x, n = T.n;
  double* x = lds;
  double* x2 = x + n;
  double* prod = x2 + n;
  double* crhs = prod + 3 * n;
  double* ckx = crhs + T.nc;
  const bool live = t < n;
  int st = status_of(sc);
  double rhs = live ? T.rhs[t] : 0.0;
  const double* __restrict__ pd = T.d + t;
  const int* __restrict__ pi = T.i + t;
  double v[kTailRegW];
  int b[kTailRegW / 3], rf[kPackR];
#pragma unroll
  for (int k = 0; k < kTailRegW; ++k) v[k] = pd[(kPdRow + k) * kTailThreads];
#pragma unroll
  for (int k = 0; k < kTailRegW / 3; ++k) b[k] = pi[(kPiCol + k) * kTailThreads];
  const int w = pi[kPiW * kTailThreads];
  const int agg = pi[kPiAgg * kTailThreads];
  const int rn = pi[kPiResN * kTailThreads];
  const double dinv = pd[kPdDinv * kTailThreads];
  const double p0 = pd[kPdP * kTailThreads], p1 = pd[(kPdP + 1) * kTailThreads],
               p2 = pd[(kPdP + 2) * kTailThreads];
#pragma unroll
  for (int e = 0; e < kPackR; ++e) rf[e] = pi[(kPiRes + e) * kTailThreads];
  hold(st, rhs);
  if (halted(st)) return;
  // row sum in RegRow::sum's order
  auto row = [&](const double* xv) {
    double a = 0.0;
#pragma unroll
    for (int k = 0; k < kTailRegW; ++k)
      if (k < w) a = a + v[k] * xv[b[k / 3] + k % 3];
    return a;
  };
  double d = 0.0;
  // pre-smoothing: two Chebyshev steps from zero
  if (live) {
    const double dd = 0.0 + T.c1r * (dinv * (rhs - 0.0));
    d = dd;
    x[t] = 0.0 + dd;
  }
  __syncthreads();
  double cd[kPackD];
#pragma unroll
  for (int e = 0; e < kPackD; ++e) cd[e] = pd[(kPdDense + e) * kTailThreads];
  if (live) {
    const double y = row(x);
    const double dd = (T.c2d != 0.0 ? T.c2d * d : 0.0) + T.c2r * (dinv * (rhs - y));
    d = dd;
    x2[t] = x[t] + dd;
  }
  __syncthreads();
  if (live) {
    const double r = rhs - row(x2);
    prod[3 * t] = p0 * r;
    prod[3 * t + 1] = p1 * r;
    prod[3 * t + 2] = p2 * r;
  }
  __syncthreads();
  // restriction, kTailLanes lanes per coarse DOF
  if (t < T.nr) {
    double s = 0.0;
#pragma unroll
    for (int e = 0; e < kPackR; ++e)
      if (e < rn) s = s + prod[rf[e]];
#pragma unroll
    for (int m = 1; m < kTailLanes; m <<= 1) s += __shfl_xor(s, m, 64);
    if (t % kTailLanes == 0) crhs[t / kTailLanes] = s;
  }
  __syncthreads();
  // dense solve on the coarsest level, kTailLanes lanes per row
  if (t < T.nc * kTailLanes) {
    const int j = t % kTailLanes;
    double s = 0.0;
#pragma unroll
    for (int e = 0; e < kPackD; ++e)
      if (j + e * kTailLanes < T.nc) s = s + cd[e] * crhs[j + e * kTailLanes];
#pragma unroll
    for (int m = 1; m < kTailLanes; m <<= 1) s += __shfl_xor(s, m, 64);
    if (j == 0) ckx[t / kTailLanes] = s;
  }
  __syncthreads();
  if (live) {
    const double* ea = ckx + 3 * agg;
    x2[t] = x2[t] + (p0 * ea[0] + p1 * ea[1] + p2 * ea[2]);
  }
  __syncthreads();
  // post-smoothing
  if (live) {
    const double y = row(x2);
    const double dd = 0.0 + T.c1r * (dinv * (rhs - y));
    d = dd;
    x[t] = x2[t] + dd;
  }
  __syncthreads();
  if (live) {
    const double y = row(x);
    const double dd = (T.c2d != 0.0 ? T.c2d * d : 0.0) + T.c2r * (dinv * (rhs - y));
    T.out[t] = x[t] + dd;
  }
}

// ---- setup kernels -------------------------------------------------------------------

// ---- wavefront helpers of the set-up kernels -----------------------------------------

constexpr int kMaxCoarseDeg = 128;  // block columns of one coarse block row
constexpr int kSetupMembers = 128;  // members of an aggregate one wavefront takes (two per lane)
constexpr int kPairSlots = 8;       // (member, block) pairs per lane held in registers

// v of lane `src` (wave-uniform) in every lane
__device__ inline double lane_value(double v, int src) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(unsigned int)b, src);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ inline int wave_min(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
  return v;
}

__device__ inline int wave_inclusive_sum(int v, int lane) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int u = __shfl_up(v, off, 64);
    if (lane >= off) v += u;
  }
  return v;
}

// The (member, block) pairs of an aggregate, in member order and then block order, as one
// wavefront numbers them: member t (node node[t], blocks first[t] ..) owns the pairs
// incl[t] - (its block count) .. incl[t] - 1, so pair p of member t is block
// p + first[t] - (incl[t] - count).  LDS arrays of kSetupMembers entries (nm <= kSetupMembers);
// returns the pair count.  The caller synchronises before reading the arrays.
struct MemberPairs {
  int incl[kSetupMembers], base[kSetupMembers], node[kSetupMembers], first[kSetupMembers];
};

__device__ inline int member_pairs(int m0, int nm, const int* __restrict__ mem,
                                   const int* __restrict__ bptr, MemberPairs& mp, int lane) {
  int total = 0;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int t = lane + 64 * h;
    int d = 0, q0 = 0, i = 0;
    if (t < nm) {
      i = mem[m0 + t];
      q0 = bptr[i];
      d = bptr[i + 1] - q0;
    }
    const int in = wave_inclusive_sum(d, lane) + total;
    if (t < nm) {
      mp.incl[t] = in;
      mp.base[t] = q0 - (in - d);
      mp.node[t] = i;
      mp.first[t] = q0;
    }
    total = __shfl(in, 63, 64);
  }
  return total;
}

// member of pair p: the first t with incl[t] > p
__device__ inline int pair_member(const MemberPairs& mp, int nm, int p) {
  int lo = 0, hi = nm - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (mp.incl[mid] > p) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// Galerkin contributions P_i^T A_q P_j of one fine block q = (i, j), as 3x3 row-major C.
// Level 0: 2x2 blocks in the planar SELL-64 layout (spos = each block's position).
struct RapFine0 {
  const int64_t* __restrict__ spos;
  const double* __restrict__ fval;
  __device__ void operator()(int i, int q, int j, int, const double* __restrict__ P,
                             double (&C)[9]) const {
    const int64_t sq = spos[q];
    const double Aq[4] = {fval[sell_vidx(sq, 0)], fval[sell_vidx(sq, 1)], fval[sell_vidx(sq, 2)],
                          fval[sell_vidx(sq, 3)]};
    const double* Pi = P + (size_t)(2 * i) * 3;
    const double* Pj = P + (size_t)(2 * j) * 3;
    // T = A_q P_j (2x3), C = P_i^T T
    double T[2][3];
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
      for (int c = 0; c < 3; ++c) T[k][c] = Aq[2 * k] * Pj[c] + Aq[2 * k + 1] * Pj[3 + c];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) C[3 * c + cc] = Pi[c] * T[0][cc] + Pi[3 + c] * T[1][cc];
  }
};

// Levels >= 1: 3x3 blocks in scalar SELL-64; block q of fine row i sits at position fp of
// that row.
struct RapFineS {
  const int64_t* __restrict__ foff;
  const double* __restrict__ fval;
  __device__ void operator()(int i, int, int j, int fp, const double* __restrict__ P,
                             double (&C)[9]) const {
    double Aq[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int row = 3 * i + k, s = row >> 6, lane = row & 63;
#pragma unroll
      for (int kk = 0; kk < 3; ++kk) Aq[k][kk] = fval[foff[s] + (int64_t)(3 * fp + kk) * 64 + lane];
    }
    const double* Pi = P + (size_t)(3 * i) * 3;
    const double* Pj = P + (size_t)(3 * j) * 3;
    double T[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int c = 0; c < 3; ++c)
        T[k][c] = Aq[k][0] * Pj[c] + Aq[k][1] * Pj[3 + c] + Aq[k][2] * Pj[6 + c];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int cc = 0; cc < 3; ++cc)
        C[3 * c + cc] = Pi[c] * T[0][cc] + Pi[3 + c] * T[1][cc] + Pi[6 + c] * T[2][cc];
  }
};

// Galerkin product A_{l+1} = P^T A_l P, one wavefront (one workgroup) per coarse block row a.
// Coarse block (a, b) is the sum of the contributions of the fine blocks q = (i, j) with i a
// member of a and agg(j) = b, added one at a time in member order and then block order.
// Lanes compute the contributions of 64 pairs at a time; lane t owns the coarse blocks at
// positions t and 64 + t (kSlots = 2: rows of more than 64 blocks) of the row and adds its
// pairs in that order.
template <class Fine, int kSlots>
__global__ __launch_bounds__(64) void k_rap(const int* __restrict__ mptr,
                                            const int* __restrict__ mem,
                                            const int* __restrict__ agg,
                                            const int* __restrict__ bptr,
                                            const int* __restrict__ bcol, Fine fine,
                                            const double* __restrict__ P,
                                            const int* __restrict__ cptr,
                                            const int* __restrict__ ccol,
                                            const int64_t* __restrict__ coff,
                                            double* __restrict__ cval) {
  __shared__ MemberPairs mp;
  __shared__ int cols[kMaxCoarseDeg];
  __shared__ int order[64];
  __shared__ double cl[64 * 9];
  const int a = blockIdx.x, lane = threadIdx.x;
  const int m0 = mptr[a], nm = mptr[a + 1] - m0;
  const int c0 = cptr[a], deg = cptr[a + 1] - c0;
  // the set-up has rejected larger aggregates and coarse rows before this runs
  if (nm > kSetupMembers || deg > 64 * kSlots || deg <= 0) return;
  for (int t = lane; t < deg; t += 64) cols[t] = ccol[c0 + t];
  const int T = member_pairs(m0, nm, mem, bptr, mp, lane);
  __syncthreads();
  double acc[kSlots][9];
#pragma unroll
  for (int h = 0; h < kSlots; ++h)
#pragma unroll
    for (int e = 0; e < 9; ++e) acc[h][e] = 0.0;
  const uint64_t below = (1ull << lane) - 1;
  for (int p0 = 0; p0 < T; p0 += 64) {
    const int p = p0 + lane;
    const bool valid = p < T;
    int pos = -1;
    if (valid) {
      const int t = pair_member(mp, nm, p);
      const int i = mp.node[t];
      const int q = mp.base[t] + p;
      const int j = bcol[q];
      const int b = agg[j];
      int lo = 0, hi = deg - 1;  // position of b in the ascending row
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (cols[mid] < b) lo = mid + 1;
        else hi = mid;
      }
      pos = lo;
      double C[9];
      fine(i, q, j, q - mp.first[t], P, C);
#pragma unroll
      for (int e = 0; e < 9; ++e) cl[9 * lane + e] = C[e];
    }
    // the chunk's pairs grouped by position, each group in pair order
    int run = 0, slot = 0, start0 = 0, cnt0 = 0, start1 = 0, cnt1 = 0;
    for (int v = 0; v < deg; ++v) {
      const uint64_t mask = __ballot(valid && pos == v);
      const int c = __popcll(mask);
      if (valid && pos == v) slot = run + __popcll(mask & below);
      if (v == lane) {
        start0 = run;
        cnt0 = c;
      }
      if (v == lane + 64) {
        start1 = run;
        cnt1 = c;
      }
      run += c;
    }
    if (valid) order[slot] = lane;
    __syncthreads();
    for (int k = 0; k < cnt0; ++k) {
      const double* src = cl + 9 * order[start0 + k];
#pragma unroll
      for (int e = 0; e < 9; ++e) acc[0][e] += src[e];
    }
    if (kSlots == 2)
      for (int k = 0; k < cnt1; ++k) {
        const double* src = cl + 9 * order[start1 + k];
#pragma unroll
        for (int e = 0; e < 9; ++e) acc[kSlots - 1][e] += src[e];
      }
    __syncthreads();
  }
#pragma unroll
  for (int h = 0; h < kSlots; ++h) {
    const int pos = lane + 64 * h;
    if (pos >= deg) continue;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int row = 3 * a + c, s = row >> 6, ln = row & 63;
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) cval[coff[s] + (int64_t)(3 * pos + cc) * 64 + ln] = acc[h][3 * c + cc];
    }
  }
}

// diagonal of a scalar SELL matrix (padding entries carry the row's own column and zero)
__global__ void ks_dinv(MatS A, double* __restrict__ dinv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  const int s = i >> 6, lane = i & 63;
  const int64_t o = A.off[s];
  const int w = (int)((A.off[s + 1] - o) >> 6);
  double d = 0.0;
  for (int k = 0; k < w; ++k) {
    const int64_t q = o + (int64_t)k * 64 + lane;
    if (A.col(o, k, lane) == i) d += A.val[q];
  }
  dinv[i] = d > 0.0 ? 1.0 / d : 0.0;
}

// ---- hierarchy set-up on the device ------------------------------------------------------

// Entry (row, c) of a level's near-null space: level 0 (B == nullptr) evaluates the global
// rigid-body modes from the node coordinates (x translation, y translation, rotation about
// (cx, cy)), zero on constrained DOFs; coarser levels read the R factors of the level above.
__device__ inline double nns_entry(int row, int c, const double* __restrict__ B,
                                   const double* __restrict__ xy,
                                   const unsigned char* __restrict__ fixed, double cx, double cy) {
  if (B) return B[(size_t)row * 3 + c];
  if (fixed && fixed[row]) return 0.0;
  const int i = row >> 1;
  if ((row & 1) == 0) return c == 0 ? 1.0 : (c == 1 ? 0.0 : -(xy[2 * (size_t)i + 1] - cy));
  return c == 0 ? 0.0 : (c == 1 ? 1.0 : xy[2 * (size_t)i] - cx);
}

// sum of x over the rows k = 0 .. m-1 of a wavefront (row k in lane k % 64, slot k / 64),
// added one at a time in row order, as one thread would
__device__ inline double row_order_sum(const double (&x)[3], int m) {
  double s = 0.0;
#pragma unroll
  for (int sl = 0; sl < 3; ++sl) {
    const int cnt = min(64, m - 64 * sl);
    for (int k = 0; k < cnt; ++k) s += lane_value(x[sl], k);
  }
  return s;
}

// Modified Gram-Schmidt of the near-null space on aggregate a (rows in member order, then DOF
// order): Q into P (n x 3), R (3x3 row-major) into Rn[9a..].  Every sum runs over the rows in
// order, one addition at a time.  A column whose remaining norm is <= 1e-10 of its original
// norm (or zero) is dead: zero in Q and a zero row in R.  Two forms with the same arithmetic:
// one wavefront per aggregate (row k in lane k % 64), and one lane per aggregate with its rows
// in registers (aggregates of up to kLaneRows rows; the larger ones go to a list for the
// wavefront form).
struct MgsArgs {
  int bs;
  const int* __restrict__ mptr;
  const int* __restrict__ mem;
  const double* __restrict__ B;
  const double* __restrict__ xy;
  const unsigned char* __restrict__ fixed;
  double cx, cy;
  double* __restrict__ P;
  double* __restrict__ Rn;
};

__device__ inline void mgs_wave(int a, int lane, const MgsArgs& g) {
  const int bs = g.bs;
  const int m0 = g.mptr[a], m = bs * (g.mptr[a + 1] - m0);
  if (m > 3 * 64) return;  // the set-up rejects aggregates above 192 DOFs
  int row[3];
  double v[3][3];
#pragma unroll
  for (int sl = 0; sl < 3; ++sl) {
    const int k = lane + 64 * sl;
    row[sl] = k < m ? bs * g.mem[m0 + k / bs] + k % bs : -1;
#pragma unroll
    for (int c = 0; c < 3; ++c)
      v[sl][c] = row[sl] >= 0 ? nns_entry(row[sl], c, g.B, g.xy, g.fixed, g.cx, g.cy) : 0.0;
  }
  double R[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  bool alive[3] = {false, false, false};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double x[3];
#pragma unroll
    for (int sl = 0; sl < 3; ++sl) x[sl] = v[sl][c] * v[sl][c];
    const double n0 = sqrt(row_order_sum(x, m));
#pragma unroll
    for (int cc = 0; cc < c; ++cc) {
      if (!alive[cc]) continue;
#pragma unroll
      for (int sl = 0; sl < 3; ++sl) x[sl] = v[sl][cc] * v[sl][c];
      const double dot = row_order_sum(x, m);
      R[cc][c] = dot;
#pragma unroll
      for (int sl = 0; sl < 3; ++sl) v[sl][c] -= dot * v[sl][cc];
    }
#pragma unroll
    for (int sl = 0; sl < 3; ++sl) x[sl] = v[sl][c] * v[sl][c];
    const double nr = sqrt(row_order_sum(x, m));
    const bool ok = n0 > 0.0 && nr > 1e-10 * n0;
    alive[c] = ok;
    if (ok) R[c][c] = nr;
#pragma unroll
    for (int cc = 0; cc < 3; ++cc)
      if (!ok) R[c][cc] = 0.0;
#pragma unroll
    for (int sl = 0; sl < 3; ++sl) v[sl][c] = ok ? v[sl][c] / nr : 0.0;
  }
#pragma unroll
  for (int sl = 0; sl < 3; ++sl)
    if (row[sl] >= 0)
#pragma unroll
      for (int c = 0; c < 3; ++c) g.P[(size_t)row[sl] * 3 + c] = v[sl][c];
  double r = 0.0;
#pragma unroll
  for (int t = 0; t < 9; ++t)
    if (lane == t) r = alive[t / 3] ? R[t / 3][t % 3] : 0.0;
  if (lane < 9) g.Rn[9 * (size_t)a + lane] = r;
}

// one wavefront per aggregate: every aggregate (list == nullptr), or the *count listed ones
__global__ __launch_bounds__(kT) void k_mgs(int na, const int* __restrict__ list,
                                            const int* __restrict__ count, MgsArgs g) {
  const int nw = list ? *count : na;
  const int lane = threadIdx.x & 63;
  for (int w = (int)(((int64_t)blockIdx.x * kT + threadIdx.x) >> 6); w < nw;
       w += gridDim.x * kW)
    mgs_wave(list ? list[w] : w, lane, g);
}

constexpr int kLaneRows = 20;      // rows of an aggregate one lane holds (10 nodes of 2 DOFs)
constexpr int kMgsLaneMin = 4096;  // aggregates from which the lane form runs (level 0)
constexpr int kMgsListGrid = 256;  // workgroups of the wavefront form over a list

template <int BS>
__global__ __launch_bounds__(kT) void k_mgs_lanes(int na, MgsArgs g, int* __restrict__ big,
                                                  int* __restrict__ nbig) {
  const int a = blockIdx.x * kT + threadIdx.x;
  if (a >= na) return;
  const int m0 = g.mptr[a], m = BS * (g.mptr[a + 1] - m0);
  if (m > kLaneRows) {
    big[atomicAdd(nbig, 1)] = a;
    return;
  }
  int row[kLaneRows];
  double v[kLaneRows][3];
#pragma unroll
  for (int k = 0; k < kLaneRows; ++k) {
    row[k] = k < m ? BS * g.mem[m0 + k / BS] + k % BS : -1;
#pragma unroll
    for (int c = 0; c < 3; ++c)
      v[k][c] = row[k] >= 0 ? nns_entry(row[k], c, g.B, g.xy, g.fixed, g.cx, g.cy) : 0.0;
  }
  double R[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  bool alive[3] = {false, false, false};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double n0 = 0.0;
#pragma unroll
    for (int k = 0; k < kLaneRows; ++k)
      if (k < m) n0 += v[k][c] * v[k][c];
    n0 = sqrt(n0);
#pragma unroll
    for (int cc = 0; cc < c; ++cc) {
      if (!alive[cc]) continue;
      double dot = 0.0;
#pragma unroll
      for (int k = 0; k < kLaneRows; ++k)
        if (k < m) dot += v[k][cc] * v[k][c];
      R[cc][c] = dot;
#pragma unroll
      for (int k = 0; k < kLaneRows; ++k) v[k][c] -= dot * v[k][cc];
    }
    double nr = 0.0;
#pragma unroll
    for (int k = 0; k < kLaneRows; ++k)
      if (k < m) nr += v[k][c] * v[k][c];
    nr = sqrt(nr);
    const bool ok = n0 > 0.0 && nr > 1e-10 * n0;
    alive[c] = ok;
    if (ok) R[c][c] = nr;
#pragma unroll
    for (int cc = 0; cc < 3; ++cc)
      if (!ok) R[c][cc] = 0.0;
#pragma unroll
    for (int k = 0; k < kLaneRows; ++k) v[k][c] = ok ? v[k][c] / nr : 0.0;
  }
#pragma unroll
  for (int k = 0; k < kLaneRows; ++k)
    if (k < m)
#pragma unroll
      for (int c = 0; c < 3; ++c) g.P[(size_t)row[k] * 3 + c] = v[k][c];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) g.Rn[9 * (size_t)a + 3 * c + cc] = alive[c] ? R[c][cc] : 0.0;
}

// bandwidth of a node graph with ascending columns: max over rows of i - first and last - i.
// A bounded grid strides over the rows and each workgroup adds one atomic: atomics on one
// address serialise at the memory side, one per wavefront of 2M rows took 0.36 ms.
constexpr int kBandwidthGrid = 512;
__global__ __launch_bounds__(kT) void k_graph_bandwidth(int n, const int* __restrict__ ptr,
                                                        const int* __restrict__ col,
                                                        int* __restrict__ out) {
  __shared__ int lds[kW];
  int b = 0;
  for (int i = blockIdx.x * kT + threadIdx.x; i < n; i += gridDim.x * kT)
    if (ptr[i + 1] > ptr[i]) b = max(b, max(i - col[ptr[i]], col[ptr[i + 1] - 1] - i));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) b = max(b, __shfl_xor(b, off, 64));
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = b;
  __syncthreads();
  if (threadIdx.x == 0) {
    int m = lds[0];
#pragma unroll
    for (int w = 1; w < kW; ++w) m = max(m, lds[w]);
    if (m > 0) atomicMax(out, m);
  }
}

// y = x rounded to FP32
__global__ void k_round32(int64_t n, const double* __restrict__ x, float* __restrict__ y) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) y[i] = (float)x[i];
}

// member lists of the aggregates: counts, offsets (a scan), members in arrival order, then
// every list sorted ascending (the order of the Gram-Schmidt step, the restriction and the
// Galerkin sums); err = 1 if an aggregate holds more than 192 DOFs
__global__ void k_member_count(int n, const int* __restrict__ agg, int* __restrict__ cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) atomicAdd(&cnt[agg[i]], 1);
}

__global__ void k_member_fill(int n, const int* __restrict__ agg, const int* __restrict__ mptr,
                              int* __restrict__ cursor, int* __restrict__ mem) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int a = agg[i];
  mem[mptr[a] + atomicAdd(&cursor[a], 1)] = i;
}

__global__ void k_member_sort(int na, int bs, const int* __restrict__ mptr, int* __restrict__ mem,
                              int* __restrict__ err) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= na) return;
  const int m0 = mptr[a], m1 = mptr[a + 1];
  if (bs * (m1 - m0) > 192) atomicMax(err, 1);
  for (int k = m0 + 1; k < m1; ++k) {
    const int v = mem[k];
    int j = k - 1;
    while (j >= m0 && mem[j] > v) {
      mem[j + 1] = mem[j];
      --j;
    }
    mem[j + 1] = v;
  }
}

// Sorted distinct aggregates adjacent to aggregate a (over its members' block rows), by one
// thread into out[]; -1 if there are more than kMaxCoarseDeg.
__device__ inline int coarse_row(int a, const int* __restrict__ mptr, const int* __restrict__ mem,
                                 const int* __restrict__ bptr, const int* __restrict__ bcol,
                                 const int* __restrict__ agg, int* out) {
  int len = 0;
  for (int m = mptr[a]; m < mptr[a + 1]; ++m) {
    const int i = mem[m];
    for (int q = bptr[i]; q < bptr[i + 1]; ++q) {
      const int b = agg[bcol[q]];
      int j = len - 1;
      while (j >= 0 && out[j] > b) --j;
      if (j >= 0 && out[j] == b) continue;
      if (len == kMaxCoarseDeg) return -1;
      for (int t = len; t > j + 1; --t) out[t] = out[t - 1];
      out[j + 1] = b;
      ++len;
    }
  }
  return len;
}

// The same by one wavefront (one workgroup): the lanes take the (member, block) pairs, and the
// distinct aggregates come out in ascending order, one per step (the smallest above the last).
// Column t ends in lane t % 64 (col0 for t < 64, col1 above).  Aggregates with more pairs
// than the lanes' register slots go through coarse_row in lane 0, with out[] in LDS.
__device__ inline int coarse_row_wave(int a, const int* __restrict__ mptr,
                                      const int* __restrict__ mem, const int* __restrict__ bptr,
                                      const int* __restrict__ bcol, const int* __restrict__ agg,
                                      MemberPairs& mp, int* spill, int lane, int& col0,
                                      int& col1) {
  const int m0 = mptr[a], nm = mptr[a + 1] - m0;
  col0 = col1 = -1;
  int T = -1;
  if (nm <= kSetupMembers) T = member_pairs(m0, nm, mem, bptr, mp, lane);
  __syncthreads();
  if (T < 0 || T > 64 * kPairSlots) {
    int len = 0;
    if (lane == 0) len = coarse_row(a, mptr, mem, bptr, bcol, agg, spill);
    len = __shfl(len, 0, 64);
    __syncthreads();
    if (lane < len) col0 = spill[lane];
    if (lane + 64 < len) col1 = spill[lane + 64];
    return len;
  }
  int b[kPairSlots];
#pragma unroll
  for (int s = 0; s < kPairSlots; ++s) {
    const int p = lane + 64 * s;
    b[s] = INT_MAX;
    if (p < T) {
      const int t = pair_member(mp, nm, p);
      b[s] = agg[bcol[mp.base[t] + p]];
    }
  }
  int last = -1, len = 0;
  for (;;) {
    int mn = INT_MAX;
#pragma unroll
    for (int s = 0; s < kPairSlots; ++s)
      if (b[s] > last) mn = min(mn, b[s]);
    mn = wave_min(mn);
    if (mn == INT_MAX) break;
    if (len == kMaxCoarseDeg) return -1;
    if (lane == (len & 63)) {
      if (len < 64) col0 = mn;
      else col1 = mn;
    }
    last = mn;
    ++len;
  }
  return len;
}

// Coarse block pattern, one wavefront per aggregate: every aggregate (list == nullptr) or the
// *count listed ones.  k_cpat_count: the degrees (err = 1 above kMaxCoarseDeg);
// k_cpat_fill: the columns at cptr.
__global__ __launch_bounds__(64) void k_cpat_count(int na, const int* __restrict__ list,
                                                   const int* __restrict__ count,
                                                   const int* __restrict__ mptr,
                                                   const int* __restrict__ mem,
                                                   const int* __restrict__ bptr,
                                                   const int* __restrict__ bcol,
                                                   const int* __restrict__ agg,
                                                   int* __restrict__ deg, int* __restrict__ err) {
  __shared__ MemberPairs mp;
  __shared__ int spill[kMaxCoarseDeg];
  const int lane = threadIdx.x, nw = list ? *count : na;
  for (int w = blockIdx.x; w < nw; w += gridDim.x) {
    const int a = list ? list[w] : w;
    int col0, col1;
    const int len = coarse_row_wave(a, mptr, mem, bptr, bcol, agg, mp, spill, lane, col0, col1);
    if (lane == 0) {
      if (len < 0) atomicMax(err, 1);
      deg[a] = len < 0 ? 0 : len;
    }
  }
}

__global__ __launch_bounds__(64) void k_cpat_fill(int na, const int* __restrict__ list,
                                                  const int* __restrict__ count,
                                                  const int* __restrict__ mptr,
                                                  const int* __restrict__ mem,
                                                  const int* __restrict__ bptr,
                                                  const int* __restrict__ bcol,
                                                  const int* __restrict__ agg,
                                                  const int* __restrict__ cptr,
                                                  int* __restrict__ ccol) {
  __shared__ MemberPairs mp;
  __shared__ int spill[kMaxCoarseDeg];
  const int lane = threadIdx.x, nw = list ? *count : na;
  for (int w = blockIdx.x; w < nw; w += gridDim.x) {
    const int a = list ? list[w] : w;
    int col0, col1;
    const int len = coarse_row_wave(a, mptr, mem, bptr, bcol, agg, mp, spill, lane, col0, col1);
    const int o = cptr[a];
    if (lane < len) ccol[o + lane] = col0;
    if (lane + 64 < len) ccol[o + lane + 64] = col1;
  }
}

// The same with one lane per aggregate, for levels of many small aggregates: the lane keeps
// the row sorted in registers (up to kLaneDeg columns) and inserts each member's neighbour
// aggregates, loaded kLaneBlocks at a time; a row with more columns is listed (count form)
// for the wavefront form and skipped (fill form).
constexpr int kLaneDeg = 16;
constexpr int kLaneBlocks = 16;
constexpr int kCpatListGrid = 1024;
constexpr int kCpatLaneMin = 65536;  // aggregates from which the lane form runs

__device__ inline int coarse_row_lane(int a, const int* __restrict__ mptr,
                                      const int* __restrict__ mem, const int* __restrict__ bptr,
                                      const int* __restrict__ bcol, const int* __restrict__ agg,
                                      int (&r)[kLaneDeg]) {
  int len = 0;
#pragma unroll
  for (int t = 0; t < kLaneDeg; ++t) r[t] = INT_MAX;
  for (int m = mptr[a]; m < mptr[a + 1]; ++m) {
    const int i = mem[m];
    const int q1 = bptr[i + 1];
    for (int q0 = bptr[i]; q0 < q1; q0 += kLaneBlocks) {
      int bb[kLaneBlocks];
#pragma unroll
      for (int u = 0; u < kLaneBlocks; ++u) bb[u] = q0 + u < q1 ? bcol[q0 + u] : -1;
#pragma unroll
      for (int u = 0; u < kLaneBlocks; ++u) bb[u] = bb[u] >= 0 ? agg[bb[u]] : -1;
#pragma unroll
      for (int u = 0; u < kLaneBlocks; ++u) {
        const int b = bb[u];
        if (b < 0) continue;
        int below = 0;
        bool dup = false;
#pragma unroll
        for (int t = 0; t < kLaneDeg; ++t) {
          below += r[t] < b ? 1 : 0;  // unused slots hold INT_MAX
          dup = dup || r[t] == b;
        }
        if (dup) continue;
        if (len == kLaneDeg) return -1;
#pragma unroll
        for (int t = kLaneDeg - 1; t > 0; --t)
          if (t > below) r[t] = r[t - 1];
#pragma unroll
        for (int t = 0; t < kLaneDeg; ++t)
          if (t == below) r[t] = b;
        ++len;
      }
    }
  }
  return len;
}

__global__ __launch_bounds__(kT) void k_cpat_count_lanes(int na, const int* __restrict__ mptr,
                                                         const int* __restrict__ mem,
                                                         const int* __restrict__ bptr,
                                                         const int* __restrict__ bcol,
                                                         const int* __restrict__ agg,
                                                         int* __restrict__ deg,
                                                         int* __restrict__ big,
                                                         int* __restrict__ nbig) {
  const int a = blockIdx.x * kT + threadIdx.x;
  if (a >= na) return;
  int r[kLaneDeg];
  const int len = coarse_row_lane(a, mptr, mem, bptr, bcol, agg, r);
  if (len < 0) big[atomicAdd(nbig, 1)] = a;
  else deg[a] = len;
}

__global__ __launch_bounds__(kT) void k_cpat_fill_lanes(int na, const int* __restrict__ mptr,
                                                        const int* __restrict__ mem,
                                                        const int* __restrict__ bptr,
                                                        const int* __restrict__ bcol,
                                                        const int* __restrict__ agg,
                                                        const int* __restrict__ cptr,
                                                        int* __restrict__ ccol) {
  const int a = blockIdx.x * kT + threadIdx.x;
  if (a >= na) return;
  int r[kLaneDeg];
  const int len = coarse_row_lane(a, mptr, mem, bptr, bcol, agg, r);
  if (len < 0) return;
  const int o = cptr[a];
#pragma unroll
  for (int t = 0; t < kLaneDeg; ++t)
    if (t < len) ccol[o + t] = r[t];
}

// column indices of the scalar SELL-64 layout of a 3x3-block pattern: lane = row, block
// column t of block row a gives the columns 3 ccol[t] + c'; padding repeats the row's own
// column (value 0)
// block columns of the scalar SELL layout: row r = 3a + c of slice s takes the coarse block
// columns of aggregate a in order; padding slots (zero values) point at the row's own block
// (rows past n: the last block), so that every column index stays in range
__global__ void k_sell_cols(int n, int nslices, const int64_t* __restrict__ off,
                            const int* __restrict__ cptr, const int* __restrict__ ccol,
                            int* __restrict__ scol) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= 64 * nslices) return;
  const int s = r >> 6, lane = r & 63;
  const int wb = (int)((off[s + 1] - off[s]) >> 6) / 3;
  const int a = r / 3;
  const int nbk = r < n ? cptr[a + 1] - cptr[a] : 0;
  for (int k = 0; k < wb; ++k) {
    int b = (r < n ? r : n - 1) / 3;
    if (k < nbk) b = ccol[cptr[a] + k];
    scol[off[s] / 3 + (int64_t)k * 64 + lane] = b;
  }
}

// y = D^-1 A x, partial sums of y'y
__global__ __launch_bounds__(kT) void ks_power(MatS A, const double* __restrict__ dinv,
                                               const double* __restrict__ x,
                                               double* __restrict__ y, double* __restrict__ part) {
  __shared__ double lds[kW];
  const int i = blockIdx.x * kT + threadIdx.x;
  double v = 0.0;
  if (i < A.n) {
    v = dinv[i] * rowS(A, x, i);
    y[i] = v;
  }
  const double s = block_sum(v * v, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(kT) void k0_power(Mat0 A, const double2* __restrict__ dinv,
                                               const double2* __restrict__ x,
                                               double2* __restrict__ y, double* __restrict__ part) {
  __shared__ double lds[kW];
  const int lane = threadIdx.x & 63;
  double acc = 0.0;
  for (int s = blockIdx.x * kW + (threadIdx.x >> 6); s < A.nslices; s += gridDim.x * kW) {
    double y0, y1;
    row0_spmv(A, x, s, lane, y0, y1);
    if (lane < A.nrow[s]) {
      const int i = A.row0[s] + lane;
      const double2 di = dinv[i];
      const double2 v = make_double2(di.x * y0, di.y * y1);
      y[i] = v;
      acc += v.x * v.x + v.y * v.y;
    }
  }
  const double s = block_sum(acc, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// y = A x on level 0
__global__ __launch_bounds__(kT) void k0_apply(Mat0 A, const double2* __restrict__ x,
                                               double2* __restrict__ y) {
  const int lane = threadIdx.x & 63;
  for (int s = blockIdx.x * kW + (threadIdx.x >> 6); s < A.nslices; s += gridDim.x * kW) {
    double y0, y1;
    row0_spmv(A, x, s, lane, y0, y1);
    if (lane < A.nrow[s]) y[A.row0[s] + lane] = make_double2(y0, y1);
  }
}

// z = M y with partial sums of z'z
__global__ __launch_bounds__(kT) void k0_mapply_sq(MatM M, const double2* __restrict__ y,
                                                   double2* __restrict__ z,
                                                   double* __restrict__ part) {
  __shared__ double lds[kW];
  const int lane = threadIdx.x & 63;
  double acc = 0.0;
  for (int s = blockIdx.x * kW + (threadIdx.x >> 6); s < M.nslices; s += gridDim.x * kW) {
    double z0, z1;
    rowM_spmv(M, y, s, lane, z0, z1);
    if (lane < M.nrow[s]) {
      z[M.row0[s] + lane] = make_double2(z0, z1);
      acc += z0 * z0 + z1 * z1;
    }
  }
  const double t = block_sum(acc, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// Gershgorin bound of lambda_max(D^-1 A): per-workgroup maxima of dinv_i sum_j |a_ij|
__global__ __launch_bounds__(kT) void ks_gersh(MatS A, const double* __restrict__ dinv,
                                               double* __restrict__ part) {
  __shared__ double lds[kW];
  double m = 0.0;
  for (int i = blockIdx.x * kT + threadIdx.x; i < A.n; i += gridDim.x * kT) {
    const int s = i >> 6, lane = i & 63;
    const int64_t o = A.off[s];
    const int w = (int)((A.off[s + 1] - o) >> 6);
    double a = 0.0;
    for (int k = 0; k < w; ++k) a += fabs(A.val[o + (int64_t)k * 64 + lane]);
    m = fmax(m, dinv[i] * a);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < kW; ++i) t = fmax(t, lds[i]);
    part[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(kT) void k0_gersh(Mat0 A, const double2* __restrict__ dinv,
                                               double* __restrict__ part) {
  __shared__ double lds[kW];
  const int lane = threadIdx.x & 63;
  double m = 0.0;
  for (int s = blockIdx.x * kW + (threadIdx.x >> 6); s < A.nslices; s += gridDim.x * kW) {
    const int64_t o = A.off[s];
    const int w = (int)((A.off[s + 1] - o) >> 6);
    const v2d* v = reinterpret_cast<const v2d*>(A.val + 4 * o) + lane;
    double a0 = 0.0, a1 = 0.0;
    for (int k = 0; k < w; ++k) {
      const v2d va = v[(size_t)k * 128], vb = v[(size_t)k * 128 + 64];
      a0 += fabs(va.x) + fabs(va.y);
      a1 += fabs(vb.x) + fabs(vb.y);
    }
    if (lane < A.nrow[s]) {
      const double2 di = dinv[A.row0[s] + lane];
      m = fmax(m, fmax(di.x * a0, di.y * a1));
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < kW; ++i) t = fmax(t, lds[i]);
    part[blockIdx.x] = t;
  }
}

__global__ void k_to_f32(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

// start vector of the power iteration: 1 + 0.5 sin(0.7 i + 0.3)
__global__ void k_power_init(int n, double* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 1.0 + 0.5 * sin(0.7 * (double)i + 0.3);
}

__global__ __launch_bounds__(kT) void k_sumsq(int n, const double* __restrict__ x,
                                              double* __restrict__ part) {
  __shared__ double lds[kW];
  double v = 0.0;
  for (int i = blockIdx.x * kT + threadIdx.x; i < n; i += gridDim.x * kT) v += x[i] * x[i];
  const double s = block_sum(v, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// y /= ||y|| from the partials of y'y (summed in the same order by every workgroup).
// st[0] = ||y|| when `record`; st[1] != 0 ends the iteration (set once ||y|| is not positive,
// by workgroup 0, while every workgroup of that launch already skips the scaling itself).
__global__ __launch_bounds__(kT) void k_power_scale(int n, double* __restrict__ y,
                                                    const double* __restrict__ part, int nparts,
                                                    double* __restrict__ st, int record) {
  __shared__ double lds[kW];
  if (st[1] != 0.0) return;
  const double yn = sqrt(sum_parts(part, nparts, lds));
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (record) st[0] = yn;
    if (!(yn > 0.0)) st[1] = 1.0;
  }
  if (!(yn > 0.0)) return;
  const double f = 1.0 / yn;
  for (int i = blockIdx.x * kT + threadIdx.x; i < n; i += gridDim.x * kT) y[i] *= f;
}

// ---- host helpers ------------------------------------------------------------------------

int grid1(int64_t n) { return (int)std::max<int64_t>(1, (n + kT - 1) / kT); }

Mat0 mat0(Context& c) {
  return Mat0{c.sell_off.get(), c.sell_col.get(), c.sell_val.get(), c.sell_row0.get(),
              c.sell_nrow.get(), c.nslices};
}

MatS mats(MgLevel& L) { return MatS{L.off.get(), L.scol.get(), L.val.get(), L.n}; }
MatS32 mats32(MgLevel& L) { return MatS32{L.off.get(), L.scol.get(), L.val32.get(), L.n}; }

int slice_grid(int nslices) {
  // at most 1024 workgroups of 4 waves, as for the flexible CG's SpMV (2048 was 3% slower)
  int g = std::min(1024, (nslices + kW - 1) / kW);
  if (g >= 64) g &= ~7;  // a multiple of 8: one contiguous slice range per XCD
  return std::max(1, g);
}

// slice offsets of the scalar SELL-64 layout of a 3x3-block pattern (cptr over n / 3 block
// rows): the width of a slice is that of its longest row
void sell_offsets(MgLevel& L, const std::vector<int>& cptr) {
  const int n = L.n;
  L.nslices = (n + 63) / 64;
  L.h_off.assign(L.nslices + 1, 0);
  for (int s = 0; s < L.nslices; ++s) {
    int w = 0;
    for (int a = (64 * s) / 3; a <= (std::min(n, 64 * s + 64) - 1) / 3; ++a)
      w = std::max(w, 3 * (cptr[a + 1] - cptr[a]));
    L.h_off[s + 1] = L.h_off[s] + 64 * (int64_t)w;
  }
}

// page-locked staging for pattern downloads, kept across set-ups
int* pinned_ints(Multigrid& mg, size_t count) {
  if (count > mg.pinned_n) {
    if (mg.pinned) HIP_CHECK(hipHostFree(mg.pinned));
    mg.pinned = nullptr;
    mg.pinned_n = 0;
    HIP_CHECK(hipHostMalloc((void**)&mg.pinned, count * sizeof(int), hipHostMallocDefault));
    mg.pinned_n = count;
  }
  return mg.pinned;
}

// the page-locked buffer of the coarse pattern's download (no copy into it is in flight when
// it grows: the set-up waits for each download before the next level's)
int* pinned_pattern(Multigrid& mg, size_t count) {
  if (count > mg.pinned_pat_n) {
    if (mg.pinned_pat) HIP_CHECK(hipHostFree(mg.pinned_pat));
    mg.pinned_pat = nullptr;
    mg.pinned_pat_n = 0;
    HIP_CHECK(hipHostMalloc((void**)&mg.pinned_pat, count * sizeof(int), hipHostMallocDefault));
    mg.pinned_pat_n = count;
  }
  return mg.pinned_pat;
}

// a workspace of at least `count` elements (reallocated only to grow)
template <class T>
T* grow(DBuf<T>& b, size_t count) {
  if (!b.get() || b.n < count) b.alloc(std::max<size_t>(count, 1));
  return b.get();
}

template <class T>
void upload(DBuf<T>& d, const std::vector<T>& h) {
  d.alloc(h.size());
  if (!h.empty()) HIP_CHECK(hipMemcpy(d.get(), h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
}

constexpr int kLamSlots = 18;  // lambda estimates on the device: one per level, one spare

// lam = min(Gershgorin bound, 1.3 x power estimate): the Gershgorin partials gp[0..g) are
// maxed in a fixed order, st[0] holds the power estimate
__global__ void k_lam_final(const double* __restrict__ gp, int g, const double* __restrict__ st,
                            double* __restrict__ lam) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double m = 0.0;
  for (int i = 0; i < g; ++i) m = fmax(m, gp[i]);
  const double pw = 1.3 * st[0];
  *lam = pw < m ? pw : m;
}

// Smoothing interval upper end of level `level` on `stream`, into *lam_dev: the power estimate
// of lambda_max(D^-1 A) (20 steps from a fixed vector) with a margin, never above the
// Gershgorin bound (which holds for every eigenvalue).  `ws` is the stream's workspace.
void smoothing_max_enqueue(Context& c, int level, hipStream_t stream, DBuf<double>& ws,
                           double* lam_dev) {
  MgLevel& L = *c.mg.lv[level];
  const int n = L.n;
  const int g = level == 0 ? slice_grid(c.nslices) : grid1(n);
  const int gn = std::min(kPart, grid1(n));
  const int gg = level == 0 ? slice_grid(c.nslices) : std::min(1024, grid1(n));
  const size_t np = std::max(g, gn);
  // x[n], y[n], power partials, the two scalars of k_power_scale, Gershgorin partials
  double* x = grow(ws, 2 * (size_t)n + np + 2 + gg);
  double* y = x + n;
  double* part = y + n;
  double* st = part + np;
  double* gpart = st + 2;
  if (level == 0)
    hipLaunchKernelGGL(k0_gersh, dim3(gg), dim3(kT), 0, stream, mat0(c),
                       (const double2*)c.inv_diag.get(), gpart);
  else
    hipLaunchKernelGGL(ks_gersh, dim3(gg), dim3(kT), 0, stream, mats(L), (const double*)L.dinv.get(),
                       gpart);
  HIP_CHECK(hipMemsetAsync(st, 0, 2 * sizeof(double), stream));
  hipLaunchKernelGGL(k_power_init, dim3(grid1(n)), dim3(kT), 0, stream, n, x);
  hipLaunchKernelGGL(k_sumsq, dim3(gn), dim3(kT), 0, stream, n, (const double*)x, part);
  hipLaunchKernelGGL(k_power_scale, dim3(gn), dim3(kT), 0, stream, n, x, (const double*)part,
                     gn, st, 0);
  for (int it = 0; it < 20; ++it) {
    if (level == 0)
      hipLaunchKernelGGL(k0_power, dim3(g), dim3(kT), 0, stream, mat0(c),
                         (const double2*)c.inv_diag.get(), (const double2*)x, (double2*)y, part);
    else
      hipLaunchKernelGGL(ks_power, dim3(g), dim3(kT), 0, stream, mats(L),
                         (const double*)L.dinv.get(), (const double*)x, y, part);
    hipLaunchKernelGGL(k_power_scale, dim3(gn), dim3(kT), 0, stream, n, y, (const double*)part,
                       g, st, 1);
    std::swap(x, y);
  }
  hipLaunchKernelGGL(k_lam_final, dim3(1), dim3(64), 0, stream, (const double*)gpart, gg,
                     (const double*)st, lam_dev);
  HIP_CHECK(hipGetLastError());
}

// the same, waited for on the library stream
double smoothing_max(Context& c, int level) {
  double* lam_dev = grow(c.mg.lam_dev, kLamSlots) + kLamSlots - 1;  // slot of the waited form
  smoothing_max_enqueue(c, level, c.stream, c.mg.dwork, lam_dev);
  double lam = 0.0;
  HIP_CHECK(hipMemcpyAsync(&lam, lam_dev, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  HIP_CHECK(hipStreamSynchronize(c.stream));
  return lam;
}

// largest eigenvalue of M A (the learned smoother on level 0) by `iters` power steps, enqueued
// on stream `st` with its own workspace (mg.dwork2); the estimate is left in *lam_dev
double* learned_smoother_bound_enqueue(Context& c, int iters, hipStream_t stream) {
  const int n = c.n, gs = slice_grid(c.nslices);
  const int gn = std::min(kPart, grid1(n));
  const MatM M{c.sell_off.get(), c.sell_col.get(), c.mpnn.gsell.get(), c.sell_row0.get(),
               c.sell_nrow.get(), c.nslices};
  const size_t np = std::max(gs, gn);
  double* x = grow(c.mg.dwork2, 2 * (size_t)n + np + 2);
  double* y = x + n;
  double* part = y + n;
  double* st = part + np;
  HIP_CHECK(hipMemsetAsync(st, 0, 2 * sizeof(double), stream));
  hipLaunchKernelGGL(k_power_init, dim3(grid1(n)), dim3(kT), 0, stream, n, x);
  hipLaunchKernelGGL(k_sumsq, dim3(gn), dim3(kT), 0, stream, n, (const double*)x, part);
  hipLaunchKernelGGL(k_power_scale, dim3(gn), dim3(kT), 0, stream, n, x, (const double*)part,
                     gn, st, 0);
  for (int it = 0; it < iters; ++it) {
    hipLaunchKernelGGL(k0_apply, dim3(gs), dim3(kT), 0, stream, mat0(c), (const double2*)x,
                       (double2*)y);
    hipLaunchKernelGGL(k0_mapply_sq, dim3(gs), dim3(kT), 0, stream, M, (const double2*)y,
                       (double2*)x, part);
    hipLaunchKernelGGL(k_power_scale, dim3(gn), dim3(kT), 0, stream, n, x, (const double*)part,
                       gs, st, 1);
  }
  HIP_CHECK(hipGetLastError());
  return st;
}

// Chebyshev coefficients of the two smoothing steps on [lam / 30, lam]
struct Cheb {
  double c1r, c2d, c2r;
};
Cheb cheb(double lam) {
  const double hi = lam, lo = hi / 30.0;
  const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo);
  const double sigma = theta / delta, rho0 = 1.0 / sigma, rho1 = 1.0 / (2.0 * sigma - rho0);
  return Cheb{1.0 / theta, rho1 * rho0, 2.0 * rho1 / delta};
}

// Dense inverse of the coarsest matrix on the device, one workgroup, the matrix in LDS: rows
// assembled from the SELL level in entry order (the host form's order), dead DOFs (diagonal
// not positive) replaced by identity rows and columns, then Gauss-Jordan elimination with the
// pivots in order (positive for a symmetric positive definite matrix; *flag = 1 and the
// kernel stops at the first that is not).  cinv is row-major n x n, zero on dead rows and
// columns, as the host form leaves it.
constexpr int kInvThreads = 1024;
__global__ __launch_bounds__(kInvThreads) void k_coarsest_inverse(MatS A, int n,
                                                                  double* __restrict__ cinv,
                                                                  int* __restrict__ flag) {
  extern __shared__ double M[];  // n x n, then n dead flags (as doubles)
  double* dead = M + (size_t)n * n;
  const int t = threadIdx.x, nt = blockDim.x;
  for (int e = t; e < n * n; e += nt) M[e] = 0.0;
  __syncthreads();
  for (int r = t; r < n; r += nt) {
    const int s = r >> 6, lane = r & 63;
    const int64_t o = A.off[s];
    const int w = (int)((A.off[s + 1] - o) >> 6);
    for (int k = 0; k < w; ++k) M[(size_t)r * n + A.col(o, k, lane)] += A.val[o + (int64_t)k * 64 + lane];
  }
  __syncthreads();
  for (int i = t; i < n; i += nt) dead[i] = M[(size_t)i * n + i] > 0.0 ? 0.0 : 1.0;
  __syncthreads();
  for (int e = t; e < n * n; e += nt) {
    const int i = e / n, j = e % n;
    if (dead[i] != 0.0 || dead[j] != 0.0) M[e] = i == j ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    const double p = M[(size_t)k * n + k];  // the same value in every thread
    if (!(p > 0.0)) {
      if (t == 0) *flag = 1;
      return;
    }
    const double ip = 1.0 / p;
    // the entries off row and column k, from the old row and column k
    for (int e = t; e < n * n; e += nt) {
      const int i = e / n, j = e % n;
      if (i != k && j != k) M[e] = M[e] - M[(size_t)i * n + k] * (M[(size_t)k * n + j] * ip);
    }
    __syncthreads();
    // then row and column k
    for (int j = t; j < n; j += nt) {
      if (j != k) {
        M[(size_t)k * n + j] = M[(size_t)k * n + j] * ip;
        M[(size_t)j * n + k] = -M[(size_t)j * n + k] * ip;
      }
    }
    if (t == 0) M[(size_t)k * n + k] = ip;
    __syncthreads();
  }
  for (int e = t; e < n * n; e += nt) {
    const int i = e / n, j = e % n;
    cinv[e] = dead[i] != 0.0 || dead[j] != 0.0 ? 0.0 : M[e];
  }
}

// LDS bytes k_coarsest_inverse needs for n DOFs
size_t device_inverse_bytes(int n) { return sizeof(double) * ((size_t)n * n + n); }

// enqueues the device form when its matrix fits one workgroup's LDS (false: use the host's)
bool coarsest_inverse_device(Context& c, MgLevel& L) {
  const int n = L.n;
  int lds = 0;
  HIP_CHECK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, c.device));
  if (n <= 0 || device_inverse_bytes(n) > (size_t)lds) return false;
  c.mg.nc = n;
  c.mg.cinv.alloc((size_t)n * n);
  int* flag = grow(c.mg.inv_flag, 1);
  HIP_CHECK(hipMemsetAsync(flag, 0, sizeof(int), c.stream));
  hipLaunchKernelGGL(k_coarsest_inverse, dim3(1), dim3(kInvThreads), device_inverse_bytes(n),
                     c.stream, mats(L), n, c.mg.cinv.get(), flag);
  HIP_CHECK(hipGetLastError());
  return true;
}

// dense inverse of the coarsest matrix (dead DOFs: identity rows, zero solution), on the host
void coarsest_inverse(Context& c, MgLevel& L) {
  const int n = L.n;
  std::vector<double> val(L.val.n);
  std::vector<int> scol(L.scol.n);
  HIP_CHECK(hipMemcpy(val.data(), L.val.get(), sizeof(double) * val.size(), hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(scol.data(), L.scol.get(), sizeof(int) * scol.size(), hipMemcpyDeviceToHost));
  std::vector<double> A((size_t)n * n, 0.0);
  for (int s = 0; s < L.nslices; ++s) {
    const int w = (int)((L.h_off[s + 1] - L.h_off[s]) / 64);
    for (int lane = 0; lane < 64; ++lane) {
      const int r = 64 * s + lane;
      if (r >= n) continue;
      for (int k = 0; k < w; ++k) {
        const int64_t q = L.h_off[s] + (int64_t)k * 64 + lane;
        const int col = 3 * scol[L.h_off[s] / 3 + (int64_t)(k / 3) * 64 + lane] + k % 3;
        A[(size_t)r * n + col] += val[q];
      }
    }
  }
  std::vector<char> dead(n, 0);
  for (int i = 0; i < n; ++i)
    if (!(A[(size_t)i * n + i] > 0.0)) {
      dead[i] = 1;
      for (int j = 0; j < n; ++j) A[(size_t)i * n + j] = A[(size_t)j * n + i] = 0.0;
      A[(size_t)i * n + i] = 1.0;
    }
  // Cholesky A = L L^T (lower, in place), then the inverse column by column
  for (int j = 0; j < n; ++j) {
    double d = A[(size_t)j * n + j];
    for (int k = 0; k < j; ++k) d -= A[(size_t)j * n + k] * A[(size_t)j * n + k];
    if (!(d > 0.0)) throw Error(CSTFEM_ERR_BAD_DIAGONAL, "multilevel: coarsest matrix is not positive definite");
    const double ljj = std::sqrt(d);
    A[(size_t)j * n + j] = ljj;
    for (int i = j + 1; i < n; ++i) {
      double s = A[(size_t)i * n + j];
      for (int k = 0; k < j; ++k) s -= A[(size_t)i * n + k] * A[(size_t)j * n + k];
      A[(size_t)i * n + j] = s / ljj;
    }
  }
  std::vector<double> inv((size_t)n * n, 0.0), e(n);
  for (int col_i = 0; col_i < n; ++col_i) {
    std::fill(e.begin(), e.end(), 0.0);
    if (!dead[col_i]) e[col_i] = 1.0;
    for (int i = 0; i < n; ++i) {  // L y = e
      double s = e[i];
      for (int k = 0; k < i; ++k) s -= A[(size_t)i * n + k] * e[k];
      e[i] = s / A[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {  // L^T x = y
      double s = e[i];
      for (int k = i + 1; k < n; ++k) s -= A[(size_t)k * n + i] * e[k];
      e[i] = s / A[(size_t)i * n + i];
    }
    for (int i = 0; i < n; ++i) inv[(size_t)i * n + col_i] = dead[i] ? 0.0 : e[i];
  }
  c.mg.nc = n;
  upload(c.mg.cinv, inv);
}

// Phase times of the setup on stderr when CSTFEM_MG_TRACE is set (each mark synchronises the
// stream, so the trace itself slows the setup down a little).
struct SetupTrace {
  // CSTFEM_MG_TRACE=1: each mark waits for the stream (device time per phase); =2: host
  // timestamps only (where the host thread spends the set-up)
  const char* env = std::getenv("CSTFEM_MG_TRACE");
  bool on = env != nullptr;
  bool sync = env != nullptr && env[0] != '2';
  std::chrono::steady_clock::time_point last = std::chrono::steady_clock::now();
  void mark(Context& c, int level, const char* what) {
    if (!on) return;
    if (sync) HIP_CHECK(hipStreamSynchronize(c.stream));
    const auto now = std::chrono::steady_clock::now();
    std::fprintf(stderr, "[mg setup] level %d %-20s %8.3f ms\n", level, what,
                 std::chrono::duration<double, std::milli>(now - last).count());
    last = now;
  }
};

constexpr size_t kRRLds = 48 * 1024;  // LDS bytes of ks_resid_restrict at most

// Work split of the fused residual + restriction (ks_resid_restrict) on level L, from the
// host's aggregate map: rr_ag aggregates per workgroup, about one member row per lane group,
// and rr_cap = the most member rows one workgroup holds.  Level 0 (2x2 blocks, its own row
// sums) keeps the separate kernels, as does a level whose rows would not fit kRRLds.
void plan_resid_restrict(MgLevel& L) {
  L.rr_ag = L.rr_cap = 0;
  if (L.nagg <= 0 || L.bs != 3) return;
  std::vector<int> cnt(L.nagg, 0);
  for (int i = 0; i < L.nodes; ++i) ++cnt[L.h_agg[i]];
  const double rows_per_agg = (double)L.n / L.nagg;
  const int ag = std::max(1, (int)((kT / lanes_per_row(L.n)) / rows_per_agg));
  int cap = 0;
  for (int a0 = 0; a0 < L.nagg; a0 += ag) {
    int rows = 0;
    for (int a = a0; a < std::min(a0 + ag, L.nagg); ++a) rows += L.bs * cnt[a];
    cap = std::max(cap, rows);
  }
  if ((size_t)cap * (sizeof(double) + sizeof(int)) > kRRLds) return;
  L.rr_ag = ag;
  L.rr_cap = cap;
}

void alloc_work(Context& c, MgLevel& L) {
  for (DBuf<double>* v :
       {&L.rhs, &L.x, &L.x2, &L.d, &L.res, &L.kx, &L.kres, &L.kres2, &L.kp, &L.kq, &L.kz}) {
    v->alloc(std::max(1, L.n));
    HIP_CHECK(hipMemsetAsync(v->get(), 0, sizeof(double) * std::max(1, L.n), c.stream));
  }
  L.part.alloc(6 * kPart);
}

// ---- the cycle -----------------------------------------------------------------------------

struct PreFold;
int cycle_s(Context& c, int l, const double* rhs, double* out, const PcgScalars* sc,
            const double* dq = nullptr, double* dpart = nullptr, const PreFold* pf = nullptr);

// ---- launches with the lane count chosen per level ------------------------------------------

#define CSTFEM_LPR_SWITCH(lpr, KERNEL, grid, ...)                                               \
  switch (lpr) {                                                                              \
    case 1: hipLaunchKernelGGL(KERNEL<1>, dim3(grid), dim3(kT), 0, c.stream, __VA_ARGS__); break; \
    case 2: hipLaunchKernelGGL(KERNEL<2>, dim3(grid), dim3(kT), 0, c.stream, __VA_ARGS__); break; \
    case 4: hipLaunchKernelGGL(KERNEL<4>, dim3(grid), dim3(kT), 0, c.stream, __VA_ARGS__); break; \
    default: hipLaunchKernelGGL(KERNEL<8>, dim3(grid), dim3(kT), 0, c.stream, __VA_ARGS__); break; \
  }

// the same with the matrix value type as the kernel's second template argument: FP32 when
// the level holds an FP32 copy of its values and the learned smoother is in use; shm bytes
// of dynamic LDS
#define CSTFEM_LPR_SWITCH_MS(L, lpr, KERNEL, grid, shm, ...)                                      \
  if ((L).val32.n && c.mg.fine_smoother == 0) {                                               \
    const MatS32 A_ = mats32(L);                                                              \
    switch (lpr) {                                                                            \
      case 1: hipLaunchKernelGGL((KERNEL<1, float>), dim3(grid), dim3(kT), shm, c.stream, A_, __VA_ARGS__); break; \
      case 2: hipLaunchKernelGGL((KERNEL<2, float>), dim3(grid), dim3(kT), shm, c.stream, A_, __VA_ARGS__); break; \
      case 4: hipLaunchKernelGGL((KERNEL<4, float>), dim3(grid), dim3(kT), shm, c.stream, A_, __VA_ARGS__); break; \
      default: hipLaunchKernelGGL((KERNEL<8, float>), dim3(grid), dim3(kT), shm, c.stream, A_, __VA_ARGS__); break; \
    }                                                                                         \
  } else {                                                                                    \
    const MatS A_ = mats(L);                                                                  \
    switch (lpr) {                                                                            \
      case 1: hipLaunchKernelGGL((KERNEL<1, double>), dim3(grid), dim3(kT), shm, c.stream, A_, __VA_ARGS__); break; \
      case 2: hipLaunchKernelGGL((KERNEL<2, double>), dim3(grid), dim3(kT), shm, c.stream, A_, __VA_ARGS__); break; \
      case 4: hipLaunchKernelGGL((KERNEL<4, double>), dim3(grid), dim3(kT), shm, c.stream, A_, __VA_ARGS__); break; \
      default: hipLaunchKernelGGL((KERNEL<8, double>), dim3(grid), dim3(kT), shm, c.stream, A_, __VA_ARGS__); break; \
    }                                                                                         \
  }
#define CSTFEM_LPR_SWITCH_M(L, lpr, KERNEL, grid, ...) \
  CSTFEM_LPR_SWITCH_MS(L, lpr, KERNEL, grid, 0, __VA_ARGS__)

// grid of ks_cheb on a level (one block partial each when it also forms a dot product)
int cheb_grid(const MgLevel& L) { return grid1((int64_t)L.n * lanes_per_row(L.n)); }

void launch_cheb(Context& c, MgLevel& L, const double* xin, double* xout, const double* r,
                 double cd, double cr, const PcgScalars* sc, const double* dq = nullptr,
                 double* dpart = nullptr) {
  const int lpr = lanes_per_row(L.n);
  CSTFEM_LPR_SWITCH_M(L, lpr, ks_cheb, cheb_grid(L), xin, xout, L.d.get(), r,
                      (const double*)L.dinv.get(), cd, cr, dq, dpart, sc);
}

void launch_cheb_pre(Context& c, MgLevel& L, double* x2, const double* r, const Cheb& ch,
                     const PcgScalars* sc) {
  const int lpr = lanes_per_row(L.n);
  CSTFEM_LPR_SWITCH_M(L, lpr, ks_cheb_pre, grid1((int64_t)L.n * lpr), x2, L.d.get(), r,
                    (const double*)L.dinv.get(), ch.c1r, ch.c2d, ch.c2r, sc);
}

void launch_resid(Context& c, MgLevel& L, const double* x, const double* r, double* res,
                  const PcgScalars* sc) {
  const int lpr = lanes_per_row(L.n);
  CSTFEM_LPR_SWITCH_M(L, lpr, ks_resid, grid1((int64_t)L.n * lpr), x, r, res, sc);
}

// lanes per coarse DOF of the restriction from level L
int restrict_lanes(const MgLevel& L) {
  const int nc = 3 * L.nagg;
  return nc >= 100000 ? 1 : nc >= 20000 ? 2 : 8;
}

// restriction of level L's residual res into rc (3 per aggregate)
template <class PT>
void launch_restrict_t(Context& c, MgLevel& L, const PT* P, const double* res, double* rc,
                       const PcgScalars* sc) {
  const int lpr = restrict_lanes(L);
  const int grid = grid1((int64_t)3 * L.nagg * lpr);
  const int* mptr = L.mptr.get();
  const int* mem = L.mem.get();
  switch (lpr) {
    case 1: hipLaunchKernelGGL((k_restrict<1, PT>), dim3(grid), dim3(kT), 0, c.stream, L.nagg, L.bs, mptr, mem, P, res, rc, sc); break;
    case 2: hipLaunchKernelGGL((k_restrict<2, PT>), dim3(grid), dim3(kT), 0, c.stream, L.nagg, L.bs, mptr, mem, P, res, rc, sc); break;
    case 4: hipLaunchKernelGGL((k_restrict<4, PT>), dim3(grid), dim3(kT), 0, c.stream, L.nagg, L.bs, mptr, mem, P, res, rc, sc); break;
    default: hipLaunchKernelGGL((k_restrict<8, PT>), dim3(grid), dim3(kT), 0, c.stream, L.nagg, L.bs, mptr, mem, P, res, rc, sc); break;
  }
}

// p32: with level 0's FP32 copy of P
void launch_restrict(Context& c, MgLevel& L, const double* res, double* rc, const PcgScalars* sc,
                     bool p32 = false) {
  if (p32)
    launch_restrict_t(c, L, (const float*)L.P32.get(), res, rc, sc);
  else
    launch_restrict_t(c, L, (const double*)L.P.get(), res, rc, sc);
}

// levels >= 1 of fewer than fuse_n DOFs run the fused residual + restriction
bool fused_level(const Multigrid& mg, int l) { return mg.fuse && l >= 1 && mg.lv[l]->n < mg.fuse_n; }

// ... and, where a row sum takes 4 or more lanes (few entries per lane), the prolongation inside
// the first post-smoothing step: its gathers of x2 + P e cost 8 or 11 loads per entry instead
// of 1, which on a level with 2 lanes per row (about 11 entries each) takes longer than the
// separate k_prolong (22 against 10 us on a level of 74k DOFs)
bool fused_prolong(const Multigrid& mg, int l) {
  return fused_level(mg, l) && lanes_per_row(mg.lv[l]->n) >= 4;
}

// A K-cycle update is left to the kernel that reads its result next when that kernel's
// workgroups, each summing the step's np partials, read fewer than fold_max partials in all
// (the separate ks_update is cheaper where they would read more)
bool fold_ok(const Multigrid& mg, int64_t reader_grid, int np) {
  return mg.fuse && reader_grid * np <= mg.fold_max;
}

// residual rhs - A x of level l and its restriction into rc: one fused kernel where the level
// is small and planned for it (rr_ag > 0), else ks_resid + k_restrict
void launch_resid_restrict(Context& c, int l, const double* x, const double* rhs, double* rc,
                           const PcgScalars* sc) {
  MgLevel& L = *c.mg.lv[l];
  if (!(fused_level(c.mg, l) && L.rr_ag > 0)) {
    launch_resid(c, L, x, rhs, L.res.get(), sc);
    launch_restrict(c, L, L.res.get(), rc, sc);
    return;
  }
  const int lpr = lanes_per_row(L.n);
  const int grid = (L.nagg + L.rr_ag - 1) / L.rr_ag;
  const size_t shm = (size_t)L.rr_cap * (sizeof(double) + sizeof(int));
  CSTFEM_LPR_SWITCH_MS(L, lpr, ks_resid_restrict, grid, shm, x, rhs, L.nagg, L.rr_ag, L.rr_cap,
                       L.bs, restrict_lanes(L), (const int*)L.mptr.get(),
                       (const int*)L.mem.get(), (const double*)L.P.get(), rc, sc);
}

// grid of the K-cycle kernels of a level (all of them share it: their partials are summed
// by count)
int kcycle_grid(const MgLevel& L) {
  return std::min<int64_t>(kPart, grid1((int64_t)L.n * lanes_per_row(L.n)));
}

// LDS doubles the fused tail needs from level t down (5 vectors per cycle level, rhs and
// solution of the coarsest)
int64_t tail_lds_doubles(const Multigrid& mg, int t) {
  const int last = (int)mg.lv.size() - 1;
  int64_t s = 2 * (int64_t)mg.lv[last]->n;
  for (int l = t; l < last; ++l) s += 5 * (int64_t)mg.lv[l]->n;
  return s;
}

// the fused tail of the hierarchy (levels tail_from .. coarsest), or -1 when it does not apply
int tail_start(const Multigrid& mg) {
  const char* env = std::getenv("CSTFEM_MG_TAIL");
  if (env && env[0] == '0') return -1;
  const int last = (int)mg.lv.size() - 1;
  int t = last;
  while (t - 1 >= 1 && t - 1 > mg.kdepth && last - (t - 1) <= kTailLevels &&
         tail_lds_doubles(mg, t - 1) * 8 <= kTailLds)
    --t;
  return t < last ? t : -1;
}

// Whether the fused tail reads FP32 row values.  It does not: the tail levels keep their FP64
// values even in the learned smoother's cycle, whose other levels read FP32 copies, because
// the tail ran 3 us per call slower on FP32 values (19.8 against 16.7 us on the 4M-DOF
// plate's 972-DOF tail, same run, same results otherwise).  Kept as a switch for the two
// instantiations of k_tail.
bool tail_v32(const Multigrid& mg) {
  (void)mg;
  return false;
}

// whether the top tail level keeps its rows in registers (one row per thread, <= kTailRegW
// entries)
bool tail_reg(const Multigrid& mg) {
  const MgLevel& top = *mg.lv[mg.tail_from];
  bool reg = top.n <= kTailThreads;
  for (int s = 0; reg && s < top.nslices; ++s) reg = top.h_off[s + 1] - top.h_off[s] <= 64 * kTailRegW;
  return reg;
}

template <class T>
void launch_tail_t(Context& c, int t, const double* rhs, double* out, const PcgScalars* sc) {
  Multigrid& mg = c.mg;
  const int last = (int)mg.lv.size() - 1;
  Tail<T> T_{};
  T_.nlev = last - t;
  int off = 0;
  for (int k = 0; k < T_.nlev; ++k) {
    MgLevel& L = *mg.lv[t + k];
    const Cheb ch = cheb(L.lam);
    TailLevel<T>& q = T_.L[k];
    if constexpr (std::is_same<T, float>::value)
      q.A = mats32(L);
    else
      q.A = mats(L);
    q.dinv = L.dinv.get();
    q.mptr = L.mptr.get();
    q.mem = L.mem.get();
    q.agg = L.agg.get();
    q.P = L.P.get();
    q.nagg = L.nagg;
    q.bs = L.bs;
    q.lds = off;
    off += 5 * L.n;
    q.c1r = ch.c1r;
    q.c2d = ch.c2d;
    q.c2r = ch.c2r;
  }
  T_.cinv = mg.cinv.get();
  T_.nc = mg.lv[last]->n;
  T_.lds_c = off;
  T_.rhs = rhs;
  T_.out = out;
  const size_t bytes = sizeof(double) * (size_t)tail_lds_doubles(mg, t);
  if (tail_reg(mg))
    hipLaunchKernelGGL((k_tail<true, T>), dim3(1), dim3(kTailThreads), bytes, c.stream, T_, sc);
  else
    hipLaunchKernelGGL((k_tail<false, T>), dim3(1), dim3(kTailThreads), bytes, c.stream, T_, sc);
}

// Gathers the packed tail's operands (k_tail_pack) when the tail is one level with its rows
// in registers over a coarsest level and aggregates small enough for the pack's slots; else
// the cycle keeps k_tail.  Called wherever tail_from changes.  CSTFEM_MG_TAIL_PACK=0: k_tail.
void tail_pack_build(Context& c) {
  Multigrid& mg = c.mg;
  mg.tail_packed = -1;
  const char* env = std::getenv("CSTFEM_MG_TAIL_PACK");
  if (env && env[0] == '0') return;
  const int t = mg.tail_from, last = (int)mg.lv.size() - 1;
  if (t < 0 || last - t != 1 || tail_v32(mg) || !tail_reg(mg)) return;
  MgLevel& L = *mg.lv[t];
  const int nc = mg.lv[last]->n;
  if (L.bs != 3 || 3 * L.nagg * kTailLanes > kTailThreads || nc > kPackD * kTailLanes ||
      nc * kTailLanes > kTailThreads || nc != mg.nc || 3 * L.nagg != nc)
    return;
  std::vector<int> cnt(L.nagg, 0);
  for (int i = 0; i < L.nodes; ++i) ++cnt[L.h_agg[i]];
  for (int a = 0; a < L.nagg; ++a)
    if (L.bs * cnt[a] > kPackR * kTailLanes) return;
  mg.tail_pd.alloc((size_t)kPdSlots * kTailThreads);
  mg.tail_pi.alloc((size_t)kPiSlots * kTailThreads);
  TailLevel<double> q{};
  q.A = mats(L);
  q.dinv = L.dinv.get();
  q.mptr = L.mptr.get();
  q.mem = L.mem.get();
  q.agg = L.agg.get();
  q.P = L.P.get();
  q.nagg = L.nagg;
  q.bs = L.bs;
  hipLaunchKernelGGL(k_tail_pack, dim3(1), dim3(kTailThreads), 0, c.stream, q,
                     (const double*)mg.cinv.get(), nc, mg.tail_pd.get(), mg.tail_pi.get());
  HIP_CHECK(hipGetLastError());
  mg.tail_packed = t;
}

void launch_tail(Context& c, int t, const double* rhs, double* out, const PcgScalars* sc) {
  Multigrid& mg = c.mg;
  if (mg.tail_packed == t) {
    const MgLevel& L = *mg.lv[t];
    const Cheb ch = cheb(L.lam);
    const TailPack T{mg.tail_pd.get(), mg.tail_pi.get(), L.n, 3 * L.nagg * kTailLanes, mg.nc,
                     ch.c1r, ch.c2d, ch.c2r, rhs, out};
    const size_t bytes = sizeof(double) * (5 * (size_t)L.n + 2 * (size_t)mg.nc);
    hipLaunchKernelGGL(k_tail1, dim3(1), dim3(kTailThreads), bytes, c.stream, T, sc);
    return;
  }
  if (tail_v32(c.mg))
    launch_tail_t<float>(c, t, rhs, out, sc);
  else
    launch_tail_t<double>(c, t, rhs, out, sc);
}

// where the restriction of level l's residual goes: straight into the K-cycle's residual when
// level l+1 runs the K-cycle (its right-hand side is not needed otherwise), else its rhs
double* coarse_rhs(Multigrid& mg, int l) {
  MgLevel& C = *mg.lv[l + 1];
  const bool kc = l + 1 < (int)mg.lv.size() - 1 && l + 1 <= mg.kdepth;
  return kc ? C.kres.get() : C.rhs.get();
}

// The last update of a K-cycle, kx += alpha p with alpha from the step's partials, when it is
// left to the kernel that reads kx next (the prolongation into the level above); p null: kx
// is final
struct Fold {
  const double* p = nullptr;
  const double* part = nullptr;
  int np = 0;
};

// The K-cycle's first update folded into the second cycle's pre-smoothing (ks_cheb_pre_fold):
// the residual res - alpha q goes to res2 and kx = alpha p
struct PreFold {
  const double* res;
  const double* q;
  const double* p;
  const double* part;
  int np;
  double* res2;
  double* kx;
};

Fold kcycle(Context& c, int l, const PcgScalars* sc);

// coarse correction of level l: lv[l+1].kx from lv[l+1].rhs -- the dense solve on the
// coarsest level, the K-cycle on levels <= kdepth, one cycle (V) below.  Returns the update
// the K-cycle left to the prolongation, if any.
Fold coarse_correction(Context& c, int l, const PcgScalars* sc) {
  Multigrid& mg = c.mg;
  MgLevel& C = *mg.lv[l + 1];
  if (l + 1 == (int)mg.lv.size() - 1) {
    hipLaunchKernelGGL(k_dense, dim3(grid1(C.n)), dim3(kT), 0, c.stream, C.n,
                       (const double*)mg.cinv.get(), (const double*)C.rhs.get(), C.kx.get(), sc);
    return Fold{};
  }
  if (l + 1 <= mg.kdepth) return kcycle(c, l + 1, sc);
  cycle_s(c, l + 1, C.rhs.get(), C.kx.get(), sc);
  return Fold{};
}

// XProlong of level L from the coarse level C's kx and the update f left on it
XProlong prolong_of(const MgLevel& L, const MgLevel& C, double* x2, const Fold& f) {
  return XProlong{x2, L.agg.get(), L.P.get(), C.kx.get(), f.p, 0.0, L.bs};
}

// x2 += P e (k_prolong) on level l after its coarse correction
void launch_prolong(Context& c, int l, double* x2, const Fold& f, const PcgScalars* sc,
                    bool p32 = false) {
  MgLevel& L = *c.mg.lv[l];
  hipLaunchKernelGGL(k_prolong, dim3(grid1(L.n)), dim3(kT), 0, c.stream, L.n,
                     prolong_of(L, *c.mg.lv[l + 1], x2, f), p32 ? (const float*)L.P32.get() : nullptr,
                     f.part, f.np, x2, sc);
}

// cycle on level l >= 1: out = B_l(rhs)
// With dpart, the last smoothing step also leaves the block partials of out'dq there and the
// return value is their count; 0 when the partials were not formed (the fused tail; a level
// whose smoothing grid exceeds the 2 kPart partial slots).  With pf (the K-cycle's second
// cycle), rhs is pf->res2, which the pre-smoothing writes first.
int cycle_s(Context& c, int l, const double* rhs, double* out, const PcgScalars* sc,
            const double* dq, double* dpart, const PreFold* pf) {
  if (l == c.mg.tail_from) {
    launch_tail(c, l, rhs, out, sc);
    return 0;
  }
  MgLevel& L = *c.mg.lv[l];
  MgLevel& C = *c.mg.lv[l + 1];
  const Cheb ch = cheb(L.lam);
  const int lpr = lanes_per_row(L.n);
  if (pf) {
    CSTFEM_LPR_SWITCH_M(L, lpr, ks_cheb_pre_fold, grid1((int64_t)L.n * lpr), L.x2.get(),
                        L.d.get(), pf->res, pf->q, pf->p, pf->part, pf->np, pf->res2, pf->kx,
                        (const double*)L.dinv.get(), ch.c1r, ch.c2d, ch.c2r, sc);
  } else {
    launch_cheb_pre(c, L, L.x2.get(), rhs, ch, sc);
  }
  launch_resid_restrict(c, l, L.x2.get(), rhs, coarse_rhs(c.mg, l), sc);
  const Fold f = coarse_correction(c, l, sc);
  if (fused_prolong(c.mg, l)) {
    CSTFEM_LPR_SWITCH_M(L, lpr, ks_cheb_prolong, grid1((int64_t)L.n * lpr),
                        prolong_of(L, C, L.x2.get(), f), f.part, f.np, L.x.get(), L.d.get(), rhs,
                        (const double*)L.dinv.get(), ch.c1r, sc);
  } else {
    launch_prolong(c, l, L.x2.get(), f, sc);
    launch_cheb(c, L, L.x2.get(), L.x.get(), rhs, 0.0, ch.c1r, sc);
  }
  const bool dot = dpart && cheb_grid(L) <= 2 * kPart;
  launch_cheb(c, L, L.x.get(), out, rhs, ch.c2d, ch.c2r, sc, dot ? dq : nullptr,
              dot ? dpart : nullptr);
  return dot ? cheb_grid(L) : 0;
}

// two flexible CG steps on level l for the right-hand side in lv[l].kres, result in lv[l].kx.
// Where fold_ok allows, the first step's update of kx and the residual happens inside the
// second cycle's pre-smoothing (the residual goes to kres2), and the second step's update of
// kx is returned for the prolongation into level l - 1 to apply; otherwise ks_update does them.
Fold kcycle(Context& c, int l, const PcgScalars* sc) {
  Multigrid& mg = c.mg;
  MgLevel& L = *mg.lv[l];
  const int lpr = lanes_per_row(L.n);
  const int g = kcycle_grid(L);           // grid-stride kernels with <= kPart partials
  double* s0 = L.part.get();             // p'q, p'r of step 1
  double* s1 = L.part.get() + 2 * kPart;  // z'q_old
  double* s2 = L.part.get() + 4 * kPart;  // p'q, p'r of step 2
  const bool fuse = fold_ok(mg, grid1((int64_t)L.n * lpr), g);
  // the reader of kx: ks_cheb_prolong on a level above that fuses the prolongation, else k_prolong
  const MgLevel& U = *mg.lv[l - 1];
  const int64_t ug = fused_prolong(mg, l - 1) ? grid1((int64_t)U.n * lanes_per_row(U.n)) : grid1(U.n);
  const bool fold_out = fold_ok(mg, ug, g);
  cycle_s(c, l, L.kres.get(), L.kz.get(), sc);  // the restriction filled kres
  CSTFEM_LPR_SWITCH_M(L, lpr, ks_spmv_dir, g, (const double*)L.kz.get(), L.kp.get(), L.kq.get(),
                      (const double*)L.kres.get(), s0, (const double*)nullptr, 0,
                      (const double*)nullptr, g, 1, sc);
  double* res = L.kres.get();
  PreFold pf{};
  if (fuse) {
    pf = PreFold{L.kres.get(), L.kq.get(), L.kp.get(), s0, g, L.kres2.get(), L.kx.get()};
    res = L.kres2.get();
  } else {
    hipLaunchKernelGGL(ks_update, dim3(g), dim3(kT), 0, c.stream, L.n, L.kx.get(), L.kres.get(),
                       (const double*)L.kp.get(), (const double*)L.kq.get(), (const double*)s0, g, 1,
                       1, sc);
  }
  // z'q_old: formed by the cycle's last smoothing step where it can, else by ks_dot
  int nzq = cycle_s(c, l, res, L.kz.get(), sc, (const double*)L.kq.get(), s1, fuse ? &pf : nullptr);
  if (nzq == 0) {
    hipLaunchKernelGGL(ks_dot, dim3(g), dim3(kT), 0, c.stream, L.n, (const double*)L.kz.get(),
                       (const double*)L.kq.get(), s1, sc);
    nzq = g;
  }
  CSTFEM_LPR_SWITCH_M(L, lpr, ks_spmv_dir, g, (const double*)L.kz.get(), L.kp.get(), L.kq.get(),
                      (const double*)res, s2, (const double*)s1, nzq, (const double*)s0, g, 0, sc);
  if (fold_out) return Fold{L.kp.get(), s2, g};
  hipLaunchKernelGGL(ks_update, dim3(g), dim3(kT), 0, c.stream, L.n, L.kx.get(), L.kres.get(),
                     (const double*)L.kp.get(), (const double*)L.kq.get(), (const double*)s2, g, 0, 0,
                     sc);
  return Fold{};
}

}  // namespace

// ---------------------------------------------------------------------------------------------

int mg_enqueue_apply(Context& c, const double* r, double* z, const PcgScalars* sc,
                     const double* zq, double* zq_part) {
  Multigrid& mg = c.mg;
  if (!mg.ready) throw Error(CSTFEM_ERR_ORDER, "multilevel: not set up");
  MgLevel& L = *mg.lv[0];
  const Mat0 A = mat0(c);
  const int gs = slice_grid(c.nslices);
  const double2* r2 = (const double2*)r;
  double2* z2 = (double2*)z;
  const bool learned = mg.fine_smoother == 0;
  if (learned && !c.mpnn.built) throw Error(CSTFEM_ERR_ORDER, "multilevel: the learned smoother is not built");
  const MatM M{c.sell_off.get(), c.sell_col.get(), c.mpnn.gsell.get(), c.sell_row0.get(),
               c.sell_nrow.get(), c.nslices};
  const Cheb ch = cheb(L.lam);
  // pre-smoothing
  if (learned) {
    hipLaunchKernelGGL(k0_msmooth, dim3(gs), dim3(kT), 0, c.stream, M, r2, (const double2*)nullptr,
                       z2, (const double2*)nullptr, (double*)nullptr, sc);
  } else {
    hipLaunchKernelGGL(k0_cheb, dim3(gs), dim3(kT), 0, c.stream, A, (const double2*)nullptr,
                       (double2*)L.x.get(), (double2*)L.d.get(), r2, (const double2*)c.inv_diag.get(),
                       0.0, ch.c1r, sc);
    hipLaunchKernelGGL(k0_cheb, dim3(gs), dim3(kT), 0, c.stream, A, (const double2*)L.x.get(), z2,
                       (double2*)L.d.get(), r2, (const double2*)c.inv_diag.get(), ch.c2d, ch.c2r, sc);
  }
  // the learned smoother's cycle computes its level-0 residuals with the FP32 values too
  const bool v32 = learned && mg.coarse_fp32;
  const MatM A32{c.sell_off.get(), c.sell_col.get(), c.sell_val32.get(), c.sell_row0.get(),
                 c.sell_nrow.get(), c.nslices};
  auto resid0 = [&]() {
    if (v32)
      hipLaunchKernelGGL(k0_resid32, dim3(gs), dim3(kT), 0, c.stream, A32, (const double2*)z2, r2,
                         (double2*)L.res.get(), sc);
    else
      hipLaunchKernelGGL(k0_resid, dim3(gs), dim3(kT), 0, c.stream, A, (const double2*)z2, r2,
                         (double2*)L.res.get(), sc);
  };
  resid0();
  // ... and reads P in FP32 for its restriction and prolongation
  const bool p32 = v32 && L.p32;
  launch_restrict(c, L, L.res.get(), coarse_rhs(mg, 0), sc, p32);
  launch_prolong(c, 0, z, coarse_correction(c, 0, sc), sc, p32);
  // post-smoothing
  if (learned) {
    resid0();
    hipLaunchKernelGGL(k0_msmooth, dim3(gs), dim3(kT), 0, c.stream, M, (const double2*)L.res.get(),
                       (const double2*)z2, z2, (const double2*)zq, zq_part, sc);
  } else {
    hipLaunchKernelGGL(k0_cheb, dim3(gs), dim3(kT), 0, c.stream, A, (const double2*)z2,
                       (double2*)L.x.get(), (double2*)L.d.get(), r2, (const double2*)c.inv_diag.get(),
                       0.0, ch.c1r, sc);
    hipLaunchKernelGGL(k0_cheb, dim3(gs), dim3(kT), 0, c.stream, A, (const double2*)L.x.get(), z2,
                       (double2*)L.d.get(), r2, (const double2*)c.inv_diag.get(), ch.c2d, ch.c2r, sc);
  }
  HIP_CHECK(hipGetLastError());
  return learned && zq_part ? gs : 0;
}

// z = P C(P^T r): the cycle's coarse correction alone, without level-0 smoothing
void mg_enqueue_coarse(Context& c, const double* r, double* z) {
  Multigrid& mg = c.mg;
  if (!mg.ready) throw Error(CSTFEM_ERR_ORDER, "multilevel: not set up");
  MgLevel& L = *mg.lv[0];
  HIP_CHECK(hipMemsetAsync(z, 0, sizeof(double) * (size_t)L.n, c.stream));
  launch_restrict(c, L, r, coarse_rhs(mg, 0), nullptr);
  launch_prolong(c, 0, z, coarse_correction(c, 0, nullptr), nullptr);
  HIP_CHECK(hipGetLastError());
}

void mg_drop_prepared(Context& c) {
  c.mg.join_prepared();
  c.mg.pre_error = nullptr;
  c.mg.pre_agg.clear();
}

void mg_prepare(Context& c) {
  if (!c.have_pattern || !c.mesh_pattern || c.n != 2 * c.nb)
    throw Error(CSTFEM_ERR_ORDER, "multilevel: prepare needs a mesh pattern (build_stiffness_pattern first)");
  // from now on every mesh pattern build starts it as soon as the node graph exists
  c.mg.auto_prepare = true;
  if (c.mg.pre_started) return;  // started for this pattern (a new pattern drops it)
  mg_prepare_start(c);
}

void mg_prepare_start(Context& c) {
  Multigrid& mg = c.mg;
  mg_drop_prepared(c);
  if (!c.aux) HIP_CHECK(hipStreamCreateWithFlags(&c.aux, hipStreamNonBlocking));
  if (!mg.pre_queued) HIP_CHECK(hipEventCreateWithFlags(&mg.pre_queued, hipEventDisableTiming));
  if (!mg.pre_done) HIP_CHECK(hipEventCreateWithFlags(&mg.pre_done, hipEventDisableTiming));
  // staging: row pointers, columns, then the graph's bandwidth
  const size_t need = (size_t)c.nb + 1 + (size_t)c.nblk + 1;
  // growing the staging frees the old one, which uploads of an earlier set-up may still read
  if (need > mg.pinned_n) HIP_CHECK(hipStreamSynchronize(c.stream));
  int* pin = pinned_ints(mg, need);
  int* bw = grow(mg.iwork, 1);
  HIP_CHECK(hipMemsetAsync(bw, 0, sizeof(int), c.stream));
  hipLaunchKernelGGL(k_graph_bandwidth, dim3(std::min(grid1(c.nb), kBandwidthGrid)), dim3(kT), 0,
                     c.stream, c.nb,
                     (const int*)c.brow_ptr.get(), (const int*)c.bcol.get(), bw);
  HIP_CHECK(hipGetLastError());
  // the download follows the work queued so far on the main stream (the pattern's kernels,
  // the bandwidth, the uploads of an earlier set-up from the same staging)
  HIP_CHECK(hipEventRecord(mg.pre_queued, c.stream));
  HIP_CHECK(hipStreamWaitEvent(c.aux, mg.pre_queued, 0));
  // row pointers and bandwidth first (the worker starts on them), then the columns in pieces
  HIP_CHECK(hipMemcpyAsync(pin, c.brow_ptr.get(), sizeof(int) * (c.nb + 1), hipMemcpyDeviceToHost, c.aux));
  HIP_CHECK(hipMemcpyAsync(pin + c.nb + 1 + c.nblk, bw, sizeof(int), hipMemcpyDeviceToHost, c.aux));
  HIP_CHECK(hipEventRecord(mg.pre_done, c.aux));
  constexpr int K = Multigrid::kPrePieces;
  int64_t piece_end[K];
  for (int p = 0; p < K; ++p) {
    const int64_t lo = c.nblk * p / K, hi = c.nblk * (p + 1) / K;
    piece_end[p] = hi;
    if (!mg.pre_piece[p]) HIP_CHECK(hipEventCreateWithFlags(&mg.pre_piece[p], hipEventDisableTiming));
    if (hi > lo)
      HIP_CHECK(hipMemcpyAsync(pin + c.nb + 1 + lo, c.bcol.get() + lo, sizeof(int) * (hi - lo),
                               hipMemcpyDeviceToHost, c.aux));
    HIP_CHECK(hipEventRecord(mg.pre_piece[p], c.aux));
  }
  const int64_t nblk = c.nblk;
  const int nb = c.nb;
  const hipEvent_t done = mg.pre_done;
  std::array<hipEvent_t, K> pieces;
  std::array<int64_t, K> ends;
  for (int p = 0; p < K; ++p) {
    pieces[p] = mg.pre_piece[p];
    ends[p] = piece_end[p];
  }
  mg.pre_error = nullptr;
  mg.pre_staged = false;
  const size_t room = mg.pinned_n;
  mg.pre_worker = std::thread([&mg, pin, nb, nblk, done, room, pieces, ends]() {
    try {
      const bool trace = std::getenv("CSTFEM_MG_TRACE") != nullptr;
      const auto t0 = std::chrono::steady_clock::now();
      HIP_CHECK(hipEventSynchronize(done));
      const auto t1 = std::chrono::steady_clock::now();
      int next = 0;  // the first piece not yet waited for
      auto need = [&](int64_t m) -> int64_t {
        while (next < K && (next == 0 ? 0 : ends[next - 1]) < m) HIP_CHECK(hipEventSynchronize(pieces[next++]));
        return next == 0 ? 0 : ends[next - 1];
      };
      mg.pre_nagg = aggregate_pipelined(pin, pin + nb + 1, nb, pin[nb + 1 + nblk], mg.pre_agg, need);
      const auto t2 = std::chrono::steady_clock::now();
      // the map into the staging area for its upload, over the node graph (no longer read)
      if ((size_t)nb <= room) {
        std::copy(mg.pre_agg.begin(), mg.pre_agg.end(), pin);
        mg.pre_staged = true;
      }
      if (trace) {
        const auto t3 = std::chrono::steady_clock::now();
        auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
        std::fprintf(stderr,
                     "[mg prepare] row pointers %.3f ms, aggregation with the columns arriving "
                     "%.3f ms, staging %.3f ms\n",
                     ms(t0, t1), ms(t1, t2), ms(t2, t3));
      }
    } catch (...) {
      mg.pre_error = std::current_exception();
    }
  });
  mg.pre_started = true;
}

// The helper thread of mg_setup that enqueues the levels' spectral bounds (see there).
class BoundHelper {
 public:
  BoundHelper(Context& c, double* lam_dev) : c_(c), lam_dev_(lam_dev) {}
  ~BoundHelper() {
    try {
      finish();
    } catch (...) {  // an error already on its way out of the set-up
    }
  }
  // the bound of `level` after event `ev` of the main stream
  void push(int level, hipEvent_t ev) {
    if (!th_.joinable()) th_ = std::thread([this]() { serve(); });
    {
      std::lock_guard<std::mutex> lk(mu_);
      q_.emplace_back(level, ev);
    }
    cv_.notify_one();
  }
  // waits until every pushed bound is enqueued; rethrows the helper's error
  void finish() {
    if (th_.joinable()) {
      {
        std::lock_guard<std::mutex> lk(mu_);
        done_ = true;
      }
      cv_.notify_one();
      th_.join();
    }
    if (err_) {
      const std::exception_ptr e = err_;
      err_ = nullptr;
      std::rethrow_exception(e);
    }
  }

 private:
  void serve() {
    if (hipSetDevice(c_.device) != hipSuccess) err_ = std::make_exception_ptr(
        Error(CSTFEM_ERR_DEVICE, "multilevel: hipSetDevice failed on the bound helper"));
    for (;;) {
      std::pair<int, hipEvent_t> job;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&]() { return done_ || !q_.empty(); });
        if (q_.empty()) return;
        job = q_.front();
        q_.pop_front();
      }
      const int level = job.first;
      if (err_) continue;  // after an error: drain the queue, enqueue nothing more
      try {
        HIP_CHECK(hipStreamWaitEvent(c_.aux, job.second, 0));
        smoothing_max_enqueue(c_, level, c_.aux, c_.mg.dwork3, lam_dev_ + level);
      } catch (...) {
        err_ = std::current_exception();
      }
    }
  }
  Context& c_;
  double* lam_dev_;
  std::thread th_;
  std::mutex mu_;
  std::condition_variable cv_;
  std::deque<std::pair<int, hipEvent_t>> q_;
  bool done_ = false;
  std::exception_ptr err_;
};

void mg_setup(Context& c, int max_levels, int coarse_max, int fine_smoother, int kdepth) {
  if (!c.setup_done) throw Error(CSTFEM_ERR_ORDER, "multilevel: call cstfem_setup first");
  if (!c.mesh_pattern || c.n != 2 * c.nb)
    throw Error(CSTFEM_ERR_ORDER, "multilevel: needs an assembled mesh system");
  if (max_levels < 2 || max_levels > 16) throw Error(CSTFEM_ERR_ARGUMENT, "multilevel: max_levels must be 2..16");
  if (coarse_max < 3 || coarse_max > 4096) throw Error(CSTFEM_ERR_ARGUMENT, "multilevel: coarse_max must be 3..4096");
  if (fine_smoother != 0 && fine_smoother != 1) throw Error(CSTFEM_ERR_ARGUMENT, "multilevel: fine_smoother must be 0 or 1");
  if (kdepth < 0) throw Error(CSTFEM_ERR_ARGUMENT, "multilevel: kcycle_levels must be >= 0");
  Multigrid& mg = c.mg;
  SetupTrace tr;
  mg.clear();
  tr.mark(c, 0, "clear");
  mg.fine_smoother = fine_smoother;
  mg.kdepth = kdepth;
  mg.kdepth_set = kdepth;
  {
    // FP32 copies of the coarse values for the learned smoother's cycle (CSTFEM_MG_COARSE32=0
    // keeps FP64); the Chebyshev-smoothed cycle always reads FP64: its iteration count at the
    // FP64 residual floor moves with such perturbations (72 -> 90 on the 4M plate)
    const char* env = std::getenv("CSTFEM_MG_COARSE32");
    mg.coarse_fp32 = fine_smoother == 0 && !(env && env[0] == '0');
    const char* fuse = std::getenv("CSTFEM_MG_FUSE");
    mg.fuse = !(fuse && fuse[0] == '0');
    const char* fuse_n = std::getenv("CSTFEM_MG_FUSE_N");
    mg.fuse_n = fuse_n ? std::atoi(fuse_n) : 200000;
    const char* fold = std::getenv("CSTFEM_MG_FOLD");
    mg.fold_max = fold ? std::atoll(fold) : 200000;
    const char* lanes = std::getenv("CSTFEM_MG_SETUP_LANES");
    mg.setup_lanes = !(lanes && lanes[0] == '0');
  }
  // level 0: the context's matrix; its node graph is read on the host (aggregation) from a
  // page-locked copy.  The pattern is complete since build_pattern returned, so the copy runs
  // on the auxiliary stream and the host aggregation overlaps whatever the main stream is
  // still doing (building M, typically); the device work below follows it in stream order.
  // the check of the learned smoother needs K and M only: it runs on the auxiliary stream,
  // after the work queued so far (M), while this thread and the main stream build the levels
  double* bound_dev = nullptr;
  if (!c.aux) HIP_CHECK(hipStreamCreateWithFlags(&c.aux, hipStreamNonBlocking));
  if (fine_smoother == 0 && c.mpnn.built) {
    if (!mg.bound_ev) HIP_CHECK(hipEventCreateWithFlags(&mg.bound_ev, hipEventDisableTiming));
    HIP_CHECK(hipEventRecord(mg.bound_ev, c.stream));
    HIP_CHECK(hipStreamWaitEvent(c.aux, mg.bound_ev, 0));
    bound_dev = learned_smoother_bound_enqueue(c, kSmootherSteps, c.aux);
  }
  MgLevel* L0 = mg.take_level();
  L0->val32.release();  // level 0's FP32 values are the context's
  L0->n = c.n;
  L0->nodes = c.nb;
  L0->bs = 2;
  // a level-0 aggregation started by mg_prepare belongs to the current pattern (every
  // pattern change drops it): take its result
  const bool prepared = mg.pre_started;
  int* pin = nullptr;
  if (prepared) {
    mg.join_prepared();
    if (mg.pre_error) {
      const std::exception_ptr e = mg.pre_error;
      mg.pre_error = nullptr;
      std::rethrow_exception(e);
    }
    L0->h_agg.swap(mg.pre_agg);
    L0->nagg = mg.pre_nagg;
    tr.mark(c, 0, "join prepared aggregation");
  } else {
    pin = pinned_ints(mg, (size_t)c.nb + 1 + (size_t)c.nblk);
    HIP_CHECK(hipMemcpyAsync(pin, c.brow_ptr.get(), sizeof(int) * (c.nb + 1), hipMemcpyDeviceToHost, c.aux));
    HIP_CHECK(hipMemcpyAsync(pin + c.nb + 1, c.bcol.get(), sizeof(int) * c.nblk, hipMemcpyDeviceToHost, c.aux));
    HIP_CHECK(hipStreamSynchronize(c.aux));
    tr.mark(c, 0, "download pattern");
  }
  alloc_work(c, *L0);
  // smoothing bounds of the levels run on the auxiliary stream, each after its level's matrix
  // is built on the main stream; the set-up collects them at the end
  double* lam_dev = grow(mg.lam_dev, kLamSlots);
  // A level's bound (about 45 launches) is enqueued on the auxiliary stream by a helper thread,
  // so that this thread goes on with the next level meanwhile: each level gets an event on
  // the main stream after its matrix, and the helper makes the auxiliary stream wait for it.
  // The helper takes the levels in order, so the auxiliary stream's work is the same as when
  // this thread enqueued it.
  BoundHelper bounds(c, lam_dev);
  auto bound_async = [&](int level) {
    while ((int)mg.lvl_ev.size() <= level) {
      hipEvent_t e;
      HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      mg.lvl_ev.push_back(e);
    }
    HIP_CHECK(hipEventRecord(mg.lvl_ev[level], c.stream));
    bounds.push(level, mg.lvl_ev[level]);
  };
  if (fine_smoother == 1) bound_async(0);
  tr.mark(c, 0, "work, smoothing");
  // rigid-body rotation about the centre of the mesh's bounding box
  double xlo = c.h_xy[0], xhi = c.h_xy[0], ylo = c.h_xy[1], yhi = c.h_xy[1];
  for (int i = 0; i < c.nb; ++i) {
    xlo = std::min(xlo, c.h_xy[2 * (size_t)i]);
    xhi = std::max(xhi, c.h_xy[2 * (size_t)i]);
    ylo = std::min(ylo, c.h_xy[2 * (size_t)i + 1]);
    yhi = std::max(yhi, c.h_xy[2 * (size_t)i + 1]);
  }
  const double cx = 0.5 * (xlo + xhi), cy = 0.5 * (ylo + yhi);
  const unsigned char* fixed = c.fixed_valid && c.fixed.n >= (size_t)c.n ? c.fixed.get() : nullptr;
  bool inv_on_device = false;
  for (int l = 0;; ++l) {
    MgLevel& L = *mg.lv[l];
    const bool last = l + 1 >= max_levels - 1;
    // aggregation (host, greedy in node order) and members (ascending node order)
    if (l > 0) {
      // this level's block pattern, downloaded while the level above was built: the host
      // aggregation below overlaps that level's Galerkin product on the device
      HIP_CHECK(hipEventSynchronize(mg.pat_ev));
      L.bcol.assign(mg.pinned_pat, mg.pinned_pat + L.bptr[L.nodes]);
    }
    if (!(l == 0 && prepared)) {
      const int* hp = l == 0 ? pin : L.bptr.data();
      const int* hc = l == 0 ? pin + c.nb + 1 : L.bcol.data();
      L.nagg = aggregate(hp, hc, L.nodes, L.h_agg);
    }
    const int na = L.nagg;
    plan_resid_restrict(L);
    // members and the aggregate map go up from page-locked memory; the staging area follows
    // the node graph, which the host no longer reads once the aggregation is done
    // the aggregate map goes up from page-locked staging; the member lists are built on the
    // device from it
    int* stage;
    if (l == 0 && prepared && mg.pre_staged) {
      stage = mg.pinned;  // the worker copied the map there
    } else {
      // the upload of the level above reads the staging area: let it finish (on level 0 only
      // the node-graph download used it, and that was waited for); growing it frees it
      if ((size_t)L.nodes > mg.pinned_n) HIP_CHECK(hipStreamSynchronize(c.stream));
      else if (l > 0) HIP_CHECK(hipEventSynchronize(mg.map_ev));
      stage = pinned_ints(mg, std::max(mg.pinned_n, (size_t)L.nodes));
      std::copy(L.h_agg.begin(), L.h_agg.end(), stage);
    }
    L.agg.alloc(L.nodes);
    L.mptr.alloc(na + 1);
    L.mem.alloc(L.nodes);
    HIP_CHECK(hipMemcpyAsync(L.agg.get(), stage, sizeof(int) * L.nodes, hipMemcpyHostToDevice, c.stream));
    if (!mg.map_ev) HIP_CHECK(hipEventCreateWithFlags(&mg.map_ev, hipEventDisableTiming));
    HIP_CHECK(hipEventRecord(mg.map_ev, c.stream));
    // iwork: flags[2] (coarse degree overflow, aggregate too large), counts[na + 1] (later the
    // coarse degrees), cursors[na]
    const size_t niw = 2 + 2 * (size_t)na + 1;
    int* iw = grow(mg.iwork, niw);
    int* flags = iw;
    int* cnt = iw + 2;
    int* cursor = cnt + na + 1;
    HIP_CHECK(hipMemsetAsync(iw, 0, sizeof(int) * niw, c.stream));
    hipLaunchKernelGGL(k_member_count, dim3(grid1(L.nodes)), dim3(kT), 0, c.stream, L.nodes,
                       (const int*)L.agg.get(), cnt);
    exclusive_scan_int(c, cnt, L.mptr.get(), na + 1, nullptr);
    hipLaunchKernelGGL(k_member_fill, dim3(grid1(L.nodes)), dim3(kT), 0, c.stream, L.nodes,
                       (const int*)L.agg.get(), (const int*)L.mptr.get(), cursor, L.mem.get());
    hipLaunchKernelGGL(k_member_sort, dim3(grid1(na)), dim3(kT), 0, c.stream, na, L.bs,
                       (const int*)L.mptr.get(), L.mem.get(), flags + 1);
    HIP_CHECK(hipGetLastError());
    tr.mark(c, l, "aggregate, members");
    MgLevel* C = mg.take_level();
    C->nodes = na;
    C->n = 3 * na;
    C->bs = 3;
    // tentative prolongator and the next near-null space (device)
    L.P.alloc((size_t)L.n * 3);
    C->nns.alloc((size_t)na * 9);
    {
      const MgsArgs ga{L.bs, L.mptr.get(), L.mem.get(), l == 0 ? nullptr : L.nns.get(), c.xy.get(),
                       fixed, cx, cy, L.P.get(), C->nns.get()};
      if (mg.setup_lanes && L.bs == 2 && na >= kMgsLaneMin) {
        // a lane per aggregate; aggregates above kLaneRows rows, listed, one wavefront each
        int* big = grow(mg.big, (size_t)na + 1);
        HIP_CHECK(hipMemsetAsync(big, 0, sizeof(int), c.stream));
        hipLaunchKernelGGL(k_mgs_lanes<2>, dim3(grid1(na)), dim3(kT), 0, c.stream, na, ga, big + 1,
                           big);
        hipLaunchKernelGGL(k_mgs, dim3(kMgsListGrid), dim3(kT), 0, c.stream, na,
                           (const int*)(big + 1), (const int*)big, ga);
      } else {
        hipLaunchKernelGGL(k_mgs, dim3(grid1(64 * (int64_t)na)), dim3(kT), 0, c.stream, na,
                           (const int*)nullptr, (const int*)nullptr, ga);
      }
    }
    HIP_CHECK(hipGetLastError());
    L.p32 = l == 0 && mg.coarse_fp32;
    if (L.p32) {  // the learned cycle's FP32 copy of level 0's P (4M DOFs: 48 MB less per pass)
      L.P32.alloc(3 * (size_t)L.n);
      hipLaunchKernelGGL(k_round32, dim3(grid1(3 * (int64_t)L.n)), dim3(kT), 0, c.stream,
                         3 * (int64_t)L.n, (const double*)L.P.get(), L.P32.get());
      HIP_CHECK(hipGetLastError());
    }
    tr.mark(c, l, "gram-schmidt");
    // coarse block pattern (device): sorted distinct neighbour aggregates of each aggregate
    const int* dbp = l == 0 ? c.brow_ptr.get() : L.d_bptr.get();
    const int* dbc = l == 0 ? c.bcol.get() : L.d_bcol.get();
    int* deg = cnt;  // the counts were consumed by the scan
    // many small aggregates: a lane each, the few larger rows listed for a wavefront each
    const bool lanes = mg.setup_lanes && na >= kCpatLaneMin;
    int* big = lanes ? grow(mg.big, (size_t)na + 1) : nullptr;
    if (lanes) {
      HIP_CHECK(hipMemsetAsync(big, 0, sizeof(int), c.stream));
      hipLaunchKernelGGL(k_cpat_count_lanes, dim3(grid1(na)), dim3(kT), 0, c.stream, na,
                         (const int*)L.mptr.get(), (const int*)L.mem.get(), dbp, dbc,
                         (const int*)L.agg.get(), deg, big + 1, big);
    }
    hipLaunchKernelGGL(k_cpat_count, dim3(lanes ? kCpatListGrid : na), dim3(64), 0, c.stream, na,
                       (const int*)(lanes ? big + 1 : nullptr), (const int*)big,
                       (const int*)L.mptr.get(), (const int*)L.mem.get(), dbp, dbc,
                       (const int*)L.agg.get(), deg, flags);
    HIP_CHECK(hipGetLastError());
    // the degrees and the two flags through the context's page-locked staging (a pageable
    // destination takes the runtime's slower staged path)
    int* hdeg = reinterpret_cast<int*>(c.stage(sizeof(int) * ((size_t)na + 2)));
    HIP_CHECK(hipMemcpyAsync(hdeg, deg, sizeof(int) * na, hipMemcpyDeviceToHost, c.stream));
    HIP_CHECK(hipMemcpyAsync(hdeg + na, flags, sizeof(int) * 2, hipMemcpyDeviceToHost, c.stream));
    HIP_CHECK(hipStreamSynchronize(c.stream));
    const int hflag[2] = {hdeg[na], hdeg[na + 1]};
    if (hflag[1]) throw Error(CSTFEM_ERR_ARGUMENT, "multilevel: aggregate too large");
    if (hflag[0])
      throw Error(CSTFEM_ERR_ARGUMENT, "multilevel: an aggregate touches more than " +
                                           std::to_string(kMaxCoarseDeg) + " aggregates");
    C->bptr.assign(na + 1, 0);
    int max_deg = 0;
    for (int a = 0; a < na; ++a) {
      C->bptr[a + 1] = C->bptr[a] + hdeg[a];
      max_deg = std::max(max_deg, hdeg[a]);
    }
    const int ncblk = C->bptr[na];
    // the coarse row pointers and the slice offsets of its SELL layout go up through the
    // page-locked staging, queued on the stream (the next level's use of the staging follows
    // them in stream order, after this level's last wait)
    sell_offsets(*C, C->bptr);
    {
      const size_t nb1 = (size_t)na + 1, ns1 = (size_t)C->nslices + 1;
      unsigned char* st = c.stage(sizeof(int64_t) * ns1 + sizeof(int) * nb1);
      int64_t* s_off = reinterpret_cast<int64_t*>(st);
      int* s_bptr = reinterpret_cast<int*>(st + sizeof(int64_t) * ns1);
      std::copy(C->h_off.begin(), C->h_off.end(), s_off);
      std::copy(C->bptr.begin(), C->bptr.end(), s_bptr);
      C->d_bptr.alloc(nb1);
      C->off.alloc(ns1);
      HIP_CHECK(hipMemcpyAsync(C->d_bptr.get(), s_bptr, sizeof(int) * nb1, hipMemcpyHostToDevice, c.stream));
      HIP_CHECK(hipMemcpyAsync(C->off.get(), s_off, sizeof(int64_t) * ns1, hipMemcpyHostToDevice, c.stream));
    }
    C->d_bcol.alloc(std::max(ncblk, 1));
    if (lanes)
      hipLaunchKernelGGL(k_cpat_fill_lanes, dim3(grid1(na)), dim3(kT), 0, c.stream, na,
                         (const int*)L.mptr.get(), (const int*)L.mem.get(), dbp, dbc,
                         (const int*)L.agg.get(), (const int*)C->d_bptr.get(), C->d_bcol.get());
    hipLaunchKernelGGL(k_cpat_fill, dim3(lanes ? kCpatListGrid : na), dim3(64), 0, c.stream, na,
                       (const int*)(lanes ? big + 1 : nullptr), (const int*)big,
                       (const int*)L.mptr.get(), (const int*)L.mem.get(), dbp, dbc,
                       (const int*)L.agg.get(), (const int*)C->d_bptr.get(), C->d_bcol.get());
    HIP_CHECK(hipGetLastError());
    // the pattern goes to the host for the next level's aggregation without holding this
    // thread: page-locked buffer, event
    if (!mg.pat_ev) HIP_CHECK(hipEventCreateWithFlags(&mg.pat_ev, hipEventDisableTiming));
    HIP_CHECK(hipMemcpyAsync(pinned_pattern(mg, std::max(ncblk, 1)), C->d_bcol.get(),
                             sizeof(int) * ncblk, hipMemcpyDeviceToHost, c.stream));
    HIP_CHECK(hipEventRecord(mg.pat_ev, c.stream));
    tr.mark(c, l, "coarse pattern");
    // scalar SELL-64 layout of the coarse matrix (offsets uploaded above)
    const int64_t nent = C->h_off[C->nslices];
    C->scol.alloc(std::max<int64_t>(nent / 3, 1));
    C->val.alloc(std::max<int64_t>(nent, 1));
    HIP_CHECK(hipMemsetAsync(C->val.get(), 0, sizeof(double) * std::max<int64_t>(nent, 1), c.stream));
    hipLaunchKernelGGL(k_sell_cols, dim3(grid1(64 * (int64_t)C->nslices)), dim3(kT), 0, c.stream, C->n,
                       C->nslices, (const int64_t*)C->off.get(), (const int*)C->d_bptr.get(),
                       (const int*)C->d_bcol.get(), C->scol.get());
    HIP_CHECK(hipGetLastError());
    tr.mark(c, l, "sell layout");
    // Galerkin product
    if (ncblk > 0) {
      const bool wide = max_deg > 64;
      const RapFine0 f0{c.sell_pos.get(), c.sell_val.get()};
      const RapFineS fs{L.off.get(), L.val.get()};
      const int* fp = l == 0 ? c.brow_ptr.get() : L.d_bptr.get();
      const int* fc = l == 0 ? c.bcol.get() : L.d_bcol.get();
#define CSTFEM_RAP(F, f, S)                                                                          \
  hipLaunchKernelGGL((k_rap<F, S>), dim3(na), dim3(64), 0, c.stream, (const int*)L.mptr.get(),     \
                     (const int*)L.mem.get(), (const int*)L.agg.get(), fp, fc, f,                  \
                     (const double*)L.P.get(), (const int*)C->d_bptr.get(),                        \
                     (const int*)C->d_bcol.get(), (const int64_t*)C->off.get(), C->val.get())
      if (l == 0) {
        if (wide) CSTFEM_RAP(RapFine0, f0, 2);
        else CSTFEM_RAP(RapFine0, f0, 1);
      } else {
        if (wide) CSTFEM_RAP(RapFineS, fs, 2);
        else CSTFEM_RAP(RapFineS, fs, 1);
      }
#undef CSTFEM_RAP
      HIP_CHECK(hipGetLastError());
    }
    C->dinv.alloc(C->n);
    hipLaunchKernelGGL(ks_dinv, dim3(grid1(C->n)), dim3(kT), 0, c.stream, mats(*C), C->dinv.get());
    HIP_CHECK(hipGetLastError());
    if (mg.coarse_fp32) {  // FP32 copy of the values for the cycle's row sums (kept across
                           // set-ups of the same size: no free and allocation per step)
      C->val32.alloc(C->val.n);
      hipLaunchKernelGGL(k_to_f32, dim3(1024), dim3(kT), 0, c.stream, (int64_t)C->val.n,
                         (const double*)C->val.get(), C->val32.get());
      HIP_CHECK(hipGetLastError());
    } else {
      C->val32.release();  // the cycle reads FP32 values wherever a level holds them
    }
    alloc_work(c, *C);
    tr.mark(c, l, "galerkin, dinv, work");
    if (last || C->n <= coarse_max || na == L.nodes) {
      if (C->n > 4096)
        throw Error(CSTFEM_ERR_ARGUMENT, "multilevel: coarsest level has " + std::to_string(C->n) +
                                             " DOFs (> 4096): allow more levels");
      if (coarsest_inverse_device(c, *C)) {
        inv_on_device = true;
        HIP_CHECK(hipEventSynchronize(mg.pat_ev));  // its pattern
        C->bcol.assign(mg.pinned_pat, mg.pinned_pat + ncblk);
      } else {
        HIP_CHECK(hipStreamSynchronize(c.stream));  // its values for the inverse, its pattern
        C->bcol.assign(mg.pinned_pat, mg.pinned_pat + ncblk);
        coarsest_inverse(c, *C);
      }
      tr.mark(c, l + 1, "coarsest inverse");
      break;
    }
    bound_async(l + 1);
    tr.mark(c, l + 1, "smoothing bound");
  }
  bounds.finish();  // every bound enqueued (rethrows the helper's error)
  int inv_flag = 0;
  if (inv_on_device)
    HIP_CHECK(hipMemcpyAsync(&inv_flag, mg.inv_flag.get(), sizeof(int), hipMemcpyDeviceToHost,
                             c.stream));
  HIP_CHECK(hipStreamSynchronize(c.stream));
  if (inv_flag)
    throw Error(CSTFEM_ERR_BAD_DIAGONAL, "multilevel: coarsest matrix is not positive definite");
  {
    // the smoothing bounds (levels below the coarsest; level 0 with the Chebyshev smoother)
    const int nl = (int)mg.lv.size();
    std::vector<double> lam(nl, 0.0);
    HIP_CHECK(hipMemcpyAsync(lam.data(), lam_dev, sizeof(double) * nl, hipMemcpyDeviceToHost, c.aux));
    HIP_CHECK(hipStreamSynchronize(c.aux));
    for (int l = fine_smoother == 1 ? 0 : 1; l + 1 < nl; ++l) mg.lv[l]->lam = lam[l];
  }
  mg.tail_from = tail_start(mg);
  tail_pack_build(c);
  if (mg.coarse_fp32) ensure_vals32(c);
  if (bound_dev) {
    // A good learned smoother has lambda_max(M A) close to 1 (1.14-1.16 on the plates it was
    // trained for); one far from its training data overshoots on some modes.  Such an M is
    // replaced by the Chebyshev smoother before any iteration is spent on it.
    HIP_CHECK(hipMemcpyAsync(&mg.m_bound, bound_dev, sizeof(double), hipMemcpyDeviceToHost, c.aux));
    HIP_CHECK(hipStreamSynchronize(c.aux));
    if (!(mg.m_bound <= kSmootherAccept)) {
      mg.ready = true;
      mg_set_fine_smoother(c, 1);
    }
    tr.mark(c, 0, "smoother bound");
  }
  mg.drop_spare();  // levels of a deeper earlier hierarchy
  mg.ready = true;
}

// aggregate of every node of `level` (host copy)
void mg_download_agg(Context& c, int level, int* out, int count) {
  if (!c.mg.ready) throw Error(CSTFEM_ERR_ORDER, "multilevel: not set up");
  if (level < 0 || level >= (int)c.mg.lv.size() - 1)
    throw Error(CSTFEM_ERR_ARGUMENT, "multilevel: no aggregation on level " + std::to_string(level));
  MgLevel& L = *c.mg.lv[level];
  if (count != L.nodes) throw Error(CSTFEM_ERR_ARGUMENT, "multilevel: size mismatch");
  std::copy(L.h_agg.begin(), L.h_agg.end(), out);
}

// matrix of a level >= 1 as scalar CSR (rows 3a + c: columns 3b + c' for the blocks b of a)
void mg_download_matrix(Context& c, int level, int* row_ptr, int* cols, double* vals) {
  if (!c.mg.ready) throw Error(CSTFEM_ERR_ORDER, "multilevel: not set up");
  if (level < 1 || level >= (int)c.mg.lv.size())
    throw Error(CSTFEM_ERR_ARGUMENT, "multilevel: level " + std::to_string(level) + " has no coarse matrix");
  MgLevel& L = *c.mg.lv[level];
  std::vector<double> val(L.val.n);
  HIP_CHECK(hipStreamSynchronize(c.stream));
  HIP_CHECK(hipMemcpy(val.data(), L.val.get(), sizeof(double) * val.size(), hipMemcpyDeviceToHost));
  int64_t k = 0;
  row_ptr[0] = 0;
  for (int r = 0; r < L.n; ++r) {
    const int a = r / 3, s = r >> 6, lane = r & 63;
    for (int t = L.bptr[a]; t < L.bptr[a + 1]; ++t)
      for (int cc = 0; cc < 3; ++cc) {
        const int e = 3 * (t - L.bptr[a]) + cc;
        cols[k] = 3 * L.bcol[t] + cc;
        vals[k] = val[L.h_off[s] + (int64_t)e * 64 + lane];
        ++k;
      }
    row_ptr[r + 1] = (int)k;
  }
}

int64_t mg_level_info(Context& c, int level, int* n, int* nagg, double* lam) {
  if (!c.mg.ready) throw Error(CSTFEM_ERR_ORDER, "multilevel: not set up");
  if (level < 0 || level >= (int)c.mg.lv.size()) return -1;
  MgLevel& L = *c.mg.lv[level];
  if (n) *n = L.n;
  if (nagg) *nagg = L.nagg;
  // level 0: lambda_max(M A) when a learned smoother was checked, else that of D^-1 A
  if (lam) *lam = (level == 0 && c.mg.m_bound != 0.0) ? c.mg.m_bound : L.lam;
  return level == 0 ? c.nblk * 4 : (int64_t)L.bcol.size() * 9;
}

int mg_levels(Context& c) { return c.mg.ready ? (int)c.mg.lv.size() : 0; }

void mg_tail_info(Context& c, int* first_level, int* packed) {
  if (!c.mg.ready) throw Error(CSTFEM_ERR_ORDER, "multilevel: not set up");
  if (first_level) *first_level = c.mg.tail_from;
  if (packed) *packed = c.mg.tail_packed >= 0 && c.mg.tail_packed == c.mg.tail_from ? 1 : 0;
}

// Switches the level-0 smoother of a hierarchy set up for the learned one.  The Chebyshev
// smoother takes the K-cycle on every level with it: with V-cycles below level 3 its
// iteration count on the heterogeneous plate depends on rounding (78-189 against 72).
void mg_set_fine_smoother(Context& c, int smoother) {
  Multigrid& mg = c.mg;
  if (!mg.ready) throw Error(CSTFEM_ERR_ORDER, "multilevel: not set up");
  if (smoother == mg.fine_smoother) return;
  if (smoother == 1) {
    if (!(mg.lv[0]->lam > 0.0)) mg.lv[0]->lam = smoothing_max(c, 0);
    mg.kdepth_set = mg.kdepth;
    mg.kdepth = std::max(mg.kdepth, (int)mg.lv.size());
  } else {
    mg.kdepth = mg.kdepth_set;
  }
  mg.fine_smoother = smoother;
  mg.tail_from = tail_start(mg);
  tail_pack_build(c);
}

}  // namespace cstfem_hip
