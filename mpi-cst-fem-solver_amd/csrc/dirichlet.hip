// dirichlet.hip -- Dirichlet elimination on the device.
//
// Reference: src/bc.cpp:137-166 (apply_dirichlet).  Only the rows the elimination can change
// are visited: the rows of constrained DOFs and the rows holding a constrained column.  On a
// structurally symmetric pattern those are the block rows listed in the constrained nodes'
// own rows, so the work is proportional to the constraints, not to the matrix:
//   k_dir_mark   one thread per constrained DOF sets its bit in a DOF bitmap (and its value in
//                a dense array of prescribed values) and appends the block rows of its node's
//                row (itself included) to a compact list, each once (a block-row bitmap
//                decides who appends; one counter increment per wavefront and chunk);
//   k_dir_rows   one thread per scalar row of the list visits its blocks in ascending column
//                order, so the subtractions b_i -= K_ik u_k happen in the reference's column
//                order and every value of K and b equals the reference's; the pattern is kept
//                (eliminated entries stay stored as zeros);
//   k_dir_clear  clears the bits it set, so the bitmaps are zero again for the next call.
// A pattern that is not structurally symmetric (a generic CSR upload) visits every row.
//
// Two entry points share these kernels: apply_dirichlet_device keeps its workspace in the
// context (grown on demand), apply_dirichlet_noalloc in `scratch`, a fixed device array per
// GPU, for callers that cannot allocate device memory.
#include <algorithm>
#include <climits>
#include <mutex>

#include "internal.hpp"

// Workspace of the allocation-free entry point: counters, the sorted constrained DOFs and
// their values, the DOF and block-row bitmaps and the block-row list, at fixed offsets.
// Limits: 2^25 scalar rows, 2^18 distinct constrained DOFs per call, 2^20 listed block rows.
namespace cstfem_hip {
namespace na {
constexpr int64_t kMaxDofs = int64_t(1) << 25;
constexpr int kMaxFixed = 1 << 18;
constexpr int kMaxRows = 1 << 20;
constexpr size_t kCtr = 0;                                        // 64 ints
constexpr size_t kDofs = 256;                                     // kMaxFixed ints
constexpr size_t kVals = kDofs + 4 * (size_t)kMaxFixed;           // kMaxFixed doubles
constexpr size_t kDofBits = kVals + 8 * (size_t)kMaxFixed;        // kMaxDofs / 32 words
constexpr size_t kRowBits = kDofBits + 4 * (size_t)(kMaxDofs / 32);  // kMaxDofs / 64 words
constexpr size_t kRows = kRowBits + 4 * (size_t)(kMaxDofs / 64);  // kMaxRows ints
constexpr size_t kBytes = kRows + 4 * (size_t)kMaxRows;           // 13 MiB + 256 B
}  // namespace na
}  // namespace cstfem_hip

// zero at load; every call leaves the bitmaps at zero again
static __device__ __align__(256) unsigned char scratch[cstfem_hip::na::kBytes];

namespace cstfem_hip {

namespace {

__global__ void k_dir_begin(int* __restrict__ ctr) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctr[0] = 0;        // block rows listed
    ctr[1] = 0;        // list overflow
    ctr[2] = INT_MAX;  // smallest constrained DOF with a non-positive diagonal
  }
}

__global__ void k_zero_bytes(unsigned char* __restrict__ p, int64_t count) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

__device__ inline bool test_bit(const unsigned* __restrict__ bits, int i) {
  return (bits[i >> 5] >> (i & 31)) & 1u;
}

constexpr int kMarkChunk = 16;  // rows of a node's row a lane marks at once

// Sets the DOF bit of every constrained DOF (and its prescribed value in uval, where given),
// and appends the block rows of its node's row (its own row first, even without a stored
// diagonal) to the compact list, each row once: a flag word per block row (row_flag), or
// where there is no room for one the block-row bitmap, decides which thread appends.  A lane
// takes its rows kMarkChunk at a time: their column loads, then their flag updates, are
// issued together, and the whole wavefront's appends of a chunk take a single
// counter increment (ballots give each appending lane its slot), so a chunk costs three
// memory round trips however many rows it appends.
__global__ __launch_bounds__(256) void k_dir_mark(int m, const int* __restrict__ dofs,
                                                  const double* __restrict__ vals,
                                                  const int* __restrict__ brow_ptr,
                                                  const int* __restrict__ bcol, int all_rows,
                                                  unsigned* __restrict__ dof_bits,
                                                  unsigned* __restrict__ row_bits,
                                                  int* __restrict__ row_flag,
                                                  int* __restrict__ rows, int* __restrict__ ctr,
                                                  int cap, double* __restrict__ uval) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = k < m;
  int v = 0, q0 = 0, len = 0;
  if (live) {
    const int dof = dofs[k];
    atomicOr(&dof_bits[dof >> 5], 1u << (dof & 31));
    if (uval) uval[dof] = vals[k];
    v = dof >> 1;
    if (!all_rows) {
      q0 = brow_ptr[v];
      len = 1 + brow_ptr[v + 1] - q0;
    }
  }
  if (all_rows) return;
  int steps = len;  // the wavefront's longest walk
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) steps = max(steps, __shfl_xor(steps, off, 64));
  const int lane = threadIdx.x & 63;
  const unsigned long long below = (1ull << lane) - 1ull;
  for (int u0 = 0; u0 < steps; u0 += kMarkChunk) {
    int r[kMarkChunk];
#pragma unroll
    for (int j = 0; j < kMarkChunk; ++j) {
      const int u = u0 + j;
      r[j] = u < len ? (u == 0 ? v : bcol[q0 + u - 1]) : -1;
    }
    bool fresh[kMarkChunk];
#pragma unroll
    for (int j = 0; j < kMarkChunk; ++j) {
      fresh[j] = false;
      if (r[j] >= 0) {
        if (row_flag) {  // one word per block row: only threads naming the same row contend
          fresh[j] = atomicExch(&row_flag[r[j]], 1) == 0;
        } else {
          const unsigned bit = 1u << (r[j] & 31);
          fresh[j] = !(atomicOr(&row_bits[r[j] >> 5], bit) & bit);
        }
      }
    }
    unsigned long long mask[kMarkChunk];
    int total = 0;
#pragma unroll
    for (int j = 0; j < kMarkChunk; ++j) {
      mask[j] = __ballot(fresh[j]);
      total += __popcll(mask[j]);
    }
    if (!total) continue;
    int base = 0;
    if (lane == 0) base = atomicAdd(&ctr[0], total);
    base = __shfl(base, 0, 64);
#pragma unroll
    for (int j = 0; j < kMarkChunk; ++j) {
      if (fresh[j]) {
        const int idx = base + __popcll(mask[j] & below);
        if (idx < cap)
          rows[idx] = r[j];
        else
          ctr[1] = 1;
      }
      base += __popcll(mask[j]);
    }
  }
}

// prescribed value of constrained DOF j: from the dense array uval where there is one, else by
// binary search in the sorted list
__device__ inline double ubar(int j, const double* __restrict__ uval, const int* __restrict__ dofs,
                              const double* __restrict__ vals, int m) {
  if (uval) return uval[j];
  int lo = 0, hi = m;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (dofs[mid] < j)
      lo = mid + 1;
    else
      hi = mid;
  }
  return vals[lo];
}

constexpr int kDirChunk = 8;  // blocks of a row whose column and bitmap loads are issued together

__global__ void k_dir_rows(int nb, int n, int all_rows, const int* __restrict__ rows,
                           const int* __restrict__ ctr, int cap, const int* __restrict__ brow_ptr,
                           const int* __restrict__ bcol, const int* __restrict__ bdiag,
                           const int64_t* __restrict__ spos, double* __restrict__ val,
                           const unsigned* __restrict__ dof_bits, const int* __restrict__ dofs,
                           const double* __restrict__ vals, int m,
                           const double* __restrict__ uval, double* __restrict__ b,
                           unsigned char* __restrict__ fixed_union, int* __restrict__ err) {
  if (!all_rows && ctr[1]) return;  // the list overflowed: change nothing
  const int64_t cnt = all_rows ? nb : min(ctr[0], cap);
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < 2 * cnt;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int r = all_rows ? (int)(t >> 1) : rows[t >> 1];
    const int c = (int)(t & 1);
    const int i = 2 * r + c;
    if (i >= n) continue;
    const int q0 = brow_ptr[r], q1 = brow_ptr[r + 1];
    if (test_bit(dof_bits, i)) {
      // constrained row: off-diagonal entries to zero, b_i = K_ii u_i
      fixed_union[i] = 1;
      const int qd = bdiag[r];
      const double diag = qd >= 0 ? val[sell_vidx(spos[qd], 3 * c)] : 0.0;
      if (!(diag > 0.0)) {
        atomicMin(err, i);
        continue;
      }
      for (int q = q0; q < q1; ++q) {
        double* v = val + sell_vidx(spos[q], 2 * c);  // row c: v[0], v[1]
        const int mc = bcol[q];
        if (2 * mc != i) v[0] = 0.0;
        if (2 * mc + 1 != i) v[1] = 0.0;
      }
      b[i] = diag * ubar(i, uval, dofs, vals, m);
    } else {
      // free row: b_i -= K_ik u_k and K_ik = 0 for every constrained column, in column order
      double bi = b[i];
      for (int q = q0; q < q1; q += kDirChunk) {
        int mc[kDirChunk];
        unsigned w[kDirChunk];
#pragma unroll
        for (int u = 0; u < kDirChunk; ++u) mc[u] = q + u < q1 ? bcol[q + u] : -1;
#pragma unroll
        for (int u = 0; u < kDirChunk; ++u) w[u] = mc[u] >= 0 ? dof_bits[mc[u] >> 4] : 0u;
#pragma unroll
        for (int u = 0; u < kDirChunk; ++u) {
          if (mc[u] < 0) continue;
          const int j0 = 2 * mc[u];  // even: j0 and j0 + 1 share a bitmap word
          const unsigned f = (w[u] >> (j0 & 31)) & 3u;
          if (!f) continue;
          double* v = val + sell_vidx(spos[q + u], 2 * c);
          if (f & 1u) {
            bi -= v[0] * ubar(j0, uval, dofs, vals, m);
            v[0] = 0.0;
          }
          if (f & 2u) {
            bi -= v[1] * ubar(j0 + 1, uval, dofs, vals, m);
            v[1] = 0.0;
          }
        }
      }
      b[i] = bi;
    }
  }
}

__global__ void k_dir_clear(int m, const int* __restrict__ dofs, const int* __restrict__ ctr,
                            int cap, const int* __restrict__ rows, int64_t row_words,
                            unsigned* __restrict__ dof_bits, unsigned* __restrict__ row_bits,
                            int* __restrict__ row_flag) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  for (int64_t k = t0; k < m; k += stride) dof_bits[dofs[k] >> 5] = 0u;
  const int listed = min(ctr[0], cap);
  for (int64_t k = t0; k < listed; k += stride) {
    if (row_flag)
      row_flag[rows[k]] = 0;
    else
      row_bits[rows[k] >> 5] = 0u;
  }
  if (ctr[1])  // rows beyond the list's capacity set bits it does not name
    for (int64_t k = t0; k < row_words; k += stride) row_bits[k] = 0u;
}

// Device pointers of one elimination's workspace: the sorted constrained DOFs and values, the
// two bitmaps (zero between calls), the block-row list and three counters.
struct DirWork {
  int* ctr;
  int* dofs;
  double* vals;
  unsigned* dof_bits;
  unsigned* row_bits;
  int* rows;
  int cap;
  double* uval;   // prescribed values by DOF (n entries; only constrained ones are read), or null
  int* row_flag;  // a flag word per block row (zero between calls) in place of row_bits, or null
};

// DirichletDofs::add (bc.cpp:107-116) over the caller's pairs: the first pair in input order
// whose DOF was given before with another value is a conflict; then the range check of
// apply_dirichlet (bc.cpp:141-146) over the DOFs in ascending order.  Returns the sorted
// distinct DOFs and their values.
void collect_dofs(int n, int m, const int* dofs, const double* values, std::vector<int>& d,
                  std::vector<double>& v) {
  std::vector<int> order(m);
  for (int k = 0; k < m; ++k) order[k] = k;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return dofs[a] < dofs[b]; });
  int conflict = -1, first_of_conflict = -1;
  d.clear();
  v.clear();
  for (int g = 0; g < m;) {
    int e = g + 1;
    while (e < m && dofs[order[e]] == dofs[order[g]]) ++e;
    const int k0 = order[g];  // the group's first pair in input order (stable sort)
    for (int h = g + 1; h < e; ++h)
      if (values[order[h]] != values[k0] && (conflict < 0 || order[h] < conflict)) {
        conflict = order[h];
        first_of_conflict = k0;
      }
    d.push_back(dofs[k0]);
    v.push_back(values[k0]);
    g = e;
  }
  if (conflict >= 0) {
    char buf[256];
    const int dof = dofs[conflict];
    snprintf(buf, sizeof buf,
             "conflicting Dirichlet values for node %d (0-based) component %s: %.17g vs %.17g",
             dof / 2, dof % 2 == 0 ? "ux" : "uy", values[first_of_conflict], values[conflict]);
    throw Error(CSTFEM_ERR_CONFLICT, buf);
  }
  for (int dof : d)
    if (dof < 0 || dof >= n)
      throw Error(CSTFEM_ERR_OUT_OF_RANGE,
                  "apply_dirichlet: DOF " + std::to_string(dof) + " out of range");
}

// enqueues the elimination of the sorted constrained DOFs d (values v), whose copies are
// already in w.dofs / w.vals on the library stream, and returns the counters
void run_elimination(Context& c, int m, const DirWork& w, int out_ctr[3], bool staged) {
  const int n = c.n;
  const int all_rows = c.pattern_symmetric ? 0 : 1;
  if (!c.fixed_valid) {  // union of the constrained DOFs since the last assembly
    hipLaunchKernelGGL(k_zero_bytes, dim3(grid_for((int64_t)c.fixed.n, 256, 1024)), dim3(256), 0,
                       c.stream, c.fixed.get(), (int64_t)c.fixed.n);
    HIP_CHECK(hipGetLastError());
    c.fixed_valid = true;
  }
  hipLaunchKernelGGL(k_dir_begin, dim3(1), dim3(64), 0, c.stream, w.ctr);
  HIP_CHECK(hipGetLastError());
  if (m > 0) {
    hipLaunchKernelGGL(k_dir_mark, dim3(grid_for(m, 256, 0)), dim3(256), 0, c.stream, m, w.dofs,
                       (const double*)w.vals, c.brow_ptr.get(), c.bcol.get(), all_rows, w.dof_bits,
                       w.row_bits, w.row_flag, w.rows, w.ctr, w.cap, w.uval);
    HIP_CHECK(hipGetLastError());
    // rows to visit: at most the distinct constrained nodes times (1 + the widest row)
    const int64_t bound = all_rows ? (int64_t)c.nb
                                   : std::min<int64_t>((int64_t)w.cap,
                                                       (int64_t)m * (1 + c.max_row_blocks));
    hipLaunchKernelGGL(k_dir_rows, dim3(grid_for(2 * bound, 256, 4096)), dim3(256), 0, c.stream,
                       c.nb, n, all_rows, w.rows, w.ctr, w.cap, c.brow_ptr.get(), c.bcol.get(),
                       c.bdiag.get(), c.sell_pos.get(), c.sell_val.get(), w.dof_bits, w.dofs,
                       w.vals, m, (const double*)w.uval, c.b.get(), c.fixed.get(), w.ctr + 2);
    HIP_CHECK(hipGetLastError());
    const int64_t row_words = (c.nb + 31) / 32;
    hipLaunchKernelGGL(k_dir_clear, dim3(grid_for(std::max<int64_t>(m, 1), 256, 1024)), dim3(256), 0,
                       c.stream, m, w.dofs, w.ctr, w.cap, w.rows, row_words, w.dof_bits,
                       w.row_bits, w.row_flag);
    HIP_CHECK(hipGetLastError());
  }
  // the counters through the page-locked staging, or (staged = false: the allocation-free
  // entry point, which may not create the staging) straight into out_ctr
  int* h_ctr = staged ? reinterpret_cast<int*>(c.stage(sizeof(int) * 3)) : out_ctr;
  HIP_CHECK(hipMemcpyAsync(h_ctr, w.ctr, sizeof(int) * 3, hipMemcpyDeviceToHost, c.stream));
  HIP_CHECK(hipStreamSynchronize(c.stream));
  if (staged) std::copy(h_ctr, h_ctr + 3, out_ctr);
}

// the sorted DOFs and their values into w.dofs / w.vals, through the context's page-locked
// staging (values first: 8-byte aligned)
void upload_list(Context& c, const std::vector<int>& d, const std::vector<double>& v,
                 const DirWork& w) {
  const size_t mu = d.size();
  unsigned char* h = c.stage(mu * (sizeof(double) + sizeof(int)));
  double* hv = reinterpret_cast<double*>(h);
  int* hd = reinterpret_cast<int*>(h + mu * sizeof(double));
  std::copy(v.begin(), v.end(), hv);
  std::copy(d.begin(), d.end(), hd);
  HIP_CHECK(hipMemcpyAsync(w.vals, hv, sizeof(double) * mu, hipMemcpyHostToDevice, c.stream));
  HIP_CHECK(hipMemcpyAsync(w.dofs, hd, sizeof(int) * mu, hipMemcpyHostToDevice, c.stream));
}

void check_preconditions(Context& c) {
  if (!c.have_pattern) throw Error(CSTFEM_ERR_ORDER, "apply_dirichlet: no matrix on the device");
  if (c.b.n == 0) throw Error(CSTFEM_ERR_ORDER, "apply_dirichlet: set the right-hand side first");
}

void finish(Context& c, const int ctr[3]) {
  c.setup_done = false;
  c.vals32_valid = false;
  if (ctr[2] != INT_MAX)
    throw Error(CSTFEM_ERR_BAD_DIAGONAL,
                "apply_dirichlet: non-positive diagonal at DOF " + std::to_string(ctr[2]));
}

template <class T>
T* grow(DBuf<T>& b, size_t count) {
  if (!b.get() || b.n < count) b.alloc(std::max<size_t>(count, 1));
  return b.get();
}

}  // namespace

void apply_dirichlet_device(Context& c, int m, const int* dofs, const double* values) {
  check_preconditions(c);
  std::vector<int> d;
  std::vector<double> v;
  collect_dofs(c.n, m, dofs, values, d, v);
  const int mu = (int)d.size();
  // workspace of the context, grown on demand and kept: repeated eliminations allocate
  // nothing; the bitmaps start at zero and every call leaves them at zero
  DirWork w{};
  w.ctr = c.err_flag.get();
  w.dofs = grow(c.dir_dofs, mu);
  w.vals = grow(c.dir_vals, mu);
  const size_t dof_words = ((size_t)c.n + 31) / 32, row_words = ((size_t)c.nb + 31) / 32;
  if (c.dir_bits.n < dof_words + row_words) {
    grow(c.dir_bits, dof_words + row_words);
    HIP_CHECK(hipMemsetAsync(c.dir_bits.get(), 0, sizeof(unsigned) * c.dir_bits.n, c.stream));
  }
  w.dof_bits = c.dir_bits.get();
  w.row_bits = c.dir_bits.get() + dof_words;
  if (c.dir_flag.n < (size_t)c.nb) {
    grow(c.dir_flag, (size_t)c.nb);
    HIP_CHECK(hipMemsetAsync(c.dir_flag.get(), 0, sizeof(int) * c.dir_flag.n, c.stream));
  }
  w.row_flag = c.dir_flag.get();
  w.rows = grow(c.dir_rows, (size_t)c.nb);
  w.cap = c.nb;  // every block row fits: the list never overflows
  w.uval = grow(c.dir_uval, (size_t)c.n);  // read only where this call set the DOF's bit
  if (mu > 0) upload_list(c, d, v, w);
  int ctr[3];
  run_elimination(c, mu, w, ctr, true);
  finish(c, ctr);
}

// ---- the allocation-free entry point ---------------------------------------------------

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu[kMaxDevices];  // calls on one GPU run one at a time
void* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_init_mu;
}  // namespace

void noalloc_attach(Context& c) {
  if (c.device < 0 || c.device >= kMaxDevices)
    throw Error(CSTFEM_ERR_ARGUMENT, "device index beyond the library's table of " +
                                         std::to_string(kMaxDevices));
  std::lock_guard<std::mutex> lk(g_scratch_init_mu);
  if (!g_scratch_ptr[c.device]) {
    // the array's device address on the current device; its host-side name is not one
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw Error(CSTFEM_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch) failed: ") +
                                         hipGetErrorString(e));
    g_scratch_ptr[c.device] = p;
  }
  c.na_scratch = static_cast<unsigned char*>(g_scratch_ptr[c.device]);
}

void apply_dirichlet_noalloc(Context& c, int m, const int* dofs, const double* values) {
  check_preconditions(c);
  if (!c.na_scratch) throw Error(CSTFEM_ERR_ORDER, "apply_dirichlet_noalloc: no scratch on this device");
  if ((int64_t)c.n > na::kMaxDofs)
    throw Error(CSTFEM_ERR_ARGUMENT, "apply_dirichlet_noalloc: " + std::to_string(c.n) +
                                         " DOFs exceed the scratch limit of " +
                                         std::to_string(na::kMaxDofs));
  if (c.fixed.n < (size_t)c.n)
    throw Error(CSTFEM_ERR_ORDER, "apply_dirichlet_noalloc: build the pattern first");
  std::vector<int> d;
  std::vector<double> v;
  collect_dofs(c.n, m, dofs, values, d, v);
  const int mu = (int)d.size();
  if (mu > na::kMaxFixed)
    throw Error(CSTFEM_ERR_ARGUMENT, "apply_dirichlet_noalloc: " + std::to_string(mu) +
                                         " constrained DOFs exceed the scratch limit of " +
                                         std::to_string(na::kMaxFixed));
  std::lock_guard<std::mutex> lk(g_scratch_mu[c.device]);
  unsigned char* base = c.na_scratch;
  DirWork w{};
  w.ctr = reinterpret_cast<int*>(base + na::kCtr);
  w.dofs = reinterpret_cast<int*>(base + na::kDofs);
  w.vals = reinterpret_cast<double*>(base + na::kVals);
  w.dof_bits = reinterpret_cast<unsigned*>(base + na::kDofBits);
  w.row_bits = reinterpret_cast<unsigned*>(base + na::kRowBits);
  w.rows = reinterpret_cast<int*>(base + na::kRows);
  w.cap = na::kMaxRows;
  w.uval = nullptr;      // no room for n values: binary search in the sorted list
  w.row_flag = nullptr;  // nor for a word per block row: the block-row bitmap
  if (mu > 0) {  // pageable copies: the page-locked staging would be an allocation
    HIP_CHECK(hipMemcpyAsync(w.dofs, d.data(), sizeof(int) * mu, hipMemcpyHostToDevice, c.stream));
    HIP_CHECK(hipMemcpyAsync(w.vals, v.data(), sizeof(double) * mu, hipMemcpyHostToDevice, c.stream));
  }
  int ctr[3];
  run_elimination(c, mu, w, ctr, false);
  if (ctr[1])
    throw Error(CSTFEM_ERR_ARGUMENT,
                "apply_dirichlet_noalloc: the elimination visits more than " +
                    std::to_string(na::kMaxRows) +
                    " block rows (the scratch limit); K and b are unchanged");
  finish(c, ctr);
}

}  // namespace cstfem_hip
