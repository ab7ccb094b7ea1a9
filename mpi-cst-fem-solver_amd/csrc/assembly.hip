// assembly.hip -- stiffness pattern, element-to-slot map and atomic-free CST assembly.
//
// Reference: src/assembly.cpp:9-74, src/element.cpp:7-55.  The pattern is built on the
// device and equals build_stiffness_pattern bit for bit (node adjacency through triangles,
// self included, ascending).  Assembly sums, for every 2x2 block, the Ke sub-blocks of the
// elements that share its two nodes in ascending element order starting from 0.0 -- the
// order in which assemble_stiffness's element loop adds them -- with no FMA contraction
// (this file is compiled with -ffp-contract=off), so every value equals the reference's.
#include <algorithm>
#include <climits>

#include "internal.hpp"

namespace cstfem_hip {

constexpr int kMaxElemsPerNode = 32;  // cap on triangles around one node

// ---------------------------------------------------------------------------------------
// exclusive scan of int counts (3-phase, recursive over block totals)

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;

__global__ void k_scan_tiles(const int* __restrict__ in, int* __restrict__ out,
                             int* __restrict__ tile_sums, int64_t count) {
  __shared__ int s[kScanThreads];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int v[kScanItems];
  int sum = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    v[i] = (base + i < count) ? in[base + i] : 0;
    sum += v[i];
  }
  s[threadIdx.x] = sum;
  __syncthreads();
  // Hillis-Steele inclusive scan of per-thread sums
  for (int off = 1; off < kScanThreads; off <<= 1) {
    const int t = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  int run = s[threadIdx.x] - sum;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    if (base + i < count) out[base + i] = run;
    run += v[i];
  }
  if (threadIdx.x == kScanThreads - 1) tile_sums[blockIdx.x] = s[threadIdx.x];
}

__global__ void k_scan_add(int* __restrict__ out, const int* __restrict__ tile_off,
                           int64_t count) {
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  const int add = tile_off[blockIdx.x];
  for (int i = threadIdx.x; i < kScanTile; i += kScanThreads)
    if (base + i < count) out[base + i] += add;
}

// scratch slots of the context (tmp_ints): the scan takes two per recursion depth
constexpr int kSlotScan = 0;     // 0 .. 2 * kScanDepth - 1
constexpr int kScanDepth = 4;    // tiles of tiles: 2048^4 elements at most
constexpr int kSlotCount = 8;    // per-node / per-block counts
constexpr int kSlotCursor = 9;   // fill cursors

void scan_rec(Context& c, const int* in, int* out, int64_t count, int depth) {
  if (depth >= kScanDepth) throw Error(CSTFEM_ERR_ARGUMENT, "scan: too many elements");
  const int64_t tiles = (count + kScanTile - 1) / kScanTile;
  int* sums = tmp_ints(c, kSlotScan + 2 * depth, tiles);
  int* sums_off = tmp_ints(c, kSlotScan + 2 * depth + 1, tiles);
  hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)tiles), dim3(kScanThreads), 0, c.stream, in,
                     out, sums, count);
  HIP_CHECK(hipGetLastError());
  if (tiles > 1) {
    scan_rec(c, sums, sums_off, tiles, depth + 1);
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)tiles), dim3(kScanThreads), 0, c.stream, out,
                       sums_off, count);
    HIP_CHECK(hipGetLastError());
  }
}

}  // namespace

// out[0..count): exclusive prefix sums of in; *total = sum of all
void exclusive_scan_int(Context& c, const int* in, int* out, int64_t count, int64_t* total) {
  if (count <= 0) {
    if (total) *total = 0;
    return;
  }
  scan_rec(c, in, out, count, 0);
  if (total) {
    int last_in = 0, last_out = 0;  // read in stream order after the scan kernels
    HIP_CHECK(hipMemcpyAsync(&last_in, in + count - 1, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    HIP_CHECK(hipMemcpyAsync(&last_out, out + count - 1, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    HIP_CHECK(hipStreamSynchronize(c.stream));
    *total = (int64_t)last_in + last_out;
  }
}

int* tmp_ints(Context& c, int slot, size_t count) {
  DBuf<int>& b = c.tmp_i[slot];
  // growing frees the old buffer; hipFree waits for the kernels still using it
  if (!b.get() || b.n < count) b.alloc(std::max<size_t>(count, 1));
  return b.get();
}

int grid_for(int64_t threads, int block, int max_grid) {
  int64_t g = (threads + block - 1) / block;
  if (g < 1) g = 1;
  if (max_grid > 0 && g > max_grid) g = max_grid;
  return (int)g;
}

// ---------------------------------------------------------------------------------------
// node -> elements (ascending element ids per node)

namespace {

__global__ void k_count_node_elems(const int* __restrict__ tri, int64_t ne, int* cnt) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ne;
       e += (int64_t)gridDim.x * blockDim.x)
    for (int a = 0; a < 3; ++a) atomicAdd(&cnt[tri[3 * e + a]], 1);
}

__global__ void k_fill_node_elems(const int* __restrict__ tri, int64_t ne,
                                  const int* __restrict__ ptr, int* cursor, int* list) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ne;
       e += (int64_t)gridDim.x * blockDim.x)
    for (int a = 0; a < 3; ++a) {
      const int n = tri[3 * e + a];
      const int pos = atomicAdd(&cursor[n], 1);
      list[ptr[n] + pos] = (int)e;
    }
}

__global__ void k_sort_node_elems(int nn, const int* __restrict__ ptr, int* list, int* err) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  const int b = ptr[n], e = ptr[n + 1];
  if (e - b > kMaxElemsPerNode) atomicMin(err, n);
  for (int i = b + 1; i < e; ++i) {  // insertion sort (short lists)
    const int v = list[i];
    int j = i - 1;
    while (j >= b && list[j] > v) {
      list[j + 1] = list[j];
      --j;
    }
    list[j + 1] = v;
  }
}

}  // namespace

void build_node_elements(Context& c) {
  const int nn = c.nn;
  const int64_t ne = c.ne;
  int* cnt = tmp_ints(c, kSlotCount, (size_t)nn + 1);
  int* cursor = tmp_ints(c, kSlotCursor, (size_t)nn);
  HIP_CHECK(hipMemsetAsync(cnt, 0, sizeof(int) * (nn + 1), c.stream));
  HIP_CHECK(hipMemsetAsync(cursor, 0, sizeof(int) * nn, c.stream));
  const int g = grid_for(ne, 256, 8192);
  hipLaunchKernelGGL(k_count_node_elems, dim3(g), dim3(256), 0, c.stream, c.tri.get(), ne, cnt);
  HIP_CHECK(hipGetLastError());
  c.ne_ptr.alloc(nn + 1);
  int64_t total = 0;
  exclusive_scan_int(c, cnt, c.ne_ptr.get(), nn + 1, &total);
  c.ne_list.alloc(3 * ne);
  hipLaunchKernelGGL(k_fill_node_elems, dim3(g), dim3(256), 0, c.stream, c.tri.get(), ne,
                     c.ne_ptr.get(), cursor, c.ne_list.get());
  HIP_CHECK(hipGetLastError());
  const int big = INT_MAX;
  HIP_CHECK(hipMemcpyAsync(c.err_flag.get(), &big, sizeof(int), hipMemcpyHostToDevice, c.stream));
  hipLaunchKernelGGL(k_sort_node_elems, dim3(grid_for(nn, 256, 0)), dim3(256), 0, c.stream, nn,
                     c.ne_ptr.get(), c.ne_list.get(), c.err_flag.get());
  HIP_CHECK(hipGetLastError());
  int bad = INT_MAX;
  HIP_CHECK(hipMemcpyAsync(&bad, c.err_flag.get(), sizeof(int), hipMemcpyDeviceToHost, c.stream));
  HIP_CHECK(hipStreamSynchronize(c.stream));
  if (bad != INT_MAX)
    throw Error(CSTFEM_ERR_ARGUMENT, "node " + std::to_string(bad) + " belongs to more than " +
                                         std::to_string(kMaxElemsPerNode) + " triangles");
}

// ---------------------------------------------------------------------------------------
// pattern: node adjacency through triangles, self included, sorted, unique

namespace {

// Collects adj(n) into `out` (ascending, unique); returns its length.
__device__ int node_adjacency(int n, const int* __restrict__ tri, const int* __restrict__ ptr,
                              const int* __restrict__ list, int* out) {
  int len = 0;
  out[len++] = n;
  for (int k = ptr[n]; k < ptr[n + 1]; ++k) {
    const int e = list[k];
    for (int a = 0; a < 3; ++a) {
      const int m = tri[3 * e + a];
      // insert m keeping `out` sorted and unique
      int j = len - 1;
      bool dup = false;
      while (j >= 0 && out[j] >= m) {
        if (out[j] == m) {
          dup = true;
          break;
        }
        --j;
      }
      if (dup) continue;
      for (int q = len; q > j + 1; --q) out[q] = out[q - 1];
      out[j + 1] = m;
      ++len;
    }
  }
  return len;
}

constexpr int kMaxAdj = 2 * kMaxElemsPerNode + 1;

__global__ void k_adj_count(int nn, const int* __restrict__ tri, const int* __restrict__ ptr,
                            const int* __restrict__ list, int* __restrict__ deg) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  int buf[kMaxAdj];
  deg[n] = node_adjacency(n, tri, ptr, list, buf);
}

__global__ void k_adj_fill(int nn, const int* __restrict__ tri, const int* __restrict__ ptr,
                           const int* __restrict__ list, const int* __restrict__ brow_ptr,
                           int* __restrict__ bcol, int* __restrict__ bdiag) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  int buf[kMaxAdj];
  const int len = node_adjacency(n, tri, ptr, list, buf);
  const int o = brow_ptr[n];
  for (int i = 0; i < len; ++i) {
    bcol[o + i] = buf[i];
    if (buf[i] == n) bdiag[n] = o + i;
  }
}

}  // namespace

void build_pattern(Context& c) {
  mg_drop_prepared(c);  // a level-0 aggregation under way reads the old pattern
  build_node_elements(c);
  const int nn = c.nn;
  int* deg = tmp_ints(c, kSlotCount, (size_t)nn + 1);
  HIP_CHECK(hipMemsetAsync(deg, 0, sizeof(int) * (nn + 1), c.stream));
  hipLaunchKernelGGL(k_adj_count, dim3(grid_for(nn, 128, 0)), dim3(128), 0, c.stream, nn,
                     c.tri.get(), c.ne_ptr.get(), c.ne_list.get(), deg);
  HIP_CHECK(hipGetLastError());
  c.brow_ptr.alloc(nn + 1);
  int64_t total = 0;
  exclusive_scan_int(c, deg, c.brow_ptr.get(), nn + 1, &total);
  if (4 * total > INT_MAX)
    throw Error(CSTFEM_ERR_ARGUMENT, "pattern exceeds 32-bit nnz (" + std::to_string(4 * total) + ")");
  c.nblk = total;
  c.bcol.alloc(total);
  c.bdiag.alloc(nn);
  hipLaunchKernelGGL(k_adj_fill, dim3(grid_for(nn, 128, 0)), dim3(128), 0, c.stream, nn,
                     c.tri.get(), c.ne_ptr.get(), c.ne_list.get(), c.brow_ptr.get(),
                     c.bcol.get(), c.bdiag.get());
  HIP_CHECK(hipGetLastError());
  c.nb = nn;
  c.n = 2 * nn;
  c.have_pattern = true;
  c.mesh_pattern = true;
  c.generic = false;
  c.setup_done = false;
  c.vals32_valid = false;
  // the node graph is complete on the stream: a context that asked for the multilevel
  // set-up's prepared aggregation starts it here, so that its download and its host pass
  // overlap the rest of the pattern (the SELL layout and the slot map)
  if (c.mg.auto_prepare) mg_prepare_start(c);
  build_sell(c);
  build_slot_map(c);
  HIP_CHECK(hipStreamSynchronize(c.stream));
}

// ---------------------------------------------------------------------------------------
// SELL-64 layout of the block rows

namespace {

__global__ void k_slice_width(int nslices, const int* __restrict__ row0,
                              const int* __restrict__ nrow, const int* __restrict__ brow_ptr,
                              int* __restrict__ width) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslices) return;
  int w = 0;
  for (int l = 0; l < nrow[s]; ++l) {
    const int r = row0[s] + l;
    w = max(w, brow_ptr[r + 1] - brow_ptr[r]);
  }
  width[s] = w;
}

// one wavefront per slice; lane = row within the slice
__global__ void k_sell_fill(int nslices, const int* __restrict__ row0, const int* __restrict__ nrow,
                            const int* __restrict__ brow_ptr, const int* __restrict__ bcol,
                            const int64_t* __restrict__ off, int* __restrict__ scol,
                            int64_t* __restrict__ spos) {
  const int lane = threadIdx.x & (kSlice - 1);
  const int s = blockIdx.x * (blockDim.x / kSlice) + threadIdx.x / kSlice;
  if (s >= nslices) return;
  const bool live = lane < nrow[s];
  const int r = row0[s] + lane;
  const int64_t o = off[s];
  const int w = (int)((off[s + 1] - o) / kSlice);
  const int b = live ? brow_ptr[r] : 0;
  const int d = live ? brow_ptr[r + 1] - b : 0;
  const int self = live ? r : row0[s];
  for (int k = 0; k < w; ++k) {
    const int64_t slot = o + (int64_t)k * kSlice + lane;
    if (k < d) {
      scol[slot] = bcol[b + k];
      spos[b + k] = slot;
    } else {
      scol[slot] = self;
    }
  }
}

}  // namespace

// Slices of at most 64 block rows that never cross a segment boundary (c.h_segments: ascending
// block-row boundaries 0 = s0 < s1 < ... = nb; empty = one segment).  Rows of slice s are
// row0[s] .. row0[s] + nrow[s] - 1.
void build_sell(Context& c) {
  const int nb = c.nb;
  std::vector<int> seg = c.h_segments;
  if (seg.empty()) seg = {0, nb};
  std::vector<int> row0, nrow;
  c.h_seg_slice.assign(1, 0);
  for (size_t k = 0; k + 1 < seg.size(); ++k) {
    for (int r = seg[k]; r < seg[k + 1]; r += kSlice) {
      row0.push_back(r);
      nrow.push_back(std::min(kSlice, seg[k + 1] - r));
    }
    c.h_seg_slice.push_back(static_cast<int>(row0.size()));
  }
  const int nslices = static_cast<int>(row0.size());
  c.sell_row0.alloc(nslices + 1);
  c.sell_nrow.alloc(nslices + 1);
  HIP_CHECK(hipMemcpyAsync(c.sell_row0.get(), row0.data(), sizeof(int) * nslices,
                           hipMemcpyHostToDevice, c.stream));
  HIP_CHECK(hipMemcpyAsync(c.sell_nrow.get(), nrow.data(), sizeof(int) * nslices,
                           hipMemcpyHostToDevice, c.stream));
  int* width = tmp_ints(c, kSlotCount, (size_t)nslices);
  hipLaunchKernelGGL(k_slice_width, dim3(grid_for(nslices, 256, 0)), dim3(256), 0, c.stream,
                     nslices, c.sell_row0.get(), c.sell_nrow.get(), c.brow_ptr.get(), width);
  HIP_CHECK(hipGetLastError());
  std::vector<int> w(nslices);
  HIP_CHECK(hipMemcpyAsync(w.data(), width, sizeof(int) * nslices, hipMemcpyDeviceToHost,
                           c.stream));
  HIP_CHECK(hipStreamSynchronize(c.stream));
  c.h_sell_off.assign(nslices + 1, 0);
  for (int s = 0; s < nslices; ++s) c.h_sell_off[s + 1] = c.h_sell_off[s] + (int64_t)w[s] * kSlice;
  c.max_row_blocks = nslices ? *std::max_element(w.begin(), w.end()) : 0;
  c.pattern_symmetric = true;  // node adjacency through triangles is symmetric
  c.nslices = nslices;
  c.sell_blocks = c.h_sell_off[nslices];
  c.sell_off.alloc(nslices + 1);
  HIP_CHECK(hipMemcpyAsync(c.sell_off.get(), c.h_sell_off.data(), sizeof(int64_t) * (nslices + 1),
                           hipMemcpyHostToDevice, c.stream));
  c.sell_col.alloc(c.sell_blocks);
  c.sell_val.alloc(4 * c.sell_blocks);
  c.sell_pos.alloc(c.nblk > 0 ? c.nblk : 1);
  HIP_CHECK(hipMemsetAsync(c.sell_val.get(), 0, sizeof(double) * 4 * c.sell_blocks, c.stream));
  const int wpb = 4;  // wavefronts per block
  hipLaunchKernelGGL(k_sell_fill, dim3(grid_for(nslices, wpb, 0)), dim3(wpb * kSlice), 0,
                     c.stream, nslices, c.sell_row0.get(), c.sell_nrow.get(), c.brow_ptr.get(),
                     c.bcol.get(), c.sell_off.get(), c.sell_col.get(), c.sell_pos.get());
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(c.stream));
  c.vals32_valid = false;
}

// ---------------------------------------------------------------------------------------
// element-to-slot map: for block (n, m) the elements holding both nodes, ascending

namespace {

__device__ inline int local_index(const int* t, int node) {
  return t[0] == node ? 0 : (t[1] == node ? 1 : 2);
}

constexpr int kSlotElems = 16;   // elements of a node held in registers by k_slot_map
constexpr int kSlotBlocks = 16;  // blocks of a row held in registers by k_slot_map

// For block (n, m) the elements holding both nodes, ascending: the count of each block of row
// n (kFill false), then the contribution codes e * 9 + a * 3 + b into their slots (kFill true).
// Node n's element list and those elements' three nodes are loaded once, together, into
// registers, and each block m takes n's elements that contain m, in their ascending order
// (the intersection of n's and m's element lists, without walking m's list).  A node with
// more than kSlotElems elements or a row of more than kSlotBlocks blocks takes the same
// result by merging the two ascending element lists.
template <bool kFill>
__global__ void k_slot_map(int nb, const int* __restrict__ tri, const int* __restrict__ ptr,
                           const int* __restrict__ list, const int* __restrict__ brow_ptr,
                           const int* __restrict__ bcol, int* __restrict__ cnt_or_ptr,
                           int* __restrict__ src) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nb) return;
  const int pn0 = ptr[n], pn1 = ptr[n + 1];
  const int q0 = brow_ptr[n], q1 = brow_ptr[n + 1];
  const int ne = pn1 - pn0, d = q1 - q0;
  if (ne <= kSlotElems && d <= kSlotBlocks) {
    int el[kSlotElems], bc[kSlotBlocks], o[kSlotBlocks];
#pragma unroll
    for (int i = 0; i < kSlotElems; ++i) el[i] = i < ne ? list[pn0 + i] : 0;
#pragma unroll
    for (int k = 0; k < kSlotBlocks; ++k) {
      bc[k] = k < d ? bcol[q0 + k] : -1;
      o[k] = kFill && k < d ? cnt_or_ptr[q0 + k] : 0;
    }
    int t0[kSlotElems], t1[kSlotElems], t2[kSlotElems];
#pragma unroll
    for (int i = 0; i < kSlotElems; ++i) {
      const int* t = tri + 3 * (int64_t)el[i];
      t0[i] = i < ne ? t[0] : -1;
      t1[i] = i < ne ? t[1] : -1;
      t2[i] = i < ne ? t[2] : -1;
    }
#pragma unroll
    for (int k = 0; k < kSlotBlocks; ++k) {
      if (k < d) {  // guards, not breaks: the loops unroll and every array stays in registers
        const int m = bc[k];
        int count = 0;
#pragma unroll
        for (int i = 0; i < kSlotElems; ++i) {
          const int b = t0[i] == m ? 0 : (t1[i] == m ? 1 : (t2[i] == m ? 2 : -1));
          if (i < ne && b >= 0) {
            if (kFill) {
              const int a = t0[i] == n ? 0 : (t1[i] == n ? 1 : 2);
              src[o[k] + count] = el[i] * 9 + a * 3 + b;
            }
            ++count;
          }
        }
        if (!kFill) cnt_or_ptr[q0 + k] = count;
      }
    }
    return;
  }
  for (int q = q0; q < q1; ++q) {
    const int m = bcol[q];
    int i = pn0, j = ptr[m];
    const int j1 = ptr[m + 1];
    int count = 0;
    int o = kFill ? cnt_or_ptr[q] : 0;
    while (i < pn1 && j < j1) {  // merge-intersect two ascending element lists
      const int ei = list[i], ej = list[j];
      if (ei < ej) {
        ++i;
      } else if (ej < ei) {
        ++j;
      } else {
        if (kFill) {
          const int* t = tri + 3 * (int64_t)ei;
          src[o++] = ei * 9 + local_index(t, n) * 3 + local_index(t, m);
        }
        ++count;
        ++i;
        ++j;
      }
    }
    if (!kFill) cnt_or_ptr[q] = count;
  }
}

}  // namespace

void build_slot_map(Context& c) {
  int* cnt = tmp_ints(c, kSlotCount, (size_t)c.nblk + 1);
  HIP_CHECK(hipMemsetAsync(cnt, 0, sizeof(int) * (c.nblk + 1), c.stream));
  const dim3 g(grid_for(c.nb, 128, 0)), b(128);
  hipLaunchKernelGGL(k_slot_map<false>, g, b, 0, c.stream, c.nb, c.tri.get(), c.ne_ptr.get(),
                     c.ne_list.get(), c.brow_ptr.get(), c.bcol.get(), cnt, (int*)nullptr);
  HIP_CHECK(hipGetLastError());
  c.slot_ptr.alloc(c.nblk + 1);
  int64_t total = 0;
  exclusive_scan_int(c, cnt, c.slot_ptr.get(), c.nblk + 1, &total);
  if (total != 9 * c.ne)
    throw Error(CSTFEM_ERR_MISSING_ENTRY,
                "assemble_stiffness: slot map holds " + std::to_string(total) +
                    " contributions, expected 9 per element (" + std::to_string(9 * c.ne) + ")");
  c.slot_src.alloc(total > 0 ? total : 1);
  hipLaunchKernelGGL(k_slot_map<true>, g, b, 0, c.stream, c.nb, c.tri.get(), c.ne_ptr.get(),
                     c.ne_list.get(), c.brow_ptr.get(), c.bcol.get(), c.slot_ptr.get(),
                     c.slot_src.get());
  HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------------------
// values: Ke sub-blocks summed per slot in element order

namespace {

// element.cpp:7-15
__device__ inline void constitutive(double E, double nu, int strain, double D[9]) {
  if (!strain) {
    const double c = E / (1.0 - nu * nu);
    D[0] = c; D[1] = c * nu; D[2] = 0.0;
    D[3] = c * nu; D[4] = c; D[5] = 0.0;
    D[6] = 0.0; D[7] = 0.0; D[8] = c * (1.0 - nu) / 2.0;
  } else {
    const double c = E / ((1.0 + nu) * (1.0 - 2.0 * nu));
    D[0] = c * (1.0 - nu); D[1] = c * nu; D[2] = 0.0;
    D[3] = c * nu; D[4] = c * (1.0 - nu); D[5] = 0.0;
    D[6] = 0.0; D[7] = 0.0; D[8] = c * (1.0 - 2.0 * nu) / 2.0;
  }
}

// Rows 2a..2a+1, columns 2b..2b+1 of cst_stiffness (element.cpp:21-55), computed as the
// reference computes the upper triangle (entry (i,j) with i > j is the mirrored (j,i)).
__device__ inline double pick3(int i, double v0, double v1, double v2) {
  return i == 0 ? v0 : (i == 1 ? v1 : v2);
}

// Column q (0..5) of the 3x6 strain-displacement matrix B of element.cpp:21-35, built with
// selects rather than an indexed local array, so that nothing is placed in scratch memory.
__device__ inline void b_column(int q, double bi, double ci, double inv2A, double col[3]) {
  const double vb = bi * inv2A, vc = ci * inv2A;
  const bool ux = (q & 1) == 0;
  col[0] = ux ? vb : 0.0;
  col[1] = ux ? 0.0 : vc;
  col[2] = ux ? vc : vb;
}

// Rows 2a..2a+1, columns 2b..2b+1 of cst_stiffness (element.cpp:21-55), computed as the
// reference computes the upper triangle (entry (i,j) with i > j is the mirrored (j,i)).
__device__ inline void ke_block(const double xy[6], const double D[9], double t, int a, int b,
                                double out[4]) {
  const double x1 = xy[0], y1 = xy[1], x2 = xy[2], y2 = xy[3], x3 = xy[4], y3 = xy[5];
  const double area = 0.5 * ((x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1));
  const double inv2A = 1.0 / (2.0 * area);
  const double b0 = y2 - y3, b1 = y3 - y1, b2 = y1 - y2;
  const double c0 = x3 - x2, c1 = x1 - x3, c2 = x2 - x1;
  const double scale = t * fabs(area);
#pragma unroll
  for (int ci = 0; ci < 2; ++ci)
#pragma unroll
    for (int cj = 0; cj < 2; ++cj) {
      const int i0 = 2 * a + ci, j0 = 2 * b + cj;
      const int p = i0 < j0 ? i0 : j0, q = i0 < j0 ? j0 : i0;
      double Bp[3], Bq[3];
      b_column(p, pick3(p >> 1, b0, b1, b2), pick3(p >> 1, c0, c1, c2), inv2A, Bp);
      b_column(q, pick3(q >> 1, b0, b1, b2), pick3(q >> 1, c0, c1, c2), inv2A, Bq);
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double db = 0.0;  // DB[k][q] = sum_m D[k][m] B[m][q]
#pragma unroll
        for (int m = 0; m < 3; ++m) db += D[k * 3 + m] * Bq[m];
        s += Bp[k] * db;
      }
      out[ci * 2 + cj] = scale * s;
    }
}

typedef double v2d __attribute__((ext_vector_type(2)));

// one wavefront per SELL slice, lane = block row; writes are coalesced
constexpr int kAsmGroup = 2;    // contributions whose loads k_assemble issues together
constexpr int kAsmBlocks = 16;  // blocks of a row whose slot ranges k_assemble keeps in registers

// The element data of one contribution e * 9 + a * 3 + b: its block's local indices, the
// element's coordinates and its modulus and Poisson ratio (D is formed at use)
struct Contribution {
  int la, lb;
  double exy[6];
  double E, nu;
};

// Values of the block row at `lane` of slice s: every 2x2 block sums its contributions in
// slot order (ascending element) from 0.0, as the reference's element loop adds them.  The
// row's contributions are consecutive in slot_src (block after block), so they are taken in
// groups of kAsmGroup across block boundaries: a group's codes, then its triangles, then
// their coordinates and moduli are loaded together (three memory round trips per group
// instead of three per contribution); the block ends come from registers.  Rows of more than
// kAsmBlocks blocks go block by block.
__global__ void k_assemble(const int* __restrict__ row0, const int* __restrict__ nrow,
                           int nslices, const int64_t* __restrict__ off,
                           const int* __restrict__ brow_ptr, const int* __restrict__ slot_ptr,
                           const int* __restrict__ slot_src, const int* __restrict__ tri,
                           const double* __restrict__ xy, const double* __restrict__ E_elem,
                           const double* __restrict__ nu_elem, Mat3 Dm, double nu, int strain, double t, double* __restrict__ val) {
  const int lane = threadIdx.x & (kSlice - 1);
  const int s = blockIdx.x * (blockDim.x / kSlice) + threadIdx.x / kSlice;
  if (s >= nslices) return;
  if (lane >= nrow[s]) return;
  const int r = row0[s] + lane;
  const int64_t o = off[s];
  const int b0 = brow_ptr[r], d = brow_ptr[r + 1] - b0;
  const double* Du = Dm.d;
  auto store = [&](int k, const double acc[4]) {
    // planar block (sell_vidx): the wavefront writes the 1 KB of row-0 pairs, then the
    // 1 KB of row-1 pairs, each as one coalesced 16-byte-per-lane store
    double* vk = val + 4 * (o + (int64_t)k * kSlice) + 2 * lane;
    *reinterpret_cast<v2d*>(vk) = v2d{acc[0], acc[1]};
    *reinterpret_cast<v2d*>(vk + 2 * kSlice) = v2d{acc[2], acc[3]};
  };
  auto load = [&](int code, Contribution& q) {
    const int e = code / 9, ab = code - 9 * e;
    q.la = ab / 3;
    q.lb = ab - 3 * q.la;
    const int n0 = tri[3 * e], n1 = tri[3 * e + 1], n2 = tri[3 * e + 2];
    q.exy[0] = xy[2 * n0];
    q.exy[1] = xy[2 * n0 + 1];
    q.exy[2] = xy[2 * n1];
    q.exy[3] = xy[2 * n1 + 1];
    q.exy[4] = xy[2 * n2];
    q.exy[5] = xy[2 * n2 + 1];
    q.E = E_elem ? E_elem[e] : 0.0;
    q.nu = nu_elem ? nu_elem[e] : nu;
  };
  // Ke block (la, lb) of a loaded contribution: D from its E and nu, or the uniform D
  auto block = [&](const Contribution& q, double kb[4]) {
    double De[9];
    if (E_elem) {
      constitutive(q.E, q.nu, strain, De);
    } else {
#pragma unroll
      for (int i = 0; i < 9; ++i) De[i] = Du[i];
    }
    ke_block(q.exy, De, t, q.la, q.lb, kb);
  };
  if (d <= kAsmBlocks) {
    int se[kAsmBlocks + 1];  // slot range of block k: [se[k], se[k + 1])
#pragma unroll
    for (int k = 0; k <= kAsmBlocks; ++k) se[k] = k <= d ? slot_ptr[b0 + k] : 0;
    const int c0 = se[0];
    int c1 = 0;
#pragma unroll
    for (int k = 0; k <= kAsmBlocks; ++k) c1 = k == d ? se[k] : c1;
    int k = 0;          // current block
    int end = se[1];    // its end
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int c = c0; c < c1; c += kAsmGroup) {
      int code[kAsmGroup];
#pragma unroll
      for (int u = 0; u < kAsmGroup; ++u) code[u] = c + u < c1 ? slot_src[c + u] : slot_src[c];
      Contribution q[kAsmGroup];
#pragma unroll
      for (int u = 0; u < kAsmGroup; ++u) load(code[u], q[u]);
#pragma unroll
      for (int u = 0; u < kAsmGroup; ++u) {
        if (c + u >= c1) break;
        while (c + u >= end) {  // the block is complete (a block holds at least one)
          store(k, acc);
          acc[0] = acc[1] = acc[2] = acc[3] = 0.0;
          ++k;
#pragma unroll
          for (int j = 1; j <= kAsmBlocks; ++j) end = j == k + 1 ? se[j] : end;
        }
        double kb[4];
        block(q[u], kb);
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i] += kb[i];
      }
    }
    if (d > 0) store(k, acc);
    return;
  }
  for (int k = 0; k < d; ++k) {
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    const int q = b0 + k;
    for (int c = slot_ptr[q]; c < slot_ptr[q + 1]; ++c) {
      Contribution cb;
      load(slot_src[c], cb);
      double kb[4];
      block(cb, kb);
#pragma unroll
      for (int i = 0; i < 4; ++i) acc[i] += kb[i];
    }
    store(k, acc);
  }
}

// cst_stiffness (element.cpp:37-55) of a batch of triangles: one thread per 2x2 block (a, b)
// of one element; Ke is row-major 6x6 per element
__global__ void k_element_stiffness(int64_t count, const double* __restrict__ xy, Mat3 Dm,
                                    double t, double* __restrict__ Ke) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= 9 * count) return;
  const int64_t e = g / 9;
  const int ab = (int)(g - 9 * e), a = ab / 3, b = ab - 3 * a;
  double exy[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) exy[i] = xy[6 * e + i];
  double kb[4];
  ke_block(exy, Dm.d, t, a, b, kb);
  double* o = Ke + 36 * e;
  o[(2 * a) * 6 + 2 * b] = kb[0];
  o[(2 * a) * 6 + 2 * b + 1] = kb[1];
  o[(2 * a + 1) * 6 + 2 * b] = kb[2];
  o[(2 * a + 1) * 6 + 2 * b + 1] = kb[3];
}

// SELL values -> the reference's CSR order
__global__ void k_export_csr(int nb, const int* __restrict__ brow_ptr,
                             const int64_t* __restrict__ spos, const double* __restrict__ val,
                             double* __restrict__ out) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nb) return;
  const int b0 = brow_ptr[n], d = brow_ptr[n + 1] - b0;
  const int64_t r0 = 4 * (int64_t)b0, r1 = r0 + 2 * d;
  for (int k = 0; k < d; ++k) {
    const int64_t q = spos[b0 + k];
    out[r0 + 2 * k] = val[sell_vidx(q, 0)];
    out[r0 + 2 * k + 1] = val[sell_vidx(q, 1)];
    out[r1 + 2 * k] = val[sell_vidx(q, 2)];
    out[r1 + 2 * k + 1] = val[sell_vidx(q, 3)];
  }
}

}  // namespace

void assemble_values(Context& c) {
  if (!c.have_pattern || !c.mesh_pattern)
    throw Error(CSTFEM_ERR_ORDER, "assemble_stiffness: build the pattern of this mesh first");
  const int wpb = 4;
  HIP_CHECK(hipMemsetAsync(c.sell_val.get(), 0, sizeof(double) * 4 * c.sell_blocks, c.stream));
  hipLaunchKernelGGL(k_assemble, dim3(grid_for(c.nslices, wpb, 0)), dim3(wpb * kSlice), 0,
                     c.stream, c.sell_row0.get(), c.sell_nrow.get(), c.nslices, c.sell_off.get(),
                     c.brow_ptr.get(), c.slot_ptr.get(), c.slot_src.get(), c.tri.get(), c.xy.get(),
                     c.has_E_elem ? c.E_elem.get() : (const double*)nullptr,
                     c.has_nu_elem ? c.nu_elem.get() : (const double*)nullptr, c.D, c.nu,
                     c.plane_strain, c.thick, c.sell_val.get());
  HIP_CHECK(hipGetLastError());
  c.setup_done = false;
  c.vals32_valid = false;
  c.fixed_valid = false;  // a fresh matrix has no constrained DOFs
  c.mpnn.built = false;
  c.mpnn.npairs = 0;  // training pairs belong to the matrix they were formed with
}

void element_stiffness(Context& c, int64_t count, const double* xy, const double* D, double t,
                       double* Ke) {
  if (count == 0) return;
  Mat3 Dm;
  for (int i = 0; i < 9; ++i) Dm.d[i] = D[i];
  // staging in the context's buffers, grown on demand and kept
  if (c.ke_io.n < (size_t)(42 * count)) c.ke_io.alloc(42 * (size_t)count);
  double* d_xy = c.ke_io.get();
  double* d_ke = d_xy + 6 * count;
  HIP_CHECK(hipMemcpyAsync(d_xy, xy, sizeof(double) * 6 * count, hipMemcpyHostToDevice, c.stream));
  hipLaunchKernelGGL(k_element_stiffness, dim3(grid_for(9 * count, 256, 0)), dim3(256), 0,
                     c.stream, count, (const double*)d_xy, Dm, t, d_ke);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(Ke, d_ke, sizeof(double) * 36 * count, hipMemcpyDeviceToHost, c.stream));
  HIP_CHECK(hipStreamSynchronize(c.stream));
}

void download_values(Context& c, double* vals) {
  if (!c.have_pattern) throw Error(CSTFEM_ERR_ORDER, "no matrix on the device");
  if (c.generic) {
    std::vector<double> all(4 * c.sell_blocks);
    HIP_CHECK(hipMemcpyAsync(all.data(), c.sell_val.get(), sizeof(double) * all.size(),
                             hipMemcpyDeviceToHost, c.stream));
    HIP_CHECK(hipStreamSynchronize(c.stream));
    for (size_t k = 0; k < c.h_generic_map.size(); ++k) vals[k] = all[c.h_generic_map[k]];
    return;
  }
  DBuf<double> out;
  out.alloc(4 * c.nblk);
  hipLaunchKernelGGL(k_export_csr, dim3(grid_for(c.nb, 256, 0)), dim3(256), 0, c.stream, c.nb,
                     c.brow_ptr.get(), c.sell_pos.get(), c.sell_val.get(), out.get());
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(vals, out.get(), sizeof(double) * 4 * c.nblk, hipMemcpyDeviceToHost,
                           c.stream));
  HIP_CHECK(hipStreamSynchronize(c.stream));
}

}  // namespace cstfem_hip
