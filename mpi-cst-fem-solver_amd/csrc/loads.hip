// loads.hip -- consistent nodal loads on the device: boundary tractions and the body force.
//
// Traction (reference: src/bc.cpp:190-253, add_traction).  For one line group of E edges:
//   k_trac_edges   one thread per edge: the triangles holding both nodes (through the
//                  device node -> elements map, ascending element order, as the reference's
//                  scan over all triangles meets them), the opposite node of the last one, the
//                  edge length (hypot_libm: the C library's hypot, as the reference calls it),
//                  the outward normal and the two Gauss points;
//   (tractions at the Gauss points: a constant, or a host function of (x, y, nx, ny) -- the
//    reference's TractionFn -- evaluated on the 2E points the device computed)
//   k_trac_sum     one thread per touched node adds its contributions to f in the reference's
//                  order: edges ascending, Gauss points -g then +g.
// Every floating-point operation is the reference's, in its order, so f equals the
// reference's bit for bit.  An edge adjacent to other than one triangle leaves f unchanged
// and raises the reference's message.
#include <algorithm>
#include <climits>
#include <cmath>
#include <unordered_map>

#define CSTFEM_HD __host__ __device__
#include "hypot_libm.h"
#include "internal.hpp"

namespace cstfem_hip {

namespace {

constexpr int kGeo = 7;  // per edge: len, nx, ny, x(-g), y(-g), x(+g), y(+g)

__global__ void k_trac_edges(int nedge, const int* __restrict__ edges,
                             const int* __restrict__ ne_ptr, const int* __restrict__ ne_list,
                             const int* __restrict__ tri, const double* __restrict__ xy, double g,
                             double* __restrict__ geo, int* __restrict__ count,
                             int* __restrict__ err) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nedge) return;
  const int a = edges[2 * k], b = edges[2 * k + 1];
  int cnt = 0, opp = -1;
  if (a != b) {
    for (int q = ne_ptr[a]; q < ne_ptr[a + 1]; ++q) {
      const int* nd = tri + 3 * (int64_t)ne_list[q];
      const int n0 = nd[0], n1 = nd[1], n2 = nd[2];
      if (n0 != b && n1 != b && n2 != b) continue;
      ++cnt;
      opp = (n2 != a && n2 != b) ? n2 : (n1 != a && n1 != b) ? n1 : n0;
    }
  }
  count[k] = cnt;
  if (cnt != 1) {
    atomicMin(err, k);
    return;
  }
  const double xa = xy[2 * a], ya = xy[2 * a + 1];
  const double xb = xy[2 * b], yb = xy[2 * b + 1];
  const double len = hypot_libm(xb - xa, yb - ya);
  double nx = (yb - ya) / len, ny = -(xb - xa) / len;
  if (nx * (xy[2 * opp] - xa) + ny * (xy[2 * opp + 1] - ya) > 0.0) {
    nx = -nx;
    ny = -ny;
  }
  double* o = geo + kGeo * (int64_t)k;
  o[0] = len;
  o[1] = nx;
  o[2] = ny;
  for (int qi = 0; qi < 2; ++qi) {
    const double q = qi ? g : -g;
    const double Na = 0.5 * (1.0 - q), Nb = 0.5 * (1.0 + q);
    o[3 + 2 * qi] = Na * xa + Nb * xb;
    o[4 + 2 * qi] = Na * ya + Nb * yb;
  }
}

// tv == NULL: the constant traction (t0, t1); else tv[4k + 2qi + c] at Gauss point qi of edge k
__global__ void k_trac_sum(int nnodes, const int* __restrict__ node, const int* __restrict__ ptr,
                           const int* __restrict__ ent, const double* __restrict__ geo,
                           const double* __restrict__ tv, double t0, double t1, double thick,
                           double g, const int* __restrict__ err, double* __restrict__ f) {
  if (err && *err != INT_MAX) return;  // a bad edge: f stays as it was (err NULL: checked)
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= nnodes) return;
  const int v = node[u];
  double fx = f[2 * v], fy = f[2 * v + 1];
  for (int e = ptr[u]; e < ptr[u + 1]; ++e) {
    const int k = ent[e] >> 1, role = ent[e] & 1;  // role 0: first node of the edge
    const double w = thick * 0.5 * geo[kGeo * (int64_t)k];
    for (int qi = 0; qi < 2; ++qi) {
      const double q = qi ? g : -g;
      const double N = role ? 0.5 * (1.0 + q) : 0.5 * (1.0 - q);
      const double tx = tv ? tv[4 * (int64_t)k + 2 * qi] : t0;
      const double ty = tv ? tv[4 * (int64_t)k + 2 * qi + 1] : t1;
      fx += w * N * tx;
      fy += w * N * ty;
    }
  }
  f[2 * v] = fx;
  f[2 * v + 1] = fy;
}

// The body force t*|A|/3*(bx, by) on each node of every element, per node in ascending
// element order, in two passes: the weight t*|A|/3 of every element (coalesced over the
// elements), then per node the sum over its element list (the same products in the same
// order as one pass per node, which would walk a chain list -> triangle -> coordinates for
// each of its elements)
__global__ void k_body_weight(int64_t ne, const int* __restrict__ tri, const double* __restrict__ xy,
                              double t, double* __restrict__ w) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int n0 = tri[3 * e], n1 = tri[3 * e + 1], n2 = tri[3 * e + 2];
  const double x1 = xy[2 * n0], y1 = xy[2 * n0 + 1], x2 = xy[2 * n1], y2 = xy[2 * n1 + 1];
  const double x3 = xy[2 * n2], y3 = xy[2 * n2 + 1];
  const double area = 0.5 * ((x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1));
  w[e] = t * fabs(area) / 3.0;
}

__global__ void k_body_force(int nn, const int* __restrict__ ptr, const int* __restrict__ list,
                             const double* __restrict__ w, double bx, double by,
                             double* __restrict__ f) {
  const int nd = blockIdx.x * blockDim.x + threadIdx.x;
  if (nd >= nn) return;
  const int k0 = ptr[nd], k1 = ptr[nd + 1];
  double fx = f[2 * nd], fy = f[2 * nd + 1];
  for (int k = k0; k < k1; k += 4) {
    double we[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) we[u] = k + u < k1 ? w[list[k + u]] : 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (k + u < k1) {
        fx += we[u] * bx;
        fy += we[u] * by;
      }
  }
  f[2 * nd] = fx;
  f[2 * nd + 1] = fy;
}

template <class T>
T* grow(DBuf<T>& b, size_t count) {
  if (!b.get() || b.n < count) b.alloc(std::max<size_t>(count, 1));
  return b.get();
}

}  // namespace

// kirsch.cpp:7-45: the Kirsch stress at (x, y) and the traction sigma.n
void kirsch_traction(double sigma_inf, double a, double x, double y, double nx, double ny,
                     double t[2]) {
  const double r = std::hypot(x, y), theta = std::atan2(y, x);
  const double q2 = (a * a) / (r * r), q4 = q2 * q2;
  const double c2 = std::cos(2.0 * theta), s2 = std::sin(2.0 * theta);
  const double s = sigma_inf;
  const double rr = 0.5 * s * (1.0 - q2) + 0.5 * s * (1.0 - 4.0 * q2 + 3.0 * q4) * c2;
  const double tt = 0.5 * s * (1.0 + q2) - 0.5 * s * (1.0 + 3.0 * q4) * c2;
  const double rt = -0.5 * s * (1.0 + 2.0 * q2 - 3.0 * q4) * s2;
  const double c = std::cos(theta), n = std::sin(theta);
  const double sxx = rr * c * c + tt * n * n - 2.0 * rt * n * c;
  const double syy = rr * n * n + tt * c * c + 2.0 * rt * n * c;
  const double txy = (rr - tt) * n * c + rt * (c * c - n * n);
  t[0] = sxx * nx + txy * ny;
  t[1] = txy * nx + syy * ny;
}

void add_traction_device(Context& c, int nedge, const int* edges, const TractionSource& src) {
  if (c.b.n == 0) throw Error(CSTFEM_ERR_ORDER, "add_traction: set the right-hand side first");
  if (c.ne_ptr.n == 0) build_node_elements(c);
  for (int k = 0; k < 2 * nedge; ++k)
    if (edges[k] < 0 || edges[k] >= c.nn)
      throw Error(CSTFEM_ERR_ARGUMENT, "cstfem_add_traction: edge node out of range");
  if (nedge == 0) return;
  const bool host_fn = src.kind != TractionSource::kConstant;
  const double g = 1.0 / std::sqrt(3.0);
  if (!host_fn) {
    // a constant traction on a group applied before on this mesh: its index lists are on the
    // device and its edges passed the check, so the two kernels alone, with nothing to wait for
    for (const auto& grp : c.trac_groups) {
      if (grp->edges.size() != 2 * (size_t)nedge ||
          !std::equal(grp->edges.begin(), grp->edges.end(), edges))
        continue;
      const int* di = grp->idx.get();
      const int U = grp->U;
      const size_t o_cnt = 2 * (size_t)nedge, o_node = o_cnt + nedge, o_ptr = o_node + U,
                   o_ent = o_ptr + U + 1;
      double* geo = grow(c.trac_d, (size_t)nedge * kGeo);
      hipLaunchKernelGGL(k_trac_edges, dim3(grid_for(nedge, 128, 0)), dim3(128), 0, c.stream, nedge,
                         di, c.ne_ptr.get(), c.ne_list.get(), c.tri.get(), c.xy.get(), g, geo,
                         grp->idx.get() + o_cnt, c.err_flag.get());
      hipLaunchKernelGGL(k_trac_sum, dim3(grid_for(U, 128, 0)), dim3(128), 0, c.stream, U,
                         di + o_node, di + o_ptr, di + o_ent, (const double*)geo,
                         (const double*)nullptr, src.p0, src.p1, c.thick, g, (const int*)nullptr,
                         c.b.get());
      HIP_CHECK(hipGetLastError());
      return;
    }
  }
  // touched nodes in first-touch order, each with its (edge, role) entries in edge order
  std::unordered_map<int, int> slot;
  slot.reserve(2 * (size_t)nedge);
  std::vector<int> node, cnt;
  for (int k = 0; k < 2 * nedge; ++k) {
    auto it = slot.emplace(edges[k], (int)node.size());
    if (it.second) {
      node.push_back(edges[k]);
      cnt.push_back(0);
    }
    ++cnt[it.first->second];
  }
  const int U = (int)node.size();
  // packed index data: edges[2E] | count[E] | node[U] | ptr[U+1] | ent[2E]
  const size_t o_cnt = 2 * (size_t)nedge, o_node = o_cnt + nedge, o_ptr = o_node + U,
               o_ent = o_ptr + U + 1, total = o_ent + 2 * (size_t)nedge;
  // host side in the context's page-locked staging: the function traction's Gauss points and
  // values (doubles first), then the packed index data and two flag words
  const size_t nd_h = host_fn ? (size_t)(kGeo + 4) * nedge : 0;
  unsigned char* stage = c.stage(sizeof(double) * nd_h + sizeof(int) * (total + 2));
  double* h_geo = reinterpret_cast<double*>(stage);
  double* h_t = h_geo + (host_fn ? (size_t)kGeo * nedge : 0);
  int* pack = reinterpret_cast<int*>(stage + sizeof(double) * nd_h);
  int* h_flag = pack + total;  // [0] = INT_MAX (upload), [1] = the flag read back
  std::fill(pack, pack + total, 0);
  std::copy(edges, edges + 2 * (size_t)nedge, pack);
  std::copy(node.begin(), node.end(), pack + o_node);
  int* ptr = pack + o_ptr;
  ptr[0] = 0;
  for (int u = 0; u < U; ++u) ptr[u + 1] = ptr[u] + cnt[u];
  std::vector<int> fill(ptr, ptr + U);
  for (int k = 0; k < 2 * nedge; ++k) pack[o_ent + fill[slot[edges[k]]]++] = k;  // 2k + role
  // a constant traction's group keeps its index lists for later calls (once its check passed)
  std::unique_ptr<Context::TractionGroup> keep;
  int* di;
  if (host_fn) {
    di = grow(c.trac_i, total);
  } else {
    keep = std::make_unique<Context::TractionGroup>();
    keep->idx.alloc(total);
    keep->U = U;
    di = keep->idx.get();
  }
  double* dd = grow(c.trac_d, (size_t)nedge * (kGeo + (host_fn ? 4 : 0)));
  double* geo = dd;
  double* tv = host_fn ? dd + (size_t)kGeo * nedge : nullptr;
  int* err = c.err_flag.get();
  const int big = INT_MAX;
  h_flag[0] = big;
  HIP_CHECK(hipMemcpyAsync(di, pack, sizeof(int) * total, hipMemcpyHostToDevice, c.stream));
  HIP_CHECK(hipMemcpyAsync(err, h_flag, sizeof(int), hipMemcpyHostToDevice, c.stream));
  hipLaunchKernelGGL(k_trac_edges, dim3(grid_for(nedge, 128, 0)), dim3(128), 0, c.stream, nedge,
                     di, c.ne_ptr.get(), c.ne_list.get(), c.tri.get(), c.xy.get(), g, geo,
                     di + o_cnt, err);
  HIP_CHECK(hipGetLastError());
  auto check_edges = [&](int bad) {
    if (bad == INT_MAX) return;
    int n_adj = 0;
    HIP_CHECK(hipMemcpyAsync(&n_adj, di + o_cnt + bad, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    HIP_CHECK(hipStreamSynchronize(c.stream));
    throw Error(CSTFEM_ERR_TRACTION_EDGE,
                "traction group: edge (" + std::to_string(edges[2 * bad]) + ", " +
                    std::to_string(edges[2 * bad + 1]) + ") is adjacent to " +
                    std::to_string(n_adj) + " triangles (expected 1)");
  };
  if (host_fn) {
    // the traction function runs on the host at the Gauss points the device computed
    std::vector<double> pts(8 * (size_t)nedge);
    double* t = h_t;
    HIP_CHECK(hipMemcpyAsync(h_geo, geo, sizeof(double) * kGeo * nedge, hipMemcpyDeviceToHost, c.stream));
    HIP_CHECK(hipMemcpyAsync(h_flag + 1, err, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    HIP_CHECK(hipStreamSynchronize(c.stream));
    check_edges(h_flag[1]);
    for (int k = 0; k < nedge; ++k) {
      const double* o = &h_geo[(size_t)kGeo * k];
      for (int qi = 0; qi < 2; ++qi) {
        double* p = &pts[8 * (size_t)k + 4 * qi];
        p[0] = o[3 + 2 * qi];
        p[1] = o[4 + 2 * qi];
        p[2] = o[1];
        p[3] = o[2];
      }
    }
    if (src.kind == TractionSource::kKirsch) {
      for (int j = 0; j < 2 * nedge; ++j)
        kirsch_traction(src.p0, src.p1, pts[4 * j], pts[4 * j + 1], pts[4 * j + 2],
                        pts[4 * j + 3], &t[2 * j]);
    } else {
      src.fn(src.user, 2 * nedge, pts.data(), t);
    }
    HIP_CHECK(hipMemcpyAsync(tv, t, sizeof(double) * 4 * nedge, hipMemcpyHostToDevice, c.stream));
  }
  hipLaunchKernelGGL(k_trac_sum, dim3(grid_for(U, 128, 0)), dim3(128), 0, c.stream, U,
                     di + o_node, di + o_ptr, di + o_ent, geo, (const double*)tv, src.p0, src.p1,
                     c.thick, g, err, c.b.get());
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(h_flag + 1, err, sizeof(int), hipMemcpyDeviceToHost, c.stream));
  HIP_CHECK(hipStreamSynchronize(c.stream));
  check_edges(h_flag[1]);
  if (keep) {
    keep->edges.assign(edges, edges + 2 * (size_t)nedge);
    if (c.trac_groups.size() >= Context::kTractionGroups) c.trac_groups.erase(c.trac_groups.begin());
    c.trac_groups.push_back(std::move(keep));
  }
}

void add_body_force_device(Context& c, double bx, double by) {
  if (c.ne_ptr.n == 0) build_node_elements(c);
  if (c.b.n == 0) throw Error(CSTFEM_ERR_ORDER, "add_body_force: set the right-hand side first");
  double* w = grow(c.body_w, (size_t)std::max<int64_t>(c.ne, 1));
  if (c.ne > 0)
    hipLaunchKernelGGL(k_body_weight, dim3(grid_for(c.ne, 256, 0)), dim3(256), 0, c.stream, c.ne,
                       c.tri.get(), c.xy.get(), c.thick, w);
  hipLaunchKernelGGL(k_body_force, dim3(grid_for(c.nn, 256, 0)), dim3(256), 0, c.stream, c.nn,
                     c.ne_ptr.get(), c.ne_list.get(), (const double*)w, bx, by, c.b.get());
  HIP_CHECK(hipGetLastError());  // nothing left to report: no wait for the kernel
}

}  // namespace cstfem_hip
