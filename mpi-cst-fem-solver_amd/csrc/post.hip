// post.hip -- stress recovery and Kirsch error metrics on the device.
//
// Reference: src/postprocess.cpp:35-60 (compute_element_stresses), :81-150 (kirsch_partial,
// set_relative_errors, set_scf), src/element.cpp:57-81 (cst_strain, stress_from_strain,
// sigma_zz, von_mises), src/kirsch.cpp:7-59 (analytic Kirsch field).  One thread per element;
// per-element stresses follow the reference's operation order (bit-identical); the metric sums
// are reduced per workgroup and added on the host in workgroup order, so they differ from the
// reference's sequential sums only by rounding.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "internal.hpp"

namespace cstfem_hip {

namespace {

constexpr int kT = 256;

__device__ inline void constitutive_d(double E, double nu, int strain, double D[9]) {
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

__device__ inline double signed_area(const double xy[6]) {
  return 0.5 * ((xy[2] - xy[0]) * (xy[5] - xy[1]) - (xy[4] - xy[0]) * (xy[3] - xy[1]));
}

// element.cpp:57-69 on the element's coordinates and displacements
__device__ inline void element_stress(const double xy[6], const double ue[6], const double D[9],
                                      double sig[3]) {
  const double inv2A = 1.0 / (2.0 * signed_area(xy));
  const double b[3] = {xy[3] - xy[5], xy[5] - xy[1], xy[1] - xy[3]};
  const double c[3] = {xy[4] - xy[2], xy[0] - xy[4], xy[2] - xy[0]};
  double B[18];
  for (int i = 0; i < 18; ++i) B[i] = 0.0;
  for (int i = 0; i < 3; ++i) {
    B[0 * 6 + 2 * i] = b[i] * inv2A;
    B[1 * 6 + 2 * i + 1] = c[i] * inv2A;
    B[2 * 6 + 2 * i] = c[i] * inv2A;
    B[2 * 6 + 2 * i + 1] = b[i] * inv2A;
  }
  double e[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 6; ++j) e[i] += B[i * 6 + j] * ue[j];
  for (int i = 0; i < 3; ++i) {
    double s = 0.0;
    for (int j = 0; j < 3; ++j) s += D[i * 3 + j] * e[j];
    sig[i] = s;
  }
}

__device__ inline void gather(const int* __restrict__ tri, const double* __restrict__ xy,
                              const double* __restrict__ u, int64_t e, double exy[6], double ue[6],
                              int n[3]) {
  for (int k = 0; k < 3; ++k) {
    n[k] = tri[3 * e + k];
    exy[2 * k] = xy[2 * n[k]];
    exy[2 * k + 1] = xy[2 * n[k] + 1];
    ue[2 * k] = u[2 * n[k]];
    ue[2 * k + 1] = u[2 * n[k] + 1];
  }
}

__global__ __launch_bounds__(kT) void k_stress(int64_t ne, const int* __restrict__ tri,
                                               const double* __restrict__ xy,
                                               const double* __restrict__ u,
                                               const double* __restrict__ E_elem,
                                               const double* __restrict__ nu_elem, Mat3 Dm,
                                               double nu, int strain, double* __restrict__ sig,
                                               double* __restrict__ vm, double* __restrict__ szz,
                                               double* __restrict__ cxy,
                                               double* __restrict__ area) {
  const int64_t e = blockIdx.x * (int64_t)kT + threadIdx.x;
  if (e >= ne) return;
  double exy[6], ue[6], D[9], s[3];
  int n[3];
  gather(tri, xy, u, e, exy, ue, n);
  double nue = nu;
  if (E_elem) {
    if (nu_elem) nue = nu_elem[e];
    constitutive_d(E_elem[e], nue, strain, D);
  } else {
    for (int k = 0; k < 9; ++k) D[k] = Dm.d[k];
  }
  element_stress(exy, ue, D, s);
  sig[3 * e] = s[0];
  sig[3 * e + 1] = s[1];
  sig[3 * e + 2] = s[2];
  // element.cpp:72-81
  const double sz = strain ? nue * (s[0] + s[1]) : 0.0;
  if (szz) szz[e] = sz;
  if (vm) {
    const double j2 = (s[0] - s[1]) * (s[0] - s[1]) + (s[1] - sz) * (s[1] - sz) + (sz - s[0]) * (sz - s[0]);
    vm[e] = sqrt(0.5 * j2 + 3.0 * s[2] * s[2]);
  }
  if (cxy) {
    cxy[2 * e] = (exy[0] + exy[2] + exy[4]) / 3.0;
    cxy[2 * e + 1] = (exy[1] + exy[3] + exy[5]) / 3.0;
  }
  if (area) area[e] = fabs(signed_area(exy));
}

// kirsch.cpp:7-45: Cartesian Kirsch stress at (x, y)
__device__ inline void kirsch_stress(double s_inf, double a, double x, double y, double out[3]) {
  const double r = hypot(x, y), th = atan2(y, x);
  const double q2 = (a * a) / (r * r), q4 = q2 * q2;
  const double c2 = cos(2.0 * th), s2 = sin(2.0 * th);
  const double rr = 0.5 * s_inf * (1.0 - q2) + 0.5 * s_inf * (1.0 - 4.0 * q2 + 3.0 * q4) * c2;
  const double tt = 0.5 * s_inf * (1.0 + q2) - 0.5 * s_inf * (1.0 + 3.0 * q4) * c2;
  const double rt = -0.5 * s_inf * (1.0 + 2.0 * q2 - 3.0 * q4) * s2;
  const double c = cos(th), n = sin(th);
  out[0] = rr * c * c + tt * n * n - 2.0 * rt * n * c;
  out[1] = rr * n * n + tt * c * c + 2.0 * rt * n * c;
  out[2] = (rr - tt) * n * c + rt * (c * c - n * n);
}

// kirsch.cpp:47-59: Cartesian Kirsch displacement at (x, y)
__device__ inline void kirsch_disp(double s_inf, double a, double E, double nu, int strain,
                                   double x, double y, double out[2]) {
  const double mu = E / (2.0 * (1.0 + nu));
  const double kappa = !strain ? (3.0 - nu) / (1.0 + nu) : 3.0 - 4.0 * nu;
  const double r = hypot(x, y), th = atan2(y, x);
  const double a2 = a * a, a4 = a2 * a2, r3 = r * r * r;
  const double ur = s_inf / (8.0 * mu) *
                    ((kappa - 1.0) * r + 2.0 * a2 / r + 2.0 * (r + (kappa + 1.0) * a2 / r - a4 / r3) * cos(2.0 * th));
  const double ut = -s_inf / (4.0 * mu) * (r + (kappa - 1.0) * a2 / r + a4 / r3) * sin(2.0 * th);
  const double c = cos(th), s = sin(th);
  out[0] = ur * c - ut * s;
  out[1] = ur * s + ut * c;
}

// per-workgroup sums of the six error integrals and the element nearest to (0, a)
__global__ __launch_bounds__(kT) void k_kirsch(int64_t ne, const int* __restrict__ tri,
                                               const double* __restrict__ xy,
                                               const double* __restrict__ u,
                                               const double* __restrict__ sig,
                                               const double* __restrict__ cxy,
                                               const double* __restrict__ area, double s_inf,
                                               double a, double E, double nu, int strain,
                                               Mat3 Dm, double* __restrict__ part,
                                               double* __restrict__ best_d,
                                               long long* __restrict__ best_e) {
  __shared__ double lds[6][kT];
  __shared__ double ld[kT];
  __shared__ long long le[kT];
  const double det = Dm.d[0] * Dm.d[4] - Dm.d[1] * Dm.d[3];
  const double C00 = Dm.d[4] / det, C01 = -Dm.d[1] / det, C11 = Dm.d[0] / det, C22 = 1.0 / Dm.d[8];
  const double bary[3][3] = {{2.0 / 3, 1.0 / 6, 1.0 / 6}, {1.0 / 6, 2.0 / 3, 1.0 / 6},
                             {1.0 / 6, 1.0 / 6, 2.0 / 3}};
  double acc[6] = {0, 0, 0, 0, 0, 0};
  double bd = DBL_MAX;
  long long be = -1;
  for (int64_t e = blockIdx.x * (int64_t)kT + threadIdx.x; e < ne; e += (int64_t)gridDim.x * kT) {
    const double sx = sig[3 * e], sy = sig[3 * e + 1], txy = sig[3 * e + 2];
    double ex[3];
    kirsch_stress(s_inf, a, cxy[2 * e], cxy[2 * e + 1], ex);
    const double d0 = sx - ex[0], d1 = sy - ex[1], d2 = txy - ex[2];
    acc[0] += area[e] * (d0 * d0 + d1 * d1 + 2 * d2 * d2);
    acc[1] += area[e] * (ex[0] * ex[0] + ex[1] * ex[1] + 2 * ex[2] * ex[2]);
    const double d = hypot(cxy[2 * e], cxy[2 * e + 1] - a);
    if (d < bd) {  // ascending e per thread: the first of equal distances wins
      bd = d;
      be = e;
    }
    const int* n = tri + 3 * e;
    const double w = area[e] / 3.0;
    for (int q = 0; q < 3; ++q) {
      double x = 0, y = 0, uh = 0, vh = 0;
      for (int i = 0; i < 3; ++i) {
        x += bary[q][i] * xy[2 * n[i]];
        y += bary[q][i] * xy[2 * n[i] + 1];
        uh += bary[q][i] * u[2 * n[i]];
        vh += bary[q][i] * u[2 * n[i] + 1];
      }
      double ua[2], sq[3];
      kirsch_disp(s_inf, a, E, nu, strain, x, y, ua);
      acc[2] += w * ((uh - ua[0]) * (uh - ua[0]) + (vh - ua[1]) * (vh - ua[1]));
      acc[3] += w * (ua[0] * ua[0] + ua[1] * ua[1]);
      kirsch_stress(s_inf, a, x, y, sq);
      const double v0 = sx - sq[0], v1 = sy - sq[1], v2 = txy - sq[2];
      acc[4] += w * (C00 * v0 * v0 + 2.0 * C01 * v0 * v1 + C11 * v1 * v1 + C22 * v2 * v2);
      acc[5] += w * (C00 * sq[0] * sq[0] + 2.0 * C01 * sq[0] * sq[1] + C11 * sq[1] * sq[1] +
                     C22 * sq[2] * sq[2]);
    }
  }
  for (int k = 0; k < 6; ++k) lds[k][threadIdx.x] = acc[k];
  ld[threadIdx.x] = bd;
  le[threadIdx.x] = be;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s[6] = {0, 0, 0, 0, 0, 0};
    double d = DBL_MAX;
    long long eb = -1;
    for (int t = 0; t < kT; ++t) {
      for (int k = 0; k < 6; ++k) s[k] += lds[k][t];
      if (le[t] >= 0 && (ld[t] < d || (ld[t] == d && le[t] < eb))) {
        d = ld[t];
        eb = le[t];
      }
    }
    for (int k = 0; k < 6; ++k) part[6 * blockIdx.x + k] = s[k];
    best_d[blockIdx.x] = d;
    best_e[blockIdx.x] = eb;
  }
}

void stresses_device(Context& c, DBuf<double>& sig, DBuf<double>* vm, DBuf<double>* szz,
                     DBuf<double>& cxy, DBuf<double>& area) {
  if (!c.setup_done || c.x.n < (size_t)c.n || !c.mesh_pattern)
    throw Error(CSTFEM_ERR_ORDER, "element_stresses: needs a solved mesh system");
  const int64_t ne = c.ne;
  sig.alloc(3 * ne);
  cxy.alloc(2 * ne);
  area.alloc(ne);
  if (vm) vm->alloc(ne);
  if (szz) szz->alloc(ne);
  hipLaunchKernelGGL(k_stress, dim3((unsigned)((ne + kT - 1) / kT)), dim3(kT), 0, c.stream, ne,
                     c.tri.get(), c.xy.get(), c.x.get(),
                     c.has_E_elem ? c.E_elem.get() : (const double*)nullptr,
                     c.has_nu_elem ? c.nu_elem.get() : (const double*)nullptr, c.D, c.nu,
                     c.plane_strain, sig.get(), vm ? vm->get() : nullptr,
                     szz ? szz->get() : nullptr, cxy.get(), area.get());
  HIP_CHECK(hipGetLastError());
}

}  // namespace

void element_stresses(Context& c, double* sig, double* vm, double* szz, double* cxy,
                      double* area) {
  DBuf<double> d_sig, d_vm, d_szz, d_cxy, d_area;
  stresses_device(c, d_sig, vm ? &d_vm : nullptr, szz ? &d_szz : nullptr, d_cxy, d_area);
  const int64_t ne = c.ne;
  HIP_CHECK(hipMemcpyAsync(sig, d_sig.get(), sizeof(double) * 3 * ne, hipMemcpyDeviceToHost, c.stream));
  if (vm) HIP_CHECK(hipMemcpyAsync(vm, d_vm.get(), sizeof(double) * ne, hipMemcpyDeviceToHost, c.stream));
  if (szz) HIP_CHECK(hipMemcpyAsync(szz, d_szz.get(), sizeof(double) * ne, hipMemcpyDeviceToHost, c.stream));
  if (cxy) HIP_CHECK(hipMemcpyAsync(cxy, d_cxy.get(), sizeof(double) * 2 * ne, hipMemcpyDeviceToHost, c.stream));
  if (area) HIP_CHECK(hipMemcpyAsync(area, d_area.get(), sizeof(double) * ne, hipMemcpyDeviceToHost, c.stream));
  HIP_CHECK(hipStreamSynchronize(c.stream));
}

void kirsch_metrics(Context& c, double sigma_inf, double a, cstfem_kirsch_result* out) {
  if (c.has_E_elem) throw Error(CSTFEM_ERR_ARGUMENT, "kirsch_metrics: needs a uniform modulus");
  if (!(a > 0.0)) throw Error(CSTFEM_ERR_ARGUMENT, "kirsch_metrics: a must be positive");
  DBuf<double> sig, cxy, area;
  stresses_device(c, sig, nullptr, nullptr, cxy, area);
  const int64_t ne = c.ne;
  const int g = (int)std::min<int64_t>(1024, std::max<int64_t>(1, (ne + kT - 1) / kT));
  DBuf<double> part, bd;
  DBuf<long long> be;
  part.alloc(6 * (size_t)g);
  bd.alloc(g);
  be.alloc(g);
  hipLaunchKernelGGL(k_kirsch, dim3(g), dim3(kT), 0, c.stream, ne, c.tri.get(), c.xy.get(),
                     c.x.get(), (const double*)sig.get(), (const double*)cxy.get(),
                     (const double*)area.get(), sigma_inf, a, c.E, c.nu, c.plane_strain, c.D,
                     part.get(), bd.get(), be.get());
  HIP_CHECK(hipGetLastError());
  std::vector<double> hp(6 * (size_t)g), hd(g);
  std::vector<long long> he(g);
  HIP_CHECK(hipMemcpyAsync(hp.data(), part.get(), sizeof(double) * hp.size(), hipMemcpyDeviceToHost, c.stream));
  HIP_CHECK(hipMemcpyAsync(hd.data(), bd.get(), sizeof(double) * g, hipMemcpyDeviceToHost, c.stream));
  HIP_CHECK(hipMemcpyAsync(he.data(), be.get(), sizeof(long long) * g, hipMemcpyDeviceToHost, c.stream));
  HIP_CHECK(hipStreamSynchronize(c.stream));
  double s[6] = {0, 0, 0, 0, 0, 0};
  double d = DBL_MAX;
  long long e = -1;
  for (int b = 0; b < g; ++b) {
    for (int k = 0; k < 6; ++k) s[k] += hp[6 * (size_t)b + k];
    if (he[b] >= 0 && (hd[b] < d || (hd[b] == d && he[b] < e))) {
      d = hd[b];
      e = he[b];
    }
  }
  *out = cstfem_kirsch_result{};
  out->l2_rel_stress_error = std::sqrt(s[0] / s[1]);
  out->l2_rel_displacement_error = std::sqrt(s[2] / s[3]);
  out->energy_rel_error = std::sqrt(s[4] / s[5]);
  out->scf_element = (int)e;
  out->scf_theoretical = 3.0;
  if (e >= 0) {
    double es[3], ec[2];
    HIP_CHECK(hipMemcpy(es, sig.get() + 3 * e, sizeof es, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(ec, cxy.get() + 2 * e, sizeof ec, hipMemcpyDeviceToHost));
    out->scf_x = ec[0];
    out->scf_y = ec[1];
    out->scf_numerical = es[0] / sigma_inf;
    // kirsch.cpp:7-45 at the centroid (host evaluation of the analytic field)
    const double x = ec[0], y = ec[1];
    const double r = std::hypot(x, y), th = std::atan2(y, x);
    const double q2 = (a * a) / (r * r), q4 = q2 * q2;
    const double c2 = std::cos(2.0 * th), s2 = std::sin(2.0 * th);
    const double rr = 0.5 * sigma_inf * (1.0 - q2) + 0.5 * sigma_inf * (1.0 - 4.0 * q2 + 3.0 * q4) * c2;
    const double tt = 0.5 * sigma_inf * (1.0 + q2) - 0.5 * sigma_inf * (1.0 + 3.0 * q4) * c2;
    const double rt = -0.5 * sigma_inf * (1.0 + 2.0 * q2 - 3.0 * q4) * s2;
    const double cs = std::cos(th), sn = std::sin(th);
    const double sxx = rr * cs * cs + tt * sn * sn - 2.0 * rt * sn * cs;
    out->scf_analytical_centroid = sxx / sigma_inf;
  }
}

}  // namespace cstfem_hip
