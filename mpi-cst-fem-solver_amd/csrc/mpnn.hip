// mpnn.hip -- learned preconditioner: a message-passing network that predicts a sparse
// approximate inverse with the stiffness pattern, trained on residual/error pairs of the
// plain solver, used inside flexible CG.
//
// Network (width H = 64, BF16 weights and activations, FP32 accumulation on MFMA):
//   features f_i (9 live of 16, zero-padded, all invariant under scaling E): 1, spread and
//     contrast of ln E over the node's elements, Dirichlet flags, diagonal-block coupling
//     a01/sqrt(a00 a11), ln(a00/a11), degree/6, ln tr K_ii - mean over neighbours ln tr K_jj
//   h0 = relu(f W0 + b0)                               (16 -> 64)
//   3 rounds: m_i = mean_{j ~ i} h_j ; h = relu([h, m] Wl + bl)    (128 -> 64)
//   decoder per stored 2x2 block (i, j): [a, b, d, c] = Wd/o^T (h_i * h_j) + bd/o
//     (Wd for i = j, Wo otherwise), G^_ij = [[a, b], [b, d]] + c S_i A_ij S_j
//   M = S G^ S, S = diag(A)^(-1/2), i.e. M_ij = S_i [[a, b], [b, d]] S_j + c D_i^-1 A_ij D_j^-1:
//   symmetric by construction (c and [a, b, d] are symmetric in i, j; A_ji = A_ij^T); the
//   c-term lets the network express polynomial preconditioners such as D^-1 (I - beta A D^-1).
//   M = Jacobi when Wd = Wo = 0, bd = (1, 0, 1, 0), bo = 0, which is the initial state.
// Training loss over pairs (r_k, e_k), r_k = A e_k:  (1/K) sum_k |M r_k - e_k|_A^2 / |e_k|_A^2.
// Dense layers run on __builtin_amdgcn_mfma_f32_16x16x16bf16_1k: one wavefront computes 16
// rows x all outputs; weights are kept both as [out x in] and [in x out] BF16 copies so that
// every MFMA operand fragment is a contiguous 8-byte load.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <random>

#include "internal.hpp"

namespace cstfem_hip {

namespace {

constexpr int kH = 64;     // hidden width
constexpr int kF = 16;     // padded feature width
constexpr int kLiveFeatures = 9;
constexpr int kRounds = kMpnnRounds;
constexpr int kT = 256;

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// parameter offsets in the FP32 master vector theta
constexpr int kW0 = 0;                        // [kF x kH]  (in x out, row-major)
constexpr int kB0 = kW0 + kF * kH;
constexpr int kWl = kB0 + kH;                 // kRounds x [2kH x kH]
constexpr int kBl = kWl + kRounds * 2 * kH * kH;
constexpr int kDec = 4;                       // decoder outputs per block: a, b, d, c
constexpr int kWd = kBl + kRounds * kH;       // [kH x kDec]
constexpr int kBd = kWd + kH * kDec;
constexpr int kWo = kBd + kDec;               // [kH x kDec]
constexpr int kBo = kWo + kH * kDec;
constexpr int kParams = kBo + kDec;

__host__ __device__ inline uint16_t f2bf(float f) {  // round to nearest even
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) {
  return __builtin_bit_cast(float, (uint32_t)h << 16);
}

// ---------------------------------------------------------------------------------------
// node features

__global__ void k_features(int nb, const int* __restrict__ brow_ptr, const int* __restrict__ bcol,
                           const int* __restrict__ bdiag, const int64_t* __restrict__ spos,
                           const double* __restrict__ val, const int* __restrict__ ne_ptr,
                           const int* __restrict__ ne_list, const double* __restrict__ E_elem,
                           const unsigned char* __restrict__ fixed, uint16_t* __restrict__ feat) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  // every feature is invariant under a uniform scaling of the modulus
  float f[kF] = {0};
  f[0] = 1.f;
  const int e0 = ne_ptr[i], e1 = ne_ptr[i + 1];
  if (E_elem && e1 > e0) {
    double s = 0.0, s2 = 0.0, lo = 1e300, hi = -1e300;
    for (int k = e0; k < e1; ++k) {
      const double l = log(E_elem[ne_list[k]]);
      s += l;
      s2 += l * l;
      lo = fmin(lo, l);
      hi = fmax(hi, l);
    }
    const double mean = s / (e1 - e0);
    f[1] = (float)sqrt(fmax(s2 / (e1 - e0) - mean * mean, 0.0));  // spread of ln E
    f[2] = (float)(hi - lo);                                      // contrast ln(Emax/Emin)
  }
  f[3] = fixed ? (float)fixed[2 * i] : 0.f;
  f[4] = fixed ? (float)fixed[2 * i + 1] : 0.f;
  const int64_t sd = spos[bdiag[i]];
  const double a00 = val[sell_vidx(sd, 0)], a01 = val[sell_vidx(sd, 1)], a11 = val[sell_vidx(sd, 3)];
  f[5] = (float)(a01 / sqrt(a00 * a11));
  f[6] = (float)log(a00 / a11);
  f[7] = (float)(brow_ptr[i + 1] - brow_ptr[i] - 1) / 6.f;
  // stiffness of the node against its neighbours: ln(tr K_ii) - mean_j ln(tr K_jj)
  double ln_nb = 0.0;
  int cnt = 0;
  for (int q = brow_ptr[i]; q < brow_ptr[i + 1]; ++q) {
    const int j = bcol[q];
    if (j == i) continue;
    const int64_t sj = spos[bdiag[j]];
    ln_nb += log(val[sell_vidx(sj, 0)] + val[sell_vidx(sj, 3)]);
    ++cnt;
  }
  f[8] = cnt ? (float)(log(a00 + a11) - ln_nb / cnt) : 0.f;
  for (int c = 0; c < kF; ++c) feat[(size_t)i * kF + c] = f2bf(f[c]);
}

// ---------------------------------------------------------------------------------------
// dense layer on MFMA: Y[N x M] = act(X[N x K] B^T + bias), B = [M x K] (rows = outputs).
// One wavefront: 16 rows x M outputs; K, M multiples of 16.  kOut: 0 = BF16, 1 = FP32.

template <int K, int M, bool kRelu, int kOut>
__global__ __launch_bounds__(kT) void k_dense(int N, const uint16_t* __restrict__ X,
                                              const uint16_t* __restrict__ B,
                                              const float* __restrict__ bias, void* __restrict__ Yv) {
  const int lane = threadIdx.x & 63;
  const int kq = 4 * (lane >> 4);          // A/B fragment k offset within a 16-step
  // every B fragment of the layer stays in registers while the wavefront walks its row tiles
  v4s bf[M / 16][K / 16];
  float bb[M / 16];
#pragma unroll
  for (int t = 0; t < M / 16; ++t) {
    const int bn = t * 16 + (lane & 15);
#pragma unroll
    for (int k = 0; k < K / 16; ++k)
      bf[t][k] = *reinterpret_cast<const v4s*>(B + (size_t)bn * K + 16 * k + kq);
    bb[t] = bias ? bias[bn] : 0.f;
  }
  const int tiles = (N + 15) / 16;
  for (int tile = blockIdx.x * (kT / 64) + (threadIdx.x >> 6); tile < tiles;
       tile += gridDim.x * (kT / 64)) {
    const int row0 = tile * 16;
    const int ar = row0 + (lane & 15);      // A fragment row
    v4f acc[M / 16];
#pragma unroll
    for (int t = 0; t < M / 16; ++t) acc[t] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < K / 16; ++k) {
      v4s a = {0, 0, 0, 0};
      if (ar < N) a = *reinterpret_cast<const v4s*>(X + (size_t)ar * K + 16 * k + kq);
#pragma unroll
      for (int t = 0; t < M / 16; ++t)
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bf[t][k], acc[t], 0, 0, 0);
    }
    // lane holds D[m = 4*(lane/16) + j][n = lane % 16] of each 16x16 tile
#pragma unroll
    for (int t = 0; t < M / 16; ++t) {
      const int col = t * 16 + (lane & 15);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int row = row0 + 4 * (lane >> 4) + j;
        if (row >= N) continue;
        float v = acc[t][j] + bb[t];
        if (kRelu) v = v > 0.f ? v : 0.f;
        if (kOut == 0)
          reinterpret_cast<uint16_t*>(Yv)[(size_t)row * M + col] = f2bf(v);
        else
          reinterpret_cast<float*>(Yv)[(size_t)row * M + col] = v;
      }
    }
  }
}

// dW partials: P[g][K x M] = sum over the rows n of workgroup g's chunks of 64 rows
// (g, g + grid, ...): X[n][k] * G[n][m] on MFMA; wave w computes output tiles w, w+4, ...
template <int K, int M>
__global__ __launch_bounds__(kT) void k_dw_partial(int N, const uint16_t* __restrict__ X,
                                                   const uint16_t* __restrict__ Gd,
                                                   float* __restrict__ P) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int kTiles = (K / 16) * (M / 16);
  for (int tile = wave; tile < kTiles; tile += kT / 64) {
    const int tk = (tile / (M / 16)) * 16, tm = (tile % (M / 16)) * 16;
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    for (int n0 = blockIdx.x * 64; n0 < N; n0 += gridDim.x * 64) {
#pragma unroll
      for (int ns = 0; ns < 64; ns += 16) {
        v4s a, b;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int n = n0 + ns + 4 * (lane >> 4) + j;
          a[j] = n < N ? (short)X[(size_t)n * K + tk + (lane & 15)] : (short)0;
          b[j] = n < N ? (short)Gd[(size_t)n * M + tm + (lane & 15)] : (short)0;
        }
        // A = X^T tile [k x n], B = G tile [n x m]
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = tk + 4 * (lane >> 4) + j, m = tm + (lane & 15);
      P[(size_t)blockIdx.x * K * M + (size_t)k * M + m] = acc[j];
    }
  }
}

// fixed-order sum of `chunks` partial matrices of `len` floats into out (+=)
__global__ void k_sum_chunks(int chunks, int len, const float* __restrict__ P,
                             float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  float s = 0.f;
  for (int c = 0; c < chunks; ++c) s += P[(size_t)c * len + i];
  out[i] += s;
}

// column sums of G[N x M] (bias gradients): workgroup b sums rows [b*rows, (b+1)*rows) of
// column threadIdx.x into P[b][M]; k_sum_chunks then adds the chunks in order
__global__ void k_colsum(int N, int M, int rows, const uint16_t* __restrict__ Gd,
                         float* __restrict__ P) {
  const int m = threadIdx.x;
  if (m >= M) return;
  const int n0 = blockIdx.x * rows, n1 = min(N, n0 + rows);
  float s = 0.f;
  for (int n = n0; n < n1; ++n) s += bf2f(Gd[(size_t)n * M + m]);
  P[(size_t)blockIdx.x * M + m] = s;
}

// X[i] = [H[i], mean_{j ~ i, j != i} H[j]]  (BF16 in, FP32 mean, BF16 out)
// X[i] = [H[i], mean over neighbours j != i of H[j]] (rows of 2 kH BF16): 8 lanes per node,
// 8 channels (one 16-byte load) per lane; the neighbour indices of a row are read 4 at a time
// ahead of their loads, and every channel sums its neighbours in pattern order
__global__ __launch_bounds__(kT) void k_aggregate(int nb, const int* __restrict__ brow_ptr,
                                                  const int* __restrict__ bcol,
                                                  const uint16_t* __restrict__ Hh,
                                                  uint16_t* __restrict__ X) {
  const int t = blockIdx.x * kT + threadIdx.x;
  const int i = t >> 3, part = t & 7;
  if (i >= nb) return;
  const uint4* H4 = reinterpret_cast<const uint4*>(Hh);  // row j: H4[8 j + part]
  float s[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) s[k] = 0.f;
  int cnt = 0;
  const int q1 = brow_ptr[i + 1];
  for (int q = brow_ptr[i]; q < q1; q += 4) {
    int jj[4];
    uint4 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) jj[u] = q + u < q1 ? bcol[q + u] : i;  // i: not summed
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = H4[8 * (size_t)jj[u] + part];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (jj[u] == i) continue;
      const uint32_t w[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        s[2 * k] += __builtin_bit_cast(float, w[k] << 16);
        s[2 * k + 1] += __builtin_bit_cast(float, w[k] & 0xffff0000u);
      }
      ++cnt;
    }
  }
  uint4* X4 = reinterpret_cast<uint4*>(X);  // row i: X4[16 i + 0..7] = H[i], [8..15] = mean
  X4[16 * (size_t)i + part] = H4[8 * (size_t)i + part];
  uint32_t m[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float lo = cnt ? s[2 * k] / cnt : 0.f, hi = cnt ? s[2 * k + 1] / cnt : 0.f;
    m[k] = (uint32_t)f2bf(lo) | ((uint32_t)f2bf(hi) << 16);
  }
  X4[16 * (size_t)i + 8 + part] = make_uint4(m[0], m[1], m[2], m[3]);
}

// backward of k_aggregate: dH[j] += dX[j][:kH] + sum_{i ~ j, i != j} dX[i][kH:] / deg_i
__global__ void k_aggregate_bwd(int nb, const int* __restrict__ brow_ptr,
                                const int* __restrict__ bcol, const float* __restrict__ dX,
                                float* __restrict__ dH) {
  const int j = blockIdx.x * (blockDim.x / kH) + threadIdx.x / kH;
  const int c = threadIdx.x % kH;
  if (j >= nb) return;
  float s = dX[(size_t)j * 2 * kH + c];
  for (int q = brow_ptr[j]; q < brow_ptr[j + 1]; ++q) {
    const int i = bcol[q];
    if (i == j) continue;
    const int deg = brow_ptr[i + 1] - brow_ptr[i] - 1;
    s += dX[(size_t)i * 2 * kH + kH + c] / (float)deg;
  }
  dH[(size_t)j * kH + c] += s;
}

// dZ = dH * (H > 0) as BF16 (the ReLU mask), dH zeroed for reuse
__global__ void k_relu_bwd(int64_t n, const float* __restrict__ dH, const uint16_t* __restrict__ Hh,
                           uint16_t* __restrict__ dZ) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  dZ[i] = bf2f(Hh[i]) > 0.f ? f2bf(dH[i]) : (uint16_t)0;
}

// decoder output of one block: o = b + W^T (h_i * h_j), channels in ascending order; h_j comes
// in 16-byte loads of 8 channels
__device__ inline void decode_block(const float* __restrict__ hi, const uint4* __restrict__ hj,
                                    const float* __restrict__ W, const float* __restrict__ bv,
                                    float o[kDec]) {
#pragma unroll
  for (int u = 0; u < kDec; ++u) o[u] = bv[u];
#pragma unroll
  for (int c8 = 0; c8 < kH / 8; ++c8) {
    const uint4 v = hj[c8];
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int c = 8 * c8 + k;
      const float h = __builtin_bit_cast(float, (k & 1) ? (w[k / 2] & 0xffff0000u) : (w[k / 2] << 16));
      const float p = hi[c] * h;
#pragma unroll
      for (int u = 0; u < kDec; ++u) o[u] += W[c * kDec + u] * p;
    }
  }
}

// decoder: per stored block q = (i, j) of row i: [a, b, d, c] and the SELL values of M.
// One thread per row; the off-diagonal blocks share the weights Wo (uniform loads), the
// diagonal block uses Wd
__global__ void k_decode(int nb, const int* __restrict__ brow_ptr, const int* __restrict__ bcol,
                         const int64_t* __restrict__ spos, const uint16_t* __restrict__ Hh,
                         const float* __restrict__ theta, const double* __restrict__ invd,
                         const double* __restrict__ aval, float* __restrict__ ghat,
                         float* __restrict__ gsell) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  const uint4* H4 = reinterpret_cast<const uint4*>(Hh);
  float hi[kH];
#pragma unroll
  for (int c8 = 0; c8 < kH / 8; ++c8) {
    const uint4 v = H4[8 * (size_t)i + c8];
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      hi[8 * c8 + 2 * k] = __builtin_bit_cast(float, w[k] << 16);
      hi[8 * c8 + 2 * k + 1] = __builtin_bit_cast(float, w[k] & 0xffff0000u);
    }
  }
  const double di0 = invd[2 * i], di1 = invd[2 * i + 1];
  const float si0 = (float)sqrt(di0), si1 = (float)sqrt(di1);
  auto store = [&](int q, int j, const float o[kDec]) {
#pragma unroll
    for (int u = 0; u < kDec; ++u) ghat[kDec * (size_t)q + u] = o[u];
    const double dj0 = invd[2 * j], dj1 = invd[2 * j + 1];
    const float sj0 = (float)sqrt(dj0), sj1 = (float)sqrt(dj1);
    const int64_t sq = spos[q];
    const double a[4] = {aval[sell_vidx(sq, 0)], aval[sell_vidx(sq, 1)], aval[sell_vidx(sq, 2)],
                         aval[sell_vidx(sq, 3)]};
    float* g = gsell + 4 * sq;
    g[0] = si0 * o[0] * sj0 + o[3] * (float)(di0 * a[0] * dj0);
    g[1] = si0 * o[1] * sj1 + o[3] * (float)(di0 * a[1] * dj1);
    g[2] = si1 * o[1] * sj0 + o[3] * (float)(di1 * a[2] * dj0);
    g[3] = si1 * o[2] * sj1 + o[3] * (float)(di1 * a[3] * dj1);
  };
  int qd = -1;
  for (int q = brow_ptr[i]; q < brow_ptr[i + 1]; ++q) {
    const int j = bcol[q];
    if (j == i) {
      qd = q;
      continue;
    }
    float o[kDec];
    decode_block(hi, H4 + 8 * (size_t)j, theta + kWo, theta + kBo, o);
    store(q, j, o);
  }
  if (qd >= 0) {
    float o[kDec];
    decode_block(hi, H4 + 8 * (size_t)i, theta + kWd, theta + kBd, o);
    store(qd, i, o);
  }
}

// y = M x with M in the SELL layout (FP32 values), FP32 vectors
__global__ __launch_bounds__(kT) void k_gspmv(int nslices, const int* __restrict__ row0,
                                              const int* __restrict__ nrow,
                                              const int64_t* __restrict__ off,
                                              const int* __restrict__ col,
                                              const float* __restrict__ val,
                                              const float2* __restrict__ x, float2* __restrict__ y) {
  const int lane = threadIdx.x & 63;
  for (int s = blockIdx.x * (kT / 64) + (threadIdx.x >> 6); s < nslices; s += gridDim.x * (kT / 64)) {
    const int64_t o = off[s];
    const int w = (int)((off[s + 1] - o) >> 6);
    float a0 = 0.f, a1 = 0.f;
    for (int k = 0; k < w; ++k) {
      const int64_t slot = o + (int64_t)k * 64 + lane;
      const int m = col[slot];
      const float4 v = reinterpret_cast<const float4*>(val)[slot];
      const float2 xm = x[m];
      a0 = a0 + v.x * xm.x;
      a0 = a0 + v.y * xm.y;
      a1 = a1 + v.z * xm.x;
      a1 = a1 + v.w * xm.y;
    }
    if (lane < nrow[s]) y[row0[s] + lane] = make_float2(a0, a1);
  }
}

// FP64-accumulated dot partials of two float2 vectors (one partial per workgroup)
__global__ __launch_bounds__(kT) void k_dot32(int n2, const float2* __restrict__ a,
                                              const float2* __restrict__ b, double* __restrict__ part) {
  __shared__ double lds[kT / 64];
  double s = 0.0;
  for (int i = blockIdx.x * kT + threadIdx.x; i < n2; i += gridDim.x * kT)
    s += (double)a[i].x * b[i].x + (double)a[i].y * b[i].y;
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kT / 64; ++w) t += lds[w];
    part[blockIdx.x] = t;
  }
}

__global__ void k_sub32(int n2, const float2* __restrict__ a, const float2* __restrict__ b,
                        float2* __restrict__ c) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n2) c[i] = make_float2(a[i].x - b[i].x, a[i].y - b[i].y);
}

// dG^ for every stored block from the pairs: dM_ij = sum_k w_k v_k,i r_k,j^T (2x2), mapped
// through M_ij = S_i [[a, b], [b, d]] S_j + c D_i^-1 A_ij D_j^-1 onto [a, b, d, c].
// V: K x n2 (v_k = A (M r_k - e_k)), R: K x n2.
__global__ void k_dghat(int nb, int K, int n2, const int* __restrict__ brow_ptr,
                        const int* __restrict__ bcol, const int64_t* __restrict__ spos,
                        const double* __restrict__ invd, const double* __restrict__ aval,
                        const float2* __restrict__ V, const float2* __restrict__ R,
                        const double* __restrict__ wk, float* __restrict__ dghat) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  const double di0 = invd[2 * i], di1 = invd[2 * i + 1];
  const double si0 = (float)sqrt(di0), si1 = (float)sqrt(di1);
  for (int q = brow_ptr[i]; q < brow_ptr[i + 1]; ++q) {
    const int j = bcol[q];
    const double dj0 = invd[2 * j], dj1 = invd[2 * j + 1];
    const double sj0 = (float)sqrt(dj0), sj1 = (float)sqrt(dj1);
    double g00 = 0, g01 = 0, g10 = 0, g11 = 0;
    for (int k = 0; k < K; ++k) {
      const float2 v = V[(size_t)k * n2 + i], r = R[(size_t)k * n2 + j];
      const double w = wk[k];
      g00 += w * v.x * r.x;
      g01 += w * v.x * r.y;
      g10 += w * v.y * r.x;
      g11 += w * v.y * r.y;
    }
    const int64_t sq = spos[q];
    const double a[4] = {aval[sell_vidx(sq, 0)], aval[sell_vidx(sq, 1)], aval[sell_vidx(sq, 2)],
                         aval[sell_vidx(sq, 3)]};
    float* d = dghat + kDec * (size_t)q;
    d[0] = (float)(g00 * si0 * sj0);
    d[1] = (float)(g01 * si0 * sj1 + g10 * si1 * sj0);
    d[2] = (float)(g11 * si1 * sj1);
    d[3] = (float)(g00 * (di0 * a[0] * dj0) + g01 * (di0 * a[1] * dj1) +
                   g10 * (di1 * a[2] * dj0) + g11 * (di1 * a[3] * dj1));
  }
}

// decoder backward, part 1: dH[i] = sum_{q = (i,j)} (W (dG^_q + dG^_tr(q))) * h_j
__global__ void k_decode_bwd_h(int nb, const int* __restrict__ brow_ptr,
                               const int* __restrict__ bcol, const int* __restrict__ btrans,
                               const uint16_t* __restrict__ Hh, const float* __restrict__ theta,
                               const float* __restrict__ dghat, float* __restrict__ dH) {
  const int i = blockIdx.x * (blockDim.x / kH) + threadIdx.x / kH;
  const int c = threadIdx.x % kH;
  if (i >= nb) return;
  float s = 0.f;
  for (int q = brow_ptr[i]; q < brow_ptr[i + 1]; ++q) {
    const int j = bcol[q], t = btrans[q];
    const float* W = theta + (j == i ? kWd : kWo);
    float g = 0.f;
#pragma unroll
    for (int o = 0; o < kDec; ++o)
      g += W[c * kDec + o] * (dghat[kDec * (size_t)q + o] + dghat[kDec * (size_t)t + o]);
    s += g * bf2f(Hh[(size_t)j * kH + c]);
  }
  dH[(size_t)i * kH + c] = s;
}

// decoder backward, part 2: dW (64 x kDec) and db (kDec) of the diagonal (which = 0) or off-diagonal
// (which = 1) decoder; one workgroup per hidden channel c, fixed-order reductions
__global__ __launch_bounds__(kT) void k_decode_bwd_w(int nb, const int* __restrict__ brow_ptr,
                                                     const int* __restrict__ bcol,
                                                     const uint16_t* __restrict__ Hh,
                                                     const float* __restrict__ dghat,
                                                     float* __restrict__ grad) {
  __shared__ double lds[kT][2 * kDec];
  const int c = blockIdx.x;  // 0..kH-1, and kH = bias
  double acc[2 * kDec] = {};
  for (int i = threadIdx.x; i < nb; i += kT) {
    const float hic = c < kH ? bf2f(Hh[(size_t)i * kH + c]) : 1.f;
    for (int q = brow_ptr[i]; q < brow_ptr[i + 1]; ++q) {
      const int j = bcol[q];
      const float p = c < kH ? hic * bf2f(Hh[(size_t)j * kH + c]) : 1.f;
      const int base = j == i ? 0 : kDec;
#pragma unroll
      for (int o = 0; o < kDec; ++o) acc[base + o] += (double)p * dghat[kDec * (size_t)q + o];
    }
  }
  for (int o = 0; o < 2 * kDec; ++o) lds[threadIdx.x][o] = acc[o];
  __syncthreads();
  for (int stride = kT / 2; stride > 0; stride >>= 1) {
    if ((int)threadIdx.x < stride)
      for (int o = 0; o < 2 * kDec; ++o) lds[threadIdx.x][o] += lds[threadIdx.x + stride][o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    for (int o = 0; o < kDec; ++o) {
      if (c < kH) {
        grad[kWd + c * kDec + o] += (float)lds[0][o];
        grad[kWo + c * kDec + o] += (float)lds[0][kDec + o];
      } else {
        grad[kBd + o] += (float)lds[0][o];
        grad[kBo + o] += (float)lds[0][kDec + o];
      }
    }
  }
}

__global__ void k_adam(int n, float* __restrict__ theta, const float* __restrict__ g,
                       float* __restrict__ m, float* __restrict__ v, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  m[i] = b1 * m[i] + (1.f - b1) * gi;
  v[i] = b2 * v[i] + (1.f - b2) * gi * gi;
  theta[i] -= lr * (m[i] / c1) / (sqrtf(v[i] / c2) + eps);
}

// BF16 copies of one dense layer: Wt [out x in] and W [in x out] from theta (in x out)
__global__ void k_weights_bf16(int in, int out, const float* __restrict__ w, uint16_t* __restrict__ wt,
                               uint16_t* __restrict__ wn) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= in * out) return;
  const int r = i / out, c = i % out;
  const uint16_t h = f2bf(w[i]);
  wn[i] = h;
  wt[(size_t)c * in + r] = h;
}

// harvest: e = x* - x (FP32), pair slot k
__global__ void k_harvest(int n2, const double2* __restrict__ xs, const double2* __restrict__ x,
                          float2* __restrict__ e) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n2) e[i] = make_float2((float)(xs[i].x - x[i].x), (float)(xs[i].y - x[i].y));
}

__global__ void k_to_f32(int n2, const double2* __restrict__ a, float2* __restrict__ b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n2) b[i] = make_float2((float)a[i].x, (float)a[i].y);
}

__global__ void k_to_f64(int n2, const float2* __restrict__ a, double2* __restrict__ b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n2) b[i] = make_double2(a[i].x, a[i].y);
}

// block transpose map: btrans[q] = position of (j, i) for q = (i, j)
__global__ void k_btrans(int nb, const int* __restrict__ brow_ptr, const int* __restrict__ bcol,
                         int* __restrict__ btrans) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  for (int q = brow_ptr[i]; q < brow_ptr[i + 1]; ++q) {
    const int j = bcol[q];
    int lo = brow_ptr[j], hi = brow_ptr[j + 1];
    while (lo < hi) {
      const int mid = (lo + hi) / 2;
      if (bcol[mid] < i) lo = mid + 1; else hi = mid;
    }
    btrans[q] = lo;
  }
}

}  // namespace

// ---------------------------------------------------------------------------------------
// host side

int mpnn_num_params() { return kParams; }

void mpnn_init(Context& c, int seed) {
  std::vector<float> th(kParams, 0.f);
  std::mt19937 gen(static_cast<unsigned>(seed));
  auto fill = [&](int off, int in, int out) {
    std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (float)in));
    for (int i = 0; i < in * out; ++i) th[off + i] = nd(gen);
  };
  fill(kW0, kLiveFeatures, kH);  // rows of the zero padding stay 0
  for (int l = 0; l < kRounds; ++l) fill(kWl + l * 2 * kH * kH, 2 * kH, kH);
  th[kBd + 0] = 1.f;  // decoder = identity: M = Jacobi
  th[kBd + 2] = 1.f;
  Mpnn& m = c.mpnn;
  m.theta.alloc(kParams);
  m.grad.alloc(kParams);
  m.am.alloc(kParams);
  m.av.alloc(kParams);
  HIP_CHECK(hipMemcpyAsync(m.theta.get(), th.data(), sizeof(float) * kParams, hipMemcpyHostToDevice, c.stream));
  HIP_CHECK(hipMemsetAsync(m.am.get(), 0, sizeof(float) * kParams, c.stream));
  HIP_CHECK(hipMemsetAsync(m.av.get(), 0, sizeof(float) * kParams, c.stream));
  HIP_CHECK(hipStreamSynchronize(c.stream));
  m.step = 0;
  m.initialized = true;
}

void mpnn_set_params(Context& c, const float* th, int n) {
  if (n != kParams) throw Error(CSTFEM_ERR_ARGUMENT, "mpnn: expected " + std::to_string(kParams) + " parameters");
  if (!c.mpnn.initialized) mpnn_init(c, 0);
  HIP_CHECK(hipMemcpyAsync(c.mpnn.theta.get(), th, sizeof(float) * kParams, hipMemcpyHostToDevice, c.stream));
  HIP_CHECK(hipStreamSynchronize(c.stream));
}

void mpnn_get_params(Context& c, float* th, int n) {
  if (n != kParams) throw Error(CSTFEM_ERR_ARGUMENT, "mpnn: expected " + std::to_string(kParams) + " parameters");
  if (!c.mpnn.initialized) throw Error(CSTFEM_ERR_ORDER, "mpnn: not initialized");
  HIP_CHECK(hipMemcpyAsync(th, c.mpnn.theta.get(), sizeof(float) * kParams, hipMemcpyDeviceToHost, c.stream));
  HIP_CHECK(hipStreamSynchronize(c.stream));
}

namespace {

template <class T>
T* grow_ws(DBuf<T>& b, size_t count) {
  if (!b.get() || b.n < count) b.alloc(std::max<size_t>(count, 1));
  return b.get();
}

void check_ready(Context& c) {
  if (!c.mpnn.initialized) throw Error(CSTFEM_ERR_ORDER, "mpnn: call cstfem_mpnn_init first");
  if (!c.have_pattern || !c.mesh_pattern)
    throw Error(CSTFEM_ERR_ORDER, "mpnn: needs an assembled mesh system");
  if (!c.setup_done) throw Error(CSTFEM_ERR_ORDER, "mpnn: call cstfem_setup first (Jacobi scaling)");
}

void weights_bf16(Context& c) {
  Mpnn& m = c.mpnn;
  m.wt[0].alloc(kF * kH);
  m.wn[0].alloc(kF * kH);
  hipLaunchKernelGGL(k_weights_bf16, dim3((kF * kH + 255) / 256), dim3(256), 0, c.stream, kF, kH,
                     (const float*)m.theta.get() + kW0, m.wt[0].get(), m.wn[0].get());
  for (int l = 0; l < kRounds; ++l) {
    m.wt[l + 1].alloc(2 * kH * kH);
    m.wn[l + 1].alloc(2 * kH * kH);
    hipLaunchKernelGGL(k_weights_bf16, dim3((2 * kH * kH + 255) / 256), dim3(256), 0, c.stream,
                       2 * kH, kH, (const float*)m.theta.get() + kWl + l * 2 * kH * kH,
                       m.wt[l + 1].get(), m.wn[l + 1].get());
  }
  HIP_CHECK(hipGetLastError());
}

// workgroups of k_dense: 4 wavefronts each, every wavefront walks 16-row tiles; capped so that
// a wavefront's registers of weights serve several tiles
int rows_grid(int N) { return std::max(1, std::min((N + 63) / 64, 2048)); }

}  // namespace

// features + forward; fills c.mpnn.ghat and the SELL values of M
void mpnn_forward(Context& c) {
  check_ready(c);
  Mpnn& m = c.mpnn;
  const int N = c.nb;
  if (c.ne_ptr.n == 0) build_node_elements(c);
  m.feat.alloc((size_t)N * kF);
  hipLaunchKernelGGL(k_features, dim3((N + 127) / 128), dim3(128), 0, c.stream, N, c.brow_ptr.get(),
                     c.bcol.get(), c.bdiag.get(), c.sell_pos.get(), c.sell_val.get(), c.ne_ptr.get(),
                     c.ne_list.get(), c.has_E_elem ? c.E_elem.get() : (const double*)nullptr,
                     c.fixed_valid ? c.fixed.get() : (const unsigned char*)nullptr, m.feat.get());
  weights_bf16(c);
  for (int l = 0; l <= kRounds; ++l) m.H[l].alloc((size_t)N * kH);
  for (int l = 1; l <= kRounds; ++l) m.X[l].alloc((size_t)N * 2 * kH);
  hipLaunchKernelGGL((k_dense<kF, kH, true, 0>), dim3(rows_grid(N)), dim3(kT), 0, c.stream, N,
                     (const uint16_t*)m.feat.get(), (const uint16_t*)m.wt[0].get(),
                     (const float*)m.theta.get() + kB0, (void*)m.H[0].get());
  for (int l = 1; l <= kRounds; ++l) {
    hipLaunchKernelGGL(k_aggregate, dim3((8 * (int64_t)N + kT - 1) / kT), dim3(kT), 0, c.stream, N,
                       c.brow_ptr.get(), c.bcol.get(), (const uint16_t*)m.H[l - 1].get(), m.X[l].get());
    hipLaunchKernelGGL((k_dense<2 * kH, kH, true, 0>), dim3(rows_grid(N)), dim3(kT), 0, c.stream, N,
                       (const uint16_t*)m.X[l].get(), (const uint16_t*)m.wt[l].get(),
                       (const float*)m.theta.get() + kBl + (l - 1) * kH, (void*)m.H[l].get());
  }
  m.ghat.alloc(kDec * (size_t)c.nblk);
  m.gsell.alloc(4 * (size_t)c.sell_blocks);
  HIP_CHECK(hipMemsetAsync(m.gsell.get(), 0, sizeof(float) * 4 * c.sell_blocks, c.stream));
  hipLaunchKernelGGL(k_decode, dim3((N + 127) / 128), dim3(128), 0, c.stream, N, c.brow_ptr.get(),
                     c.bcol.get(), c.sell_pos.get(), (const uint16_t*)m.H[kRounds].get(),
                     (const float*)m.theta.get(), (const double*)c.inv_diag.get(),
                     (const double*)c.sell_val.get(), m.ghat.get(), m.gsell.get());
  HIP_CHECK(hipGetLastError());
  m.built = true;  // M is ready in stream order; no host synchronisation (set-up may overlap)
}

void mpnn_set_pairs(Context& c, int K, const float* R, const float* E) {
  check_ready(c);
  Mpnn& m = c.mpnn;
  const size_t n = 2 * (size_t)c.nb;
  m.R.alloc(K * n);
  m.E.alloc(K * n);
  HIP_CHECK(hipMemsetAsync(m.R.get(), 0, sizeof(float) * K * n, c.stream));
  HIP_CHECK(hipMemsetAsync(m.E.get(), 0, sizeof(float) * K * n, c.stream));
  for (int k = 0; k < K; ++k) {
    if (R) HIP_CHECK(hipMemcpyAsync(m.R.get() + k * n, R + k * (size_t)c.n, sizeof(float) * c.n, hipMemcpyHostToDevice, c.stream));
    if (E) HIP_CHECK(hipMemcpyAsync(m.E.get() + k * n, E + k * (size_t)c.n, sizeof(float) * c.n, hipMemcpyHostToDevice, c.stream));
  }
  HIP_CHECK(hipStreamSynchronize(c.stream));
  m.npairs = K;
}

// Runs plain FP64 Jacobi-CG from x = 0 to rtol 1e-12 for x*, then again recording, after
// iterations k_first, k_first + stride, ... <= k_last, the pairs e = x* - x_k and r = A e.
int mpnn_harvest(Context& c, int k_first, int k_last, int stride) {
  check_ready(c);
  if (k_first < 1 || k_last < k_first || stride < 1)
    throw Error(CSTFEM_ERR_ARGUMENT, "mpnn_harvest: need 1 <= k_first <= k_last, stride >= 1");
  Mpnn& m = c.mpnn;
  const int n2 = c.nb;
  const size_t n = 2 * (size_t)c.nb;
  cstfem_solver_options o{};
  o.rtol = 1e-12;
  o.max_iter = 1000000;
  o.method = CSTFEM_METHOD_CG_FP64_REFINED;
  cstfem_solver_stats st{};
  solve_refined(c, o, st, false);
  DBuf<double> xs;
  xs.alloc(n);
  HIP_CHECK(hipMemcpyAsync(xs.get(), c.x.get(), sizeof(double) * n, hipMemcpyDeviceToDevice, c.stream));
  const int K = (k_last - k_first) / stride + 1;
  m.R.alloc(K * n);
  m.E.alloc(K * n);
  int got = 0;
  // replay: the plain recurrence (reference semantics), eager, recording pairs
  cstfem_solver_options p{};
  p.rtol = 1e-300;
  p.max_iter = 0;
  for (int k = k_first; k <= k_last && got < K; k += stride) {
    p.max_iter = k;
    cstfem_solver_stats s2{};
    solve_cg_fp64(c, p, s2, nullptr);  // x after k iterations (deterministic recurrence)
    hipLaunchKernelGGL(k_harvest, dim3((n2 + 255) / 256), dim3(256), 0, c.stream, n2,
                       (const double2*)xs.get(), (const double2*)c.x.get(),
                       (float2*)(m.E.get() + got * n));
    ++got;
  }
  // r = A e in FP64 from the FP32 errors
  for (int k = 0; k < got; ++k) {
    hipLaunchKernelGGL(k_to_f64, dim3((n2 + 255) / 256), dim3(256), 0, c.stream, n2,
                       (const float2*)(m.E.get() + k * n), (double2*)c.work.get());
    spmv_fp64(c, c.work.get(), c.Ap.get());
    hipLaunchKernelGGL(k_to_f32, dim3((n2 + 255) / 256), dim3(256), 0, c.stream, n2,
                       (const double2*)c.Ap.get(), (float2*)(m.R.get() + k * n));
  }
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(c.stream));
  m.npairs = got;
  return got;
}

namespace {
// e = e - z (FP64) and its FP32 copy
__global__ void k_sub_to_f32(int64_t n, double* __restrict__ e, const double* __restrict__ z,
                             float* __restrict__ e32) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double v = e[i] - z[i];
    e[i] = v;
    e32[i] = (float)v;
  }
}
}  // namespace

namespace {
// splitmix64 of a counter: independent uniform 64-bit words per (seed, index)
__device__ inline uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// e[i] ~ N(0, 1) by the Box-Muller transform of two uniforms drawn from (seed, row, i)
__global__ void k_white(int64_t n, uint64_t key, double* __restrict__ e64, float* __restrict__ e32) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t w = mix64(key ^ mix64((uint64_t)i));
    const double u1 = ((double)(w >> 11) + 0.5) * 0x1p-53;                // (0, 1)
    const double u2 = (double)(mix64(w) >> 11) * 0x1p-53;                 // [0, 1)
    const double v = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    e64[i] = v;
    e32[i] = (float)v;
  }
}
}  // namespace

// Appends `count` white-noise pairs: e ~ N(0, I) drawn on the device from `seed`, r = K e
// (FP64 SpMV); pairs already present are kept.
void mpnn_add_white_pairs(Context& c, int count, int seed) {
  check_ready(c);
  if (count <= 0) throw Error(CSTFEM_ERR_ARGUMENT, "mpnn_add_white_pairs: count must be positive");
  Mpnn& m = c.mpnn;
  const size_t n = 2 * (size_t)c.nb;
  const int old = m.npairs > 0 && m.R.n >= (size_t)m.npairs * n ? m.npairs : 0;
  const int K = old + count;
  DBuf<float> R2, E2;
  R2.alloc((size_t)K * n);
  E2.alloc((size_t)K * n);
  if (old) {
    HIP_CHECK(hipMemcpyAsync(R2.get(), m.R.get(), sizeof(float) * old * n, hipMemcpyDeviceToDevice, c.stream));
    HIP_CHECK(hipMemcpyAsync(E2.get(), m.E.get(), sizeof(float) * old * n, hipMemcpyDeviceToDevice, c.stream));
  }
  DBuf<double> e64, r64;
  e64.alloc(n);
  r64.alloc(n);
  const int g = std::max(1, std::min(4096, (int)((n + 255) / 256)));
  for (int k = 0; k < count; ++k) {
    const uint64_t key = ((uint64_t)(uint32_t)seed << 32) ^ (uint64_t)(uint32_t)(old + k);
    hipLaunchKernelGGL(k_white, dim3(g), dim3(256), 0, c.stream, (int64_t)n, key, e64.get(),
                       E2.get() + (size_t)(old + k) * n);
    spmv_fp64(c, e64.get(), r64.get());
    hipLaunchKernelGGL(k_to_f32, dim3((c.nb + 255) / 256), dim3(256), 0, c.stream, c.nb,
                       (const double2*)r64.get(), (float2*)(R2.get() + (size_t)(old + k) * n));
  }
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(c.stream));
  std::swap(m.R.p, R2.p);
  std::swap(m.R.n, R2.n);
  std::swap(m.E.p, E2.p);
  std::swap(m.E.n, E2.n);
  m.npairs = K;
}

// Every pair (r, e) becomes (K e', e') with e' = e - P C(P^T K e): the part of the error the
// hierarchy's coarse correction leaves, which is all a fine-level smoother has to reduce.
void mpnn_project_pairs(Context& c) {
  Mpnn& m = c.mpnn;
  if (m.npairs <= 0) throw Error(CSTFEM_ERR_ORDER, "mpnn: no training pairs");
  if (!c.mg.ready) throw Error(CSTFEM_ERR_ORDER, "mpnn: projecting pairs needs the multilevel hierarchy (cstfem_mg_setup)");
  const int n2 = c.nb;
  const size_t n = 2 * (size_t)c.nb;
  DBuf<double> e64, r64, z64;
  e64.alloc(n);
  r64.alloc(n);
  z64.alloc(n);
  const int g = std::max(1, std::min(4096, (n2 + 255) / 256));
  for (int k = 0; k < m.npairs; ++k) {
    float* E = m.E.get() + k * n;
    float* R = m.R.get() + k * n;
    hipLaunchKernelGGL(k_to_f64, dim3((n2 + 255) / 256), dim3(256), 0, c.stream, n2,
                       (const float2*)E, (double2*)e64.get());
    spmv_fp64(c, e64.get(), r64.get());
    mg_enqueue_coarse(c, r64.get(), z64.get());
    hipLaunchKernelGGL(k_sub_to_f32, dim3(g), dim3(256), 0, c.stream, (int64_t)n, e64.get(),
                       (const double*)z64.get(), E);
    spmv_fp64(c, e64.get(), r64.get());
    hipLaunchKernelGGL(k_to_f32, dim3((n2 + 255) / 256), dim3(256), 0, c.stream, n2,
                       (const double2*)r64.get(), (float2*)R);
  }
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(c.stream));
}

// forward + loss over the pairs + backward into c.mpnn.grad (zeroed first); returns the loss
double mpnn_loss_grad(Context& c) {
  Mpnn& m = c.mpnn;
  if (m.npairs <= 0) throw Error(CSTFEM_ERR_ORDER, "mpnn: no training pairs");
  mpnn_forward(c);
  const int N = c.nb, K = m.npairs, n2 = c.nb;
  const size_t n = 2 * (size_t)N;
  ensure_vals32(c);
  // the step's workspace is kept in the context (grown on demand): at millions of nodes and
  // tens of pairs it is gigabytes, which a step must not allocate and free
  float* const Ug = grow_ws(m.wsU, (size_t)K * n);
  float* const Vg = grow_ws(m.wsV, (size_t)K * n);
  const int gd = std::max(1, std::min(1024, (n2 + kT - 1) / kT));
  double* const partg = grow_ws(m.wsPart, 4 * (size_t)gd * K + 8);
  std::vector<double> h_part(4 * (size_t)gd * K);
  // u_k = M r_k ; d_k = u_k - e_k ; v_k = A d_k (FP32 values of A)
  const int gs = std::max(1, std::min(2048, (c.nslices + 3) / 4));
  for (int k = 0; k < K; ++k) {
    const float2* r = (const float2*)(m.R.get() + k * n);
    const float2* e = (const float2*)(m.E.get() + k * n);
    float2* u = (float2*)(Ug + k * n);
    float2* v = (float2*)(Vg + k * n);
    hipLaunchKernelGGL(k_gspmv, dim3(gs), dim3(kT), 0, c.stream, c.nslices, c.sell_row0.get(),
                       c.sell_nrow.get(), c.sell_off.get(), c.sell_col.get(), (const float*)m.gsell.get(), r, u);
    hipLaunchKernelGGL(k_sub32, dim3((n2 + 255) / 256), dim3(256), 0, c.stream, n2, (const float2*)u, e, u);
    hipLaunchKernelGGL(k_gspmv, dim3(gs), dim3(kT), 0, c.stream, c.nslices, c.sell_row0.get(),
                       c.sell_nrow.get(), c.sell_off.get(), c.sell_col.get(), (const float*)c.sell_val32.get(),
                       (const float2*)u, v);
    // d'Ad and e'Ae (= e'r)
    hipLaunchKernelGGL(k_dot32, dim3(gd), dim3(kT), 0, c.stream, n2, (const float2*)u, (const float2*)v,
                       partg + (size_t)(2 * k) * gd);
    hipLaunchKernelGGL(k_dot32, dim3(gd), dim3(kT), 0, c.stream, n2, e, r, partg + (size_t)(2 * k + 1) * gd);
  }
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(h_part.data(), partg, sizeof(double) * 2 * gd * K, hipMemcpyDeviceToHost, c.stream));
  HIP_CHECK(hipStreamSynchronize(c.stream));
  std::vector<double> wk(K);
  double loss = 0.0;
  for (int k = 0; k < K; ++k) {
    double dad = 0.0, eae = 0.0;
    for (int b = 0; b < gd; ++b) {
      dad += h_part[(size_t)(2 * k) * gd + b];
      eae += h_part[(size_t)(2 * k + 1) * gd + b];
    }
    const double ce = eae > 0 ? 1.0 / eae : 0.0;
    loss += dad * ce / K;
    wk[k] = 2.0 * ce / K;  // dL/dM = sum_k (2 / (K e'Ae)) A d_k r_k^T
  }
  double* const d_wk = grow_ws(m.wsW, (size_t)K);
  HIP_CHECK(hipMemcpyAsync(d_wk, wk.data(), sizeof(double) * K, hipMemcpyHostToDevice, c.stream));
  // backward
  HIP_CHECK(hipMemsetAsync(m.grad.get(), 0, sizeof(float) * kParams, c.stream));
  m.dghat.alloc(kDec * (size_t)c.nblk);
  hipLaunchKernelGGL(k_dghat, dim3((N + 127) / 128), dim3(128), 0, c.stream, N, K, n2, c.brow_ptr.get(),
                     c.bcol.get(), c.sell_pos.get(), (const double*)c.inv_diag.get(),
                     (const double*)c.sell_val.get(), (const float2*)Vg,
                     (const float2*)m.R.get(), (const double*)d_wk, m.dghat.get());
  if (m.btrans.n != (size_t)c.nblk) {
    m.btrans.alloc(c.nblk);
    hipLaunchKernelGGL(k_btrans, dim3((N + 127) / 128), dim3(128), 0, c.stream, N, c.brow_ptr.get(),
                       c.bcol.get(), m.btrans.get());
  }
  float* const dH = grow_ws(m.wsdH, (size_t)N * kH);
  float* const dX = grow_ws(m.wsdX, (size_t)N * 2 * kH);
  uint16_t* const dZ = grow_ws(m.wsdZ, (size_t)N * kH);
  hipLaunchKernelGGL(k_decode_bwd_h, dim3((N + 3) / 4), dim3(4 * kH), 0, c.stream, N, c.brow_ptr.get(),
                     c.bcol.get(), (const int*)m.btrans.get(), (const uint16_t*)m.H[kRounds].get(),
                     (const float*)m.theta.get(), (const float*)m.dghat.get(), dH);
  hipLaunchKernelGGL(k_decode_bwd_w, dim3(kH + 1), dim3(kT), 0, c.stream, N, c.brow_ptr.get(),
                     c.bcol.get(), (const uint16_t*)m.H[kRounds].get(), (const float*)m.dghat.get(),
                     m.grad.get());
  const int chunks = std::min((N + 63) / 64, 256);  // workgroups of the dW reductions
  const int crows = (N + chunks - 1) / chunks;
  float* const CP = grow_ws(m.wsCP, (size_t)chunks * kH);
  auto colsum = [&](const uint16_t* G, float* out) {
    hipLaunchKernelGGL(k_colsum, dim3(chunks), dim3(kH), 0, c.stream, N, kH, crows, G, CP);
    hipLaunchKernelGGL(k_sum_chunks, dim3(1), dim3(kH), 0, c.stream, chunks, kH,
                       (const float*)CP, out);
  };
  float* const P = grow_ws(m.wsP, (size_t)chunks * 2 * kH * kH);
  for (int l = kRounds; l >= 1; --l) {
    // dZ = dH * relu'(H_l)
    hipLaunchKernelGGL(k_relu_bwd, dim3((N * kH + 255) / 256), dim3(256), 0, c.stream, (int64_t)N * kH,
                       (const float*)dH, (const uint16_t*)m.H[l].get(), dZ);
    // dW_l = X_l^T dZ, db_l = colsum dZ
    hipLaunchKernelGGL((k_dw_partial<2 * kH, kH>), dim3(chunks), dim3(kT), 0, c.stream, N,
                       (const uint16_t*)m.X[l].get(), (const uint16_t*)dZ, P);
    hipLaunchKernelGGL(k_sum_chunks, dim3((2 * kH * kH + 255) / 256), dim3(256), 0, c.stream, chunks,
                       2 * kH * kH, (const float*)P, m.grad.get() + kWl + (l - 1) * 2 * kH * kH);
    colsum(dZ, m.grad.get() + kBl + (l - 1) * kH);
    // dX_l = dZ W_l^T  ([N x kH] x [kH x 2kH]): B = W_l as [in x out] = [2kH rows x kH]
    hipLaunchKernelGGL((k_dense<kH, 2 * kH, false, 1>), dim3(rows_grid(N)), dim3(kT), 0, c.stream, N,
                       (const uint16_t*)dZ, (const uint16_t*)m.wn[l].get(), (const float*)nullptr,
                       (void*)dX);
    // dH_{l-1} = dX[:, :kH] + aggregation backward of dX[:, kH:]
    HIP_CHECK(hipMemsetAsync(dH, 0, sizeof(float) * (size_t)N * kH, c.stream));
    hipLaunchKernelGGL(k_aggregate_bwd, dim3((N + 3) / 4), dim3(4 * kH), 0, c.stream, N,
                       c.brow_ptr.get(), c.bcol.get(), (const float*)dX, dH);
  }
  // encoder: dW0 = F^T dZ0, db0
  hipLaunchKernelGGL(k_relu_bwd, dim3((N * kH + 255) / 256), dim3(256), 0, c.stream, (int64_t)N * kH,
                     (const float*)dH, (const uint16_t*)m.H[0].get(), dZ);
  hipLaunchKernelGGL((k_dw_partial<kF, kH>), dim3(chunks), dim3(kT), 0, c.stream, N,
                     (const uint16_t*)m.feat.get(), (const uint16_t*)dZ, P);
  hipLaunchKernelGGL(k_sum_chunks, dim3((kF * kH + 255) / 256), dim3(256), 0, c.stream, chunks, kF * kH,
                     (const float*)P, m.grad.get() + kW0);
  colsum(dZ, m.grad.get() + kB0);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(c.stream));
  return loss;
}

double mpnn_train_step(Context& c, float lr) {
  Mpnn& m = c.mpnn;
  const double loss = mpnn_loss_grad(c);
  ++m.step;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)m.step), c2 = 1.f - std::pow(b2, (float)m.step);
  hipLaunchKernelGGL(k_adam, dim3((kParams + 255) / 256), dim3(256), 0, c.stream, kParams, m.theta.get(),
                     (const float*)m.grad.get(), m.am.get(), m.av.get(), lr, b1, b2, 1e-8f, c1, c2);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(c.stream));
  m.built = false;
  return loss;
}

void mpnn_download(Context& c, int what, float* out, int64_t count) {
  Mpnn& m = c.mpnn;
  HIP_CHECK(hipStreamSynchronize(c.stream));  // the arrays are written on the library stream
  const float* src = nullptr;
  int64_t n = 0;
  DBuf<float> tmp;
  switch (what) {
    case 0:  // features (FP32 of the BF16 values), N x 16
    case 1: {  // final hidden state, N x 64
      const DBuf<uint16_t>& b = what == 0 ? m.feat : m.H[kRounds];
      n = (int64_t)b.n;
      if (count != n) throw Error(CSTFEM_ERR_ARGUMENT, "mpnn_download: size mismatch");
      std::vector<uint16_t> h(n);
      HIP_CHECK(hipMemcpy(h.data(), b.get(), sizeof(uint16_t) * n, hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < n; ++i) out[i] = bf2f(h[i]);
      return;
    }
    case 2:
      src = m.ghat.get();
      n = (int64_t)m.ghat.n;
      break;
    case 3:
      src = m.grad.get();
      n = kParams;
      break;
    case 4:  // training residuals r_k, npairs x n
    case 5:  // training errors e_k, npairs x n
      src = what == 4 ? m.R.get() : m.E.get();
      n = (int64_t)m.npairs * 2 * c.nb;
      break;
    default:
      throw Error(CSTFEM_ERR_ARGUMENT, "mpnn_download: unknown array");
  }
  if (count != n) throw Error(CSTFEM_ERR_ARGUMENT, "mpnn_download: size mismatch");
  HIP_CHECK(hipMemcpy(out, src, sizeof(float) * n, hipMemcpyDeviceToHost));
}

}  // namespace cstfem_hip
