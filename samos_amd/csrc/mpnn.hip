// Warm-start network for batched FIRE relaxation, trained and evaluated on the
// GPU with bfloat16 MFMA matrix products (FP32 accumulation).
//
// The network predicts, for each particle of a 2-D soft-disk configuration,
// the displacement FIRE relaxation applies to it (no counterpart in the
// reference, which relaxes with IntegratorFIRE only,
// src/integrators/integrator_fire.cpp:34-148).  It runs on a replica ensemble:
// every replica is one configuration and the ensemble's neighbour list (pairs
// within rcut + pad) is the message-passing graph.  With H = 64:
//
//   edges e = (i -> j), d_e = min-image(x_j - x_i), rho_e = |d_e|, dhat_e = d_e / rho_e,
//   ov_e = r_i + r_j - rho_e, f_e = [rho_e / r_l, ov_e, r_i - 1, r_j - 1]
//   u_i = [r_i - 1, sum_e max(ov_e, 0), #{ov_e > 0} / 6]
//   h_i = silu(u_i W0 + b0)
//   3 x: Z_e = [h_i, h_j, f_e] We + be;  A_i = (1/8) sum_{e in row i} silu(Z_e);
//        h_i += silu([h_i, A_i] Wn + bn)
//   U_e = silu([h_i, h_j, f_e] Wo + bo);  w_e = U_e . v + c;  delta_i = sum_e w_e dhat_e
//   loss = (1 / 2N) sum_i |delta_i - delta*_i|^2
//
// Matrix operands ([h_i, h_j, f_e] rows, [h_i, A_i] rows, weights, and the
// backward's output gradients) are rounded to bfloat16 (nearest even); every
// product accumulates in FP32 on v_mfma_f32_16x16x32_bf16.  Sums over a row of
// edges run in list order, so forward and backward are deterministic.  The
// flat parameter order is: W0 [32 x 64], b0, then per layer We [160 x 64], be,
// Wn [128 x 64], bn, then Wo [160 x 64], bo, v [64], c (rows beyond the used
// inputs are zero padding).
#include "ensemble_impl.hpp"
#include "kernels.hpp"
#include "samos_hip.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

namespace samos {
namespace {

constexpr int H = 64;
constexpr int KN = 32;    // node-feature operand width
constexpr int KE = 160;   // [h_i, h_j, f_e] operand width
constexpr int KU = 128;   // [h_i, A_i] operand width
constexpr int NLAYER = 3;
constexpr float kAgg = 0.125f;  // A_i = kAgg * sum over row i of the messages
constexpr int kThreads = 256;
constexpr int kTnRows = 2048;  // rows of one weight-gradient workgroup

typedef uint16_t bf16_t;
typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f4;

__device__ __forceinline__ bf16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
}

__device__ __forceinline__ float bf2f(bf16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ float silu(float z) { return z / (1.0f + expf(-z)); }

__device__ __forceinline__ float dsilu(float z) {
  const float s = 1.0f / (1.0f + expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

__device__ __forceinline__ bf16x8 load8(const bf16_t* p) {
  const uint4 q = *reinterpret_cast<const uint4*>(p);
  return __builtin_bit_cast(bf16x8, q);
}

// ---------------------------------------------------------------------------
// C[M x NO] = A[M x KI] * B, with B given transposed (Bt [NO x KI], inner
// dimension contiguous) and staged in LDS.  Each wavefront computes 16-row
// tiles.  Lane l of v_mfma_f32_16x16x32_bf16 holds A[l % 16][8 (l / 16) + e],
// B[8 (l / 16) + e][l % 16] and returns C[4 (l / 16) + e][l % 16].
//   EPI 0: Y = C (FP32)
//   EPI 1: Z = C + bias, Y = silu(Z)
//   EPI 2: Z = C + bias, Y += silu(Z)     (residual update)
// ---------------------------------------------------------------------------
template <int KI, int NO, int EPI>
__global__ void __launch_bounds__(kThreads) k_gemm(const bf16_t* __restrict__ A, int M, const bf16_t* __restrict__ Bt,
                                                   const float* __restrict__ bias, float* __restrict__ Z,
                                                   float* __restrict__ Y) {
  static_assert(KI % 32 == 0 && NO % 16 == 0, "tile shape");
  __shared__ __attribute__((aligned(16))) bf16_t sB[NO * KI];
  for (int i = threadIdx.x * 8; i < NO * KI; i += kThreads * 8)
    *reinterpret_cast<uint4*>(sB + i) = *reinterpret_cast<const uint4*>(Bt + i);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int ntile = (M + 15) / 16;
  for (int tile = blockIdx.x * 4 + wave; tile < ntile; tile += gridDim.x * 4) {
    const int r0 = tile * 16;
    f4 acc[NO / 16];
#pragma unroll
    for (int t = 0; t < NO / 16; ++t) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
    const int ra = r0 + li;
#pragma unroll
    for (int kk = 0; kk < KI / 32; ++kk) {
      bf16x8 a;
      if (ra < M) a = load8(A + (size_t)ra * KI + kk * 32 + lk * 8);
      else a = __builtin_bit_cast(bf16x8, uint4{0u, 0u, 0u, 0u});
#pragma unroll
      for (int t = 0; t < NO / 16; ++t) {
        const bf16x8 b = load8(sB + (16 * t + li) * KI + kk * 32 + lk * 8);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
      }
    }
#pragma unroll
    for (int t = 0; t < NO / 16; ++t) {
      const int col = 16 * t + li;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = r0 + 4 * lk + e;
        if (row >= M) continue;
        const size_t o = (size_t)row * NO + col;
        if (EPI == 0) {
          Y[o] = acc[t][e];
        } else {
          const float z = acc[t][e] + bias[col];
          Z[o] = z;
          if (EPI == 1) Y[o] = silu(z);
          else Y[o] += silu(z);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Weight gradient: partial[blk] = sum over the block's rows m of
// A[m][k] G[m][o] (KI x NO) followed by the column sums of G (NO), FP32.
// Rows are staged 32 at a time in LDS; each wavefront owns a set of 16 x 16
// output tiles and reads its MFMA operands transposed from LDS.
// ---------------------------------------------------------------------------
// Edge products C[E x H] = X[E x KE] * B with the operand rows gathered inside
// the kernel: row e of X is [bf(h_src[e]), bf(h_col[e]), bf(f_e), 0 ...]
// (round to nearest even, zero padding to 160 columns), assembled per lane
// from the FP32 node features (8 columns = two 16-byte loads) and edge
// features, so no packed operand matrix is written and read back.  Lane mapping as in k_gemm (chunk c = 4 kk
// + lk of 8 columns: c < 8 source node, c < 16 target node, c = 16 the four
// edge features).  PACK also stores the rows (training's weight gradients read
// them).
//   EPI 1: Z = C + bias, Y = silu(Z)
//   EPI 3: Z = C + bias (the aggregation applies silu itself)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t bf2pack(float lo, float hi) {
  return (uint32_t)f2bf(lo) | ((uint32_t)f2bf(hi) << 16);
}

template <int EPI, bool PACK>
__global__ void __launch_bounds__(kThreads) k_gemm_edges(const float* __restrict__ h, const float* __restrict__ fe,
                                                         const int* __restrict__ src, const int* __restrict__ col,
                                                         int M, const bf16_t* __restrict__ Bt,
                                                         const float* __restrict__ bias, float* __restrict__ Z,
                                                         float* __restrict__ Y, bf16_t* __restrict__ Xout) {
  constexpr int KI = KE, NO = H;
  static_assert(KI == 160 && NO == 64, "edge operand layout");
  __shared__ __attribute__((aligned(16))) bf16_t sB[NO * KI];
  for (int i = threadIdx.x * 8; i < NO * KI; i += kThreads * 8)
    *reinterpret_cast<uint4*>(sB + i) = *reinterpret_cast<const uint4*>(Bt + i);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int ntile = (M + 15) / 16;
  for (int tile = blockIdx.x * 4 + wave; tile < ntile; tile += gridDim.x * 4) {
    const int r0 = tile * 16;
    f4 acc[NO / 16];
#pragma unroll
    for (int t = 0; t < NO / 16; ++t) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
    const int ra = r0 + li;
    const bool live = ra < M;
    const int is = live ? src[ra] : 0, ic = live ? col[ra] : 0;
#pragma unroll
    for (int kk = 0; kk < KI / 32; ++kk) {
      uint4 w{0u, 0u, 0u, 0u};
      if (live) {
        if (kk < 4) {
          const float* p = h + (size_t)(kk < 2 ? is : ic) * H + (kk & 1) * 32 + lk * 8;
          const float4 q0 = *reinterpret_cast<const float4*>(p);
          const float4 q1 = *reinterpret_cast<const float4*>(p + 4);
          w.x = bf2pack(q0.x, q0.y); w.y = bf2pack(q0.z, q0.w);
          w.z = bf2pack(q1.x, q1.y); w.w = bf2pack(q1.z, q1.w);
        } else if (lk == 0) {
          const float4 q = *reinterpret_cast<const float4*>(fe + 4 * (size_t)ra);
          w.x = bf2pack(q.x, q.y); w.y = bf2pack(q.z, q.w);
        }
        if (PACK) *reinterpret_cast<uint4*>(Xout + (size_t)ra * KI + kk * 32 + lk * 8) = w;
      }
      const bf16x8 a = __builtin_bit_cast(bf16x8, w);
#pragma unroll
      for (int t = 0; t < NO / 16; ++t) {
        const bf16x8 b = load8(sB + (16 * t + li) * KI + kk * 32 + lk * 8);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
      }
    }
#pragma unroll
    for (int t = 0; t < NO / 16; ++t) {
      const int c = 16 * t + li;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = r0 + 4 * lk + e;
        if (row >= M) continue;
        const size_t o = (size_t)row * NO + c;
        const float z = acc[t][e] + bias[c];
        Z[o] = z;
        if (EPI == 1) Y[o] = silu(z);
      }
    }
  }
}

// ---------------------------------------------------------------------------
template <int KI, int NO>
__global__ void __launch_bounds__(kThreads) k_gemm_tn(const bf16_t* __restrict__ A, const bf16_t* __restrict__ G,
                                                      int M, float* __restrict__ partial) {
  constexpr int TI = KI / 16, TJ = NO / 16, NT = TI * TJ, PER = NT / 4;
  static_assert(NT % 4 == 0, "output tiles split evenly over the four wavefronts");
  __shared__ __attribute__((aligned(16))) bf16_t sA[32 * KI];
  __shared__ __attribute__((aligned(16))) bf16_t sG[32 * NO];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  f4 acc[PER];
#pragma unroll
  for (int q = 0; q < PER; ++q) acc[q] = f4{0.f, 0.f, 0.f, 0.f};
  float colsum = 0.f;
  const int m_begin = blockIdx.x * kTnRows, m_end = min(M, m_begin + kTnRows);
  for (int m0 = m_begin; m0 < m_end; m0 += 32) {
    __syncthreads();
    for (int i = threadIdx.x * 8; i < 32 * KI; i += kThreads * 8) {
      const int r = m0 + i / KI;
      uint4 v = uint4{0u, 0u, 0u, 0u};
      if (r < m_end) v = *reinterpret_cast<const uint4*>(A + (size_t)r * KI + (i % KI));
      *reinterpret_cast<uint4*>(sA + i) = v;
    }
    for (int i = threadIdx.x * 8; i < 32 * NO; i += kThreads * 8) {
      const int r = m0 + i / NO;
      uint4 v = uint4{0u, 0u, 0u, 0u};
      if (r < m_end) v = *reinterpret_cast<const uint4*>(G + (size_t)r * NO + (i % NO));
      *reinterpret_cast<uint4*>(sG + i) = v;
    }
    __syncthreads();
    if (threadIdx.x < NO)
      for (int k = 0; k < 32; ++k) colsum += bf2f(sG[k * NO + threadIdx.x]);
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int t = wave + 4 * q;
      const int ti = t / TJ, tj = t % TJ;
      bf16x8 a, b;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        a[e] = __builtin_bit_cast(__bf16, sA[(8 * lk + e) * KI + 16 * ti + li]);
        b[e] = __builtin_bit_cast(__bf16, sG[(8 * lk + e) * NO + 16 * tj + li]);
      }
      acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[q], 0, 0, 0);
    }
  }
  float* out = partial + (size_t)blockIdx.x * (KI * NO + NO);
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int t = wave + 4 * q;
    const int ti = t / TJ, tj = t % TJ;
#pragma unroll
    for (int e = 0; e < 4; ++e) out[(16 * ti + 4 * lk + e) * NO + 16 * tj + li] = acc[q][e];
  }
  if (threadIdx.x < NO) out[KI * NO + threadIdx.x] = colsum;
}

// Sum of the per-block partials in block order into the gradient slot.
__global__ void k_reduce_partials(const float* partial, int nblk, int len, float* out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= len) return;
  float s = 0.f;
  for (int b = 0; b < nblk; ++b) s += partial[(size_t)b * len + j];
  out[j] = s;
}

// First level of a two-level reduction: group g sums partial rows
// [g * kGroup, (g + 1) * kGroup) in order.
constexpr int kGroup = 32;
__global__ void k_reduce_groups(const float* partial, int nblk, int len, float* tmp) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int g = blockIdx.y;
  if (j >= len) return;
  const int b1 = min(nblk, (g + 1) * kGroup);
  float s = 0.f;
  for (int b = g * kGroup; b < b1; ++b) s += partial[(size_t)b * len + j];
  tmp[(size_t)g * len + j] = s;
}

// ---- graph construction ----------------------------------------------------
__global__ void k_csr_fill(NListD nl, const int* row, int* col, int* src, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int a = row[s], k = nl.nnb[s];
  for (int q = 0; q < k; ++q) {
    col[a + q] = nl.nbr[nbr_at(nl, q, s)];
    src[a + q] = s;
  }
}

// rev[e] = index of the reverse edge (j -> i) of e = (i -> j).
__global__ void k_csr_rev(const int* row, const int* col, int* rev, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  for (int e = row[s]; e < row[s + 1]; ++e) {
    const int t = col[e];
    int r = -1;
    for (int q = row[t]; q < row[t + 1]; ++q)
      if (col[q] == s) { r = q; break; }
    rev[e] = r;
  }
}

// Edge geometry and features, node features (bf16 operand rows of width 32).
__global__ void k_geom(EState e, int b, const RepD* __restrict__ reps, const int* row, const int* col, float* dh,
                       float* fe, bf16_t* u, int n, float inv_rl) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const RepD& R = reps[e.rep[s]];
  const D4 pi = e.p[b][s];
  float S = 0.f, C = 0.f;
  for (int q = row[s]; q < row[s + 1]; ++q) {
    const D4 pj = e.p[b][col[q]];
    double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
    min_image(R.box, dx, dy, dz);
    const double rho = sqrt(dx * dx + dy * dy);
    const float ov = (float)(pi.w + pj.w - rho);
    dh[2 * q] = (float)(dx / rho);
    dh[2 * q + 1] = (float)(dy / rho);
    fe[4 * q] = (float)rho * inv_rl;
    fe[4 * q + 1] = ov;
    fe[4 * q + 2] = (float)(pi.w - 1.0);
    fe[4 * q + 3] = (float)(pj.w - 1.0);
    S += fmaxf(ov, 0.f);
    C += ov > 0.f ? 1.f : 0.f;
  }
  bf16_t* o = u + (size_t)s * KN;
  o[0] = f2bf((float)(pi.w - 1.0));
  o[1] = f2bf(S);
  o[2] = f2bf(C / 6.0f);
  for (int k = 3; k < KN; ++k) o[k] = 0;
}

// Node operand rows [bf(h_i), bf(A_i)].
__global__ void k_pack_nodes(const float* h, const float* A, bf16_t* X, int n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)n * KU) return;
  const int i = (int)(t / KU), c = (int)(t % KU);
  X[t] = f2bf(c < H ? h[(size_t)i * H + c] : A[(size_t)i * H + c - H]);
}

// A_i = kAgg * sum over row i of silu(Z_e) (row order).
__global__ void k_aggregate(const float* Z, const int* row, float* A, int n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)n * H) return;
  const int i = (int)(t / H), c = (int)(t % H);
  float s = 0.f;
  for (int e = row[i]; e < row[i + 1]; ++e) s += silu(Z[(size_t)e * H + c]);
  A[t] = kAgg * s;
}

// w_e = U_e . v + c
__global__ void k_head_out(const float* U, const float* v, const float* c, float* w, int E) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  float s = 0.f;
  const float* u = U + (size_t)e * H;
  for (int k = 0; k < H; ++k) s += u[k] * v[k];
  w[e] = s + c[0];
}

__global__ void k_delta(const float* w, const float* dh, const int* row, float* delta, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float x = 0.f, y = 0.f;
  for (int e = row[i]; e < row[i + 1]; ++e) {
    x += w[e] * dh[2 * e];
    y += w[e] * dh[2 * e + 1];
  }
  delta[2 * i] = x;
  delta[2 * i + 1] = y;
}

// ---- backward ---------------------------------------------------------------
// g_delta = (delta - target) / N and per-block partial sums of the squared error.
__global__ void k_loss(const float* delta, const float* target, float* gdelta, int n, float inv_n, double* partial) {
  __shared__ double red[kThreads];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double sq = 0.0;
  if (i < n) {
    const float dx = delta[2 * i] - target[2 * i], dy = delta[2 * i + 1] - target[2 * i + 1];
    gdelta[2 * i] = dx * inv_n;
    gdelta[2 * i + 1] = dy * inv_n;
    sq = (double)dx * dx + (double)dy * dy;
  }
  red[threadIdx.x] = sq;
  __syncthreads();
  for (int o = kThreads / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// g_w per edge and the head's output gradient bf(g_w v . silu'(Z_o)).
__global__ void k_head_back(const float* gdelta, const float* dh, const int* src, const float* Zo, const float* v,
                            float* gw, bf16_t* gZ, int E) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)E * H) return;
  const int e = (int)(t / H), c = (int)(t % H);
  const int i = src[e];
  const float g = gdelta[2 * i] * dh[2 * e] + gdelta[2 * i + 1] * dh[2 * e + 1];
  if (c == 0) gw[e] = g;
  gZ[t] = f2bf(g * v[c] * dsilu(Zo[t]));
}

// Partial sums over chunks of kVRows edges of g_w U (64 columns) and g_w
// (column 64): four 64-lane groups take every fourth edge, then combine in a
// fixed order.
constexpr int kVRows = 256;
__global__ void __launch_bounds__(kThreads) k_head_vgrad(const float* gw, const float* U, int E, float* partial) {
  __shared__ float red[4][H + 1];
  const int c = threadIdx.x & 63, grp = threadIdx.x >> 6;
  const int e0 = blockIdx.x * kVRows, e1 = min(E, e0 + kVRows);
  float s = 0.f, sw = 0.f;
  for (int e = e0 + grp; e < e1; e += 4) {
    const float g = gw[e];
    s += g * U[(size_t)e * H + c];
    sw += g;
  }
  red[grp][c] = s;
  if (c == 0) red[grp][H] = sw;
  __syncthreads();
  if (threadIdx.x <= H)
    partial[(size_t)blockIdx.x * (H + 1) + threadIdx.x] =
        ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) + red[3][threadIdx.x];
}

// gh[i] (+)= sum over row i of gX[e][0:64] + gX[rev e][64:128]  (gX rows of KU).
__global__ void k_edge_to_node(const float* gX, const int* row, const int* rev, float* gh, int n, int accumulate) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)n * H) return;
  const int i = (int)(t / H), c = (int)(t % H);
  float s = accumulate ? gh[t] : 0.f;
  for (int e = row[i]; e < row[i + 1]; ++e) {
    s += gX[(size_t)e * KU + c];
    if (rev[e] >= 0) s += gX[(size_t)rev[e] * KU + H + c];
  }
  gh[t] = s;
}

// g = bf(gin . silu'(Z)) for node rows.
__global__ void k_act_back(const float* gin, const float* Z, bf16_t* g, long len) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= len) return;
  g[t] = f2bf(gin[t] * dsilu(Z[t]));
}

// gh += gXn[:, 0:64]; gA = gXn[:, 64:128]
__global__ void k_node_split(const float* gXn, float* gh, float* gA, int n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)n * H) return;
  const int i = (int)(t / H), c = (int)(t % H);
  gh[t] += gXn[(size_t)i * KU + c];
  gA[t] = gXn[(size_t)i * KU + H + c];
}

// gZe = bf(gA[src e] . silu'(Z_e))
__global__ void k_edge_back(const float* gA, const int* src, const float* Ze, bf16_t* gZ, int E) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)E * H) return;
  const int e = (int)(t / H), c = (int)(t % H);
  gZ[t] = f2bf(kAgg * gA[(size_t)src[e] * H + c] * dsilu(Ze[t]));
}

// ---- parameters ---------------------------------------------------------------
// bf16 copies of one weight matrix W [K x 64]: Wt [64 x K] (forward operand)
// and Wk [K x 64] (input-gradient operand).
__global__ void k_weights_bf16(const float* W, int K, bf16_t* Wt, bf16_t* Wk) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= K * H) return;
  const int k = t / H, o = t % H;
  const bf16_t q = f2bf(W[t]);
  Wk[t] = q;
  Wt[o * K + k] = q;
}

__global__ void k_adam(float* p, const float* g, float* m, float* v, int len, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---- dataset and warm start -------------------------------------------------
// Start positions, radii and box sides of every replica, in slot order.
__global__ void k_take_start(EState e, int b, const RepD* __restrict__ reps, double* x0, double* rad, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const D4 q = e.p[b][s];
  x0[2 * s] = q.x;
  x0[2 * s + 1] = q.y;
  rad[s] = q.w;
}

// target = min-image(x_relaxed - x0) appended at `base`.
__global__ void k_take_target(EState e, int b, const RepD* __restrict__ reps, const double* x0, float* tgt, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const RepD& R = reps[e.rep[s]];
  const D4 q = e.p[b][s];
  double dx = q.x - x0[2 * s], dy = q.y - x0[2 * s + 1], dz = 0.0;
  min_image(R.box, dx, dy, dz);
  tgt[2 * s] = (float)dx;
  tgt[2 * s + 1] = (float)dy;
}

// Replica r of a work ensemble <- dataset configuration ids[r].
__global__ void k_load_configs(EState e, int b, const RepD* __restrict__ reps, const double* x0, const double* rad,
                               const float* tgt, const int* ids, int per, float* target, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int r = e.rep[s];
  const RepD& R = reps[r];
  const size_t src = (size_t)ids[r] * per + (s - R.n0);
  D4 q = e.p[b][s];
  q.x = x0[2 * src];
  q.y = x0[2 * src + 1];
  q.w = rad[src];
  e.p[b][s] = q;
  e.p[b ^ 1][s].w = q.w;
  target[2 * s] = tgt[2 * src];
  target[2 * s + 1] = tgt[2 * src + 1];
  e.vx[s] = 0.0; e.vy[s] = 0.0; e.vz[s] = 0.0;
}

__global__ void k_apply_delta(EState e, int b, const float* delta, double scale, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  D4 q = e.p[b][s];
  q.x += scale * (double)delta[2 * s];
  q.y += scale * (double)delta[2 * s + 1];
  e.p[b][s] = q;
}

int blocks(long n) { return (int)std::max<long>(1, (n + kThreads - 1) / kThreads); }

struct Block {
  long off;   // offset in the flat parameter vector
  int K;      // rows of the weight (0 for vectors)
};

}  // namespace
}  // namespace samos

using namespace samos;

struct samos_mpnn {
  std::string err;
  int device = 0;
  long nparam = 0;
  // flat FP32 parameters, gradient, Adam moments (device)
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  long adam_t = 0;
  // weight blocks (offsets) and their bf16 copies
  Block W0, b0, We[NLAYER], be[NLAYER], Wn[NLAYER], bn[NLAYER], Wo, bo, vv, cc;
  bf16_t *W0t = nullptr, *W0k = nullptr, *Wet[NLAYER] = {}, *Wek[NLAYER] = {}, *Wnt[NLAYER] = {},
         *Wnk[NLAYER] = {}, *Wot = nullptr, *Wok = nullptr;
  bool bf_stale = true;
  // graph and activations, grown on demand
  long capN = 0, capE = 0;
  int curN = 0, curE = 0;
  bool train_bufs = false;
  int *row = nullptr, *col = nullptr, *src = nullptr, *rev = nullptr;
  float *dh = nullptr, *fe = nullptr;
  bf16_t* u = nullptr;
  float *Z0 = nullptr, *h[NLAYER + 1] = {}, *Zn[NLAYER] = {}, *A = nullptr, *delta = nullptr, *target = nullptr;
  bf16_t *Xn[NLAYER] = {}, *Xe[NLAYER + 1] = {};
  float *Ze[NLAYER + 1] = {}, *U = nullptr, *w = nullptr;
  // backward
  float *gdelta = nullptr, *gw = nullptr, *gX = nullptr, *gh = nullptr, *gXn = nullptr, *gA = nullptr;
  bf16_t *gZ = nullptr, *gZn = nullptr;
  float* partial = nullptr;
  long cap_partial = 0;
  float* partial2 = nullptr;     // first level of two-level reductions
  long cap_partial2 = 0;
  double* loss_part = nullptr;
  void* scan_tmp = nullptr;
  size_t scan_bytes = 0;
  // dataset (device): start positions, radii, targets; host: box sides
  double *ds_x0 = nullptr, *ds_r = nullptr;
  float* ds_t = nullptr;
  long ds_cap = 0;
  int ds_per = 0, ds_count = 0;
  std::vector<double> ds_L;           // Lx, Ly, Lz per configuration
  double *st_x0 = nullptr, *st_r = nullptr;  // staging of the current begin()
  long st_cap = 0;
  std::vector<double> st_L;
  int* d_ids = nullptr;
  int cap_ids = 0;
  // timing of the MFMA products (HIP events around each product launch)
  bool prof = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_gemm;
  double gemm_flops = 0.0;
  std::vector<void*> owned;
};

namespace {

thread_local std::string g_mpnn_err;

#define MCHK(expr)                                                                \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) {                                                       \
      M->err = std::string("HIP error: ") + hipGetErrorString(e_) + " at " #expr; \
      return -2;                                                                  \
    }                                                                             \
  } while (0)

template <class T>
hipError_t m_alloc(samos_mpnn* M, T** p, size_t count) {
  void* q = nullptr;
  hipError_t r = hipMalloc(&q, sizeof(T) * std::max<size_t>(count, 1));
  if (r == hipSuccess) {
    M->owned.push_back(q);
    *p = static_cast<T*>(q);
  }
  return r;
}

template <class T>
void m_free(samos_mpnn* M, T*& p) {
  if (!p) return;
  auto it = std::find(M->owned.begin(), M->owned.end(), (void*)p);
  if (it != M->owned.end()) M->owned.erase(it);
  hipFree(p);
  p = nullptr;
}

void layout(samos_mpnn* M) {
  long o = 0;
  auto take = [&](Block& b, int K, int cols) {
    b.off = o;
    b.K = K;
    o += (long)(K ? K : 1) * cols;
  };
  take(M->W0, KN, H);
  take(M->b0, 0, H);
  for (int l = 0; l < NLAYER; ++l) {
    take(M->We[l], KE, H);
    take(M->be[l], 0, H);
    take(M->Wn[l], KU, H);
    take(M->bn[l], 0, H);
  }
  take(M->Wo, KE, H);
  take(M->bo, 0, H);
  take(M->vv, 0, H);
  take(M->cc, 0, 1);
  M->nparam = o;
}

int refresh_bf16(samos_mpnn* M, hipStream_t st) {
  if (!M->bf_stale) return 0;
  auto conv = [&](const Block& b, bf16_t* Wt, bf16_t* Wk) {
    hipLaunchKernelGGL(k_weights_bf16, dim3(blocks((long)b.K * H)), dim3(kThreads), 0, st, M->p + b.off, b.K, Wt, Wk);
  };
  conv(M->W0, M->W0t, M->W0k);
  for (int l = 0; l < NLAYER; ++l) {
    conv(M->We[l], M->Wet[l], M->Wek[l]);
    conv(M->Wn[l], M->Wnt[l], M->Wnk[l]);
  }
  conv(M->Wo, M->Wot, M->Wok);
  MCHK(hipGetLastError());
  M->bf_stale = false;
  return 0;
}

// Buffers for n nodes and E edges (train: keep every layer's operands).
int ensure(samos_mpnn* M, int n, int E, bool train) {
  const bool grow = n > M->capN || E > M->capE || (train && !M->train_bufs);
  if (!grow) return 0;
  // nothing queued may still use the buffers about to be replaced
  MCHK(hipDeviceSynchronize());
  const long N = std::max<long>(n, M->capN), EE = std::max<long>(E, M->capE) + E / 8 + 64;
  m_free(M, M->row); m_free(M, M->col); m_free(M, M->src); m_free(M, M->rev);
  m_free(M, M->dh); m_free(M, M->fe); m_free(M, M->u); m_free(M, M->Z0); m_free(M, M->A);
  m_free(M, M->delta); m_free(M, M->target); m_free(M, M->U); m_free(M, M->w);
  m_free(M, M->gdelta); m_free(M, M->gw); m_free(M, M->gX); m_free(M, M->gh); m_free(M, M->gXn);
  m_free(M, M->gA); m_free(M, M->gZ); m_free(M, M->gZn); m_free(M, M->loss_part);
  for (int l = 0; l <= NLAYER; ++l) { m_free(M, M->h[l]); m_free(M, M->Xe[l]); m_free(M, M->Ze[l]); }
  for (int l = 0; l < NLAYER; ++l) { m_free(M, M->Zn[l]); m_free(M, M->Xn[l]); }
  MCHK(m_alloc(M, &M->row, N + 1));
  MCHK(m_alloc(M, &M->col, EE));
  MCHK(m_alloc(M, &M->src, EE));
  MCHK(m_alloc(M, &M->rev, EE));
  MCHK(m_alloc(M, &M->dh, 2 * EE));
  MCHK(m_alloc(M, &M->fe, 4 * EE));
  MCHK(m_alloc(M, &M->u, N * KN));
  MCHK(m_alloc(M, &M->Z0, N * H));
  MCHK(m_alloc(M, &M->A, N * H));
  MCHK(m_alloc(M, &M->delta, 2 * N));
  MCHK(m_alloc(M, &M->target, 2 * N));
  MCHK(m_alloc(M, &M->U, EE * H));
  MCHK(m_alloc(M, &M->w, EE));
  for (int l = 0; l <= NLAYER; ++l) MCHK(m_alloc(M, &M->h[l], N * H));
  const int layers = train ? NLAYER + 1 : 1;
  for (int l = 0; l < layers; ++l) {
    if (train) MCHK(m_alloc(M, &M->Xe[l], EE * KE));  // operand rows kept for the weight gradients
    MCHK(m_alloc(M, &M->Ze[l], EE * H));
  }
  for (int l = 0; l < (train ? NLAYER : 1); ++l) {
    MCHK(m_alloc(M, &M->Zn[l], N * H));
    MCHK(m_alloc(M, &M->Xn[l], N * KU));
  }
  if (train) {
    MCHK(m_alloc(M, &M->gdelta, 2 * N));
    MCHK(m_alloc(M, &M->gw, EE));
    MCHK(m_alloc(M, &M->gX, EE * KU));
    MCHK(m_alloc(M, &M->gh, N * H));
    MCHK(m_alloc(M, &M->gXn, N * KU));
    MCHK(m_alloc(M, &M->gA, N * H));
    MCHK(m_alloc(M, &M->gZ, EE * H));
    MCHK(m_alloc(M, &M->gZn, N * H));
    MCHK(m_alloc(M, &M->loss_part, N / kThreads + 2));
    // per-workgroup partials of the largest weight gradient (edge rows, KE x H + H)
    m_free(M, M->partial);
    const long rows = std::max(EE, N);
    M->cap_partial = (rows / kTnRows + 2) * ((long)KE * H + H);
    MCHK(m_alloc(M, &M->partial, (size_t)M->cap_partial));
    m_free(M, M->partial2);
    M->cap_partial2 = (rows / kVRows / kGroup + 2) * ((long)KE * H + H);
    MCHK(m_alloc(M, &M->partial2, (size_t)M->cap_partial2));
  }
  M->capN = N;
  M->capE = EE;
  M->train_bufs = train;
  return 0;
}

std::pair<hipEvent_t, hipEvent_t> gprof_begin(samos_mpnn* M, hipStream_t st) {
  std::pair<hipEvent_t, hipEvent_t> ev{nullptr, nullptr};
  if (!M->prof) return ev;
  hipEventCreate(&ev.first);
  hipEventCreate(&ev.second);
  hipEventRecord(ev.first, st);
  return ev;
}

void gprof_end(samos_mpnn* M, hipStream_t st, std::pair<hipEvent_t, hipEvent_t> ev, double flops) {
  if (!ev.first) return;
  hipEventRecord(ev.second, st);
  M->ev_gemm.push_back(ev);
  M->gemm_flops += flops;
}

template <int KI, int NO, int EPI>
void gemm(samos_mpnn* M, hipStream_t st, const bf16_t* A, int rows, const bf16_t* Bt, const float* bias, float* Z,
          float* Y) {
  const int tiles = (rows + 15) / 16;
  const int grid = std::max(1, std::min((tiles + 3) / 4, 4096));
  auto ev = gprof_begin(M, st);
  hipLaunchKernelGGL((k_gemm<KI, NO, EPI>), dim3(grid), dim3(kThreads), 0, st, A, rows, Bt, bias, Z, Y);
  gprof_end(M, st, ev, 2.0 * rows * KI * NO);
}

// Edge products with operand rows gathered in the kernel (k_gemm_edges);
// Xout (training) receives the rows.
template <int EPI>
void gemm_edges(samos_mpnn* M, hipStream_t st, const float* h, int rows, const bf16_t* Bt, const float* bias,
                float* Z, float* Y, bf16_t* Xout) {
  const int tiles = (rows + 15) / 16;
  const int grid = std::max(1, std::min((tiles + 3) / 4, 4096));
  auto ev = gprof_begin(M, st);
  if (Xout)
    hipLaunchKernelGGL((k_gemm_edges<EPI, true>), dim3(grid), dim3(kThreads), 0, st, h, M->fe, M->src, M->col, rows,
                       Bt, bias, Z, Y, Xout);
  else
    hipLaunchKernelGGL((k_gemm_edges<EPI, false>), dim3(grid), dim3(kThreads), 0, st, h, M->fe, M->src, M->col, rows,
                       Bt, bias, Z, Y, Xout);
  gprof_end(M, st, ev, 2.0 * rows * KE * H);
}

// out[0:len] = sum over the nblk rows of M->partial, in row order (two levels
// when there are many rows).
int reduce_rows(samos_mpnn* M, hipStream_t st, int nblk, int len, float* out) {
  if (nblk <= 2 * kGroup) {
    hipLaunchKernelGGL(k_reduce_partials, dim3(blocks(len)), dim3(kThreads), 0, st, M->partial, nblk, len, out);
    return 0;
  }
  const int groups = (nblk + kGroup - 1) / kGroup;
  if ((long)groups * len > M->cap_partial2) {
    M->err = "internal: reduction buffer too small";
    return -1;
  }
  hipLaunchKernelGGL(k_reduce_groups, dim3(blocks(len), groups), dim3(kThreads), 0, st, M->partial, nblk, len,
                     M->partial2);
  hipLaunchKernelGGL(k_reduce_partials, dim3(blocks(len)), dim3(kThreads), 0, st, M->partial2, groups, len, out);
  return 0;
}

// Weight gradient of one layer into grad + off (matrix, then bias).
template <int KI, int NO>
int gemm_tn(samos_mpnn* M, hipStream_t st, const bf16_t* A, const bf16_t* G, int rows, long off) {
  const int nblk = std::max(1, (rows + kTnRows - 1) / kTnRows);
  const long len = (long)KI * NO + NO;
  if ((long)nblk * len > M->cap_partial) {
    M->err = "internal: weight-gradient partial buffer too small";
    return -1;
  }
  auto ev = gprof_begin(M, st);
  hipLaunchKernelGGL((k_gemm_tn<KI, NO>), dim3(nblk), dim3(kThreads), 0, st, A, G, rows, M->partial);
  gprof_end(M, st, ev, 2.0 * rows * KI * NO);
  return reduce_rows(M, st, nblk, (int)len, M->g + off);
}

// Graph (CSR of the neighbour list), features and the forward pass on the
// ensemble's current configuration.
int forward(samos_mpnn* M, samos_ens* E, bool train) {
  hipStream_t st = E->st;
  if (!E->has_nlist) { M->err = "the ensemble has no neighbour list"; return -1; }
  const int n = E->n;
  // row offsets: row[0] = 0, row[i + 1] = row[i] + nnb[i]
  size_t need = 0;
  hipcub::DeviceScan::InclusiveSum(nullptr, need, E->nl.nnb, (int*)nullptr, n, st);
  if (need > M->scan_bytes) {
    char* t = static_cast<char*>(M->scan_tmp);
    m_free(M, t);
    MCHK(m_alloc(M, &t, need));
    M->scan_tmp = t;
    M->scan_bytes = need;
  }
  if (int rc = ensure(M, n, M->capE > 0 ? (int)M->capE : 1, train)) return rc;
  MCHK(hipMemsetAsync(M->row, 0, sizeof(int), st));
  MCHK(hipcub::DeviceScan::InclusiveSum(M->scan_tmp, need, E->nl.nnb, M->row + 1, n, st));
  int Etot = 0;
  MCHK(hipMemcpyAsync(&Etot, M->row + n, sizeof(int), hipMemcpyDeviceToHost, st));
  MCHK(hipStreamSynchronize(st));
  if (Etot > M->capE) {
    if (int rc = ensure(M, n, Etot, train)) return rc;
    // the buffers were reallocated: recompute the offsets
    MCHK(hipMemsetAsync(M->row, 0, sizeof(int), st));
    MCHK(hipcub::DeviceScan::InclusiveSum(M->scan_tmp, need, E->nl.nnb, M->row + 1, n, st));
  }
  M->curN = n;
  M->curE = Etot;
  if (int rc = refresh_bf16(M, st)) return rc;
  const float inv_rl = (float)(1.0 / (E->rcut + E->pad));
  hipLaunchKernelGGL(k_csr_fill, dim3(blocks(n)), dim3(kThreads), 0, st, E->nl, M->row, M->col, M->src, n);
  hipLaunchKernelGGL(k_csr_rev, dim3(blocks(n)), dim3(kThreads), 0, st, M->row, M->col, M->rev, n);
  hipLaunchKernelGGL(k_geom, dim3(blocks(n)), dim3(kThreads), 0, st, E->e, E->cur, E->d_rep, M->row, M->col, M->dh,
                     M->fe, M->u, n, inv_rl);
  float* P = M->p;
  gemm<KN, H, 1>(M, st, M->u, n, M->W0t, P + M->b0.off, M->Z0, M->h[0]);
  for (int l = 0; l < NLAYER; ++l) {
    float* Ze = M->Ze[train ? l : 0];
    bf16_t* Xn = M->Xn[train ? l : 0];
    float* Zn = M->Zn[train ? l : 0];
    gemm_edges<3>(M, st, M->h[l], Etot, M->Wet[l], P + M->be[l].off, Ze, nullptr, train ? M->Xe[l] : nullptr);
    hipLaunchKernelGGL(k_aggregate, dim3(blocks((long)n * H)), dim3(kThreads), 0, st, Ze, M->row, M->A, n);
    hipLaunchKernelGGL(k_pack_nodes, dim3(blocks((long)n * KU)), dim3(kThreads), 0, st, M->h[l], M->A, Xn, n);
    MCHK(hipMemcpyAsync(M->h[l + 1], M->h[l], sizeof(float) * (size_t)n * H, hipMemcpyDeviceToDevice, st));
    gemm<KU, H, 2>(M, st, Xn, n, M->Wnt[l], P + M->bn[l].off, Zn, M->h[l + 1]);
  }
  float* Zo = M->Ze[train ? NLAYER : 0];
  gemm_edges<1>(M, st, M->h[NLAYER], Etot, M->Wot, P + M->bo.off, Zo, M->U, train ? M->Xe[NLAYER] : nullptr);
  hipLaunchKernelGGL(k_head_out, dim3(blocks(Etot)), dim3(kThreads), 0, st, M->U, P + M->vv.off, P + M->cc.off,
                     M->w, Etot);
  hipLaunchKernelGGL(k_delta, dim3(blocks(n)), dim3(kThreads), 0, st, M->w, M->dh, M->row, M->delta, n);
  MCHK(hipGetLastError());
  return 0;
}

// Loss and gradient (into M->g) against M->target for the last forward pass.
int backward(samos_mpnn* M, samos_ens* E, double* loss) {
  hipStream_t st = E->st;
  const int n = M->curN, Et = M->curE;
  float* P = M->p;
  const int nb = blocks(n);
  hipLaunchKernelGGL(k_loss, dim3(nb), dim3(kThreads), 0, st, M->delta, M->target, M->gdelta, n, 1.0f / n,
                     M->loss_part);
  // head
  hipLaunchKernelGGL(k_head_back, dim3(blocks((long)Et * H)), dim3(kThreads), 0, st, M->gdelta, M->dh, M->src,
                     M->Ze[NLAYER], P + M->vv.off, M->gw, M->gZ, Et);
  {
    const int nblk = std::max(1, (Et + kVRows - 1) / kVRows);
    if ((long)nblk * (H + 1) > M->cap_partial) {
      M->err = "internal: head partial buffer too small";
      return -1;
    }
    hipLaunchKernelGGL(k_head_vgrad, dim3(nblk), dim3(kThreads), 0, st, M->gw, M->U, Et, M->partial);
    // v then c are adjacent in the flat order
    if (int rc = reduce_rows(M, st, nblk, H + 1, M->g + M->vv.off)) return rc;
  }
  if (int rc = gemm_tn<KE, H>(M, st, M->Xe[NLAYER], M->gZ, Et, M->Wo.off)) return rc;
  // input gradients of the [h_i, h_j] columns only (the edge features need none)
  gemm<H, KU, 0>(M, st, M->gZ, Et, M->Wok, nullptr, nullptr, M->gX);
  hipLaunchKernelGGL(k_edge_to_node, dim3(blocks((long)n * H)), dim3(kThreads), 0, st, M->gX, M->row, M->rev, M->gh,
                     n, 0);
  for (int l = NLAYER - 1; l >= 0; --l) {
    hipLaunchKernelGGL(k_act_back, dim3(blocks((long)n * H)), dim3(kThreads), 0, st, M->gh, M->Zn[l], M->gZn,
                       (long)n * H);
    if (int rc = gemm_tn<KU, H>(M, st, M->Xn[l], M->gZn, n, M->Wn[l].off)) return rc;
    gemm<H, KU, 0>(M, st, M->gZn, n, M->Wnk[l], nullptr, nullptr, M->gXn);
    hipLaunchKernelGGL(k_node_split, dim3(blocks((long)n * H)), dim3(kThreads), 0, st, M->gXn, M->gh, M->gA, n);
    hipLaunchKernelGGL(k_edge_back, dim3(blocks((long)Et * H)), dim3(kThreads), 0, st, M->gA, M->src, M->Ze[l], M->gZ,
                       Et);
    if (int rc = gemm_tn<KE, H>(M, st, M->Xe[l], M->gZ, Et, M->We[l].off)) return rc;
    gemm<H, KU, 0>(M, st, M->gZ, Et, M->Wek[l], nullptr, nullptr, M->gX);
    hipLaunchKernelGGL(k_edge_to_node, dim3(blocks((long)n * H)), dim3(kThreads), 0, st, M->gX, M->row, M->rev,
                       M->gh, n, 1);
  }
  hipLaunchKernelGGL(k_act_back, dim3(blocks((long)n * H)), dim3(kThreads), 0, st, M->gh, M->Z0, M->gZn, (long)n * H);
  if (int rc = gemm_tn<KN, H>(M, st, M->u, M->gZn, n, M->W0.off)) return rc;
  MCHK(hipGetLastError());
  if (loss) {
    std::vector<double> part(nb);
    MCHK(hipMemcpyAsync(part.data(), M->loss_part, sizeof(double) * nb, hipMemcpyDeviceToHost, st));
    MCHK(hipStreamSynchronize(st));
    double s = 0.0;
    for (double x : part) s += x;
    *loss = 0.5 * s / n;
  }
  return 0;
}

uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

double unif(uint64_t& s) { return (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace

extern "C" {

const char* samos_mpnn_last_error(const samos_mpnn* M) { return M ? M->err.c_str() : g_mpnn_err.c_str(); }

int samos_mpnn_num_params(void) {
  samos_mpnn tmp;
  layout(&tmp);
  return (int)tmp.nparam;
}

int samos_mpnn_create(samos_mpnn** out, int device, unsigned long long seed) {
  if (!out) { g_mpnn_err = "null output handle"; return -1; }
  *out = nullptr;
  if (hipSetDevice(device) != hipSuccess) { g_mpnn_err = "hipSetDevice failed"; return -2; }
  samos_mpnn* M = new samos_mpnn();
  M->device = device;
  layout(M);
  auto fail = [&](const char* what) {
    g_mpnn_err = std::string("device allocation failed: ") + what;
    samos_mpnn_destroy(M);
    return -2;
  };
  if (m_alloc(M, &M->p, M->nparam) || m_alloc(M, &M->g, M->nparam) || m_alloc(M, &M->m, M->nparam) ||
      m_alloc(M, &M->v, M->nparam))
    return fail("parameters");
  auto wbuf = [&](const Block& b, bf16_t** t, bf16_t** k) {
    return m_alloc(M, t, (size_t)b.K * H) == hipSuccess && m_alloc(M, k, (size_t)b.K * H) == hipSuccess;
  };
  bool ok = wbuf(M->W0, &M->W0t, &M->W0k) && wbuf(M->Wo, &M->Wot, &M->Wok);
  for (int l = 0; l < NLAYER; ++l) ok = ok && wbuf(M->We[l], &M->Wet[l], &M->Wek[l]) && wbuf(M->Wn[l], &M->Wnt[l], &M->Wnk[l]);
  if (!ok) return fail("weights");
  // Uniform(+-sqrt(6 / (fan_in + fan_out))) on the used rows; zero biases and padding.
  std::vector<float> h(M->nparam, 0.f);
  uint64_t s = seed;
  auto fill = [&](const Block& b, int used) {
    const double a = std::sqrt(6.0 / (used + H));
    for (int k = 0; k < used; ++k)
      for (int o = 0; o < H; ++o) h[b.off + (long)k * H + o] = (float)((2.0 * unif(s) - 1.0) * a);
  };
  fill(M->W0, 3);
  for (int l = 0; l < NLAYER; ++l) {
    fill(M->We[l], 2 * H + 4);
    fill(M->Wn[l], KU);
  }
  fill(M->Wo, 2 * H + 4);
  for (int o = 0; o < H; ++o) h[M->vv.off + o] = (float)((2.0 * unif(s) - 1.0) * 0.1);
  if (hipMemcpy(M->p, h.data(), sizeof(float) * M->nparam, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(M->m, 0, sizeof(float) * M->nparam) != hipSuccess ||
      hipMemset(M->v, 0, sizeof(float) * M->nparam) != hipSuccess ||
      hipMemset(M->g, 0, sizeof(float) * M->nparam) != hipSuccess)
    return fail("initialisation");
  *out = M;
  return 0;
}

int samos_mpnn_destroy(samos_mpnn* M) {
  if (!M) return 0;
  hipDeviceSynchronize();
  for (auto& p : M->ev_gemm) { hipEventDestroy(p.first); hipEventDestroy(p.second); }
  for (void* q : M->owned) hipFree(q);
  delete M;
  return 0;
}

int samos_mpnn_get_params(samos_mpnn* M, float* out) {
  MCHK(hipMemcpy(out, M->p, sizeof(float) * M->nparam, hipMemcpyDeviceToHost));
  return 0;
}

int samos_mpnn_set_params(samos_mpnn* M, const float* in) {
  MCHK(hipMemcpy(M->p, in, sizeof(float) * M->nparam, hipMemcpyHostToDevice));
  MCHK(hipMemset(M->m, 0, sizeof(float) * M->nparam));
  MCHK(hipMemset(M->v, 0, sizeof(float) * M->nparam));
  M->adam_t = 0;
  M->bf_stale = true;
  return 0;
}

int samos_mpnn_predict(samos_mpnn* M, samos_ens* E, double* dxdy) {
  if (int rc = forward(M, E, false)) return rc;
  if (dxdy) {
    std::vector<float> d(2 * (size_t)E->n);
    MCHK(hipMemcpyAsync(d.data(), M->delta, sizeof(float) * 2 * E->n, hipMemcpyDeviceToHost, E->st));
    MCHK(hipStreamSynchronize(E->st));
    for (int s = 0; s < E->n; ++s) {
      const int g = E->h_gid[s];
      dxdy[2 * g] = d[2 * s];
      dxdy[2 * g + 1] = d[2 * s + 1];
    }
  }
  MCHK(hipStreamSynchronize(E->st));
  return 0;
}

int samos_mpnn_gradients(samos_mpnn* M, samos_ens* E, const double* target, float* grad, double* loss) {
  if (int rc = forward(M, E, true)) return rc;
  std::vector<float> t(2 * (size_t)E->n);
  for (int s = 0; s < E->n; ++s) {
    const int g = E->h_gid[s];
    t[2 * s] = (float)target[2 * g];
    t[2 * s + 1] = (float)target[2 * g + 1];
  }
  MCHK(hipMemcpyAsync(M->target, t.data(), sizeof(float) * t.size(), hipMemcpyHostToDevice, E->st));
  if (int rc = backward(M, E, loss)) return rc;
  if (grad) {
    MCHK(hipMemcpyAsync(grad, M->g, sizeof(float) * M->nparam, hipMemcpyDeviceToHost, E->st));
    MCHK(hipStreamSynchronize(E->st));
  }
  return 0;
}

int samos_mpnn_warm_start(samos_mpnn* M, samos_ens* E, double scale) {
  if (int rc = forward(M, E, false)) return rc;
  hipLaunchKernelGGL(k_apply_delta, dim3(blocks(E->n)), dim3(kThreads), 0, E->st, E->e, E->cur, M->delta, scale,
                     E->n);
  ens_enforce_async(E);
  MCHK(hipGetLastError());
  if (int rc = ens_build_all(E)) { M->err = E->err; return rc; }
  return 0;
}

int samos_mpnn_dataset_reserve(samos_mpnn* M, int configs, int per_config) {
  if (configs <= 0 || per_config <= 0) { M->err = "dataset needs a positive size"; return -1; }
  if (M->ds_count > 0 && per_config != M->ds_per) { M->err = "dataset configurations differ in size"; return -1; }
  const long cap = (long)configs * per_config;
  if (cap <= M->ds_cap) return 0;
  double *x0 = nullptr, *r = nullptr;
  float* t = nullptr;
  MCHK(m_alloc(M, &x0, 2 * cap));
  MCHK(m_alloc(M, &r, cap));
  MCHK(m_alloc(M, &t, 2 * cap));
  const long used = (long)M->ds_count * M->ds_per;
  if (used) {
    MCHK(hipMemcpy(x0, M->ds_x0, sizeof(double) * 2 * used, hipMemcpyDeviceToDevice));
    MCHK(hipMemcpy(r, M->ds_r, sizeof(double) * used, hipMemcpyDeviceToDevice));
    MCHK(hipMemcpy(t, M->ds_t, sizeof(float) * 2 * used, hipMemcpyDeviceToDevice));
  }
  m_free(M, M->ds_x0); m_free(M, M->ds_r); m_free(M, M->ds_t);
  M->ds_x0 = x0; M->ds_r = r; M->ds_t = t;
  M->ds_cap = cap;
  M->ds_per = per_config;
  return 0;
}

int samos_mpnn_dataset_begin(samos_mpnn* M, samos_ens* E) {
  const int per = E->h_rep[0].n;
  for (const auto& R : E->h_rep)
    if (R.n != per) { M->err = "dataset replicas must have equal particle counts"; return -1; }
  if (M->ds_count > 0 && per != M->ds_per) { M->err = "dataset configurations differ in size"; return -1; }
  if (E->n > M->st_cap) {
    m_free(M, M->st_x0); m_free(M, M->st_r);
    MCHK(m_alloc(M, &M->st_x0, 2 * (size_t)E->n));
    MCHK(m_alloc(M, &M->st_r, (size_t)E->n));
    M->st_cap = E->n;
  }
  hipLaunchKernelGGL(k_take_start, dim3(blocks(E->n)), dim3(kThreads), 0, E->st, E->e, E->cur, E->d_rep, M->st_x0,
                     M->st_r, E->n);
  M->st_L.clear();
  for (const auto& R : E->h_rep) { M->st_L.push_back(R.box.Lx); M->st_L.push_back(R.box.Ly); M->st_L.push_back(R.box.Lz); }
  MCHK(hipStreamSynchronize(E->st));
  return 0;
}

int samos_mpnn_dataset_end(samos_mpnn* M, samos_ens* E) {
  if (M->st_L.size() != 3 * (size_t)E->nrep) { M->err = "dataset_end without a matching dataset_begin"; return -1; }
  const int per = E->h_rep[0].n;
  if (M->ds_per == 0) M->ds_per = per;
  const long need = (long)(M->ds_count + E->nrep) * per;
  if (need > M->ds_cap) { M->err = "dataset capacity exceeded (samos_mpnn_dataset_reserve)"; return -1; }
  const long base = (long)M->ds_count * per;
  MCHK(hipMemcpyAsync(M->ds_x0 + 2 * base, M->st_x0, sizeof(double) * 2 * E->n, hipMemcpyDeviceToDevice, E->st));
  MCHK(hipMemcpyAsync(M->ds_r + base, M->st_r, sizeof(double) * E->n, hipMemcpyDeviceToDevice, E->st));
  hipLaunchKernelGGL(k_take_target, dim3(blocks(E->n)), dim3(kThreads), 0, E->st, E->e, E->cur, E->d_rep, M->st_x0,
                     M->ds_t + 2 * base, E->n);
  MCHK(hipGetLastError());
  MCHK(hipStreamSynchronize(E->st));
  M->ds_L.insert(M->ds_L.end(), M->st_L.begin(), M->st_L.end());
  M->ds_count += E->nrep;
  M->st_L.clear();
  return 0;
}

int samos_mpnn_dataset_size(const samos_mpnn* M) { return M ? M->ds_count : -1; }

int samos_mpnn_train(samos_mpnn* M, samos_ens* W, int steps, double lr, unsigned long long seed, int first_config,
                     int num_configs, double* losses) {
  if (M->ds_count == 0) { M->err = "empty dataset"; return -1; }
  if (!W->has_nlist) { M->err = "the work ensemble has no neighbour list"; return -1; }
  for (const auto& R : W->h_rep)
    if (R.n != M->ds_per) { M->err = "work ensemble replicas must match the dataset configuration size"; return -1; }
  if (num_configs <= 0) num_configs = M->ds_count - first_config;
  if (first_config < 0 || first_config + num_configs > M->ds_count || num_configs < W->nrep) {
    M->err = "training range outside the dataset";
    return -1;
  }
  if (W->nrep > M->cap_ids) {
    m_free(M, M->d_ids);
    MCHK(m_alloc(M, &M->d_ids, W->nrep));
    M->cap_ids = W->nrep;
  }
  uint64_t rs = seed;
  std::vector<int> order(num_configs);
  std::iota(order.begin(), order.end(), first_config);
  size_t pos = order.size();
  std::vector<int> ids(W->nrep);
  for (int step = 0; step < steps; ++step) {
    for (int r = 0; r < W->nrep; ++r) {
      if (pos >= order.size()) {
        for (size_t i = order.size() - 1; i > 0; --i) std::swap(order[i], order[splitmix(rs) % (i + 1)]);
        pos = 0;
      }
      ids[r] = order[pos++];
    }
    for (int r = 0; r < W->nrep; ++r) {
      RepD& R = W->h_rep[r];
      const double* L = &M->ds_L[3 * (size_t)ids[r]];
      BoxD& B = R.box;
      B.Lx = L[0]; B.Ly = L[1]; B.Lz = L[2];
      B.xlo = -0.5 * B.Lx; B.xhi = 0.5 * B.Lx; B.ylo = -0.5 * B.Ly; B.yhi = 0.5 * B.Ly;
      B.zlo = -0.5 * B.Lz; B.zhi = 0.5 * B.Lz;
      const double cut = W->rcut + W->pad;
      R.ncx = (int)(B.Lx / cut); R.ncy = (int)(B.Ly / cut); R.ncz = (int)(B.Lz / cut);
      if (R.ncx < 3 || R.ncy < 3 || R.ncz < 3 || R.ncx > 1024 || R.ncy > 1024 || R.ncz > 1024) {
        M->err = "dataset box does not fit the work ensemble's cell rules";
        return -1;
      }
      R.wx = B.Lx / R.ncx; R.wy = B.Ly / R.ncy; R.wz = B.Lz / R.ncz;
    }
    MCHK(hipMemcpyAsync(W->d_rep, W->h_rep.data(), sizeof(RepD) * W->nrep, hipMemcpyHostToDevice, W->st));
    MCHK(hipMemcpyAsync(M->d_ids, ids.data(), sizeof(int) * W->nrep, hipMemcpyHostToDevice, W->st));
    if (int rc = ensure(M, W->n, M->capE > 0 ? (int)M->capE : 1, true)) return rc;
    hipLaunchKernelGGL(k_load_configs, dim3(blocks(W->n)), dim3(kThreads), 0, W->st, W->e, W->cur, W->d_rep,
                       M->ds_x0, M->ds_r, M->ds_t, M->d_ids, M->ds_per, M->target, W->n);
    MCHK(hipGetLastError());
    if (int rc = ens_build_all(W)) { M->err = W->err; return rc; }
    if (int rc = forward(M, W, true)) return rc;
    // forward() may have grown the buffers: reload the targets
    hipLaunchKernelGGL(k_load_configs, dim3(blocks(W->n)), dim3(kThreads), 0, W->st, W->e, W->cur, W->d_rep,
                       M->ds_x0, M->ds_r, M->ds_t, M->d_ids, M->ds_per, M->target, W->n);
    double loss = 0.0;
    if (int rc = backward(M, W, losses ? &loss : nullptr)) return rc;
    M->adam_t++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)M->adam_t), c2 = 1.f - std::pow(b2, (float)M->adam_t);
    hipLaunchKernelGGL(k_adam, dim3(blocks(M->nparam)), dim3(kThreads), 0, W->st, M->p, M->g, M->m, M->v,
                       (int)M->nparam, (float)lr, b1, b2, 1e-8f, c1, c2);
    MCHK(hipGetLastError());
    M->bf_stale = true;
    if (losses) losses[step] = loss;
  }
  MCHK(hipStreamSynchronize(W->st));
  return 0;
}

// Mean loss of the network over dataset configurations [first, first + count),
// evaluated in batches of the work ensemble's size.
int samos_mpnn_evaluate(samos_mpnn* M, samos_ens* W, int first_config, int num_configs, double* mean_loss) {
  if (num_configs < W->nrep || first_config < 0 || first_config + num_configs > M->ds_count) {
    M->err = "evaluation range outside the dataset";
    return -1;
  }
  if (W->nrep > M->cap_ids) {
    m_free(M, M->d_ids);
    MCHK(m_alloc(M, &M->d_ids, W->nrep));
    M->cap_ids = W->nrep;
  }
  double total = 0.0;
  int batches = 0;
  std::vector<int> ids(W->nrep);
  for (int c0 = first_config; c0 + W->nrep <= first_config + num_configs; c0 += W->nrep) {
    for (int r = 0; r < W->nrep; ++r) {
      ids[r] = c0 + r;
      RepD& R = W->h_rep[r];
      const double* L = &M->ds_L[3 * (size_t)ids[r]];
      BoxD& B = R.box;
      B.Lx = L[0]; B.Ly = L[1]; B.Lz = L[2];
      B.xlo = -0.5 * B.Lx; B.xhi = 0.5 * B.Lx; B.ylo = -0.5 * B.Ly; B.yhi = 0.5 * B.Ly;
      B.zlo = -0.5 * B.Lz; B.zhi = 0.5 * B.Lz;
      const double cut = W->rcut + W->pad;
      R.ncx = (int)(B.Lx / cut); R.ncy = (int)(B.Ly / cut); R.ncz = (int)(B.Lz / cut);
      R.wx = B.Lx / R.ncx; R.wy = B.Ly / R.ncy; R.wz = B.Lz / R.ncz;
    }
    MCHK(hipMemcpyAsync(W->d_rep, W->h_rep.data(), sizeof(RepD) * W->nrep, hipMemcpyHostToDevice, W->st));
    MCHK(hipMemcpyAsync(M->d_ids, ids.data(), sizeof(int) * W->nrep, hipMemcpyHostToDevice, W->st));
    if (int rc = ensure(M, W->n, M->capE > 0 ? (int)M->capE : 1, true)) return rc;
    hipLaunchKernelGGL(k_load_configs, dim3(blocks(W->n)), dim3(kThreads), 0, W->st, W->e, W->cur, W->d_rep,
                       M->ds_x0, M->ds_r, M->ds_t, M->d_ids, M->ds_per, M->target, W->n);
    if (int rc = ens_build_all(W)) { M->err = W->err; return rc; }
    if (int rc = forward(M, W, true)) return rc;
    hipLaunchKernelGGL(k_load_configs, dim3(blocks(W->n)), dim3(kThreads), 0, W->st, W->e, W->cur, W->d_rep,
                       M->ds_x0, M->ds_r, M->ds_t, M->d_ids, M->ds_per, M->target, W->n);
    double loss = 0.0;
    if (int rc = backward(M, W, &loss)) return rc;
    total += loss;
    ++batches;
  }
  *mean_loss = batches ? total / batches : 0.0;
  return 0;
}

int samos_mpnn_profile(samos_mpnn* M, int enable) {
  hipDeviceSynchronize();
  for (auto& p : M->ev_gemm) { hipEventDestroy(p.first); hipEventDestroy(p.second); }
  M->ev_gemm.clear();
  M->gemm_flops = 0.0;
  M->prof = enable != 0;
  return 0;
}

int samos_mpnn_profile_read(samos_mpnn* M, double* ms_total, double* flops, long long* launches) {
  MCHK(hipDeviceSynchronize());
  double t = 0.0;
  for (auto& p : M->ev_gemm) {
    float ms = 0.f;
    MCHK(hipEventElapsedTime(&ms, p.first, p.second));
    t += ms;
  }
  *ms_total = t;
  *flops = M->gemm_flops;
  *launches = (long long)M->ev_gemm.size();
  return 0;
}

}  // extern "C"
