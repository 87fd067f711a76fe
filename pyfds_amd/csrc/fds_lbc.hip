// Learned absorbing boundary: BF16 MFMA (v_mfma_f32_16x16x16_bf16) kernels with FP32
// accumulation for training (1D convolution forward, backward-data, backward-weight, MSE loss,
// Adam) and the fused per-step inference kernel that writes the boundary layer.
//
// MFMA 16x16x16 lane layout used throughout (l = lane, m = l % 16, g = l / 16):
//   A (16 x 16, rows x k):  lane holds A[m][4g .. 4g+3]
//   B (16 x 16, k x cols):  lane holds B[4g .. 4g+3][m]
//   D (16 x 16, f32):       lane holds D[4g + v][m], v = 0..3
// A 1D convolution y[p, co] = sum_k sum_ci w[co, k, ci] x[p + k - K/2, ci] (zero outside the
// strip) is a GEMM over (k, ci): rows are positions, and for a fixed tap the 16 ci of a k-chunk
// are contiguous in the [position][channel] layout, so each lane loads 4 bf16 (8 bytes).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pyfds_amd.h"
#include "fds_lbc.h"

namespace fds {
namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);   // round to nearest even
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ s4 ld4(const uint16_t *p) { return *reinterpret_cast<const s4 *>(p); }

__device__ __forceinline__ s4 ld4_strided(const uint16_t *p, int stride, int n_valid) {
  s4 r;
  r[0] = n_valid > 0 ? (short)p[0] : (short)0;
  r[1] = n_valid > 1 ? (short)p[stride] : (short)0;
  r[2] = n_valid > 2 ? (short)p[2 * stride] : (short)0;
  r[3] = n_valid > 3 ? (short)p[3 * stride] : (short)0;
  return r;
}

__device__ __forceinline__ f4 mfma(s4 a, s4 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------- training kernels
// x [B][P][Ci] bf16, w [S][Co][K][Ci] bf16, side [B] -> y [B][P][Co] f32 (+ bf16 copy)
__global__ __launch_bounds__(64) void k_conv_fwd(const uint16_t *__restrict__ x,
                                                 const uint16_t *__restrict__ w,
                                                 const int32_t *__restrict__ side,
                                                 float *__restrict__ y, uint16_t *__restrict__ yb,
                                                 int P, int Ci, int Co, int K) {
  const int p0 = blockIdx.x * 16, co0 = blockIdx.y * 16, b = blockIdx.z;
  const int l = threadIdx.x, m = l & 15, g = l >> 4, R = K / 2;
  const uint16_t *xb = x + (size_t)b * P * Ci;
  const uint16_t *ws = w + (size_t)side[b] * Co * K * Ci;
  f4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int k = 0; k < K; ++k) {
    const int pa = p0 + m + k - R;
    const bool in = pa >= 0 && pa < P;
    for (int c0 = 0; c0 < Ci; c0 += 16) {
      const s4 av = in ? ld4(xb + (size_t)pa * Ci + c0 + 4 * g) : s4{0, 0, 0, 0};
      const s4 bv = ld4(ws + ((size_t)(co0 + m) * K + k) * Ci + c0 + 4 * g);
      acc = mfma(av, bv, acc);
    }
  }
  for (int v = 0; v < 4; ++v) {
    const int p = p0 + 4 * g + v;
    if (p < P) {
      const size_t o = ((size_t)b * P + p) * Co + co0 + m;
      y[o] = acc[v];
      if (yb) yb[o] = f2bf(acc[v]);
    }
  }
}

// dy [B][P][Co] bf16, wt [S][Ci][K][Co] bf16 -> dx [B][P][Ci] f32 (+ bf16 copy)
__global__ __launch_bounds__(64) void k_conv_bwd_data(const uint16_t *__restrict__ dy,
                                                      const uint16_t *__restrict__ wt,
                                                      const int32_t *__restrict__ side,
                                                      float *__restrict__ dx,
                                                      uint16_t *__restrict__ dxb, int P, int Ci,
                                                      int Co, int K) {
  const int q0 = blockIdx.x * 16, ci0 = blockIdx.y * 16, b = blockIdx.z;
  const int l = threadIdx.x, m = l & 15, g = l >> 4, R = K / 2;
  const uint16_t *dyb = dy + (size_t)b * P * Co;
  const uint16_t *ws = wt + (size_t)side[b] * Ci * K * Co;
  f4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int k = 0; k < K; ++k) {
    const int pa = q0 + m - k + R;
    const bool in = pa >= 0 && pa < P;
    for (int c0 = 0; c0 < Co; c0 += 16) {
      const s4 av = in ? ld4(dyb + (size_t)pa * Co + c0 + 4 * g) : s4{0, 0, 0, 0};
      const s4 bv = ld4(ws + ((size_t)(ci0 + m) * K + k) * Co + c0 + 4 * g);
      acc = mfma(av, bv, acc);
    }
  }
  for (int v = 0; v < 4; ++v) {
    const int q = q0 + 4 * g + v;
    if (q < P) {
      const size_t o = ((size_t)b * P + q) * Ci + ci0 + m;
      dx[o] = acc[v];
      if (dxb) dxb[o] = f2bf(acc[v]);
    }
  }
}

// dw partial per sample: part[b][co][k][ci] = sum_p dy[b, p, co] * x[b, p + k - R, ci]
__global__ __launch_bounds__(64) void k_conv_bwd_weight(const uint16_t *__restrict__ dy,
                                                        const uint16_t *__restrict__ x,
                                                        float *__restrict__ part, int P, int Ci,
                                                        int Co, int K) {
  const int k = blockIdx.x, b = blockIdx.z;
  const int nci = Ci / 16;
  const int co0 = (blockIdx.y / nci) * 16, ci0 = (blockIdx.y % nci) * 16;
  const int l = threadIdx.x, m = l & 15, g = l >> 4, R = K / 2;
  const uint16_t *dyb = dy + (size_t)b * P * Co;
  const uint16_t *xb = x + (size_t)b * P * Ci;
  f4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int p0 = 0; p0 < P; p0 += 16) {
    // A[m][kk] = dy[p0 + kk, co0 + m]; B[kk][n] = x[p0 + kk + k - R, ci0 + n]
    const int pa = p0 + 4 * g;
    const s4 av = ld4_strided(dyb + (size_t)pa * Co + co0 + m, Co, P - pa);
    s4 bv;
    for (int v = 0; v < 4; ++v) {
      const int px = pa + v + k - R;
      bv[v] = (pa + v < P && px >= 0 && px < P) ? (short)xb[(size_t)px * Ci + ci0 + m] : (short)0;
    }
    acc = mfma(av, bv, acc);
  }
  float *pb = part + (size_t)b * Co * K * Ci;
  for (int v = 0; v < 4; ++v)
    pb[((size_t)(co0 + 4 * g + v) * K + k) * Ci + ci0 + m] = acc[v];
}

// dw[s][i] = sum over samples b with side[b] == s of part[b][i]
__global__ void k_reduce_weight(const float *__restrict__ part, const int32_t *__restrict__ side,
                                float *__restrict__ dw, int B, int S, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = blockIdx.y;
  if (i >= n) return;
  float acc = 0.f;
  for (int b = 0; b < B; ++b)
    if (side[b] == s) acc += part[(size_t)b * n + i];
  dw[(size_t)s * n + i] = acc;
}

// Weighted MSE over positions [lo[b], hi[b]) of each sample: loss partial sums per block and
// dy = 2 w_c (y - t) * norm as bf16.
__global__ __launch_bounds__(256) void k_mse(const float *__restrict__ y,
                                             const float *__restrict__ t,
                                             const int32_t *__restrict__ lo,
                                             const int32_t *__restrict__ hi,
                                             const float *__restrict__ wch,
                                             uint16_t *__restrict__ dyb,
                                             float *__restrict__ loss_part, int B, int P, int C,
                                             float norm) {
  __shared__ float red[256];
  const size_t total = (size_t)B * P * C;
  float acc = 0.f;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
    const int c = (int)(i % C);
    const size_t bp = i / C;
    const int p = (int)(bp % P), b = (int)(bp / P);
    float d = 0.f;
    if (p >= lo[b] && p < hi[b]) {
      const float e = y[i] - t[i];
      acc += wch[c] * e * e;
      d = 2.f * wch[c] * e * norm;
    }
    dyb[i] = f2bf(d);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0] * norm;
}

__global__ void k_adam(float *__restrict__ w, float *__restrict__ m1, float *__restrict__ m2,
                       const float *__restrict__ gr, int n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = gr[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float v = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = v;
  w[i] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

// master f32 [S][Co][K][Ci] -> bf16 same layout and bf16 transposed [S][Ci][K][Co]
__global__ void k_pack(const float *__restrict__ w, uint16_t *__restrict__ wb,
                       uint16_t *__restrict__ wtb, int S, int Co, int K, int Ci) {
  const int n = S * Co * K * Ci;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int ci = i % Ci, k = (i / Ci) % K, co = (i / (Ci * K)) % Co, s = i / (Ci * K * Co);
  const uint16_t h = f2bf(w[i]);
  wb[i] = h;
  if (wtb) wtb[(((size_t)s * Ci + ci) * K + k) * Co + co] = h;
}

// frames [F][3][4][D][G] f32 (F = runs * T steps).  Sample b = (run, step, side):
//   x[b][a][h*24 + c*8 + j] = frames[run*T + step - h][c][side][L + j][a] * cs[c] * scale[b]
//                             (zero for a outside [L, G-L))
//   t[b][a][c*16 + d]       = frames[run*T + step][c][side][d][a] * cs[c] * scale[b]
__global__ __launch_bounds__(256) void k_gather(const float *__restrict__ frames,
                                                const int32_t *__restrict__ run,
                                                const int32_t *__restrict__ step,
                                                const int32_t *__restrict__ side,
                                                const float *__restrict__ scale,
                                                uint16_t *__restrict__ x, float *__restrict__ t,
                                                int T, int D, int G, int L, float cs0, float cs1,
                                                float cs2) {
  const int b = blockIdx.y;
  const int a = blockIdx.x * 256 + threadIdx.x;
  if (a >= G) return;
  const float cs[3] = {cs0, cs1, cs2};
  const int s = side[b];
  const size_t f0 = (size_t)run[b] * T + step[b];
  const float sc = scale[b];
  const bool band = a >= L && a < G - L;
  uint16_t *xo = x + ((size_t)b * G + a) * kLbcCin;
  for (int h = 0; h < 2; ++h)
    for (int c = 0; c < 3; ++c)
      for (int j = 0; j < kLbcBand; ++j) {
        float v = 0.f;
        if (band)
          v = frames[((((f0 - h) * 3 + c) * 4 + s) * D + L + j) * G + a] * cs[c] * sc;
        xo[h * 24 + c * 8 + j] = f2bf(v);
      }
  float *to = t + ((size_t)b * G + a) * kLbcCout;
  for (int c = 0; c < 3; ++c)
    for (int d = 0; d < kLbcDepth; ++d)
      to[c * kLbcDepth + d] = frames[(((f0 * 3 + c) * 4 + s) * D + d) * G + a] * cs[c] * sc;
}

// RMS of the pressure band of every frame and side: rms[f][s]
__global__ __launch_bounds__(256) void k_band_rms(const float *__restrict__ frames,
                                                  float *__restrict__ rms, int D, int G, int L) {
  __shared__ float red[256];
  const int f = blockIdx.x, s = blockIdx.y;
  float acc = 0.f;
  const int n = kLbcBand * (G - 2 * L);
  for (int e = threadIdx.x; e < n; e += 256) {
    const int j = e / (G - 2 * L), a = L + e % (G - 2 * L);
    const float v = frames[((((size_t)f * 3 + 0) * 4 + s) * D + L + j) * G + a];
    acc += v * v;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) rms[(size_t)f * 4 + s] = sqrtf(red[0] / (float)n);
}

// ---------------------------------------------------------------- fused inference, one step
__global__ __launch_bounds__(256) void k_lbc_step(LbcStep A) {
  constexpr int R = kLbcTaps / 2;
  constexpr int NX = kLbcSeg + 4 * R;                       // input positions
  constexpr int NH = (kLbcSeg + 2 * R + 15) / 16 * 16;      // hidden rows (whole MFMA tiles)
  constexpr int NXP = NH + 2 * R;                           // input rows the tiles read
  constexpr int NW1 = kLbcHidden * kLbcTaps * kLbcCin;
  constexpr int NW2 = kLbcCout * kLbcTaps * kLbcHidden;
  // input rows past NX feed only hidden rows past kLbcSeg + 2R, which layer 2 never reads;
  // they are zeroed so that every tile reads its operands without a bounds test
  __shared__ __attribute__((aligned(16))) uint16_t xin[NXP * kLbcCin];
  __shared__ __attribute__((aligned(16))) uint16_t hid[NH * kLbcHidden];
  __shared__ __attribute__((aligned(16))) uint16_t sw1[NW1];
  __shared__ __attribute__((aligned(16))) uint16_t sw2[NW2];
  __shared__ float outs[kLbcSeg * kLbcCout];
  const LbcStep &st = A;
  const int G = st.g, L = st.l;
  const int s = blockIdx.y, a0 = blockIdx.x * kLbcSeg;
  const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63, m = l & 15, g = l >> 4;
  // top and bottom rows run along a (contiguous in memory), left and right columns along the
  // depth: the gather and the scatter let consecutive lanes walk the contiguous direction
  const bool along = s < 2;

  // this side's weights and the band inputs into LDS: every global load is issued before the
  // first LDS store, so the workgroup waits for one memory round trip, not one per element.
  // Components are picked by selects among the argument block's pointers: indexing the block
  // with a lane-varying component makes the compiler load the pointer itself first, one
  // dependent round trip per element
  constexpr int N1 = NW1 / 4, N2 = NW2 / 4, NI = NX * kLbcCin;
  constexpr int I1 = (N1 + 255) / 256, I2 = (N2 + 255) / 256, II = (NI + 255) / 256;
  s4 r1[I1], r2[I2];
  float rx[II];
  {
    const s4 *g1 = reinterpret_cast<const s4 *>(st.w1 + (size_t)s * NW1);
    const s4 *g2 = reinterpret_cast<const s4 *>(st.w2 + (size_t)s * NW2);
    // loads past the end re-read the last element (the LDS store below skips them): a load
    // under a branch would make the compiler wait for it where the branches join
#pragma unroll
    for (int i = 0; i < I1; ++i) r1[i] = g1[min(tid + 256 * i, N1 - 1)];
#pragma unroll
    for (int i = 0; i < I2; ++i) r2[i] = g2[min(tid + 256 * i, N2 - 1)];
    const float *cur0 = st.cur[0], *cur1 = st.cur[1], *cur2 = st.cur[2];
    const float *prv0 = st.prev[0], *prv1 = st.prev[1], *prv2 = st.prev[2];
#pragma unroll
    for (int i = 0; i < II; ++i) {
      const int e = tid + 256 * i;
      float v = 0.f;
      if (e < NI) {
        const int row = along ? e % NX : e / kLbcCin, ci = along ? e / NX : e % kLbcCin;
        const int a = a0 - 2 * R + row;
        if (a >= L && a < G - L) {
          const int h = ci / 24, c = (ci % 24) / kLbcBand, j = ci % kLbcBand;
          const float *src = h == 0 ? (c == 0 ? cur0 : c == 1 ? cur1 : cur2)
                                    : (c == 0 ? prv0 : c == 1 ? prv1 : prv2);
          v = src[lbc_cell(s, L + j, a, G)];
        }
      }
      rx[i] = v;
    }
    s4 *l1 = reinterpret_cast<s4 *>(sw1);
    s4 *l2 = reinterpret_cast<s4 *>(sw2);
#pragma unroll
    for (int i = 0; i < I1; ++i) {
      const int e = tid + 256 * i;
      if (e < N1) l1[e] = r1[i];
    }
#pragma unroll
    for (int i = 0; i < I2; ++i) {
      const int e = tid + 256 * i;
      if (e < N2) l2[e] = r2[i];
    }
    const float cs0 = A.cs[0], cs1 = A.cs[1], cs2 = A.cs[2];
    for (int e = NI + tid; e < NXP * kLbcCin; e += 256) xin[e] = 0;
#pragma unroll
    for (int i = 0; i < II; ++i) {
      const int e = tid + 256 * i;
      if (e < NI) {
        const int row = along ? e % NX : e / kLbcCin, ci = along ? e / NX : e % kLbcCin;
        const int c = (ci % 24) / kLbcBand;
        xin[row * kLbcCin + ci] = f2bf(rx[i] * (c == 0 ? cs0 : c == 1 ? cs1 : cs2));
      }
    }
  }
  __syncthreads();

  // layer 1: hid[hp][co], hidden position a = a0 - R + hp (zero outside the strip)
  const uint16_t *w1 = sw1;
  for (int t = wave; t < (NH / 16) * (kLbcHidden / 16); t += 4) {
    const int mt = t / (kLbcHidden / 16), nt = t % (kLbcHidden / 16);
    // all operands of the tile first, then the MFMA chain: one LDS round trip per tile
    constexpr int NC = kLbcCin / 16;
    s4 av[kLbcTaps * NC], bv[kLbcTaps * NC];
#pragma unroll
    for (int k = 0; k < kLbcTaps; ++k)
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        av[k * NC + c] = ld4(xin + (16 * mt + m + k) * kLbcCin + 16 * c + 4 * g);
        bv[k * NC + c] = ld4(w1 + ((16 * nt + m) * kLbcTaps + k) * kLbcCin + 16 * c + 4 * g);
      }
    f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int i = 0; i < kLbcTaps * NC; ++i) acc = mfma(av[i], bv[i], acc);
    for (int v = 0; v < 4; ++v) {
      const int hp = 16 * mt + 4 * g + v;
      const int a = a0 - R + hp;
      hid[hp * kLbcHidden + 16 * nt + m] = (a >= 0 && a < G) ? f2bf(acc[v]) : (uint16_t)0;
    }
  }
  __syncthreads();

  // layer 2: outs[op][co] at a = a0 + op
  const uint16_t *w2 = sw2;
  for (int t = wave; t < (kLbcSeg / 16) * (kLbcCout / 16); t += 4) {
    const int mt = t / (kLbcCout / 16), nt = t % (kLbcCout / 16);
    constexpr int NC = kLbcHidden / 16;
    s4 av[kLbcTaps * NC], bv[kLbcTaps * NC];
#pragma unroll
    for (int k = 0; k < kLbcTaps; ++k)
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        av[k * NC + c] = ld4(hid + (16 * mt + m + k) * kLbcHidden + 16 * c + 4 * g);
        bv[k * NC + c] = ld4(w2 + ((16 * nt + m) * kLbcTaps + k) * kLbcHidden + 16 * c + 4 * g);
      }
    f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int i = 0; i < kLbcTaps * NC; ++i) acc = mfma(av[i], bv[i], acc);
    for (int v = 0; v < 4; ++v) outs[(16 * mt + 4 * g + v) * kLbcCout + 16 * nt + m] = acc[v];
  }
  __syncthreads();

  static_assert(kLbcSeg * kLbcCout % 256 == 0, "scatter covers the outputs in whole passes");
  float *const cur[3] = {st.cur[0], st.cur[1], st.cur[2]};
  float *const comp[3] = {st.cur_comp[0], st.cur_comp[1], st.cur_comp[2]};
  const float cs[3] = {A.cs[0], A.cs[1], A.cs[2]};
#pragma unroll
  for (int i = 0; i < kLbcSeg * kLbcCout / 256; ++i) {
    const int e = tid + 256 * i;
    const int op = along ? e % kLbcSeg : e / kLbcCout, co = along ? e / kLbcSeg : e % kLbcCout;
    const int a = a0 + op;
    if (a >= G || (!along && (a < L || a >= G - L))) continue;
    const int c = co / kLbcDepth, d = co % kLbcDepth;
    const int cell = lbc_cell(s, d, a, G);
    const float v = outs[op * kLbcCout + co];
    // one branch per component with a compile-time index keeps the pointers in scalar
    // registers (along the top and bottom rows c is uniform over each wave)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (c != k) continue;
      cur[k][cell] = v / cs[k];
      if (comp[k]) comp[k][cell] = 0.f;
    }
  }
}

}  // namespace

void launch_lbc_step(const LbcStep &st, hipStream_t s) {
  const int segs = (st.g + kLbcSeg - 1) / kLbcSeg;
  hipLaunchKernelGGL(k_lbc_step, dim3(segs, 4), dim3(256), 0, s, st);
}

}  // namespace fds

// ---------------------------------------------------------------- C-ABI (training side)
namespace {

hipStream_t as_stream(void *s) { return static_cast<hipStream_t>(s); }

int last_launch() { return hipGetLastError() == hipSuccess ? 0 : FDS_E_HIP; }

bool bad_dims(int Ci, int Co) { return Ci <= 0 || Co <= 0 || Ci % 16 || Co % 16; }

}  // namespace

extern "C" {

int fds_lbc_conv_fwd(const uint16_t *x, const uint16_t *w, const int32_t *side, float *y,
                     uint16_t *yb, int B, int P, int Ci, int Co, int K, void *stream) {
  if (bad_dims(Ci, Co) || B < 1 || B > 65535 || P < 1 || K < 1 || K % 2 == 0) return FDS_E_ARG;
  hipLaunchKernelGGL(fds::k_conv_fwd, dim3((P + 15) / 16, Co / 16, B), dim3(64), 0,
                     as_stream(stream), x, w, side, y, yb, P, Ci, Co, K);
  return last_launch();
}

int fds_lbc_conv_bwd_data(const uint16_t *dy, const uint16_t *wt, const int32_t *side,
                          float *dx, uint16_t *dxb, int B, int P, int Ci, int Co, int K,
                          void *stream) {
  if (bad_dims(Ci, Co) || B < 1 || B > 65535 || P < 1 || K < 1 || K % 2 == 0) return FDS_E_ARG;
  hipLaunchKernelGGL(fds::k_conv_bwd_data, dim3((P + 15) / 16, Ci / 16, B), dim3(64), 0,
                     as_stream(stream), dy, wt, side, dx, dxb, P, Ci, Co, K);
  return last_launch();
}

int fds_lbc_conv_bwd_weight(const uint16_t *dy, const uint16_t *x, float *part, int B, int P,
                            int Ci, int Co, int K, void *stream) {
  if (bad_dims(Ci, Co) || B < 1 || B > 65535 || P < 1 || K < 1 || K % 2 == 0) return FDS_E_ARG;
  hipLaunchKernelGGL(fds::k_conv_bwd_weight, dim3(K, (Co / 16) * (Ci / 16), B), dim3(64), 0,
                     as_stream(stream), dy, x, part, P, Ci, Co, K);
  return last_launch();
}

int fds_lbc_reduce_weight(const float *part, const int32_t *side, float *dw, int B, int S,
                          int n, void *stream) {
  if (B < 1 || S < 1 || n < 1) return FDS_E_ARG;
  hipLaunchKernelGGL(fds::k_reduce_weight, dim3((n + 255) / 256, S), dim3(256), 0,
                     as_stream(stream), part, side, dw, B, S, n);
  return last_launch();
}

int fds_lbc_mse(const float *y, const float *t, const int32_t *lo, const int32_t *hi,
                const float *wch, uint16_t *dyb, float *loss_part, int nblocks, int B, int P,
                int C, float norm, void *stream) {
  if (B < 1 || P < 1 || C < 1 || nblocks < 1) return FDS_E_ARG;
  hipLaunchKernelGGL(fds::k_mse, dim3(nblocks), dim3(256), 0, as_stream(stream), y, t, lo, hi,
                     wch, dyb, loss_part, B, P, C, norm);
  return last_launch();
}

int fds_lbc_adam(float *w, float *m1, float *m2, const float *g, int n, float lr, float b1,
                 float b2, float eps, int step, void *stream) {
  if (n < 1 || step < 1) return FDS_E_ARG;
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(fds::k_adam, dim3((n + 255) / 256), dim3(256), 0, as_stream(stream), w, m1,
                     m2, g, n, lr, b1, b2, eps, c1, c2);
  return last_launch();
}

int fds_lbc_pack(const float *w, uint16_t *wb, uint16_t *wtb, int S, int Co, int K, int Ci,
                 void *stream) {
  const int n = S * Co * K * Ci;
  if (n < 1) return FDS_E_ARG;
  hipLaunchKernelGGL(fds::k_pack, dim3((n + 255) / 256), dim3(256), 0, as_stream(stream), w, wb,
                     wtb, S, Co, K, Ci);
  return last_launch();
}

int fds_lbc_gather(const float *frames, const int32_t *run, const int32_t *step,
                   const int32_t *side, const float *scale, uint16_t *x, float *t, int B, int T,
                   int D, int G, int L, const float *comp_scale, void *stream) {
  if (B < 1 || B > 65535 || T < 2 || G < 2 * L + 1 || L != fds::kLbcDepth ||
      D < L + fds::kLbcBand)
    return FDS_E_ARG;
  hipLaunchKernelGGL(fds::k_gather, dim3((G + 255) / 256, B), dim3(256), 0, as_stream(stream),
                     frames, run, step, side, scale, x, t, T, D, G, L, comp_scale[0],
                     comp_scale[1], comp_scale[2]);
  return last_launch();
}

int fds_lbc_band_rms(const float *frames, float *rms, int F, int D, int G, int L, void *stream) {
  if (F < 1 || F > 65535 || G < 2 * L + 1 || D < L + fds::kLbcBand) return FDS_E_ARG;
  hipLaunchKernelGGL(fds::k_band_rms, dim3(F, 4), dim3(256), 0, as_stream(stream), frames, rms,
                     D, G, L);
  return last_launch();
}

}  // extern "C"
