// Learned pilot-to-grid channel estimator: a 1-D convolutional network over the
// pilot positions of one (receive, transmit) stream pair, bfloat16 weights and
// activations on the MFMA units with float32 accumulation, and the kernels that
// train it on drops the simulator generates (backward-data on the same MFMA
// convolution, weight gradients, NMSE loss, Adam).
//
// Layout: sample s = (t * num_drops + d) * num_rx + r, positions p = 0..63,
// channels innermost: activations [S][64][C].  Weights [Cout][taps][Cin].  The
// positions are cyclic (a stream's pilots on a grid without DC roll sample the
// periodic frequency response uniformly), so every convolution is circular:
// out[p] = sum_k w[k] in[(p + (k - taps/2) dil) mod 64], with dilations 1, 4, 16, 1
// in the four layers (a receptive field covering the whole axis) and 1 in the skip.  The network is the learned counterpart of
// OrthogonalLeastSquaresChannelEstimation's interpolation
// (hermespy/modem/waveforms/orthogonal/channel.py:29-66): it maps the LS values at
// a stream's pilots to the channel of every resource element.
#include <math.h>

#include <algorithm>

#include "hpx_common.hpp"

using namespace hpx;

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int kPos = 64;   // pilot positions per sample
constexpr int kTaps = 5;       // the four layers
constexpr int kSkipTaps = 17;  // the linear skip convolution
constexpr int kDil1 = 1, kDil2 = 4, kDil3 = 16, kDil4 = 1;

inline hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return static_cast<uint16_t>(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

// GELU, tanh form
__device__ __forceinline__ float gelu(float x) {
  const float k0 = 0.7978845608028654f, k1 = 0.044715f;
  return 0.5f * x * (1.0f + tanhf(k0 * (x + k1 * x * x * x)));
}
__device__ __forceinline__ float gelu_grad(float x) {
  const float k0 = 0.7978845608028654f, k1 = 0.044715f;
  const float u = k0 * (x + k1 * x * x * x);
  const float th = tanhf(u);
  return 0.5f * (1.0f + th) + 0.5f * x * (1.0f - th * th) * k0 * (1.0f + 3.0f * k1 * x * x);
}

// ---------------------------------------------------------------- convolution (MFMA)
// out[s][p][co] = bias[co] + sum_{k, ci} w[co][k][ci] in[s][p + k - 2][ci]
// One workgroup per sample; wave v computes rows 32 (v & 1) .. +32 of output column
// tiles v / 2, v / 2 + 2, ...  mode 0: out_f32 = acc (+ bias), out_bf16 = bf16(that);
// mode 1: out_f32 = pre-activation, out_bf16 = bf16(gelu(pre)); mode 2: out_f32 += acc (+ bias).
template <int CIN, int TAPS>
__global__ void __launch_bounds__(256) conv_fwd_kernel(const uint16_t* __restrict__ in,
                                                       const uint16_t* __restrict__ w,
                                                       const float* __restrict__ bias, int cout, int dil, int mode,
                                                       float* __restrict__ out_f32, uint16_t* __restrict__ out_bf16) {
  constexpr int kStride = CIN + 4;  // padded LDS row: conflict-free 8-byte reads across a half wave
  __shared__ __align__(16) uint16_t xs[kPos * kStride];
  const int s = blockIdx.x;
  const int tid = threadIdx.x;
  const uint16_t* xin = in + static_cast<int64_t>(s) * kPos * CIN;
  for (int i = tid; i < kPos * CIN; i += 256) {
    const int row = i / CIN, c = i - row * CIN;
    xs[row * kStride + c] = xin[i];
  }
  __syncthreads();
  const int wave = tid >> 6, lane = tid & 63;
  const int half = wave & 1;
  const int m = lane & 31, kq = lane >> 5;
  for (int ct = wave >> 1; ct < cout / 32; ct += 2) {
    floatx16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
    const int co = ct * 32 + m;
    const uint16_t* wrow = w + static_cast<int64_t>(co) * TAPS * CIN + 4 * kq;
#pragma unroll
    for (int k = 0; k < TAPS; ++k) {
      const uint16_t* arow = xs + ((32 * half + m + (k - TAPS / 2) * dil) & (kPos - 1)) * kStride + 4 * kq;
#pragma unroll 4
      for (int c0 = 0; c0 < CIN; c0 += 8) {
        const bf16x4 a = *reinterpret_cast<const bf16x4*>(arow + c0);
        const bf16x4 b = *reinterpret_cast<const bf16x4*>(wrow + k * CIN + c0);
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
      }
    }
    const float bv = bias ? bias[co] : 0.0f;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int p = 32 * half + (r >> 2) * 8 + kq * 4 + (r & 3);
      const int64_t o = (static_cast<int64_t>(s) * kPos + p) * cout + co;
      const float v = acc[r] + bv;
      if (mode == 2) {
        out_f32[o] += v;
        continue;
      }
      if (out_f32) out_f32[o] = v;
      if (out_bf16) out_bf16[o] = f2bf(mode == 1 ? gelu(v) : v);
    }
  }
}

// ---------------------------------------------------------------- elementwise
__global__ void __launch_bounds__(256) gelu_bwd_kernel(const float* __restrict__ grad, const float* __restrict__ pre,
                                                       int64_t n, uint16_t* __restrict__ dz) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dz[i] = f2bf(grad[i] * gelu_grad(pre[i]));
}

__global__ void __launch_bounds__(256) to_bf16_kernel(const float* __restrict__ x, int64_t n, uint16_t* __restrict__ y) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

// ---------------------------------------------------------------- weight gradient (float32 VALU)
// part[chunk][co][k][ci] = sum over the chunk's rows of dz[s][p][co] in[s][p + k - 2][ci]
// part_bias[chunk][co] = sum dz.  Workgroup: one chunk of samples x a 32 (co) x 64 (k, ci) tile.
constexpr int kWgSamples = 32;

__global__ void __launch_bounds__(256) wgrad_kernel(const uint16_t* __restrict__ dz, const uint16_t* __restrict__ in,
                                                    int num_samples, int cout, int cin, int taps, int dil,
                                                    float* __restrict__ part, float* __restrict__ part_bias) {
  __shared__ float dzs[kPos][33];
  __shared__ float xs[kPos][65];
  const int chunk = blockIdx.x;
  const int co0 = blockIdx.y * 32;
  const int kcols = taps * cin;
  const int halo = taps / 2;
  const int kc0 = blockIdx.z * 64;
  const int tid = threadIdx.x;
  // thread: co = co0 + (tid & 31), columns kc0 + (tid >> 5) + 8 j, j = 0..7
  const int tco = tid & 31, tcol = tid >> 5;
  float acc[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) acc[j] = 0.0f;
  float accb = 0.0f;
  const int s_end = std::min(num_samples, (chunk + 1) * kWgSamples);
  for (int s = chunk * kWgSamples; s < s_end; ++s) {
    for (int i = tid; i < kPos * 32; i += 256) {
      const int p = i >> 5, c = i & 31;
      dzs[p][c] = co0 + c < cout ? bf2f(dz[(static_cast<int64_t>(s) * kPos + p) * cout + co0 + c]) : 0.0f;
    }
    // input columns kc0 .. kc0 + 63 of the im2col row: (k, ci) = divmod(col, cin)
    for (int i = tid; i < kPos * 64; i += 256) {
      const int p = i >> 6, c = i & 63;
      const int col = kc0 + c;
      float v = 0.0f;
      if (col < kcols) {
        const int k = col / cin, ci = col - k * cin;
        const int q = (p + (k - halo) * dil) & (kPos - 1);  // input position feeding output p through tap k
        v = bf2f(in[(static_cast<int64_t>(s) * kPos + q) * cin + ci]);
      }
      xs[p][c] = v;
    }
    __syncthreads();
    for (int p = 0; p < kPos; ++p) {
      const float g = dzs[p][tco];
      if (tcol == 0) accb += g;
#pragma unroll
      for (int j = 0; j < 8; ++j) acc[j] = fmaf(g, xs[p][tcol + 8 * j], acc[j]);
    }
    __syncthreads();
  }
  const int co = co0 + tco;
  if (co < cout) {
    float* prow = part + (static_cast<int64_t>(chunk) * cout + co) * kcols;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int col = kc0 + tcol + 8 * j;
      if (col < kcols) prow[col] = acc[j];
    }
    if (tcol == 0 && blockIdx.z == 0) part_bias[static_cast<int64_t>(chunk) * cout + co] = accb;
  }
}

__global__ void __launch_bounds__(256) reduce_parts_kernel(const float* __restrict__ part, int chunks, int64_t n,
                                                           float* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.0f;
  for (int c = 0; c < chunks; ++c) s += part[static_cast<int64_t>(c) * n + i];
  out[i] = s;
}

// ---------------------------------------------------------------- pilots in, grid out
// in[s][p][c]: c = 2 v + {0: re, 1: im} of LS value v * 64 + p (v = pilot word), c = 2 V: +-1 stream flag
// -log10(noise power) / 3: the SNR the receiver knows (0 dB -> 0, 30 dB -> 1 at unit signal power)
__device__ __forceinline__ float snr_feature(const double* noise_power, int64_t d) {
  return noise_power ? static_cast<float>(-log10(noise_power[d]) / 3.0) : 0.0f;
}

__global__ void __launch_bounds__(256) net_input_kernel(const float2* __restrict__ ls, int64_t num_samples,
                                                        int num_drops, int num_rx, int pilot_words, int cin,
                                                        const double* __restrict__ noise_power,
                                                        uint16_t* __restrict__ x) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= num_samples * kPos) return;
  const int64_t s = i / kPos;
  const int p = static_cast<int>(i - s * kPos);
  const int t = static_cast<int>(s / (static_cast<int64_t>(num_drops) * num_rx));
  const float2* row = ls + s * pilot_words * kPos;
  uint16_t* o = x + i * cin;
  for (int v = 0; v < pilot_words; ++v) {
    const float2 h = row[v * kPos + p];
    o[2 * v] = f2bf(h.x);
    o[2 * v + 1] = f2bf(h.y);
  }
  o[2 * pilot_words] = f2bf(t ? 1.0f : -1.0f);
  o[2 * pilot_words + 1] = f2bf(snr_feature(noise_power, (s % (static_cast<int64_t>(num_drops) * num_rx)) / num_rx));
  for (int c = 2 * pilot_words + 2; c < cin; ++c) o[c] = 0;
}

// out[s][p][c], c = (w * B + u) * 2 + {re, im} -> est[s][w][B p + u]  (B = N / 64 subcarriers per position)
__global__ void __launch_bounds__(256) net_output_kernel(const float* __restrict__ out, int64_t num_samples, int words,
                                                         int n, float2* __restrict__ est) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t wn = static_cast<int64_t>(words) * n;
  if (i >= num_samples * wn) return;
  const int64_t s = i / wn;
  const int e = static_cast<int>(i - s * wn);
  const int w = e / n, q = e - w * n;
  const int blk = n / kPos;
  const int p = q / blk, u = q - p * blk;
  const int cout = 2 * words * blk;
  const float* o = out + (s * kPos + p) * cout + 2 * (w * blk + u);
  est[i] = make_float2(o[0], o[1]);
}

// ---------------------------------------------------------------- NMSE against the ideal CSI
// H[s][w][q] = S[d][r][t] c[d][w][q];  acc[0] += |est - H|^2, acc[1] += |H|^2 (per drop when per_drop)
__device__ __forceinline__ float2 ideal_h(const float2* csi, const float2* spatial, int64_t s, int64_t e, int num_drops,
                                          int num_rx, int num_tx, int64_t wn) {
  const int64_t dr = s % (static_cast<int64_t>(num_drops) * num_rx);
  const int t = static_cast<int>(s / (static_cast<int64_t>(num_drops) * num_rx));
  const int64_t d = dr / num_rx;
  const int r = static_cast<int>(dr - d * num_rx);
  return cmul(spatial[(d * num_rx + r) * num_tx + t], csi[d * wn + e]);
}

__global__ void __launch_bounds__(256) estimate_error_kernel(const float2* __restrict__ est,
                                                             const float2* __restrict__ csi,
                                                             const float2* __restrict__ spatial, int num_drops,
                                                             int num_rx, int num_tx, int64_t wn,
                                                             double* __restrict__ err, double* __restrict__ pow) {
  // one workgroup per (sample); per-drop sums
  __shared__ double red[2][4];
  const int64_t s = blockIdx.x;
  double e2 = 0.0, h2 = 0.0;
  for (int64_t e = threadIdx.x; e < wn; e += blockDim.x) {
    const float2 h = ideal_h(csi, spatial, s, e, num_drops, num_rx, num_tx, wn);
    const float2 x = est[s * wn + e];
    e2 += static_cast<double>(cabs2(csub(x, h)));
    h2 += static_cast<double>(cabs2(h));
  }
  for (int off = 32; off > 0; off >>= 1) {
    e2 += __shfl_down(e2, off, 64);
    h2 += __shfl_down(h2, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = e2;
    red[1][threadIdx.x >> 6] = h2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int64_t d = (s % (static_cast<int64_t>(num_drops) * num_rx)) / num_rx;
    atomicAdd(err + d, red[0][0] + red[0][1] + red[0][2] + red[0][3]);
    atomicAdd(pow + d, red[1][0] + red[1][1] + red[1][2] + red[1][3]);
  }
}

// Loss of a batch: mean over SNR groups of the NMSE in dB, 10 log10(sum |out - H|^2 / sum |H|^2) per
// group, so that every SNR point weighs alike (a batch NMSE would be dominated by the noisiest
// drops).  Pass 1 accumulates per-group error and power (one workgroup per sample), pass 2 turns
// them into the loss and per-group gradient scales, pass 3 writes d loss / d out.
__device__ __forceinline__ int64_t out_index(int64_t s, int e, int words, int n) {
  const int w = e / n, q = e - w * n;
  const int blk = n / kPos;
  const int p = q / blk, u = q - p * blk;
  return (s * kPos + p) * (2 * words * blk) + 2 * (w * blk + u);
}

__device__ __forceinline__ int sample_group(int64_t s, int num_drops, int num_rx, const int32_t* group) {
  const int64_t d = (s % (static_cast<int64_t>(num_drops) * num_rx)) / num_rx;
  return group ? group[d] : 0;
}

__global__ void __launch_bounds__(256) loss_accumulate_kernel(const float* __restrict__ out,
                                                              const float2* __restrict__ csi,
                                                              const float2* __restrict__ spatial, int num_drops,
                                                              int num_rx, int num_tx, int words, int n,
                                                              const int32_t* __restrict__ group, int num_groups,
                                                              double* __restrict__ work) {
  __shared__ double red[2][4];
  const int64_t s = blockIdx.x;
  const int64_t wn = static_cast<int64_t>(words) * n;
  double e2 = 0.0, h2 = 0.0;
  for (int e = threadIdx.x; e < wn; e += blockDim.x) {
    const float2 h = ideal_h(csi, spatial, s, e, num_drops, num_rx, num_tx, wn);
    const int64_t o = out_index(s, e, words, n);
    const float dr = out[o] - h.x, di = out[o + 1] - h.y;
    e2 += static_cast<double>(dr * dr + di * di);
    h2 += static_cast<double>(cabs2(h));
  }
  for (int off = 32; off > 0; off >>= 1) {
    e2 += __shfl_down(e2, off, 64);
    h2 += __shfl_down(h2, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = e2;
    red[1][threadIdx.x >> 6] = h2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int g = sample_group(s, num_drops, num_rx, group);
    atomicAdd(work + g, red[0][0] + red[0][1] + red[0][2] + red[0][3]);
    atomicAdd(work + num_groups + g, red[1][0] + red[1][1] + red[1][2] + red[1][3]);
  }
}

__global__ void loss_finish_kernel(int num_groups, double* __restrict__ work, double* __restrict__ loss) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double l = 0.0;
  int used = 0;
  for (int g = 0; g < num_groups; ++g) {
    const double err = work[g], pw = work[num_groups + g];
    if (err > 0.0 && pw > 0.0) {
      l += 10.0 * log10(err / pw);
      ++used;
    }
  }
  // d/d out of mean_g 10 log10(err_g / pow_g) = (10 / ln 10) / (used err_g) * 2 (out - H)
  for (int g = 0; g < num_groups; ++g)
    work[2 * num_groups + g] = (work[g] > 0.0 && used) ? 2.0 * 4.342944819032518 / (used * work[g]) : 0.0;
  loss[0] = used ? l / used : 0.0;
}

__global__ void __launch_bounds__(256) loss_grad_kernel(const float* __restrict__ out, const float2* __restrict__ csi,
                                                        const float2* __restrict__ spatial, int num_drops, int num_rx,
                                                        int num_tx, int words, int n, const int32_t* __restrict__ group,
                                                        int num_groups, const double* __restrict__ work,
                                                        float* __restrict__ grad) {
  const int64_t wn = static_cast<int64_t>(words) * n;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t num_samples = static_cast<int64_t>(num_tx) * num_drops * num_rx;
  if (i >= num_samples * wn) return;
  const int64_t s = i / wn;
  const int e = static_cast<int>(i - s * wn);
  const float2 h = ideal_h(csi, spatial, s, e, num_drops, num_rx, num_tx, wn);
  const int64_t o = out_index(s, e, words, n);
  const float scale = static_cast<float>(work[2 * num_groups + sample_group(s, num_drops, num_rx, group)]);
  grad[o] = scale * (out[o] - h.x);
  grad[o + 1] = scale * (out[o + 1] - h.y);
}

// ---------------------------------------------------------------- Adam + weight copies
// params float32 [cout][5][cin] (+ bias [cout] separately); writes w_bf16 [cout][5][cin] and the
// transposed, flipped copy wt_bf16 [cin][5][cout] (wt[ci][k][co] = w[co][4 - k][ci]) used by
// the backward-data convolution.
__global__ void __launch_bounds__(256) adam_kernel(float* __restrict__ p, const float* __restrict__ g,
                                                   float* __restrict__ m1, float* __restrict__ m2, int64_t n, float lr,
                                                   float b1, float b2, float eps, float c1, float c2) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.0f - b1) * gi;
  const float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  p[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

__global__ void __launch_bounds__(256) weight_copies_kernel(const float* __restrict__ p, int cout, int cin, int taps,
                                                            uint16_t* __restrict__ w, uint16_t* __restrict__ wt) {
  const int64_t n = static_cast<int64_t>(cout) * taps * cin;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int co = static_cast<int>(i / (taps * cin));
  const int rest = static_cast<int>(i - static_cast<int64_t>(co) * taps * cin);
  const int k = rest / cin, ci = rest - k * cin;
  const uint16_t v = f2bf(p[i]);
  if (w) w[i] = v;
  if (wt) wt[(static_cast<int64_t>(ci) * taps + (taps - 1 - k)) * cout + co] = v;
}

// ---------------------------------------------------------------- fused inference
// The whole estimator for G samples per workgroup: LS at the sample's pilots (y_p / x_p), the four
// layers and the skip convolution, with every activation in LDS (bfloat16, zero halo rows), and
// the output written straight into the estimate grid.  Numerically the layer-by-layer path: the
// same float32 accumulation, bias, GELU and bfloat16 rounding between layers; the 64-channel
// layers use the 32x32x16 MFMA (8 channels per lane and step).
constexpr int kFusedSamples = 4;  // samples per workgroup: 78 KB of LDS, two workgroups per CU
constexpr int kS8 = 8 + 4, kS64 = 64 + 8;  // padded LDS rows

// 8-channel input (32x32x8 steps): tap k reads row (p + (k - TAPS/2) dil) mod 64
template <int TAPS>
__device__ __forceinline__ void fused_tile8(const uint16_t (*in)[kPos * kS8], const uint16_t* __restrict__ w, int co,
                                            int dil, int half, int m, int kq, floatx16 (&acc)[kFusedSamples]) {
  const uint16_t* wrow = w + static_cast<int64_t>(co) * TAPS * 8 + 4 * kq;
#pragma unroll 1
  for (int k = 0; k < TAPS; ++k) {
    const bf16x4 b = *reinterpret_cast<const bf16x4*>(wrow + k * 8);
    const int row = (32 * half + m + (k - TAPS / 2) * dil) & (kPos - 1);
#pragma unroll
    for (int g = 0; g < kFusedSamples; ++g) {
      const bf16x4 a = *reinterpret_cast<const bf16x4*>(&in[g][row * kS8 + 4 * kq]);
      acc[g] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc[g], 0, 0, 0);
    }
  }
}

// 64-channel input (32x32x16 steps: lanes 0-31 take channels c0..c0+7, lanes 32-63 c0+8..c0+15)
__device__ __forceinline__ void fused_tile64(const uint16_t (*in)[kPos * kS64], const uint16_t* __restrict__ w, int co,
                                             int dil, int half, int m, int kq, floatx16 (&acc)[kFusedSamples]) {
  const uint16_t* wrow = w + static_cast<int64_t>(co) * kTaps * 64 + 8 * kq;
#pragma unroll 1
  for (int k = 0; k < kTaps; ++k) {
    const int row = (32 * half + m + (k - kTaps / 2) * dil) & (kPos - 1);
#pragma unroll
    for (int c0 = 0; c0 < 64; c0 += 16) {
      const bf16x8 b = *reinterpret_cast<const bf16x8*>(wrow + k * 64 + c0);
#pragma unroll
      for (int g = 0; g < kFusedSamples; ++g) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(&in[g][row * kS64 + 8 * kq + c0]);
        acc[g] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[g], 0, 0, 0);
      }
    }
  }
}

__device__ __forceinline__ void zero_acc(floatx16 (&acc)[kFusedSamples]) {
#pragma unroll
  for (int g = 0; g < kFusedSamples; ++g)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[g][r] = 0.0f;
}

__device__ __forceinline__ void store_hidden(const floatx16 (&acc)[kFusedSamples], const float* __restrict__ bias,
                                             int co, int half, int kq, uint16_t (*out)[kPos * kS64]) {
  const float bv = bias[co];
#pragma unroll
  for (int g = 0; g < kFusedSamples; ++g)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int p = 32 * half + (r >> 2) * 8 + kq * 4 + (r & 3);
      out[g][p * kS64 + co] = f2bf(gelu(acc[g][r] + bv));
    }
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) learned_estimate_kernel(
    const float2* __restrict__ grid, int num_drops, int num_rx, int words, int n, const int32_t* __restrict__ pidx,
    const float2* __restrict__ pval, int num_pilots, const double* __restrict__ noise_power, int64_t num_samples,
    const uint16_t* __restrict__ w1, const float* __restrict__ b1, const uint16_t* __restrict__ w2,
    const float* __restrict__ b2, const uint16_t* __restrict__ w3, const float* __restrict__ b3,
    const uint16_t* __restrict__ w4, const float* __restrict__ b4, const uint16_t* __restrict__ wsk,
    const float* __restrict__ bsk, int cout, float* __restrict__ est) {
  __shared__ __align__(16) uint16_t xs[kFusedSamples][kPos * kS8];
  __shared__ __align__(16) uint16_t a0[kFusedSamples][kPos * kS64];
  __shared__ __align__(16) uint16_t a1[kFusedSamples][kPos * kS64];
  const int tid = threadIdx.x;
  const int64_t s0 = static_cast<int64_t>(blockIdx.x) * kFusedSamples;
  const int64_t wn = static_cast<int64_t>(words) * n;
  const int64_t dr_count = static_cast<int64_t>(num_drops) * num_rx;
  const int pilot_words = num_pilots / kPos;
  // input: LS at the sample's own pilots, two channels per pilot word, the stream flag, the SNR
  for (int i = tid; i < kFusedSamples * kPos; i += 256) {
    const int g = i / kPos, p = i - g * kPos;
    const int64_t s = s0 + g;
    uint16_t* row = &xs[g][p * kS8];
    if (s >= num_samples) {
      for (int c = 0; c < 8; ++c) row[c] = 0;
      continue;
    }
    const int t = static_cast<int>(s / dr_count);
    const int64_t dr = s - t * dr_count;
    for (int v = 0; v < pilot_words; ++v) {
      const int j = v * kPos + p;
      const float2 x = pval[t * num_pilots + j];
      const float2 y = grid[dr * wn + pidx[t * num_pilots + j]];
      const float2 h = cscale(cmulc(x, y), 1.0f / cabs2(x));
      row[2 * v] = f2bf(h.x);
      row[2 * v + 1] = f2bf(h.y);
    }
    row[2 * pilot_words] = f2bf(t ? 1.0f : -1.0f);
    row[2 * pilot_words + 1] = f2bf(snr_feature(noise_power, dr / num_rx));
    for (int c = 2 * pilot_words + 2; c < 8; ++c) row[c] = 0;
  }
  __syncthreads();
  const int wave = tid >> 6, lane = tid & 63;
  const int half = wave & 1;
  const int m = lane & 31, kq = lane >> 5;
  floatx16 acc[kFusedSamples];
  {  // layer 1: 8 -> 64 (column tile wave / 2)
    const int co = (wave >> 1) * 32 + m;
    zero_acc(acc);
    fused_tile8<kTaps>(xs, w1, co, kDil1, half, m, kq, acc);
    store_hidden(acc, b1, co, half, kq, a0);
  }
  __syncthreads();
  {
    const int co = (wave >> 1) * 32 + m;
    zero_acc(acc);
    fused_tile64(a0, w2, co, kDil2, half, m, kq, acc);
    store_hidden(acc, b2, co, half, kq, a1);
  }
  __syncthreads();
  {
    const int co = (wave >> 1) * 32 + m;
    zero_acc(acc);
    fused_tile64(a1, w3, co, kDil3, half, m, kq, acc);
    store_hidden(acc, b3, co, half, kq, a0);
  }
  __syncthreads();
  // last layer + skip convolution straight into the estimate grid:
  // channel co = 2 (w B + u) + {re, im} of position p -> est[s][w][B p + u]
  const int blk = n / kPos;
  for (int ct = wave >> 1; ct < cout / 32; ct += 2) {
    const int co = ct * 32 + m;
    zero_acc(acc);
    fused_tile64(a0, w4, co, kDil4, half, m, kq, acc);
    floatx16 sk[kFusedSamples];
    zero_acc(sk);
    fused_tile8<kSkipTaps>(xs, wsk, co, 1, half, m, kq, sk);
    const float bv = b4[co], bs = bsk[co];
    const int c2 = co >> 1, reim = co & 1;
    const int w = c2 / blk, u = c2 - w * blk;
#pragma unroll
    for (int g = 0; g < kFusedSamples; ++g) {
      const int64_t s = s0 + g;
      if (s >= num_samples) continue;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int p = 32 * half + (r >> 2) * 8 + kq * 4 + (r & 3);
        est[2 * (s * wn + static_cast<int64_t>(w) * n + blk * p + u) + reim] = (acc[g][r] + bv) + (sk[g][r] + bs);
      }
    }
  }
}

unsigned blocks_for(int64_t n) { return static_cast<unsigned>((n + 255) / 256); }

}  // namespace

extern "C" int hpx_conv1d_forward(const uint16_t* in, int64_t num_samples, int cin, int taps, int dilation,
                                  const uint16_t* weight, const float* bias, int cout, int mode, float* out_f32,
                                  uint16_t* out_bf16, void* stream) {
  HPX_REQUIRE(dilation >= 1 && dilation < kPos, "hpx_conv1d_forward: dilation must be in [1, 63]");
  HPX_REQUIRE(num_samples >= 0 && num_samples <= 0x7fffffff, "hpx_conv1d_forward: invalid sample count");
  HPX_REQUIRE(cout >= 32 && cout % 32 == 0, "hpx_conv1d_forward: cout must be a multiple of 32 (got %d)", cout);
  HPX_REQUIRE(mode >= 0 && mode <= 2, "hpx_conv1d_forward: mode must be 0 (linear), 1 (GELU) or 2 (accumulate)");
  if (num_samples == 0) return HPX_OK;
  HPX_REQUIRE(in && weight && (out_f32 || out_bf16) && (mode != 2 || out_f32), "hpx_conv1d_forward: null pointer");
  const dim3 g(static_cast<unsigned>(num_samples));
  const hipStream_t s = as_stream(stream);
#define HPX_CONV(CIN, TAPS) \
  hipLaunchKernelGGL((conv_fwd_kernel<CIN, TAPS>), g, dim3(256), 0, s, in, weight, bias, cout, dilation, mode, \
                     out_f32, out_bf16)
  if (cin == 8 && taps == kTaps) HPX_CONV(8, kTaps);
  else if (cin == 8 && taps == kSkipTaps) HPX_CONV(8, kSkipTaps);
  else if (cin == 64 && taps == kTaps) HPX_CONV(64, kTaps);
  else if (cin == 224 && taps == kTaps) HPX_CONV(224, kTaps);
  else {
    set_error("hpx_conv1d_forward: (cin %d, taps %d) not instantiated ((8|64|224, 5), (8, 17))", cin, taps);
    return HPX_ERR_UNSUPPORTED;
  }
#undef HPX_CONV
  return check_launch("hpx_conv1d_forward");
}

extern "C" int64_t hpx_conv1d_wgrad_workspace(int64_t num_samples, int cin, int cout, int taps) {
  const int64_t chunks = (num_samples + kWgSamples - 1) / kWgSamples;
  return chunks * static_cast<int64_t>(cout) * (static_cast<int64_t>(taps) * cin + 1) *
         static_cast<int64_t>(sizeof(float));
}

extern "C" int hpx_conv1d_wgrad(const uint16_t* dz, const uint16_t* in, int64_t num_samples, int cin, int cout,
                                int taps, int dilation, float* workspace, float* grad_weight, float* grad_bias,
                                void* stream) {
  HPX_REQUIRE(dilation >= 1 && dilation < kPos, "hpx_conv1d_wgrad: dilation must be in [1, 63]");
  HPX_REQUIRE(num_samples >= 1 && num_samples <= 0x7fffffff && cin >= 1 && cout >= 1 && taps >= 1 && (taps & 1) &&
                  taps < kPos,
              "hpx_conv1d_wgrad: invalid sizes");
  HPX_REQUIRE(dz && in && workspace && grad_weight && grad_bias, "hpx_conv1d_wgrad: null pointer");
  const int chunks = static_cast<int>((num_samples + kWgSamples - 1) / kWgSamples);
  const int64_t nw = static_cast<int64_t>(cout) * taps * cin;
  float* part_w = workspace;
  float* part_b = workspace + static_cast<int64_t>(chunks) * nw;
  const hipStream_t s = as_stream(stream);
  const dim3 g(static_cast<unsigned>(chunks), static_cast<unsigned>((cout + 31) / 32),
               static_cast<unsigned>((taps * cin + 63) / 64));
  hipLaunchKernelGGL(wgrad_kernel, g, dim3(256), 0, s, dz, in, static_cast<int>(num_samples), cout, cin, taps,
                     dilation, part_w, part_b);
  hipLaunchKernelGGL(reduce_parts_kernel, dim3(blocks_for(nw)), dim3(256), 0, s, part_w, chunks, nw, grad_weight);
  hipLaunchKernelGGL(reduce_parts_kernel, dim3(blocks_for(cout)), dim3(256), 0, s, part_b, chunks,
                     static_cast<int64_t>(cout), grad_bias);
  return check_launch("hpx_conv1d_wgrad");
}

extern "C" int hpx_gelu_backward(const float* grad, const float* pre, int64_t n, uint16_t* dz, void* stream) {
  HPX_REQUIRE(n >= 0, "hpx_gelu_backward: negative size");
  if (n == 0) return HPX_OK;
  HPX_REQUIRE(grad && pre && dz, "hpx_gelu_backward: null pointer");
  hipLaunchKernelGGL(gelu_bwd_kernel, dim3(blocks_for(n)), dim3(256), 0, as_stream(stream), grad, pre, n, dz);
  return check_launch("hpx_gelu_backward");
}

extern "C" int hpx_to_bf16(const float* x, int64_t n, uint16_t* y, void* stream) {
  HPX_REQUIRE(n >= 0, "hpx_to_bf16: negative size");
  if (n == 0) return HPX_OK;
  HPX_REQUIRE(x && y, "hpx_to_bf16: null pointer");
  hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(n)), dim3(256), 0, as_stream(stream), x, n, y);
  return check_launch("hpx_to_bf16");
}

extern "C" int hpx_net_input(const hpx_c64* ls, int64_t num_drops, int num_rx, int num_tx, int pilot_words, int cin,
                             const double* noise_power, uint16_t* x, void* stream) {
  HPX_REQUIRE(num_drops >= 0 && num_drops <= 0x7fffffff && num_rx >= 1 && num_tx >= 1 && pilot_words >= 1 &&
                  cin >= 2 * pilot_words + 2,
              "hpx_net_input: invalid sizes");
  const int64_t samples = static_cast<int64_t>(num_tx) * num_drops * num_rx;
  if (samples == 0) return HPX_OK;
  HPX_REQUIRE(ls && x, "hpx_net_input: null pointer");
  hipLaunchKernelGGL(net_input_kernel, dim3(blocks_for(samples * kPos)), dim3(256), 0, as_stream(stream),
                     reinterpret_cast<const float2*>(ls), samples, static_cast<int>(num_drops), num_rx, pilot_words,
                     cin, noise_power, x);
  return check_launch("hpx_net_input");
}

extern "C" int hpx_net_output(const float* out, int64_t num_samples, int words, int num_subcarriers, hpx_c64* est,
                              void* stream) {
  HPX_REQUIRE(num_samples >= 0 && words >= 1 && num_subcarriers >= kPos && num_subcarriers % kPos == 0,
              "hpx_net_output: subcarriers must be a multiple of %d", kPos);
  if (num_samples == 0) return HPX_OK;
  HPX_REQUIRE(out && est, "hpx_net_output: null pointer");
  const int64_t n = num_samples * words * num_subcarriers;
  hipLaunchKernelGGL(net_output_kernel, dim3(blocks_for(n)), dim3(256), 0, as_stream(stream), out, num_samples, words,
                     num_subcarriers, reinterpret_cast<float2*>(est));
  return check_launch("hpx_net_output");
}

extern "C" int hpx_estimate_error(const hpx_c64* est, const hpx_c64* csi, const hpx_c64* spatial, int64_t num_drops,
                                  int num_rx, int num_tx, int64_t grid_elements, double* err, double* power,
                                  void* stream) {
  HPX_REQUIRE(num_drops >= 0 && num_drops <= 0x7fffffff && num_rx >= 1 && num_tx >= 1 && grid_elements >= 1,
              "hpx_estimate_error: invalid sizes");
  const int64_t samples = static_cast<int64_t>(num_tx) * num_drops * num_rx;
  if (samples == 0) return HPX_OK;
  HPX_REQUIRE(est && csi && spatial && err && power, "hpx_estimate_error: null pointer");
  HPX_REQUIRE(samples <= 0x7fffffff, "hpx_estimate_error: too many samples");
  hipLaunchKernelGGL(estimate_error_kernel, dim3(static_cast<unsigned>(samples)), dim3(256), 0, as_stream(stream),
                     reinterpret_cast<const float2*>(est), reinterpret_cast<const float2*>(csi),
                     reinterpret_cast<const float2*>(spatial), static_cast<int>(num_drops), num_rx, num_tx,
                     grid_elements, err, power);
  return check_launch("hpx_estimate_error");
}

extern "C" int hpx_nmse_grad(const float* out, const hpx_c64* csi, const hpx_c64* spatial, int64_t num_drops,
                             int num_rx, int num_tx, int words, int num_subcarriers, const int32_t* group,
                             int num_groups, double* work, float* grad, double* loss, void* stream) {
  HPX_REQUIRE(num_drops >= 1 && num_drops <= 0x7fffffff && num_rx >= 1 && num_tx >= 1 && words >= 1 &&
                  num_subcarriers % kPos == 0 && num_groups >= 1 && (group || num_groups == 1),
              "hpx_nmse_grad: invalid sizes");
  HPX_REQUIRE(out && csi && spatial && work && grad && loss, "hpx_nmse_grad: null pointer");
  const int64_t wn = static_cast<int64_t>(words) * num_subcarriers;
  const int64_t samples = static_cast<int64_t>(num_tx) * num_drops * num_rx;
  HPX_REQUIRE(samples <= 0x7fffffff, "hpx_nmse_grad: too many samples");
  const hipStream_t s = as_stream(stream);
  const auto* c = reinterpret_cast<const float2*>(csi);
  const auto* sp = reinterpret_cast<const float2*>(spatial);
  const int nd = static_cast<int>(num_drops);
  // work[0 .. 2 num_groups) must be zero on entry
  hipLaunchKernelGGL(loss_accumulate_kernel, dim3(static_cast<unsigned>(samples)), dim3(256), 0, s, out, c, sp, nd,
                     num_rx, num_tx, words, num_subcarriers, group, num_groups, work);
  hipLaunchKernelGGL(loss_finish_kernel, dim3(1), dim3(64), 0, s, num_groups, work, loss);
  hipLaunchKernelGGL(loss_grad_kernel, dim3(blocks_for(samples * wn)), dim3(256), 0, s, out, c, sp, nd, num_rx,
                     num_tx, words, num_subcarriers, group, num_groups, work, grad);
  return check_launch("hpx_nmse_grad");
}

extern "C" int hpx_adam_step(float* param, const float* grad, float* m1, float* m2, int64_t n, float lr, float beta1,
                             float beta2, float eps, int64_t step, void* stream) {
  HPX_REQUIRE(n >= 0 && step >= 1, "hpx_adam_step: invalid arguments");
  if (n == 0) return HPX_OK;
  HPX_REQUIRE(param && grad && m1 && m2, "hpx_adam_step: null pointer");
  const float c1 = static_cast<float>(1.0 - pow(static_cast<double>(beta1), static_cast<double>(step)));
  const float c2 = static_cast<float>(1.0 - pow(static_cast<double>(beta2), static_cast<double>(step)));
  hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(n)), dim3(256), 0, as_stream(stream), param, grad, m1, m2, n, lr,
                     beta1, beta2, eps, c1, c2);
  return check_launch("hpx_adam_step");
}

extern "C" int hpx_conv1d_weight_copies(const float* param, int cout, int cin, int taps, uint16_t* weight,
                                        uint16_t* weight_transposed, void* stream) {
  HPX_REQUIRE(cout >= 1 && cin >= 1 && taps >= 1, "hpx_conv1d_weight_copies: invalid sizes");
  HPX_REQUIRE(param && (weight || weight_transposed), "hpx_conv1d_weight_copies: null pointer");
  const int64_t n = static_cast<int64_t>(cout) * taps * cin;
  hipLaunchKernelGGL(weight_copies_kernel, dim3(blocks_for(n)), dim3(256), 0, as_stream(stream), param, cout, cin,
                     taps, weight, weight_transposed);
  return check_launch("hpx_conv1d_weight_copies");
}

extern "C" int hpx_learned_estimate(const hpx_c64* grid, int64_t num_drops, int num_rx, int num_tx, int words,
                                    int num_subcarriers, const int32_t* pilot_index, const hpx_c64* pilot_values,
                                    int num_pilots, const double* noise_power, const uint16_t* const* weights,
                                    const float* const* biases, int cout, hpx_c64* est, void* stream) {
  HPX_REQUIRE(num_drops >= 0 && num_drops <= 0x7fffffff && num_rx >= 1 && num_tx >= 1 && words >= 1,
              "hpx_learned_estimate: invalid sizes");
  HPX_REQUIRE(num_subcarriers >= kPos && num_subcarriers % kPos == 0 && num_pilots >= kPos && num_pilots % kPos == 0 &&
                  2 * (num_pilots / kPos) + 2 <= 8,
              "hpx_learned_estimate: needs 64-position pilot words (at most 3) and a multiple of 64 subcarriers");
  HPX_REQUIRE(cout == 2 * words * (num_subcarriers / kPos) && cout % 32 == 0,
              "hpx_learned_estimate: output channels %d do not match the grid", cout);
  const int64_t samples = static_cast<int64_t>(num_tx) * num_drops * num_rx;
  if (samples == 0) return HPX_OK;
  HPX_REQUIRE(grid && pilot_index && pilot_values && weights && biases && est, "hpx_learned_estimate: null pointer");
  for (int i = 0; i < 5; ++i) HPX_REQUIRE(weights[i] && biases[i], "hpx_learned_estimate: null layer %d", i);
  const int64_t blocks = (samples + kFusedSamples - 1) / kFusedSamples;
  HPX_REQUIRE(blocks <= 0x7fffffff, "hpx_learned_estimate: too many samples");
  hipLaunchKernelGGL(learned_estimate_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, as_stream(stream),
                     reinterpret_cast<const float2*>(grid), static_cast<int>(num_drops), num_rx, words,
                     num_subcarriers, pilot_index, reinterpret_cast<const float2*>(pilot_values), num_pilots, noise_power,
                     samples, weights[0], biases[0], weights[1], biases[1], weights[2], biases[2], weights[3], biases[3],
                     weights[4], biases[4], cout, reinterpret_cast<float*>(est));
  return check_launch("hpx_learned_estimate");
}
