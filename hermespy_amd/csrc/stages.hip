// Stage kernels of the link-drop path and their C ABI entry points
// (declared in include/hermespy_amd.h).  Each stage is the MI355X
// counterpart of one reference function; the fused batch kernel in
// batch.hip reuses the same device functions.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>

#include "hpx_common.hpp"
#include "hpx_detect.hpp"
#include "hpx_eig.hpp"
#include "hpx_fading.hpp"
#include "hpx_fft.hpp"
#include "hpx_mapping.hpp"

// ------------------------------------------------------------------ errors
namespace hpx {
namespace {
thread_local std::string g_last_error;
}

void set_error(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

int check_launch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s: %s", what, hipGetErrorString(e));
    return HPX_ERR_DEVICE;
  }
  return HPX_OK;
}

// Largest power-of-two Taylor block (log2) keeping the truncation bound
// sqrt(terms) x^(P+1) / (P+1)! <= 1e-6 with x = max_rate * B / 2; 0 = direct.
int taylor_log2_block(double max_rate, int terms, int max_log2) {
  double fact = 1.0;
  for (int k = 2; k <= kTaylorCoeffs; ++k) fact *= k;
  const double amp = sqrt(static_cast<double>(terms > 0 ? terms : 1));
  int best = 0;
  for (int lb = 1; lb <= max_log2; ++lb) {
    const double x = fabs(max_rate) * 0.5 * static_cast<double>(1 << lb);
    if (amp * pow(x, kTaylorCoeffs) / fact <= 1e-6) best = lb;
  }
  return best;
}

int ilog2_exact(int n) {
  if (n <= 0 || (n & (n - 1))) return -1;
  int l = 0;
  while ((1 << l) < n) ++l;
  return l;
}

}  // namespace hpx

using namespace hpx;

extern "C" int hpx_abi_version(void) { return HPX_ABI_VERSION; }

extern "C" int hpx_last_error(char* buf, size_t size) {
  if (!buf || size == 0) return HPX_ERR_INVALID;
  strncpy(buf, g_last_error.c_str(), size - 1);
  buf[size - 1] = '\0';
  return HPX_OK;
}

#define HPX_STREAM(s) static_cast<hipStream_t>(s)

// ------------------------------------------------------------------ taps (direct)
__global__ void __launch_bounds__(256) fading_taps_kernel(const double* __restrict__ sos, int terms, int64_t m_len,
                                                          float2* __restrict__ taps) {
  const int64_t m = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int path = blockIdx.y;
  if (m >= m_len) return;
  taps[path * m_len + m] = tap_direct(sos + static_cast<int64_t>(path) * terms * 3, terms, m);
}

extern "C" int hpx_fading_taps(const double* sos, int num_paths, int num_terms, int64_t num_samples, hpx_c64* taps,
                               void* stream) {
  HPX_REQUIRE(num_paths > 0 && num_paths <= 65535 && num_terms > 0 && num_samples >= 0,
              "hpx_fading_taps: invalid sizes (paths %d, terms %d, samples %lld)", num_paths, num_terms,
              (long long)num_samples);
  if (num_samples == 0) return HPX_OK;
  HPX_REQUIRE(sos && taps, "hpx_fading_taps: null pointer");
  const dim3 grid(static_cast<unsigned>((num_samples + 255) / 256), num_paths);
  hipLaunchKernelGGL(fading_taps_kernel, grid, dim3(256), 0, HPX_STREAM(stream), sos, num_terms, num_samples,
                     reinterpret_cast<float2*>(taps));
  return check_launch("hpx_fading_taps");
}

// ------------------------------------------------------------------ TDL propagation
constexpr int kMaxAntennas = 8;
constexpr int kTdlThreads = 256;

// One output sample per thread; Taylor coefficients of the block range the
// workgroup touches live in LDS (direct evaluation when use_taylor == 0).
__global__ void __launch_bounds__(kTdlThreads) tdl_propagate_kernel(
    const float2* __restrict__ x, int ntx, int64_t ld_x, int64_t m_len, const double* __restrict__ sos, int paths,
    int terms, int log2b, const int32_t* __restrict__ delays, int32_t max_delay, const float2* __restrict__ spatial,
    int nrx, float2* __restrict__ y, int64_t ld_y, int use_taylor) {
  extern __shared__ float2 coeff[];  // [paths][nb][kTaylorCoeffs]
  const int64_t n0 = static_cast<int64_t>(blockIdx.x) * kTdlThreads;
  const int64_t out_len = m_len + max_delay;
  const TaylorBlocks tb{log2b};
  int64_t b_lo = 0;
  int nb = 0;
  if (use_taylor) {
    const int64_t m_lo = std::max<int64_t>(0, n0 - max_delay);
    const int64_t m_hi = std::min<int64_t>(m_len - 1, n0 + kTdlThreads - 1);
    if (m_hi >= m_lo) {
      b_lo = tb.block_of(m_lo);
      nb = static_cast<int>(tb.block_of(m_hi) - b_lo + 1);
    }
    for (int i = threadIdx.x; i < paths * nb; i += kTdlThreads) {
      const int p = i / nb, b = i - p * nb;
      taylor_coeffs(sos + static_cast<int64_t>(p) * terms * 3, terms, tb.centre(b_lo + b), tb.half_width(),
                    coeff + (p * nb + b) * kTaylorCoeffs);
    }
    __syncthreads();
  }
  const int64_t n = n0 + threadIdx.x;
  if (n >= out_len) return;
  float2 z[kMaxAntennas];
#pragma unroll
  for (int t = 0; t < kMaxAntennas; ++t) z[t] = make_float2(0.f, 0.f);
  for (int p = 0; p < paths; ++p) {
    const int d = delays[p];
    const int64_t m = n - d;
    if (d < 0 || d > max_delay || m < 0 || m >= m_len) continue;
    float2 h;
    if (use_taylor) {
      const int b = static_cast<int>(tb.block_of(m) - b_lo);
      h = taylor_eval(coeff + (p * nb + b) * kTaylorCoeffs, tb.offset(m));
    } else {
      h = tap_direct(sos + static_cast<int64_t>(p) * terms * 3, terms, m);
    }
#pragma unroll
    for (int t = 0; t < kMaxAntennas; ++t)
      if (t < ntx) z[t] = cfma(h, x[t * ld_x + m], z[t]);
  }
  for (int r = 0; r < nrx; ++r) {
    float2 acc = make_float2(0.f, 0.f);
#pragma unroll
    for (int t = 0; t < kMaxAntennas; ++t)
      if (t < ntx) acc = cfma(spatial[r * ntx + t], z[t], acc);
    y[r * ld_y + n] = acc;
  }
}

extern "C" int hpx_tdl_propagate(const hpx_c64* x, int num_tx, int64_t ld_x, int64_t num_samples, const double* sos,
                                 int num_paths, int num_terms, double max_abs_rate, const int32_t* delays,
                                 int32_t max_delay, const hpx_c64* spatial, int num_rx, hpx_c64* y, int64_t ld_y,
                                 void* stream) {
  HPX_REQUIRE(x && sos && delays && spatial && y, "hpx_tdl_propagate: null pointer");
  HPX_REQUIRE(num_tx >= 1 && num_tx <= kMaxAntennas && num_rx >= 1 && num_rx <= kMaxAntennas,
              "hpx_tdl_propagate: antenna counts must be in [1, %d] (tx %d, rx %d)", kMaxAntennas, num_tx, num_rx);
  HPX_REQUIRE(num_paths >= 1 && num_terms >= 1 && max_delay >= 0 && num_samples >= 0,
              "hpx_tdl_propagate: invalid sizes");
  HPX_REQUIRE(ld_x >= num_samples && ld_y >= num_samples + max_delay, "hpx_tdl_propagate: leading dimension too small");
  const int64_t out_len = num_samples + max_delay;
  if (out_len == 0) return HPX_OK;
  int log2b = taylor_log2_block(max_abs_rate, num_terms, 11);
  // Taylor blocks only pay when a block spans a workgroup's samples.
  int use_taylor = log2b >= 8;
  size_t lds = 0;
  if (use_taylor) {
    const int64_t nb = (kTdlThreads + max_delay) / (1 << log2b) + 2;
    lds = static_cast<size_t>(num_paths) * nb * kTaylorCoeffs * sizeof(float2);
    if (lds > 48 * 1024) {
      use_taylor = 0;
      lds = 0;
    }
  }
  const dim3 grid(static_cast<unsigned>((out_len + kTdlThreads - 1) / kTdlThreads));
  hipLaunchKernelGGL(tdl_propagate_kernel, grid, dim3(kTdlThreads), lds, HPX_STREAM(stream),
                     reinterpret_cast<const float2*>(x), num_tx, ld_x, num_samples, sos, num_paths, num_terms, log2b,
                     delays, max_delay, reinterpret_cast<const float2*>(spatial), num_rx, reinterpret_cast<float2*>(y),
                     ld_y, use_taylor);
  return check_launch("hpx_tdl_propagate");
}

// ------------------------------------------------------------------ dense impulse response
// csi[r][t][n][d] = S[r][t] * sum over paths of delay d of h_p(n); one thread per (n, d).
__global__ void __launch_bounds__(256) fading_state_kernel(const double* __restrict__ sos, int paths, int terms,
                                                           const int32_t* __restrict__ delays, int num_taps,
                                                           const float2* __restrict__ spatial, int nrx, int ntx,
                                                           int64_t m_len, float2* __restrict__ csi) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m_len * num_taps) return;
  const int64_t n = i / num_taps;
  const int d = static_cast<int>(i - n * num_taps);
  float2 h = make_float2(0.f, 0.f);
  for (int p = 0; p < paths; ++p)
    if (delays[p] == d) h = cadd(h, tap_direct(sos + static_cast<int64_t>(p) * terms * 3, terms, n));
  const int64_t plane = m_len * num_taps;
  for (int e = 0; e < nrx * ntx; ++e) csi[e * plane + i] = cmul(spatial[e], h);
}

extern "C" int hpx_fading_state(const double* sos, int num_paths, int num_terms, const int32_t* delays,
                                int32_t num_taps, const hpx_c64* spatial, int num_rx, int num_tx, int64_t num_samples,
                                hpx_c64* csi, void* stream) {
  HPX_REQUIRE(sos && delays && spatial && csi, "hpx_fading_state: null pointer");
  HPX_REQUIRE(num_paths >= 1 && num_terms >= 1 && num_taps >= 1 && num_samples >= 0 && num_rx >= 1 && num_tx >= 1,
              "hpx_fading_state: invalid sizes");
  const int64_t total = num_samples * num_taps;
  if (total == 0) return HPX_OK;
  HPX_REQUIRE(total / 256 < (int64_t(1) << 31), "hpx_fading_state: too many samples");
  hipLaunchKernelGGL(fading_state_kernel, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0,
                     HPX_STREAM(stream), sos, num_paths, num_terms, delays, num_taps,
                     reinterpret_cast<const float2*>(spatial), num_rx, num_tx, num_samples,
                     reinterpret_cast<float2*>(csi));
  return check_launch("hpx_fading_state");
}

// y[r][n] = sum_{t, d} csi[r][t][n - d][d] x[t][n - d]; one thread per (r, n).
__global__ void __launch_bounds__(256) csi_propagate_kernel(const float2* __restrict__ csi, int nrx, int ntx,
                                                            int64_t csi_samples, int num_taps,
                                                            const float2* __restrict__ x, int64_t ld_x, int64_t m_len,
                                                            float2* __restrict__ y, int64_t ld_y) {
  const int64_t n = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  const int64_t out_len = m_len + num_taps - 1;
  if (n >= out_len) return;
  float2 acc = make_float2(0.f, 0.f);
  for (int t = 0; t < ntx; ++t) {
    const float2* h = csi + (static_cast<int64_t>(r) * ntx + t) * csi_samples * num_taps;
    for (int d = 0; d < num_taps; ++d) {
      const int64_t m = n - d;
      if (m < 0 || m >= m_len) continue;
      acc = cfma(h[m * num_taps + d], x[t * ld_x + m], acc);
    }
  }
  y[r * ld_y + n] = acc;
}

extern "C" int hpx_csi_propagate(const hpx_c64* csi, int num_rx, int num_tx, int64_t csi_samples, int32_t num_taps,
                                 const hpx_c64* x, int64_t ld_x, int64_t num_samples, hpx_c64* y, int64_t ld_y,
                                 void* stream) {
  HPX_REQUIRE(csi && x && y, "hpx_csi_propagate: null pointer");
  HPX_REQUIRE(num_rx >= 1 && num_rx <= 65535 && num_tx >= 1 && num_taps >= 1 && num_samples >= 0,
              "hpx_csi_propagate: invalid sizes");
  HPX_REQUIRE(csi_samples >= num_samples, "hpx_csi_propagate: the channel state covers %lld samples, the signal %lld",
              static_cast<long long>(csi_samples), static_cast<long long>(num_samples));
  HPX_REQUIRE(ld_x >= num_samples && ld_y >= num_samples + num_taps - 1, "hpx_csi_propagate: leading dimension too small");
  const int64_t out_len = num_samples + num_taps - 1;
  if (num_samples == 0) return HPX_OK;
  hipLaunchKernelGGL(csi_propagate_kernel, dim3(static_cast<unsigned>((out_len + 255) / 256), num_rx), dim3(256), 0,
                     HPX_STREAM(stream), reinterpret_cast<const float2*>(csi), num_rx, num_tx, csi_samples, num_taps,
                     reinterpret_cast<const float2*>(x), ld_x, num_samples, reinterpret_cast<float2*>(y), ld_y);
  return check_launch("hpx_csi_propagate");
}

// ------------------------------------------------------------------ AWGN
__global__ void __launch_bounds__(256) awgn_kernel(float2* __restrict__ sig, int64_t ld, int64_t len, uint64_t seed,
                                                   uint32_t drop, float sigma) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;  // sample pair
  const int r = blockIdx.y;
  const int64_t n = 2 * i;
  if (n >= len) return;
  const uint4 w = stream_call(seed, drop, HPX_STREAM_NOISE + r, static_cast<uint32_t>(i));
  float2* row = sig + r * ld;
  const float2 a = complex_normal_f32(w.x, w.y);
  row[n] = make_float2(fmaf(sigma, a.x, row[n].x), fmaf(sigma, a.y, row[n].y));
  if (n + 1 < len) {
    const float2 b = complex_normal_f32(w.z, w.w);
    row[n + 1] = make_float2(fmaf(sigma, b.x, row[n + 1].x), fmaf(sigma, b.y, row[n + 1].y));
  }
}

extern "C" int hpx_awgn(hpx_c64* signal, int num_streams, int64_t ld, int64_t num_samples, uint64_t seed,
                        uint32_t drop, double power, void* stream) {
  HPX_REQUIRE(num_streams >= 1 && num_streams <= 65535 && ld >= num_samples && num_samples >= 0,
              "hpx_awgn: invalid sizes");
  HPX_REQUIRE(power >= 0.0, "hpx_awgn: Noise power of a noise realization must be non-negative.");
  HPX_REQUIRE(num_samples <= (int64_t(1) << 33), "hpx_awgn: too many samples for the 32-bit counter");
  if (num_samples == 0 || power == 0.0) return HPX_OK;
  HPX_REQUIRE(signal, "hpx_awgn: null pointer");
  const int64_t pairs = (num_samples + 1) / 2;
  const dim3 grid(static_cast<unsigned>((pairs + 255) / 256), num_streams);
  hipLaunchKernelGGL(awgn_kernel, grid, dim3(256), 0, HPX_STREAM(stream), reinterpret_cast<float2*>(signal), ld,
                     num_samples, seed, drop, static_cast<float>(sqrt(0.5 * power)));
  return check_launch("hpx_awgn");
}

// ------------------------------------------------------------------ mapping
namespace hpx {
inline bool supported_modulation(int order, int is_complex) {
  Modulation m;
  return modulation(order, is_complex, &m);
}
}  // namespace hpx

extern "C" int hpx_constellation_host(int modulation_order, int is_complex, hpx_c128* points_host) {
  HPX_REQUIRE(points_host, "hpx_constellation_host: null pointer");
  if (!supported_modulation(modulation_order, is_complex)) {
    set_error("Modulation (%d) not supported", modulation_order);
    return HPX_ERR_UNSUPPORTED;
  }
  for (int l = 0; l < modulation_order; ++l) {
    const double2 p = constellation_point(modulation_order, is_complex, static_cast<uint32_t>(l));
    points_host[l].re = p.x;
    points_host[l].im = p.y;
  }
  return HPX_OK;
}

extern "C" int hpx_demap_hard_host(const hpx_c128* symbols_host, int64_t num_symbols, int modulation_order,
                                   int is_complex, uint8_t* bits_host) {
  HPX_REQUIRE(symbols_host && bits_host, "hpx_demap_hard_host: null pointer");
  Modulation m;
  if (!modulation(modulation_order, is_complex, &m)) {
    set_error("Unsupported modulation scheme");
    return HPX_ERR_UNSUPPORTED;
  }
  for (int64_t i = 0; i < num_symbols; ++i) {
    const uint32_t b = demap_symbol(m, symbols_host[i].re, symbols_host[i].im);
    for (int k = 0; k < m.bps; ++k) bits_host[i * m.bps + k] = static_cast<uint8_t>((b >> k) & 1u);
  }
  return HPX_OK;
}

extern "C" int hpx_demap_hard_f32_host(const hpx_c64* symbols_host, int64_t num_symbols, int modulation_order,
                                       uint8_t* bits_host) {
  HPX_REQUIRE(symbols_host && bits_host, "hpx_demap_hard_f32_host: null pointer");
  Modulation m;
  if (modulation_order == 2 || !modulation(modulation_order, 1, &m)) {
    set_error("Unsupported modulation scheme");
    return HPX_ERR_UNSUPPORTED;
  }
  const DemapBounds d = demap_bounds(m);
  for (int64_t i = 0; i < num_symbols; ++i) {
    const float re = symbols_host[i].re, im = symbols_host[i].im;
    const uint32_t b = demap_symbol_f32(m, &d, re, im);
    // the kernels' compile-time-order form must decide the same
    uint32_t bt;
    switch (m.pam) {
      case 2: bt = demap_symbol_f32_reg<2>(re, im, d.a, d.b); break;
      case 4: bt = demap_symbol_f32_reg<4>(re, im, d.a, d.b); break;
      case 8: bt = demap_symbol_f32_reg<8>(re, im, d.a, d.b); break;
      default: bt = demap_symbol_f32_reg<16>(re, im, d.a, d.b); break;
    }
    if (bt != b) {
      set_error("hpx_demap_hard_f32_host: demapper forms disagree at symbol %lld", static_cast<long long>(i));
      return HPX_ERR_INVALID;
    }
    for (int k = 0; k < m.bps; ++k) bits_host[i * m.bps + k] = static_cast<uint8_t>((b >> k) & 1u);
  }
  return HPX_OK;
}

__global__ void __launch_bounds__(256) map_kernel(const uint8_t* __restrict__ bits, int64_t nsym, int order,
                                                  int is_complex, int bps, float2* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nsym) return;
  uint32_t label = 0;
  for (int k = 0; k < bps; ++k) label = (label << 1) | (bits[i * bps + k] & 1u);
  const double2 p = constellation_point(order, is_complex, label);
  out[i] = make_float2(static_cast<float>(p.x), static_cast<float>(p.y));
}

extern "C" int hpx_map_symbols(const uint8_t* bits, int64_t num_symbols, int modulation_order, int is_complex,
                               hpx_c64* symbols, void* stream) {
  Modulation m;
  if (!modulation(modulation_order, is_complex, &m)) {
    set_error("Modulation (%d) not supported", modulation_order);
    return HPX_ERR_UNSUPPORTED;
  }
  HPX_REQUIRE(num_symbols >= 0, "hpx_map_symbols: negative length");
  if (num_symbols == 0) return HPX_OK;
  HPX_REQUIRE(bits && symbols, "hpx_map_symbols: null pointer");
  hipLaunchKernelGGL(map_kernel, dim3(static_cast<unsigned>((num_symbols + 255) / 256)), dim3(256), 0,
                     HPX_STREAM(stream), bits, num_symbols, modulation_order, is_complex, m.bps,
                     reinterpret_cast<float2*>(symbols));
  return check_launch("hpx_map_symbols");
}

__global__ void __launch_bounds__(256) demap_kernel(const float2* __restrict__ sym, int64_t nsym, Modulation m,
                                                    uint8_t* __restrict__ bits) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nsym) return;
  const float2 s = sym[i];
  const uint32_t b = demap_symbol(m, static_cast<double>(s.x), static_cast<double>(s.y));
  for (int k = 0; k < m.bps; ++k) bits[i * m.bps + k] = static_cast<uint8_t>((b >> k) & 1u);
}

extern "C" int hpx_demap_hard(const hpx_c64* symbols, int64_t num_symbols, int modulation_order, int is_complex,
                              uint8_t* bits, void* stream) {
  Modulation m;
  if (!modulation(modulation_order, is_complex, &m)) {
    set_error("Unsupported modulation scheme");
    return HPX_ERR_UNSUPPORTED;
  }
  HPX_REQUIRE(num_symbols >= 0, "hpx_demap_hard: negative length");
  if (num_symbols == 0) return HPX_OK;
  HPX_REQUIRE(symbols && bits, "hpx_demap_hard: null pointer");
  hipLaunchKernelGGL(demap_kernel, dim3(static_cast<unsigned>((num_symbols + 255) / 256)), dim3(256), 0,
                     HPX_STREAM(stream), reinterpret_cast<const float2*>(symbols), num_symbols, m, bits);
  return check_launch("hpx_demap_hard");
}

// ------------------------------------------------------------------ OFDM
constexpr int kOfdmThreads = 256;

__device__ void fill_twiddles(float2* tw, int n) {
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    double s, c;
    sincospi(-2.0 * static_cast<double>(k) / static_cast<double>(n), &s, &c);
    tw[k] = make_float2(static_cast<float>(c), static_cast<float>(s));
  }
}

// One workgroup per (stream, word).
__global__ void __launch_bounds__(kOfdmThreads) ofdm_modulate_kernel(const float2* __restrict__ grid, int words, int n,
                                                                     int log2n, int cp, int dc, float2* __restrict__ out,
                                                                     int64_t ld_out) {
  extern __shared__ float2 lds[];
  float2* buf = lds;
  float2* tw = lds + n;
  const int w = blockIdx.x, s = blockIdx.y;
  fill_twiddles(tw, n);
  const float scale = static_cast<float>(1.0 / sqrt(static_cast<double>(n)));
  const float2* row = grid + (static_cast<int64_t>(s) * words + w) * n;
  for (int q = threadIdx.x; q < n; q += kOfdmThreads) buf[subcarrier_bin(q, n, dc)] = cscale(row[q], scale);
  __syncthreads();
  lds_fft<+1>(buf, 1, n, log2n, tw);
  float2* dst = out + static_cast<int64_t>(s) * ld_out + static_cast<int64_t>(w) * (n + cp);
  for (int i = threadIdx.x; i < n + cp; i += kOfdmThreads) dst[i] = buf[i < cp ? n - cp + i : i - cp];
}

extern "C" int hpx_ofdm_modulate(const hpx_c64* grid, int num_streams, int words, int num_subcarriers, int prefix,
                                 int dc_suppression, hpx_c64* out, int64_t ld_out, void* stream) {
  HPX_REQUIRE(grid && out, "hpx_ofdm_modulate: null pointer");
  const int log2n = ilog2_exact(num_subcarriers);
  if (log2n < 1 || num_subcarriers > 4096) {
    set_error("hpx_ofdm_modulate: num_subcarriers must be a power of two in [2, 4096] (got %d)", num_subcarriers);
    return HPX_ERR_UNSUPPORTED;
  }
  HPX_REQUIRE(num_streams >= 1 && num_streams <= 65535 && words >= 1 && prefix >= 0 && prefix <= num_subcarriers,
              "hpx_ofdm_modulate: invalid sizes");
  HPX_REQUIRE(ld_out >= static_cast<int64_t>(words) * (num_subcarriers + prefix), "hpx_ofdm_modulate: ld_out too small");
  const size_t lds = 2 * static_cast<size_t>(num_subcarriers) * sizeof(float2);
  hipLaunchKernelGGL(ofdm_modulate_kernel, dim3(words, num_streams), dim3(kOfdmThreads), lds, HPX_STREAM(stream),
                     reinterpret_cast<const float2*>(grid), words, num_subcarriers, log2n, prefix, dc_suppression,
                     reinterpret_cast<float2*>(out), ld_out);
  return check_launch("hpx_ofdm_modulate");
}

__global__ void __launch_bounds__(kOfdmThreads) ofdm_demodulate_kernel(const float2* __restrict__ frame, int64_t ld_in,
                                                                       int words, int n, int log2n, int cp, int dc,
                                                                       float2* __restrict__ grid) {
  extern __shared__ float2 lds[];
  float2* buf = lds;
  float2* tw = lds + n;
  const int w = blockIdx.x, s = blockIdx.y;
  fill_twiddles(tw, n);
  const float2* src = frame + static_cast<int64_t>(s) * ld_in + static_cast<int64_t>(w) * (n + cp) + cp;
  for (int i = threadIdx.x; i < n; i += kOfdmThreads) buf[i] = src[i];
  __syncthreads();
  lds_fft<-1>(buf, 1, n, log2n, tw);
  const float scale = static_cast<float>(1.0 / sqrt(static_cast<double>(n)));
  float2* row = grid + (static_cast<int64_t>(s) * words + w) * n;
  for (int q = threadIdx.x; q < n; q += kOfdmThreads) row[q] = cscale(buf[subcarrier_bin(q, n, dc)], scale);
}

extern "C" int hpx_ofdm_demodulate(const hpx_c64* frame, int num_streams, int64_t ld_in, int words,
                                   int num_subcarriers, int prefix, int dc_suppression, hpx_c64* grid,
                                   void* stream) {
  HPX_REQUIRE(frame && grid, "hpx_ofdm_demodulate: null pointer");
  const int log2n = ilog2_exact(num_subcarriers);
  if (log2n < 1 || num_subcarriers > 4096) {
    set_error("hpx_ofdm_demodulate: num_subcarriers must be a power of two in [2, 4096] (got %d)", num_subcarriers);
    return HPX_ERR_UNSUPPORTED;
  }
  HPX_REQUIRE(num_streams >= 1 && num_streams <= 65535 && words >= 1 && prefix >= 0 && prefix <= num_subcarriers,
              "hpx_ofdm_demodulate: invalid sizes");
  HPX_REQUIRE(ld_in >= static_cast<int64_t>(words) * (num_subcarriers + prefix), "hpx_ofdm_demodulate: ld_in too small");
  const size_t lds = 2 * static_cast<size_t>(num_subcarriers) * sizeof(float2);
  hipLaunchKernelGGL(ofdm_demodulate_kernel, dim3(words, num_streams), dim3(kOfdmThreads), lds, HPX_STREAM(stream),
                     reinterpret_cast<const float2*>(frame), ld_in, words, num_subcarriers, log2n, prefix,
                     dc_suppression, reinterpret_cast<float2*>(grid));
  return check_launch("hpx_ofdm_demodulate");
}

// ------------------------------------------------------------------ ideal CSI
constexpr int kCsiThreads = 256;

// One workgroup per word: word means of every path (float64 Dirichlet form),
// then one subcarrier per thread.
__global__ void __launch_bounds__(kCsiThreads) ideal_csi_kernel(const double* __restrict__ sos, int paths, int terms,
                                                                const int32_t* __restrict__ delays, int keep_below,
                                                                const float2* __restrict__ spatial, int nrx, int ntx,
                                                                int words, int n, int cp, int dc,
                                                                float2* __restrict__ csi) {
  extern __shared__ float2 lds[];
  float2* abar = lds;           // [paths]
  float2* tw = lds + paths;     // [n]
  const int w = blockIdx.x;
  fill_twiddles(tw, n);
  const double m0 = static_cast<double>(w) * (n + cp) + cp;
  for (int p = threadIdx.x; p < paths; p += kCsiThreads) {
    const double2 a = tap_mean(sos + static_cast<int64_t>(p) * terms * 3, terms, m0, n);
    abar[p] = make_float2(static_cast<float>(a.x), static_cast<float>(a.y));
  }
  __syncthreads();
  for (int q = threadIdx.x; q < n; q += kCsiThreads) {
    const int bin = subcarrier_bin(q, n, dc);
    float2 c = make_float2(0.f, 0.f);
    for (int p = 0; p < paths; ++p) {
      const int d = delays[p];
      if (d < 0 || d >= keep_below) continue;
      const int idx = static_cast<int>((static_cast<int64_t>(bin) * d) & (n - 1));
      c = cfma(abar[p], tw[idx], c);
    }
    for (int r = 0; r < nrx; ++r)
      for (int t = 0; t < ntx; ++t)
        csi[((static_cast<int64_t>(r) * ntx + t) * words + w) * n + q] = cmul(spatial[r * ntx + t], c);
  }
}

extern "C" int hpx_ideal_csi(const double* sos, int num_paths, int num_terms, const int32_t* delays, int32_t num_taps,
                             const hpx_c64* spatial, int num_rx, int num_tx, int words, int num_subcarriers,
                             int prefix, int dc_suppression, hpx_c64* csi, void* stream) {
  HPX_REQUIRE(sos && delays && spatial && csi, "hpx_ideal_csi: null pointer");
  const int log2n = ilog2_exact(num_subcarriers);
  if (log2n < 1 || num_subcarriers > 4096) {
    set_error("hpx_ideal_csi: num_subcarriers must be a power of two in [2, 4096] (got %d)", num_subcarriers);
    return HPX_ERR_UNSUPPORTED;
  }
  HPX_REQUIRE(num_paths >= 1 && num_paths <= 4096 && num_terms >= 1 && num_rx >= 1 && num_tx >= 1 && words >= 1 &&
                  prefix >= 0 && num_taps >= 1,
              "hpx_ideal_csi: invalid sizes");
  const int keep_below = std::min<int>(num_taps, num_subcarriers);
  const size_t lds = (static_cast<size_t>(num_paths) + num_subcarriers) * sizeof(float2);
  hipLaunchKernelGGL(ideal_csi_kernel, dim3(words), dim3(kCsiThreads), lds, HPX_STREAM(stream), sos, num_paths,
                     num_terms, delays, keep_below, reinterpret_cast<const float2*>(spatial), num_rx, num_tx, words,
                     num_subcarriers, prefix, dc_suppression, reinterpret_cast<float2*>(csi));
  return check_launch("hpx_ideal_csi");
}

// ------------------------------------------------------------------ equalisation
template <int NT>
__global__ void __launch_bounds__(256) equalize_kernel(const float2* __restrict__ y, const float2* __restrict__ h,
                                                       int nrx, int64_t k_len, float s2, float2* __restrict__ x) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= k_len) return;
  float2 out[NT];
  detect_rows<NT>(
      nrx, [&](int r, int t) { return h[(static_cast<int64_t>(r) * NT + t) * k_len + k]; },
      [&](int r) { return y[static_cast<int64_t>(r) * k_len + k]; }, s2, out);
#pragma unroll
  for (int t = 0; t < NT; ++t) x[t * k_len + k] = out[t];
}

// num_rx == 1: y / sum_t h (waveform.py:242-244)
__global__ void __launch_bounds__(256) equalize_siso_kernel(const float2* __restrict__ y, const float2* __restrict__ h,
                                                            int ntx, int64_t k_len, float2* __restrict__ x) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= k_len) return;
  float2 s = make_float2(0.f, 0.f);
  for (int t = 0; t < ntx; ++t) s = cadd(s, h[t * k_len + k]);
  const float2 v = y[k];
  const float inv = 1.0f / cabs2(s);
  x[k] = cscale(cmulc(s, v), inv);
}

extern "C" int hpx_equalize(const hpx_c64* y, const hpx_c64* h, int num_rx, int num_tx, int64_t num_re, int mode,
                            double noise_variance, hpx_c64* x, void* stream) {
  HPX_REQUIRE(mode == HPX_DETECT_ZF || mode == HPX_DETECT_MMSE, "hpx_equalize: unknown mode %d", mode);
  HPX_REQUIRE(num_rx >= 1 && num_tx >= 1 && num_re >= 0, "hpx_equalize: invalid sizes");
  HPX_REQUIRE(noise_variance >= 0.0, "hpx_equalize: noise variance must be non-negative");
  if (num_re == 0) return HPX_OK;
  HPX_REQUIRE(y && h && x, "hpx_equalize: null pointer");
  const dim3 grid(static_cast<unsigned>((num_re + 255) / 256));
  const float2* yy = reinterpret_cast<const float2*>(y);
  const float2* hh = reinterpret_cast<const float2*>(h);
  float2* xx = reinterpret_cast<float2*>(x);
  const float s2 = mode == HPX_DETECT_MMSE ? static_cast<float>(noise_variance) : 0.0f;
  if (num_rx == 1 && mode == HPX_DETECT_ZF) {
    if (num_tx != 1) {
      set_error("hpx_equalize: a single receive stream equalises one transmit stream");
      return HPX_ERR_UNSUPPORTED;
    }
    hipLaunchKernelGGL(equalize_siso_kernel, grid, dim3(256), 0, HPX_STREAM(stream), yy, hh, num_tx, num_re, xx);
    return check_launch("hpx_equalize");
  }
  if (num_tx > num_rx || num_rx > kMaxAntennas) {
    set_error("hpx_equalize: requires num_tx <= num_rx <= %d (tx %d, rx %d)", kMaxAntennas, num_tx, num_rx);
    return HPX_ERR_UNSUPPORTED;
  }
  switch (num_tx) {
    case 1: hipLaunchKernelGGL(equalize_kernel<1>, grid, dim3(256), 0, HPX_STREAM(stream), yy, hh, num_rx, num_re, s2, xx); break;
    case 2: hipLaunchKernelGGL(equalize_kernel<2>, grid, dim3(256), 0, HPX_STREAM(stream), yy, hh, num_rx, num_re, s2, xx); break;
    case 3: hipLaunchKernelGGL(equalize_kernel<3>, grid, dim3(256), 0, HPX_STREAM(stream), yy, hh, num_rx, num_re, s2, xx); break;
    case 4: hipLaunchKernelGGL(equalize_kernel<4>, grid, dim3(256), 0, HPX_STREAM(stream), yy, hh, num_rx, num_re, s2, xx); break;
    case 8: hipLaunchKernelGGL(equalize_kernel<8>, grid, dim3(256), 0, HPX_STREAM(stream), yy, hh, num_rx, num_re, s2, xx); break;
    default:
      set_error("hpx_equalize: num_tx %d not instantiated (1, 2, 3, 4, 8)", num_tx);
      return HPX_ERR_UNSUPPORTED;
  }
  return check_launch("hpx_equalize");
}

// ------------------------------------------------------------------ bit errors
__global__ void __launch_bounds__(256) count_errors_kernel(const uint8_t* __restrict__ a, const uint8_t* __restrict__ b,
                                                           int64_t n, unsigned long long* __restrict__ errors) {
  unsigned long long local = 0;
  const int64_t words = n / 4;
  const uint32_t* aw = reinterpret_cast<const uint32_t*>(a);
  const uint32_t* bw = reinterpret_cast<const uint32_t*>(b);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  const int64_t tid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool aligned = ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b)) & 3u) == 0;
  if (aligned) {
    for (int64_t i = tid; i < words; i += stride) local += __popc((aw[i] ^ bw[i]) & 0x01010101u);
    for (int64_t i = words * 4 + tid; i < n; i += stride) local += ((a[i] ^ b[i]) & 1u);
  } else {
    for (int64_t i = tid; i < n; i += stride) local += ((a[i] ^ b[i]) & 1u);
  }
  // wavefront reduction then one atomic per wave
  for (int off = 32; off > 0; off >>= 1) local += __shfl_down(local, off, 64);
  if ((threadIdx.x & 63) == 0 && local) atomicAdd(errors, local);
}

extern "C" int hpx_count_errors(const uint8_t* a, const uint8_t* b, int64_t num_bits, unsigned long long* errors,
                                void* stream) {
  HPX_REQUIRE(num_bits >= 0, "hpx_count_errors: negative length");
  if (num_bits == 0) return HPX_OK;
  HPX_REQUIRE(a && b && errors, "hpx_count_errors: null pointer");
  const int64_t work = std::max<int64_t>(1, num_bits / 4);
  const unsigned blocks = static_cast<unsigned>(std::min<int64_t>((work + 255) / 256, 4096));
  hipLaunchKernelGGL(count_errors_kernel, dim3(blocks), dim3(256), 0, HPX_STREAM(stream), a, b, num_bits, errors);
  return check_launch("hpx_count_errors");
}

// ------------------------------------------------------------------ per-drop detector (host twin)
extern "C" int hpx_spatial_detector_host(int num_rx, int num_tx, const hpx_c128* spatial_host, double* lambda_host,
                                         hpx_c128* v_host, hpx_c128* a_host) {
  HPX_REQUIRE(spatial_host && lambda_host && v_host && a_host, "hpx_spatial_detector_host: null pointer");
  HPX_REQUIRE(num_tx >= 1 && num_tx <= kEigMax && num_rx >= 1 && num_rx <= kEigMax,
              "hpx_spatial_detector_host: sizes must be in [1, %d]", kEigMax);
  double work[kEigWork];
  spatial_detector(num_rx, num_tx, reinterpret_cast<const double*>(spatial_host), lambda_host,
                   reinterpret_cast<double*>(v_host), reinterpret_cast<double*>(a_host), work);
  return HPX_OK;
}

// ------------------------------------------------------------------ random streams
__global__ void __launch_bounds__(256) stream_bits_kernel(uint64_t seed, uint32_t drop, uint32_t stream_id,
                                                          int64_t nbits, uint8_t* __restrict__ bits) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;  // one Philox call = 128 bits
  if (c * 128 >= nbits) return;
  const uint4 w = stream_call(seed, drop, stream_id, static_cast<uint32_t>(c));
  const uint32_t comp[4] = {w.x, w.y, w.z, w.w};
  for (int k = 0; k < 128; ++k) {
    const int64_t i = c * 128 + k;
    if (i < nbits) bits[i] = static_cast<uint8_t>((comp[k >> 5] >> (k & 31)) & 1u);
  }
}

extern "C" int hpx_stream_bits(uint64_t seed, uint32_t drop, uint32_t stream_id, int64_t num_bits, uint8_t* bits,
                               void* stream) {
  HPX_REQUIRE(num_bits >= 0 && num_bits <= (int64_t(1) << 38), "hpx_stream_bits: invalid length");
  if (num_bits == 0) return HPX_OK;
  HPX_REQUIRE(bits, "hpx_stream_bits: null pointer");
  const int64_t calls = (num_bits + 127) / 128;
  hipLaunchKernelGGL(stream_bits_kernel, dim3(static_cast<unsigned>((calls + 255) / 256)), dim3(256), 0,
                     HPX_STREAM(stream), seed, drop, stream_id, num_bits, bits);
  return check_launch("hpx_stream_bits");
}

__global__ void __launch_bounds__(256) realization_normals_kernel(uint64_t seed, uint32_t drop, int64_t count,
                                                                  double* __restrict__ out) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (2 * c >= count) return;
  const double2 z = normal_pair_f64(stream_call(seed, drop, HPX_STREAM_REALIZATION, static_cast<uint32_t>(c)));
  out[2 * c] = z.x;
  if (2 * c + 1 < count) out[2 * c + 1] = z.y;
}

extern "C" int hpx_realization_normals(uint64_t seed, uint32_t drop, int64_t count, double* normals, void* stream) {
  HPX_REQUIRE(count >= 0 && count <= (int64_t(1) << 32), "hpx_realization_normals: invalid count");
  if (count == 0) return HPX_OK;
  HPX_REQUIRE(normals, "hpx_realization_normals: null pointer");
  const int64_t calls = (count + 1) / 2;
  hipLaunchKernelGGL(realization_normals_kernel, dim3(static_cast<unsigned>((calls + 255) / 256)), dim3(256), 0,
                     HPX_STREAM(stream), seed, drop, count, normals);
  return check_launch("hpx_realization_normals");
}

// grid (calls, drops): normals of several drops, each row as hpx_realization_normals writes it
__global__ void __launch_bounds__(256) realization_normals_batch_kernel(uint64_t seed, const uint32_t* __restrict__ ids,
                                                                        int64_t first, int64_t count,
                                                                        double* __restrict__ out) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t d = blockIdx.y;
  if (2 * c >= count) return;
  const uint32_t drop = ids ? ids[d] : static_cast<uint32_t>(first + d);
  const double2 z = normal_pair_f64(stream_call(seed, drop, HPX_STREAM_REALIZATION, static_cast<uint32_t>(c)));
  double* row = out + d * count;
  row[2 * c] = z.x;
  if (2 * c + 1 < count) row[2 * c + 1] = z.y;
}

extern "C" int hpx_realization_normals_batch(uint64_t seed, const uint32_t* drop_index, int64_t first_drop,
                                             int64_t num_drops, int64_t count, double* normals, void* stream) {
  HPX_REQUIRE(count >= 0 && count <= (int64_t(1) << 24), "hpx_realization_normals_batch: invalid count");
  HPX_REQUIRE(num_drops >= 0 && num_drops <= 65535, "hpx_realization_normals_batch: at most 65535 drops per call");
  HPX_REQUIRE(drop_index || (first_drop >= 0 && first_drop + num_drops <= (int64_t(1) << 32)),
              "hpx_realization_normals_batch: drop index exceeds 32 bits");
  if (count == 0 || num_drops == 0) return HPX_OK;
  HPX_REQUIRE(normals, "hpx_realization_normals_batch: null pointer");
  const int64_t calls = (count + 1) / 2;
  hipLaunchKernelGGL(realization_normals_batch_kernel,
                     dim3(static_cast<unsigned>((calls + 255) / 256), static_cast<unsigned>(num_drops)), dim3(256), 0,
                     HPX_STREAM(stream), seed, drop_index, first_drop, count, normals);
  return check_launch("hpx_realization_normals_batch");
}

// ------------------------------------------------------------------ pilot LS estimation
__global__ void __launch_bounds__(256) ls_pilots_kernel(const float2* __restrict__ grid, int64_t ld, int npil,
                                                        const int32_t* __restrict__ pidx,
                                                        const float2* __restrict__ psym, float2* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = blockIdx.y;
  if (i >= npil) return;
  const int64_t k = s * ld + pidx[i];
  const float2 x = psym[i], y = grid[k];
  const float inv = 1.0f / cabs2(x);
  out[k] = cscale(cmulc(x, y), inv);  // y / x
}

__global__ void __launch_bounds__(256) ls_holes_kernel(int64_t ld, int nhole, const int32_t* __restrict__ hidx,
                                                       const int32_t* __restrict__ pidx,
                                                       const int32_t* __restrict__ verts,
                                                       const float* __restrict__ w, float2* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = blockIdx.y;
  if (i >= nhole) return;
  const int v0 = verts[3 * i];
  if (v0 < 0) return;  // outside the pilots' hull: filled by the nearest element
  float2* row = out + s * ld;
  float2 acc = make_float2(0.f, 0.f);
  for (int k = 0; k < 3; ++k) {
    const float2 h = row[pidx[verts[3 * i + k]]];
    acc.x = fmaf(w[3 * i + k], h.x, acc.x);
    acc.y = fmaf(w[3 * i + k], h.y, acc.y);
  }
  row[hidx[i]] = acc;
}

__global__ void __launch_bounds__(256) ls_fill_kernel(int64_t ld, int nfill, const int32_t* __restrict__ fidx,
                                                      const int32_t* __restrict__ fsrc, float2* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = blockIdx.y;
  if (i >= nfill) return;
  float2* row = out + s * ld;
  row[fidx[i]] = row[fsrc[i]];
}

extern "C" int hpx_ls_estimate(const hpx_c64* grid, int num_streams, int64_t grid_size, int num_pilots,
                               const int32_t* pilot_index, const hpx_c64* pilot_symbols, int num_holes,
                               const int32_t* hole_index, const int32_t* hole_vertices, const float* hole_weights,
                               int num_fill, const int32_t* fill_index, const int32_t* fill_source, hpx_c64* estimate,
                               void* stream) {
  HPX_REQUIRE(num_streams >= 1 && num_streams <= 65535 && grid_size >= 1 && num_pilots >= 1 && num_holes >= 0 &&
                  num_fill >= 0 && static_cast<int64_t>(num_pilots) + num_holes <= grid_size,
              "hpx_ls_estimate: invalid sizes");
  HPX_REQUIRE(grid && pilot_index && pilot_symbols && estimate, "hpx_ls_estimate: null pointer");
  HPX_REQUIRE((num_holes == 0 || (hole_index && hole_vertices && hole_weights)) &&
                  (num_fill == 0 || (fill_index && fill_source)),
              "hpx_ls_estimate: null pointer");
  const hipStream_t s = HPX_STREAM(stream);
  float2* out = reinterpret_cast<float2*>(estimate);
  hipLaunchKernelGGL(ls_pilots_kernel, dim3((num_pilots + 255) / 256, num_streams), dim3(256), 0, s,
                     reinterpret_cast<const float2*>(grid), grid_size, num_pilots, pilot_index,
                     reinterpret_cast<const float2*>(pilot_symbols), out);
  if (num_holes > 0)
    hipLaunchKernelGGL(ls_holes_kernel, dim3((num_holes + 255) / 256, num_streams), dim3(256), 0, s, grid_size,
                       num_holes, hole_index, pilot_index, hole_vertices, hole_weights, out);
  if (num_fill > 0)
    hipLaunchKernelGGL(ls_fill_kernel, dim3((num_fill + 255) / 256, num_streams), dim3(256), 0, s, grid_size,
                       num_fill, fill_index, fill_source, out);
  return check_launch("hpx_ls_estimate");
}
