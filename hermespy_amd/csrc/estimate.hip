// Pilot-based channel estimation and detection on estimated channels.
//
//   hpx_pilot_ls          least squares at every stream's pilots
//                         (hermespy/modem/waveforms/orthogonal/channel.py:44-48)
//   hpx_lmmse_apply       linear MMSE interpolation of the pilot LS values over the
//                         whole grid with a precomputed filter (a complex float32 GEMM)
//   hpx_detect_estimated  ZF/MMSE detection of the data elements from estimated channels,
//                         hard demapping and per-drop bit errors against the Philox bits
//                         (waveform.py:235-250, psk_qam_mapping.py:171-259, evaluators.py:239-258)
#include <math.h>

#include <algorithm>

#include "hpx_common.hpp"
#include "hpx_detect.hpp"
#include "hpx_mapping.hpp"

using namespace hpx;

namespace {

constexpr int kMaxRx = 8;

inline hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

__global__ void __launch_bounds__(256) pilot_ls_kernel(const float2* __restrict__ grid, int64_t total, int num_drops,
                                                       int num_rx, int64_t grid_elements, int num_pilots,
                                                       const int32_t* __restrict__ pidx,
                                                       const float2* __restrict__ pval, float2* __restrict__ ls) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= total) return;
  // ls[t][d][r][j]
  const int j = static_cast<int>(i % num_pilots);
  const int64_t rest = i / num_pilots;
  const int64_t dr = rest % (static_cast<int64_t>(num_drops) * num_rx);
  const int t = static_cast<int>(rest / (static_cast<int64_t>(num_drops) * num_rx));
  const float2 x = pval[t * num_pilots + j];
  const float2 y = grid[dr * grid_elements + pidx[t * num_pilots + j]];
  ls[i] = cscale(cmulc(x, y), 1.0f / cabs2(x));  // y / x
}

// est[t][row][e] = sum_j W[t][e][j] ls[t][row][j] for the rows of one filter set.
// Tile: 32 rows x 128 elements per workgroup, 4 x 4 per thread, K chunks of 32 in LDS.
constexpr int kGemmRows = 32, kGemmCols = 128, kGemmK = 32;

__global__ void __launch_bounds__(256) lmmse_kernel(const float2* __restrict__ ls, int64_t ls_tx_stride,
                                                    int num_pilots, const float2* __restrict__ filt,
                                                    int64_t filt_tx_stride, int64_t elements,
                                                    const int32_t* __restrict__ rows, int count,
                                                    float2* __restrict__ est, int64_t est_tx_stride) {
  __shared__ float2 as[kGemmRows][kGemmK];
  __shared__ float2 bs[kGemmCols][kGemmK + 1];
  const int t = blockIdx.z;
  const int r0 = blockIdx.x * kGemmRows;
  const int64_t e0 = static_cast<int64_t>(blockIdx.y) * kGemmCols;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const float2* lst = ls + t * ls_tx_stride;
  const float2* ft = filt + t * filt_tx_stride;
  float2 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[i][c] = make_float2(0.f, 0.f);

  for (int k0 = 0; k0 < num_pilots; k0 += kGemmK) {
    for (int v = threadIdx.x; v < kGemmRows * kGemmK; v += 256) {
      const int rr = v / kGemmK, kk = v - rr * kGemmK;
      const int row = r0 + rr;
      as[rr][kk] = (row < count && k0 + kk < num_pilots)
                       ? lst[static_cast<int64_t>(rows[row]) * num_pilots + k0 + kk]
                       : make_float2(0.f, 0.f);
    }
    for (int v = threadIdx.x; v < kGemmCols * kGemmK; v += 256) {
      const int cc = v / kGemmK, kk = v - cc * kGemmK;
      const int64_t e = e0 + cc;
      bs[cc][kk] = (e < elements && k0 + kk < num_pilots) ? ft[e * num_pilots + k0 + kk] : make_float2(0.f, 0.f);
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < kGemmK; ++kk) {
      float2 a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = as[ty + 8 * i][kk];
#pragma unroll
      for (int c = 0; c < 4; ++c) b[c] = bs[tx + 32 * c][kk];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[i][c] = cfma(a[i], b[c], acc[i][c]);
    }
    __syncthreads();
  }
  float2* et = est + t * est_tx_stride;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int row = r0 + ty + 8 * i;
    if (row >= count) continue;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int64_t e = e0 + tx + 32 * c;
      if (e < elements) et[static_cast<int64_t>(rows[row]) * elements + e] = acc[i][c];
    }
  }
}

// bps bits of a stream starting at bit `pos` (LSB-first 32-bit words, four per Philox call)
__device__ __forceinline__ uint32_t stream_symbol_bits(uint64_t seed, uint32_t drop, uint32_t stream, int64_t pos,
                                                       int bps) {
  const int64_t wi = pos >> 5;
  const uint4 r = stream_call(seed, drop, stream, static_cast<uint32_t>(wi >> 2));
  const uint32_t comp[4] = {r.x, r.y, r.z, r.w};
  uint64_t v = comp[wi & 3];
  const int sh = static_cast<int>(pos & 31);
  if (sh + bps > 32) {
    const int64_t wn = wi + 1;
    uint32_t next;
    if ((wn & 3) != 0) {
      next = comp[wn & 3];
    } else {
      const uint4 q = stream_call(seed, drop, stream, static_cast<uint32_t>(wn >> 2));
      next = q.x;
    }
    v |= static_cast<uint64_t>(next) << 32;
  }
  return static_cast<uint32_t>(v >> sh) & ((1u << bps) - 1u);
}

template <int NT>
__global__ void __launch_bounds__(256) detect_estimated_kernel(
    const float2* __restrict__ grid, const float2* __restrict__ est, int num_drops, int num_rx, int words, int n,
    const int32_t* __restrict__ elements, const int32_t* __restrict__ word_data, Modulation mod, int mmse,
    const double* __restrict__ noise_power, uint64_t seed, int64_t first_drop, int data_total,
    int32_t* __restrict__ errors, float2* __restrict__ equalized) {
  __shared__ int red[4];
  const int d = blockIdx.x;
  const uint32_t drop = static_cast<uint32_t>(first_drop + d);
  const int64_t wn = static_cast<int64_t>(words) * n;
  const float s2 = mmse ? static_cast<float>(noise_power[d]) : 0.0f;
  const float2* yd = grid + static_cast<int64_t>(d) * num_rx * wn;
  int err = 0;
  for (int64_t e = threadIdx.x; e < wn; e += blockDim.x) {
    const int el = elements ? elements[e] : static_cast<int>(e % n);
    if (el < 0) continue;
    const int w = static_cast<int>(e / n);
    const int gidx = (word_data ? word_data[w] : w * n) + el;
    float2 x[NT];
    detect_rows<NT>(
        num_rx,
        [&](int r, int t) { return est[((static_cast<int64_t>(t) * num_drops + d) * num_rx + r) * wn + e]; },
        [&](int r) { return yd[r * wn + e]; }, s2, x);
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const uint32_t rx = demap_symbol(mod, static_cast<double>(x[t].x), static_cast<double>(x[t].y));
      const uint32_t tx =
          stream_symbol_bits(seed, drop, HPX_STREAM_BITS + t, static_cast<int64_t>(gidx) * mod.bps, mod.bps);
      err += __popc(rx ^ tx);
      if (equalized) equalized[(static_cast<int64_t>(d) * NT + t) * data_total + gidx] = x[t];
    }
  }
  for (int off = 32; off > 0; off >>= 1) err += __shfl_down(err, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = err;
  __syncthreads();
  if (threadIdx.x == 0) errors[d] = red[0] + red[1] + red[2] + red[3];
}

}  // namespace

extern "C" int hpx_pilot_ls(const hpx_c64* grid, int64_t num_drops, int num_rx, int num_tx, int64_t grid_elements,
                            int num_pilots, const int32_t* pilot_index, const hpx_c64* pilot_values, hpx_c64* ls,
                            void* stream) {
  HPX_REQUIRE(num_drops >= 0 && num_rx >= 1 && num_tx >= 1 && grid_elements >= 1 && num_pilots >= 1,
              "hpx_pilot_ls: invalid sizes");
  const int64_t total = static_cast<int64_t>(num_tx) * num_drops * num_rx * num_pilots;
  if (total == 0) return HPX_OK;
  HPX_REQUIRE(grid && pilot_index && pilot_values && ls, "hpx_pilot_ls: null pointer");
  HPX_REQUIRE(num_drops <= 0x7fffffff, "hpx_pilot_ls: too many drops");
  hipLaunchKernelGGL(pilot_ls_kernel, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0,
                     as_stream(stream), reinterpret_cast<const float2*>(grid), total, static_cast<int>(num_drops),
                     num_rx, grid_elements, num_pilots, pilot_index, reinterpret_cast<const float2*>(pilot_values),
                     reinterpret_cast<float2*>(ls));
  return check_launch("hpx_pilot_ls");
}

extern "C" int hpx_lmmse_apply(const hpx_c64* ls, int num_tx, int64_t num_rows, int num_pilots,
                               const hpx_c64* filter, int64_t elements, const int32_t* rows, int count, hpx_c64* est,
                               void* stream) {
  HPX_REQUIRE(num_tx >= 1 && num_tx <= 65535 && num_rows >= 0 && num_pilots >= 1 && elements >= 1 && count >= 0,
              "hpx_lmmse_apply: invalid sizes");
  if (count == 0) return HPX_OK;
  HPX_REQUIRE(ls && filter && rows && est, "hpx_lmmse_apply: null pointer");
  HPX_REQUIRE(count <= num_rows, "hpx_lmmse_apply: more rows than the LS array holds");
  HPX_REQUIRE((elements + kGemmCols - 1) / kGemmCols <= 65535, "hpx_lmmse_apply: grid too large");
  const dim3 g(static_cast<unsigned>((count + kGemmRows - 1) / kGemmRows),
               static_cast<unsigned>((elements + kGemmCols - 1) / kGemmCols), static_cast<unsigned>(num_tx));
  hipLaunchKernelGGL(lmmse_kernel, g, dim3(256), 0, as_stream(stream), reinterpret_cast<const float2*>(ls),
                     num_rows * num_pilots, num_pilots, reinterpret_cast<const float2*>(filter),
                     elements * num_pilots, elements, rows, count, reinterpret_cast<float2*>(est),
                     num_rows * elements);
  return check_launch("hpx_lmmse_apply");
}

extern "C" int hpx_detect_estimated(const hpx_c64* grid, const hpx_c64* est, int64_t num_drops, int num_rx,
                                    int num_tx, int words, int num_subcarriers, const int32_t* elements,
                                    const int32_t* word_data, int num_data, int modulation_order, int detector,
                                    const double* noise_power, uint64_t seed, int64_t first_drop, int32_t* errors,
                                    hpx_c64* equalized, void* stream) {
  HPX_REQUIRE(num_drops >= 0 && num_drops <= 0x7fffffff && words >= 1 && num_subcarriers >= 1 && num_data >= 0,
              "hpx_detect_estimated: invalid sizes");
  HPX_REQUIRE(detector == HPX_DETECT_ZF || detector == HPX_DETECT_MMSE, "hpx_detect_estimated: unknown detector");
  HPX_REQUIRE(num_rx >= num_tx && num_rx <= kMaxRx, "hpx_detect_estimated: needs num_tx <= num_rx <= %d", kMaxRx);
  HPX_REQUIRE((elements == nullptr) == (word_data == nullptr), "hpx_detect_estimated: give elements and word_data");
  HPX_REQUIRE(elements || num_data == words * num_subcarriers, "hpx_detect_estimated: all-data frame size mismatch");
  HPX_REQUIRE(first_drop >= 0 && first_drop + num_drops <= (int64_t(1) << 32),
              "hpx_detect_estimated: drop index exceeds 32 bits");
  Modulation mod;
  if (modulation_order == 2 || !modulation(modulation_order, 1, &mod)) {
    set_error("hpx_detect_estimated: modulation order %d not supported (4, 16, 64, 256)", modulation_order);
    return HPX_ERR_UNSUPPORTED;
  }
  if (num_drops == 0) return HPX_OK;
  HPX_REQUIRE(grid && est && noise_power && errors, "hpx_detect_estimated: null pointer");
  const dim3 g(static_cast<unsigned>(num_drops));
  const hipStream_t s = as_stream(stream);
  const float2* gy = reinterpret_cast<const float2*>(grid);
  const float2* ge = reinterpret_cast<const float2*>(est);
  float2* eq = reinterpret_cast<float2*>(equalized);
  const int nd = static_cast<int>(num_drops);
  const int mm = detector == HPX_DETECT_MMSE;
#define HPX_DETECT_EST(NT)                                                                                        \
  hipLaunchKernelGGL(detect_estimated_kernel<NT>, g, dim3(256), 0, s, gy, ge, nd, num_rx, words, num_subcarriers, \
                     elements, word_data, mod, mm, noise_power, seed, first_drop, num_data, errors, eq)
  switch (num_tx) {
    case 1: HPX_DETECT_EST(1); break;
    case 2: HPX_DETECT_EST(2); break;
    case 4: HPX_DETECT_EST(4); break;
    case 8: HPX_DETECT_EST(8); break;
    default:
      set_error("hpx_detect_estimated: num_tx %d not instantiated (1, 2, 4, 8)", num_tx);
      return HPX_ERR_UNSUPPORTED;
  }
#undef HPX_DETECT_EST
  return check_launch("hpx_detect_estimated");
}
