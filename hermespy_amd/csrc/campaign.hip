// Monte Carlo section statistics on the device.
//
//   hpx_section_accumulate  running sum, squared sum and count of the per-drop bit error
//                           rate per grid section (ScalarEvaluationResult.add_artifact,
//                           hermespy/core/pymonte/scalar.py:100-125, fed by
//                           BitErrorEvaluation.artifact, hermespy/modem/evaluators.py:231-233),
//                           plus pooled error and bit counts.
//
// The per-drop error counts stay on the device; a campaign reads back only these few
// numbers per section to apply the stopping rule between launches.
#include <stdint.h>

#include "hpx_common.hpp"

using namespace hpx;

namespace {

constexpr int kThreads = 256;

// One workgroup per section.  Thread j visits the section's drops in increasing drop order
// with stride kThreads (over the section's own drops when the section of drop i is
// (first + i) % S, over all drops otherwise); the partial sums are combined in a fixed tree.
__global__ void __launch_bounds__(kThreads) section_accumulate_kernel(const int32_t* __restrict__ errors,
                                                                      const int32_t* __restrict__ section,
                                                                      int64_t first, const int32_t* __restrict__ bpd,
                                                                      int64_t bits, int64_t n, int num_sections,
                                                                      hpx_section_stats st) {
  __shared__ double s_sum[kThreads], s_sq[kThreads];
  __shared__ long long s_cnt[kThreads], s_err[kThreads], s_bits[kThreads];
  const int s = blockIdx.x;
  const int tid = threadIdx.x;
  double sum = 0.0, sq = 0.0;
  long long cnt = 0, err = 0, nb = 0;
  auto add = [&](int64_t i) {
    const int64_t b = bpd ? static_cast<int64_t>(bpd[i]) : bits;
    const int64_t e = errors[i];
    const double ber = b > 0 ? static_cast<double>(e) / static_cast<double>(b) : 0.0;
    sum += ber;
    sq += ber * ber;
    cnt += 1;
    err += e;
    nb += b;
  };
  if (section) {
    for (int64_t i = tid; i < n; i += kThreads)
      if (section[i] == s) add(i);
  } else {
    // first drop i >= 0 with (first + i) % S == s
    const int64_t i0 = ((static_cast<int64_t>(s) - first % num_sections) + num_sections) % num_sections;
    for (int64_t i = i0 + static_cast<int64_t>(tid) * num_sections; i < n; i += static_cast<int64_t>(kThreads) * num_sections)
      add(i);
  }
  s_sum[tid] = sum;
  s_sq[tid] = sq;
  s_cnt[tid] = cnt;
  s_err[tid] = err;
  s_bits[tid] = nb;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (tid < w) {
      s_sum[tid] += s_sum[tid + w];
      s_sq[tid] += s_sq[tid + w];
      s_cnt[tid] += s_cnt[tid + w];
      s_err[tid] += s_err[tid + w];
      s_bits[tid] += s_bits[tid + w];
    }
    __syncthreads();
  }
  if (tid == 0) {
    st.sum[s] += s_sum[0];
    st.sum_sq[s] += s_sq[0];
    st.count[s] += s_cnt[0];
    st.errors[s] += s_err[0];
    st.bits[s] += s_bits[0];
  }
}

}  // namespace

extern "C" int hpx_section_accumulate(const int32_t* errors, const int32_t* section, int64_t first_drop,
                                      const int32_t* bits_per_drop, int64_t bits, int64_t num_drops,
                                      int32_t num_sections, const hpx_section_stats* stats_host, void* stream) {
  HPX_REQUIRE(stats_host, "hpx_section_accumulate: null statistics descriptor");
  HPX_REQUIRE(num_sections > 0 && num_sections <= 65535, "hpx_section_accumulate: 1 to 65535 sections");
  HPX_REQUIRE(num_drops >= 0 && first_drop >= 0, "hpx_section_accumulate: invalid drop range");
  HPX_REQUIRE(bits_per_drop || bits > 0, "hpx_section_accumulate: bits per drop must be positive");
  if (num_drops == 0) return HPX_OK;
  const hpx_section_stats st = *stats_host;
  HPX_REQUIRE(errors && st.sum && st.sum_sq && st.count && st.errors && st.bits, "hpx_section_accumulate: null pointer");
  hipLaunchKernelGGL(section_accumulate_kernel, dim3(static_cast<unsigned>(num_sections)), dim3(kThreads), 0,
                     static_cast<hipStream_t>(stream), errors, section, first_drop, bits_per_drop, bits, num_drops,
                     static_cast<int>(num_sections), st);
  return check_launch("hpx_section_accumulate");
}
