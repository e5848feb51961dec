// Device helpers shared by every kernel of libhermespy_amd: complex float
// arithmetic, the counter-based Philox4x32-10 generator, Box-Muller normals
// and the error channel of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hermespy_amd.h"

namespace hpx {

// ---------------------------------------------------------------- complex
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
// acc + a * b
__device__ __forceinline__ float2 cfma(float2 a, float2 b, float2 acc) {
  acc.x = fmaf(a.x, b.x, acc.x);
  acc.x = fmaf(-a.y, b.y, acc.x);
  acc.y = fmaf(a.x, b.y, acc.y);
  acc.y = fmaf(a.y, b.x, acc.y);
  return acc;
}
// acc + conj(a) * b
__device__ __forceinline__ float2 cfmac(float2 a, float2 b, float2 acc) {
  acc.x = fmaf(a.x, b.x, acc.x);
  acc.x = fmaf(a.y, b.y, acc.x);
  acc.y = fmaf(a.x, b.y, acc.y);
  acc.y = fmaf(-a.y, b.x, acc.y);
  return acc;
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float cabs2(float2 a) { return fmaf(a.x, a.x, a.y * a.y); }

// ---------------------------------------------------------------- Philox4x32-10
// Salmon et al., "Parallel random numbers: as easy as 1, 2, 3" (SC'11).
constexpr uint32_t kPhiloxM0 = 0xD2511F53u;
constexpr uint32_t kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u;
constexpr uint32_t kPhiloxW1 = 0xBB67AE85u;

__device__ __forceinline__ uint4 philox(uint4 c, uint2 k) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k.x += kPhiloxW0;
      k.y += kPhiloxW1;
    }
    // the full 64-bit products: one v_mad_u64_u32 each, which issues faster on gfx950 than the
    // v_mul_hi_u32 + v_mul_lo_u32 pair of the separate halves (same bits)
    const uint64_t p0 = static_cast<uint64_t>(kPhiloxM0) * c.x;
    const uint64_t p1 = static_cast<uint64_t>(kPhiloxM1) * c.z;
    const uint32_t hi0 = static_cast<uint32_t>(p0 >> 32), lo0 = static_cast<uint32_t>(p0);
    const uint32_t hi1 = static_cast<uint32_t>(p1 >> 32), lo1 = static_cast<uint32_t>(p1);
    c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
  }
  return c;
}

__device__ __forceinline__ uint2 seed_key(uint64_t seed) {
  return make_uint2(static_cast<uint32_t>(seed & 0xffffffffu), static_cast<uint32_t>(seed >> 32));
}

__device__ __forceinline__ uint4 stream_call(uint64_t seed, uint32_t drop, uint32_t stream, uint32_t index) {
  return philox(make_uint4(drop, stream, index, 0u), seed_key(seed));
}

// 53-bit uniform on (0, 1) from two words.
__device__ __forceinline__ double uniform53(uint32_t hi, uint32_t lo) {
  const double a = static_cast<double>(hi >> 5);
  const double b = static_cast<double>(lo >> 6);
  return (a * 67108864.0 + b + 0.5) * (1.0 / 9007199254740992.0);
}

// Pair of float64 standard normals from one Philox call (Box-Muller).
__device__ __forceinline__ double2 normal_pair_f64(uint4 w) {
  const double u1 = uniform53(w.x, w.y);
  const double u2 = uniform53(w.z, w.w);
  const double rho = sqrt(-2.0 * log(u1));
  double s, c;
  sincospi(2.0 * u2, &s, &c);
  return make_double2(rho * c, rho * s);
}

// Complex normal with unit-variance real and imaginary parts from two words
// (24-bit uniforms, float32 Box-Muller).
__device__ __forceinline__ float2 complex_normal_f32(uint32_t a, uint32_t b) {
  const float u1 = (static_cast<float>(a >> 8) + 0.5f) * 5.9604644775390625e-08f;  // 2^-24
  const float u2 = static_cast<float>(b >> 8) * 5.9604644775390625e-08f;
  // hardware transcendentals: v_log_f32 is log2 (-2 ln u = -2 ln2 log2 u) and
  // v_sin_f32 / v_cos_f32 take revolutions, i.e. exactly the angle 2 pi u2
  const float rho = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u1));
  const float s = __builtin_amdgcn_sinf(u2), c = __builtin_amdgcn_cosf(u2);
  return make_float2(rho * c, rho * s);
}

// Standard normal CDF (the uniform transform of ConsistentUniform).  Out of line: inlined into a
// kernel's prologue loop, erfc's float64 polynomial constants are hoisted into ~60 registers for
// the whole loop, which pushes a 128-register kernel into spilling.
__device__ __noinline__ double normal_cdf(double z) { return 0.5 * erfc(-z * 0.70710678118654752440); }

}  // namespace hpx

// ---------------------------------------------------------------- host error channel
namespace hpx {
void set_error(const char* fmt, ...);
int check_launch(const char* what);
}  // namespace hpx

#define HPX_REQUIRE(cond, ...)        \
  do {                                \
    if (!(cond)) {                    \
      ::hpx::set_error(__VA_ARGS__);  \
      return HPX_ERR_INVALID;         \
    }                                 \
  } while (0)
