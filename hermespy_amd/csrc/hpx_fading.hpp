// Sum-of-sinusoids taps of hermespy/channel/fading/fading.py:293-343.
//
// A path is a sum of terms A_n exp(j (rate_n m + phase_n)) at sample m.  Two
// evaluators:
//  * direct: float64 phase rate*m + phase, reduced to [-pi, pi], float32 sincos;
//  * block Taylor: over a block of B samples centred at mc, with s = (m - mc) / (B/2),
//      h(m) = sum_k C_k s^k,  C_k = sum_n A_n e^{j(rate_n mc + phase_n)} (j rate_n B/2)^k / k!,
//    evaluated by Horner in float32.  The coefficients are formed in float64.
//    Truncation after degree kTaylorDegree leaves |error| <= sum|A_n| x^(P+1)/(P+1)!,
//    x = max|rate| B/2; the host picks B so that this stays below 1e-6 rms, a tenth of the
//    1e-5 relative tolerance of the float32 signals (the cubic at 1e-6 takes the same blocks
//    as a quartic at 1e-8, with one coefficient and two FMAs fewer per evaluation).
#pragma once

#include "hpx_common.hpp"

namespace hpx {

constexpr int kTaylorDegree = 3;
constexpr int kTaylorCoeffs = kTaylorDegree + 1;

__device__ __forceinline__ double reduce_phase(double p) {
  // p - 2 pi rint(p / 2 pi) in float64
  const double k = rint(p * 0.15915494309189535);
  return fma(-k, 6.283185307179586, p);
}

// Direct evaluation of one path at sample m.  sos: [terms][3] (rate, phase, amplitude).
__device__ inline float2 tap_direct(const double* sos, int terms, int64_t m) {
  float2 acc = make_float2(0.f, 0.f);
  const double md = static_cast<double>(m);
  for (int n = 0; n < terms; ++n) {
    const double ph = reduce_phase(fma(sos[3 * n], md, sos[3 * n + 1]));
    float s, c;
    sincosf(static_cast<float>(ph), &s, &c);
    const float a = static_cast<float>(sos[3 * n + 2]);
    acc.x = fmaf(a, c, acc.x);
    acc.y = fmaf(a, s, acc.y);
  }
  return acc;
}

// Taylor coefficients of one path over the block centred at mc with half width hw.  The phase
// argument rate * mc + phase is accumulated and reduced in float64; each term's sine and cosine
// and the sums are float32 (the reduced phase lies in [-pi, pi], so float32 sincos keeps a
// relative error near 1e-7 per term, far below the float32 signal path's 1e-5).
__device__ inline void taylor_coeffs(const double* sos, int terms, double mc, double hw, float2* out) {
  float cr[kTaylorCoeffs], ci[kTaylorCoeffs];
#pragma unroll
  for (int k = 0; k < kTaylorCoeffs; ++k) cr[k] = ci[k] = 0.0f;
  for (int n = 0; n < terms; ++n) {
    const double rate = sos[3 * n], amp = sos[3 * n + 2];
    if (amp == 0.0) continue;
    const float ph = static_cast<float>(reduce_phase(fma(rate, mc, sos[3 * n + 1])));
    float s, c;
    sincosf(ph, &s, &c);
    const float a = static_cast<float>(amp);
    float tr = a * c, ti = a * s;
    const float u = static_cast<float>(rate * hw);
#pragma unroll
    for (int k = 0; k < kTaylorCoeffs; ++k) {
      cr[k] += tr;
      ci[k] += ti;
      // term *= j u / (k + 1)
      const float f = u * (1.0f / static_cast<float>(k + 1));
      const float nr = -ti * f, ni = tr * f;
      tr = nr;
      ti = ni;
    }
  }
#pragma unroll
  for (int k = 0; k < kTaylorCoeffs; ++k) out[k] = make_float2(cr[k], ci[k]);
}

__device__ __forceinline__ float2 taylor_eval(const float2* c, float s) {
  float2 acc = c[kTaylorDegree];
#pragma unroll
  for (int k = kTaylorDegree - 1; k >= 0; --k) {
    acc.x = fmaf(acc.x, s, c[k].x);
    acc.y = fmaf(acc.y, s, c[k].y);
  }
  return acc;
}

// Block geometry: block b covers [b B, (b+1) B), centre b B + (B-1)/2.
struct TaylorBlocks {
  int log2b;
  __device__ __forceinline__ int64_t block_of(int64_t m) const { return m >> log2b; }
  __device__ __forceinline__ double centre(int64_t b) const {
    return static_cast<double>(b << log2b) + 0.5 * static_cast<double>((1 << log2b) - 1);
  }
  __device__ __forceinline__ double half_width() const { return 0.5 * static_cast<double>(1 << log2b); }
  // normalised offset (2 r - (B - 1)) / B of sample m in its block (r = m mod B), exact in
  // float32: r 2^(1-log2b) and 2^-log2b - 1 are exact, and so is their sum (an odd integer
  // of magnitude < 2^12 over B), which the single rounding of the FMA therefore keeps
  __device__ __forceinline__ float offset(int64_t m) const {
    const int r = static_cast<int>(m & ((int64_t(1) << log2b) - 1));  // 0 .. B-1
    return fmaf(static_cast<float>(r), ldexpf(2.0f, -log2b), ldexpf(1.0f, -log2b) - 1.0f);
  }
  // the same for a sample index that fits 32 bits
  __device__ __forceinline__ float offset32(int m) const {
    const int r = m & ((1 << log2b) - 1);
    return fmaf(static_cast<float>(r), ldexpf(2.0f, -log2b), ldexpf(1.0f, -log2b) - 1.0f);
  }
};

// sin(n rate / 2) / (n sin(rate / 2)): the mean of exp(j rate k) over k = 0 .. n-1 relative to its
// value at the window centre
__device__ inline double dirichlet_factor(double rate, int n) {
  const double half = 0.5 * rate;
  const double den = static_cast<double>(n) * sin(half);
  return (den != 0.0) ? sin(static_cast<double>(n) * half) / den : 1.0;
}

// Word-mean of one path over N samples starting at m0 (the Dirichlet closed form
// of the per-symbol average of hermespy/simulation/modem/channel_estimation.py:242).
__device__ inline double2 tap_mean(const double* sos, int terms, double m0, int n) {
  const double centre = m0 + 0.5 * static_cast<double>(n - 1);
  double ar = 0.0, ai = 0.0;
  for (int k = 0; k < terms; ++k) {
    const double rate = sos[3 * k], amp = sos[3 * k + 2];
    if (amp == 0.0) continue;
    const double dir = dirichlet_factor(rate, n);
    const double ph = reduce_phase(fma(rate, centre, sos[3 * k + 1]));
    double s, c;
    sincos(ph, &s, &c);
    ar += amp * dir * c;
    ai += amp * dir * s;
  }
  return make_double2(ar, ai);
}

// tap_mean with the per-term weights amp_dir[k] = amp * dirichlet_factor(rate, n) computed once
// for every word; the phase is reduced in float64, its sine and cosine and the sum are float32
// (as taylor_coeffs).
__device__ inline float2 tap_mean_weighted(const double* sos, const double* amp_dir, int terms, double m0, int n) {
  const double centre = m0 + 0.5 * static_cast<double>(n - 1);
  float ar = 0.0f, ai = 0.0f;
  for (int k = 0; k < terms; ++k) {
    if (sos[3 * k + 2] == 0.0) continue;
    const float w = static_cast<float>(amp_dir[k]);
    const float ph = static_cast<float>(reduce_phase(fma(sos[3 * k], centre, sos[3 * k + 1])));
    float s, c;
    sincosf(ph, &s, &c);
    ar = fmaf(w, c, ar);
    ai = fmaf(w, s, ai);
  }
  return make_float2(ar, ai);
}

}  // namespace hpx
