// 3GPP Gray mapping tables and exact hard decisions of the piecewise-linear
// LLRs of hermespy/modem/tools/psk_qam_mapping.py (get_llr_3gpp :301-497,
// detect_bits :171-259).
//
// A hard decision is `llr > 0`.  Every branch of the reference's piecewise LLR
// is a positive constant times (c +/- r) on an interval, so the sign of the
// LLR is decided by comparing r with integer thresholds; r itself is formed
// with the reference's float64 operations (x * sqrt(2) * sqrt(M')), so the
// decisions are bit-identical to the reference for identical inputs.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

namespace hpx {

// Bit 0 of a PAM axis, i.e. sign(-r), evaluated with the reference's own
// operation sequence so that values that underflow to an LLR of exactly zero
// (float64 subnormals) decide 0 as they do there.
__host__ __device__ inline uint32_t pam_sign_bit(int pam, double r, int complex_axis) {
  double t;
  switch (pam) {
    case 2: t = -2.0 * r; break;                 // -2 r / nv            (:335)
    case 4: t = (2.0 * (-2.0 * r)) / 5.0; break;  // 2 (-2 r) / nv / 5    (:342-361)
    case 8: t = (4.0 * (-r)) / 21.0; break;       // 4 (-r) / nv / 21     (:368-405)
    default: t = (8.0 * (-(0.5 * r))) / 85.0;     // 8 (-0.5 r) / nv / 85 (:412-489)
  }
  if (complex_axis) t = t / 2.0;  // (:494-495)
  return t > 0.0 ? 1u : 0u;
}

// Bits of one PAM axis (bit 0 first) for a scaled amplitude r.
// pam: 2, 4, 8 or 16 levels.  A non-finite r makes every reference LLR NaN,
// which decides 0.
__host__ __device__ inline uint32_t pam_decide(int pam, double r, int complex_axis) {
  if (!isfinite(r)) return 0u;
  uint32_t b = pam_sign_bit(pam, r, complex_axis);
  switch (pam) {
    case 2:
      return b;
    case 4:
      b |= ((r < -2.0) || (r > 2.0)) ? 2u : 0u;
      return b;
    case 8:
      b |= ((r < -4.0) || (r > 4.0)) ? 2u : 0u;
      b |= ((r < -6.0) || (r > -2.0 && r < 2.0) || (r > 6.0)) ? 4u : 0u;
      return b;
    default:  // 16
      b |= ((r < -8.0) || (r > 8.0)) ? 2u : 0u;
      b |= ((r < -12.0) || (r > -4.0 && r < 4.0) || (r > 12.0)) ? 4u : 0u;
      b |= ((r < -14.0) || (r > -10.0 && r < -6.0) || (r > -2.0 && r < 2.0) || (r > 6.0 && r < 10.0) ||
            (r > 14.0))
               ? 8u
               : 0u;
      return b;
  }
}

// Per-axis PAM order and amplitude scale of a modulation.
struct Modulation {
  int bps;      // bits per symbol
  int pam;      // PAM levels per axis
  int complex;  // 1: I/Q interleaved bits
  double pre;   // sqrt(2) for complex constellations, 1 otherwise (get_llr_3gpp :331-332)
  double scale; // sqrt(5), sqrt(21), sqrt(85) or 1 (get_llr_3gpp :340,366,410)
};

__host__ __device__ inline bool modulation(int order, int is_complex, Modulation* m) {
  const double s2 = 1.4142135623730951;  // float64 sqrt(2)
  if (order == 2) {
    *m = {1, 2, 0, 1.0, 1.0};
    return true;
  }
  if (is_complex) {
    switch (order) {
      case 4: *m = {2, 2, 1, s2, 1.0}; return true;
      case 16: *m = {4, 4, 1, s2, 2.23606797749979}; return true;           // sqrt(5)
      case 64: *m = {6, 8, 1, s2, 4.58257569495584}; return true;           // sqrt(21)
      case 256: *m = {8, 16, 1, s2, 9.219544457292887}; return true;        // sqrt(85)
      default: return false;
    }
  }
  switch (order) {
    case 4: *m = {2, 4, 0, 1.0, 2.23606797749979}; return true;
    case 8: *m = {3, 8, 0, 1.0, 4.58257569495584}; return true;
    case 16: *m = {4, 16, 0, 1.0, 9.219544457292887}; return true;
    default: return false;
  }
}

// Bits of the two axes of a complex symbol interleaved: bit i of the real axis to bit 2 i, of the
// imaginary axis to bit 2 i + 1 (an axis carries at most 4 bits); no loop over the bit count.
__host__ __device__ __forceinline__ uint32_t interleave_axis_bits(uint32_t br, uint32_t bi) {
  auto spread = [](uint32_t x) { return (x & 1u) | ((x & 2u) << 1) | ((x & 4u) << 2) | ((x & 8u) << 3); };
  return spread(br) | (spread(bi) << 1);
}

// Hard bits of one symbol, bit i of the result = i-th bit of the symbol
// (the order in which detect_bits emits them).
__host__ __device__ inline uint32_t demap_symbol(const Modulation& m, double re, double im) {
  if (!m.complex) {
    if (m.pam == 2) return pam_decide(2, re, 0);  // BPSK: llr = -2 re (:221)
    return pam_decide(m.pam, re * m.scale, 0);
  }
  // QPSK has no second scale: r = x sqrt(2) (:331-335)
  const double rr = m.pam == 2 ? re * m.pre : (re * m.pre) * m.scale;
  const double ri = m.pam == 2 ? im * m.pre : (im * m.pre) * m.scale;
  const uint32_t br = pam_decide(m.pam, rr, 1);
  const uint32_t bi = pam_decide(m.pam, ri, 1);
  return interleave_axis_bits(br, bi);
}

// ---------------------------------------------------------------- float32 inputs
// For a float32 component x of a complex constellation, r(x) = (x * pre) * scale (or x * pre for
// QPSK) is formed in float64 and rounded monotonically, so every decision "r > T" or "r < T" of
// pam_decide is "x > a_T" or "x < b_T" for two float32 boundaries: a_T the largest x >= 0 with
// r(x) <= T, b_T the smallest x >= 0 with r(x) >= T (r is odd, so -T uses -a_T and -b_T).  The
// sign decision is x < 0: for a finite float32 x the reference's t of pam_sign_bit cannot underflow
// to zero unless x is zero.  demap_symbol_f32 with these boundaries therefore decides exactly what
// demap_symbol decides on the float64 value of x, using float32 comparisons only.
struct DemapBounds {
  float a[8], b[8];  // thresholds T = 2, 4, ..., 16 (index T / 2 - 1)
};

__host__ __device__ inline double pam_amplitude_of(const Modulation& m, float x) {
  return m.pam == 2 ? static_cast<double>(x) * m.pre : (static_cast<double>(x) * m.pre) * m.scale;
}

// Host: boundaries of a complex modulation by bisection over the float32 bit patterns of x >= 0.
inline DemapBounds demap_bounds(const Modulation& m) {
  DemapBounds d{};
  auto as_float = [](uint32_t u) {
    float f;
    memcpy(&f, &u, sizeof(f));
    return f;
  };
  const uint32_t top = 0x7f7fffffu;  // largest finite float32
  for (int i = 0; i < 8; ++i) {
    const double t = 2.0 * (i + 1);
    // a_T: largest x with r(x) <= T
    uint32_t lo = 0, hi = top;
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo + 1) / 2;
      if (pam_amplitude_of(m, as_float(mid)) <= t) lo = mid;
      else hi = mid - 1;
    }
    d.a[i] = as_float(lo);
    // b_T: smallest x with r(x) >= T
    lo = 0;
    hi = top;
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (pam_amplitude_of(m, as_float(mid)) >= t) hi = mid;
      else lo = mid + 1;
    }
    d.b[i] = as_float(lo);
  }
  return d;
}

// Bits of one PAM axis of a complex constellation for a float32 component x (as pam_decide on r(x));
// d points at a DemapBounds (in LDS on the device, read where a comparison needs a boundary).
__host__ __device__ inline uint32_t pam_decide_f32(int pam, float x, const DemapBounds* d) {
  if (!isfinite(x)) return 0u;
  // r > T: x > a[T/2-1]; r < T: x < b[T/2-1]; r < -T: x < -a; r > -T: x > -b
#define HPX_GT(T) (x > d->a[(T) / 2 - 1])
#define HPX_LT(T) (x < d->b[(T) / 2 - 1])
#define HPX_LTN(T) (x < -d->a[(T) / 2 - 1])
#define HPX_GTN(T) (x > -d->b[(T) / 2 - 1])
  uint32_t b = x < 0.0f ? 1u : 0u;
  switch (pam) {
    case 2:
      break;
    case 4:
      b |= (HPX_LTN(2) || HPX_GT(2)) ? 2u : 0u;
      break;
    case 8:
      b |= (HPX_LTN(4) || HPX_GT(4)) ? 2u : 0u;
      b |= (HPX_LTN(6) || (HPX_GTN(2) && HPX_LT(2)) || HPX_GT(6)) ? 4u : 0u;
      break;
    default:  // 16
      b |= (HPX_LTN(8) || HPX_GT(8)) ? 2u : 0u;
      b |= (HPX_LTN(12) || (HPX_GTN(4) && HPX_LT(4)) || HPX_GT(12)) ? 4u : 0u;
      b |= (HPX_LTN(14) || (HPX_GTN(10) && HPX_LTN(6)) || (HPX_GTN(2) && HPX_LT(2)) || (HPX_GT(6) && HPX_LT(10)) ||
            HPX_GT(14))
               ? 8u
               : 0u;
      break;
  }
#undef HPX_GT
#undef HPX_LT
#undef HPX_LTN
#undef HPX_GTN
  return b;
}

// pam_decide_f32 for a compile-time PAM order with the boundaries of T = 2, 4, .., PAM - 2 in
// registers (ta[k] = a[k], tb[k] = b[k]): every comparison is evaluated (bitwise | and &, no
// short-circuit branches), so a wavefront stays on one path; the same decisions.
template <int PAM>
__host__ __device__ __forceinline__ uint32_t pam_decide_f32_reg(float x, const float* ta, const float* tb) {
#define HPX_GT(T) (x > ta[(T) / 2 - 1])
#define HPX_LT(T) (x < tb[(T) / 2 - 1])
#define HPX_LTN(T) (x < -ta[(T) / 2 - 1])
#define HPX_GTN(T) (x > -tb[(T) / 2 - 1])
  uint32_t b = x < 0.0f ? 1u : 0u;
  if constexpr (PAM == 4) {
    b |= (HPX_LTN(2) | HPX_GT(2)) ? 2u : 0u;
  } else if constexpr (PAM == 8) {
    b |= (HPX_LTN(4) | HPX_GT(4)) ? 2u : 0u;
    b |= (HPX_LTN(6) | (HPX_GTN(2) & HPX_LT(2)) | HPX_GT(6)) ? 4u : 0u;
  } else if constexpr (PAM == 16) {
    b |= (HPX_LTN(8) | HPX_GT(8)) ? 2u : 0u;
    b |= (HPX_LTN(12) | (HPX_GTN(4) & HPX_LT(4)) | HPX_GT(12)) ? 4u : 0u;
    b |= (HPX_LTN(14) | (HPX_GTN(10) & HPX_LTN(6)) | (HPX_GTN(2) & HPX_LT(2)) | (HPX_GT(6) & HPX_LT(10)) |
          HPX_GT(14))
             ? 8u
             : 0u;
  }
#undef HPX_GT
#undef HPX_LT
#undef HPX_LTN
#undef HPX_GTN
  return __builtin_isfinite(x) ? b : 0u;
}

// demap_symbol_f32 of a complex constellation of compile-time PAM order (2, 4, 8, 16: QPSK to
// 256-QAM), boundaries in registers as for pam_decide_f32_reg.
template <int PAM>
__host__ __device__ __forceinline__ uint32_t demap_symbol_f32_reg(float re, float im, const float* ta, const float* tb) {
  const uint32_t br = pam_decide_f32_reg<PAM>(re, ta, tb);
  const uint32_t bi = pam_decide_f32_reg<PAM>(im, ta, tb);
  return interleave_axis_bits(br, bi);
}

// demap_symbol of a complex constellation for float32 components (bit-identical decisions).
__host__ __device__ inline uint32_t demap_symbol_f32(const Modulation& m, const DemapBounds* d, float re, float im) {
  const uint32_t br = pam_decide_f32(m.pam, re, d);
  const uint32_t bi = pam_decide_f32(m.pam, im, d);
  return interleave_axis_bits(br, bi);
}

// Label (MSB first) of a symbol's bits given bit i of `bits` = i-th bit.
__host__ __device__ inline uint32_t bits_to_label(uint32_t bits, int bps) {
  uint32_t label = 0;
  for (int i = 0; i < bps; ++i) label |= ((bits >> i) & 1u) << (bps - 1 - i);
  return label;
}

// Unnormalised 3GPP PAM amplitude of a label's axis bits (generate_pam_symbol_3gpp :283-298),
// bit i of `b` = i-th bit of the axis.
__host__ __device__ inline double pam_amplitude(int pam, uint32_t b) {
  const double b0 = (b & 1u) ? 1.0 : 0.0, b1 = (b & 2u) ? 1.0 : 0.0, b2 = (b & 4u) ? 1.0 : 0.0,
               b3 = (b & 8u) ? 1.0 : 0.0;
  switch (pam) {
    case 2: return 1.0 - 2 * b0;
    case 4: return (1 - 2 * b0) * (1 + 2 * b1);
    case 8: return (2 * b0 - 1) * ((1 - 2 * b1) * (1 + 2 * b2) - 4);
    default: return ((((1 - 2 * b2) * (1 + 2 * b3) - 4) * (-1 + 2 * b1)) - 8) * (-1 + 2 * b0);
  }
}

// Constellation point of an MSB-first label, float64, rounded as the reference
// rounds it: complex constellations are (I + jQ) / sqrt(M') evaluated by NumPy
// as a complex division by a real, i.e. I * (1 / sqrt(M')) (psk_qam_mapping.py:141-163);
// real PAM is a real division (:144-155).
__host__ __device__ inline double2 constellation_point(int order, int is_complex, uint32_t label) {
  Modulation m;
  modulation(order, is_complex, &m);
  uint32_t bits = 0;  // bit i = i-th transmitted bit
  for (int i = 0; i < m.bps; ++i) bits |= ((label >> (m.bps - 1 - i)) & 1u) << i;
  if (!m.complex) {
    if (order == 2) return make_double2(pam_amplitude(2, bits), 0.0);
    const double norm = m.pam == 4 ? 5.0 : (m.pam == 8 ? 21.0 : 85.0);
    return make_double2(pam_amplitude(m.pam, bits) / sqrt(norm), 0.0);
  }
  uint32_t bi = 0, bq = 0;
  for (int i = 0; i < m.bps / 2; ++i) {
    bi |= ((bits >> (2 * i)) & 1u) << i;
    bq |= ((bits >> (2 * i + 1)) & 1u) << i;
  }
  const double norm = order == 4 ? 2.0 : (order == 16 ? 10.0 : (order == 64 ? 42.0 : 170.0));
  const double scl = 1.0 / sqrt(norm);
  return make_double2(pam_amplitude(m.pam, bi) * scl, pam_amplitude(m.pam, bq) * scl);
}

}  // namespace hpx
