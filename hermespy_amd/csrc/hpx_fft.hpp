// In-place complex FFT of several power-of-two vectors held in LDS, executed
// cooperatively by one workgroup.  Stockham radix-4 passes (one radix-2 pass
// when log2 N is odd); every pass reads all of a thread's inputs into
// registers, synchronises, then writes, so no second LDS buffer is needed.
// Twiddles come from a table tw[k] = exp(-2 pi j k / N) (float32, correctly
// rounded from float64).  No scaling is applied.
#pragma once

#include "hpx_common.hpp"

namespace hpx {

// Default upper bound of complex values per thread and pass (count * N <= MAXV * blockDim).
// It sets the registers a pass holds across its barrier (2 MAXV); callers with little
// register headroom pass a smaller bound and transform their vectors in more groups.
constexpr int kFftMaxPerThread = 16;

template <int SIGN>
__device__ __forceinline__ float2 mul_sign_i(float2 a) {
  // a * (SIGN * j): SIGN = -1 -> a * (-j) = (a.y, -a.x); SIGN = +1 -> (-a.y, a.x)
  return SIGN < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

template <int SIGN>
__device__ __forceinline__ float2 twiddle(const float2* tw, int idx) {
  const float2 w = tw[idx];
  return SIGN < 0 ? w : make_float2(w.x, -w.y);
}

template <int SIGN, int R, int MAXV>
__device__ __forceinline__ void fft_pass(float2* data, int count, int stride, int log2n, int log2ns,
                                         const float2* tw) {
  constexpr int kLog2R = R == 4 ? 2 : 1;
  constexpr int kItems = MAXV / R;
  const int log2nr = log2n - kLog2R;
  const int nr = 1 << log2nr;
  const int ns = 1 << log2ns;
  const int total = count << log2nr;
  float2 v[kItems][R];
  int dst[kItems];
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const int item = threadIdx.x + it * blockDim.x;
    if (item < total) {
      const int vec = item >> log2nr;
      const int j = item & (nr - 1);
      const float2* base = data + vec * stride;
      const int pos = j & (ns - 1);
      const int twstep = pos << (log2n - log2ns - kLog2R);  // pos * N / (ns R)
      if (R == 4) {
        float2 a0 = base[j], a1 = base[j + nr], a2 = base[j + 2 * nr], a3 = base[j + 3 * nr];
        if (ns > 1) {
          a1 = cmul(a1, twiddle<SIGN>(tw, twstep));
          a2 = cmul(a2, twiddle<SIGN>(tw, 2 * twstep));
          a3 = cmul(a3, twiddle<SIGN>(tw, 3 * twstep));
        }
        const float2 t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3);
        const float2 t3 = mul_sign_i<SIGN>(csub(a1, a3));
        v[it][0] = cadd(t0, t2);
        v[it][1 % R] = cadd(t1, t3);
        v[it][2 % R] = csub(t0, t2);
        v[it][3 % R] = csub(t1, t3);
      } else {
        float2 a0 = base[j], a1 = base[j + nr];
        if (ns > 1) a1 = cmul(a1, twiddle<SIGN>(tw, twstep));
        v[it][0] = cadd(a0, a1);
        v[it][1 % R] = csub(a0, a1);
      }
      dst[it] = vec * stride + ((j - pos) << kLog2R) + pos;
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const int item = threadIdx.x + it * blockDim.x;
    if (item < total) {
      float2* out = data + dst[it];
#pragma unroll
      for (int r = 0; r < R; ++r) out[r * ns] = v[it][r];
    }
  }
  __syncthreads();
}

// data: vector v at data + v * stride; tw: N entries (LDS or global).
// Must be called by every thread of the workgroup; ends with a barrier.
// Vectors are transformed in groups of at most MAXV * blockDim.x / N.
template <int SIGN, int MAXV = kFftMaxPerThread>
__device__ __forceinline__ void lds_fft(float2* data, int count, int stride, int log2n, const float2* tw) {
  static_assert(MAXV % 4 == 0, "MAXV must be a multiple of 4");
  const int group = max(1, (MAXV * static_cast<int>(blockDim.x)) >> log2n);
  for (int g0 = 0; g0 < count; g0 += group) {
    const int cnt = min(group, count - g0);
    float2* base = data + g0 * stride;
    int log2ns = 0;
    while (log2n - log2ns >= 2) {
      fft_pass<SIGN, 4, MAXV>(base, cnt, stride, log2n, log2ns, tw);
      log2ns += 2;
    }
    if (log2n - log2ns == 1) fft_pass<SIGN, 2, MAXV>(base, cnt, stride, log2n, log2ns, tw);
  }
}

// ---------------------------------------------------------------- mixed-radix variant
// Stockham passes of radix 16 (16-point DFTs in registers as 4 x 4), then one radix-4 and/or one
// radix-2 pass for the remaining factor (N = 1024: 16 16 4; 4096: 16 16 16): fewer passes, LDS
// round trips and barriers than radix 4 throughout.  A radix-R pass of stride ns reads
// x[j + k N/R] (consecutive lanes, consecutive elements), multiplies by the twiddles
// exp(-2 pi j k pos / (R ns)), pos = j mod ns, and writes output r to (j - pos) R + pos + r ns;
// for small ns those writes put consecutive lanes 16 elements apart, so the intermediate
// vectors live in a padded layout, element i at i + (i >> 5), which spreads them over the LDS
// banks (the first pass reads and the last pass writes the natural layout).  Twiddles come from
// a pass table: pass p's entries (R - 1) ns follow the previous passes', entry (k - 1) ns + pos,
// so consecutive lanes read consecutive entries; N - 1 entries in all.  A vector's padded
// extent is N + N / 32 elements: the stride between vectors must cover it (kFftNoPad: N).

constexpr int kFftPadShift = 5;
constexpr int kFftNoPad = 30;  // i >> 30 == 0 for every index of a vector

__device__ __forceinline__ int fft_pad(int i, int shift) { return i + (i >> shift); }

// Radix of pass p of an N = 2^log2n transform (16 while 4 bits remain, then 4, then 2).
__host__ __device__ __forceinline__ int fft_pass_log2_radix(int log2n, int p) {
  const int full = log2n >> 2;
  if (p < full) return 4;
  const int rem = log2n - 4 * full;
  if (p == full) return rem >= 2 ? 2 : rem;
  return (p == full + 1 && rem == 3) ? 1 : 0;
}

__host__ __device__ __forceinline__ int fft_num_passes(int log2n) {
  int p = 0;
  while (fft_pass_log2_radix(log2n, p) > 0) ++p;
  return p;
}

// Entries of the pass table of an N-point transform (N - 1).
__host__ __device__ __forceinline__ int fft_pass_table_size(int n) { return n - 1; }

// Fills the pass table of an N-point transform (N a power of two >= 2), float32 from float64;
// called by every thread of the workgroup (no barrier).
__device__ __forceinline__ void fill_fft_pass_table(float2* tab, int log2n) {
  int ns = 1, off = 0;
  for (int p = 0;; ++p) {
    const int lr = fft_pass_log2_radix(log2n, p);
    if (lr == 0) break;
    const int r = 1 << lr;
    for (int e = threadIdx.x; e < (r - 1) * ns; e += blockDim.x) {
      const int k = 1 + e / ns, pos = e - (k - 1) * ns;
      double sn, cs;
      sincospi(-2.0 * static_cast<double>(k * pos) / static_cast<double>(r * ns), &sn, &cs);
      tab[off + e] = make_float2(static_cast<float>(cs), static_cast<float>(sn));
    }
    off += (r - 1) * ns;
    ns *= r;
  }
}

template <int SIGN>
__device__ __forceinline__ float2 pass_twiddle(const float2* tab, int idx) {
  const float2 w = tab[idx];
  return SIGN < 0 ? w : make_float2(w.x, -w.y);
}

template <int SIGN>
__device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
  const float2 t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3);
  const float2 t3 = mul_sign_i<SIGN>(csub(a1, a3));
  a0 = cadd(t0, t2);
  a1 = cadd(t1, t3);
  a2 = csub(t0, t2);
  a3 = csub(t1, t3);
}

// x * exp(SIGN 2 pi j m / 16) for the products m = n2 k1 of the 4 x 4 split (0, 1, 2, 3, 4, 6, 9)
template <int SIGN, int M>
__device__ __forceinline__ float2 w16(float2 x) {
  constexpr float c[10] = {1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508977f, 0.0f,
                           -0.38268343236508977f, -0.70710678118654752f, -0.92387953251128674f, -1.0f,
                           -0.92387953251128674f};
  constexpr float sn[10] = {0.0f, 0.38268343236508977f, 0.70710678118654752f, 0.92387953251128674f, 1.0f,
                            0.92387953251128674f, 0.70710678118654752f, 0.38268343236508977f, 0.0f,
                            -0.38268343236508977f};
  if (M == 0) return x;
  if (M == 4) return mul_sign_i<SIGN>(x);
  return cmul(x, make_float2(c[M], SIGN * sn[M]));
}

// In-place 16-point DFT of v[0..15] (sign SIGN), output X[k] in v[k]: X[k1 + 4 k2] =
// sum_n2 W4^(n2 k2) W16^(n2 k1) sum_n1 x[4 n1 + n2] W4^(n1 k1).
template <int SIGN>
__device__ __forceinline__ void dft16(float2* v) {
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) dft4<SIGN>(v[n2], v[4 + n2], v[8 + n2], v[12 + n2]);  // v[4 k1 + n2]
  v[4 + 1] = w16<SIGN, 1>(v[4 + 1]);
  v[4 + 2] = w16<SIGN, 2>(v[4 + 2]);
  v[4 + 3] = w16<SIGN, 3>(v[4 + 3]);
  v[8 + 1] = w16<SIGN, 2>(v[8 + 1]);
  v[8 + 2] = w16<SIGN, 4>(v[8 + 2]);
  v[8 + 3] = w16<SIGN, 6>(v[8 + 3]);
  v[12 + 1] = w16<SIGN, 3>(v[12 + 1]);
  v[12 + 2] = w16<SIGN, 6>(v[12 + 2]);
  v[12 + 3] = w16<SIGN, 9>(v[12 + 3]);
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) dft4<SIGN>(v[4 * k1], v[4 * k1 + 1], v[4 * k1 + 2], v[4 * k1 + 3]);  // -> v[4 k1 + k2]
  // transpose 4 x 4: X[k1 + 4 k2] currently sits at v[4 k1 + k2]
  float2 t;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = i + 1; j < 4; ++j) {
      t = v[4 * i + j];
      v[4 * i + j] = v[4 * j + i];
      v[4 * j + i] = t;
    }
}

template <int SIGN, int LOG2R, int MAXV, bool PIN, bool POUT>
__device__ __forceinline__ void fft_pass_mixed(float2* data, int count, int stride, int log2n, int log2ns,
                                               const float2* tab, int shift) {
  constexpr int R = 1 << LOG2R;
  constexpr int kItems = MAXV / R > 0 ? MAXV / R : 1;
  const int log2nr = log2n - LOG2R;
  const int nr = 1 << log2nr;
  const int ns = 1 << log2ns;
  const int total = count << log2nr;
  // the thread index passes through an opaque copy: the per-thread index arithmetic is redone in
  // every pass instead of being hoisted out of the caller's loops into registers held throughout
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  float2 v[kItems][R];
  int dst[kItems];
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const int item = tid + it * blockDim.x;
    if (item < total) {
      const float2* base = data + (item >> log2nr) * stride;
      const int j = item & (nr - 1);
      const int pos = j & (ns - 1);
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const int i = j + k * nr;
        v[it][k] = base[PIN ? fft_pad(i, shift) : i];
      }
      if (ns > 1) {
        const float2* ptab = tab + (ns - 1);  // passes before this one hold ns - 1 entries
#pragma unroll
        for (int k = 1; k < R; ++k) v[it][k] = cmul(v[it][k], pass_twiddle<SIGN>(ptab, (k - 1) * ns + pos));
      }
      if (R == 16) {
        dft16<SIGN>(v[it]);
      } else if (R == 4) {
        dft4<SIGN>(v[it][0], v[it][1 % R], v[it][2 % R], v[it][3 % R]);
      } else {
        const float2 a0 = v[it][0], a1 = v[it][1 % R];
        v[it][0] = cadd(a0, a1);
        v[it][1 % R] = csub(a0, a1);
      }
      dst[it] = ((j - pos) << LOG2R) + pos;  // within the vector
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const int item = tid + it * blockDim.x;
    if (item < total) {
      float2* row = data + (item >> log2nr) * stride;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int i = dst[it] + r * ns;
        row[POUT ? fft_pad(i, shift) : i] = v[it][r];
      }
    }
  }
  __syncthreads();
}

template <int SIGN, int LOG2R, int MAXV>
__device__ __forceinline__ void fft_pass_layout(float2* data, int count, int stride, int log2n, int log2ns,
                                                const float2* tab, int shift, bool first, bool last) {
  if (first && last) fft_pass_mixed<SIGN, LOG2R, MAXV, false, false>(data, count, stride, log2n, log2ns, tab, shift);
  else if (first) fft_pass_mixed<SIGN, LOG2R, MAXV, false, true>(data, count, stride, log2n, log2ns, tab, shift);
  else if (last) fft_pass_mixed<SIGN, LOG2R, MAXV, true, false>(data, count, stride, log2n, log2ns, tab, shift);
  else fft_pass_mixed<SIGN, LOG2R, MAXV, true, true>(data, count, stride, log2n, log2ns, tab, shift);
}

// Strength-reduced passes for N >= 512 in the padded layout of shift kFftPadShift, where every
// pass's access pattern is fixed by its position: pass 0 reads the natural layout and writes
// outputs r = 0..15 of an item at dst + r (dst a multiple of 16: one padding group, so the padded
// offsets are pad(dst) + r); pass 1 reads at j + k nr (nr = N / 16, a multiple of 32: padded
// pad(j) + k (nr + nr / 32)) and writes at dst + 16 r (the even r at pad(dst) + 16 r + r / 2, the
// odd r at pad(dst + 16) + 16 (r - 1) + (r - 1) / 2); later passes read and write at multiples
// of 32 apart (base plus a padded stride), the last one in the natural layout.  Each access is
// then one address add or an immediate offset instead of the per-element padding arithmetic.
enum { kFastInNatural = 0, kFastInPadded = 1 };
enum { kFastOutUnit = 0, kFastOutHalf = 1, kFastOutPadded = 2, kFastOutNatural = 3 };

template <int SIGN, int LOG2R, int MAXV, int IN, int OUT>
__device__ __forceinline__ void fft_pass_fast(float2* data, int count, int stride, int log2n, int log2ns,
                                              const float2* tab) {
  constexpr int R = 1 << LOG2R;
  constexpr int S = kFftPadShift;
  constexpr int kItems = MAXV / R > 0 ? MAXV / R : 1;
  const int log2nr = log2n - LOG2R;
  const int nr = 1 << log2nr;
  const int ns = 1 << log2ns;
  const int total = count << log2nr;
  const int in_step = IN == kFastInNatural ? nr : nr + (nr >> S);
  const int out_step = OUT == kFastOutPadded ? ns + (ns >> S) : ns;
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  float2 v[kItems][R];
  int dst[kItems];
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const int item = tid + it * blockDim.x;
    if (item < total) {
      const int j = item & (nr - 1);
      const int pos = j & (ns - 1);
      const float2* bp = data + (item >> log2nr) * stride + (IN == kFastInNatural ? j : fft_pad(j, S));
#pragma unroll
      for (int k = 0; k < R; ++k) v[it][k] = bp[k * in_step];
      if (ns > 1) {
        const float2* ptab = tab + (ns - 1) + pos;  // passes before this one hold ns - 1 entries
#pragma unroll
        for (int k = 1; k < R; ++k) v[it][k] = cmul(v[it][k], pass_twiddle<SIGN>(ptab, (k - 1) * ns));
      }
      if (R == 16) {
        dft16<SIGN>(v[it]);
      } else if (R == 4) {
        dft4<SIGN>(v[it][0], v[it][1 % R], v[it][2 % R], v[it][3 % R]);
      } else {
        const float2 a0 = v[it][0], a1 = v[it][1 % R];
        v[it][0] = cadd(a0, a1);
        v[it][1 % R] = csub(a0, a1);
      }
      dst[it] = ((j - pos) << LOG2R) + pos;
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const int item = tid + it * blockDim.x;
    if (item < total) {
      float2* row = data + (item >> log2nr) * stride;
      const int d = dst[it];
      if (OUT == kFastOutUnit) {
        float2* b = row + fft_pad(d, S);
#pragma unroll
        for (int r = 0; r < R; ++r) b[r] = v[it][r];
      } else if (OUT == kFastOutHalf) {  // ns = 2^(S - 1)
        float2* b0 = row + fft_pad(d, S);
        float2* b1 = row + fft_pad(d + (1 << (S - 1)), S);
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if (r & 1) b1[(r - 1) * (1 << (S - 1)) + (r - 1) / 2] = v[it][r];
          else b0[r * (1 << (S - 1)) + r / 2] = v[it][r];
        }
      } else {
        float2* b = row + (OUT == kFastOutPadded ? fft_pad(d, S) : d);
#pragma unroll
        for (int r = 0; r < R; ++r) b[r * out_step] = v[it][r];
      }
    }
  }
  __syncthreads();
}

template <int SIGN, int LOG2R, int MAXV, int IN>
__device__ __forceinline__ void fft_pass_fast_out(float2* data, int count, int stride, int log2n, int log2ns,
                                                  const float2* tab, bool last) {
  if (last) fft_pass_fast<SIGN, LOG2R, MAXV, IN, kFastOutNatural>(data, count, stride, log2n, log2ns, tab);
  else fft_pass_fast<SIGN, LOG2R, MAXV, IN, kFastOutPadded>(data, count, stride, log2n, log2ns, tab);
}

// In-place complex FFT (sign SIGN, no scaling) of `count` vectors at data + v * stride in LDS,
// natural order in and out, with the pass table `tab` (fill_fft_pass_table) and the padded
// intermediate layout of pad shift `shift`.  Called by every thread; ends with a barrier.
// Vectors are transformed in groups of at most MAXV * blockDim.x / N.
template <int SIGN, int MAXV = kFftMaxPerThread>
__device__ __forceinline__ void lds_fft_padded(float2* data, int count, int stride, int log2n, const float2* tab,
                                               int shift = kFftPadShift) {
  static_assert(MAXV % 4 == 0, "MAXV must be a multiple of 4");
  const int group = max(1, (MAXV * static_cast<int>(blockDim.x)) >> log2n);
  const int passes = fft_num_passes(log2n);
  const bool fast = shift == kFftPadShift && log2n >= 9;  // N >= 512: passes 0 and 1 are radix 16
  for (int g0 = 0; g0 < count; g0 += group) {
    const int cnt = min(group, count - g0);
    float2* base = data + g0 * stride;
    if (fast) {
      fft_pass_fast<SIGN, 4, MAXV, kFastInNatural, kFastOutUnit>(base, cnt, stride, log2n, 0, tab);
      fft_pass_fast<SIGN, 4, MAXV, kFastInPadded, kFastOutHalf>(base, cnt, stride, log2n, 4, tab);
      int log2ns = 8;
      for (int p = 2; p < passes; ++p) {
        const int lr = fft_pass_log2_radix(log2n, p);
        const bool last = p == passes - 1;
        if (lr == 4) fft_pass_fast_out<SIGN, 4, MAXV, kFastInPadded>(base, cnt, stride, log2n, log2ns, tab, last);
        else if (lr == 2) fft_pass_fast_out<SIGN, 2, MAXV, kFastInPadded>(base, cnt, stride, log2n, log2ns, tab, last);
        else fft_pass_fast_out<SIGN, 1, MAXV, kFastInPadded>(base, cnt, stride, log2n, log2ns, tab, last);
        log2ns += lr;
      }
      continue;
    }
    int log2ns = 0;
    for (int p = 0; p < passes; ++p) {
      const int lr = fft_pass_log2_radix(log2n, p);
      const bool first = p == 0, last = p == passes - 1;
      if (lr == 4) fft_pass_layout<SIGN, 4, MAXV>(base, cnt, stride, log2n, log2ns, tab, shift, first, last);
      else if (lr == 2) fft_pass_layout<SIGN, 2, MAXV>(base, cnt, stride, log2n, log2ns, tab, shift, first, last);
      else fft_pass_layout<SIGN, 1, MAXV>(base, cnt, stride, log2n, log2ns, tab, shift, first, last);
      log2ns += lr;
    }
  }
}

// tw[k] = exp(-2 pi j k / N), k < N, float32 from float64, called by every thread of the
// workgroup (no barrier).  Computes the N / 8 + 1 angles of the first octant and fills the
// other seven by the symmetries cos(pi/2 - x) = sin(x) and exp(-j (x + q pi/2)) = (-j)^q exp(-j x),
// which cuts the float64 sine/cosine evaluations eightfold (N >= 8).
__device__ __forceinline__ void fill_twiddles_octants(float2* tw, int n) {
  const int quarter = n >> 2, eighth = n >> 3;
  for (int r = threadIdx.x; r <= eighth; r += blockDim.x) {
    double sd, cd;
    sincospi(2.0 * static_cast<double>(r) / static_cast<double>(n), &sd, &cd);
    const float c = static_cast<float>(cd), sn = static_cast<float>(sd);
    // angles r and quarter - r of the first quadrant: exp(-j x) = (c, -s), exp(-j (pi/2 - x)) = (s, -c)
    const int idx[2] = {r, quarter - r};
    const float cc[2] = {c, sn}, ss[2] = {sn, c};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if (h == 1 && (r == 0 || idx[1] == idx[0])) continue;  // r = 0: quarter - r lies in the next quadrant
      const int k = idx[h];
      const float x = cc[h], y = ss[h];
      tw[k] = make_float2(x, -y);                  // q = 0
      tw[k + quarter] = make_float2(-y, -x);       // q = 1: times -j
      tw[k + 2 * quarter] = make_float2(-x, y);    // q = 2: times -1
      tw[k + 3 * quarter] = make_float2(y, x);     // q = 3: times +j
    }
  }
}

// Grid column q of an N-subcarrier OFDM symbol -> FFT bin (composition of the
// DC roll and (i)fftshift of hermespy/modem/waveforms/orthogonal/ofdm.py:116-128,144-147).
__host__ __device__ __forceinline__ int subcarrier_bin(int q, int n, int dc_suppression) {
  int p = q;
  const int dc = n / 2;
  if (dc_suppression && q >= dc) {
    const int len = n - dc;
    int t = q - dc + 1;
    if (t >= len) t -= len;
    p = dc + t;
  }
  int b = p - n / 2;
  if (b < 0) b += n;
  return b;
}

}  // namespace hpx
