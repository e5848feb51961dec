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
