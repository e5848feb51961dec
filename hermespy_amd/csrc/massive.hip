// Massive-MIMO uplink drops: many receive antennas (up to 64), single-antenna users
// (up to 8), large OFDM grids with a centred active band.  One drop does not fit a
// CU, so the drop runs as three kernels over caller-owned per-drop buffers:
//
//   K0 prologue  (workgroup per drop)       realisation (Philox -> Phi -> S = R_rx exp(2 pi j U) R_tx,
//                                           SoS table), Taylor coefficients of the taps, per-word tap
//                                           means, and the per-drop detector V, lambda of S^H S,
//                                           A = V^H S^H and B = A S (fading.py:468-515,
//                                           channel_estimation.py:215-295)
//   K1 transmit  (workgroup per drop and user pair, stepping through the words)
//                                           bits -> 256-QAM -> OFDM + CP -> the user's tapped delay
//                                           line over each word's data window (modem.py:797-888,
//                                           ofdm.py:97-130, fading.py:371-406)
//   K3 receive + detect (workgroup per drop, word)
//                                           the matched-filter outputs u = A y of every active
//                                           subcarrier straight from the users' delay-line outputs
//                                           and the antennas' noise (below), the ideal-CSI factor c,
//                                           then x = V diag(conj(c) / (|c|^2 lambda + s2)) u, exact
//                                           3GPP hard decisions and bit errors (ofdm.py:132-161,
//                                           waveform.py:235-250, psk_qam_mapping.py:171-259,
//                                           evaluators.py:239-258)
//
// K3 never forms the per-antenna grid y_r = FFT(sum_t S[r][t] z_t + sigma n_r): the FFT is linear,
// so u_i = sum_r A[i][r] y_r = FFT(sum_t B[i][t] z_t + sigma sum_r A[i][r] n_r) with the very noise
// samples n_r each antenna adds in the time domain.  That is one transform per user instead of
// one per antenna, and no grid round trip through HBM.  K2 (massive_rx_kernel) writes the
// per-antenna grid and the CSI factor only when the caller asks for them.
//
// The reference caps arrays at 10 antennas (its spatial response slices one 10 x 10 block of
// uniforms, fading.py:476-489); here the block is (max(10, Nrx), max(10, Ntx)), which keeps the
// reference's draw layout for links of at most 10 antennas.
#include <math.h>

#include <algorithm>
#include <type_traits>

#include "hpx_common.hpp"
#include "hpx_eig.hpp"
#include "hpx_fading.hpp"
#include "hpx_fft.hpp"
#include "hpx_mapping.hpp"

namespace hpx {
int taylor_log2_block(double max_rate, int terms, int max_log2);
int ilog2_exact(int n);
}  // namespace hpx

using namespace hpx;

namespace {

constexpr int kMaxRx = 64, kMaxUsers = 8;
constexpr int kTxThreads = 1024, kRxThreads = 1024;
constexpr int kTxUsers = 2;      // users per transmit workgroup (they share the tap evaluations)
constexpr int kTxSamples = 4;    // data-window samples per thread of K1 (N <= 4 kTxThreads)
constexpr int kRxGroup = 8;      // antennas per receive workgroup
constexpr int kRxRows = 2;       // antennas transformed together
typedef float floatx4 __attribute__((ext_vector_type(4)));
constexpr int kXcds = 8;         // MI355X dispatches consecutive workgroups round robin over 8 XCDs
constexpr int kLinkThreads = 1024;
constexpr int kLinkPasses = 8;   // 16-pair column tiles per wave of K3 (N <= 2 * 16 * 8 * 16 waves)
constexpr int kLinkBins = 4;     // active subcarriers per thread of K3 (active <= 4 kLinkThreads)
constexpr int kLinkRows = 3;     // rows per transform round of K3 (users, then the CSI row)

inline hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

struct MassiveArgs {
  int nr, nt, n, log2n, cp, words, active, active_lo, bps, order, dc;
  Modulation mod;
  int paths, sinusoids, terms, nnorm, blk_cols;
  int log2b, nb, maxd, dx, cpl, x0, xs, keep_below, bits_words;
  int64_t frame_len;
  int64_t num_pairs;  // (drop, word) pairs of the receive launch
  double bandwidth, doppler, los_doppler, tx_scale;
  const int32_t* delays;
  const double* path_gains;
  const double* corr_rx;
  const double* corr_tx;
  const double* noise_power;
  uint64_t seed;
  int64_t first_drop;
  float2* spatial;  // [drop][nr][nt]
  float2* coef;     // [drop][L][nb][kTaylorCoeffs]
  float2* abar;     // [drop][L][words]
  float2* det_v;    // [drop][nt][nt]
  float2* det_a;    // [drop][nt][nr] A, then [nt][nt] B = A S
  int da_stride;    // nt nr + nt nt
  float* det_l;     // [drop][nt]
  float2* z;        // [drop][nt][words][n]
  float2* grid;     // [drop][words][nr][active] or null (then csi is null too)
  float2* csi;      // [drop][words][active]
  int32_t* errors;  // [drop], accumulated
  float2* equalized;  // [drop][nt][words][active] or null
  int k0_uni, k0_sos, k0_e, k0_f, k0_s64, k0_det, k0_lds;
  int k1_x, k1_tab, k1_coef, k1_bits, k1_cons, k1_dly, k1_lds;
  int k2_y, k2_tw, k2_lds;
  int k3_rows, k3_rstride, k3_tab, k3_atab, k3_v, k3_l, k3_dmb, k3_red, k3_lds;
  DemapBounds dbounds;  // float32 decision boundaries of mod (exact hard decisions)
};

__device__ __forceinline__ uint32_t symbol_bits_at(const uint32_t* bw, int bit, int bps) {
  uint64_t v = bw[bit >> 5];
  if ((bit & 31) + bps > 32) v |= static_cast<uint64_t>(bw[(bit >> 5) + 1]) << 32;
  return static_cast<uint32_t>(v >> (bit & 31)) & ((1u << bps) - 1u);
}

// ---------------------------------------------------------------- K0: per-drop prologue
__global__ void __launch_bounds__(256) massive_prologue_kernel(MassiveArgs a) {
  extern __shared__ __align__(16) unsigned char lds[];
  double* uni = reinterpret_cast<double*>(lds + a.k0_uni);
  double* sos = reinterpret_cast<double*>(lds + a.k0_sos);
  double* e = reinterpret_cast<double*>(lds + a.k0_e);    // exp(2 pi j U) [nr][nt]
  double* f = reinterpret_cast<double*>(lds + a.k0_f);    // E R_tx
  double* s64 = reinterpret_cast<double*>(lds + a.k0_s64);
  double* det = reinterpret_cast<double*>(lds + a.k0_det);
  const int tid = threadIdx.x, nth = blockDim.x;
  const int64_t ld = blockIdx.x;
  const uint32_t drop = static_cast<uint32_t>(a.first_drop + ld);
  const int nr = a.nr, nt = a.nt, L = a.paths;
  for (int c = tid; 2 * c < a.nnorm; c += nth) {
    const double2 zz = normal_pair_f64(stream_call(a.seed, drop, HPX_STREAM_REALIZATION, static_cast<uint32_t>(c)));
    uni[2 * c] = normal_cdf(zz.x);
    if (2 * c + 1 < a.nnorm) uni[2 * c + 1] = normal_cdf(zz.y);
  }
  __syncthreads();
  for (int i = tid; i < nr * nt; i += nth) {
    const int p = i / nt, q = i - p * nt;
    double sn, cs;
    sincospi(2.0 * uni[p * a.blk_cols + q], &sn, &cs);
    e[2 * i] = cs;
    e[2 * i + 1] = sn;
  }
  __syncthreads();
  // F = E R_tx, S = R_rx F (fading.py:476-489)
  for (int i = tid; i < nr * nt; i += nth) {
    const int p = i / nt, t = i - p * nt;
    double fr = e[2 * i], fi = e[2 * i + 1];
    if (a.corr_tx) {
      fr = fi = 0.0;
      for (int q = 0; q < nt; ++q) {
        const double er = e[2 * (p * nt + q)], ei = e[2 * (p * nt + q) + 1];
        const double br = a.corr_tx[2 * (q * nt + t)], bi = a.corr_tx[2 * (q * nt + t) + 1];
        fr += er * br - ei * bi;
        fi += er * bi + ei * br;
      }
    }
    f[2 * i] = fr;
    f[2 * i + 1] = fi;
  }
  __syncthreads();
  for (int i = tid; i < nr * nt; i += nth) {
    const int r = i / nt, t = i - r * nt;
    double sr = f[2 * i], si = f[2 * i + 1];
    if (a.corr_rx) {
      sr = si = 0.0;
      for (int p = 0; p < nr; ++p) {
        const double ar = a.corr_rx[2 * (r * nr + p)], ai = a.corr_rx[2 * (r * nr + p) + 1];
        const double br = f[2 * (p * nt + t)], bi = f[2 * (p * nt + t) + 1];
        sr += ar * br - ai * bi;
        si += ar * bi + ai * br;
      }
    }
    s64[2 * i] = sr;
    s64[2 * i + 1] = si;
    a.spatial[ld * nr * nt + i] = make_float2(static_cast<float>(sr), static_cast<float>(si));
  }
  // SoS table (fading.py:326-342, angles/phases :492-499): NLOS terms, then LOS
  {
    const int ns = a.sinusoids;
    const int o0 = (nr > 10 ? nr : 10) * a.blk_cols;
    const int o_los_ang = o0, o_nlos_ang = o0 + L, o_los_ph = o0 + L + L * ns, o_nlos_ph = o0 + 2 * L + L * ns;
    const double inv_fs = 1.0 / a.bandwidth;
    for (int i = tid; i < L * a.terms; i += nth) {
      const int p = i / a.terms, k = i - p * a.terms;
      double* row = sos + 3 * i;
      if (k < ns) {
        const double ang = -M_PI + 2.0 * M_PI * uni[o_nlos_ang + p * ns + k];
        row[0] = a.doppler * cos((2.0 * M_PI * (k + 1) + ang) / ns) * inv_fs;
        row[1] = -M_PI + 2.0 * M_PI * uni[o_nlos_ph + p * ns + k];
        row[2] = a.path_gains[L + p];
      } else {
        const double ang = 2.0 * M_PI * uni[o_los_ang + p];
        row[0] = a.los_doppler * cos(ang) * inv_fs;
        row[1] = -M_PI + 2.0 * M_PI * uni[o_los_ph + p];
        row[2] = a.path_gains[p];
      }
    }
  }
  __syncthreads();
  const TaylorBlocks tb{a.log2b};
  float2* coef = a.coef + ld * L * a.nb * kTaylorCoeffs;
  for (int i = tid; i < L * a.nb; i += nth) {
    const int p = i / a.nb, b = i - p * a.nb;
    taylor_coeffs(sos + 3 * p * a.terms, a.terms, tb.centre(b), tb.half_width(), coef + i * kTaylorCoeffs);
  }
  for (int i = tid; i < L * a.words; i += nth) {
    const int p = i / a.words, w = i - p * a.words;
    const double2 m = tap_mean(sos + 3 * p * a.terms, a.terms, static_cast<double>(w) * (a.n + a.cp) + a.cp, a.n);
    a.abar[ld * L * a.words + i] = make_float2(static_cast<float>(m.x), static_cast<float>(m.y));
  }
  // The per-drop detector of spatial_detector (hpx_eig.hpp) spread over the workgroup: the
  // Gram matrix S^H S by nt * nt threads, its Jacobi eigen-decomposition by one wavefront,
  // A = V^H S^H by all threads (same sums in the same order).
  double* work = det + nt + 2 * nt * nt + 2 * nt * nr;
  double* gre = work;
  double* gim = work + kEigMax * kEigMax;
  double* vre = work + 2 * kEigMax * kEigMax;
  double* vim = work + 3 * kEigMax * kEigMax;
  for (int e = tid; e < nt * nt; e += nth) {
    const int i = e / nt, j = e - i * nt;
    double sr = 0.0, si = 0.0;
    for (int r = 0; r < nr; ++r) {  // conj(S_ri) S_rj
      const double ar = s64[2 * (r * nt + i)], ai = s64[2 * (r * nt + i) + 1];
      const double br = s64[2 * (r * nt + j)], bi = s64[2 * (r * nt + j) + 1];
      sr += ar * br + ai * bi;
      si += ar * bi - ai * br;
    }
    gre[e] = sr;
    gim[e] = si;
  }
  __syncthreads();
  if (tid < 64) hermitian_jacobi_wave(nt, gre, gim, vre, vim);
  __syncthreads();
  if (tid == 0) {
    double lmax = 0.0;
    for (int i = 0; i < nt; ++i) lmax = fmax(lmax, gre[i * nt + i]);
    // numpy.linalg.pinv's cut: singular values below max(M, N) eps sigma_max are dropped
    const double rt = static_cast<double>(nr > nt ? nr : nt) * 2.220446049250313e-16;
    for (int i = 0; i < nt; ++i) {
      const double lam = gre[i * nt + i];
      a.det_l[ld * nt + i] = lam > rt * rt * lmax ? static_cast<float>(lam) : 0.0f;
    }
  }
  __syncthreads();
  for (int i = tid; i < nt * nt; i += nth)
    a.det_v[ld * nt * nt + i] = make_float2(static_cast<float>(vre[i]), static_cast<float>(vim[i]));
  double* a64 = det + nt + 2 * nt * nt;  // A [nt][nr], float64
  float2* da = a.det_a + ld * a.da_stride;
  for (int e = tid; e < nt * nr; e += nth) {
    const int i = e / nr, r = e - i * nr;
    double ar = 0.0, ai = 0.0;
    for (int k = 0; k < nt; ++k) {  // A[i][r] = sum_k conj(V[k][i]) conj(S[r][k])
      const double vr = vre[k * nt + i], vi = -vim[k * nt + i];
      const double sr = s64[2 * (r * nt + k)], si = -s64[2 * (r * nt + k) + 1];
      ar += vr * sr - vi * si;
      ai += vr * si + vi * sr;
    }
    a64[2 * e] = ar;
    a64[2 * e + 1] = ai;
    da[e] = make_float2(static_cast<float>(ar), static_cast<float>(ai));
  }
  __syncthreads();
  for (int e = tid; e < nt * nt; e += nth) {
    const int i = e / nt, t = e - i * nt;
    double br = 0.0, bi = 0.0;
    for (int r = 0; r < nr; ++r) {  // B[i][t] = sum_r A[i][r] S[r][t]
      const double xr = a64[2 * (i * nr + r)], xi = a64[2 * (i * nr + r) + 1];
      const double sr = s64[2 * (r * nt + t)], si = s64[2 * (r * nt + t) + 1];
      br += xr * sr - xi * si;
      bi += xr * si + xi * sr;
    }
    da[nt * nr + e] = make_float2(static_cast<float>(br), static_cast<float>(bi));
  }
}

// ---------------------------------------------------------------- K1: transmitter + delay line
// One workgroup per (drop, kTxUsers users), stepping through the words: bits, mapping and IFFT of
// word w of each user, the cyclic prefix, then the tapped delay line over the word's data window,
// with each tap value evaluated once for the group's users.  The delay line reaches dx samples
// back into word w - 1; those samples stay in LDS from the previous step (silence before the
// frame).  Per user, x holds [word w - 1's last dx samples | prefix | word w], the word in the
// FFT's padded extent.
__global__ void __launch_bounds__(kTxThreads) massive_tx_kernel(MassiveArgs a) {
  extern __shared__ __align__(16) unsigned char lds[];
  float2* x = reinterpret_cast<float2*>(lds + a.k1_x);        // [kTxUsers][xs]
  float2* ptab = reinterpret_cast<float2*>(lds + a.k1_tab);   // FFT pass table
  float2* coef = reinterpret_cast<float2*>(lds + a.k1_coef);  // [L][nb][5]
  uint32_t* bits = reinterpret_cast<uint32_t*>(lds + a.k1_bits);  // [kTxUsers][bits_words]
  float2* cons = reinterpret_cast<float2*>(lds + a.k1_cons);
  int* dly = reinterpret_cast<int*>(lds + a.k1_dly);  // path delays
  const int tid = threadIdx.x, nth = blockDim.x;
  const int groups = (a.nt + kTxUsers - 1) / kTxUsers;
  const int64_t ld = blockIdx.x / groups;
  const int t0 = static_cast<int>(blockIdx.x % groups) * kTxUsers;
  const int nu = min(kTxUsers, a.nt - t0);
  const uint32_t drop = static_cast<uint32_t>(a.first_drop + ld);
  const int n = a.n, L = a.paths, x0 = a.x0, xs = a.xs, bw = a.bits_words;
  const int ncoef = L * a.nb * kTaylorCoeffs;
  const int lane = tid & 63;
  const int wbase = __builtin_amdgcn_readfirstlane(tid);  // the wavefront's first thread
  for (int i = tid; i < ncoef; i += nth) coef[i] = a.coef[ld * ncoef + i];
  for (int p = tid; p < L; p += nth) dly[p] = a.delays[p];
  for (int v = tid; v < a.order; v += nth) {
    const double2 p = constellation_point(a.order, 1, bits_to_label(static_cast<uint32_t>(v), a.bps));
    cons[v] = make_float2(static_cast<float>(p.x * a.tx_scale), static_cast<float>(p.y * a.tx_scale));
  }
  for (int e = tid; e < kTxUsers * a.dx; e += nth) {
    const int u = e / a.dx;
    x[u * xs + (e - u * a.dx)] = make_float2(0.f, 0.f);
  }
  fill_fft_pass_table(ptab, a.log2n);
  const TaylorBlocks tb{a.log2b};
  for (int w = 0; w < a.words; ++w) {
    // bits of word w of each user (streams HPX_STREAM_BITS + t)
    const int64_t bit0 = static_cast<int64_t>(w) * a.active * a.bps;
    const int bo = static_cast<int>(bit0 & 31);
    const int fw = static_cast<int>(bit0 >> 5);
    const int lw = static_cast<int>((bit0 + static_cast<int64_t>(a.active) * a.bps - 1) >> 5);
    const int calls = (lw >> 2) - (fw >> 2) + 1;
    for (int e = tid; e < nu * calls; e += nth) {
      const int u = e / calls, c = (fw >> 2) + (e - u * calls);
      const uint4 r = stream_call(a.seed, drop, HPX_STREAM_BITS + t0 + u, static_cast<uint32_t>(c));
      const uint32_t comp[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int wi = 4 * c + k - fw;
        if (wi >= 0 && wi <= lw - fw) bits[u * bw + wi] = comp[k];
      }
    }
    __syncthreads();  // also orders the previous step's tail copies before the mapping
    for (int e = tid; e < nu * n; e += nth) {
      const int u = e >> a.log2n, q = e & (n - 1);
      const int el = q - a.active_lo;
      x[u * xs + x0 + subcarrier_bin(q, n, a.dc)] =
          (el >= 0 && el < a.active) ? cons[symbol_bits_at(bits + u * bw, bo + el * a.bps, a.bps)]
                                     : make_float2(0.f, 0.f);
    }
    __syncthreads();
    lds_fft_padded<+1>(x + x0, nu, xs, a.log2n, ptab);
    for (int e = tid; e < nu * a.cpl; e += nth) {
      const int u = e / a.cpl, i = e - u * a.cpl;
      x[u * xs + x0 - a.cpl + i] = x[u * xs + x0 + n - a.cpl + i];
    }
    __syncthreads();
    // tapped delay line over the data window (the receiver keeps nothing else)
    const int win0 = w * (n + a.cp) + a.cp;
    // Wavefront v filters samples 64 kTxSamples v + 64 s + lane (s < kTxSamples): when they all
    // lie in one Taylor block (blocks hold >= 64 samples), one broadcast read of the block's
    // coefficients serves the wavefront's kTxSamples samples; otherwise each lane reads its own.
    // All kTxUsers rows are filtered (a missing user's row is never stored).
    float2 acc[kTxUsers][kTxSamples];
#pragma unroll
    for (int u = 0; u < kTxUsers; ++u)
#pragma unroll
      for (int s = 0; s < kTxSamples; ++s) acc[u][s] = make_float2(0.f, 0.f);
    const int nlw = wbase * kTxSamples;  // the wavefront's first sample (wave-uniform)
    if (nlw < n) {
      for (int p = 0; p < L; ++p) {
        const int d = dly[p];
        if (d < 0 || d > a.maxd) continue;
        const int mw = win0 + nlw - d;
        const int mlast = mw + min(64 * kTxSamples, n - nlw) - 1;
        const float2* cp = coef + p * a.nb * kTaylorCoeffs;
        if (mw >= 0 && (mw >> a.log2b) == (mlast >> a.log2b)) {
          const float2* cb = cp + (mw >> a.log2b) * kTaylorCoeffs;
          float2 cc[kTaylorCoeffs];
#pragma unroll
          for (int k = 0; k < kTaylorCoeffs; ++k) cc[k] = cb[k];
#pragma unroll
          for (int s = 0; s < kTxSamples; ++s) {
            const int nl = nlw + 64 * s + lane;
            const float2 h = taylor_eval(cc, tb.offset32(mw + 64 * s + lane));
#pragma unroll
            for (int u = 0; u < kTxUsers; ++u) acc[u][s] = cfma(h, x[u * xs + x0 + nl - d], acc[u][s]);
          }
        } else {
#pragma unroll
          for (int s = 0; s < kTxSamples; ++s) {
            const int nl = nlw + 64 * s + lane;
            const int m = mw + 64 * s + lane;
            if (nl >= n || m < 0) continue;
            const float2 h = taylor_eval(cp + (m >> a.log2b) * kTaylorCoeffs, tb.offset32(m));
#pragma unroll
            for (int u = 0; u < kTxUsers; ++u) acc[u][s] = cfma(h, x[u * xs + x0 + nl - d], acc[u][s]);
          }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < kTxUsers; ++u) {
      if (u >= nu) continue;
      float2* zt = a.z + ((ld * a.nt + t0 + u) * a.words + w) * static_cast<int64_t>(n);
#pragma unroll
      for (int s = 0; s < kTxSamples; ++s) {
        const int nl = nlw + 64 * s + lane;
        if (nl < n) zt[nl] = acc[u][s];
      }
    }
    __syncthreads();
    // word w's last dx samples become the next word's tail
    for (int e = tid; e < nu * a.dx; e += nth) {
      const int u = e / a.dx, i = e - u * a.dx;
      x[u * xs + i] = x[u * xs + x0 + n - a.dx + i];
    }
  }
}

// ---------------------------------------------------------------- K2: receive antennas
__global__ void __launch_bounds__(kRxThreads) massive_rx_kernel(MassiveArgs a) {
  extern __shared__ __align__(16) unsigned char lds[];
  float2* y = reinterpret_cast<float2*>(lds + a.k2_y);  // [kRxRows + 1][n]
  float2* tw = reinterpret_cast<float2*>(lds + a.k2_tw);
  const int tid = threadIdx.x, nth = blockDim.x;
  const int groups = (a.nr + kRxGroup - 1) / kRxGroup;
  // Workgroups b, b + 8, b + 16, ... share an XCD and its L2.  All antenna groups of one
  // (drop, word) get block indices of one residue mod 8, so the users' delay-line outputs z
  // they all read come from one L2: (drop, word) pair = x + 8 (q / groups), group = q % groups
  // for b = x + 8 q.  The grid is padded to whole sets of 8 pairs.
  const int64_t b = blockIdx.x;
  const int64_t q = b / kXcds;
  const int64_t pair = b % kXcds + kXcds * (q / groups);
  if (pair >= a.num_pairs) return;
  const int g = static_cast<int>(q % groups);
  const int w = static_cast<int>(pair % a.words);
  const int64_t ld = pair / a.words;
  const uint32_t drop = static_cast<uint32_t>(a.first_drop + ld);
  const int n = a.n, nt = a.nt;
  fill_twiddles_octants(tw, n);
  const float sigma = static_cast<float>(sqrt(0.5 * a.noise_power[ld]));
  const float rx_scale = static_cast<float>(1.0 / sqrt(static_cast<double>(n)));
  const int64_t win0 = static_cast<int64_t>(w) * (n + a.cp) + a.cp;
  const float2* zd = a.z + ld * nt * static_cast<int64_t>(a.words) * n + static_cast<int64_t>(w) * n;
  const float2* sp = a.spatial + ld * a.nr * nt;
  const int r_end = std::min(a.nr, (g + 1) * kRxGroup);
  const bool with_csi = g == 0;
  for (int r0 = g * kRxGroup; r0 < r_end; r0 += kRxRows) {
    const int rows = std::min(kRxRows, r_end - r0);
    __syncthreads();
    // two samples per thread and every row of the round at once: one load of the users' z feeds
    // all rows, one Philox call serves both noise samples of an aligned pair
    const int half = n >> 1;
    for (int i = tid; i < half; i += nth) {
      const int nl = 2 * i;
      float2 acc[kRxRows][2];
#pragma unroll
      for (int rr = 0; rr < kRxRows; ++rr) acc[rr][0] = acc[rr][1] = make_float2(0.f, 0.f);
      for (int t = 0; t < nt; ++t) {
        const float4 zz = *reinterpret_cast<const float4*>(zd + static_cast<int64_t>(t) * a.words * n + nl);
#pragma unroll
        for (int rr = 0; rr < kRxRows; ++rr) {
          const float2 st = rr < rows ? sp[(r0 + rr) * nt + t] : make_float2(0.f, 0.f);
          acc[rr][0] = cfma(st, make_float2(zz.x, zz.y), acc[rr][0]);
          acc[rr][1] = cfma(st, make_float2(zz.z, zz.w), acc[rr][1]);
        }
      }
      const int64_t nabs = win0 + nl;
#pragma unroll
      for (int rr = 0; rr < kRxRows; ++rr) {
        if (rr >= rows) break;
        const int r = r0 + rr;
        const float2 acc0 = acc[rr][0], acc1 = acc[rr][1];
        // noise sample nabs: words (0, 1) or (2, 3) of Philox call nabs / 2 (stream HPX_STREAM_NOISE + r)
        float2 n0, n1;
        const uint4 wa = stream_call(a.seed, drop, HPX_STREAM_NOISE + r, static_cast<uint32_t>(nabs >> 1));
        if ((nabs & 1) == 0) {
          n0 = complex_normal_f32(wa.x, wa.y);
          n1 = complex_normal_f32(wa.z, wa.w);
        } else {
          const uint4 wb = stream_call(a.seed, drop, HPX_STREAM_NOISE + r, static_cast<uint32_t>((nabs + 1) >> 1));
          n0 = complex_normal_f32(wa.z, wa.w);
          n1 = complex_normal_f32(wb.x, wb.y);
        }
        y[rr * n + nl] = make_float2(fmaf(sigma, n0.x, acc0.x), fmaf(sigma, n0.y, acc0.y));
        y[rr * n + nl + 1] = make_float2(fmaf(sigma, n1.x, acc1.x), fmaf(sigma, n1.y, acc1.y));
      }
    }
    const bool csi_now = with_csi && r0 == g * kRxGroup;
    if (csi_now) {
      for (int i = tid; i < n; i += nth) y[rows * n + i] = make_float2(0.f, 0.f);
      __syncthreads();
      if (tid == 0) {
        const float2* ab = a.abar + ld * a.paths * a.words;
        for (int p = 0; p < a.paths; ++p) {
          const int d = a.delays[p];
          if (d >= 0 && d < a.keep_below) y[rows * n + d] = cadd(y[rows * n + d], ab[p * a.words + w]);
        }
      }
    }
    __syncthreads();
    lds_fft<-1>(y, rows + (csi_now ? 1 : 0), n, a.log2n, tw);
    for (int i = tid; i < rows * a.active; i += nth) {
      const int rr = i / a.active, k = i - rr * a.active;
      const int bin = subcarrier_bin(a.active_lo + k, n, a.dc);
      a.grid[((ld * a.words + w) * a.nr + r0 + rr) * static_cast<int64_t>(a.active) + k] =
          cscale(y[rr * n + bin], rx_scale);
    }
    if (csi_now) {
      for (int k = tid; k < a.active; k += nth)
        a.csi[(ld * a.words + w) * static_cast<int64_t>(a.active) + k] =
            y[rows * n + subcarrier_bin(a.active_lo + k, n, a.dc)];
    }
  }
}

// ---------------------------------------------------------------- K3: receive + detection
// One workgroup per (drop, word).  The user rows
//   row_i = sum_t B[i][t] z_t + sigma sum_r A[i][r] n_r          (B = A S; all over sqrt(N))
// are one real product on the matrix cores (v_mfma_f32_16x16x4_f32, float32 operands and
// accumulation): output rows Re row_0..7, Im row_0..7; the contraction runs over chunks of four
// inputs (antennas' noise, then users' delay-line outputs), two MFMAs per chunk and column tile
// (real, then imaginary parts of the inputs against [Re a; Im a] and [-Im a; Re a]).  A column is
// a sample pair, so the lane that supplies operand (k, column) draws exactly that antenna's two
// noise samples from one Philox call (stream HPX_STREAM_NOISE + r, the samples massive_rx_kernel
// adds) and feeds the even sample to one accumulator tile and the odd one to another: no operand
// ever moves between lanes.  The rows and the ideal-CSI row (tap means at the path delays) are
// then transformed kLinkRows at a time in LDS; each thread keeps the active subcarriers
// tid + b * blockDim (b < kLinkBins) of every row in registers and detects them after the last
// round.
__global__ void __launch_bounds__(kLinkThreads) massive_link_kernel(MassiveArgs a) {
  extern __shared__ __align__(16) unsigned char lds[];
  float2* rows = reinterpret_cast<float2*>(lds + a.k3_rows);  // [kLinkRows][rstride]
  float2* ptab = reinterpret_cast<float2*>(lds + a.k3_tab);   // FFT pass table
  float2* atab = reinterpret_cast<float2*>(lds + a.k3_atab);  // [chunk][lane] MFMA A operands
  float2* vm = reinterpret_cast<float2*>(lds + a.k3_v);       // [nt][nt] V
  float* lam = reinterpret_cast<float*>(lds + a.k3_l);
  DemapBounds* dmb = reinterpret_cast<DemapBounds*>(lds + a.k3_dmb);
  int* red = reinterpret_cast<int*>(lds + a.k3_red);
  uint32_t* bits = reinterpret_cast<uint32_t*>(lds + a.k3_rows);  // [nt][bits_words], after the transforms
  const int tid = threadIdx.x, nth = blockDim.x;
  const int64_t blk = blockIdx.x;
  const int w = static_cast<int>(blk % a.words);
  const int64_t ld = blk / a.words;
  const uint32_t drop = static_cast<uint32_t>(a.first_drop + ld);
  const int n = a.n, nt = a.nt, nr = a.nr, rs = a.k3_rstride, half = a.n >> 1;
  const int noise_chunks = (nr + 3) >> 2, chunks = noise_chunks + ((nt + 3) >> 2);
  const double np = a.noise_power[ld];
  const float rx_scale = static_cast<float>(1.0 / sqrt(static_cast<double>(n)));
  const float sigma_scale = static_cast<float>(sqrt(0.5 * np) / sqrt(static_cast<double>(n)));
  const float s2 = static_cast<float>(np);
  fill_fft_pass_table(ptab, a.log2n);
  // A operand of chunk c at lane l: row i = l % 16, input k = l / 16; entry (a1, a2) for the real
  // and the imaginary parts of the inputs, a = the row's user's coefficient of input 4 c + k.
  // Output row i = 4 g + j (D register j of lanes 16 g ..) carries slot 4 j + g = 2 user + comp,
  // so that register j holds users 2 j and 2 j + 1: the transform rounds (users 0-2, 3-5, 6-7)
  // release registers 0, then 1 and 2, then 3.
  const float2* da = a.det_a + ld * a.da_stride;
  for (int e = tid; e < chunks * 64; e += nth) {
    const int c = e >> 6, l = e & 63;
    const int i = l & 15, k = l >> 4;
    const int slot = 4 * (i & 3) + (i >> 2), user = slot >> 1, comp = slot & 1;
    float2 co = make_float2(0.f, 0.f);
    if (c < noise_chunks) {
      const int r = 4 * c + k;
      if (r < nr && user < nt) co = cscale(da[user * nr + r], sigma_scale);
    } else {
      const int t = 4 * (c - noise_chunks) + k;
      if (t < nt && user < nt) co = cscale(da[nt * nr + user * nt + t], rx_scale);
    }
    atab[e] = comp == 0 ? make_float2(co.x, -co.y) : make_float2(co.y, co.x);
  }
  for (int e = tid; e < nt * nt; e += nth) vm[e] = a.det_v[ld * nt * nt + e];
  for (int i = tid; i < nt; i += nth) lam[i] = a.det_l[ld * nt + i];
  if (tid == 0) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      dmb->a[i] = a.dbounds.a[i];
      dmb->b[i] = a.dbounds.b[i];
    }
  }
  __syncthreads();
  // column tile q of wave v: sample pairs P = (q * waves + v) * 16 + column, i.e. samples 2 P
  // (accumulator 0) and 2 P + 1 (accumulator 1) of the data window; D register j of lane l is
  // output row 4 (l / 16) + j at column l % 16: component (l / 16) % 2 of user (4 j + l / 16) / 2
  const int lane = tid & 63, wave = tid >> 6, waves = nth >> 6;
  const int kq = lane >> 4, col = lane & 15;
  const int64_t win0 = static_cast<int64_t>(w) * (n + a.cp) + a.cp;
  const bool aligned = (win0 & 1) == 0;
  const float2* zd = a.z + (ld * nt * a.words + w) * static_cast<int64_t>(n);
  floatx4 acc[kLinkPasses][2];
#pragma unroll
  for (int q = 0; q < kLinkPasses; ++q) {
    acc[q][0] = floatx4{0.f, 0.f, 0.f, 0.f};
    acc[q][1] = floatx4{0.f, 0.f, 0.f, 0.f};
    const int pbase = (q * waves + wave) * 16;
    if (pbase >= half) continue;  // whole tiles: N / 2 is a multiple of 16
    const int pair = pbase + col;
    for (int c = 0; c < noise_chunks; ++c) {
      const int r = 4 * c + kq;
      float2 n0 = make_float2(0.f, 0.f), n1 = make_float2(0.f, 0.f);
      if (r < nr) {
        // noise sample m of antenna r: words (0, 1) or (2, 3) of Philox call m / 2
        const int64_t m = win0 + 2 * pair;
        const uint4 wa = stream_call(a.seed, drop, HPX_STREAM_NOISE + r, static_cast<uint32_t>(m >> 1));
        if (aligned) {
          n0 = complex_normal_f32(wa.x, wa.y);
          n1 = complex_normal_f32(wa.z, wa.w);
        } else {
          const uint4 wb = stream_call(a.seed, drop, HPX_STREAM_NOISE + r, static_cast<uint32_t>((m + 1) >> 1));
          n0 = complex_normal_f32(wa.z, wa.w);
          n1 = complex_normal_f32(wb.x, wb.y);
        }
      }
      const float2 av = atab[c * 64 + lane];
      acc[q][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.x, n0.x, acc[q][0], 0, 0, 0);
      acc[q][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.x, n1.x, acc[q][1], 0, 0, 0);
      acc[q][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.y, n0.y, acc[q][0], 0, 0, 0);
      acc[q][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.y, n1.y, acc[q][1], 0, 0, 0);
    }
    for (int c = noise_chunks; c < chunks; ++c) {
      const int t = 4 * (c - noise_chunks) + kq;
      float4 zz = make_float4(0.f, 0.f, 0.f, 0.f);
      if (t < nt) zz = *reinterpret_cast<const float4*>(zd + static_cast<int64_t>(t) * a.words * n + 2 * pair);
      const float2 av = atab[c * 64 + lane];
      acc[q][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.x, zz.x, acc[q][0], 0, 0, 0);
      acc[q][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.x, zz.z, acc[q][1], 0, 0, 0);
      acc[q][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.y, zz.y, acc[q][0], 0, 0, 0);
      acc[q][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.y, zz.w, acc[q][1], 0, 0, 0);
    }
  }
  // transforms: rows v0 .. v0 + kLinkRows - 1 of (users 0 .. nt - 1, the CSI row) per round
  float2 u[kMaxUsers][kLinkBins];
  float2 c[kLinkBins];
#pragma unroll
  for (int b = 0; b < kLinkBins; ++b) {
    c[b] = make_float2(0.f, 0.f);
#pragma unroll
    for (int i = 0; i < kMaxUsers; ++i) u[i][b] = make_float2(0.f, 0.f);
  }
  // one round per compile-time index, so that acc, u and c stay in registers
  auto round = [&](auto qc) {
    constexpr int v0 = decltype(qc)::value * kLinkRows;
    if (v0 > nt) return;
    const int cnt = min(kLinkRows, nt + 1 - v0);
    const bool csi_here = nt < v0 + kLinkRows;
    // lane (kq, col), register j: component kq % 2 of user (4 j + kq) / 2
    {
      float* rowf = reinterpret_cast<float*>(rows) + (kq & 1);
#pragma unroll
      for (int j = v0 / 2; j <= (v0 + kLinkRows - 1) / 2 && j < 4; ++j) {
        const int user = (4 * j + kq) >> 1;
        if (user < v0 || user >= v0 + kLinkRows || user >= nt) continue;
        float* rb = rowf + 2 * (user - v0) * rs;
#pragma unroll
        for (int q = 0; q < kLinkPasses; ++q) {
          const int pbase = (q * waves + wave) * 16;
          if (pbase >= half) continue;
          const int sm = 2 * (pbase + col);
          rb[2 * sm] = acc[q][0][j];
          rb[2 * sm + 2] = acc[q][1][j];
        }
      }
    }
    if (csi_here) {
      float2* row = rows + (nt - v0) * rs;
      for (int k = tid; k < n; k += nth) row[k] = make_float2(0.f, 0.f);
      __syncthreads();
      if (tid == 0) {
        const float2* ab = a.abar + ld * a.paths * a.words;
        for (int p = 0; p < a.paths; ++p) {
          const int d = a.delays[p];
          if (d >= 0 && d < a.keep_below) row[d] = cadd(row[d], ab[p * a.words + w]);
        }
      }
    }
    __syncthreads();
    lds_fft_padded<-1>(rows, cnt, rs, a.log2n, ptab);
#pragma unroll
    for (int b = 0; b < kLinkBins; ++b) {
      const int k = tid + b * nth;
      if (k >= a.active) continue;
      const int bin = subcarrier_bin(a.active_lo + k, n, a.dc);
#pragma unroll
      for (int i = v0; i < v0 + kLinkRows && i < kMaxUsers; ++i)
        if (i < nt) u[i][b] = rows[(i - v0) * rs + bin];
      if (csi_here) c[b] = rows[(nt - v0) * rs + bin];
    }
    __syncthreads();
  };
  static_assert((kMaxUsers + 1 + kLinkRows - 1) / kLinkRows == 3, "three transform rounds");
  round(std::integral_constant<int, 0>{});
  round(std::integral_constant<int, 1>{});
  round(std::integral_constant<int, 2>{});
  // transmitted bits of word w of every user (streams HPX_STREAM_BITS + t), over the rows
  const int64_t bit0 = static_cast<int64_t>(w) * a.active * a.bps;
  const int bo = static_cast<int>(bit0 & 31);
  const int fw = static_cast<int>(bit0 >> 5);
  const int lw = static_cast<int>((bit0 + static_cast<int64_t>(a.active) * a.bps - 1) >> 5);
  const int calls = (lw >> 2) - (fw >> 2) + 1;
  const int bw = a.bits_words;
  for (int e = tid; e < nt * calls; e += nth) {
    const int t = e / calls, cc = (fw >> 2) + (e - t * calls);
    const uint4 r = stream_call(a.seed, drop, HPX_STREAM_BITS + t, static_cast<uint32_t>(cc));
    const uint32_t comp[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int wi = 4 * cc + k - fw;
      if (wi >= 0 && wi <= lw - fw) bits[t * bw + wi] = comp[k];
    }
  }
  __syncthreads();
  // per subcarrier: x = V diag(conj(c) / (|c|^2 lambda + s2)) u, decisions, bit errors; one loop
  // per PAM order, with the demapper's boundaries in registers
  int err = 0;
  auto detect = [&](auto pam_c) {
    constexpr int PAM = decltype(pam_c)::value;
    float ta[PAM / 2], tbd[PAM / 2];
#pragma unroll
    for (int k = 0; k < PAM / 2; ++k) {
      ta[k] = dmb->a[k];
      tbd[k] = dmb->b[k];
    }
#pragma unroll
    for (int b = 0; b < kLinkBins; ++b) {
      const int k = tid + b * nth;
      if (k >= a.active) continue;
      const float c2 = cabs2(c[b]);
      float2 g[kMaxUsers];
#pragma unroll
      for (int i = 0; i < kMaxUsers; ++i) {
        g[i] = make_float2(0.f, 0.f);
        if (i < nt) {
          const float den = fmaf(c2, lam[i], s2);
          const float gi = (((lam[i] > 0.0f) | (s2 > 0.0f)) & (den > 0.0f)) ? 1.0f / den : 0.0f;
          g[i] = cscale(cmulc(c[b], u[i][b]), gi);
        }
      }
      const int sbit = bo + k * a.bps;
      for (int t = 0; t < nt; ++t) {
        float2 xh = make_float2(0.f, 0.f);
#pragma unroll
        for (int i = 0; i < kMaxUsers; ++i)
          if (i < nt) xh = cfma(vm[t * nt + i], g[i], xh);
        const uint32_t rx = demap_symbol_f32_reg<PAM>(xh.x, xh.y, ta, tbd);
        const uint32_t tx = symbol_bits_at(bits + t * bw, sbit, a.bps);
        err += __popc(rx ^ tx);
        if (a.equalized) a.equalized[((ld * nt + t) * a.words + w) * static_cast<int64_t>(a.active) + k] = xh;
      }
    }
  };
  switch (a.mod.pam) {
    case 2: detect(std::integral_constant<int, 2>{}); break;
    case 4: detect(std::integral_constant<int, 4>{}); break;
    case 8: detect(std::integral_constant<int, 8>{}); break;
    default: detect(std::integral_constant<int, 16>{}); break;
  }
  for (int off = 32; off > 0; off >>= 1) err += __shfl_down(err, off, 64);
  if ((tid & 63) == 0) red[tid >> 6] = err;
  __syncthreads();
  if (tid == 0) {
    int s = 0;
    for (int i = 0; i < (nth >> 6); ++i) s += red[i];
    if (s) atomicAdd(a.errors + ld, s);
  }
}

int64_t align16(int64_t v) { return (v + 15) & ~int64_t(15); }

int make_massive(const hpx_massive_link* l, MassiveArgs* out) {
  HPX_REQUIRE(l, "hpx_massive_run: null link");
  MassiveArgs& a = *out;
  a = MassiveArgs{};
  HPX_REQUIRE(l->num_rx >= 1 && l->num_rx <= kMaxRx && l->num_users >= 1 && l->num_users <= kMaxUsers &&
                  l->num_users <= l->num_rx,
              "hpx_massive_run: needs 1 <= users <= %d, users <= antennas <= %d", kMaxUsers, kMaxRx);
  a.log2n = ilog2_exact(l->num_subcarriers);
  if (a.log2n < 6 || l->num_subcarriers > 4096) {
    set_error("hpx_massive_run: num_subcarriers must be a power of two in [64, 4096]");
    return HPX_ERR_UNSUPPORTED;
  }
  HPX_REQUIRE(l->active >= 1 && l->active <= l->num_subcarriers && l->words >= 1 && l->prefix >= 0 &&
                  l->prefix <= l->num_subcarriers && l->num_paths >= 1 && l->num_sinusoids >= 1 && l->max_delay >= 0 &&
                  l->bandwidth > 0.0 && l->device_power >= 0.0,
              "hpx_massive_run: invalid link");
  if (l->modulation_order == 2 || !modulation(l->modulation_order, 1, &a.mod)) {
    set_error("hpx_massive_run: modulation order %d not supported (4, 16, 64, 256)", l->modulation_order);
    return HPX_ERR_UNSUPPORTED;
  }
  a.nr = l->num_rx;
  a.nt = l->num_users;
  a.n = l->num_subcarriers;
  a.cp = l->prefix;
  a.words = l->words;
  a.active = l->active;
  a.active_lo = l->num_subcarriers / 2 - l->active / 2;
  a.bps = a.mod.bps;
  a.order = l->modulation_order;
  a.dc = l->dc_suppression ? 1 : 0;
  a.paths = l->num_paths;
  a.sinusoids = l->num_sinusoids;
  a.terms = l->num_sinusoids + 1;
  a.blk_cols = std::max(10, l->num_users);
  a.nnorm = std::max(10, l->num_rx) * a.blk_cols + 2 * l->num_paths + 2 * l->num_paths * l->num_sinusoids;
  a.frame_len = static_cast<int64_t>(l->words) * (l->num_subcarriers + l->prefix);
  a.bandwidth = l->bandwidth;
  a.doppler = l->doppler;
  a.los_doppler = l->los_doppler;
  a.tx_scale = sqrt(l->device_power) / sqrt(static_cast<double>(l->num_subcarriers));
  const double max_rate = std::max(fabs(l->doppler), fabs(l->los_doppler)) / l->bandwidth;
  a.log2b = taylor_log2_block(max_rate, a.terms, 11);
  if (a.log2b < 6) {
    set_error("hpx_massive_run: Doppler rate %.3g rad/sample too high for the block Taylor taps", max_rate);
    return HPX_ERR_UNSUPPORTED;
  }
  a.nb = static_cast<int>((a.frame_len + (int64_t(1) << a.log2b) - 1) >> a.log2b);
  a.maxd = l->max_delay;
  a.dx = std::max(0, l->max_delay - a.cp);
  HPX_REQUIRE(a.dx <= a.n, "hpx_massive_run: delay spread too long");
  a.cpl = std::min(a.cp, a.maxd);
  a.x0 = a.dx + a.cpl;
  a.xs = a.x0 + a.n + (a.n >> kFftPadShift);  // the word in the FFT's padded extent
  a.keep_below = static_cast<int>(std::min<int64_t>(std::min<int64_t>(1 + l->max_delay, a.frame_len), a.n));
  a.bits_words = (a.active * a.bps + 63) >> 5;
  // K0 LDS
  int64_t o = 0;
  a.k0_uni = 0;
  o += align16(static_cast<int64_t>(a.nnorm) * 8);
  a.k0_sos = static_cast<int>(o);
  o += align16(static_cast<int64_t>(a.paths) * a.terms * 3 * 8);
  a.k0_e = static_cast<int>(o);
  o += align16(static_cast<int64_t>(a.nr) * a.nt * 16);
  a.k0_f = static_cast<int>(o);
  o += align16(static_cast<int64_t>(a.nr) * a.nt * 16);
  a.k0_s64 = static_cast<int>(o);
  o += align16(static_cast<int64_t>(a.nr) * a.nt * 16);
  a.k0_det = static_cast<int>(o);
  o += align16(static_cast<int64_t>(a.nt + 2 * a.nt * a.nt + 2 * a.nt * a.nr + kEigWork) * 8);
  a.k0_lds = static_cast<int>(o);
  // K1 LDS
  o = 0;
  a.k1_x = 0;
  o += align16(static_cast<int64_t>(kTxUsers) * a.xs * 8);
  a.k1_tab = static_cast<int>(o);
  o += align16(static_cast<int64_t>(fft_pass_table_size(a.n)) * 8);
  a.k1_coef = static_cast<int>(o);
  o += align16(static_cast<int64_t>(a.paths) * a.nb * kTaylorCoeffs * 8);
  a.k1_bits = static_cast<int>(o);
  o += align16(static_cast<int64_t>(kTxUsers) * a.bits_words * 4);
  a.k1_cons = static_cast<int>(o);
  o += align16(static_cast<int64_t>(a.order) * 8);
  a.k1_dly = static_cast<int>(o);
  o += align16(static_cast<int64_t>(a.paths) * 4);
  a.k1_lds = static_cast<int>(o);
  // K2 LDS
  a.k2_y = 0;
  a.k2_tw = static_cast<int>(align16(static_cast<int64_t>(kRxRows + 1) * a.n * 8));
  a.k2_lds = static_cast<int>(a.k2_tw + align16(static_cast<int64_t>(a.n) * 8));
  // K3 LDS: the transform rows (the padded extent N + N / 32 each; later the word's bits), the
  // pass table, the MFMA A operands, V, lambda, the demapper's boundaries, the error partials
  a.da_stride = a.nt * a.nr + a.nt * a.nt;
  a.dbounds = demap_bounds(a.mod);
  a.k3_rstride = a.n + (a.n >> kFftPadShift);
  o = 0;
  a.k3_rows = 0;
  o += align16(std::max<int64_t>(static_cast<int64_t>(kLinkRows) * a.k3_rstride * 8,
                                 static_cast<int64_t>(a.nt) * a.bits_words * 4));
  a.k3_tab = static_cast<int>(o);
  o += align16(static_cast<int64_t>(fft_pass_table_size(a.n)) * 8);
  a.k3_atab = static_cast<int>(o);
  o += align16(static_cast<int64_t>((a.nr + 3) / 4 + (a.nt + 3) / 4) * 64 * 8);
  a.k3_v = static_cast<int>(o);
  o += align16(static_cast<int64_t>(a.nt) * a.nt * 8);
  a.k3_l = static_cast<int>(o);
  o += align16(static_cast<int64_t>(a.nt) * 4);
  a.k3_dmb = static_cast<int>(o);
  o += align16(sizeof(DemapBounds));
  a.k3_red = static_cast<int>(o);
  o += align16((kLinkThreads / 64) * 4);
  a.k3_lds = static_cast<int>(o);
  if (a.k0_lds > 160 * 1024 || a.k1_lds > 160 * 1024 || a.k2_lds > 160 * 1024 || a.k3_lds > 160 * 1024) {
    set_error("hpx_massive_run: the link needs more than 160 KB of LDS per workgroup");
    return HPX_ERR_UNSUPPORTED;
  }
  if (a.n > kFftMaxPerThread * kRxThreads || a.n > 2 * 16 * kLinkPasses * (kLinkThreads / 64) || a.n > kTxSamples * kTxThreads ||
      a.active > kLinkBins * kLinkThreads) {
    set_error("hpx_massive_run: %d subcarriers exceed one workgroup's FFT capacity", a.n);
    return HPX_ERR_UNSUPPORTED;
  }
  return HPX_OK;
}

bool raise_lds(const void* fn, int bytes) {
  if (bytes <= 64 * 1024) return true;
  return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) == hipSuccess;
}

}  // namespace

extern "C" int hpx_massive_sizes(const hpx_massive_link* link_host, int64_t* sizes_host) {
  HPX_REQUIRE(sizes_host, "hpx_massive_sizes: null pointer");
  MassiveArgs a;
  const int st = make_massive(link_host, &a);
  if (st != HPX_OK) return st;
  sizes_host[0] = static_cast<int64_t>(a.nr) * a.nt;                        // spatial
  sizes_host[1] = static_cast<int64_t>(a.paths) * a.nb * kTaylorCoeffs;     // coef
  sizes_host[2] = static_cast<int64_t>(a.paths) * a.words;                  // abar
  sizes_host[3] = static_cast<int64_t>(a.nt) * a.nt;                        // det_v
  sizes_host[4] = a.da_stride;                                              // det_a (A, then B)
  sizes_host[5] = a.nt;                                                     // det_l (float32)
  sizes_host[6] = static_cast<int64_t>(a.nt) * a.words * a.n;               // z
  sizes_host[7] = static_cast<int64_t>(a.words) * a.nr * a.active;          // grid
  sizes_host[8] = static_cast<int64_t>(a.words) * a.active;                 // csi
  return HPX_OK;
}

extern "C" int hpx_massive_run(const hpx_massive_link* link_host, const int32_t* delays, const double* path_gains,
                               const hpx_c128* corr_rx, const hpx_c128* corr_tx, uint64_t seed, int64_t first_drop,
                               int64_t num_drops, const double* noise_power, const hpx_massive_work* work_host,
                               int32_t* errors, hpx_c64* equalized, void* stream) {
  MassiveArgs a;
  const int st = make_massive(link_host, &a);
  if (st != HPX_OK) return st;
  HPX_REQUIRE(num_drops >= 0 && num_drops <= 65535, "hpx_massive_run: 0..65535 drops per call");
  HPX_REQUIRE(first_drop >= 0 && first_drop + num_drops <= (int64_t(1) << 32), "hpx_massive_run: drop index exceeds 32 bits");
  if (num_drops == 0) return HPX_OK;
  HPX_REQUIRE(delays && path_gains && noise_power && work_host && errors, "hpx_massive_run: null pointer");
  const hpx_massive_work& wk = *work_host;
  HPX_REQUIRE(wk.spatial && wk.coef && wk.abar && wk.det_v && wk.det_a && wk.det_l && wk.z,
              "hpx_massive_run: null work buffer");
  HPX_REQUIRE((wk.grid == nullptr) == (wk.csi == nullptr), "hpx_massive_run: grid and csi are both given or both null");
  a.delays = delays;
  a.path_gains = path_gains;
  a.corr_rx = reinterpret_cast<const double*>(corr_rx);
  a.corr_tx = reinterpret_cast<const double*>(corr_tx);
  a.noise_power = noise_power;
  a.seed = seed;
  a.first_drop = first_drop;
  a.spatial = reinterpret_cast<float2*>(wk.spatial);
  a.coef = reinterpret_cast<float2*>(wk.coef);
  a.abar = reinterpret_cast<float2*>(wk.abar);
  a.det_v = reinterpret_cast<float2*>(wk.det_v);
  a.det_a = reinterpret_cast<float2*>(wk.det_a);
  a.det_l = wk.det_l;
  a.z = reinterpret_cast<float2*>(wk.z);
  a.grid = reinterpret_cast<float2*>(wk.grid);
  a.csi = reinterpret_cast<float2*>(wk.csi);
  a.errors = errors;
  a.equalized = reinterpret_cast<float2*>(equalized);
  const hipStream_t s = as_stream(stream);
  if (!raise_lds(reinterpret_cast<const void*>(massive_prologue_kernel), a.k0_lds) ||
      !raise_lds(reinterpret_cast<const void*>(massive_tx_kernel), a.k1_lds) ||
      !raise_lds(reinterpret_cast<const void*>(massive_rx_kernel), a.k2_lds) ||
      !raise_lds(reinterpret_cast<const void*>(massive_link_kernel), a.k3_lds)) {
    set_error("hpx_massive_run: LDS limit not available");
    return HPX_ERR_UNSUPPORTED;
  }
  const unsigned nd = static_cast<unsigned>(num_drops);
  hipLaunchKernelGGL(massive_prologue_kernel, dim3(nd), dim3(256), a.k0_lds, s, a);
  const unsigned tx_groups = static_cast<unsigned>((a.nt + kTxUsers - 1) / kTxUsers);
  hipLaunchKernelGGL(massive_tx_kernel, dim3(nd * tx_groups), dim3(kTxThreads), a.k1_lds, s, a);
  if (a.grid) {  // the per-antenna grid and the CSI factor, on request
    const int64_t groups = (a.nr + kRxGroup - 1) / kRxGroup;
    a.num_pairs = static_cast<int64_t>(nd) * a.words;
    const int64_t pair_sets = (a.num_pairs + kXcds - 1) / kXcds;
    hipLaunchKernelGGL(massive_rx_kernel, dim3(static_cast<unsigned>(pair_sets * kXcds * groups)), dim3(kRxThreads),
                       a.k2_lds, s, a);
  }
  hipLaunchKernelGGL(massive_link_kernel, dim3(nd * a.words), dim3(kLinkThreads), a.k3_lds, s, a);
  return check_launch("hpx_massive_run");
}
