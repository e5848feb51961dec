// Fused Monte Carlo link drops: one workgroup simulates one drop end to end.
//
// Per drop (prologue):
//   Philox normals -> Phi -> spatial response S = R_rx exp(2 pi j U) R_tx and
//   the SoS table (fading.py:468-499, consistent.py:410-485), Taylor
//   coefficients of every path over the frame, per-word tap means (ideal CSI),
//   twiddles and the constellation table.
// Per OFDM word w (main loop; nothing but the error count leaves the CU):
//   bits -> 3GPP map -> IFFT -> CP                      (modem.py:797-888, ofdm.py:97-130)
//   time-variant TDL + S over the word's N data samples (fading.py:371-406)
//     + AWGN                                             (rf/noise/model.py:143-160)
//   FFT -> ideal CSI -> ZF/MMSE -> hard demap -> XOR-popcount against the bits
//     (ofdm.py:132-161, channel_estimation.py:215-295, waveform.py:235-250,
//      psk_qam_mapping.py:171-259, evaluators.py:239-258)
// Only the data windows the receiver keeps are propagated: the cyclic prefixes
// and the delay tail are discarded by the reference's demodulation
// (waveform.py:567-614) and default synchronisation (waveform.py:90-99).
#include <math.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "hpx_common.hpp"
#include "hpx_detect.hpp"
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

constexpr int kBatchMaxThreads = 512;
constexpr int kTailPerThread = 4;  // registers holding the previous word's tail
// Complex values per thread in one FFT pass (the rows a group of the transform holds is this
// times the threads over N): by default 12 where the kernel runs in 128 registers (512 threads,
// at most 4 antennas), which keeps the FFT's registers from spilling (5 rows of 1024 points
// still form one group); 16 where registers are ample.  Heterogeneous batches may take the
// alternative instead, a tuning choice per class: 16 where the default is 12 (fewer groups,
// more registers), else 8 for 8 antennas (fewer registers, more groups) and 12 otherwise.
constexpr int fft_values_for(int nt, int maxt) { return maxt >= 512 && nt <= 4 ? 12 : 16; }
constexpr int alt_fft_values_for(int nt, int maxt) {
  return fft_values_for(nt, maxt) == 12 ? 16 : (nt == 8 ? 8 : 12);
}
template <int NT, int MAXT>
constexpr int fft_values() {
  return fft_values_for(NT, MAXT);
}
// Single-link kernels of at most 256 threads keep only the spatial realisation uniforms in a table
// (heterogeneous-batch kernels always do).
constexpr bool compact_uniforms(int maxt) { return maxt <= 256; }

struct BatchArgs {
  // link
  int n, log2n, cp, words, bps, order, detector, dc;
  int paths, sinusoids, terms, nnorm;
  int log2b, nb, dx, maxd, keep_below, xs, x0, cpl, bits_words;
  int64_t frame_len;
  double bandwidth, doppler, los_doppler, tx_scale;
  Modulation mod;
  DemapBounds dbounds;  // float32 decision boundaries of mod (exact hard decisions)
  int pad_shift;        // FFT intermediate layout: kFftPadShift, or kFftNoPad where LDS is short
  // arrays
  const int32_t* delays;
  const double* path_gains;
  const double* corr_rx;  // interleaved complex, may be null
  const double* corr_tx;
  const double* noise_power;
  const double* doppler_drop;  // [drop][2] NLOS, LOS Doppler, or null (the link's)
  const uint32_t* drop_index;  // [drop] Philox drop counters, or null (first_drop + drop)
  const double* normals;       // [drop][nnorm] recorded realisation normals, or null (drawn)
  int32_t* errors;
  float2* equalized;
  // frame (null elements: every element is data)
  const int32_t* elements;
  const int32_t* word_data;
  const float2* pilots;
  int nref, data_total;
  // exports, or null
  float2* out_grid;
  float2* out_csi;
  float2* out_spatial;
  uint64_t seed;
  int64_t first_drop;
  // LDS byte offsets
  int off_x, off_coef, off_abar, off_tw, off_bits, off_cons, off_s, off_det, off_red, off_dly, off_tmp_sos, off_tmp_s64;
  int off_dmb;     // the demapper's boundaries
  int off_chain;   // per path: next path of the same delay (+1, 0: none) | 1 << 16 for the first one
  // lean drops (working set beyond one CU's LDS): Taylor table and twiddles in device memory
  float2* ws_coef;
  float2* ws_tw;    // natural twiddles (the lean ideal-CSI sum)
  float2* ws_ptab;  // FFT pass table
  int off_tmp_dir;  // prologue only: per-term word-mean weights
};

__device__ __forceinline__ bool path_in_range(const BatchArgs& a, int p) {
  const int d = a.delays[p];
  return d >= 0 && d <= a.maxd;
}

// bps bits starting at bit `bit` of an LSB-first word window (bit i of the result = i-th bit)
__device__ __forceinline__ uint32_t symbol_bits(const uint32_t* bw, int bit, int bps) {
  uint64_t v = bw[bit >> 5];
  if ((bit & 31) + bps > 32) v |= static_cast<uint64_t>(bw[(bit >> 5) + 1]) << 32;
  return static_cast<uint32_t>(v >> (bit & 31)) & ((1u << bps) - 1u);
}

// One drop, simulated by the whole workgroup.  local_drop indexes the per-drop arrays
// (noise power, Doppler, errors, exports); the Philox drop index is drop_index[local_drop]
// or first_drop + local_drop.
// COMPACT: only the 100 spatial uniforms get a table, which takes up to ~8 KB off the LDS of links
// of at most 512 subcarriers, whose prologue tables bound it.  The heterogeneous-batch kernels and
// the single-link kernels of at most 256 threads take it; the 512-thread single-link kernels keep
// the full table and their code as it is (the compact form measured 5 registers more and 2.4 %
// slower on the 4x4 1024-subcarrier link).
template <int NT, int NR, int PAIRS, int FFTV, bool LEAN = false, bool COMPACT = false>
__device__ __forceinline__ void drop_body(const BatchArgs& a, const int64_t local_drop) {
  extern __shared__ __align__(16) unsigned char lds[];
  float2* xbuf = reinterpret_cast<float2*>(lds + a.off_x);
  // a lean drop keeps the Taylor table and the twiddles in its device workspace
  float2* coef = LEAN ? a.ws_coef : reinterpret_cast<float2*>(lds + a.off_coef);
  float2* abar = reinterpret_cast<float2*>(lds + a.off_abar);
  float2* ptab = LEAN ? a.ws_ptab : reinterpret_cast<float2*>(lds + a.off_tw);  // FFT pass table
  float2* tw = a.ws_tw;  // lean drops only: natural twiddles of the per-subcarrier ideal CSI
  uint32_t* bits = reinterpret_cast<uint32_t*>(lds + a.off_bits);
  float2* cons = reinterpret_cast<float2*>(lds + a.off_cons);
  DemapBounds* dmb = reinterpret_cast<DemapBounds*>(lds + a.off_dmb);
  float2* smat = reinterpret_cast<float2*>(lds + a.off_s);
  int* red = reinterpret_cast<int*>(lds + a.off_red);
  int* dly = reinterpret_cast<int*>(lds + a.off_dly);  // path delays: LDS broadcasts in the channel loop
  int* chain = reinterpret_cast<int*>(lds + a.off_chain);  // paths sharing a delay (ideal-CSI row)
  constexpr int kRows = NT > NR ? NT : NR;
  float2* csirow = xbuf + kRows * a.xs + a.x0;  // [N] per-word CSI over the taps (not for lean drops)
  double* uni = reinterpret_cast<double*>(lds + a.off_x);        // prologue only
  double* sos = reinterpret_cast<double*>(lds + a.off_tmp_sos);  // prologue only
  double* s64 = reinterpret_cast<double*>(lds + a.off_tmp_s64);  // prologue only: S in float64
  double* amp_dir = reinterpret_cast<double*>(lds + a.off_tmp_dir);  // prologue only: [L][terms]
  float2* det_v = reinterpret_cast<float2*>(lds + a.off_det);     // [NT][NT] eigenvectors of S^H S
  float2* det_a = det_v + NT * NT;                                 // [NT][NR] V^H S^H
  float* det_l = reinterpret_cast<float*>(det_a + NT * NR);        // [NT] eigenvalues

  const int tid = threadIdx.x;
  const int nthreads = blockDim.x;
  const uint32_t drop = a.drop_index ? a.drop_index[local_drop] : static_cast<uint32_t>(a.first_drop + local_drop);
  const int n = a.n;
  const int L = a.paths;

  // ---------------------------------------------------------------- prologue
  // realisation uniforms Phi(z), z ~ N(0,1) (consistent.py:432,485), or z of a recorded
  // StaticConsistentRealization (consistent.py:244-260)
  if constexpr (COMPACT) {
    // Uniforms 0..99 (the spatial phases, read by several S entries) go to uni; every later one is
    // read once and goes straight into its field of the SoS table (fading.py:326-342,
    // angles/phases :492-499; rows [path][NLOS terms, then LOS] of {rate, phase, gain}): per
    // path, LOS angles, NLOS angles, LOS phases, NLOS phases, in the generator's variable order.
    const int ns = a.sinusoids, nlos = L * ns;
    const double inv_fs = 1.0 / a.bandwidth;
    const double dop = a.doppler_drop ? a.doppler_drop[2 * local_drop] : a.doppler;
    const double los_dop = a.doppler_drop ? a.doppler_drop[2 * local_drop + 1] : a.los_doppler;
    auto put = [&](int i, double u) {
      if (i < 100) {
        uni[i] = u;
        return;
      }
      int j = i - 100;
      if (j < L) {  // LOS angle of path j
        const double ang = 2.0 * M_PI * u;
        sos[3 * (j * a.terms + ns)] = los_dop * cos(ang) * inv_fs;
        return;
      }
      j -= L;
      if (j < nlos) {  // NLOS angle of sinusoid k of path p
        const int p = j / ns, k = j - p * ns;
        const double ang = -M_PI + 2.0 * M_PI * u;
        sos[3 * (p * a.terms + k)] = dop * cos((2.0 * M_PI * (k + 1) + ang) / ns) * inv_fs;
        return;
      }
      j -= nlos;
      if (j < L) {  // LOS phase of path j
        sos[3 * (j * a.terms + ns) + 1] = -M_PI + 2.0 * M_PI * u;
        return;
      }
      j -= L;
      const int p = j / ns, k = j - p * ns;  // NLOS phase
      sos[3 * (p * a.terms + k) + 1] = -M_PI + 2.0 * M_PI * u;
    };
    if (a.normals) {
      const double* z = a.normals + local_drop * a.nnorm;
      for (int c = tid; c < a.nnorm; c += nthreads) put(c, normal_cdf(z[c]));
    } else {
      for (int c = tid; 2 * c < a.nnorm; c += nthreads) {
        const double2 z = normal_pair_f64(stream_call(a.seed, drop, HPX_STREAM_REALIZATION, static_cast<uint32_t>(c)));
        put(2 * c, normal_cdf(z.x));
        if (2 * c + 1 < a.nnorm) put(2 * c + 1, normal_cdf(z.y));
      }
    }
    for (int e = tid; e < L * a.terms; e += nthreads) {
      const int p = e / a.terms, k = e - p * a.terms;
      sos[3 * e + 2] = a.path_gains[k < ns ? L + p : p];
    }
  } else {
    // all uniforms in uni; the SoS table is formed from them below
    if (a.normals) {
      const double* z = a.normals + local_drop * a.nnorm;
      for (int c = tid; c < a.nnorm; c += nthreads) uni[c] = normal_cdf(z[c]);
    } else {
      for (int c = tid; 2 * c < a.nnorm; c += nthreads) {
        const double2 z = normal_pair_f64(stream_call(a.seed, drop, HPX_STREAM_REALIZATION, static_cast<uint32_t>(c)));
        uni[2 * c] = normal_cdf(z.x);
        if (2 * c + 1 < a.nnorm) uni[2 * c + 1] = normal_cdf(z.y);
      }
    }
  }
  if (tid == 0) red[0] = 0;
  // a path whose delay lies outside [0, max delay] contributes nothing: it gets delay 0 and zero
  // taps (Taylor coefficients and word means), so the channel loop needs no per-path check
  for (int p = tid; p <= L; p += nthreads) dly[p] = p < L && path_in_range(a, p) ? a.delays[p] : 0;
  __syncthreads();
  // chains of paths sharing a delay, in path order: the CSI row sums their word means in the
  // order the paths are listed
  for (int p = tid; p < L; p += nthreads) {
    const int d = dly[p];
    bool first = true;
    for (int q = 0; q < p; ++q) first = first && dly[q] != d;
    int next = -1;
    for (int q = L - 1; q > p; --q)
      if (dly[q] == d) next = q;
    chain[p] = (next + 1) | (first && d < a.keep_below ? 1 << 16 : 0);
  }

  // spatial response S = R_rx exp(2 pi j U)[:NR, :NT] R_tx (fading.py:476-489)
  for (int e = tid; e < NR * NT; e += nthreads) {
    const int r = e / NT, t = e - r * NT;
    double sr = 0.0, si = 0.0;
    if (a.corr_rx) {
      for (int p = 0; p < NR; ++p)
        for (int q = 0; q < NT; ++q) {
          double s0r, s0i;
          sincospi(2.0 * uni[10 * p + q], &s0i, &s0r);
          const double ar = a.corr_rx[2 * (r * NR + p)], ai = a.corr_rx[2 * (r * NR + p) + 1];
          const double br = a.corr_tx[2 * (q * NT + t)], bi = a.corr_tx[2 * (q * NT + t) + 1];
          const double mr = ar * s0r - ai * s0i, mi = ar * s0i + ai * s0r;
          sr += mr * br - mi * bi;
          si += mr * bi + mi * br;
        }
    } else {
      sincospi(2.0 * uni[10 * r + t], &si, &sr);
    }
    smat[e] = make_float2(static_cast<float>(sr), static_cast<float>(si));
    if (a.out_spatial) a.out_spatial[local_drop * NR * NT + e] = smat[e];
    s64[2 * e] = sr;
    s64[2 * e + 1] = si;
  }
  if constexpr (COMPACT) {
    // the word means' Dirichlet weight depends on the term only, not on the word
    for (int e = tid; e < L * a.terms; e += nthreads) amp_dir[e] = sos[3 * e + 2] * dirichlet_factor(sos[3 * e], n);
  } else {
    // SoS table (fading.py:326-342, angles/phases :492-499): NLOS terms, then LOS
    const int ns = a.sinusoids;
    const int o_los_ang = 100, o_nlos_ang = 100 + L, o_los_ph = 100 + L + L * ns, o_nlos_ph = 100 + 2 * L + L * ns;
    const double inv_fs = 1.0 / a.bandwidth;
    const double dop = a.doppler_drop ? a.doppler_drop[2 * local_drop] : a.doppler;
    const double los_dop = a.doppler_drop ? a.doppler_drop[2 * local_drop + 1] : a.los_doppler;
    for (int e = tid; e < L * a.terms; e += nthreads) {
      const int p = e / a.terms, k = e - p * a.terms;
      double* row = sos + 3 * e;
      if (k < ns) {
        const double ang = -M_PI + 2.0 * M_PI * uni[o_nlos_ang + p * ns + k];
        row[0] = dop * cos((2.0 * M_PI * (k + 1) + ang) / ns) * inv_fs;
        row[1] = -M_PI + 2.0 * M_PI * uni[o_nlos_ph + p * ns + k];
        row[2] = a.path_gains[L + p];
      } else {
        const double ang = 2.0 * M_PI * uni[o_los_ang + p];
        row[0] = los_dop * cos(ang) * inv_fs;
        row[1] = -M_PI + 2.0 * M_PI * uni[o_los_ph + p];
        row[2] = a.path_gains[p];
      }
      // the word means' Dirichlet weight depends on the term only, not on the word
      amp_dir[e] = row[2] * dirichlet_factor(row[0], n);
    }
  }
  // constellation indexed by transmitted bit order (scaled by sqrt(P_tx / N)); twiddles
  for (int v = tid; v < a.order; v += nthreads) {
    const double2 p = constellation_point(a.order, 1, bits_to_label(static_cast<uint32_t>(v), a.bps));
    cons[v] = make_float2(static_cast<float>(p.x * a.tx_scale), static_cast<float>(p.y * a.tx_scale));
  }
  fill_fft_pass_table(ptab, a.log2n);
  if (tid == 0) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      dmb->a[i] = a.dbounds.a[i];
      dmb->b[i] = a.dbounds.b[i];
    }
  }
  if (LEAN) {
    for (int k = tid; k < n; k += nthreads) {
      double s, c;
      sincospi(-2.0 * static_cast<double>(k) / static_cast<double>(n), &s, &c);
      tw[k] = make_float2(static_cast<float>(c), static_cast<float>(s));
    }
  }
  __syncthreads();

  // Taylor coefficients of every (path, block) and word means of every (path, word)
  const TaylorBlocks tb{a.log2b};
  // The last wavefront computes the per-drop detector of the rank-one-in-frequency ideal-CSI
  // channel on an LDS workspace while the others compute the Taylor coefficients (one
  // wavefront does both in turn).
  const bool split = nthreads >= 128;
  const int det_first = split ? nthreads - 64 : 0;
  if (tid >= det_first) {
    double* lam = s64 + 2 * NR * NT;
    double* v = lam + NT;
    double* am = v + 2 * NT * NT;
    spatial_detector_wave(NR, NT, s64, lam, v, am, am + 2 * NT * NR);
    const int lane = tid - det_first;
    for (int i = lane; i < NT * NT; i += 64)
      det_v[i] = make_float2(static_cast<float>(v[2 * i]), static_cast<float>(v[2 * i + 1]));
    for (int i = lane; i < NT * NR; i += 64)
      det_a[i] = make_float2(static_cast<float>(am[2 * i]), static_cast<float>(am[2 * i + 1]));
    double lmax = 0.0;
    for (int i = 0; i < NT; ++i) lmax = fmax(lmax, lam[i]);
    // numpy.linalg.pinv's cut: singular values below max(M, N) eps sigma_max are dropped
    const double rt = static_cast<double>(NR > NT ? NR : NT) * 2.220446049250313e-16;
    for (int i = lane; i < NT; i += 64) det_l[i] = lam[i] > rt * rt * lmax ? static_cast<float>(lam[i]) : 0.0f;
  }
  const int taylor_threads = split ? det_first : nthreads;
  if (tid < taylor_threads) {
    for (int e = tid; e < L * a.nb; e += taylor_threads) {
      const int p = e / a.nb, b = e - p * a.nb;
      if (path_in_range(a, p)) {
        taylor_coeffs(sos + 3 * p * a.terms, a.terms, tb.centre(b), tb.half_width(), coef + e * kTaylorCoeffs);
      } else {
#pragma unroll
        for (int k = 0; k < kTaylorCoeffs; ++k) coef[e * kTaylorCoeffs + k] = make_float2(0.f, 0.f);
      }
    }
  }
  for (int e = tid; e < L * a.words; e += nthreads) {
    const int p = e / a.words, w = e - p * a.words;
    abar[e] = path_in_range(a, p) ? tap_mean_weighted(sos + 3 * p * a.terms, amp_dir + p * a.terms, a.terms,
                                                      static_cast<double>(w) * (n + a.cp) + a.cp, n)
                                  : make_float2(0.f, 0.f);
  }
  __syncthreads();
  // the signal buffer aliased the prologue tables; the frame starts after silence
  for (int e = tid; e < NT * a.dx; e += nthreads) xbuf[(e / a.dx) * a.xs + (e % a.dx)] = make_float2(0.f, 0.f);

  const double np = a.noise_power[local_drop];
  const float sigma = static_cast<float>(sqrt(0.5 * np));
  const float s2 = a.detector == HPX_DETECT_MMSE ? static_cast<float>(np) : 0.0f;
  const float rx_scale = static_cast<float>(1.0 / sqrt(static_cast<double>(n)));
  const int period = n + a.cp;
  const int x0 = a.x0;  // LDS column of the first data sample
  int errors = 0;

  // ---------------------------------------------------------------- words
  for (int w = 0; w < a.words; ++w) {
    // The thread index passes through an opaque copy per word: the per-thread index arithmetic
    // of the word's loops is redone every word instead of being hoisted out of this loop into
    // registers held across it (which made the channel loop reload spilled values).
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    // bits of this word's data symbols for every stream (stream HPX_STREAM_BITS + t):
    // LSB-first words fw .. fw + nw - 1 of the stream; the word's first bit sits at bit bo
    const int wd0 = a.word_data ? a.word_data[w] : w * n;
    const int wd1 = a.word_data ? a.word_data[w + 1] : (w + 1) * n;
    const int64_t bit0 = static_cast<int64_t>(wd0) * a.bps;
    const int bo = static_cast<int>(bit0 & 31);
    if (wd1 > wd0) {
      const int fw = static_cast<int>(bit0 >> 5);
      const int lw = static_cast<int>((static_cast<int64_t>(wd1) * a.bps - 1) >> 5);
      const int c0 = fw >> 2, c1 = lw >> 2;
      const int ncalls = c1 - c0 + 1;
      for (int e = tid; e < NT * ncalls; e += nthreads) {
        const int t = e / ncalls, c = c0 + (e - t * ncalls);
        const uint4 r = stream_call(a.seed, drop, HPX_STREAM_BITS + t, static_cast<uint32_t>(c));
        const uint32_t comp[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int wi = 4 * c + k - fw;
          if (wi >= 0 && wi <= lw - fw) bits[t * a.bits_words + wi] = comp[k];
        }
      }
    }
    __syncthreads();
    // map + place at the IFFT bins; clear the CSI row (row kRows: tap means at their delays)
    // (element map, bin and bit offset once per subcarrier for all streams)
    for (int q = tid; q < n; q += nthreads) {
      const int el = a.elements ? a.elements[w * n + q] : q;
      const int bin = subcarrier_bin(q, n, a.dc);
      const int bit = bo + el * a.bps;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        float2 v;
        if (el >= 0) {
          v = cons[symbol_bits(bits + t * a.bits_words, bit, a.bps)];
        } else if (el == HPX_ELEMENT_NULL) {
          v = make_float2(0.f, 0.f);
        } else {
          const float2 pv = a.pilots[t * a.nref + (-1 - el)];
          v = make_float2(static_cast<float>(static_cast<double>(pv.x) * a.tx_scale),
                          static_cast<float>(static_cast<double>(pv.y) * a.tx_scale));
        }
        xbuf[t * a.xs + x0 + bin] = v;
      }
      if (!LEAN) csirow[q] = make_float2(0.f, 0.f);
    }
    __syncthreads();
    lds_fft_padded<+1, FFTV>(xbuf + x0, NT, a.xs, a.log2n, ptab, a.pad_shift);
    // cyclic prefix (only the last cpl = min(CP, max delay) prefix samples are ever read by the
    // delayed paths)
    for (int e = tid; e < NT * a.cpl; e += nthreads) {
      const int t = e / a.cpl, i = e - t * a.cpl;
      xbuf[t * a.xs + x0 - a.cpl + i] = xbuf[t * a.xs + x0 + n - a.cpl + i];
    }
    // CSI row: word means of the taps at their delays, one thread per first path of a delay
    // (paths sharing a delay add up, taps beyond min(num_taps, N) are cut as fft(n=N) cuts them;
    // out-of-range paths have delay 0 and zero means)
    if (!LEAN) {
      for (int p = tid; p < L; p += nthreads) {
        int link = chain[p];
        if (link & (1 << 16)) {
          float2 c = abar[p * a.words + w];
          for (int q = (link & 0xffff) - 1; q >= 0; q = (chain[q] & 0xffff) - 1) c = cadd(c, abar[q * a.words + w]);
          csirow[dly[p]] = c;
        }
      }
    }
    __syncthreads();

    // channel over the data window + AWGN; y stays in registers
    const int64_t win0 = static_cast<int64_t>(w) * period + a.cp;  // absolute index of the first data sample
    float2 y[PAIRS][2][NR];
#pragma unroll
    for (int k = 0; k < PAIRS; ++k) {
      const int i = tid + k * nthreads;  // sample pair of the data window
      if (2 * i < n) {
        const int nl = 2 * i;
        const int64_t nabs = win0 + nl;
        float2 z0[NT], z1[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) z0[t] = z1[t] = make_float2(0.f, 0.f);
        const float step = 2.0f / static_cast<float>(1 << a.log2b);  // offset of the next sample in a block
        if (win0 >= a.maxd) {
          // no delayed index reaches before the frame start: no per-lane checks
          // (the next path's delay is read one iteration ahead, off the loop's critical path)
          int d_next = dly[0];
          for (int p = 0; p < L; ++p) {
            const int d = d_next;
            d_next = dly[p + 1];
            const int m = static_cast<int>(nabs) - d;
            const int b0 = m >> a.log2b;
            // (24-bit multiply: full rate, where a 32-bit one takes four passes)
            const float2* c0 = coef + __umul24(static_cast<unsigned>(p * a.nb + b0), kTaylorCoeffs);
            const float s0 = tb.offset32(m);
            const float2 h0 = taylor_eval(c0, s0);
            // the pair's second sample opens the next block when m ends one
            const bool wrap = ((m + 1) >> a.log2b) != b0;
            const float2 h1 = wrap ? taylor_eval(c0 + kTaylorCoeffs, tb.offset32(m + 1)) : taylor_eval(c0, s0 + step);
            const float2* xc = xbuf + x0 + nl - d;
#pragma unroll
            for (int t = 0; t < NT; ++t) {
              z0[t] = cfma(h0, xc[t * a.xs], z0[t]);
              z1[t] = cfma(h1, xc[t * a.xs + 1], z1[t]);
            }
          }
        } else {
          for (int p = 0; p < L; ++p) {
            const int d = dly[p];
            const int64_t m = nabs - d;
            const float2* xc = xbuf + x0 + nl - d;
            if (m >= 0) {
              const float2 h0 = taylor_eval(coef + (p * a.nb + static_cast<int>(tb.block_of(m))) * kTaylorCoeffs,
                                            tb.offset(m));
#pragma unroll
              for (int t = 0; t < NT; ++t) z0[t] = cfma(h0, xc[t * a.xs], z0[t]);
            }
            if (m + 1 >= 0) {
              const float2 h1 = taylor_eval(
                  coef + (p * a.nb + static_cast<int>(tb.block_of(m + 1))) * kTaylorCoeffs, tb.offset(m + 1));
#pragma unroll
              for (int t = 0; t < NT; ++t) z1[t] = cfma(h1, xc[t * a.xs + 1], z1[t]);
            }
          }
        }
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          float2 a0 = make_float2(0.f, 0.f), a1 = make_float2(0.f, 0.f);
#pragma unroll
          for (int t = 0; t < NT; ++t) {
            a0 = cfma(smat[r * NT + t], z0[t], a0);
            a1 = cfma(smat[r * NT + t], z1[t], a1);
          }
          y[k][0][r] = a0;
          y[k][1][r] = a1;
        }
        // noise sample n uses words (0,1) (n even) or (2,3) (n odd) of call n/2
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          const int64_t nabs = win0 + 2 * i;
          if ((nabs & 1) == 0) {
            const uint4 wn = stream_call(a.seed, drop, HPX_STREAM_NOISE + r, static_cast<uint32_t>(nabs >> 1));
            const float2 n0 = complex_normal_f32(wn.x, wn.y), n1 = complex_normal_f32(wn.z, wn.w);
            y[k][0][r] = make_float2(fmaf(sigma, n0.x, y[k][0][r].x), fmaf(sigma, n0.y, y[k][0][r].y));
            y[k][1][r] = make_float2(fmaf(sigma, n1.x, y[k][1][r].x), fmaf(sigma, n1.y, y[k][1][r].y));
          } else {
            const uint4 wa = stream_call(a.seed, drop, HPX_STREAM_NOISE + r, static_cast<uint32_t>(nabs >> 1));
            const uint4 wb = stream_call(a.seed, drop, HPX_STREAM_NOISE + r, static_cast<uint32_t>((nabs + 1) >> 1));
            const float2 n0 = complex_normal_f32(wa.z, wa.w), n1 = complex_normal_f32(wb.x, wb.y);
            y[k][0][r] = make_float2(fmaf(sigma, n0.x, y[k][0][r].x), fmaf(sigma, n0.y, y[k][0][r].y));
            y[k][1][r] = make_float2(fmaf(sigma, n1.x, y[k][1][r].x), fmaf(sigma, n1.y, y[k][1][r].y));
          }
        }
      }
    }
    // last dx transmitted samples of this word feed the next word's delayed paths
    float2 tail[kTailPerThread];
#pragma unroll
    for (int k = 0; k < kTailPerThread; ++k) {
      const int e = tid + k * nthreads;
      if (e < NT * a.dx) tail[k] = xbuf[(e / a.dx) * a.xs + x0 + n - a.dx + (e % a.dx)];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kTailPerThread; ++k) {
      const int e = tid + k * nthreads;
      if (e < NT * a.dx) xbuf[(e / a.dx) * a.xs + (e % a.dx)] = tail[k];
    }
#pragma unroll
    for (int k = 0; k < PAIRS; ++k) {
      const int i = tid + k * nthreads;
      if (2 * i < n) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          xbuf[r * a.xs + x0 + 2 * i] = y[k][0][r];
          xbuf[r * a.xs + x0 + 2 * i + 1] = y[k][1][r];
        }
      }
    }
    __syncthreads();
    // receive FFT of every antenna and, in the same passes, the CSI row's FFT over the taps
    lds_fft_padded<-1, FFTV>(xbuf + x0, kRows + (LEAN ? 0 : 1), a.xs, a.log2n, ptab, a.pad_shift);

    // ideal CSI, detection, demapping and bit errors per subcarrier
    for (int q = tid; q < n; q += nthreads) {
      const int bin = subcarrier_bin(q, n, a.dc);
      float2 c;
      if (LEAN) {
        // the same DFT over the taps, path by path: c = sum_p abar_p W^(bin d_p)
        c = make_float2(0.f, 0.f);
        for (int p = 0; p < L; ++p) {
          const int d = dly[p];
          if (d >= 0 && d < a.keep_below) c = cfma(abar[p * a.words + w], tw[(bin * d) & (n - 1)], c);
        }
      } else {
        c = csirow[bin];
      }
      // opaque zero offset: keeps the detector matrices as per-iteration LDS broadcast reads
      // instead of ~70 registers hoisted for the whole kernel
      int dofs = 0;
      asm volatile("" : "+v"(dofs));
      const float2* dva = det_v + dofs;
      const float2* daa = det_a + dofs;
      const float* dla = det_l + dofs;
      // W = V diag(conj(c) / (|c|^2 lambda_i + s2)) V^H S^H applied to y
      const float2* ycol = xbuf + x0 + bin;
      float2 yv[NR];
#pragma unroll
      for (int r = 0; r < NR; ++r) yv[r] = ycol[r * a.xs];
      if (a.out_grid) {
#pragma unroll
        for (int r = 0; r < NR; ++r)
          a.out_grid[((local_drop * NR + r) * a.words + w) * static_cast<int64_t>(n) + q] = cscale(yv[r], rx_scale);
      }
      if (a.out_csi) a.out_csi[(local_drop * a.words + w) * static_cast<int64_t>(n) + q] = c;
      const int el = a.elements ? a.elements[w * n + q] : q;
      if (el < 0) continue;  // reference and null elements carry no bits
      const float c2 = cabs2(c);
      float2 u[NT];
#pragma unroll
      for (int i = 0; i < NT; ++i) {
        float2 acc = make_float2(0.f, 0.f);
#pragma unroll
        for (int r = 0; r < NR; ++r) acc = cfma(daa[i * NR + r], yv[r], acc);
        const float den = fmaf(c2, dla[i], s2);
        // rx_scale folds the receiver's 1/sqrt(N); a zero gain realises pinv's dropped singular values
        const float g = (dla[i] > 0.0f || s2 > 0.0f) && den > 0.0f ? rx_scale * __builtin_amdgcn_rcpf(den) : 0.0f;
        u[i] = cscale(cmulc(c, acc), g);
      }
      float2 xh[NT];
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        float2 acc = make_float2(0.f, 0.f);
#pragma unroll
        for (int i = 0; i < NT; ++i) acc = cfma(dva[t * NT + i], u[i], acc);
        xh[t] = acc;
      }
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const uint32_t rx = demap_symbol_f32(a.mod, dmb, xh[t].x, xh[t].y);
        const uint32_t tx = symbol_bits(bits + t * a.bits_words, bo + el * a.bps, a.bps);
        errors += __popc(tx ^ rx);
        if (a.equalized) a.equalized[(local_drop * NT + t) * static_cast<int64_t>(a.data_total) + wd0 + el] = xh[t];
      }
    }
    __syncthreads();
  }

  // ---------------------------------------------------------------- epilogue
  for (int off = 32; off > 0; off >>= 1) errors += __shfl_down(errors, off, 64);
  if ((tid & 63) == 0) atomicAdd(red, errors);
  __syncthreads();
  if (tid == 0) a.errors[local_drop] = red[0];
}

// 512-thread workgroups of up to 4 antennas target 4 waves per SIMD (128 VGPRs), i.e. two
// drops per CU, which the LDS footprint of the 4x4 link allows; measured 1.6x faster than
// letting the kernel take 256 VGPRs at one drop per CU.  8x8 drops fill a CU's LDS alone.
template <int NT, int MAXT>
constexpr int batch_waves_per_eu() {
  return MAXT >= 512 ? (NT <= 4 ? 4 : 2) : 1;
}

template <int NT, int NR, int PAIRS, int MAXT>
__global__ void __launch_bounds__(MAXT) __attribute__((amdgpu_waves_per_eu(batch_waves_per_eu<NT, MAXT>())))
batch_kernel(BatchArgs a) {
  drop_body<NT, NR, PAIRS, fft_values<NT, MAXT>(), false, compact_uniforms(MAXT)>(a, blockIdx.x);
}

// Heterogeneous batches: workgroup i simulates the drop plans[i] describes; its first_drop is
// the drop's Philox index and its per-drop pointers already point at the drop's entries.
template <int NT, int NR, int PAIRS, int MAXT, bool LEAN = false, int FFTV = fft_values<NT, MAXT>()>
__global__ void __launch_bounds__(MAXT) __attribute__((amdgpu_waves_per_eu(batch_waves_per_eu<NT, MAXT>())))
batch_kernel_multi(const BatchArgs* __restrict__ plans) {
  drop_body<NT, NR, PAIRS, FFTV, LEAN, true>(plans[blockIdx.x], 0);
}

// Raises a kernel's dynamic LDS limit once per process and size (so that launches recorded
// into a graph after a first eager run make no attribute calls).
bool ensure_lds(const void* fn, int64_t bytes) {
  static std::mutex mu;
  static std::map<const void*, int64_t> granted;
  if (bytes <= 64 * 1024) return true;
  std::lock_guard<std::mutex> lock(mu);
  int64_t& g = granted[fn];
  if (g >= bytes) return true;
  if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes)) != hipSuccess)
    return false;
  g = bytes;
  return true;
}

template <int NT, int MAXT, bool LEAN, int FFTV>
int launch_multi_kernel(const BatchArgs* plans, const hpx_uneven_class& c, hipStream_t stream) {
  const void* fn = nullptr;
  switch (c.pairs) {
    case 1: fn = reinterpret_cast<const void*>(batch_kernel_multi<NT, NT, 1, MAXT, LEAN, FFTV>); break;
    case 2: fn = reinterpret_cast<const void*>(batch_kernel_multi<NT, NT, 2, MAXT, LEAN, FFTV>); break;
    case 4: fn = reinterpret_cast<const void*>(batch_kernel_multi<NT, NT, 4, MAXT, LEAN, FFTV>); break;
    default: set_error("hpx_uneven_run: unsupported samples per thread"); return HPX_ERR_UNSUPPORTED;
  }
  if (!ensure_lds(fn, c.lds_bytes)) {
    set_error("hpx_uneven_run: %lld bytes of LDS per workgroup not available", static_cast<long long>(c.lds_bytes));
    return HPX_ERR_UNSUPPORTED;
  }
  const BatchArgs* first = plans + c.first;
  const dim3 grid(static_cast<unsigned>(c.count)), block(c.threads);
  switch (c.pairs) {
    case 1: hipLaunchKernelGGL((batch_kernel_multi<NT, NT, 1, MAXT, LEAN, FFTV>), grid, block, c.lds_bytes, stream, first); break;
    case 2: hipLaunchKernelGGL((batch_kernel_multi<NT, NT, 2, MAXT, LEAN, FFTV>), grid, block, c.lds_bytes, stream, first); break;
    default: hipLaunchKernelGGL((batch_kernel_multi<NT, NT, 4, MAXT, LEAN, FFTV>), grid, block, c.lds_bytes, stream, first); break;
  }
  return check_launch("hpx_uneven_run");
}

// One class launch with the FFT values per thread `fftv` (lean classes exist only for 8x8 links,
// the one shape beyond one CU's LDS).
template <int NT, int MAXT, int FFTV>
int launch_multi_fftv(const BatchArgs* plans, const hpx_uneven_class& c, hipStream_t stream) {
  if constexpr (NT == 8) {
    if (c.lean) return launch_multi_kernel<NT, MAXT, true, FFTV>(plans, c, stream);
  }
  if (c.lean) {
    set_error("hpx_uneven_run: lean class with %d antennas", NT);
    return HPX_ERR_INVALID;
  }
  return launch_multi_kernel<NT, MAXT, false, FFTV>(plans, c, stream);
}

}  // namespace

#ifdef HPX_BATCH_FFT_ALT
// This translation unit holds the heterogeneous-batch kernels with the alternative FFT values per
// thread; batch.hip launches them through this function.
namespace hpx {
int batch_launch_multi_alt(const void* plans_dev, const hpx_uneven_class& c, int maxt, hipStream_t s) {
  const BatchArgs* p = static_cast<const BatchArgs*>(plans_dev);
  const bool big = maxt > 256;
  switch (c.num_tx) {
    case 1: return big ? launch_multi_fftv<1, 512, alt_fft_values_for(1, 512)>(p, c, s)
                       : launch_multi_fftv<1, 256, alt_fft_values_for(1, 256)>(p, c, s);
    case 2: return big ? launch_multi_fftv<2, 512, alt_fft_values_for(2, 512)>(p, c, s)
                       : launch_multi_fftv<2, 256, alt_fft_values_for(2, 256)>(p, c, s);
    case 4: return big ? launch_multi_fftv<4, 512, alt_fft_values_for(4, 512)>(p, c, s)
                       : launch_multi_fftv<4, 256, alt_fft_values_for(4, 256)>(p, c, s);
    case 8: return big ? launch_multi_fftv<8, 512, alt_fft_values_for(8, 512)>(p, c, s)
                       : launch_multi_fftv<8, 256, alt_fft_values_for(8, 256)>(p, c, s);
  }
  set_error("hpx_uneven_run: unsupported antenna count %d", c.num_tx);
  return HPX_ERR_UNSUPPORTED;
}
}  // namespace hpx
#else
namespace hpx {
int batch_launch_multi_alt(const void* plans_dev, const hpx_uneven_class& c, int maxt, hipStream_t s);
}  // namespace hpx

namespace {

int64_t align16(int64_t v) { return (v + 15) & ~int64_t(15); }

struct Plan {
  int threads = 0, pairs = 0;
  int fftv = 0;  // FFT values per thread and pass
  int64_t lds = 0;
  bool lean = false;      // Taylor table and twiddles in device memory, ideal CSI per subcarrier
  int64_t ws_elems = 0;   // complex64 workspace elements of a lean drop
  BatchArgs args{};
};

constexpr int64_t kMaxLds = 160 * 1024;

// LDS layout of a plan; a lean layout leaves out the Taylor table, the FFT pass table and the
// CSI row.  pad: rows long enough for the FFT's padded intermediate layout (N + N / 8 after the
// first data sample).
int64_t layout_lds(BatchArgs& a, int nr, int nt, bool lean, bool pad, bool compact) {
  const int rows = std::max(nr, nt) + (lean ? 0 : 1);  // signal rows (+ the CSI row)
  a.pad_shift = pad ? kFftPadShift : kFftNoPad;
  a.xs = a.x0 + a.n + (pad ? (a.n >> kFftPadShift) : 0);
  a.xs += a.xs & 1;
  int64_t o = 0;
  a.off_x = 0;
  const int64_t xbytes = static_cast<int64_t>(rows) * a.xs * sizeof(float2);
  const int64_t uni_bytes = align16(static_cast<int64_t>(compact ? std::min(a.nnorm, 100) : a.nnorm) * sizeof(double));
  const int64_t sos_bytes = align16(static_cast<int64_t>(a.paths) * a.terms * 3 * sizeof(double));
  // S (float64), then the detector's outputs and Jacobi workspace
  const int64_t s64_bytes =
      static_cast<int64_t>(2 * nr * nt + nt + 2 * nt * nt + 2 * nt * nr + kEigWork) * sizeof(double);
  const int64_t s64_aligned = align16(s64_bytes);
  const int64_t dir_bytes = align16(static_cast<int64_t>(a.paths) * a.terms * sizeof(double));
  o += align16(std::max(xbytes, uni_bytes + sos_bytes + s64_aligned + dir_bytes));
  a.off_tmp_sos = static_cast<int>(uni_bytes);
  a.off_tmp_s64 = static_cast<int>(uni_bytes + sos_bytes);
  a.off_tmp_dir = static_cast<int>(uni_bytes + sos_bytes + s64_aligned);
  a.off_coef = static_cast<int>(o);
  if (!lean) o += align16(static_cast<int64_t>(a.paths) * a.nb * kTaylorCoeffs * sizeof(float2));
  a.off_abar = static_cast<int>(o);
  o += align16(static_cast<int64_t>(a.paths) * a.words * sizeof(float2));
  a.off_tw = static_cast<int>(o);
  if (!lean) o += align16(static_cast<int64_t>(a.n) * sizeof(float2));
  a.off_bits = static_cast<int>(o);
  o += align16(static_cast<int64_t>(nt) * a.bits_words * sizeof(uint32_t));
  a.off_cons = static_cast<int>(o);
  o += align16(static_cast<int64_t>(a.order) * sizeof(float2));
  a.off_s = static_cast<int>(o);
  o += align16(static_cast<int64_t>(nr) * nt * sizeof(float2));
  a.off_det = static_cast<int>(o);
  o += align16(static_cast<int64_t>(nt * nt + nt * nr) * sizeof(float2) + nt * sizeof(float));
  a.off_red = static_cast<int>(o);
  o += 16;
  a.off_dly = static_cast<int>(o);
  o += align16(static_cast<int64_t>(a.paths + 1) * sizeof(int32_t));  // + a pad entry for the prefetch
  a.off_dmb = static_cast<int>(o);
  o += align16(sizeof(DemapBounds));
  a.off_chain = static_cast<int>(o);
  o += align16(static_cast<int64_t>(a.paths) * sizeof(int32_t));
  return o;
}

// Validates the link (and frame, if any) and computes the launch geometry and LDS layout.
// allow_lean: an 8x8 link beyond one CU's LDS gets the lean layout (heterogeneous batches,
// which provide its workspace).
// exact_taps: the caller evaluates every tap sample directly (the float64 path), so no Taylor
// block size bounds the Doppler rate.
int make_plan(const hpx_link* l, const hpx_frame* f, Plan* out, bool allow_lean = false, bool exact_taps = false) {
  HPX_REQUIRE(l, "hpx_run_batch: null link");
  if (f) {
    HPX_REQUIRE(f->elements && f->word_data, "hpx_run_batch_frames: null frame arrays");
    HPX_REQUIRE(f->num_references >= 0 && (f->num_references == 0 || f->pilots),
                "hpx_run_batch_frames: reference elements need pilot symbols");
    HPX_REQUIRE(f->num_data >= 0 && f->max_word_data >= 0 && f->max_word_data <= l->num_subcarriers,
                "hpx_run_batch_frames: invalid data element counts");
  }
  const int log2n = ilog2_exact(l->num_subcarriers);
  if (log2n < 5 || l->num_subcarriers > 4096) {
    set_error("hpx_run_batch: num_subcarriers must be a power of two in [32, 4096] (got %d)", l->num_subcarriers);
    return HPX_ERR_UNSUPPORTED;
  }
  HPX_REQUIRE(l->words >= 1 && l->prefix >= 0 && l->prefix <= l->num_subcarriers, "hpx_run_batch: invalid frame");
  HPX_REQUIRE(l->num_paths >= 1 && l->num_sinusoids >= 1 && l->max_delay >= 0,
              "hpx_run_batch: invalid path configuration");
  HPX_REQUIRE(l->bandwidth > 0.0, "hpx_run_batch: bandwidth must be positive");
  HPX_REQUIRE(l->device_power >= 0.0, "hpx_run_batch: device power must be non-negative");
  HPX_REQUIRE(l->detector == HPX_DETECT_ZF || l->detector == HPX_DETECT_MMSE, "hpx_run_batch: unknown detector");
  const bool known = l->num_tx == 1 || l->num_tx == 2 || l->num_tx == 4 || l->num_tx == 8;
  if (l->num_rx != l->num_tx || !known) {
    set_error("hpx_run_batch: antenna configuration %dx%d not instantiated (1x1, 2x2, 4x4, 8x8)", l->num_rx,
              l->num_tx);
    return HPX_ERR_UNSUPPORTED;
  }
  Modulation mod;
  if (l->modulation_order == 2 || !modulation(l->modulation_order, 1, &mod)) {
    set_error("hpx_run_batch: modulation order %d not supported (4, 16, 64, 256)", l->modulation_order);
    return HPX_ERR_UNSUPPORTED;
  }
  BatchArgs& a = out->args;
  a.n = l->num_subcarriers;
  a.log2n = log2n;
  a.cp = l->prefix;
  a.words = l->words;
  a.bps = mod.bps;
  a.order = l->modulation_order;
  a.mod = mod;
  a.dbounds = demap_bounds(mod);
  a.detector = l->detector;
  a.dc = l->dc_suppression ? 1 : 0;
  a.paths = l->num_paths;
  a.sinusoids = l->num_sinusoids;
  a.terms = l->num_sinusoids + 1;
  a.nnorm = 100 + 2 * l->num_paths + 2 * l->num_paths * l->num_sinusoids;
  a.frame_len = static_cast<int64_t>(l->words) * (l->num_subcarriers + l->prefix);
  a.bandwidth = l->bandwidth;
  a.doppler = l->doppler;
  a.los_doppler = l->los_doppler;
  a.tx_scale = sqrt(l->device_power) / sqrt(static_cast<double>(l->num_subcarriers));
  const double max_rate = std::max(fabs(l->doppler), fabs(l->los_doppler)) / l->bandwidth;
  a.log2b = taylor_log2_block(max_rate, a.terms, 11);
  if (exact_taps) a.log2b = std::max(a.log2b, 6);
  if (a.log2b < 6) {
    set_error("hpx_run_batch: Doppler rate %.3g rad/sample too high for the block Taylor taps", max_rate);
    return HPX_ERR_UNSUPPORTED;
  }
  a.nb = static_cast<int>((a.frame_len + (int64_t(1) << a.log2b) - 1) >> a.log2b);
  // one word's data bits start anywhere in an LSB-first 32-bit word
  const int max_word_data = f ? f->max_word_data : l->num_subcarriers;
  a.bits_words = (max_word_data * mod.bps + 63) >> 5;
  a.data_total = f ? f->num_data : l->words * l->num_subcarriers;
  a.nref = f ? f->num_references : 0;
  a.maxd = l->max_delay;
  a.dx = std::max(0, l->max_delay - a.cp);
  if (a.dx > a.n) {
    set_error("hpx_run_batch: delay spread of %d samples exceeds one OFDM word", l->max_delay);
    return HPX_ERR_UNSUPPORTED;
  }
  // LDS row: [previous word's last dx samples | last min(CP, D) prefix samples | N data samples
  // (| the FFT's padding)]; the row stride a.xs is set with the layout
  a.cpl = std::min(a.cp, l->max_delay);
  a.x0 = a.dx + a.cpl;  // == max delay
  a.keep_below = static_cast<int>(std::min<int64_t>(std::min<int64_t>(1 + l->max_delay, a.frame_len), a.n));

  const int nt = l->num_tx, nr = l->num_rx;
  int threads = std::min(kBatchMaxThreads, std::max(64, l->num_subcarriers / 2));
  if (l->workgroup_size != 0) {
    const int w = l->workgroup_size;
    if (w < 64 || w > kBatchMaxThreads || (w & (w - 1))) {
      set_error("hpx_run_batch: workgroup_size must be 0 or a power of two in [64, %d]", kBatchMaxThreads);
      return HPX_ERR_INVALID;
    }
    threads = w;
  }
  const int pairs = std::max(1, l->num_subcarriers / (2 * threads));
  if (pairs != 1 && pairs != 2 && pairs != 4) {
    set_error("hpx_run_batch: %d subcarriers need %d sample pairs per thread at %d threads (1, 2 or 4)",
              l->num_subcarriers, pairs, threads);
    return HPX_ERR_UNSUPPORTED;
  }
  if (l->num_subcarriers > kFftMaxPerThread * threads) {
    set_error("hpx_run_batch: %d subcarriers exceed one workgroup's FFT capacity", l->num_subcarriers);
    return HPX_ERR_UNSUPPORTED;
  }
  if (nt * a.dx > kTailPerThread * threads) {
    set_error("hpx_run_batch: inter-word delay tail too long (%d samples)", a.dx);
    return HPX_ERR_UNSUPPORTED;
  }
  out->threads = threads;
  out->pairs = pairs;
  {
    const int maxt = threads <= 256 ? 256 : 512;
    const int def = fft_values_for(nt, maxt), alt = alt_fft_values_for(nt, maxt);
    if (l->fft_values != 0 && l->fft_values != def && l->fft_values != alt) {
      set_error("hpx_run_batch: fft_values must be 0, %d or %d for %d antennas at %d threads (got %d)", def, alt, nt,
                threads, l->fft_values);
      return HPX_ERR_INVALID;
    }
    out->fftv = l->fft_values != 0 ? l->fft_values : def;
  }

  // LDS layout; the prologue tables (uniforms, SoS table) alias the signal buffer.  The FFT's
  // padded layout is taken where it keeps as many drops per CU as the plain one.
  auto per_cu = [&](int64_t bytes) {
    const int64_t by_waves = static_cast<int64_t>(threads >= 512 ? (nt <= 4 ? 4 : 2) : 4) * 4 * 64 / threads;
    return bytes > kMaxLds ? int64_t(0) : std::min<int64_t>(kMaxLds / bytes, std::max<int64_t>(1, by_waves));
  };
  auto best_layout = [&](bool lean) {
    BatchArgs padded = a;
    // the kernel these threads launch: heterogeneous batches (allow_lean) always take the compact form
    const bool compact = allow_lean || compact_uniforms(threads <= 256 ? 256 : 512);
    const int64_t op = layout_lds(padded, nr, nt, lean, true, compact);
    const int64_t on = layout_lds(a, nr, nt, lean, false, compact);
    if (per_cu(op) >= per_cu(on) && op <= kMaxLds) {
      a = padded;
      return op;
    }
    return on;
  };
  int64_t o = best_layout(false);
  if (o > kMaxLds && allow_lean && nt == 8) {
    o = best_layout(true);
    out->lean = true;
    // [Taylor table | natural twiddles | FFT pass table]
    out->ws_elems = static_cast<int64_t>(a.paths) * a.nb * kTaylorCoeffs + a.n + fft_pass_table_size(a.n);
  }
  out->lds = o;
  if (o > kMaxLds) {
    set_error("hpx_run_batch: link needs %lld bytes of LDS per drop (max 163840)", static_cast<long long>(o));
    return HPX_ERR_UNSUPPORTED;
  }
  return HPX_OK;
}

template <class K>
int launch(K kernel, const Plan& plan, int64_t num_drops, hipStream_t stream) {
  if (!ensure_lds(reinterpret_cast<const void*>(kernel), plan.lds)) {
    set_error("hpx_run_batch: %lld bytes of LDS per workgroup not available", static_cast<long long>(plan.lds));
    return HPX_ERR_UNSUPPORTED;
  }
  hipLaunchKernelGGL(kernel, dim3(static_cast<unsigned>(num_drops)), dim3(plan.threads), plan.lds, stream, plan.args);
  return check_launch("hpx_run_batch");
}

template <int NT, int MAXT>
int launch_pairs(const Plan& plan, int64_t num_drops, hipStream_t stream) {
  switch (plan.pairs) {
    case 1: return launch(batch_kernel<NT, NT, 1, MAXT>, plan, num_drops, stream);
    case 2: return launch(batch_kernel<NT, NT, 2, MAXT>, plan, num_drops, stream);
    case 4: return launch(batch_kernel<NT, NT, 4, MAXT>, plan, num_drops, stream);
  }
  set_error("hpx_run_batch: unsupported samples per thread");
  return HPX_ERR_UNSUPPORTED;
}

// A class launch: the default FFT values per thread here, the alternative through the
// translation unit that holds those kernels.
template <int NT, int MAXT>
int launch_multi(const BatchArgs* plans, const hpx_uneven_class& c, hipStream_t stream) {
  if (c.fft_values == fft_values<NT, MAXT>()) return launch_multi_fftv<NT, MAXT, fft_values<NT, MAXT>()>(plans, c, stream);
  if (c.fft_values == alt_fft_values_for(NT, MAXT)) return hpx::batch_launch_multi_alt(plans, c, MAXT, stream);
  set_error("hpx_uneven_run: %d FFT values per thread not available for %d antennas at %d threads", c.fft_values, NT,
            c.threads);
  return HPX_ERR_INVALID;
}

}  // namespace

extern "C" int64_t hpx_uneven_workspace_elements(const hpx_link* link_host) {
  if (!link_host) {
    set_error("hpx_uneven_workspace_elements: null pointer");
    return -1;
  }
  Plan plan;
  if (make_plan(link_host, nullptr, &plan, true) != HPX_OK) return -1;
  return plan.lean ? plan.ws_elems : 0;
}

extern "C" int64_t hpx_uneven_plan_bytes(void) { return static_cast<int64_t>(sizeof(BatchArgs)); }

namespace {

// Plans of a heterogeneous batch.  Lean drops take their workspace from tables_host[i].workspace,
// or, with scratch_base set, consecutive slices of scratch_base (scratch_elems complex64 values).
int uneven_prepare(const hpx_link* links_host, const hpx_drop_tables* tables_host, const uint32_t* drop_index_host,
                   int64_t num_drops, uint64_t seed, const double* noise_power, int32_t* errors, void* plans_host,
                   int64_t* order_host, hpx_uneven_class* classes_host, int max_classes, int* num_classes_host,
                   float2* scratch_base, int64_t scratch_elems) {
  HPX_REQUIRE(num_drops >= 0 && num_drops <= 0x7fffffff && max_classes >= 1, "hpx_uneven_prepare: invalid sizes");
  HPX_REQUIRE(num_classes_host, "hpx_uneven_prepare: null pointer");
  *num_classes_host = 0;
  if (num_drops == 0) return HPX_OK;
  HPX_REQUIRE(links_host && tables_host && drop_index_host && noise_power && errors && plans_host && order_host &&
                  classes_host,
              "hpx_uneven_prepare: null pointer");
  std::vector<Plan> plans(static_cast<size_t>(num_drops));
  int64_t scratch_used = 0;
  for (int64_t i = 0; i < num_drops; ++i) {
    const hpx_drop_tables& tb = tables_host[i];
    HPX_REQUIRE(tb.delays && tb.path_gains, "hpx_uneven_prepare: drop %lld has no path tables",
                static_cast<long long>(i));
    HPX_REQUIRE((tb.corr_rx == nullptr) == (tb.corr_tx == nullptr),
                "hpx_uneven_prepare: give both correlation matrices or none");
    const int st = make_plan(&links_host[i], nullptr, &plans[i], true);
    if (st != HPX_OK) return st;
    BatchArgs& a = plans[i].args;
    if (plans[i].lean) {
      float2* ws = reinterpret_cast<float2*>(tb.workspace);
      if (scratch_base) {
        if (scratch_used + plans[i].ws_elems > scratch_elems) {
          set_error("hpx_uneven_prepare_noalloc: the lean drops need more than the %lld complex64 values of the "
                    "library's scratch array",
                    static_cast<long long>(scratch_elems));
          return HPX_ERR_INVALID;
        }
        ws = scratch_base + scratch_used;
        scratch_used += plans[i].ws_elems;
      }
      HPX_REQUIRE(ws, "hpx_uneven_prepare: drop %lld needs a workspace of %lld elements", static_cast<long long>(i),
                  static_cast<long long>(plans[i].ws_elems));
      a.ws_coef = ws;
      a.ws_tw = ws + static_cast<int64_t>(a.paths) * a.nb * kTaylorCoeffs;
      a.ws_ptab = a.ws_tw + a.n;
    }
    a.delays = tb.delays;
    a.path_gains = tb.path_gains;
    a.equalized = reinterpret_cast<float2*>(tb.equalized);
    a.corr_rx = reinterpret_cast<const double*>(tb.corr_rx);
    a.corr_tx = reinterpret_cast<const double*>(tb.corr_tx);
    a.noise_power = noise_power + i;
    a.errors = errors + i;
    a.seed = seed;
    a.first_drop = drop_index_host[i];
  }
  // kernel class = (antennas, threads, samples per thread, lean, FFT values per thread), split by
  // subcarrier count so that a class holds one link shape (which a tuner times per shape; with
  // the default workgroup sizes the kernel already determines it); heavier classes and drops first
  auto key = [&](int64_t i) {
    return std::make_tuple(links_host[i].num_tx, plans[i].threads, plans[i].pairs, plans[i].lean ? 1 : 0,
                           plans[i].fftv, links_host[i].num_subcarriers);
  };
  auto cost = [&](int64_t i) {
    const hpx_link& l = links_host[i];
    return static_cast<double>(l.num_tx) * l.num_rx * l.words * (l.num_subcarriers + l.prefix) * l.num_paths;
  };
  std::vector<int64_t> idx(static_cast<size_t>(num_drops));
  for (int64_t i = 0; i < num_drops; ++i) idx[i] = i;
  std::map<std::tuple<int, int, int, int, int, int>, double> class_cost;
  for (int64_t i = 0; i < num_drops; ++i) class_cost[key(i)] = std::max(class_cost[key(i)], cost(i));
  std::stable_sort(idx.begin(), idx.end(), [&](int64_t x, int64_t y) {
    const auto kx = key(x), ky = key(y);
    if (kx != ky) {
      const double cx = class_cost[kx], cy = class_cost[ky];
      return cx != cy ? cx > cy : kx < ky;
    }
    return cost(x) > cost(y);
  });
  BatchArgs* out = static_cast<BatchArgs*>(plans_host);
  int nc = 0;
  for (int64_t k = 0; k < num_drops; ++k) {
    const int64_t i = idx[k];
    out[k] = plans[i].args;
    order_host[k] = i;
    const bool fresh = k == 0 || key(idx[k - 1]) != key(i);
    if (fresh) {
      if (nc == max_classes) {
        set_error("hpx_uneven_prepare: more than %d kernel classes", max_classes);
        return HPX_ERR_INVALID;
      }
      hpx_uneven_class& c = classes_host[nc++];
      c.first = k;
      c.count = 0;
      c.num_tx = links_host[i].num_tx;
      c.threads = plans[i].threads;
      c.pairs = plans[i].pairs;
      c.lean = plans[i].lean ? 1 : 0;
      c.lds_bytes = 0;
      c.fft_values = plans[i].fftv;
      c.reserved = 0;
    }
    hpx_uneven_class& c = classes_host[nc - 1];
    c.count += 1;
    c.lds_bytes = std::max<int64_t>(c.lds_bytes, plans[i].lds);
  }
  *num_classes_host = nc;
  return HPX_OK;
}

}  // namespace

// Scratch of hpx_uneven_prepare_noalloc: lean drops' Taylor and twiddle tables.  Each device holds
// its own copy until the process ends; host code never uses the name itself, only the address
// hipGetSymbolAddress gives for the current device.
constexpr int64_t kNoAllocScratchElems = int64_t(1) << 22;  // complex64 values: 32 MiB
static __device__ __align__(256) float scratch[2 * kNoAllocScratchElems];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};

// The current device's address of `scratch`, looked up once per device.
int scratch_address(float2** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxScratchDevices) {
    set_error("hpx_uneven_prepare_noalloc: no current HIP device");
    return HPX_ERR_DEVICE;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("hpx_uneven_prepare_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
      return HPX_ERR_DEVICE;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = static_cast<float2*>(g_scratch_ptr[dev]);
  return HPX_OK;
}

}  // namespace

extern "C" int hpx_uneven_prepare(const hpx_link* links_host, const hpx_drop_tables* tables_host,
                                  const uint32_t* drop_index_host, int64_t num_drops, uint64_t seed,
                                  const double* noise_power, int32_t* errors, void* plans_host, int64_t* order_host,
                                  hpx_uneven_class* classes_host, int max_classes, int* num_classes_host) {
  return uneven_prepare(links_host, tables_host, drop_index_host, num_drops, seed, noise_power, errors, plans_host,
                        order_host, classes_host, max_classes, num_classes_host, nullptr, 0);
}

extern "C" int64_t hpx_uneven_noalloc_capacity(void) { return kNoAllocScratchElems; }

extern "C" int hpx_uneven_prepare_noalloc(const hpx_link* links_host, const hpx_drop_tables* tables_host,
                                          const uint32_t* drop_index_host, int64_t num_drops, uint64_t seed,
                                          const double* noise_power, int32_t* errors, void* plans_host,
                                          int64_t* order_host, hpx_uneven_class* classes_host, int max_classes,
                                          int* num_classes_host) {
  int64_t need = 0;
  if (links_host && num_drops > 0) {
    for (int64_t i = 0; i < num_drops; ++i) {
      Plan p;
      if (make_plan(&links_host[i], nullptr, &p, true) == HPX_OK && p.lean) need += p.ws_elems;
    }
  }
  if (need > kNoAllocScratchElems) {
    set_error("hpx_uneven_prepare_noalloc: the lean drops need %lld complex64 values of workspace, the library's "
              "scratch array holds %lld",
              static_cast<long long>(need), static_cast<long long>(kNoAllocScratchElems));
    return HPX_ERR_INVALID;
  }
  float2* base = nullptr;
  if (need > 0) {
    const int st = scratch_address(&base);
    if (st != HPX_OK) return st;
  }
  return uneven_prepare(links_host, tables_host, drop_index_host, num_drops, seed, noise_power, errors, plans_host,
                        order_host, classes_host, max_classes, num_classes_host, base, kNoAllocScratchElems);
}

extern "C" int hpx_uneven_run(const void* plans, const hpx_uneven_class* classes_host, int num_classes,
                              void* stream) {
  HPX_REQUIRE(num_classes >= 0, "hpx_uneven_run: invalid class count");
  if (num_classes == 0) return HPX_OK;
  HPX_REQUIRE(plans && classes_host, "hpx_uneven_run: null pointer");
  const BatchArgs* p = static_cast<const BatchArgs*>(plans);
  const hipStream_t s = static_cast<hipStream_t>(stream);
  for (int k = 0; k < num_classes; ++k) {
    const hpx_uneven_class& c = classes_host[k];
    HPX_REQUIRE(c.count >= 0 && c.count <= 0x7fffffff && c.first >= 0, "hpx_uneven_run: invalid class");
    if (c.count == 0) continue;
    int st = HPX_ERR_UNSUPPORTED;
    const bool big = c.threads > 256;
    switch (c.num_tx) {
      case 1: st = big ? launch_multi<1, 512>(p, c, s) : launch_multi<1, 256>(p, c, s); break;
      case 2: st = big ? launch_multi<2, 512>(p, c, s) : launch_multi<2, 256>(p, c, s); break;
      case 4: st = big ? launch_multi<4, 512>(p, c, s) : launch_multi<4, 256>(p, c, s); break;
      case 8: st = big ? launch_multi<8, 512>(p, c, s) : launch_multi<8, 256>(p, c, s); break;
      default: set_error("hpx_uneven_run: antenna count %d not instantiated", c.num_tx);
    }
    if (st != HPX_OK) return st;
  }
  return HPX_OK;
}

extern "C" int64_t hpx_run_batch_lds_bytes(const hpx_link* link_host) {
  Plan plan;
  if (make_plan(link_host, nullptr, &plan) != HPX_OK) return 0;
  return plan.lds;
}

extern "C" int hpx_run_batch(const hpx_link* link_host, const int32_t* delays, const double* path_gains,
                             const hpx_c128* corr_rx, const hpx_c128* corr_tx, uint64_t seed, int64_t first_drop,
                             int64_t num_drops, const double* noise_power, int32_t* errors, hpx_c64* equalized,
                             void* stream) {
  hpx_batch_out out{};
  out.errors = errors;
  out.equalized = equalized;
  return hpx_run_batch_frames(link_host, nullptr, delays, path_gains, corr_rx, corr_tx, nullptr, seed, first_drop,
                              num_drops, noise_power, &out, stream);
}

extern "C" int hpx_run_batch_frames(const hpx_link* link_host, const hpx_frame* frame_host, const int32_t* delays,
                                    const double* path_gains, const hpx_c128* corr_rx, const hpx_c128* corr_tx,
                                    const double* doppler, uint64_t seed, int64_t first_drop, int64_t num_drops,
                                    const double* noise_power, const hpx_batch_out* out_host, void* stream) {
  hpx_drop_inputs in{};
  in.noise_power = noise_power;
  in.doppler = doppler;
  return hpx_run_batch_drops(link_host, frame_host, delays, path_gains, corr_rx, corr_tx, &in, seed, first_drop,
                             num_drops, out_host, stream);
}

extern "C" int hpx_run_batch_drops(const hpx_link* link_host, const hpx_frame* frame_host, const int32_t* delays,
                                   const double* path_gains, const hpx_c128* corr_rx, const hpx_c128* corr_tx,
                                   const hpx_drop_inputs* in_host, uint64_t seed, int64_t first_drop,
                                   int64_t num_drops, const hpx_batch_out* out_host, void* stream) {
  HPX_REQUIRE(out_host && in_host, "hpx_run_batch: null input or output descriptor");
  int32_t* errors = out_host->errors;
  const double* noise_power = in_host->noise_power;
  const double* doppler = in_host->doppler;
  HPX_REQUIRE(delays && path_gains && noise_power && errors, "hpx_run_batch: null pointer");
  HPX_REQUIRE((corr_rx == nullptr) == (corr_tx == nullptr), "hpx_run_batch: give both correlation matrices or none");
  HPX_REQUIRE(num_drops >= 0 && num_drops <= 0x7fffffff, "hpx_run_batch: invalid drop count");
  HPX_REQUIRE(in_host->drop_index || (first_drop >= 0 && first_drop + num_drops <= (int64_t(1) << 32)),
              "hpx_run_batch: drop index exceeds 32 bits");
  Plan plan;
  const int st = make_plan(link_host, frame_host, &plan);
  if (st != HPX_OK) return st;
  if (plan.fftv != fft_values_for(link_host->num_tx, plan.threads <= 256 ? 256 : 512)) {
    set_error("hpx_run_batch: the alternative FFT values per thread apply to heterogeneous batches only");
    return HPX_ERR_UNSUPPORTED;
  }
  if (num_drops == 0) return HPX_OK;
  BatchArgs& a = plan.args;
  a.delays = delays;
  a.path_gains = path_gains;
  a.corr_rx = reinterpret_cast<const double*>(corr_rx);
  a.corr_tx = reinterpret_cast<const double*>(corr_tx);
  a.noise_power = noise_power;
  a.errors = errors;
  a.equalized = reinterpret_cast<float2*>(out_host->equalized);
  a.doppler_drop = doppler;
  a.drop_index = in_host->drop_index;
  a.normals = in_host->normals;
  a.elements = frame_host ? frame_host->elements : nullptr;
  a.word_data = frame_host ? frame_host->word_data : nullptr;
  a.pilots = frame_host ? reinterpret_cast<const float2*>(frame_host->pilots) : nullptr;
  a.out_grid = reinterpret_cast<float2*>(out_host->grid);
  a.out_csi = reinterpret_cast<float2*>(out_host->csi);
  a.out_spatial = reinterpret_cast<float2*>(out_host->spatial);
  a.seed = seed;
  a.first_drop = first_drop;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  switch (link_host->num_tx) {
    case 1: return plan.threads <= 256 ? launch_pairs<1, 256>(plan, num_drops, s) : launch_pairs<1, 512>(plan, num_drops, s);
    case 2: return plan.threads <= 256 ? launch_pairs<2, 256>(plan, num_drops, s) : launch_pairs<2, 512>(plan, num_drops, s);
    case 4: return plan.threads <= 256 ? launch_pairs<4, 256>(plan, num_drops, s) : launch_pairs<4, 512>(plan, num_drops, s);
    case 8: return plan.threads <= 256 ? launch_pairs<8, 256>(plan, num_drops, s) : launch_pairs<8, 512>(plan, num_drops, s);
  }
  set_error("hpx_run_batch: unreachable antenna configuration");
  return HPX_ERR_UNSUPPORTED;
}

// ---------------------------------------------------------------- float64 single-antenna drops
// The float64 counterpart of the fused drop for single-antenna links of up to 256 subcarriers:
// the same counter-based streams and the same steps, every value a double, the taps evaluated
// exactly (sum of sinusoids per sample, fading.py:293-343) instead of from Taylor blocks, and
// the OFDM transforms as direct DFTs over a twiddle table.  One 256-thread workgroup per drop.
namespace {

constexpr int kF64Threads = 256;
constexpr int kF64MaxSubcarriers = 256;

struct F64Layout {
  int off_uni, off_sos, off_tw, off_sym, off_x, off_y, off_abar, off_bits, off_red;
  int dx, frame, bits_words;
  int64_t bytes;
};

__global__ void __launch_bounds__(kF64Threads) f64_drop_kernel(BatchArgs a, F64Layout g, double2* equalized) {
  extern __shared__ __align__(16) unsigned char lds[];
  double* uni = reinterpret_cast<double*>(lds + g.off_uni);
  double* sos = reinterpret_cast<double*>(lds + g.off_sos);
  double2* tw = reinterpret_cast<double2*>(lds + g.off_tw);    // [N] exp(-2 pi j k / N)
  double2* sym = reinterpret_cast<double2*>(lds + g.off_sym);  // [N] one word's symbols at their bins
  double2* x = reinterpret_cast<double2*>(lds + g.off_x);      // [dx + frame] transmitted, dx leading zeros
  double2* y = reinterpret_cast<double2*>(lds + g.off_y);      // [frame] received
  double2* abar = reinterpret_cast<double2*>(lds + g.off_abar);  // [L][words] word means of the taps
  uint32_t* bits = reinterpret_cast<uint32_t*>(lds + g.off_bits);
  int* red = reinterpret_cast<int*>(lds + g.off_red);
  const int tid = threadIdx.x;
  const int64_t local_drop = blockIdx.x;
  const uint32_t drop = static_cast<uint32_t>(a.first_drop + local_drop);
  const int n = a.n, L = a.paths, period = a.n + a.cp;

  // realisation uniforms, the data bits (stream HPX_STREAM_BITS, LSB first), twiddles
  for (int c = tid; 2 * c < a.nnorm; c += kF64Threads) {
    const double2 z = normal_pair_f64(stream_call(a.seed, drop, HPX_STREAM_REALIZATION, static_cast<uint32_t>(c)));
    uni[2 * c] = normal_cdf(z.x);
    if (2 * c + 1 < a.nnorm) uni[2 * c + 1] = normal_cdf(z.y);
  }
  for (int c = tid; 4 * c < g.bits_words; c += kF64Threads) {
    const uint4 r = stream_call(a.seed, drop, HPX_STREAM_BITS, static_cast<uint32_t>(c));
    const uint32_t comp[4] = {r.x, r.y, r.z, r.w};
    for (int k = 0; k < 4 && 4 * c + k < g.bits_words; ++k) bits[4 * c + k] = comp[k];
  }
  for (int k = tid; k < n; k += kF64Threads) {
    double s, c;
    sincospi(-2.0 * static_cast<double>(k) / static_cast<double>(n), &s, &c);
    tw[k] = make_double2(c, s);
  }
  for (int i = tid; i < g.dx; i += kF64Threads) x[i] = make_double2(0.0, 0.0);
  if (tid == 0) red[0] = 0;
  __syncthreads();

  // SoS table exactly as the fused kernel builds it (fading.py:326-342, :492-499)
  {
    const int ns = a.sinusoids;
    const int o_los_ang = 100, o_nlos_ang = 100 + L, o_los_ph = 100 + L + L * ns, o_nlos_ph = 100 + 2 * L + L * ns;
    const double inv_fs = 1.0 / a.bandwidth;
    for (int e = tid; e < L * a.terms; e += kF64Threads) {
      const int p = e / a.terms, k = e - p * a.terms;
      double* row = sos + 3 * e;
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
  // spatial response of the single antenna pair: exp(2 pi j u_00) (fading.py:476-489)
  double2 S;
  sincospi(2.0 * uni[0], &S.y, &S.x);
  __syncthreads();
  for (int e = tid; e < L * a.words; e += kF64Threads) {
    const int p = e / a.words, w = e - p * a.words;
    abar[e] = path_in_range(a, p)
                  ? tap_mean(sos + 3 * p * a.terms, a.terms, static_cast<double>(w) * period + a.cp, n)
                  : make_double2(0.0, 0.0);
  }

  // transmit: map, inverse DFT (unnormalised, the constellation carries sqrt(P / N)), prefix
  for (int w = 0; w < a.words; ++w) {
    for (int q = tid; q < n; q += kF64Threads) {
      const int bin = subcarrier_bin(q, n, a.dc);
      const uint32_t v = symbol_bits(bits, (w * n + q) * a.bps, a.bps);
      const double2 p = constellation_point(a.order, 1, bits_to_label(v, a.bps));
      sym[bin] = make_double2(p.x * a.tx_scale, p.y * a.tx_scale);
    }
    __syncthreads();
    double2* xw = x + g.dx + w * period;
    for (int k = tid; k < n; k += kF64Threads) {
      double re = 0.0, im = 0.0;
      for (int b = 0; b < n; ++b) {
        const double2 t = tw[(b * k) & (n - 1)];  // exp(+2 pi j b k / N) = conj(t)
        re += sym[b].x * t.x + sym[b].y * t.y;
        im += sym[b].y * t.x - sym[b].x * t.y;
      }
      xw[a.cp + k] = make_double2(re, im);
    }
    __syncthreads();
    for (int i = tid; i < a.cp; i += kF64Threads) xw[i] = xw[n + i];
    __syncthreads();
  }

  // channel: y_n = S sum_l h_l(n - d_l) x_(n - d_l) + sqrt(P / 2) (N + jN) over the frame
  const double np = a.noise_power[local_drop];
  const double sigma = sqrt(0.5 * np);
  for (int i = tid; i < g.frame; i += kF64Threads) {
    double zr = 0.0, zi = 0.0;
    for (int p = 0; p < L; ++p) {
      if (!path_in_range(a, p)) continue;
      const int d = a.delays[p];
      const int m = i - d;
      if (m < 0) continue;
      double hr = 0.0, hi = 0.0;
      const double* row = sos + 3 * p * a.terms;
      for (int k = 0; k < a.terms; ++k) {
        if (row[3 * k + 2] == 0.0) continue;
        double s, c;
        sincos(reduce_phase(fma(row[3 * k], static_cast<double>(m), row[3 * k + 1])), &s, &c);
        hr += row[3 * k + 2] * c;
        hi += row[3 * k + 2] * s;
      }
      const double2 xv = x[g.dx + m];
      zr += hr * xv.x - hi * xv.y;
      zi += hr * xv.y + hi * xv.x;
    }
    // noise sample i: words (0, 1) of call i / 2 for even i, (2, 3) for odd i
    const uint4 wn = stream_call(a.seed, drop, HPX_STREAM_NOISE, static_cast<uint32_t>(i >> 1));
    const uint32_t w0 = (i & 1) ? wn.z : wn.x, w1 = (i & 1) ? wn.w : wn.y;
    const double u1 = (static_cast<double>(w0 >> 8) + 0.5) * 5.9604644775390625e-08;
    const double u2 = static_cast<double>(w1 >> 8) * 5.9604644775390625e-08;
    const double rho = sqrt(-2.0 * log(u1));
    double ns, nc;
    sincospi(2.0 * u2, &ns, &nc);
    y[i] = make_double2(S.x * zr - S.y * zi + sigma * rho * nc, S.x * zi + S.y * zr + sigma * rho * ns);
  }
  __syncthreads();

  // receive: DFT of each word's data window, ideal CSI, ZF / MMSE, demap, bit errors
  const double rx_scale = 1.0 / sqrt(static_cast<double>(n));
  const double s2 = a.detector == HPX_DETECT_MMSE ? np : 0.0;
  int errors = 0;
  for (int w = 0; w < a.words; ++w) {
    const double2* yw = y + w * period + a.cp;
    for (int q = tid; q < n; q += kF64Threads) {
      const int bin = subcarrier_bin(q, n, a.dc);
      double yr = 0.0, yi = 0.0;
      for (int k = 0; k < n; ++k) {
        const double2 t = tw[(bin * k) & (n - 1)];
        yr += yw[k].x * t.x - yw[k].y * t.y;
        yi += yw[k].x * t.y + yw[k].y * t.x;
      }
      yr *= rx_scale;
      yi *= rx_scale;
      double cr = 0.0, ci = 0.0;
      for (int p = 0; p < L; ++p) {
        const int d = a.delays[p];
        if (!path_in_range(a, p) || d >= a.keep_below) continue;
        const double2 m = abar[p * a.words + w], t = tw[(bin * d) & (n - 1)];
        cr += m.x * t.x - m.y * t.y;
        ci += m.x * t.y + m.y * t.x;
      }
      const double hr = S.x * cr - S.y * ci, hi = S.x * ci + S.y * cr;
      const double den = hr * hr + hi * hi + s2;
      const double gain = den > 0.0 ? 1.0 / den : 0.0;  // pinv of a zero channel is zero
      const double xr = (hr * yr + hi * yi) * gain, xi = (hr * yi - hi * yr) * gain;
      const uint32_t rx = demap_symbol(a.mod, xr, xi);
      const uint32_t tx = symbol_bits(bits, (w * n + q) * a.bps, a.bps);
      errors += __popc(tx ^ rx);
      if (equalized) equalized[local_drop * static_cast<int64_t>(a.words) * n + w * n + q] = make_double2(xr, xi);
    }
  }
  for (int off = 32; off > 0; off >>= 1) errors += __shfl_down(errors, off, 64);
  if ((tid & 63) == 0) atomicAdd(red, errors);
  __syncthreads();
  if (tid == 0) a.errors[local_drop] = red[0];
}

}  // namespace

extern "C" int hpx_run_batch_f64(const hpx_link* link_host, const int32_t* delays, const double* path_gains,
                                 uint64_t seed, int64_t first_drop, int64_t num_drops, const double* noise_power,
                                 int32_t* errors, hpx_c128* equalized, void* stream) {
  HPX_REQUIRE(link_host && delays && path_gains && noise_power && errors, "hpx_run_batch_f64: null pointer");
  HPX_REQUIRE(num_drops >= 0 && num_drops <= 0x7fffffff && first_drop >= 0 &&
                  first_drop + num_drops <= (int64_t(1) << 32),
              "hpx_run_batch_f64: invalid drop range");
  if (link_host->num_tx != 1 || link_host->num_rx != 1) {
    set_error("hpx_run_batch_f64: single-antenna links only (got %dx%d)", link_host->num_rx, link_host->num_tx);
    return HPX_ERR_UNSUPPORTED;
  }
  if (link_host->num_subcarriers > kF64MaxSubcarriers) {
    set_error("hpx_run_batch_f64: at most %d subcarriers (got %d)", kF64MaxSubcarriers, link_host->num_subcarriers);
    return HPX_ERR_UNSUPPORTED;
  }
  Plan plan;
  const int st = make_plan(link_host, nullptr, &plan, false, true);
  if (st != HPX_OK) return st;
  if (num_drops == 0) return HPX_OK;
  BatchArgs& a = plan.args;
  a.delays = delays;
  a.path_gains = path_gains;
  a.noise_power = noise_power;
  a.errors = errors;
  a.seed = seed;
  a.first_drop = first_drop;
  F64Layout g{};
  g.dx = a.maxd;
  g.frame = a.words * (a.n + a.cp);
  g.bits_words = static_cast<int>((static_cast<int64_t>(a.words) * a.n * a.bps + 31) / 32) + 1;
  int64_t o = 0;
  auto take = [&](int64_t bytes) {
    const int at = static_cast<int>(o);
    o += align16(bytes);
    return at;
  };
  g.off_uni = take(static_cast<int64_t>(a.nnorm) * sizeof(double));
  g.off_sos = take(static_cast<int64_t>(a.paths) * a.terms * 3 * sizeof(double));
  g.off_tw = take(static_cast<int64_t>(a.n) * sizeof(double2));
  g.off_sym = take(static_cast<int64_t>(a.n) * sizeof(double2));
  g.off_x = take(static_cast<int64_t>(g.dx + g.frame) * sizeof(double2));
  g.off_y = take(static_cast<int64_t>(g.frame) * sizeof(double2));
  g.off_abar = take(static_cast<int64_t>(a.paths) * a.words * sizeof(double2));
  g.off_bits = take(static_cast<int64_t>(g.bits_words) * sizeof(uint32_t));
  g.off_red = take(16);
  g.bytes = o;
  if (g.bytes > kMaxLds) {
    set_error("hpx_run_batch_f64: %lld bytes of LDS per drop exceed a CU's %lld", static_cast<long long>(g.bytes),
              static_cast<long long>(kMaxLds));
    return HPX_ERR_UNSUPPORTED;
  }
  if (!ensure_lds(reinterpret_cast<const void*>(f64_drop_kernel), g.bytes)) {
    set_error("hpx_run_batch_f64: %lld bytes of LDS per workgroup not available", static_cast<long long>(g.bytes));
    return HPX_ERR_UNSUPPORTED;
  }
  hipLaunchKernelGGL(f64_drop_kernel, dim3(static_cast<unsigned>(num_drops)), dim3(kF64Threads),
                     static_cast<size_t>(g.bytes), static_cast<hipStream_t>(stream), a,
                     g, reinterpret_cast<double2*>(equalized));
  return check_launch("hpx_run_batch_f64");
}

#endif  // HPX_BATCH_FFT_ALT
