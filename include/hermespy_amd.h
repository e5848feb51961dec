/*
 * hermespy_amd.h -- C ABI of libhermespy_amd.so, the MI355X (gfx950) kernels of
 * the Monte Carlo link-drop path (fading channel -> OFDM -> ZF/MMSE -> bit errors).
 *
 * Conventions of every entry point:
 *  - Returns HPX_OK (0) or a negative hpx_status; hpx_last_error() gives the
 *    message of the calling thread's last failure.
 *  - Array arguments are DEVICE pointers owned by the caller, unless the
 *    parameter name ends in _host.  Complex samples are interleaved
 *    (re, im) float32 pairs, hpx_c64; float64 pairs are hpx_c128.
 *  - Launches are asynchronous on `stream` (a hipStream_t; NULL = legacy stream).
 *    No entry point allocates device memory or synchronises the stream.
 *  - Matrices are row-major; signals are [stream][sample] with the given
 *    leading dimension.
 *
 * Random numbers: Philox4x32-10 keyed by `seed` (key = seed low/high words),
 * counter = (drop, stream id, index, 0); stream ids HPX_STREAM_*.
 *
 * Reference citations are relative to the HermesPy 1.6.0 source tree.
 */
#ifndef HERMESPY_AMD_H
#define HERMESPY_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HPX_ABI_VERSION 1

typedef enum hpx_status {
  HPX_OK = 0,
  HPX_ERR_INVALID = -1,     /* bad argument: maps to Python ValueError */
  HPX_ERR_UNSUPPORTED = -2, /* valid but not implemented on this path: ValueError */
  HPX_ERR_DEVICE = -3       /* HIP runtime failure: RuntimeError */
} hpx_status;

typedef struct hpx_c64 {
  float re, im;
} hpx_c64;

typedef struct hpx_c128 {
  double re, im;
} hpx_c128;

#define HPX_STREAM_REALIZATION 0x000u
#define HPX_STREAM_BITS 0x100u  /* + transmit stream index */
#define HPX_STREAM_NOISE 0x200u /* + receive antenna index */

#define HPX_DETECT_ZF 0
#define HPX_DETECT_MMSE 1

/* ABI version (HPX_ABI_VERSION). */
int hpx_abi_version(void);

/* Copies the calling thread's last error message into buf (NUL-terminated). */
int hpx_last_error(char* buf, size_t size);

/*
 * Sum-of-sinusoids tables ("sos", float64, [num_paths][num_terms][3]): term k of
 * path l is (rate [rad/sample], phase [rad], amplitude); the sinusoid is
 * amplitude * exp(j (rate * m + phase)) at sample m.  For the reference's
 * MultipathFadingSample (hermespy/channel/fading/fading.py:293-343) a path has
 * num_sinusoids NLOS terms, rate = doppler * cos((2 pi n + angle) / N) / fs,
 * amplitude = nlos_gain * sqrt(gain * power), followed by one LOS term.
 */

/* Taps h[l][m], m in [0, num_samples), by direct evaluation (float64 phase,
 * float32 sincos).  Replaces the generator __path_impulse_generator,
 * hermespy/channel/fading/fading.py:293-343. */
int hpx_fading_taps(const double* sos, int num_paths, int num_terms, int64_t num_samples, hpx_c64* taps,
                    void* stream);

/* Time-variant tapped delay line: y[r][n] = sum_t S[r][t] sum_l h_l[n - d_l] x[t][n - d_l],
 * n in [0, num_samples + max_delay).  Taps are evaluated on the fly from a
 * per-block Taylor expansion of the sos table whose truncation error is below
 * 1e-8 of the path's rms gain for |rate| <= max_abs_rate.
 * Replaces MultipathFadingSample._propagate, hermespy/channel/fading/fading.py:371-406.
 * x: [num_tx][ld_x], y: [num_rx][ld_y] with ld_y >= num_samples + max_delay. */
int hpx_tdl_propagate(const hpx_c64* x, int num_tx, int64_t ld_x, int64_t num_samples, const double* sos,
                      int num_paths, int num_terms, double max_abs_rate, const int32_t* delays,
                      int32_t max_delay, const hpx_c64* spatial, int num_rx, hpx_c64* y, int64_t ld_y,
                      void* stream);

/* Adds complex AWGN of total power `power` (sqrt(power/2) per component) to
 * signal[r][n] for n in [0, num_samples), drawing stream HPX_STREAM_NOISE + r of `drop`.
 * Replaces AWGNRealization.add_to, hermespy/simulation/rf/noise/model.py:143-160. */
int hpx_awgn(hpx_c64* signal, int num_streams, int64_t ld, int64_t num_samples, uint64_t seed, uint32_t drop,
             double power, void* stream);

/* Data bits of transmit stream `stream_id` - HPX_STREAM_BITS of `drop` (one uint8 0/1 per
 * bit; bit k = bit k % 32 of word k / 32).  Replaces RandomBitsSource.generate_bits,
 * hermespy/modem/bits_source.py:62-63 (a counter-based stream instead of PCG64). */
int hpx_stream_bits(uint64_t seed, uint32_t drop, uint32_t stream_id, int64_t num_bits, uint8_t* bits,
                    void* stream);

/* float64 standard normals of one drop's channel realisation (Box-Muller on
 * stream HPX_STREAM_REALIZATION).  Replaces the static branch of
 * ConsistentGenerator.realize, hermespy/channel/consistent.py:431-432. */
int hpx_realization_normals(uint64_t seed, uint32_t drop, int64_t count, double* normals, void* stream);

/* Bits (one uint8 0/1 per bit) -> symbols, 3GPP Gray mapping.
 * Replaces PskQamMapping.get_symbols, hermespy/modem/tools/psk_qam_mapping.py:104-169. */
int hpx_map_symbols(const uint8_t* bits, int64_t num_symbols, int modulation_order, int is_complex,
                    hpx_c64* symbols, void* stream);

/* Hard decisions (LLR > 0 of the piecewise 3GPP LLR) -> one uint8 per bit.
 * Replaces PskQamMapping.detect_bits, hermespy/modem/tools/psk_qam_mapping.py:171-259. */
int hpx_demap_hard(const hpx_c64* symbols, int64_t num_symbols, int modulation_order, int is_complex,
                   uint8_t* bits, void* stream);

/* Host-side twins of the two functions above (CPU, no device): the float64
 * constellation table in label order and hard decisions on float64 inputs. */
int hpx_constellation_host(int modulation_order, int is_complex, hpx_c128* points_host);
int hpx_demap_hard_host(const hpx_c128* symbols_host, int64_t num_symbols, int modulation_order, int is_complex,
                        uint8_t* bits_host);

/* OFDM modulation of a [num_streams][words][num_subcarriers] grid into
 * [num_streams][ld_out] frames of words * (num_subcarriers + prefix) samples,
 * oversampling factor 1.  Replaces OFDMWaveform._forward_transformation +
 * SymbolSection.place_samples, hermespy/modem/waveforms/orthogonal/ofdm.py:97-130,
 * hermespy/modem/waveforms/orthogonal/waveform.py:520-565. */
int hpx_ofdm_modulate(const hpx_c64* grid, int num_streams, int words, int num_subcarriers, int prefix,
                      int dc_suppression, hpx_c64* out, int64_t ld_out, void* stream);

/* Inverse of hpx_ofdm_modulate (prefix removal, FFT, shift, DC roll, 1/sqrt(N)).
 * Replaces SymbolSection.pick_samples + OFDMWaveform._backward_transformation,
 * hermespy/modem/waveforms/orthogonal/waveform.py:567-614, ofdm.py:132-161. */
int hpx_ofdm_demodulate(const hpx_c64* frame, int num_streams, int64_t ld_in, int words, int num_subcarriers,
                        int prefix, int dc_suppression, hpx_c64* grid, void* stream);

/* Ideal CSI csi[r][t][w][q] = S[r][t] * sum_l mean_w(h_l) * exp(-2 pi j bin(q) d_l / N)
 * over taps d_l < min(num_taps, N), mean_w = mean over word w's N data samples.
 * Replaces OFDMIdealChannelEstimation.estimate_channel,
 * hermespy/simulation/modem/channel_estimation.py:215-295 (with fading.py:345-369). */
int hpx_ideal_csi(const double* sos, int num_paths, int num_terms, const int32_t* delays, int32_t num_taps,
                  const hpx_c64* spatial, int num_rx, int num_tx, int words, int num_subcarriers, int prefix,
                  int dc_suppression, hpx_c64* csi, void* stream);

/* Pilot least-squares channel estimation of num_streams grids of grid_size elements:
 * estimate = grid / pilot at the pilots, the barycentric combination of three pilot
 * estimates (hole_vertices index the pilot list, -1 = outside the pilots' hull) at the
 * holes, then estimate[fill_index] = estimate[fill_source].  The plan (indices,
 * weights) is the Delaunay-linear + nearest interpolation of scipy.interpolate.griddata
 * for the frame's pilot pattern.  Replaces OrthogonalLeastSquaresChannelEstimation.
 * estimate_channel, hermespy/modem/waveforms/orthogonal/channel.py:29-66 (per stream). */
int hpx_ls_estimate(const hpx_c64* grid, int num_streams, int64_t grid_size, int num_pilots,
                    const int32_t* pilot_index, const hpx_c64* pilot_symbols, int num_holes,
                    const int32_t* hole_index, const int32_t* hole_vertices, const float* hole_weights,
                    int num_fill, const int32_t* fill_index, const int32_t* fill_source, hpx_c64* estimate,
                    void* stream);

/* Per resource element equalisation: y [num_rx][K], h [num_rx][num_tx][K] -> x [num_tx][K].
 * HPX_DETECT_ZF: pseudo-inverse (num_rx == 1: y / sum_t h), replaces
 * ZeroForcingChannelEqualization.equalize_channel, hermespy/modem/waveform.py:235-250.
 * HPX_DETECT_MMSE: (H^H H + noise_variance I)^{-1} H^H y (no reference counterpart).
 * num_tx <= num_rx <= 8 for num_rx > 1. */
int hpx_equalize(const hpx_c64* y, const hpx_c64* h, int num_rx, int num_tx, int64_t num_re, int mode,
                 double noise_variance, hpx_c64* x, void* stream);

/* errors[0] += sum_i (a[i] != b[i]) over uint8 bit arrays (XOR-popcount on packed words).
 * Replaces BitErrorEvaluator.evaluate, hermespy/modem/evaluators.py:239-258. */
int hpx_count_errors(const uint8_t* a, const uint8_t* b, int64_t num_bits, unsigned long long* errors,
                     void* stream);

/*
 * Batched drops: one fused launch simulates num_drops independent drops of one
 * link (drop index first_drop + i): Philox channel realisation, bits, 3GPP
 * mapping, OFDM + CP, SoS fading with spatial response, AWGN of power
 * noise_power[i], CP removal + FFT, ideal CSI, ZF/MMSE, hard demapping and
 * the bit error count errors[i].  Replaces one pass of the stage loop
 * MonteCarloActor.__execute_stages (hermespy/core/pymonte/actors.py:316-338)
 * over SimulationRunner's stages (hermespy/simulation/simulation.py:69-181).
 */
typedef struct hpx_link {
  int32_t num_rx, num_tx;
  int32_t num_subcarriers, words, prefix; /* all-data frame, uniform cyclic prefix */
  int32_t modulation_order;               /* 4, 16, 64, 256 (complex) */
  int32_t num_paths, num_sinusoids;
  int32_t max_delay; /* max(delays) in samples; delays outside [0, max_delay] are ignored */
  int32_t detector;  /* HPX_DETECT_ZF or HPX_DETECT_MMSE */
  int32_t dc_suppression;
  double bandwidth;    /* sampling rate [Hz], oversampling factor 1 */
  double doppler;      /* NLOS doppler_frequency as used by the reference (rad/s) */
  double los_doppler;  /* LOS doppler (rad/s) */
  double gain;         /* linear channel gain */
  double device_power; /* transmit device power */
  int32_t workgroup_size; /* threads per drop: 0 = library default, else a power of two in [64, 512] */
} hpx_link;

/* Device arrays of hpx_run_batch:
 *  delays      int32 [num_paths]   rint(delay * bandwidth), sorted ascending
 *  path_gains  float64 [2][num_paths]  {los_gain * sqrt(gain * P_l)}, {nlos_gain * sqrt(gain * P_l)}
 *  corr_rx     hpx_c128 [num_rx][num_rx] or NULL (no antenna correlation)
 *  corr_tx     hpx_c128 [num_tx][num_tx] or NULL
 *  noise_power float64 [num_drops]
 *  errors      int32 [num_drops]
 *  equalized   hpx_c64 [num_drops][num_tx][words * num_subcarriers] or NULL (debug output) */
int hpx_run_batch(const hpx_link* link_host, const int32_t* delays, const double* path_gains,
                  const hpx_c128* corr_rx, const hpx_c128* corr_tx, uint64_t seed, int64_t first_drop,
                  int64_t num_drops, const double* noise_power, int32_t* errors, hpx_c64* equalized,
                  void* stream);

/* Host twin of the per-drop detector hpx_run_batch builds for ideal CSI, where every
 * resource element's channel is H_k = c_k S: S^H S = V diag(lambda) V^H (cyclic
 * Jacobi, float64) and A = V^H S^H, so that the MMSE/ZF filter of element k is
 * V diag(conj(c_k) / (|c_k|^2 lambda_i + s2)) A.  S: [num_rx][num_tx]; outputs
 * lambda [num_tx], V [num_tx][num_tx], A [num_tx][num_rx]; sizes <= 8.
 * No reference counterpart (the reference applies numpy.linalg.pinv per element,
 * hermespy/modem/waveform.py:247). */
int hpx_spatial_detector_host(int num_rx, int num_tx, const hpx_c128* spatial_host, double* lambda_host,
                              hpx_c128* v_host, hpx_c128* a_host);

/* Dynamic LDS bytes one hpx_run_batch workgroup needs for `link` (0 if unsupported). */
int64_t hpx_run_batch_lds_bytes(const hpx_link* link_host);

#ifdef __cplusplus
}
#endif

#endif /* HERMESPY_AMD_H */
