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

#define HPX_ABI_VERSION 8

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

/* Dense impulse response csi[r][t][n][d] = S[r][t] sum_{p: delays[p] == d} h_p(n) of a sampled
 * link, n < num_samples, d < num_taps (paths beyond num_taps - 1 are left out), with h_p the
 * sum-of-sinusoids taps of hpx_fading_taps (direct evaluation).
 * Replaces MultipathFadingSample.state, hermespy/channel/fading/fading.py:345-369. */
int hpx_fading_state(const double* sos, int num_paths, int num_terms, const int32_t* delays, int32_t num_taps,
                     const hpx_c64* spatial, int num_rx, int num_tx, int64_t num_samples, hpx_c64* csi,
                     void* stream);

/* Propagation through a dense impulse response csi [num_rx][num_tx][csi_samples][num_taps]:
 * y[r][n] = sum_{t, d} csi[r][t][n - d][d] x[t][n - d], n < num_samples + num_taps - 1.
 * Replaces ChannelStateInformation.propagate, hermespy/core/channel.py:493-528. */
int hpx_csi_propagate(const hpx_c64* csi, int num_rx, int num_tx, int64_t csi_samples, int32_t num_taps,
                      const hpx_c64* x, int64_t ld_x, int64_t num_samples, hpx_c64* y, int64_t ld_y, void* stream);

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

/* Host twin of the float32 decision path the detection kernels use for complex constellations
 * (4, 16, 64, 256): every threshold of the reference's piecewise LLRs replaced by the float32
 * boundary of the float64 amplitude it is compared in (psk_qam_mapping.py:301-497); the same
 * decisions as hpx_demap_hard_host on the float64 values of the inputs.  Both forms of the
 * decision the kernels compile (run-time and compile-time constellation order) are evaluated;
 * HPX_ERR_INVALID if they ever disagree. */
int hpx_demap_hard_f32_host(const hpx_c64* symbols_host, int64_t num_symbols, int modulation_order, uint8_t* bits_host);

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

/* Least squares at every stream's own pilots, y_p / x_p
 * (hermespy/modem/waveforms/orthogonal/channel.py:44-48, per transmit stream):
 * ls[t][d][r][j] = grid[d][r][pilot_index[t][j]] / pilot_values[t][j] for grids
 * [num_drops][num_rx][grid_elements]. */
int hpx_pilot_ls(const hpx_c64* grid, int64_t num_drops, int num_rx, int num_tx, int64_t grid_elements,
                 int num_pilots, const int32_t* pilot_index, const hpx_c64* pilot_values, hpx_c64* ls,
                 void* stream);

/* Linear MMSE interpolation of pilot LS values over the grid with one precomputed
 * filter set: est[t][rows[i]][e] = sum_j filter[t][e][j] ls[t][rows[i]][j] for the
 * `count` listed rows (row = drop * num_rx + r) of ls [num_tx][num_rows][num_pilots];
 * est is [num_tx][num_rows][elements].  The fallback estimator of the learned one;
 * the reference has no LMMSE estimator (its OFDM estimators are ideal CSI and LS,
 * channel_estimation.py:215-295, orthogonal/channel.py:25-66). */
int hpx_lmmse_apply(const hpx_c64* ls, int num_tx, int64_t num_rows, int num_pilots, const hpx_c64* filter,
                    int64_t elements, const int32_t* rows, int count, hpx_c64* est, void* stream);

/* Detection of every data element from estimated channels est [num_tx][num_drops][num_rx][words * N]
 * and received grids grid [num_drops][num_rx][words * N] (ZF or MMSE with noise_power[d]), hard
 * demapping and errors[d] = bit errors of drop first_drop + d against its Philox bits.
 * elements / word_data as in hpx_frame (both NULL: all-data frame).  equalized
 * [num_drops][num_tx][num_data] or NULL.  Replaces ChannelEqualization.equalize_channel +
 * unmap + BitErrorEvaluator.evaluate on estimated CSI (waveform.py:235-250,
 * psk_qam_mapping.py:171-259, evaluators.py:239-258). */
int hpx_detect_estimated(const hpx_c64* grid, const hpx_c64* est, int64_t num_drops, int num_rx, int num_tx,
                         int words, int num_subcarriers, const int32_t* elements, const int32_t* word_data,
                         int num_data, int modulation_order, int detector, const double* noise_power, uint64_t seed,
                         int64_t first_drop, int32_t* errors, hpx_c64* equalized, void* stream);

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
  int32_t fft_values;     /* complex values per thread and FFT pass: 0 = library default (12 at 512
                             threads and at most 4 antennas, else 16); heterogeneous batches
                             (hpx_uneven_prepare) also take the alternative, 16 where the default
                             is 12, 8 for 8 antennas and 12 otherwise; the single-link entry
                             points take only the default */
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

/* Element value of a null (unused) resource element in hpx_frame.elements. */
#define HPX_ELEMENT_NULL (-2147483647 - 1)

/* OFDM frame with reference and null elements for hpx_run_batch_frames: the
 * reference's resource mask (hermespy/modem/waveforms/orthogonal/waveform.py:1195-1211)
 * with its place/pick order, row-major over (word, subcarrier) (:361-406,1222-1290).
 * Reference k of stream t transmits pilots[t][k]; streams whose pilot is zero
 * stay silent there (orthogonal pilots for per-stream estimation; the reference
 * places the same pilots on every stream, which a zero-free table reproduces). */
typedef struct hpx_frame {
  const int32_t* elements;  /* int32 [words][N]: >= 0 data symbol index within its word,
                               -1 - k reference element k, or HPX_ELEMENT_NULL */
  const int32_t* word_data; /* int32 [words + 1]: data symbols preceding each word */
  const hpx_c64* pilots;    /* [num_tx][num_references] */
  int32_t num_references;
  int32_t num_data;       /* word_data[words] */
  int32_t max_word_data;  /* largest data symbol count of one word */
} hpx_frame;

/* Outputs of hpx_run_batch_frames (device arrays; every one but errors may be NULL). */
typedef struct hpx_batch_out {
  int32_t* errors;    /* int32 [num_drops] bit errors over the data elements */
  hpx_c64* equalized; /* [num_drops][num_tx][num_data] detected data symbols */
  hpx_c64* grid;      /* [num_drops][num_rx][words][N] received grid: CP removal, FFT, 1/sqrt(N) (ofdm.py:132-161) */
  hpx_c64* csi;       /* [num_drops][words][N] c of the ideal CSI H[r][t][w][k] = S[r][t] c[w][k]
                         (channel_estimation.py:215-295) */
  hpx_c64* spatial;   /* [num_drops][num_rx][num_tx] spatial response S */
} hpx_batch_out;

/* hpx_run_batch over a frame with reference/null elements (NULL frame: all data),
 * with per-drop Doppler and optional exports of the received grid and the ideal
 * CSI (the inputs of pilot-based estimators, orthogonal/channel.py:29-66).
 * doppler: float64 [num_drops][2] {NLOS, LOS} doppler_frequency per drop, or NULL
 * for the link's; |values| must not exceed the link's doppler / los_doppler, which
 * size the Taylor blocks of the taps. */
int hpx_run_batch_frames(const hpx_link* link_host, const hpx_frame* frame_host, const int32_t* delays,
                         const double* path_gains, const hpx_c128* corr_rx, const hpx_c128* corr_tx,
                         const double* doppler, uint64_t seed, int64_t first_drop, int64_t num_drops,
                         const double* noise_power, const hpx_batch_out* out_host, void* stream);

/* Per-drop inputs of hpx_run_batch_drops (device arrays). */
typedef struct hpx_drop_inputs {
  const double* noise_power;  /* float64 [num_drops] (required) */
  const double* doppler;      /* float64 [num_drops][2] {NLOS, LOS}, or NULL (as hpx_run_batch_frames) */
  const uint32_t* drop_index; /* [num_drops] Philox drop counters of bits, noise and realisation,
                                 or NULL for first_drop + i */
  const double* normals;      /* float64 [num_drops][100 + 2 L + 2 L num_sinusoids]: the realisation's
                                 standard normals z, i.e. a recorded StaticConsistentRealization's
                                 scalar_samples (hermespy/channel/consistent.py:244-260) in the
                                 generator's variable order (fading.py:468-515); NULL draws them */
} hpx_drop_inputs;

/* hpx_run_batch_frames with drops named by an index array and, optionally, channel
 * realisations replayed from recorded normals.  The drop loop over grid sections of
 * MonteCarloActor.__run_batch (hermespy/core/pymonte/actors.py:316-452) with the
 * section order of MonteCarloQueueManager.next_batch (actors.py:77-108), and the
 * realisation replay of Scenario.replay / drop (hermespy/core/scenario.py:441-506,817-833). */
int hpx_run_batch_drops(const hpx_link* link_host, const hpx_frame* frame_host, const int32_t* delays,
                        const double* path_gains, const hpx_c128* corr_rx, const hpx_c128* corr_tx,
                        const hpx_drop_inputs* in_host, uint64_t seed, int64_t first_drop, int64_t num_drops,
                        const hpx_batch_out* out_host, void* stream);

/* Realisation normals of many drops: normals[i][c] = hpx_realization_normals(seed,
 * drop_index ? drop_index[i] : first_drop + i)[c], c < count (device arrays). */
int hpx_realization_normals_batch(uint64_t seed, const uint32_t* drop_index, int64_t first_drop, int64_t num_drops,
                                  int64_t count, double* normals, void* stream);

/* ---------------------------------------------------------------- Monte Carlo sections
 * Per-section statistics of the per-drop bit error rate e_i / b_i, the artifact that
 * BitErrorEvaluation.artifact hands to ScalarEvaluationResult.add_artifact
 * (hermespy/modem/evaluators.py:231-233, hermespy/core/pymonte/scalar.py:100-125):
 * the running sum, squared sum and count behind the mean and the Bayer et al.
 * stopping rule, plus the pooled error and bit counts. */
typedef struct hpx_section_stats {
  double* sum;     /* [num_sections] sum of e_i / b_i */
  double* sum_sq;  /* [num_sections] sum of (e_i / b_i)^2 */
  int64_t* count;  /* [num_sections] drops */
  int64_t* errors; /* [num_sections] sum of e_i */
  int64_t* bits;   /* [num_sections] sum of b_i */
} hpx_section_stats;

/* Adds num_drops drops to the statistics (device arrays, accumulated in place; clear them
 * to start).  Drop i belongs to section[i], or to (first_drop + i) % num_sections when
 * section is NULL; drops whose section lies outside [0, num_sections) are skipped.  b_i is
 * bits_per_drop[i], or bits when bits_per_drop is NULL.  One workgroup per section sums its
 * drops in a fixed order, so results do not depend on timing. */
int hpx_section_accumulate(const int32_t* errors, const int32_t* section, int64_t first_drop,
                           const int32_t* bits_per_drop, int64_t bits, int64_t num_drops, int32_t num_sections,
                           const hpx_section_stats* stats_host, void* stream);

/* ---------------------------------------------------------------- learned estimator
 * A 1-D convolutional network over the 64 pilot positions of each (transmit,
 * receive) stream pair, replacing the interpolation of
 * OrthogonalLeastSquaresChannelEstimation (orthogonal/channel.py:29-66); no reference
 * counterpart otherwise.  Activations are bfloat16 (uint16 bit patterns)
 * [samples][64][channels], sample = (t * num_drops + d) * num_rx + r; weights
 * bfloat16 [cout][taps][cin] (odd taps): 5 for the four layers (dilations 1, 4, 16, 1),
 * 17 for the linear skip convolution from the input (dilation 1).  The 64 positions
 * are cyclic (a stream's comb pilots on a grid without DC roll sample the periodic
 * frequency response uniformly), so every convolution is circular. */

/* out[s][p][co] = bias[co] + sum_{k, ci} weight[co][k][ci] in[s][(p + (k - taps/2) dilation) mod 64][ci]
 * on the MFMA units (float32 accumulation).  mode 0: out_f32 = that, out_bf16 = bf16(that); mode 1
 * (GELU, tanh form): out_f32 = the pre-activation, out_bf16 = bf16(gelu(pre)); mode 2: out_f32 +=
 * that (a skip path).  bias may be NULL; either output may be NULL.  (cin, taps) in
 * {(8|64|224, 5), (8, 17)}, cout a multiple of 32. */
int hpx_conv1d_forward(const uint16_t* in, int64_t num_samples, int cin, int taps, int dilation,
                       const uint16_t* weight, const float* bias, int cout, int mode, float* out_f32,
                       uint16_t* out_bf16, void* stream);

/* float32 workspace bytes of hpx_conv1d_wgrad. */
int64_t hpx_conv1d_wgrad_workspace(int64_t num_samples, int cin, int cout, int taps);

/* grad_weight[co][k][ci] = sum_{s, p} dz[s][p][co] in[s][(p + (k - taps/2) dilation) mod 64][ci],
 * grad_bias[co] = sum dz (float32). */
int hpx_conv1d_wgrad(const uint16_t* dz, const uint16_t* in, int64_t num_samples, int cin, int cout, int taps,
                     int dilation, float* workspace, float* grad_weight, float* grad_bias, void* stream);

/* dz = bf16(grad * gelu'(pre)) elementwise. */
int hpx_gelu_backward(const float* grad, const float* pre, int64_t n, uint16_t* dz, void* stream);

/* y = bf16(x), round to nearest even. */
int hpx_to_bf16(const float* x, int64_t n, uint16_t* y, void* stream);

/* Network input from pilot LS values ls [num_tx][num_drops][num_rx][pilot_words * 64] (hpx_pilot_ls):
 * x[s][p][2 v + {0, 1}] = {re, im} ls[..][v * 64 + p], x[s][p][2 pilot_words] = -1 / +1 for stream
 * 0 / others, x[s][p][2 pilot_words + 1] = -log10(noise_power[d]) / 3 (0 if noise_power is NULL),
 * zero up to cin channels. */
int hpx_net_input(const hpx_c64* ls, int64_t num_drops, int num_rx, int num_tx, int pilot_words, int cin,
                  const double* noise_power, uint16_t* x, void* stream);

/* Channel estimates est [num_samples][words * N] from the network output out [num_samples][64][2 words B]
 * (B = N / 64 subcarriers per position): est[s][w][B p + u] = out[s][p][2 (w B + u) + {0, 1}]. */
int hpx_net_output(const float* out, int64_t num_samples, int words, int num_subcarriers, hpx_c64* est,
                   void* stream);

/* Per-drop squared estimation error and channel power against the ideal CSI H[r][t] = S[r][t] c:
 * err[d] += sum |est - H|^2, power[d] += sum |H|^2 (float64; zero them first).  est as in
 * hpx_lmmse_apply, csi / spatial as exported by hpx_run_batch_frames. */
int hpx_estimate_error(const hpx_c64* est, const hpx_c64* csi, const hpx_c64* spatial, int64_t num_drops,
                       int num_rx, int num_tx, int64_t grid_elements, double* err, double* power, void* stream);

/* Training loss of network outputs against the ideal CSI: the mean over drop groups (SNR points;
 * group[d] in [0, num_groups), or NULL for one group) of 10 log10(sum |out - H|^2 / sum |H|^2),
 * into loss[0], and its gradient grad (same layout as out).  work: float64 [3 num_groups], the first
 * 2 num_groups zero on entry. */
int hpx_nmse_grad(const float* out, const hpx_c64* csi, const hpx_c64* spatial, int64_t num_drops, int num_rx,
                  int num_tx, int words, int num_subcarriers, const int32_t* group, int num_groups, double* work,
                  float* grad, double* loss, void* stream);

/* The learned estimator end to end in one launch: LS at every stream's pilots of received grids
 * grid [num_drops][num_rx][words * N] (pilot_index / pilot_values [num_tx][num_pilots] as in
 * hpx_pilot_ls, num_pilots = 64 x pilot words), the four layers and the skip convolution with the
 * activations in LDS, and the estimates est [num_tx][num_drops][num_rx][words * N].  weights /
 * biases: host arrays of 5 device pointers (layers 1-4, then the skip convolution), bfloat16
 * [cout][taps][cin] with (cin, taps) = (8, 5), (64, 5), (64, 5), (64, 5), (8, 17) and float32 [cout];
 * cout = 64 for the hidden layers and 2 words N / 64 for the last two; noise_power [num_drops] feeds
 * the SNR input channel.  Same arithmetic as hpx_net_input + hpx_conv1d_forward (x4, then mode 2)
 * + hpx_net_output. */
int hpx_learned_estimate(const hpx_c64* grid, int64_t num_drops, int num_rx, int num_tx, int words,
                         int num_subcarriers, const int32_t* pilot_index, const hpx_c64* pilot_values, int num_pilots,
                         const double* noise_power, const uint16_t* const* weights, const float* const* biases,
                         int cout, hpx_c64* est, void* stream);

/* One Adam step (bias-corrected, step >= 1) on float32 parameters. */
int hpx_adam_step(float* param, const float* grad, float* m1, float* m2, int64_t n, float lr, float beta1, float beta2,
                  float eps, int64_t step, void* stream);

/* bfloat16 copies of float32 weights [cout][taps][cin]: weight (same layout) and weight_transposed
 * [cin][taps][cout] with weight_transposed[ci][k][co] = w[co][taps - 1 - k][ci] (the backward-data
 * convolution's weights); either may be NULL. */
int hpx_conv1d_weight_copies(const float* param, int cout, int cin, int taps, uint16_t* weight,
                             uint16_t* weight_transposed, void* stream);

/* ---------------------------------------------------------------- heterogeneous batches
 * Thousands of drops with different antenna counts, FFT sizes, path counts, Doppler,
 * modulation and SNR in one submission: the host groups the drops by kernel class
 * (antennas, threads, samples per thread) into per-drop plans, the device runs one
 * launch per class (capturable into one graph).  Replaces the reference's drop loop
 * over independently configured simulations (MonteCarloActor.__execute_stages,
 * hermespy/core/pymonte/actors.py:316-452), which runs one shape at a time. */
typedef struct hpx_drop_tables {
  const int32_t* delays;     /* device arrays of hpx_run_batch for this drop's link */
  const double* path_gains;
  const hpx_c128* corr_rx;   /* or NULL */
  const hpx_c128* corr_tx;
  hpx_c64* workspace;        /* hpx_uneven_workspace_elements(link) elements of device memory for a
                                drop whose link exceeds the LDS of one CU, else NULL */
  hpx_c64* equalized;        /* [num_tx][words * N] the drop's detected data symbols, or NULL */
} hpx_drop_tables;

typedef struct hpx_uneven_class {
  int64_t first, count;      /* contiguous plan slots of the class */
  int32_t num_tx, threads, pairs, lean; /* lean: drops that keep the Taylor table and the
                                            twiddles in their workspace and compute the ideal
                                            CSI per subcarrier (8x8 at 2048 subcarriers) */
  int64_t lds_bytes;         /* dynamic LDS per workgroup (max over the class) */
  int32_t fft_values;        /* complex values per thread and FFT pass of the class's kernel */
  int32_t reserved;          /* 0 */
} hpx_uneven_class;

/* Complex64 elements of device workspace a drop of `link` needs in a heterogeneous batch: 0 when
 * its working set fits one CU's LDS, the Taylor table plus the twiddle table when it fits only
 * without them (and without the ideal-CSI row), -1 (with hpx_last_error) when unsupported. */
int64_t hpx_uneven_workspace_elements(const hpx_link* link_host);

/* Bytes of one drop plan. */
int64_t hpx_uneven_plan_bytes(void);

/* Builds the plans on the host (no device work): drop i has link links_host[i], tables
 * tables_host[i], Philox drop index drop_index_host[i], noise power noise_power[i] and error
 * count errors[i] (device arrays [num_drops]).  plans_host (num_drops plan bytes) receives the
 * plans grouped by class (one kernel and one subcarrier count), heavier classes and drops first;
 * order_host[k] = the drop at slot k; classes_host (capacity max_classes) the classes.  Copy
 * plans_host to the device for hpx_uneven_run. */
int hpx_uneven_prepare(const hpx_link* links_host, const hpx_drop_tables* tables_host, const uint32_t* drop_index_host,
                       int64_t num_drops, uint64_t seed, const double* noise_power, int32_t* errors, void* plans_host,
                       int64_t* order_host, hpx_uneven_class* classes_host, int max_classes, int* num_classes_host);

/* One launch per class from the device copy of the plans. */
int hpx_uneven_run(const void* plans, const hpx_uneven_class* classes_host, int num_classes, void* stream);

/* hpx_uneven_prepare without caller workspace: the lean drops' Taylor and twiddle tables live in
 * the library's own scratch array (a file-scope device array of hpx_uneven_noalloc_capacity()
 * complex64 values, 32 MiB, held on every device the library runs on until the process ends);
 * tables_host[i].workspace is ignored.  No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops).  Allocates no device
 * memory; rejects (HPX_ERR_INVALID) a batch whose lean drops need more than the capacity.  All
 * batches planned this way on one device share that array: launches of two such batches that
 * overlap in time on one device race (the later plan overwrites the earlier one's tables).  The
 * first call on a device looks up the array's address there (that device must be current). */
int hpx_uneven_prepare_noalloc(const hpx_link* links_host, const hpx_drop_tables* tables_host,
                               const uint32_t* drop_index_host, int64_t num_drops, uint64_t seed,
                               const double* noise_power, int32_t* errors, void* plans_host, int64_t* order_host,
                               hpx_uneven_class* classes_host, int max_classes, int* num_classes_host);

/* Capacity of the scratch array of hpx_uneven_prepare_noalloc in complex64 values. */
int64_t hpx_uneven_noalloc_capacity(void);

/* ---------------------------------------------------------------- massive MIMO
 * Uplink drops with many receive antennas (<= 64) and single-antenna users (<= 8): beyond the
 * reference's 10-antenna arrays (its spatial response slices a 10 x 10 block of uniforms,
 * hermespy/channel/fading/fading.py:476-489; here the block is (max(10, Nrx), max(10, Ntx)),
 * identical for arrays of at most 10).  OFDM grid of num_subcarriers with `active` data
 * subcarriers centred on DC (grid columns N/2 - active/2 ..), NULL elsewhere.  LMMSE detection
 * on ideal CSI (the MMSE extension of waveform.py:235-250) with the matched filter V^H S^H y on
 * the MFMA units (bfloat16 hi/lo split, float32 accumulation). */
typedef struct hpx_massive_link {
  int32_t num_rx, num_users, num_subcarriers, active, words, prefix, modulation_order;
  int32_t num_paths, num_sinusoids, max_delay, dc_suppression;
  double bandwidth, doppler, los_doppler, gain, device_power;
} hpx_massive_link;

/* Caller-owned per-drop device buffers; element counts per drop from hpx_massive_sizes (in this
 * order; det_l is float32, the rest hpx_c64). */
typedef struct hpx_massive_work {
  hpx_c64* spatial;  /* [nr][users] */
  hpx_c64* coef;     /* [paths][blocks][5] Taylor coefficients of the taps */
  hpx_c64* abar;     /* [paths][words] tap means */
  hpx_c64* det_v;    /* [users][users] */
  hpx_c64* det_a;    /* [users][nr] A = V^H S^H, then [users][users] B = A S */
  float* det_l;      /* [users] eigenvalues of S^H S */
  hpx_c64* z;        /* [users][words][N] delay-line outputs over the data windows */
  hpx_c64* grid;     /* [words][nr][active] received active subcarriers, or NULL */
  hpx_c64* csi;      /* [words][active] ideal-CSI factor c (H = S c), NULL with grid */
} hpx_massive_work;

/* Per-drop element counts of the nine hpx_massive_work buffers. */
int hpx_massive_sizes(const hpx_massive_link* link_host, int64_t* sizes_host);

/* num_drops (<= 65535) drops first_drop.. with noise_power[d]; errors[d] += the drop's bit errors
 * (zero them first); equalized [drops][users][words][active] or NULL.  delays / path_gains /
 * corr_rx / corr_tx as in hpx_run_batch (corr_tx may be NULL with corr_rx given: identity).
 * The detection never needs the per-antenna grid: work->grid and work->csi are exports, written
 * when given (both or neither). */
int hpx_massive_run(const hpx_massive_link* link_host, const int32_t* delays, const double* path_gains,
                    const hpx_c128* corr_rx, const hpx_c128* corr_tx, uint64_t seed, int64_t first_drop,
                    int64_t num_drops, const double* noise_power, const hpx_massive_work* work_host, int32_t* errors,
                    hpx_c64* equalized, void* stream);

/* Host twin of the per-drop detector hpx_run_batch builds for ideal CSI, where every
 * resource element's channel is H_k = c_k S: S^H S = V diag(lambda) V^H (cyclic
 * Jacobi, float64) and A = V^H S^H, so that the MMSE/ZF filter of element k is
 * V diag(conj(c_k) / (|c_k|^2 lambda_i + s2)) A.  S: [num_rx][num_tx]; outputs
 * lambda [num_tx], V [num_tx][num_tx], A [num_tx][num_rx]; sizes <= 8.
 * No reference counterpart (the reference applies numpy.linalg.pinv per element,
 * hermespy/modem/waveform.py:247). */
int hpx_spatial_detector_host(int num_rx, int num_tx, const hpx_c128* spatial_host, double* lambda_host,
                              hpx_c128* v_host, hpx_c128* a_host);

/* Float64 drops of a single-antenna link of at most 256 subcarriers without a frame or antenna
 * correlation: the drops of hpx_run_batch (same seed, drop indices and streams; the noise from the
 * same 24-bit uniforms) with every value a double, the taps summed exactly per sample
 * (hermespy/channel/fading/fading.py:293-343 in float64, as the reference computes them) and the
 * OFDM transforms as direct DFTs.  Replaces the same call chain as hpx_run_batch for the float64
 * measurement of the small link.  equalized: complex128 [num_drops][words * num_subcarriers]
 * (element q of word w at w * N + q) or NULL. */
int hpx_run_batch_f64(const hpx_link* link_host, const int32_t* delays, const double* path_gains, uint64_t seed,
                      int64_t first_drop, int64_t num_drops, const double* noise_power, int32_t* errors,
                      hpx_c128* equalized, void* stream);

/* Dynamic LDS bytes one hpx_run_batch workgroup needs for `link` (0 if unsupported). */
int64_t hpx_run_batch_lds_bytes(const hpx_link* link_host);

#ifdef __cplusplus
}
#endif

#endif /* HERMESPY_AMD_H */
