"""Link configurations of the drop path (plain parameters, as a reference user
would pass them to ``MultipathFadingChannel`` / ``OFDMWaveform`` / ``SNR``).

Doppler: the reference multiplies ``doppler_frequency`` directly with time
(``hermespy/channel/fading/fading.py:303-304,330``) without a 2 pi, so the
parameter acts as an angular rate; a Doppler shift f_D in Hz is passed as
``2 pi f_D``.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

SPEED_OF_LIGHT = 299792458.0


@dataclass
class LinkConfig:
    num_rx: int
    num_tx: int
    num_subcarriers: int
    words: int
    prefix_ratio: float
    modulation_order: int
    bandwidth: float
    delays: np.ndarray  # seconds
    power_profile: np.ndarray
    rice_factors: np.ndarray
    doppler: float  # reference doppler_frequency (angular rate, rad/s)
    num_sinusoids: int = 20
    corr_rx: np.ndarray | None = None
    corr_tx: np.ndarray | None = None
    detector: str = "mmse"
    dc_suppression: bool = True
    gain: float = 1.0
    device_power: float = 1.0
    los_doppler: float | None = None
    name: str = "link"
    snr_db: tuple = field(default=(0.0,))
    workgroup_size: int = 0  # threads per drop in the fused kernel (0: library default)
    # frame with reference / null elements: (words, N) ElementType codes (REFERENCE 0, DATA 1, NULL 2)
    frame_mask: np.ndarray | None = None
    pilots: np.ndarray | None = None  # (num_tx, num_references) reference symbols of each stream
    doppler_hz_range: tuple | None = None  # per-drop Doppler shift drawn uniformly from this range [Hz]

    def __post_init__(self) -> None:
        order = np.argsort(np.asarray(self.delays, dtype=np.float64))  # MultipathFadingChannel sorts, fading.py:707-711
        self.delays = np.asarray(self.delays, dtype=np.float64)[order]
        self.power_profile = np.asarray(self.power_profile, dtype=np.float64)[order]
        self.rice_factors = np.asarray(self.rice_factors, dtype=np.float64)[order]
        if self.frame_mask is not None:
            self.frame_mask = np.asarray(self.frame_mask, dtype=np.int8)
            if self.frame_mask.shape != (self.words, self.num_subcarriers):
                raise ValueError("frame_mask must have shape (words, num_subcarriers)")
            nref = int(np.count_nonzero(self.frame_mask == 0))
            if nref and (self.pilots is None or np.shape(self.pilots) != (self.num_tx, nref)):
                raise ValueError("reference elements need (num_tx, num_references) pilot symbols")

    @classmethod
    def from_channel(cls, channel, num_rx: int, num_tx: int, num_subcarriers: int, words: int, prefix_ratio: float,
                     modulation_order: int, bandwidth: float, **kwargs) -> "LinkConfig":
        """The link of a configured ``MultipathFadingChannel`` (or one of its parameterisations:
        ``TDL``, ``Cost259``, ``Exponential``) between a transmitter of ``num_tx`` and a receiver of
        ``num_rx`` antennas, sampled at ``bandwidth``: its sorted delays, powers and Rice factors,
        sinusoid count, NLOS and LOS doppler_frequency, gain and the covariances its antenna
        correlation gives the two ends (``fading.py:476-489``; a ``StandardAntennaCorrelation`` treats
        the transmitter as base station and the receiver as terminal, ``correlation.py:72-106``)."""

        corr = getattr(channel, "antenna_correlation", None)
        if corr is not None:
            if hasattr(corr, "covariance_for"):
                kwargs.setdefault("corr_rx", corr.covariance_for(num_rx, True))
                kwargs.setdefault("corr_tx", corr.covariance_for(num_tx, False))
            else:
                kwargs.setdefault("corr_rx", corr.sample_covariance(num_rx))
                kwargs.setdefault("corr_tx", corr.sample_covariance(num_tx))
        kwargs.setdefault("num_sinusoids", int(channel.num_sinusoids))
        kwargs.setdefault("gain", float(channel.gain))
        kwargs.setdefault("los_doppler", float(channel.los_doppler_frequency))
        return cls(num_rx, num_tx, num_subcarriers, words, prefix_ratio, modulation_order, bandwidth,
                   np.array(channel.delays, dtype=np.float64), np.array(channel.power_profile, dtype=np.float64),
                   np.array(channel.rice_factors, dtype=np.float64), float(channel.doppler_frequency), **kwargs)

    # ---------------------------------------------------------------- derived
    @property
    def prefix(self) -> int:
        return int(self.num_subcarriers * 1 * self.prefix_ratio)  # waveform.py:538

    @property
    def samples_per_frame(self) -> int:
        return self.words * (self.num_subcarriers + self.prefix)

    @property
    def bits_per_symbol(self) -> int:
        return int(np.log2(self.modulation_order))

    @property
    def num_data(self) -> int:
        """Data symbols per stream and frame."""

        if self.frame_mask is None:
            return self.words * self.num_subcarriers
        return int(np.count_nonzero(self.frame_mask == 1))

    @property
    def num_references(self) -> int:
        return 0 if self.frame_mask is None else int(np.count_nonzero(self.frame_mask == 0))

    @property
    def bits_per_drop(self) -> int:
        return self.num_tx * self.num_data * self.bits_per_symbol

    @property
    def algorithmic_flops(self) -> float:
        """Per-drop work of the reference algorithm, the roofline's numerator: 8 flops per complex
        MAC of the tapped delay line (paths x frame samples x transmit streams), the spatial mix
        (Nrx Ntx (frame + max delay)), the ideal-CSI DFT over the taps (paths x words x N) and the
        per-element MMSE detection (Gram Ntx^2 Nrx, Cholesky Ntx^3 / 3, matched filter Ntx Nrx and two
        triangular solves 2 Ntx^2 complex MACs: 133 at 4x4), plus 5 N log2 N flops per FFT of every
        transmit and receive stream and word."""

        m = self.samples_per_frame
        nt, nr, n = self.num_tx, self.num_rx, self.num_subcarriers
        tdl = self.num_paths * m * nt
        smix = nr * nt * (m + self.max_delay)
        csi = self.num_paths * self.words * n
        det = (nt * nt * nr + nt ** 3 // 3 + nt * nr + 2 * nt * nt) * self.words * n
        fft = 5 * n * np.log2(n) * self.words * (nt + nr)
        return 8.0 * (tdl + smix + csi + det) + fft

    def draw_doppler(self, first_drop: int, count: int) -> np.ndarray | None:
        """Per-drop doppler_frequency (2 pi f_D, NLOS and LOS alike) of drops first_drop.., or None.

        Drop i's shift is a function of i alone (a splitmix64 hash to a uniform), so any drop can
        be re-simulated on its own."""

        if self.doppler_hz_range is None:
            return None
        lo, hi = self.doppler_hz_range
        i = np.arange(first_drop, first_drop + count, dtype=np.uint64)
        with np.errstate(over="ignore"):
            z = i * np.uint64(0x9E3779B97F4A7C15) + np.uint64(0x632BE59BD9B4E019)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        w = 2 * np.pi * (lo + (hi - lo) * u)
        return np.stack([w, w], axis=-1)

    @property
    def num_paths(self) -> int:
        return int(self.delays.size)

    @property
    def num_normals(self) -> int:
        """Standard normals of one realisation: the 10 x 10 spatial block, then the LOS angles,
        NLOS angles, LOS phases and NLOS phases (``consistent.py:410-432``, ``fading.py:468-515``)."""

        return 100 + 2 * self.num_paths + 2 * self.num_paths * self.num_sinusoids

    @property
    def delay_samples(self) -> np.ndarray:
        return np.rint(self.delays * self.bandwidth).astype(np.int32)  # fading.py:297

    @property
    def max_delay(self) -> int:
        return int(round(float(self.delays.max()) * self.bandwidth))  # fading.py:372

    @property
    def los_gains(self) -> np.ndarray:
        k = self.rice_factors
        g = np.ones(k.size)
        fin = ~np.isposinf(k)
        g[fin] = np.sqrt(k[fin] / (1 + k[fin]))  # fading.py:726-729
        return g

    @property
    def nlos_gains(self) -> np.ndarray:
        k = self.rice_factors
        g = np.zeros(k.size)
        fin = ~np.isposinf(k)
        g[fin] = np.sqrt(1 / ((1 + k[fin]) * self.num_sinusoids))  # fading.py:730-733
        return g

    @property
    def path_gains(self) -> np.ndarray:
        """[2][L]: LOS and NLOS term amplitudes including sqrt(gain * P_l) (fading.py:342)."""

        s = np.sqrt(self.gain * self.power_profile)
        return np.stack([self.los_gains * s, self.nlos_gains * s])

    def noise_power(self, snr_db) -> np.ndarray:
        """``SNR.get_power``: P_ref / snr * expected_energy_scale (rf/noise/level.py:257-259)."""

        snr = 10.0 ** (np.asarray(snr_db, dtype=np.float64) / 10.0)
        return self.device_power / snr * (self.gain * float(np.sum(self.power_profile)))


def exponential_correlation(n: int, rho: float) -> np.ndarray:
    idx = np.arange(n)
    return (rho ** np.abs(np.subtract.outer(idx, idx))).astype(np.complex128)


def siso_64() -> LinkConfig:
    """SISO 64-subcarrier OFDM, 5-tap Rayleigh, 3 dB/tap, 50 Hz Doppler at 1 MHz, QPSK, ZF."""

    fs = 1e6
    p = 10.0 ** (-0.3 * np.arange(5))
    return LinkConfig(1, 1, 64, 14, 0.25, 4, fs, np.arange(5) / fs, p / p.sum(), np.zeros(5), 2 * np.pi * 50.0,
                      detector="zf", name="siso-64sc-qpsk-zf", snr_db=(0.0, 5.0, 10.0, 15.0, 20.0))


def mimo4x4_1024() -> LinkConfig:
    """4x4 correlated (rho=0.5) 23-tap TDL over 3 us at 30.72 MHz, K=0 dB on tap 0, 120 km/h at 3.5 GHz,
    1024 subcarriers, CP 72, 14 words, 16-QAM, MMSE with ideal CSI."""

    fs = 30.72e6
    tau = np.arange(23) * 3e-6 / 22
    p = np.exp(-tau / 1e-6)
    k = np.zeros(23)
    k[0] = 1.0
    fd = (120.0 / 3.6) * 3.5e9 / SPEED_OF_LIGHT
    r = exponential_correlation(4, 0.5)
    return LinkConfig(4, 4, 1024, 14, 72 / 1024, 16, fs, tau, p / p.sum(), k, 2 * np.pi * fd, corr_rx=r, corr_tx=r,
                      detector="mmse", name="mimo4x4-1024sc-16qam-mmse",
                      snr_db=tuple(float(s) for s in range(0, 31, 3)))


def comb_pilot_frame(num_subcarriers: int, words: int, pilot_words, spacing: int, num_tx: int, seed: int = 20261202):
    """Mask and per-stream pilots of a comb frame: ``GridResource`` [REF 1, DATA spacing-1] in the pilot
    words, all-data words elsewhere.  Reference element j of a pilot word belongs to stream j % num_tx
    (the others stay silent there), so every stream is estimated from its own pilots; the pilot values
    are a fixed QPSK sequence (a ``CustomPilotSymbolSequence``)."""

    mask = np.ones((words, num_subcarriers), dtype=np.int8)
    for w in pilot_words:
        mask[w, ::spacing] = 0
    nref = int(np.count_nonzero(mask == 0))
    rng = np.random.default_rng(seed)
    qpsk = ((2 * rng.integers(0, 2, nref) - 1) + 1j * (2 * rng.integers(0, 2, nref) - 1)) / np.sqrt(2.0)
    owner = np.tile(np.arange(num_subcarriers // spacing) % num_tx, len(pilot_words))
    pilots = np.zeros((num_tx, nref), dtype=np.complex128)
    for t in range(num_tx):
        pilots[t, owner == t] = qpsk[owner == t]
    return mask, pilots


def mimo2x2_512_pilots() -> LinkConfig:
    """2x2 correlated (rho=0.5) 23-tap TDL over 3 us (the 4x4 link's channel) at 15.36 MHz, 512 subcarriers,
    CP 64 (N/8), 14 words with comb QPSK pilots on every 4th subcarrier of words 2 and 11, Doppler
    uniform in 10..500 Hz per drop, 16-QAM, MMSE.  The grid has no DC roll (``dc_suppression=False``,
    an OFDMWaveform option), so each stream's comb samples the frequency response uniformly."""

    fs = 15.36e6
    tau = np.arange(23) * 3e-6 / 22
    p = np.exp(-tau / 1e-6)
    k = np.zeros(23)
    k[0] = 1.0
    r = exponential_correlation(2, 0.5)
    mask, pilots = comb_pilot_frame(512, 14, (2, 11), 4, 2)
    return LinkConfig(2, 2, 512, 14, 64 / 512, 16, fs, tau, p / p.sum(), k, 2 * np.pi * 500.0, corr_rx=r, corr_tx=r,
                      detector="mmse", name="mimo2x2-512sc-16qam-comb-pilots",
                      snr_db=tuple(float(s) for s in range(0, 31, 3)), frame_mask=mask, pilots=pilots,
                      doppler_hz_range=(10.0, 500.0), dc_suppression=False)


CONFIGS = {"siso-64": siso_64, "mimo4x4-1024": mimo4x4_1024, "mimo2x2-512-pilots": mimo2x2_512_pilots}
