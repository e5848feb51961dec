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

    def __post_init__(self) -> None:
        order = np.argsort(np.asarray(self.delays, dtype=np.float64))  # MultipathFadingChannel sorts, fading.py:707-711
        self.delays = np.asarray(self.delays, dtype=np.float64)[order]
        self.power_profile = np.asarray(self.power_profile, dtype=np.float64)[order]
        self.rice_factors = np.asarray(self.rice_factors, dtype=np.float64)[order]

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
    def bits_per_drop(self) -> int:
        return self.num_tx * self.words * self.num_subcarriers * self.bits_per_symbol

    @property
    def num_paths(self) -> int:
        return int(self.delays.size)

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


CONFIGS = {"siso-64": siso_64, "mimo4x4-1024": mimo4x4_1024}
