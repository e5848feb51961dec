"""Massive-MIMO uplink drop (many receive antennas, single-antenna users), float64.

Test infrastructure: only ``tests/`` use this module, to check the
massive-MIMO device path against it.

The reference caps antenna arrays at 10 elements: ``MultipathFadingRealization``
draws one 10 x 10 block of uniforms for the spatial response and slices it
(``hermespy/channel/fading/fading.py:476-489``, ``consistent.py:423-432``).  The
restatement here extends the block to (max(10, Nrx), max(10, Ntx)) so that links
of at most 10 antennas keep the reference's draw layout exactly
(``test_massive.py`` checks that against :func:`oracle.fading.sample_from_normals`).
Everything else is the reference chain of :mod:`oracle.drop`: the same tap model
(``fading.py:293-406``), OFDM (``ofdm.py:97-161``) with the active subcarriers
centred and NULL elements elsewhere, AWGN (``rf/noise/model.py:143-160``), the
ideal CSI H = S c (``channel_estimation.py:215-295``) and LMMSE detection
x = (H^H H + s2 I)^-1 H^H y per resource element (the MMSE extension of
``waveform.py:235-250``), written through the eigen-decomposition of S^H S.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
from scipy.special import ndtr

from . import csi, fading, mapping, ofdm, philox


def spatial_block(num_rx: int, num_tx: int) -> tuple:
    return max(10, num_rx), max(10, num_tx)


@dataclass
class MassiveLink:
    num_rx: int
    num_users: int
    num_subcarriers: int
    active: int
    words: int
    prefix: int
    modulation_order: int
    bandwidth: float
    delays: np.ndarray
    power_profile: np.ndarray
    rice_factors: np.ndarray
    doppler: float
    num_sinusoids: int = 20
    corr_rx: np.ndarray | None = None
    gain: float = 1.0
    device_power: float = 1.0
    dc_suppression: bool = True
    params: fading.FadingParameters = field(init=False)
    frame: ofdm.Frame = field(init=False)

    def __post_init__(self) -> None:
        corr_tx = None if self.corr_rx is None else np.eye(self.num_users, dtype=np.complex128)
        self.params = fading.make_parameters(self.delays, self.power_profile, self.rice_factors, self.num_sinusoids,
                                             self.doppler, None, self.gain, self.corr_rx, corr_tx)
        mask = np.full((self.words, self.num_subcarriers), ofdm.NULL, dtype=np.int8)
        lo = self.active_first
        mask[:, lo:lo + self.active] = ofdm.DATA
        self.frame = ofdm.Frame(self.num_subcarriers, np.full(self.words, self.prefix, dtype=int), mask,
                                self.dc_suppression)

    @property
    def active_first(self) -> int:
        """First active grid column: the band is centred on the grid (DC at column N / 2)."""

        return self.num_subcarriers // 2 - self.active // 2

    @property
    def num_normals(self) -> int:
        r, c = spatial_block(self.num_rx, self.num_users)
        L, n = self.params.num_paths, self.num_sinusoids
        return r * c + 2 * L + 2 * L * n

    @property
    def bits_per_user(self) -> int:
        return self.frame.num_data_symbols * int(np.log2(self.modulation_order))

    def noise_power(self, snr_db: float) -> float:
        return self.device_power / 10 ** (snr_db / 10) * (self.gain * float(np.sum(self.params.power_profile)))


def sample_from_normals(link: MassiveLink, z: np.ndarray) -> fading.FadingSample:
    """``MultipathFadingRealization._sample`` with the extended spatial block."""

    p = link.params
    L, N = p.num_paths, p.num_sinusoids
    rows, cols = spatial_block(link.num_rx, link.num_users)
    u = ndtr(np.asarray(z, dtype=np.float64))
    o = 0
    corr = u[o:o + rows * cols].reshape(rows, cols); o += rows * cols
    los_u = u[o:o + L]; o += L
    nlos_ang_u = u[o:o + L * N].reshape(L, N); o += L * N
    los_ph_u = u[o:o + L]; o += L
    nlos_ph_u = u[o:o + L * N].reshape(L, N); o += L * N
    s = np.exp(2j * np.pi * corr)[:link.num_rx, :link.num_users]
    if p.corr_rx is not None:
        s = p.corr_rx[:link.num_rx, :link.num_rx] @ s @ p.corr_tx[:link.num_users, :link.num_users]
    return fading.FadingSample(p, s, 2 * np.pi * los_u, -np.pi + 2 * np.pi * nlos_ang_u, -np.pi + 2 * np.pi * los_ph_u,
                               -np.pi + 2 * np.pi * nlos_ph_u, float(link.bandwidth))


def realize(link: MassiveLink, seed: int, drop: int) -> fading.FadingSample:
    return sample_from_normals(link, philox.realization_normals(seed, drop, link.num_normals))


def csi_factor(sample: fading.FadingSample, frame: ofdm.Frame) -> np.ndarray:
    """c (words, N) of the ideal CSI H[r, t, w, k] = S[r, t] c[w, k] (the closed form of
    :func:`oracle.csi.closed_form_ideal_csi` without the spatial outer product)."""

    h = csi.closed_form_ideal_csi(_siso(sample), frame)
    return h[0, 0]


def _siso(sample: fading.FadingSample) -> fading.FadingSample:
    return fading.FadingSample(sample.params, np.ones((1, 1), dtype=np.complex128), sample.los_angles,
                               sample.nlos_angles, sample.los_phases, sample.nlos_phases, sample.bandwidth)


@dataclass
class MassiveResult:
    tx_bits: np.ndarray  # (users, bits)
    grid: np.ndarray  # (Nrx, words, N) received grid
    csi: np.ndarray  # (words, N) c
    spatial: np.ndarray  # (Nrx, users)
    equalized: np.ndarray  # (users, data symbols)
    rx_bits: np.ndarray
    errors: int


def run_drop(link: MassiveLink, seed: int, drop: int, snr_db: float) -> MassiveResult:
    sample = realize(link, seed, drop)
    bits = np.stack([philox.stream_bits(seed, drop, t, link.bits_per_user) for t in range(link.num_users)])
    x = np.stack([ofdm.modulate(link.frame, ofdm.place(link.frame, mapping.get_symbols(bits[t], link.modulation_order)))
                  for t in range(link.num_users)]) * link.device_power ** 0.5
    y = fading.propagate(sample, x)
    p = link.noise_power(snr_db)
    for r in range(link.num_rx):
        y[r] += (0.5 * p) ** 0.5 * philox.noise_normals(seed, drop, r, y.shape[1])
    grid = ofdm.demodulate(link.frame, y[:, :link.frame.samples_per_frame])
    c = csi_factor(sample, link.frame)
    s = sample.spatial_response
    yd = ofdm.pick(link.frame, grid)  # (Nrx, K)
    cd = ofdm.pick(link.frame, c)  # (K,)
    # x = (|c|^2 S^H S + p I)^-1 conj(c) S^H y, via S^H S = V diag(lam) V^H
    lam, v = np.linalg.eigh(s.conj().T @ s)
    b = v.conj().T @ (s.conj().T @ yd)  # (users, K)
    u = b * (np.conj(cd) / (np.abs(cd)[None, :] ** 2 * lam[:, None] + p))
    xh = v @ u
    rx = np.stack([mapping.detect_bits(xh[t], link.modulation_order).astype(np.uint8) for t in range(link.num_users)])
    errors = int(np.sum(rx != bits))
    return MassiveResult(bits, grid, c, s, xh, rx, errors)


def lmmse_direct(yd: np.ndarray, s: np.ndarray, cd: np.ndarray, p: float) -> np.ndarray:
    """(H^H H + p I)^-1 H^H y solved element by element with H = c S (the definition the eigen form
    of :func:`run_drop` must equal)."""

    out = np.empty((s.shape[1], yd.shape[1]), dtype=np.complex128)
    for k in range(yd.shape[1]):
        h = cd[k] * s
        out[:, k] = np.linalg.solve(h.conj().T @ h + p * np.eye(s.shape[1]), h.conj().T @ yd[:, k])
    return out
