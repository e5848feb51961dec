"""Sum-of-sinusoids multipath fading, restated in float64.

Follows ``hermespy/channel/fading/fading.py``:

* channel set-up (delay sort, LOS/NLOS gains, RNG variable offsets) :606-754,
* realisation sampling (spatial response R_rx S R_tx, angles, phases) :468-499,
  with the uniform transform Phi(z) of ``hermespy/channel/consistent.py:475-485``,
* tap generation ``__path_impulse_generator`` :293-343 (no 2*pi on the Doppler
  argument; time restarts at 0 every call),
* ``_propagate`` :371-406 and the dense impulse response ``state`` :345-369.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy.special import ndtr


@dataclass
class FadingParameters:
    """Static parameters of a ``MultipathFadingChannel`` after its constructor ran."""

    delays: np.ndarray  # seconds, sorted
    power_profile: np.ndarray
    rice_factors: np.ndarray
    num_sinusoids: int
    doppler: float  # the reference's ``doppler_frequency`` (used as rad/s)
    los_doppler: float
    gain: float = 1.0
    corr_rx: np.ndarray | None = None
    corr_tx: np.ndarray | None = None

    @property
    def num_paths(self) -> int:
        return int(self.delays.size)

    @property
    def num_normals(self) -> int:
        """Scalars one static realisation consumes, ``consistent.py:423-432`` + ``fading.py:742-754``."""

        L, N = self.num_paths, self.num_sinusoids
        return 100 + 2 * L + 2 * L * N

    def los_gains(self) -> np.ndarray:
        """``fading.py:721-729``."""

        k = self.rice_factors
        g = np.empty(k.size)
        inf = np.isposinf(k)
        g[inf] = 1.0
        g[~inf] = np.sqrt(k[~inf] / (1 + k[~inf]))
        return g

    def nlos_gains(self) -> np.ndarray:
        """``fading.py:730-733``."""

        k = self.rice_factors
        g = np.empty(k.size)
        inf = np.isposinf(k)
        g[~inf] = np.sqrt(1 / ((1 + k[~inf]) * self.num_sinusoids))
        g[inf] = 0.0
        return g


def make_parameters(delays, power_profile, rice_factors, num_sinusoids=20, doppler=0.0, los_doppler=None,
                    gain=1.0, corr_rx=None, corr_tx=None) -> FadingParameters:
    """Constructor semantics of ``MultipathFadingChannel.__init__`` (``fading.py:669-716``)."""

    d = np.asarray(delays, dtype=np.float64)
    p = np.asarray(power_profile, dtype=np.float64)
    k = np.asarray(rice_factors, dtype=np.float64)
    if d.ndim != 1 or p.ndim != 1 or k.ndim != 1:
        raise ValueError("Delays, power profile and rice factors must be vectors")
    if d.size < 1:
        raise ValueError("Configuration must contain at least one delay tap")
    if d.size != p.size or p.size != k.size:
        raise ValueError("Delays, power profile and rice factor vectors must be of equal length")
    if np.any(d < 0) or np.any(p < 0) or np.any(k < 0):
        raise ValueError("Delays, power profile and rice factors must be non-negative")
    order = np.argsort(d)
    return FadingParameters(d[order], p[order], k[order], int(num_sinusoids), float(doppler),
                            float(doppler if los_doppler is None else los_doppler), float(gain), corr_rx, corr_tx)


@dataclass
class FadingSample:
    """Exported realisation of one link: what ``MultipathFadingSample`` holds (``fading.py:156-215``)."""

    params: FadingParameters
    spatial_response: np.ndarray  # (Nrx, Ntx) complex
    los_angles: np.ndarray  # (L,)
    nlos_angles: np.ndarray  # (L, N)
    los_phases: np.ndarray  # (L,)
    nlos_phases: np.ndarray  # (L, N)
    bandwidth: float


def sample_from_normals(params: FadingParameters, z: np.ndarray, num_rx: int, num_tx: int, bandwidth: float) -> FadingSample:
    """``MultipathFadingRealization._sample`` (``fading.py:468-499``) on a static realisation ``z``."""

    L, N = params.num_paths, params.num_sinusoids
    u = ndtr(np.asarray(z, dtype=np.float64))  # ConsistentUniform.sample, consistent.py:485
    o = 0
    corr = u[o:o + 100].reshape(10, 10); o += 100
    los_u = u[o:o + L]; o += L
    nlos_ang_u = u[o:o + L * N].reshape(L, N); o += L * N
    los_ph_u = u[o:o + L]; o += L
    nlos_ph_u = u[o:o + L * N].reshape(L, N); o += L * N
    s = np.exp(2j * np.pi * corr)[:num_rx, :num_tx]
    if params.corr_rx is not None or params.corr_tx is not None:
        rrx = params.corr_rx[:num_rx, :num_rx]
        rtx = params.corr_tx[:num_tx, :num_tx]
        s = rrx @ s @ rtx
    return FadingSample(params, s, 2 * np.pi * los_u, -np.pi + 2 * np.pi * nlos_ang_u,
                        -np.pi + 2 * np.pi * los_ph_u, -np.pi + 2 * np.pi * nlos_ph_u, float(bandwidth))


def delay_samples(sample: FadingSample) -> np.ndarray:
    """``fading.py:297``: ``rint(delay * bandwidth)``."""

    return np.rint(sample.params.delays * sample.bandwidth).astype(int)


def max_delay_samples(sample: FadingSample) -> int:
    """``fading.py:372``: ``round(max_delay * bandwidth)``."""

    return int(round(float(sample.params.delays.max()) * sample.bandwidth))


def path_impulses(sample: FadingSample, num_samples: int):
    """``__path_impulse_generator`` (``fading.py:293-343``), path by path."""

    p = sample.params
    n = 1 + np.arange(p.num_sinusoids)
    t = np.arange(num_samples) / sample.bandwidth
    nlos_time = p.doppler * t
    los_time = p.los_doppler * t
    lg, ng = p.los_gains(), p.nlos_gains()
    for d, pw, lgl, ngl, la, na, lp, nph in zip(delay_samples(sample), p.power_profile, lg, ng, sample.los_angles,
                                               sample.nlos_angles, sample.los_phases, sample.nlos_phases):
        imp = ngl * np.sum(np.exp(1j * (np.outer(nlos_time, np.cos((2 * np.pi * n + na) / p.num_sinusoids)) + nph[None, :])), axis=1)
        imp += lgl * np.exp(1j * (los_time * np.cos(la) + lp))
        imp *= (p.gain * pw) ** 0.5
        yield imp, int(d)


def taps(sample: FadingSample, num_samples: int) -> np.ndarray:
    """All path impulses stacked, (L, num_samples)."""

    return np.stack([imp for imp, _ in path_impulses(sample, num_samples)])


def propagate(sample: FadingSample, x: np.ndarray) -> np.ndarray:
    """``MultipathFadingSample._propagate`` (``fading.py:371-406``); x: (Ntx, M) -> (Nrx, M + D)."""

    x = np.atleast_2d(x)
    m = x.shape[1]
    out_len = m + max_delay_samples(sample)
    s = sample.spatial_response
    z = np.zeros((s.shape[1], out_len), dtype=np.complex128)
    for imp, d in path_impulses(sample, m):
        z[:, d:d + m] += x * imp[None, :]
    return s @ z


def state(sample: FadingSample, num_samples: int, max_num_taps: int) -> np.ndarray:
    """Dense impulse response (Nrx, Ntx, num_samples, taps), ``fading.py:345-369``."""

    num_taps = min(1 + round(float(sample.params.delays.max()) * sample.bandwidth), max_num_taps)
    siso = np.zeros((num_samples, num_taps), dtype=np.complex128)
    for imp, d in path_impulses(sample, num_samples):
        if d > num_taps:
            continue
        siso[:, d] += imp
    return np.einsum("ij,kl->ijkl", sample.spatial_response, siso)


def propagate_state(csi: np.ndarray, x: np.ndarray) -> np.ndarray:
    """Propagation through a dense impulse response: y[r, n] = sum_{t,d} csi[r, t, n - d, d] x[t, n - d].

    The convention of ``ChannelStateInformation.propagate`` for impulse responses
    (the reference's invariant ``propagate == state().propagate``,
    ``tests/unit_tests/channel/test_fading.py:167-175``).
    """

    nrx, ntx, m, ntaps = csi.shape
    y = np.zeros((nrx, m + ntaps - 1), dtype=np.complex128)
    for d in range(ntaps):
        y[:, d:d + m] += np.einsum("rtn,tn->rn", csi[:, :, :, d], x)
    return y
