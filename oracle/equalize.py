"""Per-resource-element equalisers, float64.

``zero_forcing`` restates ``ZeroForcingChannelEqualization.equalize_channel``
(``hermespy/modem/waveform.py:235-250``): SISO divides by the sum over transmit
states, MIMO applies ``numpy.linalg.pinv`` per resource element.

``mmse`` has no counterpart in the reference (its OFDM path only offers ZF);
it is the linear MMSE detector W = (H^H H + sigma^2 I)^{-1} H^H.  Parity
against the reference is unpinned; it is pinned by this float64 closed form
and by its sigma -> 0 limit, which is ZF for full-column-rank H.
"""

from __future__ import annotations

import numpy as np


def zero_forcing(y: np.ndarray, h: np.ndarray) -> np.ndarray:
    """y: (Nrx, K), h: (Nrx, Ntx, K) -> (Ntx, K)."""

    if y.shape[0] < 2:
        return y / np.sum(h, axis=1)
    eq = np.linalg.pinv(h.transpose((2, 0, 1)))  # (K, Ntx, Nrx)
    return np.einsum("ktr,rk->tk", eq, y)


def mmse(y: np.ndarray, h: np.ndarray, noise_variance: float) -> np.ndarray:
    """y: (Nrx, K), h: (Nrx, Ntx, K) -> (Ntx, K)."""

    hk = h.transpose((2, 0, 1))  # (K, Nrx, Ntx)
    hh = np.conj(hk.transpose((0, 2, 1)))
    g = hh @ hk + noise_variance * np.eye(hk.shape[2])[None]
    b = hh @ y.T[:, :, None]
    return np.linalg.solve(g, b)[:, :, 0].T
