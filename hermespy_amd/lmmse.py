"""Linear MMSE pilot-to-grid channel estimation, the fallback of the learned estimator.

The reference's OFDM estimators are ideal CSI and pilot LS with linear
interpolation (``hermespy/simulation/modem/channel_estimation.py:215-295``,
``hermespy/modem/waveforms/orthogonal/channel.py:25-66``).  This estimator
keeps the LS front end (``y_p / x_p`` at every stream's own pilots) and
replaces the interpolation by the linear MMSE filter of the channel's second
order statistics:

    h_hat = R_hp (R_pp + sigma^2 I)^-1 h_ls

with R(w1, q1; w2, q2) = sigma_h^2 rho_t((w1 - w2) T) sum_l P_l exp(-2 pi j (b(q1) - b(q2)) d_l / N),
where b(q) is the FFT bin of grid column q, d_l the tap delays in samples, T the
word duration and rho_t the time correlation of the sum-of-sinusoids taps,
J0(omega t) (``fading.py:293-343``), averaged over the link's Doppler
distribution.  The filters depend only on the frame, the channel statistics
and the SNR, so they are built once on the host (float64) per SNR point and
applied on the device as one complex GEMM per launch (``hpx_lmmse_apply``).
"""

from __future__ import annotations

import numpy as np
import torch
from scipy.special import j0

from . import _lib
from .configs import LinkConfig


def subcarrier_bins(n: int, dc_suppression: bool = True) -> np.ndarray:
    """FFT bin of every grid column: the DC roll and the (i)fftshift of ``ofdm.py:97-161``."""

    q = np.arange(n)
    p = q.copy()
    dc = n // 2
    if dc_suppression:
        upper = q >= dc
        p[upper] = dc + (q[upper] - dc + 1) % (n - dc)
    return (p - n // 2) % n


def stream_pilots(cfg: LinkConfig):
    """Flat element index and value of every stream's own (non-zero) pilots, (num_tx, P) each."""

    if cfg.frame_mask is None or cfg.num_references == 0:
        raise ValueError("the link has no reference elements")
    ref_flat = np.flatnonzero(cfg.frame_mask.reshape(-1) == 0)
    own = [np.flatnonzero(cfg.pilots[t] != 0) for t in range(cfg.num_tx)]
    p = {o.size for o in own}
    if len(p) != 1:
        raise ValueError("every stream needs the same number of pilots")
    idx = np.stack([ref_flat[o] for o in own]).astype(np.int32)
    val = np.stack([cfg.pilots[t][own[t]] for t in range(cfg.num_tx)])
    return idx, val


def channel_power(cfg: LinkConfig) -> np.ndarray:
    """E|H[r, t]|^2 per stream pair: (sum_p |R_rx[r, p]|^2)(sum_q |R_tx[q, t]|^2) g sum_l P_l, the spatial
    response S = R_rx exp(2 pi j U) R_tx having independent unit-modulus phases (fading.py:476-489)."""

    g = cfg.gain * float(np.sum(cfg.power_profile))
    if cfg.corr_rx is None:
        return np.full((cfg.num_rx, cfg.num_tx), g)
    rr = np.abs(np.asarray(cfg.corr_rx)[: cfg.num_rx, : cfg.num_rx]) ** 2
    rt = np.abs(np.asarray(cfg.corr_tx)[: cfg.num_tx, : cfg.num_tx]) ** 2
    return np.outer(rr.sum(axis=1), rt.sum(axis=0)) * g


def time_correlation(cfg: LinkConfig, lags_s: np.ndarray, points: int = 2001) -> np.ndarray:
    """rho_t(tau) = E_omega[J0(omega tau)] over the link's Doppler distribution (uniform in Hz when
    the link draws its Doppler per drop, else the fixed doppler_frequency)."""

    lags_s = np.asarray(lags_s, dtype=np.float64)
    if cfg.doppler_hz_range is None:
        return j0(cfg.doppler * lags_s)
    lo, hi = cfg.doppler_hz_range
    f = np.linspace(lo, hi, points)
    return np.mean(j0(2 * np.pi * f[:, None] * lags_s.reshape(1, -1)), axis=0).reshape(lags_s.shape)


def lmmse_filters(cfg: LinkConfig, snr_db) -> np.ndarray:
    """Filters (num_snr, num_tx, words * N, P) complex128 of every SNR point."""

    n, words = cfg.num_subcarriers, cfg.words
    idx, _ = stream_pilots(cfg)
    bins = subcarrier_bins(n, cfg.dc_suppression)
    d = cfg.delay_samples
    p = cfg.power_profile / np.sum(cfg.power_profile)
    # frequency correlation over bin differences, with paths sharing a delay adding their power
    db = np.arange(-(n - 1), n)
    rf = np.exp(-2j * np.pi * np.outer(db, d) / n) @ p  # (2n - 1,)
    period = (n + cfg.prefix) / cfg.bandwidth
    rt = time_correlation(cfg, np.arange(-(words - 1), words) * period)  # (2 words - 1,)
    sh2 = float(np.mean(channel_power(cfg)))
    ew, eq = np.divmod(np.arange(words * n), n)
    out = np.empty((len(snr_db), cfg.num_tx, words * n, idx.shape[1]), dtype=np.complex128)
    for t in range(cfg.num_tx):
        pw, pq = np.divmod(idx[t], n)
        r_ep = sh2 * rt[ew[:, None] - pw[None, :] + words - 1] * rf[bins[eq][:, None] - bins[pq][None, :] + n - 1]
        r_pp = r_ep[idx[t]]
        for s, snr in enumerate(snr_db):
            sigma2 = float(cfg.noise_power(snr))
            a = r_pp + sigma2 * np.eye(r_pp.shape[0])
            # W = R_ep A^-1  <=>  A^T W^T = R_ep^T  (A Hermitian)
            out[s, t] = np.linalg.solve(a.T, r_ep.T).T
    return out


class LMMSEChannelEstimation:
    """Device-resident LMMSE filters of one pilot link; estimates [num_tx][drops][num_rx][words * N]."""

    def __init__(self, cfg: LinkConfig, snr_db=None, device: str | torch.device = "cuda") -> None:
        self.cfg = cfg
        self.snr_db = np.asarray(cfg.snr_db if snr_db is None else snr_db, dtype=np.float64)
        dev = torch.device(device)
        self.device = dev
        idx, val = stream_pilots(cfg)
        self.num_pilots = idx.shape[1]
        self.pilot_index = torch.as_tensor(idx, device=dev)
        self.pilot_values = torch.as_tensor(val.astype(np.complex64), device=dev)
        self.filters = torch.as_tensor(lmmse_filters(cfg, self.snr_db).astype(np.complex64), device=dev)

    def pilot_ls(self, grid: torch.Tensor, num_drops: int) -> torch.Tensor:
        """LS values [num_tx][drops][num_rx][P] of received grids [drops][num_rx][words][N]."""

        c = self.cfg
        ls = torch.empty((c.num_tx, num_drops, c.num_rx, self.num_pilots), dtype=torch.complex64, device=self.device)
        _lib.check(_lib.load().hpx_pilot_ls(_lib.ptr(grid), num_drops, c.num_rx, c.num_tx, c.words * c.num_subcarriers,
                                            self.num_pilots, _lib.ptr(self.pilot_index), _lib.ptr(self.pilot_values),
                                            _lib.ptr(ls), _lib.stream_handle()))
        return ls

    def apply(self, ls: torch.Tensor, snr_index: np.ndarray, out: torch.Tensor | None = None,
              rows_by_set: list | None = None) -> torch.Tensor:
        """Interpolate LS values; drop d uses the filter of SNR point snr_index[d]."""

        c = self.cfg
        num_drops = ls.shape[1]
        e = c.words * c.num_subcarriers
        if out is None:
            out = torch.empty((c.num_tx, num_drops, c.num_rx, e), dtype=torch.complex64, device=self.device)
        if rows_by_set is None:
            rows_by_set = self.rows_by_set(snr_index)
        lib, st = _lib.load(), _lib.stream_handle()
        for s, rows in enumerate(rows_by_set):
            if rows is None or rows.numel() == 0:
                continue
            _lib.check(lib.hpx_lmmse_apply(_lib.ptr(ls), c.num_tx, num_drops * c.num_rx, self.num_pilots,
                                           _lib.ptr(self.filters[s]), e, _lib.ptr(rows), int(rows.numel()),
                                           _lib.ptr(out), st))
        return out

    def rows_by_set(self, snr_index: np.ndarray) -> list:
        """Device lists of the LS rows (drop * num_rx + r) of every SNR point."""

        snr_index = np.asarray(snr_index)
        nr = self.cfg.num_rx
        out = []
        for s in range(self.snr_db.size):
            drops = np.flatnonzero(snr_index == s)
            rows = (drops[:, None] * nr + np.arange(nr)[None, :]).reshape(-1).astype(np.int32)
            out.append(torch.as_tensor(rows, device=self.device) if rows.size else None)
        return out
