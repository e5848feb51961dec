"""Massive-MIMO uplink drops on the device: 64 receive antennas, 8 single-antenna users,
4096-point OFDM with 3276 active subcarriers (273 resource blocks of 12, centred on DC),
CP 288, 14 words, 256-QAM, the correlated 23-tap TDL (rho = 0.7 at the receive array),
LMMSE detection on ideal CSI.

Beyond the reference's 10-antenna arrays (``hermespy/channel/fading/fading.py:476-489``);
``hpx_massive_run`` documents how the spatial draw is extended.  A drop does not fit one
CU, so it runs as three kernels over per-drop device buffers held here (prologue, transmit +
delay line, receive + detection per (drop, word)); the receiver forms the matched-filter
outputs of the 64 antennas directly, one transform per user, and writes the per-antenna grid
only on request (a fourth kernel); see ``hermespy_amd/csrc/massive.hip``.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .configs import SPEED_OF_LIGHT, LinkConfig, exponential_correlation


@dataclass
class MassiveConfig:
    num_rx: int = 64
    num_users: int = 8
    num_subcarriers: int = 4096
    active: int = 3276
    words: int = 14
    prefix: int = 288
    modulation_order: int = 256
    bandwidth: float = 4096 * 30e3
    delays: np.ndarray = field(default_factory=lambda: np.arange(23) * 3e-6 / 22)
    power_profile: np.ndarray | None = None
    rice_factors: np.ndarray | None = None
    doppler: float = 2 * np.pi * (120.0 / 3.6) * 3.5e9 / SPEED_OF_LIGHT
    num_sinusoids: int = 20
    rho_rx: float = 0.7
    gain: float = 1.0
    device_power: float = 1.0
    dc_suppression: bool = True
    snr_db: tuple = tuple(float(s) for s in range(10, 41, 3))
    name: str = "massive-64x8-4096sc-256qam-lmmse"

    def __post_init__(self) -> None:
        self.delays = np.asarray(self.delays, dtype=np.float64)
        if self.power_profile is None:
            p = np.exp(-self.delays / 1e-6)
            self.power_profile = p / p.sum()
        if self.rice_factors is None:
            k = np.zeros(self.delays.size)
            k[0] = 1.0
            self.rice_factors = k
        self.power_profile = np.asarray(self.power_profile, dtype=np.float64)
        self.rice_factors = np.asarray(self.rice_factors, dtype=np.float64)

    @property
    def corr_rx(self) -> np.ndarray:
        return exponential_correlation(self.num_rx, self.rho_rx)

    def link(self) -> LinkConfig:
        """The channel parameters as a LinkConfig (delay samples, path amplitudes, noise power)."""

        return LinkConfig(self.num_rx, self.num_users, self.num_subcarriers, self.words,
                          self.prefix / self.num_subcarriers, self.modulation_order, self.bandwidth, self.delays,
                          self.power_profile, self.rice_factors, self.doppler, self.num_sinusoids, gain=self.gain,
                          device_power=self.device_power, dc_suppression=self.dc_suppression)

    @property
    def bits_per_drop(self) -> int:
        return self.num_users * self.words * self.active * int(np.log2(self.modulation_order))

    def noise_power(self, snr_db) -> np.ndarray:
        return self.link().noise_power(snr_db)

    @property
    def mfma_flops_per_drop(self) -> float:
        """Matrix-core work the receiver executes per drop: per word, N / 32 column tiles of 16
        sample pairs, each over ceil(Nrx / 4) noise chunks and ceil(users / 4) user chunks of four
        16x16x4 float32 MFMAs (2 x 16 x 16 x 4 flops each); every product is a term of the
        matched filter A n or of the user mix A S z."""

        chunks = -(-self.num_rx // 4) + -(-self.num_users // 4)
        return float(self.words * (self.num_subcarriers // 32) * chunks * 4 * 2 * 16 * 16 * 4)

    @property
    def hbm_bytes_per_drop(self) -> float:
        """Compulsory HBM traffic of a drop: the users' delay-line outputs (complex64, users x words
        x N) written by the transmitter and read once by the receiver, the per-drop tables and the
        error count."""

        z = self.num_users * self.words * self.num_subcarriers * 8
        return float(2 * z)

    @property
    def reference_flops_per_drop(self) -> float:
        """Work of the reference's form of the link (per-antenna receive FFTs, the per-element
        Gram, Cholesky, matched filter and triangular solves, the full-frame delay line and spatial
        mix): the figure a drop would cost without the per-drop eigen form and the FFT's linearity,
        8 flops per complex MAC and 5 N log2 N per FFT."""

        lk = self.link()
        m, nt, nr, n = lk.samples_per_frame, self.num_users, self.num_rx, self.num_subcarriers
        tdl = lk.num_paths * m * nt
        smix = nr * nt * (m + lk.max_delay)
        csi = lk.num_paths * self.words * n
        det = (nt * nt * nr + nt ** 3 // 3 + nt * nr + 2 * nt * nt) * self.words * self.active
        fft = 5 * n * np.log2(n) * self.words * (nt + nr)
        return 8.0 * (tdl + smix + csi + det) + fft

    def c_link(self) -> _lib.HpxMassiveLink:
        lk = self.link()
        return _lib.HpxMassiveLink(self.num_rx, self.num_users, self.num_subcarriers, self.active, self.words,
                                   self.prefix, self.modulation_order, lk.num_paths, self.num_sinusoids, lk.max_delay,
                                   int(bool(self.dc_suppression)), float(self.bandwidth), float(self.doppler),
                                   float(self.doppler), float(self.gain), float(self.device_power))


WORK_FIELDS = ("spatial", "coef", "abar", "det_v", "det_a", "det_l", "z", "grid", "csi")
EXPORT_FIELDS = ("grid", "csi")


class MassiveBatch:
    """Device tables and per-drop buffers for launches of up to `chunk` drops.

    The received per-antenna grid (``buffers["grid"]``, [drop][words][nr][active]) and the
    ideal-CSI factor (``buffers["csi"]``) are diagnostic exports: with ``keep_grid=True`` every
    run writes them, with ``keep_grid=None`` (default) the runs that export equalized symbols
    do, and with ``False`` none.  They are allocated on first use."""

    def __init__(self, cfg: MassiveConfig, chunk: int = 32, device: str | torch.device = "cuda",
                 keep_grid: bool | None = None) -> None:
        self.cfg = cfg
        self.chunk = int(chunk)
        self.device = torch.device(device)
        self.keep_grid = keep_grid
        lib = _lib.load()
        self.c_link = cfg.c_link()
        sizes = np.zeros(9, dtype=np.int64)
        _lib.check(lib.hpx_massive_sizes(ctypes.byref(self.c_link), sizes.ctypes.data))
        self.sizes = sizes
        dev = self.device
        self.buffers = {}
        for name, count in zip(WORK_FIELDS, sizes):
            if name in EXPORT_FIELDS:
                continue
            dt = torch.float32 if name == "det_l" else torch.complex64
            self.buffers[name] = torch.empty((self.chunk, int(count)), dtype=dt, device=dev)
        lk = cfg.link()
        self.delays = torch.as_tensor(lk.delay_samples, dtype=torch.int32, device=dev)
        self.path_gains = torch.as_tensor(lk.path_gains, dtype=torch.float64, device=dev).contiguous()
        self.corr_rx = torch.as_tensor(np.ascontiguousarray(cfg.corr_rx), device=dev)

    def _work(self, export: bool) -> _lib.HpxMassiveWork:
        if export and "grid" not in self.buffers:
            for name in EXPORT_FIELDS:
                count = int(self.sizes[WORK_FIELDS.index(name)])
                self.buffers[name] = torch.empty((self.chunk, count), dtype=torch.complex64, device=self.device)
        return _lib.HpxMassiveWork(*[self.buffers[n].data_ptr() if (export or n not in EXPORT_FIELDS) else None
                                     for n in WORK_FIELDS])

    def run(self, seed: int, first_drop: int, noise_power: torch.Tensor, errors: torch.Tensor,
            equalized: torch.Tensor | None = None) -> None:
        """Drops first_drop .. first_drop + len(noise_power) (at most `chunk`); errors[d] += bit errors."""

        n = int(noise_power.numel())
        if n > self.chunk:
            raise ValueError(f"at most {self.chunk} drops per launch")
        c = self.cfg
        if noise_power.dtype != torch.float64:
            raise ValueError("noise_power must be float64")
        if errors.dtype != torch.int32 or errors.numel() < n:
            raise ValueError("errors must be an int32 tensor with one entry per drop")
        if equalized is not None and (equalized.dtype != torch.complex64 or
                                      equalized.numel() < n * c.num_users * c.words * c.active):
            raise ValueError("equalized must be complex64 [drops][users][words][active]")
        export = self.keep_grid if self.keep_grid is not None else equalized is not None
        work = self._work(bool(export))
        _lib.check(_lib.load().hpx_massive_run(
            ctypes.byref(self.c_link), _lib.ptr(self.delays), _lib.ptr(self.path_gains), _lib.ptr(self.corr_rx), None,
            ctypes.c_uint64(seed & (2**64 - 1)), int(first_drop), n, _lib.ptr(noise_power), ctypes.byref(work),
            _lib.ptr(errors), _lib.ptr(equalized), _lib.stream_handle()))
