"""Batched Monte Carlo drops on the GPU: the ``Simulation.run()`` drop loop of
the reference (``hermespy/simulation/simulation.py:408-457`` driving
``hermespy/core/pymonte/actors.py:340-452``) as fused launches of
``hpx_run_batch``.

Every drop is independent (the reference realises a fresh channel per drop,
``simulation.py:69-70``), so a launch simulates ``drops_per_launch`` drops in
parallel, one workgroup each, and returns their bit error counts.  The
campaign BER of an SNR point is the mean of the per-drop BERs, as
``BitErrorEvaluation.artifact`` / ``ScalarEvaluationResult`` aggregate it
(``hermespy/modem/evaluators.py:231-233``, ``hermespy/core/pymonte/scalar.py:100-125``).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .configs import LinkConfig


@dataclass
class BatchResult:
    snr_db: np.ndarray  # per SNR point
    errors: np.ndarray  # per SNR point, summed bit errors
    drops: np.ndarray  # per SNR point
    bits_per_drop: int

    @property
    def ber(self) -> np.ndarray:
        return self.errors / np.maximum(self.drops * self.bits_per_drop, 1)


class BatchedLink:
    """Device-resident parameters of one link for ``hpx_run_batch``."""

    def __init__(self, link: LinkConfig, device: str | torch.device = "cuda") -> None:
        self.link = link
        self.device = torch.device(device)
        lib = _lib.load()
        det = {"zf": _lib.HPX_DETECT_ZF, "mmse": _lib.HPX_DETECT_MMSE}.get(link.detector)
        if det is None:
            raise ValueError(f"unknown detector {link.detector}")
        self.c_link = _lib.HpxLink(
            link.num_rx, link.num_tx, link.num_subcarriers, link.words, link.prefix, link.modulation_order,
            link.num_paths, link.num_sinusoids, link.max_delay, det, int(bool(link.dc_suppression)),
            float(link.bandwidth), float(link.doppler),
            float(link.doppler if link.los_doppler is None else link.los_doppler), float(link.gain),
            float(link.device_power), int(link.workgroup_size))
        self.lds_bytes = int(lib.hpx_run_batch_lds_bytes(ctypes.byref(self.c_link)))
        if self.lds_bytes <= 0:
            # surface the validation message
            _lib.check(_lib.HPX_ERR_UNSUPPORTED)
        dev = self.device
        self.delays = torch.as_tensor(link.delay_samples, dtype=torch.int32, device=dev)
        self.path_gains = torch.as_tensor(link.path_gains, dtype=torch.float64, device=dev).contiguous()
        if link.corr_rx is not None:
            rrx = np.asarray(link.corr_rx, dtype=np.complex128)[: link.num_rx, : link.num_rx]
            rtx = np.asarray(link.corr_tx, dtype=np.complex128)[: link.num_tx, : link.num_tx]
            self.corr_rx = torch.as_tensor(np.ascontiguousarray(rrx), device=dev)
            self.corr_tx = torch.as_tensor(np.ascontiguousarray(rtx), device=dev)
        else:
            self.corr_rx = self.corr_tx = None
        self.c_frame = None
        if link.frame_mask is not None:
            mask = link.frame_mask
            data, ref = mask == 1, mask == 0
            el = np.full(mask.shape, _lib.ELEMENT_NULL, dtype=np.int64)
            el[data] = (np.cumsum(data, axis=1) - 1)[data]  # data index within the word
            el[ref] = -1 - np.arange(int(ref.sum()))  # reference elements in frame (row-major) order
            per_word = data.sum(axis=1)
            word_data = np.concatenate([[0], np.cumsum(per_word)]).astype(np.int32)
            self.elements = torch.as_tensor(el.astype(np.int32), device=dev)
            self.word_data = torch.as_tensor(word_data, device=dev)
            self.pilots = (torch.as_tensor(np.asarray(link.pilots, dtype=np.complex64), device=dev)
                           if link.num_references else None)
            self.c_frame = _lib.HpxFrame(_lib.ptr(self.elements), _lib.ptr(self.word_data), _lib.ptr(self.pilots),
                                         link.num_references, int(word_data[-1]), int(per_word.max()))

    def run(self, seed: int, first_drop: int, noise_power: torch.Tensor, errors: torch.Tensor,
            equalized: torch.Tensor | None = None, stream: int | None = None, doppler=None,
            grid: torch.Tensor | None = None, csi: torch.Tensor | None = None,
            spatial: torch.Tensor | None = None, drop_index: torch.Tensor | None = None,
            normals: torch.Tensor | None = None) -> None:
        """Launch ``noise_power.numel()`` drops (asynchronous on the current stream).

        drop_index: int32 tensor holding the uint32 Philox drop counter of each drop (bit
        pattern), instead of first_drop + i.  normals: float64 (num_drops, num_normals)
        realisation normals (recorded ``scalar_samples``) replacing the Philox draw.

        doppler: per-drop (NLOS, LOS) doppler_frequency, (num_drops, 2) float64 (host array or device
        tensor), bounded by the link's Doppler.  grid / csi / spatial: optional complex64 exports of the
        received grid (drops, Nrx, words, N), the ideal-CSI factor c (drops, words, N) and the spatial
        response S (drops, Nrx, Ntx); the ideal CSI of element (w, k) is S c[w, k]."""

        lk = self.link
        n = int(noise_power.numel())
        if errors.numel() < n or errors.dtype != torch.int32:
            raise ValueError("errors must be an int32 tensor with one entry per drop")
        if noise_power.dtype != torch.float64:
            raise ValueError("noise_power must be float64")
        sizes = {"equalized": (equalized, lk.num_tx * lk.num_data), "grid": (grid, lk.num_rx * lk.words * lk.num_subcarriers),
                 "csi": (csi, lk.words * lk.num_subcarriers), "spatial": (spatial, lk.num_rx * lk.num_tx)}
        for name, (t, per_drop) in sizes.items():
            if t is not None and (t.dtype != torch.complex64 or not t.is_contiguous() or t.numel() < n * per_drop):
                raise ValueError(f"{name} must be a contiguous complex64 tensor with {per_drop} entries per drop")
        if doppler is not None:
            if isinstance(doppler, np.ndarray):
                bound = np.array([lk.doppler, lk.doppler if lk.los_doppler is None else lk.los_doppler])
                if doppler.shape != (n, 2) or np.any(np.abs(doppler) > bound * (1 + 1e-12)):
                    raise ValueError("per-drop Doppler must be (num_drops, 2) and bounded by the link's Doppler")
                doppler = torch.as_tensor(np.ascontiguousarray(doppler, dtype=np.float64), device=self.device)
            elif doppler.dtype != torch.float64 or doppler.numel() != 2 * n:
                raise ValueError("per-drop Doppler must be a float64 (num_drops, 2) tensor")
            self._doppler = doppler  # held until the next launch
        if drop_index is not None and (drop_index.dtype != torch.int32 or drop_index.numel() < n
                                       or not drop_index.is_contiguous()):
            raise ValueError("drop_index must be a contiguous int32 tensor (uint32 bit patterns) per drop")
        if normals is not None and (normals.dtype != torch.float64 or not normals.is_contiguous()
                                    or normals.numel() < n * lk.num_normals):
            raise ValueError(f"normals must be a contiguous float64 tensor of {lk.num_normals} per drop")
        lib = _lib.load()
        out = _lib.HpxBatchOut(_lib.ptr(errors), _lib.ptr(equalized), _lib.ptr(grid), _lib.ptr(csi), _lib.ptr(spatial))
        inputs = _lib.HpxDropInputs(_lib.ptr(noise_power), _lib.ptr(doppler), _lib.ptr(drop_index), _lib.ptr(normals))
        st = lib.hpx_run_batch_drops(
            ctypes.byref(self.c_link), ctypes.byref(self.c_frame) if self.c_frame is not None else None,
            _lib.ptr(self.delays), _lib.ptr(self.path_gains), _lib.ptr(self.corr_rx), _lib.ptr(self.corr_tx),
            ctypes.byref(inputs), ctypes.c_uint64(seed & (2**64 - 1)), int(first_drop), n,
            ctypes.byref(out), _lib.stream_handle() if stream is None else stream)
        _lib.check(st)


class BatchedSimulation:
    """Monte Carlo BER campaign over an SNR grid, ``drops_per_launch`` drops per fused launch.

    Drop ``i`` of the campaign uses SNR point ``i % len(snr_db)`` and Philox
    counter ``i``, so any drop can be re-simulated on its own.
    """

    def __init__(self, link: LinkConfig, seed: int = 20261202, device: str = "cuda") -> None:
        self.link = link
        self.seed = int(seed)
        self.batched = BatchedLink(link, device)
        self.device = self.batched.device

    def noise_powers(self, first_drop: int, count: int, snr_db) -> torch.Tensor:
        snr = np.asarray(snr_db, dtype=np.float64)
        idx = (first_drop + np.arange(count)) % snr.size
        return torch.as_tensor(self.link.noise_power(snr[idx]), dtype=torch.float64, device=self.device)

    def simulate(self, num_samples: int, evaluator=None, snr_db=None, max_drops_per_launch: int = 16384,
                 record=None):
        """``Simulation.run()`` with the reference's confidence-based stopping per SNR section
        (``hermespy/core/pymonte/scalar.py:100-125``) on device counters; returns a
        :class:`hermespy_amd.campaign.MonteCarloResult`.  Sample k of SNR section s is drop
        k * len(snr_db) + s, the drop :meth:`run` simulates at that index."""

        from .campaign import run_campaign

        snr = self.link.snr_db if snr_db is None else snr_db
        return run_campaign(self.batched, self.seed, snr, num_samples, evaluator, max_drops_per_launch, record)

    def run(self, num_drops: int, snr_db=None, drops_per_launch: int = 4096, first_drop: int = 0) -> BatchResult:
        snr = np.asarray(self.link.snr_db if snr_db is None else snr_db, dtype=np.float64)
        errors = torch.empty(num_drops, dtype=torch.int32, device=self.device)
        for start in range(0, num_drops, drops_per_launch):
            cnt = min(drops_per_launch, num_drops - start)
            npw = self.noise_powers(first_drop + start, cnt, snr)
            self.batched.run(self.seed, first_drop + start, npw, errors[start:start + cnt])
        err = errors.cpu().numpy().astype(np.int64)
        idx = (first_drop + np.arange(num_drops)) % snr.size
        per_err = np.bincount(idx, weights=err, minlength=snr.size).astype(np.int64)
        per_drops = np.bincount(idx, minlength=snr.size).astype(np.int64)
        return BatchResult(snr, per_err, per_drops, self.link.bits_per_drop)


class SingleDropLink:
    """The plain GPU path: one drop at a time through the stage entry points.

    Per drop: realisation normals (device) -> host realisation sample (the
    reference's ``MultipathFadingRealization._sample``) -> bits -> mapping ->
    OFDM modulation -> TDL propagation -> AWGN -> demodulation -> ideal CSI ->
    equalisation -> hard demapping -> bit error count, each a separate launch,
    consuming the same Philox streams as the fused batch kernel.
    """

    def __init__(self, link: LinkConfig, seed: int = 20261202, device: str = "cuda") -> None:
        from .channel import CustomAntennaCorrelation, MultipathFadingChannel
        from .modem import (BitErrorEvaluator, GridElement, GridResource, MMSEChannelEqualization, OFDMIdealChannelEstimation,
                            OFDMWaveform, SymbolSection, ZeroForcingChannelEqualization)

        self.link = link
        self.seed = int(seed)
        self.device = torch.device(device)
        corr = CustomAntennaCorrelation(link.corr_rx) if link.corr_rx is not None else None
        if link.corr_rx is not None and not np.allclose(link.corr_rx[:link.num_tx, :link.num_tx], link.corr_tx):
            raise ValueError("SingleDropLink models one correlation matrix for both ends")
        self.channel = MultipathFadingChannel(link.delays, link.power_profile, link.rice_factors,
                                              num_sinusoids=link.num_sinusoids, doppler_frequency=link.doppler,
                                              los_doppler_frequency=link.los_doppler, antenna_correlation=corr,
                                              gain=link.gain, los_angle=0.0)
        res = GridResource(1, "CYCLIC", link.prefix_ratio, [GridElement("DATA", link.num_subcarriers)])
        self.waveform = OFDMWaveform([res], [SymbolSection(link.words, [0])], link.num_subcarriers,
                                     link.dc_suppression, link.modulation_order)
        self.estimator = OFDMIdealChannelEstimation(self.waveform)
        self.zf = ZeroForcingChannelEqualization()
        self.mmse = MMSEChannelEqualization()
        self.ber = BitErrorEvaluator()

    def run_drop(self, drop: int, snr_db: float, errors: torch.Tensor | None = None,
                 return_equalized: bool = False):
        from .channel import MultipathFadingRealization

        lk, lib, st = self.link, _lib.load(), _lib.stream_handle()
        dev = self.device
        # channel realisation: the drop's normals, sampled on the host like the reference
        z = torch.empty(self.channel.num_scalars, dtype=torch.float64, device=dev)
        _lib.check(lib.hpx_realization_normals(self.seed, drop, z.numel(), _lib.ptr(z), st))
        sample = MultipathFadingRealization(z.cpu().numpy(), self.channel).sample(lk.num_tx, lk.num_rx, lk.bandwidth, dev)
        # transmitter
        nbits = self.waveform.bits_per_frame()
        bits = torch.empty((lk.num_tx, nbits), dtype=torch.uint8, device=dev)
        for t in range(lk.num_tx):
            _lib.check(lib.hpx_stream_bits(self.seed, drop, _lib.STREAM_BITS + t, nbits, _lib.ptr(bits[t]), st))
        symbols = self.waveform.map(bits.reshape(-1)).reshape(lk.num_tx, -1)
        x = self.waveform.modulate(self.waveform.place(symbols))
        if lk.device_power != 1.0:
            x = x * float(np.sqrt(lk.device_power))
        # channel and noise
        y = sample.propagate(x)
        p = float(lk.noise_power(snr_db))
        _lib.check(lib.hpx_awgn(_lib.ptr(y), y.shape[0], y.shape[1], y.shape[1], self.seed, drop, p, st))
        # receiver
        grid = self.waveform.demodulate(y)
        h = self.estimator.estimate_channel(sample)
        yd = self.waveform.pick(grid)
        hd = self.waveform.pick(h)
        if lk.detector == "zf":
            xh = self.zf.equalize_channel(yd, hd)
        else:
            self.mmse.noise_variance = p
            xh = self.mmse.equalize_channel(yd, hd)
        rx_bits = self.waveform.unmap(xh)
        acc = self.ber.evaluate(bits, rx_bits, errors)
        return (acc, xh) if return_equalized else acc
