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
from enum import IntEnum

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
        self._f32_error = None
        if self.lds_bytes <= 0:
            msg = _lib.last_error()
            # a single-antenna link beyond the float32 path's Taylor taps may still run in float64
            if not ("Doppler rate" in msg and link.num_rx == link.num_tx == 1):
                raise ValueError(msg)
            self._f32_error = msg
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
        if self._f32_error is not None:
            raise ValueError(self._f32_error)
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
            else:
                # the launch's Taylor blocks are sized for the link's Doppler: one host read per launch
                bound = torch.tensor([lk.doppler, lk.doppler if lk.los_doppler is None else lk.los_doppler],
                                     dtype=torch.float64, device=doppler.device)
                if n and bool((doppler.reshape(n, 2).abs() > bound.abs() * (1 + 1e-12)).any()):
                    raise ValueError("per-drop Doppler must be (num_drops, 2) and bounded by the link's Doppler")
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

    def run_f64(self, seed: int, first_drop: int, noise_power: torch.Tensor, errors: torch.Tensor,
                equalized: torch.Tensor | None = None) -> None:
        """The float64 counterpart of ``run`` for single-antenna links of up to 256 subcarriers
        without a frame or antenna correlation (``hpx_run_batch_f64``): the same drops with every
        value a double and the taps summed exactly per sample.  equalized: optional complex128
        (drops, words * N) in grid order."""

        lk = self.link
        n = int(noise_power.numel())
        if errors.numel() < n or errors.dtype != torch.int32:
            raise ValueError("errors must be an int32 tensor with one entry per drop")
        if noise_power.dtype != torch.float64:
            raise ValueError("noise_power must be float64")
        if self.c_frame is not None or self.corr_rx is not None:
            raise ValueError("the float64 path takes links without a frame or antenna correlation")
        per_drop = lk.words * lk.num_subcarriers
        if equalized is not None and (equalized.dtype != torch.complex128 or not equalized.is_contiguous()
                                      or equalized.numel() < n * per_drop):
            raise ValueError(f"equalized must be a contiguous complex128 tensor with {per_drop} entries per drop")
        lib = _lib.load()
        _lib.check(lib.hpx_run_batch_f64(ctypes.byref(self.c_link), _lib.ptr(self.delays), _lib.ptr(self.path_gains),
                                         ctypes.c_uint64(seed & (2**64 - 1)), int(first_drop), n,
                                         _lib.ptr(noise_power), _lib.ptr(errors), _lib.ptr(equalized),
                                         _lib.stream_handle()))


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
        self._mode = ScenarioMode.DEFAULT
        self._recorder = None
        self._replay = None
        self._replay_file = None
        self._drop_counter = 0
        self.replay_mismatches = 0
        self._num_samples = 100  # Simulation default (simulation.py:260-264)
        self._snr_db: tuple | None = None
        self._evaluators: list = []

    # ---------------------------------------------------------------- the reference's grid form
    @property
    def num_samples(self) -> int:
        """Drops per grid section (``MonteCarlo.num_samples``, ``monte_carlo.py:587-605``)."""

        return self._num_samples

    @num_samples.setter
    def num_samples(self, value: int) -> None:
        if value < 1:
            raise ValueError("Number of samples must be greater than zero")
        self._num_samples = int(value)

    def new_dimension(self, dimension: str, sample_points, *args, **kwargs):
        """``MonteCarlo.new_dimension`` (``monte_carlo.py:503-534``) for the dimension this path
        sweeps, the receiver's ``noise_level`` (``snr`` is accepted as well): sample points are SNRs as
        power ratios, or logarithmic values carrying ``value_db`` (the reference's ``dB(...)``).
        Further arguments (the devices the dimension applies to) are accepted and not needed."""

        if dimension not in ("noise_level", "snr"):
            raise ValueError(f"only the noise level (SNR) dimension runs on the batched path, not '{dimension}'")
        pts = list(sample_points)
        if not pts:
            raise ValueError("a dimension needs at least one sample point")
        snr = []
        for v in pts:
            if hasattr(v, "value_db"):
                snr.append(float(v.value_db))
            elif float(v) > 0.0:
                snr.append(10.0 * float(np.log10(float(v))))
            else:
                raise ValueError("SNR sample points must be positive power ratios")
        self._snr_db = tuple(snr)
        from .campaign import snr_dimension

        return snr_dimension(self._snr_db)

    def add_evaluator(self, evaluator) -> None:
        """``MonteCarlo.add_evaluator`` (``monte_carlo.py:572-580``): the bit error evaluator."""

        if getattr(evaluator, "abbreviation", None) != "BER":
            raise ValueError("the batched drop path evaluates bit errors (BitErrorEvaluator)")
        self._evaluators.append(evaluator)

    # ---------------------------------------------------------------- record / replay
    @property
    def mode(self) -> "ScenarioMode":
        return self._mode

    def record(self, file, campaign: str | None = None, overwrite: bool = False, normals: bool = False) -> None:
        """Record every following drop (``Scenario.record``, ``hermespy/core/scenario.py:368-439``):
        drop index, noise power, errors, and with ``normals`` the realisation normals."""

        from .record import DropRecorder

        if self._mode != ScenarioMode.DEFAULT:
            raise RuntimeError("Initialize a recording is only possible in default mode. "
                               "Please stop before starting a new recording.")
        self._recorder = DropRecorder(file, self.link, self.seed, campaign, overwrite, normals)
        self._mode = ScenarioMode.RECORD
        self._drop_counter = 0

    def replay(self, file=None, campaign: str | None = None, drop_offset: int = 0) -> int:
        """Replay recorded drops on the following :meth:`run` calls (``Scenario.replay``,
        ``scenario.py:441-506``); returns the number of drops to replay."""

        from .record import DropReplay

        if file is None:
            if self._replay_file is None:
                raise ValueError("Initial replay requires a valid file path argument")
            file = self._replay_file
        rp = DropReplay(file, campaign, drop_offset)
        self.stop()
        self._replay, self._replay_file = rp, file
        self._drop_counter = 0
        self._mode = ScenarioMode.REPLAY
        return rp.num_drops

    @classmethod
    def Replay(cls, file, campaign: str | None = None, drop_offset: int = 0, device: str = "cuda") -> tuple:
        """(simulation in replay mode, number of recorded drops) from a recording alone
        (``Scenario.Replay``, ``scenario.py:508-548``)."""

        from .record import DropReplay

        rp = DropReplay(file, campaign)
        link, seed = rp.link, rp.seed
        rp.close()
        sim = cls(link, seed, device)
        return sim, sim.replay(file, campaign, drop_offset)

    def stop(self) -> None:
        """End a recording or replay (``scenario.py:550-566``)."""

        if self._recorder is not None:
            self._recorder.close()
            self._recorder = None
        if self._replay is not None:
            self._replay.close()
            self._replay = None
        self._drop_counter = 0
        self._mode = ScenarioMode.DEFAULT

    def _record_launch(self, drop_index, noise_power, errors, section=None) -> None:
        rec = self._recorder
        normals = None
        if rec.normals:
            n = int(np.asarray(drop_index).size)
            normals = torch.empty((n, self.link.num_normals), dtype=torch.float64, device=self.device)
            ids = torch.as_tensor(np.asarray(drop_index, dtype=np.uint32).view(np.int32), device=self.device)
            _lib.check(_lib.load().hpx_realization_normals_batch(
                ctypes.c_uint64(self.seed & (2**64 - 1)), _lib.ptr(ids), 0, n, self.link.num_normals,
                _lib.ptr(normals), _lib.stream_handle()))
        rec.record(drop_index, noise_power, errors, section=section, normals=normals)

    # ---------------------------------------------------------------- campaigns
    def noise_powers(self, first_drop: int, count: int, snr_db) -> torch.Tensor:
        snr = np.asarray(snr_db, dtype=np.float64)
        idx = (first_drop + np.arange(count)) % snr.size
        return torch.as_tensor(self.link.noise_power(snr[idx]), dtype=torch.float64, device=self.device)

    def simulate(self, num_samples: int, evaluator=None, snr_db=None, max_drops_per_launch: int = 16384):
        """``Simulation.run()`` with the reference's confidence-based stopping per SNR section
        (``hermespy/core/pymonte/scalar.py:100-125``) on device counters; returns a
        :class:`hermespy_amd.campaign.MonteCarloResult`.  Sample k of SNR section s is drop
        k * len(snr_db) + s, the drop :meth:`run` simulates at that index.  In record mode every
        launch is recorded with its sections."""

        from .campaign import run_campaign

        if self._mode == ScenarioMode.REPLAY:
            raise RuntimeError("campaigns with early stopping generate drops; use run() to replay")
        snr = np.asarray(self.link.snr_db if snr_db is None else snr_db, dtype=np.float64)
        hook = None
        if self._mode == ScenarioMode.RECORD:
            noise = self.link.noise_power(snr)

            def hook(launch, errors):
                self._record_launch(launch.drop_index, noise[launch.sections], errors, launch.sections)

        return run_campaign(self.batched, self.seed, snr, num_samples, evaluator, max_drops_per_launch, hook)

    def run(self, num_drops: int | None = None, snr_db=None, drops_per_launch: int = 4096,
            first_drop: int = 0) -> BatchResult:
        """Without arguments (default mode): ``Simulation.run()``, the campaign over the grid set by
        :meth:`new_dimension` (the link's SNR points otherwise), ``num_samples`` drops per section
        with the stopping rule of the evaluator added by :meth:`add_evaluator`, returning a
        ``MonteCarloResult``.  Otherwise num_drops drops, drop i at SNR point i % len(snr_db).  In record mode the drops are
        recorded; in replay mode the next num_drops recorded drops (all remaining ones if None)
        are simulated again from their recorded inputs, and ``replay_mismatches`` counts drops
        whose error count differs from the recorded one."""

        if num_drops is None and snr_db is None and self._mode == ScenarioMode.DEFAULT:
            # Simulation.run() (simulation.py:408-457): the configured grid, num_samples drops per
            # section with the evaluator's stopping rule, as a MonteCarloResult
            return self.simulate(self.num_samples, self._evaluators[0] if self._evaluators else None,
                                 self._snr_db)
        snr = np.asarray(self.link.snr_db if snr_db is None else snr_db, dtype=np.float64)
        if self._mode == ScenarioMode.REPLAY:
            return self._run_replay(num_drops, snr, drops_per_launch)
        if num_drops is None:
            raise ValueError("give the number of drops")
        errors = torch.empty(num_drops, dtype=torch.int32, device=self.device)
        for start in range(0, num_drops, drops_per_launch):
            cnt = min(drops_per_launch, num_drops - start)
            npw = self.noise_powers(first_drop + start, cnt, snr)
            self.batched.run(self.seed, first_drop + start, npw, errors[start:start + cnt])
            if self._mode == ScenarioMode.RECORD:
                ids = np.arange(first_drop + start, first_drop + start + cnt, dtype=np.uint32)
                self._record_launch(ids, npw, errors[start:start + cnt],
                                    ((first_drop + start + np.arange(cnt)) % snr.size).astype(np.int32))
        err = errors.cpu().numpy().astype(np.int64)
        idx = (first_drop + np.arange(num_drops)) % snr.size
        per_err = np.bincount(idx, weights=err, minlength=snr.size).astype(np.int64)
        per_drops = np.bincount(idx, minlength=snr.size).astype(np.int64)
        return BatchResult(snr, per_err, per_drops, self.link.bits_per_drop)

    def replay_realizations(self, file, campaign: str | None = None, snr_db=None,
                            drops_per_launch: int = 4096) -> BatchResult:
        """Drops over the channel realisations of a recording in the reference's layout
        (``drops/NNNN/channel_realizations/0000``, see :func:`hermespy_amd.record.write_reference_drops`):
        drop k runs over recorded realisation k (its ``scalar_samples``), with Philox drop k's bits
        and noise at SNR point k % len(snr_db)."""

        from .record import read_reference_drops

        z = read_reference_drops(file, self.link, campaign)
        snr = np.asarray(self.link.snr_db if snr_db is None else snr_db, dtype=np.float64)
        n = z.shape[0]
        errors = torch.empty(n, dtype=torch.int32, device=self.device)
        for start in range(0, n, drops_per_launch):
            cnt = min(drops_per_launch, n - start)
            normals = torch.as_tensor(np.ascontiguousarray(z[start:start + cnt]), device=self.device)
            self.batched.run(self.seed, start, self.noise_powers(start, cnt, snr), errors[start:start + cnt],
                             normals=normals)
        err = errors.cpu().numpy().astype(np.int64)
        idx = np.arange(n) % snr.size
        return BatchResult(snr, np.bincount(idx, weights=err, minlength=snr.size).astype(np.int64),
                           np.bincount(idx, minlength=snr.size).astype(np.int64), self.link.bits_per_drop)

    def _run_replay(self, num_drops, snr, drops_per_launch) -> BatchResult:
        rp = self._replay
        left = rp.num_drops - self._drop_counter
        num = left if num_drops is None else int(num_drops)
        if num > left:
            raise ValueError(f"only {left} recorded drops are left to replay")
        per_err = np.zeros(snr.size, dtype=np.int64)
        per_drops = np.zeros(snr.size, dtype=np.int64)
        errors = torch.empty(min(num, drops_per_launch), dtype=torch.int32, device=self.device)
        dev = self.device
        for start in range(0, num, drops_per_launch):
            cnt = min(drops_per_launch, num - start)
            cols = rp.read(self._drop_counter, cnt)
            ids = torch.as_tensor(cols["drop_index"].view(np.int32), device=dev)
            npw = torch.as_tensor(cols["noise_power"], device=dev)
            dop = torch.as_tensor(np.ascontiguousarray(cols["doppler"]), device=dev) if "doppler" in cols else None
            z = torch.as_tensor(np.ascontiguousarray(cols["normals"]), device=dev) if "normals" in cols else None
            self.batched.run(self.seed, 0, npw, errors[:cnt], doppler=dop, drop_index=ids, normals=z)
            err = errors[:cnt].cpu().numpy()
            self.replay_mismatches += int(np.count_nonzero(err != cols["errors"]))
            sec = cols["section"] if "section" in cols else cols["drop_index"].astype(np.int64) % snr.size
            per_err += np.bincount(sec, weights=err.astype(np.int64), minlength=snr.size).astype(np.int64)
            per_drops += np.bincount(sec, minlength=snr.size).astype(np.int64)
            self._drop_counter += cnt
        return BatchResult(snr, per_err, per_drops, self.link.bits_per_drop)


class ScenarioMode(IntEnum):
    """``hermespy/core/scenario.py:49-70``."""

    DEFAULT = 0
    RECORD = 1
    REPLAY = 2


class SingleDropLink:
    """The plain GPU path: one drop at a time through the stage entry points, written against
    the reference's typed interfaces.

    Per drop: realisation normals (device) -> ``realization.sample(transmitter, receiver)``
    (the reference's ``MultipathFadingRealization._sample``, on the host; the ideal estimator
    caches the sample through the channel's sample hook) -> bits -> ``map`` / ``place`` /
    ``modulate`` -> ``sample.propagate(Signal)`` -> AWGN -> ``demodulate`` ->
    ``estimate_channel(Symbols) -> StatedSymbols`` -> ``pick`` -> ``equalize_channel`` ->
    ``unmap`` -> ``BitErrorEvaluator.evaluate``, each a separate launch, consuming the same
    Philox streams as the fused batch kernel.
    """

    def __init__(self, link: LinkConfig, seed: int = 20261202, device: str = "cuda") -> None:
        from .channel import CustomAntennaCorrelation, MultipathFadingChannel
        from .core import SimulatedDevice
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
        self.transmitter = SimulatedDevice(link.num_tx, sampling_rate=link.bandwidth, power=link.device_power)
        self.receiver = SimulatedDevice(link.num_rx, sampling_rate=link.bandwidth)
        self.estimator = OFDMIdealChannelEstimation(self.channel, self.transmitter, self.receiver, self.waveform)
        self.equalizer = (ZeroForcingChannelEqualization() if link.detector == "zf" else MMSEChannelEqualization())
        self.ber = BitErrorEvaluator()

    def run_drop(self, drop: int, snr_db: float, errors: torch.Tensor | None = None,
                 return_equalized: bool = False):
        from .channel import MultipathFadingRealization
        from .core import Signal
        from .modem import Symbols

        lk, lib, st = self.link, _lib.load(), _lib.stream_handle()
        dev = self.device
        # channel realisation: the drop's normals, sampled on the host like the reference
        z = torch.empty(self.channel.num_scalars, dtype=torch.float64, device=dev)
        _lib.check(lib.hpx_realization_normals(self.seed, drop, z.numel(), _lib.ptr(z), st))
        sample = MultipathFadingRealization(z.cpu().numpy(), self.channel).sample(self.transmitter, self.receiver,
                                                                                  device=dev)
        # transmitter
        nbits = self.waveform.bits_per_frame()
        bits = torch.empty((lk.num_tx, nbits), dtype=torch.uint8, device=dev)
        for t in range(lk.num_tx):
            _lib.check(lib.hpx_stream_bits(self.seed, drop, _lib.STREAM_BITS + t, nbits, _lib.ptr(bits[t]), st))
        symbols = self.waveform.map(bits.reshape(-1)).raw.reshape(lk.num_tx, -1)
        x = self.waveform.modulate(self.waveform.place(Symbols(symbols)))
        if self.transmitter.power != 1.0:
            x = x * float(np.sqrt(self.transmitter.power))  # simulated_device.py:1688-1692
        # channel and noise
        received = sample.propagate(Signal(x, lk.bandwidth))
        y = received.samples
        p = float(lk.noise_power(snr_db))
        _lib.check(lib.hpx_awgn(_lib.ptr(y), y.shape[0], y.shape[1], y.shape[1], self.seed, drop, p, st))
        # receiver
        stated = self.estimator.estimate_channel(self.waveform.demodulate(y), lk.bandwidth, 1)
        if hasattr(self.equalizer, "noise_variance"):
            self.equalizer.noise_variance = p
        equalized = self.equalizer.equalize_channel(self.waveform.pick(stated))
        rx_bits = self.waveform.unmap(equalized)
        acc = self.ber.evaluate(bits, rx_bits, errors).errors
        return (acc, equalized.raw.reshape(lk.num_tx, -1)) if return_equalized else acc
