"""Heterogeneous batches: thousands of drops with different antenna counts, FFT
sizes, path counts, power-delay profiles, Doppler, modulation and SNR in one
graph-captured submission.

The reference simulates one configuration at a time: every shape is its own
``Simulation`` whose ``MonteCarloActor`` loops over drops
(``hermespy/core/pymonte/actors.py:316-452``).  Here the host sorts a mixed
population into kernel classes (antennas, threads per drop, samples per
thread), writes one plan per drop (``hpx_uneven_prepare``) and replays one
launch per class from a captured graph (``hpx_uneven_run``); drop i's result
depends only on (seed, its drop index, its link), exactly as in
``hpx_run_batch``.  A per-class autotuner picks the workgroup size.  8x8 drops
at 2048 subcarriers, whose working set exceeds one CU's LDS, run in lean
classes: their Taylor table and twiddles live in a per-drop device workspace
and the ideal CSI is summed path by path per subcarrier.
"""

from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .configs import LinkConfig

MAX_CLASSES = 64
SUBCARRIER_SPACING = 30e3  # sampling rate = N x 30 kHz


@dataclass
class DropShape:
    """One drop of a mixed population."""

    antennas: int
    num_subcarriers: int
    num_paths: int
    decay_db: float
    doppler_hz: float
    modulation_order: int
    snr_db: float

    def link(self) -> LinkConfig:
        n = self.num_subcarriers
        fs = n * SUBCARRIER_SPACING
        p = 10.0 ** (-self.decay_db * np.arange(self.num_paths) / 10.0)
        return LinkConfig(self.antennas, self.antennas, n, 14, 1.0 / 8.0, self.modulation_order, fs,
                          np.arange(self.num_paths) / fs, p / p.sum(), np.zeros(self.num_paths),
                          2 * np.pi * self.doppler_hz, detector="mmse", name="uneven")


def draw_population(num_drops: int, seed: int = 20261202, exclude=()) -> list:
    """The mixed population: antennas {1, 2, 4, 8} square, N in {128, ..., 2048} with CP N/8, 5..23
    sample-spaced paths with an exponential profile of 1..6 dB per tap, Doppler 10..900 Hz,
    QPSK / 16-QAM / 64-QAM, SNR from 0..30 dB in 3 dB steps.  Shapes in `exclude` (antennas, N)
    are redrawn."""

    rng = np.random.default_rng(seed)
    out = []
    while len(out) < num_drops:
        a = int(rng.choice([1, 2, 4, 8]))
        n = int(rng.choice([128, 256, 512, 1024, 2048]))
        shape = DropShape(a, n, int(rng.integers(5, 24)), float(rng.uniform(1.0, 6.0)), float(rng.uniform(10.0, 900.0)),
                          int(rng.choice([4, 16, 64])), float(rng.choice(np.arange(0.0, 31.0, 3.0))))
        if (a, n) in exclude:
            continue
        out.append(shape)
    return out


def _c_link(cfg: LinkConfig, workgroup_size: int = 0, fft_values: int = 0) -> _lib.HpxLink:
    return _lib.HpxLink(cfg.num_rx, cfg.num_tx, cfg.num_subcarriers, cfg.words, cfg.prefix, cfg.modulation_order,
                        cfg.num_paths, cfg.num_sinusoids, cfg.max_delay, _lib.HPX_DETECT_MMSE,
                        int(bool(cfg.dc_suppression)), float(cfg.bandwidth), float(cfg.doppler),
                        float(cfg.doppler if cfg.los_doppler is None else cfg.los_doppler), float(cfg.gain),
                        float(cfg.device_power), int(workgroup_size), int(fft_values))


def _choice(value) -> tuple:
    """A class's launch configuration: a workgroup size, or (workgroup size, FFT values per thread);
    0 stands for the library default."""

    if isinstance(value, (tuple, list)):
        return int(value[0]), int(value[1])
    return int(value), 0


def schedule_launches(classes, num_classes: int, cost, num_streams: int) -> tuple:
    """Launch pieces per stream.  A class costlier than half a stream's share of the total is cut
    into equal drop ranges (a piece is the class's descriptor over fewer plan slots); the pieces
    then go, longest first, to the stream with the least work so far (longest-processing-time
    scheduling), so that the streams end together instead of one stream's last long class running
    alone.  Returns (pieces per stream in launch order, summed cost per stream)."""

    cost = np.asarray(cost, dtype=np.float64)
    share = float(cost[:num_classes].sum()) / (2 * num_streams)
    pieces = []
    for k in range(num_classes):
        c = classes[k]
        parts = 1 if share <= 0 else int(min(c.count, max(1, np.ceil(cost[k] / share))))
        cuts = np.rint(np.linspace(0, c.count, parts + 1)).astype(np.int64)
        for lo, hi in zip(cuts[:-1], cuts[1:]):
            if hi > lo:
                pc = _lib.HpxUnevenClass(c.first + int(lo), int(hi - lo), c.num_tx, c.threads, c.pairs, c.lean,
                                         c.lds_bytes, c.fft_values, 0)
                pieces.append((cost[k] * (hi - lo) / c.count, pc))
    pieces.sort(key=lambda t: -t[0])
    load = np.zeros(num_streams)
    lanes = [[] for _ in range(num_streams)]
    for w, pc in pieces:
        j = int(np.argmin(load))
        load[j] += w
        lanes[j].append(pc)
    return lanes, load.tolist()


class UnevenBatch:
    """Device-resident plans of one mixed population (drop k of the population is Philox drop
    first_drop + k)."""

    def __init__(self, shapes: list, seed: int = 20261202, first_drop: int = 0, device="cuda",
                 workgroup_sizes: dict | None = None, streams: int = 4, no_alloc: bool = False,
                 export_equalized: bool = False) -> None:
        """no_alloc: plan with ``hpx_uneven_prepare_noalloc`` (the lean drops' tables in the
        library's own scratch array, no workspace allocated here).  export_equalized: keep every
        drop's detected data symbols, ``equalized[i]`` (num_tx, words * N) complex64."""

        self.shapes = shapes
        self.seed = int(seed)
        self.first_drop = int(first_drop)
        self.device = torch.device(device)
        self.lib = _lib.load()
        self.links = [s.link() for s in shapes]
        n = len(shapes)
        dev = self.device
        # per-drop path tables, concatenated
        lens = np.array([c.num_paths for c in self.links])
        offs = np.concatenate([[0], np.cumsum(lens)[:-1]])
        delays = np.concatenate([c.delay_samples for c in self.links]).astype(np.int32)
        gains = np.concatenate([c.path_gains.reshape(-1) for c in self.links])
        self.delays = torch.as_tensor(delays, device=dev)
        self.gains = torch.as_tensor(gains, dtype=torch.float64, device=dev)
        self.tables = (_lib.HpxDropTables * n)()
        for i in range(n):
            self.tables[i].delays = self.delays.data_ptr() + 4 * int(offs[i])
            self.tables[i].path_gains = self.gains.data_ptr() + 8 * 2 * int(offs[i])
        # drops beyond one CU's LDS (8x8 at 2048 subcarriers) keep their Taylor and twiddle
        # tables in device workspaces
        ws = np.array([_lib.check_count(self.lib.hpx_uneven_workspace_elements(ctypes.byref(_c_link(c))))
                       for c in self.links], dtype=np.int64)
        self.no_alloc = bool(no_alloc)
        self.workspace = None
        if not self.no_alloc:
            self.workspace = torch.empty(max(1, int(ws.sum())), dtype=torch.complex64, device=dev)
            ws_offs = np.concatenate([[0], np.cumsum(ws)[:-1]])
            for i in np.flatnonzero(ws):
                self.tables[i].workspace = self.workspace.data_ptr() + 8 * int(ws_offs[i])
        self.num_lean = int(np.count_nonzero(ws))
        self.equalized = None
        if export_equalized:
            self.equalized = [torch.empty((c.num_tx, c.num_data), dtype=torch.complex64, device=dev)
                              for c in self.links]
            for i, t in enumerate(self.equalized):
                self.tables[i].equalized = t.data_ptr()
        self.noise_power = torch.as_tensor([c.noise_power(s.snr_db) for c, s in zip(self.links, shapes)],
                                           dtype=torch.float64, device=dev)
        self.errors = torch.zeros(n, dtype=torch.int32, device=dev)
        self.bits = np.array([c.bits_per_drop for c in self.links], dtype=np.int64)
        # Classes of a few hundred drops each fill a fraction of the GPU; launched on several
        # streams (default priority) they run side by side, also inside a captured graph.
        self.streams = [torch.cuda.Stream(device=dev) for _ in range(max(1, int(streams)))]
        self.prepare(workgroup_sizes or {})

    def class_key(self, i: int) -> tuple:
        return (self.links[i].num_tx, self.links[i].num_subcarriers)

    def prepare(self, workgroup_sizes: dict) -> None:
        """(Re)build the plans; workgroup_sizes maps (antennas, N) to a workgroup size or to a pair
        (workgroup size, FFT values per thread), 0 meaning the library default."""

        n = len(self.links)
        self.workgroup_sizes = dict(workgroup_sizes)
        links = (_lib.HpxLink * n)(*[_c_link(c, *_choice(self.workgroup_sizes.get(self.class_key(i), 0)))
                                     for i, c in enumerate(self.links)])
        idx = np.arange(self.first_drop, self.first_drop + n, dtype=np.uint32)
        plan_bytes = int(self.lib.hpx_uneven_plan_bytes())
        plans = np.empty(n * plan_bytes, dtype=np.uint8)
        order = np.empty(n, dtype=np.int64)
        classes = (_lib.HpxUnevenClass * MAX_CLASSES)()
        nc = _lib.c_int(0)
        prepare = self.lib.hpx_uneven_prepare_noalloc if self.no_alloc else self.lib.hpx_uneven_prepare
        _lib.check(prepare(links, self.tables, idx.ctypes.data, n, ctypes.c_uint64(self.seed),
                           _lib.ptr(self.noise_power, "noise_power"), _lib.ptr(self.errors, "errors"), plans.ctypes.data,
                           order.ctypes.data, classes, MAX_CLASSES, ctypes.byref(nc)))
        self.plans = torch.as_tensor(plans, device=self.device)
        self.order = order
        self.num_classes = nc.value
        self.classes = classes
        # class costs for the stream schedule: the population's algorithmic flops until tune()
        # replaces them with measured launch times
        flops = np.array([c.algorithmic_flops for c in self.links])
        self.class_cost = np.array([flops[order[classes[k].first:classes[k].first + classes[k].count]].sum()
                                    for k in range(self.num_classes)])
        self.schedule()
        self.graph = None

    def schedule(self) -> list:
        """Spread the class launches over the streams (schedule_launches with the classes' costs);
        returns the streams' summed costs."""

        lanes, load = schedule_launches(self.classes, self.num_classes, self.class_cost, len(self.streams))
        self.lanes = [(_lib.HpxUnevenClass * len(lane))(*lane) for lane in lanes if lane]
        self.stream_cost = load
        self.graph = None
        return load

    def run(self) -> None:
        """All classes, asynchronous on the current stream: forked over the side streams as
        scheduled (schedule()) and joined back."""

        cur = torch.cuda.current_stream(self.device)
        if len(self.streams) <= 1 or self.num_classes <= 1:
            _lib.check(self.lib.hpx_uneven_run(self.plans.data_ptr(), self.classes, self.num_classes,
                                               cur.cuda_stream))
            return
        fork = torch.cuda.Event()
        fork.record(cur)
        for st in self.streams:
            st.wait_event(fork)
        for st, lane in zip(self.streams, self.lanes):
            _lib.check(self.lib.hpx_uneven_run(self.plans.data_ptr(), lane, len(lane), st.cuda_stream))
        for st in self.streams:
            join = torch.cuda.Event()
            join.record(st)
            cur.wait_event(join)

    def run_class(self, k: int) -> None:
        _lib.check(self.lib.hpx_uneven_run(self.plans.data_ptr(), ctypes.byref(self.classes[k]), 1,
                                           _lib.stream_handle()))

    def capture(self) -> None:
        """Capture every class launch into one graph (after one eager run has set the kernels' LDS limits)."""

        self.run()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run()
        self.graph = g

    def replay(self) -> None:
        if self.graph is None:
            self.capture()
        self.graph.replay()

    def class_table(self) -> list:
        out = []
        for k in range(self.num_classes):
            c = self.classes[k]
            i = int(self.order[c.first])
            out.append({"antennas": c.num_tx, "num_subcarriers": self.links[i].num_subcarriers, "drops": c.count,
                        "threads": c.threads, "pairs": c.pairs, "lean": bool(c.lean), "lds_bytes": c.lds_bytes})
        return out

    def ber_by_snr(self) -> dict:
        err = self.errors.cpu().numpy().astype(np.int64)
        snr = np.array([s.snr_db for s in self.shapes])
        return {float(v): float(err[snr == v].sum() / self.bits[snr == v].sum()) for v in np.unique(snr)}


def workgroup_candidates(n: int) -> list:
    """Workgroup sizes of 64..512 threads with 1, 2 or 4 sample pairs per thread, and N threads
    (half of them idle in the sample-pair loops, all of them in the prologue, the transforms and
    detection), which shortens the drops of the smallest links."""

    return [t for t in (n, n // 2, n // 4, n // 8) if 64 <= t <= 512]


def fft_value_candidates(num_tx: int, threads: int) -> list:
    """The FFT values per thread and pass a class's kernel comes in (batch.hip's fft_values_for and
    alt_fft_values_for): the default, 12 at 512 threads and at most 4 antennas, else 16; and the
    alternative, 16 where the default is 12, 8 for 8 antennas, else 12."""

    maxt = 256 if threads <= 256 else 512
    default = 12 if maxt >= 512 and num_tx <= 4 else 16
    alt = 16 if default == 12 else (8 if num_tx == 8 else 12)
    return [default, alt]


def tune(batch: UnevenBatch, reps: int = 3, verbose: bool = False) -> dict:
    """Per (antennas, N) class, the launch configuration (workgroup size, i.e. sample pairs per thread,
    and FFT values per thread) with the shortest launch; re-prepares the batch."""

    keys = sorted({batch.class_key(i) for i in range(len(batch.links))})
    best = {}
    timings = {}
    for key in keys:
        cands = [(t, f) for t in workgroup_candidates(key[1]) for f in fft_value_candidates(key[0], t)]
        times = {}
        for t in cands:
            batch.prepare({**best, key: t})
            k = next(j for j in range(batch.num_classes)
                     if batch.class_key(int(batch.order[batch.classes[j].first])) == key)
            batch.run_class(k)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                batch.run_class(k)
            torch.cuda.synchronize()
            times[t] = (time.perf_counter() - t0) / reps
        best[key] = min(times, key=times.get)
        timings[key] = times
        if verbose:
            print(f"tune {key}: " + ", ".join(f"{t}: {1e3 * v:.3f} ms" for t, v in times.items()) + f" -> {best[key]}",
                  flush=True)
    batch.prepare(best)
    # the chosen configurations' launch times become the classes' costs in the stream schedule
    batch.class_cost = np.array([timings[batch.class_key(int(batch.order[batch.classes[k].first]))]
                                 [best[batch.class_key(int(batch.order[batch.classes[k].first]))]]
                                 for k in range(batch.num_classes)])
    batch.schedule()
    return {"choice": {f"{a}x{a}-{n}": list(t) for (a, n), t in best.items()},
            "ms": {f"{a}x{a}-{n}": {f"{t[0]}/{t[1]}": 1e3 * v for t, v in tt.items()} for (a, n), tt in timings.items()}}
