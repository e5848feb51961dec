"""Monte Carlo campaigns on device counters: grid sections, confidence-based early stopping
and the campaign result.

Mirrors of the reference's pymonte layer for the bit error rate of the drop path:

* :class:`ScalarEvaluationResult` keeps, per grid section, the running sum, squared sum and
  count of the per-drop BER (``hermespy/core/pymonte/scalar.py:37-125``) in device memory;
  ``hpx_section_accumulate`` adds a whole launch of drops at once.  The stopping rule is the
  reference's (Bayer et al., Algorithm 1, ``scalar.py:117-123``), evaluated whenever a
  section's count reaches a multiple of ``min_num_samples``.
* :class:`SectionQueue` is ``MonteCarloQueueManager`` (``actors.py:20-148``) issuing blocks
  of drops instead of single tasks: every active section receives, per launch, the drops up
  to its next multiple of ``min_num_samples`` (or up to ``num_samples``), so each stopping
  test sees exactly the first ``m * min_num_samples`` drops of the section, as the sequential
  ``add_artifact`` calls of the reference do when a batch holds each section once.
* :class:`MonteCarloResult` (``monte_carlo.py:30-148``): ``save_to_matlab``, ``print``, ``plot``.

Drop ``k`` of section ``s`` of an ``S``-section grid is Philox drop ``k * S + s``, the same
drop that :meth:`BatchedSimulation.run` simulates at index ``k * S + s``, so a campaign
with early stopping simulates a prefix of every section's drops of the full campaign.
"""

from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass
from typing import Sequence

import numpy as np
import torch
from scipy.stats import norm

from . import _lib


# ---------------------------------------------------------------- grid
class SamplePoint:
    """``hermespy/core/pymonte/grid.py:140-187``."""

    def __init__(self, value: object, title: str | None = None) -> None:
        self.value = value
        self._title = title

    @property
    def title(self) -> str:
        if self._title is not None:
            return self._title
        if type(self.value).__str__ is not object.__str__:
            return str(self.value)
        if isinstance(self.value, float):
            return f"{self.value:.2g}"
        if isinstance(self.value, int):
            return f"{self.value}"
        return self.value.__class__.__name__


class GridDimensionInfo:
    """``grid.py:190-247``: sample points, title, plot scale and tick format of one dimension."""

    def __init__(self, sample_points: Sequence[SamplePoint], title: str, plot_scale: str = "linear",
                 tick_format: str = "lin") -> None:
        self.sample_points = list(sample_points)
        self.title = title
        self.plot_scale = plot_scale
        self.tick_format = tick_format

    @property
    def num_sample_points(self) -> int:
        return len(self.sample_points)


def snr_dimension(snr_db: Sequence[float]) -> GridDimensionInfo:
    """The SNR grid of a drop campaign; titles as ``Logarithmic.__str__`` (``core/logarithmic.py:220-232``)."""

    def title(v: float) -> str:
        return f"{int(v)}dB" if math.isclose(int(v), v) else f"{v:.2g}dB"

    return GridDimensionInfo([SamplePoint(float(v), title(float(v))) for v in snr_db], "SNR", "linear", "db")


# ---------------------------------------------------------------- results
def _check_arguments(accuracy: float, confidence: float, min_num_samples: int) -> None:
    # scalar.py:56-66
    if confidence > 1.0 or confidence < 0.0:
        raise ValueError("Confidence requirement must be between zero and one")
    if accuracy < 0.0:
        raise ValueError("Confidence bound must be non-negative")
    if min_num_samples < 2:
        raise ValueError("Minimum number of samples must be at least two")


class ScalarEvaluationResult:
    """Per-section BER statistics in device memory (``scalar.py:37-125``).

    ``sum``, ``sum_sq`` (float64) and ``count``, ``errors``, ``bits`` (int64) are tensors of
    one entry per grid section, sections in ``numpy.ndindex`` order of the grid."""

    def __init__(self, grid: Sequence[GridDimensionInfo], evaluator=None, accuracy: float = 0.0,
                 confidence: float = 1.0, min_num_samples: int = 1024, plot_surface: bool = True,
                 base_dimension_index: int = 0, device: str | torch.device = "cuda") -> None:
        _check_arguments(accuracy, confidence, min_num_samples)
        self.grid = list(grid)
        self.evaluator = evaluator
        self.accuracy = float(accuracy)
        self.confidence = float(confidence)
        self.min_num_samples = int(min_num_samples)
        self.plot_surface = plot_surface
        self.base_dimension_index = base_dimension_index
        self.shape = tuple(d.num_sample_points for d in self.grid) or (1,)
        self.num_sections = int(np.prod(self.shape))
        dev = torch.device(device)
        self.sum = torch.zeros(self.num_sections, dtype=torch.float64, device=dev)
        self.sum_sq = torch.zeros_like(self.sum)
        self.count = torch.zeros(self.num_sections, dtype=torch.int64, device=dev)
        self.errors = torch.zeros_like(self.count)
        self.bits = torch.zeros_like(self.count)

    @property
    def title(self) -> str:
        return "Bit Error Rate Evaluation" if self.evaluator is None else self.evaluator.title

    def section_index(self, coordinates) -> int:
        return int(np.ravel_multi_index(tuple(np.atleast_1d(coordinates)), self.shape))

    def accumulate(self, errors: torch.Tensor, bits: int | None = None, section: torch.Tensor | None = None,
                   first_drop: int = 0, bits_per_drop: torch.Tensor | None = None, stream: int | None = None) -> None:
        """Add a launch's drops (asynchronous): drop i has ``errors[i]`` bit errors out of ``bits``
        (or ``bits_per_drop[i]``) and belongs to section ``section[i]`` (int32), or to
        ``(first_drop + i) % num_sections``."""

        n = int(errors.numel())
        if errors.dtype != torch.int32:
            raise ValueError("errors must be int32")
        for name, t in (("section", section), ("bits_per_drop", bits_per_drop)):
            if t is not None and (t.dtype != torch.int32 or t.numel() < n or not t.is_contiguous()):
                raise ValueError(f"{name} must be a contiguous int32 tensor with one entry per drop")
        if bits_per_drop is None and (bits is None or bits <= 0):
            raise ValueError("give the bits per drop")
        stats = _lib.HpxSectionStats(*[_lib.ptr(t, "section counter") for t in (self.sum, self.sum_sq, self.count,
                                                                                self.errors, self.bits)])
        _lib.check(_lib.load().hpx_section_accumulate(
            _lib.ptr(errors, "errors"), _lib.ptr(section, "section"), int(first_drop),
            _lib.ptr(bits_per_drop, "bits_per_drop"), int(bits or 0), n, self.num_sections, ctypes.byref(stats),
            _lib.stream_handle() if stream is None else stream))

    def add_artifact(self, coordinates, artifact, compute_confidence: bool = True) -> bool:
        """One artifact (a per-drop BER) at a time, as ``scalar.py:100-125``; reads the counters back."""

        s = self.section_index(coordinates)
        scalar = float(artifact.to_scalar() if hasattr(artifact, "to_scalar") else artifact)
        self.sum[s] += scalar
        self.sum_sq[s] += scalar**2
        self.count[s] += 1
        count = int(self.count[s].item())
        if compute_confidence and count % self.min_num_samples == 0:
            return bool(self.stopping_test(np.array([s]))[0])
        return False

    def host_counters(self) -> tuple:
        return (self.sum.cpu().numpy(), self.sum_sq.cpu().numpy(), self.count.cpu().numpy(),
                self.errors.cpu().numpy(), self.bits.cpu().numpy())

    def stopping_test(self, sections: np.ndarray, counters: tuple | None = None) -> np.ndarray:
        """The reference's test (``scalar.py:117-123``) for the given sections on their current
        counters; sections whose count is not a positive multiple of ``min_num_samples`` are
        not tested (False)."""

        total, squared, count = (counters or self.host_counters())[:3]
        out = np.zeros(len(sections), dtype=bool)
        for j, s in enumerate(np.asarray(sections, dtype=np.int64)):
            c = int(count[s])
            if c == 0 or c % self.min_num_samples:
                continue
            std = ((squared[s] - (total[s] ** 2 / c)) / (c - 1)) ** 0.5
            if std > 0.0:
                out[j] = 2 * (1 - norm.cdf(c**0.5 * self.accuracy / std)) < self.confidence
        return out

    def to_array(self) -> np.ndarray:
        """Mean per-drop BER per section (``scalar.py:171-174``; NaN where no drop landed)."""

        total, _, count = self.host_counters()[:3]
        with np.errstate(invalid="ignore", divide="ignore"):
            return (total / count).reshape(self.shape)

    def runtime_estimates(self) -> np.ndarray:
        return self.to_array()

    def pooled_ber(self) -> np.ndarray:
        """Bit errors over bits per section (equal to :meth:`to_array` when every drop carries
        the same number of bits)."""

        _, _, _, errors, bits = self.host_counters()
        with np.errstate(invalid="ignore", divide="ignore"):
            return (errors / bits).reshape(self.shape)

    def num_drops(self) -> np.ndarray:
        return self.count.cpu().numpy().reshape(self.shape)

    def _value_str(self, value) -> str:
        # evaluation.py:166-176: log-scale evaluators print 10 log10(value) dB
        log = self.evaluator is not None and (getattr(self.evaluator, "plot_scale", "linear") == "log" or
                                               getattr(self.evaluator, "tick_format", "lin") == "db")
        if log:
            with np.errstate(divide="ignore", invalid="ignore"):
                return f"{10 * np.log10(value):.2g} dB"
        return f"{value:.2g}"

    def to_str(self, grid_coordinates) -> str:
        return self._value_str(float(self.to_array()[tuple(grid_coordinates)]))

    def print(self, console=None) -> None:
        from rich.console import Console
        from rich.table import Table

        console = Console() if console is None else console
        table = Table()
        for d in self.grid:
            table.add_column(d.title, style="cyan")
        table.add_column("?" if self.evaluator is None else self.evaluator.abbreviation)
        for section in np.ndindex(*self.shape):
            table.add_row(*[g.sample_points[s].title for s, g in zip(section, self.grid)], self.to_str(section))
        console.print(table)


class MonteCarloResult:
    """``hermespy/core/pymonte/monte_carlo.py:30-148``."""

    def __init__(self, grid: Sequence[GridDimensionInfo], evaluators: Sequence, results: Sequence[ScalarEvaluationResult],
                 performance_time: float) -> None:
        self._grid = list(grid)
        self._evaluators = list(evaluators)
        self._results = list(results)
        self._performance_time = float(performance_time)

    @property
    def performance_time(self) -> float:
        return self._performance_time

    @property
    def evaluation_results(self) -> Sequence[ScalarEvaluationResult]:
        return self._results

    def save_to_matlab(self, file: str) -> None:
        """The reference's .mat layout (``monte_carlo.py:80-106``); sample points are stored by value."""

        from scipy.io import savemat

        mat = {"dimensions": [d.title for d in self._grid], "evaluators": [e.abbreviation for e in self._evaluators],
               "performance_time": self._performance_time}
        for r, result in enumerate(self._results):
            mat[f"result_{r}"] = result.to_array()
        for d in self._grid:
            mat[d.title] = np.asarray([p.value for p in d.sample_points], dtype=np.float64)
        savemat(file, mat)

    def print(self, console=None) -> None:
        from rich.console import Console
        from rich.table import Table

        console = Console() if console is None else console
        table = Table()
        for d in self._grid:
            table.add_column(d.title, style="cyan")
        for e in self._evaluators:
            table.add_column(e.abbreviation)
        shape = tuple(d.num_sample_points for d in self._grid)
        for section in np.ndindex(*shape):
            row = [g.sample_points[s].title for s, g in zip(section, self._grid)]
            row += [r.to_str(section) for r in self._results]
            table.add_row(*row)
        console.print(table)

    def plot(self) -> list:
        """One figure per evaluator: result over the first grid dimension (log axis for log-scale
        evaluators); needs matplotlib."""

        import matplotlib

        matplotlib.use("Agg", force=False)
        import matplotlib.pyplot as plt

        figures = []
        for e, r in zip(self._evaluators, self._results):
            fig, ax = plt.subplots(1, 1, squeeze=False)
            d = self._grid[0]
            x = np.asarray([p.value for p in d.sample_points], dtype=np.float64)
            data = np.moveaxis(r.to_array(), r.base_dimension_index, 0).reshape(len(x), -1)
            ax[0, 0].plot(x, data)
            ax[0, 0].set_xlabel(d.title)
            ax[0, 0].set_ylabel(e.abbreviation)
            if getattr(e, "plot_scale", "linear") == "log":
                ax[0, 0].set_yscale("log")
            figures.append((fig, ax))
        return figures


# ---------------------------------------------------------------- queue
class SectionQueue:
    """``MonteCarloQueueManager`` (``actors.py:20-148``) issuing blocks of drops per launch.

    Per launch, every active section (in grid order) receives
    ``min(num_samples - issued, min_num_samples - issued % min_num_samples, cap)`` drops, with
    ``cap = max(1, max_drops // active)``; with more active sections than ``max_drops`` a launch
    serves ``max_drops`` of them, one drop each, round robin; a section leaves the queue when ``num_samples``
    drops were issued (``actors.py:102-106``) or when :meth:`disable_section` is called."""

    def __init__(self, num_sections: int, num_samples: int, min_num_samples: int = 1024,
                 max_drops: int = 16384) -> None:
        if num_sections < 1 or num_samples < 1 or max_drops < 1:
            raise ValueError("sections, samples and drops per launch must be positive")
        self.num_sections = int(num_sections)
        self.num_samples = int(num_samples)
        self.min_num_samples = int(min_num_samples)
        self.max_drops = int(max_drops)
        self.issued = np.zeros(self.num_sections, dtype=np.int64)
        self.active = np.ones(self.num_sections, dtype=bool)
        self._cursor = 0  # round-robin start when more sections are active than drops fit a launch

    @property
    def num_active(self) -> int:
        return int(self.active.sum())

    def disable_section(self, section: int) -> None:
        self.active[section] = False

    def next_launch(self) -> tuple:
        """(sections int32, sample index within the section int64) of the next launch, in
        section-major order; empty once no section is active."""

        act = np.flatnonzero(self.active)
        if act.size == 0:
            return np.zeros(0, dtype=np.int32), np.zeros(0, dtype=np.int64)
        if act.size > self.max_drops:
            # more active sections than drops per launch: one drop each for the next max_drops
            # sections in round-robin order, starting after the last section served
            start = int(np.searchsorted(act, self._cursor)) % act.size
            chosen = np.roll(act, -start)[: self.max_drops]
            self._cursor = int(chosen[-1]) + 1
            act = np.sort(chosen)
        cap = max(1, self.max_drops // act.size)
        issued = self.issued[act]
        block = np.minimum(np.minimum(self.num_samples - issued, self.min_num_samples - issued % self.min_num_samples),
                           cap)
        sections = np.repeat(act, block).astype(np.int32)
        starts = np.repeat(issued - np.concatenate([[0], np.cumsum(block)[:-1]]), block)
        samples = starts + np.arange(int(block.sum()), dtype=np.int64)
        self.issued[act] += block
        self.active[act[self.issued[act] >= self.num_samples]] = False
        return sections, samples

    def query_progress(self) -> tuple:
        """(progress in [0, 1], active flags), ``actors.py:133-148``: a disabled section counts as done."""

        done = np.where(self.active, np.minimum(self.issued, self.num_samples), self.num_samples)
        return float(done.sum()) / (self.num_sections * self.num_samples), self.active.copy()


@dataclass
class CampaignLaunch:
    sections: np.ndarray
    samples: np.ndarray
    drop_index: np.ndarray  # uint32


def drop_indices(sections: np.ndarray, samples: np.ndarray, num_sections: int) -> np.ndarray:
    """Philox drop of sample k of section s: k * S + s (the drop BatchedSimulation.run gives index k * S + s)."""

    d = samples.astype(np.int64) * num_sections + sections.astype(np.int64)
    if d.size and int(d.max()) >= 2**32:
        raise ValueError("drop index exceeds 32 bits")
    return d.astype(np.uint32)


def run_campaign(batched, seed: int, snr_db: Sequence[float], num_samples: int, evaluator=None,
                 max_drops_per_launch: int = 16384, record=None) -> MonteCarloResult:
    """Monte Carlo BER campaign of one link over an SNR grid with the reference's stopping rule.

    batched: a :class:`hermespy_amd.simulation.BatchedLink`; evaluator: a
    :class:`hermespy_amd.modem.BitErrorEvaluator` (its confidence, tolerance and
    min_num_samples configure the stopping rule; defaults never stop early).
    record: optional callable(launch, errors) invoked after every launch (drop recording)."""

    from .modem import BitErrorEvaluator

    ev = BitErrorEvaluator() if evaluator is None else evaluator
    start = time.perf_counter()
    snr = np.asarray(snr_db, dtype=np.float64)
    grid = [snr_dimension(snr)]
    result = ev.initialize_result(grid, device=batched.device)
    queue = SectionQueue(result.num_sections, num_samples, result.min_num_samples, max_drops_per_launch)
    link = batched.link
    noise = link.noise_power(snr)
    dev = batched.device
    errors = torch.empty(max_drops_per_launch, dtype=torch.int32, device=dev)
    while queue.num_active:
        sections, samples = queue.next_launch()
        n = sections.size
        ids = drop_indices(sections, samples, result.num_sections)
        sec_t = torch.as_tensor(sections, device=dev)
        ids_t = torch.as_tensor(ids.view(np.int32), device=dev)
        npw = torch.as_tensor(noise[sections], dtype=torch.float64, device=dev)
        batched.run(seed, 0, npw, errors[:n], drop_index=ids_t)
        result.accumulate(errors[:n], bits=link.bits_per_drop, section=sec_t)
        if record is not None:
            record(CampaignLaunch(sections, samples, ids), errors[:n])
        counters = result.host_counters()
        tested = np.unique(sections)
        stop = result.stopping_test(tested, counters)
        for s in tested[stop]:
            queue.disable_section(int(s))
    if dev.type == "cuda":
        torch.cuda.synchronize(dev)
    return MonteCarloResult(grid, [ev], [result], time.perf_counter() - start)
