"""Launch-shape autotuner of the fused drop kernel.

For a link it times ``hpx_run_batch`` over the candidate workgroup sizes (the
kernel's tile shape: threads per drop, which sets samples per thread, FFT
values per thread and how many drops share a CU) and returns the fastest.
Results are cached per link shape for the process.
"""

from __future__ import annotations

import dataclasses
import sys

import torch

from .configs import LinkConfig

_CACHE: dict[tuple, int] = {}


def _shape_key(link: LinkConfig) -> tuple:
    return (link.num_rx, link.num_tx, link.num_subcarriers, link.words, link.prefix, link.modulation_order,
            link.num_paths, link.max_delay, link.detector)


def candidates(link: LinkConfig) -> list[int]:
    out = []
    for w in (64, 128, 256, 512):
        pairs = max(1, link.num_subcarriers // (2 * w))
        if pairs in (1, 2, 4) and link.num_subcarriers * max(link.num_rx, link.num_tx) <= 16 * w:
            out.append(w)
    return out


def tune(link: LinkConfig, drops: int = 2048, reps: int = 2, seed: int = 1, verbose: bool = False) -> int:
    """Fastest workgroup size for ``link`` (timed with HIP events on the current stream)."""

    from .simulation import BatchedLink

    key = _shape_key(link)
    if key in _CACHE:
        return _CACHE[key]
    best, best_ms = 0, float("inf")
    timings = {}
    npw = torch.full((drops,), float(link.noise_power(10.0)), dtype=torch.float64, device="cuda")
    err = torch.empty(drops, dtype=torch.int32, device="cuda")
    for w in candidates(link):
        bl = BatchedLink(dataclasses.replace(link, workgroup_size=w))
        bl.run(seed, 0, npw, err)  # warm-up
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for r in range(reps):
            bl.run(seed, (r + 1) * drops, npw, err)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        timings[w] = ms
        if ms < best_ms:
            best, best_ms = w, ms
    if verbose:
        print("autotune", link.name, {k: round(v, 3) for k, v in timings.items()}, "->", best, file=sys.stderr,
              flush=True)
    _CACHE[key] = best
    return best


def tuned(link: LinkConfig, **kwargs) -> LinkConfig:
    """``link`` with the autotuned workgroup size."""

    return dataclasses.replace(link, workgroup_size=tune(link, **kwargs))


__all__ = ["candidates", "tune", "tuned"]
