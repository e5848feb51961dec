"""Oracle drops in worker processes, for GPU parity tests that compare many drops.

The workers are spawned (the test process holds a GPU context, which a forked
child must not inherit) and run the float64 oracle single-threaded, one drop
per task.
"""

from __future__ import annotations

import multiprocessing as mp
import os
from concurrent.futures import ProcessPoolExecutor


def _drop(args):
    link, seed, drop, snr_db, faithful = args[:5]
    normals = args[5] if len(args) > 5 else None
    from threadpoolctl import threadpool_limits

    from oracle import drop as odrop

    with threadpool_limits(1):
        r = odrop.run_drop(link, seed, drop, snr_db, faithful=faithful, normals=normals)
    return r.errors, r.bits, r.equalized, r.tx_bits


def oracle_drops(link, seed: int, drops, snrs, faithful: bool = False, workers: int | None = None,
                 normals=None) -> list:
    """[(errors, bits, equalized (Ntx, K) complex128, tx_bits (Ntx, bits))] of the oracle's drops
    (normals: optional per-drop realisation scalars)."""

    tasks = [(link, seed, int(d), float(s), faithful, None if normals is None else normals[i])
             for i, (d, s) in enumerate(zip(drops, snrs))]
    n = workers or max(1, min(16, os.cpu_count() or 1, len(tasks)))
    if n == 1:
        return [_drop(t) for t in tasks]
    with ProcessPoolExecutor(n, mp_context=mp.get_context("spawn")) as pool:
        return list(pool.map(_drop, tasks))


def oracle_links(links, seed: int, drops, snrs, faithful: bool = False, workers: int | None = None) -> list:
    """As oracle_drops with one oracle link per drop."""

    tasks = [(lk, seed, int(d), float(s), faithful) for lk, d, s in zip(links, drops, snrs)]
    n = workers or max(1, min(16, os.cpu_count() or 1, len(tasks)))
    if n == 1:
        return [_drop(t) for t in tasks]
    with ProcessPoolExecutor(n, mp_context=mp.get_context("spawn")) as pool:
        return list(pool.map(_drop, tasks))
