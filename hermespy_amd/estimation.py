"""Pilot least-squares channel estimation with the reference's interpolation, on the GPU.

Mirrors ``OrthogonalLeastSquaresChannelEstimation``
(``hermespy/modem/waveforms/orthogonal/channel.py:25-66``): LS at the reference
elements, then ``scipy.interpolate.griddata`` -- linear on the Delaunay
triangulation of the pilot positions, then nearest-neighbour fill outside
their convex hull.  Both interpolants depend only on the frame's pilot
pattern, so the host builds them once per waveform with the same SciPy
objects griddata uses (``Delaunay``/barycentric weights, ``NearestNDInterpolator``
over the same 4-D index points) and every drop runs three gather kernels
(``hpx_ls_estimate``).  The reference estimates SISO links only
(``channel.py:32-35``); here each receive stream is estimated on its own.
"""

from __future__ import annotations

import numpy as np
import torch
from scipy.interpolate import NearestNDInterpolator
from scipy.spatial import Delaunay

from . import _lib
from .modem import ElementType, OFDMWaveform


def pilot_symbols(sequence: np.ndarray, count: int) -> np.ndarray:
    """``ConfigurablePilotWaveform.pilot_symbols`` (``hermespy/modem/waveform.py:805-820``)."""

    sequence = np.asarray(sequence)
    reps = int(np.ceil(count / len(sequence)))
    return np.tile(sequence, reps)[:count] if reps > 1 else sequence[:count]


def ls_plan(ref_mask: np.ndarray) -> dict:
    """Static interpolation plan of a (words, N) reference mask.

    Returns flat grid indices of the pilots, of the holes with their Delaunay
    simplex vertices (pilot numbers, -1 outside the hull) and barycentric
    weights, and of the elements filled by their nearest estimated element.
    """

    words, n = ref_mask.shape
    stems = np.nonzero(ref_mask)
    holes = np.nonzero(~ref_mask)
    pilot_flat = (stems[0] * n + stems[1]).astype(np.int32)
    hole_flat = (holes[0] * n + holes[1]).astype(np.int32)
    verts = np.full((hole_flat.size, 3), -1, dtype=np.int32)
    weights = np.zeros((hole_flat.size, 3), dtype=np.float64)
    if hole_flat.size and pilot_flat.size >= 3:
        pts = np.stack(stems, axis=-1).astype(np.float64)
        xi = np.stack(holes, axis=-1).astype(np.float64)
        try:
            tri = Delaunay(pts)
        except Exception:  # degenerate pilot sets (collinear) interpolate nothing, as griddata
            tri = None
        if tri is not None:
            simplex = tri.find_simplex(xi, tol=100 * np.finfo(float).eps)
            inside = simplex >= 0
            tr = tri.transform[simplex[inside]]
            b = np.einsum("hij,hj->hi", tr[:, :2, :], xi[inside] - tr[:, 2, :])
            weights[inside, :2] = b
            weights[inside, 2] = 1.0 - b.sum(axis=1)
            verts[inside] = tri.simplices[simplex[inside]]
    # elements still undefined after the linear stage (channel.py:59-64)
    defined = np.zeros(words * n, dtype=bool)
    defined[pilot_flat] = True
    defined[hole_flat[verts[:, 0] >= 0]] = True
    nan_flat = np.flatnonzero(~defined).astype(np.int32)
    src = np.zeros(0, dtype=np.int32)
    if nan_flat.size:
        if not defined.any():
            raise ValueError("no reference elements to estimate from")
        stem4 = np.stack([np.zeros(defined.sum()), np.zeros(defined.sum()), *np.divmod(np.flatnonzero(defined), n)], -1)
        nan4 = np.stack([np.zeros(nan_flat.size), np.zeros(nan_flat.size), *np.divmod(nan_flat, n)], -1)
        src_rank = NearestNDInterpolator(stem4, np.arange(stem4.shape[0]))(nan4).astype(np.int64)
        src = np.flatnonzero(defined)[src_rank].astype(np.int32)
    return {"pilot": pilot_flat, "hole": hole_flat, "verts": verts, "weights": weights, "fill": nan_flat,
            "fill_src": src}


def apply_plan(plan: dict, ls_values: np.ndarray, words: int, n: int) -> np.ndarray:
    """Host application of a plan (used to validate it against griddata)."""

    est = np.zeros(words * n, dtype=np.complex128)
    est[plan["pilot"]] = ls_values
    inside = plan["verts"][:, 0] >= 0
    est[plan["hole"][inside]] = np.sum(plan["weights"][inside] * ls_values[plan["verts"][inside]], axis=1)
    est[plan["fill"]] = est[plan["fill_src"]]
    return est.reshape(words, n)


class OrthogonalLeastSquaresChannelEstimation:
    """LS channel estimation of an OFDM frame (``channel.py:25-66``)."""

    def __init__(self, waveform: OFDMWaveform, pilot_sequence: np.ndarray | None = None,
                 device: str | torch.device = "cuda") -> None:
        self.waveform = waveform
        mask = waveform.resource_mask()[ElementType.REFERENCE.value]
        if not mask.any():
            raise ValueError("the waveform has no reference elements")
        self.plan = ls_plan(mask)
        seq = waveform.mapping.get_mapping() if pilot_sequence is None else np.asarray(pilot_sequence)
        self.pilots = pilot_symbols(seq, self.plan["pilot"].size)
        dev = torch.device(device)
        self._dev = {k: torch.as_tensor(np.ascontiguousarray(v), device=dev) for k, v in self.plan.items()}
        self._dev["weights"] = self._dev["weights"].to(torch.float32)
        self._dev["pilots"] = torch.as_tensor(self.pilots.astype(np.complex64), device=dev)

    def estimate_channel(self, grid: torch.Tensor) -> torch.Tensor:
        """(streams, words, N) received grids -> (streams, words, N) channel estimates."""

        grid = grid.to(torch.complex64).contiguous()
        streams = grid.shape[0]
        words, n = self.waveform.words_per_frame, self.waveform.num_subcarriers
        out = torch.empty_like(grid)
        d = self._dev
        _lib.check(_lib.load().hpx_ls_estimate(
            _lib.ptr(grid), streams, words * n, d["pilot"].numel(), _lib.ptr(d["pilot"]), _lib.ptr(d["pilots"]),
            d["hole"].numel(), _lib.ptr(d["hole"]), _lib.ptr(d["verts"]), _lib.ptr(d["weights"]), d["fill"].numel(),
            _lib.ptr(d["fill"]), _lib.ptr(d["fill_src"]), _lib.ptr(out), _lib.stream_handle()))
        return out
