"""Pilot least-squares channel estimation, float64.

Restates ``OrthogonalLeastSquaresChannelEstimation.estimate_channel``
(``hermespy/modem/waveforms/orthogonal/channel.py:29-66``): LS at the
reference elements, ``scipy.interpolate.griddata`` linear interpolation over
the holes in (word, subcarrier) index space, then nearest-neighbour fill of
what lies outside the pilots' convex hull.  The pilot sequence follows
``ConfigurablePilotWaveform.pilot_symbols`` (``hermespy/modem/waveform.py:805-820``):
the configured sequence tiled to length.
"""

from __future__ import annotations

import numpy as np
from scipy.interpolate import griddata

from . import ofdm


def pilot_symbols(sequence: np.ndarray, count: int) -> np.ndarray:
    reps = int(np.ceil(count / len(sequence)))
    return np.tile(sequence, reps)[:count] if reps > 1 else np.asarray(sequence)[:count]


def ls_estimate(frame: ofdm.Frame, grid: np.ndarray, pilots: np.ndarray) -> np.ndarray:
    """grid: (words, N) received symbols of one stream -> (words, N) channel estimate."""

    ref = frame.mask == ofdm.REFERENCE
    prop = grid[ref]
    est_ref = prop / pilots
    est = np.zeros((1, 1) + grid.shape, dtype=complex)
    est[0, 0, ref] = est_ref
    holes = np.where(~ref)
    if holes[0].size != 0:
        stems = np.where(ref)
        est[0, 0, holes[0], holes[1]] = griddata(stems, est_ref, holes, method="linear")
    nan_idx = np.where(np.isnan(est))
    stem_idx = np.where(~np.isnan(est))
    est[nan_idx] = griddata(stem_idx, est[stem_idx], nan_idx, method="nearest")
    return est[0, 0]
