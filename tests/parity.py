"""Per-drop parity checks of GPU drops against float64 oracle drops."""

from __future__ import annotations

import numpy as np

from oracle import mapping


def axis_thresholds(order: int) -> np.ndarray:
    """Decision thresholds of one axis of square 3GPP QAM (midpoints of the unit-energy PAM levels
    (2i - m + 1) / sqrt(2 (M - 1) / 3), ``psk_qam_mapping.py:262-298``)."""

    m = int(round(np.sqrt(order)))
    return (2.0 * np.arange(1, m) - m) / np.sqrt(2.0 * (order - 1) / 3.0)


def boundary_distance(x: np.ndarray, order: int) -> tuple:
    """Distance of every symbol's real and imaginary part to the nearest threshold of its axis."""

    th = axis_thresholds(order)
    return (np.min(np.abs(x.real[..., None] - th), axis=-1), np.min(np.abs(x.imag[..., None] - th), axis=-1))


def rel_err(a, b) -> float:
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def check_drop(gpu_eq: np.ndarray, gpu_errors: int, oracle_result, order: int, tol: float, label="") -> dict:
    """gpu_eq (Ntx, K) equalised data symbols and the kernel's error count against one oracle drop
    (errors, bits, equalized, tx_bits):

    * the symbols agree with the float64 oracle's within ``tol`` (norm-wise relative);
    * the kernel's count is exactly the count of the oracle's demapper on the kernel's own symbols
      (bit-exact demapping and counting);
    * the two error counts differ by at most the bits of symbols whose float32 deviation from the
      oracle reaches across a decision threshold on some axis (the decision-boundary slack), and
      such symbols are rare.
    """

    errors, bits, eq, tx_bits = oracle_result
    bps = int(np.log2(order))
    e = rel_err(gpu_eq, eq)
    assert e < tol, (label, "equalised symbols", e)
    rx = np.stack([mapping.detect_bits(gpu_eq[t].astype(np.complex128), order).astype(np.uint8)
                   for t in range(gpu_eq.shape[0])])
    own = int(np.sum(rx != tx_bits))
    assert int(gpu_errors) == own, (label, "count of the kernel's own decisions", int(gpu_errors), own)
    dev = gpu_eq.astype(np.complex128) - eq
    dre, dim = boundary_distance(eq, order)
    near = (dre <= np.abs(dev.real)) | (dim <= np.abs(dev.imag))
    slack = bps * int(np.count_nonzero(near))
    assert abs(int(gpu_errors) - int(errors)) <= slack, (label, "error counts", int(gpu_errors), int(errors), slack)
    assert slack <= max(2 * bps, 1e-3 * bits), (label, "symbols within float32 reach of a threshold", slack, bits)
    return {"rel_err": e, "delta_errors": int(gpu_errors) - int(errors), "slack_bits": slack}
