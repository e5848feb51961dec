"""3GPP Gray PSK/QAM mapping and piecewise-linear LLR hard demapping (float64).

Restates ``hermespy/modem/tools/psk_qam_mapping.py``:
``get_symbols`` :104-169, ``detect_bits`` :171-259, ``generate_pam_symbol_3gpp``
:262-298, ``get_llr_3gpp`` :301-497.  Only the default 3GPP constellations
(BPSK, QPSK, 16/64/256-QAM, 4/8/16-PAM) are restated; custom tables are not on
the drop path.
"""

from __future__ import annotations

import numpy as np


def pam_3gpp(order: int, bits: np.ndarray) -> np.ndarray:
    """Unnormalised PAM amplitudes, ``psk_qam_mapping.py:283-298``. ``bits``: (log2 order, K)."""

    b = bits.astype(np.float64)
    if order == 2:
        s = 1.0 - 2 * b
    elif order == 4:
        s = (1 - 2 * b[0, :]) * (1 + 2 * b[1, :])
    elif order == 8:
        s = (2 * b[0, :] - 1) * ((1 - 2 * b[1, :]) * (1 + 2 * b[2, :]) - 4)
    elif order == 16:
        s = ((((1 - 2 * b[2, :]) * (1 + 2 * b[3, :]) - 4) * (-1 + 2 * b[1, :])) - 8) * (-1 + 2 * b[0, :])
    else:
        raise ValueError(f"Unsupported modulation order ({order})")
    return s.flatten()


def get_symbols(bits: np.ndarray, order: int, is_complex: bool = True) -> np.ndarray:
    """Bits -> unit-energy symbols, ``psk_qam_mapping.py:121-169``."""

    bps = int(np.log2(order))
    n = int(bits.size / bps)
    sb = np.asarray(bits).reshape((bps, n), order="F")
    if order == 2:
        sym = pam_3gpp(2, sb) + 1j * 0
    elif order == 4 and is_complex:
        sym = (pam_3gpp(2, sb[0:1, :]) + 1j * pam_3gpp(2, sb[1:2, :])) / np.sqrt(2)
    elif order == 4:
        sym = pam_3gpp(4, sb) / np.sqrt(5) + 1j * 0
    elif order == 8 and not is_complex:
        sym = pam_3gpp(8, sb) / np.sqrt(21) + 1j * 0
    elif order == 16 and is_complex:
        sym = (pam_3gpp(4, sb[[0, 2], :]) + 1j * pam_3gpp(4, sb[[1, 3], :])) / np.sqrt(10)
    elif order == 16:
        sym = pam_3gpp(16, sb) / np.sqrt(85) + 1j * 0
    elif order == 64 and is_complex:
        sym = (pam_3gpp(8, sb[[0, 2, 4], :]) + 1j * pam_3gpp(8, sb[[1, 3, 5], :])) / np.sqrt(42)
    elif order == 256 and is_complex:
        sym = (pam_3gpp(16, sb[[0, 2, 4, 6], :]) + 1j * pam_3gpp(16, sb[[1, 3, 5, 7], :])) / np.sqrt(170)
    else:
        raise RuntimeError(f"Modulation ({order}) not supported")
    return np.ravel(sym)


def _band(r: np.ndarray, lo: float, hi: float) -> np.ndarray:
    return np.bitwise_and(r > lo, r <= hi)


def llr_3gpp(order: int, rx: np.ndarray, nv: np.ndarray, is_complex: bool) -> np.ndarray:
    """Piecewise-linear LLRs, literally as ``psk_qam_mapping.py:331-497``."""

    if is_complex:
        rx = rx * np.sqrt(2)
    if order == 2:
        llr = -2 * rx / nv
    elif order == 4:
        llr = np.zeros([2, rx.size])
        r = rx * np.sqrt(5)
        llr[0, :] = 2 * ((r <= -2) * (-4 * (1 + r)) + np.logical_and(r > -2, r <= 2) * (-2 * r) + (r > 2) * (4 * (1 - r))) / nv
        llr[1, :] = 2 * ((r <= 0) * (-2 * (2 + r)) + (r > 0) * (-2 * (2 - r))) / nv
        llr = llr / 5
    elif order == 8:
        llr = np.zeros([3, rx.size])
        r = rx * np.sqrt(21)
        llr[0, :] = 4 * (
            (r <= -6) * (-4 * (3 + r)) + _band(r, -6, -4) * (-3 * (2 + r)) + _band(r, -4, -2) * (-2 * (1 + r))
            + _band(r, -2, 2) * (-r) + _band(r, 2, 4) * (2 * (1 - r)) + _band(r, 4, 6) * (3 * (2 - r)) + (r > 6) * (4 * (3 - r))
        ) / nv
        llr[1, :] = 4 * (
            (r <= -6) * (-2 * (5 + r)) + _band(r, -6, -2) * (-(4 + r)) + _band(r, -2, 0) * (-2 * (3 + r))
            + _band(r, 0, 2) * (-2 * (3 - r)) + _band(r, 2, 6) * (-(4 - r)) + (r > 6) * (-2 * (5 - r))
        ) / nv
        llr[2, :] = 4 * ((r <= -4) * (-(6 + r)) + _band(r, -4, 0) * (2 + r) + _band(r, 0, 4) * (2 - r) + (r > 4) * (-(6 - r))) / nv
        llr = llr / 21
    elif order == 16:
        llr = np.zeros([4, rx.size])
        r = rx * np.sqrt(85)
        llr[0, :] = 8 * (
            (r <= -14) * (-4 * (7 + r)) - _band(r, -14, -12) * 3.5 * (6 + r) - _band(r, -12, -10) * 3 * (5 + r)
            - _band(r, -10, -8) * 2.5 * (4 + r) - _band(r, -8, -6) * 2 * (3 + r) - _band(r, -6, -4) * 1.5 * (2 + r)
            - _band(r, -4, -2) * (1 + r) - _band(r, -2, 2) * 0.5 * r + _band(r, 2, 4) * (1 - r)
            + _band(r, 4, 6) * 1.5 * (2 - r) + _band(r, 6, 8) * 2 * (3 - r) + _band(r, 8, 10) * 2.5 * (4 - r)
            + _band(r, 10, 12) * 3 * (5 - r) + _band(r, 12, 14) * 3.5 * (6 - r) + (r > 14) * 4 * (7 - r)
        ) / nv
        llr[1, :] = 8 * (
            (r <= -14) * (-2 * (11 + r)) - _band(r, -14, -12) * 1.5 * (10 + r) - _band(r, -12, -10) * (9 + r)
            - _band(r, -10, -6) * 0.5 * (8 + r) - _band(r, -6, -4) * (7 + r) - _band(r, -4, -2) * 1.5 * (6 + r)
            - _band(r, -2, 0) * 2 * (5 + r) - _band(r, 0, 2) * 2 * (5 - r) - _band(r, 2, 4) * 1.5 * (6 - r)
            - _band(r, 4, 6) * (7 - r) - _band(r, 6, 10) * 0.5 * (8 - r) - _band(r, 10, 12) * (9 - r)
            - _band(r, 12, 14) * 1.5 * (10 - r) - (r > 14) * 2 * (11 - r)
        ) / nv
        llr[2, :] = 8 * (
            (r <= -14) * (-13 - r) - _band(r, -14, -10) * 0.5 * (12 + r) - _band(r, -10, -8) * (11 + r)
            + _band(r, -8, -6) * (5 + r) + _band(r, -6, -2) * 0.5 * (4 + r) + _band(r, -2, 0) * (3 + r)
            + _band(r, 0, 2) * (3 - r) + _band(r, 2, 6) * 0.5 * (4 - r) + _band(r, 6, 8) * (5 - r)
            - _band(r, 8, 10) * (11 - r) - _band(r, 10, 14) * 0.5 * (12 - r) - (r > 14) * (13 - r)
        ) / nv
        llr[3, :] = 8 * (
            (r <= -12) * (-0.5 * (14 + r)) + _band(r, -12, -8) * 0.5 * (10 + r) - _band(r, -8, -4) * 0.5 * (6 + r)
            + _band(r, -4, 0) * 0.5 * (2 + r) + _band(r, 0, 4) * 0.5 * (2 - r) - _band(r, 4, 8) * 0.5 * (6 - r)
            + _band(r, 8, 12) * 0.5 * (10 - r) - (r > 12) * 0.5 * (14 - r)
        ) / nv
        llr = llr / 85
    else:
        raise ValueError(f"unsupported modulation order ({order})")
    if is_complex:
        llr = llr / 2
    return llr.ravel("F")


def detect_bits(rx: np.ndarray, order: int, is_complex: bool = True, soft: bool = False) -> np.ndarray:
    """Hard (``llr > 0``) or soft bits, ``psk_qam_mapping.py:171-259`` with unit noise variance."""

    rx = np.asarray(rx)
    nbits = rx.size * int(np.log2(order))
    llr = np.zeros(nbits)
    nv = np.ones(rx.shape)
    if order == 2:
        llr = llr_3gpp(2, np.real(rx), nv, False)
    elif order == 4 and is_complex:
        llr[0::2] = llr_3gpp(2, np.real(rx), nv, True)
        llr[1::2] = llr_3gpp(2, np.imag(rx), nv, True)
    elif order in (4, 8, 16) and not is_complex:
        llr = llr_3gpp(order, np.real(rx), nv, False)
    elif order in (16, 64, 256) and is_complex:
        pam = {16: 4, 64: 8, 256: 16}[order]
        llr[0::2] = llr_3gpp(pam, np.real(rx), nv, True)
        llr[1::2] = llr_3gpp(pam, np.imag(rx), nv, True)
    else:
        raise RuntimeError("Unsupported modulation scheme")
    return llr if soft else llr > 0


def constellation(order: int, is_complex: bool = True) -> np.ndarray:
    """All constellation points indexed by their MSB-first bit label, ``psk_qam_mapping.py:499-524``."""

    bps = int(np.log2(order))
    labels = np.arange(order)
    bits = ((labels[:, None] >> (bps - 1 - np.arange(bps))[None, :]) & 1).astype(np.uint8)
    return get_symbols(bits.reshape(-1), order, is_complex)
