"""OFDM frame restated in float64 (oversampling factor 1).

Follows ``hermespy/modem/waveforms/orthogonal/ofdm.py`` (forward :97-130,
backward :132-161) and ``hermespy/modem/waveforms/orthogonal/waveform.py``
(``GridSection.place_symbols``/``pick_symbols`` :361-406, ``SymbolSection.place_samples``
:520-565, ``pick_samples`` :567-614, ``OrthogonalWaveform.place/pick/modulate/demodulate``
:1222-1364).  A frame is one ``SymbolSection`` whose words each carry a cyclic
prefix of ``int(N * prefix_ratio)`` samples and an element-type mask.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy.fft import fft, fftshift, ifft, ifftshift

REFERENCE, DATA, NULL = 0, 1, 2


@dataclass
class Frame:
    num_subcarriers: int
    prefix: np.ndarray  # (words,) cyclic prefix lengths
    mask: np.ndarray  # (words, N) element types (REFERENCE/DATA/NULL)
    dc_suppression: bool = True

    @property
    def words(self) -> int:
        return int(self.mask.shape[0])

    @property
    def num_data_symbols(self) -> int:
        return int(np.count_nonzero(self.mask == DATA))

    @property
    def num_references(self) -> int:
        return int(np.count_nonzero(self.mask == REFERENCE))

    @property
    def samples_per_frame(self) -> int:
        return int(self.words * self.num_subcarriers + self.prefix.sum())

    def word_starts(self) -> np.ndarray:
        """First sample of each word's prefix."""

        lengths = self.prefix + self.num_subcarriers
        return np.concatenate([[0], np.cumsum(lengths)[:-1]]).astype(int)


def simple_frame(num_subcarriers: int, words: int, prefix_ratio: float, dc_suppression: bool = True) -> Frame:
    """All-data frame: ``GridResource(N, prefix_ratio, [DATA])`` repeated ``words`` times."""

    cp = int(num_subcarriers * 1 * prefix_ratio)  # waveform.py:538
    return Frame(num_subcarriers, np.full(words, cp, dtype=int), np.full((words, num_subcarriers), DATA, dtype=np.int8),
                 dc_suppression)


def place(frame: Frame, data: np.ndarray, references: np.ndarray | None = None) -> np.ndarray:
    """Symbols -> (words, N) grid, ``waveform.py:361-379,1222-1258``."""

    if data.size != frame.num_data_symbols:
        raise ValueError("Number of provided data symbols does not match the number of symbols in the frame")
    grid = np.zeros((frame.words, frame.num_subcarriers), dtype=np.complex128)
    if frame.num_references:
        grid[frame.mask == REFERENCE] = references
    grid[frame.mask == DATA] = data
    return grid


def pick(frame: Frame, grid: np.ndarray) -> np.ndarray:
    """Data symbols of a (..., words, N) grid, ``waveform.py:381-406``."""

    return grid[..., frame.mask == DATA]


def forward_transformation(frame: Frame, grid: np.ndarray) -> np.ndarray:
    """``ofdm.py:97-130`` with oversampling 1."""

    n = frame.num_subcarriers
    padded = grid / np.sqrt(n)
    padded = np.array(padded, dtype=np.complex128)
    if frame.dc_suppression:
        dc = int(0.5 * n)
        padded[:, dc:] = np.roll(padded[:, dc:], 1, axis=-1)
    return np.asarray(ifft(ifftshift(padded, axes=-1), n, axis=-1, norm="forward"))


def backward_transformation(frame: Frame, signal_grid: np.ndarray, normalize: bool = True) -> np.ndarray:
    """``ofdm.py:132-161`` with oversampling 1 (FFT over the last axis, zero-padded/truncated to N)."""

    n = frame.num_subcarriers
    sym = fft(signal_grid, n=n, axis=-1, norm="backward")
    sym = fftshift(sym, axes=-1)
    if frame.dc_suppression:
        dc = int(0.5 * n)
        sym[..., dc:] = np.roll(sym[..., dc:], -1, axis=-1)
    if normalize:
        sym = sym * n ** -0.5
    return sym


def modulate(frame: Frame, grid: np.ndarray) -> np.ndarray:
    """Grid (words, N) -> frame samples with cyclic prefixes, ``waveform.py:520-565,1293-1323``."""

    sig = forward_transformation(frame, grid)
    out = np.empty(frame.samples_per_frame, dtype=np.complex128)
    i = 0
    for w in range(frame.words):
        cp = int(frame.prefix[w])
        if cp > 0:
            out[i:i + cp] = sig[w, -cp:]
            i += cp
        out[i:i + frame.num_subcarriers] = sig[w]
        i += frame.num_subcarriers
    return out


def pick_samples(frame: Frame, signal: np.ndarray) -> np.ndarray:
    """Remove prefixes: (..., samples) -> (..., words, N), ``waveform.py:567-614``."""

    n = frame.num_subcarriers
    out = np.empty(signal.shape[:-1] + (frame.words, n), dtype=np.complex128)
    i = 0
    for w in range(frame.words):
        i += int(frame.prefix[w])
        out[..., w, :] = signal[..., i:i + n]
        i += n
    return out


def demodulate(frame: Frame, signal: np.ndarray) -> np.ndarray:
    """Frame samples (..., samples) -> symbol grid (..., words, N), ``waveform.py:1326-1364``."""

    return backward_transformation(frame, pick_samples(frame, signal))


def subcarrier_bins(n: int, dc_suppression: bool = True) -> np.ndarray:
    """FFT bin that carries grid column ``q`` (the composition of the DC roll and (i)fftshift)."""

    q = np.arange(n)
    p = q.copy()
    if dc_suppression:
        dc = n // 2
        p[dc:] = dc + ((q[dc:] - dc + 1) % (n - dc))
    return (p - n // 2) % n
