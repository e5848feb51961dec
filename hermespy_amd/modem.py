"""OFDM modem of the drop path with the reference's API, computing on the GPU.

Mirrors ``hermespy/modem/waveforms/orthogonal/waveform.py`` (``ElementType``,
``PrefixType``, ``GridElement`` :65-90, ``GridResource`` :102-222,
``SymbolSection`` :463-650, ``OrthogonalWaveform`` :950-1364),
``hermespy/modem/waveforms/orthogonal/ofdm.py`` (``OFDMWaveform`` :33-161),
``hermespy/modem/tools/psk_qam_mapping.py`` (``PskQamMapping`` :20-259),
``hermespy/modem/waveform.py`` (``ZeroForcingChannelEqualization`` :235-250),
``hermespy/simulation/modem/channel_estimation.py`` (``OFDMIdealChannelEstimation``
:182-295) and ``hermespy/modem/evaluators.py`` (``BitErrorEvaluator`` :236-270).

Arrays are torch tensors on the device (complex64 samples, uint8 bits).  The
device path covers oversampling factor 1 and one cyclic-prefix length per frame.
"""

from __future__ import annotations

import ctypes
from enum import Enum

import numpy as np
import torch

from . import _lib


class ElementType(Enum):
    REFERENCE = 0
    DATA = 1
    NULL = 2


class PrefixType(Enum):
    CYCLIC = 0
    ZEROPAD = 1
    NONE = 2


class Symbols:
    """Communication symbols (streams, blocks, symbols per block) on the device
    (``hermespy/modem/symbols.py:136-318``); 1-D and 2-D inputs are expanded as the reference
    expands them."""

    def __init__(self, symbols=None, device="cuda") -> None:
        if symbols is None:
            t = torch.zeros((0, 0, 0), dtype=torch.complex64, device=device)
        elif isinstance(symbols, torch.Tensor):
            t = symbols
        else:
            t = torch.as_tensor(np.asarray(symbols, dtype=np.complex64), device=device)
        if t.ndim > 3:
            raise ValueError("Symbols initialization array may have a maximum of three dimensions")
        if t.ndim == 1:
            t = t[None, :, None]
        elif t.ndim == 2:
            t = t[:, :, None]
        self._raw = t

    @property
    def raw(self) -> torch.Tensor:
        return self._raw

    @raw.setter
    def raw(self, value: torch.Tensor) -> None:
        self._raw = value

    @property
    def num_streams(self) -> int:
        return int(self._raw.shape[0])

    @property
    def num_blocks(self) -> int:
        return int(self._raw.shape[1])

    @property
    def num_symbols(self) -> int:
        return int(self._raw.shape[2])

    def copy(self) -> "Symbols":
        return Symbols(self._raw.clone())


class StatedSymbols(Symbols):
    """Symbols with their channel states (Nrx, Ntx, blocks, symbols per block)
    (``hermespy/modem/symbols.py:321-410``)."""

    def __init__(self, symbols, states) -> None:
        Symbols.__init__(self, symbols)
        self.states = states

    @property
    def states(self) -> torch.Tensor:
        return self._states

    @states.setter
    def states(self, value) -> None:
        if not isinstance(value, torch.Tensor):
            value = torch.as_tensor(np.asarray(value, dtype=np.complex64), device=self._raw.device)
        if value.ndim != 4:
            raise ValueError("State must be a four-dimensional numpy array")
        if value.shape[0] != self.num_streams:
            raise ValueError(f"Number of received streams don't match, expected {self.num_streams} instead of "
                             f"{value.shape[0]}")
        if value.shape[2] != self.num_blocks:
            raise ValueError(f"Number of received blocks don't match, expected {self.num_blocks} instead of "
                             f"{value.shape[2]}")
        if value.shape[3] != self.num_symbols:
            raise ValueError(f"Symbol block sizes don't match, expected {self.num_symbols} instead of "
                             f"{value.shape[3]}")
        self._states = value

    def dense_states(self) -> torch.Tensor:
        return self._states

    @property
    def num_transmit_streams(self) -> int:
        return int(self._states.shape[1])

    def copy(self) -> "StatedSymbols":
        return StatedSymbols(self._raw.clone(), self._states.clone())


class GridElement:
    def __init__(self, type: ElementType | str, repetitions: int = 1) -> None:
        self.type = ElementType[type] if isinstance(type, str) else type
        self.repetitions = int(repetitions)


class GridResource:
    """One frequency-domain resource pattern (``waveform.py:102-222``)."""

    def __init__(self, repetitions: int = 1, prefix_type: PrefixType | str = PrefixType.CYCLIC,
                 prefix_ratio: float = 0.0, elements: list[GridElement] | None = None) -> None:
        if repetitions < 1:
            raise ValueError("Number of frame resource repetitions must be greater or equal to one")
        if prefix_ratio < 0.0:
            raise ValueError(f"Cyclic prefix ratio must be non-negative (not {prefix_ratio})")
        self.repetitions = int(repetitions)
        self.prefix_type = PrefixType[prefix_type] if isinstance(prefix_type, str) else prefix_type
        self.prefix_ratio = float(prefix_ratio)
        self.elements = elements if elements is not None else []

    @property
    def num_subcarriers(self) -> int:
        return self.repetitions * sum(e.repetitions for e in self.elements)

    @property
    def mask(self) -> np.ndarray:
        """(len(ElementType), num_subcarriers) boolean mask (``waveform.py:204-222``)."""

        m = np.zeros((len(ElementType), sum(e.repetitions for e in self.elements)), dtype=bool)
        i = 0
        for e in self.elements:
            m[e.type.value, i:i + e.repetitions] = True
            i += e.repetitions
        return np.tile(m, (1, self.repetitions))


class SymbolSection:
    """Time-domain section of words (``waveform.py:463-650``)."""

    def __init__(self, num_repetitions: int = 1, pattern: list[int] | None = None, sample_offset: int = 0) -> None:
        if num_repetitions < 1:
            raise ValueError("OFDM frame number of repetitions must be greater or equal to one")
        if sample_offset != 0:
            raise ValueError("sample offsets are not supported on the device path")
        self.num_repetitions = int(num_repetitions)
        self.pattern = pattern if pattern is not None else []
        self.sample_offset = sample_offset

    @property
    def num_words(self) -> int:
        return self.num_repetitions * len(self.pattern)


class PskQamMapping:
    """3GPP Gray mapping (``psk_qam_mapping.py:20-259``) on the device."""

    def __init__(self, modulation_order: int, is_complex: bool = True) -> None:
        if modulation_order <= 0 or (modulation_order & (modulation_order - 1)) != 0:
            raise ValueError("modulation_order must be a power of two")
        self.modulation_order = int(modulation_order)
        self.is_complex = bool(is_complex)
        self.bits_per_symbol = int(np.log2(modulation_order))

    def get_symbols(self, bits: torch.Tensor) -> torch.Tensor:
        n = bits.numel() // self.bits_per_symbol
        bits = bits.to(torch.uint8).contiguous()
        out = torch.empty(n, dtype=torch.complex64, device=bits.device)
        _lib.check(_lib.load().hpx_map_symbols(_lib.ptr(bits), n, self.modulation_order, int(self.is_complex),
                                               _lib.ptr(out), _lib.stream_handle()))
        return out

    def detect_bits(self, symbols: torch.Tensor) -> torch.Tensor:
        symbols = symbols.to(torch.complex64).contiguous().reshape(-1)
        out = torch.empty(symbols.numel() * self.bits_per_symbol, dtype=torch.uint8, device=symbols.device)
        _lib.check(_lib.load().hpx_demap_hard(_lib.ptr(symbols), symbols.numel(), self.modulation_order,
                                              int(self.is_complex), _lib.ptr(out), _lib.stream_handle()))
        return out

    def get_mapping(self) -> np.ndarray:
        """Constellation in label order (``psk_qam_mapping.py:499-524``), float64 as the reference rounds it."""

        pts = np.zeros(2 * self.modulation_order)
        _lib.check(_lib.load().hpx_constellation_host(self.modulation_order, int(self.is_complex), pts.ctypes.data))
        return pts[0::2] + 1j * pts[1::2]


class OFDMWaveform:
    """OFDM frame of ``SymbolSection`` words (``ofdm.py:33-161``, ``waveform.py:950-1364``)."""

    def __init__(self, grid_resources: list[GridResource], grid_structure: list[SymbolSection],
                 num_subcarriers: int = 1024, dc_suppression: bool = True, modulation_order: int = 16,
                 oversampling_factor: int = 1) -> None:
        if num_subcarriers < 1:
            raise ValueError("Number of subcarriers must be greater or equal to one")
        if oversampling_factor != 1:
            raise ValueError("the device path runs at oversampling factor 1")
        self.num_subcarriers = int(num_subcarriers)
        self.grid_resources = grid_resources
        self.grid_structure = grid_structure
        self.dc_suppression = bool(dc_suppression)
        self.mapping = PskQamMapping(modulation_order)
        prefixes = set()
        for sec in grid_structure:
            for r in sec.pattern:
                res = grid_resources[r]
                cp = int(self.num_subcarriers * 1 * res.prefix_ratio)  # waveform.py:538
                if cp > 0 and res.prefix_type == PrefixType.ZEROPAD:
                    raise ValueError("zero-padding prefixes are not supported on the device path")
                prefixes.add(cp if res.prefix_type != PrefixType.NONE else 0)
        if len(prefixes) > 1:
            raise ValueError("the device path needs one prefix length per frame")
        self.prefix = prefixes.pop() if prefixes else 0
        self._mask = self.resource_mask()
        self._data_index = torch.as_tensor(np.flatnonzero(self._mask[ElementType.DATA.value].ravel()))

    @property
    def modulation_order(self) -> int:
        return self.mapping.modulation_order

    @property
    def bits_per_symbol(self) -> int:
        return self.mapping.bits_per_symbol

    @property
    def words_per_frame(self) -> int:
        return sum(s.num_words for s in self.grid_structure)

    def resource_mask(self) -> np.ndarray:
        """(len(ElementType), words, num_subcarriers) (``waveform.py:1194-1211``)."""

        mask = np.zeros((len(ElementType), self.words_per_frame, self.num_subcarriers), dtype=bool)
        w = 0
        for sec in self.grid_structure:
            for _ in range(sec.num_repetitions):
                for r in sec.pattern:
                    m = self.grid_resources[r].mask
                    mask[:, w, :m.shape[1]] = m
                    w += 1
        return mask

    @property
    def num_data_symbols(self) -> int:
        return int(self._mask[ElementType.DATA.value].sum())

    def bits_per_frame(self) -> int:
        return self.num_data_symbols * self.bits_per_symbol

    def samples_per_frame(self, bandwidth: float = 0.0, oversampling_factor: int = 1) -> int:
        return self.words_per_frame * (self.num_subcarriers + self.prefix)

    # ------------------------------------------------------------ Tx
    def map(self, bits) -> Symbols:
        """``OrthogonalWaveform.map`` (``waveform.py:1214-1215``): bits -> Symbols (1, K, 1)."""

        if not isinstance(bits, torch.Tensor):
            bits = torch.as_tensor(np.asarray(bits, dtype=np.uint8), device="cuda")
        return Symbols(self.mapping.get_symbols(bits.reshape(-1))[None, :, None])

    def place(self, symbols):
        """Data symbols -> grids (``waveform.py:1222-1258``); reference elements stay zero.  A
        ``Symbols`` of (streams, K, 1) or (streams, 1, K) gives ``Symbols`` (streams, words, N);
        a (..., K) tensor gives a (..., words, N) tensor."""

        typed = isinstance(symbols, Symbols)
        data = symbols.raw.reshape(symbols.num_streams, -1) if typed else symbols
        if data.shape[-1] != self.num_data_symbols:
            raise ValueError("Number of provided data symbols does not match the number of symbols in the frame "
                             f"({data.shape[-1]} != {self.num_data_symbols})")
        lead = data.shape[:-1]
        grid = torch.zeros(lead + (self.words_per_frame * self.num_subcarriers,), dtype=torch.complex64,
                           device=data.device)
        grid[..., self._data_index.to(data.device)] = data.to(torch.complex64)
        grid = grid.reshape(lead + (self.words_per_frame, self.num_subcarriers))
        return Symbols(grid) if typed else grid

    def modulate(self, grid, bandwidth: float | None = None, oversampling_factor: int = 1) -> torch.Tensor:
        """(streams, words, N) grids (tensor or ``Symbols``) -> (streams, samples_per_frame) frames
        (``waveform.py:1293-1323``)."""

        if oversampling_factor != 1:
            raise ValueError("the device path runs at oversampling factor 1")
        if isinstance(grid, Symbols):
            grid = grid.raw
        grid = grid.to(torch.complex64).contiguous()
        streams = grid.shape[0] if grid.ndim == 3 else 1
        m = self.samples_per_frame()
        out = torch.empty((streams, m), dtype=torch.complex64, device=grid.device)
        _lib.check(_lib.load().hpx_ofdm_modulate(_lib.ptr(grid), streams, self.words_per_frame, self.num_subcarriers,
                                                 self.prefix, int(self.dc_suppression), _lib.ptr(out), m,
                                                 _lib.stream_handle()))
        return out

    # ------------------------------------------------------------ Rx
    def demodulate(self, signal, bandwidth: float | None = None, oversampling_factor: int = 1) -> Symbols:
        """(streams, >= samples_per_frame) samples -> ``Symbols`` (streams, words, N), frame at
        offset 0 (``waveform.py:1326-1364``, ``waveform.py:90-99``)."""

        if oversampling_factor != 1:
            raise ValueError("the device path runs at oversampling factor 1")
        if not isinstance(signal, torch.Tensor):
            signal = getattr(signal, "samples", signal)
        if not isinstance(signal, torch.Tensor):
            signal = torch.as_tensor(np.asarray(signal, dtype=np.complex64), device="cuda")
        signal = signal.to(torch.complex64)
        if signal.ndim == 1:
            signal = signal[None]
        m = self.samples_per_frame()
        if signal.shape[1] < m:
            signal = torch.nn.functional.pad(signal, (0, m - signal.shape[1]))
        signal = signal.contiguous()
        out = torch.empty((signal.shape[0], self.words_per_frame, self.num_subcarriers), dtype=torch.complex64,
                          device=signal.device)
        _lib.check(_lib.load().hpx_ofdm_demodulate(_lib.ptr(signal), signal.shape[0], signal.shape[1],
                                                   self.words_per_frame, self.num_subcarriers, self.prefix,
                                                   int(self.dc_suppression), _lib.ptr(out), _lib.stream_handle()))
        return Symbols(out)

    def pick(self, grid):
        """Data resource elements (``waveform.py:1261-1290``): ``StatedSymbols`` (Nrx, words, N)
        with states (Nrx, Ntx, words, N) -> ``StatedSymbols`` (Nrx, 1, K) with states
        (Nrx, Ntx, 1, K); a (..., words, N) tensor -> a (..., K) tensor."""

        if isinstance(grid, StatedSymbols):
            idx = self._data_index.to(grid.raw.device)
            raw = grid.raw.reshape(grid.num_streams, -1)[:, idx]
            st = grid.states.reshape(grid.states.shape[0], grid.states.shape[1], -1)[:, :, idx]
            return StatedSymbols(raw[:, None, :], st[:, :, None, :])
        flat = grid.reshape(grid.shape[:-2] + (-1,))
        return flat[..., self._data_index.to(grid.device)]

    def unmap(self, symbols) -> torch.Tensor:
        """Hard decisions of every symbol, stream-major (``waveform.py:1218-1219``)."""

        if isinstance(symbols, Symbols):
            symbols = symbols.raw
        return self.mapping.detect_bits(symbols)


class OFDMIdealChannelEstimation:
    """Ideal CSI of an OFDM frame (``channel_estimation.py:182-295``): the state of the most recent
    sample of ``channel`` from ``transmitter`` to ``receiver`` (cached through the channel's
    sample hook, ``channel_estimation.py:30-78``), symbol-averaged and transformed per word."""

    def __init__(self, channel=None, transmitter=None, receiver=None, waveform: "OFDMWaveform | None" = None) -> None:
        if isinstance(channel, OFDMWaveform):  # estimation of explicitly given samples only
            channel, waveform = None, channel
        self.waveform = waveform
        self.channel = channel
        self.transmitter = transmitter
        self.receiver = receiver
        self._cached_sample = None
        if channel is not None:
            channel.add_sample_hook(self._sample_callback)

    def _sample_callback(self, sample, transmitter_id: int, receiver_id: int) -> None:
        if ((self.transmitter is None or transmitter_id == self.transmitter.device_id)
                and (self.receiver is None or receiver_id == self.receiver.device_id)):
            self._cached_sample = sample

    def csi(self, sample) -> torch.Tensor:
        """(Nrx, Ntx, words, N) ideal CSI of a fading sample (``hpx_ideal_csi``)."""

        wf = self.waveform
        if wf is None:
            raise RuntimeError("Ideal channel state estimation routine floating")
        d = sample._device_tables()
        nrx, ntx = sample.num_receive_antennas, sample.num_transmit_antennas
        out = torch.empty((nrx, ntx, wf.words_per_frame, wf.num_subcarriers), dtype=torch.complex64,
                          device=sample.device)
        m = wf.samples_per_frame()
        num_taps = min(1 + round(float(sample.delay_profile.max()) * sample.bandwidth), m)  # fading.py:351
        _lib.check(_lib.load().hpx_ideal_csi(_lib.ptr(d["sos"]), sample.num_paths, sample.num_sinusoids + 1,
                                             _lib.ptr(d["delays"]), num_taps, _lib.ptr(d["spatial"]), nrx, ntx,
                                             wf.words_per_frame, wf.num_subcarriers, wf.prefix,
                                             int(wf.dc_suppression), _lib.ptr(out), _lib.stream_handle()))
        return out

    def estimate_channel(self, symbols, bandwidth: float | None = None, oversampling_factor: int = 1,
                         delay: float = 0.0):
        """``estimate_channel(symbols, bandwidth, oversampling_factor, delay) -> StatedSymbols``
        (``channel_estimation.py:252-295``): the received ``Symbols`` (Nrx, words, N) with the ideal
        states of the cached sample.  Given a fading sample instead, its (Nrx, Ntx, words, N) CSI."""

        if not isinstance(symbols, Symbols):
            return self.csi(symbols)
        if oversampling_factor != 1:
            raise ValueError("the device path runs at oversampling factor 1")
        if self._cached_sample is None:
            raise RuntimeError("No channel sample available from which to estimate the ideal channel state "
                               "information")
        h = self.csi(self._cached_sample)[: symbols.num_streams]
        return StatedSymbols(symbols.raw, h)


def _equalize(y: torch.Tensor, h: torch.Tensor, mode: int, noise_variance: float) -> torch.Tensor:
    y = y.to(torch.complex64).contiguous()
    h = h.to(torch.complex64).contiguous()
    nrx, ntx, k = h.shape
    out = torch.empty((ntx, k), dtype=torch.complex64, device=y.device)
    _lib.check(_lib.load().hpx_equalize(_lib.ptr(y), _lib.ptr(h), nrx, ntx, k, mode, float(noise_variance),
                                        _lib.ptr(out), _lib.stream_handle()))
    return out


def _equalize_stated(stated: StatedSymbols, mode: int, noise_variance: float) -> Symbols:
    y = stated.raw.reshape(stated.num_streams, -1)
    h = stated.dense_states()
    h = h.reshape(h.shape[0], h.shape[1], -1)
    if stated.num_streams < 2:
        # a single receive stream divides by the sum of its transmit states (waveform.py:240-243)
        h = h.sum(dim=1, keepdim=True)
    x = _equalize(y, h, mode, noise_variance)
    return Symbols(x.reshape(x.shape[0], stated.num_blocks, stated.num_symbols))


class ZeroForcingChannelEqualization:
    """``waveform.py:235-250``: ``equalize_channel(StatedSymbols) -> Symbols`` (pinv of every
    element's (Nrx, Ntx) state; one receive stream: y / sum_t H_t); also y (Nrx, K), h (Nrx, Ntx, K)
    -> (Ntx, K) tensors."""

    def __init__(self, waveform=None) -> None:
        self.waveform = waveform

    def equalize_channel(self, stated_symbols, h: torch.Tensor | None = None):
        if isinstance(stated_symbols, StatedSymbols):
            return _equalize_stated(stated_symbols, _lib.HPX_DETECT_ZF, 0.0)
        return _equalize(stated_symbols, h, _lib.HPX_DETECT_ZF, 0.0)


class MMSEChannelEqualization:
    """Linear MMSE (H^H H + sigma^2 I)^{-1} H^H y; no reference counterpart for OFDM.  Same call
    forms as ``ZeroForcingChannelEqualization``."""

    def __init__(self, noise_variance: float = 0.0, waveform=None) -> None:
        self.noise_variance = float(noise_variance)
        self.waveform = waveform

    def equalize_channel(self, stated_symbols, h: torch.Tensor | None = None):
        if isinstance(stated_symbols, StatedSymbols):
            return _equalize_stated(stated_symbols, _lib.HPX_DETECT_MMSE, self.noise_variance)
        return _equalize(stated_symbols, h, _lib.HPX_DETECT_MMSE, self.noise_variance)


class BitErrorArtifact:
    """``evaluators.py:204-210``: the bit error rate of one evaluation."""

    def __init__(self, value: float) -> None:
        self.value = float(value)

    def to_scalar(self) -> float:
        return self.value

    def __str__(self) -> str:
        return f"{self.value:.3E}"


class BitErrorEvaluation:
    """``evaluators.py:213-233``: bit errors of one transmission/reception pair, kept as the
    device count of differing bits (``errors``, int64) over ``num_bits``."""

    title = "Bit Error Evaluation"

    def __init__(self, errors: torch.Tensor, num_bits: int, transmitted: torch.Tensor, received: torch.Tensor) -> None:
        self.errors = errors
        self.num_bits = int(num_bits)
        self._tx, self._rx = transmitted, received

    @property
    def evaluation(self) -> torch.Tensor:
        """The per-bit error indicator |tx - rx| of the padded sequences."""

        return (self._tx != self._rx).to(torch.int8)

    def artifact(self) -> BitErrorArtifact:
        return BitErrorArtifact(float(self.errors.sum().item()) / max(self.num_bits, 1))


class BitErrorEvaluator:
    """``evaluators.py:236-270``: bit errors between transmitted and received bits.

    confidence, tolerance and min_num_samples configure the early stopping of the campaign
    result (``CommunicationEvaluator.__init__``, ``evaluators.py:54-85``, and
    ``ScalarEvaluator``, ``hermespy/core/pymonte/scalar.py:263-372``); the modems are accepted
    for signature compatibility (the fused drop path evaluates its own link)."""

    abbreviation = "BER"
    title = "Bit Error Rate Evaluation"

    def __init__(self, transmitting_modem=None, receiving_modem=None, confidence: float = 1.0,
                 tolerance: float = 0.0, min_num_samples: int = 1024, plot_scale: str = "log",
                 tick_format: str = "lin", plot_surface: bool = True) -> None:
        self.transmitting_modem = transmitting_modem
        self.receiving_modem = receiving_modem
        self.confidence = confidence
        self.tolerance = tolerance
        self.min_num_samples = min_num_samples
        self.plot_scale = plot_scale
        self.tick_format = tick_format
        self.plot_surface = plot_surface

    @property
    def confidence(self) -> float:
        return self._confidence

    @confidence.setter
    def confidence(self, value: float) -> None:
        if value < 0.0 or value > 1.0:
            raise ValueError("Confidence level must be in the interval between zero and one")
        self._confidence = float(value)

    @property
    def tolerance(self) -> float:
        return self._tolerance

    @tolerance.setter
    def tolerance(self, value: float) -> None:
        if value < 0.0:
            raise ValueError("Tolerance must be greater or equal to zero")
        self._tolerance = float(value)

    @property
    def min_num_samples(self) -> int:
        return self._min_num_samples

    @min_num_samples.setter
    def min_num_samples(self, value: int) -> None:
        if value < 2:
            raise ValueError("Minimum number of samples must be at least two")
        self._min_num_samples = int(value)

    def initialize_result(self, grid, device="cuda"):
        """``ScalarEvaluator.initialize_result`` (``scalar.py:286-290``) on device counters."""

        from .campaign import ScalarEvaluationResult

        return ScalarEvaluationResult(grid, self, self.tolerance, self.confidence, self.min_num_samples,
                                      self.plot_surface, device=device)

    def evaluate(self, transmitted=None, received=None, out: torch.Tensor | None = None) -> BitErrorEvaluation:
        """``BitErrorEvaluator.evaluate`` (``evaluators.py:239-258``): with no arguments, the bits of
        the transmitting modem's last transmission and the receiving modem's last reception
        (``transmitting_modem.transmission.bits`` / ``receiving_modem.reception.bits``); the
        shorter sequence is zero-padded.  ``out`` (int64 device tensor) accumulates the count."""

        if transmitted is None:
            transmitted = self.transmitting_modem.transmission.bits
        if received is None:
            received = self.receiving_modem.reception.bits
        dev = out.device if out is not None else (transmitted.device if isinstance(transmitted, torch.Tensor)
                                                  else torch.device("cuda"))
        a = torch.as_tensor(transmitted, device=dev).reshape(-1).to(torch.uint8)
        b = torch.as_tensor(received, device=dev).reshape(-1).to(torch.uint8)
        n = max(a.numel(), b.numel())
        if a.numel() < n:  # zero-padding of the shorter sequence (evaluators.py:246-252)
            a = torch.nn.functional.pad(a, (0, n - a.numel()))
        if b.numel() < n:
            b = torch.nn.functional.pad(b, (0, n - b.numel()))
        a, b = a.contiguous(), b.contiguous()
        acc = torch.zeros(1, dtype=torch.int64, device=a.device) if out is None else out
        if acc.dtype != torch.int64 or acc.device != a.device or not acc.is_contiguous():
            raise ValueError("out must be a contiguous int64 tensor on the bits' device")
        _lib.require_device(a, "bits")
        _lib.check(_lib.load().hpx_count_errors(_lib.ptr(a), _lib.ptr(b), n, _lib.ptr(acc), _lib.stream_handle()))
        return BitErrorEvaluation(acc, n, a, b)
