"""One Monte Carlo link drop end to end, float64.

Restates the stage sequence the reference runs per drop
(``hermespy/simulation/simulation.py:69-181``, stages listed at :226-246):
channel realisation, bits -> mapping -> OFDM modulation (``modem.py:797-888``),
device power scaling (``simulated_device.py:1688-1692``), fading propagation
(``fading.py:371-406``), AWGN with ``SNR.get_power`` (``rf/noise/level.py:257-259``,
``rf/noise/model.py:151-157``), default synchronisation at offset 0
(``waveform.py:90-99``, ``modem.py:1085-1146``), demodulation, ideal CSI,
data pick, equalisation, hard demapping (``modem.py:1167-1190``) and the bit
error count of ``BitErrorEvaluator.evaluate`` (``evaluators.py:239-258``).

Random numbers come from the counter-based streams of :mod:`oracle.philox`
(the HIP path consumes the same streams).  ``faithful=True`` keeps the
reference's own data structures (dense impulse response, per-path loop,
``pinv``); ``faithful=False`` uses the closed-form CSI.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import csi, equalize, fading, mapping, ofdm, philox


@dataclass
class Link:
    """A link configuration (plain parameters, shared with the tests)."""

    num_rx: int
    num_tx: int
    num_subcarriers: int
    words: int
    prefix_ratio: float
    modulation_order: int
    bandwidth: float
    delays: np.ndarray
    power_profile: np.ndarray
    rice_factors: np.ndarray
    doppler: float
    num_sinusoids: int = 20
    corr_rx: np.ndarray | None = None
    corr_tx: np.ndarray | None = None
    detector: str = "mmse"  # "zf" | "mmse"
    dc_suppression: bool = True
    gain: float = 1.0
    device_power: float = 1.0
    frame: ofdm.Frame | None = None  # None: all-data frame of `words` words
    pilots: np.ndarray | None = None  # (num_tx, num_references) reference symbols of each stream
    los_doppler: float | None = None  # LOS doppler_frequency; None: the NLOS one (fading.py:829-843)
    params: fading.FadingParameters = field(init=False)

    def __post_init__(self) -> None:
        self.params = fading.make_parameters(self.delays, self.power_profile, self.rice_factors, self.num_sinusoids,
                                             self.doppler, self.los_doppler, self.gain, self.corr_rx, self.corr_tx)
        if self.frame is None:
            self.frame = ofdm.simple_frame(self.num_subcarriers, self.words, self.prefix_ratio, self.dc_suppression)
        if self.frame.num_references and (self.pilots is None or np.shape(self.pilots) !=
                                          (self.num_tx, self.frame.num_references)):
            raise ValueError("a frame with reference elements needs (num_tx, num_references) pilot symbols")

    def with_doppler(self, doppler: float) -> "Link":
        """The same link at another (NLOS and LOS) Doppler."""

        other = Link(**{f: getattr(self, f) for f in self.__dataclass_fields__ if f != "params"})
        other.params = fading.make_parameters(self.delays, self.power_profile, self.rice_factors, self.num_sinusoids,
                                              doppler, None, self.gain, self.corr_rx, self.corr_tx)
        return other

    @property
    def bits_per_symbol(self) -> int:
        return int(np.log2(self.modulation_order))

    @property
    def bits_per_stream(self) -> int:
        return self.frame.num_data_symbols * self.bits_per_symbol

    @property
    def bits_per_drop(self) -> int:
        return self.bits_per_stream * self.num_tx

    def noise_power(self, snr_db: float) -> float:
        """``SNR.get_power``: P_ref / snr * (gain * sum(power_profile))."""

        return self.device_power / 10 ** (snr_db / 10) * (self.gain * float(np.sum(self.params.power_profile)))


@dataclass
class DropResult:
    errors: int
    bits: int
    tx_bits: np.ndarray
    rx_bits: np.ndarray
    equalized: np.ndarray
    received: np.ndarray  # time-domain frame after noise (Nrx, M)
    grid: np.ndarray | None = None  # (Nrx, words, N) demodulated grid
    csi: np.ndarray | None = None  # (Nrx, Ntx, words, N) ideal CSI


def realize(link: Link, seed: int, drop: int, normals: np.ndarray | None = None) -> fading.FadingSample:
    """The drop's channel sample: from its Philox realisation normals, or from given normals (the
    scalars of a recorded ``StaticConsistentRealization``, ``consistent.py:244-260``)."""

    z = philox.realization_normals(seed, drop, link.params.num_normals) if normals is None else normals
    return fading.sample_from_normals(link.params, z, link.num_rx, link.num_tx, link.bandwidth)


def transmit(link: Link, seed: int, drop: int):
    """Bits and the (Ntx, M) transmitted frame."""

    bits = np.stack([philox.stream_bits(seed, drop, t, link.bits_per_stream) for t in range(link.num_tx)])
    refs = (lambda t: None) if link.pilots is None else (lambda t: link.pilots[t])
    x = np.stack([ofdm.modulate(link.frame, ofdm.place(link.frame, mapping.get_symbols(bits[t], link.modulation_order),
                                                       refs(t)))
                  for t in range(link.num_tx)])
    return bits, x * link.device_power ** 0.5


def run_drop(link: Link, seed: int, drop: int, snr_db: float, faithful: bool = True,
             noise: bool = True, normals: np.ndarray | None = None) -> DropResult:
    sample = realize(link, seed, drop, normals)
    tx_bits, x = transmit(link, seed, drop)
    y = fading.propagate(sample, x)  # (Nrx, M + D)
    p = link.noise_power(snr_db)
    if noise and p > 0:
        for r in range(link.num_rx):
            y[r] += (0.5 * p) ** 0.5 * philox.noise_normals(seed, drop, r, y.shape[1])
    m = link.frame.samples_per_frame
    rx_frame = y[:, :m]
    grid = ofdm.demodulate(link.frame, rx_frame)  # (Nrx, words, N)
    h = csi.reference_ideal_csi(sample, link.frame) if faithful else csi.closed_form_ideal_csi(sample, link.frame)
    yd = ofdm.pick(link.frame, grid)  # (Nrx, K)
    hd = ofdm.pick(link.frame, h)  # (Nrx, Ntx, K)
    if link.detector == "zf":
        xh = equalize.zero_forcing(yd, hd)
    else:
        xh = equalize.mmse(yd, hd, p if noise else 0.0)
    rx_bits = np.stack([mapping.detect_bits(xh[t], link.modulation_order).astype(np.uint8) for t in range(xh.shape[0])])
    errors = int(np.sum(np.abs(tx_bits.astype(np.int8) - rx_bits.astype(np.int8))))
    return DropResult(errors, int(tx_bits.size), tx_bits, rx_bits, xh, rx_frame, grid, h)
