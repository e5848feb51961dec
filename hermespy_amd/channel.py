"""Statistical multipath fading channel with the reference's API, propagating on the GPU.

Mirrors ``hermespy/channel/fading/fading.py``: ``CustomAntennaCorrelation``
(:92-133), ``MultipathFadingChannel`` (:585-910), ``MultipathFadingRealization``
(:425-515) and ``MultipathFadingSample`` (:147-406).  Realisations are drawn on
the host exactly as the reference draws them -- one ``los_angle`` uniform at
construction (:713) and ``standard_normal(100 + 2L + 2LN)`` per ``realize()``
from the channel's NumPy generator (``hermespy/channel/consistent.py:410-432``)
-- so a seeded channel yields the reference's realisation.  Taps, propagation
and the dense impulse response run in HIP through libhermespy_amd.
"""

from __future__ import annotations

import ctypes
from enum import Enum

import numpy as np
import torch
from scipy.special import ndtr

from . import _lib
from .core import (ChannelStateFormat, ChannelStateInformation, InterpolationMode, LinkState, Signal,
                   SimulatedDevice, device_state)


class AntennaCorrelation:
    """Base class of antenna correlation models (``fading.py:44-89``)."""

    def sample_covariance(self, num_antennas: int) -> np.ndarray:  # pragma: no cover - interface
        raise NotImplementedError


class CustomAntennaCorrelation(AntennaCorrelation):
    """Fixed Hermitian positive definite covariance (``fading.py:92-133``)."""

    def __init__(self, covariance: np.ndarray) -> None:
        self.covariance = covariance

    @property
    def covariance(self) -> np.ndarray:
        return self.__covariance

    @covariance.setter
    def covariance(self, value: np.ndarray) -> None:
        value = np.asarray(value)
        if value.ndim != 2 or not np.allclose(value, value.T.conj()):
            raise ValueError("Antenna correlation must be a hermitian matrix")
        if np.any(np.linalg.eigvals(value) <= 0.0):
            raise ValueError("Antenna correlation matrix must be positive definite")
        self.__covariance = value

    def sample_covariance(self, num_antennas: int) -> np.ndarray:
        if self.__covariance.shape[0] < num_antennas:
            raise ValueError("Antenna correlation matrix does not match the number of antennas")
        return self.__covariance[:num_antennas, :num_antennas]


def _as_device_signal(signal, device) -> torch.Tensor:
    if isinstance(signal, torch.Tensor):
        t = signal.to(device=device, dtype=torch.complex64)
    else:
        t = torch.as_tensor(np.asarray(signal, dtype=np.complex64), device=device)
    if t.ndim == 1:
        t = t[None, :]
    return t.contiguous()


class MultipathFadingSample:
    """One sampled link of a fading realisation (``fading.py:147-406``)."""

    def __init__(self, power_profile, delay_profile, los_angles, nlos_angles, los_phases, nlos_phases, los_gains,
                 nlos_gains, los_doppler, nlos_doppler, spatial_response, gain, bandwidth, device="cuda",
                 carrier_frequency: float = 0.0) -> None:
        self.power_profile = np.asarray(power_profile, dtype=np.float64)
        self.delay_profile = np.asarray(delay_profile, dtype=np.float64)
        self.los_angles = np.asarray(los_angles, dtype=np.float64)
        self.nlos_angles = np.asarray(nlos_angles, dtype=np.float64)
        self.los_phases = np.asarray(los_phases, dtype=np.float64)
        self.nlos_phases = np.asarray(nlos_phases, dtype=np.float64)
        self.los_gains = np.asarray(los_gains, dtype=np.float64)
        self.nlos_gains = np.asarray(nlos_gains, dtype=np.float64)
        self.los_doppler = float(los_doppler)
        self.nlos_doppler = float(nlos_doppler)
        self.spatial_response = np.asarray(spatial_response, dtype=np.complex128)
        self.gain = float(gain)
        self.bandwidth = float(bandwidth)
        self.carrier_frequency = float(carrier_frequency)
        self.device = torch.device(device)
        self._dev_cache: dict | None = None

    # ------------------------------------------------------------ properties
    @property
    def num_paths(self) -> int:
        return int(self.delay_profile.size)

    @property
    def num_sinusoids(self) -> int:
        return int(self.nlos_angles.shape[1])

    @property
    def num_receive_antennas(self) -> int:
        return int(self.spatial_response.shape[0])

    @property
    def num_transmit_antennas(self) -> int:
        return int(self.spatial_response.shape[1])

    @property
    def expected_energy_scale(self) -> float:
        """``fading.py:289-291``."""

        return self.gain * float(np.sum(self.power_profile))

    @property
    def delay_samples(self) -> np.ndarray:
        return np.rint(self.delay_profile * self.bandwidth).astype(np.int32)  # fading.py:297

    @property
    def max_delay_samples(self) -> int:
        return int(round(float(self.delay_profile.max()) * self.bandwidth))  # fading.py:372

    def sos_table(self) -> np.ndarray:
        """[L][N + 1][3] (rate rad/sample, phase, amplitude): the NLOS sinusoids of
        ``fading.py:326-336`` followed by the specular term of :339, scaled as :342."""

        L, N = self.num_paths, self.num_sinusoids
        n = 1 + np.arange(N)
        scale = np.sqrt(self.gain * self.power_profile)
        t = np.empty((L, N + 1, 3), dtype=np.float64)
        t[:, :N, 0] = self.nlos_doppler * np.cos((2 * np.pi * n[None, :] + self.nlos_angles) / N) / self.bandwidth
        t[:, :N, 1] = self.nlos_phases
        t[:, :N, 2] = (self.nlos_gains * scale)[:, None]
        t[:, N, 0] = self.los_doppler * np.cos(self.los_angles) / self.bandwidth
        t[:, N, 1] = self.los_phases
        t[:, N, 2] = self.los_gains * scale
        return t

    def _device_tables(self):
        if self._dev_cache is None:
            tab = self.sos_table()
            self._dev_cache = {
                "sos": torch.as_tensor(tab, device=self.device).contiguous(),
                "delays": torch.as_tensor(self.delay_samples, device=self.device),
                "spatial": torch.as_tensor(np.ascontiguousarray(self.spatial_response, dtype=np.complex64),
                                           device=self.device),
                "max_rate": float(np.max(np.abs(tab[:, :, 0]))) if tab.size else 0.0,
            }
        return self._dev_cache

    # ------------------------------------------------------------ kernels
    def taps(self, num_samples: int) -> torch.Tensor:
        """All path impulses (L, num_samples) complex64, direct SoS evaluation (``fading.py:293-343``)."""

        d = self._device_tables()
        out = torch.empty((self.num_paths, int(num_samples)), dtype=torch.complex64, device=self.device)
        _lib.check(_lib.load().hpx_fading_taps(_lib.ptr(d["sos"]), self.num_paths, self.num_sinusoids + 1,
                                               int(num_samples), _lib.ptr(out), _lib.stream_handle()))
        return out

    def _propagate(self, x: torch.Tensor) -> torch.Tensor:
        """``_propagate`` (``fading.py:371-406``): (Ntx, M) -> (Nrx, M + max delay), complex64 on the device."""

        d = self._device_tables()
        m = int(x.shape[1])
        dmax = self.max_delay_samples
        y = torch.empty((self.num_receive_antennas, m + dmax), dtype=torch.complex64, device=self.device)
        if y.numel() == 0:
            return y
        _lib.check(_lib.load().hpx_tdl_propagate(
            _lib.ptr(x), self.num_transmit_antennas, m, m, _lib.ptr(d["sos"]), self.num_paths, self.num_sinusoids + 1,
            d["max_rate"], _lib.ptr(d["delays"]), dmax, _lib.ptr(d["spatial"]), self.num_receive_antennas,
            _lib.ptr(y), m + dmax, _lib.stream_handle()))
        return y

    def propagate(self, signal, interpolation_mode: InterpolationMode = InterpolationMode.NEAREST):
        """``ChannelSample.propagate`` (``channel.py:306-379``) over ``_propagate``: a ``Signal``
        gives a ``Signal`` at the sample's bandwidth (empty when the expected energy scale is
        zero); a bare (Ntx, M) array or tensor gives the (Nrx, M + max delay) tensor."""

        if isinstance(signal, Signal):
            if self.expected_energy_scale <= 0.0:
                return Signal.Empty(signal.sampling_rate, self.num_receive_antennas, 0, signal.carrier_frequency,
                                    signal.noise_power, signal.delay, device=self.device)
            if signal.num_streams != self.num_transmit_antennas:
                raise ValueError("Number of signal streams to be propagated does not match the number of transmitter "
                                 f"antennas ({signal.num_streams} != {self.num_transmit_antennas}))")
            x = _as_device_signal(signal.samples, self.device)
            return Signal(self._propagate(x), self.bandwidth, self.carrier_frequency, signal.noise_power, signal.delay)
        x = _as_device_signal(signal, self.device)
        if x.shape[0] != self.num_transmit_antennas:
            raise ValueError(f"Number of signal streams ({x.shape[0]}) does not match the number of transmit "
                             f"antennas ({self.num_transmit_antennas})")
        return self._propagate(x)

    def state(self, num_samples: int, max_num_taps: int,
              interpolation_mode: InterpolationMode = InterpolationMode.NEAREST) -> ChannelStateInformation:
        """Dense impulse response (Nrx, Ntx, num_samples, taps) (``fading.py:345-369``): taps =
        min(1 + max delay, max_num_taps); a path whose delay exceeds the last tap is left out
        (``hpx_fading_state``)."""

        num_taps = min(1 + round(float(self.delay_profile.max()) * self.bandwidth), int(max_num_taps))
        if num_taps < 1:
            raise ValueError("the channel state needs at least one tap")
        d = self._device_tables()
        nrx, ntx = self.num_receive_antennas, self.num_transmit_antennas
        csi = torch.empty((nrx, ntx, int(num_samples), num_taps), dtype=torch.complex64, device=self.device)
        _lib.check(_lib.load().hpx_fading_state(_lib.ptr(d["sos"]), self.num_paths, self.num_sinusoids + 1,
                                                _lib.ptr(d["delays"]), num_taps, _lib.ptr(d["spatial"]), nrx, ntx,
                                                int(num_samples), _lib.ptr(csi), _lib.stream_handle()))
        return ChannelStateInformation(ChannelStateFormat.IMPULSE_RESPONSE, csi, num_delay_taps=num_taps)


class MultipathFadingRealization:
    """Realisation of a fading channel (``fading.py:425-515``)."""

    def __init__(self, scalars: np.ndarray, channel: "MultipathFadingChannel") -> None:
        self.scalars = np.asarray(scalars, dtype=np.float64)
        self.channel = channel

    def sample(self, transmitter, receiver, timestamp: float = 0.0, *args, device="cuda") -> MultipathFadingSample:
        """``ChannelRealization.sample`` (``channel.py:530-577``): the link from ``transmitter`` to
        ``receiver`` (``SimulatedDevice`` or ``SimulatedDeviceState``, both of one kind) at
        ``timestamp``; optional positional carrier frequency and bandwidth default to the
        transmitter's carrier frequency and sampling rate."""

        if isinstance(transmitter, SimulatedDevice) != isinstance(receiver, SimulatedDevice):
            raise ValueError("Invalid input argument types for channel sampling.")
        tx = device_state(transmitter, timestamp)
        rx = device_state(receiver, timestamp)
        carrier = float(args[0]) if len(args) > 0 and args[0] is not None else tx.carrier_frequency
        bandwidth = float(args[1]) if len(args) > 1 and args[1] is not None else tx.sampling_rate
        sample = self._sample(LinkState(tx, rx, carrier, bandwidth, timestamp), device)
        for hook in list(self.channel.sample_hooks):
            hook(sample, tx.device_id, rx.device_id)
        return sample

    def _sample(self, state: LinkState, device="cuda") -> MultipathFadingSample:
        """``MultipathFadingRealization._sample`` (``fading.py:468-515``)."""

        ch = self.channel
        num_transmit_antennas = state.transmitter.num_digital_transmit_ports
        num_receive_antennas = state.receiver.num_digital_receive_ports
        L, N = ch.num_resolvable_paths, ch.num_sinusoids
        if num_receive_antennas > 10 or num_transmit_antennas > 10:
            raise ValueError("the fading spatial response is drawn on a 10 x 10 grid (at most 10 antennas)")
        u = ndtr(self.scalars)  # ConsistentUniform.sample, consistent.py:485
        o = 0
        corr = u[o:o + 100].reshape(10, 10); o += 100
        los_u = u[o:o + L]; o += L
        nlos_ang_u = u[o:o + L * N].reshape(L, N); o += L * N
        los_ph_u = u[o:o + L]; o += L
        nlos_ph_u = u[o:o + L * N].reshape(L, N)
        spatial = np.exp(2j * np.pi * corr)[:num_receive_antennas, :num_transmit_antennas]
        if ch.antenna_correlation is not None:
            corr = ch.antenna_correlation
            if isinstance(corr, StandardAntennaCorrelation):
                rrx = corr.covariance_for(num_receive_antennas, True)
                rtx = corr.covariance_for(num_transmit_antennas, False)
            else:
                rrx = corr.sample_covariance(num_receive_antennas)
                rtx = corr.sample_covariance(num_transmit_antennas)
            spatial = rrx @ spatial @ rtx
        return MultipathFadingSample(ch.power_profile, ch.delays, 2 * np.pi * los_u, -np.pi + 2 * np.pi * nlos_ang_u,
                                     -np.pi + 2 * np.pi * los_ph_u, -np.pi + 2 * np.pi * nlos_ph_u, ch.los_gains,
                                     ch.nlos_gains, ch.los_doppler_frequency, ch.doppler_frequency, spatial, ch.gain,
                                     state.bandwidth, device, state.carrier_frequency)

    def reciprocal_sample(self, sample: MultipathFadingSample, transmitter, receiver, carrier_frequency=None,
                          bandwidth=None) -> MultipathFadingSample:
        """``_reciprocal_sample`` (``fading.py:517-538``): the same paths over the transposed
        spatial response (antennas are not re-drawn for the reverse direction)."""

        device_state(transmitter)
        device_state(receiver)
        return MultipathFadingSample(sample.power_profile, sample.delay_profile, sample.los_angles,
                                     sample.nlos_angles, sample.los_phases, sample.nlos_phases, sample.los_gains,
                                     sample.nlos_gains, sample.los_doppler, sample.nlos_doppler,
                                     sample.spatial_response.T, sample.gain,
                                     sample.bandwidth if bandwidth is None else float(bandwidth), sample.device,
                                     sample.carrier_frequency if carrier_frequency is None else float(carrier_frequency))


class MultipathFadingChannel:
    """Statistical multipath fading channel (``fading.py:585-910``)."""

    _DEFAULT_NUM_SINUSOIDS = 20

    def __init__(self, delays, power_profile, rice_factors, correlation_distance: float = float("inf"),
                 num_sinusoids: int = _DEFAULT_NUM_SINUSOIDS, los_angle: float | None = None,
                 doppler_frequency: float = 0.0, los_doppler_frequency: float | None = None,
                 antenna_correlation: AntennaCorrelation | None = None, gain: float = 1.0,
                 seed: int | None = None) -> None:
        d = np.array(delays) if isinstance(delays, list) else np.asarray(delays)
        p = np.array(power_profile) if isinstance(power_profile, list) else np.asarray(power_profile)
        k = np.array(rice_factors) if isinstance(rice_factors, list) else np.asarray(rice_factors)
        if d.ndim != 1 or p.ndim != 1 or k.ndim != 1:
            raise ValueError("Delays, power profile and rice factors must be vectors")
        if len(d) < 1:
            raise ValueError("Configuration must contain at least one delay tap")
        if len(d) != len(p) or len(p) != len(k):
            raise ValueError("Delays, power profile and rice factor vectors must be of equal length")
        if np.any(d < 0.0):
            raise ValueError("Delays must be greater or equal to zero")
        if np.any(p < 0.0):
            raise ValueError("Power profile factors must be greater or equal to zero")
        if np.any(k < 0.0):
            raise ValueError("Rice factors must be greater or equal to zero")
        if correlation_distance < 0:
            raise ValueError("Correlation distance must be greater or equal to zero")
        if correlation_distance != float("inf"):
            raise ValueError("only spatially static channels (correlation_distance = inf) run on this path")
        self.gain = float(gain)
        self.seed = seed  # RandomNode: a fresh generator, core/random_node.py
        order = np.argsort(d)
        self.delays = d[order].astype(np.float64)
        self.power_profile = p[order].astype(np.float64)
        self.rice_factors = k[order].astype(np.float64)
        self.num_sinusoids = int(num_sinusoids)
        self.los_angle = self._rng.uniform(-np.pi, np.pi) if los_angle is None else los_angle  # fading.py:713
        self.doppler_frequency = float(doppler_frequency)
        self.__los_doppler = los_doppler_frequency
        self.antenna_correlation = antenna_correlation
        self.correlation_distance = correlation_distance
        inf = np.isposinf(self.rice_factors)
        self.los_gains = np.ones(self.delays.size)
        self.los_gains[~inf] = np.sqrt(self.rice_factors[~inf] / (1 + self.rice_factors[~inf]))
        self.nlos_gains = np.zeros(self.delays.size)
        self.nlos_gains[~inf] = np.sqrt(1 / ((1 + self.rice_factors[~inf]) * self.num_sinusoids))
        self.realization: MultipathFadingRealization | None = None
        self.sample_hooks: set = set()

    def add_sample_hook(self, hook) -> None:
        """Call ``hook(sample, transmitter_id, receiver_id)`` for every sample of this channel's
        realisations (the sample hooks of ``channel.py:569-575``)."""

        self.sample_hooks.add(hook)

    def remove_sample_hook(self, hook) -> None:
        self.sample_hooks.discard(hook)

    @property
    def seed(self) -> int | None:
        return self.__seed

    @seed.setter
    def seed(self, value: int | None) -> None:
        self.__seed = value
        self._rng = np.random.default_rng(value)

    @property
    def num_resolvable_paths(self) -> int:
        return int(self.delays.size)

    @property
    def max_delay(self) -> float:
        return float(self.delays.max())

    @property
    def los_doppler_frequency(self) -> float:
        return self.doppler_frequency if self.__los_doppler is None else float(self.__los_doppler)

    @property
    def num_scalars(self) -> int:
        L, N = self.num_resolvable_paths, self.num_sinusoids
        return 100 + 2 * L + 2 * L * N

    def realize(self, cache: bool = True) -> MultipathFadingRealization:
        """``Channel.realize`` -> ``_realize`` (``channel.py:877-898``, ``fading.py:893-910``)."""

        real = MultipathFadingRealization(self._rng.standard_normal(self.num_scalars), self)
        if cache:
            self.realization = real
        return real

    def propagate(self, signal, transmitter, receiver, timestamp: float = 0.0,
                  interpolation_mode: InterpolationMode = InterpolationMode.NEAREST) -> Signal:
        """``Channel.propagate`` (``channel.py:910-973``): a new realisation, sampled between the
        devices at the signal's carrier frequency and sampling rate, propagating ``signal``."""

        if not isinstance(signal, Signal):
            raise ValueError("Signal is of unsupported type")
        sample = self.realize().sample(transmitter, receiver, timestamp, signal.carrier_frequency,
                                       signal.sampling_rate, device=signal.device)
        return sample.propagate(signal, interpolation_mode)


# ---------------------------------------------------------------------- parameterisations
_TABLES: dict | None = None


def _tables() -> dict:
    global _TABLES
    if _TABLES is None:
        import json
        from pathlib import Path

        _TABLES = json.loads((Path(__file__).parent / "data" / "fading_tables.json").read_text())
    return _TABLES


class TDLType(Enum):
    """5G TDL model types (``tdl.py:18-25``)."""

    A = 0
    B = 1
    C = 2
    D = 4
    E = 5


class TDL(MultipathFadingChannel):
    """3GPP TR 38.901 TDL-A..E (``tdl.py:28-397``): normalised delays scaled by ``rms_delay``,
    powers normalised to sum one; D and E carry a specular first tap (K 13.3 / 22, LOS Doppler 0.7)."""

    def __init__(self, model_type: TDLType = TDLType.A, rms_delay: float = 0.0, **kwargs) -> None:
        if rms_delay < 0.0:
            raise ValueError("Root-Mean-Squared delay must be greater or equal to zero")
        model_type = TDLType(model_type)
        t = _tables()["TDL"][model_type.name]
        los_doppler = kwargs.pop("los_doppler_frequency", None)
        if t["los_doppler_frequency"] is not None:
            if los_doppler is not None:
                raise ValueError(f"Model type {model_type.name} does not support line of sight doppler frequency "
                                 f"configuration")
            los_doppler = t["los_doppler_frequency"]
        p = 10.0 ** (np.asarray(t["power_db"], dtype=np.float64) / 10)
        p /= np.sum(p)
        k = np.zeros(p.size)
        k[0] = t["rice_factor_first"]
        self.model_type = model_type
        self.rms_delay = float(rms_delay)
        MultipathFadingChannel.__init__(self, rms_delay * np.asarray(t["normalized_delays"], dtype=np.float64), p, k,
                                        los_doppler_frequency=los_doppler, **kwargs)


class Cost259Type(Enum):
    URBAN = 0
    RURAL = 1
    HILLY = 2


class Cost259(MultipathFadingChannel):
    """COST 259 typical urban / rural area / hilly terrain (``cost259.py:35-262``)."""

    def __init__(self, model_type: Cost259Type = Cost259Type.URBAN, **kwargs) -> None:
        model_type = Cost259Type(model_type)
        t = _tables()["COST259"][model_type.name]
        k = np.zeros(len(t["delays_s"]))
        if t["los_first"]:
            if kwargs.get("los_angle") is not None:
                raise ValueError("Model type HILLY does not support line of sight angle configuration")
            k[0] = np.inf
            kwargs["los_angle"] = float(np.arccos(t["los_angle_cos"]))
        p = 10.0 ** (np.asarray(t["power_db"], dtype=np.float64) / 10)
        p /= np.sum(p)
        self.model_type = model_type
        MultipathFadingChannel.__init__(self, np.asarray(t["delays_s"], dtype=np.float64), p, k, **kwargs)


class Exponential(MultipathFadingChannel):
    """Exponential power delay profile (``exponential.py:19-117``): taps every ``tap_interval``,
    decay alpha from the rms delay of a geometric profile, truncated below 1e-5."""

    _TRUNCATION = 1e-5

    def __init__(self, tap_interval: float, rms_delay: float, **kwargs) -> None:
        if tap_interval <= 0.0:
            raise ValueError("Tap interval must be greater than zero")
        if rms_delay <= 0.0:
            raise ValueError("Root-Mean-Squared delay must be greater than zero")
        rms_norm = rms_delay / tap_interval
        alpha = -2 * np.log((-1 + np.sqrt(1 + 4 * rms_norm ** 2)) / (2 * rms_norm))
        taps = int(-np.ceil(np.log(self._TRUNCATION) / alpha))
        self.tap_interval = float(tap_interval)
        self.rms_delay = float(rms_delay)
        MultipathFadingChannel.__init__(self, np.arange(taps + 1) * tap_interval, np.exp(-alpha * np.arange(taps + 1)),
                                        np.zeros(taps + 1), **kwargs)


class CorrelationType(Enum):
    """3GPP correlation levels (base station, terminal) (``correlation.py:30-45``)."""

    LOW = (0.0, 0.0)
    MEDIUM = (0.3, 0.3)
    MEDIUM_A = (0.3, 0.3874)
    HIGH = (0.9, 0.9)


class StandardAntennaCorrelation(AntennaCorrelation):
    """3GPP standard covariance for 1, 2 or 4 antennas (``correlation.py:48-106``).  The transmitter
    is taken as the base station and the receiver as the terminal, as in the reference."""

    def __init__(self, correlation: CorrelationType | str) -> None:
        self.correlation = CorrelationType[correlation] if isinstance(correlation, str) else CorrelationType(correlation)

    def covariance_for(self, num_antennas: int, receive: bool) -> np.ndarray:
        f = self.correlation.value[1 if receive else 0]
        n = num_antennas
        if n == 1:
            return np.ones((1, 1), dtype=complex)
        if n == 2:
            return np.array([[1, f], [f, 1]], dtype=complex)
        if n == 4:
            a, b = f ** (1 / 9), f ** (4 / 9)
            return np.array([[1, a, b, f], [a, 1, a, b], [b, a, 1, a], [f, b, a, 1]], dtype=complex)
        raise RuntimeError(f"3GPP standard antenna covariance is only defined for 1, 2 and 4 antennas, device has "
                           f"{n} antennas")
