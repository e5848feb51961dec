"""Device-resident counterparts of the reference's core value types on the drop path.

* ``Signal`` / ``SignalBlock`` (``hermespy/core/signal_model.py:395-480``, ``Signal.Create``):
  a (streams, samples) complex64 device tensor with its sampling rate, carrier frequency,
  noise power and delay; one block at offset 0.
* ``ChannelStateFormat``, ``ChannelStateInformation`` (``hermespy/core/channel.py:26-545``):
  the dense (Nrx, Ntx, samples, taps) impulse response of a channel sample, propagating
  signals through ``hpx_csi_propagate``.
* ``SimulatedDevice`` / ``SimulatedDeviceState`` and ``LinkState``
  (``hermespy/simulation/simulated_device.py``, ``hermespy/channel/channel.py:107-200``): the
  device facts a channel sample reads -- antenna counts, carrier frequency and sampling rate --
  so that ``realization.sample(transmitter, receiver, ...)`` keeps the reference's call form.
* ``InterpolationMode`` (``hermespy/core/channel.py``; the fading path rounds delays to the
  nearest sample).
"""

from __future__ import annotations

import itertools
from enum import Enum

import numpy as np
import torch

from . import _lib


class InterpolationMode(Enum):
    """Delay interpolation of channel propagation; the fading path rounds to the nearest sample."""

    NEAREST = 0
    SINC = 1


class SignalBlock:
    """One contiguous block of a signal: (streams, samples) samples at an integer offset."""

    def __init__(self, samples: torch.Tensor, offset: int = 0) -> None:
        self.samples = samples
        self.offset = int(offset)

    @property
    def num_streams(self) -> int:
        return int(self.samples.shape[0])

    @property
    def num_samples(self) -> int:
        return int(self.samples.shape[1])


class Signal:
    """Base-band samples of a device, (streams, samples) complex64 on the device."""

    def __init__(self, samples: torch.Tensor, sampling_rate: float, carrier_frequency: float = 0.0,
                 noise_power: float = 0.0, delay: float = 0.0) -> None:
        if samples.ndim == 1:
            samples = samples[None, :]
        if samples.ndim != 2:
            raise ValueError("Signal samples must be a (streams, samples) array")
        if sampling_rate <= 0.0:
            raise ValueError("Sampling rate must be greater than zero")
        self.samples = samples
        self.sampling_rate = float(sampling_rate)
        self.carrier_frequency = float(carrier_frequency)
        self.noise_power = float(noise_power)
        self.delay = float(delay)

    @classmethod
    def Create(cls, samples, sampling_rate: float, carrier_frequency: float = 0.0, noise_power: float = 0.0,
               delay: float = 0.0, device="cuda") -> "Signal":
        """``Signal.Create`` (``signal_model.py``) from an array or tensor of samples."""

        if isinstance(samples, torch.Tensor):
            t = samples.to(dtype=torch.complex64)
        else:
            t = torch.as_tensor(np.asarray(samples, dtype=np.complex64), device=device)
        return cls(t.contiguous(), sampling_rate, carrier_frequency, noise_power, delay)

    @classmethod
    def Empty(cls, sampling_rate: float, num_streams: int = 0, num_samples: int = 0, carrier_frequency: float = 0.0,
              noise_power: float = 0.0, delay: float = 0.0, device="cuda") -> "Signal":
        return cls(torch.zeros((num_streams, num_samples), dtype=torch.complex64, device=device), sampling_rate,
                   carrier_frequency, noise_power, delay)

    @property
    def num_streams(self) -> int:
        return int(self.samples.shape[0])

    @property
    def num_samples(self) -> int:
        return int(self.samples.shape[1])

    @property
    def blocks(self) -> list:
        return [SignalBlock(self.samples, 0)]

    @property
    def device(self) -> torch.device:
        return self.samples.device


class ChannelStateFormat(Enum):
    """``hermespy/core/channel.py:26-31``."""

    IMPULSE_RESPONSE = 0
    FREQUENCY_SELECTIVITY = 1


class ChannelStateInformation:
    """Dense channel state (Nrx, Ntx, samples, taps) of one link in impulse-response format
    (``hermespy/core/channel.py:42-545``)."""

    def __init__(self, state_format: ChannelStateFormat, state: torch.Tensor, num_delay_taps: int | None = None,
                 num_frequency_bins: int | None = None) -> None:
        if state.ndim != 4:
            raise ValueError("Channel state tensors must be four-dimensional")
        if state_format != ChannelStateFormat.IMPULSE_RESPONSE:
            raise ValueError("the device channel state is kept as an impulse response")
        self.__state_format = state_format
        self.__state = state
        self.__num_delay_taps = int(state.shape[3] if num_delay_taps is None else num_delay_taps)
        self.__num_frequency_bins = num_frequency_bins

    @property
    def state_format(self) -> ChannelStateFormat:
        return self.__state_format

    @property
    def state(self) -> torch.Tensor:
        return self.__state

    def dense_state(self) -> torch.Tensor:
        return self.__state

    def to_impulse_response(self) -> "ChannelStateInformation":
        return self

    @property
    def num_receive_streams(self) -> int:
        return int(self.__state.shape[0])

    @property
    def num_transmit_streams(self) -> int:
        return int(self.__state.shape[1])

    @property
    def num_samples(self) -> int:
        return int(self.__state.shape[2])

    @property
    def num_symbols(self) -> int:
        return int(self.__state.shape[2])

    @property
    def num_delay_taps(self) -> int:
        return self.__num_delay_taps

    def propagate(self, signal: Signal) -> Signal:
        """``ChannelStateInformation.propagate`` (``channel.py:493-528``): y[r, n] = sum over
        transmit streams and taps d of state[r, t, n - d, d] x[t, n - d]."""

        st = self.__state.to(torch.complex64).contiguous()
        x = signal.samples.to(torch.complex64).contiguous()
        _lib.require_device(x, "signal samples")
        _lib.require_device(st, "channel state")
        if x.shape[0] != st.shape[1]:
            raise ValueError(f"signal has {x.shape[0]} streams, the channel state {st.shape[1]} transmit streams")
        m, taps = int(x.shape[1]), int(st.shape[3])
        y = torch.empty((st.shape[0], m + taps - 1), dtype=torch.complex64, device=x.device)
        _lib.check(_lib.load().hpx_csi_propagate(_lib.ptr(st), st.shape[0], st.shape[1], st.shape[2], taps,
                                                 _lib.ptr(x), m, m, _lib.ptr(y), y.shape[1],
                                                 _lib.stream_handle(x.device)))
        return Signal(y, signal.sampling_rate, signal.carrier_frequency, signal.noise_power, signal.delay)

    def reciprocal(self) -> "ChannelStateInformation":
        """``channel.py:530-539``: transmit and receive streams swapped."""

        return ChannelStateInformation(self.__state_format, self.__state.transpose(0, 1).contiguous(),
                                       self.__num_delay_taps, self.__num_frequency_bins)


class AntennaMode(Enum):
    TX = 0
    RX = 1
    DUPLEX = 2


class SimulatedDeviceState:
    """What a channel sample reads of a device at one time (``simulated_device.py``,
    ``SimulatedDeviceState``): identity, antenna counts, carrier frequency, sampling rate."""

    def __init__(self, device_id: int, num_transmit_antennas: int, num_receive_antennas: int,
                 carrier_frequency: float, sampling_rate: float, power: float = 1.0, timestamp: float = 0.0) -> None:
        self.device_id = int(device_id)
        self.num_transmit_antennas = int(num_transmit_antennas)
        self.num_receive_antennas = int(num_receive_antennas)
        self.carrier_frequency = float(carrier_frequency)
        self.sampling_rate = float(sampling_rate)
        self.power = float(power)
        self.timestamp = float(timestamp)

    @property
    def num_digital_transmit_ports(self) -> int:
        return self.num_transmit_antennas

    @property
    def num_digital_receive_ports(self) -> int:
        return self.num_receive_antennas


_device_ids = itertools.count()


class SimulatedDevice:
    """A simulated device as far as the drop path reads it: a uniform array of ``antennas``
    elements used for transmission and reception, carrier frequency, sampling rate and
    transmit power (``simulated_device.py:1688-1692`` scales the transmission by
    sqrt(power))."""

    def __init__(self, antennas: int = 1, carrier_frequency: float = 0.0, sampling_rate: float = 1.0,
                 power: float = 1.0) -> None:
        if antennas < 1:
            raise ValueError("a device needs at least one antenna")
        if sampling_rate <= 0.0:
            raise ValueError("Sampling rate must be greater than zero")
        if power < 0.0:
            raise ValueError("Transmit power must be non-negative")
        self.num_antennas = int(antennas)
        self.carrier_frequency = float(carrier_frequency)
        self.sampling_rate = float(sampling_rate)
        self.power = float(power)
        self.device_id = next(_device_ids)

    @property
    def num_transmit_antennas(self) -> int:
        return self.num_antennas

    @property
    def num_receive_antennas(self) -> int:
        return self.num_antennas

    def state(self, timestamp: float = 0.0) -> SimulatedDeviceState:
        return SimulatedDeviceState(self.device_id, self.num_antennas, self.num_antennas, self.carrier_frequency,
                                    self.sampling_rate, self.power, timestamp)


def device_state(device, timestamp: float = 0.0) -> SimulatedDeviceState:
    """The state of a ``SimulatedDevice`` or a ``SimulatedDeviceState`` as given; anything else
    is rejected as the reference rejects it (``channel.py:530-577``)."""

    if isinstance(device, SimulatedDevice):
        return device.state(timestamp)
    if isinstance(device, SimulatedDeviceState):
        return device
    raise ValueError("Invalid input argument types for channel sampling.")


class LinkState:
    """Physical parameters of a link at one time (``hermespy/channel/channel.py:107-200``)."""

    def __init__(self, transmitter: SimulatedDeviceState, receiver: SimulatedDeviceState, carrier_frequency: float,
                 bandwidth: float, time: float) -> None:
        self.transmitter = transmitter
        self.receiver = receiver
        self.carrier_frequency = float(carrier_frequency)
        self.bandwidth = float(bandwidth)
        self.time = float(time)
