"""The reference call forms of the mirror classes, on the host (no GPU): typed symbols and
their checks, ``realization.sample(transmitter, receiver, ...)`` with its defaults, hooks and
reciprocal, and the equalisers' reshaping of ``StatedSymbols`` (``modem/waveform.py:235-250``)
with the device call replaced by a float64 host twin of the same contract."""

import numpy as np
import pytest
import torch

from oracle import equalize as oeq
from oracle import fading


def _channel(**kw):
    from hermespy_amd.channel import MultipathFadingChannel

    return MultipathFadingChannel([0.0, 2e-7, 5e-7], [0.6, 0.3, 0.1], [2.0, 0.0, 0.0], doppler_frequency=300.0,
                                  los_doppler_frequency=120.0, seed=4, **kw)


def test_symbols_expand_and_states_are_checked():
    from hermespy_amd.modem import StatedSymbols, Symbols

    s = Symbols(np.arange(6, dtype=complex), device="cpu")
    assert (s.num_streams, s.num_blocks, s.num_symbols) == (1, 6, 1)
    s2 = Symbols(np.ones((2, 3), dtype=complex), device="cpu")
    assert s2.raw.shape == (2, 3, 1)
    with pytest.raises(ValueError, match="maximum of three dimensions"):
        Symbols(np.ones((1, 1, 1, 1)), device="cpu")
    raw = torch.ones((2, 3, 4), dtype=torch.complex64)
    st = StatedSymbols(raw, torch.ones((2, 2, 3, 4), dtype=torch.complex64))
    assert st.num_transmit_streams == 2 and st.dense_states().shape == (2, 2, 3, 4)
    with pytest.raises(ValueError, match="four-dimensional"):
        StatedSymbols(raw, torch.ones((2, 3, 4), dtype=torch.complex64))
    with pytest.raises(ValueError, match="received streams"):
        StatedSymbols(raw, torch.ones((3, 2, 3, 4), dtype=torch.complex64))
    with pytest.raises(ValueError, match="received blocks"):
        StatedSymbols(raw, torch.ones((2, 2, 2, 4), dtype=torch.complex64))
    with pytest.raises(ValueError, match="block sizes"):
        StatedSymbols(raw, torch.ones((2, 2, 3, 5), dtype=torch.complex64))


def test_sample_between_devices_matches_oracle():
    from hermespy_amd.core import SimulatedDevice

    ch = _channel()
    real = ch.realize()
    tx = SimulatedDevice(2, carrier_frequency=3.5e9, sampling_rate=10e6)
    rx = SimulatedDevice(3, carrier_frequency=3.5e9, sampling_rate=10e6)
    seen = []
    ch.add_sample_hook(lambda s, t, r: seen.append((t, r)))
    s = real.sample(tx, rx, 0.0, device="cpu")
    assert seen == [(tx.device_id, rx.device_id)]
    assert s.bandwidth == 10e6 and s.carrier_frequency == 3.5e9  # the transmitter's by default
    params = fading.make_parameters(ch.delays, ch.power_profile, ch.rice_factors, ch.num_sinusoids,
                                    ch.doppler_frequency, ch.los_doppler_frequency, ch.gain)
    ref = fading.sample_from_normals(params, real.scalars, 3, 2, 10e6)
    np.testing.assert_array_equal(s.spatial_response, ref.spatial_response)
    np.testing.assert_array_equal(s.los_angles, ref.los_angles)
    np.testing.assert_array_equal(s.nlos_phases, ref.nlos_phases)
    assert s.los_doppler == 120.0 and s.nlos_doppler == 300.0
    # positional carrier frequency and bandwidth override the transmitter's (channel.py:538-539)
    s2 = real.sample(tx.state(), rx.state(), 0.0, 2.4e9, 20e6, device="cpu")
    assert s2.bandwidth == 20e6 and s2.carrier_frequency == 2.4e9
    np.testing.assert_array_equal(s2.spatial_response, s.spatial_response)
    # the reciprocal sample transposes the spatial response (fading.py:517-538)
    back = real.reciprocal_sample(s, rx, tx)
    np.testing.assert_array_equal(back.spatial_response, s.spatial_response.T)
    assert back.num_transmit_antennas == 3 and back.num_receive_antennas == 2
    with pytest.raises(ValueError, match="Invalid input argument types"):
        real.sample(tx, rx.state(), device="cpu")
    with pytest.raises(ValueError, match="Invalid input argument types"):
        real.sample(2, 3, device="cpu")


def test_channel_propagate_rejects_bare_arrays():
    from hermespy_amd.core import SimulatedDevice

    with pytest.raises(ValueError, match="unsupported type"):
        _channel().propagate(np.zeros((1, 8)), SimulatedDevice(1), SimulatedDevice(1))


def _host_equalize(y, h, mode, noise_variance):
    """float64 host twin of hpx_equalize: y (Nrx, K), h (Nrx, Ntx, K) -> (Ntx, K)."""

    yy = y.numpy().astype(np.complex128)
    hh = h.numpy().astype(np.complex128)
    from hermespy_amd import _lib

    if mode == _lib.HPX_DETECT_ZF:
        x = oeq.zero_forcing(yy, hh)
    else:
        x = oeq.mmse(yy, hh, noise_variance)
    return torch.as_tensor(x.astype(np.complex64))


@pytest.mark.parametrize("nrx,ntx", [(1, 1), (1, 2), (2, 2), (4, 2)])
def test_equalizers_reshape_stated_symbols(monkeypatch, nrx, ntx):
    from hermespy_amd import modem
    from hermespy_amd.modem import MMSEChannelEqualization, StatedSymbols, Symbols, ZeroForcingChannelEqualization

    monkeypatch.setattr(modem, "_equalize", _host_equalize)
    rng = np.random.default_rng(nrx * 10 + ntx)
    blocks, syms = 3, 5
    y = (rng.standard_normal((nrx, blocks, syms)) + 1j * rng.standard_normal((nrx, blocks, syms)))
    h = (rng.standard_normal((nrx, ntx, blocks, syms)) + 1j * rng.standard_normal((nrx, ntx, blocks, syms)))
    stated = StatedSymbols(torch.as_tensor(y.astype(np.complex64)), torch.as_tensor(h.astype(np.complex64)))
    out = ZeroForcingChannelEqualization().equalize_channel(stated)
    assert isinstance(out, Symbols)
    if nrx < 2:
        # one receive stream: y / sum_t H_t (waveform.py:240-243)
        exp = (y / h.sum(axis=1))
    else:
        # pinv of every element's state, einsum "ijkl,lij->kij" (waveform.py:245-248)
        eq = np.linalg.pinv(h.transpose((2, 3, 0, 1)))
        exp = np.einsum("ijkl,lij->kij", eq, y)
    assert out.raw.shape == exp.shape
    np.testing.assert_allclose(out.raw.numpy(), exp, rtol=1e-4, atol=1e-5)
    mm = MMSEChannelEqualization(0.0).equalize_channel(stated)
    assert mm.raw.shape == exp.shape


def test_simulation_grid_form_arguments():
    """``new_dimension`` / ``num_samples`` / ``add_evaluator`` of the reference's ``MonteCarlo``
    (``monte_carlo.py:503-605``): SNR points as power ratios or dB-carrying values, and the
    reference's argument errors."""

    from hermespy_amd.configs import siso_64
    from hermespy_amd.modem import BitErrorEvaluator
    from hermespy_amd.simulation import BatchedSimulation

    sim = BatchedSimulation(siso_64(), device="cpu")
    dim = sim.new_dimension("noise_level", [1.0, 10.0, 100.0])
    assert np.allclose(sim._snr_db, [0.0, 10.0, 20.0]) and dim.num_sample_points == 3

    class _Db(float):
        value_db = 7.0

    sim.new_dimension("noise_level", [_Db(5.0)])
    assert sim._snr_db == (7.0,)
    with pytest.raises(ValueError, match="noise level"):
        sim.new_dimension("carrier_frequency", [1e9])
    with pytest.raises(ValueError, match="positive"):
        sim.new_dimension("noise_level", [0.0])
    with pytest.raises(ValueError, match="greater than zero"):
        sim.num_samples = 0
    sim.num_samples = 64
    assert sim.num_samples == 64
    with pytest.raises(ValueError, match="bit errors"):
        sim.add_evaluator(object())
    sim.add_evaluator(BitErrorEvaluator())
