"""GPU parity of the fading parameterisations and the link parameters the measurement links
leave at their defaults: TDL-D/E (specular first tap, K 13.3 / 22, their LOS
doppler_frequency 0.7, ``channel/fading/tdl.py:101-397``), COST-259 hilly terrain (K = inf
on the first path, ``cost259.py:35-262``), ``StandardAntennaCorrelation`` (MEDIUM_A at 2 and
4 antennas, ``correlation.py:46-106``), an LOS Doppler that differs from the NLOS one, a
channel gain and a device power other than one (``simulated_device.py:1688-1692``).

Each link runs fused drops through ``hpx_run_batch`` and is compared with the float64
oracle on the same random streams (equalised symbols 1e-4, exact demapping of the kernel's
own symbols, error counts within the decision-boundary slack).  ``MultipathFadingSample.state``
is compared with the oracle's dense impulse response, and its ``propagate`` with the
reference's invariant ``propagate == state().propagate``.
"""

import numpy as np
import pytest
import torch

from conftest import oracle_link
from oracle_pool import oracle_drops
from parity import check_drop, rel_err

pytestmark = pytest.mark.gpu

SEED = 20261202
SNRS = [0.0, 15.0, 30.0]


def _links():
    from hermespy_amd.channel import (TDL, Cost259, Cost259Type, CorrelationType, MultipathFadingChannel,
                                      StandardAntennaCorrelation, TDLType)
    from hermespy_amd.configs import LinkConfig

    two_pi = 2 * np.pi
    out = {}
    tdl_d = TDL(TDLType.D, rms_delay=100e-9, doppler_frequency=two_pi * 300.0, seed=1)
    out["tdl-d-2x2"] = LinkConfig.from_channel(tdl_d, 2, 2, 512, 14, 0.07, 16, 15.36e6, detector="mmse")
    tdl_e = TDL(TDLType.E, rms_delay=300e-9, doppler_frequency=two_pi * 500.0, seed=2,
                antenna_correlation=StandardAntennaCorrelation(CorrelationType.MEDIUM_A))
    out["tdl-e-4x4-medium-a"] = LinkConfig.from_channel(tdl_e, 4, 4, 256, 14, 0.125, 64, 7.68e6, detector="mmse")
    hilly = Cost259(Cost259Type.HILLY, doppler_frequency=two_pi * 100.0, seed=3)
    out["cost259-hilly-1x1"] = LinkConfig.from_channel(hilly, 1, 1, 256, 14, 0.125, 4, 3.84e6, detector="zf")
    medium_a = TDL(TDLType.A, rms_delay=50e-9, doppler_frequency=two_pi * 200.0, seed=4,
                   antenna_correlation=StandardAntennaCorrelation("MEDIUM_A"))
    out["tdl-a-2x2-medium-a"] = LinkConfig.from_channel(medium_a, 2, 2, 256, 14, 0.125, 16, 7.68e6, detector="zf")
    k = np.zeros(8)
    k[:3] = [4.0, 1.0, np.inf]
    los = MultipathFadingChannel(np.arange(8) * 65e-9, np.exp(-np.arange(8) / 3.0), k,
                                 doppler_frequency=two_pi * 200.0, los_doppler_frequency=two_pi * 700.0,
                                 gain=0.5, seed=5)
    out["los-doppler-gain-power-2x2"] = LinkConfig.from_channel(los, 2, 2, 512, 14, 0.07, 16, 15.36e6,
                                                                device_power=2.0, detector="mmse")
    return out


LINKS = ["tdl-d-2x2", "tdl-e-4x4-medium-a", "cost259-hilly-1x1", "tdl-a-2x2-medium-a", "los-doppler-gain-power-2x2"]


def test_link_parameters_reach_the_kernel():
    """The links differ from the defaults where they should (so the GPU cases below test
    what they name)."""

    links = _links()
    d = links["tdl-d-2x2"]
    assert d.rice_factors[0] == pytest.approx(13.3) and d.los_doppler == pytest.approx(0.7)
    assert links["tdl-e-4x4-medium-a"].rice_factors[0] == pytest.approx(22.0)
    assert not np.allclose(links["tdl-e-4x4-medium-a"].corr_rx, links["tdl-e-4x4-medium-a"].corr_tx)
    assert np.isposinf(links["cost259-hilly-1x1"].rice_factors[0])
    lk = links["los-doppler-gain-power-2x2"]
    assert lk.gain == 0.5 and lk.device_power == 2.0 and lk.los_doppler == pytest.approx(3.5 * lk.doppler)


@pytest.mark.parametrize("name", LINKS)
def test_link_matches_oracle(cuda, name):
    from hermespy_amd.simulation import BatchedLink

    cfg = _links()[name]
    print(f"{name}: {cfg.num_rx}x{cfg.num_tx}, {cfg.num_subcarriers} subcarriers, {cfg.num_paths} paths", flush=True)
    n = len(SNRS)
    first = 11
    npw = torch.as_tensor(cfg.noise_power(np.asarray(SNRS)), dtype=torch.float64, device="cuda")
    err = torch.empty(n, dtype=torch.int32, device="cuda")
    eq = torch.empty((n, cfg.num_tx, cfg.num_data), dtype=torch.complex64, device="cuda")
    BatchedLink(cfg).run(SEED, first, npw, err, eq)
    torch.cuda.synchronize()
    err, eq = err.cpu().numpy(), eq.cpu().numpy()
    ora = oracle_drops(oracle_link(cfg), SEED, range(first, first + n), SNRS, faithful=False)
    for i in range(n):
        check_drop(eq[i], err[i], ora[i], cfg.modulation_order, 1e-4, (name, SNRS[i]))


def test_state_and_propagation_match_oracle(cuda):
    """Dense impulse response of a Rician, LOS-Doppler, correlated 2x2 sample against the
    oracle's (``fading.py:345-369``), and propagation through it against the sample's own
    time-variant delay line (``test_fading.py:167-175``)."""

    from hermespy_amd.channel import (CorrelationType, MultipathFadingChannel, MultipathFadingRealization,
                                      StandardAntennaCorrelation)
    from hermespy_amd.core import ChannelStateFormat, ChannelStateInformation, Signal, SimulatedDevice
    from oracle import fading

    ch = MultipathFadingChannel([0.0, 1e-7, 2.5e-7, 4e-7], [1.0, 0.5, 0.25, 0.1], [3.0, 0.0, 0.0, np.inf],
                                doppler_frequency=2 * np.pi * 3e3, los_doppler_frequency=2 * np.pi * 5e3, gain=0.7,
                                antenna_correlation=StandardAntennaCorrelation(CorrelationType.MEDIUM_A), seed=9)
    real = ch.realize()
    tx = SimulatedDevice(antennas=2, sampling_rate=20e6)
    rx = SimulatedDevice(antennas=2, sampling_rate=20e6)
    sample = real.sample(tx, rx)
    m, taps = 700, 12
    csi = sample.state(m, taps)
    assert isinstance(csi, ChannelStateInformation) and csi.state_format == ChannelStateFormat.IMPULSE_RESPONSE
    params = fading.make_parameters(ch.delays, ch.power_profile, ch.rice_factors, ch.num_sinusoids,
                                    ch.doppler_frequency, ch.los_doppler_frequency, ch.gain,
                                    StandardAntennaCorrelation(CorrelationType.MEDIUM_A).covariance_for(2, True),
                                    StandardAntennaCorrelation(CorrelationType.MEDIUM_A).covariance_for(2, False))
    osample = fading.sample_from_normals(params, real.scalars, 2, 2, 20e6)
    ref = fading.state(osample, m, taps)
    got = csi.dense_state().cpu().numpy()
    assert got.shape == ref.shape == (2, 2, m, min(taps, 1 + 8))
    assert rel_err(got, ref) < 1e-5
    rng = np.random.default_rng(3)
    x = (rng.standard_normal((2, m)) + 1j * rng.standard_normal((2, m))) / np.sqrt(2)
    sig = Signal.Create(torch.as_tensor(x.astype(np.complex64), device="cuda"), 20e6)
    y_state = csi.propagate(sig).samples.cpu().numpy()
    y_line = sample.propagate(sig).samples.cpu().numpy()
    y_ref = fading.propagate(osample, x)
    assert rel_err(y_state, y_ref) < 1e-5
    assert rel_err(y_line, y_ref) < 1e-5
    # the reciprocal sample (fading.py:517-538) is the transposed link
    back = real.reciprocal_sample(sample, rx, tx)
    assert np.allclose(back.spatial_response, sample.spatial_response.T)
    assert rel_err(back.state(m, taps).dense_state().cpu().numpy(), np.transpose(ref, (1, 0, 2, 3))) < 1e-5
