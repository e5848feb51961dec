"""GPU parity of the massive-MIMO drops (hpx_massive_run) against the float64 oracle: spatial
response, received grid and ideal-CSI factor within float32 tolerance given the same Philox
streams; equalised symbols within the bfloat16 hi/lo MFMA matched filter's tolerance (stated
in the assertion); error counts exactly what the oracle's demapper gives on the kernel's symbols."""

import numpy as np
import pytest
import torch

from conftest import rel_err

pytestmark = pytest.mark.gpu

SEED = 77


def oracle_link(cfg):
    from oracle import massive

    return massive.MassiveLink(cfg.num_rx, cfg.num_users, cfg.num_subcarriers, cfg.active, cfg.words, cfg.prefix,
                               cfg.modulation_order, cfg.bandwidth, cfg.delays, cfg.power_profile, cfg.rice_factors,
                               cfg.doppler, cfg.num_sinusoids, cfg.corr_rx, cfg.gain, cfg.device_power,
                               cfg.dc_suppression)


def run(cfg, first, snrs):
    from hermespy_amd.massive import MassiveBatch

    mb = MassiveBatch(cfg, chunk=len(snrs))
    npw = torch.as_tensor(cfg.noise_power(np.asarray(snrs)), dtype=torch.float64, device="cuda")
    err = torch.zeros(len(snrs), dtype=torch.int32, device="cuda")
    eq = torch.empty((len(snrs), cfg.num_users, cfg.words, cfg.active), dtype=torch.complex64, device="cuda")
    mb.run(SEED, first, npw, err, eq)
    torch.cuda.synchronize()
    return mb, err.cpu().numpy(), eq.cpu().numpy()


@pytest.mark.parametrize("shape", ["small", "full"])
def test_massive_matches_oracle(cuda, shape):
    from hermespy_amd.massive import MassiveConfig
    from oracle import mapping
    from oracle import massive as om
    from oracle import ofdm

    if shape == "small":
        cfg = MassiveConfig(num_rx=16, num_users=4, num_subcarriers=256, active=200, prefix=36,
                            bandwidth=256 * 30e3, delays=np.arange(6) / (256 * 30e3), snr_db=(30.0,))
        snrs = [20.0, 35.0]
    else:
        cfg = MassiveConfig()
        snrs = [40.0]
    link = oracle_link(cfg)
    mb, err, eq = run(cfg, 11, snrs)
    sp = mb.buffers["spatial"].cpu().numpy()
    grid = mb.buffers["grid"].cpu().numpy()
    csi = mb.buffers["csi"].cpu().numpy()
    for i, snr in enumerate(snrs):
        ref = om.run_drop(link, SEED, 11 + i, snr)
        assert rel_err(sp[i].reshape(cfg.num_rx, cfg.num_users), ref.spatial) < 1e-6
        g_ref = ofdm.pick(link.frame, ref.grid).reshape(cfg.num_rx, cfg.words, cfg.active).transpose(1, 0, 2)
        assert rel_err(grid[i].reshape(cfg.words, cfg.num_rx, cfg.active), g_ref) < 1e-5
        c_ref = ofdm.pick(link.frame, ref.csi).reshape(cfg.words, cfg.active)
        assert rel_err(csi[i].reshape(cfg.words, cfg.active), c_ref) < 1e-5
        e = rel_err(eq[i].reshape(cfg.num_users, -1), ref.equalized)
        assert e < 1e-4, ("equalized", shape, i, e)
        rx = np.stack([mapping.detect_bits(eq[i, t].reshape(-1).astype(np.complex128), cfg.modulation_order)
                       .astype(np.uint8) for t in range(cfg.num_users)])
        assert int(err[i]) == int(np.sum(rx != ref.tx_bits)), (shape, i)
        assert abs(int(err[i]) - ref.errors) <= max(4, int(1e-4 * ref.tx_bits.size))


def test_learned_fused_estimate_matches_layer_path(cuda):
    """The one-launch learned estimator equals pilot LS + the layer-by-layer network."""

    from hermespy_amd.configs import mimo2x2_512_pilots
    from hermespy_amd.learned import ConvNetEstimator, init_params
    from hermespy_amd.lmmse import LMMSEChannelEstimation
    from hermespy_amd.simulation import BatchedLink

    cfg = mimo2x2_512_pilots()
    params = init_params(cfg, seed=4)
    rng = np.random.default_rng(2)
    w, b = params[-1]
    params[-1] = (rng.uniform(-0.2, 0.2, w.shape).astype(np.float32), rng.uniform(-0.1, 0.1, b.shape).astype(np.float32))
    net = ConvNetEstimator(cfg, params)
    n = 9
    npw = torch.as_tensor(cfg.noise_power(np.full(n, 12.0)), dtype=torch.float64, device=cuda)
    bl = BatchedLink(cfg)
    err = torch.empty(n, dtype=torch.int32, device=cuda)
    grid = torch.empty((n, 2, 14, 512), dtype=torch.complex64, device=cuda)
    bl.run(SEED, 0, npw, err, doppler=cfg.draw_doppler(0, n), grid=grid)
    lm = LMMSEChannelEstimation(cfg, snr_db=[12.0])
    a = net.estimate(lm.pilot_ls(grid, n), npw).cpu().numpy()
    f = net.estimate_grid(grid, lm.pilot_index, lm.pilot_values, npw).cpu().numpy()
    assert rel_err(f, a) < 1e-5
