"""GPU parity of fused drops over frames with pilots, per-drop Doppler and exports.

The comb-pilot 2x2 link (reference elements, orthogonal per-stream pilots,
Doppler drawn per drop) runs through hpx_run_batch_frames; each drop's
exported received grid, ideal-CSI factor c and spatial response S, and its
equalised data symbols, must match the float64 oracle run at that drop's
Doppler, and its error count must be exactly what the oracle's demapper
gives on the kernel's own symbols.
"""

import numpy as np
import pytest
import torch

from conftest import oracle_link, rel_err
from oracle import drop as odrop
from oracle import mapping

pytestmark = pytest.mark.gpu

SEED = 20261202


def run_frames(cfg, first, snrs):
    from hermespy_amd.simulation import BatchedLink

    bl = BatchedLink(cfg)
    n = len(snrs)
    dev = "cuda"
    npw = torch.as_tensor(cfg.noise_power(np.asarray(snrs)), dtype=torch.float64, device=dev)
    err = torch.empty(n, dtype=torch.int32, device=dev)
    eq = torch.empty((n, cfg.num_tx, cfg.num_data), dtype=torch.complex64, device=dev)
    grid = torch.empty((n, cfg.num_rx, cfg.words, cfg.num_subcarriers), dtype=torch.complex64, device=dev)
    csi = torch.empty((n, cfg.words, cfg.num_subcarriers), dtype=torch.complex64, device=dev)
    spatial = torch.empty((n, cfg.num_rx, cfg.num_tx), dtype=torch.complex64, device=dev)
    dop = cfg.draw_doppler(first, n)
    bl.run(SEED, first, npw, err, eq, doppler=dop, grid=grid, csi=csi, spatial=spatial)
    torch.cuda.synchronize()
    return dop, err.cpu().numpy(), eq.cpu().numpy(), grid.cpu().numpy(), csi.cpu().numpy(), spatial.cpu().numpy()


def test_pilot_frames_match_oracle(cuda):
    from hermespy_amd.configs import mimo2x2_512_pilots

    cfg = mimo2x2_512_pilots()
    link = oracle_link(cfg)
    first, snrs = 5, [0.0, 12.0, 30.0]
    dop, err, eq, grid, csi, spatial = run_frames(cfg, first, snrs)
    assert np.all(dop[:, 0] >= 2 * np.pi * 10) and np.all(dop[:, 0] <= 2 * np.pi * 500)
    for i, snr in enumerate(snrs):
        ref = odrop.run_drop(link.with_doppler(float(dop[i, 0])), SEED, first + i, snr, faithful=False)
        assert rel_err(grid[i], ref.grid) < 1e-4, ("grid", i, rel_err(grid[i], ref.grid))
        h = spatial[i][:, :, None, None] * csi[i][None, None]
        assert rel_err(h, ref.csi) < 1e-4, ("csi", i, rel_err(h, ref.csi))
        assert rel_err(eq[i], ref.equalized) < 1e-4, ("equalized", i, rel_err(eq[i], ref.equalized))
        rx = np.stack([mapping.detect_bits(eq[i, t].astype(np.complex128), cfg.modulation_order).astype(np.uint8)
                       for t in range(cfg.num_tx)])
        assert int(err[i]) == int(np.sum(rx != ref.tx_bits))


def test_all_data_frame_equals_plain_batch(cuda):
    """An explicit all-data frame reproduces the frameless launch exactly."""

    from hermespy_amd.configs import siso_64
    from hermespy_amd.simulation import BatchedLink

    plain = siso_64()
    framed = siso_64()
    framed.frame_mask = np.ones((plain.words, plain.num_subcarriers), dtype=np.int8)
    n = 32
    npw = torch.as_tensor(plain.noise_power(np.full(n, 5.0)), dtype=torch.float64, device="cuda")
    outs = []
    for cfg in (plain, framed):
        err = torch.empty(n, dtype=torch.int32, device="cuda")
        eq = torch.empty((n, 1, cfg.num_data), dtype=torch.complex64, device="cuda")
        BatchedLink(cfg).run(SEED, 100, npw, err, eq)
        torch.cuda.synchronize()
        outs.append((err.cpu().numpy(), eq.cpu().numpy()))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])
