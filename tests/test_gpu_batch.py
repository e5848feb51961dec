"""GPU parity of the fused batched drops (hpx_run_batch) against whole oracle drops.

Given the same counter-based random streams, each drop's equalised symbols
must match the float64 oracle (reference-faithful dense-CIR path for the small
link, closed-form CSI for the 4x4 link), and the drop's error count must be
exactly the count the oracle's demapper gives on the kernel's own equalised
symbols (bit-exact demapping and XOR-popcount).
"""

import numpy as np
import pytest
import torch

from conftest import oracle_link, rel_err
from oracle import drop as odrop
from oracle import mapping

pytestmark = pytest.mark.gpu

SEED = 20261202


def run_gpu(cfg, drops, snr_db):
    from hermespy_amd.simulation import BatchedLink

    bl = BatchedLink(cfg)
    n = len(drops)
    npw = torch.as_tensor(cfg.noise_power(np.asarray(snr_db)), dtype=torch.float64, device="cuda")
    err = torch.empty(n, dtype=torch.int32, device="cuda")
    eq = torch.empty((n, cfg.num_tx, cfg.words * cfg.num_subcarriers), dtype=torch.complex64, device="cuda")
    assert list(drops) == list(range(drops[0], drops[0] + n))
    bl.run(SEED, drops[0], npw, err, eq)
    torch.cuda.synchronize()
    return err.cpu().numpy(), eq.cpu().numpy()


@pytest.mark.parametrize("name,drops,snrs,faithful,tol", [
    ("siso-64", [0, 1, 2, 3, 4], [0.0, 5.0, 10.0, 15.0, 20.0], True, 1e-5),
    ("mimo4x4-1024", [7, 8, 9], [0.0, 15.0, 30.0], False, 1e-4),
])
def test_batch_matches_oracle(cuda, name, drops, snrs, faithful, tol):
    from hermespy_amd.configs import CONFIGS

    cfg = CONFIGS[name]()
    link = oracle_link(cfg)
    err, eq = run_gpu(cfg, drops, snrs)
    for i, (d, snr) in enumerate(zip(drops, snrs)):
        ref = odrop.run_drop(link, SEED, d, snr, faithful=faithful)
        e = rel_err(eq[i], ref.equalized)
        assert e < tol, (name, d, snr, e)
        # bit-exact demap + count on the kernel's own symbols
        rx = np.stack([mapping.detect_bits(eq[i, t].astype(np.complex128), cfg.modulation_order).astype(np.uint8)
                       for t in range(cfg.num_tx)])
        assert int(err[i]) == int(np.sum(rx != ref.tx_bits)), (name, d)
        # and the decisions agree with the float64 oracle except for symbols within rounding of a boundary
        assert abs(int(err[i]) - ref.errors) <= max(2, int(2e-4 * ref.bits)), (int(err[i]), ref.errors)


def test_batch_independent_of_launch_split(cuda):
    """Drop i's result depends only on (seed, i): one launch == two launches."""

    from hermespy_amd.configs import CONFIGS

    cfg = CONFIGS["siso-64"]()
    e1, _ = run_gpu(cfg, list(range(0, 64)), [10.0] * 64)
    a, _ = run_gpu(cfg, list(range(0, 20)), [10.0] * 20)
    b, _ = run_gpu(cfg, list(range(20, 64)), [10.0] * 44)
    assert np.array_equal(e1, np.concatenate([a, b]))


def test_batch_ber_within_oracle_ci(cuda):
    """Seeded campaign BER of the fused kernel lies in the 99% binomial interval of the oracle's."""

    from hermespy_amd.configs import CONFIGS
    from hermespy_amd.simulation import BatchedSimulation

    cfg = CONFIGS["siso-64"]()
    n = 300
    sim = BatchedSimulation(cfg, seed=SEED)
    res = sim.run(n, snr_db=[5.0], drops_per_launch=128)
    link = oracle_link(cfg)
    oerr = sum(odrop.run_drop(link, SEED, d, 5.0, faithful=False).errors for d in range(n))
    nbits = n * cfg.bits_per_drop
    p = oerr / nbits
    half = 2.576 * np.sqrt(p * (1 - p) / nbits)
    assert abs(res.ber[0] - p) <= half + 1e-12, (res.ber[0], p, half)


def test_batch_rejects_unsupported(cuda):
    from hermespy_amd.configs import CONFIGS
    from hermespy_amd.simulation import BatchedLink

    cfg = CONFIGS["siso-64"]()
    cfg.num_rx = 2  # 2x1 is not instantiated
    with pytest.raises(ValueError, match="not instantiated"):
        BatchedLink(cfg)
