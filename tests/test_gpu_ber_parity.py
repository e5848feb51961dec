"""BER parity of the fused kernel on the 4x4 1024-subcarrier MMSE link (the throughput
benchmark's link) at every SNR point of its campaign.

176 drops, 16 at each of the 11 SNR points 0..30 dB (drop i at point i % 11, as in the
benchmark), run through ``hpx_run_batch`` and compared with the float64 oracle on the same
counter-based random streams: per drop the equalised symbols (1e-4), the error count of the
kernel's own decisions (exact) and the error count against the oracle's (within the
decision-boundary slack); per SNR point the BER inside the oracle's 99 % binomial interval
(``BitErrorEvaluator``, ``modem/evaluators.py:236-258``, aggregated as
``ScalarEvaluationResult``, ``core/pymonte/scalar.py:100-125``).  One drop per SNR point is
also compared with the reference-faithful oracle (dense impulse response, pinv).
"""

import numpy as np
import pytest
import torch

from conftest import oracle_link
from oracle_pool import oracle_drops
from parity import check_drop

pytestmark = pytest.mark.gpu

SEED = 20261202
POINTS = 11
PER_POINT = 16


def _gpu_drops(cfg, first, n):
    from hermespy_amd.simulation import BatchedLink

    snr = np.asarray(cfg.snr_db, dtype=np.float64)
    idx = (first + np.arange(n)) % snr.size
    npw = torch.as_tensor(cfg.noise_power(snr[idx]), dtype=torch.float64, device="cuda")
    err = torch.empty(n, dtype=torch.int32, device="cuda")
    eq = torch.empty((n, cfg.num_tx, cfg.num_data), dtype=torch.complex64, device="cuda")
    BatchedLink(cfg).run(SEED, first, npw, err, eq)
    torch.cuda.synchronize()
    return snr[idx], err.cpu().numpy(), eq.cpu().numpy()


def test_campaign_ber_parity_4x4_1024(cuda):
    from hermespy_amd.configs import mimo4x4_1024

    cfg = mimo4x4_1024()
    assert len(cfg.snr_db) == POINTS
    n = POINTS * PER_POINT
    print(f"4x4 1024-subcarrier link: {n} drops through hpx_run_batch", flush=True)
    snrs, err, eq = _gpu_drops(cfg, 0, n)
    link = oracle_link(cfg)
    ora = oracle_drops(link, SEED, range(n), snrs, faithful=False)
    stats = [check_drop(eq[i], err[i], ora[i], cfg.modulation_order, 1e-4, (i, snrs[i])) for i in range(n)]
    assert max(s["rel_err"] for s in stats) < 1e-4
    for k in range(POINTS):
        sel = np.arange(k, n, POINTS)
        bits = sum(ora[i][1] for i in sel)
        p_ref = sum(ora[i][0] for i in sel) / bits
        p_gpu = float(err[sel].sum()) / bits
        half = 2.576 * np.sqrt(max(p_ref * (1 - p_ref), 1.0 / bits) / bits)
        assert abs(p_gpu - p_ref) <= half, (cfg.snr_db[k], p_gpu, p_ref, half)
    # the reference-faithful oracle (dense impulse response, pinv) on one drop per SNR point
    faithful = oracle_drops(link, SEED, range(POINTS), snrs[:POINTS], faithful=True)
    for i in range(POINTS):
        check_drop(eq[i], err[i], faithful[i], cfg.modulation_order, 1e-4, ("faithful", i))
