"""Recording drops of the fused path and replaying them from the file alone: identical
per-drop errors, with and without recorded realisation normals, for plain runs and for
campaigns with early stopping; the reference's mode errors."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

SEED = 20261202


@pytest.fixture(scope="module", autouse=True)
def h5lib():
    from hermespy_amd import record

    if not record.H5_LIB_PATH.exists():
        pytest.skip("drop store not built (no libhdf5)")
    return record.load_h5()


@pytest.mark.parametrize("normals", [False, True])
def test_run_record_and_replay(cuda, tmp_path, normals):
    from hermespy_amd.configs import mimo4x4_1024
    from hermespy_amd.simulation import BatchedSimulation, ScenarioMode

    print(f"record + replay mimo4x4-1024 normals={normals}", flush=True)
    sim = BatchedSimulation(mimo4x4_1024(), seed=31)
    path = tmp_path / "drops.h5"
    sim.record(path, campaign="c1", normals=normals)
    assert sim.mode == ScenarioMode.RECORD
    with pytest.raises(RuntimeError, match="default mode"):
        sim.record(path)
    first = sim.run(150, drops_per_launch=64)
    second = sim.run(40, drops_per_launch=64, first_drop=150)
    sim.stop()
    replayed, n = BatchedSimulation.Replay(path, "c1")
    assert n == 190 and replayed.mode == ScenarioMode.REPLAY and replayed.seed == 31
    r1 = replayed.run(150, drops_per_launch=50)
    r2 = replayed.run()
    assert replayed.replay_mismatches == 0
    assert np.array_equal(r1.errors, first.errors) and np.array_equal(r1.drops, first.drops)
    assert np.array_equal(r2.errors, second.errors)
    with pytest.raises(ValueError, match="left"):
        replayed.run(1)
    # an offset replays the tail
    assert replayed.replay(campaign="c1", drop_offset=150) == 40
    assert np.array_equal(replayed.run().errors, second.errors)
    replayed.stop()
    with pytest.raises(ValueError, match="Drop offset"):
        BatchedSimulation.Replay(path, "c1", drop_offset=190)


def test_campaign_record_and_replay(cuda, tmp_path):
    from hermespy_amd.configs import mimo2x2_512_pilots
    from hermespy_amd.modem import BitErrorEvaluator
    from hermespy_amd.simulation import BatchedSimulation

    print("record + replay an early-stopping campaign (mimo2x2-512-pilots, per-drop Doppler off)", flush=True)
    cfg = mimo2x2_512_pilots()
    sim = BatchedSimulation(cfg, seed=4)
    path = tmp_path / "campaign.h5"
    sim.record(path, normals=True)
    res = sim.simulate(48, BitErrorEvaluator(confidence=0.1, tolerance=5e-3, min_num_samples=16),
                       max_drops_per_launch=200)
    sim.stop()
    r = res.evaluation_results[0]
    replayed, n = BatchedSimulation.Replay(path)
    assert n == int(r.num_drops().sum())
    out = replayed.run()
    assert replayed.replay_mismatches == 0
    _, _, count, errors, _ = r.host_counters()
    assert np.array_equal(out.drops, count) and np.array_equal(out.errors, errors)


def test_replay_requires_a_file(cuda):
    from hermespy_amd.configs import siso_64
    from hermespy_amd.simulation import BatchedSimulation

    sim = BatchedSimulation(siso_64())
    with pytest.raises(ValueError, match="valid file path"):
        sim.replay()


def test_reference_realizations_replay_against_oracle(cuda, tmp_path):
    """Channel realisations recorded in the reference's layout (normals drawn by NumPy, not by the
    library's Philox stream) replay through ``hpx_run_batch_drops`` and match the float64 oracle
    sampling the same normals (``fading.sample_from_normals``) drop by drop."""

    from conftest import oracle_link
    from oracle_pool import oracle_drops
    from parity import check_drop

    from hermespy_amd import record
    from hermespy_amd.configs import mimo4x4_1024
    from hermespy_amd.simulation import BatchedLink, BatchedSimulation

    cfg = mimo4x4_1024()
    snrs = np.array([0.0, 12.0, 24.0, 30.0])
    z = np.random.default_rng(11).standard_normal((snrs.size, cfg.num_normals))
    path = tmp_path / "reference_layout.h5"
    record.write_reference_drops(path, cfg, z, campaign="replay")
    back = record.read_reference_drops(path, cfg, campaign="replay")
    assert np.array_equal(back, z)
    print("reference-layout realisations through hpx_run_batch_drops", flush=True)
    n = snrs.size
    npw = torch.as_tensor(cfg.noise_power(snrs), dtype=torch.float64, device="cuda")
    err = torch.empty(n, dtype=torch.int32, device="cuda")
    eq = torch.empty((n, cfg.num_tx, cfg.num_data), dtype=torch.complex64, device="cuda")
    BatchedLink(cfg).run(SEED, 0, npw, err, eq, normals=torch.as_tensor(back, device="cuda"))
    torch.cuda.synchronize()
    err, eq = err.cpu().numpy(), eq.cpu().numpy()
    ora = oracle_drops(oracle_link(cfg), SEED, range(n), snrs, normals=back)
    for i in range(n):
        check_drop(eq[i], err[i], ora[i], cfg.modulation_order, 1e-4, ("recorded realisation", i))
    # the simulation's replay of the file runs the same drops
    sim = BatchedSimulation(cfg, SEED)
    res = sim.replay_realizations(path, "replay", snr_db=snrs)
    assert res.errors.tolist() == err.astype(np.int64).tolist()
