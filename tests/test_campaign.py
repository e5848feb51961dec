"""Monte Carlo campaign control on the host: the stopping rule of the oracle pinned by its
mathematics, the section queue, the campaign loop against the sequential oracle, and the
result outputs (.mat, table, plot).  No GPU: device accumulation is replaced by a NumPy sum
in the loop test only."""

import io

import numpy as np
import pytest
import torch
from scipy.stats import norm

from oracle import montecarlo


def gaussian_sections(mu, sigma, seed):
    rng = np.random.default_rng(seed)
    cache = {}

    def scalar_of(s, k):
        if s not in cache:
            cache[s] = mu + sigma * rng.standard_normal(200_000)
        return cache[s][k]

    return scalar_of


def test_oracle_stopping_count_follows_the_clt_bound():
    """Gaussian artifacts stop at the first multiple of min_num_samples past
    n* = (z_{1 - delta / 2} sigma / TOL)^2 (the count where the CLT half-width reaches TOL);
    a dropped factor 2, sqrt or a wrong Phi moves the stop by several multiples."""

    delta, sigma, m = 0.05, 1.0, 256
    tol = norm.ppf(1 - delta / 2) * sigma / np.sqrt(1957.0)  # n* = 1957: between 7 m = 1792 and 8 m = 2048
    counts, means, _, _ = montecarlo.run_sections(gaussian_sections(3.0, sigma, 1), 16, 100_000, tol, delta, m)
    assert np.all(counts == 8 * m), counts
    assert np.all(np.abs(means - 3.0) < 5 * sigma / np.sqrt(counts))


def test_oracle_stopping_rule_coverage():
    """P(|mean - E X| > TOL) at the stop stays near delta (Bayer et al.'s guarantee)."""

    delta, tol, m = 0.2, 0.05, 64
    counts, means, _, _ = montecarlo.run_sections(gaussian_sections(0.5, 1.0, 2), 400, 100_000, tol, delta, m)
    miss = np.mean(np.abs(means - 0.5) > tol)
    # n* = (z_0.9 / 0.05)^2 = 657: stops near the multiples 640 .. 768
    assert counts.min() >= 512 and counts.max() <= 832 and miss < 0.3, (counts.min(), counts.max(), miss)


def test_oracle_special_cases():
    # constant artifacts: std == 0 never stops; default accuracy 0 and confidence 1 never stop
    c, mu, _, _ = montecarlo.run_sections(lambda s, k: 0.25, 2, 3000, 0.1, 0.5, 100)
    assert np.all(c == 3000) and np.all(mu == 0.25)
    c, _, _, _ = montecarlo.run_sections(gaussian_sections(0.0, 1.0, 3), 2, 5000, 0.0, 1.0, 100)
    assert np.all(c == 5000)
    # running sums: mean = sum / count, squared sum of the artifacts
    c, mu, s, q = montecarlo.run_sections(lambda s, k: float(k), 1, 10, 0.0, 1.0, 2)
    assert c[0] == 10 and s[0] == 45.0 and q[0] == 285.0 and mu[0] == 4.5


@pytest.mark.parametrize("args,match", [((0.1, 1.5, 10), "between zero and one"), ((0.1, -0.1, 10), "between zero"),
                                        ((-1.0, 0.5, 10), "non-negative"), ((0.1, 0.5, 1), "at least two")])
def test_argument_checks(args, match):
    from hermespy_amd.campaign import ScalarEvaluationResult, snr_dimension

    with pytest.raises(ValueError, match=match):
        montecarlo.check_arguments(*args)
    with pytest.raises(ValueError):
        ScalarEvaluationResult([snr_dimension([0.0])], None, *args, device="cpu")


def test_evaluator_properties_validate():
    from hermespy_amd.modem import BitErrorEvaluator

    ev = BitErrorEvaluator(confidence=0.1, tolerance=1e-3, min_num_samples=64)
    assert (ev.confidence, ev.tolerance, ev.min_num_samples, ev.abbreviation) == (0.1, 1e-3, 64, "BER")
    for name, bad in (("confidence", 1.2), ("tolerance", -1.0), ("min_num_samples", 1)):
        with pytest.raises(ValueError):
            setattr(ev, name, bad)


def test_result_stopping_test_matches_oracle():
    from hermespy_amd.campaign import ScalarEvaluationResult, snr_dimension

    rng = np.random.default_rng(4)
    res = ScalarEvaluationResult([snr_dimension(np.arange(5.0))], None, 0.02, 0.2, 32, device="cpu")
    xs = [rng.uniform(0, 0.3 + 0.1 * s, size=64 * (s + 1)) for s in range(5)]
    decisions = []
    for s, x in enumerate(xs):
        sec = montecarlo.Section(0.02, 0.2, 32)
        for v in x:
            got, ref = res.add_artifact((s,), v), sec.add_artifact(v)
            assert got == ref
            decisions.append(got)
    assert any(decisions) and not all(decisions)
    assert np.allclose(res.to_array(), [x.mean() for x in xs], rtol=1e-13, atol=0)


def test_section_queue_blocks():
    from hermespy_amd.campaign import SectionQueue

    q = SectionQueue(4, 2500, 512, 1000)
    issued = {s: [] for s in range(4)}
    launches = 0
    while q.num_active:
        sec, k = q.next_launch()
        launches += 1
        for s in np.unique(sec):
            ks = k[sec == s]
            # consecutive samples that never run past the next multiple of min_num_samples
            assert np.array_equal(ks, np.arange(ks[0], ks[0] + ks.size))
            assert ks[0] // 512 == ks[-1] // 512
            issued[s].extend(ks.tolist())
        if launches == 3:
            q.disable_section(2)
    for s in range(4):
        assert issued[s] == list(range(len(issued[s])))
    assert len(issued[2]) < 2500 and all(len(issued[s]) == 2500 for s in (0, 1, 3))
    assert q.query_progress()[0] == 1.0


def test_section_queue_more_sections_than_drops_per_launch():
    """With more active sections than drops per launch, every launch fits max_drops and the
    sections are served round robin, each in consecutive samples."""

    from hermespy_amd.campaign import SectionQueue

    q = SectionQueue(7, 5, 2, 3)
    seen = {s: [] for s in range(7)}
    while q.num_active:
        sec, k = q.next_launch()
        assert 1 <= sec.size <= 3
        for s, kk in zip(sec, k):
            seen[int(s)].append(int(kk))
    assert all(seen[s] == list(range(5)) for s in range(7))


class _FakeLink:
    """Per-drop errors from a fixed function of the Philox drop index (no GPU)."""

    def __init__(self, bits, device="cpu"):
        from hermespy_amd.configs import mimo4x4_1024

        self.link = mimo4x4_1024()
        self.bits = bits
        self.device = torch.device(device)

    @staticmethod
    def errors_of(drop, section, bits):
        rng = np.random.default_rng(drop)
        return int(rng.binomial(bits, 0.02 + 0.03 * section))

    def run(self, seed, first, npw, errors, drop_index=None):
        ids = drop_index.numpy().view(np.uint32)
        for i, d in enumerate(ids):
            errors[i] = self.errors_of(int(d), int(d) % 11, self.bits)


def test_campaign_loop_matches_sequential_oracle(monkeypatch):
    from hermespy_amd import campaign
    from hermespy_amd.modem import BitErrorEvaluator

    bits = 229376
    fake = _FakeLink(bits)

    def accumulate(self, errors, bits=None, section=None, **kw):
        e = errors.numpy().astype(np.int64)
        s = section.numpy()
        for i in range(e.size):
            ber = e[i] / bits
            self.sum[s[i]] += ber
            self.sum_sq[s[i]] += ber * ber
            self.count[s[i]] += 1
            self.errors[s[i]] += e[i]
            self.bits[s[i]] += bits

    monkeypatch.setattr(campaign.ScalarEvaluationResult, "accumulate", accumulate)
    ev = BitErrorEvaluator(confidence=0.05, tolerance=4e-4, min_num_samples=16)
    snr = np.arange(0.0, 31.0, 3.0)
    res = campaign.run_campaign(fake, 1, snr, 400, ev, max_drops_per_launch=100)
    r = res.evaluation_results[0]
    counts, means, _, _ = montecarlo.run_sections(
        lambda s, k: fake.errors_of(k * 11 + s, s, bits) / bits, 11, 400, 4e-4, 0.05, 16)
    assert np.array_equal(r.num_drops(), counts)
    assert np.allclose(r.to_array(), means, rtol=1e-12, atol=0)
    assert counts.min() < 400 and counts.max() > counts.min()  # some sections stop early, at different counts


def test_result_outputs(tmp_path):
    from scipy.io import loadmat
    from rich.console import Console

    from hermespy_amd.campaign import MonteCarloResult, ScalarEvaluationResult, snr_dimension
    from hermespy_amd.modem import BitErrorEvaluator

    ev = BitErrorEvaluator()
    grid = [snr_dimension([0.0, 10.0, 20.0])]
    r = ScalarEvaluationResult(grid, ev, device="cpu")
    for s, v in enumerate([0.1, 0.01, 0.001]):
        r.add_artifact((s,), v, False)
    mc = MonteCarloResult(grid, [ev], [r], 1.25)
    f = tmp_path / "campaign.mat"
    mc.save_to_matlab(str(f))
    m = loadmat(str(f))
    assert np.allclose(m["result_0"].ravel(), [0.1, 0.01, 0.001])
    assert np.allclose(m["SNR"].ravel(), [0.0, 10.0, 20.0]) and float(m["performance_time"].ravel()[0]) == 1.25
    assert str(m["evaluators"][0]).strip() == "BER"
    con = Console(file=io.StringIO(), width=80)
    mc.print(con)
    text = con.file.getvalue()
    assert "10dB" in text and "-20 dB" in text and "-30 dB" in text  # log-scale BER printed in dB
    figs = mc.plot()
    assert len(figs) == 1
