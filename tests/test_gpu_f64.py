"""The float64 drop path (hpx_run_batch_f64) of the small single-antenna link: equalised
symbols within 1e-11 of the float64 oracle's (reference-faithful dense impulse response,
pinv) and error counts identical, for ZF and MMSE; the float32 fused path's error against it
stays within its stated 1e-5."""

import dataclasses

import numpy as np
import pytest
import torch

from conftest import oracle_link, rel_err
from oracle import drop as odrop

pytestmark = pytest.mark.gpu

SEED = 20261202


@pytest.mark.parametrize("detector", ["zf", "mmse"])
def test_f64_matches_oracle(cuda, detector):
    from hermespy_amd.configs import siso_64
    from hermespy_amd.simulation import BatchedLink

    cfg = dataclasses.replace(siso_64(), detector=detector)
    print(f"float64 siso-64 {detector} vs oracle", flush=True)
    drops = list(range(40, 50))
    snrs = np.array([0.0, 5.0, 10.0, 15.0, 20.0] * 2)
    bl = BatchedLink(cfg)
    npw = torch.as_tensor(cfg.noise_power(snrs), dtype=torch.float64, device=cuda)
    err = torch.empty(len(drops), dtype=torch.int32, device=cuda)
    eq = torch.empty((len(drops), cfg.words * cfg.num_subcarriers), dtype=torch.complex128, device=cuda)
    bl.run_f64(SEED, drops[0], npw, err, eq)
    err32 = torch.empty(len(drops), dtype=torch.int32, device=cuda)
    eq32 = torch.empty((len(drops), 1, cfg.words * cfg.num_subcarriers), dtype=torch.complex64, device=cuda)
    bl.run(SEED, drops[0], npw, err32, eq32)
    torch.cuda.synchronize()
    err, eq, eq32 = err.cpu().numpy(), eq.cpu().numpy(), eq32.cpu().numpy()
    link = oracle_link(cfg)
    for i, (d, snr) in enumerate(zip(drops, snrs)):
        ref = odrop.run_drop(link, SEED, d, float(snr), faithful=True)
        e = rel_err(eq[i], ref.equalized[0])
        assert e < 1e-11, (detector, d, snr, e)
        assert int(err[i]) == ref.errors, (detector, d, int(err[i]), ref.errors)
        assert rel_err(eq32[i, 0], eq[i]) < 1e-5
    assert err.sum() > 0


def test_f64_rician_wide_link_matches_oracle(cuda):
    """256 subcarriers without DC suppression, 9 paths with a line-of-sight first tap (K = 0 dB),
    16-QAM, MMSE: the float64 path against the float64 oracle."""

    from hermespy_amd.simulation import BatchedLink
    from hermespy_amd.uneven import DropShape

    base = DropShape(1, 256, 9, 2.0, 300.0, 16, 15.0).link()
    cfg = dataclasses.replace(base, rice_factors=np.array([1.0] + [0.0] * 8), dc_suppression=False)
    print("float64 1x1-256 Rician vs oracle", flush=True)
    snrs = np.array([6.0, 12.0, 18.0, 24.0])
    bl = BatchedLink(cfg)
    npw = torch.as_tensor(cfg.noise_power(snrs), dtype=torch.float64, device=cuda)
    err = torch.empty(snrs.size, dtype=torch.int32, device=cuda)
    eq = torch.empty((snrs.size, cfg.words * cfg.num_subcarriers), dtype=torch.complex128, device=cuda)
    bl.run_f64(SEED, 7, npw, err, eq)
    torch.cuda.synchronize()
    err, eq = err.cpu().numpy(), eq.cpu().numpy()
    link = oracle_link(cfg)
    for i, snr in enumerate(snrs):
        ref = odrop.run_drop(link, SEED, 7 + i, float(snr), faithful=True)
        e = rel_err(eq[i], ref.equalized[0])
        assert e < 1e-11, (i, snr, e)
        assert int(err[i]) == ref.errors, (i, int(err[i]), ref.errors)


def test_f64_beyond_the_float32_doppler_limit(cuda):
    """2 kHz Doppler at 1 MHz sampling: beyond the float32 path's Taylor taps (which refuse the
    link), the float64 path evaluates every tap exactly and matches the float64 oracle."""

    from hermespy_amd.configs import siso_64
    from hermespy_amd.simulation import BatchedLink

    cfg = dataclasses.replace(siso_64(), doppler=2 * np.pi * 2000.0)
    print("float64 siso-64 at 2 kHz Doppler vs oracle", flush=True)
    bl = BatchedLink(cfg)
    snrs = np.array([5.0, 15.0, 25.0])
    npw = torch.as_tensor(cfg.noise_power(snrs), dtype=torch.float64, device=cuda)
    err = torch.empty(snrs.size, dtype=torch.int32, device=cuda)
    eq = torch.empty((snrs.size, cfg.words * cfg.num_subcarriers), dtype=torch.complex128, device=cuda)
    with pytest.raises(ValueError, match="Doppler rate"):
        bl.run(SEED, 3, npw, err)
    bl.run_f64(SEED, 3, npw, err, eq)
    torch.cuda.synchronize()
    err, eq = err.cpu().numpy(), eq.cpu().numpy()
    link = oracle_link(cfg)
    for i, snr in enumerate(snrs):
        ref = odrop.run_drop(link, SEED, 3 + i, float(snr), faithful=True)
        assert rel_err(eq[i], ref.equalized[0]) < 1e-11, (i, snr)
        assert int(err[i]) == ref.errors, (i, int(err[i]), ref.errors)


def test_f64_rejects_other_links(cuda):
    from hermespy_amd.configs import mimo2x2_512_pilots
    from hermespy_amd.simulation import BatchedLink

    bl = BatchedLink(mimo2x2_512_pilots())
    npw = torch.ones(2, dtype=torch.float64, device=cuda)
    err = torch.empty(2, dtype=torch.int32, device=cuda)
    with pytest.raises(ValueError):
        bl.run_f64(1, 0, npw, err)
