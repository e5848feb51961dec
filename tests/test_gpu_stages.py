"""GPU parity of every stage entry point of libhermespy_amd against the oracle.

Tolerances: float32 signal paths within 1e-5 relative (norm-wise) of the
float64 oracle given identical realisations and noise; mapping, demapping and
error counting bit-exact.
"""

import numpy as np
import pytest
import torch

from conftest import rel_err
from oracle import csi as ocsi
from oracle import equalize as oeq
from oracle import fading, mapping, ofdm, philox

pytestmark = pytest.mark.gpu

SEED = 20261202


def mirror(s: fading.FadingSample):
    from hermespy_amd.channel import MultipathFadingSample

    p = s.params
    return MultipathFadingSample(p.power_profile, p.delays, s.los_angles, s.nlos_angles, s.los_phases, s.nlos_phases,
                                 p.los_gains(), p.nlos_gains(), p.los_doppler, p.doppler, s.spatial_response, p.gain,
                                 s.bandwidth)


def c1_sample(seed=3, nrx=4, ntx=4):
    fs = 30.72e6
    tau = np.arange(23) * 3e-6 / 22
    pw = np.exp(-tau / 1e-6)
    k = np.zeros(23)
    k[0] = 1.0
    r = 0.5 ** np.abs(np.subtract.outer(np.arange(4), np.arange(4))).astype(complex)
    params = fading.make_parameters(tau, pw / pw.sum(), k, 20, 2 * np.pi * 389.16, corr_rx=r, corr_tx=r)
    z = np.random.default_rng(seed).standard_normal(params.num_normals)
    return fading.sample_from_normals(params, z, nrx, ntx, fs)


def c0_sample(seed=4, doppler=2 * np.pi * 50):
    fs = 1e6
    pw = 10 ** (-0.3 * np.arange(5))
    params = fading.make_parameters(np.arange(5) / fs, pw / pw.sum(), np.zeros(5), 20, doppler)
    z = np.random.default_rng(seed).standard_normal(params.num_normals)
    return fading.sample_from_normals(params, z, 1, 1, fs)


def dev(a, dtype=None):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda") if dtype is None else \
        torch.as_tensor(np.ascontiguousarray(a), device="cuda", dtype=dtype)


def test_taps_direct(cuda):
    s = c1_sample()
    m = 15344
    got = mirror(s).taps(m).cpu().numpy()
    exp = fading.taps(s, m)
    assert rel_err(got, exp) < 2e-6


@pytest.mark.parametrize("which", ["c0", "c1", "c0-fast-doppler"])
def test_tdl_propagate(cuda, which):
    s = {"c0": c0_sample, "c1": c1_sample, "c0-fast-doppler": lambda: c0_sample(doppler=2 * np.pi * 20000)}[which]()
    ntx = s.spatial_response.shape[1]
    m = 15344 if which == "c1" else 1120
    rng = np.random.default_rng(1)
    x = (rng.standard_normal((ntx, m)) + 1j * rng.standard_normal((ntx, m))) / np.sqrt(2)
    got = mirror(s).propagate(x.astype(np.complex64)).cpu().numpy()
    exp = fading.propagate(s, x.astype(np.complex64).astype(np.complex128))
    assert got.shape == exp.shape
    assert rel_err(got, exp) < 1e-5, rel_err(got, exp)


def test_awgn(cuda):
    from hermespy_amd import _lib

    nrx, n, power = 3, 5001, 0.37
    y = torch.zeros((nrx, n), dtype=torch.complex64, device="cuda")
    _lib.check(_lib.load().hpx_awgn(_lib.ptr(y), nrx, n, n, SEED, 17, power, _lib.stream_handle()))
    exp = np.stack([np.sqrt(0.5 * power) * philox.noise_normals(SEED, 17, r, n) for r in range(nrx)])
    assert rel_err(y.cpu().numpy(), exp) < 1e-5


@pytest.mark.parametrize("order,is_complex", [(2, 1), (4, 1), (16, 1), (64, 1), (256, 1), (4, 0), (8, 0), (16, 0)])
def test_map_demap_bit_exact(cuda, order, is_complex):
    from hermespy_amd import _lib

    lib = _lib.load()
    bps = int(np.log2(order))
    rng = np.random.default_rng(order)
    bits = rng.integers(0, 2, 4096 * bps).astype(np.uint8)
    sym = torch.empty(4096, dtype=torch.complex64, device="cuda")
    tbits = dev(bits)
    _lib.check(lib.hpx_map_symbols(_lib.ptr(tbits), 4096, order, is_complex, _lib.ptr(sym), _lib.stream_handle()))
    exp = mapping.get_symbols(bits, order, bool(is_complex)).astype(np.complex64)
    assert np.array_equal(sym.cpu().numpy().view(np.float32), exp.view(np.float32))
    # demap: noisy points + thresholds (float32 inputs; the oracle sees the same values in float64)
    noisy = (exp + 0.2 * (rng.standard_normal(4096) + 1j * rng.standard_normal(4096))).astype(np.complex64)
    thr = (np.arange(-16, 17) / 8.0 / np.sqrt(order)).astype(np.float32)
    extra = (thr[:, None] + 1j * thr[None, :]).reshape(-1).astype(np.complex64)
    pts = np.concatenate([noisy, extra, np.array([0, -0.0, np.float32(1e-40)], np.complex64)])
    out = torch.empty(pts.size * bps, dtype=torch.uint8, device="cuda")
    tpts = dev(pts)
    _lib.check(lib.hpx_demap_hard(_lib.ptr(tpts), pts.size, order, is_complex, _lib.ptr(out), _lib.stream_handle()))
    ref = mapping.detect_bits(pts.astype(np.complex128) if is_complex else pts.real.astype(np.float64) + 0j, order,
                              bool(is_complex)).astype(np.uint8)
    assert np.array_equal(out.cpu().numpy(), ref)


@pytest.mark.parametrize("n,cp,words,streams", [(64, 16, 14, 1), (1024, 72, 14, 4), (2048, 256, 3, 2), (4096, 288, 2, 1)])
def test_ofdm_modulate_demodulate(cuda, n, cp, words, streams):
    from hermespy_amd import _lib

    lib = _lib.load()
    frame = ofdm.simple_frame(n, words, cp / n)
    rng = np.random.default_rng(n)
    grid = ((rng.standard_normal((streams, words, n)) + 1j * rng.standard_normal((streams, words, n))) / np.sqrt(2))
    grid = grid.astype(np.complex64)
    m = frame.samples_per_frame
    out = torch.empty((streams, m), dtype=torch.complex64, device="cuda")
    tgrid = dev(grid)
    _lib.check(lib.hpx_ofdm_modulate(_lib.ptr(tgrid), streams, words, n, cp, 1, _lib.ptr(out), m,
                                     _lib.stream_handle()))
    exp = np.stack([ofdm.modulate(frame, grid[s].astype(np.complex128)) for s in range(streams)])
    assert rel_err(out.cpu().numpy(), exp) < 1e-5
    back = torch.empty((streams, words, n), dtype=torch.complex64, device="cuda")
    _lib.check(lib.hpx_ofdm_demodulate(_lib.ptr(out), streams, m, words, n, cp, 1, _lib.ptr(back),
                                       _lib.stream_handle()))
    exp_back = ofdm.demodulate(frame, exp)
    assert rel_err(back.cpu().numpy(), exp_back) < 1e-5


@pytest.mark.parametrize("which", ["c0", "c1"])
def test_ideal_csi(cuda, which):
    from hermespy_amd import _lib

    s = c0_sample() if which == "c0" else c1_sample()
    n, cp, words = (64, 16, 14) if which == "c0" else (1024, 72, 14)
    frame = ofdm.simple_frame(n, words, cp / n)
    ms = mirror(s)
    d = ms._device_tables()
    nrx, ntx = s.spatial_response.shape
    out = torch.empty((nrx, ntx, words, n), dtype=torch.complex64, device="cuda")
    num_taps = min(1 + ms.max_delay_samples, frame.samples_per_frame)
    _lib.check(_lib.load().hpx_ideal_csi(_lib.ptr(d["sos"]), ms.num_paths, ms.num_sinusoids + 1, _lib.ptr(d["delays"]),
                                         num_taps, _lib.ptr(d["spatial"]), nrx, ntx, words, n, cp, 1, _lib.ptr(out),
                                         _lib.stream_handle()))
    exp = ocsi.reference_ideal_csi(s, frame) if which == "c0" else ocsi.closed_form_ideal_csi(s, frame)
    assert rel_err(out.cpu().numpy(), exp) < 1e-5


@pytest.mark.parametrize("nrx,ntx,mode", [(1, 1, 0), (2, 2, 0), (4, 4, 0), (4, 4, 1), (8, 8, 1), (4, 2, 1), (8, 4, 0)])
def test_equalize(cuda, nrx, ntx, mode):
    from hermespy_amd import _lib

    rng = np.random.default_rng(nrx * 10 + ntx)
    k = 20000
    h = ((rng.standard_normal((nrx, ntx, k)) + 1j * rng.standard_normal((nrx, ntx, k))) / np.sqrt(2)).astype(np.complex64)
    y = ((rng.standard_normal((nrx, k)) + 1j * rng.standard_normal((nrx, k))) / np.sqrt(2)).astype(np.complex64)
    s2 = 0.05
    out = torch.empty((ntx, k), dtype=torch.complex64, device="cuda")
    ty, th = dev(y), dev(h)  # keep the uploads alive until the kernel has run
    _lib.check(_lib.load().hpx_equalize(_lib.ptr(ty), _lib.ptr(th), nrx, ntx, k, mode, s2, _lib.ptr(out),
                                        _lib.stream_handle()))
    h64, y64 = h.astype(np.complex128), y.astype(np.complex128)
    exp = oeq.zero_forcing(y64, h64) if mode == 0 else oeq.mmse(y64, h64, s2)
    got = out.cpu().numpy()
    # float32 normal equations + one refinement step; error grows with cond(H)^2 for ZF
    cond = np.linalg.cond(h64.transpose(2, 0, 1))
    well = cond < 30
    assert rel_err(got[:, well], exp[:, well]) < (1e-5 if mode == 1 else 2e-5)


def test_count_errors(cuda):
    from hermespy_amd import _lib

    rng = np.random.default_rng(5)
    for n in [0, 1, 3, 4, 1000, 229377]:
        a = rng.integers(0, 2, n).astype(np.uint8)
        b = rng.integers(0, 2, n).astype(np.uint8)
        acc = torch.zeros(1, dtype=torch.int64, device="cuda")
        ta, tb = dev(a, torch.uint8), dev(b, torch.uint8)
        _lib.check(_lib.load().hpx_count_errors(_lib.ptr(ta), _lib.ptr(tb), n, _lib.ptr(acc), _lib.stream_handle()))
        assert int(acc.item()) == int(np.abs(a.astype(np.int16) - b.astype(np.int16)).sum())


def test_channel_mirror_realization_matches_reference_draw_order(cuda):
    """A seeded channel draws los_angle, then the 100 + 2L + 2LN normals per realisation."""

    from hermespy_amd.channel import MultipathFadingChannel
    from hermespy_amd.core import SimulatedDevice

    ch = MultipathFadingChannel([0.0, 1e-6], [0.5, 0.5], [0.0, 1.0], num_sinusoids=5, doppler_frequency=10.0, seed=42)
    rng = np.random.default_rng(42)
    assert ch.los_angle == rng.uniform(-np.pi, np.pi)
    real = ch.realize()
    assert np.array_equal(real.scalars, rng.standard_normal(100 + 4 + 20))
    s = real.sample(SimulatedDevice(2, sampling_rate=1e6), SimulatedDevice(2, sampling_rate=1e6))
    params = fading.make_parameters([0.0, 1e-6], [0.5, 0.5], [0.0, 1.0], 5, 10.0)
    os_ = fading.sample_from_normals(params, real.scalars, 2, 2, 1e6)
    np.testing.assert_array_equal(s.nlos_angles, os_.nlos_angles)
    np.testing.assert_array_equal(s.spatial_response, os_.spatial_response)


def test_ls_estimate(cuda):
    """Pilot LS + griddata interpolation (hpx_ls_estimate) against the oracle's scipy path, 2 streams."""

    from hermespy_amd.estimation import OrthogonalLeastSquaresChannelEstimation
    from hermespy_amd.modem import ElementType, GridElement, GridResource, OFDMWaveform, SymbolSection
    from oracle import estimation as oest

    n = 512
    res = [GridResource(1, "CYCLIC", 0.0, [GridElement("DATA", n)]),
           GridResource(n // 4, "CYCLIC", 0.0, [GridElement("REFERENCE", 1), GridElement("DATA", 3)])]
    pattern = [0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0]
    wf = OFDMWaveform(res, [SymbolSection(1, pattern)], n, modulation_order=16)
    est = OrthogonalLeastSquaresChannelEstimation(wf, pilot_sequence=mapping.constellation(4))
    rng = np.random.default_rng(7)
    grid = ((rng.standard_normal((2, 14, n)) + 1j * rng.standard_normal((2, 14, n))) / np.sqrt(2)).astype(np.complex64)
    tgrid = dev(grid)
    got = est.estimate_channel(tgrid).cpu().numpy()
    mask = wf.resource_mask()
    frame = ofdm.Frame(n, np.zeros(14, int), np.where(mask[ElementType.REFERENCE.value], ofdm.REFERENCE,
                                                      ofdm.DATA).astype(np.int8))
    for s in range(2):
        ref = oest.ls_estimate(frame, grid[s].astype(np.complex128), est.pilots)
        assert rel_err(got[s], ref) < 1e-6
