"""GPU parity of the pilot-based estimation path: LS at the pilots, the LMMSE GEMM, the bfloat16
MFMA convolutions of the learned estimator (forward, backward-data, weight gradients) against
torch float64 on the same bfloat16-rounded operands, detection on estimated channels, and a
short training run of the learned estimator on simulator drops."""

import numpy as np
import pytest
import torch
import torch.nn.functional as F

from conftest import rel_err

pytestmark = pytest.mark.gpu

SEED = 20261202


def bf16(x):
    from hermespy_amd.learned import bf16_round

    return bf16_round(np.asarray(x, dtype=np.float32))


def test_pilot_ls_and_lmmse_apply(cuda):
    from hermespy_amd import _lib
    from hermespy_amd.configs import mimo2x2_512_pilots
    from hermespy_amd.lmmse import LMMSEChannelEstimation

    cfg = mimo2x2_512_pilots()
    est = LMMSEChannelEstimation(cfg, snr_db=[0.0, 20.0])
    rng = np.random.default_rng(1)
    d = 5
    g = (rng.standard_normal((d, 2, 14, 512)) + 1j * rng.standard_normal((d, 2, 14, 512))).astype(np.complex64)
    ls = est.pilot_ls(torch.as_tensor(g, device=cuda), d)
    idx, val = est.pilot_index.cpu().numpy(), est.pilot_values.cpu().numpy()
    ref = np.stack([g.reshape(d, 2, -1)[:, :, idx[t]] / val[t] for t in range(2)])
    assert rel_err(ls.cpu().numpy(), ref) < 1e-6
    snr_index = np.array([0, 1, 1, 0, 1])
    out = est.apply(ls, snr_index).cpu().numpy()
    w = est.filters.cpu().numpy().astype(np.complex128)
    lsn = ls.cpu().numpy().astype(np.complex128)
    for t in range(2):
        for dd in range(d):
            for r in range(2):
                want = w[snr_index[dd], t] @ lsn[t, dd, r]
                assert rel_err(out[t, dd, r], want) < 1e-5


def circular_conv(x, w, b, dil):
    """float64 reference: x [S][64][cin], w [cout][taps][cin] -> [S][64][cout], circular, dilated."""

    h = (w.shape[1] // 2) * dil
    xt = F.pad(torch.as_tensor(x, dtype=torch.float64).permute(0, 2, 1), (h, h), mode="circular")
    bt = None if b is None else torch.as_tensor(b, dtype=torch.float64)
    return F.conv1d(xt, torch.as_tensor(w, dtype=torch.float64).permute(0, 2, 1), bt, dilation=dil).permute(0, 2, 1)


@pytest.mark.parametrize("cin,cout,mode,taps,dil", [(8, 64, 1, 5, 1), (64, 64, 1, 5, 4), (64, 64, 1, 5, 16),
                                                    (64, 224, 0, 5, 1), (224, 64, 0, 5, 16), (8, 224, 0, 17, 1)])
def test_conv1d_forward_matches_torch(cuda, cin, cout, mode, taps, dil):
    from hermespy_amd import _lib

    rng = np.random.default_rng(cin + cout + taps)
    s = 3
    x = bf16(rng.standard_normal((s, 64, cin)))
    w = bf16(rng.standard_normal((cout, taps, cin)) / np.sqrt(taps * cin))
    b = rng.standard_normal(cout).astype(np.float32)
    xd = torch.as_tensor(x, device=cuda).to(torch.bfloat16)
    wd = torch.as_tensor(w, device=cuda).to(torch.bfloat16)
    bd = torch.as_tensor(b, device=cuda)
    of = torch.empty((s, 64, cout), dtype=torch.float32, device=cuda)
    ob = torch.empty((s, 64, cout), dtype=torch.bfloat16, device=cuda)
    _lib.check(_lib.load().hpx_conv1d_forward(_lib.ptr(xd), s, cin, taps, dil, _lib.ptr(wd), _lib.ptr(bd), cout,
                                              mode, _lib.ptr(of), _lib.ptr(ob), _lib.stream_handle()))
    torch.cuda.synchronize()
    ref = circular_conv(x, w, b, dil).numpy()
    assert rel_err(of.cpu().numpy(), ref) < 1e-5
    act = F.gelu(torch.as_tensor(ref), approximate="tanh").numpy() if mode == 1 else ref
    got = ob.float().cpu().numpy()
    assert np.all(np.abs(got - act) <= 2.0 ** -7 * np.abs(act) + 1e-6)


def test_conv1d_backward_matches_torch(cuda):
    """GELU layer: dz = bf16(g * gelu'(pre)); dW, db by hpx_conv1d_wgrad; dX by the forward kernel
    on the transposed, flipped weights (circular, dilation 4)."""

    from hermespy_amd import _lib

    lib, st = _lib.load(), _lib.stream_handle()
    rng = np.random.default_rng(3)
    s, cin, cout, dil = 70, 64, 224, 4
    x = bf16(rng.standard_normal((s, 64, cin)))
    wf = (rng.standard_normal((cout, 5, cin)) / np.sqrt(5 * cin)).astype(np.float32)
    pre = rng.standard_normal((s, 64, cout)).astype(np.float32)
    g = rng.standard_normal((s, 64, cout)).astype(np.float32)
    dev = lambda a, dt=torch.float32: torch.as_tensor(a, device=cuda).to(dt)
    xd, pd, gd, wfd = dev(x, torch.bfloat16), dev(pre), dev(g), dev(wf)
    dz = torch.empty((s, 64, cout), dtype=torch.bfloat16, device=cuda)
    _lib.check(lib.hpx_gelu_backward(_lib.ptr(gd), _lib.ptr(pd), gd.numel(), _lib.ptr(dz), st))
    w16 = torch.empty((cout, 5, cin), dtype=torch.bfloat16, device=cuda)
    wt16 = torch.empty((cin, 5, cout), dtype=torch.bfloat16, device=cuda)
    _lib.check(lib.hpx_conv1d_weight_copies(_lib.ptr(wfd), cout, cin, 5, _lib.ptr(w16), _lib.ptr(wt16), st))
    ws = torch.empty(int(lib.hpx_conv1d_wgrad_workspace(s, cin, cout, 5)) // 4 + 1, dtype=torch.float32, device=cuda)
    gw = torch.empty((cout, 5, cin), dtype=torch.float32, device=cuda)
    gb = torch.empty(cout, dtype=torch.float32, device=cuda)
    _lib.check(lib.hpx_conv1d_wgrad(_lib.ptr(dz), _lib.ptr(xd), s, cin, cout, 5, dil, _lib.ptr(ws), _lib.ptr(gw),
                                    _lib.ptr(gb), st))
    gx = torch.empty((s, 64, cin), dtype=torch.float32, device=cuda)
    _lib.check(lib.hpx_conv1d_forward(_lib.ptr(dz), s, cout, 5, dil, _lib.ptr(wt16), None, cin, 0, _lib.ptr(gx), None,
                                      st))
    torch.cuda.synchronize()

    pre64 = torch.as_tensor(pre, dtype=torch.float64).requires_grad_(True)
    F.gelu(pre64, approximate="tanh").backward(torch.as_tensor(g, dtype=torch.float64))
    dz_ref = bf16(pre64.grad.numpy())
    got_dz = dz.float().cpu().numpy()
    assert np.all(np.abs(got_dz - dz_ref) <= 2.0 ** -7 * np.abs(dz_ref) + 1e-6)
    # autograd of the circular dilated convolution on the same bf16 operands
    xt = torch.as_tensor(x, dtype=torch.float64).requires_grad_(True)
    wt = torch.as_tensor(bf16(wf), dtype=torch.float64).requires_grad_(True)
    h = 2 * dil
    out = F.conv1d(F.pad(xt.permute(0, 2, 1), (h, h), mode="circular"), wt.permute(0, 2, 1), dilation=dil)
    out.backward(torch.as_tensor(got_dz, dtype=torch.float64).permute(0, 2, 1))
    assert rel_err(gw.cpu().numpy(), wt.grad.numpy()) < 1e-5
    assert rel_err(gb.cpu().numpy(), got_dz.sum(axis=(0, 1))) < 1e-5
    assert rel_err(gx.cpu().numpy(), xt.grad.numpy()) < 1e-5
    assert np.array_equal(wt16.float().cpu().numpy(), np.flip(bf16(wf), axis=1).transpose(2, 1, 0))


def test_detect_on_ideal_estimates_matches_fused_kernel(cuda):
    """Detection from exported ideal CSI reproduces the fused kernel's decisions: bit-exact demapping
    of its own symbols, and the same error counts up to symbols within rounding of a boundary."""

    from hermespy_amd import _lib
    from hermespy_amd.configs import mimo2x2_512_pilots
    from hermespy_amd.simulation import BatchedLink
    from oracle import mapping

    cfg = mimo2x2_512_pilots()
    n = 16
    snr = np.asarray(cfg.snr_db)[np.arange(n) % len(cfg.snr_db)]
    npw = torch.as_tensor(cfg.noise_power(snr), dtype=torch.float64, device=cuda)
    bl = BatchedLink(cfg)
    err = torch.empty(n, dtype=torch.int32, device=cuda)
    grid = torch.empty((n, 2, 14, 512), dtype=torch.complex64, device=cuda)
    csi = torch.empty((n, 14, 512), dtype=torch.complex64, device=cuda)
    sp = torch.empty((n, 2, 2), dtype=torch.complex64, device=cuda)
    bl.run(SEED, 40, npw, err, doppler=cfg.draw_doppler(40, n), grid=grid, csi=csi, spatial=sp)
    h = (sp.permute(2, 0, 1)[:, :, :, None, None] * csi[None, :, None]).reshape(2, n, 2, -1).contiguous()
    err2 = torch.empty(n, dtype=torch.int32, device=cuda)
    eq = torch.empty((n, 2, cfg.num_data), dtype=torch.complex64, device=cuda)
    _lib.check(_lib.load().hpx_detect_estimated(_lib.ptr(grid), _lib.ptr(h), n, 2, 2, 14, 512, _lib.ptr(bl.elements),
                                                _lib.ptr(bl.word_data), cfg.num_data, 16, _lib.HPX_DETECT_MMSE,
                                                _lib.ptr(npw), SEED, 40, _lib.ptr(err2), _lib.ptr(eq),
                                                _lib.stream_handle()))
    torch.cuda.synchronize()
    e1, e2, eqn = err.cpu().numpy(), err2.cpu().numpy(), eq.cpu().numpy()
    from oracle import philox

    for i in range(n):
        tx = np.stack([philox.stream_bits(SEED, 40 + i, t, cfg.num_data * 4) for t in range(2)])
        rx = np.stack([mapping.detect_bits(eqn[i, t].astype(np.complex128), 16).astype(np.uint8) for t in range(2)])
        assert int(e2[i]) == int(np.sum(rx != tx))
        assert abs(int(e1[i]) - int(e2[i])) <= max(2, int(2e-4 * tx.size)), (i, e1[i], e2[i])


def test_learned_estimator_trains(cuda):
    """A few Adam steps on simulator drops lower the SNR-balanced loss (mean NMSE in dB over the SNR
    points), and that loss equals the NMSE of the estimate path computed independently per drop."""

    from hermespy_amd import _lib
    from hermespy_amd.configs import mimo2x2_512_pilots
    from hermespy_amd.learned import ConvNetEstimator, Trainer

    cfg = mimo2x2_512_pilots()
    net = ConvNetEstimator(cfg)
    assert 135_000 < net.num_params < 155_000
    tr = Trainer(net, drops_per_step=66, lr=3e-3)
    first = 1 << 30
    before = float(tr.step(first, lr=0.0))
    for k in range(30):
        tr.step(first + 66 * (k + 1))
    after = float(tr.step(first, lr=0.0))
    assert np.isfinite(before) and after < before - 1.0, (before, after)
    # the loss from the estimate path: per-drop error and power, grouped by SNR point
    ls = tr.simulate(first)
    est = net.estimate(ls, tr.noise_power)
    e = torch.zeros(66, dtype=torch.float64, device=cuda)
    p = torch.zeros(66, dtype=torch.float64, device=cuda)
    _lib.check(_lib.load().hpx_estimate_error(_lib.ptr(est), _lib.ptr(tr.csi), _lib.ptr(tr.spatial), 66, 2, 2,
                                              14 * 512, _lib.ptr(e), _lib.ptr(p), _lib.stream_handle()))
    g = (first + np.arange(66)) % len(cfg.snr_db)
    en, pn = e.cpu().numpy(), p.cpu().numpy()
    want = np.mean([10 * np.log10(en[g == k].sum() / pn[g == k].sum()) for k in range(len(cfg.snr_db))])
    assert abs(want - after) < 1e-3, (want, after)


def test_skip_wgrad_17_taps_matches_torch(cuda):
    from hermespy_amd import _lib

    lib, st = _lib.load(), _lib.stream_handle()
    rng = np.random.default_rng(8)
    s, cin, cout, taps = 40, 8, 224, 17
    x = bf16(rng.standard_normal((s, 64, cin)))
    dz = bf16(rng.standard_normal((s, 64, cout)))
    xd = torch.as_tensor(x, device=cuda).to(torch.bfloat16)
    dzd = torch.as_tensor(dz, device=cuda).to(torch.bfloat16)
    ws = torch.empty(int(lib.hpx_conv1d_wgrad_workspace(s, cin, cout, taps)) // 4 + 1, dtype=torch.float32, device=cuda)
    gw = torch.empty((cout, taps, cin), dtype=torch.float32, device=cuda)
    gb = torch.empty(cout, dtype=torch.float32, device=cuda)
    _lib.check(lib.hpx_conv1d_wgrad(_lib.ptr(dzd), _lib.ptr(xd), s, cin, cout, taps, 1, _lib.ptr(ws), _lib.ptr(gw),
                                    _lib.ptr(gb), st))
    torch.cuda.synchronize()
    xt = torch.as_tensor(x, dtype=torch.float64)
    wt = torch.zeros((cout, taps, cin), dtype=torch.float64, requires_grad=True)
    circular_conv_t = F.conv1d(F.pad(xt.permute(0, 2, 1), (taps // 2, taps // 2), mode="circular"),
                               wt.permute(0, 2, 1))
    circular_conv_t.backward(torch.as_tensor(dz, dtype=torch.float64).permute(0, 2, 1))
    ref = wt.grad.numpy()
    assert rel_err(gw.cpu().numpy(), ref) < 1e-5
    assert rel_err(gb.cpu().numpy(), dz.sum(axis=(0, 1))) < 1e-5


def test_trained_estimator_tracks_lmmse(cuda):
    """The committed weights: on 440 fresh drops (40 per SNR point) the learned estimator's NMSE is
    within 0.5 dB of the LMMSE fallback's at every SNR point (it is trained to beat it)."""

    from hermespy_amd import _lib
    from hermespy_amd.configs import mimo2x2_512_pilots
    from hermespy_amd.learned import ConvNetEstimator
    from hermespy_amd.lmmse import LMMSEChannelEstimation
    from hermespy_amd.simulation import BatchedLink

    cfg = mimo2x2_512_pilots()
    n = 440
    snr = np.asarray(cfg.snr_db)
    sidx = np.arange(n) % snr.size
    npw = torch.as_tensor(cfg.noise_power(snr[sidx]), dtype=torch.float64, device=cuda)
    bl = BatchedLink(cfg)
    err = torch.empty(n, dtype=torch.int32, device=cuda)
    grid = torch.empty((n, 2, 14, 512), dtype=torch.complex64, device=cuda)
    csi = torch.empty((n, 14, 512), dtype=torch.complex64, device=cuda)
    sp = torch.empty((n, 2, 2), dtype=torch.complex64, device=cuda)
    first = 3_000_000  # away from the training drops (from 2^30) and the bench drops (from 0)
    bl.run(SEED, first, npw, err, doppler=cfg.draw_doppler(first, n), grid=grid, csi=csi, spatial=sp)
    lm = LMMSEChannelEstimation(cfg)
    net = ConvNetEstimator.load(cfg)
    lib, st = _lib.load(), _lib.stream_handle()

    def nmse(est):
        e = torch.zeros(n, dtype=torch.float64, device=cuda)
        p = torch.zeros(n, dtype=torch.float64, device=cuda)
        _lib.check(lib.hpx_estimate_error(_lib.ptr(est), _lib.ptr(csi), _lib.ptr(sp), n, 2, 2, 14 * 512, _lib.ptr(e),
                                          _lib.ptr(p), st))
        en, pn = e.cpu().numpy(), p.cpu().numpy()
        return np.array([10 * np.log10(en[sidx == k].sum() / pn[sidx == k].sum()) for k in range(snr.size)])

    a = nmse(lm.apply(lm.pilot_ls(grid, n), sidx))
    b = nmse(net.estimate_grid(grid, lm.pilot_index, lm.pilot_values, npw))
    assert np.all(b - a < 0.5), (a.round(2), b.round(2))
