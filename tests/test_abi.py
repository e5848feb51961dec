"""C ABI checks that need no GPU: the library loads, exports every symbol the
header declares, and its host-side mapping/demapping twins are bit-identical to
the oracle's restatement of the reference."""

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from oracle import mapping

HEADER = Path(__file__).resolve().parents[1] / "include" / "hermespy_amd.h"
MODS = [(2, 1), (4, 1), (16, 1), (64, 1), (256, 1), (4, 0), (8, 0), (16, 0)]


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(hpx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for required in ["hpx_fading_taps", "hpx_tdl_propagate", "hpx_awgn", "hpx_map_symbols", "hpx_demap_hard",
                     "hpx_ofdm_modulate", "hpx_ofdm_demodulate", "hpx_ideal_csi", "hpx_equalize",
                     "hpx_count_errors", "hpx_run_batch"]:
        assert required in names


def test_library_exports_every_declared_symbol(native_lib):
    from hermespy_amd import _lib

    for name in declared_functions():
        assert hasattr(native_lib, name), name
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"
    assert native_lib.hpx_abi_version() == _lib.ABI_VERSION == 8


def test_error_channel(native_lib):
    from hermespy_amd import _lib

    with pytest.raises(ValueError, match="not supported"):
        _lib.check(native_lib.hpx_constellation_host(12, 1, np.zeros(24).ctypes.data))
    with pytest.raises(ValueError, match="null pointer"):
        _lib.check(native_lib.hpx_fading_taps(None, 1, 1, 10, None, None))  # non-empty work needs pointers


@pytest.mark.parametrize("order,is_complex", MODS)
def test_host_constellation_bit_exact(native_lib, order, is_complex):
    pts = np.zeros(2 * order, dtype=np.float64)
    assert native_lib.hpx_constellation_host(order, is_complex, pts.ctypes.data) == 0
    got = pts[0::2] + 1j * pts[1::2]
    exp = mapping.constellation(order, bool(is_complex))
    assert np.array_equal(got.view(np.float64), exp.astype(np.complex128).view(np.float64))


@pytest.mark.parametrize("order,is_complex", MODS)
def test_host_demap_bit_exact(native_lib, order, is_complex):
    """Random points, every decision threshold and its float64 neighbours."""

    rng = np.random.default_rng(order + 7 * is_complex)
    pts = mapping.constellation(order, bool(is_complex))
    lim = 1.5 * np.max(np.abs(pts.real))
    x = rng.uniform(-lim, lim, 4000) + 1j * rng.uniform(-lim, lim, 4000)
    # thresholds in the scaled domain r = x * sqrt(2) * sqrt(M') are integers
    pam = {2: 2, 4: 2 if is_complex else 4, 16: 4 if is_complex else 16, 64: 8, 256: 16, 8: 8}[order]
    scale = (np.sqrt(2) if (is_complex and order > 2) else 1.0) * {2: 1.0, 4: np.sqrt(5), 8: np.sqrt(21),
                                                                   16: np.sqrt(85)}[pam]
    thr = np.arange(-pam, pam + 1, dtype=np.float64) / scale
    edge = np.concatenate([thr, np.nextafter(thr, np.inf), np.nextafter(thr, -np.inf)])
    x = np.concatenate([x, edge + 1j * edge[::-1], edge + 0.3j, 0.3 + 1j * edge])
    bits = np.zeros(x.size * int(np.log2(order)), dtype=np.uint8)
    xx = np.ascontiguousarray(np.stack([x.real, x.imag], -1))
    assert native_lib.hpx_demap_hard_host(xx.ctypes.data, x.size, order, is_complex, bits.ctypes.data) == 0
    exp = mapping.detect_bits(x if is_complex else x.real + 0j, order, bool(is_complex)).astype(np.uint8)
    assert np.array_equal(bits, exp)


@pytest.mark.parametrize("nr,nt", [(1, 1), (2, 2), (4, 4), (8, 8), (4, 2), (8, 4)])
def test_host_spatial_detector(native_lib, nr, nt):
    """Jacobi eigen-decomposition of S^H S against numpy.linalg.eigh, and the per-element
    filter V diag(conj(c) / (|c|^2 lambda + s2)) V^H S^H against the MMSE closed form."""

    from oracle import equalize

    rng = np.random.default_rng(nr * 8 + nt)
    r = 0.5 ** np.abs(np.subtract.outer(np.arange(nr), np.arange(nr)))
    rt = 0.5 ** np.abs(np.subtract.outer(np.arange(nt), np.arange(nt)))
    s = r @ np.exp(2j * np.pi * rng.random((nr, nt))) @ rt
    lam = np.zeros(nt)
    v = np.zeros((nt, nt), np.complex128)
    a = np.zeros((nt, nr), np.complex128)
    sc = np.ascontiguousarray(s)
    assert native_lib.hpx_spatial_detector_host(nr, nt, sc.ctypes.data, lam.ctypes.data, v.ctypes.data,
                                                a.ctypes.data) == 0
    g = s.conj().T @ s
    np.testing.assert_allclose(np.sort(lam), np.linalg.eigvalsh(g), rtol=1e-10, atol=1e-12 * np.abs(g).max())
    np.testing.assert_allclose(v @ np.diag(lam) @ v.conj().T, g, atol=1e-12 * np.abs(g).max())
    np.testing.assert_allclose(v.conj().T @ v, np.eye(nt), atol=1e-12)
    np.testing.assert_allclose(a, v.conj().T @ s.conj().T, atol=1e-12)
    c = 0.7 - 0.2j
    y = rng.standard_normal(nr) + 1j * rng.standard_normal(nr)
    for s2 in (0.0, 0.1):
        if s2 == 0.0 and nr < nt:
            continue
        x = v @ ((np.conj(c) / (abs(c) ** 2 * lam + s2)) * (a @ y))
        h = (c * s)[:, :, None]
        ref = equalize.mmse(y[:, None], h, s2)[:, 0] if s2 > 0 else equalize.zero_forcing(y[:, None], h)[:, 0]
        np.testing.assert_allclose(x, ref, rtol=1e-9, atol=1e-12)


def test_new_entry_points_validate_without_a_device(native_lib):
    """Argument checks of the drop-index / recorded-normals and section-statistics entry points
    (no device call happens when they fail)."""
    import ctypes

    from hermespy_amd import _lib

    stats = _lib.HpxSectionStats()
    with pytest.raises(ValueError, match="sections"):
        _lib.check(native_lib.hpx_section_accumulate(None, None, 0, None, 1, 10, 0, ctypes.byref(stats), None))
    with pytest.raises(ValueError, match="bits per drop"):
        _lib.check(native_lib.hpx_section_accumulate(None, None, 0, None, 0, 10, 4, ctypes.byref(stats), None))
    with pytest.raises(ValueError, match="null"):
        _lib.check(native_lib.hpx_section_accumulate(None, None, 0, None, 8, 10, 4, ctypes.byref(stats), None))
    assert native_lib.hpx_section_accumulate(None, None, 0, None, 8, 0, 4, ctypes.byref(stats), None) == 0
    with pytest.raises(ValueError, match="65535"):
        _lib.check(native_lib.hpx_realization_normals_batch(1, None, 0, 70000, 100, None, None))
    with pytest.raises(ValueError, match="32 bits"):
        _lib.check(native_lib.hpx_realization_normals_batch(1, None, 2**32 - 1, 2, 100, None, None))
    assert native_lib.hpx_realization_normals_batch(1, None, 0, 0, 100, None, None) == 0
    link = _lib.HpxLink()
    out = _lib.HpxBatchOut()
    with pytest.raises(ValueError, match="null input or output"):
        _lib.check(native_lib.hpx_run_batch_drops(ctypes.byref(link), None, None, None, None, None, None, 1, 0, 1,
                                                  ctypes.byref(out), None))


def test_float64_path_validates_without_a_device(native_lib):
    """hpx_run_batch_f64 refuses links it does not cover before any device call."""
    import ctypes

    from hermespy_amd import _lib
    from hermespy_amd.configs import mimo4x4_1024, siso_64
    from hermespy_amd.uneven import _c_link

    big = _c_link(mimo4x4_1024())
    with pytest.raises(ValueError, match="single-antenna"):
        _lib.check(native_lib.hpx_run_batch_f64(ctypes.byref(big), 0x10, 0x20, 1, 0, 4, 0x30, 0x40, None, None))
    cfg = siso_64()
    wide = _c_link(cfg)
    wide.num_subcarriers = 512
    with pytest.raises(ValueError, match="at most 256"):
        _lib.check(native_lib.hpx_run_batch_f64(ctypes.byref(wide), 0x10, 0x20, 1, 0, 4, 0x30, 0x40, None, None))
    small = _c_link(cfg)
    with pytest.raises(ValueError, match="null"):
        _lib.check(native_lib.hpx_run_batch_f64(ctypes.byref(small), None, 0x20, 1, 0, 4, 0x30, 0x40, None, None))
    with pytest.raises(ValueError, match="drop range"):
        _lib.check(native_lib.hpx_run_batch_f64(ctypes.byref(small), 0x10, 0x20, 1, 2**32 - 1, 4, 0x30, 0x40, None,
                                                None))
    assert native_lib.hpx_run_batch_f64(ctypes.byref(small), 0x10, 0x20, 1, 0, 0, 0x30, 0x40, None, None) == 0


def test_wrappers_reject_host_tensors(native_lib):
    """Every array argument of the library is a device pointer: a host tensor, or a view that
    is not contiguous, is refused with ValueError before any launch."""

    import torch

    from hermespy_amd import _lib
    from hermespy_amd.configs import siso_64
    from hermespy_amd.simulation import BatchedLink

    with pytest.raises(ValueError, match="device tensor"):
        _lib.ptr(torch.zeros(4), "noise_power")
    link = BatchedLink(siso_64(), device="cpu")
    with pytest.raises(ValueError, match="device tensor"):
        link.run(1, 0, torch.zeros(4, dtype=torch.float64), torch.zeros(4, dtype=torch.int32))
    with pytest.raises(ValueError, match="one entry per drop"):
        link.run(1, 0, torch.zeros(4, dtype=torch.float64), torch.zeros(2, dtype=torch.int32))


def test_massive_run_checks_its_buffers(native_lib):
    import torch

    from hermespy_amd.massive import MassiveBatch, MassiveConfig

    cfg = MassiveConfig()
    batch = MassiveBatch.__new__(MassiveBatch)  # buffers are not needed to reach the checks
    batch.cfg, batch.chunk = cfg, 4
    with pytest.raises(ValueError, match="float64"):
        batch.run(1, 0, torch.zeros(2, dtype=torch.float32), torch.zeros(2, dtype=torch.int32))
    with pytest.raises(ValueError, match="one entry per drop"):
        batch.run(1, 0, torch.zeros(2, dtype=torch.float64), torch.zeros(1, dtype=torch.int32))


@pytest.mark.parametrize("order", [4, 16, 64, 256])
def test_host_demap_f32_matches_f64(native_lib, order):
    """The float32 decision path (boundaries found by bisection of the float64 amplitude) decides
    exactly as the float64 demapper on the same float32 inputs: random points, every threshold's
    float32 neighbours (8 ulps each side), signed zeros, the largest floats and non-finite values."""

    rng = np.random.default_rng(order)
    pts = mapping.constellation(order, True)
    lim = 1.5 * np.max(np.abs(pts.real))
    x = (rng.uniform(-lim, lim, 20000) + 1j * rng.uniform(-lim, lim, 20000)).astype(np.complex64)
    pam = {4: 2, 16: 4, 64: 8, 256: 16}[order]
    scale = np.sqrt(2) * {2: 1.0, 4: np.sqrt(5), 8: np.sqrt(21), 16: np.sqrt(85)}[pam]
    thr = (np.arange(-pam, pam + 1, dtype=np.float64) / scale).astype(np.float32)
    edge = [thr]
    for _ in range(8):
        edge.append(np.nextafter(edge[-1], np.float32(np.inf)))
    down = [thr]
    for _ in range(8):
        down.append(np.nextafter(down[-1], np.float32(-np.inf)))
    e = np.concatenate(edge + down + [np.array([0.0, -0.0, 3.4e38, -3.4e38, np.inf, -np.inf, np.nan, 1e-45,
                                                -1e-45], dtype=np.float32)])
    x = np.concatenate([x, (e + 1j * e[::-1]).astype(np.complex64), (e + 0.3j).astype(np.complex64),
                        (0.3 + 1j * e).astype(np.complex64)])
    bps = int(np.log2(order))
    got = np.zeros(x.size * bps, dtype=np.uint8)
    xx = np.ascontiguousarray(np.stack([x.real, x.imag], -1).astype(np.float32))
    assert native_lib.hpx_demap_hard_f32_host(xx.ctypes.data, x.size, order, got.ctypes.data) == 0
    exp = np.zeros_like(got)
    x64 = np.ascontiguousarray(np.stack([x.real, x.imag], -1).astype(np.float64))
    assert native_lib.hpx_demap_hard_host(x64.ctypes.data, x.size, order, 1, exp.ctypes.data) == 0
    assert np.array_equal(got, exp)
    # and the float64 demapper is the oracle's restatement of the reference (finite inputs)
    fin = np.isfinite(x.real) & np.isfinite(x.imag)
    ref = mapping.detect_bits(x[fin].astype(np.complex128), order, True).astype(np.uint8)
    assert np.array_equal(got.reshape(-1, bps)[fin].reshape(-1), ref)
