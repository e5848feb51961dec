"""The massive-MIMO oracle: its extended spatial draw keeps the reference's layout for arrays of at
most 10 antennas, and its eigen-form LMMSE equals the per-element solve of the definition."""

import numpy as np


def small_massive(num_rx=16, users=4):
    from oracle import massive

    fs = 256 * 30e3
    tau = np.arange(6) / fs
    p = np.exp(-np.arange(6) / 2.0)
    k = np.zeros(6)
    k[0] = 1.0
    r = (0.7 ** np.abs(np.subtract.outer(np.arange(num_rx), np.arange(num_rx)))).astype(complex)
    return massive.MassiveLink(num_rx, users, 256, 200, 14, 36, 256, fs, tau, p / p.sum(), k, 2 * np.pi * 400, corr_rx=r)


def test_extended_draw_matches_reference_layout_up_to_ten_antennas():
    from oracle import fading, massive, philox

    link = small_massive(num_rx=8, users=4)
    z = philox.realization_normals(3, 9, link.num_normals)
    a = massive.sample_from_normals(link, z)
    b = fading.sample_from_normals(link.params, z, 8, 4, link.bandwidth)
    assert link.num_normals == link.params.num_normals
    np.testing.assert_array_equal(a.spatial_response, b.spatial_response)
    np.testing.assert_array_equal(a.nlos_phases, b.nlos_phases)
    np.testing.assert_array_equal(a.los_angles, b.los_angles)


def test_eigen_lmmse_equals_direct_solve():
    from oracle import massive, ofdm

    link = small_massive()
    res = massive.run_drop(link, 5, 3, 25.0)
    yd = ofdm.pick(link.frame, res.grid)
    cd = ofdm.pick(link.frame, res.csi)
    sel = np.arange(0, yd.shape[1], 97)
    want = massive.lmmse_direct(yd[:, sel], res.spatial, cd[sel], link.noise_power(25.0))
    np.testing.assert_allclose(res.equalized[:, sel], want, rtol=0, atol=1e-12)
    assert res.errors < 0.05 * res.tx_bits.size


def test_centred_band_sits_on_the_middle_bins():
    link = small_massive()
    from oracle import ofdm

    data_cols = np.flatnonzero(link.frame.mask[0] == ofdm.DATA)
    assert data_cols[0] == 128 - 100 and data_cols[-1] == 128 + 99
