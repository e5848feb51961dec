"""The HDF5 drop store without a GPU: element types, attributes, groups, growth along the
first axis, error behaviour, the file read by HDF5's own h5dump, and the campaign layout
(link state, drop columns, offsets)."""

import os
import shutil
import subprocess

import numpy as np
import pytest

from hermespy_amd import record
from hermespy_amd.configs import CONFIGS

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)


@pytest.fixture(scope="module", autouse=True)
def h5lib():
    if not record.H5_LIB_PATH.exists():
        from hermespy_amd.build import build_h5

        if build_h5() is None:
            pytest.skip("no libhdf5 to build the drop store against")
    return record.load_h5()


@pytest.mark.parametrize("dtype", [np.uint8, np.int8, np.int32, np.uint32, np.int64, np.float32, np.float64,
                                   np.complex64, np.complex128])
def test_dataset_round_trip(tmp_path, dtype):
    rng = np.random.default_rng(1)
    a = rng.standard_normal((7, 3, 5)) * 100
    if np.dtype(dtype).kind == "c":
        a = a + 1j * rng.standard_normal(a.shape)
    a = a.astype(dtype)
    f = record.H5File(tmp_path / "t.h5", record.CREATE)
    f.append("/g/h/x", a[:4])
    f.append("/g/h/x", a[4:])
    f.append("/g/h/x", a[:0])
    assert f.shape("/g/h/x") == (np.dtype(dtype).type, (7, 3, 5))
    f.close()
    with record.H5File(tmp_path / "t.h5") as f:
        assert np.array_equal(f.read("/g/h/x"), a)
        assert np.array_equal(f.read("/g/h/x", 2, 3), a[2:5])
        if np.dtype(dtype).kind != "c":
            assert np.array_equal(f.read("/g/h/x", dtype=np.float64), a.astype(np.float64))
        assert f.list("/g") == ["h"] and f.exists("/g/h/x") and not f.exists("/g/nope/x")


def test_attributes_and_errors(tmp_path):
    p = tmp_path / "a.h5"
    with record.H5File(p, record.CREATE) as f:
        f.create_group("/c")
        f.set_attr("/c", "s", "ünïcode text")
        f.set_attr("/c", "i", 2**40 + 3)
        f.set_attr("/c", "x", 0.125)
        f.set_attr("/c", "v", np.array([1.5, 2.5, 3.5]))
        f.set_attr("/c", "x", 0.25)  # replaced
        assert f.get_str("/c", "s") == "ünïcode text"
        assert f.get_attr("/c", "i", np.int64)[0] == 2**40 + 3
        assert f.get_attr("/c", "x")[0] == 0.25 and np.array_equal(f.get_attr("/c", "v"), [1.5, 2.5, 3.5])
        f.append("/c/d", np.zeros((3, 2), np.float32))
        with pytest.raises(ValueError, match="row shape"):
            f.append("/c/d", np.zeros((3, 4), np.float32))
        with pytest.raises(ValueError, match="element type"):
            f.append("/c/d", np.zeros((3, 2), np.float64))
        with pytest.raises(ValueError, match="beyond"):
            f.read("/c/d", 2, 5)
        with pytest.raises(OSError, match="no dataset"):
            f.read("/c/missing")
        with pytest.raises(OSError, match="no attribute"):
            f.get_attr("/c", "missing")
    with pytest.raises(OSError, match="cannot open"):
        record.H5File(p, record.CREATE)  # exists
    with pytest.raises(OSError, match="cannot open"):
        record.H5File(tmp_path / "absent.h5", record.READ)


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not installed")
def test_hdf5_tools_read_the_store(tmp_path):
    """HDF5's own h5dump reads the file: attributes, the {r, i} complex compound, values."""

    p = tmp_path / "d.h5"
    with record.H5File(p, record.CREATE) as f:
        f.set_attr("/", "version", "v1")
        f.append("/z", np.array([1 + 2j, -3.5 - 0.25j], dtype=np.complex64))
        f.append("/n", np.array([7, 8, 9], dtype=np.uint32))
    out = subprocess.run([H5DUMP, str(p)], capture_output=True, text=True, check=True).stdout
    assert '"v1"' in out and 'H5T_STD_U32LE' in out and "7, 8, 9" in out
    assert 'H5T_IEEE_F32LE "r";' in out and 'H5T_IEEE_F32LE "i";' in out
    assert "1," in out and "2" in out and "-3.5," in out and "-0.25" in out


@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_campaign_state_and_columns(tmp_path, name):
    cfg = CONFIGS[name]()
    p = tmp_path / "c.h5"
    rng = np.random.default_rng(2)
    rec = record.DropRecorder(p, cfg, 2**63 + 5, campaign="run", normals=True)
    cols = []
    for start, n in ((0, 5), (5, 3)):
        c = dict(drop_index=np.arange(start, start + n, dtype=np.uint32), noise_power=rng.uniform(size=n),
                 errors=rng.integers(0, 100, n).astype(np.int32), section=np.arange(n, dtype=np.int32) % 2,
                 normals=rng.standard_normal((n, cfg.num_normals)))
        rec.record(**c)
        cols.append(c)
    rec.close()
    full = {k: np.concatenate([c[k] for c in cols]) for k in cols[0]}
    rp = record.DropReplay(p, "run", drop_offset=2)
    assert rp.seed == 2**63 + 5 and rp.num_drops == 6 and rp.has_normals
    assert record._same_link(rp.link, cfg)
    got = rp.read(0, 6)
    for k, v in full.items():
        assert np.array_equal(got[k], v[2:]), k
    rp.close()
    # appending to the campaign continues it; another link is refused
    rec = record.DropRecorder(p, cfg, 2**63 + 5, campaign="run")
    assert rec.count == 8
    rec.close()
    other = CONFIGS["siso-64" if name != "siso-64" else "mimo4x4-1024"]()
    with pytest.raises(ValueError, match="another link"):
        record.DropRecorder(p, other, 2**63 + 5, campaign="run")
    with pytest.raises(ValueError, match="Drop offset"):
        record.DropReplay(p, "run", drop_offset=8)
    with pytest.raises(ValueError, match="no recorded drops"):
        record.DropReplay(p, "absent")


def test_store_library_exports_every_declared_symbol():
    import re
    from pathlib import Path

    header = Path(__file__).resolve().parents[1] / "include" / "hermespy_amd_h5.h"
    text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
    names = sorted(set(re.findall(r"\b(hpx_h5_[a-z0-9_]+)\s*\(", text)))
    lib = record.load_h5()
    assert len(names) == 11
    for name in names:
        assert hasattr(lib, name), name


def test_reference_realization_layout_round_trip(tmp_path):
    """Drops written in the reference's recording layout (SimulatedDrop groups holding a
    MultipathFadingRealization, ``fading.py:540-556``, ``consistent.py:93-95,254-255``) read back
    to the same realisation scalars, and the mirror samples the same channel from them."""

    from hermespy_amd.channel import CustomAntennaCorrelation, MultipathFadingChannel, MultipathFadingRealization
    from hermespy_amd.core import SimulatedDevice

    cfg = CONFIGS["mimo4x4-1024"]()
    rng = np.random.default_rng(5)
    z = rng.standard_normal((3, cfg.num_normals))
    path = tmp_path / "ref.h5"
    record.write_reference_drops(path, cfg, z, campaign="c1")
    record.write_reference_drops(path, cfg, z[:1], campaign="c1")  # appends drop 0003
    back = record.read_reference_drops(path, cfg, campaign="c1")
    assert np.array_equal(back, np.concatenate([z, z[:1]]))
    with record.H5File(path) as f:
        assert f.list("/c1/drops") == ["0000", "0001", "0002", "0003"]
        d = "/c1/drops/0001"
        assert f.get_str(d, "type") == "hermespy.simulation.drop.SimulatedDrop"
        r = f"{d}/channel_realizations/0000"
        assert f.get_str(r, "type") == "hermespy.channel.fading.fading.MultipathFadingRealization"
        assert f.get_str(f"{r}/random_realization", "type") == "hermespy.channel.consistent.StaticConsistentRealization"
        assert np.array_equal(f.read(f"{r}/random_realization/scalar_samples"), z[1])
        L, N = cfg.num_paths, cfg.num_sinusoids
        for name, shape, offset in [("antenna_correlation_variable", (10, 10), 0), ("los_angles_variable", (L,), 100),
                                    ("path_angles_variable", (L, N), 100 + L),
                                    ("los_phases_variable", (L,), 100 + L + L * N),
                                    ("path_phases_variable", (L, N), 100 + 2 * L + L * N)]:
            assert f.get_str(f"{r}/{name}", "type") == "hermespy.channel.consistent.ConsistentUniform"
            assert tuple(f.read(f"{r}/{name}/shape")) == shape
            assert int(f.get_attr(f"{r}/{name}", "offset", np.int64)[0]) == offset
        assert np.allclose(f.read(f"{r}/power_profile"), cfg.power_profile)
        assert float(f.get_attr(r, "nlos_doppler")[0]) == cfg.doppler
        assert np.allclose(f.read(f"{r}/antenna_correlation/covariance_matrix"), cfg.corr_rx)
    # one realisation read on its own
    with record.H5File(path, record.APPEND) as f:
        r = "/c1/drops/0002/channel_realizations/0000"
        assert np.array_equal(record.read_realization(f, r, cfg), z[2])
    # the mirror's sample of the replayed scalars is the original realisation's
    ch = MultipathFadingChannel(cfg.delays, cfg.power_profile, cfg.rice_factors, doppler_frequency=cfg.doppler,
                                antenna_correlation=CustomAntennaCorrelation(cfg.corr_rx), seed=1)
    tx, rx = SimulatedDevice(4, sampling_rate=cfg.bandwidth), SimulatedDevice(4, sampling_rate=cfg.bandwidth)
    a = MultipathFadingRealization(z[1], ch).sample(tx, rx, device="cpu")
    b = MultipathFadingRealization(back[1], ch).sample(tx, rx, device="cpu")
    assert np.array_equal(a.spatial_response, b.spatial_response) and np.array_equal(a.nlos_angles, b.nlos_angles)
    # a link with other profiles is refused
    other = CONFIGS["siso-64"]()
    with pytest.raises(ValueError):
        record.read_reference_drops(path, other, campaign="c1")


def test_reference_realization_offsets_are_honoured(tmp_path):
    """A realisation whose scalars list the variables in another order (offsets permuted) reads
    back in this library's order."""

    cfg = CONFIGS["siso-64"]()
    rng = np.random.default_rng(2)
    z = rng.standard_normal(cfg.num_normals)
    path = tmp_path / "perm.h5"
    with record.H5File(path, record.CREATE) as f:
        record.write_realization(f, "/r", z, cfg)
    # rewrite: LOS angles stored last, everything after them moved forward
    vars_ = record.realization_variables(cfg.num_paths, cfg.num_sinusoids)
    parts = {name: z[off:off + int(np.prod(shape))] for name, shape, off in vars_}
    order = ["antenna_correlation_variable", "path_angles_variable", "los_phases_variable", "path_phases_variable",
             "los_angles_variable"]
    scalars = np.concatenate([parts[n] for n in order])
    offsets = np.concatenate([[0], np.cumsum([parts[n].size for n in order])[:-1]])
    path2 = tmp_path / "perm2.h5"
    with record.H5File(path2, record.CREATE) as f:
        f.create_group("/r")
        f.create_group("/r/random_realization")
        f.append("/r/random_realization/scalar_samples", scalars)
        f.set_attr("/r/random_realization", "type", record.STATIC_REALIZATION_TAG)
        for (name, shape, _), off in zip(sorted(vars_, key=lambda v: order.index(v[0])), offsets):
            f.create_group(f"/r/{name}")
            f.append(f"/r/{name}/shape", np.asarray(shape, dtype=np.int64))
            f.set_attr(f"/r/{name}", "offset", int(off))
            f.set_attr(f"/r/{name}", "type", record.UNIFORM_TAG)
        for name, v in (("power_profile", cfg.power_profile), ("delay_profile", cfg.delays),
                        ("los_gains", cfg.los_gains), ("nlos_gains", cfg.nlos_gains)):
            f.append(f"/r/{name}", np.asarray(v, dtype=np.float64))
        f.set_attr("/r", "type", record.REALIZATION_TAG)
        assert np.array_equal(record.read_realization(f, "/r", cfg), z)
