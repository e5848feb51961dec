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
