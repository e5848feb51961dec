"""Drop record / replay in HDF5 files (``libhermespy_amd_h5.so``, host C++ over libhdf5).

The reference records every drop of a scenario through its HDF5 serialization backend
(``Scenario.record / replay / Replay / stop``, ``hermespy/core/scenario.py:368-566``;
``HDFSerializationProcess``, ``hermespy/core/factory.py:838-1240``).  The store here keeps
that backend's conventions -- root attributes ``version`` and ``timestamp``, an optional
campaign group holding an ``arrays`` group, a ``state`` group tagged with a ``type``
attribute, scalars as attributes, complex data as the compound ``{r, i}`` -- and lays the
drops out column-wise: one dataset per quantity with one row per drop, appended a launch at
a time.  Layout of a campaign group::

    state/                 attrs: type, every scalar LinkConfig field, seed
      delays, power_profile, rice_factors[, corr_rx, corr_tx, frame_mask, pilots]
    drops/                 attrs: count
      drop_index   uint32 [n]   Philox drop counter (bits, noise, realisation)
      noise_power  float64 [n]
      errors       int32 [n]    bit errors of the drop as simulated when recorded
      [section     int32 [n]]   grid section (campaigns)
      [doppler     float64 [n, 2]]
      [channel_realizations/random_realization/scalar_samples  float64 [n, num_normals]]
      [grid complex64 [n, Nrx, words, N], csi complex64 [n, words, N], spatial complex64 [n, Nrx, Ntx]]

``scalar_samples`` carries the name and the meaning of the reference's
``StaticConsistentRealization`` record (``hermespy/channel/consistent.py:244-260``, written
under a ``MultipathFadingRealization`` as ``random_realization``, ``fading.py:540-556``):
the standard normals a channel realisation is drawn from.  Replaying them through
``hpx_run_batch_drops`` rebuilds the very same channel.
"""

from __future__ import annotations

import ctypes
import dataclasses
import os
import time
from pathlib import Path

import numpy as np
import torch

from .configs import LinkConfig

H5_LIB_PATH = Path(__file__).resolve().parent / "_native" / "libhermespy_amd_h5.so"
FORMAT_VERSION = "hermespy_amd-drops-1"

U8, I8, I32, U32, I64, F32, F64, C64, C128, STR = range(10)
READ, CREATE, TRUNCATE, APPEND = range(4)
_NP = {U8: np.uint8, I8: np.int8, I32: np.int32, U32: np.uint32, I64: np.int64, F32: np.float32, F64: np.float64,
       C64: np.complex64, C128: np.complex128}
_CODE = {np.dtype(v): k for k, v in _NP.items()}

_h5: ctypes.CDLL | None = None


def load_h5() -> ctypes.CDLL:
    global _h5
    if _h5 is not None:
        return _h5
    path = Path(os.environ.get("HPX_H5_LIBRARY", str(H5_LIB_PATH)))
    if not path.exists():
        raise RuntimeError(f"{path} is missing: build it with `python -m hermespy_amd.build` (needs libhdf5)")
    lib = ctypes.CDLL(str(path))
    c_vp, c_i64, c_int, c_str = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_char_p
    sig = {
        "hpx_h5_last_error": [c_str, ctypes.c_size_t],
        "hpx_h5_open": [c_str, c_int, ctypes.POINTER(c_i64)],
        "hpx_h5_close": [c_i64],
        "hpx_h5_create_group": [c_i64, c_str],
        "hpx_h5_exists": [c_i64, c_str],
        "hpx_h5_set_attr": [c_i64, c_str, c_str, c_int, c_i64, c_vp],
        "hpx_h5_get_attr": [c_i64, c_str, c_str, c_int, c_i64, c_vp, ctypes.POINTER(c_i64)],
        "hpx_h5_append": [c_i64, c_str, c_int, c_int, c_vp, c_i64, c_vp],
        "hpx_h5_shape": [c_i64, c_str, ctypes.POINTER(c_int), ctypes.POINTER(c_int), c_vp],
        "hpx_h5_read": [c_i64, c_str, c_i64, c_i64, c_int, c_vp],
        "hpx_h5_list": [c_i64, c_str, c_vp, c_i64, ctypes.POINTER(c_i64)],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = c_int
    _h5 = lib
    return lib


def _check(status: int) -> None:
    if status >= 0:
        return
    buf = ctypes.create_string_buffer(512)
    load_h5().hpx_h5_last_error(buf, 512)
    msg = buf.value.decode(errors="replace")
    raise ValueError(msg) if status == -1 else OSError(msg)


def _b(s: str) -> bytes:
    return s.encode()


class H5File:
    """A file of the drop store (context manager)."""

    def __init__(self, path: str | os.PathLike, mode: int = READ) -> None:
        self.path = str(path)
        h = ctypes.c_int64(-1)
        _check(load_h5().hpx_h5_open(_b(self.path), mode, ctypes.byref(h)))
        self.handle = h.value

    def close(self) -> None:
        if self.handle >= 0:
            _check(load_h5().hpx_h5_close(self.handle))
            self.handle = -1

    def __enter__(self) -> "H5File":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def create_group(self, path: str) -> None:
        _check(load_h5().hpx_h5_create_group(self.handle, _b(path)))

    def exists(self, path: str) -> bool:
        r = load_h5().hpx_h5_exists(self.handle, _b(path))
        _check(r)
        return r == 1

    def set_attr(self, obj: str, name: str, value) -> None:
        lib = load_h5()
        if isinstance(value, str):
            _check(lib.hpx_h5_set_attr(self.handle, _b(obj), _b(name), STR, 1, ctypes.c_char_p(_b(value))))
            return
        if isinstance(value, (bool, int, np.integer)):
            arr = np.asarray([int(value)], dtype=np.int64)
        elif isinstance(value, (float, np.floating)):
            arr = np.asarray([float(value)], dtype=np.float64)
        else:
            arr = np.ascontiguousarray(value)
            if arr.dtype.kind == "f":
                arr = arr.astype(np.float64)
            elif arr.dtype.kind in "iub":
                arr = arr.astype(np.int64)
            arr = arr.reshape(-1)
        _check(lib.hpx_h5_set_attr(self.handle, _b(obj), _b(name), _CODE[arr.dtype], arr.size, arr.ctypes.data))

    def get_str(self, obj: str, name: str) -> str:
        buf = ctypes.create_string_buffer(4096)
        n = ctypes.c_int64(0)
        _check(load_h5().hpx_h5_get_attr(self.handle, _b(obj), _b(name), STR, 4096, buf, ctypes.byref(n)))
        return buf.value.decode()

    def get_attr(self, obj: str, name: str, dtype=np.float64) -> np.ndarray:
        out = np.zeros(4096, dtype=dtype)
        n = ctypes.c_int64(0)
        _check(load_h5().hpx_h5_get_attr(self.handle, _b(obj), _b(name), _CODE[np.dtype(dtype)], out.size,
                                          out.ctypes.data, ctypes.byref(n)))
        return out[:n.value]

    def append(self, path: str, data) -> None:
        """Appends rows (first axis) to the dataset at path, creating it on first use."""

        if isinstance(data, torch.Tensor):
            data = data.detach().cpu().numpy()
        arr = np.ascontiguousarray(data)
        if arr.ndim == 0:
            arr = arr.reshape(1)
        code = _CODE.get(arr.dtype)
        if code is None:
            raise ValueError(f"unsupported element type {arr.dtype}")
        shape = np.asarray(arr.shape[1:], dtype=np.int64)
        _check(load_h5().hpx_h5_append(self.handle, _b(path), code, shape.size, shape.ctypes.data if shape.size else None,
                                       arr.shape[0], arr.ctypes.data))

    def shape(self, path: str) -> tuple:
        dt, rank = ctypes.c_int(0), ctypes.c_int(0)
        dims = np.zeros(8, dtype=np.int64)
        _check(load_h5().hpx_h5_shape(self.handle, _b(path), ctypes.byref(dt), ctypes.byref(rank), dims.ctypes.data))
        return _NP.get(dt.value), tuple(int(d) for d in dims[:rank.value])

    def read(self, path: str, row0: int = 0, rows: int | None = None, dtype=None) -> np.ndarray:
        stored, shape = self.shape(path)
        dtype = np.dtype(stored if dtype is None else dtype)
        if not shape:
            out = np.zeros(1, dtype=dtype)
            _check(load_h5().hpx_h5_read(self.handle, _b(path), 0, 1, _CODE[dtype], out.ctypes.data))
            return out[0]
        rows = shape[0] - row0 if rows is None else rows
        out = np.empty((rows,) + shape[1:], dtype=dtype)
        _check(load_h5().hpx_h5_read(self.handle, _b(path), row0, rows, _CODE[dtype], out.ctypes.data))
        return out

    def list(self, path: str) -> list:
        size = ctypes.c_int64(0)
        lib = load_h5()
        lib.hpx_h5_list(self.handle, _b(path), None, 0, ctypes.byref(size))
        if size.value <= 0:
            _check(-4)
        buf = ctypes.create_string_buffer(int(size.value))
        _check(lib.hpx_h5_list(self.handle, _b(path), buf, size.value, ctypes.byref(size)))
        return [n for n in buf.value.decode().split("\n") if n]


# ---------------------------------------------------------------- link state
_ARRAYS = ("delays", "power_profile", "rice_factors", "corr_rx", "corr_tx", "frame_mask", "pilots")


def write_state(f: H5File, group: str, link: LinkConfig, seed: int) -> None:
    """The link as the campaign's ``state`` (the reference serializes its scenario there,
    ``scenario.py:425-429``), tagged with a ``type`` attribute (``factory.py:940-943``)."""

    g = f"{group}/state"
    f.create_group(g)
    for fld in dataclasses.fields(link):
        v = getattr(link, fld.name)
        if fld.name in _ARRAYS:
            if v is not None:
                f.append(f"{g}/{fld.name}", np.asarray(v, dtype=np.int8 if fld.name == "frame_mask" else None))
        elif v is None:
            continue
        elif isinstance(v, tuple):
            f.set_attr(g, fld.name, np.asarray(v, dtype=np.float64))
        else:
            f.set_attr(g, fld.name, v)
    f.set_attr(g, "seed", int(np.uint64(seed).astype(np.int64)))
    f.set_attr(g, "type", "hermespy_amd.configs.LinkConfig")


def read_state(f: H5File, group: str) -> tuple:
    """(LinkConfig, seed) of a campaign written by :func:`write_state`."""

    g = f"{group}/state"
    if not f.exists(g) or f.get_str(g, "type") != "hermespy_amd.configs.LinkConfig":
        raise ValueError(f"no link state under '{g}'")
    kw = {}
    for fld in dataclasses.fields(LinkConfig):
        name = fld.name
        if name in _ARRAYS:
            if f.exists(f"{g}/{name}"):
                kw[name] = f.read(f"{g}/{name}")
            continue
        try:
            if name in _STR_FIELDS:
                kw[name] = f.get_str(g, name)
            elif name in _TUPLE_FIELDS:
                kw[name] = tuple(float(x) for x in f.get_attr(g, name))
            elif name in _INT_FIELDS:
                kw[name] = int(f.get_attr(g, name, np.int64)[0])
            elif name == "dc_suppression":
                kw[name] = bool(f.get_attr(g, name, np.int64)[0])
            else:
                kw[name] = float(f.get_attr(g, name)[0])
        except OSError:
            continue  # an optional field that was None
    seed = int(np.int64(f.get_attr(g, "seed", np.int64)[0]).astype(np.uint64))
    return LinkConfig(**kw), seed


_STR_FIELDS = ("detector", "name")
_TUPLE_FIELDS = ("snr_db", "doppler_hz_range")
_INT_FIELDS = ("num_rx", "num_tx", "num_subcarriers", "words", "modulation_order", "num_sinusoids", "workgroup_size")


# ---------------------------------------------------------------- recorder / replay
class DropRecorder:
    """Writes launches of drops to a campaign (``Scenario.record``, ``scenario.py:368-439``)."""

    def __init__(self, file: str | os.PathLike, link: LinkConfig, seed: int, campaign: str | None = None,
                 overwrite: bool = False, normals: bool = False, exports: tuple = ()) -> None:
        path = Path(file)
        exists = path.exists()
        if overwrite and exists:
            path.unlink()
            exists = False
        self.file = H5File(path, APPEND if exists else CREATE)
        self.group = f"/{campaign}" if campaign else ""
        self.link, self.seed = link, int(seed)
        self.normals = bool(normals)
        self.exports = tuple(exports)
        if not exists:
            self.file.set_attr("/", "version", FORMAT_VERSION)
            self.file.set_attr("/", "timestamp", time.time())
        if self.group:
            self.file.create_group(self.group)
        self.file.create_group(f"{self.group}/arrays")
        if self.file.exists(f"{self.group}/state"):
            stored, stored_seed = read_state(self.file, self.group or "")
            if stored_seed != self.seed or not _same_link(stored, link):
                raise ValueError("the campaign holds drops of another link or seed")
        else:
            write_state(self.file, self.group, link, self.seed)
        self.drops = f"{self.group}/drops"
        self.file.create_group(self.drops)
        self.count = num_drops(self.file, self.drops)
        self.file.set_attr(self.drops, "count", self.count)

    def record(self, drop_index, noise_power, errors, section=None, doppler=None, normals=None, **exports) -> None:
        """Appends one launch: host or device arrays with one entry (row) per drop."""

        def host(x, dtype=None):
            if isinstance(x, torch.Tensor):
                x = x.detach().cpu().numpy()
            x = np.asarray(x)
            return x if dtype is None else x.astype(dtype, copy=False)

        ids = host(drop_index)
        ids = ids.view(np.uint32) if ids.dtype == np.int32 else ids.astype(np.uint32)
        n = ids.size
        f, d = self.file, self.drops
        f.append(f"{d}/drop_index", ids)
        f.append(f"{d}/noise_power", host(noise_power, np.float64)[:n])
        f.append(f"{d}/errors", host(errors, np.int32)[:n])
        if section is not None:
            f.append(f"{d}/section", host(section, np.int32)[:n])
        if doppler is not None:
            f.append(f"{d}/doppler", host(doppler, np.float64).reshape(-1, 2)[:n])
        if normals is not None:
            z = host(normals, np.float64).reshape(n, -1)
            f.append(f"{d}/channel_realizations/random_realization/scalar_samples", z)
        for name, value in exports.items():
            if value is not None:
                f.append(f"{d}/{name}", host(value)[:n])
        self.count += n
        f.set_attr(d, "count", self.count)

    def close(self) -> None:
        self.file.close()


def num_drops(f: H5File, drops: str) -> int:
    return f.shape(f"{drops}/drop_index")[1][0] if f.exists(f"{drops}/drop_index") else 0


def _same_link(a: LinkConfig, b: LinkConfig) -> bool:
    for fld in dataclasses.fields(LinkConfig):
        if fld.name in ("name", "snr_db", "workgroup_size"):
            continue
        x, y = getattr(a, fld.name), getattr(b, fld.name)
        if (x is None) != (y is None):
            return False
        if x is not None and not np.array_equal(np.asarray(x), np.asarray(y)):
            return False
    return True


class DropReplay:
    """Reads a recorded campaign back (``Scenario.replay``, ``scenario.py:441-506``)."""

    def __init__(self, file: str | os.PathLike, campaign: str | None = None, drop_offset: int = 0) -> None:
        self.file = H5File(file, READ)
        self.group = f"/{campaign}" if campaign else ""
        self.drops = f"{self.group}/drops"
        if not self.file.exists(self.drops):
            self.file.close()
            raise ValueError(f"no recorded drops in campaign '{campaign}'")
        self.link, self.seed = read_state(self.file, self.group)
        self.available = num_drops(self.file, self.drops)
        if drop_offset >= self.available:
            self.file.close()
            raise ValueError("Drop offset is not smaller than the number of available drops")
        self.offset = int(drop_offset)

    @property
    def num_drops(self) -> int:
        return self.available - self.offset

    @property
    def has_normals(self) -> bool:
        return self.file.exists(f"{self.drops}/channel_realizations/random_realization/scalar_samples")

    def read(self, start: int, count: int) -> dict:
        """Recorded columns of drops offset + start .. offset + start + count - 1."""

        f, d = self.file, self.drops
        r0 = self.offset + start
        out = {"drop_index": f.read(f"{d}/drop_index", r0, count), "noise_power": f.read(f"{d}/noise_power", r0, count),
               "errors": f.read(f"{d}/errors", r0, count)}
        for name, path in (("section", "section"), ("doppler", "doppler"),
                           ("normals", "channel_realizations/random_realization/scalar_samples")):
            if f.exists(f"{d}/{path}"):
                out[name] = f.read(f"{d}/{path}", r0, count)
        return out

    def close(self) -> None:
        self.file.close()


# ---------------------------------------------------------------- reference realisation records
# Serialization tags (``Serializable.serialization_tag``: module + class name, factory.py:88-93)
REALIZATION_TAG = "hermespy.channel.fading.fading.MultipathFadingRealization"
STATIC_REALIZATION_TAG = "hermespy.channel.consistent.StaticConsistentRealization"
UNIFORM_TAG = "hermespy.channel.consistent.ConsistentUniform"
CUSTOM_CORRELATION_TAG = "hermespy.channel.fading.fading.CustomAntennaCorrelation"
DROP_TAG = "hermespy.simulation.drop.SimulatedDrop"


def realization_variables(num_paths: int, num_sinusoids: int) -> list:
    """(name, shape, offset) of the uniform variables a ``MultipathFadingChannel`` registers with
    its ``ConsistentGenerator``, in registration order (``fading.py:737-754``; offsets accumulate,
    ``consistent.py:410-432``): the scalars of a realisation are their concatenation."""

    L, N = int(num_paths), int(num_sinusoids)
    out, offset = [], 0
    for name, shape in (("antenna_correlation_variable", (10, 10)), ("los_angles_variable", (L,)),
                        ("path_angles_variable", (L, N)), ("los_phases_variable", (L,)),
                        ("path_phases_variable", (L, N))):
        out.append((name, shape, offset))
        offset += int(np.prod(shape))
    return out


def write_realization(f: H5File, path: str, normals, link: LinkConfig, doppler=None) -> None:
    """A ``MultipathFadingRealization`` group at ``path`` in the reference's layout
    (``fading.py:540-556``): the ``StaticConsistentRealization`` of its scalars
    (``random_realization/scalar_samples``, ``consistent.py:254-255``), the five
    ``ConsistentUniform`` variables (``shape`` dataset, ``offset`` attribute, ``consistent.py:93-95``),
    the profiles and gains as datasets, the Doppler frequencies and ``gain`` as attributes and, for
    a custom antenna correlation, its ``covariance_matrix``; ``type`` tags are written after the
    members (``factory.py:940-943``).  doppler: (NLOS, LOS) of this realisation, default the link's."""

    z = np.ascontiguousarray(np.asarray(normals, dtype=np.float64).reshape(-1))
    if z.size != link.num_normals:
        raise ValueError(f"a realisation of this link has {link.num_normals} scalars, not {z.size}")
    nlos, los = (link.doppler, link.doppler if link.los_doppler is None else link.los_doppler) if doppler is None \
        else (float(doppler[0]), float(doppler[1]))
    f.create_group(path)
    f.create_group(f"{path}/random_realization")
    f.append(f"{path}/random_realization/scalar_samples", z)
    f.set_attr(f"{path}/random_realization", "type", STATIC_REALIZATION_TAG)
    for name, shape, offset in realization_variables(link.num_paths, link.num_sinusoids):
        g = f"{path}/{name}"
        f.create_group(g)
        f.append(f"{g}/shape", np.asarray(shape, dtype=np.int64))
        f.set_attr(g, "offset", int(offset))
        f.set_attr(g, "type", UNIFORM_TAG)
    f.append(f"{path}/power_profile", np.asarray(link.power_profile, dtype=np.float64))
    f.append(f"{path}/delay_profile", np.asarray(link.delays, dtype=np.float64))
    f.append(f"{path}/los_gains", np.asarray(link.los_gains, dtype=np.float64))
    f.append(f"{path}/nlos_gains", np.asarray(link.nlos_gains, dtype=np.float64))
    f.set_attr(path, "los_doppler", float(los))
    f.set_attr(path, "nlos_doppler", float(nlos))
    if link.corr_rx is not None:
        if link.corr_tx is None or not np.array_equal(np.asarray(link.corr_rx)[:link.num_tx, :link.num_tx],
                                                      np.asarray(link.corr_tx)):
            raise ValueError("the reference records one custom covariance for both ends of a link")
        g = f"{path}/antenna_correlation"
        f.create_group(g)
        f.append(f"{g}/covariance_matrix", np.asarray(link.corr_rx, dtype=np.complex128))
        f.set_attr(g, "type", CUSTOM_CORRELATION_TAG)
    f.set_attr(path, "gain", float(link.gain))
    f.set_attr(path, "type", REALIZATION_TAG)


def read_realization(f: H5File, path: str, link: LinkConfig) -> np.ndarray:
    """The scalars of the ``MultipathFadingRealization`` group at ``path`` in this library's order
    (the five variables concatenated), gathered by each variable's recorded ``offset`` and
    ``shape`` from ``random_realization/scalar_samples``; the recorded profiles and gains must be
    the link's (ValueError otherwise)."""

    if f.get_str(path, "type") != REALIZATION_TAG:
        raise ValueError(f"'{path}' is not a MultipathFadingRealization")
    if f.get_str(f"{path}/random_realization", "type") != STATIC_REALIZATION_TAG:
        raise ValueError("only static (spatially invariant) realisations replay on this path")
    scalars = np.asarray(f.read(f"{path}/random_realization/scalar_samples"), dtype=np.float64).reshape(-1)
    for name, expect in (("power_profile", link.power_profile), ("delay_profile", link.delays),
                         ("los_gains", link.los_gains), ("nlos_gains", link.nlos_gains)):
        got = np.asarray(f.read(f"{path}/{name}"), dtype=np.float64)
        if got.shape != np.shape(expect) or not np.allclose(got, expect, rtol=1e-12, atol=0.0):
            raise ValueError(f"recorded {name} differs from the link's")
    parts = []
    for name, shape, _ in realization_variables(link.num_paths, link.num_sinusoids):
        g = f"{path}/{name}"
        rec_shape = tuple(int(v) for v in f.read(f"{g}/shape", dtype=np.int64))
        offset = int(f.get_attr(g, "offset", np.int64)[0])
        if rec_shape != tuple(shape):
            raise ValueError(f"recorded {name} has shape {rec_shape}, the link needs {tuple(shape)}")
        count = int(np.prod(shape))
        if offset < 0 or offset + count > scalars.size:
            raise ValueError(f"recorded {name} lies outside the realisation's scalars")
        parts.append(scalars[offset:offset + count])
    return np.concatenate(parts)


def _names(f: H5File, path: str) -> list:
    """Members of a group (none for an empty or missing group)."""

    if not f.exists(path):
        return []
    try:
        return f.list(path)
    except OSError:
        return []


def write_reference_drops(file: str | os.PathLike, link: LinkConfig, normals, campaign: str | None = None,
                          overwrite: bool = False, doppler=None, timestamps=None) -> None:
    """Drops in the reference's recording layout (``Scenario.record`` + ``SimulatedDrop.serialize``,
    ``scenario.py:368-439,840-851``, ``simulation/drop.py:54-56``, ``core/drop.py:92-95``):
    ``drops/0000, 0001, ...`` tagged ``SimulatedDrop`` with a ``timestamp`` attribute, empty
    ``transmissions`` / ``receptions`` sequences (signals are re-simulated here, not stored) and
    the link's realisation as ``channel_realizations/0000``.  normals: (drops, num_normals)."""

    z = np.asarray(normals, dtype=np.float64).reshape(-1, link.num_normals)
    mode = TRUNCATE if overwrite else (APPEND if Path(file).exists() else CREATE)
    with H5File(file, mode) as f:
        f.set_attr("/", "version", FORMAT_VERSION)
        f.set_attr("/", "timestamp", time.time())
        group = "/" if campaign is None else f"/{campaign}"
        if campaign is not None:
            f.create_group(group)
        base = group.rstrip("/")
        f.create_group(f"{base}/arrays")
        f.create_group(f"{base}/drops")
        first = len(_names(f, f"{base}/drops"))
        for k in range(z.shape[0]):
            d = f"{base}/drops/{first + k:04d}"
            f.create_group(d)
            f.set_attr(d, "timestamp", float(0.0 if timestamps is None else timestamps[k]))
            f.create_group(f"{d}/transmissions")
            f.create_group(f"{d}/receptions")
            f.create_group(f"{d}/channel_realizations")
            write_realization(f, f"{d}/channel_realizations/0000", z[k], link,
                              None if doppler is None else doppler[k])
            f.set_attr(d, "type", DROP_TAG)


def read_reference_drops(file: str | os.PathLike, link: LinkConfig, campaign: str | None = None) -> np.ndarray:
    """(drops, num_normals) realisation scalars of the drops of a reference-layout recording, in
    drop order (``Scenario.replay`` reads ``drops`` as a sequence, ``scenario.py:486-506``)."""

    with H5File(file, READ) as f:
        base = "" if campaign is None else f"/{campaign}"
        names = sorted(_names(f, f"{base}/drops"), key=int)
        out = np.empty((len(names), link.num_normals), dtype=np.float64)
        for k, name in enumerate(names):
            d = f"{base}/drops/{name}"
            if f.get_str(d, "type") != DROP_TAG:
                raise ValueError(f"'{d}' is not a SimulatedDrop")
            out[k] = read_realization(f, f"{d}/channel_realizations/0000", link)
    return out
