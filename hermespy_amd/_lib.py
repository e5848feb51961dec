"""ctypes binding of libhermespy_amd.so (the C ABI declared in include/hermespy_amd.h).

torch is imported first so that the library binds to the HIP runtime torch
already loaded (both carry the soname libamdhip64.so.7): device pointers and
streams from torch are then valid in the library.  There is no fallback: if
the library is missing, every entry point raises.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

import torch  # noqa: F401  (must precede the library load, see module docstring)

LIB_PATH = Path(__file__).resolve().parent / "_native" / "libhermespy_amd.so"

ABI_VERSION = 8

HPX_OK = 0
HPX_ERR_INVALID = -1
HPX_ERR_UNSUPPORTED = -2
HPX_ERR_DEVICE = -3

HPX_DETECT_ZF = 0
HPX_DETECT_MMSE = 1

STREAM_REALIZATION = 0x000
STREAM_BITS = 0x100
STREAM_NOISE = 0x200

ELEMENT_NULL = -2147483648

c_int = ctypes.c_int
c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_u32 = ctypes.c_uint32
c_u64 = ctypes.c_uint64
c_dbl = ctypes.c_double
c_vp = ctypes.c_void_p


class HpxLink(ctypes.Structure):
    """Mirror of ``hpx_link``."""

    _fields_ = [
        ("num_rx", c_i32), ("num_tx", c_i32),
        ("num_subcarriers", c_i32), ("words", c_i32), ("prefix", c_i32),
        ("modulation_order", c_i32),
        ("num_paths", c_i32), ("num_sinusoids", c_i32),
        ("max_delay", c_i32),
        ("detector", c_i32),
        ("dc_suppression", c_i32),
        ("bandwidth", c_dbl), ("doppler", c_dbl), ("los_doppler", c_dbl), ("gain", c_dbl),
        ("device_power", c_dbl),
        ("workgroup_size", c_i32), ("fft_values", c_i32),
    ]


class HpxFrame(ctypes.Structure):
    """Mirror of ``hpx_frame`` (device pointers)."""

    _fields_ = [("elements", c_vp), ("word_data", c_vp), ("pilots", c_vp), ("num_references", c_i32),
                ("num_data", c_i32), ("max_word_data", c_i32)]


class HpxBatchOut(ctypes.Structure):
    """Mirror of ``hpx_batch_out`` (device pointers)."""

    _fields_ = [("errors", c_vp), ("equalized", c_vp), ("grid", c_vp), ("csi", c_vp), ("spatial", c_vp)]


class HpxDropInputs(ctypes.Structure):
    """Mirror of ``hpx_drop_inputs`` (device pointers)."""

    _fields_ = [("noise_power", c_vp), ("doppler", c_vp), ("drop_index", c_vp), ("normals", c_vp)]


class HpxSectionStats(ctypes.Structure):
    """Mirror of ``hpx_section_stats`` (device pointers)."""

    _fields_ = [("sum", c_vp), ("sum_sq", c_vp), ("count", c_vp), ("errors", c_vp), ("bits", c_vp)]


class HpxDropTables(ctypes.Structure):
    """Mirror of ``hpx_drop_tables`` (device pointers)."""

    _fields_ = [("delays", c_vp), ("path_gains", c_vp), ("corr_rx", c_vp), ("corr_tx", c_vp), ("workspace", c_vp),
                ("equalized", c_vp)]


class HpxUnevenClass(ctypes.Structure):
    """Mirror of ``hpx_uneven_class``."""

    _fields_ = [("first", c_i64), ("count", c_i64), ("num_tx", c_i32), ("threads", c_i32), ("pairs", c_i32),
                ("lean", c_i32), ("lds_bytes", c_i64), ("fft_values", c_i32), ("reserved", c_i32)]


class HpxMassiveLink(ctypes.Structure):
    """Mirror of ``hpx_massive_link``."""

    _fields_ = [("num_rx", c_i32), ("num_users", c_i32), ("num_subcarriers", c_i32), ("active", c_i32),
                ("words", c_i32), ("prefix", c_i32), ("modulation_order", c_i32), ("num_paths", c_i32),
                ("num_sinusoids", c_i32), ("max_delay", c_i32), ("dc_suppression", c_i32), ("bandwidth", c_dbl),
                ("doppler", c_dbl), ("los_doppler", c_dbl), ("gain", c_dbl), ("device_power", c_dbl)]


class HpxMassiveWork(ctypes.Structure):
    """Mirror of ``hpx_massive_work`` (device pointers)."""

    _fields_ = [(name, c_vp) for name in ("spatial", "coef", "abar", "det_v", "det_a", "det_l", "z", "grid", "csi")]


# name -> (restype, argtypes)
SIGNATURES = {
    "hpx_abi_version": (c_int, []),
    "hpx_last_error": (c_int, [ctypes.c_char_p, ctypes.c_size_t]),
    "hpx_fading_taps": (c_int, [c_vp, c_int, c_int, c_i64, c_vp, c_vp]),
    "hpx_tdl_propagate": (c_int, [c_vp, c_int, c_i64, c_i64, c_vp, c_int, c_int, c_dbl, c_vp, c_i32, c_vp, c_int,
                                  c_vp, c_i64, c_vp]),
    "hpx_fading_state": (c_int, [c_vp, c_int, c_int, c_vp, c_i32, c_vp, c_int, c_int, c_i64, c_vp, c_vp]),
    "hpx_csi_propagate": (c_int, [c_vp, c_int, c_int, c_i64, c_i32, c_vp, c_i64, c_i64, c_vp, c_i64, c_vp]),
    "hpx_awgn": (c_int, [c_vp, c_int, c_i64, c_i64, c_u64, c_u32, c_dbl, c_vp]),
    "hpx_map_symbols": (c_int, [c_vp, c_i64, c_int, c_int, c_vp, c_vp]),
    "hpx_demap_hard": (c_int, [c_vp, c_i64, c_int, c_int, c_vp, c_vp]),
    "hpx_constellation_host": (c_int, [c_int, c_int, c_vp]),
    "hpx_demap_hard_host": (c_int, [c_vp, c_i64, c_int, c_int, c_vp]),
    "hpx_demap_hard_f32_host": (c_int, [c_vp, c_i64, c_int, c_vp]),
    "hpx_ofdm_modulate": (c_int, [c_vp, c_int, c_int, c_int, c_int, c_int, c_vp, c_i64, c_vp]),
    "hpx_ofdm_demodulate": (c_int, [c_vp, c_int, c_i64, c_int, c_int, c_int, c_int, c_vp, c_vp]),
    "hpx_ideal_csi": (c_int, [c_vp, c_int, c_int, c_vp, c_i32, c_vp, c_int, c_int, c_int, c_int, c_int, c_int, c_vp,
                              c_vp]),
    "hpx_equalize": (c_int, [c_vp, c_vp, c_int, c_int, c_i64, c_int, c_dbl, c_vp, c_vp]),
    "hpx_count_errors": (c_int, [c_vp, c_vp, c_i64, c_vp, c_vp]),
    "hpx_run_batch": (c_int, [ctypes.POINTER(HpxLink), c_vp, c_vp, c_vp, c_vp, c_u64, c_i64, c_i64, c_vp, c_vp, c_vp,
                              c_vp]),
    "hpx_run_batch_f64": (c_int, [ctypes.POINTER(HpxLink), c_vp, c_vp, c_u64, c_i64, c_i64, c_vp, c_vp, c_vp, c_vp]),
    "hpx_run_batch_frames": (c_int, [ctypes.POINTER(HpxLink), ctypes.POINTER(HpxFrame), c_vp, c_vp, c_vp, c_vp,
                                     c_vp, c_u64, c_i64, c_i64, c_vp, ctypes.POINTER(HpxBatchOut), c_vp]),
    "hpx_uneven_plan_bytes": (c_i64, []),
    "hpx_uneven_workspace_elements": (c_i64, [ctypes.POINTER(HpxLink)]),
    "hpx_uneven_prepare": (c_int, [ctypes.POINTER(HpxLink), ctypes.POINTER(HpxDropTables), c_vp, c_i64, c_u64, c_vp,
                                   c_vp, c_vp, c_vp, ctypes.POINTER(HpxUnevenClass), c_int, ctypes.POINTER(c_int)]),
    "hpx_uneven_run": (c_int, [c_vp, ctypes.POINTER(HpxUnevenClass), c_int, c_vp]),
    "hpx_uneven_prepare_noalloc": (c_int, [ctypes.POINTER(HpxLink), ctypes.POINTER(HpxDropTables), c_vp, c_i64, c_u64,
                                           c_vp, c_vp, c_vp, c_vp, ctypes.POINTER(HpxUnevenClass), c_int,
                                           ctypes.POINTER(c_int)]),
    "hpx_uneven_noalloc_capacity": (c_i64, []),
    "hpx_run_batch_lds_bytes": (c_i64, [ctypes.POINTER(HpxLink)]),
    "hpx_spatial_detector_host": (c_int, [c_int, c_int, c_vp, c_vp, c_vp, c_vp]),
    "hpx_stream_bits": (c_int, [c_u64, c_u32, c_u32, c_i64, c_vp, c_vp]),
    "hpx_realization_normals": (c_int, [c_u64, c_u32, c_i64, c_vp, c_vp]),
    "hpx_realization_normals_batch": (c_int, [c_u64, c_vp, c_i64, c_i64, c_i64, c_vp, c_vp]),
    "hpx_run_batch_drops": (c_int, [ctypes.POINTER(HpxLink), ctypes.POINTER(HpxFrame), c_vp, c_vp, c_vp, c_vp,
                                    ctypes.POINTER(HpxDropInputs), c_u64, c_i64, c_i64, ctypes.POINTER(HpxBatchOut),
                                    c_vp]),
    "hpx_section_accumulate": (c_int, [c_vp, c_vp, c_i64, c_vp, c_i64, c_i64, c_i32,
                                       ctypes.POINTER(HpxSectionStats), c_vp]),
    "hpx_pilot_ls": (c_int, [c_vp, c_i64, c_int, c_int, c_i64, c_int, c_vp, c_vp, c_vp, c_vp]),
    "hpx_lmmse_apply": (c_int, [c_vp, c_int, c_i64, c_int, c_vp, c_i64, c_vp, c_int, c_vp, c_vp]),
    "hpx_detect_estimated": (c_int, [c_vp, c_vp, c_i64, c_int, c_int, c_int, c_int, c_vp, c_vp, c_int, c_int, c_int,
                                     c_vp, c_u64, c_i64, c_vp, c_vp, c_vp]),
    "hpx_conv1d_forward": (c_int, [c_vp, c_i64, c_int, c_int, c_int, c_vp, c_vp, c_int, c_int, c_vp, c_vp, c_vp]),
    "hpx_conv1d_wgrad_workspace": (c_i64, [c_i64, c_int, c_int, c_int]),
    "hpx_conv1d_wgrad": (c_int, [c_vp, c_vp, c_i64, c_int, c_int, c_int, c_int, c_vp, c_vp, c_vp, c_vp]),
    "hpx_gelu_backward": (c_int, [c_vp, c_vp, c_i64, c_vp, c_vp]),
    "hpx_to_bf16": (c_int, [c_vp, c_i64, c_vp, c_vp]),
    "hpx_net_input": (c_int, [c_vp, c_i64, c_int, c_int, c_int, c_int, c_vp, c_vp, c_vp]),
    "hpx_net_output": (c_int, [c_vp, c_i64, c_int, c_int, c_vp, c_vp]),
    "hpx_estimate_error": (c_int, [c_vp, c_vp, c_vp, c_i64, c_int, c_int, c_i64, c_vp, c_vp, c_vp]),
    "hpx_nmse_grad": (c_int, [c_vp, c_vp, c_vp, c_i64, c_int, c_int, c_int, c_int, c_vp, c_int, c_vp, c_vp, c_vp,
                              c_vp]),
    "hpx_adam_step": (c_int, [c_vp, c_vp, c_vp, c_vp, c_i64, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                              ctypes.c_float, c_i64, c_vp]),
    "hpx_conv1d_weight_copies": (c_int, [c_vp, c_int, c_int, c_int, c_vp, c_vp, c_vp]),
    "hpx_learned_estimate": (c_int, [c_vp, c_i64, c_int, c_int, c_int, c_int, c_vp, c_vp, c_int, c_vp, c_vp, c_vp,
                                     c_int, c_vp, c_vp]),
    "hpx_massive_sizes": (c_int, [ctypes.POINTER(HpxMassiveLink), c_vp]),
    "hpx_massive_run": (c_int, [ctypes.POINTER(HpxMassiveLink), c_vp, c_vp, c_vp, c_vp, c_u64, c_i64, c_i64, c_vp,
                                ctypes.POINTER(HpxMassiveWork), c_vp, c_vp, c_vp]),
    "hpx_ls_estimate": (c_int, [c_vp, c_int, c_i64, c_int, c_vp, c_vp, c_int, c_vp, c_vp, c_vp, c_int, c_vp, c_vp,
                                c_vp, c_vp]),
}

_lib: ctypes.CDLL | None = None


def load() -> ctypes.CDLL:
    """Load the library (once).  Raises RuntimeError when it was not built."""

    global _lib
    if _lib is not None:
        return _lib
    path = Path(os.environ.get("HPX_LIBRARY", str(LIB_PATH)))
    if not path.exists():
        raise RuntimeError(f"libhermespy_amd.so not found at {path}: run `python -m hermespy_amd.build` first")
    lib = ctypes.CDLL(str(path))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.hpx_abi_version() != ABI_VERSION:
        raise RuntimeError("libhermespy_amd ABI version mismatch")
    _lib = lib
    return lib


def last_error() -> str:
    buf = ctypes.create_string_buffer(512)
    load().hpx_last_error(buf, 512)
    return buf.value.decode(errors="replace")


def check(status: int) -> None:
    """Raise the reference's exception class for a failed call."""

    if status == HPX_OK:
        return
    msg = last_error()
    if status in (HPX_ERR_INVALID, HPX_ERR_UNSUPPORTED):
        raise ValueError(msg)
    raise RuntimeError(msg)


def check_count(value: int) -> int:
    """A count returned by the library, or its error (negative counts) raised as ValueError."""

    if value < 0:
        raise ValueError(last_error())
    return int(value)


def ptr(t, name: str = "tensor") -> int | None:
    """Device address of a tensor for an array argument of the library, None for None.

    Every such argument is a device pointer: a host tensor would hand the kernels a host
    address, so it is rejected (ValueError) before any launch, as is a non-contiguous view."""

    if t is None:
        return None
    require_device(t, name)
    return t.data_ptr()


def stream_handle(device=None) -> int:
    """Raw hipStream_t of torch's current stream."""

    return torch.cuda.current_stream(device).cuda_stream


def require_device(t, name: str) -> None:
    if not t.is_cuda:
        raise ValueError(f"{name} must be a device tensor")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
