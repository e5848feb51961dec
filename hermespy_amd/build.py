"""Build libhermespy_amd.so (gfx950) from hermespy_amd/csrc with hipcc.

Usage: ``python -m hermespy_amd.build`` (or ``__graft_entry__.build()``).
The library is written next to the package, in ``hermespy_amd/_native/``, so
that it ships with the source tree.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
ROOT = PACKAGE.parent
CSRC = PACKAGE / "csrc"
INCLUDE = ROOT / "include"
OUT_DIR = PACKAGE / "_native"
LIB_NAME = "libhermespy_amd.so"
ARCH = os.environ.get("HPX_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    # packed FP32 instructions stay disabled
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found: the HIP toolchain is required to build libhermespy_amd")
    return path


def build(verbose: bool = False) -> Path:
    sources = sorted(CSRC.glob("*.hip"))
    obj_dir = ROOT / "_build"
    obj_dir.mkdir(exist_ok=True)
    OUT_DIR.mkdir(exist_ok=True)
    cc = hipcc()

    def compile_one(src: Path) -> Path:
        obj = obj_dir / (src.stem + ".o")
        cmd = [cc, *HIPCC_FLAGS, "-c", str(src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src.name}:\n{res.stdout}\n{res.stderr}")
        return obj

    with ThreadPoolExecutor(max_workers=min(8, len(sources))) as pool:
        objects = list(pool.map(compile_one, sources))
    lib = OUT_DIR / LIB_NAME
    tmp = OUT_DIR / (LIB_NAME + ".tmp")
    cmd = [cc, "-shared", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objects)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, lib)
    build_h5(verbose)
    return lib


HDF5_ROOT = Path(os.environ.get("HPX_HDF5_ROOT", "/opt/conda"))
H5_LIB_NAME = "libhermespy_amd_h5.so"


def build_h5(verbose: bool = False) -> Path | None:
    """The drop record/replay store (host C++ over libhdf5); skipped when no HDF5 is installed."""

    if not (HDF5_ROOT / "include" / "hdf5.h").exists():
        if verbose:
            print(f"no HDF5 under {HDF5_ROOT}: {H5_LIB_NAME} not built", flush=True)
        return None
    out = OUT_DIR / H5_LIB_NAME
    tmp = OUT_DIR / (H5_LIB_NAME + ".tmp")
    cmd = [shutil.which("g++") or "g++", "-O2", "-std=c++17", "-fPIC", "-shared", f"-I{INCLUDE}",
           f"-I{HDF5_ROOT / 'include'}", str(CSRC / "record.cpp"), "-o", str(tmp), f"-L{HDF5_ROOT / 'lib'}",
           "-lhdf5", f"-Wl,-rpath,{HDF5_ROOT / 'lib'}"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"g++ failed on record.cpp:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
