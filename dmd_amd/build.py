"""Build dmd_amd/libdmdx.so (HIP, gfx950) and dmd_amd/libdmdx_h5md.so in place.

    python -m dmd_amd.build

Every .hip / .cpp file under dmd_amd/csrc is compiled by hipcc for gfx950
with contraction off (the FP64 kernels keep the reference's rounding) and
packed FP32 instructions disabled, then linked into one shared library.
dmd_amd/h5md/h5md.cpp (host code: H5MD trajectory files) is built by g++
against the HDF5 C library into libdmdx_h5md.so, which links libdmdx.so;
without HDF5 headers it is skipped with a message.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libdmdx.so")
OBJ = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
H5MD_OUT = os.path.join(HERE, "libdmdx_h5md.so")
HDF5_ROOT = os.environ.get("DMDX_HDF5_ROOT", "/opt/conda")
SYSTEM_LIBDIR = "/usr/lib/x86_64-linux-gnu"
ARCH = os.environ.get("DMDX_ARCH", "gfx950")

FLAGS = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-munsafe-fp-atomics", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith((".hip", ".cpp")))


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    lang = ["-x", "hip"]
    cmd = [HIPCC, *lang, *FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    build_h5md(verbose)
    return OUT


def build_h5md(verbose: bool = True):
    """libdmdx_h5md.so: g++ on the host-only writer, linked to HDF5 and to
    libdmdx.so (found next to it through $ORIGIN)."""
    inc = os.path.join(HDF5_ROOT, "include")
    lib = os.path.join(HDF5_ROOT, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        if verbose:
            print(f"skipped {H5MD_OUT}: no hdf5.h under {inc}")
        return None
    tmp = H5MD_OUT + ".tmp"
    cmd = ["g++", "-O2", "-fPIC", "-std=c++17", "-Wall", "-shared",
           "-I", os.path.join(os.path.dirname(HERE), "include"), "-I", inc,
           os.path.join(HERE, "h5md", "h5md.cpp"), "-o", tmp,
           "-L", lib, "-lhdf5", "-L", HERE, "-ldmdx",
           # the system C++ runtime before HDF5's directory, which holds an
           # older libstdc++ than the ROCm runtime needs
           f"-Wl,-rpath,$ORIGIN:{SYSTEM_LIBDIR}:{lib}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed on h5md.cpp:\n{r.stderr[-6000:]}")
    os.replace(tmp, H5MD_OUT)
    if verbose:
        print(f"built {H5MD_OUT}")
    return H5MD_OUT


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
