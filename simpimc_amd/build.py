"""Build recipe of the engine library (hipcc, gfx950, in-tree)."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "spimc_engine.hip")
OUT = os.path.join(HERE, "libspimc.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed FP32 instructions are disabled for all device code.
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]


HOST_SRC = os.path.join(HERE, "host", "spimc_main.cpp")
HOST_OUT = os.path.join(HERE, "spimc")


def hdf5_flags() -> list[str]:
    """HDF5 support of the driver: the header at build time; the library is
    loaded at run time from the directory beside it (or $SPIMC_HDF5_LIB)."""
    for root in (os.environ.get("SPIMC_HDF5_ROOT", ""), "/opt/conda", "/usr"):
        if root and os.path.exists(os.path.join(root, "include", "hdf5.h")):
            return ["-DSPIMC_HAVE_HDF5=1", f"-I{root}/include", f'-DSPIMC_HDF5_LIB_DIR="{root}/lib"']
    return ["-DSPIMC_HAVE_HDF5=0"]


def build_host() -> str:
    """The XML driver `spimc` (host C++ over the C ABI), linked to libspimc.so."""
    tmp = HOST_OUT + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-Wno-unused-parameter", *hdf5_flags(), "-o", tmp, HOST_SRC,
           f"-L{HERE}", "-lspimc", "-L/opt/rocm/lib", "-lamdhip64", "-ldl", "-Wl,-rpath,$ORIGIN",
           "-Wl,-rpath,/opt/rocm/lib"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("host driver build failed:\n" + res.stderr[-4000:])
    os.replace(tmp, HOST_OUT)
    return HOST_OUT


def build_engine(verbose: bool = False) -> str:
    tmp = OUT + ".tmp"
    cmd = [HIPCC, *FLAGS, "-o", tmp, SRC]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("engine build failed:\n" + res.stderr[-4000:])
    if verbose and res.stderr:
        print(res.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build_engine(verbose=True))
    print(build_host())
