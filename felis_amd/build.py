"""Builds libfelis_hip.so (gfx950) in-tree with hipcc.

Usage: python -m felis_amd.build   (or felis_amd.build.build() from Python)
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libfelis_hip.so"

HIP_SOURCES = ["felis_api.hip", "felis_run.hip", "felis_load.hip", "felis_deck.cpp", "felis_mlp.hip", "felis_sur.hip",
               "felis_diag.hip", "felis_wake.hip", "felis_fno.hip"]

# -packed-fp32-ops off: packed FP32 forms are kept out of generated code.
HIPFLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-munsafe-fp-atomics",
    "-ffp-contract=fast",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (Path(cand).exists() or cand == "hipcc"):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> Path:
    """Compiles every source to an object in parallel, then links the library."""
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    inc = [f"-I{ROOT / 'include'}", f"-I{CSRC}"]
    flags = [f for f in HIPFLAGS if f != "-shared"]
    procs = []
    objs = []
    for s in HIP_SOURCES:
        obj = objdir / (Path(s).stem + ".o")
        cmd = [hipcc(), *flags, *inc, "-c", str(CSRC / s), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(str(obj))
    failed = []
    for s, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            failed.append(s)
    if failed:
        raise RuntimeError(f"hipcc failed on {', '.join(failed)}")
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"link failed ({res.returncode})")
    os.replace(tmp, LIB)
    return LIB


# HDF5 dump/restore library: host code on the image's serial HDF5
HDF5_DIR = Path(os.environ.get("FELIS_HDF5_DIR", "/opt/conda"))
HDF_LIB = PKG / "libfelis_hdf.so"


def build_hdf(verbose: bool = False):
    """Builds libfelis_hdf.so (g++, linked to libfelis_hip.so and HDF5) when
    the HDF5 headers are present; returns its path or None. libstdc++ and
    libgcc are linked statically so that the HDF5 directory on the runpath
    supplies HDF5 only."""
    if not (HDF5_DIR / "include" / "hdf5.h").exists():
        if verbose:
            print(f"no HDF5 under {HDF5_DIR}: libfelis_hdf.so not built", flush=True)
        return None
    tmp = HDF_LIB.with_suffix(".so.tmp")
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", str(CSRC / "felis_hdf.cpp"),
           f"-I{ROOT / 'include'}", "-isystem", str(HDF5_DIR / "include"),
           f"-L{PKG}", "-lfelis_hip", f"-L{HDF5_DIR / 'lib'}", "-lhdf5",
           "-static-libstdc++", "-static-libgcc",
           f"-Wl,-rpath,$ORIGIN:{HDF5_DIR / 'lib'}", "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"g++ failed on felis_hdf.cpp ({res.returncode})")
    os.replace(tmp, HDF_LIB)
    return HDF_LIB


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_hdf(verbose=True))
