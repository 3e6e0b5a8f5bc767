"""Build recipe of libnwcch.so (HIP for gfx950 + host C++), in-tree.

``python -m nwcch._build`` (from ``negative-weight-cch_amd/``) or
``__graft_entry__.build()`` compiles it.  Objects go to ``csrc/_build/``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = ROOT / "csrc"
LIB = PKG / "libnwcch.so"
ARCH = os.environ.get("NWCCH_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # keep packed-FP32 instructions out of the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-march=x86-64-v2"]

SOURCES_HIP = ["device.hip", "update.hip", "batch.hip", "symbolic.hip", "query_blk.hip", "rows.hip"]
SOURCES_CPP = ["capi.cpp", "host_prep.cpp", "contract_host.cpp"]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and Path(c).exists():
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    return r


def build(verbose: bool = False) -> Path:
    out = CSRC / "_build"
    out.mkdir(exist_ok=True)
    objs = []
    hc = hipcc()
    for src in SOURCES_CPP:
        o = out / (src + ".o")
        cmd = ["g++", *HOST_FLAGS, "-c", str(CSRC / src), "-o", str(o)]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(str(o))
    for src in SOURCES_HIP:
        if not (CSRC / src).exists():
            continue
        o = out / (src + ".o")
        cmd = [hc, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(o)]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(str(o))
    # the synthetic-graph generator's host helpers: their own library, no HIP
    gen = PKG / "libnwgen.so"
    gtmp = gen.with_suffix(".so.tmp")
    cmd = ["g++", *HOST_FLAGS, "-shared", str(CSRC / "gen_host.cpp"), "-o", str(gtmp), "-pthread"]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(gtmp, gen)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
