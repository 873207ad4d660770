"""Builds libswx.so (HIP, gfx950) in-tree from csrc/*.hip.

Usage: python seismicwaves.jl_amd/build.py [--debug]

--debug builds seismicwaves/libswx_debug.so instead: -O1 -g and SWX_DEBUG,
which turns on device-side asserts on the indices of the shared accessors
(point-table rows, C-PML coefficient vectors, the elastic field accessor).
Load it with SWX_LIBRARY=<path> (seismicwaves/_native.py).
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "seismicwaves", "libswx.so")
OUT_DEBUG = os.path.join(HERE, "seismicwaves", "libswx_debug.so")
ARCH = "gfx950"

# -ffp-contract=on fuses a*b+c into an FMA only inside one source
# expression.  Kernel variants that spell a cell update with the same
# expressions (interior/general tiles, one or two steps per launch,
# z-marching strips) then round identically whatever the surrounding control
# flow, so their results are bitwise equal.
HIPFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=on", "-Wall", "-Wno-unused-result",
    f"-I{os.path.join(REPO, 'include')}", f"-I{CSRC}",
]


def build(verbose: bool = False, jobs: int = 4, debug: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    objdir = os.path.join(HERE, "_build_debug" if debug else "_build")
    flags = [f if f != "-O3" else "-O1" for f in HIPFLAGS] + ["-g", "-DSWX_DEBUG"] if debug else HIPFLAGS
    out = OUT_DEBUG if debug else OUT
    os.makedirs(objdir, exist_ok=True)
    procs, objs = [], []
    for src in srcs:
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        objs.append(obj)
        cmd = [hipcc, *flags, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd))
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len(procs) >= jobs:
            _wait(procs.pop(0))
    for p in procs:
        _wait(p)
    tmp = out + ".tmp"
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link of {os.path.basename(out)} failed:\n" + r.stdout)
    os.replace(tmp, out)
    return out


PROBE_SRC = os.path.join(HERE, "probe", "hbm_triad.hip")
PROBE_OUT = os.path.join(HERE, "probe", "libswx_probe.so")


def build_probe(verbose: bool = False) -> str:
    """Builds the HBM bandwidth probe (probe/hbm_triad.hip) used by
    bench.py --hbm-triad; it is not part of libswx."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = PROBE_OUT + ".tmp"
    cmd = [hipcc, *HIPFLAGS, "-shared", "-o", tmp, PROBE_SRC]
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed on hbm_triad.hip:\n" + r.stdout)
    os.replace(tmp, PROBE_OUT)
    return PROBE_OUT


def _wait(item):
    src, p = item
    out, _ = p.communicate()
    if p.returncode != 0:
        raise RuntimeError(f"hipcc failed on {os.path.basename(src)}:\n{out.decode(errors='replace')}")
    if out:
        sys.stderr.write(out.decode(errors="replace"))


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, debug="--debug" in sys.argv))
    if "--debug" not in sys.argv:
        print(build_probe(verbose="-v" in sys.argv))
