"""Builds the HIP/CDNA4 kriging library ``krigekit_amd/libkrigekit_amd.so``.

Plain hipcc invocations for gfx950, in-tree, no build system:
  * every HIP translation unit is compiled with the packed-FP32 target
    feature disabled;
  * search.hip is compiled with ``-ffp-contract=off`` so squared distances
    round exactly like the reference build (no fused multiply-add);
  * capi.cpp holds the host side and the exported C ABI.

``python -m krigekit_amd.build`` rebuilds unconditionally.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libkrigekit_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-command-line-argument", "-fvisibility=hidden", "-I", CSRC]

# (source, extra flags, object name).  solve.hip is compiled five times: the
# shared kernels and dispatch, then the k_solve instances of each coordinate
# count (1-3 spatial, 4 space-time) in a unit of their own, in parallel;
# cokrige.hip likewise, its k_solve_mv instances split by coordinate count
# and weights mode.
UNITS = [
    ("search.hip", ["-ffp-contract=off"], "search.o"),
    ("solve.hip", [], "solve.o"),
    ("solve.hip", ["-DKK_SOLVE_DIM=1"], "solve_d1.o"),
    ("solve.hip", ["-DKK_SOLVE_DIM=2"], "solve_d2.o"),
    ("solve.hip", ["-DKK_SOLVE_DIM=3"], "solve_d3.o"),
    ("solve.hip", ["-DKK_SOLVE_DIM=4"], "solve_d4.o"),
    ("loograd.hip", [], "loograd.o"),
    ("cokrige.hip", [], "cokrige.o"),
    ("cokrige.hip", ["-DKK_MV_DIM=1"], "cokrige_d1.o"),
    ("cokrige.hip", ["-DKK_MV_DIM=2"], "cokrige_d2.o"),
    ("cokrige.hip", ["-DKK_MV_DIM=3"], "cokrige_d3.o"),
    ("cokrige.hip", ["-DKK_MV_DIM=1", "-DKK_MV_W=1"], "cokrige_d1w.o"),
    ("cokrige.hip", ["-DKK_MV_DIM=2", "-DKK_MV_W=1"], "cokrige_d2w.o"),
    ("cokrige.hip", ["-DKK_MV_DIM=3", "-DKK_MV_W=1"], "cokrige_d3w.o"),
    ("sets.hip", [], "sets.o"),
    ("replay.hip", [], "replay.o"),
    ("sgsim.hip", [], "sgsim.o"),
    ("order.hip", [], "order.o"),
    ("capi.cpp", ["-x", "hip"], "capi.o"),
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ... ({r.returncode})")
    return r


def build(verbose=False):
    os.makedirs(OUT, exist_ok=True)
    objs, procs = [], []
    for src, extra, name in UNITS:
        obj = os.path.join(OUT, name)
        cmd = [HIPCC, "-c"] + COMMON + extra + [os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
        objs.append(obj)
    for cmd, p in procs:
        out, err = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out + err)
            raise RuntimeError(f"compile failed: {cmd[-3]}")
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs, verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
