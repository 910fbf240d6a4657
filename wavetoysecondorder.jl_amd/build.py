"""Build libwt2.so (HIP for gfx950) in-tree with hipcc, and next to it the
standalone A/B program wt2_mfma_ab (VALU against MFMA contractions; run by
bench.py --config mfma_ab).

Usage: python wavetoysecondorder.jl_amd/build.py   (or __graft_entry__.build()).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "wavetoy_hip", "libwt2.so")
PROBE_SRC = "wt2_mfma_ab.hip"
PROBE_OUT = os.path.join(HERE, "wavetoy_hip", "wt2_mfma_ab")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
SOURCES = ["wt2_flat.hip", "wt2_flat_lean.hip", "wt2_curved.hip", "wt2_fo.hip", "wt2_focurv.hip", "wt2_capi.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
    # packed FP32 VALU forms are switched off for every device pass
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-value",
]


def build(verbose=False, jobs=None):
    """Compile each source to an object in parallel (they share no device
    symbols), then link the shared library."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    cflags = [f for f in FLAGS if f != "-shared"]
    cmds, objs = [], []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC] + cflags + ["-c", "-o", obj, os.path.join(CSRC, src)])
    jobs = jobs or min(len(cmds), os.cpu_count() or 1)
    running = []
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        running.append(subprocess.Popen(cmd))
        while len(running) >= jobs:
            p = running.pop(0)
            if p.wait() != 0:
                raise subprocess.CalledProcessError(p.returncode, p.args)
    for p in running:
        if p.wait() != 0:
            raise subprocess.CalledProcessError(p.returncode, p.args)
    link = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", OUT + ".tmp"] + objs
    if verbose:
        print(" ".join(link), flush=True)
    subprocess.run(link, check=True)
    os.replace(OUT + ".tmp", OUT)
    probe = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-I", INCLUDE, "-I", CSRC, "-o", PROBE_OUT + ".tmp", os.path.join(CSRC, PROBE_SRC),
             "-L", os.path.dirname(OUT), "-lwt2", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(probe), flush=True)
    subprocess.run(probe, check=True)
    os.replace(PROBE_OUT + ".tmp", PROBE_OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    print("built", OUT)
    sys.exit(0)
