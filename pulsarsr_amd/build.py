"""Build recipe for libpsrgpu.so (gfx950) -- invoked by __graft_entry__.build().

Every source is compiled with hipcc for --offload-arch=gfx950 with
-ffp-contract=off (the reference's Rust never fuses multiply-add, and the
parity-critical f32/f64 expressions must round identically) and with packed
FP32 instructions disabled.  Objects go to pulsarsr_amd/_build/, the shared
library to pulsarsr_amd/libpsrgpu.so.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpsrgpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["k_dedisperse.hip", "k_fft.hip", "k_spectrum.hip", "k_fold.hip", "k_rfi.hip", "api.cpp",
           "analyze.cpp", "rfi.cpp", "host_math.cpp", "noalloc.cpp", "batch.cpp"]
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-ffp-contract=off",
    "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def _compile(src):
    obj = os.path.join(BUILD, src + ".o")
    cmd = [HIPCC] + FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj, r.stderr


SYNTH_LIB = os.path.join(HERE, "libpsrsynth.so")


def build_synth():
    """The workload generator: plain host C++ (no device code)."""
    tmp = SYNTH_LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", os.path.join(CSRC, "synth.cpp"),
           "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("synth build failed: %s" % r.stderr)
    os.replace(tmp, SYNTH_LIB)
    return SYNTH_LIB


SYNTH_GPU_LIB = os.path.join(HERE, "libpsrsynthgpu.so")


def build_synth_gpu():
    """The GPU workload generator (k_synth.hip), a library of its own: it is a
    workload tool, not part of the search path."""
    tmp = SYNTH_GPU_LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-shared", "-x", "hip", os.path.join(CSRC, "k_synth.hip"), "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("synth gpu build failed: %s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, SYNTH_GPU_LIB)
    return SYNTH_GPU_LIB


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        results = list(ex.map(_compile, SOURCES))
    objs = [o for o, _ in results]
    if verbose:
        for _, err in results:
            if err.strip():
                sys.stderr.write(err)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    build_synth()
    build_synth_gpu()
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
