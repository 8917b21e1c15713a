"""The u8 dedispersion's plan and row build, checked on the CPU.

The kernel's own planner (dd_plan_block), shape check (dd_shape) and row build
(dd_build) are compiled for the host from pulsarsr_amd/csrc/k_dedisperse.hip
by tests/dd_plan_host.cpp.  For every 64-trial block of each plan it checks
every stage record: headers within bounds, rows packed in order with every quad
mapped to its row, every read address inside its row, and each trial reading
every channel exactly once at that trial's own delay.  Those delays are the
oracle's `delays` (crates/dedisperse/src/lib.rs:9-33).  For the tiles listed it
also builds the rows from seeded u8 channel rows, reads them at the planned
addresses, and compares with the brute-force channel sums of
crates/dedisperse/src/lib.rs:62-90, over every trial of every block.  The
plans are those of the benchmark workloads (the full 4000-trial headline plan
included) and of the GPU parity tests' edge cases.  The channel rows end at an
unmapped page after the observation's read slack, so a read past it faults.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
from pulsarsr_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KERNEL = os.path.join(ROOT, "pulsarsr_amd", "csrc", "k_dedisperse.hip")
HARNESS = os.path.join(ROOT, "tests", "dd_plan_host.cpp")
DISPERSION = 4148.808  # crates/dedisperse/src/lib.rs:3


def kernel_part():
    """The planner, record helpers and row build, plus dd_shape, as host C++."""
    src = open(KERNEL).read()
    a = src.index("constexpr int DD_M = 8;")
    b = src.index("// One summing wave's 8 trials")
    part = re.sub(r"__global__ __launch_bounds__\(256\) void k_dd_plan\(.*?\n}\n", "", src[a:b], flags=re.S)
    assert "__global__" not in part, "a kernel left in the host part"
    m = re.search(r"static bool dd_shape\(.*?\n}\n", src, flags=re.S)
    return part + "\n" + m.group(0)


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    d = tmp_path_factory.mktemp("ddplan")
    (d / "kernel_part.inc").write_text(kernel_part())
    exe = d / "dd_plan_host"
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
           '-DKERNEL_PART="%s"' % (d / "kernel_part.inc"), HARNESS, "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return d, exe


def run_case(harness, name, freqs, tsamp, dms, maxval, tiles, seed=1, extra=600):
    d, exe = harness
    dms = np.sort(np.asarray(dms, dtype=np.float64))
    freqs = np.asarray(freqs, dtype=np.float64)
    D = np.stack([oracle.delays(freqs, tsamp, float(dm)) for dm in dms]).astype(np.int64)
    assert D.max() < 2 ** 31
    ref_inv = 1.0 / (freqs.max() ** 2)
    # per-channel spread bound over any 64-trial block, as psr::spread_bounds
    worst = max(dms[min(len(dms), b + 64) - 1] - dms[b] for b in range(0, len(dms), 64))
    a = np.maximum(0.0, 1.0 / (freqs * freqs) - ref_inv)
    spread = (np.ceil(DISPERSION * a * worst / tsamp) + 3).astype(np.uint32)
    max_delay = int(D.max()) + 1
    nsamples = int(D.max()) + extra
    head = np.array([len(freqs), len(dms), nsamples, maxval, seed, len(tiles)], dtype=np.uint32)
    path = d / (name + ".bin")
    with open(path, "wb") as f:
        f.write(head.tobytes())
        f.write(np.array([max_delay], dtype=np.uint64).tobytes())
        f.write(np.asarray(tiles, dtype=np.uint32).tobytes())
        f.write(spread.tobytes())
        f.write(D.astype(np.int32).tobytes())
    r = subprocess.run([str(exe), str(path)], capture_output=True, text=True, timeout=600)
    print(name, r.stdout.strip(), r.stderr.strip())
    assert r.returncode == 0 and r.stdout.startswith("OK"), r.stdout + r.stderr
    return r.stdout


def test_headline_plan(harness):
    """All 63 blocks of the 4000-trial headline plan (1024 channels, 0-999.75 step
    0.25); the rows of the first tile and a partial second one summed exactly."""
    w = synth.workload("headline-1024ch-8bit")
    out = run_case(harness, "headline", w.frequencies(), w.tsamp, w.dms, 255, [0, 1])
    assert "blocks 63 kmax 3" in out


def test_precision_plan(harness):
    """The step-doubling plan of the precision workload (3000 trials: steps 0.1,
    0.2, 0.4), 1024 channels."""
    w = synth.workload("precision-1024ch-8bit")
    run_case(harness, "precision", w.frequencies(), w.tsamp, w.dms, 255, [0])


def test_lband_plan(harness):
    """96 channels of 4-bit samples, 2000 trials of step 0.25 to DM 500."""
    w = synth.workload("lband-96ch-4bit")
    run_case(harness, "lband", w.frequencies(), w.tsamp, w.dms, 15, [0, 1], extra=1200)


def test_mixed_steps_ascending_odd(harness):
    """257 channels in ascending frequency (negative relative delays), fine steps
    (pair rows) then coarse ones (single-channel rows) and a block holding both,
    and a partial last block."""
    freqs = 1200.0 + np.arange(257)
    dms = np.concatenate([np.arange(128) * 0.25, 40.0 + np.arange(65) * 1.25])
    run_case(harness, "mixed", freqs, 64e-6, dms, 255, [0, 1], extra=1200)


def test_coarse_steps_and_tiny_counts(harness):
    """Coarse DM steps on 97 channels (mostly single-channel rows), and 1, 2 and 5
    channels."""
    run_case(harness, "coarse97", 1500.0 - 2.0 * np.arange(97), 64e-6, np.arange(80) * 3.0, 255, [0, 1])
    for nch in (1, 2, 5):
        run_case(harness, "tiny%d" % nch, 1500.0 - 20.0 * np.arange(nch), 64e-6, np.arange(70) * 5.0, 255, [0])
