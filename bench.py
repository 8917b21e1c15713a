"""Benchmark of the search hot path on one MI355X.

Workload (seeded synthetic, generated in-process): 1024 channels x 8-bit,
1200-1600 MHz, tsamp 64 us, 2^22 samples, 8 injected pulsars, narrowband RFI in
51 channels and 3 broadband bursts; 4000 DM trials 0-999.75 (step 0.25); 16
harmonics.  A step is one pass of delays -> dedisperse -> FFT -> power ->
whitening -> harmonic sums/peaks -> per-trial ranking over all 4000 trials,
with the packed filterbank already resident in HBM.  The metric counts one
unit per dedispersed, searched sample of one trial: trials x N_d, where N_d
is the per-trial series length (the CLI's discovery length).

Prints ONE JSON line.  --gpus above 1 is not supported (one GPU per box).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
VALU_PEAK_TADDS = 256 * 64 * 2.4e9 / 1e12  # 32-bit adds: 4 SIMD x 16 lanes per CU per clock


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def discovery_length(w):
    from pulsarsr_amd import dedisperse
    d = dedisperse.delays(w.frequencies(), w.tsamp, float(w.dms[-1]))
    return (w.nsamples - 2 * int(d.max())) // 2  # crates/cli/src/main.rs:433-446


def run_search(obs, w, n, precision):
    from pulsarsr_amd import search
    cfg = search.Config(min_hz=w.min_hz, max_hz=w.max_hz, threshold=w.threshold, max_candidates=32,
                        max_harmonics=w.max_harmonics, precision=precision)
    t0 = time.perf_counter()
    out, counts, cap = search.search_raw(obs, w.dms, n, cfg)
    dt = time.perf_counter() - t0
    return dt, out, counts, cap, obs.ctx.stage_times()


def top_list(out, counts, cap, k=20):
    c = []
    for t in range(len(counts)):
        for i in range(int(counts[t])):
            x = out[t * cap + i]
            c.append((x.score, x.frequency_hz, x.dm, int(x.harmonics)))
    c.sort(key=lambda v: -v[0])
    return c[:k]


def cpu_baseline(payload, w, n, ntrials, threads):
    """The oracle's per-trial loop (restated reference algorithm) on a sample of trials."""
    import oracle
    data = np.ascontiguousarray(payload.reshape(w.nsamples, w.nchans).T, dtype=np.float32)
    sel = w.dms[np.linspace(0, len(w.dms) - 1, ntrials).astype(int)]
    t0 = time.perf_counter()
    oracle.search_trials(data, w.nsamples, w.frequencies(), w.tsamp, sel, n, threads=threads,
                         min_hz=w.min_hz, max_hz=w.max_hz, threshold=w.threshold, max_candidates=32,
                         max_harmonics=w.max_harmonics)
    dt = time.perf_counter() - t0
    return dict(value=ntrials * n / dt, unit="DM-trial*samples/s", cores=threads, kind="port",
                sample="%d of %d DM trials (evenly spaced), full %d-sample series each, %.1f s"
                       % (ntrials, len(w.dms), n, dt))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="headline-1024ch-8bit")
    ap.add_argument("--cpu-trials", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fp32-steps", type=int, default=1)
    ap.add_argument("--trials", type=int, default=0, help="use the first N DM trials only (profiling runs)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: multi-GPU runs are not part of this build "
              "(DM trials would shard across GPUs; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    from pulsarsr_amd import PRECISION_FP16, PRECISION_FP32, dedisperse, synth
    w = synth.workload(a.workload)
    if a.trials:
        w.dms = w.dms[:a.trials]
    t0 = time.perf_counter()
    payload = synth.generate(w, threads=min(16, os.cpu_count() or 1))
    log("generated %.2f GB in %.1f s" % (payload.nbytes / 1e9, time.perf_counter() - t0))
    n = discovery_length(w)
    t0 = time.perf_counter()
    obs = dedisperse.Observation.from_packed(payload, w.nbits, w.nchans, w.tsamp, w.fch1, w.foff)
    ingest_s = time.perf_counter() - t0
    units = len(w.dms) * n
    log("observation resident (%.2f s upload+ingest); %d trials x %d samples" % (ingest_s, len(w.dms), n))

    for _ in range(a.warmup):
        run_search(obs, w, n, PRECISION_FP16)
    times, stages = [], None
    for i in range(a.steps):
        dt, out16, c16, cap, st = run_search(obs, w, n, PRECISION_FP16)
        times.append(dt)
        stages = st
        log("fp16 step %d: %.1f ms" % (i, dt * 1e3))
    ms16 = 1e3 * float(np.median(times))
    t32 = []
    for i in range(max(1, a.fp32_steps)):
        dt, out32, c32, cap32, st32 = run_search(obs, w, n, PRECISION_FP32)
        t32.append(dt)
        log("fp32 step %d: %.1f ms" % (i, dt * 1e3))
    ms32 = 1e3 * float(np.median(t32))
    top16, top32 = top_list(out16, c16, cap), top_list(out32, c32, cap32)
    same_top = [(round(x[1], 6), x[2], x[3]) for x in top16] == [(round(x[1], 6), x[2], x[3]) for x in top32]

    # dominant kernel: dedispersion, VALU-bound (1 add per channel per output)
    dd_ms = stages["dedisperse_ms"]
    adds = float(units) * w.nchans
    ach_tadds = adds / (dd_ms * 1e-3) / 1e12
    # spectral chain bytes (FP16 storage): series write+read 2+2, spectrum 2r+2r per unit
    r = (stages["fft_len"] // 2 + 1) / n
    spec_ms = stages["fft_ms"] + stages["whiten_ms"] + stages["harmsum_ms"]
    res = {
        "metric": "DM-trial·samples/s (dedisp+FFT+harmsum), FP16-store vs FP32 at same candidates",
        "value": units / (ms16 * 1e-3),
        "unit": "DM-trial*samples/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms16,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u8", "data": "synthetic (seeded generator, pulsars+RFI, generated in-process)",
        "config": {"workload": w.name, "nchans": w.nchans, "nbits": w.nbits, "nsamples": w.nsamples,
                   "trials": len(w.dms), "series_len": n, "fft_len": int(stages["fft_len"]),
                   "harmonics": w.max_harmonics, "storage": "fp16"},
        "fp32": {"value": units / (ms32 * 1e-3), "ms_per_step": ms32},
        "fp16_over_fp32": ms32 / ms16,
        "top20_identical_fp16_vs_fp32": same_top,
        "stages_ms": {k: stages[k] for k in ("delays_ms", "dedisperse_ms", "fft_ms", "whiten_ms", "harmsum_ms", "finalize_ms", "total_ms")},
        "ingest_s": ingest_s,
        "roofline": {"bound": "valu", "kernel": "k_dedisp_u8", "achieved": ach_tadds, "peak": VALU_PEAK_TADDS,
                     "unit": "Tadd/s", "frac": ach_tadds / VALU_PEAK_TADDS, "traffic": None},
        "roofline_spectral": {"bound": "hbm", "kernels": "fft+whiten+harmsum",
                              "achieved": float(units) * (4 + 4 * r) / (spec_ms * 1e-3) / 1e9,
                              "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": float(units) * (4 + 4 * r) / (spec_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                              "traffic": None},
    }
    if not a.no_cpu:
        res["cpu_baseline"] = cpu_baseline(payload, w, n, a.cpu_trials, min(16, os.cpu_count() or 1))
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
