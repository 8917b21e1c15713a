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
# MI355X: 256 CUs at 2.4 GHz.  LDS: 256 B per CU per clock of ds_read_b128 (16
# lanes per clock; a probe with little VALU reached 190 B/clk/CU).  VALU: 64 lanes
# per CU per clock.  The dedispersion sums channel-pair rows: a 4-byte LDS word
# holds two outputs' sums of a channel pair (one byte per add), and v_add3_u32
# adds two such words into u16-pair accumulators (eight adds per lane-op).
LDS_PEAK_GBS = 256 * 256 * 2.4e9 / 1e9
VALU_PEAK_TADDS = 256 * 64 * 8 * 2.4e9 / 1e12
LDS_BYTES_PER_ADD = 1.0


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


def all_candidates(out, counts, cap):
    """Generated candidates in trial order (the `generated` list of confirmation::analyze)."""
    c = []
    for t in range(len(counts)):
        for i in range(int(counts[t])):
            x = out[t * cap + i]
            c.append(dict(frequency_hz=x.frequency_hz, period_s=x.period_s, dm=x.dm, harmonics=int(x.harmonics),
                          score=x.score))
    return c


def top_families(cands, n, tsamp, k=20):
    """Leaders of the k strongest harmonic families (confirmation.rs:177-199 with
    analyze's resolution 1.5 / (N_d tsamp)): the report's candidate list."""
    from pulsarsr_amd import confirmation
    leaders, _ = confirmation.group(cands, 1.5 / (n * tsamp), k)
    return leaders


# HBM bytes per unit of each kernel on the FP16 path (FETCH_SIZE + WRITE_SIZE,
# calibrated against a 1 GiB device copy), from the rocprofv3 passes recorded in
# profiles/hbm_traffic_headline-1024ch-8bit.md (one 768-trial launch).
MEASURED_TRAFFIC_BYTES_PER_UNIT = {
    "k_dedisp_u8": 7.53,
    "k_fft_pass1r": 10.03,
    "k_fft_pass2r": 10.04,
    "k_whiten_fused16": 5.89,
    "k_harmsum": 1.82,
}


def traffic_per_unit():
    return MEASURED_TRAFFIC_BYTES_PER_UNIT


def fp64_error(obs, w, n, out16, c16, cap, ntrials):
    """Candidate-SNR error of the FP16 search against the FP64 oracle on the
    trials holding the strongest FP16 candidates (full-length series)."""
    import oracle
    from pulsarsr_amd import dedisperse
    best = sorted(range(len(c16)), key=lambda t: -max([out16[t * cap + i].score for i in range(int(c16[t]))] or [0]))
    worst, matched = 0.0, 0
    for t in best[:ntrials]:
        x = dedisperse.apply_dms(obs, [float(w.dms[t])], n)[0]
        ref = oracle.periodicity(x, w.tsamp, float(w.dms[t]), f64=True, min_hz=w.min_hz, max_hz=w.max_hz,
                                 threshold=w.threshold, max_candidates=32, max_harmonics=w.max_harmonics)
        for i in range(int(c16[t])):
            g = out16[t * cap + i]
            hit = [r for r in ref if r["harmonics"] == g.harmonics and abs(r["frequency_hz"] - g.frequency_hz) < 1e-6 * g.frequency_hz]
            if hit:
                worst = max(worst, abs(g.score - hit[0]["score"]) / abs(hit[0]["score"]))
                matched += 1
    return dict(max_rel_snr_error=worst, matched=matched, trials=ntrials)


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
    ap.add_argument("--cpu-trials", type=int, default=160, help="trials in the CPU baseline sample (~10 s)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fp32-steps", type=int, default=1)
    ap.add_argument("--trials", type=int, default=0, help="use the first N DM trials only (profiling runs)")
    ap.add_argument("--fp64-trials", type=int, default=2, help="trials checked against the FP64 oracle")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="first copy this many MiB device-to-device (HBM counter calibration for rocprofv3)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU only (--gpus 1): the DM trials of a multi-GPU run would be "
              "split into contiguous ranges, one process per GPU, each running this one-GPU workload",
              file=sys.stderr)
        sys.exit(2)

    from pulsarsr_amd import PRECISION_FP16, PRECISION_FP32, dedisperse, synth
    if a.copy_calibration_mib:
        import torch
        src = torch.ones(a.copy_calibration_mib << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("calibration copy: %d MiB" % a.copy_calibration_mib)
        del src, dst
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
    fam16 = top_families(all_candidates(out16, c16, cap), n, w.tsamp)
    fam32 = top_families(all_candidates(out32, c32, cap32), n, w.tsamp)
    key = lambda f: [(x["frequency_hz"], x["dm"], x["harmonics"]) for x in f]
    # leaders identical: same DM and harmonic count, frequency within 1e-8 relative (the
    # parabolic bin offset moves with the FP16 spectrum's rounding)
    same_top = len(fam16) == len(fam32) and all(
        x["dm"] == y["dm"] and x["harmonics"] == y["harmonics"] and
        abs(x["frequency_hz"] - y["frequency_hz"]) <= 1e-8 * y["frequency_hz"] for x, y in zip(fam16, fam32))
    # the same families in the same rank order: leaders' frequencies within the
    # grouping resolution (a near-tie between two members of one family may pick a
    # leader one DM step or one bin apart)
    res = 1.5 / (n * w.tsamp)
    same_fam = len(fam16) == len(fam32) and all(abs(x["frequency_hz"] - y["frequency_hz"]) <= res
                                                 for x, y in zip(fam16, fam32))
    diff = [(i, key([x])[0], key([y])[0]) for i, (x, y) in enumerate(zip(fam16, fam32))
            if x["dm"] != y["dm"] or x["harmonics"] != y["harmonics"]
            or abs(x["frequency_hz"] - y["frequency_hz"]) > 1e-8 * y["frequency_hz"]]
    if diff:
        log("top-20 leaders differing (rank, fp16, fp32): %s" % diff[:6])
    snr_err = max([abs(x["score"] - y["score"]) / abs(y["score"]) for x, y in zip(fam16, fam32)] or [0.0])

    # dominant kernel: dedispersion (nchans adds per output; their operands come
    # from LDS as channel-pair words, one byte per add)
    dd_ms = stages["dedisperse_ms"]
    adds = float(units) * w.nchans
    ach_tadds = adds / (dd_ms * 1e-3) / 1e12
    ach_lds = adds * LDS_BYTES_PER_ADD / (dd_ms * 1e-3) / 1e9
    tpu = traffic_per_unit()
    launches = int(stages.get("launches", 0)) or 1
    units_per_launch = float(units) / launches
    dd_traffic = tpu.get("k_dedisp_u8")
    # spectral chain bytes (FP16 storage): series read 2, spectrum write 2r, whitening
    # read 2r + write 2r, harmonic-sum read 2r per unit (r = bins per series sample)
    r = (stages["fft_len"] // 2 + 1) / n
    spec_ms = stages["fft_ms"] + stages["whiten_ms"] + stages["harmsum_ms"]
    spec_alg = 2 + 2 * r + 2 * r + 2 * r
    spec_traffic = sum(tpu.get(k, 0.0) for k in ("k_fft_pass1r", "k_fft_pass2r", "k_whiten_fused16", "k_harmsum")) or None
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
        "top20_leaders_identical_fp16_vs_fp32": same_top,
        "top20_families_identical_fp16_vs_fp32": same_fam,
        "top20_max_rel_snr_diff_fp16_vs_fp32": snr_err,
        "stages_ms": {k: stages[k] for k in ("delays_ms", "dedisperse_ms", "fft_ms", "whiten_ms", "harmsum_ms", "finalize_ms", "total_ms")},
        "ingest_s": ingest_s,
        "roofline": {"bound": "lds", "kernel": "k_dedisp_u8", "achieved": ach_lds, "peak": LDS_PEAK_GBS,
                     "unit": "GB/s", "frac": ach_lds / LDS_PEAK_GBS,
                     "traffic": dd_traffic * units_per_launch if dd_traffic else None,
                     "valu": {"achieved": ach_tadds, "peak": VALU_PEAK_TADDS, "unit": "Tadd/s",
                              "frac": ach_tadds / VALU_PEAK_TADDS},
                     "note": "1 LDS byte per add (channel-pair rows), adds = trials x series_len x nchans; "
                             "HIP-event time of the dedispersion stage; traffic = HBM bytes per launch"},
        "roofline_spectral": {"bound": "hbm", "kernels": "fft+whiten+harmsum",
                              "achieved": float(units) * spec_alg / (spec_ms * 1e-3) / 1e9,
                              "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": float(units) * spec_alg / (spec_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                              "traffic": spec_traffic * units_per_launch if spec_traffic else None},
        "launches_per_step": launches,
    }
    if not a.no_cpu:
        res["fp64_check"] = fp64_error(obs, w, n, out16, c16, cap, a.fp64_trials)
    if not a.no_cpu:
        res["cpu_baseline"] = cpu_baseline(payload, w, n, a.cpu_trials, min(16, os.cpu_count() or 1))
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
