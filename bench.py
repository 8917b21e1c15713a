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
LDS_PROBE_GBS = 256 * 190 * 2.4e9 / 1e9  # the ds_read_b128 rate a probe reached
VALU_PEAK_TADDS = 256 * 64 * 8 * 2.4e9 / 1e12
# FP32 VALU: 157.3 TFLOP/s with packed FP32 instructions, which this build leaves
# disabled (one FMA per lane per clock: 78.6 TFLOP/s)
FP32_PEAK_TFLOPS = 157.3
FP32_UNPACKED_TFLOPS = 256 * 64 * 2 * 2.4e9 / 1e12
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


# HBM bytes per launch of each kernel on the FP16 path of the headline step (all
# 4000 trials in one dedispersion launch; FETCH_SIZE + WRITE_SIZE calibrated
# against a 1 GiB device copy), from the rocprofv3 passes over this script's own
# launch recorded in profiles/hbm_traffic_headline-1024ch-8bit.md.  They describe
# that launch shape only: other workloads or trial counts report traffic null.
MEASURED_LAUNCH = {"workload": "headline-1024ch-8bit", "trials": 4000, "launches": 1}
MEASURED_TRAFFIC_GB_PER_LAUNCH = {
    "k_dedisp_u8": 85.65 + 17.64,
    "k_fft_pass1r": 17.02 + 67.11,
    "k_fft_pass2r": 67.25 + 16.81,
    "k_whiten_fused16": 32.54 + 17.36,
    "k_harmsum": 15.22 + 0.02,
}


def traffic_gb(kernel, w, launches):
    if (w.name, len(w.dms), launches) != (MEASURED_LAUNCH["workload"], MEASURED_LAUNCH["trials"],
                                        MEASURED_LAUNCH["launches"]):
        return None
    return MEASURED_TRAFFIC_GB_PER_LAUNCH.get(kernel)


def host_cores():
    """(threads to use, cores visible, CPU model): the box's CPU share is 16
    cores per GPU whatever the affinity mask shows."""
    visible = len(os.sched_getaffinity(0))
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return min(16, visible), visible, model


def cpu_baseline(payload, w, n, ntrials):
    """The oracle's per-trial loop (restated reference algorithm) on a sample of
    trials, on the host cores: one warm-up run, then the median of three
    (the reference's own protocol, scripts/benchmark.py:22-31)."""
    import oracle
    threads, visible, model = host_cores()
    data, _ = oracle.decode(payload, w.nbits, w.nchans)
    sel = w.dms[np.linspace(0, len(w.dms) - 1, ntrials).astype(int)]
    kw = dict(min_hz=w.min_hz, max_hz=w.max_hz, threshold=w.threshold, max_candidates=32,
              max_harmonics=w.max_harmonics)
    oracle.search_trials(data, w.nsamples, w.frequencies(), w.tsamp, sel[:threads], n, threads=threads, **kw)
    runs = []
    for _ in range(3):
        t0 = time.perf_counter()
        oracle.search_trials(data, w.nsamples, w.frequencies(), w.tsamp, sel, n, threads=threads, **kw)
        runs.append(time.perf_counter() - t0)
        log("cpu baseline run: %.2f s" % runs[-1])
    dt = float(np.median(runs))
    return dict(value=ntrials * n / dt, unit="DM-trial*samples/s", cores=threads, kind="port",
                host_cores_visible=visible, cpu=model, runs_s=runs,
                sample="%d of %d DM trials (evenly spaced), full %d-sample series each; one warm-up, median of "
                       "3 runs (%.2f s)" % (ntrials, len(w.dms), n, dt))


def fp64_report(obs, w, n, out16, c16, cap, ntrials):
    """FP16 storage against the all-f64 oracle on `ntrials` trials (the ones
    holding the strongest FP16 candidates): candidate-SNR error over matched
    candidates and the spectral-power error (normwise and worst bin relative to
    its whitening level) of the stored FP16 spectrum."""
    import oracle
    from pulsarsr_amd import PRECISION_FP16, dedisperse, search
    best = sorted(range(len(c16)), key=lambda t: -max([out16[t * cap + i].score for i in range(int(c16[t]))] or [0]))
    worst, matched, total, pnorm, pbin = 0.0, 0, 0, 0.0, 0.0
    for t in best[:ntrials]:
        x = dedisperse.apply_dms(obs, [float(w.dms[t])], n)[0]
        ref = oracle.periodicity(x, w.tsamp, float(w.dms[t]), f64=True, min_hz=w.min_hz, max_hz=w.max_hz,
                                 threshold=w.threshold, max_candidates=32, max_harmonics=w.max_harmonics)
        nfft = 2 * (1 << (n - 1).bit_length())
        bin_hz = 1.0 / (nfft * w.tsamp)
        for i in range(int(c16[t])):
            g = out16[t * cap + i]
            total += 1
            hit = [r for r in ref if r["harmonics"] == g.harmonics and abs(r["frequency_hz"] - g.frequency_hz) <= 0.05 * bin_hz]
            if hit:
                worst = max(worst, abs(g.score - hit[0]["score"]) / abs(hit[0]["score"]))
                matched += 1
        raw16, _ = search.spectrum(x, w.tsamp, search.Config(min_hz=w.min_hz, precision=PRECISION_FP16))
        raw64, white64 = oracle.spectrum_f64(x, w.tsamp, min_hz=w.min_hz)
        p16 = raw16.astype(np.float64) * n  # stored amplitudes are |X| / sqrt(n)
        pnorm = max(pnorm, float(np.linalg.norm(p16 - raw64) / np.linalg.norm(raw64)))
        ok = white64 > 0  # the whitening scale raw / white, where defined (not DC or birdies)
        level = raw64[ok] / white64[ok]
        # each bin's error relative to the larger of its power and its whitening level
        pbin = max(pbin, float(np.max(np.abs(p16 - raw64)[ok] / np.maximum(raw64[ok], level))))
    return dict(max_rel_snr_error=worst, matched=matched, candidates=total, trials=ntrials,
                max_spectral_power_error_normwise=pnorm, max_spectral_power_error_per_bin=pbin)


def measure_config(name, steps, fp64_trials):
    """One smaller workload: FP16 and FP32 rates and the FP16 errors against FP64."""
    from pulsarsr_amd import PRECISION_FP16, PRECISION_FP32, dedisperse, synth
    w = synth.workload(name)
    payload = synth.generate(w, threads=host_cores()[0])
    n = discovery_length(w)
    obs = dedisperse.Observation.from_packed(payload, w.nbits, w.nchans, w.tsamp, w.fch1, w.foff)
    res = {"nchans": w.nchans, "nbits": w.nbits, "nsamples": w.nsamples, "trials": len(w.dms), "series_len": n}
    units = len(w.dms) * n
    for prec, key in ((PRECISION_FP16, "fp16"), (PRECISION_FP32, "fp32")):
        run_search(obs, w, n, prec)
        ts = []
        for _ in range(steps):
            dt, out, cnt, cap, st = run_search(obs, w, n, prec)
            ts.append(dt)
        ms = 1e3 * float(np.median(ts))
        res[key] = {"value": units / (ms * 1e-3), "ms_per_step": ms, "dedisperse_ms": st["dedisperse_ms"],
                    "spectral_ms": st["fft_ms"] + st["whiten_ms"] + st["harmsum_ms"]}
        if key == "fp16":
            out16, c16, cap16 = out, cnt.copy(), cap
    res["fp16_over_fp32"] = res["fp32"]["ms_per_step"] / res["fp16"]["ms_per_step"]
    if fp64_trials:
        res["fp64_check"] = fp64_report(obs, w, n, out16, c16, cap16, fp64_trials)
    obs.close()
    log("%s: fp16 %.1f ms, fp32 %.1f ms" % (name, res["fp16"]["ms_per_step"], res["fp32"]["ms_per_step"]))
    return res


def measure_batching(steps):
    """The batching form: 256 observations of uneven size (channels 64-1024,
    2/4/8-bit, 2^16-2^21 samples, 50-2000 trials), generated on the GPU, made
    resident, then searched as one autotuned batch captured in a hipGraph."""
    from pulsarsr_amd import PRECISION_FP16, PRECISION_FP32, dedisperse, search, synth
    ws = synth.batch_workloads()
    t0 = time.perf_counter()
    obs = []
    for w in ws:
        p = synth.generate(w, device=True)
        obs.append(dedisperse.Observation.from_packed(p, w.nbits, w.nchans, w.tsamp, w.fch1, w.foff))
        del p
    lengths = [discovery_length(w) for w in ws]
    units = sum(len(w.dms) * n for w, n in zip(ws, lengths))
    log("batching workload resident: %d observations, %d trials, %.1f s" % (len(ws), sum(len(w.dms) for w in ws),
                                                                          time.perf_counter() - t0))
    res = {"observations": len(ws), "trials": sum(len(w.dms) for w in ws), "units": units}
    for prec, key in ((PRECISION_FP16, "fp16"), (PRECISION_FP32, "fp32")):
        cfg = search.Config(min_hz=10.0, max_hz=2000.0, threshold=8.0, max_candidates=32, max_harmonics=16,
                            precision=prec)
        t0 = time.perf_counter()
        b = search.Batch(obs, [w.dms for w in ws], cfg, lengths=lengths, graph=True, autotune=True)
        setup = time.perf_counter() - t0
        b.run_raw()
        ts = []
        for _ in range(steps):
            t0 = time.perf_counter()
            b.run_raw()
            ts.append(time.perf_counter() - t0)
        ms = 1e3 * float(np.median(ts))
        info = b.info()
        res[key] = {"value": units / (ms * 1e-3), "ms_per_step": ms, "device_ms": info["last_ms"],
                    "setup_s": setup, "rounds": info["rounds"]}
        if key == "fp16":
            res["autotune"] = {"fft": info["fft"], "dedisperse_span": info["dedisperse_span"],
                               "dedisperse_wg_per_cu": info["dedisperse_wg_per_cu"]}
        b.close()
        log("batching %s: %.1f ms per step (%d rounds)" % (key, ms, info["rounds"]))
    res["fp16_over_fp32"] = res["fp32"]["ms_per_step"] / res["fp16"]["ms_per_step"]
    for o in obs:
        o.close()
    return res


def copy_calibration(mib):
    """One device-to-device copy of `mib` MiB through the HIP runtime: a
    transfer of known size for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    nbytes = C.c_size_t(mib << 20)
    src, dst = C.c_void_p(), C.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(C.byref(p), nbytes) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, nbytes)
    if hip.hipMemcpy(dst, src, nbytes, 3) != 0:  # hipMemcpyDeviceToDevice
        raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    log("calibration copy: %d MiB" % mib)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="headline-1024ch-8bit")
    ap.add_argument("--cpu-trials", type=int, default=96, help="trials in the CPU baseline sample (~5 s per run)")
    ap.add_argument("--no-configs", action="store_true", help="skip the smaller workloads and the batching form")
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
        copy_calibration(a.copy_calibration_mib)
    w = synth.workload(a.workload)
    if a.trials:
        w.dms = w.dms[:a.trials]
    t0 = time.perf_counter()
    payload = synth.generate(w, threads=host_cores()[0])
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
    launches = int(stages.get("launches", 0)) or 1
    dd_traffic = traffic_gb("k_dedisp_u8", w, launches)
    # spectral chain bytes (FP16 storage): series read 2, spectrum write 2r, whitening
    # read 2r + write 2r, harmonic-sum read 2r per unit (r = bins per series sample)
    r = (stages["fft_len"] // 2 + 1) / n
    spec_ms = stages["fft_ms"] + stages["whiten_ms"] + stages["harmsum_ms"]
    spec_alg = 2 + 2 * r + 2 * r + 2 * r
    spec_parts = [traffic_gb(k, w, launches) for k in ("k_fft_pass1r", "k_fft_pass2r", "k_whiten_fused16", "k_harmsum")]
    spec_traffic = sum(spec_parts) if all(x is not None for x in spec_parts) else None
    nfft = float(stages["fft_len"])
    fft_flop = len(w.dms) * 2.5 * nfft * np.log2(nfft)
    res = {
        "metric": "DM-trial·samples/s (dedisp+FFT+harmsum), FP16-store vs FP32 at same candidates",
        "value": units / (ms16 * 1e-3),
        "unit": "DM-trial*samples/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms16,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp16-storage/fp32-compute (u8 samples, exact integer channel sums)",
        "data": "synthetic (seeded generator, pulsars+RFI, generated in-process)",
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
                     "probe_peak": LDS_PROBE_GBS, "frac_of_probe_peak": ach_lds / LDS_PROBE_GBS,
                     "traffic": dd_traffic * 1e9 if dd_traffic else None,
                     "valu": {"achieved": ach_tadds, "peak": VALU_PEAK_TADDS, "unit": "Tadd/s",
                              "frac": ach_tadds / VALU_PEAK_TADDS},
                     "note": "1 LDS byte per add (channel-pair rows), adds = trials x series_len x nchans; "
                             "HIP-event time of the dedispersion stage; traffic = HBM bytes of this launch "
                             "(FETCH_SIZE + WRITE_SIZE, profiles/hbm_traffic_headline-1024ch-8bit.md): 10.3 of "
                             "12.4 B/unit are reads, the stage records (0.18 MB per work item from the plan "
                             "table) and the channel bytes the row build reads, fetched again after L2 eviction",
                     "hbm": {"achieved": (units * 2.0 + w.nchans * w.nsamples) / (dd_ms * 1e-3) / 1e9,
                             "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": (units * 2.0 + w.nchans * w.nsamples) / (dd_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                             "note": "algorithmic bytes: FP16 series write + the filterbank read once"}},
        # the FFT is VALU-bound: 2.5 N log2 N flop per real transform of N points
        "roofline_fft": {"bound": "valu", "kernels": "k_fft_pass1r+k_fft_pass2r",
                         "achieved": fft_flop / (stages["fft_ms"] * 1e-3) / 1e12, "unit": "TFLOP/s",
                         "peak": FP32_PEAK_TFLOPS,
                         "frac": fft_flop / (stages["fft_ms"] * 1e-3) / 1e12 / FP32_PEAK_TFLOPS,
                         "frac_unpacked": fft_flop / (stages["fft_ms"] * 1e-3) / 1e12 / FP32_UNPACKED_TFLOPS,
                         "note": "flop = trials x 2.5 N log2 N (N = FFT length); peak 157.3 TFLOP/s counts packed "
                                 "FP32, which this build disables; frac_unpacked is against 78.6 TFLOP/s"},
        "roofline_spectral": {"bound": "hbm", "kernels": "fft+whiten+harmsum",
                              "achieved": float(units) * spec_alg / (spec_ms * 1e-3) / 1e9,
                              "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": float(units) * spec_alg / (spec_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                              "traffic": spec_traffic * 1e9 if spec_traffic else None},
        "launches_per_step": launches,
    }
    if not a.no_cpu:
        res["fp64_check"] = fp64_report(obs, w, n, out16, c16, cap, a.fp64_trials)
    obs.close()
    if not a.no_configs:
        res["configs"] = {name: measure_config(name, 3, a.fp64_trials if not a.no_cpu else 0)
                          for name in ("plain-64ch-8bit", "lband-96ch-4bit", "precision-1024ch-8bit")}
        res["configs"]["batching-256obs"] = measure_batching(3)
    if not a.no_cpu:
        res["cpu_baseline"] = cpu_baseline(payload, w, n, a.cpu_trials)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
