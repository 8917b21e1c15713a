"""`psr` drop-in for the path's commands (crates/cli/src/main.rs): search, fold,
dedisperse, header.  Run as `python -m pulsarsr_amd <command> ...`.

Arguments, defaults, validation messages, console lines and the JSON outputs
(header.json, report.json schema 2, fold.json, series.json/series.f32) follow
the reference; the PNG plots and the `inspect`/`synth` commands are not
provided.  `search` adds three flags the reference lacks (its behaviour is the
default of each): --max-harmonics, --precision and --dm-plan.  Every stage runs on the GPU through libpsrgpu: packed ingest,
rfi::mask, the DM-trial search, confirmation::analyze, dedispersion and folds.
"""
import argparse
import hashlib
import math
import os
import struct
import subprocess
import sys
import time

import numpy as np

from . import PRECISION_FP16, PRECISION_FP32, confirmation, dedisperse, fbank, fold, report, rfi, search
from .core import PsrError

VERSION = "0.1.0"
SCORE_DEFINITION = ("discovery spectral ranking and approximate profile statistics; confirmation fixed-window "
                    "on/off mean difference / sample standard error on disjoint second half; colored noise/RFI "
                    "may violate independent-noise assumptions")

# crates/fbank/src/lib.rs:87-99: keyword -> value type
_TEXT = {"rawdatafile", "source_name"}
_INT = {"machine_id", "telescope_id", "data_type", "barycentric", "pulsarcentric", "nbits", "nsamples", "nchans",
        "nifs", "nbeams", "ibeam", "signed"}
_FLOAT = {"az_start", "za_start", "src_raj", "src_dej", "tstart", "tsamp", "fch1", "foff", "refdm", "period"}


def header_fields(raw):
    """The header's keyword map (fbank::Header::fields), after the C ABI's strict parse."""
    pos = 0

    def string():
        nonlocal pos
        n = struct.unpack_from("<i", raw, pos)[0]
        pos += 4
        s = bytes(raw[pos:pos + n]).decode()
        pos += n
        return s
    string()  # HEADER_START
    fields = {}
    while True:
        k = string()
        if k == "HEADER_END":
            break
        if k in _TEXT:
            fields[k] = string()
        elif k in _INT:
            fields[k] = struct.unpack_from("<i", raw, pos)[0]
            pos += 4
        elif k in _FLOAT:
            fields[k] = float(struct.unpack_from("<d", raw, pos)[0])
            pos += 8
    return dict(sorted(fields.items()))


def header_json(path, hdr, raw):
    return {"fields": header_fields(raw), "nchans": hdr.nchans, "nbits": hdr.nbits, "tsamp": hdr.tsamp,
            "fch1": hdr.fch1, "foff": hdr.foff, "data_offset": hdr.data_offset}


def _open(path, max_samples=None):
    hdr, payload = fbank.read_packed(path)
    raw = memoryview(payload.base)[:hdr.data_offset] if payload.base is not None else open(path, "rb").read()
    frame_bits = hdr.nchans * hdr.nbits
    if (payload.nbytes * 8) % frame_bits:
        raise PsrError(-5, "invalid or unsupported filterbank: truncated frame or packed padding unsupported")
    obs = dedisperse.Observation.from_packed(payload, hdr.nbits, hdr.nchans, hdr.tsamp, hdr.fch1, hdr.foff,
                                             nsamples=max_samples or 0)
    return hdr, obs, raw


def _profile(p):
    return {"period_s": p["period_s"], "snr": p["snr"], "reduced_chi_square": p["reduced_chi_square"],
            "bins": [report.F32(v) for v in np.asarray(p["bins"], np.float32)],
            "counts": [int(v) for v in p["counts"]]}


def _confirmed(c):
    return {
        "candidate": c["candidate"], "refined_dm": c["refined_dm"], "profile": _profile(c["profile"]),
        "window": c["window"], "validation": c["validation"], "zero_dm_validation": c["zero_dm_validation"],
        "zero_dm_period_s": c["zero_dm_period_s"], "zero_dm_window": c["zero_dm_window"],
        "dm_contrast": c["dm_contrast"], "dispersion_widths": c["dispersion_widths"],
        "zero_dm_related_spectral_ratio": c["zero_dm_related_spectral_ratio"],
        "time_parts": c["time_parts"], "subbands": c["subbands"],
        "time_profiles": [[report.F32(v) for v in np.asarray(p, np.float32)] for p in c["time_profiles"]],
        "subband_profiles": [[report.F32(v) for v in np.asarray(p, np.float32)] for p in c["subband_profiles"]],
        "dm_response": c["dm_response"], "harmonic_members": c["harmonic_members"],
        "accepted": c["accepted"], "rejection_reasons": c["rejection_reasons"]}


def cmd_search(a):
    # crates/cli/src/main.rs:395-565
    if (not math.isfinite(a.min_hz) or not math.isfinite(a.max_hz) or a.min_hz <= 0 or a.max_hz <= a.min_hz
            or not math.isfinite(a.threshold) or a.threshold <= 0 or a.top == 0 or a.bins < 8):
        raise SystemExit("Error: \"invalid search settings\"")
    if a.threads is not None and a.threads == 0:
        raise SystemExit("Error: \"threads must be positive\"")
    start = time.perf_counter()
    os.makedirs(a.out, exist_ok=True)
    hdr, obs, raw = _open(a.file, a.max_samples)
    freqs = hdr.frequencies()
    tsamp = hdr.tsamp
    report.write(os.path.join(a.out, "header.json"), header_json(a.file, hdr, raw))
    mask = None if a.no_rfi else rfi.mask(obs, max(int(0.1 / tsamp), 32), 8.0)
    if a.dm_plan is not None:
        trials, step = a.dm_plan, a.dm_plan_step
        a.dm_min, a.dm_max = float(trials[0]), float(trials[-1])
    else:
        step = a.dm_step if a.dm_step is not None else dedisperse.dm_step(freqs, tsamp)
        trials = dedisperse.plan(a.dm_min, a.dm_max, step)
    max_delay = int(dedisperse.delays(freqs, tsamp, a.dm_max).max())
    if obs.nsamples < max_delay or obs.nsamples - max_delay < 512:
        raise SystemExit("Error: \"DM range leaves fewer than 512 samples\"")
    length = obs.nsamples - max_delay
    if length < max_delay or length - max_delay < 256:
        raise SystemExit("Error: \"DM range leaves no independent discovery/confirmation split\"")
    discovery = (length - max_delay) // 2
    cfg = search.Config(min_hz=a.min_hz, max_hz=a.max_hz, threshold=a.threshold, max_candidates=32,
                        birdies=list(a.birdie), max_harmonics=a.max_harmonics,
                        precision=PRECISION_FP16 if a.precision == "fp16" else PRECISION_FP32)
    # rayon's default pool: one worker per available core (crates/cli/src/main.rs:407-414)
    threads = a.threads or len(os.sched_getaffinity(0))
    print("%d samples, %d channels, %d DM trials (step %.6f), %d threads"
          % (obs.nsamples, hdr.nchans, len(trials), step, threads))
    lists = search.search(obs, trials, discovery, cfg)
    generated = [c for lst in lists for c in lst]
    conf = confirmation.analyze(obs, trials, generated, a.top, a.bins)
    confirmed = [_confirmed(c) for c in conf["candidates"]]
    folded = []
    for index, c in enumerate(confirmed):
        best = c["profile"]
        print("#%d: P=%.6f ms DM=%.3f spectral score=%.2f held-out SNR=%.2f %s"
              % (index, best["period_s"] * 1000.0, c["refined_dm"], c["candidate"]["score"],
                 c["validation"]["snr"], "ACCEPTED for follow-up" if c["accepted"] else "REJECTED"))
        folded.append({"candidate": c["candidate"], "refined_dm": c["refined_dm"], "profile": best,
                       "dm_curve": [[p["dm"], p["discovery_profile_response"]] for p in c["dm_response"]],
                       "confirmation": c})
    sha = hashlib.sha256()
    with open(a.file, "rb") as f:
        for chunk in iter(lambda: f.read(1 << 20), b""):
            sha.update(chunk)
    try:
        rev = subprocess.run(["git", "rev-parse", "HEAD"], capture_output=True, text=True, timeout=10)
        code_revision = rev.stdout.strip() if rev.returncode == 0 else None
    except (OSError, subprocess.SubprocessError):
        code_revision = None
    outcome = {"rules": {"version": "split-confirm-v1", **conf["rules"]}}
    for k in ("discovery_samples", "validation_start_sample", "validation_samples", "common_samples", "trial_count"):
        outcome[k] = conf[k]
    outcome["candidates"] = confirmed
    elapsed = time.perf_counter() - start
    rep = {"schema_version": 2, "input_sha256": sha.hexdigest(), "code_revision": code_revision,
           "confirmation": outcome, "input": a.file, "version": VERSION, "sample_count": obs.nsamples,
           "tsamp": tsamp, "dm_min": a.dm_min, "dm_max": a.dm_max, "dm_step": step, "trial_count": len(trials),
           "common_samples": length, "threads": threads, "elapsed_seconds": elapsed, "min_hz": a.min_hz,
           "max_hz": a.max_hz, "threshold": a.threshold, "birdies": [list(b) for b in a.birdie],
           "score_definition": SCORE_DEFINITION, "rfi": mask,
           "candidates": [c["candidate"] for c in confirmed], "folded": folded}
    report.write(os.path.join(a.out, "report.json"), rep)
    print("Saved %s in %.2fs" % (a.out, elapsed))
    return rep


def cmd_fold(a):
    # crates/cli/src/main.rs:295-330
    if not math.isfinite(a.period) or a.period <= 0 or a.bins < 8:
        raise SystemExit("Error: \"positive period and >=8 bins required\"")
    hdr, obs, _ = _open(a.file)
    x = dedisperse.apply(obs, a.dm)
    p = fold.profile(x, hdr.tsamp, a.period, a.bins)
    os.makedirs(a.out, exist_ok=True)
    prof = _profile({"period_s": p.period_s, "snr": p.snr, "reduced_chi_square": p.reduced_chi_square,
                     "bins": p.bins, "counts": p.counts})
    out = {"dm": a.dm, "input": a.file, "profile": prof, "searched": False}  # json! keys, sorted
    report.write(os.path.join(a.out, "fold.json"), out)
    return out


def cmd_dedisperse(a):
    # crates/cli/src/main.rs:371-388
    os.makedirs(a.out, exist_ok=True)
    hdr, obs, _ = _open(a.file)
    d = dedisperse.delays(hdr.frequencies(), hdr.tsamp, a.dm)
    x = dedisperse.apply(obs, a.dm)
    x.astype("<f4").tofile(os.path.join(a.out, "series.f32"))
    out = {"delays": [int(v) for v in d], "dm": a.dm, "dtype": "little-endian float32",
           "normalization": "raw channel sum", "samples": len(x), "tsamp": hdr.tsamp}
    report.write(os.path.join(a.out, "series.json"), out)
    return out


def cmd_header(a):
    hdr, payload = fbank.read_packed(a.file)
    raw = memoryview(payload.base)[:hdr.data_offset]
    print(report.dumps(header_json(a.file, hdr, raw)))


def _birdie(s):
    # crates/cli/src/main.rs:110-118
    a, sep, b = s.partition(":")
    if not sep:
        raise argparse.ArgumentTypeError("expected HZ:HALF_WIDTH_HZ")
    try:
        f, w = float(a), float(b)
    except ValueError:
        raise argparse.ArgumentTypeError("invalid frequency or width")
    if not math.isfinite(f) or not math.isfinite(w) or f < 0 or w < 0:
        raise argparse.ArgumentTypeError("nonnegative finite birdie required")
    return (f, w)


def _dm_plan(s):
    """--dm-plan LO:HI:STEP[,LO:HI:STEP...]: dedisperse::plan over each segment,
    concatenated (a step-doubling plan); the trials must increase."""
    out, steps = [], []
    for seg in s.split(","):
        parts = seg.split(":")
        if len(parts) != 3:
            raise argparse.ArgumentTypeError("expected LO:HI:STEP segments separated by commas")
        try:
            lo, hi, st = (float(v) for v in parts)
        except ValueError:
            raise argparse.ArgumentTypeError("invalid DM plan segment %r" % seg)
        try:
            out.extend(dedisperse.plan(lo, hi, st).tolist())
        except PsrError as e:
            raise argparse.ArgumentTypeError(str(e))
        steps.append(st)
    if any(b <= a for a, b in zip(out, out[1:])):
        raise argparse.ArgumentTypeError("DM plan segments must increase")
    return np.array(out, dtype=np.float64), steps[0]


def main(argv=None):
    ap = argparse.ArgumentParser(prog="psr", description="GPU filterbank pulsar search (PulsarSr drop-in)")
    sub = ap.add_subparsers(dest="command", required=True)
    f = sub.add_parser("fold", help="Fold a specified period/DM, without searching or declaring a detection.")
    f.add_argument("file")
    f.add_argument("--period", type=float, required=True)
    f.add_argument("--dm", type=float, required=True)
    f.add_argument("--bins", type=int, default=64)
    f.add_argument("--out", default="output_fold")
    h = sub.add_parser("header", help="Print parsed SIGPROC header as JSON.")
    h.add_argument("file")
    d = sub.add_parser("dedisperse", help="Dedisperse one DM, exporting little-endian f32 time series.")
    d.add_argument("file")
    d.add_argument("--dm", type=float, required=True)
    d.add_argument("--out", default="output")
    s = sub.add_parser("search", help="Search trial DMs, sift candidates, refine and fold the strongest candidates.")
    s.add_argument("file")
    s.add_argument("--dm-min", type=float, default=0.0)
    s.add_argument("--dm-max", type=float, default=500.0)
    s.add_argument("--dm-step", type=float, default=None)
    s.add_argument("--min-hz", type=float, default=15.0)
    s.add_argument("--max-hz", type=float, default=2000.0)
    s.add_argument("--threshold", type=float, default=8.0)
    s.add_argument("--top", type=int, default=5)
    s.add_argument("--bins", type=int, default=64)
    s.add_argument("--no-rfi", action="store_true")
    s.add_argument("--birdie", type=_birdie, action="append", default=[])
    s.add_argument("--max-samples", type=int, default=None)
    s.add_argument("--threads", type=int, default=None)
    s.add_argument("--out", default="output")
    # knobs the reference hard-codes (16 harmonics, FP32, one uniform plan)
    s.add_argument("--max-harmonics", type=int, choices=[1, 2, 4, 8, 16], default=16,
                   help="highest harmonic sum (the reference always sums 1, 2, 4, 8 and 16)")
    s.add_argument("--precision", choices=["fp32", "fp16"], default="fp32",
                   help="storage of dedispersed series and spectra (fp16: per-trial scaled half precision)")
    s.add_argument("--dm-plan", type=_dm_plan, default=None, metavar="LO:HI:STEP[,...]",
                   help="segmented DM plan (e.g. step doubling); overrides --dm-min/--dm-max/--dm-step")
    a = ap.parse_args(argv)
    if a.command == "search":
        a.dm_plan, a.dm_plan_step = a.dm_plan if a.dm_plan is not None else (None, None)
    try:
        return {"search": cmd_search, "fold": cmd_fold, "dedisperse": cmd_dedisperse, "header": cmd_header}[a.command](a)
    except PsrError as e:
        raise SystemExit("Error: %s" % e)


if __name__ == "__main__":
    main()
    sys.exit(0)
