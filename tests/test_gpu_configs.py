"""The benchmark workloads end to end on the GPU, at full size, against the oracle.

Each test generates one named workload (pulsarsr_amd.synth.workload), makes its
packed filterbank resident on the GPU, searches every DM trial through
psr_search, and checks the result against the oracle's per-trial loop (the
restated crates/cli/src/main.rs:462-469) on the same decoded samples:

- plain-64ch-8bit (64 ch x 8-bit, 2^17 samples, 61 trials, RFI mask, 8 harmonics,
  top 2): the `psr search` drop-in on a SIGPROC file of the workload -- mask,
  every trial's candidates and the confirmation report against the oracle's
  pipeline;
- lband-96ch-4bit (96 ch x 4-bit, 2^19 samples, 2000 trials, 16 harmonics): all
  2000 trials' candidate lists, series bit-exact on every 8th trial;
- precision-1024ch-8bit (2^21 samples, 3000 step-doubling trials) and
  headline-1024ch-8bit (2^22 samples, 4000 trials): the full search of every
  trial in FP32 and in FP16 storage; on 64 evenly spaced trials, every
  top-family leader's trial and a sample of the families' member trials (at
  most 160 trials, so the oracle's CPU loop finishes in about a minute): series
  bit-exact, FP32 lists against the oracle, FP16 lists against the all-f64
  oracle (match counts asserted, SNR within 1e-2), and the FP16 spectral power
  against the f64 spectrum.

The FP32 and FP16 top-20 family leaders (confirmation::group, the report's
candidate list) must also be identical in DM and harmonics.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle
from matching import match_fp16_vs_f64, match_lists
from pulsarsr_amd import PRECISION_FP16, PRECISION_FP32, cli, confirmation, dedisperse, rfi, search, synth

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))
SCORE_RTOL = 1e-4


def discovery_length(w, dms=None):
    d = oracle.delays(w.frequencies(), w.tsamp, float((w.dms if dms is None else dms)[-1]))
    return (w.nsamples - 2 * int(d.max())) // 2  # crates/cli/src/main.rs:433-446


def prepare(name):
    w = synth.workload(name)
    print("workload %s: %d ch x %d-bit, %d samples, %d trials" % (name, w.nchans, w.nbits, w.nsamples,
                                                                   len(w.dms)), flush=True)
    payload = synth.generate(w, threads=THREADS)
    obs = dedisperse.Observation.from_packed(payload, w.nbits, w.nchans, w.tsamp, w.fch1, w.foff)
    data, ns = oracle.decode(payload, w.nbits, w.nchans)
    assert ns == w.nsamples
    return w, payload, obs, data


def config_of(w, precision):
    return search.Config(min_hz=w.min_hz, max_hz=w.max_hz, threshold=w.threshold, max_candidates=32,
                         max_harmonics=w.max_harmonics, precision=precision)


def oracle_cfg(w):
    return dict(min_hz=w.min_hz, max_hz=w.max_hz, threshold=w.threshold, max_candidates=32,
                max_harmonics=w.max_harmonics)


def top_families(lists, n, tsamp, k=20):
    cands = [c for lst in lists for c in lst]
    leaders, members = confirmation.group(cands, 1.5 / (n * tsamp), k)
    trial_of = []
    for t, lst in enumerate(lists):
        trial_of.extend([t] * len(lst))
    return leaders, members, trial_of


def sample_trials(ntrials, lists, n, tsamp, limit=160):
    """64 evenly spaced trials, every top-20 leader's trial, then member trials
    of the families spread evenly, up to `limit` trials in all."""
    leaders, members, trial_of = top_families(lists, n, tsamp)
    lead_dm = {(round(l["dm"], 9), round(l["frequency_hz"], 9)) for l in leaders}
    sel = set(np.linspace(0, ntrials - 1, 64).round().astype(int).tolist())
    cands = [c for lst in lists for c in lst]
    for fam, idx in members:
        c = cands[idx]
        if (round(c.dm, 9), round(c.frequency_hz, 9)) in lead_dm:
            sel.add(trial_of[idx])
    rest = sorted({trial_of[idx] for _, idx in members} - sel)
    room = max(0, limit - len(sel))
    if rest and room:
        sel.update(rest[i] for i in np.linspace(0, len(rest) - 1, min(room, len(rest))).round().astype(int))
    return np.array(sorted(sel)), leaders


def check_large(name):
    w, payload, obs, data = prepare(name)
    n = discovery_length(w)
    freqs = w.frequencies()
    got32 = search.search(obs, w.dms, n, config_of(w, PRECISION_FP32))
    st32 = obs.ctx.stage_times()
    got16 = search.search(obs, w.dms, n, config_of(w, PRECISION_FP16))
    st16 = obs.ctx.stage_times()
    print("%s: series %d, fp32 %.1f ms, fp16 %.1f ms" % (name, n, st32["total_ms"], st16["total_ms"]), flush=True)
    sel, leaders32 = sample_trials(len(w.dms), got32, n, w.tsamp)
    dms = w.dms[sel]
    print("checking %d trials against the oracle" % len(sel), flush=True)
    # series bit-exact (the GPU's own dedispersion of the sampled trials)
    ours = dedisperse.apply_dms(obs, dms, n)
    ref_series = np.empty((len(sel), n), np.float32)
    ref32 = oracle.search_trials(data, w.nsamples, freqs, w.tsamp, dms, n, threads=THREADS, series=ref_series,
                                 **oracle_cfg(w))
    bad = [float(dms[i]) for i in range(len(sel)) if not (ours[i].view(np.uint32) == ref_series[i].view(np.uint32)).all()]
    assert not bad, "series differ at DMs %r" % bad[:8]
    # FP32 candidate lists against the oracle
    worst32 = match_lists([got32[t] for t in sel], ref32, w.threshold)
    assert sum(len(x) for x in ref32) > len(sel)
    assert worst32 < SCORE_RTOL
    # FP16 candidate lists against the all-f64 oracle
    ref64 = oracle.search_trials(data, w.nsamples, freqs, w.tsamp, dms, n, threads=THREADS, f64=True,
                                 **oracle_cfg(w))
    nfft = 2 * (1 << (n - 1).bit_length())
    m = match_fp16_vs_f64([got16[t] for t in sel], ref64, w.threshold, 1.0 / (nfft * w.tsamp), 1.0 / (n * w.tsamp))
    print("fp16 vs f64 oracle: %s" % json.dumps(m), flush=True)
    assert m["total"] >= len(sel) and m["matched"] >= 0.9 * m["total"]
    assert m["top_matched"] >= 0.9 * m["trials"]
    assert m["max_rel_snr_error_sharp"] <= 1e-2 and m["max_rel_snr_error"] <= 5e-2
    assert m["flat_peaks"] <= max(2, 0.01 * m["matched"])
    # FP16 spectral power against the f64 spectrum on four trials (the strongest leaders' and two spaced)
    lead_trials = [int(np.nonzero(w.dms == l["dm"])[0][0]) for l in leaders32[:2]]
    worst_norm = 0.0
    for t in lead_trials + [int(sel[0]), int(sel[len(sel) // 2])]:
        i = int(np.nonzero(sel == t)[0][0]) if t in sel else None
        x = ref_series[i] if i is not None else dedisperse.apply_dms(obs, [w.dms[t]], n)[0]
        # FP16 storage keeps amplitudes |X| / sqrt(n): the power it searches is n times smaller
        raw16, _ = search.spectrum(x, w.tsamp, search.Config(min_hz=w.min_hz, precision=PRECISION_FP16))
        raw64, _ = oracle.spectrum_f64(x, w.tsamp, min_hz=w.min_hz)
        p16 = raw16.astype(np.float64) * n
        norm = float(np.linalg.norm(p16 - raw64) / np.linalg.norm(raw64))
        worst_norm = max(worst_norm, norm)
    print("fp16 spectral power vs f64: normwise %.3g" % worst_norm, flush=True)
    assert worst_norm <= 5e-3
    # the report's candidate list: top-20 family leaders, FP16 against FP32
    lead16, _, _ = top_families(got16, n, w.tsamp)
    same = compare_leaders(lead16, leaders32, 1.5 / (n * w.tsamp))
    print("top-20 leaders identical fp16 vs fp32: %d of %d" % (same, len(leaders32)), flush=True)
    obs.close()
    return same, len(leaders32)


def compare_leaders(l16, l32, res, tie=1e-2):
    """Rank by rank, the FP16 and FP32 family leaders agree in DM, harmonics and
    frequency (within the grouping resolution), except where two families' scores
    are within `tie` of each other: FP16 rounding may then swap their order, or
    which of them takes the last place.  Returns the number of identical ranks."""
    assert len(l16) == len(l32) > 0
    same = 0
    for i, (a, b) in enumerate(zip(l16, l32)):
        if a["dm"] == b["dm"] and a["harmonics"] == b["harmonics"] and abs(a["frequency_hz"] - b["frequency_hz"]) <= res:
            same += 1
            continue
        assert abs(a["score"] - b["score"]) <= tie * abs(b["score"]), ("leaders differ beyond a near tie", i, a, b)
    return same


def test_precision_workload(ctx):
    check_large("precision-1024ch-8bit")


def test_headline_workload(ctx):
    same, total = check_large("headline-1024ch-8bit")
    assert same == total  # the headline's top-20 list is identical in FP16 and FP32


def test_lband_workload_all_trials(ctx):
    w, payload, obs, data = prepare("lband-96ch-4bit")
    n = discovery_length(w)
    got = search.search(obs, w.dms, n, config_of(w, PRECISION_FP32))
    ref = oracle.search_trials(data, w.nsamples, w.frequencies(), w.tsamp, w.dms, n, threads=THREADS,
                               **oracle_cfg(w))
    assert sum(len(x) for x in ref) > len(w.dms)
    assert match_lists(got, ref, w.threshold) < SCORE_RTOL
    every = w.dms[::8]
    ours = dedisperse.apply_dms(obs, every, n)
    ref_series = np.empty((len(every), n), np.float32)
    oracle.search_trials(data, w.nsamples, w.frequencies(), w.tsamp, every, n, threads=THREADS, series=ref_series,
                         **oracle_cfg(w))
    assert (ours.view(np.uint32) == ref_series.view(np.uint32)).all()
    # the injected 4.62 ms pulsar leads the families
    leaders, _, _ = top_families(got, n, w.tsamp)
    f0 = 1.0 / 0.00462
    assert any(abs(l["frequency_hz"] / f0 - round(l["frequency_hz"] / f0)) < 0.01 for l in leaders[:3])
    obs.close()


def test_plain_workload_psr_search(ctx, tmp_path):
    """psr search on the plain workload's SIGPROC file with its settings
    (--dm-max 120 --dm-step 2 --min-hz 10 --top 2 --max-harmonics 8, mask on)."""
    from test_gpu_cli import write_fil
    w = synth.workload("plain-64ch-8bit")
    payload = synth.generate(w, threads=THREADS)
    path = str(tmp_path / "plain.fil")
    write_fil(path, w, payload)
    out = str(tmp_path / "out")
    rep = cli.main(["search", path, "--dm-max", "120", "--dm-step", "2", "--min-hz", "10", "--top", "2",
                    "--max-harmonics", "8", "--out", out])
    written = json.load(open(os.path.join(out, "report.json")))
    data, ns = oracle.decode(payload, w.nbits, w.nchans)
    block = max(int(0.1 / w.tsamp), 32)
    chans, blocks, frac = oracle.mask(data, ns, block, 8.0)
    assert written["rfi"] == {"channels": chans, "blocks": blocks, "block_samples": block, "masked_fraction": frac}
    trials = oracle.plan(0.0, 120.0, 2.0)
    assert written["trial_count"] == len(trials) == 61
    n = discovery_length(w, trials)
    ocfg = dict(min_hz=10.0, max_hz=2000.0, threshold=8.0, max_candidates=32, max_harmonics=8)
    per = oracle.search_trials(data, ns, w.frequencies(), w.tsamp, trials, n, threads=THREADS, **ocfg)
    # every trial's list, through the GPU (same masked observation)
    cfg = search.Config(min_hz=10.0, max_hz=2000.0, threshold=8.0, max_candidates=32, max_harmonics=8)
    obs = dedisperse.Observation.from_packed(payload, w.nbits, w.nchans, w.tsamp, w.fch1, w.foff)
    gm = rfi.mask(obs, block, 8.0)
    assert gm["channels"] == chans and gm["blocks"] == blocks
    got = search.search(obs, trials, n, cfg)
    assert sum(len(x) for x in per) > 0
    assert match_lists(got, per, 8.0) < SCORE_RTOL
    # the report: members are oracle candidates, and analyze over them is the oracle's
    conf = written["confirmation"]
    assert conf["discovery_samples"] == n
    members = [m for c in conf["candidates"] for m in c["harmonic_members"]]
    by_dm = {round(float(trials[t]), 9): per[t] for t in range(len(trials))}
    for m in members:
        hit = [c for c in by_dm[round(m["dm"], 9)] if c["harmonics"] == m["harmonics"]
               and abs(c["frequency_hz"] - m["frequency_hz"]) <= 1e-6 * m["frequency_hz"]]
        assert hit and hit[0]["score"] == pytest.approx(m["score"], rel=SCORE_RTOL)
    gen = sorted(members, key=lambda m: (m["dm"], -m["score"]))
    want = oracle.analyze(data, ns, w.frequencies(), w.tsamp, trials, gen, 2, 64)
    assert len(conf["candidates"]) == len(want["candidates"]) > 0
    for g, o in zip(conf["candidates"], want["candidates"]):
        for k in ("candidate", "refined_dm", "window", "validation", "zero_dm_validation", "dm_contrast",
                  "time_parts", "subbands", "accepted"):
            assert g[k] == o[k], k
    # the injected 32 ms pulsar at DM 62 is the leading family
    best = conf["candidates"][0]["candidate"]
    assert abs(best["frequency_hz"] * 0.032 - round(best["frequency_hz"] * 0.032)) < 0.02
    assert rep["trial_count"] == 61
    obs.close()
