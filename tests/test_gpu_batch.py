"""Many observations in one device schedule (psr_batch) against psr_search and the oracle.

- With the default shapes every observation's candidate lists equal, bit for
  bit, what psr_search gives that observation alone (same kernels, same
  arithmetic), including when the memory budget splits the batch into many
  rounds that mix observations.
- The captured hipGraph, replayed several times with other GPU work between
  the replays, gives the uncaptured schedule's results every time.
- Autotuned shapes (FFT split, FFT round size, dedispersion tile span and
  workgroups per CU) change only the FFT's rounding: candidate lists match the
  oracle's per-trial loop (crates/cli/src/main.rs:462-469) on sampled trials
  within the FP32 parity bar, and FP16 storage matches the all-f64 oracle.
- The full batching workload (256 observations of uneven size, generated on
  the GPU) against the oracle on a sample of observations and trials.
"""
import os

import numpy as np
import pytest

import oracle
from matching import match_fp16_vs_f64, match_lists
from pulsarsr_amd import PRECISION_FP16, PRECISION_FP32, dedisperse, search, synth

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))
SCORE_RTOL = 1e-4


def discovery(w):
    d = oracle.delays(w.frequencies(), w.tsamp, float(w.dms[-1]))
    return (w.nsamples - 2 * int(d.max())) // 2


def small_batch(count=12, device=False):
    ws = synth.batch_workloads(count=count, seed=777, log2_len=(15, 17.5), trials=(50, 200))
    obs, pays = [], []
    for w in ws:
        p = synth.generate(w, threads=THREADS, device=device)
        pays.append(p)
        obs.append(dedisperse.Observation.from_packed(p, w.nbits, w.nchans, w.tsamp, w.fch1, w.foff))
    return ws, obs, pays


def cfg(precision=PRECISION_FP32):
    return search.Config(min_hz=10.0, max_hz=2000.0, threshold=8.0, max_candidates=32, max_harmonics=16,
                         precision=precision)


def as_tuples(lists):
    return [[(c.frequency_hz, c.period_s, c.dm, c.harmonics, c.score) for c in l] for l in lists]


def test_batch_equals_per_observation_search(ctx):
    ws, obs, _ = small_batch()
    lengths = [discovery(w) for w in ws]
    print("batch of %d observations, %d trials" % (len(ws), sum(len(w.dms) for w in ws)), flush=True)
    for precision in (PRECISION_FP32, PRECISION_FP16):
        # a 192 MiB budget cuts the batch into many rounds mixing observations
        b = search.Batch(obs, [w.dms for w in ws], cfg(precision), lengths=lengths, graph=False,
                         memory_budget=192 << 20)
        info = b.info()
        print("rounds %d, classes %d" % (info["rounds"], info["classes"]), flush=True)
        assert info["rounds"] > info["classes"] and any(r["segments"] > 1 for r in info["round_shapes"])
        got = b.run()
        for i, (w, o) in enumerate(zip(ws, obs)):
            want = search.search(o, w.dms, lengths[i], cfg(precision))
            assert as_tuples(got[i]) == as_tuples(want), (precision, w.name)
        b.close()


def test_batch_graph_replays(ctx):
    ws, obs, _ = small_batch(count=8)
    lengths = [discovery(w) for w in ws]
    eager = search.Batch(obs, [w.dms for w in ws], cfg(PRECISION_FP16), lengths=lengths, graph=False)
    want = as_tuples([l for per in eager.run() for l in per])
    eager.close()
    g = search.Batch(obs, [w.dms for w in ws], cfg(PRECISION_FP16), lengths=lengths, graph=True)
    assert g.info()["graph"] is True
    for rep in range(3):
        got = as_tuples([l for per in g.run() for l in per])
        assert got == want, "replay %d differs" % rep
        # eager GPU work between replays
        search.search(obs[0], ws[0].dms[:7], lengths[0], cfg(PRECISION_FP32))
    g.close()


def check_against_oracle(ws, obs, pays, b, precision, per_obs_trials=8, obs_idx=None):
    got = b.run()
    worst, matched, total, flat = 0.0, 0, 0, 0
    for i in (obs_idx if obs_idx is not None else range(len(ws))):
        w = ws[i]
        n = discovery(w)
        data, ns = oracle.decode(pays[i], w.nbits, w.nchans)
        sel = sorted(set(np.linspace(0, len(w.dms) - 1, per_obs_trials).round().astype(int).tolist()))
        best = max(range(len(w.dms)), key=lambda t: max([c.score for c in got[i][t]] or [0.0]))
        sel = sorted(set(sel) | {best})
        ocfg = dict(min_hz=10.0, max_hz=2000.0, threshold=8.0, max_candidates=32, max_harmonics=16)
        if precision == PRECISION_FP32:
            ref = oracle.search_trials(data, ns, w.frequencies(), w.tsamp, w.dms[sel], n, threads=THREADS, **ocfg)
            worst = max(worst, match_lists([got[i][t] for t in sel], ref, 8.0))
        else:
            ref = oracle.search_trials(data, ns, w.frequencies(), w.tsamp, w.dms[sel], n, threads=THREADS, f64=True,
                                       **ocfg)
            nfft = 2 * (1 << (n - 1).bit_length())
            m = match_fp16_vs_f64([got[i][t] for t in sel], ref, 8.0, 1.0 / (nfft * w.tsamp), 1.0 / (n * w.tsamp))
            worst = max(worst, m["max_rel_snr_error_sharp"])
            print("  %s: %s" % (w.name, m), flush=True)
            assert m["max_rel_snr_error"] <= 5e-2
            matched += m["matched"]
            total += m["total"]
            flat += m["flat_peaks"]
    assert flat <= max(2, 0.01 * matched), "%d flat-peak candidates of %d" % (flat, matched)
    return worst, matched, total


def test_batch_autotuned_matches_oracle(ctx):
    ws, obs, pays = small_batch(count=10)
    lengths = [discovery(w) for w in ws]
    b = search.Batch(obs, [w.dms for w in ws], cfg(PRECISION_FP32), lengths=lengths, graph=True, autotune=True)
    info = b.info()
    print("autotune: %s" % info.get("autotune"), flush=True)
    assert info["graph"] and "autotune" in info
    worst, _, _ = check_against_oracle(ws, obs, pays, b, PRECISION_FP32)
    assert worst < SCORE_RTOL
    b.close()
    b = search.Batch(obs, [w.dms for w in ws], cfg(PRECISION_FP16), lengths=lengths, graph=True, autotune=True)
    worst, matched, total = check_against_oracle(ws, obs, pays, b, PRECISION_FP16)
    print("fp16 vs f64: matched %d of %d, worst SNR error %.3g" % (matched, total, worst), flush=True)
    assert total > 0 and matched >= 0.9 * total and worst <= 1e-2
    b.close()


def test_batching_workload_full(ctx):
    """256 observations: channels 64-1024, 2/4/8-bit, 2^16-2^21 samples, 50-2000
    trials, FP16 storage, tuned and captured; 12 observations checked."""
    ws = synth.batch_workloads()
    obs, keep = [], {}
    check = set(np.linspace(0, len(ws) - 1, 12).round().astype(int).tolist())
    for i, w in enumerate(ws):
        p = synth.generate(w, device=True)
        obs.append(dedisperse.Observation.from_packed(p, w.nbits, w.nchans, w.tsamp, w.fch1, w.foff))
        if i in check:
            keep[i] = p
    print("resident: %d observations, %d trials" % (len(ws), sum(len(w.dms) for w in ws)), flush=True)
    lengths = [discovery(w) for w in ws]
    b = search.Batch(obs, [w.dms for w in ws], cfg(PRECISION_FP16), lengths=lengths, graph=True, autotune=True)
    pays = [keep.get(i) for i in range(len(ws))]
    worst, matched, total = check_against_oracle(ws, obs, pays, b, PRECISION_FP16, per_obs_trials=6,
                                                 obs_idx=sorted(check))
    info = b.info()
    print("batch run %.1f ms, rounds %d; fp16 vs f64: matched %d of %d, worst %.3g"
          % (info["last_ms"], info["rounds"], matched, total, worst), flush=True)
    assert total > 0 and matched >= 0.9 * total and worst <= 1e-2
    b.close()
    for o in obs:
        o.close()
