"""Pins the oracle (oracle/oracle.c) against the reference's golden vectors.

Every oracle function is checked against at least one independent source:
known-answer tests from the reference's own unit tests, the reference's saved
run outputs (tests/golden/*.json, extracted by tests/golden/make_fixtures.py),
closed forms (numpy's FP64 FFT), or invariants.
"""
import json
import math
import os
import struct

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, reference_file


def golden(name):
    return json.load(open(os.path.join(GOLDEN, name)))


# ---------------------------------------------------------------- dedisperse
def test_delays_known_answer():
    # crates/dedisperse/src/lib.rs:111
    assert oracle.delays([1000.0, 2000.0], 0.001, 1.0).tolist() == [3, 0]


def test_delays_invalid_inputs():
    # crates/dedisperse/src/lib.rs:112-114
    with pytest.raises(ValueError):
        oracle.plan(0.0, 5.0, 0.0)
    with pytest.raises(ValueError):
        oracle.delays([0.0], 1.0, 1.0)
    with pytest.raises(ValueError):
        oracle.apply(np.array([1.0, 2.0], np.float32), 2, [3])


def test_dm62_delay_table_matches_saved_run():
    # output_dm62/series.json: 96 GBT channel delays at DM 62 and the output length
    g = golden("dm62_delays.json")
    h = golden("headers.json")["gbt_header"]
    freqs = h["fch1"] + np.arange(h["nchans"]) * h["foff"]
    d = oracle.delays(freqs, g["tsamp"], g["dm"])
    assert d.tolist() == g["delays"]
    nsamples = (h["file_bytes"] - h["data_offset"]) * 8 // (h["nchans"] * h["nbits"])
    assert nsamples - int(d.max()) == g["samples"]


def test_dm62_rounding_rule_is_half_away_from_zero():
    # a plausible mistake: banker's rounding.  It changes at least one entry of
    # the saved table, so the golden check above would catch it.
    g = golden("dm62_delays.json")
    h = golden("headers.json")["gbt_header"]
    freqs = h["fch1"] + np.arange(h["nchans"]) * h["foff"]
    fmax = freqs.max()
    x = 4148.808 * g["dm"] * (1.0 / (freqs * freqs) - 1.0 / (fmax * fmax)) / g["tsamp"]
    assert np.floor(x + 0.5).astype(int).tolist() == g["delays"]
    truncated = np.floor(x).astype(int).tolist()
    assert truncated != g["delays"]


def test_dm_step_and_plan_match_gbt_report():
    # output_confirm_gbt/report.json: dm_step 0.25005711444624, 2000 trials over 0..500
    g = golden("gbt_report.json")["settings"]
    h = golden("headers.json")["gbt_header"]
    freqs = h["fch1"] + np.arange(h["nchans"]) * h["foff"]
    step = oracle.dm_step(freqs, h["tsamp"])
    assert step == g["dm_step"]
    assert len(oracle.plan(g["dm_min"], g["dm_max"], step)) == g["trial_count"]
    # the maximum is inclusive: plan(0, 500, 0.25) has 2001 trials
    assert len(oracle.plan(0.0, 500.0, 0.25)) == 2001


def test_dispersed_delta_collapses_for_both_frequency_orders():
    # crates/dedisperse/src/lib.rs:95-108
    for freqs in ([1400.0, 1300.0, 1200.0], [1200.0, 1300.0, 1400.0]):
        d = oracle.delays(freqs, 0.001, 20.0)
        n = 100
        data = np.zeros(n * 3, np.float32)
        for c, dl in enumerate(d):
            data[c * n + 10 + int(dl)] = 1.0
        out = oracle.apply(data, n, d)
        assert out[10] == 3.0 and out.sum() == 3.0


def test_apply_is_channel_ordered_f32_sum():
    # non-integer data: the sum must be accumulated in channel order in f32
    rng = np.random.default_rng(1)
    nch, n = 7, 50
    data = (rng.standard_normal(nch * n) * 1e3).astype(np.float32)
    d = np.array([0, 1, 2, 3, 4, 5, 6], np.uint64)
    out = oracle.apply(data, n, d, 40)
    ref = np.zeros(40, np.float32)
    for c in range(nch):
        ref = (ref + data[c * n + c:c * n + c + 40]).astype(np.float32)
    assert (out.view(np.uint32) == ref.view(np.uint32)).all()


# --------------------------------------------------------------------- fbank
def _fixture(bits, payload):
    # crates/fbank/src/lib.rs:295-313 (header built the same way)
    def put(s):
        return struct.pack("<i", len(s)) + s.encode()
    b = put("HEADER_START")
    for k, v in (("nchans", 2), ("nbits", bits), ("nifs", 1), ("data_type", 1)):
        b += put(k) + struct.pack("<i", v)
    for k, v in (("tsamp", 0.001), ("fch1", 1400.0), ("foff", -1.0)):
        b += put(k) + struct.pack("<d", v)
    return b + put("HEADER_END") + bytes(payload)


BIT_DEPTH_CASES = [  # crates/fbank/src/lib.rs:317-327
    (1, [0b10100110], [0., 1., 0., 0., 1., 0., 1., 1.]),
    (2, [0b11100100], [0., 2., 1., 3.]),
    (4, [0x21, 0x43], [1., 3., 2., 4.]),
    (8, [1, 2, 3, 4], [1., 3., 2., 4.]),
    (16, [1, 0, 2, 1, 3, 0, 4, 2], [1., 3., 258., 516.]),
    (32, [0, 0, 128, 63, 0, 0, 0, 64, 0, 0, 64, 64, 0, 0, 128, 64], [1., 3., 2., 4.]),
]


@pytest.mark.parametrize("bits,payload,expected", BIT_DEPTH_CASES)
def test_decode_bit_depth_fixtures(bits, payload, expected):
    data, ns = oracle.decode(bytes(payload), bits, 2)
    assert data.tolist() == expected
    assert oracle.decode(bytes(payload), bits, 2, limit=2)[1] == 2


def test_decode_rejects_nonfinite_and_partial_frames():
    # crates/fbank/src/lib.rs:333-340
    with pytest.raises(ValueError):
        oracle.decode(bytes([0, 0, 128, 127] + [0] * 12), 32, 2)
    with pytest.raises(ValueError):
        oracle.decode(bytes([1, 0, 2]), 16, 2)


# ----------------------------------------------------------------------- rfi
def test_median_is_upper_median():
    # crates/rfi/src/lib.rs:4-12: v[len/2] of the sorted slice
    assert oracle.median([3.0, 1.0, 2.0, 4.0]) == 3.0
    assert oracle.median([5.0, 1.0, 3.0]) == 3.0
    assert oracle.median([]) == 0.0
    rng = np.random.default_rng(3)
    for n in (1, 2, 63, 64, 320, 1001):
        v = rng.exponential(size=n).astype(np.float32)
        assert oracle.median(v) == np.sort(v)[n // 2]


def test_modest_broadband_burst_on_high_baseline_is_masked():
    # crates/rfi/src/lib.rs:100-114 (multiplicative-hash data)
    n, chans = 4000, 16
    i = np.arange(n * chans, dtype=np.uint64)
    data = (100.0 + ((i * np.uint64(2654435761)) % np.uint64(1000)).astype(np.float32) / 500.0 - 1.0).astype(np.float32)
    for c in range(chans):
        data[c * n + 1600:c * n + 1680] += 30.0
    _, blocks, _ = oracle.mask(data, n, 400, 8.0)
    assert 4 in blocks


def test_planted_channel_and_block_are_masked():
    # crates/rfi/src/lib.rs:115-129
    n = 1000
    data = (10.0 + (np.arange(n * 16) % 7)).astype(np.float32)
    data[:n] = 1000.0
    for c in range(1, 16):
        data[c * n + 400:c * n + 450] += 500.0
    chans, blocks, frac = oracle.mask(data, n, 50, 6.0)
    assert 0 in chans and 8 in blocks and frac > 0.05


# ---------------------------------------------------------------------- FFT
@pytest.mark.parametrize("N", [4, 16, 64, 1024, 1 << 14])
def test_rfft64_matches_numpy(N):
    x = np.random.default_rng(N).standard_normal(N)
    ref = np.fft.rfft(x)
    got = oracle.rfft64(x)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref)) * math.log2(N)


# ------------------------------------------------------------------- search
def _sine_series():
    tsamp, f = 0.001, 31.25
    return np.sin(2 * np.pi * f * np.arange(8192) * tsamp).astype(np.float32), tsamp, f


def test_finds_sine_and_respects_birdies():
    # crates/search/src/lib.rs:180-203
    x, tsamp, f = _sine_series()
    c = oracle.periodicity(x, tsamp, 0.0, min_hz=10.0)
    assert abs(c[0]["frequency_hz"] - f) < 0.1
    c = oracle.periodicity(x, tsamp, 0.0, min_hz=10.0, birdies=[(f, 1.0)])
    assert all(abs(k["frequency_hz"] - f) > 1.0 for k in c)


def xorshift_pulse_train():
    # crates/search/src/lib.rs:228-247 (xorshift64, seed 123456789)
    dt, period = 0.00025, 0.02371
    state = 123456789
    M = (1 << 64) - 1
    out = np.zeros(16384, np.float32)
    for i in range(16384):
        phase = (i * dt / period) % 1.0
        distance = abs(phase - 0.25)
        state ^= (state << 13) & M
        state ^= state >> 7
        state ^= (state << 17) & M
        noise = np.float32(state / float(M) - 0.5)
        out[i] = np.float32(math.exp(-0.5 * (distance / 0.013) ** 2)) + noise * np.float32(0.4)
    return out, dt, period


def test_off_bin_pulse_train_retains_fundamental():
    x, dt, period = xorshift_pulse_train()
    c = oracle.periodicity(x, dt, 18.0)
    kept = oracle.sift(c, 1.5 / (len(x) * dt), 10, False)
    assert abs(kept[0]["frequency_hz"] - 1.0 / period) < 0.2


def test_sifts_dm_duplicates_and_harmonics():
    # crates/search/src/lib.rs:204-226
    def mk(f, dm, s):
        return dict(frequency_hz=f, period_s=1 / f, dm=dm, harmonics=1, score=s)
    c = oracle.sift([mk(100.0, 20.0, 10.0), mk(200.0, 21.0, 8.0), mk(100.01, 19.0, 9.0), mk(137.0, 0.0, 20.0)],
                    0.1, 10, True)
    assert len(c) == 1 and c[0]["dm"] == 20.0


def test_whitening_is_scale_invariant():
    x, tsamp, _ = _sine_series()
    x = x + np.random.default_rng(5).standard_normal(len(x)).astype(np.float32)
    _, w1 = oracle.spectrum(x, tsamp)
    _, w2 = oracle.spectrum(x * np.float32(4.0), tsamp)
    assert np.allclose(w1, w2, rtol=1e-5, atol=1e-6)


def _match_members(per_trial, members, step):
    by = {}
    for lst in per_trial:
        if lst:
            by[lst[0]["dm"]] = lst
    worst, checked = 0.0, 0
    for f, dm, h, score in members:
        if dm not in by:
            continue
        hit = [c for c in by[dm] if c["harmonics"] == h and abs(c["frequency_hz"] - f) < 1e-5 * max(1.0, f)]
        assert hit, "golden member (%r, dm %r, h %r, score %r) missing" % (f, dm, h, score)
        worst = max(worst, abs(hit[0]["score"] - score) / abs(score))
        checked += 1
    return worst, checked


def test_periodicity_reproduces_synthetic_report_members():
    # output_confirm_synth/report.json: all 549 per-trial family members (61 trials)
    path = reference_file("synthetic.fil")
    g = golden("synthetic_report.json")
    s = g["settings"]
    raw = open(path, "rb").read()
    off = golden("headers.json")["synthetic_header"]["data_offset"]
    data, ns = oracle.decode(raw[off:], 32, 32)
    freqs = 1450.0 - 4.0 * np.arange(32)
    dms = oracle.plan(s["dm_min"], s["dm_max"], s["dm_step"])
    res = oracle.search_trials(data, ns, freqs, s["tsamp"], dms, s["discovery_samples"] - 0, threads=4,
                               min_hz=s["min_hz"], max_hz=s["max_hz"], threshold=s["threshold"])
    worst, checked = _match_members(res, g["members"], s["dm_step"])
    assert checked == len(g["members"]) == 549
    assert worst < 1e-5


def test_rfi_mask_and_members_reproduce_gbt_report():
    # output_confirm_gbt/report.json: RFI mask and a sample of the 4,682 members
    path = reference_file("GBT_Lband_PSR.fil")
    g = golden("gbt_report.json")
    s = g["settings"]
    h = golden("headers.json")["gbt_header"]
    raw = np.fromfile(path, dtype=np.uint8)
    data, ns = oracle.decode(raw[h["data_offset"]:], h["nbits"], h["nchans"])
    chans, blocks, frac = oracle.mask(data, ns, int(max(0.1 / s["tsamp"], 32)), 8.0)
    assert chans == g["rfi"]["channels"] and blocks == g["rfi"]["blocks"]
    assert frac == g["rfi"]["masked_fraction"]
    freqs = h["fch1"] + np.arange(h["nchans"]) * h["foff"]
    dms = oracle.plan(s["dm_min"], s["dm_max"], s["dm_step"])
    d = oracle.delays(freqs, s["tsamp"], s["dm_max"])
    disc = (ns - 2 * int(d.max())) // 2  # crates/cli/src/main.rs:433-446
    trials = sorted(set(m[1] for m in g["members"]))[::40]
    idx = [int(round(t / s["dm_step"])) for t in trials]
    res = oracle.search_trials(data, ns, freqs, s["tsamp"], dms[idx], disc, threads=4,
                               min_hz=s["min_hz"], max_hz=s["max_hz"], threshold=s["threshold"])
    worst, checked = _match_members(res, g["members"], s["dm_step"])
    assert checked > 100 and worst < 1e-5


def test_fold_reproduces_saved_profile():
    # output_fold/fold.json: GBT, no RFI mask, DM 62.389, P 4.621398 ms, 64 bins
    path = reference_file("GBT_Lband_PSR.fil")
    g = golden("gbt_fold_dm62.389.json")
    h = golden("headers.json")["gbt_header"]
    raw = np.fromfile(path, dtype=np.uint8)
    data, ns = oracle.decode(raw[h["data_offset"]:], h["nbits"], h["nchans"])
    freqs = h["fch1"] + np.arange(h["nchans"]) * h["foff"]
    x = oracle.apply(data, ns, oracle.delays(freqs, h["tsamp"], g["dm"]))
    p = oracle.profile_at(x, h["tsamp"], g["period_s"], 64, 0)
    assert p["counts"].tolist() == g["counts"]
    assert (p["bins"] == np.array(g["bins"], np.float32)).all()
    assert p["snr"] == g["snr"] and p["reduced_chi_square"] == g["reduced_chi_square"]


# --------------------------------------------------------------------- fold
def test_true_period_concentrates_pulses():
    # crates/fold/src/lib.rs:101-116
    i = np.arange(20000)
    x = np.where(i % 100 < 5, 10.0, ((i * 17 % 23) - 11.0) / 10.0).astype(np.float32)
    a = oracle.profile_at(x, 0.001, 0.1, 32)
    b = oracle.profile_at(x, 0.001, 0.1037, 32)
    assert a["snr"] > b["snr"] * 2 and int(a["counts"].sum()) == len(x)


def test_refine_keeps_last_maximum():
    # Iterator::max_by returns the last of equal maxima (crates/fold/src/lib.rs:94-96):
    # a constant series gives chi2 == 0 for every trial frequency.
    x = np.ones(1000, np.float32)
    r = oracle.refine(x, 0.001, 10.0, 1.0, 4, 8)
    assert r["period_s"] == 1.0 / (10.0 - 1.0 + 2.0 * 1.0 * 4 / 4)


# ------------------------------------------------------- FP64 oracle pins
def _round_half_away(p):
    f = math.floor(p)
    return int(f) + (1 if p - f >= 0.5 else 0)


def numpy_periodicity_f64(x, tsamp, dm, min_hz=1.0, max_hz=math.inf, threshold=6.0, max_candidates=32,
                          birdies=(), max_harmonics=16):
    """crates/search/src/lib.rs:33-145 in float64, built on numpy's FFT (pocketfft)
    and numpy's sort for the upper median -- independent of the C oracle's
    transform and selection."""
    n = len(x)
    size = 2 * (1 << (n - 1).bit_length())
    xd = x.astype(np.float64)
    mean = float(np.sum(xd)) / n
    inp = np.zeros(size)
    inp[:n] = xd - mean
    X = np.fft.rfft(inp)
    raw = X.real * X.real + X.imag * X.imag
    nb = len(raw)
    floor = float(np.sum(raw)) / nb * 1e-6
    pw = raw.copy()
    for b in range(0, nb, 64):
        lo, hi = max(b - 128, 1), min(b + 192, nb)
        med = np.sort(raw[lo:hi])[(hi - lo) // 2]
        pw[b:b + 64] /= max(max(med / math.log(2.0), floor), 1e-20)
    pw[0] = 0.0
    bin_hz = 1.0 / (size * tsamp)
    f = np.arange(nb) * bin_hz
    for c, w in birdies:
        pw[np.abs(f - c) <= w] = 0.0
    first = max(math.ceil(min_hz / bin_hz), 1)
    last = min(int(math.floor(min(max_hz, 0.5 / tsamp) / bin_hz)), nb - 2)

    def offset(k):
        curv = raw[k - 1] - 2.0 * raw[k] + raw[k + 1]
        return min(max(0.5 * (raw[k - 1] - raw[k + 1]) / curv, -0.5), 0.5) if curv < 0.0 else 0.0

    cands = []
    for h in (1, 2, 4, 8, 16):
        if h > max_harmonics:
            break
        end = min(last, (nb - 2) // h)
        if end <= first + 1:
            continue
        sc = np.full(end - first + 1, -math.inf)
        for k in range(first, end + 1):
            if h > 1 and pw[k] < 4.0:
                continue
            o = offset(k)
            s = 0.0
            for j in range(1, h + 1):
                s += pw[min(_round_half_away((k + o) * j), nb - 1)]
            sc[k - first] = (s - h) / math.sqrt(h)
        for i in range(1, len(sc) - 1):
            if sc[i] >= threshold and sc[i] > sc[i - 1] and sc[i] >= sc[i + 1]:
                k = first + i
                fr = (k + offset(k)) * bin_hz
                cands.append(dict(frequency_hz=fr, period_s=1.0 / fr, dm=dm, harmonics=h, score=float(sc[i])))
    cands.sort(key=lambda c: -c["score"])  # stable, as sort_by with total_cmp
    kept = []
    for c in cands:
        if any(abs(k["frequency_hz"] - c["frequency_hz"]) < 1.0 / (n * tsamp) for k in kept):
            continue
        kept.append(c)
        if len(kept) >= max_candidates:
            break
    return kept


def test_median64_is_upper_median():
    # crates/rfi/src/lib.rs:4-12 in double: v[len/2] of the sorted slice
    rng = np.random.default_rng(64)
    assert oracle.median64([]) == 0.0
    assert oracle.median64([2.0, 1.0]) == 2.0 and oracle.median64([5.0, 1.0, 3.0]) == 3.0
    for n in (1, 2, 3, 63, 64, 255, 256, 320, 1001):
        v = rng.exponential(size=n)
        assert oracle.median64(v) == np.sort(v)[n // 2], n
    v = np.array([1.0, 1.0, 0.5, 1.0, 2.0, 2.0])  # ties
    assert oracle.median64(v) == np.sort(v)[3]


def test_spectrum_f64_matches_numpy():
    x = (np.random.default_rng(11).standard_normal(3000) * 16 + 2000).astype(np.float32)
    raw, white = oracle.spectrum_f64(x, 64e-6, min_hz=10.0, birdies=[(500.0, 20.0)])
    xd = x.astype(np.float64)
    inp = np.zeros(8192)
    inp[:3000] = xd - xd.sum() / 3000
    X = np.fft.rfft(inp)
    ref = X.real ** 2 + X.imag ** 2
    assert np.max(np.abs(raw - ref)) <= 1e-12 * ref.max()
    f = np.arange(len(raw)) / (8192 * 64e-6)
    assert white[0] == 0.0 and np.all(white[np.abs(f - 500.0) <= 20.0] == 0.0)


@pytest.mark.parametrize("case", ["pulse_train", "sine_noise_birdie", "pulse_train_h8"])
def test_periodicity_f64_matches_numpy_restatement(case):
    if case.startswith("pulse_train"):
        x, dt, _ = xorshift_pulse_train()
        kw = dict(max_harmonics=8) if case.endswith("h8") else {}
        args = (x, dt, 18.0)
    else:
        x, dt, f = _sine_series()
        x = x + np.random.default_rng(6).standard_normal(len(x)).astype(np.float32) * np.float32(0.5)
        kw = dict(min_hz=10.0, birdies=[(100.0, 2.0)], threshold=4.0)
        args = (x, dt, 0.0)
    got = oracle.periodicity(*args, f64=True, **kw)
    ref = numpy_periodicity_f64(*args, **kw)
    assert len(ref) > 3 and len(got) == len(ref)
    for g, r in zip(got, ref):
        assert g["harmonics"] == r["harmonics"]
        assert abs(g["frequency_hz"] - r["frequency_hz"]) <= 1e-9 * r["frequency_hz"]
        assert abs(g["score"] - r["score"]) <= 1e-9 * abs(r["score"])


def test_periodicity_f64_reproduces_synthetic_report_members():
    # the FP64 twin on synthetic.fil: every one of the 549 saved members is found
    # (same trial, harmonics, frequency within 1e-5) with its f32-computed score
    # within 1e-4 (the reference rounds its spectrum to f32; the twin does not)
    path = reference_file("synthetic.fil")
    g = golden("synthetic_report.json")
    s = g["settings"]
    raw = open(path, "rb").read()
    off = golden("headers.json")["synthetic_header"]["data_offset"]
    data, ns = oracle.decode(raw[off:], 32, 32)
    freqs = 1450.0 - 4.0 * np.arange(32)
    dms = oracle.plan(s["dm_min"], s["dm_max"], s["dm_step"])
    res = oracle.search_trials(data, ns, freqs, s["tsamp"], dms, s["discovery_samples"], threads=4, f64=True,
                               min_hz=s["min_hz"], max_hz=s["max_hz"], threshold=s["threshold"])
    worst, checked = _match_members(res, g["members"], s["dm_step"])
    assert checked == len(g["members"]) == 549
    assert worst < 1e-4


def test_periodicity_f64_reproduces_gbt_report_members():
    # a sample of the GBT report's members (after its RFI mask), FP64 twin
    path = reference_file("GBT_Lband_PSR.fil")
    g = golden("gbt_report.json")
    s = g["settings"]
    h = golden("headers.json")["gbt_header"]
    raw = np.fromfile(path, dtype=np.uint8)
    data, ns = oracle.decode(raw[h["data_offset"]:], h["nbits"], h["nchans"])
    oracle.mask(data, ns, int(max(0.1 / s["tsamp"], 32)), 8.0)
    freqs = h["fch1"] + np.arange(h["nchans"]) * h["foff"]
    dms = oracle.plan(s["dm_min"], s["dm_max"], s["dm_step"])
    d = oracle.delays(freqs, s["tsamp"], s["dm_max"])
    disc = (ns - 2 * int(d.max())) // 2
    trials = sorted(set(m[1] for m in g["members"]))[7::40]
    idx = [int(round(t / s["dm_step"])) for t in trials]
    res = oracle.search_trials(data, ns, freqs, s["tsamp"], dms[idx], disc, threads=4, f64=True,
                               min_hz=s["min_hz"], max_hz=s["max_hz"], threshold=s["threshold"])
    worst, checked = _match_members(res, g["members"], s["dm_step"])
    assert checked > 100 and worst < 1e-4
