"""Parity of the gfx950 path (through the C ABI) against the oracle.

Integer work (delay tables, dedispersed integer channel sums, fold bins and
counts) must be bit-exact.  FP32 spectra are compared normwise (the FFT
rounding of the reference, rustfft, is not reproducible; see the oracle header)
with a 1e-5 relative bound, and candidate lists must agree on (harmonics,
frequency) with scores within 1e-4 relative, allowing only threshold-edge flips.
"""
import math

import numpy as np
import pytest

import oracle
from pulsarsr_amd import PRECISION_FP16, PRECISION_FP32, dedisperse, fold, search, synth
from pulsarsr_amd.core import PsrError
from matching import match_fp16_vs_f64, match_lists

pytestmark = pytest.mark.gpu

SCORE_RTOL = 1e-4


def small_workload(nbits, nchans=48, nsamples=1 << 15, seed=3, pulsar_amp=None):
    mean, sigma = {1: (0.5, 0.5), 2: (1.5, 0.8), 4: (8.0, 2.0), 8: (128.0, 16.0), 16: (30000.0, 3000.0),
                   32: (100.0, 1.0)}[nbits]
    amp = pulsar_amp if pulsar_amp is not None else synth.amplitude_for_snr(40.0, nchans, nsamples, 0.05)
    return synth.Workload("test-%dbit" % nbits, seed, nsamples, nchans, nbits, 64e-6, 1500.0, -2.0, mean, sigma,
                          pulsars=[(0.0123, 40.0, 0.05, amp, 0.3)], dms=np.arange(0, 80.0, 2.5),
                          min_hz=10.0, threshold=7.0)


def setup(w):
    payload = synth.generate(w)
    obs = dedisperse.Observation.from_packed(payload, w.nbits, w.nchans, w.tsamp, w.fch1, w.foff)
    data, ns = oracle.decode(payload, w.nbits, w.nchans)
    return payload, obs, data, ns


def disc_len(w):
    d = oracle.delays(w.frequencies(), w.tsamp, float(w.dms[-1]))
    return (w.nsamples - 2 * int(d.max())) // 2


@pytest.mark.parametrize("nbits", [1, 2, 4, 8, 16, 32])
def test_dedisperse_bit_exact(ctx, nbits):
    w = small_workload(nbits)
    _, obs, data, ns = setup(w)
    n = disc_len(w)
    got = dedisperse.apply_dms(obs, w.dms, n)
    for i, dm in enumerate(w.dms):
        ref = oracle.apply(data, ns, oracle.delays(w.frequencies(), w.tsamp, dm), n)
        assert (got[i].view(np.uint32) == ref.view(np.uint32)).all(), "DM %g" % dm


@pytest.mark.parametrize("nchans,lo,hi", [(601, 200, 256), (1023, 240, 256), (4200, 250, 256)])
def test_dedisperse_many_channels_exact(ctx, nchans, lo, hi):
    """Channel counts past one 16-bit phase (256 channels of 8-bit samples), odd
    counts, totals whose low halves wrap several times, and (4200 x 255 >= 2^20)
    the channel-ordered f32 kernel taking over: all bit-exact."""
    rng = np.random.default_rng(nchans)
    ns = 1 << 13
    payload = rng.integers(lo, hi, (ns, nchans), dtype=np.uint8).tobytes()
    foff = -300.0 / nchans
    obs = dedisperse.Observation.from_packed(payload, 8, nchans, 64e-6, 1500.0, foff)
    data, ns2 = oracle.decode(payload, 8, nchans)
    freqs = 1500.0 + foff * np.arange(nchans)
    dms = np.array([0.0, 3.0, 41.0, 97.5, 150.0])
    n = ns - int(oracle.delays(freqs, 64e-6, 150.0).max())
    print("dedisperse %d channels, samples %d..%d, %d outputs" % (nchans, lo, hi - 1, n), flush=True)
    got = dedisperse.apply_dms(obs, dms, n)
    for i, dm in enumerate(dms):
        ref = oracle.apply(data, ns2, oracle.delays(freqs, 64e-6, dm), n)
        assert ref.max() > 65535.0
        assert (got[i].view(np.uint32) == ref.view(np.uint32)).all(), "DM %g" % dm
    obs.close()


def test_dedisperse_pair_rows_mixed_steps(ctx):
    """Channel-pair rows against the oracle: an odd channel count in ascending
    frequency order, fine DM steps (every adjacent channel pair has one to three
    relative delays per 64-trial block, summed from pair rows), coarse steps
    (pairs with more relative delays, summed as single-channel rows), a block
    that mixes both and a partial last block -- bit-exact."""
    rng = np.random.default_rng(257)
    nchans, ns = 257, 1 << 14
    payload = rng.integers(0, 256, (ns, nchans), dtype=np.uint8).tobytes()
    obs = dedisperse.Observation.from_packed(payload, 8, nchans, 64e-6, 1200.0, 1.0)
    data, ns2 = oracle.decode(payload, 8, nchans)
    freqs = 1200.0 + np.arange(nchans)
    dms = np.concatenate([np.arange(128) * 0.25, 40.0 + np.arange(65) * 1.25])
    n = ns - int(oracle.delays(freqs, 64e-6, float(dms.max())).max())
    print("dedisperse %d channels, %d trials, %d outputs" % (nchans, len(dms), n), flush=True)
    got = dedisperse.apply_dms(obs, dms, n)
    for i, dm in enumerate(dms):
        ref = oracle.apply(data, ns2, oracle.delays(freqs, 64e-6, float(dm)), n)
        assert (got[i].view(np.uint32) == ref.view(np.uint32)).all(), "DM %g" % dm
    obs.close()


def test_dedisperse_full_length_and_unsorted_dms(ctx):
    w = small_workload(8)
    _, obs, data, ns = setup(w)
    dms = np.array([30.0, 0.0, 77.5, 12.5])
    got = dedisperse.apply_dms(obs, dms)  # None: full available length of the largest DM
    d = oracle.delays(w.frequencies(), w.tsamp, 77.5)
    assert got.shape[1] == ns - int(d.max())
    for i, dm in enumerate(dms):
        ref = oracle.apply(data, ns, oracle.delays(w.frequencies(), w.tsamp, dm), got.shape[1])
        assert (got[i] == ref).all()


@pytest.mark.parametrize("nchans,step", [(97, 0.25), (97, 3.0), (96, 0.5)])
def test_dedisperse_reads_to_the_end_of_the_data(ctx, nchans, step):
    """The row build reads each channel a few hundred samples past the last
    delayed sample it sums: with descending frequencies the last channel has the
    largest delays, and at the full available output length (the largest DM's
    delay plus n reaching the last sample, nsamples a multiple of 256 so the row
    pitch has its minimum slack) its reads run to the end of the allocation.
    Odd channel counts leave the last channel as a single-channel row; fine and
    coarse DM steps give pair rows and split pairs -- bit-exact."""
    rng = np.random.default_rng(nchans * 7 + int(step * 4))
    ns = 3 << 13
    payload = rng.integers(0, 256, (ns, nchans), dtype=np.uint8).tobytes()
    foff = -200.0 / nchans
    obs = dedisperse.Observation.from_packed(payload, 8, nchans, 64e-6, 1500.0, foff)
    data, ns2 = oracle.decode(payload, 8, nchans)
    freqs = 1500.0 + foff * np.arange(nchans)
    dms = np.arange(80) * step
    print("dedisperse %d channels, %d trials of step %g to the last sample" % (nchans, len(dms), step), flush=True)
    got = dedisperse.apply_dms(obs, dms)  # None: the full available length
    dmax = int(oracle.delays(freqs, 64e-6, float(dms[-1])).max())
    assert got.shape[1] == ns - dmax
    for i, dm in enumerate(dms):
        ref = oracle.apply(data, ns2, oracle.delays(freqs, 64e-6, float(dm)), got.shape[1])
        assert (got[i].view(np.uint32) == ref.view(np.uint32)).all(), "DM %g" % dm
    obs.close()


def test_apply_delay_table_matches_oracle(ctx):
    """psr_apply, 1:1 with dedisperse::apply(data, nsamples, delays, length): the
    full band with a DM's delays, subband channel slices with the full band's
    delays (confirmation.rs:362-372), all-zero tables (confirmation.rs:253,321),
    arbitrary non-monotonic delays, u8 and float observations -- bit-exact."""
    rng = np.random.default_rng(62)
    nch, ns = 48, 20000
    payload = rng.integers(0, 256, (ns, nch), dtype=np.uint8).tobytes()
    obs = dedisperse.Observation.from_packed(payload, 8, nch, 64e-6, 1500.0, -2.0)
    data, _ = oracle.decode(payload, 8, nch)
    freqs = 1500.0 - 2.0 * np.arange(nch)
    d = oracle.delays(freqs, 64e-6, 55.0)
    x = dedisperse.apply_delays(obs, d)
    assert (x == oracle.apply(data, ns, d)).all()
    n = 9000
    for lo, hi in ((0, 12), (12, 24), (24, 36), (36, 48), (5, 6)):
        sub = data[lo * ns:hi * ns]
        got = dedisperse.apply_delays(obs, d[lo:hi], n, channels=(lo, hi))
        assert (got.view(np.uint32) == oracle.apply(sub, ns, d[lo:hi], n).view(np.uint32)).all(), (lo, hi)
    z = np.zeros(nch, np.uint64)
    assert (dedisperse.apply_delays(obs, z, n) == oracle.apply(data, ns, z, n)).all()
    arb = rng.integers(0, 3000, nch).astype(np.uint64)
    assert (dedisperse.apply_delays(obs, arb) == oracle.apply(data, ns, arb)).all()
    f = (rng.standard_normal(nch * 5000) * 3.0 + 7.0).astype(np.float32)
    fobs = dedisperse.Observation.from_f32(f, 5000, 64e-6, 1500.0, -2.0)
    arb = rng.integers(0, 700, nch).astype(np.uint64)
    got = dedisperse.apply_delays(fobs, arb[10:30], channels=(10, 30))
    ref = oracle.apply(f[10 * 5000:30 * 5000], 5000, arb[10:30])
    assert (got.view(np.uint32) == ref.view(np.uint32)).all()
    # apply()'s errors (crates/dedisperse/src/lib.rs:67-78)
    with pytest.raises(PsrError, match="array shape mismatch"):
        dedisperse.apply_delays(obs, d[:3], channels=(0, 4))
    with pytest.raises(PsrError, match="array shape mismatch"):
        dedisperse.apply_delays(obs, [], channels=(4, 4))
    with pytest.raises(PsrError, match="delay exceeds observation"):
        dedisperse.apply_delays(obs, np.full(nch, ns + 1, np.uint64))
    with pytest.raises(PsrError, match="invalid output length"):
        dedisperse.apply_delays(obs, d, ns)
    with pytest.raises(PsrError, match="invalid output length"):
        dedisperse.apply_delays(obs, d, 1)
    obs.close()
    fobs.close()


def test_dedisperse_longer_than_a_search_series(ctx):
    """apply() has no length limit: a series past 2^23 samples (longer than any
    search takes) is dedispersed in output segments, bit-exact."""
    rng = np.random.default_rng(23)
    nch, ns = 6, (1 << 23) + 70000
    payload = rng.integers(0, 16, (ns, nch), dtype=np.uint8)
    packed = (payload[:, 0::2] | (payload[:, 1::2] << 4)).astype(np.uint8).tobytes()
    obs = dedisperse.Observation.from_packed(packed, 4, nch, 64e-6, 1500.0, -20.0)
    data, ns2 = oracle.decode(packed, 4, nch)
    assert ns2 == ns
    freqs = 1500.0 - 20.0 * np.arange(nch)
    for dm in (0.0, 300.0):
        got = dedisperse.apply(obs, dm)
        ref = oracle.apply(data, ns, oracle.delays(freqs, 64e-6, dm))
        assert len(got) == len(ref) > (1 << 23)
        assert (got == ref).all(), dm
    obs.close()


def test_dispersed_delta_collapses(ctx):
    # crates/dedisperse/src/lib.rs:95-108 through the GPU
    for freqs in ([1400.0, 1300.0, 1200.0], [1200.0, 1300.0, 1400.0]):
        d = dedisperse.delays(freqs, 0.001, 20.0)
        n = 100
        data = np.zeros(n * 3, np.float32)
        for c, dl in enumerate(d):
            data[c * n + 10 + int(dl)] = 1.0
        obs = dedisperse.Observation.from_f32(data, n, 0.001, freqs[0], freqs[1] - freqs[0])
        out = dedisperse.apply(obs, 20.0)
        assert out[10] == 3.0 and out.sum() == 3.0


def test_f32_channel_order_sums(ctx):
    rng = np.random.default_rng(9)
    nch, n = 24, 5000
    data = (rng.standard_normal(nch * n) * 37.0 + 1000.0).astype(np.float32)
    freqs = 1400.0 - np.arange(nch) * 3.0
    obs = dedisperse.Observation.from_f32(data, n, 1e-3, 1400.0, -3.0)
    dms = [0.0, 35.0, 70.0]
    got = dedisperse.apply_dms(obs, dms, 4000)
    for i, dm in enumerate(dms):
        ref = oracle.apply(data, n, oracle.delays(freqs, 1e-3, dm), 4000)
        assert (got[i].view(np.uint32) == ref.view(np.uint32)).all()


# M = next_pow2(n) complex points: one pass up to 4096, else rows of m1 = M / 512
# (m1 = 16, 64, 128, 256, 512, 1024, 2048, 4096) and m1 = 4096 with m2 = 1024, 2048
@pytest.mark.parametrize("n", [1000, 4096, 8000, 32583, 50000, 100000, 200000, 263500, 600000, 1500000, 2097152,
                               2300000, 4194305, 8388608])
def test_spectrum_matches_oracle(ctx, n):
    rng = np.random.default_rng(n)
    t = np.arange(n)
    x = (rng.standard_normal(n) * 16 + 2000 + 30 * np.sin(2 * np.pi * t / 97.3)).round().astype(np.float32)
    raw, white = search.spectrum(x, 64e-6, search.Config(min_hz=10.0))
    oraw, owhite = oracle.spectrum(x, 64e-6, min_hz=10.0)
    assert len(raw) == len(oraw)
    assert np.linalg.norm(raw - oraw) <= 1e-5 * np.linalg.norm(oraw)
    # the median-based scales are exact order statistics: compare bin by bin against
    # the local whitening level (1e-5 of it)
    assert np.max(np.abs(white - owhite)) <= 1e-4 * max(1.0, np.max(np.abs(owhite))) + 1e-5
    assert np.linalg.norm(white - owhite) <= 1e-5 * np.linalg.norm(owhite)


# The four-step splits psr_batch's autotuner chooses among (m2 = 1024 and 2048
# rows of m1 = M / m2): row lengths m1 = 32 .. 2048 with the other pass-1
# lengths, FP32 against the oracle as above; FP16 storage (amplitudes |X| /
# sqrt(n): n times the power it returns) normwise within its half-precision
# rounding (5e-3 of the power, the bound the configs hold it to)
@pytest.mark.parametrize("n,m2", [(50000, 1024), (50000, 2048), (200000, 1024), (200000, 2048), (263500, 2048),
                                  (600000, 1024), (600000, 2048), (1500000, 1024), (1500000, 2048),
                                  (2300000, 2048)])
def test_spectrum_splits_match_oracle(ctx, n, m2):
    print("split n=%d m2=%d" % (n, m2), flush=True)
    rng = np.random.default_rng(n + m2)
    t = np.arange(n)
    x = (rng.standard_normal(n) * 16 + 2000 + 30 * np.sin(2 * np.pi * t / 97.3)).round().astype(np.float32)
    oraw, owhite = oracle.spectrum(x, 64e-6, min_hz=10.0)
    raw, white = search.spectrum(x, 64e-6, search.Config(min_hz=10.0), m2=m2)
    assert len(raw) == len(oraw)
    assert np.linalg.norm(raw - oraw) <= 1e-5 * np.linalg.norm(oraw)
    assert np.max(np.abs(white - owhite)) <= 1e-4 * max(1.0, np.max(np.abs(owhite))) + 1e-5
    raw16, _ = search.spectrum(x, 64e-6, search.Config(min_hz=10.0, precision=PRECISION_FP16), m2=m2)
    p16 = raw16.astype(np.float64) * n
    assert np.linalg.norm(p16 - oraw) <= 5e-3 * np.linalg.norm(oraw)


def test_spectrum_split_errors(ctx):
    x = np.zeros(600000, dtype=np.float32)
    with pytest.raises(PsrError, match="m2 must be"):
        search.spectrum(x, 64e-6, search.Config(min_hz=10.0), m2=777)
    with pytest.raises(PsrError, match="no four-step split"):  # 4096 points: one pass
        search.spectrum(x[:4000], 64e-6, search.Config(min_hz=10.0), m2=512)


# (columns per whitening group: 32 at m2 = 512, 16 at m2 = 1024, 8 at m2 = 2048)
@pytest.mark.parametrize("n,m2", [(263500, 0), (1500000, 0), (2097152, 0), (2300000, 0), (4194305, 0),
                                  (8388608, 0), (600000, 1024)])
def test_fp16_whitening_scales_are_exact_window_medians(ctx, n, m2):
    # FP16 storage (packed 16-bit block sort, windowed cut search): every 64-bin
    # block's scale is max(max(upper median / ln2, floor), 1e-20) of the stored
    # powers a * a over bins [64b - 128, 64b + 192) clipped to [1, nbins)
    # (crates/search/src/lib.rs:45-60), so the whitened power must equal
    # raw / scale recomputed from the returned raw spectrum with the oracle's
    # median, bit for bit (the floor, mean * 1e-6, is far below every median here)
    print("fp16 whitening n=%d m2=%d" % (n, m2), flush=True)
    rng = np.random.default_rng(n + 7)
    t = np.arange(n)
    x = (rng.standard_normal(n) * 16 + 2000 + 30 * np.sin(2 * np.pi * t / 97.3)).round().astype(np.float32)
    raw, white = search.spectrum(x, 64e-6, search.Config(min_hz=10.0, precision=PRECISION_FP16), m2=m2)
    nb = len(raw)
    assert np.all(np.isfinite(raw)) and raw[1:].min() >= 0.0
    floor = np.float32(raw.astype(np.float64).sum() / nb * 1e-6)
    ln2 = np.float32(math.log(2.0))
    exp = np.empty(nb, np.float32)
    for b in range((nb + 63) // 64):
        lo, hi = max(64 * b - 128, 1), min(64 * b + 192, nb)
        med = np.float32(oracle.median(raw[lo:hi]))
        assert med / ln2 > floor
        scale = max(max(med / ln2, floor), np.float32(1e-20))
        exp[64 * b:64 * b + 64] = raw[64 * b:64 * b + 64] / scale
    exp[0] = 0.0
    bad = np.nonzero(white != exp)[0]
    assert bad.size == 0, (bad[:8], white[bad[:8]], exp[bad[:8]])


@pytest.mark.parametrize("nbits", [4, 8])
def test_search_matches_oracle(ctx, nbits):
    w = small_workload(nbits, nsamples=1 << 16)
    _, obs, data, ns = setup(w)
    n = disc_len(w)
    cfg = search.Config(min_hz=w.min_hz, max_hz=w.max_hz, threshold=w.threshold, max_harmonics=16)
    got = search.search(obs, w.dms, n, cfg)
    ref = oracle.search_trials(data, ns, w.frequencies(), w.tsamp, w.dms, n, threads=4, min_hz=w.min_hz,
                               max_hz=w.max_hz, threshold=w.threshold)
    assert sum(len(x) for x in ref) > 0
    assert match_lists(got, ref, w.threshold) < SCORE_RTOL
    # the injected pulsar is the strongest family
    best = max((c for x in got for c in x), key=lambda c: c.score)
    assert abs(best.frequency_hz * 0.0123 - round(best.frequency_hz * 0.0123)) < 0.05


def test_periodicity_finds_sine_and_respects_birdies(ctx):
    # crates/search/src/lib.rs:180-203
    tsamp, f = 0.001, 31.25
    x = np.sin(2 * np.pi * f * np.arange(8192) * tsamp).astype(np.float32)
    c = search.periodicity(x, tsamp, 0.0, search.Config(min_hz=10.0))
    assert abs(c[0].frequency_hz - f) < 0.1
    ref = oracle.periodicity(x, tsamp, 0.0, min_hz=10.0)
    assert match_lists([c], [ref], 6.0) < SCORE_RTOL
    c = search.periodicity(x, tsamp, 0.0, search.Config(min_hz=10.0, birdies=[(f, 1.0)]))
    assert all(abs(k.frequency_hz - f) > 1.0 for k in c)


def test_off_bin_pulse_train_retains_fundamental(ctx):
    # crates/search/src/lib.rs:228-247
    from test_oracle_golden import xorshift_pulse_train
    x, dt, period = xorshift_pulse_train()
    c = search.periodicity(x, dt, 18.0)
    kept = search.sift(c, 1.5 / (len(x) * dt), 10, False)
    assert abs(kept[0].frequency_hz - 1.0 / period) < 0.2
    ref = oracle.periodicity(x, dt, 18.0)
    assert match_lists([c], [ref], 6.0) < SCORE_RTOL


@pytest.mark.parametrize("precision", [PRECISION_FP32, PRECISION_FP16])
def test_batched_search_matches_one_batch(precision):
    """Trial batches overlapped on two streams (and run serially) give the
    candidates of a single batch, scattered back to the callers' unsorted
    trial order."""
    from pulsarsr_amd import Context
    w = small_workload(8, nsamples=1 << 16)
    payload = synth.generate(w)
    dms = np.random.default_rng(9).permutation(np.arange(200) * 0.4)
    d = oracle.delays(w.frequencies(), w.tsamp, float(dms.max()))
    n = (w.nsamples - 2 * int(d.max())) // 2
    cfg = search.Config(min_hz=w.min_hz, threshold=w.threshold, precision=precision)
    res = {}
    for name, overlap, cap in (("one", False, 0), ("overlapped", True, 120 << 20), ("serial", False, 120 << 20)):
        c = Context(0, overlap=overlap)
        obs = dedisperse.Observation.from_packed(payload, w.nbits, w.nchans, w.tsamp, w.fch1, w.foff, ctx=c)
        c.set_memory_cap(cap)
        out, counts, _ = search.search_raw(obs, dms, n, cfg)
        st = c.stage_times()
        print("%s: %d batches, %.2f ms" % (name, st["launches"], st["total_ms"]), flush=True)
        res[name] = (bytes(out), counts.copy(), st["launches"])
        obs.close()
        c.close()
    assert res["one"][2] == 1 and res["overlapped"][2] >= 2 and res["serial"][2] >= 2
    assert res["one"][1].sum() > 0
    for k in ("overlapped", "serial"):
        assert (res[k][1] == res["one"][1]).all() and res[k][0] == res["one"][0], k


def test_search_is_deterministic(ctx):
    w = small_workload(8, nsamples=1 << 16)
    _, obs, _, _ = setup(w)
    n = disc_len(w)
    cfg = search.Config(min_hz=w.min_hz, threshold=w.threshold)
    a, ca, _ = search.search_raw(obs, w.dms, n, cfg)
    b, cb, _ = search.search_raw(obs, w.dms, n, cfg)
    assert (ca == cb).all() and bytes(a) == bytes(b)


def test_fp16_storage_error_vs_fp64_oracle(ctx):
    """FP16 storage against the all-f64 oracle on every trial: every candidate
    matched (or one the FP16 rounding can flip), at least 90 % matched, the
    strongest candidate matched in at least 90 % of the trials, SNR within 1e-2."""
    w = small_workload(8, nsamples=1 << 16)
    _, obs, data, ns = setup(w)
    n = disc_len(w)
    cfg = search.Config(min_hz=w.min_hz, threshold=w.threshold, precision=PRECISION_FP16)
    got = search.search(obs, w.dms, n, cfg)
    ref = oracle.search_trials(data, ns, w.frequencies(), w.tsamp, w.dms, n, threads=4, f64=True,
                               min_hz=w.min_hz, threshold=w.threshold)
    nfft = 2 * (1 << (n - 1).bit_length())
    m = match_fp16_vs_f64(got, ref, w.threshold, 1.0 / (nfft * w.tsamp), 1.0 / (n * w.tsamp))
    print("fp16 vs f64: %r" % m, flush=True)
    assert m["total"] >= 2 * len(w.dms) and m["matched"] >= 0.9 * m["total"]
    assert m["trials"] >= len(w.dms) // 2 and m["top_matched"] >= 0.9 * m["trials"]
    assert m["max_rel_snr_error_sharp"] <= 1e-2 and m["max_rel_snr_error"] <= 5e-2
    assert m["flat_peaks"] <= max(2, 0.01 * m["matched"])


def test_fold_bins_bit_exact(ctx):
    rng = np.random.default_rng(4)
    x = (rng.integers(0, 4000, 50000)).astype(np.float32)
    periods = [0.0123, 0.0456, 0.1, 0.00462]
    offs = [0, 17, 1000, 3]
    got = fold.profiles_at(x, 64e-6, periods, 64, offs)
    for g, p, o in zip(got, periods, offs):
        r = oracle.profile_at(x, 64e-6, p, 64, o)
        assert (g.counts == r["counts"]).all()
        assert (g.bins == r["bins"]).all()
        assert g.snr == r["snr"] and g.reduced_chi_square == r["reduced_chi_square"]


def test_refine_matches_oracle(ctx):
    i = np.arange(20000)
    x = np.where(i % 100 < 5, 10.0, ((i * 17 % 23) - 11.0) / 10.0).astype(np.float32)
    g = fold.refine(x, 0.001, 10.0, 0.05, 40, 32)
    r = oracle.refine(x, 0.001, 10.0, 0.05, 40, 32)
    assert g.period_s == r["period_s"] and g.reduced_chi_square == r["reduced_chi_square"]
    a = fold.profile(x, 0.001, 0.1, 32)
    b = fold.profile(x, 0.001, 0.1037, 32)
    assert a.snr > b.snr * 2 and int(a.counts.sum()) == len(x)


def test_error_paths(ctx):
    w = small_workload(8)
    _, obs, _, _ = setup(w)
    with pytest.raises(PsrError, match="nonnegative DM"):
        dedisperse.apply_dms(obs, [-1.0], 100)
    with pytest.raises(PsrError, match="invalid output length"):
        dedisperse.apply_dms(obs, [10.0], w.nsamples)
    with pytest.raises(PsrError, match="delay exceeds observation"):
        dedisperse.apply_dms(obs, [1e6], 100)
    with pytest.raises(PsrError) as e:
        search.periodicity(np.zeros(1, np.float32), 1e-3, 0.0)
    assert e.value.code == -6


def test_memory_cap_forces_oom(ctx):
    from pulsarsr_amd import Context
    c = Context(0)
    w = small_workload(8)
    payload = synth.generate(w)
    obs = dedisperse.Observation.from_packed(payload, w.nbits, w.nchans, w.tsamp, w.fch1, w.foff, ctx=c)
    c.set_memory_cap(1024)
    with pytest.raises(PsrError) as e:
        dedisperse.apply_dms(obs, w.dms, 1000)
    assert e.value.code == -4
    c.set_memory_cap(0)
    assert dedisperse.apply_dms(obs, w.dms[:2], 1000).shape == (2, 1000)
