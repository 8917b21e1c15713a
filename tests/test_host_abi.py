"""The C-ABI library on the CPU: it loads, exports every symbol of
include/psr_gpu.h, and its host-side numerics (no GPU) agree with the oracle.
No compute call here touches a GPU."""
import json
import math
import os
import re
import struct

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, ROOT
from pulsarsr_amd import _ffi, dedisperse, fbank, search
from pulsarsr_amd.core import PsrError


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "psr_gpu.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(psr_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = _ffi.lib()
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert {n for n, _, _ in _ffi.EXPORTS} == set(names)


def test_version_and_error_channel():
    assert b"gfx950" in _ffi.lib().psr_version()
    with pytest.raises(PsrError) as e:
        dedisperse.delays([0.0], 1.0, 1.0)
    assert e.value.code == _ffi.PSR_E_INVALID
    assert "positive frequencies/tsamp and nonnegative DM required" in str(e.value)


def test_delays_kat_and_errors():
    assert dedisperse.delays([1000.0, 2000.0], 0.001, 1.0).tolist() == [3, 0]
    with pytest.raises(PsrError):
        dedisperse.plan(0.0, 5.0, 0.0)
    with pytest.raises(PsrError, match="more than 100000 DM trials"):
        dedisperse.plan(0.0, 1e6, 1.0)


def test_delays_match_oracle_bit_exactly():
    rng = np.random.default_rng(0)
    for nch, lo, hi, ts in ((96, 1352.0, 1448.0, 72e-6), (1024, 1200.0, 1600.0, 64e-6), (33, 300.0, 400.0, 1e-4)):
        freqs = np.sort(rng.uniform(lo, hi, nch))[::-1].copy()
        for dm in rng.uniform(0, 1000, 20):
            assert (dedisperse.delays(freqs, ts, dm) == oracle.delays(freqs, ts, dm)).all()


def test_dm_step_plan_match_golden():
    g = json.load(open(os.path.join(GOLDEN, "gbt_report.json")))["settings"]
    h = json.load(open(os.path.join(GOLDEN, "headers.json")))["gbt_header"]
    freqs = dedisperse.frequencies(h["fch1"], h["foff"], h["nchans"])
    assert dedisperse.dm_step(freqs, h["tsamp"]) == g["dm_step"]
    p = dedisperse.plan(g["dm_min"], g["dm_max"], g["dm_step"])
    assert len(p) == g["trial_count"] and (p == oracle.plan(g["dm_min"], g["dm_max"], g["dm_step"])).all()
    d = json.load(open(os.path.join(GOLDEN, "dm62_delays.json")))
    assert dedisperse.delays(freqs, d["tsamp"], d["dm"]).tolist() == d["delays"]


def _put(s):
    return struct.pack("<i", len(s)) + s.encode()


def _header(bits=8, extra=b""):
    b = _put("HEADER_START")
    for k, v in (("nchans", 2), ("nbits", bits), ("nifs", 1), ("data_type", 1)):
        b += _put(k) + struct.pack("<i", v)
    for k, v in (("tsamp", 0.001), ("fch1", 1400.0), ("foff", -1.0)):
        b += _put(k) + struct.pack("<d", v)
    return b + extra + _put("HEADER_END")


def test_parse_header_fields():
    b = _header(4) + bytes([0x21, 0x43])
    h = fbank.parse_header(b)
    assert (h.nchans, h.nbits, h.tsamp, h.fch1, h.foff) == (2, 4, 0.001, 1400.0, -1.0)
    assert h.data_offset == len(b) - 2 and h.nsamples == 2
    assert h.frequencies().tolist() == [1400.0, 1399.0]


def test_parse_header_rejects_unknown_and_duplicate_keywords():
    # crates/fbank/src/lib.rs:281-292
    with pytest.raises(PsrError, match="unknown keyword"):
        fbank.parse_header(_put("HEADER_START") + _put("surprise"))
    with pytest.raises(PsrError, match="duplicate nbits"):
        fbank.parse_header(_header(8, _put("nbits") + struct.pack("<i", 8)))
    with pytest.raises(PsrError, match="missing HEADER_START"):
        fbank.parse_header(_put("HEADER_STOP"))


def test_sift_matches_oracle():
    rng = np.random.default_rng(2)
    cands = []
    for _ in range(300):
        f = float(rng.uniform(1, 400))
        cands.append(search.Candidate(f, 1 / f, float(rng.choice([0.0, 10.0, 20.0])), int(rng.choice([1, 2, 4])),
                                      float(np.round(rng.uniform(6, 50), 1))))
    got = search.sift(cands, 0.05, 40, True)
    ref = oracle.sift([c.__dict__ for c in cands], 0.05, 40, True)
    assert [(c.frequency_hz, c.dm, c.score) for c in got] == [(c["frequency_hz"], c["dm"], c["score"]) for c in ref]


def test_sift_reference_case():
    # crates/search/src/lib.rs:204-226
    mk = lambda f, dm, s: search.Candidate(f, 1 / f, dm, 1, s)  # noqa: E731
    c = search.sift([mk(100.0, 20.0, 10.0), mk(200.0, 21.0, 8.0), mk(100.01, 19.0, 9.0), mk(137.0, 0.0, 20.0)],
                    0.1, 10, True)
    assert len(c) == 1 and c[0].dm == 20.0


def test_group_matches_oracle_and_rules_default():
    # confirmation::group (crates/search/src/confirmation.rs:177-199) is host code
    from pulsarsr_amd import confirmation
    rng = np.random.default_rng(7)
    base = rng.uniform(1.0, 50.0, 12)
    cands = []
    for i in range(400):
        f = base[i % 12] * rng.integers(1, 5) / rng.integers(1, 3) + rng.normal(0, 1e-3)
        s = float(np.float32(rng.uniform(6, 40)))  # f32-valued scores give exact ties
        cands.append(dict(frequency_hz=f, period_s=1.0 / f, dm=float(rng.integers(0, 40)) * 0.5, harmonics=1,
                          score=s if i % 7 else 12.0))
    for limit in (1, 5, 50):
        got = confirmation.group(cands, 0.01, limit)
        want = oracle.group(cands, 0.01, limit)
        assert got == want
    assert confirmation.default_rules() == oracle.DEFAULT_RULES


def test_measure_matches_oracle_and_closed_form():
    # confirmation::measure (confirmation.rs:107-141) through psr_measure
    from pulsarsr_amd import confirmation
    rng = np.random.default_rng(107)
    x = (rng.standard_normal(30000) * 5.0 + 100.0).astype(np.float32)
    tsamp, period = 64e-6, 0.0123
    for win, off in (({"start_bin": 10, "width_bins": 4, "bins": 64}, 0),
                     ({"start_bin": 62, "width_bins": 8, "bins": 64}, 12345),  # wraps past bin 63
                     ({"start_bin": 0, "width_bins": 1, "bins": 32}, 7)):
        got = confirmation.measure(x, off, tsamp, period, win)
        assert got == oracle.measure(x, off, tsamp, period, win)
        # closed form: Welch's two-sample statistic over the phase-binned samples
        ph = (np.arange(len(x)) + off) * tsamp / period
        b = np.minimum(((ph - np.trunc(ph)) * win["bins"]).astype(np.int64), win["bins"] - 1)
        on = (b + win["bins"] - win["start_bin"]) % win["bins"] < win["width_bins"]
        a, o = x[on].astype(np.float64), x[~on].astype(np.float64)
        d = a.mean() - o.mean()
        se = math.sqrt(a.var(ddof=1) / len(a) + o.var(ddof=1) / len(o))
        assert got["on_samples"] == len(a) and got["off_samples"] == len(o)
        assert got["mean_difference"] == pytest.approx(d, rel=1e-9, abs=1e-9)
        assert got["standard_error"] == pytest.approx(se, rel=1e-6)
        assert got["snr"] == pytest.approx(d / se, rel=1e-6, abs=1e-6)
    # fewer than two on-samples: the zero measurement (confirmation.rs:118-126)
    z = confirmation.measure(x[:3], 0, tsamp, period, {"start_bin": 0, "width_bins": 1, "bins": 64})
    assert z["snr"] == 0.0 and z["mean_difference"] == 0.0 and z["on_samples"] + z["off_samples"] == 3
    with pytest.raises(PsrError):
        confirmation.measure(x, 0, tsamp, period, {"start_bin": 0, "width_bins": 1, "bins": 0})
