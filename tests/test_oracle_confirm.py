"""The oracle's confirmation stage (oracle/analyze.c) against the reference's saved reports.

output_confirm_gbt/report.json and output_confirm_synth/report.json hold, per
confirmed family, everything confirmation::analyze computed
(crates/search/src/confirmation.rs:203-496).  The inputs are the reference's
own filterbanks; the generated candidates are the union of the report's family
members in generation order (trial, then score), which reproduces the grouping
of every reported family.  Everything is exact arithmetic on the same data
except the zero-DM control's spectral ratio, whose FFT is rustfft's.
"""
import json
import os

import numpy as np
import pytest

import oracle
from conftest import reference_file

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def golden(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def _generated(members):
    # generation order of crates/cli/src/main.rs:462-469: trial (ascending DM), then
    # each trial's ranked list (descending score)
    gen = sorted(members, key=lambda m: (m[1], -m[3]))
    return [dict(frequency_hz=f, period_s=1.0 / f, dm=dm, harmonics=h, score=s) for f, dm, h, s in gen]


def _check(got, want, outcome_want):
    for k, v in outcome_want.items():
        assert got[k] == v, k
    for fam, w in zip(got["candidates"], want):
        assert fam["candidate"] == w["candidate"]
        assert fam["refined_dm"] == w["refined_dm"]
        p = w["profile"]
        assert fam["profile"]["period_s"] == p["period_s"]
        assert fam["profile"]["snr"] == p["snr"]
        assert fam["profile"]["reduced_chi_square"] == p["reduced_chi_square"]
        assert fam["profile"]["counts"].tolist() == p["counts"]
        assert (fam["profile"]["bins"] == np.array(p["bins"], np.float32)).all()
        for k in ("window", "validation", "zero_dm_validation", "zero_dm_period_s", "zero_dm_window",
                  "dm_contrast", "dispersion_widths", "time_parts", "subbands", "accepted", "rejection_reasons"):
            assert fam[k] == w[k], k
        assert fam["zero_dm_related_spectral_ratio"] == pytest.approx(w["zero_dm_related_spectral_ratio"], rel=1e-6)
        for a, b in zip(fam["time_profiles"] + fam["subband_profiles"], w["time_profiles"] + w["subband_profiles"]):
            assert (a == np.array(b, np.float32)).all()
        for i, pt in zip(w["dm_response_index"], w["dm_response"]):
            g = fam["dm_response"][i]
            assert g["dm"] == pt["dm"]
            assert g["spectral_score"] == pt["spectral_score"]
            assert g["spectral_frequency_hz"] == pt["spectral_frequency_hz"]
            assert g["discovery_profile_response"] == pt["discovery_profile_response"]
        assert len(fam["harmonic_members"]) == w["nmembers"]


def test_analyze_reproduces_synthetic_confirmation():
    path = reference_file("synthetic.fil")
    rep = golden("synthetic_report.json")
    conf = golden("synthetic_confirm.json")
    s = rep["settings"]
    off = golden("headers.json")["synthetic_header"]["data_offset"]
    data, ns = oracle.decode(open(path, "rb").read()[off:], 32, 32)
    freqs = 1450.0 - 4.0 * np.arange(32)
    trials = oracle.plan(s["dm_min"], s["dm_max"], s["dm_step"])
    got = oracle.analyze(data, ns, freqs, s["tsamp"], trials, _generated(rep["members"]), 2, 64)
    _check(got, conf["families"], conf["outcome"])


def test_analyze_reproduces_gbt_confirmation():
    path = reference_file("GBT_Lband_PSR.fil")
    rep = golden("gbt_report.json")
    conf = golden("gbt_confirm.json")
    s = rep["settings"]
    h = golden("headers.json")["gbt_header"]
    raw = np.fromfile(path, dtype=np.uint8)
    data, ns = oracle.decode(raw[h["data_offset"]:], h["nbits"], h["nchans"])
    oracle.mask(data, ns, int(max(0.1 / s["tsamp"], 32)), 8.0)
    freqs = h["fch1"] + np.arange(h["nchans"]) * h["foff"]
    trials = oracle.plan(s["dm_min"], s["dm_max"], s["dm_step"])
    got = oracle.analyze(data, ns, freqs, s["tsamp"], trials, _generated(rep["members"]), 2, 64)
    _check(got, conf["families"], conf["outcome"])


def test_measure_window_and_group_invariants():
    # window.contains wraps around (confirmation.rs:47-51); measure splits on/off by phase bin
    x = np.where(np.arange(4000) % 100 < 10, 5.0, 1.0).astype(np.float32)
    w = oracle.select_window(oracle.profile_at(x, 1.0, 100.0, 16)["bins"])
    assert w["width_bins"] == 2 and w["bins"] == 16 and w["start_bin"] == 0
    m = oracle.measure(x, 0, 1.0, 100.0, w)
    # bins 0 and 1 hold phases [0, 12.5) of 100 samples: 13 samples per period,
    # 10 of them in the pulse
    assert m["on_samples"] == 13 * 40 and m["off_samples"] == 87 * 40
    assert m["mean_difference"] == pytest.approx((5.0 * 10 + 1.0 * 3) / 13 - 1.0, rel=1e-12)
    # group: harmonically related candidates join the strongest family's leader
    mk = lambda f, s: dict(frequency_hz=f, period_s=1.0 / f, dm=1.0, harmonics=1, score=s)
    leaders, members = oracle.group([mk(10.0, 5.0), mk(30.0, 9.0), mk(7.3, 4.0), mk(7.31, 3.0)], 0.02, 2)
    assert [l["frequency_hz"] for l in leaders] == [30.0, 7.3]
    assert members == [(0, 1), (0, 0), (1, 2), (1, 3)]
