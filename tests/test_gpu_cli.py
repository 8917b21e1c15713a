"""The `psr` drop-in (python -m pulsarsr_amd) end to end on a synthetic SIGPROC file.

search: header.json and report.json must have the reference's schema-2 layout
(key order of output_confirm_gbt/report.json, tests/golden/report_schema.json),
and the report's content must equal the oracle's pipeline on the same file:
rfi::mask exactly; every reported family member is a candidate of the oracle's
per-trial search (scores within 1e-4: the FFT is not the oracle's); and
confirmation::analyze over those members equals the report exactly (except the
zero-DM spectral ratio, again FFT-dependent).  fold and dedisperse must equal the oracle's apply
and profile.
"""
import json
import os
import struct

import numpy as np
import pytest

import oracle
from conftest import GOLDEN
from pulsarsr_amd import cli, synth

pytestmark = pytest.mark.gpu


def _str(s):
    return struct.pack("<i", len(s)) + s.encode()


def write_fil(path, w, payload):
    h = _str("HEADER_START") + _str("rawdatafile") + _str(os.path.basename(path))
    h += _str("source_name") + _str("Synthetic")
    for k, v in (("machine_id", 10), ("telescope_id", 4), ("data_type", 1), ("nchans", w.nchans),
                 ("nbits", w.nbits), ("nifs", 1)):
        h += _str(k) + struct.pack("<i", v)
    for k, v in (("tstart", 60000.0), ("tsamp", w.tsamp), ("fch1", w.fch1), ("foff", w.foff)):
        h += _str(k) + struct.pack("<d", v)
    h += _str("HEADER_END")
    with open(path, "wb") as f:
        f.write(h)
        f.write(bytes(payload))


def keys(v, ref):
    """The key structure of v matches the reference report's skeleton."""
    if isinstance(ref, dict):
        assert isinstance(v, dict) and list(v.keys()) == list(ref.keys()), (list(v.keys()), list(ref.keys()))
        for k in ref:
            keys(v[k], ref[k])
    elif isinstance(ref, list) and ref and v:
        keys(v[0], ref[0])


@pytest.fixture(scope="module")
def fil(tmp_path_factory):
    w = synth.Workload("cli-8bit", 21, 1 << 16, 64, 8, 64e-6, 1500.0, -2.0, 128.0, 16.0,
                       pulsars=[(0.0093, 30.0, 0.05, synth.amplitude_for_snr(30.0, 64, 1 << 16, 0.05), 0.2)],
                       bursts=[(1.2, 0.02, 3.0)], rfi_chans=[9, 33], rfi_amp=3.0, dms=np.arange(0, 60.0, 2.0))
    payload = synth.generate(w)
    d = tmp_path_factory.mktemp("cli")
    path = str(d / "synthetic.fil")
    write_fil(path, w, payload)
    return w, payload, path, d


def test_search_report_matches_oracle_pipeline(ctx, fil):
    w, payload, path, d = fil
    out = str(d / "out")
    print("psr search on %s" % path, flush=True)
    rep = cli.main(["search", path, "--dm-max", "58", "--dm-step", "2", "--min-hz", "10", "--top", "3",
                    "--out", out])
    written = json.load(open(os.path.join(out, "report.json")))
    keys(written, json.load(open(os.path.join(GOLDEN, "report_schema.json"))))
    hdr = json.load(open(os.path.join(out, "header.json")))
    assert hdr["nchans"] == 64 and hdr["fields"]["source_name"] == "Synthetic" and list(hdr["fields"]) == sorted(hdr["fields"])
    # the oracle's pipeline on the same file (crates/cli/src/main.rs:416-487)
    data, ns = oracle.decode(payload, w.nbits, w.nchans)
    chans, blocks, frac = oracle.mask(data, ns, max(int(0.1 / w.tsamp), 32), 8.0)
    assert written["rfi"] == {"channels": chans, "blocks": blocks, "block_samples": max(int(0.1 / w.tsamp), 32),
                              "masked_fraction": frac}
    freqs = w.frequencies()
    trials = oracle.plan(0.0, 58.0, 2.0)
    maxd = int(oracle.delays(freqs, w.tsamp, 58.0).max())
    disc = (ns - 2 * maxd) // 2
    per = oracle.search_trials(data, ns, freqs, w.tsamp, trials, disc, threads=4, min_hz=10.0, max_hz=2000.0,
                               threshold=8.0)
    conf = written["confirmation"]
    assert conf["discovery_samples"] == disc and conf["trial_count"] == len(trials)
    # every reported family member is a per-trial candidate of the oracle's search
    members = [m for c in conf["candidates"] for m in c["harmonic_members"]]
    assert members
    by_dm = {round(float(trials[t]), 9): per[t] for t in range(len(trials))}
    for m in members:
        hit = [c for c in by_dm[round(m["dm"], 9)] if c["harmonics"] == m["harmonics"]
               and abs(c["frequency_hz"] - m["frequency_hz"]) <= 1e-6 * m["frequency_hz"]]
        assert hit and hit[0]["score"] == pytest.approx(m["score"], rel=1e-4)
    # analyze on the same generated candidates (the members, in generation order) is exact
    gen = sorted(members, key=lambda m: (m["dm"], -m["score"]))
    want = oracle.analyze(data, ns, freqs, w.tsamp, trials, gen, 3, 64)
    assert len(conf["candidates"]) == len(want["candidates"]) > 0
    for g, o in zip(conf["candidates"], want["candidates"]):
        for k in ("candidate", "refined_dm", "window", "validation", "zero_dm_validation", "zero_dm_period_s",
                  "dm_contrast", "time_parts", "subbands", "accepted"):
            assert g[k] == o[k], k
        assert (np.array(g["profile"]["bins"], np.float32) == o["profile"]["bins"]).all()
        assert [p["discovery_profile_response"] for p in g["dm_response"]] == \
            [p["discovery_profile_response"] for p in o["dm_response"]]
        assert g["zero_dm_related_spectral_ratio"] == pytest.approx(o["zero_dm_related_spectral_ratio"], rel=1e-4)
    assert [c["candidate"] for c in conf["candidates"]] == written["candidates"]
    assert rep["trial_count"] == len(trials) and written["input_sha256"] == rep["input_sha256"]


def test_fold_and_dedisperse_commands(ctx, fil):
    w, payload, path, d = fil
    data, ns = oracle.decode(payload, w.nbits, w.nchans)
    dl = oracle.delays(w.frequencies(), w.tsamp, 30.0)
    x = oracle.apply(data, ns, dl)
    out = str(d / "fold")
    cli.main(["fold", path, "--period", "0.0093", "--dm", "30", "--out", out])
    f = json.load(open(os.path.join(out, "fold.json")))
    p = oracle.profile_at(x, w.tsamp, 0.0093, 64, 0)
    assert list(f) == ["dm", "input", "profile", "searched"] and f["searched"] is False
    assert f["profile"]["counts"] == p["counts"].tolist()
    assert (np.array(f["profile"]["bins"], np.float32) == p["bins"]).all()
    assert f["profile"]["snr"] == p["snr"] and f["profile"]["reduced_chi_square"] == p["reduced_chi_square"]
    out = str(d / "dd")
    cli.main(["dedisperse", path, "--dm", "30", "--out", out])
    s = json.load(open(os.path.join(out, "series.json")))
    assert s["delays"] == dl.tolist() and s["samples"] == len(x)
    assert (np.fromfile(os.path.join(out, "series.f32"), "<f4") == x).all()
