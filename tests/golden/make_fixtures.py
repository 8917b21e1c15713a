"""Extract small golden fixtures from the reference's saved run outputs.

Reads only data files of the reference tree (JSON reports and SIGPROC headers);
writes compact JSON under tests/golden/.  Run from the repo root:
    python tests/golden/make_fixtures.py [/root/reference]
"""
import json
import os
import struct
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))


def header(path):
    """Numerical SIGPROC header fields (enough for the fixtures)."""
    with open(path, "rb") as f:
        b = f.read(4096)
    pos = 0

    def s():
        nonlocal pos
        n = struct.unpack_from("<i", b, pos)[0]
        pos += 4
        v = b[pos:pos + n].decode()
        pos += n
        return v
    assert s() == "HEADER_START"
    h = {}
    while True:
        k = s()
        if k == "HEADER_END":
            break
        if k in ("rawdatafile", "source_name"):
            h[k] = s()
        elif k in ("az_start", "za_start", "src_raj", "src_dej", "tstart", "tsamp", "fch1", "foff", "refdm", "period"):
            h[k] = struct.unpack_from("<d", b, pos)[0]
            pos += 8
        else:
            h[k] = struct.unpack_from("<i", b, pos)[0]
            pos += 4
    h["data_offset"] = pos
    h["file_bytes"] = os.path.getsize(path)
    return h


def members(report):
    r = json.load(open(report))
    c = r["confirmation"]
    mem = [[m["frequency_hz"], m["dm"], m["harmonics"], m["score"]] for cc in c["candidates"] for m in cc["harmonic_members"]]
    leaders = [[x["frequency_hz"], x["dm"], x["harmonics"], x["score"]] for x in r["candidates"]]
    return dict(
        settings=dict(dm_min=r["dm_min"], dm_max=r["dm_max"], dm_step=r["dm_step"], trial_count=r["trial_count"],
                      min_hz=r["min_hz"], max_hz=r["max_hz"], threshold=r["threshold"], birdies=r["birdies"],
                      common_samples=r["common_samples"], sample_count=r["sample_count"], tsamp=r["tsamp"],
                      discovery_samples=c["discovery_samples"], validation_start_sample=c["validation_start_sample"]),
        rfi=r["rfi"], members=mem, leaders=leaders,
        profiles=[dict(counts=cc["profile"]["counts"], bins=cc["profile"]["bins"], period_s=cc["profile"]["period_s"],
                       snr=cc["profile"]["snr"], reduced_chi_square=cc["profile"]["reduced_chi_square"],
                       refined_dm=cc["refined_dm"]) for cc in c["candidates"]])


def confirmation(report, families=2, dm_every=8):
    """The first `families` confirmed candidates of a report, for pinning analyze.

    dm_response keeps every `dm_every`-th trial (and the last) to stay small."""
    r = json.load(open(report))
    c = r["confirmation"]
    keep = ("candidate", "refined_dm", "profile", "window", "validation", "zero_dm_validation", "zero_dm_period_s",
            "zero_dm_window", "dm_contrast", "dispersion_widths", "zero_dm_related_spectral_ratio", "time_parts",
            "subbands", "time_profiles", "subband_profiles", "accepted", "rejection_reasons")
    out = []
    for cc in c["candidates"][:families]:
        d = {k: cc[k] for k in keep}
        n = len(cc["dm_response"])
        idx = sorted(set(list(range(0, n, dm_every)) + [n - 1]))
        d["dm_response_index"] = idx
        d["dm_response"] = [cc["dm_response"][i] for i in idx]
        d["nmembers"] = len(cc["harmonic_members"])
        out.append(d)
    head = {k: c[k] for k in ("discovery_samples", "validation_start_sample", "validation_samples",
                              "common_samples", "trial_count")}
    return dict(outcome=head, rules=c["rules"], families=out)


def skeleton(v):
    """Key structure of a JSON value: dicts keep key order, lists keep their first element's shape."""
    if isinstance(v, dict):
        return {k: skeleton(x) for k, x in v.items()}
    if isinstance(v, list):
        return [skeleton(v[0])] if v else []
    return type(v).__name__


def main():
    gbt = header(os.path.join(REF, "GBT_Lband_PSR.fil"))
    syn = header(os.path.join(REF, "synthetic.fil"))
    dm62 = json.load(open(os.path.join(REF, "output_dm62/series.json")))
    fold = json.load(open(os.path.join(REF, "output_fold/fold.json")))
    json.dump(dict(gbt_header=gbt, synthetic_header=syn), open(os.path.join(OUT, "headers.json"), "w"), indent=1)
    json.dump(dict(dm=dm62["dm"], tsamp=dm62["tsamp"], samples=dm62["samples"], delays=dm62["delays"]),
              open(os.path.join(OUT, "dm62_delays.json"), "w"))
    json.dump(dict(dm=fold["dm"], period_s=fold["profile"]["period_s"], bins=fold["profile"]["bins"],
                   counts=fold["profile"]["counts"], snr=fold["profile"]["snr"],
                   reduced_chi_square=fold["profile"]["reduced_chi_square"]),
              open(os.path.join(OUT, "gbt_fold_dm62.389.json"), "w"))
    json.dump(members(os.path.join(REF, "output_confirm_synth/report.json")),
              open(os.path.join(OUT, "synthetic_report.json"), "w"), separators=(",", ":"))
    json.dump(members(os.path.join(REF, "output_confirm_gbt/report.json")),
              open(os.path.join(OUT, "gbt_report.json"), "w"), separators=(",", ":"))
    json.dump(confirmation(os.path.join(REF, "output_confirm_gbt/report.json")),
              open(os.path.join(OUT, "gbt_confirm.json"), "w"), separators=(",", ":"))
    json.dump(confirmation(os.path.join(REF, "output_confirm_synth/report.json")),
              open(os.path.join(OUT, "synthetic_confirm.json"), "w"), separators=(",", ":"))
    json.dump(skeleton(json.load(open(os.path.join(REF, "output_confirm_gbt/report.json")))),
              open(os.path.join(OUT, "report_schema.json"), "w"), indent=1)


if __name__ == "__main__":
    main()
