"""The serde_json-compatible writer (pulsarsr_amd/report.py): byte-identical to the
reference's own saved JSON outputs, and ryu's float layout on known cases."""
import json

import numpy as np
import pytest

from conftest import reference_file
from pulsarsr_amd import report


@pytest.mark.parametrize("rel", ["output_confirm_gbt/report.json", "output_confirm_synth/report.json",
                                 "output_fold/fold.json", "output_confirm_gbt/header.json",
                                 "output_dm62/series.json"])
def test_reference_outputs_round_trip_byte_for_byte(rel):
    raw = open(reference_file(rel)).read()
    assert report.dumps(json.loads(raw)) == raw.rstrip("\n")


def test_ryu_float_layout():
    f = report.format_float
    assert f(6.0) == "6.0" and f(-0.0) == "-0.0" and f(1e16) == "1e16" and f(1234567.0) == "1234567.0"
    assert f(1e-6) == "1e-6" and f(1e-5) == "0.00001" and f(7.2e-5) == "0.000072" and f(1.5e-7) == "1.5e-7"
    assert f(0.1) == "0.1" and f(123.456) == "123.456" and f(2.5e20) == "2.5e20" and f(float("nan")) == "null"
    assert f(float(np.float32(0.1)), f32=True) == "0.1"
    assert report.dumps({"a": [], "b": {}, "c": [1, 2.0], "d": None, "e": True, "f": "x\"y"}) == \
        '{\n  "a": [],\n  "b": {},\n  "c": [\n    1,\n    2.0\n  ],\n  "d": null,\n  "e": true,\n  "f": "x\\"y"\n}'
