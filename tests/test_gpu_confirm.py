"""psr_analyze (confirmation::analyze on the GPU) against the oracle's analyze.

The oracle's analyze reproduces the reference's saved confirmation reports
(tests/test_oracle_confirm.py).  On integer-valued samples every dedispersion,
fold and DM-response value of the GPU path is exact, so the outcome must match
the oracle field for field; the one FFT-dependent value, the zero-DM spectral
ratio (its zero-DM candidates come from the GPU search), gets a 1e-4 relative
bound.
"""
import numpy as np
import pytest

import oracle
from pulsarsr_amd import confirmation, dedisperse, synth

pytestmark = pytest.mark.gpu


def workload(nbits, nchans):
    mean, sigma = {4: (8.0, 2.0), 8: (128.0, 16.0)}[nbits]
    amp = synth.amplitude_for_snr(40.0, nchans, 1 << 15, 0.05)
    return synth.Workload("confirm-%dbit" % nbits, 11, 1 << 15, nchans, nbits, 64e-6, 1500.0, -2.0, mean, sigma,
                          pulsars=[(0.0123, 40.0, 0.05, amp, 0.3)], dms=np.arange(0, 80.0, 2.5),
                          min_hz=10.0, threshold=7.0)


def compare(got, want):
    for k in ("discovery_samples", "validation_start_sample", "validation_samples", "common_samples",
              "trial_count"):
        assert got[k] == want[k], k
    assert len(got["candidates"]) == len(want["candidates"]) > 0
    for g, w in zip(got["candidates"], want["candidates"]):
        assert g["candidate"] == w["candidate"]
        assert g["refined_dm"] == w["refined_dm"]
        for k in ("period_s", "snr", "reduced_chi_square"):
            assert g["profile"][k] == w["profile"][k], k
        assert (g["profile"]["bins"] == w["profile"]["bins"]).all()
        assert (g["profile"]["counts"] == w["profile"]["counts"]).all()
        for k in ("window", "validation", "zero_dm_validation", "zero_dm_period_s", "zero_dm_window",
                  "dm_contrast", "dispersion_widths", "time_parts", "subbands", "accepted",
                  "rejection_reasons"):
            assert g[k] == w[k], k
        assert g["zero_dm_related_spectral_ratio"] == pytest.approx(w["zero_dm_related_spectral_ratio"], rel=1e-4)
        for a, b in zip(g["time_profiles"] + g["subband_profiles"], w["time_profiles"] + w["subband_profiles"]):
            assert (a == b).all()
        assert g["dm_response"] == w["dm_response"]
        assert g["harmonic_members"] == w["harmonic_members"]


@pytest.mark.parametrize("nbits,nchans", [(8, 48), (4, 50)])
def test_analyze_matches_oracle(ctx, nbits, nchans):
    w = workload(nbits, nchans)
    payload = synth.generate(w)
    obs = dedisperse.Observation.from_packed(payload, w.nbits, w.nchans, w.tsamp, w.fch1, w.foff)
    data, ns = oracle.decode(payload, w.nbits, w.nchans)
    freqs = w.frequencies()
    d = oracle.delays(freqs, w.tsamp, float(w.dms[-1]))
    disc = (ns - 2 * int(d.max())) // 2
    per_trial = oracle.search_trials(data, ns, freqs, w.tsamp, w.dms, disc, threads=4, min_hz=w.min_hz,
                                     threshold=w.threshold)
    generated = [c for lst in per_trial for c in lst]
    print("analyze: %d-bit, %d channels, %d generated candidates" % (nbits, nchans, len(generated)), flush=True)
    got = confirmation.analyze(obs, w.dms, generated, 3, 32)
    want = oracle.analyze(data, ns, freqs, w.tsamp, w.dms, generated, 3, 32)
    compare(got, want)


def test_analyze_error_paths(ctx):
    w = workload(8, 48)
    obs = dedisperse.Observation.from_packed(synth.generate(w), w.nbits, w.nchans, w.tsamp, w.fch1, w.foff)
    with pytest.raises(Exception, match="DM trials must increase"):
        confirmation.analyze(obs, [0.0, 5.0, 5.0], [], 2, 32)
    with pytest.raises(Exception, match=">=8 bins"):
        confirmation.analyze(obs, [0.0, 5.0], [], 2, 4)
