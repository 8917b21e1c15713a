"""Candidate-list comparison rules shared by the GPU parity tests.

FP32 lists (the reference's arithmetic) must agree with the oracle on
(harmonics, frequency), scores within a relative bound, allowing only
threshold-edge flips and cap-edge drops.  FP16-storage lists are compared with
the all-f64 oracle: frequencies within half a bin (the same spectral peak; the
parabolic offset moves with the stored spectrum's rounding, most where the
peak is flat), SNR within 1e-2, and
unmatched candidates only where the FP16 rounding can legitimately decide:
scores within 2 % of the threshold, of the list's cap edge, or a dedup pair
(two candidates closer than the per-trial resolution with scores within 2 %).
"""
import math


def match_lists(got, ref, threshold, max_candidates=32):
    """Per-trial candidate lists agree up to threshold-edge flips; returns worst rel score error."""
    worst = 0.0
    for t, (g, r) in enumerate(zip(got, ref)):
        rs = [(c["harmonics"], c["frequency_hz"], c["score"]) for c in r]
        gs = [(c.harmonics, c.frequency_hz, c.score) for c in g]
        used = set()
        for h, f, s in gs:
            hit = None
            for j, (h2, f2, s2) in enumerate(rs):
                if j not in used and h2 == h and abs(f2 - f) <= 1e-6 * max(1.0, f):
                    hit = j
                    break
            if hit is None:
                assert abs(s - threshold) < 1e-3 * threshold or len(gs) >= max_candidates, (t, h, f, s)
                continue
            used.add(hit)
            worst = max(worst, abs(s - rs[hit][2]) / abs(rs[hit][2]))
        for j, (h2, f2, s2) in enumerate(rs):
            if j not in used:
                assert abs(s2 - threshold) < 1e-3 * threshold or len(rs) >= max_candidates, (t, h2, f2, s2)
    return worst


def match_fp16_vs_f64(got, ref, threshold, bin_hz, resolution_hz, max_candidates=32, ftol_bins=0.5,
                      edge=2e-2, sharp_bins=0.05):
    """FP16-storage candidate lists `got` (psr_candidate objects) against the f64
    oracle's lists `ref` (dicts), trial by trial.  Returns a dict with the
    matched and total counts, the worst relative SNR error over matched
    candidates, and the number of trials whose strongest FP16 candidate matched
    the oracle's strongest.  Asserts that every unmatched candidate, on either
    side, is one the FP16 rounding can flip (see the module docstring).

    Matched candidates whose frequencies differ by more than `sharp_bins` sit
    on flat spectral peaks, where the parabolic offset (and with it the bins
    the harmonic sum reads) is ill-conditioned; they are counted as flat_peaks
    and their SNR errors are kept apart from max_rel_snr_error_sharp."""
    worst, matched, total, top_ok, trials, fworst = 0.0, 0, 0, 0, 0, 0.0
    worst_sharp, flat = 0.0, 0
    for t, (g, r) in enumerate(zip(got, ref)):
        used = set()
        cap_edge_g = g[-1].score if len(g) >= max_candidates else -math.inf
        cap_edge_r = r[-1]["score"] if len(r) >= max_candidates else -math.inf

        def flippable(s, f, others, cap_edge):
            if abs(s - threshold) <= edge * threshold or s <= cap_edge * (1 + edge):
                return True
            return any(abs(o[1] - f) < resolution_hz + ftol_bins * bin_hz and abs(o[2] - s) <= edge * abs(s)
                       for o in others)
        gs = [(c.harmonics, c.frequency_hz, c.score) for c in g]
        rs = [(c["harmonics"], c["frequency_hz"], c["score"]) for c in r]
        for i, (h, f, s) in enumerate(gs):
            total += 1
            hit = None
            for j, (h2, f2, s2) in enumerate(rs):
                if j not in used and h2 == h and abs(f2 - f) <= ftol_bins * bin_hz:
                    hit = j
                    break
            if hit is None:
                near = sorted(((abs(f2 - f) / bin_hz, h2, s2) for h2, f2, s2 in rs if h2 == h))[:2]
                assert flippable(s, f, gs[:i] + rs, cap_edge_r), ("unmatched fp16 candidate", t, h, f, s,
                                                                  "nearest f64 (bins, h, score)", near)
                continue
            used.add(hit)
            matched += 1
            fe = abs(rs[hit][1] - f) / bin_hz
            err = abs(s - rs[hit][2]) / abs(rs[hit][2])
            fworst = max(fworst, fe)
            worst = max(worst, err)
            if fe <= sharp_bins:
                worst_sharp = max(worst_sharp, err)
            else:  # a flat peak: the interpolated offset, hence the harmonic bins, moved
                flat += 1
            if i == 0 and hit == 0:
                top_ok += 1
        for j, (h2, f2, s2) in enumerate(rs):
            if j not in used:
                assert flippable(s2, f2, rs[:j] + gs, cap_edge_g), ("missed f64 candidate", t, h2, f2, s2)
        if g or r:
            trials += 1
    return dict(matched=matched, total=total, max_rel_snr_error=worst, top_matched=top_ok, trials=trials,
                max_freq_error_bins=fworst, max_rel_snr_error_sharp=worst_sharp, flat_peaks=flat)
