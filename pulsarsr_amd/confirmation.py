"""crates/search/src/confirmation.rs on the GPU: group and analyze.

`analyze` returns the reference's Outcome as a dict with the field names of
its serde serialisation (the confirmation block of report.json).
"""
import ctypes as C

import numpy as np

from . import _ffi
from .core import check, default_context
from .search import Candidate

# confirmation.rs:434-466, in bit order of psr_confirmed.rejection_mask
REJECTION_REASONS = [
    "held-out pulse SNR below threshold",
    "pulse does not persist in enough held-out time segments",
    "pulse does not persist in enough held-out subbands",
    "insufficient improvement over zero-DM interference control",
    "dispersion sweep too small relative to selected pulse width to distinguish local interference",
    "stronger rationally related periodic signal in zero-DM discovery control",
]


def default_rules():
    """Rules::default() (confirmation.rs:28-40) as a dict."""
    r = _ffi.Rules()
    _ffi.lib().psr_rules_default(C.byref(r))
    return {k: getattr(r, k) for k, _ in _ffi.Rules._fields_}


def _cand_array(cands):
    arr = (_ffi.Candidate * max(1, len(cands)))()
    for i, c in enumerate(cands):
        get = (lambda k: c[k]) if isinstance(c, dict) else (lambda k: getattr(c, k))
        arr[i] = _ffi.Candidate(get("frequency_hz"), get("period_s"), get("dm"), int(get("harmonics")),
                                get("score"))
    return arr


def _cand(c):
    return dict(frequency_hz=c.frequency_hz, period_s=c.period_s, dm=c.dm, harmonics=int(c.harmonics),
                score=c.score)


def measure(x, offset, tsamp, period, window):
    """confirmation::measure (confirmation.rs:107-141) -> Measurement dict.

    `window` is a dict with start_bin, width_bins and bins."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    w = _ffi.Window(int(window["start_bin"]), int(window["width_bins"]), int(window["bins"]))
    m = _ffi.Measurement()
    check(_ffi.lib().psr_measure(a.ctypes.data_as(C.POINTER(C.c_float)), a.size, int(offset), float(tsamp),
                                 float(period), C.byref(w), C.byref(m)))
    return {k: (int(getattr(m, k)) if k.endswith("samples") else getattr(m, k)) for k, _ in _ffi.Measurement._fields_}


def group(candidates, resolution, limit):
    """confirmation::group (confirmation.rs:177-199).

    Returns (leaders, members): members lists (family, index into candidates)
    in the order they joined."""
    n = len(candidates)
    arr = _cand_array(candidates)
    leaders = (_ffi.Candidate * max(1, limit))()
    members = (_ffi.Member * max(1, n))()
    nf, nm = C.c_size_t(), C.c_size_t()
    check(_ffi.lib().psr_group(arr, n, float(resolution), int(limit), leaders, C.byref(nf), members, C.byref(nm)))
    return ([_cand(leaders[i]) for i in range(nf.value)],
            [(int(members[i].family), int(members[i].index)) for i in range(nm.value)])


def _meas(m):
    return dict(snr=m.snr, on_samples=int(m.on_samples), off_samples=int(m.off_samples),
                mean_difference=m.mean_difference, standard_error=m.standard_error)


def _win(w):
    return dict(start_bin=int(w.start_bin), width_bins=int(w.width_bins), bins=int(w.bins))


def analyze(obs, trials, generated, limit, bins, rules=None, ctx=None):
    """confirmation::analyze (confirmation.rs:203-496) on a device-resident Observation."""
    ctx = ctx or obs.ctx or default_context()
    tr = np.ascontiguousarray(trials, dtype=np.float64)
    r = dict(default_rules(), **(rules or {}))
    R = _ffi.Rules(**r)
    gen = _cand_array(generated)
    out = _ffi.Outcome()
    conf = (_ffi.Confirmed * limit)()
    pb = np.zeros(limit * bins, np.float32)
    pc = np.zeros(limit * bins, np.uint64)
    tp = np.zeros(limit * 4 * bins, np.float32)
    sp = np.zeros(limit * 4 * bins, np.float32)
    dmr = (_ffi.DmPoint * (limit * len(tr)))()
    mem = (_ffi.Member * max(1, len(generated)))()
    nm = C.c_size_t()
    fp, up = C.POINTER(C.c_float), C.POINTER(C.c_uint64)
    check(_ffi.lib().psr_analyze(ctx.handle, obs.handle, tr.ctypes.data_as(C.POINTER(C.c_double)), len(tr), gen,
                                 len(generated), int(limit), int(bins), C.byref(R), C.byref(out), conf,
                                 pb.ctypes.data_as(fp), pc.ctypes.data_as(up), tp.ctypes.data_as(fp),
                                 sp.ctypes.data_as(fp), dmr, mem, C.byref(nm)))
    members = [(int(mem[i].family), int(mem[i].index)) for i in range(nm.value)]
    nt = len(tr)
    cands = []
    for k in range(int(out.ncandidates)):
        o = conf[k]
        cands.append(dict(
            candidate=_cand(o.candidate), refined_dm=o.refined_dm,
            profile=dict(period_s=o.profile.period_s, snr=o.profile.snr,
                         reduced_chi_square=o.profile.reduced_chi_square,
                         bins=pb[k * bins:(k + 1) * bins].copy(), counts=pc[k * bins:(k + 1) * bins].copy()),
            window=_win(o.window), validation=_meas(o.validation), zero_dm_validation=_meas(o.zero_dm_validation),
            zero_dm_period_s=o.zero_dm_period_s, zero_dm_window=_win(o.zero_dm_window), dm_contrast=o.dm_contrast,
            dispersion_widths=o.dispersion_widths, zero_dm_related_spectral_ratio=o.zero_dm_related_spectral_ratio,
            time_parts=[_meas(o.time_parts[i]) for i in range(4)], subbands=[_meas(o.subbands[i]) for i in range(4)],
            time_profiles=[tp[(k * 4 + i) * bins:(k * 4 + i + 1) * bins].copy() for i in range(4)],
            subband_profiles=[sp[(k * 4 + i) * bins:(k * 4 + i + 1) * bins].copy() for i in range(4)],
            dm_response=[dict(dm=p.dm, spectral_score=p.spectral_score if p.has_spectral else None,
                              spectral_frequency_hz=p.spectral_frequency_hz if p.has_spectral else None,
                              discovery_profile_response=p.discovery_profile_response)
                         for p in dmr[k * nt:(k + 1) * nt]],
            harmonic_members=[_cand(gen[i]) for fam, i in members if fam == k],
            accepted=bool(o.accepted),
            rejection_reasons=[REJECTION_REASONS[b] for b in range(6) if o.rejection_mask >> b & 1]))
    return dict(rules=r, discovery_samples=int(out.discovery_samples),
                validation_start_sample=int(out.validation_start_sample),
                validation_samples=int(out.validation_samples), common_samples=int(out.common_samples),
                trial_count=int(out.trial_count), candidates=cands)


__all__ = ["REJECTION_REASONS", "default_rules", "group", "analyze", "Candidate"]
