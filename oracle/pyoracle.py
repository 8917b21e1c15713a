"""ctypes binding of oracle/liboracle.so (see oracle.h for each function's reference citation)."""
import ctypes as C
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


class OCand(C.Structure):
    _fields_ = [("frequency_hz", C.c_double), ("period_s", C.c_double), ("dm", C.c_double),
                ("harmonics", C.c_uint64), ("score", C.c_double)]


class OConfig(C.Structure):
    _fields_ = [("min_hz", C.c_double), ("max_hz", C.c_double), ("threshold", C.c_double),
                ("max_candidates", C.c_uint64), ("birdies", C.POINTER(C.c_double)), ("nbirdies", C.c_uint64),
                ("max_harmonics", C.c_uint32), ("pad_", C.c_uint32)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        dp, fp, up = C.POINTER(C.c_double), C.POINTER(C.c_float), C.POINTER(C.c_uint64)
        sig = {
            "oracle_delays": (C.c_int, [dp, C.c_size_t, C.c_double, C.c_double, up]),
            "oracle_dm_step": (C.c_int, [dp, C.c_size_t, C.c_double, dp]),
            "oracle_plan": (C.c_long, [C.c_double, C.c_double, C.c_double, dp, C.c_size_t]),
            "oracle_apply": (C.c_int, [fp, C.c_size_t, C.c_size_t, up, C.c_size_t, fp, C.POINTER(C.c_size_t)]),
            "oracle_decode": (C.c_int, [C.c_void_p, C.c_size_t, C.c_int, C.c_size_t, C.c_size_t, fp, C.POINTER(C.c_size_t)]),
            "oracle_median": (C.c_float, [fp, C.c_size_t]),
            "oracle_mask": (C.c_double, [fp, C.c_size_t, C.c_size_t, C.c_size_t, C.c_float, C.c_void_p, C.c_void_p]),
            "oracle_rfft64": (None, [dp, C.c_size_t, dp, dp]),
            "oracle_spectrum": (C.c_size_t, [fp, C.c_size_t, C.c_double, C.POINTER(OConfig), fp, fp]),
            "oracle_periodicity": (C.c_size_t, [fp, C.c_size_t, C.c_double, C.c_double, C.POINTER(OConfig), C.POINTER(OCand), C.c_size_t]),
            "oracle_periodicity_f64": (C.c_size_t, [fp, C.c_size_t, C.c_double, C.c_double, C.POINTER(OConfig), C.POINTER(OCand), C.c_size_t]),
            "oracle_sift": (C.c_size_t, [C.POINTER(OCand), C.c_size_t, C.c_double, C.c_size_t, C.c_int, C.POINTER(OCand)]),
            "oracle_profile_at": (None, [fp, C.c_size_t, C.c_double, C.c_double, C.c_size_t, C.c_size_t, dp, dp, fp, up]),
            "oracle_refine": (C.c_double, [fp, C.c_size_t, C.c_double, C.c_double, C.c_double, C.c_size_t, C.c_size_t, dp, dp, fp, up]),
            "oracle_search_trials": (C.c_int, [fp, C.c_size_t, dp, C.c_size_t, C.c_double, dp, C.c_size_t, C.c_size_t,
                                               C.POINTER(OConfig), C.POINTER(OCand), C.c_size_t, C.POINTER(C.c_uint32), C.c_int]),
            "oracle_search_trials_ex": (C.c_int, [fp, C.c_size_t, dp, C.c_size_t, C.c_double, dp, C.c_size_t,
                                                  C.c_size_t, C.POINTER(OConfig), C.POINTER(OCand), C.c_size_t,
                                                  C.POINTER(C.c_uint32), fp, C.c_int, C.c_int]),
            "oracle_median64": (C.c_double, [dp, C.c_size_t]),
            "oracle_spectrum_f64": (C.c_size_t, [fp, C.c_size_t, C.c_double, C.POINTER(OConfig), dp, dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _u(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint64))


def delays(freqs, tsamp, dm):
    f = np.ascontiguousarray(freqs, dtype=np.float64)
    out = np.zeros(len(f), dtype=np.uint64)
    if lib().oracle_delays(_d(f), len(f), tsamp, dm, _u(out)):
        raise ValueError("invalid dedispersion input")
    return out


def dm_step(freqs, tsamp):
    f = np.ascontiguousarray(freqs, dtype=np.float64)
    out = C.c_double()
    if lib().oracle_dm_step(_d(f), len(f), tsamp, C.byref(out)):
        raise ValueError("invalid dedispersion input")
    return out.value


def plan(lo, hi, step):
    cap = 100001
    out = np.zeros(cap, dtype=np.float64)
    n = lib().oracle_plan(lo, hi, step, _d(out), cap)
    if n < 0:
        raise ValueError("invalid DM bounds/step")
    return out[:n]


def apply(data, nsamples, dl, length=None):
    d = np.ascontiguousarray(data, dtype=np.float32).reshape(-1)
    dl = np.ascontiguousarray(dl, dtype=np.uint64)
    out = np.zeros(nsamples, dtype=np.float32)
    n = C.c_size_t()
    if lib().oracle_apply(_f(d), nsamples, len(dl), _u(dl), int(length or 0), _f(out), C.byref(n)):
        raise ValueError("invalid dedispersion input")
    return out[:n.value].copy()


def decode(payload, nbits, nchans, limit=0):
    p = np.ascontiguousarray(np.frombuffer(bytes(payload), dtype=np.uint8) if not isinstance(payload, np.ndarray) else payload)
    total = p.nbytes * 8 // (nchans * nbits) if nchans * nbits else 0
    ns_cap = min(limit, total) if limit else total
    out = np.empty(max(1, ns_cap * nchans), dtype=np.float32)
    ns = C.c_size_t()
    if lib().oracle_decode(p.ctypes.data, p.nbytes, nbits, nchans, int(limit), _f(out), C.byref(ns)):
        raise ValueError("invalid or unsupported filterbank")
    return out[:ns.value * nchans], ns.value


def median(v):
    a = np.ascontiguousarray(v, dtype=np.float32)
    return float(lib().oracle_median(_f(a), len(a)))


def mask(data, nsamples, block, threshold):
    """rfi::mask in place on a float32 channel-major array; returns (channels, blocks, fraction)."""
    assert data.dtype == np.float32 and data.flags.c_contiguous
    nch = data.size // nsamples
    bad = np.zeros(nch, dtype=np.uint8)
    nblk = (nsamples + block - 1) // block
    badb = np.zeros(nblk, dtype=np.uint8)
    frac = lib().oracle_mask(_f(data), nsamples, nch, block, threshold, bad.ctypes.data, badb.ctypes.data)
    return np.nonzero(bad)[0].tolist(), np.nonzero(badb)[0].tolist(), frac


def _cfg(min_hz=1.0, max_hz=math.inf, threshold=6.0, max_candidates=32, birdies=(), max_harmonics=16):
    b = np.ascontiguousarray(np.array(birdies, dtype=np.float64).reshape(-1))
    c = OConfig(min_hz, max_hz, threshold, max_candidates, _d(b) if b.size else None, len(birdies), max_harmonics, 0)
    return c, b


def _cands(arr, n):
    return [dict(frequency_hz=arr[i].frequency_hz, period_s=arr[i].period_s, dm=arr[i].dm,
                 harmonics=int(arr[i].harmonics), score=arr[i].score) for i in range(n)]


def periodicity(series, tsamp, dm, f64=False, **cfg):
    x = np.ascontiguousarray(series, dtype=np.float32)
    c, keep = _cfg(**cfg)
    cap = max(1, int(c.max_candidates))
    out = (OCand * cap)()
    fn = lib().oracle_periodicity_f64 if f64 else lib().oracle_periodicity
    n = fn(_f(x), len(x), tsamp, dm, C.byref(c), out, cap)
    del keep
    return _cands(out, n)


def spectrum(series, tsamp, **cfg):
    x = np.ascontiguousarray(series, dtype=np.float32)
    size = 1
    while size < len(x):
        size <<= 1
    nb = size + 1
    raw = np.zeros(nb, dtype=np.float32)
    white = np.zeros(nb, dtype=np.float32)
    c, keep = _cfg(**cfg)
    lib().oracle_spectrum(_f(x), len(x), tsamp, C.byref(c), _f(raw), _f(white))
    del keep
    return raw, white


def rfft64(x):
    a = np.ascontiguousarray(x, dtype=np.float64)
    N = len(a)
    re = np.zeros(N // 2 + 1)
    im = np.zeros(N // 2 + 1)
    lib().oracle_rfft64(_d(a), N, _d(re), _d(im))
    return re + 1j * im


def sift(cands, resolution_hz, limit, reject_dm_zero):
    n = len(cands)
    arr = (OCand * max(1, n))()
    for i, c in enumerate(cands):
        arr[i] = OCand(c["frequency_hz"], c["period_s"], c["dm"], c["harmonics"], c["score"])
    out = (OCand * max(1, n))()
    k = lib().oracle_sift(arr, n, resolution_hz, limit, int(bool(reject_dm_zero)), out)
    return _cands(out, k)


def profile_at(series, tsamp, period, bins, offset=0):
    x = np.ascontiguousarray(series, dtype=np.float32)
    snr, chi = C.c_double(), C.c_double()
    vals = np.zeros(bins, dtype=np.float32)
    cnts = np.zeros(bins, dtype=np.uint64)
    lib().oracle_profile_at(_f(x), len(x), tsamp, period, bins, offset, C.byref(snr), C.byref(chi), _f(vals), _u(cnts))
    return dict(period_s=period, snr=snr.value, reduced_chi_square=chi.value, bins=vals, counts=cnts)


def refine(series, tsamp, frequency, span_hz, steps, bins):
    x = np.ascontiguousarray(series, dtype=np.float32)
    snr, chi = C.c_double(), C.c_double()
    vals = np.zeros(bins, dtype=np.float32)
    cnts = np.zeros(bins, dtype=np.uint64)
    p = lib().oracle_refine(_f(x), len(x), tsamp, frequency, span_hz, steps, bins, C.byref(snr), C.byref(chi),
                            _f(vals), _u(cnts))
    return dict(period_s=p, snr=snr.value, reduced_chi_square=chi.value, bins=vals, counts=cnts)


def median64(v):
    a = np.ascontiguousarray(v, dtype=np.float64)
    return lib().oracle_median64(_d(a), len(a))


def spectrum_f64(series, tsamp, **cfg):
    """oracle_spectrum_f64: (raw, white) float64 arrays of next_pow2(n) + 1 bins."""
    x = np.ascontiguousarray(series, dtype=np.float32)
    size = 1
    while size < len(x):
        size <<= 1
    raw = np.zeros(size + 1, dtype=np.float64)
    white = np.zeros(size + 1, dtype=np.float64)
    c, keep = _cfg(**cfg)
    lib().oracle_spectrum_f64(_f(x), len(x), tsamp, C.byref(c), _d(raw), _d(white))
    del keep
    return raw, white


def search_trials(data, nsamples, freqs, tsamp, dms, length, threads=1, series=None, f64=False, **cfg):
    """The per-trial loop over f32 channel-major data -> list of candidate lists (trial order).

    `series`, when given, is a float32 array [len(dms), length] that receives
    each trial's dedispersed series; f64 selects the all-f64 periodicity."""
    d = np.ascontiguousarray(data, dtype=np.float32).reshape(-1)
    f = np.ascontiguousarray(freqs, dtype=np.float64)
    dm = np.ascontiguousarray(dms, dtype=np.float64)
    c, keep = _cfg(**cfg)
    cap = max(1, int(c.max_candidates))
    out = (OCand * (cap * len(dm)))()
    counts = np.zeros(len(dm), dtype=np.uint32)
    sp = None
    if series is not None:
        assert series.dtype == np.float32 and series.flags.c_contiguous and series.shape == (len(dm), length)
        sp = _f(series)
    rc = lib().oracle_search_trials_ex(_f(d), nsamples, _d(f), len(f), tsamp, _d(dm), len(dm), length,
                                       C.byref(c), out, cap, counts.ctypes.data_as(C.POINTER(C.c_uint32)), sp,
                                       1 if f64 else 0, threads)
    if rc:
        raise ValueError("invalid dedispersion input")
    del keep
    res = []
    for t in range(len(dm)):
        res.append([dict(frequency_hz=out[t * cap + i].frequency_hz, period_s=out[t * cap + i].period_s,
                         dm=out[t * cap + i].dm, harmonics=int(out[t * cap + i].harmonics), score=out[t * cap + i].score)
                    for i in range(int(counts[t]))])
    return res


# ------------------------------------------------ confirmation (analyze.c)
class OWindow(C.Structure):
    _fields_ = [("start_bin", C.c_uint64), ("width_bins", C.c_uint64), ("bins", C.c_uint64)]


class OMeasurement(C.Structure):
    _fields_ = [("snr", C.c_double), ("on_samples", C.c_uint64), ("off_samples", C.c_uint64),
                ("mean_difference", C.c_double), ("standard_error", C.c_double)]


class ORules(C.Structure):
    _fields_ = [("minimum_validation_snr", C.c_double), ("minimum_part_snr", C.c_double),
                ("minimum_time_parts", C.c_uint64), ("minimum_subbands", C.c_uint64),
                ("minimum_dm_contrast", C.c_double), ("minimum_dispersion_widths", C.c_double),
                ("maximum_zero_dm_spectral_ratio", C.c_double)]


class ODmPoint(C.Structure):
    _fields_ = [("dm", C.c_double), ("spectral_score", C.c_double), ("spectral_frequency_hz", C.c_double),
                ("discovery_profile_response", C.c_double), ("has_spectral", C.c_uint32), ("pad_", C.c_uint32)]


class OMember(C.Structure):
    _fields_ = [("family", C.c_uint32), ("pad_", C.c_uint32), ("index", C.c_uint64)]


class OConfirmed(C.Structure):
    _fields_ = [("candidate", OCand), ("refined_dm", C.c_double), ("profile_period_s", C.c_double),
                ("profile_snr", C.c_double), ("profile_chi2", C.c_double), ("window", OWindow),
                ("validation", OMeasurement), ("zero_dm_validation", OMeasurement), ("zero_dm_period_s", C.c_double),
                ("zero_dm_window", OWindow), ("dm_contrast", C.c_double), ("dispersion_widths", C.c_double),
                ("zero_dm_related_spectral_ratio", C.c_double), ("time_parts", OMeasurement * 4),
                ("subbands", OMeasurement * 4), ("nmembers", C.c_uint64), ("accepted", C.c_uint32),
                ("rejection_mask", C.c_uint32)]


class OOutcome(C.Structure):
    _fields_ = [("discovery_samples", C.c_uint64), ("validation_start_sample", C.c_uint64),
                ("validation_samples", C.c_uint64), ("common_samples", C.c_uint64), ("trial_count", C.c_uint64),
                ("ncandidates", C.c_uint64)]


# crates/search/src/confirmation.rs:434-466, in bit order of rejection_mask
REJECTION_REASONS = [
    "held-out pulse SNR below threshold",
    "pulse does not persist in enough held-out time segments",
    "pulse does not persist in enough held-out subbands",
    "insufficient improvement over zero-DM interference control",
    "dispersion sweep too small relative to selected pulse width to distinguish local interference",
    "stronger rationally related periodic signal in zero-DM discovery control",
]
DEFAULT_RULES = dict(minimum_validation_snr=6.0, minimum_part_snr=1.0, minimum_time_parts=3, minimum_subbands=3,
                     minimum_dm_contrast=1.25, minimum_dispersion_widths=1.5, maximum_zero_dm_spectral_ratio=1.0)


def _bind_confirmation(L):
    fp, up, dp = C.POINTER(C.c_float), C.POINTER(C.c_uint64), C.POINTER(C.c_double)
    L.oracle_measure.restype = None
    L.oracle_measure.argtypes = [fp, C.c_size_t, C.c_size_t, C.c_double, C.c_double, C.POINTER(OWindow),
                                 C.POINTER(OMeasurement)]
    L.oracle_select_window.restype = None
    L.oracle_select_window.argtypes = [fp, C.c_size_t, C.POINTER(OWindow)]
    L.oracle_group.restype = C.c_size_t
    L.oracle_group.argtypes = [C.POINTER(OCand), C.c_size_t, C.c_double, C.c_size_t, C.POINTER(OCand),
                               C.POINTER(OMember), C.POINTER(C.c_size_t)]
    L.oracle_analyze.restype = C.c_int
    L.oracle_analyze.argtypes = [fp, C.c_size_t, dp, C.c_size_t, C.c_double, dp, C.c_size_t, C.POINTER(OCand),
                                 C.c_size_t, C.c_size_t, C.c_size_t, C.POINTER(ORules), C.POINTER(OOutcome),
                                 C.POINTER(OConfirmed), fp, up, fp, fp, C.POINTER(ODmPoint), C.POINTER(OMember),
                                 C.POINTER(C.c_size_t), C.c_char_p, C.c_size_t]


def _clib():
    L = lib()
    if not getattr(L, "_conf_bound", False):
        _bind_confirmation(L)
        L._conf_bound = True
    return L


def _cand_array(cands):
    arr = (OCand * max(1, len(cands)))()
    for i, c in enumerate(cands):
        arr[i] = OCand(c["frequency_hz"], c["period_s"], c["dm"], int(c["harmonics"]), c["score"])
    return arr


def _meas(m):
    return dict(snr=m.snr, on_samples=int(m.on_samples), off_samples=int(m.off_samples),
                mean_difference=m.mean_difference, standard_error=m.standard_error)


def _win(w):
    return dict(start_bin=int(w.start_bin), width_bins=int(w.width_bins), bins=int(w.bins))


def measure(x, offset, tsamp, period, window):
    a = np.ascontiguousarray(x, dtype=np.float32)
    w = OWindow(window["start_bin"], window["width_bins"], window["bins"])
    m = OMeasurement()
    _clib().oracle_measure(_f(a), len(a), offset, tsamp, period, C.byref(w), C.byref(m))
    return _meas(m)


def select_window(values):
    a = np.ascontiguousarray(values, dtype=np.float32)
    w = OWindow()
    _clib().oracle_select_window(_f(a), len(a), C.byref(w))
    return _win(w)


def group(cands, resolution, limit):
    """-> (leaders, members): members lists (family, index into cands) in append order."""
    arr = _cand_array(cands)
    leaders = (OCand * max(1, limit))()
    members = (OMember * max(1, len(cands)))()
    nm = C.c_size_t()
    nf = _clib().oracle_group(arr, len(cands), resolution, limit, leaders, members, C.byref(nm))
    return _cands(leaders, nf), [(int(members[i].family), int(members[i].index)) for i in range(nm.value)]


def analyze(data, nsamples, freqs, tsamp, trials, generated, limit, bins, rules=None):
    """confirmation::analyze over f32 channel-major data -> the Outcome as a dict."""
    d = np.ascontiguousarray(data, dtype=np.float32).reshape(-1)
    f = np.ascontiguousarray(freqs, dtype=np.float64)
    tr = np.ascontiguousarray(trials, dtype=np.float64)
    r = dict(DEFAULT_RULES, **(rules or {}))
    R = ORules(**r)
    gen = _cand_array(generated)
    out = OOutcome()
    conf = (OConfirmed * limit)()
    pb = np.zeros(limit * bins, np.float32)
    pc = np.zeros(limit * bins, np.uint64)
    tp = np.zeros(limit * 4 * bins, np.float32)
    sp = np.zeros(limit * 4 * bins, np.float32)
    dmr = (ODmPoint * (limit * len(tr)))()
    mem = (OMember * max(1, len(generated)))()
    nm = C.c_size_t()
    err = C.create_string_buffer(256)
    rc = _clib().oracle_analyze(_f(d), nsamples, _d(f), len(f), tsamp, _d(tr), len(tr), gen, len(generated), limit,
                                bins, C.byref(R), C.byref(out), conf, _f(pb), _u(pc), _f(tp), _f(sp), dmr, mem,
                                C.byref(nm), err, 256)
    if rc:
        raise ValueError(err.value.decode())
    members = [(int(mem[i].family), int(mem[i].index)) for i in range(nm.value)]
    cands = []
    for k in range(int(out.ncandidates)):
        o = conf[k]
        cands.append(dict(
            candidate=dict(frequency_hz=o.candidate.frequency_hz, period_s=o.candidate.period_s,
                           dm=o.candidate.dm, harmonics=int(o.candidate.harmonics), score=o.candidate.score),
            refined_dm=o.refined_dm,
            profile=dict(period_s=o.profile_period_s, snr=o.profile_snr, reduced_chi_square=o.profile_chi2,
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
                         for p in dmr[k * len(tr):(k + 1) * len(tr)]],
            harmonic_members=[generated[i] for fam, i in members if fam == k],
            accepted=bool(o.accepted),
            rejection_reasons=[REJECTION_REASONS[b] for b in range(6) if o.rejection_mask >> b & 1]))
    return dict(discovery_samples=int(out.discovery_samples), validation_start_sample=int(out.validation_start_sample),
                validation_samples=int(out.validation_samples), common_samples=int(out.common_samples),
                trial_count=int(out.trial_count), candidates=cands)
