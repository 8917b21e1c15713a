"""crates/search on the GPU: periodicity (1:1), the batched per-trial search, sift."""
import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _ffi
from .core import check, default_context


@dataclass
class Config:
    """search::Config (crates/search/src/lib.rs:14-32) plus max_harmonics/precision."""
    min_hz: float = 1.0
    max_hz: float = math.inf
    threshold: float = 6.0
    max_candidates: int = 32
    birdies: list = field(default_factory=list)
    max_harmonics: int = 16
    precision: int = _ffi.PRECISION_FP32

    def _c(self):
        b = np.ascontiguousarray(np.array(self.birdies, dtype=np.float64).reshape(-1))
        cfg = _ffi.SearchCfg(self.min_hz, self.max_hz, self.threshold, self.max_candidates,
                             b.ctypes.data_as(C.POINTER(C.c_double)) if b.size else None,
                             len(self.birdies), self.max_harmonics, self.precision)
        return cfg, b


@dataclass
class Candidate:
    """search::Candidate (crates/search/src/lib.rs:6-13)."""
    frequency_hz: float
    period_s: float
    dm: float
    harmonics: int
    score: float


def _cands(arr, n):
    return [Candidate(arr[i].frequency_hz, arr[i].period_s, arr[i].dm, int(arr[i].harmonics), arr[i].score)
            for i in range(n)]


def periodicity(series, tsamp, dm, config=None, ctx=None):
    """search::periodicity (crates/search/src/lib.rs:33-145) on the GPU."""
    config = config or Config()
    ctx = ctx or default_context()
    x = np.ascontiguousarray(series, dtype=np.float32)
    cap = max(1, int(config.max_candidates))
    out = (_ffi.Candidate * cap)()
    n = C.c_size_t()
    cfg, keep = config._c()
    check(_ffi.lib().psr_periodicity(ctx.handle, x.ctypes.data_as(C.POINTER(C.c_float)), len(x), float(tsamp),
                                     float(dm), C.byref(cfg), out, cap, C.byref(n)))
    del keep
    return _cands(out, n.value)


def periodicity_noalloc(series, tsamp, dm, config=None, ctx=None):
    """periodicity through psr_periodicity_noalloc: no device allocation (static
    per-device scratch; series of at most 131072 samples, 256 candidates)."""
    config = config or Config()
    ctx = ctx or default_context()
    x = np.ascontiguousarray(series, dtype=np.float32)
    cap = max(1, int(config.max_candidates))
    out = (_ffi.Candidate * cap)()
    n = C.c_size_t()
    cfg, keep = config._c()
    check(_ffi.lib().psr_periodicity_noalloc(ctx.handle, x.ctypes.data_as(C.POINTER(C.c_float)), len(x),
                                             float(tsamp), float(dm), C.byref(cfg), out, cap, C.byref(n)))
    del keep
    return _cands(out, n.value)


def spectrum(series, tsamp, config=None, ctx=None):
    """(raw |X|^2, whitened power) of periodicity's spectral stages (lib.rs:35-72)."""
    config = config or Config()
    ctx = ctx or default_context()
    x = np.ascontiguousarray(series, dtype=np.float32)
    nb = C.c_size_t()
    cfg, keep = config._c()
    check(_ffi.lib().psr_spectrum(ctx.handle, x.ctypes.data_as(C.POINTER(C.c_float)), len(x), float(tsamp),
                                  C.byref(cfg), None, None, C.byref(nb)))
    raw = np.zeros(nb.value, dtype=np.float32)
    white = np.zeros(nb.value, dtype=np.float32)
    check(_ffi.lib().psr_spectrum(ctx.handle, x.ctypes.data_as(C.POINTER(C.c_float)), len(x), float(tsamp),
                                  C.byref(cfg), raw.ctypes.data_as(C.POINTER(C.c_float)),
                                  white.ctypes.data_as(C.POINTER(C.c_float)), C.byref(nb)))
    del keep
    return raw, white


def search_raw(obs, dms, length, config, cap_per_trial=None):
    """The per-trial loop (crates/cli/src/main.rs:462-469) -> (Candidate ctypes array, counts)."""
    d = np.ascontiguousarray(dms, dtype=np.float64)
    cap = int(cap_per_trial or max(1, config.max_candidates))
    out = (_ffi.Candidate * (cap * len(d)))()
    counts = np.zeros(len(d), dtype=np.uint32)
    cfg, keep = config._c()
    check(_ffi.lib().psr_search(obs.ctx.handle, obs.handle, d.ctypes.data_as(C.POINTER(C.c_double)), len(d),
                                int(length or 0), C.byref(cfg), out, cap,
                                counts.ctypes.data_as(C.POINTER(C.c_uint32))))
    del keep
    return out, counts, cap


def search(obs, dms, length, config):
    """Candidates per trial, in trial order, as the reference's `lists` (main.rs:462-473)."""
    out, counts, cap = search_raw(obs, dms, length, config)
    res = []
    for t in range(len(counts)):
        res.append([Candidate(out[t * cap + i].frequency_hz, out[t * cap + i].period_s, out[t * cap + i].dm,
                              int(out[t * cap + i].harmonics), out[t * cap + i].score)
                    for i in range(int(counts[t]))])
    return res


def sift(candidates, resolution_hz, limit, reject_dm_zero):
    """search::sift (crates/search/src/lib.rs:147-176)."""
    n = len(candidates)
    arr = (_ffi.Candidate * max(n, 1))()
    for i, c in enumerate(candidates):
        arr[i] = _ffi.Candidate(c.frequency_hz, c.period_s, c.dm, c.harmonics, c.score)
    out = (_ffi.Candidate * max(n, 1))()
    k = C.c_size_t()
    check(_ffi.lib().psr_sift(arr, n, float(resolution_hz), int(limit), int(bool(reject_dm_zero)), out,
                              C.byref(k)))
    return _cands(out, k.value)
