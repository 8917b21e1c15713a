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


def spectrum(series, tsamp, config=None, ctx=None, m2=0):
    """(raw |X|^2, whitened power) of periodicity's spectral stages (lib.rs:35-72).
    m2 = 512, 1024 or 2048 forces that four-step FFT split (psr_spectrum_split)."""
    config = config or Config()
    ctx = ctx or default_context()
    x = np.ascontiguousarray(series, dtype=np.float32)
    nb = C.c_size_t()
    cfg, keep = config._c()
    lib = _ffi.lib()

    def call(raw, white):
        xp = x.ctypes.data_as(C.POINTER(C.c_float))
        if m2:
            return lib.psr_spectrum_split(ctx.handle, xp, len(x), float(tsamp), C.byref(cfg), int(m2), raw, white,
                                          C.byref(nb))
        return lib.psr_spectrum(ctx.handle, xp, len(x), float(tsamp), C.byref(cfg), raw, white, C.byref(nb))

    check(call(None, None))
    raw = np.zeros(nb.value, dtype=np.float32)
    white = np.zeros(nb.value, dtype=np.float32)
    check(call(raw.ctypes.data_as(C.POINTER(C.c_float)), white.ctypes.data_as(C.POINTER(C.c_float))))
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


class Batch:
    """psr_batch: the per-trial search of many device-resident observations in
    one device schedule (one launch per stage per round, a persistent
    dedispersion kernel over per-XCD work queues, optionally captured in a
    hipGraph and tuned on the batch's own data).

    observations: Observation list; dms: one DM array per observation;
    lengths: series length per observation (None/0 = the full available length).
    run() returns, per observation, its candidate lists in its trial order."""

    def __init__(self, observations, dms, config, lengths=None, cap_per_trial=None, graph=True, autotune=False,
                 memory_budget=0, ctx=None):
        self.ctx = ctx or observations[0].ctx
        self.obs = list(observations)
        self._dms = [np.ascontiguousarray(d, dtype=np.float64) for d in dms]
        self.ntrials = [len(d) for d in self._dms]
        self.cap = int(cap_per_trial or max(1, config.max_candidates))
        n = len(self.obs)
        handles = (C.c_void_p * max(1, n))(*[o.handle for o in self.obs])
        dptr = (C.POINTER(C.c_double) * max(1, n))(*[d.ctypes.data_as(C.POINTER(C.c_double)) for d in self._dms])
        nt = (C.c_size_t * max(1, n))(*self.ntrials)
        ln = (C.c_size_t * max(1, n))(*[int(x or 0) for x in (lengths or [0] * n)])
        opts = _ffi.BatchOpts((_ffi.PSR_BATCH_GRAPH if graph else 0) | (_ffi.PSR_BATCH_AUTOTUNE if autotune else 0), 0,
                              int(memory_budget))
        cfg, keep = config._c()
        self._h = C.c_void_p()
        check(_ffi.lib().psr_batch_create(self.ctx.handle, handles, n, dptr, nt, ln, C.byref(cfg), self.cap,
                                          C.byref(opts), C.byref(self._h)))
        del keep
        self.rows = sum(self.ntrials)
        self._out = (_ffi.Candidate * max(1, self.rows * self.cap))()
        self._counts = np.zeros(max(1, self.rows), dtype=np.uint32)

    def run_raw(self):
        """Run once; returns (ctypes candidate array, counts) in row order."""
        check(_ffi.lib().psr_batch_run(self._h, self._out, self._counts.ctypes.data_as(C.POINTER(C.c_uint32))))
        return self._out, self._counts

    def run(self):
        out, counts = self.run_raw()
        res, row = [], 0
        for k in self.ntrials:
            lists = []
            for t in range(k):
                lists.append([Candidate(out[(row + t) * self.cap + i].frequency_hz,
                                        out[(row + t) * self.cap + i].period_s, out[(row + t) * self.cap + i].dm,
                                        int(out[(row + t) * self.cap + i].harmonics),
                                        out[(row + t) * self.cap + i].score) for i in range(int(counts[row + t]))])
            res.append(lists)
            row += k
        return res

    def info(self):
        import json
        n = C.c_size_t()
        check(_ffi.lib().psr_batch_info(self._h, None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value + 1)
        check(_ffi.lib().psr_batch_info(self._h, buf, n.value + 1, C.byref(n)))
        return json.loads(buf.value.decode())

    def close(self):
        if self._h:
            _ffi.lib().psr_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
