"""crates/fold on the GPU: profile, profile_at, refine."""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _ffi
from .core import check, default_context


@dataclass
class Profile:
    """fold::Profile (crates/fold/src/lib.rs:3-11)."""
    period_s: float
    snr: float
    reduced_chi_square: float
    bins: np.ndarray
    counts: np.ndarray


def profiles_at(series, tsamp, periods, bins, offsets=None, ctx=None):
    """fold::profile_at for many (period, sample_offset) pairs (lib.rs:16-80)."""
    ctx = ctx or default_context()
    x = np.ascontiguousarray(series, dtype=np.float32)
    p = np.ascontiguousarray(periods, dtype=np.float64)
    o = np.ascontiguousarray(offsets if offsets is not None else np.zeros(len(p)), dtype=np.uint64)
    st = (_ffi.ProfileStats * max(1, len(p)))()
    vals = np.zeros((len(p), bins), dtype=np.float32)
    cnts = np.zeros((len(p), bins), dtype=np.uint64)
    check(_ffi.lib().psr_fold(ctx.handle, x.ctypes.data_as(C.POINTER(C.c_float)), len(x), float(tsamp),
                              p.ctypes.data_as(C.POINTER(C.c_double)), o.ctypes.data_as(C.POINTER(C.c_uint64)),
                              len(p), int(bins), st, vals.ctypes.data_as(C.POINTER(C.c_float)),
                              cnts.ctypes.data_as(C.POINTER(C.c_uint64))))
    return [Profile(st[i].period_s, st[i].snr, st[i].reduced_chi_square, vals[i], cnts[i]) for i in range(len(p))]


def profile_at(series, tsamp, period, bins, sample_offset, ctx=None):
    return profiles_at(series, tsamp, [period], bins, [sample_offset], ctx)[0]


def profile(series, tsamp, period, bins, ctx=None):
    """fold::profile (crates/fold/src/lib.rs:12-14)."""
    return profile_at(series, tsamp, period, bins, 0, ctx)


def refine(series, tsamp, frequency, span_hz, steps, bins, ctx=None):
    """fold::refine (crates/fold/src/lib.rs:81-97): the last maximum chi-square wins."""
    ctx = ctx or default_context()
    x = np.ascontiguousarray(series, dtype=np.float32)
    st = _ffi.ProfileStats()
    vals = np.zeros(bins, dtype=np.float32)
    cnts = np.zeros(bins, dtype=np.uint64)
    check(_ffi.lib().psr_refine(ctx.handle, x.ctypes.data_as(C.POINTER(C.c_float)), len(x), float(tsamp),
                                float(frequency), float(span_hz), int(steps), int(bins), C.byref(st),
                                vals.ctypes.data_as(C.POINTER(C.c_float)), cnts.ctypes.data_as(C.POINTER(C.c_uint64))))
    return Profile(st.period_s, st.snr, st.reduced_chi_square, vals, cnts)
