"""crates/dedisperse on the GPU: delays/dm_step/plan (host FP64) and apply.

Errors raise PsrError whose text mirrors dedisperse::Error
(crates/dedisperse/src/lib.rs:5-7).
"""
import ctypes as C

import numpy as np

from . import _ffi
from .core import PsrError, check, default_context

DISPERSION_SECONDS_MHZ = 4148.808  # crates/dedisperse/src/lib.rs:3


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def delays(frequencies, tsamp, dm):
    """dedisperse::delays (crates/dedisperse/src/lib.rs:9-34) -> uint64 array."""
    f = np.ascontiguousarray(frequencies, dtype=np.float64)
    out = np.zeros(len(f), dtype=np.uint64)
    check(_ffi.lib().psr_delays(_dp(f), len(f), float(tsamp), float(dm),
                                out.ctypes.data_as(C.POINTER(C.c_uint64))))
    return out


def dm_step(frequencies, tsamp):
    """dedisperse::dm_step (crates/dedisperse/src/lib.rs:35-44)."""
    f = np.ascontiguousarray(frequencies, dtype=np.float64)
    out = C.c_double()
    check(_ffi.lib().psr_dm_step(_dp(f), len(f), float(tsamp), C.byref(out)))
    return out.value


def plan(dm_min, dm_max, step):
    """dedisperse::plan (crates/dedisperse/src/lib.rs:45-60)."""
    n = C.c_size_t()
    check(_ffi.lib().psr_plan(float(dm_min), float(dm_max), float(step), None, 0, C.byref(n)))
    out = np.zeros(n.value, dtype=np.float64)
    check(_ffi.lib().psr_plan(float(dm_min), float(dm_max), float(step), _dp(out), len(out), C.byref(n)))
    return out


def frequencies(fch1, foff, nchans):
    """fbank::Header::frequencies (crates/fbank/src/lib.rs:40-44)."""
    out = np.zeros(nchans, dtype=np.float64)
    check(_ffi.lib().psr_frequencies(float(fch1), float(foff), nchans, _dp(out)))
    return out


class Observation:
    """A filterbank resident in HBM (channel-major), the `data` of apply()."""

    def __init__(self, handle, ctx, nchans, nsamples, tsamp, fch1, foff, nbits):
        self._h = handle
        self.ctx = ctx
        self.nchans = nchans
        self.nsamples = nsamples
        self.tsamp = tsamp
        self.fch1 = fch1
        self.foff = foff
        self.nbits = nbits

    @property
    def handle(self):
        return self._h

    @property
    def frequencies(self):
        return frequencies(self.fch1, self.foff, self.nchans)

    @classmethod
    def from_packed(cls, payload, nbits, nchans, tsamp, fch1, foff, nsamples=0, ctx=None):
        """Upload a packed SIGPROC payload (time-major, LSB-first below 8 bits)."""
        ctx = ctx or default_context()
        buf = np.frombuffer(payload, dtype=np.uint8) if not isinstance(payload, np.ndarray) else payload
        buf = np.ascontiguousarray(buf.view(np.uint8).reshape(-1))
        fb = _ffi.Filterbank(buf.ctypes.data, buf.nbytes, nbits, nchans, nsamples, tsamp, fch1, foff)
        h = C.c_void_p()
        check(_ffi.lib().psr_obs_create(ctx.handle, C.byref(fb), C.byref(h)))
        frame_bits = nchans * nbits
        total = buf.nbytes * 8 // frame_bits
        ns = min(nsamples, total) if nsamples else total
        return cls(h, ctx, nchans, ns, tsamp, fch1, foff, nbits)

    @classmethod
    def from_f32(cls, data, nsamples, tsamp, fch1, foff, ctx=None):
        """Upload f32 channel-major data: data[channel * nsamples + t]."""
        ctx = ctx or default_context()
        a = np.ascontiguousarray(data, dtype=np.float32).reshape(-1)
        if nsamples == 0 or a.size % nsamples:
            raise PsrError(_ffi.PSR_E_SHAPE, "invalid dedispersion input: array shape mismatch")
        nch = a.size // nsamples
        h = C.c_void_p()
        check(_ffi.lib().psr_obs_create_f32(ctx.handle, a.ctypes.data_as(C.POINTER(C.c_float)),
                                            nsamples, nch, tsamp, fch1, foff, C.byref(h)))
        return cls(h, ctx, nch, nsamples, tsamp, fch1, foff, 32)

    def download(self):
        out = np.zeros(self.nchans * self.nsamples, dtype=np.float32)
        check(_ffi.lib().psr_obs_download_f32(self._h, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def close(self):
        if self._h:
            _ffi.lib().psr_obs_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def apply_dms(obs, dms, length=None):
    """delays + apply for every DM on the GPU -> float32 [ntrials, n]."""
    d = np.ascontiguousarray(dms, dtype=np.float64)
    n = C.c_size_t()
    check(_ffi.lib().psr_dedisperse(obs.ctx.handle, obs.handle, _dp(d), len(d), int(length or 0),
                                    None, C.byref(n)))
    out = np.zeros((len(d), n.value), dtype=np.float32)
    check(_ffi.lib().psr_dedisperse(obs.ctx.handle, obs.handle, _dp(d), len(d), int(length or 0),
                                    out.ctypes.data_as(C.POINTER(C.c_float)), C.byref(n)))
    return out


def apply_delays(obs, delays, length=None, channels=None):
    """dedisperse::apply(data, nsamples, delays, length) (lib.rs:62-90), one to one.

    `data` is the device observation restricted to the channel range
    `channels` = (lo, hi) (default: all channels); `delays` has one entry per
    channel of that range.  Returns the float32 channel-ordered sums."""
    lo, hi = channels if channels is not None else (0, obs.nchans)
    d = np.ascontiguousarray(delays, dtype=np.uint64)
    if len(d) != max(hi - lo, 0):
        raise PsrError(_ffi.PSR_E_SHAPE, "invalid dedispersion input: array shape mismatch")
    dp = d.ctypes.data_as(C.POINTER(C.c_uint64))
    n = C.c_size_t()
    check(_ffi.lib().psr_apply(obs.ctx.handle, obs.handle, dp, int(lo), int(hi), int(length or 0), None,
                               C.byref(n)))
    out = np.zeros(n.value, dtype=np.float32)
    check(_ffi.lib().psr_apply(obs.ctx.handle, obs.handle, dp, int(lo), int(hi), int(length or 0),
                               out.ctypes.data_as(C.POINTER(C.c_float)), C.byref(n)))
    return out


def apply(obs, dm, length=None):
    """dedisperse::apply(delays(freqs, tsamp, dm)) for one DM (lib.rs:62-90)."""
    return apply_dms(obs, [dm], length)[0]
