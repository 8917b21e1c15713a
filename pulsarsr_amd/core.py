"""Context and error handling around libpsrgpu.so."""
import ctypes as C

from . import _ffi


class PsrError(RuntimeError):
    """A non-zero status from the C ABI; ``code`` is the PSR_E_* value."""

    def __init__(self, code, message):
        super().__init__(message)
        self.code = code


def check(rc):
    if rc != _ffi.PSR_OK:
        msg = _ffi.lib().psr_last_error().decode(errors="replace")
        raise PsrError(rc, msg or "error %d" % rc)
    return rc


class Context:
    """A psr_ctx: its HIP streams and the device buffers reused across calls.

    ``overlap=True`` (PSR_CTX_OVERLAP) makes psr_search dedisperse the next
    trial batch on a second stream while the current batch's spectral stages
    run; by default batches run one after another and each reported stage time
    is that stage's own kernel time.
    """

    def __init__(self, device=0, overlap=False):
        self._h = C.c_void_p()
        check(_ffi.lib().psr_ctx_create(device, _ffi.PSR_CTX_OVERLAP if overlap else 0, C.byref(self._h)))

    @property
    def handle(self):
        if not self._h:
            raise PsrError(_ffi.PSR_E_INVALID, "context is closed")
        return self._h

    def set_memory_cap(self, nbytes):
        check(_ffi.lib().psr_ctx_set_memory_cap(self.handle, int(nbytes)))

    def stage_times(self):
        t = _ffi.StageTimes()
        check(_ffi.lib().psr_last_stage_times(self.handle, C.byref(t)))
        return {f: getattr(t, f) for f, _ in _ffi.StageTimes._fields_}

    def close(self):
        if self._h:
            _ffi.lib().psr_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default = None


def default_context():
    global _default
    if _default is None:
        _default = Context(0)
    return _default
