"""pulsarsr_amd -- MI355X (gfx950) search hot path of PulsarSr behind a C ABI.

Python mirror of the reference crates' interface for the path
(crates/dedisperse, crates/search incl. confirmation.rs, crates/fold, crates/rfi), calling
libpsrgpu.so.
There is no CPU fallback: GPU calls raise PsrError when no gfx950 device or
no built library is present.
"""
from ._ffi import PRECISION_FP16, PRECISION_FP32  # noqa: F401
from .core import Context, PsrError, default_context  # noqa: F401
from . import confirmation, dedisperse, fbank, fold, rfi, search, synth  # noqa: F401

__all__ = ["Context", "PsrError", "default_context", "confirmation", "dedisperse", "fbank", "fold", "rfi", "search", "synth",
           "PRECISION_FP32", "PRECISION_FP16"]
