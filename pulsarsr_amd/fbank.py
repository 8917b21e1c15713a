"""SIGPROC filterbank files (crates/fbank): strict header parse, packed payload."""
import ctypes as C
import mmap
from dataclasses import dataclass

import numpy as np

from . import _ffi
from .core import check
from .dedisperse import Observation


@dataclass
class Header:
    nchans: int
    nbits: int
    tsamp: float
    fch1: float
    foff: float
    data_offset: int
    nsamples: int

    def frequencies(self):
        return self.fch1 + np.arange(self.nchans, dtype=np.float64) * self.foff


def parse_header(data):
    """fbank::parse_header (crates/fbank/src/lib.rs:72-147) via the C ABI."""
    h = _ffi.Header()
    b = bytes(data[:65536]) if len(data) > 65536 else bytes(data)
    rc = _ffi.lib().psr_parse_header(b, len(b), C.byref(h))
    check(rc)
    # nsamples counts whole frames of the full payload
    frame_bits = h.nchans * h.nbits
    ns = (len(data) - h.data_offset) * 8 // frame_bits
    return Header(h.nchans, h.nbits, h.tsamp, h.fch1, h.foff, h.data_offset, ns)


def read_packed(path):
    """(Header, payload ndarray[uint8]) of a SIGPROC file, memory-mapped."""
    with open(path, "rb") as f:
        mm = mmap.mmap(f.fileno(), 0, access=mmap.ACCESS_READ)
    hdr = parse_header(memoryview(mm))
    payload = np.frombuffer(mm, dtype=np.uint8, offset=hdr.data_offset)
    return hdr, payload


def observe(path, limit=None, ctx=None):
    """Read a SIGPROC file straight into a device-resident Observation."""
    hdr, payload = read_packed(path)
    frame_bits = hdr.nchans * hdr.nbits
    if (payload.nbytes * 8) % frame_bits:
        from .core import PsrError
        raise PsrError(_ffi.PSR_E_FORMAT, "invalid or unsupported filterbank: truncated frame or packed padding unsupported")
    return hdr, Observation.from_packed(payload, hdr.nbits, hdr.nchans, hdr.tsamp, hdr.fch1, hdr.foff,
                                        nsamples=limit or 0, ctx=ctx)
