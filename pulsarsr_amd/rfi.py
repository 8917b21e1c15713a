"""crates/rfi on the GPU: the channel and broadband-block mask, in place on an Observation."""
import ctypes as C

import numpy as np

from . import _ffi
from .core import check


def mask(obs, block_samples, threshold):
    """rfi::mask (crates/rfi/src/lib.rs:31-96) applied to the device-resident samples.

    Returns rfi::Mask as a dict: channels, blocks, block_samples, masked_fraction."""
    nblocks = -(-obs.nsamples // int(block_samples)) if block_samples else 0
    chans = np.zeros(max(1, obs.nchans), np.uint64)
    blocks = np.zeros(max(1, nblocks), np.uint64)
    m = _ffi.Mask()
    check(_ffi.lib().psr_obs_mask(obs.ctx.handle, obs.handle, int(block_samples), float(threshold), C.byref(m),
                                  chans.ctypes.data_as(C.POINTER(C.c_uint64)),
                                  blocks.ctypes.data_as(C.POINTER(C.c_uint64))))
    return dict(channels=chans[:m.nchannels].tolist(), blocks=blocks[:m.nblocks].tolist(),
                block_samples=int(m.block_samples), masked_fraction=m.masked_fraction)
