"""psr_obs_mask (rfi::mask on the GPU) against the oracle's restatement.

The oracle's mask reproduces the GBT report's channels, blocks and masked
fraction exactly (tests/test_oracle_golden.py).  The GPU mask must agree on
the channel and block lists and the fraction exactly, and leave the samples
bit-identical to the oracle's masked f32 array: the reference's order-dependent
f32 sums are made in the same order, and the fill value is the exact median.
"""
import numpy as np
import pytest

import oracle
from pulsarsr_amd import dedisperse, rfi, synth

pytestmark = pytest.mark.gpu


def check_against_oracle(obs, data, ns, block, threshold):
    got = rfi.mask(obs, block, threshold)
    ref = data.copy()
    chans, blocks, frac = oracle.mask(ref, ns, block, threshold)
    print("mask: %d channels x %d samples, block %d: channels %s blocks %s"
          % (obs.nchans, ns, block, chans[:8], blocks[:8]), flush=True)
    assert got["channels"] == chans and got["blocks"] == blocks
    assert got["masked_fraction"] == frac and got["block_samples"] == block
    after = obs.download()
    assert (after.view(np.uint32) == ref.view(np.uint32)).all()
    return got


def test_mask_u8_planted_channels_and_bursts(ctx):
    w = synth.Workload("rfi-8bit", 5, 1 << 15, 64, 8, 64e-6, 1500.0, -2.0, 128.0, 16.0,
                       pulsars=[(0.0123, 40.0, 0.05, 0.2, 0.3)], bursts=[(0.5, 0.02, 3.0), (1.6, 0.01, 4.0)],
                       rfi_chans=[3, 17, 40], rfi_amp=3.0, dms=np.arange(0, 20.0, 5.0), min_hz=10.0)
    payload = synth.generate(w)
    obs = dedisperse.Observation.from_packed(payload, w.nbits, w.nchans, w.tsamp, w.fch1, w.foff)
    data, ns = oracle.decode(payload, w.nbits, w.nchans)
    got = check_against_oracle(obs, data, ns, int(max(0.1 / w.tsamp, 32)), 8.0)
    assert got["channels"] and got["blocks"]
    # the masked observation dedisperses exactly like the oracle's masked array
    ref = data.copy()
    oracle.mask(ref, ns, int(max(0.1 / w.tsamp, 32)), 8.0)
    x = dedisperse.apply_dms(obs, [12.5], 20000)[0]
    assert (x == oracle.apply(ref, ns, oracle.delays(w.frequencies(), w.tsamp, 12.5), 20000)).all()


def test_mask_reference_unit_cases_on_gpu(ctx):
    # crates/rfi/src/lib.rs:100-129, through the device path (f32 samples)
    n, chans = 4000, 16
    i = np.arange(n * chans, dtype=np.uint64)
    data = (100.0 + ((i * np.uint64(2654435761)) % np.uint64(1000)).astype(np.float32) / 500.0 - 1.0).astype(np.float32)
    for c in range(chans):
        data[c * n + 1600:c * n + 1680] += 30.0
    obs = dedisperse.Observation.from_f32(data, n, 64e-6, 1500.0, -2.0)
    got = check_against_oracle(obs, data, n, 400, 8.0)
    assert 4 in got["blocks"]
    n = 1000
    data = (10.0 + (np.arange(n * 16) % 7)).astype(np.float32)
    data[:n] = 1000.0
    for c in range(1, 16):
        data[c * n + 400:c * n + 450] += 500.0
    obs = dedisperse.Observation.from_f32(data, n, 64e-6, 1500.0, -2.0)
    got = check_against_oracle(obs, data, n, 50, 6.0)
    assert 0 in got["channels"] and 8 in got["blocks"] and got["masked_fraction"] > 0.05


@pytest.mark.parametrize("seed", [1, 2])
def test_mask_f32_noninteger_and_u16(ctx, seed):
    rng = np.random.default_rng(seed)
    nch, n = 40, 20000
    data = rng.normal(50.0, 3.0, (nch, n)).astype(np.float32)
    data[7] *= 4.0                       # a noisy channel
    data[:, 9000:9300] += 2.5            # a broadband burst
    data[21] = 50.0                      # a dead (constant) channel
    flat = data.reshape(-1)
    obs = dedisperse.Observation.from_f32(flat, n, 64e-6, 1500.0, -2.0)
    got = check_against_oracle(obs, flat, n, 250, 6.0)
    assert 7 in got["channels"] and 21 in got["channels"] and got["blocks"]
    # 16-bit samples (packed little-endian, time-major)
    v = np.clip(np.rint(rng.normal(30000.0, 3000.0, (n, nch))), 0, 65535).astype(np.uint16)
    v[:, 5] = 1000
    obs16 = dedisperse.Observation.from_packed(v.tobytes(), 16, nch, 64e-6, 1500.0, -2.0)
    d16, ns = oracle.decode(v.tobytes(), 16, nch)
    check_against_oracle(obs16, d16, ns, 500, 6.0)


def test_mask_contract_errors(ctx):
    obs = dedisperse.Observation.from_f32(np.ones(64 * 4, np.float32), 64, 64e-6, 1500.0, -2.0)
    with pytest.raises(Exception, match="block_samples > 0"):
        rfi.mask(obs, 0, 8.0)
    with pytest.raises(Exception, match="finite positive threshold"):
        rfi.mask(obs, 16, float("nan"))
