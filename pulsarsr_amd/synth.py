"""Seeded synthetic workloads (SIGPROC packed filterbanks) for tests and benchmarks.

The reference generator (crates/synth/src/lib.rs:38-85) only writes 32-bit
floats with one pulsar; these workloads need 1/2/4/8-bit data with several
pulsars, RFI and red noise, so the project generates them itself
(csrc/synth.cpp).  Pulse amplitudes come from a target SNR through the
reference's strength proxy A = SNR / sqrt(nch * N * (duty/2.355) * sqrt(pi))
(crates/synth/examples/recovery.rs:82-89).
"""
import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _ffi


@dataclass
class Workload:
    name: str
    seed: int
    nsamples: int
    nchans: int
    nbits: int
    tsamp: float
    fch1: float
    foff: float
    mean: float
    sigma: float
    pulsars: list = field(default_factory=list)  # (period_s, dm, duty, amplitude, phase)
    bursts: list = field(default_factory=list)   # (start_s, width_s, amplitude)
    sine: tuple = (0.0, 0.0)                     # (hz, amplitude)
    rfi_chans: list = field(default_factory=list)
    rfi_amp: float = 0.0
    red_noise: float = 0.0                       # common-mode amplitude at 1 Hz (sigma units)
    dms: np.ndarray = None
    min_hz: float = 15.0
    max_hz: float = 2000.0
    threshold: float = 8.0
    max_harmonics: int = 16
    top: int = 5
    rfi_mask: bool = False

    @property
    def payload_bytes(self):
        return (self.nsamples * self.nchans * self.nbits + 7) // 8

    def frequencies(self):
        return self.fch1 + np.arange(self.nchans, dtype=np.float64) * self.foff


def amplitude_for_snr(snr, nchans, nsamples, duty):
    return snr / math.sqrt(nchans * nsamples * (duty / 2.355) * math.sqrt(math.pi))


def _red_noise(seed, nsamples, tsamp, amp):
    rng = np.random.default_rng(seed ^ 0x5EED)
    w = rng.standard_normal(nsamples)
    spec = np.fft.rfft(w)
    f = np.fft.rfftfreq(nsamples, tsamp)
    gain = np.zeros_like(f)
    gain[1:] = amp * np.sqrt(1.0 / f[1:])
    return np.fft.irfft(spec * gain, n=nsamples).astype(np.float32)


def generate(w, threads=0, device=False):
    """Packed payload (uint8 array) of workload `w`.

    device=True generates it on the GPU (csrc/k_synth.hip: the same model,
    deterministic in the seed, much faster for large batches; 1/2/4/8-bit
    samples in whole-byte frames only) and copies it back."""
    ps = (_ffi.SynthPulsar * max(1, len(w.pulsars)))(*[_ffi.SynthPulsar(*p) for p in w.pulsars])
    b = np.ascontiguousarray(np.array(w.bursts, dtype=np.float64).reshape(-1))
    rc_ = np.ascontiguousarray(np.array(w.rfi_chans, dtype=np.uint32))
    common = _red_noise(w.seed, w.nsamples, w.tsamp, w.red_noise) if w.red_noise else None
    cfg = _ffi.SynthConfig(
        w.seed, w.nsamples, w.nchans, w.nbits, w.tsamp, w.fch1, w.foff, w.mean, w.sigma,
        len(w.pulsars), len(w.bursts), ps,
        b.ctypes.data_as(C.POINTER(C.c_double)) if b.size else None,
        w.sine[0], w.sine[1],
        rc_.ctypes.data_as(C.POINTER(C.c_uint32)) if rc_.size else None, len(w.rfi_chans), w.rfi_amp,
        common.ctypes.data_as(C.POINTER(C.c_float)) if common is not None else None, threads)
    out = np.empty(w.payload_bytes, dtype=np.uint8) if device else np.zeros(w.payload_bytes, dtype=np.uint8)
    if device:
        rc = _ffi.synth_gpu_lib().psr_synth_generate_gpu(C.byref(cfg), out.ctypes.data, out.nbytes)
    else:
        rc = _ffi.synth_lib().psr_synth_generate(C.byref(cfg), out.ctypes.data, out.nbytes)
    if rc != 0:
        raise RuntimeError("psr_synth_generate failed (%d)" % rc)
    return out


def _plan(lo, hi, step):
    n = int(math.floor((hi - lo) / step)) + 1
    return lo + np.arange(n, dtype=np.float64) * step


# the precision workload's PRESTO-style plan: 1500 + 750 + 750 trials, the step
# doubling at 150 and 300 pc cm^-3
PRECISION_DM_PLAN = ((0.0, 149.95, 0.1), (150.0, 299.9, 0.2), (300.0, 599.8, 0.4))


def workload(name, nsamples=None):
    """The named benchmark workloads; `nsamples` overrides the length (tests)."""
    if name == "plain-64ch-8bit":  # CPU-runnable reference case
        n = nsamples or 1 << 17
        return Workload(name, 42, n, 64, 8, 64e-6, 1449.21875, -1.5625, 128.0, 16.0,
                        pulsars=[(0.032, 62.0, 0.05, 0.5, 0.25)], dms=_plan(0.0, 120.0, 2.0),
                        min_hz=10.0, max_harmonics=8, top=2, rfi_mask=True)
    if name == "lband-96ch-4bit":
        n = nsamples or 1 << 19
        amp = amplitude_for_snr(30.0, 96, n, 0.03)
        return Workload(name, 7, n, 96, 4, 72e-6, 1447.5, -1.0, 8.0, 1.0,
                        pulsars=[(0.00462, 62.4, 0.03, amp, 0.25)],
                        bursts=[(0.25 * n * 72e-6, 0.05, 3.0), (0.7 * n * 72e-6, 0.05, 3.0)],
                        sine=(28.76, 0.05), dms=_plan(0.0, 499.75, 0.25), max_harmonics=16)
    if name in ("precision-1024ch-8bit", "headline-1024ch-8bit"):
        headline = name.startswith("headline")
        n = nsamples or (1 << 22 if headline else 1 << 21)
        rng = np.random.default_rng(2024 if headline else 2023)
        if headline:
            P = np.exp(rng.uniform(np.log(1e-3), np.log(0.5), 8))
            DM = rng.uniform(5, 800, 8)
            duty = rng.uniform(0.01, 0.10, 8)
            snr = rng.uniform(6, 30, 8)
        else:
            P = np.array([0.0017, 0.0053, 0.033, 0.410])
            DM = np.array([15.0, 62.0, 180.0, 430.0])
            duty = np.array([0.02, 0.05, 0.08, 0.10])
            snr = np.array([8.0, 15.0, 25.0, 40.0])
        pulsars = [(float(P[i]), float(DM[i]), float(duty[i]), amplitude_for_snr(snr[i], 1024, n, duty[i]),
                    float(rng.uniform(0, 1))) for i in range(len(P))]
        tobs = n * 64e-6
        if headline:
            rfi = sorted(rng.choice(1024, 51, replace=False).tolist())
            bursts = [(float(s), 0.05, 2.0) for s in rng.uniform(0.05 * tobs, 0.95 * tobs, 3)]
            dms = _plan(0.0, 999.75, 0.25)
            red = 0.0
        else:
            rfi, bursts, red = [], [], 0.2
            # step doubling (psr search --dm-plan 0:149.95:0.1,150:299.9:0.2,300:599.8:0.4)
            dms = np.concatenate([_plan(*seg) for seg in PRECISION_DM_PLAN])
        return Workload(name, 11 if headline else 10, n, 1024, 8, 64e-6, 1599.8046875, -0.390625, 128.0, 16.0,
                        pulsars=pulsars, bursts=bursts, rfi_chans=rfi, rfi_amp=3.0 if rfi else 0.0,
                        red_noise=red, dms=dms, max_harmonics=16, top=20 if headline else 5)
    raise KeyError(name)


def batch_workloads(count=256, seed=4096, log2_len=(16, 21), trials=(50, 2000), controls=0.2):
    """The batching workload: `count` seeded observations of uneven size.

    Observation i (seed + i) draws its channel count from {64, 96, 256, 512,
    1024}, its bit depth from {2, 4, 8}, tsamp uniform in 50-100 us, its
    length as 2^U(log2_len), and its DM plan as 50-2000 trials from DM 0 at the
    reference's default step (dedisperse::dm_step of its band).  A fraction
    `controls` of the observations are noise/RFI-only controls; the others
    carry one or two pulsars (period log-uniform 1-500 ms, duty 1-10 %, SNR
    6-30, DM within the plan).  Bands: 64 channels over 1350-1450 MHz, 96
    over 1352-1448 MHz, 256-1024 over 1200-1600 MHz (descending channels)."""
    from .dedisperse import dm_step
    out = []
    for i in range(count):
        rng = np.random.default_rng(seed + i)
        nch = int(rng.choice([64, 96, 256, 512, 1024]))
        nbits = int(rng.choice([2, 4, 8]))
        tsamp = float(rng.uniform(50e-6, 100e-6))
        ns = int(round(2.0 ** rng.uniform(*log2_len)))
        lo, hi = {64: (1350.0, 1450.0), 96: (1352.0, 1448.0)}.get(nch, (1200.0, 1600.0))
        foff = -(hi - lo) / nch
        fch1 = hi + foff / 2
        mean, sigma = {2: (1.5, 0.8), 4: (8.0, 1.5), 8: (128.0, 16.0)}[nbits]
        freqs = fch1 + np.arange(nch) * foff
        step = dm_step(freqs, tsamp)
        ntr = int(rng.integers(trials[0], trials[1] + 1))
        # keep the largest delay (about ntr samples at the default step) below a quarter of the length
        ntr = max(trials[0], min(ntr, ns // 4))
        dms = np.arange(ntr, dtype=np.float64) * step
        pulsars = []
        control = rng.uniform() < controls
        if not control:
            for _ in range(int(rng.integers(1, 3))):
                P = float(np.exp(rng.uniform(np.log(1e-3), np.log(0.5))))
                duty = float(rng.uniform(0.01, 0.10))
                snr = float(rng.uniform(6, 30))
                dm = float(rng.uniform(0.05, 0.9) * dms[-1])
                pulsars.append((P, dm, duty, amplitude_for_snr(snr, nch, ns, duty), float(rng.uniform(0, 1))))
        rfi = sorted(rng.choice(nch, max(1, nch // 20), replace=False).tolist()) if rng.uniform() < 0.3 else []
        out.append(Workload("batch-%03d%s" % (i, "-control" if control else ""), seed + i, ns, nch, nbits, tsamp,
                            fch1, foff, mean, sigma, pulsars=pulsars, rfi_chans=rfi, rfi_amp=2.0 if rfi else 0.0,
                            dms=dms, min_hz=10.0, max_hz=2000.0, threshold=8.0, max_harmonics=16, top=5))
    return out
