"""psr_periodicity_noalloc (no device allocation, static per-device scratch)
against psr_periodicity: identical candidate lists, bit for bit, on two calls in
a row (the second on a shorter series, so it runs over the first call's
leftover scratch contents), and its size limits."""
import numpy as np
import pytest

from pulsarsr_amd import _ffi, search
from pulsarsr_amd.core import PsrError

pytestmark = pytest.mark.gpu


def _series(n, period, seed):
    rng = np.random.default_rng(seed)
    t = np.arange(n)
    ph = (t * 64e-6 / period) % 1.0
    return (rng.standard_normal(n) * 16 + 2000 + 40 * np.exp(-0.5 * ((ph - 0.3) / 0.02) ** 2)).astype(np.float32)


def _same(a, b):
    assert len(a) == len(b) > 0
    for x, y in zip(a, b):
        assert (x.frequency_hz, x.period_s, x.dm, x.harmonics, x.score) == \
            (y.frequency_hz, y.period_s, y.dm, y.harmonics, y.score)


def test_noalloc_matches_periodicity_twice_then_smaller():
    cfg = search.Config(min_hz=10.0, threshold=6.0, birdies=[(50.0, 0.5)])
    big = _series(_ffi.PSR_NOALLOC_MAX_SERIES, 0.0123, 1)
    small = _series(20000, 0.0071, 2)
    for x, dm in ((big, 31.5), (small, 4.0)):
        print("noalloc periodicity n=%d" % len(x), flush=True)
        want = search.periodicity(x, 64e-6, dm, cfg)
        got = search.periodicity_noalloc(x, 64e-6, dm, cfg)
        _same(got, want)
    # a one-pass transform (m <= 4096) after the four-step ones
    tiny = _series(3000, 0.0051, 3)
    _same(search.periodicity_noalloc(tiny, 64e-6, 0.0, cfg), search.periodicity(tiny, 64e-6, 0.0, cfg))


def test_noalloc_limits():
    x = np.zeros(_ffi.PSR_NOALLOC_MAX_SERIES + 1, np.float32)
    with pytest.raises(PsrError) as e:
        search.periodicity_noalloc(x, 64e-6, 0.0)
    assert e.value.code == _ffi.PSR_E_SHAPE
    with pytest.raises(PsrError) as e:
        search.periodicity_noalloc(x[:1000], 64e-6, 0.0, search.Config(max_candidates=257))
    assert e.value.code == _ffi.PSR_E_INVALID
    with pytest.raises(PsrError) as e:
        search.periodicity_noalloc(x[:1], 64e-6, 0.0)
    assert e.value.code == _ffi.PSR_E_CONTRACT
