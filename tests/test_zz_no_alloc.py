"""The allocation-free Riemann entry point (mds_riemann_points_noalloc)
against the allocating one (mds_riemann_points): the same kernels on the
same inputs, so the outputs agree bit for bit.  Two calls in a row, the
second with fewer faces, show that the fixed device staging keeps no stale
state that leaks into a later call; a size beyond the maximum is refused
before any device work."""
import ctypes

import numpy as np
import pytest

from medusa_amd import mds


def _faces(rng, n):
    pl = np.zeros((n, 7))
    pr = np.zeros((n, 7))
    for p in (pl, pr):
        p[:, 0] = 10 ** rng.uniform(-1, 1, n)
        p[:, 5] = 10 ** rng.uniform(-1, 1, n)
        p[:, 2:5] = rng.uniform(-2.0, 2.0, (n, 3))
        p[:, 6] = 1.4
        p[:, 1] = p[:, 5] / (p[:, 6] - 1.0)
    fg = np.zeros((n, 5))
    fg[:, 0] = rng.uniform(0.5, 2.0, n)
    fg[:, 1] = rng.uniform(0.5, 2.0, n)
    fg[:, 2:5] = rng.uniform(0.5, 2.0, (n, 3))
    return pl, pr, fg


def _same(a, b):
    """bitwise equality, NaN payloads included (degenerate HLLE inputs give 0/0)"""
    return np.array_equal(np.asarray(a).view(np.uint64 if np.asarray(a).dtype == np.float64 else np.int32),
                          np.asarray(b).view(np.uint64 if np.asarray(b).dtype == np.float64 else np.int32))


def test_size_beyond_the_maximum_is_refused():
    n = mds.MAX_NOALLOC_FACES + 1
    a = np.zeros(7 * n)
    out = np.zeros(5 * n)
    sig = np.zeros(n)
    sh = np.zeros(n, dtype=np.int32)
    r = mds.lib().mds_riemann_points_noalloc(mds.FP64, 0, n, mds._ptr(a), mds._ptr(a), mds._ptr(a), 0,
                                             mds._ptr(out), mds._ptr(sig),
                                             sh.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    assert r == mds.ERANGE


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [mds.FP64, mds.FP32_BULK])
@pytest.mark.parametrize("solver", [0, 1])
def test_noalloc_matches_allocating_entry(precision, solver):
    rng = np.random.default_rng(11 + 2 * precision + solver)
    for n, d in ((mds.MAX_NOALLOC_FACES, 0), (3001, 2)):
        print(f"[no-alloc] precision={precision} solver={solver} nface={n} dir={d}", flush=True)
        pl, pr, fg = _faces(rng, n)
        ref = mds.riemann_points(precision, solver, pl, pr, fg, d)
        got = mds.riemann_points(precision, solver, pl, pr, fg, d, noalloc=True)
        for x, y in zip(got, ref):
            assert _same(x, y)
