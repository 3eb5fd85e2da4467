"""CPU-side checks of the C-ABI library: it loads, exports every symbol the
headers declare, its host geometry equals the oracle's restatement bit for
bit, and the problem initialisers follow their specification.  No device
calls are made here."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

from oracle import pyoracle as po
from medusa_amd import mds, problems

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module", autouse=True)
def built_library():
    if not os.path.exists(mds.LIB_PATH):
        from medusa_amd import build
        build.build(verbose=False)
    return mds.lib()


def declared_functions():
    names = set()
    for h in ("mds_hydro.h", "mds_host.h", "mds_eos.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"\b(mds_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(mds.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_splitmix64_known_answers():
    lib = mds.lib()
    lib.mds_splitmix64.argtypes = [ctypes.c_ulonglong, ctypes.c_ulonglong]
    lib.mds_splitmix64.restype = ctypes.c_ulonglong
    # published splitmix64 outputs for seed 0 (Vigna's reference implementation)
    assert lib.mds_splitmix64(0, 0) == 0xE220A8397B1DCDAF
    assert lib.mds_splitmix64(0, 1) == 0x6E789E6AA1B965F4
    assert lib.mds_splitmix64(0, 2) == 0x06C45D188009454F


def _oracle_geometry(cfg):
    s = po.Sim(cfg)
    return s, s.geometry()


def test_host_cartesian_geometry_matches_oracle_bitwise():
    n = [24, 20, 16]
    start = [-0.25, -0.25, -0.75]
    dx = [0.5 / 24, 0.5 / 20, 1.5 / 16]
    g = mds.GridBuilder(3, n, start, dx)
    s, geo = _oracle_geometry(po.make_config(3, n, start, dx, [3, 3, 1], [3, 3, 1], 1.4))
    for d in range(3):
        for a, b in zip(g.moments(d), s.moments(d)):
            assert np.array_equal(a, b)
    assert g.g.volume == geo[0, po.GEOM_VOLUME]
    for d in range(3):
        assert g.g.area[d] == geo[0, po.GEOM_AREA + d]


def test_host_spherical_geometry_matches_oracle_bitwise():
    n = 300
    x0, x1 = math.log(1e5), math.log(1e9)
    g = mds.GridBuilder(1, [n], [x0], [(x1 - x0) / n], geom=mds.GEOM_SPHERICAL, rcoord=mds.RCOORD_LOG)
    s, geo = _oracle_geometry(po.make_config(1, [n], [x0], [(x1 - x0) / n], [1], [0], 1.4,
                                             geom=po.GEOM_SPHERICAL, rcoord=po.RCOORD_LOG))
    for a, b in zip(g.moments(0), s.moments(0)):
        assert np.array_equal(a, b)
    z = g.zone()
    conn = geo[:, po.GEOM_CONN:po.GEOM_CONN + 27].reshape(-1, 3, 3, 3)
    gcov = geo[:, po.GEOM_GCOV:po.GEOM_GCOV + 12].reshape(-1, 4, 3)
    assert np.array_equal(z[:, 0], geo[:, po.GEOM_VOLUME])
    assert np.array_equal(z[2:-1, 1], geo[2:-1, po.GEOM_AREA])
    assert np.array_equal(z[:, 2:5], gcov[:, 0, :])
    assert np.array_equal(z[:, 5:8], geo[:, po.GEOM_GCON:po.GEOM_GCON + 3])
    assert np.array_equal(z[:, 8:11], gcov[:, 1, :])
    assert np.array_equal(z[:, 11], geo[:, po.GEOM_SCALE + 3])
    assert np.array_equal(z[:, 12], geo[:, po.GEOM_SCALE])
    names = [(0, 0, 0), (0, 1, 1), (1, 0, 1), (1, 1, 0), (1, 1, 1), (2, 0, 2), (2, 2, 0), (2, 1, 2), (2, 2, 1),
             (0, 2, 2), (1, 2, 2)]
    for k, (l, m, nn) in enumerate(names):
        assert np.array_equal(z[2:-2, 13 + k], conn[2:-2, l, m, nn]), (l, m, nn)
    # every other connection coefficient is zero in 1D (metric_spherical_general.c:867-918)
    mask = np.ones((3, 3, 3), bool)
    for l, m, nn in names:
        mask[l, m, nn] = False
    assert np.all(conn[:, mask] == 0.0)


def test_fp32_recon_coefficients_reproduce_uniform_ppm():
    g = mds.GridBuilder(1, [64], [0.0], [1.0 / 64])
    rc = g.rcoef(0)[2:-2]
    # uniform Cartesian PPM faces al = (2a_{i-1} + 5a_i - a_{i+1})/6 and
    # ar = (-a_{i-1} + 5a_i + 2a_{i+1})/6, i.e. in dqm = a_i - a_{i-1}, dqp = a_{i+1} - a_i:
    #   al = a - dqm/3 - dqp/6,  ar = a + dqm/6 + dqp/3
    assert np.allclose(rc[:, 0], -1.0 / 3.0, atol=1e-9)
    assert np.allclose(rc[:, 1], -1.0 / 6.0, atol=1e-9)
    assert np.allclose(rc[:, 2], 1.0 / 6.0, atol=1e-9)
    assert np.allclose(rc[:, 3], 1.0 / 3.0, atol=1e-9)
    assert np.allclose(rc[:, 4], 0.5, atol=1e-9)
    assert np.allclose(rc[:, 5], -6.0, rtol=1e-6)
    assert np.allclose(rc[:, 6], 3.0, rtol=1e-6)
    assert np.allclose(rc[:, 7], 3.0, rtol=1e-6)


def test_problem_initialisers():
    g, prob, p = problems.sod(100)
    # zone 49 straddles x=0.5 in floating point (xr < 0.5 is false): volume-fraction mix
    assert np.all(p[0, 0, 0, :49] == 1.0) and np.all(p[0, 0, 0, 50:] == 0.125)
    assert p[0, 0, 0, 49] == pytest.approx(1.0, abs=1e-12)
    g, prob, p = problems.kh(64)
    y = (np.arange(64) + 0.5) / 64
    inner = np.abs(y - 0.5) < 0.25
    assert np.all(p[0, 0][inner] == 2.0) and np.all(p[0, 0][~inner] == 1.0)
    assert np.allclose(p[1], 2.5 / (2.0 / 3.0))
    x = (np.arange(64) + 0.5) / 64
    noise = p[3, 0] - 0.01 * np.sin(4 * np.pi * x)[None, :]
    assert np.max(np.abs(noise)) <= 1e-3 and np.std(noise) > 4e-4
    p2 = problems.kh(64)[2]
    assert np.array_equal(p, p2)                  # seeded: reproducible
    par = mds.ensemble_params(500)
    assert par[:, 0].min() >= 256 and par[:, 0].max() <= 8192
    assert np.all((par[:, 1] >= 1.3) & (par[:, 1] <= 1.67))
    assert np.all((np.log10(par[:, 2]) >= 8) & (np.log10(par[:, 2]) <= 10))
    assert np.all((np.log10(par[:, 4]) >= 6) & (np.log10(par[:, 4]) <= 8))
