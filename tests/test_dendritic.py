"""Dendritic 3D spherical grids (utils.c:181-590): theta zones coarsened by
powers of two near the origin (dj), phi zones coarsened near the poles (dk),
prolongation / restriction across level changes in the flux pencils
(step.c:783-1085) and the flux fan-in of push (step.c:374-433).

Pins of the oracle's restatement: every level count is a power of two with
2:1 jumps at most between neighbours in r and theta, row 0 has nk_min = 4,
levels are symmetric about the equator; each shell's zone volumes sum to
4 pi/3 (r1^3 - r0^3) and each sphere's radial face areas to 4 pi r^2 dr/dx
(the closed forms of the volume element), so a dropped or doubled child
fails; a uniform gas at rest stays at rest to round-off (the pressure
balance needs the half-area / two-face dA rules of init_conn to match the
flux rules of push); a blast conserves mass and energy to round-off before
it reaches the outer boundary (a missing or doubled fan-in term breaks it)."""
import math

import numpy as np
import pytest

from oracle import pyoracle as po

POLY_XT, POLY_ALPHA = 1.2, 4.0
POLY_NORM = 0.5 * math.pi / (1.0 + 1.0 / (POLY_ALPHA + 1.0) / POLY_XT ** POLY_ALPHA)


def th_of_x(x):
    return POLY_NORM * x * (1.0 + (x / POLY_XT) ** POLY_ALPHA / (POLY_ALPHA + 1.0)) + 0.5 * math.pi


def rtrans_for(dr0, n1, rout):
    """rtrans of the sinh map r = rtrans sinh(x / rtrans) with r(n1 dr0) = rout (bisection)."""
    lo, hi = dr0 * 1.0e-2, rout * 1.0e2
    for _ in range(300):
        rt = math.sqrt(lo * hi)
        if rt * math.sinh(n1 * dr0 / rt) > rout:
            lo = rt
        else:
            hi = rt
    return rt


def dend_config(n, dr0=1.0, rout=400.0, gam=4.0 / 3.0, nvars=5, **kw):
    n1, n2, n3 = n
    return po.make_config(3, [n1, n2, n3], [0.0, -1.0, 0.0], [dr0, 2.0 / n2, 2 * math.pi / n3],
                          [po.BC_REFLECT, po.BC_SPHERICAL_ORIGIN, po.BC_PERIODIC],
                          [po.BC_OUTFLOW, po.BC_SPHERICAL_ORIGIN, po.BC_PERIODIC], gam, cfl=0.5, nvars=nvars,
                          geom=po.GEOM_SPHERICAL, rcoord=po.RCOORD_SINH, rtrans=rtrans_for(dr0, n1, rout),
                          dendritic=True, **kw)


def valid_mask(nj, nk, n):
    m = np.zeros((n[2], n[1], n[0]), bool)
    for i in range(n[0]):
        for j in range(nj[i]):
            m[:nk[i, j], j, i] = True
    return m


GRIDS = [(24, 16, 32), (20, 32, 64), (32, 16, 32)]


@pytest.mark.parametrize("n", GRIDS)
def test_levels_are_two_to_one(n):
    sim = po.Sim(dend_config(n))
    nj, nk, mj, mk = sim.levels()
    assert nj[0] == 2 and mj >= 1 and mk >= 1
    for i in range(n[0]):
        assert nj[i] & (nj[i] - 1) == 0 and n[1] % nj[i] == 0
        if i:
            assert nj[i] in (nj[i - 1], 2 * nj[i - 1])
        row = nk[i, :nj[i]]
        assert row[0] == 4 and np.all(row & (row - 1) == 0) and np.all(n[2] % row == 0)
        assert np.array_equal(row, row[::-1])                       # equatorial symmetry
        assert np.all(np.maximum(row[1:], row[:-1]) <= 2 * np.minimum(row[1:], row[:-1]))
        if i:
            # the zone under (i, j) on shell i-1 has at most 2:1 fewer phi zones
            jp = np.arange(nj[i]) * nj[i - 1] // nj[i]
            ratio = row / nk[i - 1, jp]
            assert np.all((ratio >= 1) & (ratio <= 2))


def test_log_radius_grid_is_rejected():
    """log r keeps r dtheta / dr fixed, so the level counts double in every inner
    shell; init_dendritic (utils.c:540-579) finds 4:1 jumps and exits."""
    n1, n2, n3 = 64, 32, 64
    x0 = math.log(1.0e5)
    cfg = po.make_config(3, [n1, n2, n3], [x0, -1.0, 0.0], [(math.log(2.0e9) - x0) / n1, 2.0 / n2, 2 * math.pi / n3],
                         [1, 5, 3], [0, 5, 3], 4.0 / 3.0, geom=po.GEOM_SPHERICAL, rcoord=po.RCOORD_LOG,
                         dendritic=True)
    with pytest.raises(ValueError):
        po.Sim(cfg)


def test_shell_volumes_and_sphere_areas_closed_form():
    n = (24, 16, 32)
    cfg = dend_config(n)
    sim = po.Sim(cfg)
    nj, nk, _, _ = sim.levels()
    arr = sim.geometry()
    N0, N1 = n[0] + 4, n[1] + 4
    rt = cfg.rtrans

    def rec(i, j, k):
        return arr[(i + 2) + N0 * ((j + 2) + N1 * (k + 2))]

    for i in range(n[0]):
        x0, x1 = i * cfg.dx[0], (i + 1) * cfg.dx[0]
        r0, r1 = rt * math.sinh(x0 / rt), rt * math.sinh(x1 / rt)
        vol = sum(rec(i, j, k)[po.GEOM_VOLUME] for j in range(nj[i]) for k in range(nk[i, j]))
        assert abs(vol / (4 * math.pi / 3 * (r1 ** 3 - r0 ** 3)) - 1) < 1e-12
        if i:
            area = sum(rec(i, j, k)[po.GEOM_AREA] for j in range(nj[i]) for k in range(nk[i, j]))
            assert abs(area / (4 * math.pi * r0 ** 2 * math.cosh(x0 / rt)) - 1) < 1e-12
        # a zone's theta extent covers dj fine zones
        dj = n[1] // nj[i]
        for j in range(nj[i]):
            t0, t1 = th_of_x(-1.0 + j * dj * cfg.dx[1]), th_of_x(-1.0 + (j + 1) * dj * cfg.dx[1])
            v = (r1 ** 3 - r0 ** 3) / 3.0 * (math.cos(t0) - math.cos(t1)) * (n[2] // nk[i, j]) * cfg.dx[2]
            assert abs(rec(i, j, 0)[po.GEOM_VOLUME] / v - 1) < 1e-12


def test_uniform_gas_stays_at_rest():
    n = (24, 16, 32)
    sim = po.Sim(dend_config(n, gam=5.0 / 3.0))
    nj, nk, _, _ = sim.levels()
    prims = np.zeros((5,) + (n[2], n[1], n[0]))
    prims[0] = 1.0
    prims[1] = 1.5
    sim.set_prims(prims)
    sim.run(max_steps=8)
    out = sim.get_prims()
    m = valid_mask(nj, nk, n)
    cs = math.sqrt(5.0 / 3.0 * (2.0 / 3.0) * 1.5)
    rt = sim.cfg.rtrans
    xc = (np.arange(n[0]) + 0.5) * sim.cfg.dx[0]
    scale_r = np.cosh(xc / rt)                               # dr/dx: physical v^r = v^1 dr/dx
    assert np.max(np.abs(out[2] * scale_r)[m]) / cs < 1e-13
    assert np.max(np.abs(out[0] - 1.0)[m]) < 1e-13
    assert np.all(out[0][~m] == 0.0)


def test_blast_conserves_mass_and_energy():
    n = (24, 16, 32)
    sim = po.Sim(dend_config(n))
    nj, nk, _, _ = sim.levels()
    rng = np.random.default_rng(5)
    prims = np.zeros((5,) + (n[2], n[1], n[0]))
    prims[0] = 1.0 + 0.05 * rng.standard_normal(prims[0].shape)
    prims[1] = 3.0
    prims[1][:, :, 3:6] *= 10.0
    sim.set_prims(prims)
    m0 = sim.totals()
    sim.run(max_steps=20)
    m1 = sim.totals()
    assert sim.shock_faces() > 0 and sim.hlle_faces() > 0
    assert np.all(np.abs(m1 / m0 - 1) < 1e-13)
    out = sim.get_prims()
    m = valid_mask(nj, nk, n)
    assert np.all(np.isfinite(out[:, m])) and np.all(out[0][m] > 0)
