"""3D spherical (r, theta, phi) grids: the oracle's geometry (metric_spherical_general.c
on a uniform angular grid), the pole boundary (bc.c SPHERICAL_ORIGIN in theta)
and the polar-axis phi-face states (step.c:1299-1397); the device path against it.

Pins: volumes and radial areas against the closed forms of the spherical
volume element; shells of volumes sum to 4 pi / 3 (r1^3 - r0^3); areas on the
axis are zero; a uniform gas at rest stays at rest to round-off (the
pressure-balanced Gamma^theta_thetatheta) in the oracle."""
import math

import numpy as np
import pytest

from oracle import pyoracle as po

POLY_XT, POLY_ALPHA = 1.2, 4.0
POLY_NORM = 0.5 * math.pi / (1.0 + 1.0 / (POLY_ALPHA + 1.0) / POLY_XT ** POLY_ALPHA)


def th_of_x(x):
    return POLY_NORM * x * (1.0 + (x / POLY_XT) ** POLY_ALPHA / (POLY_ALPHA + 1.0)) + 0.5 * math.pi


def sph3d_config(n, rmin=1.0e5, rmax=2.0e9, gam=4.0 / 3.0, nvars=5, **kw):
    n1, n2, n3 = n
    x0 = math.log(rmin)
    return po.make_config(3, [n1, n2, n3], [x0, -1.0, 0.0], [(math.log(rmax) - x0) / n1, 2.0 / n2, 2 * math.pi / n3],
                          [1, 5, 3], [0, 5, 3], gam, cfl=0.5, nvars=nvars, geom=po.GEOM_SPHERICAL,
                          rcoord=po.RCOORD_LOG, **kw)


def _geometry(sim, n):
    """zone geometry accessor (i, j, k) -> dict of volume and areas"""
    arr = sim.geometry()
    N0, N1 = n[0] + 4, n[1] + 4

    def at(i, j, k):
        rec = arr[(i + 2) + N0 * ((j + 2) + N1 * (k + 2))]
        return {"volume": rec[po.GEOM_VOLUME], "area": rec[po.GEOM_AREA:po.GEOM_AREA + 3],
                "conn": rec[po.GEOM_CONN:po.GEOM_CONN + 27].reshape(3, 3, 3),
                "gcov": rec[po.GEOM_GCOV:po.GEOM_GCOV + 12].reshape(4, 3)}
    return at


def test_spherical3d_volumes_and_areas_closed_form():
    n = (6, 8, 4)
    cfg = sph3d_config(n)
    sim = po.Sim(cfg)
    geo = _geometry(sim, n)
    dx = [cfg.dx[d] for d in range(3)]
    for i in range(n[0]):
        r0, r1 = math.exp(cfg.startx[0] + i * dx[0]), math.exp(cfg.startx[0] + (i + 1) * dx[0])
        shell = 0.0
        for j in range(n[1]):
            t0, t1 = th_of_x(-1.0 + j * dx[1]), th_of_x(-1.0 + (j + 1) * dx[1])
            dmu = math.cos(t0) - math.cos(t1)
            vol = (r1 ** 3 - r0 ** 3) / 3.0 * dmu * dx[2]
            g = geo(i, j, 0)
            assert abs(g["volume"] / vol - 1) < 1e-12
            assert abs(g["area"][0] / (r0 ** 3 * dmu * dx[2]) - 1) < 1e-12      # r^2 dr/dx, dr/dx = r
            shell += g["volume"] * n[2]
        assert abs(shell / (4 * math.pi / 3 * (r1 ** 3 - r0 ** 3)) - 1) < 1e-12
        assert geo(i, 0, 0)["area"][1] == 0.0 and geo(i, n[1], 0)["area"][1] == 0.0


def test_spherical3d_uniform_gas_stays_at_rest():
    n = (8, 8, 8)
    cfg = sph3d_config(n, gam=5.0 / 3.0)
    sim = po.Sim(cfg)
    prims = np.zeros((5, n[2], n[1], n[0]))
    prims[0] = 1.0e8
    prims[1] = 1.0e26
    sim.set_prims(prims)
    sim.run(max_steps=10)
    out = sim.get_prims()
    cs = math.sqrt(5.0 / 3.0 * (2.0 / 3.0) * 1.0e26 / 1.0e8)
    r = np.exp(cfg.startx[0] + (np.arange(n[0]) + 0.5) * cfg.dx[0])
    vr = np.abs(out[2]) * r
    assert np.max(vr) / cs < 1e-12
    assert np.max(np.abs(out[0] / 1.0e8 - 1)) < 1e-12


def test_host_geometry_matches_oracle_bitwise():
    from medusa_amd import mds
    n = (6, 8, 4)
    cfg = sph3d_config(n)
    sim = po.Sim(cfg)
    geo = _geometry(sim, n)
    g = mds.GridBuilder(3, list(n), [cfg.startx[d] for d in range(3)], [cfg.dx[d] for d in range(3)],
                        geom=mds.GEOM_SPHERICAL, rcoord=mds.RCOORD_LOG)
    z = g.zone()
    arr = sim.geometry()
    for d in range(3):
        om = sim.moments(d)
        hm = g.moments(d)
        for q in range(3):
            assert np.array_equal(om[q], hm[q]), (d, q)
    N0, N1 = n[0] + 4, n[1] + 4
    for j in range(-2, n[1] + 2):
        for i in range(-2, n[0] + 2):
            rec = arr[(i + 2) + N0 * ((j + 2) + N1 * 2)]
            h = z[j + 2, i + 2]
            assert h[mds.ZS_VOLUME] == rec[po.GEOM_VOLUME]
            for a in range(3):
                assert h[mds.ZS_AREA0 + a] == rec[po.GEOM_AREA + a], (i, j, a)
                assert h[mds.ZS_GCOV0 + a] == rec[po.GEOM_GCOV + a]
                assert h[mds.ZS_GCOVF0 + a] == rec[po.GEOM_GCOV + 3 + a]
                assert h[mds.ZS_GCOVF1 + a] == rec[po.GEOM_GCOV + 6 + a]
                assert h[mds.ZS_GCON + a] == rec[po.GEOM_GCON + a]
                assert h[mds.ZS_SCALE0 + a] == rec[po.GEOM_SCALE + a]
            assert h[mds.ZS_SCALEF0] == rec[po.GEOM_SCALE + 3]
            assert h[mds.ZS_SCALEF1] == rec[po.GEOM_SCALE + 7]
            if 0 <= i < n[0] and 0 <= j < n[1]:
                conn = rec[po.GEOM_CONN:po.GEOM_CONN + 27].reshape(3, 3, 3)
                names = {"C000": (0, 0, 0), "C011": (0, 1, 1), "C101": (1, 0, 1), "C110": (1, 1, 0),
                         "C202": (2, 0, 2), "C220": (2, 2, 0), "C212": (2, 1, 2), "C221": (2, 2, 1),
                         "C022": (0, 2, 2), "C122": (1, 2, 2), "C111": (1, 1, 1)}
                for nm, idx in names.items():
                    assert h[getattr(mds, "ZS_" + nm)] == conn[idx], (i, j, nm)
                nz = np.count_nonzero(conn)
                assert nz <= 11


# ---------------------------------------------------------------------------
# device path against the oracle
# ---------------------------------------------------------------------------
def _oracle_sph3d(prob, table=None, monopole=True):
    import sys
    from parity_util import oracle_config
    kw = dict(nvars=6, monopole=monopole)
    if table is not None:
        t = table.t
        kw["eos_tab"] = po.eos_tab(dict(nr=t.nr, nt=t.nt, ny=t.ny, lr0=t.lr0, lr1=t.lr1, lt0=t.lt0, lt1=t.lt1,
                                        ye0=t.ye0, ye1=t.ye1), table.data)
    return po.Sim(oracle_config(prob, 0.5, **kw))


def _rel(a, b, w):
    return float(np.sum(np.abs(a - b) * w) / np.sum(np.abs(b) * w))


def _blast(n):
    from medusa_amd import mds, problems
    g = problems.sph3d_grid(*n)
    prob = dict(grid=g, bc_lo=list(problems.SPH3D_BC_LO), bc_hi=list(problems.SPH3D_BC_HI), gam=5.0 / 3.0)
    prims = np.zeros((6, n[2], n[1], n[0]))
    prims[0] = 1.0e8
    prims[1] = 1.0e26
    rng = np.random.default_rng(3)
    prims[0] *= 1.0 + 0.01 * rng.uniform(-1, 1, prims[0].shape)
    # over-pressured blob off the axis and off the centre, in the last theta rows too
    prims[1][:, n[1] - 3:, n[0] // 3:n[0] // 3 + 3] *= 10.0
    prims[1][n[2] // 4:n[2] // 4 + 2, 2:4, n[0] // 2:n[0] // 2 + 2] *= 20.0
    prims[5] = 0.5
    return g, prob, prims


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_sph3d_blast_matches_oracle(precision):
    from medusa_amd import mds
    n = (24, 12, 8)
    g, prob, prims = _blast(n)
    print(f"[sph3d] blast {n} precision={precision}, 12 steps", flush=True)
    ctx = mds.Context([prob], precision=precision, rp=mds.run_params(cfl=0.5))
    ctx.upload(prims)
    ctx.update_eos()
    ctx.run(12)
    dev = ctx.download()
    sim = _oracle_sph3d(prob, monopole=False)
    sim.set_prims(prims)
    sim.run(max_steps=12)
    ref = sim.get_prims()
    vol = np.broadcast_to(g.zone()[2:-2, 2:-2, mds.ZS_VOLUME][None], ref[0].shape)
    errs = {"rho": _rel(dev[0], ref[0], vol), "u": _rel(dev[1], ref[1], vol)}
    cs = np.sqrt(5.0 / 3.0 * (2.0 / 3.0) * ref[1] / ref[0])
    sc = g.zone()[2:-2, 2:-2, mds.ZS_SCALE0:mds.ZS_SCALE0 + 3]
    for d in range(3):
        errs[f"v{d + 1}/cs"] = float(np.sum(np.abs(dev[2 + d] - ref[2 + d]) * sc[None, :, :, d] * vol)
                                     / np.sum(cs * vol))
    t, dt, ist = ctx.clock()
    hl = ctx.counters()
    print(f"[sph3d]   L1 {errs}; t {t!r} vs {sim.t!r}; shocks/HLLE {hl} vs {sim.shock_faces()}/{sim.hlle_faces()}",
          flush=True)
    tol = 1e-12 if precision == 0 else 1e-5
    for k, v in errs.items():
        assert v <= tol, (k, v)
    if precision == 0:
        assert ist == sim.istep and abs(t / sim.t - 1) < 1e-13
        assert hl[0] == sim.shock_faces() and hl[1] == sim.hlle_faces()


@pytest.mark.gpu
@pytest.mark.parametrize("use_table,precision", [(False, 0), (True, 0), (False, 1), (True, 1)])
def test_sph3d_star_matches_oracle(use_table, precision):
    """FP64: parity to 1e-12.  FP32-bulk: rho, u, Ye L1 <= 1e-5, velocity
    on the sound-speed scale <= 1e-5, total mass within 1e-6 of the oracle's."""
    from medusa_amd import mds, problems
    from test_dendritic import star_sound_speed
    T = mds.EosTable() if use_table else None
    n = (32, 8, 8)
    g, prob, prims = problems.star3d(*n, table=T)
    nsteps = 10 if precision == 0 else 20
    print(f"[sph3d] star {n} table={use_table} precision={precision}, {nsteps} steps", flush=True)
    ctx = mds.Context([prob], precision=precision, rp=mds.run_params(cfl=0.5))
    ctx.upload(prims)
    ctx.update_eos()
    m0 = ctx.totals()
    ctx.run(nsteps, graph=True)
    dev = ctx.download()
    m1 = ctx.totals()
    sim = _oracle_sph3d(prob, T)
    sim.set_prims(prims)
    sim.run(max_steps=nsteps)
    ref = sim.get_prims()
    vol = np.broadcast_to(g.zone()[2:-2, 2:-2, mds.ZS_VOLUME][None], ref[0].shape)
    errs = {"rho": _rel(dev[0], ref[0], vol), "u": _rel(dev[1], ref[1], vol)}
    if use_table:
        errs["ye"] = _rel(dev[5], ref[5], vol)
    cs = star_sound_speed(ref, g, T, 4.0 / 3.0, np.ones(ref[0].shape, bool))
    sc = g.zone()[2:-2, 2:-2, mds.ZS_SCALE0:mds.ZS_SCALE0 + 3]
    for d in range(3):
        errs[f"v{d + 1}/cs"] = float(np.sum(np.abs(dev[2 + d] - ref[2 + d]) * sc[None, :, :, d] * vol)
                                     / np.sum(cs * vol))
    # mass against the oracle's after the same steps (outflow at the outer radius moves mass in both)
    drift = abs(m1[0] - sim.totals()[0]) / m0[0]
    print(f"[sph3d]   L1 {errs}; mass vs oracle {drift:.2e}", flush=True)
    tol = 1e-12 if precision == 0 else 1e-5
    for k, v in errs.items():
        assert v <= tol, (k, v)
    assert drift <= (1e-13 if precision == 0 else 1e-6)


@pytest.mark.gpu
def test_sph3d_uniform_gas_stays_at_rest_on_device():
    from medusa_amd import mds, problems
    n = (16, 8, 8)
    g = problems.sph3d_grid(*n)
    prob = dict(grid=g, bc_lo=list(problems.SPH3D_BC_LO), bc_hi=list(problems.SPH3D_BC_HI), gam=5.0 / 3.0)
    prims = np.zeros((5, n[2], n[1], n[0]))
    prims[0] = 1.0e8
    prims[1] = 1.0e26
    for precision in (0, 1):
        ctx = mds.Context([prob], precision=precision, rp=mds.run_params(cfl=0.5))
        ctx.upload(prims)
        ctx.update_eos()
        ctx.run(10)
        out = ctx.download()
        cs = math.sqrt(5.0 / 3.0 * (2.0 / 3.0) * 1.0e26 / 1.0e8)
        sc = g.zone()[2:-2, 2:-2, mds.ZS_SCALE0:mds.ZS_SCALE0 + 3]
        vmax = max(float(np.max(np.abs(out[2 + d]) * sc[None, :, :, d])) for d in range(3))
        print(f"[sph3d] uniform gas precision={precision}: max |v|/cs {vmax / cs:.2e}", flush=True)
        assert vmax / cs < (1e-12 if precision == 0 else 1e-5)
