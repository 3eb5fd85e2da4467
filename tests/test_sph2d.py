"""2D spherical (r, theta) grids (prim_to_fluxes2, step.c:659-764, on the
spherical metric).  The host builds their geometry as the 3D records of one
phi zone spanning 2 pi with zero phi-face areas, and the device runs them
through the 3D spherical kernels without phi faces and without the phi term
of the zone dt (the reference's estimate_dt loops over NDIM directions), so
no phi flux or phi CFL limit exists, as in the reference's 2D path.

Pins: the oracle's 2D geometry against the closed forms of the volume
element (V = 2 pi d(r^3)/3 * d(cos theta); radial areas 2 pi r^2 dr/dx *
d(cos theta)); a uniform gas at rest stays at rest to round-off (pressure
balance of the theta connection terms); the host builder's records equal the
oracle's bit for bit; the device path equals the oracle (FP64 1e-12 relative
L1, every zone bounded; FP32-bulk 1e-5)."""
import math

import numpy as np
import pytest

from oracle import pyoracle as po

POLY_XT, POLY_ALPHA = 1.2, 4.0
POLY_NORM = 0.5 * math.pi / (1.0 + 1.0 / (POLY_ALPHA + 1.0) / POLY_XT ** POLY_ALPHA)


def th_of_x(x):
    return POLY_NORM * x * (1.0 + (x / POLY_XT) ** POLY_ALPHA / (POLY_ALPHA + 1.0)) + 0.5 * math.pi


def _grid(n, rmin=1.0e5, rmax=2.0e9):
    from medusa_amd import mds
    x0 = math.log(rmin)
    return mds.GridBuilder(2, list(n), [x0, -1.0], [(math.log(rmax) - x0) / n[0], 2.0 / n[1]],
                           geom=mds.GEOM_SPHERICAL, rcoord=mds.RCOORD_LOG)


def _prob(g, gam=5.0 / 3.0):
    from medusa_amd import mds
    return dict(grid=g, bc_lo=[mds.BC_REFLECT, mds.BC_SPHERICAL_ORIGIN],
                bc_hi=[mds.BC_OUTFLOW, mds.BC_SPHERICAL_ORIGIN], gam=gam)


def _oracle(prob, nvars=5):
    from parity_util import oracle_config
    return po.Sim(oracle_config(prob, 0.5, nvars=nvars))


def test_oracle_2d_volumes_closed_form_and_rest_state():
    g = _grid((8, 12))
    sim = _oracle(_prob(g))
    arr = sim.geometry()
    c = sim.cfg
    N0 = 8 + 4
    for i in range(8):
        r0, r1 = math.exp(c.startx[0] + i * c.dx[0]), math.exp(c.startx[0] + (i + 1) * c.dx[0])
        for j in range(12):
            t0, t1 = th_of_x(-1.0 + j * c.dx[1]), th_of_x(-1.0 + (j + 1) * c.dx[1])
            dmu = math.cos(t0) - math.cos(t1)
            rec = arr[(i + 2) + N0 * (j + 2)]
            assert abs(rec[po.GEOM_VOLUME] / (2 * math.pi * (r1 ** 3 - r0 ** 3) / 3.0 * dmu) - 1) < 1e-12
            assert abs(rec[po.GEOM_AREA] / (2 * math.pi * r0 ** 3 * dmu) - 1) < 1e-12   # r^2 dr/dx, dr/dx = r
    p = np.zeros((5, 1, 12, 8))
    p[0] = 1.0
    p[1] = 1.5
    sim.set_prims(p)
    sim.run(max_steps=20)
    q = sim.get_prims()
    assert np.max(np.abs(q[2:5])) < 1e-12 and np.max(np.abs(q[0] - 1.0)) < 1e-13


def test_host_2d_geometry_matches_oracle_bitwise():
    from medusa_amd import mds
    n = (6, 8)
    g = _grid(n)
    sim = _oracle(_prob(g))
    z = g.zone()
    arr = sim.geometry()
    for d in range(2):
        for q in range(3):
            assert np.array_equal(sim.moments(d)[q], g.moments(d)[q]), (d, q)
    N0 = n[0] + 4
    for j in range(-2, n[1] + 2):
        for i in range(-2, n[0] + 2):
            if not (0 <= i < n[0] or 0 <= j < n[1]):
                continue            # corner ghost records: no 2D stencil reads them
            rec = arr[(i + 2) + N0 * (j + 2)]
            h = z[j + 2, i + 2]
            assert h[mds.ZS_VOLUME] == rec[po.GEOM_VOLUME], (i, j)
            for a in range(3):
                assert h[mds.ZS_AREA0 + a] == rec[po.GEOM_AREA + a], (i, j, a)
                assert h[mds.ZS_GCOV0 + a] == rec[po.GEOM_GCOV + a]
                assert h[mds.ZS_SCALE0 + a] == rec[po.GEOM_SCALE + a]
            if 0 <= i < n[0] and 0 <= j < n[1]:
                conn = rec[po.GEOM_CONN:po.GEOM_CONN + 27].reshape(3, 3, 3)
                for nm, idx in {"C000": (0, 0, 0), "C011": (0, 1, 1), "C101": (1, 0, 1), "C110": (1, 1, 0),
                                "C022": (0, 2, 2), "C122": (1, 2, 2), "C111": (1, 1, 1)}.items():
                    assert h[getattr(mds, "ZS_" + nm)] == conn[idx], (i, j, nm)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_sph2d_blast_matches_oracle(precision):
    from medusa_amd import mds
    n = (48, 32)
    g = _grid(n)
    prob = _prob(g)
    rng = np.random.default_rng(21)
    prims = np.zeros((5, 1, n[1], n[0]))
    prims[0] = 1.0e8 * (1.0 + 0.01 * rng.uniform(-1, 1, prims[0].shape))
    prims[1] = 1.0e26
    prims[1][:, 10:14, n[0] // 3:n[0] // 3 + 3] *= 20.0          # off the axis
    prims[1][:, n[1] - 2:, n[0] // 2:n[0] // 2 + 2] *= 10.0       # at the pole
    print(f"[sph2d] blast {n} precision={precision}, 40 steps (past the step where the CFL estimate takes over)",
          flush=True)
    ctx = mds.Context([prob], precision=precision, rp=mds.run_params(cfl=0.5))
    ctx.upload(prims)
    ctx.update_eos()
    ctx.run(40)
    dev = ctx.download()
    sim = _oracle(prob)
    sim.set_prims(prims)
    sim.run(max_steps=40)
    ref = sim.get_prims()
    w = np.broadcast_to(g.zone()[2:-2, 2:-2, mds.ZS_VOLUME][None], ref[0].shape)
    errs = {v: float(np.sum(np.abs(dev[v] - ref[v]) * w) / np.sum(np.abs(ref[v]) * w)) for v in (0, 1)}
    zmax = {v: float(np.max(np.abs(dev[v] - ref[v])) / np.max(np.abs(ref[v]))) for v in (0, 1)}
    t, dt, ist = ctx.clock()
    print(f"[sph2d]   L1 {errs}; max zone {zmax}; t {t!r} vs {sim.t!r}; HLLE {ctx.counters()} vs "
          f"{sim.shock_faces()}/{sim.hlle_faces()}", flush=True)
    for v in (0, 1):
        assert errs[v] <= (1e-12 if precision == 0 else 1e-5), (v, errs[v])
        assert zmax[v] <= (1e-11 if precision == 0 else 1e-3), (v, zmax[v])
    if precision == 0:
        assert ist == sim.istep and abs(t / sim.t - 1) < 1e-13
        assert ctx.counters()[1] == sim.hlle_faces()


@pytest.mark.gpu
@pytest.mark.parametrize("use_table,precision", [(False, 0), (True, 0), (True, 1)])
def test_sph2d_star_matches_oracle(use_table, precision):
    """A polytropic star on a 2D (r, theta) grid with monopole gravity and
    (optionally) the table EOS: the shell-mass kernels over one phi zone and
    the table path on 2D records.  FP64: relative L1 <= 1e-12 in rho, u (and
    Ye); FP32-bulk: <= 1e-5, velocity on the sound-speed scale <= 1e-5."""
    from medusa_amd import mds
    from test_dendritic import star_sound_speed
    T = mds.EosTable() if use_table else None
    n = (40, 16)
    g = _grid(n)
    prob = dict(_prob(g, gam=4.0 / 3.0), monopole=True, eos=T)
    prims = mds.init_polytrope(g, table=T, sigma=0.01)
    nsteps = 10 if precision == 0 else 20
    print(f"[sph2d] star {n} table={use_table} precision={precision}, {nsteps} steps", flush=True)
    ctx = mds.Context([prob], precision=precision, rp=mds.run_params(cfl=0.5))
    ctx.upload(prims)
    ctx.update_eos()
    m0 = ctx.totals()
    ctx.run(nsteps, graph=True)
    dev = ctx.download()
    m1 = ctx.totals()
    kw = dict(nvars=6, monopole=True)
    if T is not None:
        t = T.t
        kw["eos_tab"] = po.eos_tab(dict(nr=t.nr, nt=t.nt, ny=t.ny, lr0=t.lr0, lr1=t.lr1, lt0=t.lt0, lt1=t.lt1,
                                        ye0=t.ye0, ye1=t.ye1), T.data)
    from parity_util import oracle_config
    sim = po.Sim(oracle_config(prob, 0.5, **kw))
    sim.set_prims(prims)
    sim.run(max_steps=nsteps)
    ref = sim.get_prims()
    w = np.broadcast_to(g.zone()[2:-2, 2:-2, mds.ZS_VOLUME][None], ref[0].shape)
    errs = {"rho": float(np.sum(np.abs(dev[0] - ref[0]) * w) / np.sum(np.abs(ref[0]) * w)),
            "u": float(np.sum(np.abs(dev[1] - ref[1]) * w) / np.sum(np.abs(ref[1]) * w))}
    if use_table:
        errs["ye"] = float(np.sum(np.abs(dev[5] - ref[5]) * w) / np.sum(np.abs(ref[5]) * w))
    cs = star_sound_speed(ref, g, T, 4.0 / 3.0, np.ones(ref[0].shape, bool))
    sc = g.zone()[2:-2, 2:-2, mds.ZS_SCALE0:mds.ZS_SCALE0 + 3]
    for d in range(2):
        errs[f"v{d + 1}/cs"] = float(np.sum(np.abs(dev[2 + d] - ref[2 + d]) * sc[None, :, :, d] * w)
                                     / np.sum(cs * w))
    drift = abs(m1[0] - sim.totals()[0]) / m0[0]
    print(f"[sph2d]   L1 {errs}; mass vs oracle {drift:.2e}", flush=True)
    tol = 1e-12 if precision == 0 else 1e-5
    for k, v in errs.items():
        assert v <= tol, (k, v)
    assert drift <= (1e-13 if precision == 0 else 1e-6)
