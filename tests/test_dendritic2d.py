"""2D dendritic (r, theta) grids: theta zones coarsened by powers of two near
the origin (compute_dj, utils.c:181-237; every row keeps its single phi zone,
compute_dk's 2D branch, utils.c:311-316), the 2D prolongation / restriction
of the radial pencils (prolongate0, restrict0, step.c:579-656) inside
prim_to_fluxes2 (step.c:659-764), and the radial flux fan-in of push
(step.c:398-408).  The oracle and the device run them as 3D dendritic grids
of one phi zone spanning 2 pi, without phi faces: with one phi zone the k
slopes of the 3D prolongation vanish and its restriction sums one zone per
row, which is the 2D formula.

Pins: the levels follow the 3D rule for nj and leave every row one phi zone;
shell volumes sum to 4 pi/3 (r1^3 - r0^3) and sphere areas to 4 pi r^2 dr/dx;
a uniform gas stays at rest; a blast conserves mass and energy; the host
builder equals the oracle bit for bit; the device equals the oracle (FP64
1e-12, FP32-bulk 1e-5 with velocity on the sound-speed scale)."""
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from test_dendritic import rtrans_for, th_of_x


def dend2d_config(n, dr0=1.0, rout=400.0, gam=5.0 / 3.0, nvars=5, inner=po.BC_SPHERICAL_ORIGIN, **kw):
    n1, n2 = n
    return po.make_config(2, [n1, n2], [0.0, -1.0], [dr0, 2.0 / n2],
                          [inner, po.BC_SPHERICAL_ORIGIN], [po.BC_OUTFLOW, po.BC_SPHERICAL_ORIGIN], gam, cfl=0.5,
                          nvars=nvars, geom=po.GEOM_SPHERICAL, rcoord=po.RCOORD_SINH,
                          rtrans=rtrans_for(dr0, n1, rout), dendritic=True, **kw)


GRIDS = [(24, 16), (20, 32), (32, 64)]


def _mask(nj, n):
    m = np.zeros((1, n[1], n[0]), bool)
    for i in range(n[0]):
        m[0, :nj[i], i] = True
    return m


@pytest.mark.parametrize("n", GRIDS)
def test_levels_two_to_one_and_one_phi_zone(n):
    sim = po.Sim(dend2d_config(n))
    nj, nk, mj, mk = sim.levels()
    assert nj[0] == 2 and mk == 0
    for i in range(n[0]):
        assert nj[i] & (nj[i] - 1) == 0 and nj[i] <= n[1]
        if i:
            assert nj[i] in (nj[i - 1], 2 * nj[i - 1])
        assert np.all(nk[i, :nj[i]] == 1) and np.all(nk[i, nj[i]:] == 0)
    # the 3D rule for nj: the first shell doubles when its aspect exceeds 2^0.5
    assert mj == int(round(math.log2(n[1] // nj[0])))


def test_shell_volumes_and_sphere_areas_closed_form():
    n = (24, 16)
    cfg = dend2d_config(n)
    sim = po.Sim(cfg)
    nj = sim.levels()[0]
    arr = sim.geometry()
    N0, N1 = n[0] + 4, n[1] + 4
    rt = cfg.rtrans

    def rec(i, j):
        return arr[(i + 2) + N0 * ((j + 2) + N1 * 2)]

    for i in range(n[0]):
        x0, x1 = i * cfg.dx[0], (i + 1) * cfg.dx[0]
        r0, r1 = rt * math.sinh(x0 / rt), rt * math.sinh(x1 / rt)
        vol = sum(rec(i, j)[po.GEOM_VOLUME] for j in range(nj[i]))
        assert abs(vol / (4 * math.pi / 3 * (r1 ** 3 - r0 ** 3)) - 1) < 1e-12
        if i:
            area = sum(rec(i, j)[po.GEOM_AREA] for j in range(nj[i]))
            assert abs(area / (4 * math.pi * r0 ** 2 * math.cosh(x0 / rt)) - 1) < 1e-12
        dj = n[1] // nj[i]
        for j in range(nj[i]):
            t0, t1 = th_of_x(-1.0 + j * dj * cfg.dx[1]), th_of_x(-1.0 + (j + 1) * dj * cfg.dx[1])
            v = (r1 ** 3 - r0 ** 3) / 3.0 * (math.cos(t0) - math.cos(t1)) * 2.0 * math.pi
            assert abs(rec(i, j)[po.GEOM_VOLUME] / v - 1) < 1e-12


def test_uniform_gas_stays_at_rest_and_blast_conserves():
    n = (24, 16)
    sim = po.Sim(dend2d_config(n))
    nj = sim.levels()[0]
    m = _mask(nj, n)
    prims = np.zeros((5, 1, n[1], n[0]))
    prims[0] = 1.0
    prims[1] = 1.5
    sim.set_prims(prims * m[None])
    sim.run(max_steps=10)
    out = sim.get_prims()
    cs = math.sqrt(5.0 / 3.0 * (2.0 / 3.0) * 1.5)
    rt = sim.cfg.rtrans
    scale_r = np.cosh((np.arange(n[0]) + 0.5) * sim.cfg.dx[0] / rt)
    assert np.max(np.abs(out[2] * scale_r)[m]) / cs < 1e-13
    assert np.max(np.abs(out[0] - 1.0)[m]) < 1e-13
    sim = po.Sim(dend2d_config(n))
    rng = np.random.default_rng(4)
    prims[0] = 1.0 + 0.05 * rng.standard_normal(prims[0].shape)
    prims[1] = 3.0
    prims[1][:, :, 3:6] *= 10.0
    sim.set_prims(prims * m[None])
    m0 = sim.totals()
    sim.run(max_steps=20)
    m1 = sim.totals()
    assert sim.shock_faces() > 0 and sim.hlle_faces() > 0
    assert np.all(np.abs(m1 / m0 - 1) < 1e-13)


def _host_grid(cfg):
    from medusa_amd import mds
    return mds.GridBuilder(2, [cfg.n[0], cfg.n[1]], [cfg.startx[0], cfg.startx[1]], [cfg.dx[0], cfg.dx[1]],
                           geom=mds.GEOM_SPHERICAL, rcoord=mds.RCOORD_SINH, rtrans=cfg.rtrans, dendritic=True)


@pytest.mark.parametrize("n", GRIDS)
def test_host_levels_and_geometry_match_oracle_bitwise(n):
    from medusa_amd import mds
    cfg = dend2d_config(n)
    sim = po.Sim(cfg)
    g = _host_grid(cfg)
    onj, onk, omj, omk = sim.levels()
    hnj, hnk = g.levels()
    assert np.array_equal(onj, hnj) and np.array_equal(onk[:, :n[1]], hnk[:, :n[1]])
    assert (g.g.maxjlev, g.g.maxklev) == (omj, omk)
    for lev in range(omj + 1):
        om, hm = sim.level_moments(1, lev), g.level_moments(1, lev)
        nn = n[1] // (1 << lev) + 4
        for q in range(3):
            assert np.array_equal(om[q][:nn], hm[q][:nn]), (lev, q)
    z = g.zone()
    arr = sim.geometry()
    N0, N1 = n[0] + 4, n[1] + 4
    names = {"C000": (0, 0, 0), "C011": (0, 1, 1), "C101": (1, 0, 1), "C110": (1, 1, 0),
             "C022": (0, 2, 2), "C122": (1, 2, 2), "C111": (1, 1, 1)}
    for i in range(-2, n[0] + 2):
        nj = onj[min(i, n[0] - 1)] if i >= 0 else onj[-1 - i]
        for j in range(-2, nj + 2):
            rec = arr[(i + 2) + N0 * ((j + 2) + N1 * 2)]
            h = z[j + 2, i + 2]
            assert h[mds.ZS_VOLUME] == rec[po.GEOM_VOLUME], (i, j)
            for a in range(3):
                assert h[mds.ZS_AREA0 + a] == rec[po.GEOM_AREA + a], (i, j, a)
                assert h[mds.ZS_GCOV0 + a] == rec[po.GEOM_GCOV + a]
                assert h[mds.ZS_SCALE0 + a] == rec[po.GEOM_SCALE + a]
            if 0 <= i < n[0] and 0 <= j < nj:
                conn = rec[po.GEOM_CONN:po.GEOM_CONN + 27].reshape(3, 3, 3)
                for nm, idx in names.items():
                    assert h[getattr(mds, "ZS_" + nm)] == conn[idx], (i, j, nm)


# ---------------------------------------------------------------------------
# device path against the oracle
# ---------------------------------------------------------------------------
def _problem(n, inner, gam=5.0 / 3.0, dr0=1.0, rout=400.0, **kw):
    from medusa_amd import mds, problems
    g = mds.GridBuilder(2, list(n), [0.0, -1.0], [dr0, 2.0 / n[1]], geom=mds.GEOM_SPHERICAL,
                        rcoord=mds.RCOORD_SINH, rtrans=problems.rtrans_for(dr0, n[0], rout), dendritic=True)
    bc_in = mds.BC_SPHERICAL_ORIGIN if inner == "origin" else mds.BC_REFLECT
    prob = dict(grid=g, bc_lo=[bc_in, mds.BC_SPHERICAL_ORIGIN], bc_hi=[mds.BC_OUTFLOW, mds.BC_SPHERICAL_ORIGIN],
                gam=gam, **kw)
    return g, prob


def _oracle(prob, **kw):
    from parity_util import oracle_config
    return po.Sim(oracle_config(prob, 0.5, **kw))


def _weights(g, shape):
    from medusa_amd import mds
    vol = g.zone()[2:-2, 2:-2, mds.ZS_VOLUME]
    return np.broadcast_to(vol[None], shape) * g.valid_mask()


def _rel(a, b, w):
    return float(np.sum(np.abs(a - b) * w) / np.sum(np.abs(b) * w))


@pytest.mark.gpu
@pytest.mark.parametrize("inner", ["origin", "reflect"])
@pytest.mark.parametrize("precision", [0, 1])
def test_dendritic2d_blast_matches_oracle(inner, precision):
    from medusa_amd import mds
    n = (32, 32)
    g, prob = _problem(n, inner)
    rng = np.random.default_rng(9)
    prims = np.zeros((5, 1, n[1], n[0]))
    prims[0] = 1.0 + 0.02 * rng.uniform(-1, 1, prims[0].shape)
    prims[1] = 1.5
    prims[1][:, :, 3:6] *= 10.0                       # a shell across the level changes
    prims[1][:, 2:5, n[0] // 2:n[0] // 2 + 2] *= 20.0  # off the axis
    prims *= g.valid_mask()[None]
    print(f"[dend2d] blast {n} inner={inner} precision={precision}, 16 steps", flush=True)
    ctx = mds.Context([prob], precision=precision, rp=mds.run_params(cfl=0.5))
    ctx.upload(prims)
    ctx.update_eos()
    m0 = ctx.totals()
    ctx.run(16)
    dev = ctx.download()
    m1 = ctx.totals()
    sim = _oracle(prob, nvars=5)
    sim.set_prims(prims)
    sim.run(max_steps=16)
    ref = sim.get_prims()
    w = _weights(g, ref[0].shape)
    errs = {"rho": _rel(dev[0], ref[0], w), "u": _rel(dev[1], ref[1], w)}
    cs = np.sqrt(5.0 / 3.0 * (2.0 / 3.0) * np.where(w > 0, ref[1] / np.where(ref[0] > 0, ref[0], 1.0), 0.0))
    sc = g.zone()[2:-2, 2:-2, mds.ZS_SCALE0:mds.ZS_SCALE0 + 3]
    for d in range(2):
        errs[f"v{d + 1}/cs"] = float(np.sum(np.abs(dev[2 + d] - ref[2 + d]) * sc[None, :, :, d] * w)
                                     / np.sum(cs * w))
    t, dt, ist = ctx.clock()
    hl = ctx.counters()
    print(f"[dend2d]   L1 {errs}; t {t!r} vs {sim.t!r}; shocks/HLLE {hl} vs {sim.shock_faces()}/"
          f"{sim.hlle_faces()}; mass drift {m1[0] / m0[0] - 1:.2e}", flush=True)
    tol = 1e-12 if precision == 0 else 1e-5
    for k, v in errs.items():
        assert v <= tol, (k, v)
    assert abs(m1[0] / m0[0] - 1) < (1e-13 if precision == 0 else 1e-6)
    if precision == 0:
        assert ist == sim.istep and abs(t / sim.t - 1) < 1e-13
        assert hl[0] == sim.shock_faces() and hl[1] == sim.hlle_faces()


@pytest.mark.gpu
@pytest.mark.parametrize("use_table", [False, True])
def test_dendritic2d_star_fp64_matches_oracle(use_table):
    """a 2D star with monopole gravity (and the table EOS) on a dendritic grid"""
    from medusa_amd import mds, problems
    T = mds.EosTable() if use_table else None
    n = (32, 16)
    g, prob = _problem(n, "origin", gam=4.0 / 3.0, dr0=2.0e6, rout=2.0e9, monopole=True, eos=T)
    prims = mds.init_polytrope(g, table=T, sigma=0.01) * g.valid_mask()[None]
    print(f"[dend2d] star {n} table={use_table} FP64, 10 steps", flush=True)
    ctx = mds.Context([prob], precision=mds.FP64, rp=mds.run_params(cfl=0.5))
    ctx.upload(prims)
    ctx.update_eos()
    ctx.run(10, graph=True)
    dev = ctx.download()
    kw = dict(nvars=6, monopole=True)
    if T is not None:
        t = T.t
        kw["eos_tab"] = po.eos_tab(dict(nr=t.nr, nt=t.nt, ny=t.ny, lr0=t.lr0, lr1=t.lr1, lt0=t.lt0, lt1=t.lt1,
                                        ye0=t.ye0, ye1=t.ye1), T.data)
    sim = _oracle(prob, **kw)
    sim.set_prims(prims)
    sim.run(max_steps=10)
    ref = sim.get_prims()
    w = _weights(g, ref[0].shape)
    errs = {"rho": _rel(dev[0], ref[0], w), "u": _rel(dev[1], ref[1], w)}
    if use_table:
        errs["ye"] = _rel(dev[5], ref[5], w)
    print(f"[dend2d]   L1 {errs}", flush=True)
    for k, v in errs.items():
        assert v <= 1e-12, (k, v)
