"""Inner radial boundary through the origin (SPHERICAL_ORIGIN at the inner
radius, bc.c:774-801): a ghost zone at negative radius copies the zone at the
point reflection of its centre, (r, theta, phi) -> (r, pi - theta, phi + pi),
with v^1 and v^3 flipped.

Pins of the oracle's fill: on a state built from Cartesian fields, every
radial ghost zone holds the fields at its own centre -- the point reflection
of the zone of the mirrored shell on the same ray -- expressed in its own
(e_r, e_theta, e_phi) basis.  The statement is geometric and does not restate
the index map: a dropped theta mirror, a missing phi + pi, a level mix-up on a
dendritic grid or a wrong velocity sign puts a different value there.  In 1D
with v^3 = 0 the fill is REFLECT's, so a run matches the REFLECT run bit for
bit.  The device path is held against the oracle on 1D, uniform 3D and
dendritic 3D grids from the origin, in FP64 (1e-12 relative L1) and FP32-bulk
(1e-5)."""
import math

import numpy as np
import pytest

from oracle import pyoracle as po

NG = 2


def _basis(th, ph):
    er = np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
    et = np.array([math.cos(th) * math.cos(ph), math.cos(th) * math.sin(ph), -math.sin(th)])
    ep = np.array([-math.sin(ph), math.cos(ph), 0.0])
    return er, et, ep


def _fields(x):
    """a scalar and a vector field with no symmetry under x -> -x"""
    f = 2.0 + 0.3 * x[0] - 0.2 * x[1] + 0.1 * x[2] + 0.05 * x[0] * x[1] - 0.07 * x[1] * x[2]
    v = np.array([1.0 + 0.2 * x[1], -0.5 + 0.3 * x[2] - 0.1 * x[0], 0.2 + 0.4 * x[0]])
    return f, v


def _zone_values(rc, th, ph):
    """(rho, u, v_r, v_theta, v_phi) of a zone centred at radius rc (negative:
    through the origin) on the ray (th, ph)"""
    er, et, ep = _basis(th, ph)
    x = rc * er
    f, v = _fields(x)
    return np.array([f, 1.0 + 0.5 * f, v @ er, v @ et, v @ ep])


def _check_origin_ghosts(sim, n, nj, nk):
    """fill the interior from the fields, run the ghost fill, check every
    radial ghost zone.  Angles are zone labels: theta_j = pi (j + 1/2) / nj,
    phi_k = 2 pi (k + 1/2) / nk at the row's own level."""
    prims = np.zeros((5, n[2], n[1], n[0]))
    for i in range(n[0]):
        for j in range(nj(i)):
            th = math.pi * (j + 0.5) / nj(i)
            for k in range(nk(i, j)):
                prims[:, k, j, i] = _zone_values(i + 0.5, th, 2 * math.pi * (k + 0.5) / nk(i, j))
    sim.set_prims(prims)
    sim.fill_ghosts()
    worst = 0.0
    for m in range(NG):
        i = -1 - m
        for j in range(nj(m)):
            th = math.pi * (j + 0.5) / nj(m)
            for k in range(nk(m, j)):
                want = _zone_values(-(m + 0.5), th, 2 * math.pi * (k + 0.5) / nk(m, j))
                got = sim.zone(i, j, k)[:5]
                err = float(np.max(np.abs(got - want)))
                worst = max(worst, err)
                assert err < 1e-12, (i, j, k, got, want)
    return worst


def test_origin_ghosts_hold_the_point_reflection_uniform_3d():
    from medusa_amd import problems
    n = (6, 8, 12)
    g = problems.dend3d_grid(*n, dr0=1.0, rout=50.0, dendritic=False)
    from parity_util import oracle_config
    prob = dict(grid=g, bc_lo=[po.BC_SPHERICAL_ORIGIN, po.BC_SPHERICAL_ORIGIN, po.BC_PERIODIC],
                bc_hi=[po.BC_OUTFLOW, po.BC_SPHERICAL_ORIGIN, po.BC_PERIODIC], gam=5.0 / 3.0)
    sim = po.Sim(oracle_config(prob, 0.5, nvars=5))
    _check_origin_ghosts(sim, n, lambda i: n[1], lambda i, j: n[2])


@pytest.mark.parametrize("n", [(24, 16, 32), (20, 32, 64)])
def test_origin_ghosts_hold_the_point_reflection_dendritic(n):
    from medusa_amd import problems
    from parity_util import oracle_config
    g = problems.dend3d_grid(*n, dr0=1.0, rout=400.0)
    prob = dict(grid=g, bc_lo=[po.BC_SPHERICAL_ORIGIN, po.BC_SPHERICAL_ORIGIN, po.BC_PERIODIC],
                bc_hi=[po.BC_OUTFLOW, po.BC_SPHERICAL_ORIGIN, po.BC_PERIODIC], gam=5.0 / 3.0)
    sim = po.Sim(oracle_config(prob, 0.5, nvars=5))
    NJ, NK, _, _ = sim.levels()
    assert NJ[0] < n[1]                      # the mirrored shells are coarsened in theta
    _check_origin_ghosts(sim, n, lambda i: int(NJ[i]), lambda i, j: int(NK[i, j]))


def test_origin_ghosts_differ_from_reflect():
    """the same state under REFLECT leaves theta and phi alone: the check
    above must fail for it (guards the check itself)"""
    from medusa_amd import problems
    from parity_util import oracle_config
    n = (6, 8, 12)
    g = problems.dend3d_grid(*n, dr0=1.0, rout=50.0, dendritic=False)
    prob = dict(grid=g, bc_lo=[po.BC_REFLECT, po.BC_SPHERICAL_ORIGIN, po.BC_PERIODIC],
                bc_hi=[po.BC_OUTFLOW, po.BC_SPHERICAL_ORIGIN, po.BC_PERIODIC], gam=5.0 / 3.0)
    sim = po.Sim(oracle_config(prob, 0.5, nvars=5))
    with pytest.raises(AssertionError):
        _check_origin_ghosts(sim, n, lambda i: n[1], lambda i, j: n[2])


def _col1d(n=200, dr=1.0):
    from medusa_amd import mds
    return mds.GridBuilder(1, [n], [0.0], [dr], geom=mds.GEOM_SPHERICAL, rcoord=mds.RCOORD_UNIFORM)


def _col1d_prims(n, v3):
    x = (np.arange(n) + 0.5) / n
    prims = np.zeros((5, 1, 1, n))
    prims[0, 0, 0] = 1.0 + 0.5 * np.exp(-((x - 0.1) / 0.05) ** 2)
    prims[1, 0, 0] = 1.0 + 9.0 * (x < 0.15)
    prims[2, 0, 0] = 0.05 * np.sin(3 * x)
    prims[4, 0, 0] = v3 * x * np.exp(-x)
    return prims


def test_origin_1d_equals_reflect_without_v3():
    from parity_util import oracle_config
    n = 200
    g = _col1d(n)
    prims = _col1d_prims(n, 0.0)
    out = []
    for lo in (po.BC_SPHERICAL_ORIGIN, po.BC_REFLECT):
        sim = po.Sim(oracle_config(dict(grid=g, bc_lo=[lo], bc_hi=[po.BC_OUTFLOW], gam=5.0 / 3.0), 0.5, nvars=5))
        sim.set_prims(prims)
        sim.run(max_steps=40)
        out.append(sim.get_prims())
    assert np.array_equal(out[0], out[1])


def test_origin_1d_ghosts_flip_v1_and_v3():
    from parity_util import oracle_config
    n = 64
    g = _col1d(n)
    prims = _col1d_prims(n, 1.0)
    sim = po.Sim(oracle_config(dict(grid=g, bc_lo=[po.BC_SPHERICAL_ORIGIN], bc_hi=[po.BC_OUTFLOW],
                                    gam=5.0 / 3.0), 0.5, nvars=5))
    sim.set_prims(prims)
    sim.fill_ghosts()
    for m in range(NG):
        src = prims[:, 0, 0, m]
        got = sim.zone(-1 - m)
        assert np.array_equal(got[:5], src * np.array([1, 1, -1, 1, -1]))


# ---------------------------------------------------------------------------
# device path against the oracle
# ---------------------------------------------------------------------------
def _rel(a, b, w):
    return float(np.sum(np.abs(a - b) * w) / np.sum(np.abs(b) * w))


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_origin_1d_matches_oracle(precision):
    from medusa_amd import mds
    from parity_util import oracle_config
    n = 256
    g = _col1d(n)
    prob = dict(grid=g, bc_lo=[mds.BC_SPHERICAL_ORIGIN], bc_hi=[mds.BC_OUTFLOW], gam=5.0 / 3.0)
    # FP64 columns carry v^3 (flipped by the fill); FP32-bulk columns carry v^1 only
    prims = _col1d_prims(n, 1.0 if precision == 0 else 0.0)
    print(f"[origin] 1D n={n} precision={precision}, 60 steps", flush=True)
    ctx = mds.Context([prob], precision=precision, rp=mds.run_params(cfl=0.5))
    ctx.upload(prims)
    ctx.update_eos()
    ctx.run(60)
    dev = ctx.download()
    sim = po.Sim(oracle_config(prob, 0.5, nvars=5))
    sim.set_prims(prims)
    sim.run(max_steps=60)
    ref = sim.get_prims()
    w = g.zone()[2:-2, 0]
    rows = (0, 1, 2, 4) if precision == 0 else (0, 1, 2)
    errs = {r: _rel(dev[r].ravel(), ref[r].ravel(), w) for r in rows}
    t, dt, ist = ctx.clock()
    print(f"[origin]   L1 {errs}; t {t!r} vs {sim.t!r}", flush=True)
    tol = 1e-12 if precision == 0 else 1e-5
    for k, v in errs.items():
        assert v <= tol, (k, v)


def _origin_blast(g, n, dendritic):
    """off-centre over-pressure near the origin on a random density field"""
    prims = np.zeros((6, n[2], n[1], n[0]))
    rng = np.random.default_rng(11)
    prims[0] = 1.0 + 0.05 * rng.uniform(-1, 1, prims[0].shape)
    prims[1] = 1.5
    prims[1][: n[2] // 3, : n[1] // 2, 0:3] *= 8.0
    prims[2] = 0.02 * rng.uniform(-1, 1, prims[0].shape)
    prims[4] = 0.02 * rng.uniform(-1, 1, prims[0].shape)
    prims[5] = 0.5
    return prims * g.valid_mask()[None] if dendritic else prims


@pytest.mark.gpu
@pytest.mark.parametrize("dendritic", [False, True])
@pytest.mark.parametrize("precision", [0, 1])
def test_origin_3d_matches_oracle(dendritic, precision):
    from medusa_amd import mds, problems
    from parity_util import oracle_config
    n = (24, 16, 32)
    g = problems.dend3d_grid(*n, dr0=1.0, rout=400.0, dendritic=dendritic)
    bc_lo = [mds.BC_SPHERICAL_ORIGIN, mds.BC_SPHERICAL_ORIGIN, mds.BC_PERIODIC]
    prob = dict(grid=g, bc_lo=bc_lo, bc_hi=list(problems.SPH3D_BC_HI), gam=5.0 / 3.0)
    prims = _origin_blast(g, n, dendritic)
    print(f"[origin] 3D {n} dendritic={dendritic} precision={precision}, 10 steps", flush=True)
    ctx = mds.Context([prob], precision=precision, rp=mds.run_params(cfl=0.5))
    ctx.upload(prims)
    ctx.update_eos()
    ctx.run(10)
    dev = ctx.download()
    sim = po.Sim(oracle_config(prob, 0.5, nvars=6))
    sim.set_prims(prims)
    sim.run(max_steps=10)
    ref = sim.get_prims()
    vol = g.zone()[2:-2, 2:-2, mds.ZS_VOLUME]
    w = np.broadcast_to(vol[None], ref[0].shape) * (g.valid_mask() if dendritic else 1.0)
    errs = {"rho": _rel(dev[0], ref[0], w), "u": _rel(dev[1], ref[1], w)}
    sc = g.zone()[2:-2, 2:-2, mds.ZS_SCALE0:mds.ZS_SCALE0 + 3]
    cs = np.sqrt(5.0 / 3.0 * (2.0 / 3.0) * np.where(w > 0, ref[1] / np.where(ref[0] > 0, ref[0], 1.0), 0.0))
    for d in range(3):
        errs[f"v{d + 1}/cs"] = float(np.sum(np.abs(dev[2 + d] - ref[2 + d]) * sc[None, :, :, d] * w)
                                     / np.sum(cs * w))
    t, dt, ist = ctx.clock()
    print(f"[origin]   L1 {errs}; t {t!r} vs {sim.t!r}", flush=True)
    tol = 1e-12 if precision == 0 else 1e-5
    for k, v in errs.items():
        assert v <= tol, (k, v)
    # the volume weights hide the smallest zones, those next to the origin:
    # every zone on its own, against the largest magnitude of its field
    zmax = {v: float(np.max(np.abs(dev[v] - ref[v])) / np.max(np.abs(ref[v]))) for v in range(5)}
    print(f"[origin]   max zone error / field max {zmax}", flush=True)
    for v, e in zmax.items():
        assert e <= (1e-11 if precision == 0 else 1e-3), (v, e)
    if precision == 0:
        assert ist == sim.istep and abs(t / sim.t - 1) < 1e-13
