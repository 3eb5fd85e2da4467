"""Spherical star columns: monopole gravity (gravity.c:145-498), the
tabulated EOS in the sweep and the passive Ye (riemann.c p_uflux/pflux,
fluid.c) in the CPU oracle, and the device path against it.

Pins: face gravity of a uniform sphere against -G M(r)/r^2 dr/dx with the
closed-form enclosed mass; a gamma = 4/3 polytrope in its own gravity stays
near hydrostatic equilibrium; Ye of a star at rest stays put."""
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from medusa_amd import mds, problems
from parity_util import oracle_config, compare

GNEWT = 6.67384e-8


def _otab(T):
    t = T.t
    return po.eos_tab(dict(nr=t.nr, nt=t.nt, ny=t.ny, lr0=t.lr0, lr1=t.lr1, lt0=t.lt0, lt1=t.lt1, ye0=t.ye0,
                           ye1=t.ye1), T.data)


def _star_oracle(prob, table=None, fast=False):
    kw = dict(nvars=6, monopole=True)
    if table is not None:
        kw["eos_tab"] = _otab(table)
    return po.Sim(oracle_config(prob, 0.5, **kw), fast=fast)


def test_monopole_face_gravity_of_uniform_sphere():
    n, rmin, rmax, rho = 128, 1.0e5, 1.0e9, 3.0e7
    x0, x1 = math.log(rmin), math.log(rmax)
    g = mds.GridBuilder(1, [n], [x0], [(x1 - x0) / n], geom=mds.GEOM_SPHERICAL, rcoord=mds.RCOORD_LOG)
    prob = dict(grid=g, bc_lo=[mds.BC_REFLECT], bc_hi=[mds.BC_OUTFLOW], gam=1.4)
    sim = po.Sim(oracle_config(prob, 0.5, nvars=5, monopole=True))
    prims = np.zeros((5, 1, 1, n))
    prims[0] = rho
    prims[1] = 1.0e20
    sim.set_prims(prims)
    ge = sim.monopole_edges()
    re = np.exp(x0 + np.arange(n + 1) * ((x1 - x0) / n))
    mass = rho * 4.0 * math.pi / 3.0 * (re ** 3 - rmin ** 3)
    ref = -GNEWT * mass / re ** 2 * re            # dr/dx = r for r = exp(x)
    assert ge[0] == 0.0
    np.testing.assert_allclose(ge[1:], ref[1:], rtol=1e-12)


def test_gamma_law_polytrope_stays_near_equilibrium():
    g, prob, prims = problems.star1d(192, table=None, sigma=0.0)
    sim = _star_oracle(prob)
    sim.set_prims(prims)
    sim.run(max_steps=150)
    out = sim.get_prims()
    inside = prims[0, 0, 0] > 1.0e6
    cs = np.sqrt(4.0 / 3.0 * (prims[1, 0, 0] / 3.0) / prims[0, 0, 0])
    mach = np.abs(out[2, 0, 0]) * np.exp(g.moments(0)[1][2:-2] / g.moments(0)[2][2:-2]) / cs
    assert np.max(mach[inside]) < 2e-2, np.max(mach[inside])
    assert np.max(np.abs(out[0, 0, 0][inside] / prims[0, 0, 0][inside] - 1)) < 2e-2


def test_table_star_keeps_ye_and_conserves_mass():
    T = mds.EosTable()
    g, prob, prims = problems.star1d(192, table=T, sigma=0.01)
    sim = _star_oracle(prob, T)
    sim.set_prims(prims)
    m0 = sim.totals()
    sim.run(max_steps=40)
    out = sim.get_prims()
    assert np.all(np.isfinite(out))
    inside = prims[0, 0, 0] > 1.0e6
    assert np.max(np.abs(out[5, 0, 0][inside] - prims[5, 0, 0][inside])) < 1e-6
    m1 = sim.totals()
    assert abs(m1[0] / m0[0] - 1) < 1e-9


FP64_TOL = 1.0e-12
FP32_RHO_TOL = 1.0e-5


def _dev_run(prob, prims, precision, nsteps, fuse=None, graph=False):
    ctx = mds.Context([prob], precision=precision, rp=mds.run_params(cfl=0.5))
    if fuse is not None:
        ctx.set_tile(mds.TILE_FUSE1D, fuse)
    ctx.upload(prims)
    ctx.update_eos()
    ctx.run(nsteps, graph=graph)
    return ctx


def _compare_star(dev, ref, g, table):
    """Relative L1 of rho, u (and Ye); the radial velocity of a star at rest is
    round-off sized, so its error is measured on the sound-speed scale:
    sum |dv_r| w / sum cs w, with v_r = r v^x and cs from the oracle state."""
    w = g.zone()[2:-2, 0]
    err = {"rho": _l1(dev[0], ref[0], w), "u": _l1(dev[1], ref[1], w)}
    r = np.exp(g.moments(0)[1][2:-2] / g.moments(0)[2][2:-2])
    if table is None:
        cs = np.sqrt(4.0 / 3.0 * (ref[1].ravel() / 3.0) / ref[0].ravel())
    else:
        t = table.t
        ot = _otab(table)
        cs = np.array([po.eos_walk(ot, a, b, c)[0][1] for a, b, c in zip(ref[0].ravel(), ref[1].ravel(),
                                                                         ref[5].ravel())])
        err["ye"] = _l1(dev[5], ref[5], w)
    err["v/cs"] = float(np.sum(np.abs(dev[2].ravel() - ref[2].ravel()) * r * w) / np.sum(cs * w))
    return err


def _l1(a, b, w):
    from parity_util import l1rel
    return l1rel(a.ravel(), b.ravel(), w)


@pytest.mark.gpu
@pytest.mark.parametrize("use_table", [False, True])
def test_star_fp64_matches_oracle(use_table):
    T = mds.EosTable() if use_table else None
    g, prob, prims = problems.star1d(256, table=T, sigma=0.01)
    print(f"[star] FP64 star column n=256 table={use_table} monopole, 60 steps", flush=True)
    ctx = _dev_run(prob, prims, mds.FP64, 60)
    sim = _star_oracle(prob, T)
    sim.set_prims(prims)
    sim.run(max_steps=60)
    t, dt, ist = ctx.clock()
    err = _compare_star(ctx.download(), sim.get_prims(), g, T)
    print(f"[star]   L1 {err}; t {t!r} vs {sim.t!r}", flush=True)
    assert ist == sim.istep and abs(t / sim.t - 1) < 1e-12
    for k, v in err.items():
        assert v <= FP64_TOL, (k, v)


@pytest.mark.gpu
def test_star_fp32_table_within_tolerance():
    T = mds.EosTable()
    g, prob, prims = problems.star1d(256, table=T, sigma=0.01)
    print("[star] FP32-bulk star column n=256 table monopole, 60 steps", flush=True)
    ctx = _dev_run(prob, prims, mds.FP32_BULK, 60, graph=True)
    sim = _star_oracle(prob, T)
    sim.set_prims(prims)
    sim.run(max_steps=60)
    err = _compare_star(ctx.download(), sim.get_prims(), g, T)
    print(f"[star]   L1 {err}", flush=True)
    assert err["rho"] <= FP32_RHO_TOL
    assert err["ye"] <= FP32_RHO_TOL


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [mds.FP64, mds.FP32_BULK])
def test_table_eos_two_stage_kernel_matches_single_stage(precision):
    T = mds.EosTable()
    g, prob, prims = problems.star1d(300, table=T, sigma=0.01)
    prob = dict(prob, monopole=False)
    outs = []
    for fuse in ((0, 1, 2) if precision == mds.FP32_BULK else (0, 1)):
        print(f"[star] table EOS without gravity precision={precision} two-stage={fuse}", flush=True)
        ctx = _dev_run(prob, prims, precision, 9, fuse=fuse)
        ctx.run(4, graph=True)
        outs.append((ctx.download(), ctx.clock()))
    for o in outs[1:]:
        assert np.array_equal(outs[0][0], o[0])
        assert outs[0][1] == o[1]


@pytest.mark.gpu
def test_batched_star_columns_fp64_match_oracle():
    T = mds.EosTable()
    cols = [problems.star1d(n, table=T, sigma=0.01, seed=100 + n) for n in (128, 200, 256, 333)]
    print("[star] 4 batched star columns FP64, 30 steps", flush=True)
    ctx = mds.Context([c[1] for c in cols], precision=mds.FP64, rp=mds.run_params(cfl=0.5))
    for k, c in enumerate(cols):
        ctx.upload(c[2], k)
    ctx.update_eos()
    ctx.run(30, graph=True)
    for k, (g, prob, prims) in enumerate(cols):
        sim = _star_oracle(prob, T)
        sim.set_prims(prims)
        sim.run(max_steps=30)
        err = _compare_star(ctx.download(k), sim.get_prims(), g, T)
        print(f"[star]   column {k} n={g.n[0]} L1 {err}", flush=True)
        for kk, v in err.items():
            assert v <= FP64_TOL, (k, kk, v)
