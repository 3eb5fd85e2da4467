"""hydro_interp_order (io.c:5054, reconstruct.c:246-266): 2 selects lin_recon
(reconstruct.c:89-124), 1 the piecewise-constant states, on every kernel
family of the FP64 path, held against the oracle run with the same order
(its lin_recon is pinned in test_oracle.py).  FP32-bulk reconstructs with
para_recon only and refuses the other orders."""
import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _rel(a, b, w=None):
    w = 1.0 if w is None else w
    return float(np.sum(np.abs(a - b) * w) / np.sum(np.abs(b) * w))


def _pair(prob, prims, order, steps, nvars=5, tmax=None):
    from medusa_amd import mds
    from parity_util import oracle_config
    rp = mds.run_params(cfl=0.5, hydro_interp_order=order, tmax=tmax if tmax is not None else 1.0e300)
    ctx = mds.Context([prob], precision=mds.FP64, rp=rp)
    ctx.upload(prims)
    ctx.update_eos()
    sim = po.Sim(oracle_config(prob, 0.5, nvars=nvars, interp_order=order))
    sim.set_prims(prims)
    if tmax is not None:
        ctx.run_to()
        sim.run(tmax)
    else:
        ctx.run(steps)
        sim.run(max_steps=steps)
    return ctx, sim, ctx.download(), sim.get_prims()


def _check(ctx, sim, dev, ref, rows, w=None, tag=""):
    errs = {r: _rel(dev[r], ref[r], w) for r in rows}
    zmax = {r: float(np.max(np.abs(dev[r] - ref[r])) / max(np.max(np.abs(ref[r])), 1e-300)) for r in rows}
    t, dt, ist = ctx.clock()
    print(f"[order] {tag} L1 {errs}; max zone {zmax}; t {t!r} vs {sim.t!r}", flush=True)
    for r in rows:
        assert errs[r] <= 1e-12 and zmax[r] <= 1e-11, (r, errs[r], zmax[r])
    assert ist == sim.istep and abs(t / sim.t - 1) < 1e-13


@pytest.mark.parametrize("order", [2, 1])
def test_sod_to_t02(order):
    from medusa_amd import problems
    g, prob, prims = problems.sod(1024)
    ctx, sim, dev, ref = _pair(prob, prims, order, None, tmax=0.2)
    _check(ctx, sim, dev, ref, (0, 1, 2), tag=f"Sod order {order}")
    # the orders differ: order 1 smears the contact more than order 2
    if order == 2:
        assert not np.array_equal(dev, _pair(prob, prims, 3, None, tmax=0.2)[2])


@pytest.mark.parametrize("order", [2, 1])
def test_spherical_column(order):
    from medusa_amd import problems
    par, ens = problems.ensemble(1, 300, 301)
    g, prob, prims = ens[0]
    ctx, sim, dev, ref = _pair(prob, prims, order, 60)
    _check(ctx, sim, dev, ref, (0, 1, 2), w=g.zone()[2:-2, 0], tag=f"spherical column order {order}")


@pytest.mark.parametrize("order", [2, 1])
def test_kh_2d(order):
    from medusa_amd import problems
    g, prob, prims = problems.kh(64)
    ctx, sim, dev, ref = _pair(prob, prims, order, 20)
    _check(ctx, sim, dev, ref, (0, 1, 2, 3), tag=f"KH 64^2 order {order}")


@pytest.mark.parametrize("order", [2, 1])
def test_rt_3d(order):
    from medusa_amd import problems
    g, prob, prims = problems.rt(16)
    ctx, sim, dev, ref = _pair(prob, prims, order, 10)
    _check(ctx, sim, dev, ref, (0, 1, 4), tag=f"RT 16^3 order {order}")


@pytest.mark.parametrize("dendritic", [False, True])
@pytest.mark.parametrize("order", [2, 1])
def test_spherical_3d_blast(order, dendritic):
    from medusa_amd import mds, problems
    n = (24, 16, 32)
    g = problems.dend3d_grid(*n, dr0=1.0, rout=400.0, dendritic=dendritic)
    prob = dict(grid=g, bc_lo=[mds.BC_SPHERICAL_ORIGIN, mds.BC_SPHERICAL_ORIGIN, mds.BC_PERIODIC],
                bc_hi=list(problems.SPH3D_BC_HI), gam=5.0 / 3.0)
    rng = np.random.default_rng(9)
    prims = np.zeros((6, n[2], n[1], n[0]))
    prims[0] = 1.0 + 0.05 * rng.uniform(-1, 1, prims[0].shape)
    prims[1] = 1.5
    prims[1][:, :, 3:6] *= 8.0
    # a physical phi velocity of 0.02 (the state holds coordinate components)
    sc = g.zone()[2:-2, 2:-2, mds.ZS_SCALE0 + 2]
    prims[4] = 0.02 * rng.uniform(-1, 1, prims[0].shape) / np.where(sc > 0, sc, 1.0)[None]
    if dendritic:
        prims *= g.valid_mask()[None]
    ctx, sim, dev, ref = _pair(prob, prims, order, 10, nvars=6)
    vol = g.zone()[2:-2, 2:-2, mds.ZS_VOLUME]
    w = np.broadcast_to(vol[None], ref[0].shape) * (g.valid_mask() if dendritic else 1.0)
    _check(ctx, sim, dev, ref, (0, 1), w=w, tag=f"sph3d dendritic={dendritic} order {order}")


def test_fp32_bulk_refuses_other_orders():
    from medusa_amd import mds, problems
    g, prob, prims = problems.kh(64)
    with pytest.raises(RuntimeError):
        mds.Context([prob], precision=mds.FP32_BULK, rp=mds.run_params(cfl=0.5, hydro_interp_order=2))
    with pytest.raises(RuntimeError):
        mds.Context([prob], precision=mds.FP64, rp=mds.run_params(cfl=0.5, hydro_interp_order=4))
