"""FP32-bulk spherical columns with three velocity components (nvel = 3):
a rotating spherical Riemann column against the FP64 oracle (all 11
connection terms of rad_fluid.c:195-214, the transverse metric in p <-> u
and in the Riemann solve), and the 1D launch shapes giving the same bits."""
import numpy as np
import pytest

from oracle import pyoracle as po
from medusa_amd import mds, problems
from parity_util import oracle_config

pytestmark = pytest.mark.gpu


def _rotating(n):
    """A spherical Riemann column (rho 1e9 / 1e8, P 1e27 / 1e26 at 1e7 cm)
    with theta and phi motion: physical transverse speeds of a tenth and a
    third of the inner sound speed, contravariant through the zone scales."""
    g, prob, prims = problems.sph_riemann(n, 1.4, 1.0e9, 1.0e27, 1.0e7, 10.0, 10.0)
    zs = g.zone()[2:-2]
    r = np.exp(g.g.startx[0] + (np.arange(n) + 0.5) * g.g.dx[0])
    cs = np.sqrt(1.4 * 1.0e27 / 1.0e9)
    shape = np.exp(-((np.log(r) - np.log(1.0e7)) / 1.5) ** 2)
    prims = prims.copy()
    scale = np.sqrt(zs[:, 2:5])              # sqrt of the zone's gcov[0][0..2] (MDS_ZG_GCOV00..02)
    prims[3, 0, 0] = 0.1 * cs * shape / scale[:, 1]
    prims[4, 0, 0] = 0.3 * cs * shape / scale[:, 2]
    return g, dict(prob, nvel=3), prims


def _run(prob, prims, nsteps, fuse=None):
    ctx = mds.Context([prob], precision=mds.FP32_BULK, rp=mds.run_params(cfl=0.5))
    if fuse is not None:
        ctx.set_tile(mds.TILE_FUSE1D, fuse)
    ctx.upload(prims)
    ctx.update_eos()
    ctx.run(nsteps)
    return ctx


def test_rotating_column_fp32_matches_fp64_oracle():
    n, nsteps = 600, 150
    g, prob, prims = _rotating(n)
    print(f"[rotating] FP32-bulk nvel=3 column n={n}, {nsteps} steps", flush=True)
    dev = _run(prob, prims, nsteps).download()[:, 0, 0]
    sim = po.Sim(oracle_config(prob, 0.5, nvars=5))
    sim.set_prims(prims)
    sim.run(max_steps=nsteps)
    ref = sim.get_prims()[:, 0, 0]
    zs = g.zone()[2:-2]
    w = zs[:, mds.ZG_VOLUME]
    rel = lambda a, b: float(np.sum(np.abs(a - b) * w) / np.sum(np.abs(b) * w))
    cs = np.sqrt(1.4 * 0.4 * np.abs(ref[1]) / ref[0])
    err = {"rho": rel(dev[0], ref[0]), "u": rel(dev[1], ref[1])}
    scale = np.sqrt(zs[:, 2:5])
    for d in range(3):
        err[f"v{d + 1}/cs"] = float(np.sum(np.abs(dev[2 + d] - ref[2 + d]) * scale[:, d] * w) / np.sum(cs * w))
    # the transverse motion is real: without its sources the flow would differ
    vphi = np.abs(ref[4] * scale[:, 2])
    print(f"[rotating]   L1 {err}; max |v_phi| {vphi.max():.3e}", flush=True)
    assert vphi.max() > 0.1 * np.sqrt(1.4e18)
    for k, v in err.items():
        assert v <= 1e-5, (k, v)


def test_rotating_column_launch_shapes_agree():
    g, prob, prims = _rotating(500)
    outs = []
    for fuse in (0, 1, 2):
        print(f"[rotating] launch shape two-stage={fuse}", flush=True)
        ctx = _run(prob, prims, 9, fuse)
        ctx.run(4, graph=True)
        outs.append((ctx.download(), ctx.clock()))
    for o in outs[1:]:
        assert np.array_equal(outs[0][0], o[0])
        assert outs[0][1] == o[1]


def test_rotating_star_column_table_monopole_fp32_matches_oracle():
    """A rotating star column with the table EOS and monopole gravity in
    FP32-bulk with three velocity components (the one-stage kernels with the
    monopole field) against the FP64 oracle."""
    T = mds.EosTable()
    n, nsteps = 256, 40
    g, prob, prims = problems.star1d(n, table=T, sigma=0.01)
    zs = g.zone()[2:-2]
    scale = np.sqrt(zs[:, 2:5])
    r = np.exp(g.g.startx[0] + (np.arange(n) + 0.5) * g.g.dx[0])
    prims = prims.copy()
    prims[4, 0, 0] = 2.0e8 * np.exp(-((r - 3.0e7) / 2.0e7) ** 2) / scale[:, 2]      # v_phi up to 2e8 cm/s
    prob = dict(prob, nvel=3)
    print(f"[rotating] FP32-bulk nvel=3 star column n={n} table EOS + monopole, {nsteps} steps", flush=True)
    ctx = mds.Context([prob], precision=mds.FP32_BULK, rp=mds.run_params(cfl=0.5))
    ctx.upload(prims)
    ctx.update_eos()
    ctx.run(nsteps)
    dev = ctx.download()[:, 0, 0]
    t = T.t
    ot = po.eos_tab(dict(nr=t.nr, nt=t.nt, ny=t.ny, lr0=t.lr0, lr1=t.lr1, lt0=t.lt0, lt1=t.lt1, ye0=t.ye0,
                         ye1=t.ye1), T.data)
    sim = po.Sim(oracle_config(prob, 0.5, nvars=6, monopole=True, eos_tab=ot))
    sim.set_prims(prims)
    sim.run(max_steps=nsteps)
    ref = sim.get_prims()[:, 0, 0]
    w = zs[:, mds.ZG_VOLUME]
    rel = lambda a, b: float(np.sum(np.abs(a - b) * w) / np.sum(np.abs(b) * w))
    cs = np.array([po.eos_walk(ot, a, b, c)[0][1] for a, b, c in zip(ref[0], ref[1], ref[5])])
    err = {"rho": rel(dev[0], ref[0]), "u": rel(dev[1], ref[1]), "ye": rel(dev[5], ref[5])}
    for d in range(3):
        err[f"v{d + 1}/cs"] = float(np.sum(np.abs(dev[2 + d] - ref[2 + d]) * scale[:, d] * w) / np.sum(cs * w))
    print(f"[rotating]   L1 {err}", flush=True)
    assert np.max(np.abs(ref[4] * scale[:, 2])) > 1.0e8
    for k, v in err.items():
        assert v <= 1e-5, (k, v)


def test_rotating_uniform_radius_column_reads_the_zone_records():
    """On a uniform-r column the connection coefficients change from zone to
    zone, so the rotating FP32 path reads them from the zone records (on a
    log-r column they are one set per problem): the same parity bound."""
    n, nsteps = 200, 60
    g = mds.GridBuilder(1, [n], [0.0], [1.0], geom=mds.GEOM_SPHERICAL, rcoord=mds.RCOORD_UNIFORM)
    x = (np.arange(n) + 0.5) / n
    prims = np.zeros((5, 1, 1, n))
    prims[0, 0, 0] = 1.0 + 0.5 * np.exp(-((x - 0.3) / 0.05) ** 2)
    prims[1, 0, 0] = 1.0 + 9.0 * (x < 0.35)
    prims[2, 0, 0] = 0.05 * np.sin(3 * x)
    zs = g.zone()[2:-2]
    scale = np.sqrt(zs[:, 2:5])
    prims[3, 0, 0] = 0.2 * np.exp(-((x - 0.5) / 0.1) ** 2) / scale[:, 1]
    prims[4, 0, 0] = 0.5 * x * np.exp(-x) / scale[:, 2]
    prob = dict(grid=g, bc_lo=[mds.BC_SPHERICAL_ORIGIN], bc_hi=[mds.BC_OUTFLOW], gam=5.0 / 3.0, nvel=3)
    print(f"[rotating] FP32-bulk nvel=3 uniform-r column n={n} from the origin, {nsteps} steps", flush=True)
    dev = _run(prob, prims, nsteps).download()[:, 0, 0]
    sim = po.Sim(oracle_config(prob, 0.5, nvars=5))
    sim.set_prims(prims)
    sim.run(max_steps=nsteps)
    ref = sim.get_prims()[:, 0, 0]
    w = zs[:, mds.ZG_VOLUME]
    rel = lambda a, b: float(np.sum(np.abs(a - b) * w) / np.sum(np.abs(b) * w))
    cs = np.sqrt(5.0 / 3.0 * (2.0 / 3.0) * ref[1] / ref[0])
    err = {"rho": rel(dev[0], ref[0]), "u": rel(dev[1], ref[1])}
    for d in range(3):
        err[f"v{d + 1}/cs"] = float(np.sum(np.abs(dev[2 + d] - ref[2 + d]) * scale[:, d] * w) / np.sum(cs * w))
    print(f"[rotating]   L1 {err}", flush=True)
    for k, v in err.items():
        assert v <= 1e-5, (k, v)
