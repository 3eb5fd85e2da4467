"""Multipole gravity in the oracle (gravity.c:499-1110, SPHERICAL_MULTIPOLE_GRAV):
pins by what the mathematics fixes, independently of the restatement.

  * the real spherical harmonics equal the textbook ones built from scipy's
    associated Legendre functions (Condon-Shortley phase, sqrt(2) cos(m phi)
    for m > 0, sqrt(2) sin(|m| phi) for m < 0, orthonormal on the sphere);
  * their zone integrals (the reference's theta recurrences and phi
    integrals) equal Gauss-Legendre quadrature of the same harmonics;
  * the l >= 1 potential at zone corners outside the mass equals the direct
    Newtonian potential of the zone-constant density (27-point quadrature of
    1/|x - x'| over every zone) minus its monopole -G M / r, and inside an
    empty core it equals the direct potential minus the constant l = 0 part;
    a truncation at l = 12 leaves (r_mass / r)^13 of the dipole term;
  * the momentum sources of a state at rest are -rho times the gradient of
    the same direct potential at the zone, to the second-order accuracy of
    the corner differences.
"""
import math

import numpy as np
import pytest
from scipy.special import lpmv

from oracle import pyoracle as po
from test_sph3d import sph3d_config, th_of_x

G = 6.67384e-8


def y_real(l, m, th, ph):
    am = abs(m)
    nrm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - am) / math.factorial(l + am))
    p = lpmv(am, l, np.cos(th))
    if m == 0:
        return nrm * p
    if m > 0:
        return math.sqrt(2.0) * nrm * p * np.cos(m * ph)
    return math.sqrt(2.0) * nrm * p * np.sin(am * ph)


def test_ylms_are_the_textbook_real_harmonics():
    rng = np.random.default_rng(5)
    lmax = 12
    for th, ph in zip(rng.uniform(0.01, math.pi - 0.01, 20), rng.uniform(0, 2 * math.pi, 20)):
        y = po.ylms(th, ph, lmax, lmax)
        for l in range(lmax + 1):
            for m in range(-l, l + 1):
                ref = y_real(l, m, th, ph)
                assert abs(y[l * (l + 1) + m] - ref) <= 1e-12 * max(1.0, abs(ref)), (l, m, th, ph)


def test_ylm_ints_equal_quadrature():
    rng = np.random.default_rng(6)
    lmax = 12
    xg, wg = np.polynomial.legendre.leggauss(40)
    zones = [(0.0, 0.3, 0.0, 0.4), (math.pi - 0.25, math.pi, 5.0, 6.2)]
    for _ in range(6):
        a, b = sorted(rng.uniform(0, math.pi, 2))
        c, d = sorted(rng.uniform(0, 2 * math.pi, 2))
        zones.append((a, b, c, d))
    for th0, th1, ph0, ph1 in zones:
        s = po.ylm_ints(th0, th1, ph0, ph1, lmax, lmax)
        th = 0.5 * (th1 - th0) * xg + 0.5 * (th1 + th0)
        ph = 0.5 * (ph1 - ph0) * xg + 0.5 * (ph1 + ph0)
        T, P = np.meshgrid(th, ph, indexing="ij")
        W = np.outer(wg, wg) * 0.25 * (th1 - th0) * (ph1 - ph0)
        omega = (math.cos(th0) - math.cos(th1)) * (ph1 - ph0)
        for l in range(lmax + 1):
            for m in range(-l, l + 1):
                ref = float(np.sum(W * y_real(l, m, T, P) * np.sin(T)))
                assert abs(s[l * (l + 1) + m] - ref) <= 1e-11 * omega, (l, m, th0, th1, ph0, ph1)


def test_ylm_ints_of_the_axisymmetric_2d_set():
    # 2D grids take m = 0 over the whole azimuth
    s = po.ylm_ints(0.4, 1.3, 0.0, 2 * math.pi, 8, 0)
    xg, wg = np.polynomial.legendre.leggauss(40)
    th = 0.45 * xg + 0.85
    for l in range(9):
        ref = 2 * math.pi * float(np.sum(0.45 * wg * y_real(l, 0, th, 0.0) * np.sin(th)))
        assert abs(s[l * (l + 1)] - ref) <= 1e-12


# ---------------------------------------------------------------------------
# potential and sources against the direct potential of the zone-constant density
# ---------------------------------------------------------------------------
def _grid(cfg):
    n = cfg.n[0], cfg.n[1], cfg.n[2]
    x0, dx = [cfg.startx[d] for d in range(3)], [cfg.dx[d] for d in range(3)]
    re = np.exp(x0[0] + np.arange(n[0] + 1) * dx[0])
    te = np.array([th_of_x(x0[1] + j * dx[1]) for j in range(n[1] + 1)])
    pe = x0[2] + np.arange(n[2] + 1) * dx[2]
    return n, re, te, pe


def _quad_points(re, te, pe, mask):
    """27-point Gauss quadrature of every zone in mask: positions and weights
    (dV = r^2 sin th dr dth dphi), the zone index of each point"""
    xg, wg = np.polynomial.legendre.leggauss(3)
    pts, wts, zid = [], [], []
    for k, j, i in zip(*np.nonzero(mask)):
        r = 0.5 * (re[i + 1] - re[i]) * xg + 0.5 * (re[i + 1] + re[i])
        t = 0.5 * (te[j + 1] - te[j]) * xg + 0.5 * (te[j + 1] + te[j])
        p = 0.5 * (pe[k + 1] - pe[k]) * xg + 0.5 * (pe[k + 1] + pe[k])
        R, T, P = np.meshgrid(r, t, p, indexing="ij")
        W = np.einsum("a,b,c->abc", wg, wg, wg) * 0.125 * (re[i + 1] - re[i]) * (te[j + 1] - te[j]) * \
            (pe[k + 1] - pe[k]) * R ** 2 * np.sin(T)
        pts.append(np.stack([R * np.sin(T) * np.cos(P), R * np.sin(T) * np.sin(P), R * np.cos(T)], -1).reshape(-1, 3))
        wts.append(W.ravel())
        zid.append(np.full(27, len(zid)))
    return np.concatenate(pts), np.concatenate(wts), np.concatenate(zid)


def _direct(X, pts, wq):
    """-G sum rho dV / |X - x'| for points X (m, 3); wq = rho * weight per quadrature point"""
    out = np.zeros(len(X))
    for q in range(len(X)):
        d = np.sqrt(np.sum((pts - X[q]) ** 2, axis=1))
        out[q] = -G * np.sum(wq / d)
    return out


def _cart(r, t, p):
    return np.array([r * math.sin(t) * math.cos(p), r * math.sin(t) * math.sin(p), r * math.cos(t)])


def _setup(n, inside, outside_shell=False, lmax=12):
    cfg = sph3d_config(n, rmin=1.0e7, rmax=1.0e9, monopole=True, multipole_lmax=lmax)
    sim = po.Sim(cfg)
    n, re, te, pe = _grid(cfg)
    rc = np.sqrt(re[:-1] * re[1:])
    tc = 0.5 * (te[:-1] + te[1:])
    pc = 0.5 * (pe[:-1] + pe[1:])
    K, J, I = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
    x, y, z = (rc[I] * np.sin(tc[J]) * np.cos(pc[K]), rc[I] * np.sin(tc[J]) * np.sin(pc[K]), rc[I] * np.cos(tc[J]))
    # an asymmetric lump: every l and m appears
    rho = 1.0e6 * (1.0 + 0.6 * x / rc[I] + 0.3 * (z / rc[I]) ** 2 - 0.4 * x * y / rc[I] ** 2)
    mask = inside(rc[I])
    prims = np.zeros((5, n[2], n[1], n[0]))
    prims[0] = np.where(mask, rho, 1.0e-20)
    prims[1] = 1.0e-10
    sim.set_prims(prims)
    pts, wts, zq = _quad_points(re, te, pe, mask)
    wq = wts * prims[0][mask][zq]
    return cfg, sim, n, re, te, pe, prims, pts, wts, wq, zq


def test_exterior_potential_equals_direct_sum():
    n = (20, 12, 16)
    cfg, sim, n, re, te, pe, prims, pts, wts, wq, zq = _setup(n, lambda r: r < 6.0e7)
    phi = sim.multipole_phi()
    M = float(np.sum(wq))
    Rin = 6.0e7
    sel = []
    for ii in range(n[0] + 1):
        if re[ii] < 3.5 * Rin:
            continue
        for jj in range(0, n[1] + 1, 3):
            for kk in range(0, n[2] + 1, 5):
                sel.append((ii, jj, kk))
    X = np.array([_cart(re[a], te[b], pe[c]) for a, b, c in sel])
    want = _direct(X, pts, wq) + G * M / np.linalg.norm(X, axis=1)
    got = np.array([phi[a, b, c] for a, b, c in sel])
    scale = np.max(np.abs(want))
    err = np.max(np.abs(got - want)) / scale
    print(f"exterior corners {len(sel)}: max |l>=1 potential| {scale:.3e}, max rel err {err:.2e}")
    assert scale > 1e-3 * G * M / (3.5 * Rin)          # the lump is far from spherical
    assert err < 2e-5


def test_interior_potential_of_a_shell_equals_direct_sum():
    n = (20, 12, 16)
    lo, hi = 2.0e8, 6.0e8
    cfg, sim, n, re, te, pe, prims, pts, wts, wq, zq = _setup(n, lambda r: (r > lo) & (r < hi))
    phi = sim.multipole_phi()
    sel = []
    for ii in range(n[0] + 1):
        if re[ii] > lo / 3.5 or re[ii] < 1.0e-16:
            continue
        for jj in range(0, n[1] + 1, 3):
            for kk in range(0, n[2] + 1, 5):
                sel.append((ii, jj, kk))
    X = np.array([_cart(re[a], te[b], pe[c]) for a, b, c in sel])
    phi0 = -G * float(np.sum(wq / np.linalg.norm(pts, axis=1)))      # the l = 0 part, constant inside
    want = _direct(X, pts, wq) - phi0
    got = np.array([phi[a, b, c] for a, b, c in sel])
    scale = np.max(np.abs(want))
    err = np.max(np.abs(got - want)) / scale
    print(f"interior corners {len(sel)}: max |l>=1 potential| {scale:.3e}, max rel err {err:.2e}")
    assert scale > 1e-4 * abs(phi0)
    assert err < 2e-5


def test_momentum_sources_are_the_potential_gradient():
    n = (20, 12, 16)
    Rin = 6.0e7
    cfg, sim, n, re, te, pe, prims, pts, wts, wq, zq = _setup(n, lambda r: r < Rin)
    src = sim.multipole_sources()
    M = float(np.sum(wq))
    dx = [cfg.dx[d] for d in range(3)]
    rc = np.sqrt(re[:-1] * re[1:])
    sel = [(i, j, k) for i in range(n[0]) if rc[i] > 4 * Rin for j in (2, 5, 8) for k in (1, 6, 11)]
    errs = []
    for i, j, k in sel:
        # the source is -rho dPhi/dx^d (x: the grid coordinates); take the
        # direct l >= 1 potential's derivative at the zone centre
        xc = [cfg.startx[0] + (i + 0.5) * dx[0], cfg.startx[1] + (j + 0.5) * dx[1], cfg.startx[2] + (k + 0.5) * dx[2]]
        g = []
        for d in range(3):
            h = 1e-3 * dx[d]
            vals = []
            for sgn in (-1, 1):
                x = list(xc)
                x[d] += sgn * h
                X = _cart(math.exp(x[0]), th_of_x(x[1]), x[2])
                vals.append(_direct(X[None], pts, wq)[0] + G * M / np.linalg.norm(X))
            g.append((vals[1] - vals[0]) / (2 * h))
        rho = prims[0, k, j, i]
        want = -rho * np.array(g)
        got = src[:, k, j, i]
        errs.append(np.max(np.abs(got - want)) / np.max(np.abs(want)))
    print(f"zones {len(sel)}: source vs -rho grad(direct): max rel err {max(errs):.3e}, median {np.median(errs):.3e}")
    assert max(errs) < 5e-2 and np.median(errs) < 2e-2


def test_multipole_needs_a_grid_beyond_1d():
    cfg = po.make_config(1, [32], [0.0], [1.0 / 32], [po.BC_REFLECT], [po.BC_OUTFLOW], 1.4, nvars=5,
                         geom=po.GEOM_SPHERICAL, monopole=True, multipole_lmax=4)
    sim = po.Sim(cfg)
    with pytest.raises(ValueError):
        sim.multipole_phi()


# ---------------------------------------------------------------------------
# device parity: stars with monopole + multipole gravity against the oracle
# ---------------------------------------------------------------------------
def _otab(T):
    t = T.t
    return po.eos_tab(dict(nr=t.nr, nt=t.nt, ny=t.ny, lr0=t.lr0, lr1=t.lr1, lt0=t.lt0, lt1=t.lt1, ye0=t.ye0,
                           ye1=t.ye1), T.data)


def _star(kind, use_table, lmax):
    from medusa_amd import mds, problems
    T = mds.EosTable() if use_table else None
    if kind == "dend":
        n = (32, 16, 32)
        g = problems.dend3d_grid(*n, dr0=2.0e6, rout=2.0e9)
        bc_lo = list(problems.SPH3D_BC_LO)
        bc_lo[0] = mds.BC_SPHERICAL_ORIGIN
        prob = dict(grid=g, bc_lo=bc_lo, bc_hi=list(problems.SPH3D_BC_HI), gam=4.0 / 3.0, monopole=True, eos=T)
        prims = mds.init_polytrope(g, table=T, sigma=0.05) * g.valid_mask()[None]
        valid = g.valid_mask()
    elif kind == "2d":
        n = (48, 16)
        x0, x1 = math.log(1.0e5), math.log(2.0e9)
        g = mds.GridBuilder(2, [n[0], n[1]], [x0, -1.0], [(x1 - x0) / n[0], 2.0 / n[1]],
                            geom=mds.GEOM_SPHERICAL, rcoord=mds.RCOORD_LOG)
        prob = dict(grid=g, bc_lo=[mds.BC_REFLECT, mds.BC_SPHERICAL_ORIGIN],
                    bc_hi=[mds.BC_OUTFLOW, mds.BC_SPHERICAL_ORIGIN], gam=4.0 / 3.0, monopole=True, eos=T)
        prims = mds.init_polytrope(g, table=T, sigma=0.05)
        valid = np.ones(prims[0].shape, bool)
    else:
        n = (32, 8, 8)
        g, prob, prims = problems.star3d(*n, table=T, sigma=0.05)
        valid = np.ones(prims[0].shape, bool)
    prob = dict(prob, multipole_lmax=lmax)
    return T, g, prob, prims, valid


@pytest.mark.gpu
@pytest.mark.parametrize("kind,use_table,precision,lmax", [("uniform", False, 0, 12), ("dend", True, 0, 6),
                                                           ("dend", True, 1, 6), ("2d", False, 0, 12)])
def test_multipole_star_matches_oracle(kind, use_table, precision, lmax):
    from medusa_amd import mds
    from parity_util import oracle_config
    from test_dendritic import star_sound_speed
    T, g, prob, prims, valid = _star(kind, use_table, lmax)
    nsteps = 10 if precision == 0 else 20
    print(f"[multipole] {kind} star {g.n} table={use_table} precision={precision} lmax={lmax}, {nsteps} steps",
          flush=True)
    ctx = mds.Context([prob], precision=precision, rp=mds.run_params(cfl=0.5))
    ctx.upload(prims)
    ctx.update_eos()
    ctx.run(nsteps, graph=True)
    dev = ctx.download()
    kw = dict(nvars=6, monopole=True, multipole_lmax=lmax)
    if T is not None:
        kw["eos_tab"] = _otab(T)
    sim = po.Sim(oracle_config(prob, 0.5, **kw))
    sim.set_prims(prims)
    sim.run(max_steps=nsteps)
    ref = sim.get_prims()
    zs = g.zone()[2:-2, 2:-2]
    w = np.broadcast_to(zs[..., mds.ZS_VOLUME][None], ref[0].shape) * valid
    rel = lambda a, b: float(np.sum(np.abs(a - b) * w) / np.sum(np.abs(b) * w))
    err = {"rho": rel(dev[0], ref[0]), "u": rel(dev[1], ref[1])}
    cs = star_sound_speed(ref, g, T, 4.0 / 3.0, valid)
    for d in range(3 if g.ndim == 3 else 2):
        err[f"v{d + 1}/cs"] = float(np.sum(np.abs(dev[2 + d] - ref[2 + d]) * zs[None, :, :, mds.ZS_SCALE0 + d] * w)
                                    / np.sum(cs * w))
    t, dt, ist = ctx.clock()
    print(f"[multipole]   L1 {err}; t {t!r} vs {sim.t!r}", flush=True)
    tol = 1e-12 if precision == 0 else 1e-5
    for k, v in err.items():
        assert v <= tol, (k, v)


@pytest.mark.gpu
def test_multipole_changes_the_flow():
    """guard: the same star with the monopole only moves differently"""
    from medusa_amd import mds
    T, g, prob, prims, valid = _star("uniform", False, 12)
    outs = []
    for lmax in (0, 12):
        ctx = mds.Context([dict(prob, multipole_lmax=lmax)], precision=0, rp=mds.run_params(cfl=0.5))
        ctx.upload(prims)
        ctx.update_eos()
        ctx.run(10)
        outs.append(ctx.download())
    dv = float(np.max(np.abs(outs[1][3] - outs[0][3])))         # theta velocity
    print(f"[multipole] max |dv_theta| monopole vs + multipole: {dv:.3e}", flush=True)
    assert dv > 0.0
