"""Benchmark configurations (grids, boundaries, initial states).

Each builder returns (grid, problem_dict, prims) where problem_dict is what
medusa_amd.mds.Context expects and prims is the [5][n3][n2][n1] initial state.
"""
import math

from . import mds


def sod(n=1024, gam=1.4):
    """1D Sod shock tube on [0,1], outflow both ends (problems/init_sod.c)."""
    g = mds.GridBuilder(1, [n], [0.0], [1.0 / n])
    prob = dict(grid=g, bc_lo=[mds.BC_OUTFLOW], bc_hi=[mds.BC_OUTFLOW], gam=gam)
    return g, prob, mds.init_sod(g, gam)


def kh(n=2048, gam=5.0 / 3.0, seed=0x5EED0002):
    """2D Kelvin-Helmholtz on the periodic unit square."""
    g = mds.GridBuilder(2, [n, n], [0.0, 0.0], [1.0 / n, 1.0 / n])
    per = [mds.BC_PERIODIC, mds.BC_PERIODIC]
    prob = dict(grid=g, bc_lo=per, bc_hi=per, gam=gam)
    return g, prob, mds.init_kh(g, gam, seed=seed)


def rt(n=512, gam=1.4, seed=0x5EED0003):
    """3D Rayleigh-Taylor on [-0.25,0.25]^2 x [-0.75,0.75], periodic x,y, reflecting z."""
    g = mds.GridBuilder(3, [n, n, n], [-0.25, -0.25, -0.75], [0.5 / n, 0.5 / n, 1.5 / n])
    prob = dict(grid=g, bc_lo=[mds.BC_PERIODIC, mds.BC_PERIODIC, mds.BC_REFLECT],
                bc_hi=[mds.BC_PERIODIC, mds.BC_PERIODIC, mds.BC_REFLECT], gam=gam, g0=[0.0, 0.0, -0.1])
    return g, prob, mds.init_rt(g, gam, seed=seed)


def sph_riemann(n, gam, rho_in, p_in, r0, frho, fp, rmin=1.0e5, rmax=1.0e9):
    """One spherical Riemann problem on a log-r grid, reflecting inner and outflow outer boundary."""
    x0, x1 = math.log(rmin), math.log(rmax)
    g = mds.GridBuilder(1, [int(n)], [x0], [(x1 - x0) / int(n)], geom=mds.GEOM_SPHERICAL, rcoord=mds.RCOORD_LOG)
    prob = dict(grid=g, bc_lo=[mds.BC_REFLECT], bc_hi=[mds.BC_OUTFLOW], gam=gam)
    return g, prob, mds.init_sph_riemann(g, gam, rho_in, p_in, r0, frho, fp)


def ensemble(nprob=2048, nmin=256, nmax=8192, seed=0x5EED0004, rmin=1.0e5, rmax=1.0e9):
    """The batched ensemble of spherical Riemann problems; returns (params, list of (grid, prob, prims)).
    Geometry is built with host threads (mds_grid_build_batch)."""
    par = mds.ensemble_params(nprob, nmin, nmax, seed)
    x0, x1 = math.log(rmin), math.log(rmax)
    descs = [(1, [int(r[0])], [x0], [(x1 - x0) / int(r[0])], mds.GEOM_SPHERICAL, mds.RCOORD_LOG) for r in par]
    grids = mds.GridBuilder.batch(descs)
    out = []
    for g, r in zip(grids, par):
        n, gam, rho_in, p_in, r0, frho, fp = r
        prob = dict(grid=g, bc_lo=[mds.BC_REFLECT], bc_hi=[mds.BC_OUTFLOW], gam=gam)
        out.append((g, prob, mds.init_sph_riemann(g, gam, rho_in, p_in, r0, frho, fp)))
    return par, out


def star1d(n=640, table=None, gam=4.0 / 3.0, sigma=0.01, seed=0x5EED0006, rmin=1.0e5, rmax=2.0e9, **kw):
    """A polytropic star (n=3, rho_c = 1e10) on a 1D log-r column with monopole
    gravity, reflecting inner and outflow outer boundary.  With a table
    (mds.EosTable) the EOS is tabulated and Ye is carried; otherwise gamma-law."""
    x0, x1 = math.log(rmin), math.log(rmax)
    g = mds.GridBuilder(1, [int(n)], [x0], [(x1 - x0) / int(n)], geom=mds.GEOM_SPHERICAL, rcoord=mds.RCOORD_LOG)
    prob = dict(grid=g, bc_lo=[mds.BC_REFLECT], bc_hi=[mds.BC_OUTFLOW], gam=gam, monopole=True, eos=table)
    return g, prob, mds.init_polytrope(g, gam=gam, table=table, sigma=sigma, seed=seed, **kw)


def sph3d_grid(n1, n2, n3, rmin=1.0e5, rmax=2.0e9):
    """3D spherical grid: log r over [rmin, rmax], theta over x1 in [-1, 1]
    (POLYTH map), phi over [0, 2 pi)."""
    x0, x1 = math.log(rmin), math.log(rmax)
    return mds.GridBuilder(3, [int(n1), int(n2), int(n3)], [x0, -1.0, 0.0],
                           [(x1 - x0) / int(n1), 2.0 / int(n2), 2.0 * math.pi / int(n3)],
                           geom=mds.GEOM_SPHERICAL, rcoord=mds.RCOORD_LOG)


SPH3D_BC_LO = [mds.BC_REFLECT, mds.BC_SPHERICAL_ORIGIN, mds.BC_PERIODIC]
SPH3D_BC_HI = [mds.BC_OUTFLOW, mds.BC_SPHERICAL_ORIGIN, mds.BC_PERIODIC]


def rtrans_for(dr0, n1, rout):
    """Scale of the sinh radial map r = rtrans sinh(x / rtrans) with central
    spacing dr/dx = 1 over x in [0, n1 dr0] and r(n1 dr0) = rout (bisection)."""
    lo, hi = dr0 * 1.0e-2, rout * 1.0e2
    for _ in range(300):
        rt = math.sqrt(lo * hi)
        if rt * math.sinh(n1 * dr0 / rt) > rout:
            lo = rt
        else:
            hi = rt
    return rt


def dend3d_grid(n1, n2, n3, dr0=5.0e4, rout=2.0e9, dendritic=True):
    """3D spherical dendritic grid from the origin: sinh radial map with
    central spacing dr0 reaching rout, theta over x1 in [-1, 1] (POLYTH map),
    phi over [0, 2 pi); theta zones coarsened near the origin, phi zones near
    the poles (utils.c:181-590)."""
    return mds.GridBuilder(3, [int(n1), int(n2), int(n3)], [0.0, -1.0, 0.0],
                           [float(dr0), 2.0 / int(n2), 2.0 * math.pi / int(n3)],
                           geom=mds.GEOM_SPHERICAL, rcoord=mds.RCOORD_SINH,
                           rtrans=rtrans_for(dr0, n1, rout), dendritic=dendritic)


def star3d(n1=640, n2=128, n3=256, table=None, gam=4.0 / 3.0, sigma=0.01, seed=0x5EED0006, monopole=True,
           dendritic=False, **kw):
    """The polytropic star (n=3, rho_c = 1e10) on a 3D spherical grid with
    monopole gravity; outflow outer radius, theta across the poles, phi
    periodic.  dendritic=False: log r over [1e5, 2e9] on a uniform angular
    mesh with a reflecting inner radius; True: the dendritic grid of
    dend3d_grid (the log r map cannot carry one: its zone aspect is constant,
    so the level counts double in every inner shell and the 2:1 rule fails),
    which reaches the centre, with the inner radius through the origin as in
    the reference's core-collapse setup (problems/init_ccsn.c:112)."""
    g = dend3d_grid(n1, n2, n3) if dendritic else sph3d_grid(n1, n2, n3)
    bc_lo = list(SPH3D_BC_LO)
    if dendritic:
        bc_lo[0] = mds.BC_SPHERICAL_ORIGIN
    prob = dict(grid=g, bc_lo=bc_lo, bc_hi=list(SPH3D_BC_HI), gam=gam, monopole=monopole, eos=table)
    return g, prob, mds.init_polytrope(g, gam=gam, table=table, sigma=sigma, seed=seed, **kw)
