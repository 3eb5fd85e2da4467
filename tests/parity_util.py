"""Helpers shared by the parity tests: oracle configs mirroring a device
problem, and the reference's face-weighted relative L1 convention
(tst/regression/scripts/utils/comparison.py:9-27: sum |v| * face width)."""
import numpy as np

from oracle import pyoracle as po
from medusa_amd import mds


def oracle_config(prob, rp_cfl=0.5, **kw):
    """oracle configuration mirroring a device problem description"""
    g = prob["grid"]
    nd = g.ndim
    return po.make_config(
        nd, g.n[:nd], [g.g.startx[d] for d in range(nd)], [g.g.dx[d] for d in range(nd)],
        list(prob["bc_lo"]), list(prob["bc_hi"]), prob["gam"], cfl=rp_cfl,
        geom=po.GEOM_SPHERICAL if g.g.geom == mds.GEOM_SPHERICAL else po.GEOM_CARTESIAN,
        rcoord=g.g.rcoord, g0=tuple(prob.get("g0", (0.0, 0.0, 0.0))), rtrans=g.g.rtrans,
        dendritic=bool(g.g.dendritic), **kw)


def oracle_for(prob, rp_cfl=0.5, fast=False, **kw):
    return po.Sim(oracle_config(prob, rp_cfl, **kw), fast=fast)


def l1rel(a, b, w=None):
    """face-weighted relative L1 of a against b (w: zone widths, default uniform)."""
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    if w is None:
        w = np.ones_like(b)
    w = np.broadcast_to(np.asarray(w, dtype=np.float64), a.shape).ravel() if np.ndim(w) else w
    den = np.sum(np.abs(b) * w)
    num = np.sum(np.abs(a - b) * w)
    return num / den if den > 0 else num


def pressure(prims, gam):
    return (gam - 1.0) * prims[1]


def compare(dev, ref, gam, w=None):
    """Relative L1 of density, pressure and speed between two prim arrays."""
    sp_dev = np.sqrt(np.sum(dev[2:5] ** 2, axis=0))
    sp_ref = np.sqrt(np.sum(ref[2:5] ** 2, axis=0))
    return {
        "rho": l1rel(dev[0], ref[0], w),
        "p": l1rel(pressure(dev, gam), pressure(ref, gam), w),
        "v": l1rel(sp_dev, sp_ref, w),
    }


def v_over_cs(dev, ref, gam, w=None, scale=None):
    """Velocity difference on the sound-speed scale: sum |dv_phys| w / sum cs w,
    with the reference's gamma-law sound speed and (spherical) the physical
    components v^d scale_d.  Meaningful where the flow is near rest, where a
    relative L1 of |v| divides by almost nothing."""
    dv = 0.0
    for d in range(3):
        s = 1.0 if scale is None else scale[d]
        dv = dv + (np.abs(dev[2 + d] - ref[2 + d]) * s) ** 2
    dv = np.sqrt(dv)
    cs = np.sqrt(gam * (gam - 1.0) * np.abs(ref[1]) / ref[0])
    w = np.ones_like(cs) if w is None else np.broadcast_to(w, cs.shape)
    return float(np.sum(dv * w) / np.sum(cs * w))
