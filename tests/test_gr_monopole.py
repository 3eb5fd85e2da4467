"""GR effective monopole gravity (gravity.c:142-251, 266-375 with
GR_MONOPOLE): pins of the oracle's restatement by what the mathematics
fixes in the weak field, where every correction is first order in a small
parameter and has a closed form.

  * compactness: a static pressureless uniform sphere of mass M and radius R
    (eps = G M / (R c^2)) has TOV mass m(r) = m_N(r) (1 - 3/5 eps r^2/R^2)
    and Gamma^2 = 1 - 2 G m / (r c^2), so at a face inside it
    g_GR / g_N - 1 = 1.4 eps r^2 / R^2 (+ O(eps^2)); a dropped Gamma in the
    mass integral gives 2.0, a dropped Gamma^2 below gives -0.6;
  * pressure and internal energy: with uniform P and u and negligible
    compactness, g_GR / g_N - 1 = (2 u + 4 P) / (rho c^2) (u in the TOV mass
    and in the inertia factor, P in the active mass 4 pi r^3 P / c^2 and in
    the inertia factor);
  * motion: a uniform radial velocity v raises Gamma^2 by v^2/c^2 and the
    Gamma-weighted mass by v^2/(2 c^2), so g changes by -v^2/(2 c^2);
  * the lapse outside the mass is exp(-G M / (r c^2)) (the potential is
    integrated inwards from -G M / R_out);
  * the iteration converges in a few passes and, warm-started from the kept
    Gamma, in no more passes the second time.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as po

G, C = 6.67384e-8, 2.99792458e10


def _sim(n1, rout, prims_fn, gam=5.0 / 3.0):
    cfg = po.make_config(1, [n1], [0.0], [rout / n1], [po.BC_REFLECT], [po.BC_OUTFLOW], gam, nvars=5,
                         geom=po.GEOM_SPHERICAL, rcoord=po.RCOORD_UNIFORM, monopole=2)
    sim = po.Sim(cfg)
    r = (np.arange(n1) + 0.5) * rout / n1
    sim.set_prims(prims_fn(r).reshape(5, 1, 1, n1))
    return sim, np.arange(n1 + 1) * rout / n1


def _newtonian_edges(sim, rf, rho_shell_mass):
    m = np.concatenate([[0.0], np.cumsum(rho_shell_mass)])
    g = np.zeros_like(rf)
    g[1:] = -G * m[1:] / rf[1:] ** 2
    return g, m


def _uniform_sphere(n1, rout, R, rho0, P=0.0, u=0.0, v=0.0, gam=5.0 / 3.0):
    def f(r):
        p = np.zeros((5, len(r)))
        inside = r < R
        p[0] = np.where(inside, rho0, rho0 * 1e-12)
        p[1] = np.where(inside, u, u * 1e-12) if u else 1e-30 * p[0]
        p[2] = v
        return p
    # gamma law: P = (gam - 1) u; a pressure P given on its own sets u
    if P and not u:
        u = P / (gam - 1.0)
    sim, rf = _sim(n1, rout, f, gam)
    return sim, rf


def test_gr_edges_first_order_compactness_correction():
    n1, R, rout = 400, 1.0e8, 2.0e8
    rho0 = 1.0e9                               # eps = G M / (R c^2) ~ 3e-3
    sim, rf = _uniform_sphere(n1, rout, R, rho0)
    g = sim.gr_edges()
    vsh = 4.0 * math.pi / 3.0 * (rf[1:] ** 3 - rf[:-1] ** 3)
    rho = np.where((rf[:-1] + rf[1:]) / 2 < R, rho0, rho0 * 1e-12)
    gN, m = _newtonian_edges(sim, rf, vsh * rho)
    M = m[-1]
    eps = G * M / (R * C ** 2)
    inside = (rf > 0.2 * R) & (rf < 0.95 * R)
    dev = g[inside] / gN[inside] - 1.0
    want = 1.4 * eps * (rf[inside] / R) ** 2
    print(f"eps {eps:.3e}; dev/want range {np.min(dev / want):.5f} .. {np.max(dev / want):.5f}")
    assert np.all(np.abs(dev / want - 1.0) < 0.01)
    st = sim.gr_state()
    assert st["fail"] == 0 and 1 <= st["iter"] <= 4


def test_gr_edges_pressure_and_internal_energy_terms():
    n1, R, rout = 200, 1.0e8, 2.0e8
    rho0 = 1.0e-2                              # compactness ~ 3e-16: negligible
    gam = 5.0 / 3.0
    u = 1.0e-4 * rho0 * C ** 2
    sim, rf = _uniform_sphere(n1, rout, R, rho0, u=u, gam=gam)
    P = (gam - 1.0) * u
    g = sim.gr_edges()
    vsh = 4.0 * math.pi / 3.0 * (rf[1:] ** 3 - rf[:-1] ** 3)
    rho = np.where((rf[:-1] + rf[1:]) / 2 < R, rho0, rho0 * 1e-12)
    gN, _ = _newtonian_edges(sim, rf, vsh * rho)
    inside = (rf > 0.1 * R) & (rf < 0.9 * R)
    dev = g[inside] / gN[inside] - 1.0
    want = (2.0 * u + 4.0 * P) / (rho0 * C ** 2)
    print(f"want {want:.4e}; dev {np.min(dev):.4e} .. {np.max(dev):.4e}")
    assert np.all(np.abs(dev / want - 1.0) < 2e-3)


def test_gr_edges_velocity_term():
    n1, R, rout = 200, 1.0e8, 2.0e8
    rho0, v = 1.0e-2, 3.0e8                    # v / c = 1e-2, negligible compactness
    sim0, rf = _uniform_sphere(n1, rout, R, rho0)
    sim1, _ = _uniform_sphere(n1, rout, R, rho0, v=v)
    g0, g1 = sim0.gr_edges(), sim1.gr_edges()
    sel = (rf > 0.1 * R) & (rf < 0.9 * R)
    ratio = g1[sel] / g0[sel] - 1.0
    want = -0.5 * (v / C) ** 2
    print(f"want {want:.4e}; got {np.min(ratio):.4e} .. {np.max(ratio):.4e}")
    assert np.all(np.abs(ratio / want - 1.0) < 1e-3)


def test_gr_lapse_outside_the_mass_and_warm_start():
    n1, R, rout = 400, 1.0e8, 4.0e8
    rho0 = 1.0e6                               # eps ~ 3e-6: GR terms in the potential ~ eps^2
    sim, rf = _uniform_sphere(n1, rout, R, rho0)
    sim.gr_edges()
    st = sim.gr_state()
    vsh = 4.0 * math.pi / 3.0 * (rf[1:] ** 3 - rf[:-1] ** 3)
    rho = np.where((rf[:-1] + rf[1:]) / 2 < R, rho0, rho0 * 1e-12)
    M = float(np.sum(vsh * rho))
    out = rf > 1.5 * R
    want = np.exp(-G * M / (rf[out] * C ** 2))
    rel = np.abs(st["lapse_edge"][out] / want - 1.0)
    print(f"lapse outside: max rel {rel.max():.3e}; phi/c^2 at R_out {G * M / (rout * C ** 2):.3e}")
    assert rel.max() < 1e-9                    # of an effect 1 - lapse ~ 3e-6
    assert st["lapse_edge"][-1] == pytest.approx(math.exp(-G * M / (rout * C ** 2)), rel=1e-15)
    # warm start: the kept Gamma is already converged for the same state
    it1 = st["iter"]
    sim.gr_edges()
    assert sim.gr_state()["iter"] <= it1
