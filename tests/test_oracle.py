"""Pinning of the CPU oracle (oracle/) without the reference's own tests.

The reference ships no golden vectors or numeric tests for this path
(tst/regression only compile-tests radiation setups that cannot build), so
each oracle function is pinned by what the mathematics fixes: exactness of
the reconstruction on linear and quadratic data, the closed-form limiter
branches, Riemann-solver consistency / upwinding / symmetry / contact
preservation / Toro's wave-speed formulas, the exact Sod solution,
conservation, free-stream and hydrostatic preservation, and closed-form
geometry.  Each test is chosen so that a dropped term, wrong sign or index,
or transposed operand in the restatement fails it.
"""
import ctypes
import math

import numpy as np
import pytest

from oracle import pyoracle as po
import exact_riemann as er

L = po.lib()


# --------------------------------------------------------------------------
# reconstruction (reconstruct.c:127-213)
# --------------------------------------------------------------------------
def recon(a3, alpha, beta, gam, xlo, xhi):
    a = np.array([a3[0], a3[1], a3[2], 0.0])
    al_ = np.array([alpha[0], alpha[1], alpha[2], 1.0])
    be_ = np.array([beta[0], beta[1], beta[2], 1.0])
    ga_ = np.array([gam[0], gam[1], gam[2], 1.0])
    x = np.array([0.0, xlo, xhi, 0.0])
    l = np.zeros(4)
    r = np.zeros(4)
    L.orc_para_recon(po.ptr(a), po.ptr(l), po.ptr(r), po.ptr(al_), po.ptr(be_), po.ptr(ga_), po.ptr(x), -1, -1)
    return l[1], r[1]


def cart_moments(x0, h):
    xs = x0 + h * np.arange(4)
    alpha = (xs[1:] ** 3 - xs[:-1] ** 3) / 3.0
    beta = (xs[1:] ** 2 - xs[:-1] ** 2) / 2.0
    gam = np.full(3, h)
    return alpha, beta, gam, xs


def sph_moments(x0, h):
    """moments of r^2 dr/dx dx for r = exp(x) in closed form (per zone)."""
    xs = x0 + h * np.arange(4)

    def F(x, k):   # antiderivative of x^k e^{3x}
        e = np.exp(3 * x)
        if k == 0:
            return e / 3
        if k == 1:
            return e * (x / 3 - 1 / 9)
        return e * (x * x / 3 - 2 * x / 9 + 2 / 27)
    gam = F(xs[1:], 0) - F(xs[:-1], 0)
    beta = F(xs[1:], 1) - F(xs[:-1], 1)
    alpha = F(xs[1:], 2) - F(xs[:-1], 2)
    return alpha, beta, gam, xs


@pytest.mark.parametrize("geom", ["cart", "sph"])
def test_recon_exact_on_linear_data(geom):
    rng = np.random.default_rng(1)
    for _ in range(200):
        h = rng.uniform(0.01, 0.1)
        x0 = rng.uniform(-1, 1)
        alpha, beta, gam, xs = cart_moments(x0, h) if geom == "cart" else sph_moments(x0, h)
        c0, c1 = rng.uniform(-1, 1), rng.uniform(-2, 2)
        a3 = c0 + c1 * beta / gam                    # zone means of c0 + c1 x
        l, r = recon(a3, alpha, beta, gam, xs[1], xs[2])
        assert l == pytest.approx(c0 + c1 * xs[1], abs=1e-10)
        assert r == pytest.approx(c0 + c1 * xs[2], abs=1e-10)


def lin(a3, beta, gam, xlo, xhi):
    """lin_recon (reconstruct.c:89-124) of the middle zone"""
    a = np.array([a3[0], a3[1], a3[2], 0.0])
    be_ = np.array([beta[0], beta[1], beta[2], 1.0])
    ga_ = np.array([gam[0], gam[1], gam[2], 1.0])
    x = np.array([0.0, xlo, xhi, 0.0])
    l = np.zeros(4)
    r = np.zeros(4)
    L.orc_lin_recon(po.ptr(a), po.ptr(l), po.ptr(r), po.ptr(be_), po.ptr(ga_), po.ptr(x), -1, -1)
    return l[1], r[1]


@pytest.mark.parametrize("geom", ["cart", "sph"])
def test_lin_recon_exact_on_linear_data(geom):
    """the centred slope of linear data is exact and never the largest of
    the three, so the faces carry the line's values"""
    rng = np.random.default_rng(4)
    for _ in range(200):
        h = rng.uniform(0.01, 0.1)
        alpha, beta, gam, xs = cart_moments(rng.uniform(-1, 1), h) if geom == "cart" else sph_moments(
            rng.uniform(-1, 1), h)
        c0, c1 = rng.uniform(-1, 1), rng.uniform(-2, 2)
        l, r = lin(c0 + c1 * beta / gam, beta, gam, xs[1], xs[2])
        assert l == pytest.approx(c0 + c1 * xs[1], abs=1e-10)
        assert r == pytest.approx(c0 + c1 * xs[2], abs=1e-10)


def test_lin_recon_flat_at_extrema_and_bounded_on_monotone_data():
    rng = np.random.default_rng(5)
    for _ in range(300):
        alpha, beta, gam, xs = sph_moments(rng.uniform(-1, 1), rng.uniform(0.01, 0.2))
        q = np.sort(rng.uniform(-3, 3, 3)) * rng.choice([-1.0, 1.0])
        l, r = lin(q, beta, gam, xs[1], xs[2])
        # no new extrema: each face value lies between the zone and the neighbour on its side
        assert min(q[0], q[1]) - 1e-12 <= l <= max(q[0], q[1]) + 1e-12
        assert min(q[1], q[2]) - 1e-12 <= r <= max(q[1], q[2]) + 1e-12
        peak = np.array([q[0], max(q) + 1.0, q[2]])
        assert lin(peak, beta, gam, xs[1], xs[2]) == (peak[1], peak[1])


def test_lin_recon_slope_saturates_at_the_one_sided_limit():
    """uniform zones of width 1, means 0, 1, M: once M is large the slope is
    the backward one-sided bound (the face midpoint at x_i - 1/4, 3/4 from the
    centroid), so both faces stop depending on M: 1/3 and 5/3"""
    alpha, beta, gam, xs = cart_moments(0.0, 1.0)
    for M in (3.0, 10.0, 1.0e6):
        l, r = lin([0.0, 1.0, M], beta, gam, xs[1], xs[2])
        assert l == pytest.approx(1.0 / 3.0, rel=1e-14) and r == pytest.approx(5.0 / 3.0, rel=1e-14)


@pytest.mark.parametrize("geom", ["cart", "sph"])
def test_recon_exact_on_monotone_quadratic(geom):
    rng = np.random.default_rng(2)
    checked = 0
    for _ in range(400):
        h = rng.uniform(0.01, 0.05)
        x0 = rng.uniform(-0.5, 0.5)
        alpha, beta, gam, xs = cart_moments(x0, h) if geom == "cart" else sph_moments(x0, h)
        c0, c1, c2 = rng.uniform(-1, 1), rng.uniform(1, 3), rng.uniform(-0.5, 0.5)
        f = lambda x: c0 + c1 * x + c2 * x * x
        a3 = c0 + c1 * beta / gam + c2 * alpha / gam  # zone means of f
        vertex = -c1 / (2 * c2)
        if xs[0] - 0.5 < vertex < xs[3] + 0.5:
            continue                                  # keep the limiter out of play
        # the face clamp towards 0.75*nbr+0.25*self must not be active either
        l_true, r_true = f(xs[1]), f(xs[2])
        if (a3[1] - l_true) * (l_true - (0.75 * a3[0] + 0.25 * a3[1])) < 0:
            continue
        if ((0.75 * a3[2] + 0.25 * a3[1]) - r_true) * (r_true - a3[1]) < 0:
            continue
        l, r = recon(a3, alpha, beta, gam, xs[1], xs[2])
        assert l == pytest.approx(l_true, abs=1e-9)
        assert r == pytest.approx(r_true, abs=1e-9)
        checked += 1
    assert checked > 50


def test_recon_extremum_is_flat():
    alpha, beta, gam, xs = cart_moments(0.2, 0.05)
    for a3 in ([1.0, 2.0, 1.5], [2.0, 1.0, 3.0], [1.0, 1.0, 2.0]):
        l, r = recon(np.array(a3), alpha, beta, gam, xs[1], xs[2])
        assert l == a3[1] and r == a3[1]


def test_recon_face_clamp_branch():
    """al outside [mean, a_i] is replaced by mean = 0.75*a_{i-1}+0.25*a_i exactly."""
    alpha, beta, gam, xs = cart_moments(0.0, 0.1)
    a3 = np.array([0.0, 1.0, 4.0])        # unclamped al = 1/6 lies below the mean 0.25
    l, r = recon(a3, alpha, beta, gam, xs[1], xs[2])
    mean = 0.75 * a3[0] + 0.25 * a3[1]
    assert l == mean


def test_recon_relocation_matches_ppm_closed_form():
    """Uniform Cartesian: an interior extremum is pushed onto a face, giving the
    classic PPM values al = 3a - 2ar (vertex at the right face) or ar = 3a - 2al."""
    alpha, beta, gam, xs = cart_moments(0.0, 0.1)
    hits = 0
    rng = np.random.default_rng(5)
    for _ in range(2000):
        a3 = np.sort(rng.uniform(0, 1, 3))
        if rng.uniform() < 0.5:
            a3 = a3[::-1].copy()
        # reproduce the pre-relocation state with the textbook uniform-grid fit
        l0 = (2 * a3[0] + 5 * a3[1] - a3[2]) / 6.0
        r0 = (-a3[0] + 5 * a3[1] + 2 * a3[2]) / 6.0
        m = 0.75 * a3[0] + 0.25 * a3[1]
        if (a3[1] - l0) * (l0 - m) < 0:
            l0 = m
        m = 0.75 * a3[2] + 0.25 * a3[1]
        if (m - r0) * (r0 - a3[1]) < 0:
            r0 = m
        b2 = (a3[1] - l0 - (r0 - l0) * 0.5) * (-6.0)
        b1 = (r0 - l0) - b2
        if b2 == 0:
            continue
        sz = -b1 / (2 * b2)
        if not (0.02 < sz < 0.98) or abs(sz - 0.5) < 0.02:
            continue
        l, r = recon(a3, alpha, beta, gam, xs[1], xs[2])
        if sz > 0.5:
            assert r == pytest.approx(r0, abs=1e-12)
            assert l == pytest.approx(3 * a3[1] - 2 * r0, abs=1e-10)
        else:
            assert l == pytest.approx(l0, abs=1e-12)
            assert r == pytest.approx(3 * a3[1] - 2 * l0, abs=1e-10)
        hits += 1
    assert hits > 100


# --------------------------------------------------------------------------
# Riemann solvers (riemann.c:178-499)
# --------------------------------------------------------------------------
def geom(area=1.0, scale=1.0, gcov=(1.0, 1.0, 1.0), d=0):
    g = po.ZoneGeom()
    g.area[d] = area
    g.scale[(1 + d) * 3 + d] = scale
    for q in range(3):
        g.gcov[(1 + d) * 3 + q] = gcov[q]
    return g


def solve(pl, pr, g, d=0, hlle=False):
    fn = L.orc_riemann_hlle if hlle else L.orc_riemann_hllc
    a = np.ascontiguousarray(pl, dtype=float)
    b = np.ascontiguousarray(pr, dtype=float)
    f = np.zeros(5)
    s = np.zeros(1)
    vr = np.zeros(3)
    sh = fn(5, 5, po.ptr(a), po.ptr(b), ctypes.byref(g), d, po.ptr(f), po.ptr(s), po.ptr(vr))
    return f, s[0], sh, vr


def state(rho, p, v, gam=1.4):
    return np.array([rho, p / (gam - 1), v[0], v[1], v[2], p, gam])


def phys_flux(st, gcov, d, area):
    rho, uu, v, p = st[0], st[1], st[2:5], st[5]
    vcov = np.array(gcov) * v
    E = 0.5 * rho * np.dot(v, vcov) + uu
    f = np.zeros(5)
    f[0] = rho * v[d]
    f[1] = v[d] * (E + p)
    f[2:5] = rho * v[d] * vcov
    f[2 + d] += p
    return area * f


@pytest.mark.parametrize("hlle", [False, True])
def test_riemann_consistency(hlle):
    rng = np.random.default_rng(3)
    for _ in range(100):
        d = int(rng.integers(0, 3))
        gcov = rng.uniform(0.5, 2.0, 3)
        g = geom(area=rng.uniform(0.5, 2), scale=math.sqrt(gcov[d]), gcov=gcov, d=d)
        st = state(rng.uniform(0.1, 2), rng.uniform(0.1, 2), rng.uniform(-0.5, 0.5, 3))
        f, s, _, _ = solve(st, st, g, d, hlle)
        assert np.allclose(f, phys_flux(st, gcov, d, g.area[d]), rtol=1e-12, atol=1e-13)


def test_riemann_supersonic_upwinding():
    g = geom(area=1.3)
    l = state(1.0, 1.0, [5.0, 0.2, -0.1])
    r = state(0.5, 0.7, [4.0, 0.0, 0.3])
    f, s, _, _ = solve(l, r, g)
    assert np.allclose(f, phys_flux(l, (1, 1, 1), 0, 1.3), rtol=1e-14)
    f, s, _, _ = solve(state(1.0, 1.0, [-5, 0, 0]), state(0.5, 0.7, [-4.0, 0.1, 0]), g)
    assert np.allclose(f, phys_flux(state(0.5, 0.7, [-4.0, 0.1, 0]), (1, 1, 1), 0, 1.3), rtol=1e-14)


@pytest.mark.parametrize("hlle", [False, True])
def test_riemann_mirror_symmetry(hlle):
    rng = np.random.default_rng(4)
    g = geom()
    for _ in range(100):
        l = state(rng.uniform(0.1, 2), rng.uniform(0.1, 2), rng.uniform(-1, 1, 3))
        r = state(rng.uniform(0.1, 2), rng.uniform(0.1, 2), rng.uniform(-1, 1, 3))
        lm, rm = r.copy(), l.copy()
        lm[2] *= -1
        rm[2] *= -1
        f, s, _, _ = solve(l, r, g, 0, hlle)
        fm, sm, _, _ = solve(lm, rm, g, 0, hlle)
        assert fm[0] == pytest.approx(-f[0], rel=1e-12, abs=1e-14)
        assert fm[1] == pytest.approx(-f[1], rel=1e-12, abs=1e-14)
        assert fm[2] == pytest.approx(f[2], rel=1e-12, abs=1e-14)
        assert sm == pytest.approx(s, rel=1e-12)


def test_hllc_stationary_contact_has_no_mass_flux():
    g = geom(area=2.0)
    f, s, sh, _ = solve(state(1.0, 1.0, [0, 0.3, 0]), state(0.1, 1.0, [0, -0.2, 0]), g)
    assert f[0] == 0.0 and f[1] == 0.0
    assert f[2] == 2.0 * 1.0
    assert sh == 0


def test_riemann_zero_area_face():
    f, s, sh, _ = solve(state(1, 1, [0.1, 0, 0]), state(0.5, 0.5, [0, 0, 0]), geom(area=0.0))
    assert np.all(f == 0.0) and s == 1.0e-200


def toro_speeds(l, r, gam):
    """Independent restatement of Toro's adaptive wave-speed estimate (eqs. 9.32-9.43)."""
    rl, ul, pl = l[0], l[2], l[5]
    rr, ur, pr = r[0], r[2], r[5]
    cl, cr = math.sqrt(gam * pl / rl), math.sqrt(gam * pr / rr)
    pmin, pmax = min(pl, pr), max(pl, pr)
    pv = max(0.0, 0.5 * (pl + pr) - 0.125 * (ur - ul) * (rl + rr) * (cl + cr))
    if pmax / pmin < 2 and pmin <= pv <= pmax:
        ps, kind = pv, "pvrs"
    elif pv < pmin:
        z = (gam - 1) / (2 * gam)
        plr = (pl / pr) ** z
        us = (plr * ul / cl + ur / cr + 2 * (plr - 1) / (gam - 1)) / (plr / cl + 1 / cr)
        with np.errstate(invalid="ignore"):
            ps = 0.5 * (pl * np.float64(1 + (gam - 1) / (2 * cl) * (ul - us)) ** (1 / z)
                        + pr * np.float64(1 + (gam - 1) / (2 * cr) * (us - ur)) ** (1 / z))
        ps = ps if ps > 0 else 0.0      # MAX(p_s, 0.0) (decs.h:628) also maps a vacuum NaN to 0
        kind = "trrs"
    else:
        gl = math.sqrt(2 / ((gam + 1) * rl) / (pv + (gam - 1) / (gam + 1) * pl))
        gr = math.sqrt(2 / ((gam + 1) * rr) / (pv + (gam - 1) / (gam + 1) * pr))
        ps, kind = max(0.0, (gl * pl + gr * pr - (ur - ul)) / (gl + gr)), "tsrs"
    q = lambda p, pk: 1.0 if p <= pk else math.sqrt(1 + (gam + 1) / (2 * gam) * (p / pk - 1))
    sl, sr = ul - cl * q(ps, pl), ur + cr * q(ps, pr)
    ss = (pr - pl + rl * ul * (sl - ul) - rr * ur * (sr - ur)) / (rl * (sl - ul) - rr * (sr - ur))
    if ss > sr:
        sr = 1.001 * ss
    if ss < sl:
        sl = 1.001 * ss
    return max(abs(sl), abs(sr)), kind


def test_hllc_wave_speed_branches_match_toro():
    rng = np.random.default_rng(6)
    kinds = set()
    for _ in range(600):
        l = state(rng.uniform(0.1, 3), 10 ** rng.uniform(-1, 1), [rng.uniform(-3, 3), 0, 0])
        r = state(rng.uniform(0.1, 3), 10 ** rng.uniform(-1, 1), [rng.uniform(-3, 3), 0, 0])
        _, s, sh, _ = solve(l, r, geom())
        ref, kind = toro_speeds(l, r, 1.4)
        kinds.add(kind)
        assert s == pytest.approx(ref, rel=1e-10)
        pmin, pmax = min(l[5], r[5]), max(l[5], r[5])
        assert sh == int((pmax - pmin) / pmin > 0.5)
    assert kinds == {"pvrs", "trrs", "tsrs"}


def test_p_u_roundtrip():
    rng = np.random.default_rng(8)
    for _ in range(100):
        p = np.array([rng.uniform(0.1, 2), rng.uniform(0.1, 2), *rng.uniform(-1, 1, 3)])
        g = rng.uniform(0.5, 2, 3)
        u = np.zeros(5)
        L.orc_p_to_u(po.ptr(p), po.ptr(u), po.ptr(g), 5)
        assert u[1] == pytest.approx(0.5 * p[0] * np.sum(g * p[2:] ** 2) + p[1], rel=1e-14)
        q = np.zeros(5)
        L.orc_u_to_p(po.ptr(u), po.ptr(q), po.ptr(1.0 / g), 5)
        assert np.allclose(q, p, rtol=1e-13, atol=1e-14)


# --------------------------------------------------------------------------
# whole-step behaviour
# --------------------------------------------------------------------------
def sod_sim(n):
    cfg = po.make_config(1, [n], [0.0], [1.0 / n], [0], [0], 1.4, cfl=0.5)
    s = po.Sim(cfg)
    x = (np.arange(n) + 0.5) / n
    p = np.zeros((5, 1, 1, n))
    p[0] = np.where(x < 0.5, 1.0, 0.125)
    p[1] = np.where(x < 0.5, 1.0, 0.1) / 0.4
    s.set_prims(p)
    return s


def test_sod_against_exact_solution_and_conservation():
    errs = {}
    for n in (256, 1024):
        s = sod_sim(n)
        m0 = s.totals()
        s.run(0.2)
        m1 = s.totals()
        assert abs(s.t - 0.2) < 1e-12
        assert abs(m1[0] - m0[0]) <= 1e-15 * abs(m0[0])     # no boundary flux before t=0.2
        assert abs(m1[1] - m0[1]) <= 1e-14 * abs(m0[1])
        q = s.get_prims()
        xl = np.arange(n) / n
        r, u, p = er.cell_average(xl, xl + 1.0 / n, 0.2, 0.5, (1, 0, 1), (0.125, 0, 0.1), 1.4)
        errs[n] = np.sum(np.abs(q[0, 0, 0] - r)) / np.sum(r)
    assert errs[1024] < 1.5e-3
    assert errs[256] / errs[1024] > 2.5      # converging (first order at discontinuities)


def test_periodic_2d_conservation_and_free_stream():
    n = 32
    cfg = po.make_config(2, [n, n], [0, 0], [1.0 / n, 1.0 / n], [3, 3], [3, 3], 1.4)
    s = po.Sim(cfg)
    p = np.zeros((5, 1, n, n))
    p[0], p[1], p[2], p[3] = 1.3, 2.0, 0.3, -0.2
    s.set_prims(p)
    s.run(max_steps=20)
    q = s.get_prims()
    for k in range(4):
        assert np.max(np.abs(q[k] - p[k])) < 1e-14
    rng = np.random.default_rng(9)
    p[0] = 1 + 0.3 * rng.uniform(size=(1, n, n))
    p[2] = 0.2 * rng.uniform(-1, 1, size=(1, n, n))
    s = po.Sim(cfg)
    s.set_prims(p)
    t0 = s.totals()
    s.run(max_steps=20)
    t1 = s.totals()
    assert abs(t1[0] - t0[0]) < 1e-14 * t0[0]
    assert abs(t1[1] - t0[1]) < 1e-14 * t0[1]


def test_spherical_hydrostatic_uniform_state_stays_static():
    n = 200
    x0, x1 = math.log(1e5), math.log(1e9)
    cfg = po.make_config(1, [n], [x0], [(x1 - x0) / n], [1], [0], 1.4, geom=po.GEOM_SPHERICAL,
                         rcoord=po.RCOORD_LOG)
    s = po.Sim(cfg)
    p = np.zeros((5, 1, 1, n))
    p[0], p[1] = 1.0, 1.0 / 0.4
    s.set_prims(p)
    s.run(max_steps=50)
    q = s.get_prims()
    cs = math.sqrt(1.4)
    assert np.max(np.abs(q[2] * np.exp(x0 + (np.arange(n) + 0.5) * (x1 - x0) / n))) < 1e-12 * cs
    assert np.max(np.abs(q[0] - 1.0)) < 1e-13


def test_spherical_geometry_closed_forms():
    n = 64
    x0, x1 = math.log(1e5), math.log(1e9)
    h = (x1 - x0) / n
    cfg = po.make_config(1, [n], [x0], [h], [1], [0], 1.4, geom=po.GEOM_SPHERICAL, rcoord=po.RCOORD_LOG)
    s = po.Sim(cfg)
    geo = s.geometry()
    al, be, ga = s.moments(0)
    for i in range(-2, n + 2):
        r0, r1 = math.exp(x0 + i * h), math.exp(x0 + (i + 1) * h)
        z = geo[i + 2]
        assert z[po.GEOM_VOLUME] == pytest.approx(4 * math.pi / 3 * (r1 ** 3 - r0 ** 3), rel=1e-12)
        assert ga[i + 2] == pytest.approx((r1 ** 3 - r0 ** 3) / 3, rel=1e-12)
        if 0 <= i <= n:
            assert z[po.GEOM_AREA] == pytest.approx(4 * math.pi * r0 ** 2 * r0, rel=1e-12)   # 4 pi r^2 dr/dx
        assert z[po.GEOM_GCOV + 3] == pytest.approx(r0 * r0, rel=1e-12)                   # face g_00 = (dr/dx)^2
    am, bm, gm, xs = sph_moments(x0 + 10 * h, h)
    assert al[12] == pytest.approx(am[0], rel=1e-10)
    assert be[12] == pytest.approx(bm[0], rel=1e-10)


def test_cartesian_geometry():
    cfg = po.make_config(3, [8, 8, 8], [-0.25, -0.25, -0.75], [0.5 / 8, 0.5 / 8, 1.5 / 8], [3, 3, 1], [3, 3, 1], 1.4)
    s = po.Sim(cfg)
    geo = s.geometry()
    z = geo[500]
    assert z[po.GEOM_VOLUME] == pytest.approx(0.5 / 8 * 0.5 / 8 * 1.5 / 8, rel=1e-15)
    assert z[po.GEOM_AREA + 2] == pytest.approx((0.5 / 8) ** 2, rel=1e-15)
    al, be, ga = s.moments(2)
    xs = -0.75 + 1.5 / 8 * (np.arange(12) - 2)
    assert np.allclose(ga, 1.5 / 8)
    assert np.allclose(be, (xs + 1.5 / 8) ** 2 / 2 - xs ** 2 / 2, rtol=1e-14)
