"""Pins of the oracle's dendritic pencil entries and polar-axis phi-face
states, evaluated one entry at a time (orc_dend_radial_entry,
orc_dend_theta_entry, orc_polar_face_states).

Reference: kslopes0/1 (step.c:783-896), prolongate0/1 (step.c:905-1010),
restrict0/1 (step.c:1012-1085), the polar-axis phi faces (step.c:1299-1397).

What the mathematics fixes, independently of the restatement:
  * a field linear in the zone centroid coordinates (beta/Gamma of each
    zone's own level) is reproduced exactly at the target zone's centroid
    by every entry: prolongation evaluates the limited linear fit, whose
    three slopes coincide so the limiter is idle; restriction takes the
    volume-weighted mean, which for a separable volume element equals the
    value at the parent's centroid (integrals are additive);
  * the restricted Gamma_1 entry of the radial pencil is the mean of v^3,
    not of Gamma_1 (the reference indexes p[...][GAMMA], step.c:654, 1048);
  * at a kink the slope is van Leer's monotonized-central slope,
    minmod(2 s-, 2 s+, s_c) (three-argument form), and zero at an extremum;
  * a uniform Cartesian velocity field, written in each zone's spherical
    basis, arrives at a polar phi face as its projection onto the face's
    basis divided by the face scale factors, and the other entries are the
    zone values (first order).
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from test_dendritic import dend_config
from test_sph3d import sph3d_config, th_of_x

GAM = 4.0 / 3.0


def _setup(n, nvars=5):
    cfg = dend_config(n, gam=GAM, nvars=nvars)
    sim = po.Sim(cfg)
    nj, nk, mj, mk = sim.levels()
    th = [sim.level_moments(1, l) for l in range(mj + 1)]
    ph = [sim.level_moments(2, l) for l in range(mk + 1)]
    return cfg, sim, nj, nk, th, ph


def _lev(f):
    return int(round(math.log2(f)))


class Grid:
    """Level counts and centroid coordinates of a dendritic oracle grid."""

    def __init__(self, n, nvars=5):
        self.n = n
        self.cfg, self.sim, self.nj, self.nk, self.th, self.ph = _setup(n, nvars)
        self.dx1, self.dx2 = 2.0 / n[1], 2.0 * math.pi / n[2]

    def dj(self, i):
        return self.n[1] // self.nj[i]

    def dk(self, i, j):
        return self.n[2] // self.nk[i, j]

    def x1c(self, i, j):
        """theta centroid coordinate of row j of shell i (its own level)"""
        a, b, g = self.th[_lev(self.dj(i))]
        return b[j + 2] / g[j + 2]

    def x2c_lev(self, dk, k):
        a, b, g = self.ph[_lev(dk)]
        return b[k + 2] / g[k + 2]

    def x1_range(self, i, j):
        w = self.dj(i) * self.dx1
        return -1.0 + j * w, -1.0 + (j + 1) * w

    def x2_range_lev(self, dk, k):
        w = dk * self.dx2
        return k * w, (k + 1) * w

    def rows_near(self, ii, lo, hi):
        """rows of shell ii overlapping [lo, hi] in x1 (None if it leaves [-1, 1])"""
        if lo < -1.0 or hi > 1.0:
            return None
        w = self.dj(ii) * self.dx1
        return range(max(0, int(math.floor((lo + 1.0) / w))), min(self.nj[ii], int(math.ceil((hi + 1.0) / w))))

    def phi_inside(self, ii, rows, lo, hi):
        """every row's two-zone margin around [lo, hi] stays inside [0, 2 pi)"""
        for r in rows:
            w = self.dk(ii, r) * self.dx2
            if lo - 2.0 * w < 0.0 or hi + 2.0 * w > 2.0 * math.pi:
                return False
        return True


def _linear_prims(G, coef, nvars=5):
    """every variable v = c0 + c1 x1c + c2 x2c over the valid zones"""
    n = G.n
    prims = np.zeros((nvars, n[2], n[1], n[0]))
    for i in range(n[0]):
        for j in range(G.nj[i]):
            x1 = G.x1c(i, j)
            dk = G.dk(i, j)
            for k in range(G.nk[i, j]):
                x2 = G.x2c_lev(dk, k)
                for v in range(nvars):
                    c0, c1, c2 = coef[v]
                    prims[v, k, j, i] = c0 + c1 * x1 + c2 * x2
    return prims


# rho, u, v1, v2, v3: positive rho and u, all slopes non-zero
COEF = [(5.0, 0.7, 0.3), (9.0, -1.1, 0.45), (0.2, 0.9, -0.35), (-0.4, 0.25, 0.6), (1.5, -0.8, 0.2)]


def _lin(v, x1, x2):
    c0, c1, c2 = COEF[v]
    return c0 + c1 * x1 + c2 * x2


@pytest.mark.parametrize("n", [(20, 32, 64), (24, 16, 32)])
def test_radial_entries_reproduce_linear_fields(n):
    G = Grid(n)
    G.sim.set_prims(_linear_prims(G, COEF))
    G.sim.prepare_entries()
    checked = {"prolong": 0, "restrict": 0, "prolong_interp": 0}
    for i in range(2, n[0] - 1):
        for j in range(G.nj[i]):
            lo1, hi1 = G.x1_range(i, j)
            x1 = G.x1c(i, j)
            for k in range(G.nk[i, j]):
                dk = G.dk(i, j)
                lo2, hi2 = G.x2_range_lev(dk, k)
                x2 = G.x2c_lev(dk, k)
                ok = True
                for ii in range(i - 2, i + 2):
                    w1 = G.dj(ii) * G.dx1
                    rows = G.rows_near(ii, lo1 - 2.0 * w1, hi1 + 2.0 * w1)
                    if rows is None or not G.phi_inside(ii, rows, lo2, hi2):
                        ok = False
                if not ok:
                    continue
                for ii in range(i - 2, i + 2):
                    e = G.sim.radial_entry(i, j, k, ii)
                    for v in range(5):
                        want = _lin(v, x1, x2)
                        assert abs(e[v] - want) <= 1e-12 * (1.0 + abs(want)), (i, j, k, ii, v, e[v], want)
                    want_p = (GAM - 1.0) * _lin(1, x1, x2)
                    assert abs(e[5] - want_p) <= 1e-12 * abs(want_p)
                    if G.dj(ii) < G.dj(i):
                        # restriction: Gamma_1 entry is the mean of v^3 (the reference's index quirk)
                        want_g = _lin(4, x1, x2)
                        assert abs(e[6] - want_g) <= 1e-12 * (1.0 + abs(want_g)), (i, j, k, ii, e[6], want_g)
                        checked["restrict"] += 1
                    else:
                        assert e[6] == GAM
                        checked["prolong"] += 1
                        if G.dj(ii) > G.dj(i) or G.dk(ii, j * G.dj(i) // G.dj(ii)) != dk:
                            checked["prolong_interp"] += 1
    print(f"[pencil] radial entries checked: {checked}")
    assert checked["restrict"] >= 10 and checked["prolong_interp"] >= 100 and checked["prolong"] >= 1000


@pytest.mark.parametrize("n", [(20, 32, 64), (24, 16, 32)])
def test_theta_entries_reproduce_linear_fields(n):
    G = Grid(n)
    G.sim.set_prims(_linear_prims(G, COEF))
    G.sim.prepare_entries()
    kinds = {"restrict": 0, "same": 0, "prolong": 0}
    for i in range(n[0]):
        for j in range(2, G.nj[i] - 1):
            dkf = min(G.dk(i, j), G.dk(i, j - 1))
            for kk in range(n[2] // dkf):
                lo2, hi2 = G.x2_range_lev(dkf, kk)
                if not G.phi_inside(i, range(j - 2, j + 2), lo2, hi2):
                    continue
                x2 = G.x2c_lev(dkf, kk)
                for jj in range(j - 2, j + 2):
                    e = G.sim.theta_entry(i, j, kk, jj)
                    x1 = G.x1c(i, jj)
                    for v in range(5):
                        want = _lin(v, x1, x2)
                        assert abs(e[v] - want) <= 1e-12 * (1.0 + abs(want)), (i, j, kk, jj, v, e[v], want)
                    assert abs(e[5] - (GAM - 1.0) * _lin(1, x1, x2)) <= 1e-12 * abs(e[5])
                    assert abs(e[6] - GAM) <= 1e-15       # restrict1 averages Gamma_1 itself
                    dkr = G.dk(i, jj)
                    kinds["restrict" if dkr < dkf else ("same" if dkr == dkf else "prolong")] += 1
    print(f"[pencil] theta entries checked: {kinds}")
    assert min(kinds.values()) >= 100


def _minmod3(a, b, c):
    if a > 0 and b > 0 and c > 0:
        return min(a, b, c)
    if a < 0 and b < 0 and c < 0:
        return max(a, b, c)
    return 0.0


def _mc_value(xs, fs, xp):
    """van Leer's monotonized-central limited linear value at xp from three zones"""
    sm = (fs[1] - fs[0]) / (xs[1] - xs[0])
    sp = (fs[2] - fs[1]) / (xs[2] - xs[1])
    sc = (fs[2] - fs[0]) / (xs[2] - xs[0])
    return fs[1] + _minmod3(2.0 * sm, 2.0 * sp, sc) * (xp - xs[1]), sm, sp


@pytest.mark.parametrize("s_lo,s_hi", [(0.5, 4.0), (-1.0, 2.0), (3.0, 0.25)])
def test_theta_prolongation_limits_like_monotonized_central(s_lo, s_hi):
    """Kinked rho(x1) (constant in phi): the radial entries of coarser rows
    take the MC-limited theta slope at the target centroid."""
    n = (20, 32, 64)
    G = Grid(n)
    xk = 0.137

    def f(x1):
        return 10.0 + (s_lo if x1 < xk else s_hi) * (x1 - xk)

    prims = np.zeros((5, n[2], n[1], n[0]))
    for i in range(n[0]):
        for j in range(G.nj[i]):
            prims[0, :G.nk[i, j], j, i] = f(G.x1c(i, j))
    prims[1] = 3.0
    G.sim.set_prims(prims * (prims[0] != 0))
    G.sim.prepare_entries()
    kinked = 0
    for i in range(2, n[0] - 1):
        for j in range(G.nj[i]):
            x1p = G.x1c(i, j)
            for ii in range(i - 2, i + 2):
                if G.dj(ii) <= G.dj(i):
                    continue
                w = G.dj(ii) * G.dx1
                J = int(math.floor((x1p + 1.0) / w))           # row of shell ii holding the centroid
                if J - 1 < 0 or J + 1 >= G.nj[ii]:
                    continue
                xs = [G.x1c(ii, J + q) for q in (-1, 0, 1)]
                want, sm, sp = _mc_value(xs, [f(x) for x in xs], x1p)
                e = G.sim.radial_entry(i, j, 0, ii)
                assert abs(e[0] - want) <= 1e-12 * abs(want), (i, j, ii, e[0], want, sm, sp)
                kinked += abs(sm - sp) > 1e-9
    print(f"[pencil] MC theta prolongation, {kinked} stencils across the kink")
    assert kinked >= 4


@pytest.mark.parametrize("s_lo,s_hi", [(0.5, 4.0), (-1.0, 2.0), (3.0, 0.25)])
def test_phi_prolongation_limits_like_monotonized_central(s_lo, s_hi):
    """Kinked rho(x2) (constant in theta): theta-pencil entries of rows
    coarser in phi than the face level take the MC-limited phi slope."""
    n = (20, 32, 64)
    G = Grid(n)
    xk = 2.9

    def f(x2):
        return 10.0 + (s_lo if x2 < xk else s_hi) * (x2 - xk)

    prims = np.zeros((5, n[2], n[1], n[0]))
    for i in range(n[0]):
        for j in range(G.nj[i]):
            dk = G.dk(i, j)
            for k in range(G.nk[i, j]):
                prims[0, k, j, i] = f(G.x2c_lev(dk, k))
    prims[1] = 3.0 * (prims[0] != 0)
    G.sim.set_prims(prims)
    G.sim.prepare_entries()
    kinked = checked = 0
    for i in range(n[0]):
        for j in range(1, G.nj[i]):
            dkf = min(G.dk(i, j), G.dk(i, j - 1))
            for jj in (j - 1, j):
                dkr = G.dk(i, jj)
                if dkr <= dkf:
                    continue
                for kk in range(n[2] // dkf):
                    x2p = G.x2c_lev(dkf, kk)
                    K = int(math.floor(x2p / (dkr * G.dx2)))
                    if K - 1 < 0 or K + 1 >= G.nk[i, jj]:
                        continue
                    xs = [G.x2c_lev(dkr, K + q) for q in (-1, 0, 1)]
                    want, sm, sp = _mc_value(xs, [f(x) for x in xs], x2p)
                    e = G.sim.theta_entry(i, j, kk, jj)
                    assert abs(e[0] - want) <= 1e-12 * abs(want), (i, j, kk, jj, e[0], want)
                    checked += 1
                    kinked += abs(sm - sp) > 1e-9
    print(f"[pencil] MC phi prolongation: {checked} entries, {kinked} across the kink")
    assert checked >= 50 and kinked >= 4


def test_restriction_is_volume_weighted_mean():
    """Random state: every restricted radial entry is the mean of the finer
    zones nested in the target zone, weighted by the zone volumes."""
    n = (20, 32, 64)
    G = Grid(n)
    rng = np.random.default_rng(11)
    prims = np.zeros((5, n[2], n[1], n[0]))
    prims[0] = rng.uniform(1.0, 2.0, prims[0].shape)
    prims[1] = rng.uniform(1.0, 2.0, prims[0].shape)
    prims[2:5] = rng.uniform(-1.0, 1.0, prims[2:5].shape)
    G.sim.set_prims(prims)
    G.sim.prepare_entries()
    geo = G.sim.geometry()
    N0, N1 = n[0] + 4, n[1] + 4

    def vol(i, j, k):
        return geo[(i + 2) + N0 * ((j + 2) + N1 * (k + 2)), po.GEOM_VOLUME]

    checked = 0
    for i in range(2, n[0] - 1):
        for j in range(G.nj[i]):
            lo1, hi1 = G.x1_range(i, j)
            for k in range(G.nk[i, j]):
                lo2, hi2 = G.x2_range_lev(G.dk(i, j), k)
                for ii in range(i - 2, i + 2):
                    if G.dj(ii) >= G.dj(i):
                        continue
                    # children: zones of shell ii whose theta and phi intervals lie in the target's
                    ws, vals = [], []
                    for jj in range(G.nj[ii]):
                        a1, b1 = G.x1_range(ii, jj)
                        if a1 < lo1 - 1e-12 or b1 > hi1 + 1e-12:
                            continue
                        dkc = G.dk(ii, jj)
                        for kk in range(G.nk[ii, jj]):
                            a2, b2 = G.x2_range_lev(dkc, kk)
                            if a2 < lo2 - 1e-12 or b2 > hi2 + 1e-12:
                                continue
                            ws.append(vol(ii, jj, kk))
                            vals.append(prims[:, kk, jj, ii])
                    ws = np.array(ws)
                    vals = np.array(vals)
                    assert len(ws) >= 2
                    want = (ws[:, None] * vals).sum(0) / ws.sum()
                    e = G.sim.radial_entry(i, j, k, ii)
                    assert np.allclose(e[:5], want, rtol=1e-13, atol=1e-13), (i, j, k, ii)
                    assert abs(e[5] - (GAM - 1.0) * want[1]) <= 1e-13 * abs(e[5])
                    assert abs(e[6] - want[4]) <= 1e-13                    # mean of v^3 (quirk)
                    checked += 1
    print(f"[pencil] restricted entries checked: {checked}")
    assert checked >= 50


def _basis(th, ph):
    """unit vectors e_r, e_theta, e_phi in Cartesian components"""
    st, ct, sp, cp = math.sin(th), math.cos(th), math.sin(ph), math.cos(ph)
    return np.array([[st * cp, st * sp, ct], [ct * cp, ct * sp, -st], [-sp, cp, 0.0]])


def _polar_case(sim, n, nj, nk, th_row, vcart, dx2):
    """set a uniform Cartesian velocity (and a phi-varying density) and check
    every phi face of the last theta row of each shell"""
    geo = sim.geometry()
    N0, N1 = n[0] + 4, n[1] + 4

    def rec(i, j, k):
        return geo[(i + 2) + N0 * ((j + 2) + N1 * (k + 2))]

    prims = np.zeros((5, n[2], n[1], n[0]))
    for i in range(n[0]):
        for j in range(nj[i]):
            dk = n[2] // nk[i, j]
            th = th_row(i, j)
            for k in range(nk[i, j]):
                ph = (k + 0.5) * dk * dx2
                e = _basis(th, ph)
                sc = rec(i, j, k)[po.GEOM_SCALE:po.GEOM_SCALE + 3]        # zone-centred scale factors
                prims[0, k, j, i] = 2.0 + 0.3 * math.sin(ph) + 0.01 * i
                prims[1, k, j, i] = 3.0
                for m in range(3):
                    prims[2 + m, k, j, i] = (e[m] @ vcart) / sc[m]
    sim.set_prims(prims)
    sim.prepare_entries()
    checked = 0
    for i in range(n[0]):
        j = nj[i] - 1
        dk = n[2] // nk[i, j]
        th = th_row(i, j)
        for k in range(nk[i, j] + 1):
            left, right = sim.polar_face_states(i, j, k)
            ef = _basis(th, k * dk * dx2)
            scf = rec(i, j, k)[po.GEOM_SCALE + 9:po.GEOM_SCALE + 12]     # phi-face scale factors
            want_v = np.array([(ef[l] @ vcart) / scf[l] for l in range(3)])
            for st in (left, right):
                assert np.allclose(st[2:5], want_v, rtol=1e-12, atol=1e-12 * np.max(np.abs(want_v))), \
                    (i, k, st[2:5], want_v)
            kl, kr = (k - 1) % nk[i, j], k % nk[i, j]
            assert left[0] == prims[0, kl, j, i] and right[0] == prims[0, kr, j, i]    # first order
            checked += 1
    return checked


@pytest.mark.parametrize("vcart", [(1.0, 0.0, 0.0), (0.3, -0.7, 0.5)])
def test_polar_phi_states_project_uniform_cartesian_velocity_dendritic(vcart):
    n = (20, 32, 64)
    G = Grid(n)

    def th_row(i, j):
        return th_of_x(G.x1c(i, j))
    checked = _polar_case(G.sim, n, G.nj, G.nk, th_row, np.array(vcart), G.dx2)
    assert checked >= n[0] * 5


@pytest.mark.parametrize("vcart", [(0.0, 1.0, 0.0), (0.3, -0.7, 0.5)])
def test_polar_phi_states_project_uniform_cartesian_velocity_uniform_grid(vcart):
    n = (6, 8, 8)
    cfg = sph3d_config(n, nvars=5)
    sim = po.Sim(cfg)
    nj = np.full(n[0], n[1])
    nk = np.full((n[0], n[1]), n[2])
    a, b, g = sim.moments(1)

    def th_row(i, j):
        return th_of_x(b[j + 2] / g[j + 2])
    checked = _polar_case(sim, n, nj, nk, th_row, np.array(vcart), cfg.dx[2])
    assert checked == n[0] * (n[2] + 1)


def test_polar_check_fails_for_a_transposed_rotation():
    """Guard of the polar check: the same projection with the zone basis
    transposed does not match, so the test above can tell the two apart."""
    th, ph, phf = 0.1, 0.7, 0.9
    v = np.array([0.3, -0.7, 0.5])
    e, ef = _basis(th, ph), _basis(th, phf)
    comp = e @ v
    good = ef @ (e.T @ comp)
    bad = ef @ (e @ comp)
    assert np.allclose(good, ef @ v) and not np.allclose(bad, ef @ v)
