"""Tabulated EOS: the synthetic table, the CPU oracle of the reference's
table use (eos_stuff.c:507-650 walk, eos_stellar_collapse.c:294-369
Newton/bisection) and, on the GPU, the device fill against the oracle.

Pins: the table against an independent numpy evaluation of the analytic
model with analytic derivatives; strict monotonicity of e in T; the walk's
independence from its starting temperature; e(T(e)) = e on the interpolant;
Newton/bisection landing on the walk's root; pressure at nodes against the
closed form."""
import numpy as np
import pytest

from oracle import pyoracle as po
from medusa_amd import mds

KB, MU, ARAD = 1.380649e-16, 1.66053906660e-24, 7.565723e-15


@pytest.fixture(scope="module")
def table():
    return mds.EosTable()


def _otab(T):
    t = T.t
    dims = dict(nr=t.nr, nt=t.nt, ny=t.ny, lr0=t.lr0, lr1=t.lr1, lt0=t.lt0, lt1=t.lt1, ye0=t.ye0, ye1=t.ye1)
    return po.eos_tab(dims, T.data)


def _model_numpy(prm, rho, T, ye):
    """Independent evaluation of the model with analytic derivatives: (P, e, Gamma_1)."""
    A, K53, rhox = prm
    rye = rho * ye
    x = np.cbrt(rye / rhox)
    Pd = K53 * rye ** (5.0 / 3.0) / (1 + x)
    gd = 5.0 / 3.0 - (x / (1 + x)) / 3.0
    ed = Pd / (rho * (gd - 1))
    TF = 3.0e9 * (rye / 1.0e6) ** (2.0 / 3.0) / (1 + x)
    Pe = rye / MU * KB * T * T / (T + TF)
    Pi = rho * KB * T / (A * MU)
    Pr = ARAD * T ** 4 / 3
    P = Pd + Pe + Pi + Pr
    e = ed + 1.5 * Pe / rho + 1.5 * Pi / rho + ARAD * T ** 4 / rho
    dlnx = (x / 3) / (1 + x)
    dPd = Pd * (5.0 / 3.0 - dlnx)                                 # rho dP/drho of each part
    dPe = Pe * (1 - TF / (T + TF) * (2.0 / 3.0 - dlnx))
    chi_rho = (dPd + dPe + Pi) / P
    chi_T = (Pe * (1 + TF / (T + TF)) + Pi + 4 * Pr) / P
    cv = 1.5 * Pe / (rho * T) * (1 + TF / (T + TF)) + 1.5 * Pi / (rho * T) + 4 * ARAD * T ** 3 / rho
    g1 = chi_rho + P * chi_T ** 2 / (rho * T * cv)
    return P, e, g1


def test_table_nodes_match_independent_model(table):
    prm = mds.eos_model_params()
    lr, lt, ye = table.nodes()
    rng = np.random.default_rng(1)
    for _ in range(300):
        jy, jr, jt = rng.integers(table.t.ny), rng.integers(table.t.nr), rng.integers(table.t.nt)
        P, e, g1 = _model_numpy(prm, 10.0 ** lr[jr], 10.0 ** lt[jt], ye[jy])
        assert abs(table.data[0, jy, jr, jt] / e - 1) < 2e-7
        assert abs(table.data[1, jy, jr, jt] - np.log(P)) < 1e-5
        assert abs(table.data[2, jy, jr, jt] / g1 - 1) < 1e-6


def test_model_limits():
    """Radiation-dominated corner -> Gamma_1 = 4/3; jitter within its stated bounds."""
    prm = mds.eos_model_params()
    assert 56 * 0.9 <= prm[0] <= 56 * 1.1 and 1.0036e13 * 0.95 <= prm[1] <= 1.0036e13 * 1.05
    P, e, g1 = mds.eos_model_eval(prm, 1.0e3, 1.0e12, 0.5)
    assert abs(g1 - 4.0 / 3.0) < 1e-6
    assert abs(P / (ARAD * 1e48 / 3) - 1) < 1e-6


def test_energy_strictly_increasing_in_temperature(table):
    assert np.all(np.diff(table.data[0].astype(np.float64), axis=2) > 0)


def _random_states(table, n, seed, margin=0.0):
    prm = mds.eos_model_params()
    rng = np.random.default_rng(seed)
    t = table.t
    lr = rng.uniform(t.lr0 + margin, t.lr1 - margin, n)
    lt = rng.uniform(t.lt0 + 0.05, t.lt1 - 0.05, n)
    ye = rng.uniform(t.ye0 + margin, t.ye1 - margin, n)
    rho = 10.0 ** lr
    e = np.array([mds.eos_model_eval(prm, r, 10.0 ** a, y)[1] for r, a, y in zip(rho, lt, ye)])
    return rho, rho * e, ye, lt


def test_walk_is_independent_of_start(table):
    ot = _otab(table)
    rho, u, ye, lt = _random_states(table, 200, 2)
    for k in range(rho.size):
        outs = [po.eos_walk(ot, rho[k], u[k], ye[k], tg)[0] for tg in (0.0, 1.0e9, 3.0e10, 1.0e12)]
        for o in outs[1:]:
            assert np.array_equal(o, outs[0])


def _interp_e(table, rho, ye, lt):
    """Piecewise-linear e(log T) at fixed (rho, Ye): bilinear projection of the
    energy record onto every T node, then linear in log T (numpy)."""
    t = table.t
    dr, dt, dy = (t.lr1 - t.lr0) / (t.nr - 1), (t.lt1 - t.lt0) / (t.nt - 1), (t.ye1 - t.ye0) / (t.ny - 1)
    rl = np.log10(rho)
    jr, jy = int((rl - t.lr0) / dr), int((ye - t.ye0) / dy)
    fr, fy = (rl - (t.lr0 + jr * dr)) / dr, (ye - (t.ye0 + jy * dy)) / dy
    E = table.data[0].astype(np.float64)
    col = ((1 - fr) * (1 - fy) * E[jy, jr] + fy * (1 - fr) * E[jy + 1, jr] + (1 - fy) * fr * E[jy, jr + 1]
           + fy * fr * E[jy + 1, jr + 1])
    s = (lt - t.lt0) / dt
    j = min(int(s), t.nt - 2)
    return (1 - (s - j)) * col[j] + (s - j) * col[j + 1]


def test_round_trip_energy(table):
    ot = _otab(table)
    rho, u, ye, lt_true = _random_states(table, 300, 3, margin=0.01)
    on = 0
    for k in range(rho.size):
        out, st = po.eos_walk(ot, rho[k], u[k], ye[k], 0.0)
        if st:
            # the model energy lies below the interpolated table at its lowest
            # temperature (strongly degenerate states): off the table
            continue
        on += 1
        lt = np.log10(out[2])
        assert abs(_interp_e(table, rho[k], ye[k], lt) / (u[k] / rho[k]) - 1) < 1e-9
        # against the model's own temperature where e depends on T enough for
        # the table's interpolation error in e not to swamp it (degenerate
        # states have a nearly T-independent e)
        prm = mds.eos_model_params()
        e1 = mds.eos_model_eval(prm, rho[k], 10.0 ** (lt_true[k] + 1e-3), ye[k])[1]
        sens = np.log(e1 / (u[k] / rho[k])) / (1e-3 * np.log(10.0))
        if sens > 0.5:
            assert abs(lt - lt_true[k]) < 2e-3, (k, sens)
    assert on >= 0.8 * rho.size


def test_newton_bisection_finds_the_walk_root(table):
    ot = _otab(table)
    rho, u, ye, lt_true = _random_states(table, 200, 4)
    rng = np.random.default_rng(5)
    for k in range(rho.size):
        out, st = po.eos_walk(ot, rho[k], u[k], ye[k], 0.0)
        lw = np.log10(out[2])
        guess = lw + rng.uniform(-0.3, 0.3)
        lr, status, iters = po.eos_root(ot, rho[k], u[k], ye[k], guess)
        assert status == 1
        assert abs(lr - lw) < 1e-8, (k, lr, lw, iters)


def test_pressure_at_nodes_equals_closed_form(table):
    ot = _otab(table)
    prm = mds.eos_model_params()
    lr, lt, ye = table.nodes()
    rng = np.random.default_rng(6)
    for _ in range(100):
        jy, jr, jt = rng.integers(table.t.ny - 1), rng.integers(table.t.nr - 1), rng.integers(1, table.t.nt - 1)
        rho = 10.0 ** lr[jr]
        e = float(table.data[0, jy, jr, jt])
        out, st = po.eos_walk(ot, rho, rho * e, ye[jy], 0.0)
        P, _, g1 = _model_numpy(prm, rho, 10.0 ** lt[jt], ye[jy])
        assert abs(out[0] / P - 1) < 1e-5
        assert abs(np.log10(out[2]) - lt[jt]) < 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [mds.FP64, mds.FP32_BULK])
def test_device_eos_matches_oracle(table, precision):
    ot = _otab(table)
    rho, u, ye, _ = _random_states(table, 3000, 7)
    # off-table states: density and Ye beyond the edges, energies below and above the range
    rng = np.random.default_rng(8)
    rho = np.concatenate([rho, 10.0 ** rng.uniform(1.0, 3.0, 20), 10.0 ** rng.uniform(15.0, 16.0, 20)])
    ye = np.concatenate([ye, rng.uniform(0.01, 0.05, 20), rng.uniform(0.55, 0.6, 20)])
    u = np.concatenate([u, rho[-40:] * 10.0 ** rng.uniform(17.0, 19.0, 40)])
    lo = rng.integers(0, 3000, 20)
    rho, ye = np.concatenate([rho, rho[lo], rho[lo]]), np.concatenate([ye, ye[lo], ye[lo]])
    u = np.concatenate([u, u[lo] * 1e-3, u[lo] * 1e6])
    P, cs, T, g1 = mds.eos_points(precision, table, rho, u, ye)
    walk = [po.eos_walk(ot, r, uu, y, 0.0) for r, uu, y in zip(rho, u, ye)]
    ref = np.array([w[0] for w in walk])
    on_table = np.array([w[1] == 0 for w in walk])
    # far above the table the linear extrapolation in log T overflows exp():
    # both sides must then be non-finite in the same places
    # (FP64), and far below it the extrapolated temperature underflows to 0
    dev = np.stack([P, cs, T, g1], 1)
    if precision == mds.FP64:
        assert np.array_equal(np.isfinite(dev), np.isfinite(ref))
    ok = np.all(np.isfinite(ref), axis=1) & np.all(np.isfinite(dev), axis=1)
    if precision != mds.FP64:
        # the FP32 increments are meant for interpolation; off the table (linear
        # extrapolation in log T far beyond the last node) only FP64 is compared
        ok &= on_table
    assert ok.sum() >= 2800
    dev, ref = dev[ok], ref[ok]
    zero = ref == 0.0
    assert np.array_equal(dev[zero], ref[zero])
    rel = np.abs(np.where(zero, 0.0, dev / np.where(zero, 1.0, ref) - 1))
    tol = 1e-13 if precision == mds.FP64 else 2e-6
    print(f"[eos] precision={precision} compared {ok.sum()}/{ok.size}: max rel err P {rel[:, 0].max():.2e} "
          f"cs {rel[:, 1].max():.2e} T {rel[:, 2].max():.2e} Gamma1 {rel[:, 3].max():.2e}", flush=True)
    for k in range(4):
        assert rel[:, k].max() <= (1e-13 if k == 2 else tol), k


# ---------------------------------------------------------------------------
# The table at a given temperature (mds_eos_rtx, the restart reader's
# u_given_rtx): exact at the nodes, linear in log T between them, and the
# walk's inverse.
# ---------------------------------------------------------------------------
def test_rtx_at_nodes_is_the_node_record(table):
    lr, lt, ye = table.nodes()
    rng = np.random.default_rng(5)
    d = table.data
    for _ in range(200):
        jy, jr, jt = rng.integers(table.t.ny - 1), rng.integers(table.t.nr - 1), rng.integers(table.t.nt - 1)
        rho, T = 10.0 ** lr[jr], 10.0 ** lt[jt]
        u, P, g1, cs = mds.eos_rtx(table, rho, T, ye[jy])
        # log10(10^x) and the node fractions carry a few ulps: the node values to 1e-12
        np.testing.assert_allclose(u, rho * float(d[0, jy, jr, jt]), rtol=1e-12)
        np.testing.assert_allclose(P, np.exp(float(d[1, jy, jr, jt])), rtol=1e-12)
        np.testing.assert_allclose(g1, float(d[2, jy, jr, jt]), rtol=1e-12)
        np.testing.assert_allclose(cs, np.sqrt(g1 * P / rho), rtol=1e-15)


def test_rtx_is_linear_in_log_temperature_between_nodes(table):
    lr, lt, ye = table.nodes()
    rng = np.random.default_rng(6)
    d = table.data.astype(np.float64)
    for _ in range(200):
        jy, jr, jt = rng.integers(table.t.ny - 1), rng.integers(table.t.nr - 1), rng.integers(table.t.nt - 1)
        f = rng.uniform(0.05, 0.95)
        rho, T = 10.0 ** lr[jr], 10.0 ** (lt[jt] + f * (lt[jt + 1] - lt[jt]))
        u = mds.eos_rtx(table, rho, T, ye[jy])[0]
        want = rho * ((1 - f) * d[0, jy, jr, jt] + f * d[0, jy, jr, jt + 1])
        np.testing.assert_allclose(u, want, rtol=1e-10)


def test_rtx_inverts_the_walk(table):
    """u(rho, T, Ye) from the forward table, then the oracle's walk from u:
    the walk lands on T (both use the same piecewise-linear e(log T))."""
    ot = _otab(table)
    lr, lt, ye = table.nodes()
    rng = np.random.default_rng(7)
    for _ in range(200):
        rl = rng.uniform(lr[1], lr[-2])
        tl = rng.uniform(lt[1], lt[-2])
        y = rng.uniform(ye[1], ye[-2])
        u = mds.eos_rtx(table, 10.0 ** rl, 10.0 ** tl, y)[0]
        out, st = po.eos_walk(ot, 10.0 ** rl, u, y, 0.0)
        assert st == 0
        assert abs(np.log10(out[2]) - tl) < 1e-9
