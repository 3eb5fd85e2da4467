"""Tables in the reference's Burrows binary layout (collapse_init,
eos/burrows/eos_stuff.c:76-133) and the walk's unit handling
(eos_stuff.c:523-650): the specific energy in erg/g is looked up as
e / 0.95655684e18 - const_hack MeV per baryon (:530), T comes out in MeV
(:531-532, :645) and P = exp(ln P) * 1.60217733e33 dyn/cm^2 (:618).

The test writes a small file in that layout (16 records of ny*nr*nt doubles,
T fastest, record 2 a pressure that the loader log-transforms), from an
analytic model in MeV units.  No real table is available offline.

Pins (what the layout and the unit handling fix):
  * the loader keeps records 1, 2 and 12 at their byte offsets, as FP32, with
    ln of record 2 (distinct filler values in the other records make a wrong
    offset visible);
  * a state whose energy equals a node energy, at a (rho, Ye) node, returns
    that node's temperature in MeV and P = P_node[MeV/fm^3] * 1.60217733e33;
  * half way between two node energies along T, log10 T is half way between
    the nodes (the walk's inverse is linear in log T);
  * a short or missing file is MDS_EIO.
GPU: the device EOS on the loaded table against the oracle's walk."""
import math
import os

import numpy as np
import pytest

from medusa_amd import mds
from oracle import pyoracle as po

DIMS = dict(ny=6, nr=9, nt=12)
E_SCALE, P_SCALE = 0.95655684e18, 1.60217733e33


def _axes(t):
    lr = t.lr0 + np.arange(t.nr) * ((t.lr1 - t.lr0) / (t.nr - 1))
    lt = t.lt0 + np.arange(t.nt) * ((t.lt1 - t.lt0) / (t.nt - 1))
    ye = t.ye0 + np.arange(t.ny) * ((t.ye1 - t.ye0) / (t.ny - 1))
    return lr, lt, ye


def _model(lr, lt, ye):
    """MeV units: energy per baryon strictly increasing in T, P > 0, Gamma_1 near 4/3."""
    R, T, Y = np.meshgrid(lr, lt, ye, indexing="ij")        # [nr][nt][ny]
    Tm = 10.0 ** T
    rho = 10.0 ** R
    e = -8.0 + 0.3 * (R - 8.0) + 2.0 * Y + 1.5 * Tm + 0.05 * Tm ** 2
    p = 1.0e-12 * rho ** 1.2 * (1.0 + Y) + 3.0e-14 * rho * Tm
    g = 4.0 / 3.0 + 0.1 * np.tanh(R - 10.0) + 0.02 * Y
    to = lambda a: np.ascontiguousarray(np.transpose(a, (2, 0, 1)))   # -> [ny][nr][nt]
    return to(e), to(p), to(g)


def _write(path, t, e, p, g, truncate=0):
    rec = t.ny * t.nr * t.nt
    recs = [np.full((t.ny, t.nr, t.nt), 1000.0 + k, dtype=np.float64) for k in range(16)]
    recs[0], recs[1], recs[11] = e, p, g
    blob = b"".join(np.ascontiguousarray(r, dtype="<f8").tobytes() for r in recs)
    with open(path, "wb") as f:
        f.write(blob[:len(blob) - truncate] if truncate else blob)
    assert len(blob) == 16 * 8 * rec


@pytest.fixture(scope="module")
def table(tmp_path_factory):
    t = mds.EosTableDesc()
    mds.lib().mds_eos_burrows_default(t)
    for k, v in DIMS.items():
        setattr(t, k, v)
    lr, lt, ye = _axes(t)
    e, p, g = _model(lr, lt, ye)
    path = str(tmp_path_factory.mktemp("eos") / "burrows_small.dat")
    _write(path, t, e, p, g)
    T = mds.EosTable(burrows_path=path, **DIMS)
    return dict(T=T, e=e, p=p, g=g, lr=lr, lt=lt, ye=ye, path=path)


def _otab(T):
    t = T.t
    return po.eos_tab(dict(nr=t.nr, nt=t.nt, ny=t.ny, lr0=t.lr0, lr1=t.lr1, lt0=t.lt0, lt1=t.lt1, ye0=t.ye0,
                           ye1=t.ye1, e_scale=t.e_scale, e_shift=t.e_shift, p_scale=t.p_scale), T.data)


def test_burrows_defaults_are_the_reference_constants():
    t = mds.EosTableDesc()
    mds.lib().mds_eos_burrows_default(t)
    assert (t.nr, t.nt, t.ny) == (300, 300, 50)
    assert (t.lr0, t.lr1, t.lt0, t.lt1, t.ye0, t.ye1) == (1.0, 15.5, -2.0, 2.0, 0.035, 0.56)
    assert (t.e_scale, t.e_shift, t.p_scale) == (E_SCALE, 0.0, P_SCALE)


def test_loader_keeps_records_1_2_12(table):
    T = table["T"]
    assert np.array_equal(T.data[0], table["e"].astype(np.float32))
    assert np.array_equal(T.data[1], np.log(table["p"]).astype(np.float32))
    assert np.array_equal(T.data[2], table["g"].astype(np.float32))


def test_loader_rejects_short_and_missing_files(table, tmp_path):
    t = mds.EosTableDesc()
    mds.lib().mds_eos_burrows_default(t)
    for k, v in DIMS.items():
        setattr(t, k, v)
    short = str(tmp_path / "short.dat")
    _write(short, t, table["e"], table["p"], table["g"], truncate=8)
    assert mds.lib().mds_eos_table_load_burrows(os.fsencode(short), t) == mds.EIO
    assert mds.lib().mds_eos_table_load_burrows(os.fsencode(str(tmp_path / "none.dat")), t) == mds.EIO


def test_walk_at_a_node_returns_node_temperature_and_scaled_pressure(table):
    T, ot = table["T"], _otab(table["T"])
    checked = 0
    for jy in (1, 4):
        for jr in (2, 6):
            for jt in (3, 7, 10):
                rho = 10.0 ** table["lr"][jr]
                ye = table["ye"][jy]
                e_mev = float(T.data[0, jy, jr, jt])              # the FP32 node energy the walk sees
                u = rho * (e_mev + T.t.e_shift) * E_SCALE
                out, st = po.eos_walk(ot, rho, u, ye)
                assert st == 0
                assert abs(math.log10(out[2]) - table["lt"][jt]) < 1e-9, (jy, jr, jt)
                want_p = table["p"][jy, jr, jt] * P_SCALE
                assert abs(out[0] / want_p - 1.0) < 1e-6 * (1 + abs(math.log(table["p"][jy, jr, jt])))
                assert abs(out[3] - table["g"][jy, jr, jt]) < 1e-6
                assert abs(out[1] - math.sqrt(out[3] * out[0] / rho)) <= 1e-12 * out[1]
                checked += 1
    assert checked == 12


def test_walk_is_linear_in_log_t_between_nodes(table):
    T, ot = table["T"], _otab(table["T"])
    for jy, jr, jt in ((2, 3, 4), (5, 8, 0), (0, 0, 9)):
        rho = 10.0 ** table["lr"][jr]
        e0, e1 = float(T.data[0, jy, jr, jt]), float(T.data[0, jy, jr, jt + 1])
        u = rho * 0.5 * (e0 + e1) * E_SCALE
        out, st = po.eos_walk(ot, rho, u, table["ye"][jy])
        want = 0.5 * (table["lt"][jt] + table["lt"][jt + 1])
        assert abs(math.log10(out[2]) - want) < 1e-9


def _states(table, n=400, seed=3):
    rng = np.random.default_rng(seed)
    T = table["T"]
    lr = rng.uniform(T.t.lr0 + 0.01, T.t.lr1 - 0.01, n)
    ye = rng.uniform(T.t.ye0, T.t.ye1, n)
    tm = 10.0 ** rng.uniform(T.t.lt0 + 0.01, T.t.lt1 - 0.01, n)
    rho = 10.0 ** lr
    e_mev = -8.0 + 0.3 * (lr - 8.0) + 2.0 * ye + 1.5 * tm + 0.05 * tm ** 2
    return rho, rho * e_mev * E_SCALE, ye


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [mds.FP64, mds.FP32_BULK])
def test_device_eos_on_burrows_table_matches_oracle(table, precision):
    T, ot = table["T"], _otab(table["T"])
    rho, u, ye = _states(table)
    P, cs, temp, g1 = mds.eos_points(precision, T, rho, u, ye)
    ref = np.array([po.eos_walk(ot, a, b, c)[0] for a, b, c in zip(rho, u, ye)])
    tol = 1e-13 if precision == mds.FP64 else 2e-6
    for name, dev, k in (("P", P, 0), ("cs", cs, 1), ("T", temp, 2), ("Gamma1", g1, 3)):
        err = np.max(np.abs(dev / ref[:, k] - 1.0))
        print(f"[burrows] precision={precision} {name}: max rel err {err:.2e}", flush=True)
        assert err <= tol, (name, err)
