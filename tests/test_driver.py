"""The C host program (medusa_amd/lib/mds_driver) in the shape of the
reference's main loop, on the C-ABI only: its ASCII dump (the reference's
1D layout, io.c:109-230) against the CPU oracle."""
import os
import subprocess

import numpy as np
import pytest

from oracle import pyoracle as po
from medusa_amd import mds, problems
from parity_util import oracle_for, oracle_config, l1rel

DRIVER = os.path.join(os.path.dirname(mds.LIB_PATH), "mds_driver")


def test_driver_is_built_and_checks_arguments():
    assert os.path.exists(DRIVER)
    r = subprocess.run([DRIVER], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr


def _read_dump(path, nvars):
    with open(path) as f:
        head = f.readline()
    assert head.startswith("# ")
    d = np.loadtxt(path, comments="#")
    # x, prims (nvars), eos (5), enclosed mass, two lapses
    assert d.shape[1] == 1 + nvars + 5 + 1 + 2
    return d


@pytest.mark.gpu
def test_driver_sod_c1_matches_oracle(tmp_path):
    out = str(tmp_path / "dump_sod")
    print("[driver] sod 1024 to t=0.2", flush=True)
    r = subprocess.run([DRIVER, "sod", "1024", "0.2", out], capture_output=True, text=True, timeout=120)
    print(r.stderr, flush=True)
    assert r.returncode == 0
    d = _read_dump(out, 5)
    g, prob, prims = problems.sod(1024)
    sim = oracle_for(prob)
    sim.set_prims(prims)
    sim.run(tmax=0.2)
    ref = sim.get_prims()[:, 0, 0]
    assert f"steps={sim.istep} " in r.stderr
    for k in range(3):
        # the dump prints 15 significant digits
        assert l1rel(d[:, 1 + k], ref[k]) <= 1e-14
    np.testing.assert_allclose(d[:, 0], 0.5 / 1024 + np.arange(1024) / 1024, rtol=1e-14)
    np.testing.assert_allclose(d[:, 6], 0.4 * ref[1], rtol=1e-14)           # P = (gam-1) u


@pytest.mark.gpu
def test_driver_star_column_matches_oracle(tmp_path):
    out = str(tmp_path / "dump_star")
    print("[driver] star 256, 40 steps", flush=True)
    r = subprocess.run([DRIVER, "star", "256", "40", out], capture_output=True, text=True, timeout=120)
    print(r.stderr, flush=True)
    assert r.returncode == 0
    d = _read_dump(out, 6)
    T = mds.EosTable()
    g, prob, prims = problems.star1d(256, table=T, sigma=0.01)
    t = T.t
    ot = po.eos_tab(dict(nr=t.nr, nt=t.nt, ny=t.ny, lr0=t.lr0, lr1=t.lr1, lt0=t.lt0, lt1=t.lt1, ye0=t.ye0,
                         ye1=t.ye1), T.data)
    sim = po.Sim(oracle_config(prob, 0.5, nvars=6, monopole=True, eos_tab=ot))
    sim.set_prims(prims)
    sim.run(max_steps=40)
    ref = sim.get_prims()[:, 0, 0]
    w = g.zone()[2:-2, 0]
    for k in (0, 1, 5):
        assert l1rel(d[:, 1 + k], ref[k], w) <= 1e-12, k
    # enclosed mass at the outer zone centre ~ total mass of the grid
    assert abs(d[-1, 12] / np.sum(w * ref[0]) - 1) < 1e-2


def test_driver_rejects_unknown_parameters():
    r = subprocess.run([DRIVER, "star", "64", "1", "/dev/null", "nonsense=1"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 2 and "unknown parameter" in r.stderr


@pytest.mark.gpu
def test_driver_gr_monopole_dump_prints_the_lapses(tmp_path):
    """gr_monopole=1: the dump's last two columns are the reference's
    geom lapse[0] (zone centre) and lapse[1] (lower face), io.c:198-202,
    as the oracle's GR monopole leaves them after the same steps."""
    out = str(tmp_path / "dump_gr")
    print("[driver] star 256 GR monopole, 30 steps", flush=True)
    r = subprocess.run([DRIVER, "star", "256", "30", out, "gr_monopole=1"], capture_output=True, text=True,
                       timeout=120)
    print(r.stderr, flush=True)
    assert r.returncode == 0
    d = _read_dump(out, 6)
    T = mds.EosTable()
    g, prob, prims = problems.star1d(256, table=T, sigma=0.01)
    t = T.t
    ot = po.eos_tab(dict(nr=t.nr, nt=t.nt, ny=t.ny, lr0=t.lr0, lr1=t.lr1, lt0=t.lt0, lt1=t.lt1, ye0=t.ye0,
                         ye1=t.ye1), T.data)
    sim = po.Sim(oracle_config(prob, 0.5, nvars=6, monopole=2, eos_tab=ot))
    sim.set_prims(prims)
    sim.run(max_steps=30)
    st = sim.gr_state()
    assert np.all(d[:, 13] < 1.0) and np.all(np.diff(d[:, 13]) > 0)       # deepest potential at the centre
    np.testing.assert_allclose(d[:, 13], st["lapse"], rtol=1e-12)
    np.testing.assert_allclose(d[:, 14], st["lapse_edge"][:256], rtol=1e-12)
    ref = sim.get_prims()[:, 0, 0]
    assert l1rel(d[:, 1], ref[0], g.zone()[2:-2, 0]) <= 1e-12


@pytest.mark.gpu
def test_driver_1d_restart_round_trip(tmp_path):
    """restart_create_1d (main.c:268-274): a star set expanding (vexp) passes
    the bounce test (central density below 0.95 of its maximum), runs on
    restart_delay, writes restarts/<model>.<n1>.0g.<ms>ms.restart_1d in the
    dump layout and exits with status 1.  restart_from_1d (init_ccsn.c:157-264)
    reads it back onto the same grid: rho, v1 and Ye as written, u from the
    table at the written temperature, so P, T and the enclosed mass of a
    dump taken right after the read match the restart file."""
    print("[driver] star 128 expanding, bounce detection and 1D restart", flush=True)
    r = subprocess.run([DRIVER, "star", "128", "5000", "dump_a", "vexp=300", "restart_create_1d=1",
                        "restart_delay=5e-5", "model_name=models/s15"], capture_output=True, text=True,
                       timeout=120, cwd=str(tmp_path))
    print(r.stderr[-2000:], flush=True)
    assert r.returncode == 1
    assert "Bounce detected" in r.stderr
    path = tmp_path / "restarts" / "s15.128.0g.10ms.restart_1d"
    assert path.exists()
    a = _read_dump(str(path), 6)
    assert a.shape[0] == 128
    r = subprocess.run([DRIVER, "star", "128", "0", "dump_b", f"restart_from_1d={path}"], capture_output=True,
                       text=True, timeout=120, cwd=str(tmp_path))
    print(r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    b = _read_dump(str(tmp_path / "dump_b"), 6)
    np.testing.assert_allclose(b[:, 0], a[:, 0], rtol=1e-14)                 # radii
    for c in (1, 3, 6):                                                       # rho, v1, Ye
        np.testing.assert_allclose(b[:, c], a[:, c], rtol=1e-13, atol=1e-13 * np.max(np.abs(a[:, c])))
    for c in (2, 7, 9, 12):                                                   # u, P, T, mass
        np.testing.assert_allclose(b[:, c], a[:, c], rtol=1e-9)
    assert np.max(np.abs(a[:, 3])) > 0.0                                      # the star is moving
    # a file of another length is refused, as the reference refuses it
    r = subprocess.run([DRIVER, "star", "100", "0", "dump_c", f"restart_from_1d={path}"], capture_output=True,
                       text=True, timeout=120, cwd=str(tmp_path))
    assert r.returncode == 1 and "but require 100" in r.stderr
