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
