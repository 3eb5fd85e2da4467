"""Device health counters (mds_get_health): unphysical zone updates and table
EOS lookups off the table, counted by the stage kernels of every family.

The reference has no error channel on this path: it prints off-table EOS
warnings and continues (eos/burrows/eos_stuff.c:571-604) and exits when dt
falls below dtmin (main.c:229-238).  The cases below plant finite bad data
from the host (a negative internal energy, a density below the table); no
kernel faults, the counts come back as ordinary data."""
import numpy as np
import pytest

from medusa_amd import mds, problems

pytestmark = pytest.mark.gpu


def _run(prob, prims, precision, steps, graph=False):
    ctx = mds.Context([prob], precision=precision, rp=mds.run_params(cfl=0.5))
    ctx.upload(prims)
    ctx.update_eos()
    ctx.run(steps, graph=graph)
    return ctx


@pytest.mark.parametrize("precision,fuse", [(mds.FP64, 0), (mds.FP64, 1), (mds.FP32_BULK, 0), (mds.FP32_BULK, 1),
                                            (mds.FP32_BULK, 2)])
def test_sod_negative_internal_energy_is_counted(precision, fuse):
    g, prob, prims = problems.sod(256)
    for bad in (False, True):
        p = prims.copy()
        if bad:
            p[1, 0, 0, 100] = -0.5          # finite negative internal energy (negative pressure)
        ctx = mds.Context([prob], precision=precision, rp=mds.run_params(cfl=0.5))
        ctx.set_tile(mds.TILE_FUSE1D, fuse)
        ctx.upload(p)
        ctx.update_eos()
        ctx.run(2)
        h = ctx.health()
        print(f"[health] sod precision={precision} two-stage={fuse} planted={bad}: {h}", flush=True)
        if bad:
            assert h["bad_state"] >= 1
        else:
            assert h["bad_state"] == 0 and h["eos_off_table"] == 0 and h["dt_below_min"] == 0


def test_clock_of_health_matches_get_clock():
    g, prob, prims = problems.sod(128)
    ctx = _run(prob, prims, mds.FP64, 7)
    h = ctx.health()
    t, dt, ist = ctx.clock()
    assert (h["t"], h["dt"], h["istep"]) == (t, dt, ist)


@pytest.mark.parametrize("precision", [mds.FP64, mds.FP32_BULK])
def test_star_column_density_below_table_is_counted(precision):
    T = mds.EosTable()
    g, prob, prims = problems.star1d(256, table=T, sigma=0.01)
    counts = []
    for bad in (False, True):
        p = prims.copy()
        if bad:
            p[0, 0, 0, 40] = 10.0 ** (T.t.lr0 - 1.0)     # a decade below the table's density range
        ctx = _run(prob, p, precision, 1)
        h = ctx.health()
        print(f"[health] star column precision={precision} planted={bad}: {h}", flush=True)
        counts.append(h["eos_off_table"])
    # the planted zone is looked up in both stages of the step
    assert counts[1] >= counts[0] + 2


@pytest.mark.parametrize("precision", [mds.FP64, mds.FP32_BULK])
def test_kh_and_rt_negative_internal_energy_is_counted(precision):
    g, prob, prims = problems.kh(64)
    p = prims.copy()
    ctx = _run(prob, p, precision, 1)
    assert ctx.health()["bad_state"] == 0
    p[1, 0, 20, 30] = -1.0
    ctx = _run(prob, p, precision, 1)
    h = ctx.health()
    print(f"[health] kh64 precision={precision}: {h}", flush=True)
    assert h["bad_state"] >= 1
    g, prob, prims = problems.rt(16)
    p = prims.copy()
    p[1, 8, 7, 6] = -1.0
    ctx = _run(prob, p, precision, 1)
    h = ctx.health()
    print(f"[health] rt16 precision={precision}: {h}", flush=True)
    assert h["bad_state"] >= 1


@pytest.mark.parametrize("precision", [mds.FP64, mds.FP32_BULK])
def test_sph3d_star_planted_zones_are_counted(precision):
    T = mds.EosTable()
    n = (32, 8, 8)
    g, prob, prims = problems.star3d(*n, table=T)
    ctx = _run(prob, prims, precision, 1)
    h0 = ctx.health()
    p = prims.copy()
    p[0, 3, 4, 20] = 10.0 ** (T.t.lr0 - 1.0)
    ctx = _run(prob, p, precision, 1)
    h1 = ctx.health()
    print(f"[health] sph3d star precision={precision}: clean {h0}, planted {h1}", flush=True)
    # the first stage looks the planted zone up off the table; the inflow from
    # its neighbours (a million times denser) can lift it into the table
    # before the second stage's lookup
    assert h1["eos_off_table"] >= h0["eos_off_table"] + 1
    # gamma law: a negative internal energy
    g, prob, prims = problems.star3d(*n, table=None)
    p = prims.copy()
    p[1, 3, 4, 20] = -abs(p[1, 3, 4, 20])
    ctx = _run(prob, p, precision, 1)
    h = ctx.health()
    print(f"[health] sph3d gamma-law star precision={precision}: {h}", flush=True)
    assert h["bad_state"] >= 1


def test_batched_columns_count_per_problem():
    """Counts belong to the problem whose zone went bad."""
    cols = [problems.sph_riemann(n, 1.4, 1.0e9, 1.0e27, 1.0e7, 10.0, 10.0) for n in (300, 500, 400)]
    ctx = mds.Context([c[1] for c in cols], precision=mds.FP32_BULK, rp=mds.run_params(cfl=0.5))
    for k, c in enumerate(cols):
        p = c[2].copy()
        if k == 1:
            p[1, 0, 0, 250] = -1.0e20
        ctx.upload(p, k)
    ctx.update_eos()
    ctx.run(2)
    hs = [ctx.health(k)["bad_state"] for k in range(3)]
    print(f"[health] batched columns: {hs}", flush=True)
    assert hs[0] == 0 and hs[2] == 0 and hs[1] >= 1
