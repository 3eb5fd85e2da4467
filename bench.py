#!/usr/bin/env python3
"""Benchmark of the explicit hydro sweep on one MI355X.

Default workload (the headline, the batched sweep the north-star target is
quoted on): an ensemble of 2048 independent 1D spherical Riemann problems on
log-r grids, 256-8192 zones each (seeded, 8.8M zones in all), gamma-law EOS,
volume-moment PPM + HLLC, SSP-RK2, CFL 0.5, per-problem dt, every problem
advanced by one launch per step (both RK stages) from a segment table.  A
step is one full SSP-RK2 step of every zone (both stages, boundaries, EOS,
sources and the device-side dt policy).
Other workloads (--workload): kh2d (2048^2 Kelvin-Helmholtz, 200 steps),
rt3d (512^3 Rayleigh-Taylor, 100 steps), star3d (640x128x256 dendritic star,
table EOS, monopole gravity, 200 steps), sod1d (1024 zones, to t=0.2).

Prints ONE JSON line:
  value         zone-updates/s of the FP32-bulk path (FP64 for sod1d)
  fp64          the FP64 path on the same workload and step count
  roofline      dominant kernel (the headline path's stage kernel):
                algorithmic bytes per launch / average launch duration,
                measured with HIP events on the context stream inside the
                timed region
  cpu_baseline  the CPU oracle (reference-flag build) on a bounded sample of
                the same workload: one process on one core, and (ensemble)
                independent single-thread processes on the host's cores
Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload ensemble]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, vendor figure


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def l1rel(a, b, w=None):
    w = 1.0 if w is None else w
    den = np.sum(np.abs(b) * w)
    return float(np.sum(np.abs(a - b) * w) / den) if den > 0 else float(np.sum(np.abs(a - b) * w))


# --------------------------------------------------------------------------
# workloads: (problems, NVAR, default steps, label, kernel name)
# --------------------------------------------------------------------------
def workload(name, n):
    from medusa_amd import problems
    if name == "kh2d":
        n = n or 2048
        g, prob, prims = problems.kh(n)
        return dict(probs=[(prob, prims)], nvar=4, steps=200, zones=n * n, gam=[5.0 / 3.0],
                    label=f"kh2d {n}x{n} periodic, gamma=5/3, PPM+HLLC, SSP-RK2, CFL 0.5",
                    kernel={"fp32": "sweep2d<float,2,false,false>", "fp64": "sweep2d<double,2,false,false>"}, n=n)
    if name == "rt3d":
        n = n or 512
        g, prob, prims = problems.rt(n)
        return dict(probs=[(prob, prims)], nvar=5, steps=100, zones=n ** 3, gam=[1.4],
                    label=f"rt3d {n}^3, g=-0.1 z, periodic x,y, reflecting z, gamma=1.4, CFL 0.5",
                    kernel={"fp32": "sweep3d<float,3,true,TY,false>", "fp64": "sweep3d<double,3,true,TY,false>"}, n=n)
    if name == "sod1d":
        n = n or 1024
        g, prob, prims = problems.sod(n)
        return dict(probs=[(prob, prims)], nvar=3, steps=909, zones=n, gam=[1.4],
                    label=f"sod1d {n} zones, outflow, gamma=1.4, CFL 0.5, 909 steps (t=0.2)",
                    kernel={"fp32": "sweep1d_f32<1,false>", "fp64": "sweep1d_x64<1,false>"}, n=n)
    if name == "ensemble":
        nprob = n or 2048
        par, ens = problems.ensemble(nprob)
        zones = int(sum(int(r[0]) for r in par))
        return dict(probs=[(e[1], e[2]) for e in ens], nvar=3, steps=500, zones=zones, gam=list(par[:, 1]),
                    label=f"ensemble of {nprob} spherical log-r Riemann columns, 256-8192 zones (seeded), "
                          f"{zones} zones, per-problem dt, 500 steps",
                    kernel={"fp32": "sweep1d_f32<1,true> per stage, or sweep1d2_f32<1,true> for both stages",
                            "fp64": "sweep1d_x64<3,true> per stage, or sweep1d2_x64<3,true> for both stages"},
                    n=nprob)
    if name == "star3d":
        n1 = n or 640
        n2 = 1 << max(3, round(math.log2(max(8, n1 // 5))))
        n3 = 2 * n2
        T = mds_table()
        g, prob, prims = problems.star3d(n1, n2, n3, table=T, dendritic=True)
        zones = int(g.valid_mask().sum())
        return dict(probs=[(prob, prims)], nvar=6, steps=200, zones=zones, gam=[4.0 / 3.0], table=T,
                    label=f"star3d {n1}x{n2}x{n3} dendritic spherical grid ({zones} zones; sinh r from the "
                          f"origin, 0.5 km central spacing, to 2e9 cm), n=3 polytrope rho_c=1e10, synthetic FP32 "
                          f"EOS table 180x180x50, monopole gravity",
                    kernel={"fp32": "sph3d stage: eos + 3 flux (+2 listed level-change) + update kernels (float "
                                    "state)",
                            "fp64": "sph3d stage: eos + 3 flux (+2 listed level-change) + update kernels (double "
                                    "state)"}, n=n1)
    raise SystemExit(f"unknown workload {name}")


def phys_velocity(mds, prob, prims):
    """physical velocity components v^d scale_d (spherical grids) of a prim array"""
    g = prob["grid"]
    v = np.array(prims[2:5], dtype=np.float64)
    if g.g.geom == mds.GEOM_SPHERICAL:
        if g.g.ndim == 1:
            v[0] = v[0] * g.zone()[2:-2, mds.ZG_SCALE0][None, None, :]
        else:
            sc = g.zone()[2:-2, 2:-2, mds.ZS_SCALE0:mds.ZS_SCALE0 + 3]
            for d in range(3):
                v[d] = v[d] * sc[None, :, :, d]
    return v


def mds_table():
    from medusa_amd import mds
    return mds.EosTable()


def run_path(mds, W, precision, steps, warmup, tiles=None, nvel=None):
    probs = [p for p, _ in W["probs"]]
    if nvel is not None:
        probs = [dict(p, nvel=nvel) for p in probs]
    ctx = mds.Context(probs, precision=precision, rp=mds.run_params(cfl=0.5))
    for k, (_, prims) in enumerate(W["probs"]):
        ctx.upload(prims, k)
    ctx.update_eos()
    if tiles:
        for t, v in zip((mds.TILE_ROWS2D, mds.TILE_TY3D, mds.TILE_KC3D, mds.TILE_FUSE1D, mds.TILE_GRID1D,
                         mds.TILE_HALO3D, mds.TILE_UPD3S), tiles):
            if t == mds.TILE_FUSE1D and v == 2 and precision != mds.FP32_BULK:
                v = 1           # two zones per lane is an FP32-bulk shape
            if t not in (mds.TILE_FUSE1D, mds.TILE_GRID1D) or ctx.ndim == 1:
                ctx.set_tile(t, v)
        tuned_ms = None
    else:
        tuned_ms = ctx.autotune(8)      # launch-shape autotuner (state is restored afterwards)
    m0 = np.array([ctx.totals(k) for k in range(len(probs))])
    if warmup:
        ctx.run(warmup, graph=True)
        ctx.sync()
    tot_ms, st1_ms, st2_ms = ctx.run_timed(steps)
    finals = [ctx.download(k) for k in range(len(probs))]
    m1 = np.array([ctx.totals(k) for k in range(len(probs))])
    g0 = time.perf_counter()
    ctx.run(steps, graph=True)          # graph replay rate on top of the timed region (informational)
    ctx.sync()
    g1 = time.perf_counter()
    zones = W["zones"]
    # sweep launches per step: one two-stage launch for 1D columns without periodic x, else one per stage
    launches = 1 if (ctx.ndim == 1 and ctx.get_tile(mds.TILE_FUSE1D)) else 2
    drift = np.max(np.abs((m1[:, 0] - m0[:, 0]) / m0[:, 0]))
    edrift = np.max(np.abs((m1[:, 1] - m0[:, 1]) / m0[:, 1]))
    return dict(finals=finals, ms_per_step=tot_ms / steps, value=zones * steps / (tot_ms * 1e-3),
                graph_value=zones * steps / (g1 - g0), avg_launch_ms=(st1_ms + st2_ms) / (launches * steps),
                launches=launches,
                stage_ms=(st1_ms / steps, st2_ms / steps), mass_drift=float(drift), energy_drift=float(edrift),
                counters=ctx.counters(), tuned_ms=tuned_ms,
                tiles=[ctx.get_tile(t) for t in (mds.TILE_ROWS2D, mds.TILE_TY3D, mds.TILE_KC3D, mds.TILE_FUSE1D,
                                                 mds.TILE_GRID1D, mds.TILE_HALO3D, mds.TILE_UPD3S)])


def cpu_baseline(name, W, budget_s):
    """CPU oracle, reference optimisation flags (-O3 -ffast-math), 1 thread,
    on a bounded sample: one step of every sample problem to calibrate, then
    as many more steps as fill about budget_s seconds."""
    from oracle import pyoracle as po
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from parity_util import oracle_for
    from medusa_amd import problems
    if name == "rt3d":
        # bounded sample: a 64 x 64 x n column of the same zone shape and state
        n = W["n"]
        g, prob, prims = problems.rt(n)
        sub = np.ascontiguousarray(prims[:, :, :64, :64])
        from medusa_amd import mds
        g2 = mds.GridBuilder(3, [64, 64, n], [-0.25, -0.25, -0.75], [0.5 / n, 0.5 / n, 1.5 / n])
        prob2 = dict(prob, grid=g2)
        sample = [(prob2, sub)]
        what = f"64x64x{n} column of the rt3d grid (same zone shape and state), periodic x,y"
    elif name == "ensemble":
        sample = W["probs"][:64]
        what = "first 64 columns of the ensemble"
    elif name == "star3d":
        n1 = W["n"]
        g, prob, prims = problems.star3d(n1, 16, 32, table=W["table"], dendritic=True)
        sample = [(prob, prims)]
        what = f"the same star on a {n1}x16x32 dendritic grid ({int(g.valid_mask().sum())} zones)"
    else:
        sample = W["probs"]
        what = "the full grid"
    sims, zones = [], 0
    for prob, prims in sample:
        if prob.get("eos") is not None or prob.get("monopole"):
            kw = dict(nvars=6, monopole=bool(prob.get("monopole")))
            if prob.get("eos") is not None:
                t = prob["eos"].t
                kw["eos_tab"] = po.eos_tab(dict(nr=t.nr, nt=t.nt, ny=t.ny, lr0=t.lr0, lr1=t.lr1, lt0=t.lt0,
                                                lt1=t.lt1, ye0=t.ye0, ye1=t.ye1), prob["eos"].data)
            sim = oracle_for(prob, fast=True, **kw)
        else:
            sim = oracle_for(prob, fast=True)
        sim.set_prims(prims)
        sims.append(sim)
        zones += int(prob["grid"].valid_mask().sum()) if getattr(prob["grid"], "levels", None) and \
            prob["grid"].levels() is not None else int(np.prod(prims.shape[1:]))
    t0 = time.perf_counter()
    for sim in sims:
        sim.run(max_steps=1)
    t1 = time.perf_counter() - t0
    more = int(max(0, min(W["steps"] - 1, round(budget_s / max(t1, 1e-9)) - 1)))
    t0 = time.perf_counter()
    for sim in sims:
        sim.run(max_steps=more)
    el = t1 + (time.perf_counter() - t0)
    steps = 1 + more
    return dict(value=zones * steps / el, unit="zone-updates/s", cores=1, kind="port",
                sample=f"{what}, {steps} SSP-RK2 steps from the initial state, 1 thread, oracle built with "
                       f"-O3 -ffast-math -ftree-vectorize -fpeel-loops; {el:.1f} s")


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _column_worker(args):
    """One single-thread oracle process: its columns, stepped for a time budget."""
    cols, budget_s = args
    sys.path.insert(0, ROOT)
    from oracle import pyoracle as po
    sims, zones = [], 0
    for n, x0, dx, gam, prims in cols:
        cfg = po.make_config(1, [n], [x0], [dx], [po.BC_REFLECT], [po.BC_OUTFLOW], gam, cfl=0.5,
                             geom=po.GEOM_SPHERICAL, rcoord=po.RCOORD_LOG)
        sim = po.Sim(cfg, fast=True)
        sim.set_prims(prims)
        sims.append(sim)
        zones += n
    t0 = time.perf_counter()
    steps = 0
    while True:
        for sim in sims:
            sim.run(max_steps=1)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s:
            break
    return zones * steps, el


def cpu_baseline_procs(W, nproc, budget_s):
    """Ensemble throughput of nproc independent single-thread oracle processes
    (the reference runs one process per core and has no host threading on the
    hot path), each stepping its own share of the ensemble's columns."""
    import multiprocessing as mp
    probs = W["probs"]
    order = sorted(range(len(probs)), key=lambda k: -probs[k][0]["grid"].n[0])
    shares = [[] for _ in range(nproc)]
    for r, k in enumerate(order[:8 * nproc]):          # 8 columns per process, largest first
        prob, prims = probs[k]
        g = prob["grid"]
        shares[r % nproc].append((g.n[0], g.g.startx[0], g.g.dx[0], prob["gam"], prims))
    ctx = mp.get_context("spawn")
    with ctx.Pool(nproc) as pool:
        res = pool.map(_column_worker, [(sh, budget_s) for sh in shares])
    rate = sum(z / el for z, el in res)
    return dict(value=rate, unit="zone-updates/s", processes=nproc, cores=nproc,
                sample=f"{8 * nproc} largest columns of the ensemble, 8 per process, each process stepping its "
                       f"columns for about {budget_s:.0f} s; the sum of the per-process rates")


def traffic_from_profile(name, path):
    """HBM bytes per launch from the committed counter summary, if present."""
    p = os.path.join(ROOT, "profiles", f"{name}_{path}_traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get("traffic_bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="ensemble", choices=["ensemble", "kh2d", "rt3d", "sod1d", "star3d"])
    ap.add_argument("--n", type=int, default=None, help="grid size (problem count for ensemble)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="CPU baseline time budget")
    ap.add_argument("--cpu-procs", type=int, default=16,
                    help="independent single-thread oracle processes of the multi-process CPU baseline (ensemble)")
    ap.add_argument("--valu-probe", action="store_true",
                    help="print the device's FMA issue rate (wavefront instructions/s) and exit")
    ap.add_argument("--tiles", default=None,
                    help="ROWS2D,TY3D,KC3D[,FUSE1D[,GRID1D[,HALO3D[,UPD3S]]]] launch shape to use instead of autotuning "
                         "(profiling runs)")
    ap.add_argument("--copy-probe", type=int, default=None, metavar="BYTES",
                    help="only run the device copy kernel over BYTES (counter calibration) and exit")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--only", choices=["fp32", "fp64"], default=None)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)

    from medusa_amd import mds
    if not os.path.exists(mds.LIB_PATH):
        from medusa_amd import build
        build.build(verbose=False)

    if args.copy_probe:
        ms = mds.copy_probe_ms(args.copy_probe, 5)
        print(json.dumps({"copy_probe_bytes": args.copy_probe, "ms": ms,
                          "GB/s": 2.0 * args.copy_probe / (ms * 1e-3) / 1e9}), flush=True)
        return
    if args.valu_probe:
        rates = {"fp32": mds.valu_probe_rate(False), "fp64": mds.valu_probe_rate(True)}
        for k in ("rcp32", "rcp64", "sqrt32", "sqrt64"):
            rates[k] = mds.valu_probe_rate(kind=k)
        print(json.dumps({"valu_wave_instr_per_s": rates}), flush=True)
        return
    tiles = [int(t) for t in args.tiles.split(",")] if args.tiles else None

    log(f"[bench] workload {args.workload}: building grids and initial state")
    W = workload(args.workload, args.n)
    steps = args.steps if args.steps is not None else W["steps"]
    zones = W["zones"]
    paths = [("fp64", mds.FP64), ("fp32", mds.FP32_BULK)]
    if args.workload == "sod1d" and not args.only:
        paths = [("fp64", mds.FP64)]
    res = {}
    for name, prec in paths:
        if args.only and args.only != name:
            continue
        log(f"[bench] {name}: warmup {args.warmup} steps, timed {steps} steps")
        res[name] = r = run_path(mds, W, prec, steps, args.warmup, tiles)
        log(f"[bench] {name}: {r['value']:.4e} zone-updates/s, {r['ms_per_step']:.4f} ms/step, "
            f"graph {r['graph_value']:.4e}, launch {r['avg_launch_ms'] * 1e3:.1f} us, "
            f"mass drift {r['mass_drift']:.2e}, tiles {r['tiles']}")

    if args.workload == "ensemble" and "fp32" in res and "fp64" in res:
        # FP32-bulk carrying all three velocity components, as the FP64 path
        # does (the columns' transverse velocities stay zero): the
        # like-for-like state size beside the headline's radial-only columns
        log(f"[bench] fp32 with three velocity components: timed {steps} steps")
        r3 = run_path(mds, W, mds.FP32_BULK, steps, args.warmup, tiles, nvel=3)
        res["fp32_nvel3"] = r3
        log(f"[bench] fp32_nvel3: {r3['value']:.4e} zone-updates/s, launch {r3['avg_launch_ms'] * 1e3:.1f} us, "
            f"tiles {r3['tiles']}")
    head_name = "fp32" if "fp32" in res else "fp64"
    head = res[head_name]
    err = None
    if "fp32" in res and "fp64" in res:
        a, b = res["fp32"]["finals"], res["fp64"]["finals"]
        e = {"rho": [], "p": [], "v": [], "v/cs": []}
        for k, (fa, fb) in enumerate(zip(a, b)):
            gam = W["gam"][k] if k < len(W["gam"]) else W["gam"][0]
            prob = W["probs"][k][0]
            e["rho"].append(l1rel(fa[0], fb[0]))
            e["p"].append(l1rel((gam - 1) * fa[1], (gam - 1) * fb[1]))
            va, vb = phys_velocity(mds, prob, fa), phys_velocity(mds, prob, fb)
            e["v"].append(l1rel(np.sqrt(np.sum(va ** 2, 0)), np.sqrt(np.sum(vb ** 2, 0))))
            if prob.get("eos") is None:
                # the velocity difference on the sound-speed scale: where most of a
                # column is at rest, the relative L1 of |v| divides by almost nothing
                cs = np.sqrt(gam * (gam - 1.0) * np.abs(fb[1]) / np.where(fb[0] > 0, fb[0], 1.0))
                e["v/cs"].append(float(np.sum(np.sqrt(np.sum((va - vb) ** 2, 0))) / np.sum(cs)))
        err = {k: float(np.max(v)) for k, v in e.items() if v}
        log(f"[bench] FP32-bulk vs FP64 relative L1 (max over problems) after {args.warmup + steps} steps: {err}")

    bpz = 5 * W["nvar"] * (4 if head_name == "fp32" else 8)      # algorithmic bytes per zone-update
    bpl = bpz / head["launches"] * zones                           # per sweep launch
    achieved = bpl / (head["avg_launch_ms"] * 1e-3) / 1e9
    kname = W["kernel"][head_name]
    if args.workload in ("sod1d", "ensemble"):
        # the launch the autotuner (or --tiles) chose: both RK stages in one, or one per stage
        two = head["launches"] == 1
        base = {"sod1d": ("sweep1d_f32<1,false,0>", "sweep1d_x64<1,false,0>"),
                "ensemble": ("sweep1d_f32<1,true,0>", "sweep1d_x64<3,true,0>")}[args.workload]
        kname = base[0 if head_name == "fp32" else 1]
        if two and head_name == "fp32":
            # template <NV, SPH, TAB, zones per lane>
            z = 2 if head["tiles"][3] == 2 else 1
            kname = kname.replace("sweep1d_", "sweep1d2_").replace(",0>", f",0,{z}>") + \
                f" (both RK stages per launch, {z} zone{'s' if z > 1 else ''} per lane)"
        elif two:
            kname = kname.replace("sweep1d_", "sweep1d2_") + " (both RK stages per launch)"
    elif args.workload == "rt3d":
        kname = kname.replace("TY", str(head["tiles"][1]))
        if head["tiles"][5]:
            kname = kname.replace("sweep3d<", "sweep3dh<")
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic_from_profile(args.workload, head_name),
                    kernel=kname, bytes_per_launch=bpl,
                    avg_launch_us=round(head["avg_launch_ms"] * 1e3, 2))

    cpu = None
    if not args.no_cpu_baseline:
        log(f"[bench] CPU baseline: oracle, about {args.cpu_seconds:.0f} s")
        cpu = cpu_baseline(args.workload, W, args.cpu_seconds)
        cpu["cpu_model"] = cpu_model()
        log(f"[bench] CPU baseline {cpu['value']:.4e} zone-updates/s ({cpu['cpu_model']})")
        if args.workload == "ensemble" and args.cpu_procs > 1:
            nproc = min(args.cpu_procs, os.cpu_count() or 1)
            log(f"[bench] CPU baseline: {nproc} oracle processes, about {args.cpu_seconds:.0f} s")
            cpu["n_process"] = cpu_baseline_procs(W, nproc, args.cpu_seconds)
            log(f"[bench] CPU baseline {nproc} processes {cpu['n_process']['value']:.4e} zone-updates/s")

    out = {
        "metric": "zone-updates/s on 1 MI355X, FP32-bulk vs FP64 path, %HBM peak at stated L1 err",
        "value": head["value"],
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if head_name == "fp32" else "f64",
        "data": "synthetic",
        "config": {"workload": W["label"], "zones": zones,
                   "path": "FP32-bulk" if head_name == "fp32" else "FP64", "parallelism": "single-gpu"},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if "fp64" in res:
        r = res["fp64"]
        out["fp64"] = dict(value=r["value"], ms_per_step=r["ms_per_step"], graph_value=r["graph_value"],
                           avg_launch_us=r["avg_launch_ms"] * 1e3,
                           achieved_gbs=5 * W["nvar"] * 8 / r["launches"] * zones / (r["avg_launch_ms"] * 1e-3) / 1e9,
                           mass_drift=r["mass_drift"], energy_drift=r["energy_drift"], tiles=r["tiles"])
    if "fp32" in res:
        r = res["fp32"]
        out["fp32"] = dict(graph_value=r["graph_value"], mass_drift=r["mass_drift"],
                           energy_drift=r["energy_drift"], stage_ms=r["stage_ms"], tiles=r["tiles"])
    if "fp32" in res and "fp64" in res:
        out["speedup_fp32_over_fp64"] = res["fp32"]["value"] / res["fp64"]["value"]
        out["l1_fp32_vs_fp64"] = err
    if "fp32_nvel3" in res:
        r = res["fp32_nvel3"]
        out["fp32_nvel3"] = dict(value=r["value"], avg_launch_us=r["avg_launch_ms"] * 1e3, tiles=r["tiles"],
                                 speedup_over_fp64=r["value"] / res["fp64"]["value"],
                                 note="FP32-bulk carrying v^1..v^3 like the FP64 path (5 state arrays)")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
