#!/usr/bin/env python3
"""Headline benchmark of the explicit hydro sweep on one MI355X.

Workload: 2D Kelvin-Helmholtz, 2048 x 2048 periodic box, gamma = 5/3,
volume-moment PPM + HLLC/HLLE, SSP-RK2, CFL 0.5 (synthetic seeded initial
state).  A step is one full SSP-RK2 step of every zone (both stages,
boundaries, EOS, sources and the device-side dt policy included).

Prints ONE JSON line:
  value      zone-updates/s of the FP32-bulk path (the headline)
  fp64       the FP64 path on the same workload and step count
  roofline   dominant kernel (the FP32-bulk 2D stage kernel): algorithmic
             bytes per launch / average launch duration measured with HIP
             events on the context stream inside the timed region
  cpu_baseline  the CPU oracle (reference-flag build, 1 thread) on a
             bounded number of steps of the same grid
Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, vendor figure
BYTES_PER_ZONE_UPDATE = {"fp32": 5 * 4 * 4, "fp64": 5 * 4 * 8}   # 5 state passes x NVAR=4 x size


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def l1rel(a, b):
    den = np.sum(np.abs(b))
    return float(np.sum(np.abs(a - b)) / den) if den > 0 else float(np.sum(np.abs(a - b)))


def run_path(mds, prob, prims, precision, steps, warmup, nz):
    ctx = mds.Context([prob], precision=precision, rp=mds.run_params(cfl=0.5))
    ctx.upload(prims)
    ctx.update_eos()
    tuned_ms = ctx.autotune(3)          # launch-shape autotuner (state is restored afterwards)
    rows = ctx.get_tile(mds.TILE_ROWS2D)
    m0 = ctx.totals()
    if warmup:
        ctx.run(warmup, graph=True)
        ctx.sync()
    tot_ms, st1_ms, st2_ms = ctx.run_timed(steps)
    final = ctx.download()
    m1 = ctx.totals()
    t, dt, ist = ctx.clock()
    # graph replay rate on top of the timed region (informational)
    g0 = time.perf_counter()
    ctx.run(steps, graph=True)
    ctx.sync()
    g1 = time.perf_counter()
    launches = 2 * steps
    avg_launch_ms = (st1_ms + st2_ms) / launches
    return dict(
        ctx=ctx, final=final,
        ms_per_step=tot_ms / steps,
        value=nz * steps / (tot_ms * 1e-3),
        graph_value=nz * steps / (g1 - g0),
        avg_launch_ms=avg_launch_ms,
        stage_ms=(st1_ms / steps, st2_ms / steps),
        mass_drift=float((m1[0] - m0[0]) / m0[0]),
        energy_drift=float((m1[1] - m0[1]) / m0[1]),
        t=t, istep=ist, counters=ctx.counters(), rows=rows, tuned_ms=tuned_ms)


def cpu_baseline(n, steps):
    """CPU oracle, reference optimisation flags (-O3 -ffast-math), 1 thread, same grid."""
    from oracle import pyoracle as po
    from medusa_amd import problems
    g, prob, prims = problems.kh(n)
    cfg = po.make_config(2, [n, n], [0.0, 0.0], [1.0 / n, 1.0 / n], [3, 3], [3, 3], 5.0 / 3.0, cfl=0.5)
    sim = po.Sim(cfg, fast=True)
    sim.set_prims(prims)
    t0 = time.perf_counter()
    sim.run(max_steps=steps)
    el = time.perf_counter() - t0
    return dict(value=n * n * steps / el, unit="zone-updates/s", cores=1, kind="port",
                sample=f"KH {n}x{n}, first {steps} SSP-RK2 steps from the same initial state, "
                       f"1 thread, oracle built with -O3 -ffast-math -ftree-vectorize -fpeel-loops; "
                       f"{el:.1f} s")


def traffic_from_profile():
    """HBM bytes per launch of the FP32 stage kernel from the committed counter summary, if present."""
    p = os.path.join(ROOT, "profiles", "kh2d_2048_sweep2d_f32_traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get("traffic_bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--only", choices=["fp32", "fp64"], default=None)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)

    from medusa_amd import mds, problems
    if not os.path.exists(mds.LIB_PATH):
        from medusa_amd import build
        build.build(verbose=False)

    n = args.n
    nz = n * n
    log(f"[bench] KH {n}x{n}: building grid and initial state")
    g, prob, prims = problems.kh(n)
    res = {}
    for name, prec in (("fp64", mds.FP64), ("fp32", mds.FP32_BULK)):
        if args.only and args.only != name:
            continue
        log(f"[bench] {name}: warmup {args.warmup} steps, timed {args.steps} steps")
        res[name] = run_path(mds, prob, prims, prec, args.steps, args.warmup, nz)
        r = res[name]
        log(f"[bench] {name}: {r['value']:.4e} zone-updates/s, {r['ms_per_step']:.3f} ms/step, "
            f"graph {r['graph_value']:.4e}, launch {r['avg_launch_ms'] * 1e3:.1f} us, "
            f"mass drift {r['mass_drift']:.2e}, autotuned rows/wave {r['rows']}")

    head = res.get("fp32") or res.get("fp64")
    head_name = "fp32" if "fp32" in res else "fp64"
    err = None
    if "fp32" in res and "fp64" in res:
        a, b = res["fp32"]["final"], res["fp64"]["final"]
        gam = 5.0 / 3.0
        err = dict(rho=l1rel(a[0], b[0]), p=l1rel((gam - 1) * a[1], (gam - 1) * b[1]),
                   v=l1rel(np.hypot(a[2], a[3]), np.hypot(b[2], b[3])))
        log(f"[bench] FP32-bulk vs FP64 relative L1 after {args.warmup + args.steps} steps: {err}")

    bpl = BYTES_PER_ZONE_UPDATE[head_name] / 2.0 * nz       # algorithmic bytes per stage launch
    achieved = bpl / (head["avg_launch_ms"] * 1e-3) / 1e9
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic_from_profile(),
                    kernel=f"sweep2d<{'float' if head_name == 'fp32' else 'double'},2>",
                    bytes_per_launch=bpl, avg_launch_us=round(head["avg_launch_ms"] * 1e3, 2))

    cpu = None
    if not args.no_cpu_baseline:
        log(f"[bench] CPU baseline: oracle on {n}x{n} for {args.cpu_steps} steps")
        cpu = cpu_baseline(n, args.cpu_steps)
        log(f"[bench] CPU baseline {cpu['value']:.4e} zone-updates/s")

    out = {
        "metric": "zone-updates/s on 1 MI355X, FP32-bulk vs FP64 path, %HBM peak at stated L1 err",
        "value": head["value"],
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if head_name == "fp32" else "f64",
        "data": "synthetic",
        "config": {"workload": f"kh2d {n}x{n} periodic, gamma=5/3, PPM+HLLC, SSP-RK2, CFL 0.5",
                   "zones": nz, "path": "FP32-bulk" if head_name == "fp32" else "FP64",
                   "parallelism": "single-gpu"},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if "fp64" in res:
        r = res["fp64"]
        out["fp64"] = dict(value=r["value"], ms_per_step=r["ms_per_step"], graph_value=r["graph_value"],
                           avg_launch_us=r["avg_launch_ms"] * 1e3,
                           achieved_gbs=BYTES_PER_ZONE_UPDATE["fp64"] / 2.0 * nz / (r["avg_launch_ms"] * 1e-3) / 1e9,
                           mass_drift=r["mass_drift"], energy_drift=r["energy_drift"])
    if "fp32" in res:
        r = res["fp32"]
        out["fp32"] = dict(graph_value=r["graph_value"], mass_drift=r["mass_drift"],
                           energy_drift=r["energy_drift"], stage_ms=r["stage_ms"])
    if "fp32" in res and "fp64" in res:
        out["speedup_fp32_over_fp64"] = res["fp32"]["value"] / res["fp64"]["value"]
        out["l1_fp32_vs_fp64"] = err
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
