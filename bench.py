#!/usr/bin/env python3
"""Benchmark of the MI355X-native vHLLE hydro step.

Workload (one GPU): the 3-D viscous run, grid 200x200x100, dx=dy=0.1 fm,
d(eta)=0.05, tau0=0.5 fm, dtau=0.025 fm, 300 predictor-corrector steps, smooth
Gaussian initial energy density e0=50 GeV/fm^3 with an eta plateau, zero
initial flow, p=e/3, eta/s=0.08, zeta/s=0 -- plain FP64 (faithful) step.

"step" = one Hydro::performStep (src/hdo.cpp:1016-1134) over the whole grid,
ghost cells included.  value = cells * steps / time of the timed region, with
the state resident in HBM; the timed region runs the steps as replayed HIP
graphs from the initial state at tau0.  Per-kernel durations for the roofline
come from HIP events recorded on the context's stream around every launch
during a second, identical pass.

python bench.py [--gpus 1] [--steps 300] [--warmup 5] [--no-cpu]
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "viscous cell-updates/s (plain FP64 vs mixed+learned c2p) + % of HBM peak"
HBM_PEAK_GBS = 8000.0

# Algorithmic (compulsory) HBM bytes per cell for each kernel of the FP64
# step: every input array read once and every output array written once, at
# 8 bytes per value ("Kernels and their rooflines").
ALG_BYTES = {
    "faces_x_pred": 7 * 8 + 7 * 8, "faces_y_pred": 7 * 8 + 7 * 8, "faces_z_pred": 7 * 8 + 7 * 8,
    "faces_x_corr": 7 * 8 + 7 * 8, "faces_y_corr": 7 * 8 + 7 * 8, "faces_z_corr": 7 * 8 + 7 * 8,
    "update_pred": (7 + 21 + 7) * 8, "update_corr": (7 + 7 + 21 + 7) * 8,
    "visc_prim": 7 * 8 + 12 * 8,
    "visc_relax": (8 + 4 + 7 + 7 + 10 + 1) * 8 + 22 * 8,
    "visc_advect": (5 + 22 + 11) * 8 + (22 + 1 + 4) * 8,
    "visc_flux_update": (3 + 11 + 4 + 7) * 8 + 7 * 8,
}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def run_gpu(args, cfg):
    from vhlle_amd import abi, setups
    nx, ny, nz = cfg["nx"], cfg["ny"], cfg["nz"]
    N = nx * ny * nz
    trc = abi.make_trc(etaS=cfg["etaS"])
    ctx = abi.Context(nx, ny, nz, cfg["xlim"], cfg["ylim"], cfg["zlim"], trc=trc)
    e, vx, vy, vz = setups.gaussian3d_ic(nx, ny, nz, cfg["xlim"], cfg["ylim"], cfg["zlim"],
                                         e0=cfg["e0"], sx=cfg["sx"], sy=cfg["sy"],
                                         eta_plateau=cfg["eta_plateau"],
                                         eta_sigma=cfg["eta_sigma"])
    dt, tau0 = cfg["dt"], cfg["tau0"]
    print(f"[bench] device: {abi.device_info()}", file=sys.stderr, flush=True)
    # warmup (also builds the step graphs)
    ctx.set_prim(tau0, e, vx, vy, vz)
    if args.warmup:
        st, _ = ctx.run(tau0, dt, args.warmup)
    # timed region: K steps from the initial state, resident in HBM
    ctx.set_prim(tau0, e, vx, vy, vz)
    t0 = time.perf_counter()
    st, tau = ctx.run(tau0, dt, args.steps)  # synchronises before returning
    t1 = time.perf_counter()
    elapsed = t1 - t0
    print(f"[bench] timed {args.steps} steps: {elapsed:.3f} s, status {st}, tau {tau:.4f}",
          file=sys.stderr, flush=True)
    # physical sanity of the final state
    s = ctx.download()
    tot = float(np.sum(s["Q"][0]))
    # per-kernel pass with HIP events on the context stream
    ctx.set_prim(tau0, e, vx, vy, vz)
    ctx.enable_timing(True)
    st2, _ = ctx.run(tau0, dt, args.steps)
    kt = ctx.kernel_times()
    ctx.enable_timing(False)
    ctx.close()
    return dict(N=N, elapsed=elapsed, status=st | st2, kernel_times=kt, Qt_sum=tot)


def roofline(kt, N):
    best = max(kt.items(), key=lambda kv: kv[1][1])
    name, (launches, ms) = best
    avg_s = ms / 1e3 / launches
    alg = ALG_BYTES.get(name, 0) * N
    achieved = alg / avg_s / 1e9
    return name, dict(bound="hbm", achieved=round(achieved, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                      frac=round(achieved / HBM_PEAK_GBS, 5), traffic=None,
                      kernel=name, avg_ms=round(avg_s * 1e3, 4),
                      alg_bytes_per_launch=alg)


def cpu_baseline(cfg):
    """The oracle (single-threaded CPU restatement) on a bounded sample of the
    same workload: a 200x200x10 slab of the grid (same spacings, centre of the
    eta plateau) for 5 steps (~20 s on one core)."""
    from oracle import pyoracle
    from vhlle_amd import setups
    nz = 10
    half = (nz - 1) / 2 * 0.05
    zl = (-half, half)
    nx, ny = cfg["nx"], cfg["ny"]
    e, vx, vy, vz = setups.gaussian3d_ic(nx, ny, nz, cfg["xlim"], cfg["ylim"], zl, e0=cfg["e0"],
                                         sx=cfg["sx"], sy=cfg["sy"],
                                         eta_plateau=cfg["eta_plateau"],
                                         eta_sigma=cfg["eta_sigma"])
    s = pyoracle.Solver(nx, ny, nz, cfg["xlim"], cfg["ylim"], zl,
                        trc=pyoracle.make_trc(etaS=cfg["etaS"]))
    s.set_prim(cfg["tau0"], e, vx, vy, vz)
    steps = 5
    t0 = time.perf_counter()
    rc, _ = s.run(cfg["tau0"], cfg["dt"], steps)
    dt = time.perf_counter() - t0
    return dict(value=round(nx * ny * nz * steps / dt, 1), unit="cell-updates/s", cores=1,
                kind="port", sample=f"oracle, {nx}x{ny}x{nz} slab of the workload grid, "
                f"{steps} steps, {dt:.1f} s on 1 core of {cpu_model()}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; multi-GPU sharding of the hydro "
              "grid is designed (DESIGN.md) but not built.", file=sys.stderr)
        sys.exit(2)
    from vhlle_amd import setups
    cfg = setups.VISCOUS_3D
    r = run_gpu(args, cfg)
    N = r["N"]
    value = N * args.steps / r["elapsed"]
    kname, roof = roofline(r["kernel_times"], N)
    # whole-step % of HBM peak with the compulsory 288 B per viscous cell-update
    step_frac = 288.0 * value / (HBM_PEAK_GBS * 1e9)
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["elapsed"] / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{cfg['name']}: 3-D viscous, dx=dy=0.1, deta=0.05, tau0=0.5, "
                               f"dtau=0.025, eta/s=0.08, p=e/3, plain FP64 step",
                   "cells": N, "precision": "fp64", "c2p": "reference"},
        "step_hbm_frac_288B": round(step_frac, 5),
        "status": r["status"],
        "roofline": roof,
        "kernels_ms": {k: round(v[1], 3) for k, v in sorted(r["kernel_times"].items())},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cfg)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
