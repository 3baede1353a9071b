#!/usr/bin/env python3
"""Benchmark of the MI355X-native vHLLE hydro step.

Workload (one GPU): the 3-D viscous run, grid 200x200x100, dx=dy=0.1 fm,
d(eta)=0.05, tau0=0.5 fm, dtau=0.025 fm, 300 predictor-corrector steps, smooth
Gaussian initial energy density e0=50 GeV/fm^3 with an eta plateau, zero
initial flow, p=e/3, eta/s=0.08, zeta/s=0.

Variants (the metric compares plain FP64 with mixed precision + faster
con2prim):
  fp64          faithful FP64 step (reference operation order, parity path)
  mixed_ref     FP32 fluxes/viscous sources/pi, FP64 Q + reference con2prim
  mixed_closed  as mixed_ref with the closed-form conformal con2prim
  mixed_newton  as mixed_ref, con2prim = FP64 Newton polish to |f| <= 1e-12 Q_t
                started at v = 0.5 (the ablation of mixed_learned)
  mixed_learned the same polish started from the BF16-MFMA MLP guess; the MLP is
                trained first, on the GPU, on 4M seeded pairs harvested from ideal
                Gubser runs (101x101, |v| up to 0.99 by augmentation)
Errors: separate runs of every variant are compared with the fp64 run after
20, 50 and 100 steps over the cells with e > e_crit = 0.5 GeV/fm^3 (inside
the freeze-out surface): max |T^{tau mu} - T_fp64| / max T^{tau tau}_fp64 and
the same for pi^{mu nu}.  Beside them, the same measure for FP64 runs whose
initial state was changed by 1e-12 and by 6e-8 (FP32 roundoff): the
reference scheme's discontinuous switches amplify any change about 1000-fold
every 30 steps (1e-12 -> 1e-6 at 100 steps), so the measure means something
only while these floors stay far below the bar.  The bar, 1e-6 in T^{tau mu},
is applied at 20 steps (the timed horizon, floors 4e-11 and 8e-9).

"step" = one Hydro::performStep (src/hdo.cpp:1016-1134) over the whole grid,
ghost cells included.  value = cells * steps / time of the timed region of the
headline variant (mixed_learned, the variant the metric names, if its T^{tau mu}
error after 20 steps is <= 1e-6 and its run raised no error status;
otherwise the fastest variant that passes), with the
state resident in HBM; the timed region replays the step as HIP graphs from the
initial state at tau0.  Per-kernel durations for the roofline come from HIP
events recorded on the context's stream around every launch during a second,
identical pass.

python bench.py [--gpus 1] [--steps 300] [--warmup 5] [--variants a,b] [--no-cpu]
                [--no-batch] [--no-full3d] [--no-table] [--no-vort] [--quick]

Also measured (field "batch_ebe"): 256 event-by-event boost-invariant grids of
sides 100..160 with hot-spot initial states in one batch context, against the
same events run one context at a time; (field "full_3d_freezeout") the
300x300x150 hot-spot run, 600 steps with the freeze-out surface every 5
steps, FP64 against FP32 fluxes (element count and sum of dsigma_mu); (field
"table_eos") the 200x200x100 workload with a tabulated crossover EoS, where
con2prim has no closed form: FP64 against mixed precision with the Newton
from v = 0.5 and from the learned guess, and the guess's Newton-update
ablation; (field "vorticity") the FP64 200x200x100 step with and without the
vorticity output, and the freeze-out with the beta.dat records.
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
ERR_BAR = 1e-6
ERR_STEPS = 20               # the horizon of the bar
ERR_HORIZONS = (20, 50, 100)  # reported
E_CRIT = 0.5

# Algorithmic (compulsory) HBM bytes per cell of each kernel: every input
# array read once and every output written once at its stored width (FP64 8 B,
# FP32 4 B).
ALG_BYTES = {
    "fp64": {
        "faces_x_pred": 112, "faces_y_pred": 112, "faces_z_pred": 112,
        "faces_x_corr": 112, "faces_y_corr": 112, "faces_z_corr": 112,
        "update_pred": (7 + 21 + 7) * 8, "update_corr": (7 + 7 + 21 + 7) * 8,
        "visc_prim": 7 * 8 + 12 * 8,
        "visc_relax": (8 + 4 + 7 + 7 + 10 + 1) * 8 + 22 * 8,
        "visc_advect": (5 + 22 + 11) * 8 + (22 + 1 + 4) * 8,
        "visc_flux_update": (3 + 11 + 4 + 7) * 8 + 7 * 8,
    },
    # mixed kernels on a state without charges: 4 conserved components (the
    # workload's n_b = n_q = n_s = 0), pi and Pi in FP32
    "mixed": {
        "stage_pred": 32 + 32, "stage_corr": 32 + 32 + 32, "c2p_guess": 32 + 4,
        "visc_prim": 32 + 40 + 32,
        "visc_relax": 32 + 32 + 32 + 32 + 44 + 88,
        "visc_advect": 40 + 88 + 20 + 80 + 8 + 8 + 16,
        "visc_flux_update": 32 + 44 + 16 + 32 + 32,
    },
}
# HBM bytes per launch measured by rocprofv3 (profiles/traffic_viscous_200x200x100.md):
# 2 x FETCH_SIZE + WRITE_SIZE in KiB, the factors those counters need to read
# 1 GiB for a plain copy of 1 GiB.  Keyed by (precision, kernel, stage kernel
# shape, con2prim); None where not profiled.
KIB = 1024
TRAFFIC = {
    ("fp64", "visc_relax", None, "reference"): (2 * 536633 + 843226) * KIB,
    ("fp64", "visc_advect", None, "reference"): (2 * 550046 + 843750) * KIB,
    ("fp64", "visc_flux_update", None, "reference"): (2 * 544937 + 218782) * KIB,
    ("mixed", "stage_pred", "cells", "closed"): (2 * 93173 + 130859) * KIB,
    ("mixed", "stage_corr", "cells", "closed"): (2 * 155871 + 130859) * KIB,
    ("mixed", "stage_pred", "cells", "newton"): (2 * 93182 + 130859) * KIB,
    ("mixed", "stage_corr", "cells", "newton"): (2 * 155881 + 130859) * KIB,
    ("mixed", "stage_pred", "cells", "learned"): (2 * 101135 + 130859) * KIB,
    ("mixed", "stage_corr", "cells", "learned"): (2 * 165839 + 130859) * KIB,
    ("mixed", "stage_pred", "rows", "closed"): (2 * 104965 + 137743) * KIB,
    ("mixed", "stage_corr", "rows", "closed"): (2 * 209804 + 136193) * KIB,
    ("mixed", "stage_pred", "rows", "newton"): (2 * 104800 + 137863) * KIB,
    ("mixed", "stage_corr", "rows", "newton"): (2 * 209354 + 136380) * KIB,
    ("mixed", "stage_pred", "rows", "learned"): (2 * 104832 + 137756) * KIB,
    ("mixed", "stage_corr", "rows", "learned"): (2 * 235317 + 136074) * KIB,
    ("mixed", "stage_pred", "bands", "closed"): (2 * 423052 + 135446) * KIB,
    ("mixed", "stage_corr", "bands", "closed"): (2 * 517915 + 135043) * KIB,
    ("mixed", "stage_pred", "bands", "newton"): (2 * 426944 + 134767) * KIB,
    ("mixed", "stage_corr", "bands", "newton"): (2 * 516415 + 134492) * KIB,
    ("mixed", "stage_pred", "bands", "learned"): (2 * 479110 + 135489) * KIB,
    ("mixed", "stage_corr", "bands", "learned"): (2 * 582564 + 134406) * KIB,
    ("mixed", "visc_relax", None, "closed"): (2 * 352331 + 375000) * KIB,
    ("mixed", "visc_relax", None, "newton"): (2 * 352360 + 375001) * KIB,
    ("mixed", "visc_relax", None, "learned"): (2 * 360216 + 375003) * KIB,
    ("mixed", "visc_advect", None, None): (2 * 330176 + 437500) * KIB,
    ("mixed", "visc_flux_update", None, None): (2 * 257136 + 125000) * KIB,
    ("mixed", "visc_prim", None, "closed"): (2 * 62522 + 281250) * KIB,
    ("mixed", "visc_prim", None, "newton"): (2 * 62560 + 281250) * KIB,
    ("mixed", "visc_prim", None, "learned"): (2 * 70365 + 281250) * KIB,
    ("mixed", "c2p_guess", None, "learned"): (2 * 62619 + 15625) * KIB,
}


def traffic_of(prec, kernel, stage, c2p):
    for key in ((prec, kernel, stage, c2p), (prec, kernel, stage, None), (prec, kernel, None, c2p),
                (prec, kernel, None, None)):
        if key in TRAFFIC:
            return TRAFFIC[key]
    return None


# whole-step compulsory state traffic per cell-update: 2*(n_Q*s_Q + 11*s_pi)
# at the stored widths -- FP64 288 B (7 components); mixed 152 B (the 4
# components a state without charges carries, pi and Pi in FP32)
STEP_BYTES = {"fp64": 288.0, "mixed": 152.0}

VARIANTS = {
    "fp64": ("fp64", "reference"),
    "mixed_ref": ("mixed", "reference"),
    "mixed_closed": ("mixed", "closed"),
    "mixed_newton": ("mixed", "newton"),
    "mixed_learned": ("mixed", "learned"),
}
C2P_PAIRS = 4 << 20   # training pairs of the learned con2prim
C2P_EPOCHS = 8
_MODEL = {}


def c2p_model():
    """Trains the con2prim MLP once (GPU kernels) and measures the ablation on
    held-out harvested states: Newton updates from the MLP guess vs v = 0.5."""
    if "p" in _MODEL:
        return _MODEL["p"]
    from vhlle_amd import abi
    t0 = time.perf_counter()
    Q, prim = abi.harvest(C2P_PAIRS, seed=2024, steps=200, amps=(10.0, 1.0), augment=0.5,
                          vmax=0.99)
    t1 = time.perf_counter()
    p, hist = abi.mlp_train(abi.mlp_init(7), Q, prim, batch=512, epochs=C2P_EPOCHS, lr=2e-3,
                            lr_final=1e-6, seed=7)
    t2 = time.perf_counter()
    Qh, _ = abi.harvest(1 << 20, seed=4242, steps=200, amps=(10.0, 1.0), augment=0.5, vmax=0.99)
    # exact |v| of p = e/3: the root of M v^2 - 4 Q_t v + 3 M = 0, as
    # 3x / (2 + sqrt(4 - 3x^2)) with x = M / Q_t (no cancellation)
    x = np.minimum(np.linalg.norm(Qh[:, 1:4], axis=1) / Qh[:, 0], 0.9999)
    v_ex = 3 * x / (2 + np.sqrt(4 - 3 * x * x))
    g = abi.mlp_forward(p, Qh)[:, 0]
    stats = {}
    for name, start in (("mlp_start", abi.START_MLP), ("v05_start", abi.START_HALF)):
        _, it, _ = abi.c2p_learned(p, Qh, start)
        stats[name] = dict(mean_updates=round(float(np.mean(np.where(it >= 0, it, 0))), 4),
                           fallback_rate=float(np.mean(it < 0)))
    _, rit, _ = abi.c2p(Qh)
    _MODEL["p"] = p
    _MODEL["info"] = dict(pairs=C2P_PAIRS, harvest_s=round(t1 - t0, 2),
                          train_s=round(t2 - t1, 2), epochs=C2P_EPOCHS, batch=512,
                          loss_per_epoch=[float(x) for x in hist],
                          guess_rms_v=float(np.sqrt(np.mean((g - v_ex) ** 2))),
                          held_out=int(len(Qh)), ablation=stats,
                          reference_iterations_dv_1e5=round(float(rit.mean()), 4))
    print(f"[bench] con2prim MLP: {_MODEL['info']}", file=sys.stderr, flush=True)
    return p


def set_c2p(ctx, c2p):
    from vhlle_amd import abi
    cm = {"reference": abi.C2P_REFERENCE, "closed": abi.C2P_CLOSED,
          "learned": abi.C2P_LEARNED, "newton": abi.C2P_NEWTON}[c2p]
    if cm != abi.C2P_REFERENCE:
        ctx.set_c2p(cm, c2p_model() if cm == abi.C2P_LEARNED else None)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def run_variant(args, cfg, prec, c2p, ic):
    from vhlle_amd import abi
    nx, ny, nz = cfg["nx"], cfg["ny"], cfg["nz"]
    trc = abi.make_trc(etaS=cfg["etaS"])
    p = abi.PREC_FP64 if prec == "fp64" else abi.PREC_MIXED
    ctx = abi.Context(nx, ny, nz, cfg["xlim"], cfg["ylim"], cfg["zlim"], trc=trc, precision=p)
    set_c2p(ctx, c2p)
    dt, tau0 = cfg["dt"], cfg["tau0"]
    ctx.set_prim(tau0, *ic)
    if args.warmup:
        ctx.run(tau0, dt, args.warmup)  # also captures the step graphs
    ctx.set_prim(tau0, *ic)
    t0 = time.perf_counter()
    st, tau = ctx.run(tau0, dt, args.steps)  # synchronises before returning
    elapsed = time.perf_counter() - t0
    ctx.set_prim(tau0, *ic)
    ctx.enable_timing(True)
    st2, _ = ctx.run(tau0, dt, args.steps)
    kt = ctx.kernel_times()
    ctx.enable_timing(False)
    # the whole trajectory of the workload (tau 0.5 -> 8 fm), as developed
    # flow makes con2prim and the limiter cost more than the first steps
    traj = None
    if not args.no_traj:
        ctx.set_prim(tau0, *ic)
        t0 = time.perf_counter()
        st3, _ = ctx.run(tau0, dt, cfg["steps"])
        traj = dict(steps=cfg["steps"], seconds=round(time.perf_counter() - t0, 4), status=st3)
        st2 |= st3
    stage = None
    if p == abi.PREC_MIXED:
        k, cz, ms = ctx.get_stage()
        stage = dict(kernel={1: "cells", 2: "rows", 3: "bands"}.get(k, "auto"), CZ=cz,
                     CY=ctx.get_stage_bands(), tuned_pred_corr_ms=round(ms, 4))
    ctx.close()
    print(f"[bench] {prec}/{c2p}: {args.steps} steps in {elapsed:.3f} s, status {st | st2}",
          file=sys.stderr, flush=True)
    return dict(elapsed=elapsed, status=st | st2, kernel_times=kt, tau=tau, stage=stage,
                traj=traj)


def error_run(cfg, prec, c2p, ic, horizons, trc=None):
    """{h: (Q[:4], pi, Pi, tau)} after h steps from tau0 for each horizon h
    (for the error comparison)."""
    from vhlle_amd import abi
    p = abi.PREC_FP64 if prec == "fp64" else abi.PREC_MIXED
    ctx = abi.Context(cfg["nx"], cfg["ny"], cfg["nz"], cfg["xlim"], cfg["ylim"], cfg["zlim"],
                      trc=trc or abi.make_trc(etaS=cfg["etaS"]), precision=p)
    set_c2p(ctx, c2p)
    ctx.set_prim(cfg["tau0"], *ic)
    out, tau, done = {}, cfg["tau0"], 0
    for h in horizons:
        st, tau = ctx.run(tau, cfg["dt"], h - done)
        done = h
        s = ctx.download()
        out[h] = dict(Q=s["Q"][:4].copy(), pi=s["pi"].copy(), Pi=s["Pi"].copy(), tau=tau,
                      status=st)
        del s
    ctx.close()
    return out


def perturbed(ic, eps, seed=12345):
    rng = np.random.default_rng(seed)
    return (ic[0] * (1 + eps * rng.uniform(-1, 1, ic[0].shape)),) + tuple(ic[1:])


def err_mask(state):
    from vhlle_amd import abi
    q = np.zeros((state["Q"].shape[1], 7))
    q[:, :4] = state["Q"].T / state["tau"]
    return abi.c2p(q)[0][:, 0] > E_CRIT


def state_errors(a, b, mask, with_bulk=False):
    qs = np.max(np.abs(b["Q"][0]))
    eT = float(np.max(np.abs(a["Q"][:4, mask] - b["Q"][:4, mask])) / qs)
    ps = max(float(np.max(np.abs(b["pi"][:, mask]))), 1e-300)
    epi = float(np.max(np.abs(a["pi"][:, mask] - b["pi"][:, mask])) / ps)
    if not with_bulk:
        return eT, epi
    bs = max(float(np.max(np.abs(b["Pi"][:, mask]))), 1e-300)
    ePi = float(np.max(np.abs(a["Pi"][:, mask] - b["Pi"][:, mask])) / bs)
    return eT, epi, ePi


def roofline(kt, N, prec, stage=None, c2p=None):
    name, (launches, ms) = max(kt.items(), key=lambda kv: kv[1][1])
    avg_s = ms / 1e3 / launches
    alg = ALG_BYTES[prec].get(name, 0) * N
    achieved = alg / avg_s / 1e9
    tr = traffic_of(prec, name, stage, c2p)
    return dict(bound="hbm", achieved=round(achieved, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 5), traffic=tr, kernel=name,
                avg_ms=round(avg_s * 1e3, 4), alg_bytes_per_launch=alg,
                traffic_GBs=round(tr / avg_s / 1e9, 1) if tr else None)


def _oracle_slab(cfg, nz, zc, steps):
    """Seconds the oracle takes for `steps` steps of a 200 x 200 x nz slab of
    the workload grid (same spacings) centred at eta = zc."""
    from oracle import pyoracle
    from vhlle_amd import setups
    half = (nz - 1) / 2 * 0.05
    zl = (zc - half, zc + half)
    nx, ny = cfg["nx"], cfg["ny"]
    e, vx, vy, vz = setups.gaussian3d_ic(nx, ny, nz, cfg["xlim"], cfg["ylim"], zl, e0=cfg["e0"],
                                         sx=cfg["sx"], sy=cfg["sy"],
                                         eta_plateau=cfg["eta_plateau"],
                                         eta_sigma=cfg["eta_sigma"])
    s = pyoracle.Solver(nx, ny, nz, cfg["xlim"], cfg["ylim"], zl,
                        trc=pyoracle.make_trc(etaS=cfg["etaS"]))
    s.set_prim(cfg["tau0"], e, vx, vy, vz)
    t0 = time.perf_counter()
    s.run(cfg["tau0"], cfg["dt"], steps)
    return time.perf_counter() - t0


def _slab_worker(a):
    cfg, nz, zc, steps = a
    return _oracle_slab(cfg, nz, zc, steps)


def cpu_baseline(cfg):
    """The oracle (a serial CPU restatement of the reference, which is serial
    itself) on a bounded sample of the same workload: a 200x200x10 slab of the
    grid (same spacings, centre of the eta plateau) for 5 steps on one core
    (~10 s); then every usable core (at most 16, the host share of one GPU)
    running its own slab of the grid at once (a domain decomposition without
    halo exchange: an upper bound of what the serial code gains from the
    cores).  Runs in worker processes started before any GPU call."""
    import multiprocessing as mp
    nz, steps = 10, 5
    dt1 = _oracle_slab(cfg, nz, 0.0, steps)
    cells = cfg["nx"] * cfg["ny"] * nz
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count() or 1
    ncpu = max(1, min(16, ncpu))
    zcs = [(k - (ncpu - 1) / 2) * nz * 0.05 for k in range(ncpu)]  # slabs side by side in eta
    t0 = time.perf_counter()
    with mp.get_context("spawn").Pool(ncpu) as pool:
        per = pool.map(_slab_worker, [(cfg, nz, zc, steps) for zc in zcs])
    wall = time.perf_counter() - t0
    return dict(value=round(cells * steps / dt1, 1), unit="cell-updates/s", cores=1,
                kind="port", sample=f"oracle, {cfg['nx']}x{cfg['ny']}x{nz} slab of the workload "
                f"grid, {steps} steps, {dt1:.1f} s on 1 core of {cpu_model()}",
                all_cores=dict(value=round(ncpu * cells * steps / wall, 1), cores=ncpu,
                               sample=f"{ncpu} processes, each its own {cfg['nx']}x{cfg['ny']}x"
                               f"{nz} slab (eta centres {zcs[0]:.2f}..{zcs[-1]:.2f}), {steps} "
                               f"steps; wall {wall:.1f} s including process start, slowest "
                               f"slab {max(per):.1f} s"))


def batch_bench(args, c2p):
    """Event-by-event batch: 256 boost-invariant grids of sides 100..160 with
    hot-spot initial states, all advanced by the same kernel launches (one HIP
    graph per step), against the same events run one context at a time."""
    from vhlle_amd import abi, setups
    cfg = setups.BATCH_EBE
    grids = setups.batch_grids()
    e, npart = abi.hotspot_ic(grids, seed=cfg["seed"], npart_mean=cfg["npart_mean"], R=cfg["R"],
                              a=cfg["a"], sigma=cfg["sigma"], amp=cfg["amp"])
    z = np.zeros_like(e)
    N = len(e)
    steps = cfg["steps"]
    ctx = abi.Context(grids=grids, trc=setups.batch_trc(), precision=abi.PREC_MIXED)
    set_c2p(ctx, c2p)
    ctx.set_prim(cfg["tau0"], e, z, z, z)
    ctx.run(cfg["tau0"], cfg["dt"], 3)
    ctx.set_prim(cfg["tau0"], e, z, z, z)
    t0 = time.perf_counter()
    st, _ = ctx.run(cfg["tau0"], cfg["dt"], steps)
    el = time.perf_counter() - t0
    k, cz, _ = ctx.get_stage()
    cy = ctx.get_stage_bands()
    ctx.close()
    # the first 16 events one context each, same steps
    nseq = 16
    off, seq_cells, seq_t = 0, 0, 0.0
    for g in grids[:nseq]:
        n = g[0] * g[1] * g[2]
        c = abi.Context(*g, trc=setups.batch_trc(), precision=abi.PREC_MIXED)
        set_c2p(c, c2p)
        c.set_prim(cfg["tau0"], e[off:off + n], z[:n], z[:n], z[:n])
        c.run(cfg["tau0"], cfg["dt"], 3)
        c.set_prim(cfg["tau0"], e[off:off + n], z[:n], z[:n], z[:n])
        t0 = time.perf_counter()
        c.run(cfg["tau0"], cfg["dt"], steps)
        seq_t += time.perf_counter() - t0
        c.close()
        seq_cells += n
        off += n
    v = N * steps / el
    print(f"[bench] batch {c2p}: {len(grids)} events, {N} cells, {steps} steps in {el:.3f} s, "
          f"status {st}", file=sys.stderr, flush=True)
    return dict(workload=cfg["name"], events=len(grids), cells=N, steps=steps, c2p=c2p,
                status=st, value=round(v, 1), unit="cell-updates/s",
                ms_per_step=round(el / steps * 1e3, 4),
                # vh_run replays one graph of 8 steps per 8 steps, single-step
                # graphs for the rest: every event advances in the same launches
                graph_launches_per_step=round((steps // 8 + steps % 8) / steps, 4),
                stage_kernel={1: "cells", 2: "rows", 3: "bands"}.get(k, "auto"), CZ=cz, CY=cy,
                npart_mean=float(npart.mean()), e_max=float(e.max()),
                one_context_per_event=dict(events=nseq, value=round(seq_cells * steps / seq_t, 1),
                                           batch_speedup=round(v / (seq_cells * steps / seq_t), 2)))


def full3d_bench(c2p_mixed="closed"):
    """300 x 300 x 150 hot-spot run, 600 steps, freeze-out every 5 steps
    (snapshot -> 5 steps as one graph replay -> surface), FP64 and FP32
    fluxes: throughput and the surface (elements, sum of dsigma_mu)."""
    from vhlle_amd import abi, setups
    cfg = setups.FULL_3D_FO
    ic, npart = setups.full3d_ic(cfg)
    N = cfg["nx"] * cfg["ny"] * cfg["nz"]
    res = {}
    for name, prec, c2p in (("fp64", abi.PREC_FP64, "reference"),
                            ("mixed_" + c2p_mixed, abi.PREC_MIXED, c2p_mixed)):
        ctx = abi.Context(cfg["nx"], cfg["ny"], cfg["nz"], cfg["xlim"], cfg["ylim"], cfg["zlim"],
                          trc=abi.make_trc(etaS=cfg["etaS"], zetaS=cfg["zetaS"], zetaSparam=5),
                          precision=prec)
        set_c2p(ctx, c2p)
        ctx.set_prim(cfg["tau0"], *ic)
        tau, t_step, t_fo, nel, ds, st_all, nsusp = cfg["tau0"], 0.0, 0.0, 0, np.zeros(4), 0, 0
        for _ in range(cfg["steps"] // cfg["fo_every"]):
            ctx.fo_snapshot(tau)
            t0 = time.perf_counter()
            st, tau = ctx.run(tau, cfg["dt"], cfg["fo_every"])
            t1 = time.perf_counter()
            _, fs = ctx.freezeout(tau, 0.0, cfg["e_sw"], lower=abi.FO_SNAPSHOT, max_rows=0)
            t2 = time.perf_counter()
            t_step += t1 - t0
            t_fo += t2 - t1
            st_all |= st
            nel += fs["nelements"]
            nsusp += fs["nsusp"]
            ds += np.array(fs["dsigma"])
        ctx.close()
        v = N * cfg["steps"] / (t_step + t_fo)
        res[name] = dict(value=round(v, 1), unit="cell-updates/s", status=st_all,
                         ms_per_step=round(t_step / cfg["steps"] * 1e3, 4),
                         fo_ms_per_call=round(t_fo / (cfg["steps"] // cfg["fo_every"]) * 1e3, 3),
                         elements=int(nel), suspicious=int(nsusp),
                         dsigma_sum=[float(x) for x in ds])
        print(f"[bench] full 3-D {name}: {cfg['steps']} steps {t_step:.2f} s, freeze-out "
              f"{t_fo:.2f} s, {nel} elements, status {st_all}", file=sys.stderr, flush=True)
    a, b = res["fp64"], res["mixed_" + c2p_mixed]
    da, db = np.array(a["dsigma_sum"]), np.array(b["dsigma_sum"])
    # T^{tau mu}, pi and Pi of the mixed variants against FP64 after 5 steps
    # (where a 6e-8 change of the FP64 initial state still moves T^{tau mu} by
    # less than 1e-7; after 10 steps it moves it by 2e-5), with both floors
    trc = abi.make_trc(etaS=cfg["etaS"], zetaS=cfg["zetaS"], zetaSparam=5)
    h = (5,)
    ref = error_run(cfg, "fp64", "reference", ic, h, trc)[5]
    mask = err_mask(ref)
    errs = {}
    for label, prec, c2p, init in (("fp64_ic_x(1+1e-12)", "fp64", "reference", perturbed(ic, 1e-12)),
                                   ("fp64_ic_x(1+6e-8)", "fp64", "reference", perturbed(ic, 6e-8)),
                                   ("mixed_closed", "mixed", "closed", ic),
                                   ("mixed_learned", "mixed", "learned", ic)):
        r = error_run(cfg, prec, c2p, init, h, trc)[5]
        errs[label] = dict(zip(("T", "pi", "Pi"), state_errors(r, ref, mask, with_bulk=True)))
        del r
    print(f"[bench] full 3-D errors after 5 steps: {errs}", file=sys.stderr, flush=True)
    return dict(workload=cfg["name"], cells=N, steps=cfg["steps"], fo_every=cfg["fo_every"],
                e_sw=cfg["e_sw"], npart=npart, variants=res,
                elements_rel_diff=abs(b["elements"] - a["elements"]) / max(a["elements"], 1),
                dsigma_rel_diff=[float(x) for x in np.abs(db - da) / np.abs(da).max()],
                errors_after_5_steps=dict(region="e_fp64 > 0.5 GeV/fm^3", cells=int(mask.sum()),
                                          vs_fp64=errs))


def table_eos_bench(args):
    """The 3-D viscous workload (200x200x100) with the crossover EoS table
    (p/e down to 0.10 near T = 170 MeV; no closed-form con2prim): FP64 with the
    reference iteration against mixed precision with the FP32 Newton from
    v = 0.5 and from the learned guess (an MLP trained on the GPU on states of
    the table), the Newton-update ablation of the guess on held-out states,
    and the T^{tau mu} error of each mixed run against FP64 after the timed
    steps (over e_fp64 > 0.5)."""
    from vhlle_amd import abi, setups
    cfg = setups.VISCOUS_3D
    table = setups.crossover_eos()
    N = cfg["nx"] * cfg["ny"] * cfg["nz"]
    steps = min(args.steps, 100)
    ic = setups.gaussian3d_ic(cfg["nx"], cfg["ny"], cfg["nz"], cfg["xlim"], cfg["ylim"],
                              cfg["zlim"], e0=cfg["e0"], sx=cfg["sx"], sy=cfg["sy"],
                              eta_plateau=cfg["eta_plateau"], eta_sigma=cfg["eta_sigma"])
    t0 = time.perf_counter()
    Q, prim = abi.harvest(C2P_PAIRS, seed=2024, steps=200, amps=(10.0, 1.0), augment=0.5,
                          vmax=0.99, eos=table)
    params, hist = abi.mlp_train(abi.mlp_init(7), Q, prim, batch=512, epochs=C2P_EPOCHS,
                                 lr=2e-3, lr_final=1e-6, seed=7)
    t_train = time.perf_counter() - t0
    Qh, _ = abi.harvest(1 << 18, seed=4242, steps=200, amps=(10.0, 1.0), augment=0.5,
                        vmax=0.99, eos=table)
    ablation = {}
    for name, start in (("mlp_start", abi.START_MLP), ("v05_start", abi.START_HALF)):
        _, it, _ = abi.c2p_learned(params, Qh, start, eos=table)
        ablation[name] = dict(mean_updates=round(float(np.mean(np.where(it >= 0, it, 0))), 4),
                              fallback_rate=float(np.mean(it < 0)))
    runs = {}
    for name, prec, c2p in (("fp64", abi.PREC_FP64, abi.C2P_REFERENCE),
                            ("mixed_newton", abi.PREC_MIXED, abi.C2P_NEWTON),
                            ("mixed_learned", abi.PREC_MIXED, abi.C2P_LEARNED)):
        ctx = abi.Context(cfg["nx"], cfg["ny"], cfg["nz"], cfg["xlim"], cfg["ylim"], cfg["zlim"],
                          trc=abi.make_trc(etaS=cfg["etaS"]), precision=prec, eos=table)
        if c2p != abi.C2P_REFERENCE:
            ctx.set_c2p(c2p, params if c2p == abi.C2P_LEARNED else None)
        ctx.set_prim(cfg["tau0"], *ic)
        ctx.run(cfg["tau0"], cfg["dt"], 3)
        ctx.set_prim(cfg["tau0"], *ic)
        t0 = time.perf_counter()
        st, tau = ctx.run(cfg["tau0"], cfg["dt"], steps)
        el = time.perf_counter() - t0
        runs[name] = dict(state=ctx.download(), elapsed=el, status=st, tau=tau)
        ctx.close()
        print(f"[bench] table EoS {name}: {steps} steps in {el:.3f} s, status {st}",
              file=sys.stderr, flush=True)
    ref = runs["fp64"]
    e_ref = abi.c2p((ref["state"]["Q"].T / ref["tau"]).copy(), eos=table)[0][:, 0]
    mask = e_ref > E_CRIT
    out = dict(workload="viscous_3d_200x200x100 with the crossover EoS table (setups.CROSSOVER_EOS)",
               cells=N, steps=steps, train_s=round(t_train, 2), train_loss=[float(h) for h in hist],
               ablation=ablation, variants={})
    for name, r in runs.items():
        v = N * steps / r["elapsed"]
        d = dict(value=round(v, 1), unit="cell-updates/s",
                 ms_per_step=round(r["elapsed"] / steps * 1e3, 4), status=r["status"],
                 speedup_vs_fp64=round(ref["elapsed"] / r["elapsed"], 3))
        if name != "fp64":
            d["err_T_vs_fp64"], d["err_pi_vs_fp64"] = state_errors(r["state"], ref["state"], mask)
        out["variants"][name] = d
    return out


def vorticity_bench(args):
    """The 3-D viscous workload (200x200x100, FP64) with and without the
    vorticity output (dbeta formed in every step), and the freeze-out with the
    beta.dat records: the first call after the steps also resolves the entries
    the reference inherits from earlier cells; a second call on the same state
    shows the records alone."""
    from vhlle_amd import abi, setups
    cfg = setups.VISCOUS_3D
    N = cfg["nx"] * cfg["ny"] * cfg["nz"]
    steps = min(args.steps, 100)
    ic = setups.gaussian3d_ic(cfg["nx"], cfg["ny"], cfg["nz"], cfg["xlim"], cfg["ylim"],
                              cfg["zlim"], e0=cfg["e0"], sx=cfg["sx"], sy=cfg["sy"],
                              eta_plateau=cfg["eta_plateau"], eta_sigma=cfg["eta_sigma"])
    out = dict(workload="viscous_3d_200x200x100, FP64", cells=N, steps=steps)
    for on in (False, True):
        ctx = abi.Context(cfg["nx"], cfg["ny"], cfg["nz"], cfg["xlim"], cfg["ylim"], cfg["zlim"],
                          trc=abi.make_trc(etaS=cfg["etaS"]))
        ctx.set_vorticity(on)
        ctx.set_prim(cfg["tau0"], *ic)
        ctx.run(cfg["tau0"], cfg["dt"], 3)
        ctx.set_prim(cfg["tau0"], *ic)
        t0 = time.perf_counter()
        st, tau = ctx.run(cfg["tau0"], cfg["dt"], steps)
        el = time.perf_counter() - t0
        key = "vorticity_on" if on else "vorticity_off"
        out[key] = dict(ms_per_step=round(el / steps * 1e3, 4), status=st)
        if on:
            t0 = time.perf_counter()
            rows, _ = ctx.freezeout(tau, cfg["dt"], E_CRIT)
            t_fo = time.perf_counter() - t0
            t0 = time.perf_counter()
            _, beta = ctx.freezeout_beta(tau, cfg["dt"], E_CRIT)
            t_first = time.perf_counter() - t0
            t0 = time.perf_counter()
            ctx.freezeout_beta(tau, cfg["dt"], E_CRIT)
            t_again = time.perf_counter() - t0
            out["freezeout"] = dict(elements=len(rows), ms_records=round(t_fo * 1e3, 3),
                                    ms_with_beta_first=round(t_first * 1e3, 3),
                                    ms_with_beta_again=round(t_again * 1e3, 3),
                                    max_abs_dbeta_cart=float(np.max(np.abs(beta[:, 16:32])))
                                    if len(beta) else 0.0)
        ctx.close()
        print(f"[bench] vorticity {key}: {out[key]}", file=sys.stderr, flush=True)
    out["step_overhead"] = round(out["vorticity_on"]["ms_per_step"] /
                                 out["vorticity_off"]["ms_per_step"] - 1.0, 4)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--variants", default="fp64,mixed_ref,mixed_closed,mixed_newton,mixed_learned")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-full3d", action="store_true")
    ap.add_argument("--no-table", action="store_true")
    ap.add_argument("--no-vort", action="store_true")
    ap.add_argument("--no-traj", action="store_true",
                    help="skip timing the whole 300-step trajectory of each variant")
    ap.add_argument("--quick", action="store_true",
                    help="skip the error runs (profiling); errors reported as null")
    args = ap.parse_args()
    if args.quick:  # a small training set keeps profiler runs short
        global C2P_PAIRS, C2P_EPOCHS
        C2P_PAIRS, C2P_EPOCHS = 1 << 18, 2
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; multi-GPU sharding of the hydro "
              "grid (x-slabs with halo exchanges) is not built.", file=sys.stderr)
        sys.exit(2)
    from vhlle_amd import setups
    cfg = setups.VISCOUS_3D
    # the CPU baseline first: its worker processes start before the GPU is used
    cpu = None if args.no_cpu else cpu_baseline(cfg)
    if cpu:
        print(f"[bench] CPU baseline: {cpu}", file=sys.stderr, flush=True)
    from vhlle_amd import abi
    N = cfg["nx"] * cfg["ny"] * cfg["nz"]
    ic = setups.gaussian3d_ic(cfg["nx"], cfg["ny"], cfg["nz"], cfg["xlim"], cfg["ylim"],
                              cfg["zlim"], e0=cfg["e0"], sx=cfg["sx"], sy=cfg["sy"],
                              eta_plateau=cfg["eta_plateau"], eta_sigma=cfg["eta_sigma"])
    print(f"[bench] device: {abi.device_info()}", file=sys.stderr, flush=True)
    copy_gbs = abi.copy_bandwidth(1 << 30, 10)
    print(f"[bench] plain copy of 1 GiB: {copy_gbs:.0f} GB/s", file=sys.stderr, flush=True)
    names = [v for v in args.variants.split(",") if v]
    if "fp64" not in names:
        names.insert(0, "fp64")
    results = {}
    for name in names:
        prec, c2p = VARIANTS[name]
        results[name] = run_variant(args, cfg, prec, c2p, ic)
    # error runs: every variant against FP64 after each horizon, beside the
    # floors (FP64 from an initial state changed by 1e-12 and by 6e-8)
    errs = {n: {} for n in names}
    floors = {}
    if not args.quick:
        ref = error_run(cfg, "fp64", "reference", ic, ERR_HORIZONS)
        masks = {h: err_mask(ref[h]) for h in ERR_HORIZONS}
        for eps in (1e-12, 6e-8):
            pert = error_run(cfg, "fp64", "reference", perturbed(ic, eps), ERR_HORIZONS)
            floors[f"fp64_ic_x(1+{eps:g})"] = {
                h: dict(zip(("T", "pi"), state_errors(pert[h], ref[h], masks[h])))
                for h in ERR_HORIZONS}
            del pert
        for name in names:
            prec, c2p = VARIANTS[name]
            if name == "fp64":
                continue
            r = error_run(cfg, prec, c2p, ic, ERR_HORIZONS)
            errs[name] = {h: dict(zip(("T", "pi"), state_errors(r[h], ref[h], masks[h])))
                          for h in ERR_HORIZONS}
            del r
        tau_e = ref[ERR_STEPS]["tau"]
        ncells = int(masks[ERR_STEPS].sum())
        del ref
    else:
        tau_e, ncells = None, 0
    variants = {}
    for name, r in results.items():
        prec, c2p = VARIANTS[name]
        v = N * args.steps / r["elapsed"]
        e = errs[name].get(ERR_STEPS, {}) if name != "fp64" else {"T": 0.0, "pi": 0.0}
        variants[name] = dict(
            value=round(v, 1), ms_per_step=round(r["elapsed"] / args.steps * 1e3, 4),
            precision=prec, c2p=c2p, status=r["status"],
            err_T_vs_fp64=e.get("T"), err_pi_vs_fp64=e.get("pi"),
            err_vs_fp64_by_steps={str(h): d for h, d in errs[name].items()},
            step_hbm_frac=round(STEP_BYTES[prec] * v / (HBM_PEAK_GBS * 1e9), 5),
            speedup_vs_fp64=round(v / (N * args.steps / results["fp64"]["elapsed"]), 3),
            roofline=roofline(r["kernel_times"], N, prec,
                              r["stage"]["kernel"] if r["stage"] else None, c2p),
            kernels_ms={k: round(t[1], 3) for k, t in sorted(r["kernel_times"].items())})
        if r["traj"]:
            t = r["traj"]
            variants[name]["trajectory"] = dict(
                steps=t["steps"], ms_per_step=round(t["seconds"] / t["steps"] * 1e3, 4),
                value=round(N * t["steps"] / t["seconds"], 1), status=t["status"])
        if r["stage"]:
            variants[name]["stage_kernel"] = r["stage"]
    # accepted: T^{tau mu} error after ERR_STEPS steps within 1e-6 (strict)
    ok = [n for n in variants if variants[n]["status"] == 0 and
          (args.quick or (variants[n]["err_T_vs_fp64"] is not None and
                          variants[n]["err_T_vs_fp64"] <= ERR_BAR))]
    if "mixed_learned" in ok:
        head = "mixed_learned"
    else:
        head = max(ok, key=lambda n: variants[n]["value"]) if ok else "fp64"
    H = variants[head]
    out = {
        "metric": METRIC,
        "value": H["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": H["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if H["precision"] == "fp64" else "f64+f32",
        "data": "synthetic",
        "config": {"workload": f"{cfg['name']}: 3-D viscous, dx=dy=0.1, deta=0.05, tau0=0.5, "
                               f"dtau=0.025, eta/s=0.08, p=e/3",
                   "cells": N, "variant": head, "precision": H["precision"], "c2p": H["c2p"]},
        "hbm_frac_step": H["step_hbm_frac"],
        "hbm_copy_GBs": round(copy_gbs, 1),
        "error_measure": {"steps": ERR_STEPS, "tau": tau_e and round(tau_e, 4),
                          "region": "e_fp64 > 0.5 GeV/fm^3", "cells": ncells,
                          "bar_T": ERR_BAR, "horizons": list(ERR_HORIZONS),
                          "floors": {k: {str(h): d for h, d in v.items()}
                                     for k, v in floors.items()}},
        "roofline": H["roofline"],
        "variants": variants,
    }
    if "info" in _MODEL:
        out["learned_c2p"] = _MODEL["info"]
    if not args.no_batch:
        out["batch_ebe"] = [batch_bench(args, c) for c in ("closed", "learned")]
    if not args.no_full3d:
        out["full_3d_freezeout"] = full3d_bench("closed")
    if not args.no_table:
        out["table_eos"] = table_eos_bench(args)
    if not args.no_vort:
        out["vorticity"] = vorticity_bench(args)
    out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
