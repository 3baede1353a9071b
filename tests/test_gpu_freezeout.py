"""Freeze-out hypersurface on the GPU against the oracle (oracle/fo_oracle.cpp):

* the Cornelius construction on random and structured hypercubes equals the
  oracle's (element counts, centroids, normals);
* the element records of a viscous 3-D state (FP64 path) equal the oracle's
  records of the same state;
* the snapshot mode and the diagnostics agree with the records;
* FP32 fluxes against the FP64 step: element count and sum of dsigma_mu.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from vhlle_amd import abi, setups

pytestmark = pytest.mark.gpu

DX = np.array([0.05, 0.1, 0.12, 0.08])


def test_cornelius_matches_oracle(gpu):
    rng = np.random.default_rng(17)
    h = rng.uniform(0.0, 1.0, (20000, 2, 2, 2, 2))
    # corners exactly at the level, and near-degenerate ones
    h[:2000][rng.uniform(size=(2000, 2, 2, 2, 2)) < 0.2] = 0.5
    h[2000:4000] = 0.5 + 1e-12 * rng.normal(size=(2000, 2, 2, 2, 2))
    ng, cg, mg = abi.cornelius(h, DX, 0.5)
    no, co, mo = po.cornelius(h, DX, 0.5)
    assert np.array_equal(ng, no)
    print(f"elements: {np.bincount(np.maximum(ng, 0))}, failures {np.sum(ng < 0)}", flush=True)
    assert np.array_equal(cg, co) and np.array_equal(mg, mo)


def viscous_state(prec, steps, nx=26, ny=24, nz=14, c2p=abi.C2P_REFERENCE):
    xl, yl, zl = (-2.5, 2.5), (-2.3, 2.3), (-0.65, 0.65)
    e, vx, vy, vz = setups.gaussian3d_ic(nx, ny, nz, xl, yl, zl, e0=8.0, sx=1.0, sy=1.2,
                                         eta_plateau=0.3, eta_sigma=0.3)
    ctx = abi.Context(nx, ny, nz, xl, yl, zl, trc=abi.make_trc(etaS=0.1),
                      precision=prec)
    if c2p != abi.C2P_REFERENCE:
        ctx.set_c2p(c2p)
    ctx.set_prim(0.6, e, vx, vy, vz)
    st, tau = ctx.run(0.6, 0.02, steps)
    assert st == 0
    return ctx, tau, ((nx, ny, nz), ((xl[1] - xl[0]) / (nx - 1), (yl[1] - yl[0]) / (ny - 1),
                                     (zl[1] - zl[0]) / (nz - 1)), (xl[0], yl[0], zl[0]))


def test_records_match_oracle(gpu):
    ctx, tau, (shape, spacing, mins) = viscous_state(abi.PREC_FP64, 12)
    rows, stats = ctx.freezeout(tau, 0.02, 0.5, extend=True)
    s = ctx.download()
    ctx.close()
    ref = po.freezeout(shape, spacing, mins, s["Q"], s["Qprev"], s["pi"], s["Pi"], tau, 0.02,
                       0.5, extend=True)
    print(f"{len(rows)} elements, {stats}", flush=True)
    assert len(rows) == len(ref) and len(rows) > 100 and stats["err"] == 0
    # geometry columns bit for bit; libm-dependent columns (cosh, sinh,
    # tanh, log, pow) to a few ulp
    assert np.array_equal(rows[:, :4], ref[:, :4])
    scale = np.maximum(np.abs(ref).max(axis=0), 1e-300)
    assert np.all(np.abs(rows - ref) <= 1e-12 * scale)
    assert abs(stats["dsigma"][0] - rows[:, 4].sum()) <= 1e-12 * np.abs(rows[:, 4]).sum()


def test_snapshot_mode(gpu):
    ctx, tau, _ = viscous_state(abi.PREC_FP64, 6)
    ctx.fo_snapshot(tau)
    st, tau2 = ctx.run(tau, 0.02, 5)
    rows, stats = ctx.freezeout(tau2, 0.0, 0.5, lower=abi.FO_SNAPSHOT)
    ctx.close()
    assert st == 0 and len(rows) > 50 and stats["nelements"] == len(rows)
    # centroid times lie within the time edge above tau2 (the reference's
    # tau + c0 convention)
    assert np.all(rows[:, 0] >= tau2 - 1e-12) and np.all(rows[:, 0] <= tau2 + (tau2 - tau) + 1e-12)


def test_fp32_fluxes_surface(gpu):
    out = {}
    for name, prec, c2p in (("fp64", abi.PREC_FP64, abi.C2P_REFERENCE),
                            ("mixed", abi.PREC_MIXED, abi.C2P_CLOSED)):
        ctx, tau, _ = viscous_state(prec, 20, c2p=c2p)
        rows, stats = ctx.freezeout(tau, 0.02, 0.5)
        ctx.close()
        out[name] = (len(rows), np.array(stats["dsigma"]))
    n64, d64 = out["fp64"]
    n32, d32 = out["mixed"]
    rel = np.abs(d32 - d64) / np.abs(d64).max()
    print(f"elements fp64 {n64} mixed {n32}; sum dsigma rel diff {rel}", flush=True)
    assert abs(n32 - n64) <= max(2, 0.01 * n64)
    assert np.all(rel < 1e-3)


def test_corona_matches_oracle(gpu):
    """Fluid::outputCorona's elements of an initial state (FP64 path)"""
    nx, ny, nz = 26, 24, 14
    xl, yl, zl = (-2.5, 2.5), (-2.3, 2.3), (-0.65, 0.65)
    e, vx, vy, vz = setups.gaussian3d_ic(nx, ny, nz, xl, yl, zl, e0=8.0, sx=1.0, sy=1.2,
                                         eta_plateau=0.3, eta_sigma=0.3)
    ctx = abi.Context(nx, ny, nz, xl, yl, zl, trc=abi.make_trc(etaS=0.1))
    ctx.set_prim(0.6, e, vx, vy, vz)
    rows, stats = ctx.corona(0.6, 0.5, extend=True)
    s = ctx.download()
    ctx.close()
    sp = ((xl[1] - xl[0]) / (nx - 1), (yl[1] - yl[0]) / (ny - 1), (zl[1] - zl[0]) / (nz - 1))
    ref, rst = po.corona((nx, ny, nz), sp, (xl[0], yl[0], zl[0]), s["Q"], s["pi"], s["Pi"],
                         0.6, 0.5, extend=True)
    print(f"{len(rows)} corona elements", flush=True)
    assert len(rows) == len(ref) == stats["nelements"] > 100
    assert np.array_equal(rows[:, :4], ref[:, :4])
    assert np.allclose(rows, ref, rtol=1e-13, atol=1e-15)
    assert stats["v_eff"] == pytest.approx(rst[1], rel=1e-12)
    assert stats["e_surf"] == pytest.approx(rst[2], rel=1e-12)
