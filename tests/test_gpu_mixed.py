"""Mixed precision (FP32 fluxes / viscous sources / pi storage, FP64 Q and
con2prim) against the FP64 faithful path, through the C-ABI.

Error measures (target: <= 1e-6 relative error in T^{mu nu}), over the cells
where the FP64 run has e > e_crit = 0.5 GeV/fm^3 (the reference's default
switching energy density, src/main.cpp:75), i.e. inside the freeze-out
surface:
  T^{tau mu}:  max |Q_mixed - Q_fp64| / max |Q^tau_fp64|
  pi:          max |pi_mixed - pi_fp64| / max |pi_fp64|
Outside that region the reference scheme itself is chaotic: there T < 0.1 GeV
switches tau_pi to 0 and the viscous limiter and the VMIN/UDIFF switches of
NSquant are discontinuous, so a 1e-12 perturbation of the FP64 initial state
grows to O(1) differences in the dilute corona.  test_fp64_corona_chaos shows
it.
"""
import numpy as np
import pytest

from vhlle_amd import abi, setups

pytestmark = pytest.mark.gpu


def run(prec, c2p, nx, ny, nz, xl, yl, zl, ic, tau0, dt, steps, etaS):
    ctx = abi.Context(nx, ny, nz, xl, yl, zl, trc=abi.make_trc(etaS=etaS), precision=prec)
    if c2p != abi.C2P_REFERENCE:
        abi.check(abi.lib().vh_set_c2p_mode(ctx.h, c2p), "vh_set_c2p_mode")
    e, vx, vy, vz = ic
    ctx.set_prim(tau0, e, vx, vy, vz)
    st, tau = ctx.run(tau0, dt, steps)
    assert st == 0
    s = ctx.download()
    s["tau"] = tau
    ctx.close()
    return s


E_CRIT = 0.5


def errors(a, b, tau, ecrit=E_CRIT):
    """(T error, pi error) of state a against the FP64 state b, over e_b > ecrit."""
    e = abi.c2p((b["Q"].T / tau).copy())[0][:, 0]
    m = e > ecrit
    assert m.sum() > 100
    qs = np.max(np.abs(b["Q"][0]))
    eT = np.max(np.abs(a["Q"][:4, m] - b["Q"][:4, m])) / qs
    ps = max(np.max(np.abs(b["pi"][:, m])), 1e-300)
    epi = np.max(np.abs(a["pi"][:, m] - b["pi"][:, m])) / ps
    return eT, epi


@pytest.mark.parametrize("c2p", [abi.C2P_REFERENCE, abi.C2P_CLOSED])
def test_mixed_viscous_error_budget(gpu, c2p):
    """3-D viscous workload of the benchmark at 1/4 resolution per axis
    (50x50x25, same physical box), 50 steps to tau = 3 fm."""
    cfg = setups.VISCOUS_3D
    nx, ny, nz = 50, 50, 25
    ic = setups.gaussian3d_ic(nx, ny, nz, cfg["xlim"], cfg["ylim"], cfg["zlim"], e0=cfg["e0"],
                              sx=cfg["sx"], sy=cfg["sy"], eta_plateau=cfg["eta_plateau"],
                              eta_sigma=cfg["eta_sigma"])
    args = (nx, ny, nz, cfg["xlim"], cfg["ylim"], cfg["zlim"], ic, cfg["tau0"], 0.05, 50,
            cfg["etaS"])
    ref = run(abi.PREC_FP64, abi.C2P_REFERENCE, *args)
    mix = run(abi.PREC_MIXED, c2p, *args)
    eT, epi = errors(mix, ref, ref["tau"])
    print(f"c2p={c2p}: T^tau-mu rel err {eT:.3e}, pi rel err {epi:.3e} (e > 0.5)", flush=True)
    assert eT < 1e-6 and epi < 1e-4


def test_fp64_corona_chaos(gpu):
    """The FP64 step is chaotic in the dilute corona: a 1e-12 relative
    perturbation of the initial energy density changes cells with e < e_crit at
    O(1e-3) of max T^{tau tau} after 40 steps, while cells with e > e_crit stay
    at the 1e-12 level."""
    cfg = setups.VISCOUS_3D
    nx, ny, nz = 50, 50, 25
    e, vx, vy, vz = setups.gaussian3d_ic(nx, ny, nz, cfg["xlim"], cfg["ylim"], cfg["zlim"],
                                         e0=cfg["e0"], sx=cfg["sx"], sy=cfg["sy"],
                                         eta_plateau=cfg["eta_plateau"],
                                         eta_sigma=cfg["eta_sigma"])
    ep = e * (1 + 1e-12 * np.random.default_rng(1).uniform(-1, 1, e.shape))
    base = (nx, ny, nz, cfg["xlim"], cfg["ylim"], cfg["zlim"])
    a = run(abi.PREC_FP64, abi.C2P_REFERENCE, *base, (e, vx, vy, vz), 0.5, 0.05, 40, 0.08)
    b = run(abi.PREC_FP64, abi.C2P_REFERENCE, *base, (ep, vx, vy, vz), 0.5, 0.05, 40, 0.08)
    ein, _ = errors(b, a, a["tau"])
    eall, _ = errors(b, a, a["tau"], ecrit=-1.0)
    print(f"perturbation 1e-12: e>0.5 {ein:.2e}, all cells {eall:.2e}", flush=True)
    assert ein < 1e-9 and eall > 1e-6


def test_mixed_gubser_vs_analytic(gpu):
    """Ideal Gubser flow, 101x101, 200 steps: every variant is as close to the
    analytic solution as the FP64 step is.  (Against each other, mixed+closed
    stays within 2e-4 of FP64, i.e. within the reference con2prim's own
    |dv| < 1e-5 stopping error; mixed+reference is noisier because FP32 flux
    differences flip that iteration's bisection/Newton branches.)"""
    cfg = setups.GUBSER_2D
    n = 101
    e, vx, vy, vz = setups.gubser_ic(cfg, nx=n)
    args = (n, n, 1, cfg["xlim"], cfg["ylim"], cfg["zlim"], (e, vx, vy, vz), 1.0, 0.025, 200, 0.0)
    X, Y, _ = setups.axes(n, n, 1, cfg["xlim"], cfg["ylim"], cfg["zlim"])
    l1 = {}
    states = {}
    for name, prec, cm in (("fp64", abi.PREC_FP64, abi.C2P_REFERENCE),
                           ("mixed_ref", abi.PREC_MIXED, abi.C2P_REFERENCE),
                           ("mixed_closed", abi.PREC_MIXED, abi.C2P_CLOSED)):
        s = run(prec, cm, *args)
        states[name] = s
        ea, vxa, _ = setups.gubser(s["tau"], X, Y, amp=10.0)
        ee = abi.c2p((s["Q"].T / s["tau"]).copy())[0][:, 0]
        m = (np.sqrt(X * X + Y * Y) < 4.0).ravel()
        l1[name] = np.sum(np.abs(ee[m] - ea.ravel()[m])) / np.sum(ea.ravel()[m])
    eC, _ = errors(states["mixed_closed"], states["fp64"], states["fp64"]["tau"], ecrit=0.0)
    print(f"L1 error vs analytic: {l1}; mixed_closed vs fp64 {eC:.2e}", flush=True)
    for k in ("mixed_ref", "mixed_closed"):
        assert abs(l1[k] - l1["fp64"]) < 0.01 * l1["fp64"], l1
    assert eC < 2e-4


def test_mixed_graph_equals_steps(gpu):
    nx, ny, nz = 12, 10, 6
    xl, yl, zl = (-2.75, 2.75), (-2.25, 2.25), (-1.0, 1.0)
    ic = setups.gaussian3d_ic(nx, ny, nz, xl, yl, zl, e0=20.0, sx=1.0, sy=1.2,
                              eta_plateau=0.5, eta_sigma=0.5)
    a = abi.Context(nx, ny, nz, xl, yl, zl, trc=abi.make_trc(etaS=0.1), precision=abi.PREC_MIXED)
    b = abi.Context(nx, ny, nz, xl, yl, zl, trc=abi.make_trc(etaS=0.1), precision=abi.PREC_MIXED)
    for c in (a, b):
        abi.check(abi.lib().vh_set_c2p_mode(c.h, abi.C2P_CLOSED))
        c.set_prim(0.6, *ic)
    ta = tb = 0.6
    for rep in range(3):
        _, ta = a.run(ta, 0.02, 4)
        for _ in range(4):
            _, tb = b.step(tb, 0.02)
        da, db = a.download(), b.download()
        for f in ("Q", "pi", "Pi"):
            assert np.array_equal(da[f], db[f]), (rep, f)


def test_fp64_path_refuses_non_reference_c2p(gpu):
    ctx = abi.Context(8, 8, 1, (-1, 1), (-1, 1), (-1, 1))
    assert abi.lib().vh_set_c2p_mode(ctx.h, abi.C2P_CLOSED) == abi.VH_ERR_ARG


@pytest.mark.parametrize("dims,c2p,kind", [((37, 29, 13), abi.C2P_CLOSED, "ideal"),
                                           ((41, 41, 1), abi.C2P_CLOSED, "ideal"),
                                           ((24, 20, 9), abi.C2P_LEARNED, "ideal"),
                                           ((22, 19, 7), abi.C2P_NEWTON, "table"),
                                           ((19, 23, 8), abi.C2P_NEWTON, "charged")])
def test_tile_shapes_identical(gpu, dims, c2p, kind):
    """The row stage kernel (each x- and eta-interface evaluated once; any eta
    chunk, autotuned or forced) and the band kernel (the y-interfaces inside
    a band evaluated once as well) give bit-identical states to the per-cell
    kernel; with the ideal gas, the spline-table instances (kind "table", the
    crossover EoS) and 7-component states (kind "charged", n_b = 0.05 e)."""
    nx, ny, nz = dims
    xl, yl, zl = (-4.0, 4.0), (-3.5, 3.5), (-1.2, 1.2)
    ic = setups.gaussian3d_ic(nx, ny, nz, xl, yl, zl, e0=30.0, sx=1.5, sy=1.2, eta_plateau=0.4,
                              eta_sigma=0.4)
    eos = setups.crossover_eos() if kind == "table" else None
    nb = 0.05 * ic[0] if kind == "charged" else None
    # (kind, CZ) or (STAGE_BANDS, CZ, CY): bands of 2, 4 and 8 rows, ragged
    # against ny = 29, 41 and 20, with eta chunks of 1, 3 and 13 cells
    shapes = [(abi.STAGE_CELLS, 0), (abi.STAGE_AUTO, 0), (abi.STAGE_ROWS, 1),
              (abi.STAGE_ROWS, 5), (abi.STAGE_ROWS, 32), (abi.STAGE_BANDS, 1, 2),
              (abi.STAGE_BANDS, 3, 4), (abi.STAGE_BANDS, 13, 8)]
    model = abi.mlp_init(9) if c2p == abi.C2P_LEARNED else None
    ref = None
    for sh in shapes:
        ctx = abi.Context(nx, ny, nz, xl, yl, zl, trc=abi.make_trc(etaS=0.1),
                          precision=abi.PREC_MIXED, eos=eos)
        ctx.set_c2p(c2p, model)
        if sh[0] == abi.STAGE_BANDS:
            ctx.set_stage_bands(sh[2], sh[1])
            assert ctx.get_stage_bands() == sh[2]
        else:
            ctx.set_stage(*sh)
        ctx.set_prim(0.6, *ic, nb=nb)
        st, _ = ctx.run(0.6, 0.02, 6)
        assert st == 0
        tile = ctx.get_stage()
        s = ctx.download()
        ctx.close()
        if ref is None:
            ref = s
            continue
        for f in ("Q", "pi", "Pi"):
            d = np.max(np.abs(s[f] - ref[f])) / max(np.max(np.abs(ref[f])), 1e-300)
            assert np.array_equal(s[f], ref[f]), (sh, tile, f, d)


@pytest.mark.parametrize("c2p", [abi.C2P_REFERENCE, abi.C2P_CLOSED, abi.C2P_NEWTON])
def test_vacuum_tails_below_fp32_and_fp64_squares(gpu, c2p):
    """A plateau with steep Gaussian edges (sigma 0.18 fm) leaves tails from
    1e-240 up through the FP32 range, moving with v_x = 0.2 tanh(x/3): Q^tau
    below 1e-154 (whose square underflows FP64) and below 1e-38 (zero in FP32)
    must be vacuum-like and never a NaN flux or NaN pi (status 0, checked in
    run()).  The error budget is checked on the ideal fluid: with eta/s = 0.1
    the FP64 step itself is chaotic on this edge (a 1e-12 perturbation of e
    grows to ~1e-4 of max T^{tau tau} in cells with e > 0.5 after 30 steps)."""
    nx, ny, nz = 40, 40, 8
    xl, yl, zl = (-6.0, 6.0), (-6.0, 6.0), (-1.0, 1.0)
    X, Y, Z = setups.axes(nx, ny, nz, xl, yl, zl)
    r = np.sqrt(X * X + Y * Y)
    e = 10.0 * np.exp(-np.maximum(r - 2.5, 0.0) ** 2 / (2 * 0.18 ** 2))
    assert e.min() < 1e-200 and np.sum((e > 0) & (e < 1e-154)) > 100
    vx = 0.2 * np.tanh(X / 3.0)
    z = np.zeros_like(e)
    for etaS in (0.0, 0.1):
        args = (nx, ny, nz, xl, yl, zl, (e, vx, z, z.copy()), 0.6, 0.03, 30, etaS)
        ref = run(abi.PREC_FP64, abi.C2P_REFERENCE, *args)
        mix = run(abi.PREC_MIXED, c2p, *args)
        assert np.all(np.isfinite(mix["Q"])) and np.all(np.isfinite(mix["pi"]))
        if etaS == 0.0:
            eT, _ = errors(mix, ref, ref["tau"])
            print(f"c2p={c2p} ideal: T^tau-mu rel err {eT:.3e} (e > 0.5)", flush=True)
            assert eT < 1e-6


@pytest.mark.parametrize("c2p", [abi.C2P_CLOSED, abi.C2P_NEWTON])
def test_charge_free_kernels_match_charged(gpu, c2p):
    """A state without charges runs the 4-component kernels (n_b, n_q, n_s
    are zero and stay zero); the same state with n_b = 1e-300 (rounded to 0 in
    every FP32 operation) runs the 7-component kernels.  T^{tau mu} and pi
    must be bit-identical, and the charge components of the first run zero."""
    cfg = setups.VISCOUS_3D
    nx, ny, nz = 30, 28, 12
    e, vx, vy, vz = setups.gaussian3d_ic(nx, ny, nz, cfg["xlim"], cfg["ylim"], cfg["zlim"],
                                         e0=cfg["e0"], sx=cfg["sx"], sy=cfg["sy"],
                                         eta_plateau=cfg["eta_plateau"],
                                         eta_sigma=cfg["eta_sigma"])
    vx = 0.1 * np.tanh(setups.axes(nx, ny, nz, cfg["xlim"], cfg["ylim"], cfg["zlim"])[0])
    out = []
    for nb in (None, np.full_like(e, 1e-300)):
        ctx = abi.Context(nx, ny, nz, cfg["xlim"], cfg["ylim"], cfg["zlim"],
                          trc=abi.make_trc(etaS=0.1), precision=abi.PREC_MIXED)
        ctx.set_c2p(c2p)
        ctx.set_prim(0.6, e, vx, vy, vz, nb=nb)
        st, _ = ctx.run(0.6, 0.03, 20)
        assert st == 0
        out.append(ctx.download())
        ctx.close()
    a, b = out
    assert np.array_equal(a["Q"][:4], b["Q"][:4]) and np.array_equal(a["pi"], b["pi"])
    assert not a["Q"][4:].any() and not a["Qh"][4:].any() and b["Q"][4].any()


@pytest.mark.parametrize("shape", [(1, 1, 1), (3, 3, 1), (4, 2, 3), (5, 1, 2), (6, 5, 4),
                                   (9, 7, 3), (70, 3, 2)])
def test_tiny_and_ragged_grids_mixed(gpu, shape):
    """The mixed kernels (y-strip cell order, stage tiles) on extents 1 to 9
    and a long thin grid: status clean, the same state as the per-cell stage
    kernel forced, and close to the FP64 step"""
    nx, ny, nz = shape
    lim = lambda n, h: (-(n - 1) * h / 2, (n - 1) * h / 2) if n > 1 else (-1.0, 1.0)
    xl, yl, zl = lim(nx, 0.5), lim(ny, 0.5), lim(nz, 0.2)
    rng = np.random.default_rng(sum(shape))
    e = 5.0 + 5.0 * rng.uniform(size=(nz, ny, nx))
    vx, vy, vz = (0.2 * rng.uniform(-1, 1, e.shape) for _ in range(3))
    out = {}
    for name, prec, stage in (("fp64", abi.PREC_FP64, None),
                              ("mixed", abi.PREC_MIXED, None),
                              ("mixed_cells", abi.PREC_MIXED, (abi.STAGE_CELLS, 0))):
        ctx = abi.Context(nx, ny, nz, xl, yl, zl, trc=abi.make_trc(etaS=0.1), precision=prec)
        if stage:
            ctx.set_stage(*stage)
        ctx.set_prim(1.0, e, vx, vy, vz)
        st, _ = ctx.run(1.0, 0.02, 4)
        assert st == 0, (name, st)
        out[name] = ctx.download()
        ctx.close()
    a, b = out["mixed"]["Q"], out["fp64"]["Q"]
    assert np.all(np.isfinite(a))
    assert np.max(np.abs(a - b)) <= 1e-5 * np.max(np.abs(b)), shape
    if out["mixed"]["Q"].shape == out["mixed_cells"]["Q"].shape:
        assert np.max(np.abs(out["mixed"]["Q"] - out["mixed_cells"]["Q"])) <= 1e-12 * np.max(
            np.abs(b))
