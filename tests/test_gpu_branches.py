"""Branches of the hot path that the smooth workloads never take, and the
signal-speed reduction, on the GPU against the oracle.

* Negative pressure: hlle_flux evaluates both face states again from Q (not
  Qh, even in the corrector) when the right state's pressure is negative
  (src/hdo.cpp:134-140).  With p = e/3 that cannot happen; with a table EoS
  whose nodes the reference clamps to p >= 0 (src/eos.cpp:49) the natural
  spline undershoots below zero next to the clamped region.  A bag-model
  table, p = max(e - 0.1, 0)/3, does that; the oracle's counters show the
  branch taken in both stages.
* ISformal limiter and bulk clamp (src/hdo.cpp:925-941): a state with
  |pi| > T^{mu nu} and |Pi| > p in a blob (uploaded through the state
  interface) makes the oracle rescale pi and clamp Pi.
* Signal speeds: the per-wave maxima of max(-b_l, b_r) over the flux
  kernels equal the oracle's maximum over every interface (FP64: bit for
  bit; mixed precision: FP32 speeds, within 1e-5).

FP64 bar: 1e-10 of each field's scale, as every FP64 parity test.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from vhlle_amd import abi, setups

pytestmark = pytest.mark.gpu

TOL = 1e-10
FIELDS = ("Q", "Qh", "Qprev", "pi", "piH", "Pi", "PiH", "viscCorrCut")


def assert_close(a, b, fields=FIELDS, tol=TOL, label=""):
    for f in fields:
        scale = max(float(np.max(np.abs(b[f]))), 1e-300)
        d = float(np.max(np.abs(a[f] - b[f])))
        assert np.all(np.isfinite(a[f])), f"{label} {f} not finite"
        assert d <= tol * scale, f"{label} {f}: {d:.3e} of {scale:.3e}"


def bag_table(e_bag=0.1):
    """p = max(e - e_bag, 0)/3 on 601 nodes (0, then 1e-4..1e4 geometric),
    T = (e/1809.03)^(1/4): the spline of the clamped nodes rings around zero
    below e_bag (down to |p| ~ 1e-40).  With e_bag = 0.1 those states have
    T < 0.1 GeV, where the transport coefficients vanish (src/trancoeff.cpp:
    92-100)."""
    e = np.concatenate(([0.0], np.geomspace(1e-4, 1e4, 600)))
    return abi.EosTable(e, np.maximum((e - e_bag) / 3.0, 0.0), (e / 1809.03) ** 0.25)


def smooth_ic(nx, ny, nz, xl, yl, zl, e0=20.0, floor=0.05):
    e, _, _, _ = setups.gaussian3d_ic(nx, ny, nz, xl, yl, zl, e0=e0, sx=2.0, sy=1.8,
                                      eta_plateau=0.5, eta_sigma=0.5)
    X, Y, Z = setups.axes(nx, ny, nz, xl, yl, zl)
    return e + floor, 0.3 * np.tanh(X / 2), 0.3 * np.tanh(Y / 2), 0.2 * np.tanh(Z)


GRID = (24, 22, 6, (-5.75, 5.75), (-5.25, 5.25), (-1.0, 1.0))


@pytest.mark.parametrize("trc_kw", [dict(etaS=0.08), {}])
def test_negative_pressure_reevaluation_fp64(gpu, trc_kw):
    tab = bag_table()
    nx, ny, nz, xl, yl, zl = GRID
    e, vx, vy, vz = smooth_ic(nx, ny, nz, xl, yl, zl, floor=0.03)
    # the table's spline is negative between clamped nodes
    assert po.eos_spline(tab, 0, np.linspace(0.05, 0.0999, 50)).min() < 0.0
    ctx = abi.Context(nx, ny, nz, xl, yl, zl, trc=abi.make_trc(**trc_kw), eos=tab)
    orc = po.Solver(nx, ny, nz, xl, yl, zl, trc=po.make_trc(**trc_kw), table=tab)
    ctx.set_prim(0.6, e, vx, vy, vz)
    orc.set_prim(0.6, e, vx, vy, vz)
    tg = to = 0.6
    for s in range(10):
        st, tg = ctx.step(tg, 0.02)
        rc, to = orc.run(to, 0.02, 1)
        assert st == 0 and rc == 0, (s, st, rc)
        assert_close(ctx.download(), orc.get_state(), label=f"{trc_kw} step {s}")
    n = orc.branch_counts()
    print(f"{trc_kw}: oracle re-evaluations predictor {n[0]}, corrector {n[1]}, "
          f"limiter {n[2]}, Pi clamp {n[3]}", flush=True)
    assert n[0] > 0 and n[1] > 0
    np.testing.assert_array_equal(ctx.signal_speed(), orc.signal_speed())


@pytest.mark.parametrize("c2p", [abi.C2P_REFERENCE, abi.C2P_NEWTON])
def test_negative_pressure_reevaluation_mixed(gpu, c2p):
    """The mixed kernels take the same branch (states from Q in FP64 MUSCL,
    El and Er of the first states): within 2e-5 of max T^{tau tau} of the
    FP64 path over the whole grid after 10 steps."""
    tab = bag_table()
    nx, ny, nz, xl, yl, zl = GRID
    e, vx, vy, vz = smooth_ic(nx, ny, nz, xl, yl, zl, floor=0.03)
    out = {}
    for name, prec, mode in (("fp64", abi.PREC_FP64, abi.C2P_REFERENCE),
                             ("mixed", abi.PREC_MIXED, c2p)):
        ctx = abi.Context(nx, ny, nz, xl, yl, zl, trc=abi.make_trc(etaS=0.08), precision=prec,
                          eos=tab)
        if mode != abi.C2P_REFERENCE:
            ctx.set_c2p(mode)
        ctx.set_prim(0.6, e, vx, vy, vz)
        st, tau = ctx.run(0.6, 0.02, 10)
        assert st == 0
        out[name] = (ctx.download(), ctx.signal_speed())
        ctx.close()
    (a, sa), (b, sb) = out["mixed"], out["fp64"]
    eT = np.max(np.abs(a["Q"][:4] - b["Q"][:4])) / np.max(np.abs(b["Q"][0]))
    epi = np.max(np.abs(a["pi"] - b["pi"])) / np.max(np.abs(b["pi"]))
    print(f"c2p {c2p}: error against FP64: T^tau-mu {eT:.2e}, pi {epi:.2e}", flush=True)
    assert eT < 2e-5 and epi < 1e-4
    np.testing.assert_allclose(sa, sb, rtol=1e-5)


def test_mixed_finite_where_phi7_explodes(gpu):
    """A table whose p rings around zero (|p| down to 1e-37) at T > 0.1 GeV
    (kink at e = 1): the reference's phi7 = 9/(70 p tau_pi) reaches 1e35 and
    its cross terms 1e20 and more, which its limiter scales back.  FP32 cannot
    carry those magnitudes (the kernel caps |phi7| at 1e10), so the limited
    pi of such cells differs from FP64's (even in sign); the state stays
    finite, and the hot region (e > 3 GeV/fm^3) stays within 2e-5.  The
    spread over all cells is printed."""
    tab = bag_table(1.0)
    nx, ny, nz, xl, yl, zl = GRID
    e, vx, vy, vz = smooth_ic(nx, ny, nz, xl, yl, zl)
    out = {}
    for name, prec in (("fp64", abi.PREC_FP64), ("mixed", abi.PREC_MIXED)):
        ctx = abi.Context(nx, ny, nz, xl, yl, zl, trc=abi.make_trc(etaS=0.08), precision=prec,
                          eos=tab)
        ctx.set_prim(0.6, e, vx, vy, vz)
        st, tau = ctx.run(0.6, 0.02, 10)
        assert st == 0
        out[name] = ctx.download()
        ctx.close()
    a, b = out["mixed"], out["fp64"]
    for f in ("Q", "pi", "Pi"):
        assert np.all(np.isfinite(a[f])), f
    hot = abi.c2p((b["Q"].T / tau).copy(), eos=tab)[0][:, 0] > 3.0
    scale = np.max(np.abs(b["Q"][0]))
    eT = float(np.max(np.abs(a["Q"][:4, hot] - b["Q"][:4, hot])) / scale)
    eT_all = float(np.max(np.abs(a["Q"][:4] - b["Q"][:4])) / scale)
    print(f"T^tau-mu error against FP64: {eT:.2e} over e > 3 ({hot.sum()} cells), "
          f"{eT_all:.2e} over all cells", flush=True)
    assert eT < 2e-5


def limiter_state(orc, e):
    """The IC state with a blob of |pi| up to 1.5 e (above T^{tau tau}) and
    Pi = -0.8 e (beyond -p)"""
    nx, ny, nz = orc.shape
    X, Y, _ = setups.axes(nx, ny, nz, (-3.75, 3.75), (-3.25, 3.25), (-1.0, 1.0))
    s = orc.get_state()
    ef = e.ravel()
    blob = np.exp(-(X ** 2 + Y ** 2) / 2).ravel()
    pi = np.zeros((10, ef.size))
    pi[2] = 1.5 * ef * blob * np.sign(np.sin(3 * X.ravel()))   # xx
    pi[5] = -1.2 * ef * blob                                   # yy
    pi[4] = 0.3 * ef * blob                                    # xy
    Pi = -0.8 * ef * blob
    s.update(pi=pi, piH=pi.copy(), Pi=Pi, PiH=Pi.copy())
    return s


LIM = (16, 14, 6, (-3.75, 3.75), (-3.25, 3.25), (-1.0, 1.0))
LIM_TRC = dict(etaS=0.2, zetaS=0.1, zetaSparam=2)


def test_limiter_and_bulk_clamp_fp64(gpu):
    nx, ny, nz, xl, yl, zl = LIM
    e, vx, vy, vz = smooth_ic(nx, ny, nz, xl, yl, zl, floor=0.5)
    orc = po.Solver(nx, ny, nz, xl, yl, zl, trc=po.make_trc(**LIM_TRC))
    orc.set_prim(0.6, e, vx, vy, vz)
    s = limiter_state(orc, e)
    orc.set_state(s)
    ctx = abi.Context(nx, ny, nz, xl, yl, zl, trc=abi.make_trc(**LIM_TRC))
    ctx.upload({k: s[k] for k in ("Q", "Qh", "Qprev", "pi", "piH", "Pi", "PiH", "m",
                                  "viscCorrCut")})
    tg = to = 0.6
    cut = 0
    for k in range(6):
        st, tg = ctx.step(tg, 0.02)
        rc, to = orc.run(to, 0.02, 1)
        assert st == 0 and rc == 0
        b = orc.get_state()
        cut += int(np.sum(b["viscCorrCut"] < 1.0))
        assert_close(ctx.download(), b, label=f"limiter step {k}")
    n = orc.branch_counts()
    print(f"oracle: limiter {n[2]} cells, Pi clamp {n[3]} cells, cells with "
          f"viscCorrCut < 1 summed over steps {cut}", flush=True)
    assert n[2] > 0 and n[3] > 0 and cut > 0


def test_limiter_and_bulk_clamp_mixed(gpu):
    """Mixed precision (FP32 relaxation and limiter) on the same state: the
    same cells are limited in the first step, and the state stays close to
    the FP64 path's."""
    nx, ny, nz, xl, yl, zl = LIM
    e, vx, vy, vz = smooth_ic(nx, ny, nz, xl, yl, zl, floor=0.5)
    orc = po.Solver(nx, ny, nz, xl, yl, zl, trc=po.make_trc(**LIM_TRC))
    orc.set_prim(0.6, e, vx, vy, vz)
    s = limiter_state(orc, e)
    out = {}
    for name, prec in (("fp64", abi.PREC_FP64), ("mixed", abi.PREC_MIXED)):
        ctx = abi.Context(nx, ny, nz, xl, yl, zl, trc=abi.make_trc(**LIM_TRC), precision=prec)
        ctx.upload({k: s[k] for k in ("Q", "Qh", "Qprev", "pi", "piH", "Pi", "PiH", "m",
                                      "viscCorrCut")})
        st, t = ctx.step(0.6, 0.02)
        assert st == 0
        first = ctx.download()["viscCorrCut"].copy()
        st, _ = ctx.run(t, 0.02, 5)
        assert st == 0
        out[name] = (first, ctx.download())
        ctx.close()
    (fa, a), (fb, b) = out["mixed"], out["fp64"]
    assert np.sum(fb < 1.0) > 0
    np.testing.assert_array_equal(fa < 1.0, fb < 1.0)
    np.testing.assert_allclose(fa, fb, rtol=1e-4)
    eT = np.max(np.abs(a["Q"][:4] - b["Q"][:4])) / np.max(np.abs(b["Q"][0]))
    epi = np.max(np.abs(a["pi"] - b["pi"])) / np.max(np.abs(b["pi"]))
    ePi = np.max(np.abs(a["Pi"] - b["Pi"])) / np.max(np.abs(b["Pi"]))
    print(f"mixed vs FP64 after 6 steps: T {eT:.2e}, pi {epi:.2e}, Pi {ePi:.2e}", flush=True)
    assert eT < 1e-5 and epi < 1e-4 and ePi < 1e-4


def _viscous_pair(precision, c2p=None, vacuum=False):
    nx, ny, nz = 16, 16, 6
    xl, yl, zl = (-3.75, 3.75), (-3.75, 3.75), (-1.0, 1.0)
    e, vx, vy, vz = smooth_ic(nx, ny, nz, xl, yl, zl, e0=30.0, floor=0.0 if vacuum else 0.05)
    if vacuum:
        e = np.where(e < 1.0, 0.0, e)
    ctx = abi.Context(nx, ny, nz, xl, yl, zl, trc=abi.make_trc(etaS=0.1), precision=precision)
    if c2p is not None:
        ctx.set_c2p(c2p)
    orc = po.Solver(nx, ny, nz, xl, yl, zl, trc=po.make_trc(etaS=0.1))
    ctx.set_prim(0.8, e, vx, vy, vz)
    orc.set_prim(0.8, e, vx, vy, vz)
    return ctx, orc


@pytest.mark.parametrize("vacuum", [False, True])
def test_signal_speed_fp64_equals_oracle(gpu, vacuum):
    ctx, orc = _viscous_pair(abi.PREC_FP64, vacuum=vacuum)
    tg = to = 0.8
    first = None
    for k in range(4):
        st, tg = ctx.step(tg, 0.04)
        rc, to = orc.run(to, 0.04, 1)
        assert st == 0 and rc == 0
        a, b = ctx.signal_speed(), orc.signal_speed()
        print(f"vacuum={vacuum} step {k}: {a}", flush=True)
        np.testing.assert_array_equal(a, b)
        first = a if first is None else first
    # the graph replay reduces the same maxima; a reset clears the run maxima
    st, tg = ctx.run(tg, 0.04, 3)
    rc, to = orc.run(to, 0.04, 3)
    np.testing.assert_array_equal(ctx.signal_speed(reset=True), orc.signal_speed(reset=True))
    assert np.all(ctx.signal_speed()[6:] == 0.0)
    assert np.all(ctx.signal_speed()[:6] > 0.0)
    if vacuum:  # the first step sees the vacuum boundary's b = -+1 (1/tau along eta)
        assert first[0] == first[1] == 1.0 and first[2] == 1.0 / 0.8


@pytest.mark.parametrize("c2p", [abi.C2P_REFERENCE, abi.C2P_CLOSED, abi.C2P_NEWTON])
def test_signal_speed_mixed_close_to_oracle(gpu, c2p):
    ctx, orc = _viscous_pair(abi.PREC_MIXED, c2p=c2p)
    st, tg = ctx.run(0.8, 0.04, 4)
    rc, to = orc.run(0.8, 0.04, 4)
    assert st == 0 and rc == 0
    a, b = ctx.signal_speed(), orc.signal_speed()
    print(f"c2p {c2p}: {a}\n  oracle {b}", flush=True)
    np.testing.assert_allclose(a, b, rtol=1e-5)
    # the row stage kernel (each x and eta interface evaluated once) reduces
    # the same maxima as the cell kernel, on a fresh context (set_prim does
    # not reset pi, as Cell::setPrimVar does not)
    for kind, cz in ((abi.STAGE_CELLS, 0), (abi.STAGE_ROWS, 2), (abi.STAGE_BANDS, 2)):
        c2, _ = _viscous_pair(abi.PREC_MIXED, c2p=c2p)
        c2.set_stage(kind, cz)
        st, _ = c2.run(0.8, 0.04, 4)
        assert st == 0
        np.testing.assert_array_equal(c2.signal_speed(), a)
        c2.close()
