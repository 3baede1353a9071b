"""Parity of the FP64 HIP path with the CPU oracle, through the C-ABI.

Bar: for every state field f, max|gpu - oracle| <= 1e-10 * max|oracle| (the
tolerance of the plain FP64 step); con2prim/prim2con are expected
bit-identical because sqrt and division are correctly rounded on both sides.
"""
import numpy as np
import pytest

from vhlle_amd import setups

pytestmark = pytest.mark.gpu

TOL = 1e-10
FIELDS = ("Q", "Qh", "Qprev", "pi", "piH", "Pi", "PiH", "m", "viscCorrCut")


def random_Q(rng, n):
    """Conserved states covering the con2prim branches: ordinary (|v| up to
    0.9999), at rest, superluminal (|Q_vec| > Q_t), vacuum and negative."""
    e = 10 ** rng.uniform(-6, 2.5, n)
    v = rng.uniform(0, 0.9999, n) ** 0.5
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    prim = np.zeros((n, 8))
    prim[:, 0] = e
    prim[:, 1] = e / 3.0
    prim[:, 2] = rng.uniform(0, 1, n) * (rng.uniform(size=n) < 0.5)
    prim[:, 5:8] = v[:, None] * d
    from oracle import pyoracle
    Q = pyoracle.p2c(prim)
    k = n // 10
    Q[:k, 1:4] = 0.0                                   # at rest
    Q[k:2 * k, 1:4] *= 1.0 + rng.uniform(0, 0.5, k)[:, None]  # superluminal input
    Q[2 * k:2 * k + k // 4, 0] = 0.0                   # vacuum
    Q[2 * k + k // 4:2 * k + k // 2, 0] *= -1.0        # negative energy
    return Q


def assert_state_close(a, b, fields=FIELDS, tol=TOL, label=""):
    for f in fields:
        ref = b[f]
        scale = np.max(np.abs(ref))
        diff = np.max(np.abs(a[f] - ref)) if ref.size else 0.0
        assert np.all(np.isfinite(a[f])), f"{label} {f}: non-finite values"
        assert diff <= tol * max(scale, 1e-300), f"{label} {f}: max diff {diff:.3e} scale {scale:.3e}"


def test_c2p_bitwise(gpu, oracle):
    rng = np.random.default_rng(1234)
    Q = random_Q(rng, 200_000)
    og, itg, Qg = gpu.c2p(Q)
    oo, ito, Qo = oracle.c2p(Q)
    print("iterations: mean %.2f max %d" % (ito.mean(), ito.max()), flush=True)
    assert np.array_equal(itg, ito)
    assert np.array_equal(Qg, Qo)
    nz = np.abs(oo) > 0
    rel = np.max(np.abs(og - oo)[nz] / np.abs(oo)[nz])
    print("c2p max rel diff", rel, "exact fraction", np.mean(og == oo), flush=True)
    assert np.array_equal(og, oo)


def test_pressure_of_rest_states_bitwise(gpu):
    """p = e/3 of the ideal gas (the device forms it with a multiply and two
    FMAs): a state at rest returns e = Q^tau and p = Q^tau / 3, bit for bit
    IEEE division, over the whole exponent range (subnormal, zero, huge)"""
    rng = np.random.default_rng(99)
    n = 200_000
    qt = np.ldexp(1.0 + rng.random(n), rng.integers(-1074, 1023, n))
    qt[:8] = [0.0, 5e-324, 2.2e-308, 2.0 ** -1000, 2.0 ** 1000, 1.7e308, 1.0, 3.0]
    Q = np.zeros((n, 7))
    Q[:, 0] = qt
    out, _, _ = gpu.c2p(Q)
    assert np.array_equal(out[:, 0], qt)
    assert np.array_equal(out[:, 1], qt / 3.0)


def test_p2c_bitwise(gpu, oracle):
    rng = np.random.default_rng(7)
    n = 50_000
    prim = np.zeros((n, 8))
    prim[:, 0] = 10 ** rng.uniform(-4, 2, n)
    prim[:, 1] = prim[:, 0] / 3
    prim[:, 2:5] = rng.uniform(0, 1, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    prim[:, 5:8] = rng.uniform(0, 0.99, n)[:, None] * d
    assert np.array_equal(gpu.p2c(prim), oracle.p2c(prim))


def _pair(gpu, oracle, nx, ny, nz, xl, yl, zl, trc_kw):
    ctx = gpu.Context(nx, ny, nz, xl, yl, zl, trc=gpu.make_trc(**trc_kw))
    orc = oracle.Solver(nx, ny, nz, xl, yl, zl, trc=oracle.make_trc(**trc_kw))
    return ctx, orc


def test_ideal_gubser_steps(gpu, oracle):
    cfg = setups.GUBSER_2D
    n = 41
    X, Y, _ = setups.axes(n, n, 1, cfg["xlim"], cfg["ylim"], cfg["zlim"])
    e, vx, vy = setups.gubser(1.0, X, Y, amp=10.0)
    z = np.zeros_like(e)
    ctx, orc = _pair(gpu, oracle, n, n, 1, cfg["xlim"], cfg["ylim"], cfg["zlim"], {})
    ctx.set_prim(1.0, e, vx, vy, z)
    orc.set_prim(1.0, e, vx, vy, z)
    assert_state_close(ctx.download(), orc.get_state(), label="IC")
    tau_g = tau_o = 1.0
    dt = 0.05
    for s in range(30):
        st, tau_g = ctx.step(tau_g, dt)
        rc, tau_o = orc.run(tau_o, dt, 1)
        assert st == 0 and rc == 0
    assert tau_g == tau_o
    a, b = ctx.download(), orc.get_state()
    assert_state_close(a, b, label="gubser 30 steps")
    print("Q bit-identical:", np.array_equal(a["Q"], b["Q"]), flush=True)


def _viscous_ic(nx, ny, nz, xl, yl, zl, seed, vmax=0.3, vacuum=False):
    rng = np.random.default_rng(seed)
    e, vx, vy, vz = setups.gaussian3d_ic(nx, ny, nz, xl, yl, zl, e0=30.0, sx=1.5, sy=1.2,
                                         eta_plateau=0.5, eta_sigma=0.5)
    X, Y, Z = setups.axes(nx, ny, nz, xl, yl, zl)
    vx = vmax * np.tanh(X / 2.0) + 0.05 * rng.uniform(-1, 1, e.shape)
    vy = vmax * np.tanh(Y / 2.0) + 0.05 * rng.uniform(-1, 1, e.shape)
    vz = 0.2 * np.tanh(Z) + 0.05 * rng.uniform(-1, 1, e.shape)
    e = e * (1 + 0.1 * rng.uniform(-1, 1, e.shape))
    nb = 0.1 * e / 30.0
    if vacuum:
        e = np.where(e < 1.0, 0.0, e)
    return e, vx, vy, vz, nb


@pytest.mark.parametrize("trc_kw", [
    dict(etaS=0.08),
    dict(etaS=0.12, zetaS=0.05, zetaSparam=5),
    dict(etaS=0.2, zetaS=0.1, zetaSparam=2),
])
def test_viscous_3d_steps(gpu, oracle, trc_kw):
    nx, ny, nz = 14, 12, 8
    xl, yl, zl = (-3.25, 3.25), (-2.75, 2.75), (-1.4, 1.4)
    e, vx, vy, vz, nb = _viscous_ic(nx, ny, nz, xl, yl, zl, seed=3)
    ctx, orc = _pair(gpu, oracle, nx, ny, nz, xl, yl, zl, trc_kw)
    ctx.set_prim(0.6, e, vx, vy, vz, nb=nb)
    orc.set_prim(0.6, e, vx, vy, vz, nb=nb)
    tau_g = tau_o = 0.6
    dt = 0.02
    for s in range(12):
        st, tau_g = ctx.step(tau_g, dt)
        rc, tau_o = orc.run(tau_o, dt, 1)
        assert st == 0 and rc == 0, (st, rc)
        assert_state_close(ctx.download(), orc.get_state(), label=f"{trc_kw} step {s}")
    b = orc.get_state()
    assert np.max(np.abs(b["pi"])) > 0.0  # the viscous path was exercised


def test_vacuum_front(gpu, oracle):
    """Matter next to vacuum: updateM propagation, vacuum interface rules and
    NSquant's vacuum check (src/fld.cpp:384-423, src/hdo.cpp:109-143,399-414)."""
    nx, ny, nz = 16, 16, 6
    xl, yl, zl = (-3.75, 3.75), (-3.75, 3.75), (-1.0, 1.0)
    e, vx, vy, vz, nb = _viscous_ic(nx, ny, nz, xl, yl, zl, seed=5, vacuum=True)
    assert np.any(e == 0.0) and np.any(e > 0.0)
    ctx, orc = _pair(gpu, oracle, nx, ny, nz, xl, yl, zl, dict(etaS=0.1))
    ctx.set_prim(0.8, e, vx, vy, vz, nb=nb)
    orc.set_prim(0.8, e, vx, vy, vz, nb=nb)
    tau_g = tau_o = 0.8
    dt = 0.05
    for s in range(15):
        st, tau_g = ctx.step(tau_g, dt)
        rc, tau_o = orc.run(tau_o, dt, 1)
        assert st == 0 and rc == 0
    assert_state_close(ctx.download(), orc.get_state(), label="vacuum front")


def test_graph_run_equals_steps(gpu, oracle):
    """vh_run (replayed HIP graphs) == repeated vh_step, bit for bit, and the
    oracle; replays interleaved with eager work."""
    nx, ny, nz = 12, 10, 6
    xl, yl, zl = (-2.75, 2.75), (-2.25, 2.25), (-1.0, 1.0)
    e, vx, vy, vz, nb = _viscous_ic(nx, ny, nz, xl, yl, zl, seed=11)
    a = gpu.Context(nx, ny, nz, xl, yl, zl, trc=gpu.make_trc(etaS=0.1))
    b = gpu.Context(nx, ny, nz, xl, yl, zl, trc=gpu.make_trc(etaS=0.1))
    orc = oracle.Solver(nx, ny, nz, xl, yl, zl, trc=oracle.make_trc(etaS=0.1))
    for c in (a, b):
        c.set_prim(0.7, e, vx, vy, vz, nb=nb)
    orc.set_prim(0.7, e, vx, vy, vz, nb=nb)
    ta = tb = to = 0.7
    dt = 0.02
    for rep, n in enumerate((5, 17, 3)):  # 17: two 8-step graphs and a single step
        st, ta = a.run(ta, dt, n)
        assert st == 0
        for _ in range(n):
            st, tb = b.step(tb, dt)
        rc, to = orc.run(to, dt, n)
        da, db = a.download(), b.download()
        for f in FIELDS:
            assert np.array_equal(da[f], db[f]), (rep, f)
        assert ta == tb == to
        assert_state_close(da, orc.get_state(), label=f"graph replay {rep}")


def test_nan_status(gpu, oracle):
    """A NaN in one cell's Q stops the reference with exit(1) in hlle_flux
    (src/hdo.cpp:285-298); the GPU step reports VH_ST_NAN_FLUX."""
    n = 12
    xl = (-2.75, 2.75)
    ctx, orc = _pair(gpu, oracle, n, n, 1, xl, xl, (-1, 1), {})
    e = np.full((1, n, n), 5.0)
    z = np.zeros_like(e)
    ctx.set_prim(1.0, e, z, z, z)
    orc.set_prim(1.0, e, z, z, z)
    s = ctx.download()
    s["Q"][4, 5 * n + 5] = np.nan
    ctx.upload(s)
    orc.set_state({"Q": s["Q"]})
    st, _ = ctx.step(1.0, 0.05)
    rc, _ = orc.run(1.0, 0.05, 1)
    assert rc == 1
    assert st & gpu.ST_NAN_FLUX


def test_dtau_check(gpu):
    ctx = gpu.Context(11, 11, 1, (-1, 1), (-1, 1), (-1, 1))
    assert ctx.check_dtau(0.1) == gpu.VH_OK
    assert ctx.check_dtau(0.11) == gpu.VH_ERR_DTAU


@pytest.mark.parametrize("shape", [(1, 1, 1), (2, 1, 1), (3, 3, 1), (4, 2, 3), (5, 1, 2),
                                   (6, 5, 4)])
def test_tiny_and_ragged_grids(gpu, oracle, shape):
    """extents 1 to 6, where the ghost copies (2 layers from index 2 and n-3)
    and the clamped neighbours overlap: the step still equals the oracle's"""
    nx, ny, nz = shape
    lim = lambda n, h: (-(n - 1) * h / 2, (n - 1) * h / 2) if n > 1 else (-1.0, 1.0)
    xl, yl, zl = lim(nx, 0.5), lim(ny, 0.5), lim(nz, 0.2)
    rng = np.random.default_rng(sum(shape))
    e = 5.0 + 5.0 * rng.uniform(size=(nz, ny, nx))
    vx, vy, vz = (0.2 * rng.uniform(-1, 1, e.shape) for _ in range(3))
    ctx, orc = _pair(gpu, oracle, nx, ny, nz, xl, yl, zl, dict(etaS=0.1, zetaS=0.03,
                                                                   zetaSparam=2))
    ctx.set_prim(1.0, e, vx, vy, vz)
    orc.set_prim(1.0, e, vx, vy, vz)
    tg = to = 1.0
    for s in range(4):
        st, tg = ctx.step(tg, 0.02)
        rc, to = orc.run(to, 0.02, 1)
        assert st == 0 and rc == 0
    assert_state_close(ctx.download(), orc.get_state(), label=f"grid {shape}")
