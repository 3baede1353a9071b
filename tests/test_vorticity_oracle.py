"""CPU pins of the oracle's vorticity (thermal-vorticity gradient) output:
dbeta[i][j] = d_i (u^j / T) formed in NSquant (src/hdo.cpp:420-611), stored per
cell by Cell::setDbeta (src/cll.h:173-183) and written to beta.dat by
Fluid::outputSurface (src/fld.cpp:663-834).

* Bjorken flow (uniform e, v = 0): the time row is d(1/T)/dtau with T from the
  ideal-gas Bjorken solution T ~ tau^(-1/3), dbeta[3][3] is the Christoffel
  term 1/(T tau), every other entry vanishes.
* A static-looking smooth flow after one tiny step: the space rows equal the
  analytic gradients of u^nu / T (from a 4th-order derivative of the closed
  form, not the oracle's stencil) to the grid's O(dx^2) accuracy, with
  d_eta / tau and the Christoffel terms in row 3.
* Next to vacuum, a cell whose NSquant returns at the vacuum check keeps the
  shared matrix of the cell before it in the reference's loop order (ix, iy,
  iz with iz fastest); empty cells hold zeros; stored values are clamped to
  [-100, 100].
* beta.dat records of a Bjorken surface: the Cartesian d_i beta_j of
  beta_mu = u_mu / T with u = (t, 0, 0, z) / tau, in closed form.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from vhlle_amd import setups

CNST = (16 + 0.5 * 21.0 * 2.5) * np.pi ** 2 / 30.0 / 0.197326968 ** 3  # EoSs::t, ideal gas


def T_of(e):
    return np.where(e > 0, (np.maximum(e, 0) / CNST) ** 0.25, 0.0)


def bjorken(n=8, tau0=1.0, e0=40.0, dt=0.01, steps=3, etaS=1e-4):
    xl = (-0.35, 0.35)
    s = po.Solver(n, n, n, xl, xl, xl, trc=po.make_trc(etaS=etaS))
    s.set_vorticity(True)
    e = np.full((n, n, n), e0)
    z = np.zeros_like(e)
    s.set_prim(tau0, e, z, z, z)
    rc, tau = s.run(tau0, dt, steps)
    assert rc == 0
    return s, tau


def test_bjorken_time_row_and_christoffel():
    dt = 0.01
    tau0, e0 = 1.0, 40.0
    s, tau = bjorken(tau0=tau0, e0=e0, dt=dt)
    db = s.get_dbeta()
    assert np.all(db == db[:, :1])  # uniform state, uniform gradients
    d = db[:, 0].reshape(4, 4)

    def T_an(t):  # ideal Bjorken: e ~ tau^(-4/3)
        return T_of(e0 * (tau0 / t) ** (4.0 / 3.0))

    # d(1/T)/dtau between tau - dt and tau, and 1/(T tau) at the half step
    th = tau - 0.5 * dt
    A = (1.0 / T_an(tau) - 1.0 / T_an(tau - dt)) / dt
    assert d[0, 0] == pytest.approx(A, rel=2e-3)
    assert d[0, 0] == pytest.approx(1.0 / (3.0 * T_an(th) * th), rel=2e-3)
    assert d[3, 3] == pytest.approx(1.0 / (T_an(th) * th), rel=1e-3)
    off = [(i, j) for i in range(4) for j in range(4) if (i, j) not in ((0, 0), (3, 3))]
    for i, j in off:
        assert abs(d[i, j]) <= 1e-12 * abs(d[3, 3]), (i, j, d[i, j])


def _flow(X, Y, H):
    """smooth fields: e, vx, vy, vz (vz: vHLLE's eta velocity, u^eta = gamma vz)"""
    e = 20.0 * (1.0 + 0.3 * np.sin(0.9 * X + 0.2) * np.cos(0.6 * Y) + 0.2 * np.cos(1.1 * H))
    vx = 0.15 * np.sin(0.8 * Y + 0.3) + 0.05 * np.cos(H)
    vy = 0.12 * np.cos(0.7 * X) * np.cos(0.5 * H)
    vz = 0.1 * np.sin(0.6 * X + 0.9 * Y) + 0.08 * np.sin(1.3 * H)
    return e, vx, vy, vz


def _beta_up(X, Y, H):
    e, vx, vy, vz = _flow(X, Y, H)
    g = 1.0 / np.sqrt(1.0 - vx * vx - vy * vy - vz * vz)
    T = T_of(e)
    return np.stack([g / T, g * vx / T, g * vy / T, g * vz / T])


def _grad(f, X, Y, H, axis, h=1e-3):
    """4th-order central derivative of the closed form along axis 0/1/2"""
    def sh(k):
        d = [0.0, 0.0, 0.0]
        d[axis] = k * h
        return f(X + d[0], Y + d[1], H + d[2])
    return (-sh(2) + 8 * sh(1) - 8 * sh(-1) + sh(-2)) / (12 * h)


def test_space_rows_match_analytic_gradients():
    nx, ny, nz = 16, 14, 12
    xl, yl, zl = (-0.4, 0.35), (-0.3, 0.35), (-0.275, 0.275)
    X, Y, H = setups.axes(nx, ny, nz, xl, yl, zl)
    e, vx, vy, vz = _flow(X, Y, H)
    tau0, dt = 1.0, 1e-3
    s = po.Solver(nx, ny, nz, xl, yl, zl, trc=po.make_trc(etaS=0.08))
    s.set_vorticity(True)
    s.set_prim(tau0, e, vx, vy, vz)
    rc, tau = s.run(tau0, dt, 1)
    assert rc == 0
    db = s.get_dbeta().reshape(4, 4, nz, ny, nx)
    # the two outer layers are overwritten by the boundary copies
    # (Fluid::correctImagCells), so a centred difference is clean from the third
    inner = (slice(3, -3), slice(3, -3), slice(3, -3))
    for d in (1, 2, 3):
        an = _grad(_beta_up, X, Y, H, d - 1)
        if d == 3:
            an = an / tau
        got = db[d]
        if d == 3:  # Christoffel terms at the half step (T and u of the cell)
            bu = _beta_up(X, Y, H)
            th = tau - 0.5 * dt
            got = got.copy()
            got[0] = got[0] - bu[3] / th
            got[3] = got[3] - bu[0] / th
        for j in range(4):
            a, b = got[j][inner], an[j][inner]
            scale = np.max(np.abs(an[:, 3:-3, 3:-3, 3:-3]))
            err = np.max(np.abs(a - b))
            assert err <= 4e-3 * scale, (d, j, err, scale)


def _vacuum_case():
    nx, ny, nz = 14, 13, 7
    xl, yl, zl = (-3.25, 3.25), (-3.0, 3.0), (-0.6, 0.6)
    X, Y, H = setups.axes(nx, ny, nz, xl, yl, zl)
    r2 = X ** 2 / 4.0 + Y ** 2 / 3.0 + H ** 2 / 0.5
    e = np.where(r2 < 1.0, 25.0 * (1.0 - r2) + 0.5, 0.0)
    vx = 0.2 * np.tanh(X)
    vy = 0.25 * np.tanh(Y)
    vz = 0.1 * np.tanh(H)
    s = po.Solver(nx, ny, nz, xl, yl, zl, trc=po.make_trc(etaS=0.1))
    s.set_vorticity(True)
    s.set_prim(1.0, e, vx, vy, vz)
    dt = 0.05
    rc, tau = s.run(1.0, dt, 4)
    assert rc == 0
    return s, (nx, ny, nz), tau, dt


def test_vacuum_neighbours_keep_the_previous_cells_matrix():
    s, (nx, ny, nz), tau, dt = _vacuum_case()
    N = nx * ny * nz
    st = s.get_state()
    db = s.get_dbeta()
    # H-centre energy density of every cell (ISformal's empty-cell test)
    q = (st["Qh"] / (tau - dt / 2)).T.copy()
    eH = po.c2p(q)[0][:, 0]
    maxM = st["m"].max(axis=0)

    def cid(ix, iy, iz):
        ix = min(max(ix, 0), nx - 1)
        iy = min(max(iy, 0), ny - 1)
        return ix + nx * (iy + ny * iz)

    order = [cid(ix, iy, iz) for ix in range(nx) for iy in range(ny) for iz in range(nz)]
    prev = None
    n_stale = n_empty = 0
    for c in order:
        if eH[c] <= 0.0:
            assert np.all(db[:, c] == 0.0)
            n_empty += 1
            continue
        ix, iy, iz = c % nx, (c // nx) % ny, c // (nx * ny)
        near_vac = any(maxM[cid(ix + a, iy + b, iz)] <= 0.9
                       for a, b in ((1, 0), (0, 1), (-1, 0), (0, -1), (1, 1), (1, -1), (-1, 1),
                                    (-1, -1)))
        if near_vac:
            want = db[:, prev] if prev is not None else np.zeros(16)
            assert np.array_equal(db[:, c], want), c
            n_stale += 1
        prev = c
    assert n_stale > 10 and n_empty > 10
    assert np.all(np.abs(db) <= 100.0)
    assert np.any(db != 0.0)


def test_clamp_at_100():
    nx, ny, nz = 10, 10, 5
    xl = (-0.225, 0.225)
    X, Y, H = setups.axes(nx, ny, nz, xl, xl, (-0.2, 0.2))
    e = 5.0 + 4.0 * np.sign(X) # a sharp jump: d(1/T)/dx ~ 1/dx
    z = np.zeros_like(e)
    vx = 0.6 * np.sign(Y)
    s = po.Solver(nx, ny, nz, xl, xl, (-0.2, 0.2), trc=po.make_trc(etaS=0.08))
    s.set_vorticity(True)
    s.set_prim(1.0, e, vx, z, z)
    rc, _ = s.run(1.0, 0.01, 1)
    assert rc == 0
    db = s.get_dbeta()
    assert np.max(np.abs(db)) == 100.0


def test_beta_records_of_a_bjorken_surface():
    n, dt, tau0, e0 = 8, 0.01, 1.0, 40.0
    s, tau = bjorken(n=n, tau0=tau0, e0=e0, dt=dt, steps=2)
    lo = s.get_state()
    rc, tau = s.run(tau, dt, 1)
    hi = s.get_state()
    db = s.get_dbeta()
    q_lo = (lo["Q"] / (tau - dt)).T.copy()
    q_hi = (hi["Q"] / tau).T.copy()
    e_lo, e_hi = po.c2p(q_lo)[0][0, 0], po.c2p(q_hi)[0][0, 0]
    ecrit = 0.5 * (e_lo + e_hi)
    xl = (-0.35, 0.35)
    sp = 0.7 / (n - 1)
    rows, beta = po.freezeout((n, n, n), (sp, sp, sp), (xl[0], xl[0], xl[0]), hi["Q"], lo["Q"],
                              hi["pi"], hi["Pi"], tau, dt, ecrit, extend=True, dbeta=db)
    assert len(rows) == (n - 4) ** 3
    assert np.array_equal(beta[:, :16], rows[:, :16])
    assert np.array_equal(beta[:, 32], rows[:, 27])
    d = db[:, 0].reshape(4, 4)  # uniform: d_tau(1/T) = d[0,0], 1/(T tau) = d[3,3]
    A, B = d[0, 0], d[3, 3]
    for r, b in zip(rows, beta):
        eta = r[3]
        ch, sh = np.cosh(eta), np.sinh(eta)
        # beta_t = ch/T, beta_z = -sh/T; d_t = ch d_tau - sh d_eta/tau,
        # d_z = -sh d_tau + ch d_eta/tau
        want = np.zeros((4, 4))
        want[0, 0] = ch * ch * A - sh * sh * B
        want[0, 3] = -ch * sh * A + sh * ch * B
        want[3, 0] = -sh * ch * A + ch * sh * B
        want[3, 3] = sh * sh * A - ch * ch * B
        got = b[16:32].reshape(4, 4)
        assert np.allclose(got, want, rtol=1e-12, atol=1e-12 * abs(B)), (eta, got, want)
