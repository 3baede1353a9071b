"""Pins of the oracle's functions that no golden vector fixes, each against
what the definition or the mathematics fixes independently of the
restatement's code (CPU only):

* ghost fills (src/fld.cpp:211-382): after the z, y, x passes every cell
  holds the interior cell at its clamped indices (2 for 0 and 1, n-3 for
  n-2 and n-1, on every axis at once -- corners included); a copied Q with
  Q^tau < 0 is zero (Cell::setQ, src/cll.h:75-80); the full fill copies Q, pi,
  Pi and leaves Qh alone, the ideal fill copies Q and Qh and leaves pi, Pi;
* updateM (src/fld.cpp:384-423): a planar vacuum front advances one cell
  per ceil-step count n1 = floor(0.9 / r) + 1 calls (r = dt/dx, dt/dy,
  dt/(dz tau)), a cell diagonal to matter only gets 0.707 r along x and y;
* ISformal loop #2 (src/hdo.cpp:866-962): zero velocity is the identity, a
  uniform field is preserved by any velocity, |v dt/dx| = 1 shifts by exactly
  one cell from the upwind side (the half step averages the two), the eta
  weight uses tau - dt/2; the limiter leaves max |pi| = 0.1 T^{mu nu}_max
  where it rescales and |Pi| <= p; cells with e <= 0 keep their old pi;
* hlle_flux (src/hdo.cpp:77-305) on single interfaces with closed forms:
  equal states give the physical flux; matter at rest next to vacuum gives
  F^tau = pre c_s e / (1 + c_s) and F^x = pre p / (1 + c_s) along x, and
  along eta the wave speeds carry 1/tau while the flux prefactor carries
  tauFactor = tau + dt/4 (predictor); supersonic flow takes the upwind
  flux; two vacuum cells, or two partially filled cells, give no flux;
* the eta sweep: on a 1 x 1 x nz column whose ends are at rest the interior
  sums of Q^tau and Q^eta change by exactly dt times the Milne sources of the
  half-step state (conservation form), and a 1-D profile along eta at
  tau = 1000 with deta = dx/tau evolves as the same profile along x;
* the viscous step keeps pi^{mu nu} traceless (to 1e-3 of max |pi|, a
  wrong 1/3 in the Delta-trace terms of the cross terms gives 3e-3) in a flow
  with transverse and longitudinal gradients;
* the transversality terms -(pi^{mu a} u^nu + pi^{nu a} u^mu) u^b d_b u_a
  (src/hdo.cpp:790-809, 843-862) keep pi^{mu nu} u_nu small while a uniform
  flow accelerates: within 1.1 % of max |pi| after 40 steps, where the
  opposite sign gives 3.3 % and the terms left out 1.7 % (both planted in a
  copy of the oracle on the CPU);
* the Christoffel terms of the relaxation (src/hdo.cpp:778-788, 833-842), in a
  uniform flow with a longitudinal (Milne) velocity: pi stays transverse
  within 2 % and traceless within 1e-3; all the terms with the opposite
  sign give 20 %, left out 10 %.

Not fixed by anything here: the sign of the (pi u) u.d u transversality terms
and the v_eta Christoffel terms of the relaxation (transversality of pi in
such a flow is dominated by the advection step, and planting a sign error in
either moved it by less than its own size).
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from vhlle_amd import setups

CS = np.sqrt(1.0 / 3.0)


def clamp_idx(i, n):
    i = np.where(i < 2, 2, i)
    return np.where(i > n - 3, n - 3, i)


def random_state(rng, nx, ny, nz):
    N = nx * ny * nz
    prim = np.zeros((N, 8))
    prim[:, 0] = 10 ** rng.uniform(-1, 1.5, N)
    prim[:, 1] = prim[:, 0] / 3
    prim[:, 2] = rng.uniform(0, 0.5, N)
    prim[:, 5:8] = rng.uniform(-0.4, 0.4, (N, 3))
    Q = po.p2c(prim).T.copy()
    return Q


# --------------------------------------------------------------- ghost fills
@pytest.mark.parametrize("full", [False, True])
def test_ghost_fill_copy_identities(full):
    nx, ny, nz = 9, 8, 7
    rng = np.random.default_rng(3)
    s = po.Solver(nx, ny, nz, (-2, 2), (-1.75, 1.75), (-1.5, 1.5))
    N = nx * ny * nz
    st = dict(Q=random_state(rng, nx, ny, nz), Qh=random_state(rng, nx, ny, nz),
              pi=rng.normal(size=(10, N)), Pi=rng.normal(size=N))
    # interior boundary cells with Q^tau < 0: their ghost copies are zero
    for ix, iy, iz in ((2, 4, 3), (nx - 3, 5, 4), (3, 2, 2)):
        st["Q"][:, ix + nx * (iy + ny * iz)] *= -1.0
        st["Qh"][:, ix + nx * (iy + ny * iz)] *= -1.0
    s.set_state(st)
    s.ghost(full)
    out = s.get_state()
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    src = (clamp_idx(ix, nx) + nx * (clamp_idx(iy, ny) + ny * clamp_idx(iz, nz))).ravel()

    def copied(a, zero_rule):
        b = a[:, src].copy()
        if zero_rule:
            b[:, b[0] < 0] = 0.0
        return b

    exp_Q = copied(st["Q"], True)
    interior = src == np.arange(N)
    exp_Q[:, interior] = st["Q"][:, interior]  # interior cells are not touched
    np.testing.assert_array_equal(out["Q"], exp_Q)
    # at least one corner (ghost on all three axes) and the zeroed copies
    assert np.array_equal(out["Q"][:, 0], st["Q"][:, 2 + nx * (2 + ny * 2)])
    assert not out["Q"][:, 0 + nx * (4 + ny * 3)].any()
    if full:
        exp_pi = st["pi"][:, src]
        np.testing.assert_array_equal(out["pi"], exp_pi)
        np.testing.assert_array_equal(out["Pi"][0], st["Pi"][src])
        np.testing.assert_array_equal(out["Qh"], st["Qh"])
    else:
        exp_Qh = copied(st["Qh"], True)
        exp_Qh[:, interior] = st["Qh"][:, interior]
        np.testing.assert_array_equal(out["Qh"], exp_Qh)
        np.testing.assert_array_equal(out["pi"], st["pi"])
        np.testing.assert_array_equal(out["Pi"][0], st["Pi"])


# ------------------------------------------------------------------- updateM
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_update_m_planar_front_speed(axis):
    n = 12
    shape = [1, 1, 1]
    shape[axis] = n
    nx, ny, nz = shape
    lim = (-(n - 1) * 0.25, (n - 1) * 0.25)
    s = po.Solver(nx, ny, nz, lim if axis == 0 else (-1, 1), lim if axis == 1 else (-1, 1),
                  lim if axis == 2 else (-1, 1))
    m = np.zeros((3, n))
    m[:, :3] = 1.0  # matter in the first three cells
    s.set_state({"m": m})
    tau, dt = 2.0, 0.035
    r = dt / 0.5 if axis < 2 else dt / 0.5 / tau  # dx = 0.5 (dz: / tau)
    n1 = int(np.floor(0.9 / r)) + 1  # calls until m > 0.9 (then all axes = 1)
    for call in range(1, 3 * n1 + 1):
        s.update_m(tau, dt)
        mm = s.get_state()["m"]
        full = int(np.sum(mm.max(axis=0) >= 1.0))
        assert full == 3 + call // n1, (call, full)
        # the cell at the front holds its partial fill along the axis only
        front = 3 + call // n1
        k = call % n1
        np.testing.assert_allclose(mm[axis, front], k * r, rtol=1e-12, atol=0)
        others = [a for a in range(3) if a != axis]
        assert np.all(mm[others, front] == 0.0)
        assert np.all(mm[:, front + 1:] == 0.0)


def test_update_m_diagonal_neighbour():
    nx = ny = 5
    s = po.Solver(nx, ny, 1, (-1, 1), (-1, 1), (-1, 1))
    m = np.zeros((3, nx * ny))
    m[:, 0] = 1.0  # one full cell at (0, 0): (1, 1) is its diagonal neighbour
    s.set_state({"m": m})
    s.update_m(1.0, 0.05)
    mm = s.get_state()["m"].reshape(3, ny, nx)
    r = 0.05 / 0.5
    assert np.isclose(mm[0, 1, 1], 0.707 * r, rtol=1e-15, atol=0) and mm[2, 1, 1] == 0.0
    assert np.isclose(mm[1, 1, 1], 0.707 * r, rtol=1e-15, atol=0)
    assert np.isclose(mm[0, 0, 1], r, rtol=1e-15, atol=0) and mm[1, 0, 1] == 0.0  # along x
    assert np.isclose(mm[1, 1, 0], r, rtol=1e-15, atol=0) and mm[0, 1, 0] == 0.0  # along y
    assert np.all(mm[:, 2:, :] == 0) and np.all(mm[:, :, 2:] == 0)


# ---------------------------------------------------------- ISformal loop #2
def advect_case(nx, v, tau=1.0, dt=0.05, e=5.0, aux=None, seed=0, dz_axis=False):
    """1-D row along x (or eta): Qh of a uniform state with velocity v at
    tau - dt/2, auxiliary arrays aux (or random)"""
    rng = np.random.default_rng(seed)
    if dz_axis:
        s = po.Solver(1, 1, nx, (-1, 1), (-1, 1), (-0.5 * (nx - 1) * 0.1, 0.5 * (nx - 1) * 0.1))
    else:
        s = po.Solver(nx, 1, 1, (-0.5 * (nx - 1) * 0.5, 0.5 * (nx - 1) * 0.5), (-1, 1), (-1, 1))
    prim = np.zeros((nx, 8))
    prim[:, 0] = e
    prim[:, 1] = e / 3
    prim[:, 7 if dz_axis else 5] = v
    Qh = (po.p2c(prim) * (tau - dt / 2)).T.copy()
    a = aux if aux is not None else dict(pi0=rng.normal(size=(10, nx)),
                                         piH0=rng.normal(size=(10, nx)),
                                         Pi0=rng.normal(size=nx), PiH0=rng.normal(size=nx))
    s.set_state(dict(Qh=Qh, **a))
    s.advect(tau, dt)
    return s.get_state(), a


def test_advection_identity_uniform_and_shift():
    n = 10
    out, a = advect_case(n, 0.0)
    np.testing.assert_array_equal(out["pi"], a["pi0"])
    np.testing.assert_array_equal(out["piH"], a["piH0"])
    np.testing.assert_array_equal(out["Pi"][0], a["Pi0"])
    # uniform fields survive any velocity (the weights sum to 1)
    uni = dict(pi0=np.tile(np.arange(10.0)[:, None] * 0.01, n), piH0=np.full((10, n), 0.02),
               Pi0=np.full(n, -0.03), PiH0=np.full(n, 0.04))
    out, _ = advect_case(n, 0.37, aux=uni)
    np.testing.assert_allclose(out["pi"], uni["pi0"], rtol=1e-15, atol=1e-17)
    np.testing.assert_allclose(out["Pi"][0], uni["Pi0"], rtol=1e-15)
    # |v dt/dx| = 1 (v = 0.8, dt = dx / 0.8): the full step takes the upwind
    # neighbour's value, the half step averages the cell and that neighbour
    # (to the precision of the con2prim velocity, |dv| < 1e-5 then Newton)
    for v in (0.8, -0.8):
        dt = 0.5 / 0.8
        out, a = advect_case(n, v, tau=2.0, dt=dt)
        up = -1 if v > 0 else 1
        inner = np.arange(1, n - 1)
        np.testing.assert_allclose(out["pi"][:, inner], a["pi0"][:, inner + up], rtol=1e-8,
                                   atol=1e-9)
        np.testing.assert_allclose(out["piH"][:, inner],
                                   0.5 * (a["piH0"][:, inner] + a["piH0"][:, inner + up]),
                                   rtol=1e-8, atol=1e-9)


def test_advection_eta_weight_uses_half_step_tau():
    """along eta the Courant fraction is v dt / (deta (tau - dt/2)): with
    deta = 0.1, tau = 2.05 and dt = 0.1, v = 0.8 makes it 0.4 exactly"""
    n = 8
    tau, dt = 2.05, 0.1
    out, a = advect_case(n, 0.8, tau=tau, dt=dt, dz_axis=True)
    f = 0.8 * dt / (0.1 * (tau - dt / 2))
    i = np.arange(1, n - 1)
    np.testing.assert_allclose(out["pi"][:, i], (1 - f) * a["pi0"][:, i] + f * a["pi0"][:, i - 1],
                               rtol=1e-8, atol=1e-9)


def test_limiter_bounds_and_empty_cells():
    n = 10
    rng = np.random.default_rng(9)
    e = 2.0
    aux = dict(pi0=rng.normal(size=(10, n)) * 10.0, piH0=rng.normal(size=(10, n)) * 10.0,
               Pi0=rng.normal(size=n) * 5.0, PiH0=rng.normal(size=n) * 5.0)
    out, a = advect_case(n, 0.3, e=e, aux=aux)
    p = e / 3
    for c in range(n):
        P = np.zeros((4, 4))
        for i in range(4):
            for j in range(i + 1):
                P[i, j] = P[j, i] = out["pi"][i * (i + 1) // 2 + j, c]
        maxpi = np.abs(P).max()
        vx = 0.3
        g2 = 1 / (1 - vx * vx)
        maxT0 = max((e + p) * g2 - p, (e + p) * vx * vx * g2 + p)
        assert out["viscCorrCut"][0, c] < 1.0
        assert abs(maxpi - 0.1 * maxT0) <= 1e-8 * maxT0
        assert abs(out["Pi"][0, c]) <= p * (1 + 1e-8)
    # empty cells (e = 0) keep their old pi, Pi
    s = po.Solver(n, 1, 1, (-2.25, 2.25), (-1, 1), (-1, 1))
    old = dict(pi=rng.normal(size=(10, n)), Pi=rng.normal(size=n))
    s.set_state(dict(Qh=np.zeros((7, n)), **old, **aux))
    s.advect(1.0, 0.05)
    out = s.get_state()
    np.testing.assert_array_equal(out["pi"], old["pi"])
    np.testing.assert_array_equal(out["Pi"][0], old["Pi"])


# ------------------------------------------------------------ hlle closed forms
def interface_grid(axis, left, right, n=6, spacing=0.5, m_left=1.0, m_right=0.0):
    """n cells along axis: the first n/2 with conserved state `left`, the rest
    `right` (the MUSCL slopes vanish at the middle interface), m set per side"""
    shape = [1, 1, 1]
    shape[axis] = n
    lims = [(-1.0, 1.0)] * 3
    lims[axis] = (-(n - 1) * spacing / 2, (n - 1) * spacing / 2)
    s = po.Solver(*shape, *lims)
    Q = np.zeros((7, n))
    Q[:, :n // 2] = np.asarray(left)[:, None]
    Q[:, n // 2:] = np.asarray(right)[:, None]
    m = np.zeros((3, n))
    m[:, :n // 2] = m_left
    m[:, n // 2:] = m_right
    s.set_state(dict(Q=Q, Qh=Q.copy(), m=m))
    ijk = [0, 0, 0]
    ijk[axis] = n // 2 - 1
    return s, ijk


@pytest.mark.parametrize("mode", [0, 1])
def test_hlle_rest_next_to_vacuum(mode):
    e0, tau, dt = 6.0, 1.3, 0.05
    tf = tau + (0.25 if mode == 0 else 0.5) * dt
    dta = dt / 2 if mode == 0 else dt
    left = [tau * e0, 0, 0, 0, 0, 0, 0]
    # along x: b_l = -c_s (matter at rest), b_r = 1 (vacuum)
    s, (i, j, k) = interface_grid(0, left, [0] * 7)
    rc, f = s.hlle(i, j, k, 1, mode, tau, dt)
    pre = tf * dta / 0.5
    assert rc == 1
    np.testing.assert_allclose(f[0], pre * CS * e0 / (1 + CS), rtol=1e-14)
    np.testing.assert_allclose(f[1], pre * e0 / 3 / (1 + CS), rtol=1e-14)
    assert f[2] == f[3] == 0.0
    # along eta: b_l = -c_s/tau, b_r = 1/tau; fluxes / tauFactor, prefactor
    # tauFactor dta / deta
    s, (i, j, k) = interface_grid(2, left, [0] * 7, spacing=0.1)
    rc, f = s.hlle(i, j, k, 3, mode, tau, dt)
    pre = tf * dta / 0.1
    np.testing.assert_allclose(f[0], pre * CS * e0 / (tau * (1 + CS)), rtol=1e-14)
    np.testing.assert_allclose(f[3], dta / 0.1 * e0 / 3 / (1 + CS), rtol=1e-14)
    assert f[1] == f[2] == 0.0


def perfect_fluid_Q(e, v, tau, axis):
    prim = np.zeros((1, 8))
    prim[0, 0], prim[0, 1] = e, e / 3
    prim[0, 5 + axis] = v
    return (po.p2c(prim)[0] * tau)


@pytest.mark.parametrize("axis", [0, 1])
def test_hlle_equal_and_supersonic_states(axis):
    tau, dt = 1.0, 0.05
    pre = (tau + 0.25 * dt) * (dt / 2) / 0.5
    d = axis + 1
    # equal states: the physical flux (e + p) g^2 v (tau) and (e + p) g^2 v^2 + p
    for e, v in ((3.0, 0.2), (8.0, -0.6)):
        q = perfect_fluid_Q(e, v, tau, axis)
        s, ijk = interface_grid(axis, q, q, m_right=1.0)
        rc, f = s.hlle(*ijk, d, 0, tau, dt)
        w = 4 * e / 3 / (1 - v * v)
        np.testing.assert_allclose(f[0], pre * w * v, rtol=1e-9)
        np.testing.assert_allclose(f[d], pre * (w * v * v + e / 3), rtol=1e-9)
    # both states supersonic to the right (same v > c_s): the upwind flux of
    # the left state, whatever the right state holds
    v = 0.9
    qa, qb = perfect_fluid_Q(5.0, v, tau, axis), perfect_fluid_Q(1.0, v, tau, axis)
    s, ijk = interface_grid(axis, qa, qb, m_right=1.0)
    rc, f = s.hlle(*ijk, d, 0, tau, dt)
    w = 4 * 5.0 / 3 / (1 - v * v)
    np.testing.assert_allclose(f[0], pre * w * v, rtol=1e-9)
    np.testing.assert_allclose(f[d], pre * (w * v * v + 5.0 / 3), rtol=1e-9)


def test_hlle_skips():
    tau, dt = 1.0, 0.05
    # two vacuum cells: no flux
    s, ijk = interface_grid(0, [0] * 7, [0] * 7, m_left=1.0, m_right=1.0)
    assert s.hlle(*ijk, 1, 0, tau, dt)[0] == 0
    # matter on both sides but both partially filled (m < 1): no flux
    q = perfect_fluid_Q(3.0, 0.3, tau, 0)
    s, ijk = interface_grid(0, q, q * 0.5, m_left=0.4, m_right=0.2)
    assert s.hlle(*ijk, 1, 0, tau, dt)[0] == 0
    # one side full: flux
    s, ijk = interface_grid(0, q, q * 0.5, m_left=1.0, m_right=0.2)
    assert s.hlle(*ijk, 1, 0, tau, dt)[0] == 1


# ------------------------------------------------------------------ eta sweep
@pytest.mark.parametrize("tau,dt", [(1.0, 0.01), (0.6, 0.02)])
def test_eta_sweep_conservation_with_milne_sources(tau, dt):
    """1 x 1 x 40 column, compact bump at rest near both ends: the interior
    sums of Q^tau and Q^eta change over one step by dt times the Milne
    sources S^tau = -q^tau v^2 - p (1 + v^2), S^eta = -q^eta of the half-step
    state q = Qh / (tau + dt/2)"""
    nz, dz = 40, 0.05
    zl = (-(nz - 1) * dz / 2, (nz - 1) * dz / 2)
    z = np.linspace(*zl, nz)
    L = 0.6
    bump = np.where(np.abs(z) < L, np.cos(np.pi * z / (2 * L)) ** 2, 0.0)
    e, vz = 1 + 10 * bump, 0.4 * np.sin(np.pi * z / L) * bump
    s = po.Solver(1, 1, nz, (-1, 1), (-1, 1), zl)
    zero = np.zeros(nz)
    s.set_prim(tau, e, zero, zero, vz)
    a = s.get_state()
    rc, _ = s.run(tau, dt, 1)
    assert rc == 0
    b = s.get_state()
    inner = slice(2, nz - 2)
    q = (b["Qh"].T / (tau + dt / 2)).copy()
    pr, _, _ = po.c2p(q)
    S_t = -q[:, 0] * pr[:, 7] ** 2 - pr[:, 1] * (1 + pr[:, 7] ** 2)
    S_z = -q[:, 3]
    for k, S in ((0, S_t), (3, S_z)):
        dQ = np.sum(b["Q"][k, inner] - a["Q"][k, inner])
        scale = np.sum(np.abs(a["Q"][k, inner])) + np.sum(np.abs(S[inner])) * dt
        assert abs(dQ - dt * np.sum(S[inner])) <= 1e-14 * scale, k
    # and the sources are not small: a wrong time level (tau instead of
    # tau + dt/2) changes the Q^tau source sum by dt/(2 tau) of it
    assert abs(dt * np.sum(S_t[inner])) * dt / (2 * tau) > 1e-9 * np.sum(np.abs(a["Q"][0, inner]))


def test_eta_sweep_equals_x_sweep_at_large_tau():
    """at tau = 1000 a 1-D profile along eta with deta = dx/tau is the same
    problem as along x up to O(dt/tau) = 1e-5 per step (the Milne terms)"""
    n, dx, tau, dt = 40, 0.1, 1000.0, 0.01
    x = np.linspace(-(n - 1) * dx / 2, (n - 1) * dx / 2, n)
    e = 1 + 10 * np.exp(-x ** 2 / 0.1)
    v = 0.3 * np.tanh(x / 0.3)
    z = np.zeros(n)
    out = []
    for axis in (0, 2):
        if axis == 0:
            s = po.Solver(n, 1, 1, (x[0], x[-1]), (-1, 1), (-1, 1))
            s.set_prim(tau, e, v, z, z)
        else:
            s = po.Solver(1, 1, n, (-1, 1), (-1, 1), (x[0] / tau, x[-1] / tau))
            s.set_prim(tau, e, z, z, v)
        rc, t = s.run(tau, dt, 10)
        assert rc == 0
        pr, _, _ = po.c2p((s.get_state()["Q"].T / t).copy())
        out.append((pr[:, 0], pr[:, 5 if axis == 0 else 7]))
    (ex, vx), (ez, vz) = out
    assert np.max(np.abs(ex - ez)) <= 1e-4 * ex.max()
    assert np.max(np.abs(vx - vz)) <= 1e-4
    # the profile did evolve (the comparison is not of two initial states)
    assert np.max(np.abs(ex - e)) > 1e-2 * e.max()


# --------------------------------------------------------- pi stays traceless
def test_pi_traceless_in_3d_flow():
    nx, ny, nz = 16, 14, 10
    xl, yl, zl = (-3.75, 3.75), (-3.25, 3.25), (-1.8, 1.8)
    e, _, _, _ = setups.gaussian3d_ic(nx, ny, nz, xl, yl, zl, e0=30.0, sx=1.5, sy=1.2,
                                      eta_plateau=0.5, eta_sigma=0.6)
    X, Y, Z = setups.axes(nx, ny, nz, xl, yl, zl)
    s = po.Solver(nx, ny, nz, xl, yl, zl, trc=po.make_trc(etaS=0.2))
    s.set_prim(0.8, e + 0.5, 0.3 * np.tanh(X / 2), 0.25 * np.tanh(Y / 1.5), 0.3 * np.tanh(Z))
    rc, _ = s.run(0.8, 0.02, 15)
    assert rc == 0
    pi = s.get_state()["pi"]
    tr = pi[0] - pi[2] - pi[5] - pi[9]  # pi^00 - pi^11 - pi^22 - pi^33 (index44)
    inside = np.zeros((nz, ny, nx), bool)
    inside[2:-2, 2:-2, 2:-2] = True
    scale = np.max(np.abs(pi[:, inside.ravel()]))
    assert scale > 0.5
    assert np.max(np.abs(tr[inside.ravel()])) <= 1e-3 * scale


def test_pi_stays_transverse_in_an_accelerating_flow():
    """A uniform transverse flow in the Bjorken expansion accelerates (its
    momentum density is kept while e falls), so u turns every step; the
    second-order transversality terms rotate pi with it.  With pi relaxing
    from 0 toward the (transverse) Navier-Stokes value, pi^{mu nu} u_nu stays
    within 1.1 % of max |pi|; with the terms' sign flipped it reaches 3.3 %,
    without them 1.7 % (the scheme's own residual is about 0.7 %)."""
    n = 6
    N = n ** 3
    s = po.Solver(n, n, n, (-0.5, 0.5), (-0.5, 0.5), (-0.5, 0.5), trc=po.make_trc(etaS=0.2))
    s.set_prim(0.6, np.full(N, 10.0), np.full(N, 0.7), np.full(N, 0.5), np.zeros(N))
    rc, tau = s.run(0.6, 0.01, 40)
    assert rc == 0
    st = s.get_state()
    c = N // 2
    pr = po.c2p((st["Q"][:, c] / tau)[None, :])[0][0]
    v = pr[5:8]
    g = 1.0 / np.sqrt(1.0 - v @ v)
    u = np.array([g, g * v[0], g * v[1], g * v[2]])
    assert u[0] > 1.0 / np.sqrt(1.0 - 0.7 ** 2 - 0.5 ** 2) + 0.01  # the flow accelerated
    P = np.zeros((4, 4))
    for i in range(4):
        for j in range(i + 1):
            P[i, j] = P[j, i] = st["pi"][i * (i + 1) // 2 + j, c]
    gm = np.array([1.0, -1.0, -1.0, -1.0])
    scale = np.max(np.abs(P))
    assert scale > 0.3
    assert np.max(np.abs(P @ (gm * u))) <= 0.011 * scale


def test_pi_transverse_and_traceless_with_longitudinal_flow():
    """A uniform flow with v_x, v_y and a Milne velocity v_eta: the
    relaxation's Christoffel terms (the v_eta / tau terms of the half and the
    full step) keep pi transverse to u (within 2 % of max |pi|) and traceless
    (1e-3).  With all of them of the opposite sign pi^{mu nu} u_nu reaches
    20 %, without them 10 % (planted in a copy of the oracle on the CPU); a
    single off-diagonal term, small in this flow, is not resolved."""
    n = 6
    N = n ** 3
    s = po.Solver(n, n, n, (-0.5, 0.5), (-0.5, 0.5), (-0.5, 0.5), trc=po.make_trc(etaS=0.2))
    s.set_prim(0.6, np.full(N, 10.0), np.full(N, 0.4), np.full(N, 0.3), np.full(N, 0.5))
    rc, tau = s.run(0.6, 0.01, 40)
    assert rc == 0
    st = s.get_state()
    c = N // 2
    pr = po.c2p((st["Q"][:, c] / tau)[None, :])[0][0]
    v = pr[5:8]
    assert v[2] > 0.3  # the longitudinal velocity persists
    g = 1.0 / np.sqrt(1.0 - v @ v)
    u = np.array([g, g * v[0], g * v[1], g * v[2]])
    P = np.zeros((4, 4))
    for i in range(4):
        for j in range(i + 1):
            P[i, j] = P[j, i] = st["pi"][i * (i + 1) // 2 + j, c]
    gm = np.array([1.0, -1.0, -1.0, -1.0])
    scale = np.max(np.abs(P))
    assert scale > 0.3
    assert np.max(np.abs(P @ (gm * u))) <= 0.02 * scale
    assert abs(P[0, 0] - P[1, 1] - P[2, 2] - P[3, 3]) <= 1e-3 * scale


# ------------------------------------------------ freeze-out pi, Pi columns
def test_freezeout_pi_columns_reproduce_linear_fields():
    """pi^{mu nu} and Pi linear in (x, y, eta): the element records
    (src/fld.cpp:700-726, 656-797) carry the fields at the element's centroid
    exactly (trilinear interpolation is exact for linear data), pi boosted from
    Milne to Cartesian components by the boost of rapidity eta_C, built here
    as a matrix: L pi L^T with L = [[ch, sh], [sh, ch]] in the (t, z) block."""
    nx, ny, nz = 6, 6, 7
    N = nx * ny * nz
    sp, mins = (0.1, 0.1, 0.3), (-0.25, -0.25, -0.9)
    tau, dt = 1.2, 0.05
    X, Y, Z = setups.axes(nx, ny, nz, (mins[0], mins[0] + 5 * sp[0]),
                          (mins[1], mins[1] + 5 * sp[1]), (mins[2], mins[2] + 6 * sp[2]))
    x, y, eta = X.ravel(), Y.ravel(), Z.ravel()
    # an oblique level surface off the cube centres, so that no weight is 1/2
    e = 0.5 + 0.5 * (x - 0.023) + 0.2 * y + 0.03 * eta
    prim = np.zeros((N, 8))
    prim[:, 0], prim[:, 1], prim[:, 5] = e, e / 3, 0.2
    Q = (po.p2c(prim) * tau).T.copy()
    Qlow = (po.p2c(prim) * (tau - dt)).T.copy()
    rng = np.random.default_rng(21)
    coef = rng.normal(size=(11, 4))  # 10 pi components + Pi, each a + b x + c y + d eta
    lin = lambda k, xx, yy, ee: coef[k, 0] + coef[k, 1] * xx + coef[k, 2] * yy + coef[k, 3] * ee
    pi = np.stack([lin(k, x, y, eta) for k in range(10)])
    Pi = lin(10, x, y, eta)
    rows = po.freezeout((nx, ny, nz), sp, mins, Q, Qlow, pi, Pi, tau, dt, 0.5)
    assert len(rows) >= 6  # interior cubes: ix, iy, iz in [2, n - 2)
    assert np.ptp(rows[:, 3]) > 0.5  # elements at several rapidities
    # centroids off the cube centres along every axis
    assert np.all(np.abs(((rows[:, 1] - mins[0]) / sp[0]) % 1 - 0.5) > 0.01)
    assert np.any(np.abs(((rows[:, 2] - mins[1]) / sp[1]) % 1 - 0.5) > 0.01)
    assert np.any(np.abs(((rows[:, 3] - mins[2]) / sp[2]) % 1 - 0.5) > 0.01)
    for r in rows:
        xc, yc, ec = r[1], r[2], r[3]
        P = np.zeros((4, 4))
        for i in range(4):
            for j in range(i + 1):
                P[i, j] = P[j, i] = lin(i * (i + 1) // 2 + j, xc, yc, ec)
        L = np.eye(4)
        L[0, 0] = L[3, 3] = np.cosh(ec)
        L[0, 3] = L[3, 0] = np.sinh(ec)
        C = L @ P @ L.T
        got = r[16:26]
        want = np.array([C[i, j] for i in range(4) for j in range(i + 1)])
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)
        assert abs(r[26] - lin(10, xc, yc, ec)) <= 1e-12 * (1 + abs(r[26]))


def _uniform_milne_run(vx, vy, vz, n=6, steps=40):
    N = n ** 3
    s = po.Solver(n, n, n, (-0.5, 0.5), (-0.5, 0.5), (-0.5, 0.5), trc=po.make_trc(etaS=0.2))
    s.set_prim(0.6, np.full(N, 10.0), np.full(N, vx), np.full(N, vy), np.full(N, vz))
    rc, tau = s.run(0.6, 0.01, steps)
    assert rc == 0
    st = s.get_state()
    c = N // 2
    P = np.zeros((4, 4))
    for i in range(4):
        for j in range(i + 1):
            P[i, j] = P[j, i] = st["pi"][i * (i + 1) // 2 + j, c]
    pr = po.c2p((st["Q"][:, c] / tau)[None, :])[0][0]
    v = pr[5:8]
    g = 1.0 / np.sqrt(1.0 - v @ v)
    return P, np.array([g, g * v[0], g * v[1], g * v[2]])


def test_single_christoffel_terms_by_mirror():
    """The single off-diagonal Christoffel terms of the relaxation
    (src/hdo.cpp:785-788, 838-841: pi^{x tau} and pi^{y tau} fed by
    v_eta pi^{x eta}, pi^{y eta}; pi^{eta x}, pi^{eta y} by v_eta pi^{tau x},
    pi^{tau y}).  A uniform flow and its x <-> y mirror give mirrored pi and u
    to 1e-12, which fixes the x terms against the y terms.  Planted in copies
    of the oracle on the CPU: any one of the four terms dropped, or pi^{tau x}
    read in place of pi^{tau y}, fails it (pi off by 2e-3 and 9e-3 of
    max |pi|); the eta-x and eta-y pair dropped or flipped together fails the
    transversality test above; the tau-x and tau-y pair dropped or flipped
    together passes both and fails the boosted rest frame below."""
    Pa, ua = _uniform_milne_run(0.4, 0.15, 0.5)
    Pb, ub = _uniform_milne_run(0.15, 0.4, 0.5)
    perm = [0, 2, 1, 3]
    scale = np.max(np.abs(Pa))
    assert scale > 0.3 and abs(Pa[1, 3]) > 0.02 * scale and abs(Pa[0, 2]) > 0.02 * scale
    assert np.max(np.abs(Pa - Pb[np.ix_(perm, perm)])) <= 1e-12 * scale
    assert np.max(np.abs(ua - ub[perm])) <= 1e-12


def _boosted_rest_frame_pi(nz, dt, t_run=0.3, tau0=1.0, a=0.01, etas=0.2, e0=10.0):
    """A fluid at rest in Cartesian coordinates, seen on an eta grid
    (v_eta = -tanh eta, u^tau = cosh eta), carrying a homogeneous Cartesian
    pi^{xz} that relaxes as A exp(-t/tau_pi) along the proper time
    t = tau cosh eta.  Returns the largest deviations of pi^{tau x} from
    -A(t) sinh eta and of tau pi^{eta x} from A(t) cosh eta at |eta| <= 0.4,
    relative to max A."""
    eta = np.linspace(-1.0, 1.0, nz)
    cnst = (16 + 0.5 * 21.0 * 2.5) * np.pi ** 2 / 30.0 / 0.197326968 ** 3  # T = (e/cnst)^(1/4)
    taupi = 5 * etas / 5.068 / (e0 / cnst) ** 0.25  # 5 eta/(e + p), p = e/3
    s = po.Solver(1, 1, nz, (-1, 1), (-1, 1), (-1.0, 1.0), trc=po.make_trc(etaS=etas))
    z = np.zeros(nz)
    s.set_prim(tau0, np.full(nz, e0), z, z, -np.tanh(eta))
    A = a * (e0 / 3) * np.exp(-tau0 * np.cosh(eta) / taupi)
    pi = np.zeros((10, nz))
    pi[1], pi[7] = -A * np.sinh(eta), A * np.cosh(eta)  # index44(1,0), (3,1)
    s.set_state({"pi": pi, "piH": pi.copy()})
    rc, tau = s.run(tau0, dt, int(round(t_run / dt)))
    assert rc == 0
    st = s.get_state()
    A = a * (e0 / 3) * np.exp(-tau * np.cosh(eta) / taupi)
    inner = np.abs(eta) <= 0.41
    d1 = np.max(np.abs(st["pi"][1] + A * np.sinh(eta))[inner]) / np.max(A)
    d7 = np.max(np.abs(st["pi"][7] - A * np.cosh(eta))[inner]) / np.max(A)
    return d1, d7


def test_christoffel_terms_by_a_boosted_rest_frame():
    """The tau-x and eta-x Christoffel terms of the relaxation
    (src/hdo.cpp:779-788, 832-841) against the exact solution of a boosted
    rest frame: at a fixed eta the Cartesian pi^{xz} appears as
    pi^{tau x} = -A sinh eta and tau pi^{eta x} = A cosh eta, whose changes
    along tau come from the relaxation at gamma = cosh eta, the upwind
    advection with v_eta = -tanh eta and the v_eta / tau terms together.  The
    oracle lies 3.7e-3 and 5.0e-3 of max A away at deta = 0.05, dtau = 0.01,
    and 2.0e-3 at half of both (first order).  The tau-x and tau-y pair of
    terms dropped or flipped together (planted in copies of the oracle on the
    CPU, invisible to the mirror and transversality pins) ends 0.10 and 0.21
    away.  Q holds the static fluid's ideal T^{tau mu} only, so the fluid
    stays at rest (the reference's con2prim would read a pi^{tau x} in Q as
    flow)."""
    d1, d7 = _boosted_rest_frame_pi(41, 0.01)
    print(f"boosted rest frame: pi^tx {d1:.3e}, pi^etax {d7:.3e}", flush=True)
    assert d1 <= 0.01 and d7 <= 0.01
    h1, h7 = _boosted_rest_frame_pi(81, 0.005)
    assert h1 <= 0.7 * d1 and h7 <= 0.7 * d7
