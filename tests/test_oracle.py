"""Pins the CPU oracle to mathematics that does not depend on it.

The reference cannot be built in this image (every hot-path file includes GSL
headers that are absent), and it ships no tests or golden vectors, so each
oracle function is checked here against closed forms, exact recurrences,
textbook equations and symmetries:

  con2prim  -- closed-form conformal root (golden fixture, 50-digit decimal),
               special branches, round trip with prim2con, brute-force bisection
  ideal step -- Gubser's analytic flow (second-order convergence), exact
               Bjorken recurrence of the predictor-corrector source terms,
               mirror symmetry
  viscous step -- Israel-Stewart Bjorken ODE with the DNMR coefficients
               (lambda = 38/21 and the phi_7 term), first-order convergence
  EoS / transport coefficients -- Stefan-Boltzmann gas, tau_pi T = 5 eta/s
"""
import json
import os

import numpy as np
import pytest

from vhlle_amd import setups

HERE = os.path.dirname(os.path.abspath(__file__))
CNST = (16 + 0.5 * 21.0 * 2.5) * np.pi ** 2 / 30.0 / 0.197326968 ** 3


# --------------------------------------------------------------- con2prim
def test_c2p_against_closed_form_golden(oracle):
    g = json.load(open(os.path.join(HERE, "golden", "c2p_conformal.json")))
    Q = np.array([c["Q"] for c in g["cases"]])
    prim, it, _ = oracle.c2p(Q)
    for c, row, n in zip(g["cases"], prim, it):
        ex = {k: float(v) for k, v in c["exact"].items()}
        v = np.sqrt(row[5] ** 2 + row[6] ** 2 + row[7] ** 2)
        # transformPV stops at |dv| < 1e-5 (src/rmn.cpp:96)
        assert abs(v - ex["v"]) < 1e-5, (c["Q"], v, ex["v"])
        # e = Q_t - M v and p = e/3 hold for the returned iterate
        assert abs(row[0] - ex["e"]) <= 1e-5 * c["Q"][0] + 1e-15
        assert row[1] == row[0] / 3.0
        assert 0 < n <= 100 or ex["v"] == 0.0


def test_c2p_special_branches(oracle):
    Q = np.array([
        [0.0, 1.0, 0.0, 0.0, 1.0, 0.0, 0.0],    # Q_t <= 0 -> vacuum
        [-2.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0],
        [3.0, 0.0, 0.0, 0.0, 0.5, 0.2, 0.1],    # M = 0 -> at rest, n = Q_n
        [1.0, 3.0, 4.0, 0.0, 0.0, 0.0, 0.0],    # |Q_vec| = 5 > Q_t -> rescaled
    ])
    prim, it, Qa = oracle.c2p(Q)
    assert np.all(prim[0] == 0) and np.all(prim[1] == 0)
    assert list(prim[2]) == [3.0, 1.0, 0.5, 0.2, 0.1, 0.0, 0.0, 0.0]
    M = np.hypot(Qa[3, 1], Qa[3, 2])
    assert abs(M - 0.9999) < 1e-15 and abs(Qa[3, 1] / Qa[3, 2] - 0.75) < 1e-15
    v = np.hypot(prim[3, 5], prim[3, 6])
    vex = 3 * 0.9999 / (2 + np.sqrt(4 - 3 * 0.9999 ** 2))
    assert abs(v - vex) < 1e-5


def test_c2p_round_trip_and_bisection(oracle):
    rng = np.random.default_rng(3)
    n = 2000
    prim = np.zeros((n, 8))
    prim[:, 0] = 10 ** rng.uniform(-4, 3, n)
    prim[:, 1] = prim[:, 0] / 3
    prim[:, 2] = rng.uniform(0, 2, n)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    vm = rng.uniform(0, 0.99, n)
    prim[:, 5:8] = vm[:, None] * d
    Q = oracle.p2c(prim)
    # textbook T^{tau mu} and N^tau of a perfect fluid
    g2 = 1 / (1 - vm ** 2)
    w = prim[:, 0] + prim[:, 1]
    np.testing.assert_allclose(Q[:, 0], w * g2 - prim[:, 1], rtol=1e-13)
    np.testing.assert_allclose(Q[:, 1:4], (w * g2)[:, None] * prim[:, 5:8], rtol=1e-12,
                               atol=1e-300)
    np.testing.assert_allclose(Q[:, 4], prim[:, 2] * np.sqrt(g2), rtol=1e-13)
    back, it, _ = oracle.c2p(Q)
    v = np.linalg.norm(back[:, 5:8], axis=1)
    # brute-force bisection on the same scalar equation to 1e-15
    Qt, M = Q[:, 0], np.linalg.norm(Q[:, 1:4], axis=1)
    lo, hi = np.zeros(n), np.ones(n)
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        f = (Qt + (Qt - M * mid) / 3) * mid - M
        hi = np.where(f > 0, mid, hi)
        lo = np.where(f > 0, lo, mid)
    assert np.max(np.abs(v - 0.5 * (lo + hi))) < 1e-5
    assert np.max(np.abs(v - vm)) < 1e-5


# ----------------------------------------------------------- ideal hydro
def _gubser_error(oracle, n, tau_end=2.0):
    cfg = setups.GUBSER_2D
    X, Y, _ = setups.axes(n, n, 1, cfg["xlim"], cfg["ylim"], cfg["zlim"])
    e, vx, vy = setups.gubser(1.0, X, Y, amp=10.0)
    s = oracle.Solver(n, n, 1, cfg["xlim"], cfg["ylim"], cfg["zlim"])
    s.set_prim(1.0, e, vx, vy, np.zeros_like(e))
    dx = 10.0 / (n - 1)
    dt = dx / 4
    steps = int(round((tau_end - 1.0) / dt))
    rc, tau = s.run(1.0, dt, steps)
    assert rc == 0
    Q = s.get_state()["Q"].T / tau
    prim, _, _ = oracle.c2p(Q)
    ea, vxa, vya = setups.gubser(tau, X, Y, amp=10.0)
    r = np.sqrt(X * X + Y * Y).ravel()
    m = r < 3.0
    ee, ea = prim[:, 0], ea.ravel()
    l1_e = np.sum(np.abs(ee[m] - ea[m])) / np.sum(ea[m])
    l1_v = np.mean(np.abs(prim[:, 5] - vxa.ravel())[m])
    return l1_e, l1_v


def test_gubser_convergence(oracle):
    """Gubser's analytic conformal flow (Phys. Rev. D 82, 085027), normalised
    as ICGubser (src/icGubser.cpp:32-43): the L1 errors of e and v_x within
    r < 3 fm shrink with resolution (MUSCL-minmod + HLLE: between first and
    second order; a missing source term or a wrong time level stalls this)."""
    errs = [_gubser_error(oracle, n) for n in (41, 81, 161)]
    (e1, v1), (e2, v2), (e3, v3) = errs
    assert e3 < 0.015 and v3 < 0.002, errs
    assert e1 / e2 > 1.8 and e2 / e3 > 1.8, errs
    assert v1 / v2 > 2.5 and v2 / v3 > 2.5, errs


def test_bjorken_ideal_exact_recurrence(oracle):
    """One cell, no interfaces: the step reduces to the midpoint rule on the
    Milne source S_t = -p (src/hdo.cpp:307-353): Qh = Q - dt/2 * (Q/tau)/3,
    Q' = Q - dt * (Qh/(tau + dt/2))/3, bit for bit; and e ~ tau^(-4/3)."""
    s = oracle.Solver(1, 1, 1, (-1, 1), (-1, 1), (-1, 1))
    tau, dt, e0 = 0.6, 0.01, 12.0
    s.set_prim(tau, np.array([e0]), np.zeros(1), np.zeros(1), np.zeros(1))
    Q = tau * (e0 + (e0 / 3.0) * 0.0) * 1.0
    t = tau
    for k in range(200):
        p = (Q / t) / 3.0
        Qh = Q + (-0.0 - p * (1.0 + 0.0)) * (dt / 2.0)
        ph = (Qh / (t + 0.5 * dt)) / 3.0
        Q = Q + (-0.0 - ph * (1.0 + 0.0)) * dt
        t = t + dt
    rc, tau_o = s.run(tau, dt, 200)
    assert rc == 0 and tau_o == t
    Qo = s.get_state()["Q"][0, 0]
    assert Qo == Q
    e = Q / t
    assert abs(e * t ** (4 / 3) / (e0 * tau ** (4 / 3)) - 1) < 1e-5


def test_mirror_symmetry(oracle):
    """x -> -x maps the solution onto itself with vx -> -vx."""
    nx, ny, nz = 16, 10, 6
    xl, yl, zl = (-3.75, 3.75), (-2.25, 2.25), (-1.0, 1.0)
    X, Y, Z = setups.axes(nx, ny, nz, xl, yl, zl)
    e = 20 * np.exp(-X ** 2 / 2 - Y ** 2 / 3 - Z ** 2)
    vx = 0.3 * np.tanh(X)
    vy = 0.1 * np.sin(Y)
    vz = 0.1 * np.tanh(Z)
    s = oracle.Solver(nx, ny, nz, xl, yl, zl, trc=oracle.make_trc(etaS=0.1))
    s.set_prim(0.6, e, vx, vy, vz)
    rc, _ = s.run(0.6, 0.02, 10)
    assert rc == 0
    Q = s.get_state()["Q"].reshape(7, nz, ny, nx)
    Qm = Q[:, :, :, ::-1].copy()
    Qm[1] *= -1
    assert np.max(np.abs(Q - Qm)) < 1e-12 * np.max(np.abs(Q))


# --------------------------------------------------------- viscous hydro
def _bjorken_ode(etas, tau0, e0, tend):
    from scipy.integrate import solve_ivp

    def rhs(t, y):
        e, phi = y
        p = e / 3
        T = (e / CNST) ** 0.25
        s = (e + p) / T
        eta = etas * s / 5.068
        taupi = 5 * eta / (e + p)
        return [-(e + p - phi) / t,
                -(phi - 4 / 3 * eta / t) / taupi - 38 / 21 * phi / t
                + 9 / 140 * phi ** 2 / (p * taupi)]
    return solve_ivp(rhs, (tau0, tend), [e0, 0.0], rtol=1e-11, atol=1e-13, dense_output=True)


def test_bjorken_viscous_ode(oracle):
    """One cell (pure longitudinal Bjorken expansion): the Israel-Stewart step
    (src/hdo.cpp:682-963) converges at first order in dtau to the DNMR
    Bjorken equations for phi = -pi^{eta eta} tau^2:
      de/dtau = -(e + p - phi)/tau,
      dphi/dtau = -(phi - 4 eta/3tau)/tau_pi - (38/21) phi/tau
                  + (9/140) phi^2/(p tau_pi),  tau_pi = 5 eta/(e+p),
    i.e. delta_pipi = 4/3, tau_pipi = 10/7, phi_7 = 9/(70 p) of
    src/trancoeff.h:28-33; pi^{xx} = pi^{yy} = phi/2 by tracelessness."""
    etas, tau0, e0, tend = 0.08, 0.5, 10.0, 2.5
    sol = _bjorken_ode(etas, tau0, e0, tend)
    errs = []
    for dt in (0.01, 0.005):
        s = oracle.Solver(1, 1, 1, (-1, 1), (-1, 1), (-1, 1), trc=oracle.make_trc(etaS=etas))
        s.set_prim(tau0, np.array([e0]), np.zeros(1), np.zeros(1), np.zeros(1))
        rc, tau = s.run(tau0, dt, int(round((tend - tau0) / dt)))
        assert rc == 0
        st = s.get_state()
        e, phi = st["Q"][0, 0] / tau, -st["pi"][9, 0]
        # x <-> y symmetry exactly; tracelessness to the scheme's accuracy
        assert st["pi"][2, 0] == st["pi"][5, 0]
        assert abs(st["pi"][2, 0] - phi / 2) < 1e-3 * phi
        ee, pp = sol.sol(tau)
        errs.append(max(abs(e - ee) / ee, abs(phi - pp) / pp))
    assert errs[1] < 1e-4, errs
    assert 1.6 < errs[0] / errs[1] < 2.6, errs


# ----------------------------------------------- EoS and transport coeffs
def test_eos_stefan_boltzmann(oracle):
    e = np.array([0.5, 1.0, 10.0, 100.0])
    out = oracle.eos(e)
    T = (e / CNST) ** 0.25
    np.testing.assert_allclose(out[:, 0], e / 3, rtol=0)
    np.testing.assert_allclose(out[:, 1], T, rtol=1e-15)
    np.testing.assert_allclose(out[:, 2], 4 * e / (3 * T), rtol=1e-14)  # s = (e+p)/T
    # degrees of freedom of the reference's SIMPLE EoS: 16 + 21*2.5/2 = 42.25
    assert abs(CNST * 0.197326968 ** 3 * 30 / np.pi ** 2 - 42.25) < 1e-12


def test_transport_coefficients(oracle):
    e = np.array([0.05, 1.0, 10.0])
    T = (e / CNST) ** 0.25
    s = 4 * e / (3 * T)
    inp = np.stack([e, np.zeros(3), T, np.zeros(3), s, e / 3], axis=1)
    out = oracle.trcoeff(oracle.make_trc(etaS=0.08), inp)
    # tau_pi = 5 eta / (e+p) = 5 (eta/s) / T in fm, zero below T = 0.1 GeV
    hot = T > 0.1
    np.testing.assert_allclose(out[hot, 2] * T[hot] * 5.068, 5 * 0.08, rtol=1e-13)
    assert np.all(out[~hot, 2] == 0)
    # zeta/s param 0 vanishes for a conformal EoS (1/3 - cs^2 = 0)
    out0 = oracle.trcoeff(oracle.make_trc(etaS=0.08, zetaS=0.1, zetaSparam=0), inp)
    assert np.all(out0[:, 1] == 0)
    # param 5: peak value zetaS0 at T = zetaSPeakT, falls to exp(-1/2) one width away
    Tk = np.array([0.170, 0.160, 0.200])
    inp5 = np.stack([CNST * Tk ** 4, np.zeros(3), Tk, np.zeros(3), np.ones(3), np.ones(3)], 1)
    z = oracle.trcoeff(oracle.make_trc(zetaS=0.05, zetaSparam=5), inp5)[:, 1]
    np.testing.assert_allclose(z, 0.05 * np.exp(-0.5) ** np.array([0, 1, 1]), rtol=1e-14)
