"""CPU pins of the oracle's viscous terms in their Navier-Stokes limits, where
the Israel-Stewart quantities relax, within a few relaxation times, to closed
forms:

* transverse shear: a boost-invariant fluid at uniform e with a linear
  v_x(y) drives pi^{xy} to 2 eta sigma^{xy} = -eta d_y u^x, eta = (eta/s) s /
  5.068 (hbar c in GeV fm), with the slope read off the state itself (the
  viscous longitudinal expansion steepens it by 25 % over the run; a linear
  profile makes any centred difference exact).  This fixes the sign and
  factor of the shear term, the projector's transverse components and the
  y-derivative of NSquant, which the one-cell Bjorken check does not reach.
* bulk: Bjorken flow with zeta/s constant drives Pi to -zeta theta with
  theta = 1/tau, zeta = (zeta/s) s / 5.068, and relaxation time
  tau_Pi = 6 zeta / (5.068 s T) (src/trancoeff.cpp:96).

* the charge-dependent eta/s (etaSparam 3, src/trancoeff.cpp:51-62) with the
  chiral EoS: the same shear flow at uniform n_b drives pi^{xy} to -eta d_y u^x
  with eta/s evaluated at the cell's T, mu_B, s and p, taken here from the EoS
  lookup at the cell's (e, n_b); with etaSScaleMuB = 3 mu_B raises eta/s by
  22 %, so NSquant's eta/s call with a mu_B of 0, or with mu_Q, fails
  (planted in a copy of the oracle on the CPU).

The residual differences are the second-order terms and the lag behind a
slowly moving target, O(tau_relax / tau); the tolerances leave room for them
and no more (a wrong sign or a factor 2 fails by far).
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from vhlle_amd import setups

CNST = (16 + 0.5 * 21.0 * 2.5) * np.pi ** 2 / 30.0 / 0.197326968 ** 3


def T_of(e):
    return (e / CNST) ** 0.25


def test_transverse_shear_navier_stokes_limit():
    nx, ny = 9, 21
    xl, yl = (-0.8, 0.8), (-2.0, 2.0)
    X, Y, _ = setups.axes(nx, ny, 1, xl, yl, (-1, 1))
    kappa, etaS = 0.02, 0.08
    e = np.full(X.shape, 200.0)
    z = np.zeros_like(e)
    s = po.Solver(nx, ny, 1, xl, yl, (-1, 1), trc=po.make_trc(etaS=etaS))
    s.set_prim(2.0, e, kappa * Y, z, z)
    rc, tau = s.run(2.0, 0.02, 100)
    assert rc == 0
    st = s.get_state()
    pr = po.c2p((st["Q"] / tau).T.copy())[0]
    ee = pr[:, 0].reshape(ny, nx)
    vx = pr[:, 5].reshape(ny, nx)
    ux = vx / np.sqrt(1.0 - vx ** 2)
    dux = np.gradient(ux, (yl[1] - yl[0]) / (ny - 1), axis=0)
    T = T_of(ee)
    sden = (ee + ee / 3.0) / T
    eta = etaS * sden / 5.068
    pixy = st["pi"][4].reshape(ny, nx)  # index44(1, 2) = 4
    rows = slice(ny // 2 - 5, ny // 2 + 6)
    want = -eta * dux
    rel = np.abs(pixy[rows, 2:-2] / want[rows, 2:-2] - 1.0)
    print(f"pi^xy / (-eta d_y u^x) - 1: max {rel.max():.3e}", flush=True)
    assert rel.max() < 0.03
    taupi = 5.0 * etaS * 0.197327 / T.mean()  # tau_pi T = 5 eta/s, in fm
    assert taupi < 0.05 * tau  # the limit is reached: relaxation fast against tau


def test_bulk_navier_stokes_limit_in_bjorken_flow():
    zs = 0.05
    trc = po.make_trc(etaS=0.0, zetaS=zs, zetaSparam=5, zetaSPeakT=0.5, zetaSWidthMinus=50.0,
                      zetaSWidthPlus=50.0)  # zeta/s = 0.05 at every T reached here
    s = po.Solver(1, 1, 1, (-1, 1), (-1, 1), (-1, 1), trc=trc)
    e0, tau0, dt = 300.0, 4.0, 0.01
    s.set_prim(tau0, np.full((1, 1, 1), e0), *[np.zeros((1, 1, 1))] * 3)
    rc, tau = s.run(tau0, dt, 200)
    assert rc == 0
    st = s.get_state()
    pr = po.c2p((st["Q"] / tau).T.copy())[0]
    e = pr[0, 0]
    T = T_of(e)
    sden = (e + e / 3.0) / T
    Pi = st["Pi"][0, 0]
    want = -zs * sden / 5.068 / tau
    taubulk = 6.0 * zs / (5.068 * T)
    print(f"Pi / (-zeta/tau) - 1 = {Pi / want - 1:.3e}; tau_Pi/tau = {taubulk / tau:.3e}",
          flush=True)
    assert Pi / want == pytest.approx(1.0, abs=0.06)
    assert taubulk < 0.03 * tau


def test_charged_eta_over_s_navier_stokes_limit():
    import charged_eos as ce
    tab = ce.chiral("smooth")
    nx, ny = 9, 21
    xl, yl = (-0.8, 0.8), (-2.0, 2.0)
    X, Y, _ = setups.axes(nx, ny, 1, xl, yl, (-1, 1))
    kappa = 0.02
    e = np.full(X.shape, 200.0)
    z = np.zeros_like(e)
    trc = po.make_trc(etaSparam=3, etaSMin=0.08, ah=0.0, al=0.0, T0=0.15, etaSShiftMuB=0.0,
                      etaSScaleMuB=3.0)
    s = po.Solver(nx, ny, 1, xl, yl, (-1, 1), trc=trc, table=tab)
    s.set_prim(2.0, e, kappa * Y, z, z, nb=np.full(X.shape, 5.0))
    rc, tau = s.run(2.0, 0.02, 100)
    assert rc == 0
    st = s.get_state()
    pr = po.c2p((st["Q"] / tau).T.copy(), tab)[0]
    ee, nb = pr[:, 0], pr[:, 2]
    ux = (pr[:, 5] / np.sqrt(1.0 - pr[:, 5] ** 2)).reshape(ny, nx)
    dux = np.gradient(ux, (yl[1] - yl[0]) / (ny - 1), axis=0)
    ev = po.eos_eval(tab, ee, nb)
    p, T, mub = ev[:, 0], ev[:, 1], ev[:, 2]
    sden = (ee + p - mub * nb) / T  # EoS::s with n_q = n_s = 0
    etaS = po.trcoeff(trc, np.stack([ee, nb, T, mub, sden, p], axis=1), tab)[:, 0]
    etaS0 = po.trcoeff(trc, np.stack([ee, nb, T, 0 * mub, sden, p], axis=1), tab)[:, 0]
    assert np.min(etaS / etaS0) > 1.15  # mu_B matters here
    eta = (etaS * sden / 5.068).reshape(ny, nx)
    pixy = st["pi"][4].reshape(ny, nx)  # index44(1, 2) = 4
    rows = slice(ny // 2 - 5, ny // 2 + 6)
    rel = np.abs(pixy[rows, 2:-2] / (-eta * dux)[rows, 2:-2] - 1.0)
    print(f"charged: pi^xy / (-eta d_y u^x) - 1: max {rel.max():.3e}", flush=True)
    assert rel.max() < 0.05
