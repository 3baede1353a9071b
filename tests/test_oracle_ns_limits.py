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

* the charged Bjorken flow from phi = 0: the relaxation time tau_pi and
  ISformal's own eta/s call (the sigma_NS of the tau_pipi term) evaluated at
  the cell's mu_B, which the Navier-Stokes limit cannot see.

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


def _charged_bjorken_ode(tab, trc, tau0, e0, nb0, tend):
    """DNMR Bjorken equations for phi = -pi^{eta eta} tau^2 in a charged fluid
    (n_b tau conserved), with the three eta/s evaluations of the relaxation
    kept apart:
      dphi/dtau = -(phi - 4 eta_NS/3tau)/tau_pi - (4/3) phi/tau
                  - (10/21) (eta_NS/eta_IS) phi/tau + (9/140) phi^2/(p tau_pi)
    eta_NS: NSquant's target pi_NS (src/hdo.cpp:432), eta_IS: the eta that
    ISformal divides pi_NS by to form sigma_NS for the tau_pipi term
    (src/hdo.cpp:730-737), tau_pi = (5/5.068) s eta/s / (e + p)
    (src/trancoeff.cpp:95, called at src/hdo.cpp:741).  Three systems are
    integrated together (one EoS and one eta/s call per right-hand side):
    y[0:2] the equations as stated, y[2:4] with tau_pi evaluated at mu_B = 0,
    y[4:6] with eta_IS evaluated at mu_B = 0."""
    from scipy.integrate import solve_ivp

    def rhs(t, y):
        nb = nb0 * tau0 / t
        e = y[0::2]
        ev = po.eos_eval(tab, e, np.full(3, nb))
        p, T, mub, s = ev[:, 0], ev[:, 1], ev[:, 2], ev[:, 5]
        inp = np.stack([np.tile(e, 2), np.full(6, nb), np.tile(T, 2),
                        np.concatenate([mub, 0 * mub]), np.tile(s, 2), np.tile(p, 2)], axis=1)
        es = po.trcoeff(trc, inp, tab)[:, 0]
        es_mu, es_0 = es[:3], es[3:]
        eta_ns = es_mu * s / 5.068
        eta_is = np.where([True, True, False], es_mu, es_0) * s / 5.068
        taupi = 5.0 / 5.068 * s * np.where([True, False, True], es_mu, es_0) / (e + p)
        phi = y[1::2]
        out = np.empty(6)
        out[0::2] = -(e + p - phi) / t
        out[1::2] = (-(phi - 4 / 3 * eta_ns / t) / taupi - 4 / 3 * phi / t
                     - 10 / 21 * (eta_ns / eta_is) * phi / t + 9 / 140 * phi ** 2 / (p * taupi))
        return out
    return solve_ivp(rhs, (tau0, tend), [e0, 0.0] * 3, rtol=1e-8, atol=1e-10, dense_output=True)


def test_charged_bjorken_relaxation_time_and_sigma():
    """One charged cell in Bjorken flow with the chiral EoS and eta/s raised
    17-34 % by mu_B (etaSparam 3, etaSScaleMuB 3): the oracle's Israel-Stewart
    step converges at first order in dtau to the charged DNMR Bjorken
    equations above (phi with a second-order part: error ratios 3.0, 2.65,
    2.38 for dtau halved from 0.02 to 0.0025).  Starting from phi = 0 the
    transient fixes tau_pi and the tau_pipi term's sigma_NS, so the mu_B passed
    to getTau (src/hdo.cpp:741) and to ISformal's own getEta
    (src/hdo.cpp:730) are pinned: the equations with either evaluated at
    mu_B = 0 end 1.0e-3 and 2.0e-2 away in phi, 26x and 500x the oracle's
    distance from the right ones at dtau = 0.0025 (3.9e-5).  Copies of the
    oracle with mu_B = 0 planted in either call fail this test and pass the
    charged Navier-Stokes limit above, which does not reach them."""
    import charged_eos as ce
    tab = ce.chiral("smooth")
    trc = po.make_trc(etaSparam=3, etaSMin=0.08, ah=0.0, al=0.0, T0=0.15, etaSShiftMuB=0.0,
                      etaSScaleMuB=3.0)
    tau0, e0, nb0, tend = 0.5, 50.0, 3.0, 2.5
    sol = _charged_bjorken_ode(tab, trc, tau0, e0, nb0, tend)
    errs = []
    for dt in (0.005, 0.0025):
        s = po.Solver(1, 1, 1, (-1, 1), (-1, 1), (-1, 1), trc=trc, table=tab)
        s.set_prim(tau0, np.array([e0]), np.zeros(1), np.zeros(1), np.zeros(1),
                   nb=np.array([nb0]))
        rc, tau = s.run(tau0, dt, int(round((tend - tau0) / dt)))
        assert rc == 0 and abs(tau - tend) < 1e-9
        st = s.get_state()
        e, phi = po.c2p((st["Q"] / tau).T.copy(), tab)[0][0, 0], -st["pi"][9, 0]
        y = sol.sol(tau)
        errs.append(max(abs(e - y[0]) / y[0], abs(phi - y[1]) / y[1]))
    print(f"charged Bjorken: relative errors {errs}", flush=True)
    assert errs[1] < 1e-4, errs
    assert 1.6 < errs[0] / errs[1] < 2.6, errs
    # mu_B = 0 in tau_pi or in eta_IS: far outside the oracle's error
    y = sol.sol(tend)
    for name, k in (("tau_pi", 3), ("eta_IS", 5)):
        dev = abs(y[k] - y[1]) / y[1]
        print(f"  mu_B = 0 in {name}: phi off by {dev:.3e}", flush=True)
        assert dev > 20 * errs[1], (name, dev, errs)
        assert abs(phi - y[k]) / y[k] > 10 * errs[1]
