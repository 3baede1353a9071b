"""Host-side initial states for the benchmark workloads and tests (numpy).

Arrays are returned in the cell order of the C-ABI: shape (nz, ny, nx) in C
order, so that .ravel() gives cell = ix + nx*(iy + ny*iz).  Coordinates follow
Fluid::getX/getY/getZ (src/fld.h:62-64): x_i = xmin + i*dx with
dx = (xmax-xmin)/(nx-1).
"""
import numpy as np


def axes(nx, ny, nz, xlim, ylim, zlim):
    x = xlim[0] + np.arange(nx) * ((xlim[1] - xlim[0]) / (nx - 1))
    y = ylim[0] + np.arange(ny) * ((ylim[1] - ylim[0]) / (ny - 1))
    if nz > 1:
        z = zlim[0] + np.arange(nz) * ((zlim[1] - zlim[0]) / (nz - 1))
    else:
        z = np.zeros(1)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return X, Y, Z


def gubser(tau, X, Y, q=1.0, amp=1.0):
    """Gubser's conformal ideal flow (Phys. Rev. D 82, 085027) with the
    normalisation of ICGubser (src/icGubser.cpp:32-43): amp=1 gives
    e(tau=1, r=0) = 1 GeV/fm^3.  Returns e, vx, vy."""
    r = np.sqrt(X * X + Y * Y)
    D = 1.0 + 2.0 * q * q * (tau * tau + r * r) + (q * q * (tau * tau - r * r)) ** 2
    e = amp * (4.0 * q * q / (D * tau)) ** (4.0 / 3.0)
    k = 2.0 * q * q * tau * r / (1.0 + q * q * tau * tau + q * q * r * r)
    vx = X / (r + 1e-50) * k
    vy = Y / (r + 1e-50) * k
    return e, vx, vy


# ---- workload definitions (sizes as the benchmark configurations state) ----

GUBSER_2D = dict(name="gubser_2d_101x101", nx=101, ny=101, nz=1, xlim=(-5.0, 5.0),
                 ylim=(-5.0, 5.0), zlim=(-1.0, 1.0), tau0=1.0, dt=0.025, steps=200,
                 etaS=0.0, amp=10.0)

VISCOUS_3D = dict(name="viscous_3d_200x200x100", nx=200, ny=200, nz=100, xlim=(-9.95, 9.95),
                  ylim=(-9.95, 9.95), zlim=(-2.475, 2.475), tau0=0.5, dt=0.025, steps=300,
                  etaS=0.08, e0=50.0, sx=2.5, sy=3.0, eta_plateau=1.5, eta_sigma=1.0)


def gubser_ic(cfg=GUBSER_2D, nx=None):
    """Config-1 style Gubser IC (e(tau0, 0) = amp); optional resolution nx."""
    n = nx or cfg["nx"]
    X, Y, _ = axes(n, n, 1, cfg["xlim"], cfg["ylim"], cfg["zlim"])
    e, vx, vy = gubser(cfg["tau0"], X, Y, amp=cfg["amp"])
    return e, vx, vy, np.zeros_like(e)


def gaussian3d_ic(nx, ny, nz, xlim, ylim, zlim, e0=50.0, sx=2.5, sy=3.0, eta_plateau=1.5,
                  eta_sigma=1.0):
    """Smooth 3-D state: e = e0 exp(-x^2/2sx^2 - y^2/2sy^2) H(eta), H = 1 on the
    plateau |eta| < eta_plateau and Gaussian tails of width eta_sigma; v = 0."""
    X, Y, Z = axes(nx, ny, nz, xlim, ylim, zlim)
    a = np.abs(Z)
    H = np.where(a < eta_plateau, 1.0, np.exp(-((a - eta_plateau) ** 2) / (2 * eta_sigma ** 2)))
    e = e0 * np.exp(-X * X / (2 * sx * sx) - Y * Y / (2 * sy * sy)) * H
    z = np.zeros_like(e)
    return e, z, z.copy(), z.copy()


# Event-by-event batch: 256 boost-invariant (nz = 1) grids with sides drawn
# uniformly from 100..160 cells (dx = dy = 0.1 fm), hot-spot initial states
# (N_part ~ Poisson(150), 2-D Woods-Saxon R = 6.6 fm, a = 0.55 fm, Gaussian
# spots sigma = 0.5 fm, Gamma(1) weights; amp = 5.2 GeV/fm^3 per unit weight
# gives <e_max> ~ 40 GeV/fm^3), eta/s = 0.12, zeta/s peaked at 0.05 at
# T = 170 MeV (zetaSparam 5), tau0 = 0.5 fm, dtau = 0.025 fm, 200 steps.
BATCH_EBE = dict(name="batch_ebe_256x(100..160)^2", events=256, side_lo=100, side_hi=160, dx=0.1,
                 tau0=0.5, dt=0.025, steps=200, etaS=0.12, zetaS=0.05, npart_mean=150.0,
                 R=6.6, a=0.55, sigma=0.5, amp=5.2, seed=20240501)


def batch_grids(cfg=BATCH_EBE, events=None, seed=None):
    """(nx, ny, 1, xlim, ylim, zlim) of each event: sides uniform in
    side_lo..side_hi from numpy's default_rng(seed), centred boxes."""
    rng = np.random.default_rng(cfg["seed"] if seed is None else seed)
    n = cfg["events"] if events is None else events
    sides = rng.integers(cfg["side_lo"], cfg["side_hi"] + 1, size=n)
    out = []
    for s in sides:
        h = (int(s) - 1) * cfg["dx"] / 2
        out.append((int(s), int(s), 1, (-h, h), (-h, h), (-0.5, 0.5)))
    return out


def batch_trc(cfg=BATCH_EBE):
    from vhlle_amd import abi
    return abi.make_trc(etaS=cfg["etaS"], zetaS=cfg["zetaS"], zetaSparam=5)


# Full 3-D viscous run with freeze-out: 300 x 300 x 150 cells, dx = dy = 0.08 fm,
# d(eta) = 0.04 (x, y in [-11.96, 11.96], eta in [-2.98, 2.98]), hot-spot
# initial state as the batch (same generator and amplitude) extended in eta
# with a plateau |eta| < 2 and Gaussian tails of width 1.0, eta/s = 0.08 and
# the T-peaked zeta/s, tau0 = 0.5 fm, dtau = 0.02 fm (= tau0 d(eta): no
# sub-steps), 600 steps; Cornelius surface at e_sw = 0.5 every 5 steps.
FULL_3D_FO = dict(name="full_3d_fo_300x300x150", nx=300, ny=300, nz=150, xlim=(-11.96, 11.96),
                  ylim=(-11.96, 11.96), zlim=(-2.98, 2.98), tau0=0.5, dt=0.02, steps=600,
                  fo_every=5, e_sw=0.5, etaS=0.08, zetaS=0.05, eta_plateau=2.0, eta_sigma=1.0,
                  npart_mean=150.0, R=6.6, a=0.55, sigma=0.5, amp=5.2, seed=20240502)


def full3d_ic(cfg=FULL_3D_FO):
    from vhlle_amd import abi
    g = [(cfg["nx"], cfg["ny"], cfg["nz"], cfg["xlim"], cfg["ylim"], cfg["zlim"])]
    e, npart = abi.hotspot_ic(g, seed=cfg["seed"], npart_mean=cfg["npart_mean"], R=cfg["R"],
                              a=cfg["a"], sigma=cfg["sigma"], amp=cfg["amp"],
                              eta_plateau=cfg["eta_plateau"], eta_sigma=cfg["eta_sigma"])
    z = np.zeros_like(e)
    return (e, z, z.copy(), z.copy()), int(npart[0])


# A crossover equation of state in the TABLE format of EoSs (columns e, p, T):
# entropy density s = (2 pi^2/45) g(T) T^3 / (hbar c)^3 with g rising from
# g_h (a massless pion gas) to g_q = 42.25 (the ideal gas of the SIMPLE
# branch: e = cnst T^4 at high T) as (1 + tanh((T - Tc)/dT))/2; p = int s dT,
# e = T s - p (thermodynamically consistent, so s = (e + p)/T as EoS::s
# computes it).  Rows log-spaced in e from e_min to e_max plus e = 0.
HBARC = 0.197326968
CROSSOVER_EOS = dict(Tc=0.170, dT=0.020, g_h=3.0, g_q=42.25, n=800, e_min=1e-8, e_max=1e4)


def crossover_eos_columns(cfg=CROSSOVER_EOS):
    Tc, dT, gh, gq = cfg["Tc"], cfg["dT"], cfg["g_h"], cfg["g_q"]
    k = 2 * np.pi ** 2 / 45 / HBARC ** 3
    T = np.geomspace(1e-4, 5.0, 400001)
    g = gh + (gq - gh) * 0.5 * (1 + np.tanh((T - Tc) / dT))
    s = k * g * T ** 3
    # p(T_0) of the pion gas below the grid, then the trapezoid integral of s
    p = k * gh * T[0] ** 4 / 4 + np.concatenate(([0.0], np.cumsum(0.5 * (s[1:] + s[:-1]) *
                                                                  np.diff(T))))
    e = T * s - p
    et = np.concatenate(([0.0], np.geomspace(cfg["e_min"], cfg["e_max"], cfg["n"] - 1)))
    Tt = np.interp(et, e, T, left=0.0)
    pt = np.interp(et, e, p, left=0.0)
    Tt[0] = pt[0] = 0.0
    return et, pt, Tt


def crossover_eos(cfg=CROSSOVER_EOS):
    from vhlle_amd import abi
    return abi.EosTable(*crossover_eos_columns(cfg))


def write_eos_table(path, cfg=CROSSOVER_EOS):
    """The table file EoSs reads (three columns e p T)."""
    e, p, T = crossover_eos_columns(cfg)
    np.savetxt(path, np.column_stack([e, p, T]), fmt="%.17e")
