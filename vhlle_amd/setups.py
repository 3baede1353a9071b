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
