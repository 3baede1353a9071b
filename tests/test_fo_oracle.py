"""CPU pins of the freeze-out oracle (oracle/fo_oracle.cpp: the Cornelius
construction and Fluid::outputSurface's element records).

* A level set of a linear field is a flat piece of a hyperplane: one element,
  its normal parallel to -grad e, its size the 3-volume of the cut (checked
  against the divergence theorem over the box facets, estimated by Monte
  Carlo); for a field along one axis the element is the whole cross-section
  box, exactly.
* The surface of a 4-D ball built from many hypercubes is closed: the
  element normals sum to zero.
* Mirroring a hypercube mirrors its element.
* Two diagonally opposite corners above the level give two elements.
* An element record of a uniform state carries that state (u, T, dsigma).
* The corona (Fluid::outputCorona) of a uniform static state below e_crit is
  every interior cube, each in closed form (centre, dsigma, u, T, dV_eff,
  EtotSurf); centres below e = 0.01 and cubes touching a hot cell are left out.
"""
import numpy as np
import pytest

from oracle import pyoracle as po

DX = np.array([0.05, 0.1, 0.12, 0.08])


def corners(f):
    """h[t][x][y][z] = f(t, x, y, z) at the 16 corners (coordinates from 0)."""
    h = np.zeros((2, 2, 2, 2))
    for it in range(2):
        for ix in range(2):
            for iy in range(2):
                for iz in range(2):
                    h[it, ix, iy, iz] = f(np.array([it, ix, iy, iz]) * DX)
    return h


def test_axis_plane_is_exact():
    s, x0 = 3.0, 0.037
    nel, cen, nrm = po.cornelius(corners(lambda p: 1.0 + s * (p[1] - x0)), DX, 1.0)
    assert nel[0] == 1
    vol = DX[0] * DX[2] * DX[3]
    assert np.allclose(nrm[0, 0], [0, -vol, 0, 0], rtol=0, atol=1e-15)
    assert np.allclose(cen[0, 0], [DX[0] / 2, x0, DX[2] / 2, DX[3] / 2], rtol=0, atol=1e-15)


def facet_volume_below(grad, c0, E0, axis, side, rng, n=400000):
    """3-volume of {e < E0} on the box facet x_axis = side * dx (Monte Carlo)."""
    p = rng.uniform(0, 1, (n, 4)) * DX
    p[:, axis] = side * DX[axis]
    below = (c0 + p @ grad) < E0
    return np.prod(np.delete(DX, axis)) * below.mean()


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_general_plane_direction_and_size(seed):
    rng = np.random.default_rng(seed)
    grad = rng.normal(size=4) / DX
    c0 = 0.3
    mid = c0 + grad @ (DX / 2)
    E0 = mid + 0.1 * rng.normal() * np.abs(grad) @ DX
    nel, cen, nrm = po.cornelius(corners(lambda p: c0 + grad @ p), DX, E0)
    assert nel[0] == 1
    n = nrm[0, 0]
    assert np.allclose(n / np.linalg.norm(n), -grad / np.linalg.norm(grad), atol=1e-12)
    # divergence theorem on {e < E0}: the element normal (into the region)
    # equals the sum of the outward facet normals times the facet parts inside
    expect = np.zeros(4)
    for ax in range(4):
        expect[ax] = (facet_volume_below(grad, c0, E0, ax, 1, rng) -
                      facet_volume_below(grad, c0, E0, ax, 0, rng))
    assert np.allclose(n, expect, atol=5e-3 * np.abs(expect).max())
    # the centroid lies on the plane
    assert abs(c0 + grad @ cen[0, 0] - E0) < 1e-12 * np.abs(grad) @ DX


def test_ball_surface_is_closed():
    R, m = 0.33, 10
    origin = -DX * m / 2
    hs = []
    for it in range(m):
        for ix in range(m):
            for iy in range(m):
                for iz in range(m):
                    base = origin + np.array([it, ix, iy, iz]) * DX
                    hs.append(corners(lambda p: R * R - np.sum((base + p) ** 2 / (DX / DX[1]) ** 2)))
    nel, cen, nrm = po.cornelius(np.array(hs), DX, 0.0)
    assert np.all(nel >= 0) and nel.sum() > 500
    total = np.zeros(4)
    mag = 0.0
    for i in np.nonzero(nel)[0]:
        for e in range(nel[i]):
            total += nrm[i, e]
            mag += np.linalg.norm(nrm[i, e])
    assert np.all(np.abs(total) < 1e-12 * mag), (total, mag)


def test_mirror_symmetry():
    rng = np.random.default_rng(5)
    for _ in range(200):
        h = rng.uniform(0, 1, (2, 2, 2, 2))
        nel, cen, nrm = po.cornelius(h, DX, 0.5)
        hm = h[:, ::-1, :, :].copy()
        nelm, cenm, nrmm = po.cornelius(hm, DX, 0.5)
        assert nel[0] == nelm[0]
        if nel[0] == 1:
            flip = np.array([1, -1, 1, 1])
            assert np.allclose(nrmm[0, 0], nrm[0, 0] * flip, atol=1e-14)
            assert np.allclose(cenm[0, 0], cen[0, 0] * flip + np.array([0, DX[1], 0, 0]),
                               atol=1e-14)


def test_two_separated_corners_give_two_elements():
    h = np.zeros((2, 2, 2, 2))
    h[0, 0, 0, 0] = h[1, 1, 1, 1] = 1.0
    nel, cen, nrm = po.cornelius(h, DX, 0.5)
    assert nel[0] == 2
    # each element's normal points away from its own corner
    for e in range(2):
        near0 = np.linalg.norm(cen[0, e]) < np.linalg.norm(cen[0, e] - DX)
        direction = 1.0 if near0 else -1.0
        assert np.all(direction * nrm[0, e] > 0)


def test_record_of_a_uniform_flow():
    """A uniform state with an e gradient along x: u, T and dsigma of the
    element follow from the state and the cut."""
    nx = ny = nz = 6
    N = nx * ny * nz
    tau, dt = 1.2, 0.05
    vx = 0.3
    x = np.arange(nx) * 0.1 - 0.25
    e = np.broadcast_to((0.5 + 0.5 * x)[None, None, :], (nz, ny, nx)).ravel()
    prim = np.zeros((N, 8))
    prim[:, 0] = e
    prim[:, 1] = e / 3
    prim[:, 5] = vx
    Q = (po.p2c(prim) * tau).T.copy()
    Qlow = (po.p2c(prim) * (tau - dt)).T.copy()
    pi = np.zeros((10, N))
    Pi = np.zeros(N)
    rows = po.freezeout((nx, ny, nz), (0.1, 0.1, 0.1), (-0.25, -0.25, -0.25), Q, Qlow, pi, Pi,
                        tau, dt, 0.5)
    # interior hypercubes ix, iy, iz in {2, 3}; the level crosses between
    # cells 2 and 3 in x: one element in each of the 4 (y, z) columns
    assert len(rows) == 4
    for r in rows:
        assert abs(r[1] - 0.0) < 1e-12  # x of the cut
        # the reference divides the interpolated Q (a mix of the levels
        # tau - dt and tau, here Q ~ tau) by tau + c0 instead of
        # tau - dt + c0 (src/fld.cpp:694-695): e at the centroid is scaled
        c0 = r[0] - tau
        eC = 0.5 * (tau - dt + c0) / (tau + c0)
        T = po.eos(np.array([eC]))[0, 1]
        assert abs(r[12] - T) < 1e-5 * T  # con2prim stops at |dv| < 1e-5
        # dsigma_mu = tau_C (0, -dt dy dz, 0, 0): the normal toward lower e
        assert abs(r[5] + r[0] * dt * 0.1 * 0.1) < 1e-14 and abs(r[4]) < 1e-14
        if abs(r[3]) < 1e-12:  # eta = 0: lab frame = Milne frame
            g = 1 / np.sqrt(1 - vx * vx)
            assert np.allclose(r[8:12], [g, g * vx, 0, 0], atol=1e-12)


def _uniform_Q(shape, e, tau):
    N = shape[0] * shape[1] * shape[2]
    Q = np.zeros((7, N))
    Q[0] = tau * e  # at rest: Q^tau = tau e
    return Q


def test_corona_of_a_uniform_state_in_closed_form():
    """Fluid::outputCorona on a static uniform state below e_crit: every
    interior cube is one element at its centre, dsigma = tau dV (cosh eta, 0,
    0, -sinh eta), u = (cosh eta, 0, 0, sinh eta) in the lab, T = T(e),
    dV_eff = tau dV, and the EtotSurf increment e tau dV cosh eta."""
    shape, sp, mins = (9, 8, 10), (0.3, 0.25, 0.2), (-1.2, -0.875, -0.9)
    tau, e = 1.3, 0.3
    N = np.prod(shape)
    rows, st = po.corona(shape, sp, mins, _uniform_Q(shape, e, tau), np.zeros((10, N)),
                         np.zeros(N), tau, 0.5, extend=True)
    n = (shape[0] - 4) * (shape[1] - 4) * (shape[2] - 4)
    assert len(rows) == n and st[0] == n
    dV = sp[0] * sp[1] * sp[2]
    ix, iy, iz = np.meshgrid(*[np.arange(2, s - 2) for s in shape], indexing="ij")
    x = mins[0] + (ix.ravel() + 0.5) * sp[0]
    y = mins[1] + (iy.ravel() + 0.5) * sp[1]
    eta = mins[2] + (iz.ravel() + 0.5) * sp[2]
    assert np.allclose(rows[:, 1], x, atol=1e-14) and np.allclose(rows[:, 2], y, atol=1e-14)
    assert np.allclose(rows[:, 3], eta, atol=1e-14) and np.all(rows[:, 0] == tau)
    ch, sh = np.cosh(eta), np.sinh(eta)
    assert np.allclose(rows[:, 4], tau * dV * ch, rtol=1e-14)
    assert np.allclose(rows[:, 7], -tau * dV * sh, rtol=1e-14, atol=1e-16)
    assert np.all(rows[:, 5:7] == 0.0)
    assert np.allclose(rows[:, 8], ch, rtol=1e-14) and np.allclose(rows[:, 11], sh, atol=1e-14)
    T = (e / ((16 + 0.5 * 21.0 * 2.5) * np.pi ** 2 / 30.0 / 0.197326968 ** 3)) ** 0.25
    assert np.allclose(rows[:, 12], T, rtol=1e-14)
    assert np.allclose(rows[:, 27], e, rtol=1e-14)
    assert st[1] == pytest.approx(n * tau * dV, rel=1e-12)
    assert st[2] == pytest.approx(e * tau * dV * np.sum(ch), rel=1e-12)


def test_corona_excludes_tails_and_cubes_touching_hot_cells():
    shape, sp, mins = (10, 9, 8), (0.3, 0.3, 0.3), (-1.35, -1.2, -1.05)
    tau = 1.0
    N = np.prod(shape)
    z = np.zeros((10, N)), np.zeros(N)
    tail, _ = po.corona(shape, sp, mins, _uniform_Q(shape, 0.009, tau), *z, tau, 0.5)
    assert len(tail) == 0  # centre e < 0.01
    Q = _uniform_Q(shape, 0.2, tau)
    hot = np.zeros(shape[::-1], dtype=bool)
    hot[3:5, 4:6, 4:7] = True  # [iz, iy, ix]
    Q[0, hot.ravel()] = tau * 2.0
    rows, st = po.corona(shape, sp, mins, Q, *z, tau, 0.5)
    cubes = 0
    for ix in range(2, shape[0] - 2):
        for iy in range(2, shape[1] - 2):
            for iz in range(2, shape[2] - 2):
                cubes += not hot[iz:iz + 2, iy:iy + 2, ix:ix + 2].any()
    assert len(rows) == cubes < (shape[0] - 4) * (shape[1] - 4) * (shape[2] - 4)
