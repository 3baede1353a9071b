"""CPU pins of the tabulated EoS (the TABLE branch of EoSs, src/eos.cpp:27-120):
the oracle's natural cubic splines (oracle/vh_oracle.cpp, restating GSL's
gsl_interp_cspline, which is not in the reference tree) and the library's host
table (vh_eos_table, host logic only: no GPU call).

* The oracle spline equals scipy's natural CubicSpline (an independent
  implementation) to 1e-12 of the function's scale, passes through the nodes
  exactly, is C1 at the nodes, and reproduces linear data to rounding.
* The library's spline table evaluates to the oracle's values bit for bit.
* A table sampled from the ideal gas p = e/3 gives the ideal gas's con2prim
  and, for an inviscid flow (where only p(e) enters), the same hydro steps.
"""
import numpy as np
import pytest
from scipy.interpolate import CubicSpline

from oracle import pyoracle as po
from vhlle_amd import abi, setups


class Tab:
    def __init__(self, e, p, T, mu=None):
        self.e, self.p, self.T, self.mu = e, p, T, mu


@pytest.fixture(scope="module")
def cross():
    return Tab(*setups.crossover_eos_columns())


def test_spline_matches_scipy_natural(cross):
    rng = np.random.default_rng(3)
    v = np.concatenate([rng.uniform(0, 2.0, 20000), np.exp(rng.uniform(-18, 9, 20000))])
    v = v[v <= cross.e[-1]]
    for f, y in ((0, cross.p), (1, cross.T)):
        ours = po.eos_spline(cross, f, v)
        ref = CubicSpline(cross.e, y, bc_type="natural")(v)
        scale = np.abs(y).max()
        assert np.max(np.abs(ours - ref)) <= 1e-12 * scale, (f, np.max(np.abs(ours - ref)))


def test_spline_nodes_and_continuity(cross):
    for f, y in ((0, cross.p), (1, cross.T)):
        assert np.array_equal(po.eos_spline(cross, f, cross.e), y)
        # one-sided derivatives at interior nodes by tiny steps agree (C1)
        x = cross.e[100:700:37]
        h = 1e-7 * x
        left = (po.eos_spline(cross, f, x) - po.eos_spline(cross, f, x - h)) / h
        right = (po.eos_spline(cross, f, x + h) - po.eos_spline(cross, f, x)) / h
        assert np.allclose(left, right, rtol=1e-4, atol=1e-9 * np.abs(y).max() / x.min()), f


def test_linear_data_reproduced():
    e = np.concatenate(([0.0], np.geomspace(1e-6, 1e3, 300)))
    t = Tab(e, e / 3.0, 0.1 + 0.0 * e)
    v = np.geomspace(1e-6, 1e3, 5000)
    assert np.allclose(po.eos_spline(t, 0, v), v / 3.0, rtol=1e-13, atol=1e-300)
    assert np.allclose(po.eos_spline(t, 1, v), 0.1, rtol=1e-14)


def test_library_table_equals_oracle(cross):
    tab = abi.EosTable(cross.e, cross.p, cross.T)
    v = np.exp(np.random.default_rng(4).uniform(-18, 9, 20000))
    for f in (0, 1):
        assert np.array_equal(tab.spline(f, v), po.eos_spline(cross, f, v))


def test_negative_pressure_clamped():
    e = np.linspace(0.0, 1.0, 11)
    p = e / 3.0 - 0.05  # the reference clamps p < 0 to 0 (src/eos.cpp:49)
    t = Tab(e, p, 0.1 + e)
    assert np.array_equal(po.eos_spline(t, 0, e), np.maximum(p, 0.0))
    lib = abi.EosTable(e, p, 0.1 + e)
    assert np.array_equal(lib.spline(0, e), np.maximum(p, 0.0))


def ideal_table():
    e = np.concatenate(([0.0], np.geomspace(1e-8, 1e4, 1500)))
    T = po.eos(e)[:, 1]
    return Tab(e, e / 3.0, T)


def test_ideal_table_con2prim_and_eos():
    t = ideal_table()
    rng = np.random.default_rng(6)
    prim = np.zeros((3000, 8))
    prim[:, 0] = np.exp(rng.uniform(-8, 6, 3000))
    prim[:, 1] = prim[:, 0] / 3
    v = 0.95 * rng.uniform(0, 1, 3000) ** 0.5
    d = rng.normal(size=(3000, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    prim[:, 5:8] = v[:, None] * d
    Q = po.p2c(prim)
    a, _, _ = po.c2p(Q)
    b, _, _ = po.c2p(Q, t)
    assert np.allclose(a, b, rtol=1e-12, atol=1e-14 * np.abs(a).max())
    e = np.geomspace(1e-4, 1e3, 200)
    ea, eb = po.eos(e), po.eos(e, t)
    assert np.allclose(eb[:, 0], ea[:, 0], rtol=1e-13)        # p
    assert np.allclose(eb[:, 3], 1.0 / 3.0, rtol=1e-10)       # cs2 = dpe of a linear p
    assert np.allclose(eb[:, 1], ea[:, 1], rtol=2e-4)         # T: spline of e^(1/4)


def test_ideal_table_inviscid_steps_equal_ideal():
    """Inviscid Gubser flow: only p(e) enters the step, and the spline of the
    linear p = e/3 is exact to rounding, so the two runs agree to 1e-13."""
    nx = ny = 31
    xl = yl = (-3.0, 3.0)
    X, Y, _ = setups.axes(nx, ny, 1, xl, yl, (-1, 1))
    e, vx, vy = setups.gubser(1.0, X, Y, amp=10.0)
    out = []
    for t in (None, ideal_table()):
        s = po.Solver(nx, ny, 1, xl, yl, (-1, 1), table=t)
        s.set_prim(1.0, e, vx, vy, 0 * e)
        rc, _ = s.run(1.0, 0.025, 10)
        assert rc == 0
        out.append(s.get_state()["Q"])
    assert np.allclose(out[1], out[0], rtol=0, atol=1e-13 * np.abs(out[0]).max())
