"""CPU pins of the charge-dependent EoSs: EoSChiral (bilinear tables,
src/eoChiral.cpp:99-154) and EoSHadron (log-trilinear table,
src/eoHadron.cpp:78-208), restated in the oracle (oracle/vh_oracle.cpp
ChiralAux, HadronTable, Eos) and in the library's lookups (vh_device.h
ChiralTab, HadronTab), whose host build vh_eos_eval(where=0) runs here
without a GPU.

The pins: interpolation reproduces the table at its nodes and reproduces
bilinear (chiral) / trilinear-in-the-mapped-coordinates (hadron) data
everywhere, including the extrapolation of the end cells; the chiral region
switches (small / big table / power law with its closed form) at the
reference's thresholds; p < 0 is clamped; e <= 0 gives zeros; below e_min
the hadron EoS scales the centre node of the first e row; a node with status
1 zeroes the lookups of its cell only; the entropy is (e + p - mu.n)/T; the
hadron table is symmetric in n_b -> -n_b.  con2prim with these EoSs inverts
prim2con, and a Bjorken flow (a uniform grid at rest, where the scheme is
the midpoint rule for d(tau e)/dtau = -p(e, n), tau n = const) converges at
second order to the closed-form solution for p = c e + d n_b.  Finally the
library's lookups (host build) equal the oracle's bit for bit.
"""
import numpy as np
import pytest

import charged_eos as ce
from oracle import pyoracle as po
from vhlle_amd import abi

LIN_CHIRAL = [(0.01, 0.30, 0.05, 0.02), (0.12, 0.02, 0.01, 0.003), (0.0, 0.01, 0.5, 0.01),
              (0.0, 0.002, 0.03, 0.001)]
LIN_HADRON = [(5.0, 0.3, 0.05, 0.02, 0.001), (0.2, 0.01, 0.003, 0.002, 0.0005),
              (0.0, 0.01, 0.1, 0.02, 0.001), (0.0, 0.003, 0.01, 0.05, 0.0002),
              (0.0, 0.001, 0.02, 0.01, 0.0001)]


@pytest.fixture(scope="module")
def chiral_lin():
    return ce.chiral("linear", LIN_CHIRAL)


@pytest.fixture(scope="module")
def hadron_lin():
    return ce.hadron("linear", LIN_HADRON, nnq=31)


def lin_chiral(e, nb, k):
    a, b, c, d = LIN_CHIRAL[k]
    return a + b * e + c * nb + d * e * nb


def test_chiral_bilinear_data_exact(chiral_lin):
    """at the nodes and at random points (also below n_min = 0 and in the
    last cells, where the end cell's bilinear form extrapolates) the lookup
    equals the bilinear field; p, T, mu_b, mu_s are the columns 0, 1, 2, 4."""
    rng = np.random.default_rng(1)
    t = chiral_lin.small
    E, N = t.nodes()
    ie, inn = rng.integers(0, t.ne, 300), rng.integers(0, t.nn, 300)
    e = np.concatenate([E[ie, inn], rng.uniform(0, 1.46, 3000), rng.uniform(0, 146.0, 3000)])
    nb = np.concatenate([N[ie, inn], rng.uniform(-0.3, 0.3, 3000), rng.uniform(-1.0, 6.0, 3000)])
    keep = (e < 146.0) & (nb < 6.0)
    e, nb = e[keep], nb[keep]
    out = po.eos_eval(chiral_lin, e, nb)
    for k, col in ((0, 0), (1, 1), (2, 2), (3, 4)):
        ref = lin_chiral(e, nb, k)
        if k == 0:
            ref = np.maximum(ref, 0.0)
        assert np.max(np.abs(out[:, col] - ref)) <= 1e-12 * (1 + np.abs(ref).max()), k
    assert np.all(out[:, 3] == 0.0)  # mu_q = 0 always


def test_chiral_regions_and_power_law():
    """small table where e < 1.46 and n_b < 0.3, big where e < 146 and
    n_b < 6 (strict inequalities), p = 0.2964 e, T = 0.15120476935 e^(1/4),
    mu = 0 beyond; e < 0 gives zeros"""
    small_coef = [(1.0, 0, 0, 0), (0.1, 0, 0, 0), (0.01, 0, 0, 0), (0.001, 0, 0, 0)]
    big_coef = [(2.0, 0, 0, 0), (0.2, 0, 0, 0), (0.02, 0, 0, 0), (0.002, 0, 0, 0)]
    tab = ce.chiral("linear", (small_coef, big_coef))
    e = np.array([1.0, 1.46, 1.0, 100.0, 146.0, 10.0, 200.0, -1e-3])
    nb = np.array([0.1, 0.1, 0.3, 1.0, 1.0, 6.0, 0.0, 0.1])
    out = po.eos_eval(tab, e, nb)
    assert list(out[:4, 0]) == [1.0, 2.0, 2.0, 2.0]
    assert list(out[:4, 1]) == [0.1, 0.2, 0.2, 0.2]
    for i in (4, 5, 6):
        assert out[i, 0] == 0.2964 * e[i]
        assert abs(out[i, 1] - 0.15120476935 * e[i] ** 0.25) <= 1e-16 * out[i, 1]
        assert out[i, 2] == 0.0 and out[i, 4] == 0.0
    assert np.all(out[7, :5] == 0.0) and out[7, 5] == 0.0


def test_chiral_pressure_clamp():
    """p < 0 from the interpolation is clamped to 0; T is not"""
    tab = ce.chiral("linear", [(-0.1, 0.2, 0, 0), (-0.05, 0.1, 0, 0), (0, 0, 0, 0),
                               (0, 0, 0, 0)])
    e = np.array([0.1, 0.4, 1.0])
    out = po.eos_eval(tab, e, np.full(3, 0.1))
    assert list(out[:, 0]) == [0.0, 0.0, pytest.approx(0.1, rel=1e-14)]
    assert out[0, 1] == pytest.approx(-0.05 + 0.1 * 0.1, rel=1e-12)


def _hadron_points(h, rng, n):
    """random (e, nb, nq) from mapped coordinates inside the table, and the
    coordinates"""
    dxe, dxb, dxq = h.dx()
    X = rng.uniform(h.logemin, h.logemax, n)
    Y = rng.uniform(-h.lognmax, h.lognmax, n)
    Z = rng.uniform(-h.lognmax, h.lognmax, n)
    Y[: n // 4] = rng.uniform(-1.5 * dxb, 1.5 * dxb, n // 4)  # the linear region near 0
    return (h.e0 * np.exp(X), ce.mapped_to_density(Y, h.n0, h.lognmax, dxb),
            ce.mapped_to_density(Z, h.n0, h.lognmax, dxq), X, Y, Z)


def test_hadron_trilinear_data_exact(hadron_lin):
    """lookups of data trilinear in (log(e/e0), mapped nb, mapped nq) --
    a + b X + c Y + d Z + g X Y Z -- equal the data at random points and at
    the nodes; this pins the log mapping of e, the log/linear mapping of the
    charges, the index3 layout and the weights"""
    h = hadron_lin
    rng = np.random.default_rng(2)
    e, nb, nq, X, Y, Z = _hadron_points(h, rng, 6000)
    Xn, Yn, Zn = h.axes()
    i, j, k = rng.integers(0, h.ne, 200), rng.integers(0, h.nnb, 200), rng.integers(0, h.nnq, 200)
    _, dxb, dxq = h.dx()
    e = np.concatenate([e, h.e0 * np.exp(Xn[i])])
    nb = np.concatenate([nb, ce.mapped_to_density(Yn[j], h.n0, h.lognmax, dxb)])
    nq = np.concatenate([nq, ce.mapped_to_density(Zn[k], h.n0, h.lognmax, dxq)])
    X, Y, Z = np.concatenate([X, Xn[i]]), np.concatenate([Y, Yn[j]]), np.concatenate([Z, Zn[k]])
    out = po.eos_eval(h, e, nb, nq)
    for f, col in zip(range(5), (0, 1, 2, 3, 4)):
        a, b, c, d, g = LIN_HADRON[f]
        ref = a + b * X + c * Y + d * Z + g * X * Y * Z
        err = np.max(np.abs(out[:, col] - ref))
        assert err <= 1e-11 * (1 + np.abs(ref).max()), (f, err)
    # at the nodes exactly the stored values (weights 1 and 0)
    m = slice(6000, None)
    assert np.allclose(out[m, 0], h.p[i, j, k], rtol=1e-13, atol=0)


def test_hadron_below_e_min_and_zero():
    """e < e_min: p and T scale the node (0, nnb/2, nnq/2) by e / e_min, mu = 0
    (even sizes, so that nnb/2 is not the middle of the axis); e <= 0: zeros"""
    h = ce.hadron("smooth", ne=20, nnb=12, nnq=10)
    h.p[0] = np.arange(12 * 10).reshape(12, 10) * 0.01 + 0.5
    h.T[0] = np.arange(12 * 10).reshape(12, 10) * 0.001 + 0.05
    e_min = h.e0 * np.exp(h.logemin)
    e = np.array([0.5 * e_min, 0.25 * e_min, 0.0, -1.0])
    out = po.eos_eval(h, e, np.array([0.3, -0.2, 0.1, 0.1]), np.array([0.1, 0.0, 0.0, 0.0]))
    assert out[0, 0] == 0.5 * e_min / e_min * h.p[0, 6, 5]
    assert out[1, 1] == 0.25 * e_min / e_min * h.T[0, 6, 5]
    assert np.all(out[:2, 2:5] == 0.0)
    assert np.all(out[2:, :] == 0.0)


def test_hadron_status_zeroes_its_cell_only(hadron_lin):
    """status 1 at node (ie, inb, inq) zeroes lookups whose base cell is that
    node; lookups of the cell below, which use it as a corner, are not zero"""
    h = ce.hadron("linear", LIN_HADRON)
    Xn, Yn, Zn = h.axes()
    dxe, dxb, dxq = h.dx()
    ie, ib, iq = 50, 30, 25
    h.status[ie, ib, iq] = 1
    pts = [(ie + 0.5, ib + 0.5, iq + 0.5), (ie - 0.5, ib - 0.5, iq - 0.5)]
    X = np.array([h.logemin + a * dxe for a, _, _ in pts])
    Y = np.array([-h.lognmax + b * dxb for _, b, _ in pts])
    Z = np.array([-h.lognmax + c * dxq for _, _, c in pts])
    out = po.eos_eval(h, h.e0 * np.exp(X), ce.mapped_to_density(Y, h.n0, h.lognmax, dxb),
                      ce.mapped_to_density(Z, h.n0, h.lognmax, dxq))
    assert np.all(out[0, :5] == 0.0)
    assert out[1, 0] > 1.0


def test_hadron_symmetric_in_baryon_density():
    h = ce.hadron("smooth")
    rng = np.random.default_rng(4)
    e, nb, nq, *_ = _hadron_points(h, rng, 2000)
    a = po.eos_eval(h, e, nb, nq)
    b = po.eos_eval(h, e, -nb, nq)
    assert np.max(np.abs(a[:, 0] - b[:, 0]) / a[:, 0]) <= 1e-13
    assert np.max(np.abs(a[:, 2] + b[:, 2])) <= 1e-13 * np.abs(a[:, 2]).max()


@pytest.mark.parametrize("which", ["chiral", "hadron"])
def test_entropy(which):
    """s = (e + p - mu_b n_b - mu_q n_q - mu_s n_s) / T (src/eos.cpp:18-25)"""
    tab = ce.chiral() if which == "chiral" else ce.hadron()
    rng = np.random.default_rng(5)
    e = rng.uniform(0.01, 3.0, 500)
    nb, nq, ns = rng.uniform(-0.2, 0.25, 500), rng.uniform(-0.1, 0.1, 500), rng.uniform(-0.05,
                                                                                        0.05, 500)
    o = po.eos_eval(tab, e, nb, nq, ns)
    ref = (e + o[:, 0] - o[:, 2] * nb - o[:, 3] * nq - o[:, 4] * ns) / o[:, 1]
    assert np.max(np.abs(o[:, 5] - ref) / np.abs(ref)) <= 1e-14


@pytest.mark.parametrize("which", ["chiral", "hadron"])
def test_con2prim_inverts_prim2con(which):
    """transformPV with n_b refreshed every iteration inverts transformCV to
    the iteration's tolerance; the final p is evaluated with the last
    iterate's n_b (src/rmn.cpp:122-126), so it differs from p(e, n_b) at the
    final n_b by a little (but not by more than the tolerance)"""
    tab = ce.chiral() if which == "chiral" else ce.hadron()
    rng = np.random.default_rng(6)
    n = 3000
    e = rng.uniform(0.05, 5.0, n)
    nb = rng.uniform(-0.2, 0.28, n)
    nq = rng.uniform(-0.1, 0.1, n) if which == "hadron" else np.zeros(n)
    v = rng.uniform(0, 0.95, n)
    th = rng.uniform(0, 2 * np.pi, n)
    vx, vy = v * np.cos(th), v * np.sin(th)
    p = po.eos_eval(tab, e, nb, nq)[:, 0]
    prim = np.stack([e, p, nb, nq, np.zeros(n), vx, vy, np.zeros(n)], axis=1)
    Q = po.p2c(prim)
    out, it, _ = po.c2p(Q, tab)
    assert np.all(it > 0)
    assert np.max(np.abs(out[:, 0] - e) / e) < 2e-4
    assert np.max(np.abs(out[:, 2] - nb)) < 2e-4 * np.abs(nb).max()
    assert np.max(np.abs(out[:, 5] - vx)) < 2e-5
    p_final = po.eos_eval(tab, out[:, 0], out[:, 2], out[:, 3])[:, 0]
    d = np.abs(out[:, 1] - p_final) / out[:, 1]
    assert np.max(d) < 1e-4 and np.count_nonzero(d > 1e-10) > n // 10


def _bjorken(tab, e0, n0, dt, tau1):
    s = po.Solver(5, 5, 1, (-1, 1), (-1, 1), (-0.5, 0.5), trc=po.make_trc(0.0, 0.0), table=tab)
    N = 25
    z = np.zeros(N)
    s.set_prim(1.0, np.full(N, e0), z, z, z, nb=np.full(N, n0))
    steps = int(round((tau1 - 1.0) / dt))
    rc, tau = s.run(1.0, dt, steps)
    assert rc == 0 and abs(tau - tau1) < 1e-9
    return s.get_state()["Q"][:, 12] / tau


def test_bjorken_flow_converges_to_closed_form():
    """p = c e + d n_b (exactly bilinear in the chiral small table): the
    Bjorken solution e(tau) = [e0 tau0^(1+c) - d n0 tau0 (tau^c - tau0^c)/c]
    / tau^(1+c), n = n0 tau0 / tau; the scheme converges at second order"""
    c, d = 0.3, 0.05
    tab = ce.chiral("linear", [(0.0, c, d, 0.0), (0.15, 0.0, 0.0, 0.0), (0.0, 0.0, 0.1, 0.0),
                               (0.0, 0.0, 0.0, 0.0)])
    e0, n0, tau1 = 1.2, 0.25, 1.6
    ref = (e0 - d * n0 * (tau1 ** c - 1.0) / c) / tau1 ** (1 + c)
    errs = []
    for dt in (0.04, 0.02, 0.01):
        q = _bjorken(tab, e0, n0, dt, tau1)
        assert abs(q[4] - n0 / tau1) <= 1e-14  # tau n = const (Q_nb / tau)
        errs.append(abs(q[0] - ref) / ref)
    assert errs[2] < 2e-5, errs
    assert 3.0 < errs[0] / errs[1] < 5.0 and 3.0 < errs[1] / errs[2] < 5.0, errs


@pytest.mark.parametrize("which", ["chiral", "hadron"])
def test_library_host_lookups_equal_oracle(which):
    """vh_eos_eval(where=0): the device lookups compiled for the host, on the
    descriptor's arrays, equal the oracle's restatement bit for bit (no GPU
    call)"""
    if which == "chiral":
        t = ce.chiral()
        desc = abi.EosChiral(*(abi.ChiralTable(x.emin, x.emax, x.nmin, x.nmax, x.p, x.T, x.mub,
                                               x.mus) for x in (t.small, t.big)))
        rng = np.random.default_rng(7)
        e = np.concatenate([rng.uniform(-0.1, 2.0, 4000), rng.uniform(0, 200.0, 4000)])
        nb = np.concatenate([rng.uniform(-0.4, 0.4, 4000), rng.uniform(-1.0, 7.0, 4000)])
        nq = np.zeros_like(e)
    else:
        t = ce.hadron()
        t.status[40:45, 10:12, 20] = 1
        desc = abi.EosHadron(t.e0, t.n0, t.logemin, t.logemax, t.lognmax, t.p, t.T, t.mub, t.muq,
                             t.mus, t.status)
        rng = np.random.default_rng(8)
        e, nb, nq, *_ = _hadron_points(t, rng, 8000)
        e = np.concatenate([e, rng.uniform(-1e-3, 2e-3, 500), [1e4]])
        nb = np.concatenate([nb, rng.uniform(-0.1, 0.1, 501)])
        nq = np.concatenate([nq, np.zeros(501)])
    ns = np.zeros_like(e)
    ours = abi.eos_eval(desc, e, nb, nq, ns, on_gpu=False)
    ref = po.eos_eval(t, e, nb, nq, ns)
    assert np.array_equal(ours, ref)
