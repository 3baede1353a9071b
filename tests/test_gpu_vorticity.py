"""Vorticity output on the GPU (vh_set_vorticity, vh_get_dbeta,
vh_freezeout_beta) against the oracle's restatement of the reference
(src/hdo.cpp:420-611, src/cll.h:167-204, src/fld.cpp:663-834).

Bar: the FP64 step's, max|gpu - oracle| <= 1e-10 max|oracle| per field,
including the entries the reference inherits from the cell before in its loop
order (vacuum fronts; a matter sheet one cell thick in eta, where every cell
only adds the Christoffel terms to [3][0] and [3][3], so those entries are
one running sum over the whole grid).
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from vhlle_amd import abi, setups

pytestmark = pytest.mark.gpu

TOL = 1e-10


def close(a, b, label):
    scale = max(np.max(np.abs(b)), 1e-300)
    assert np.all(np.isfinite(a)), label
    d = np.max(np.abs(a - b))
    assert d <= TOL * scale, f"{label}: max diff {d:.3e}, scale {scale:.3e}"


def pair(shape, lims, trc_kw, tau0, e, vx, vy, vz, nb=None):
    nx, ny, nz = shape
    g = abi.Context(nx, ny, nz, *lims, trc=abi.make_trc(**trc_kw))
    o = po.Solver(nx, ny, nz, *lims, trc=po.make_trc(**trc_kw))
    g.set_vorticity(True)
    o.set_vorticity(True)
    g.set_prim(tau0, e, vx, vy, vz, nb=nb)
    o.set_prim(tau0, e, vx, vy, vz, nb=nb)
    return g, o


def vacuum_ic(nx, ny, nz, lims, seed):
    rng = np.random.default_rng(seed)
    X, Y, H = setups.axes(nx, ny, nz, *lims)
    r2 = X ** 2 / 6.0 + Y ** 2 / 4.0 + H ** 2 / 0.6
    e = np.where(r2 < 1.0, 30.0 * (1.0 - r2) ** 2 + 0.2, 0.0)
    e *= 1 + 0.05 * rng.uniform(-1, 1, e.shape)
    vx = 0.3 * np.tanh(X / 2) + 0.2 * np.sin(Y)
    vy = 0.3 * np.tanh(Y / 2) - 0.2 * np.sin(X)
    vz = 0.2 * np.tanh(H) + 0.05 * rng.uniform(-1, 1, e.shape)
    return e, vx, vy, vz


def test_vacuum_front_steps(gpu):
    nx, ny, nz = 18, 16, 9
    lims = ((-4.25, 4.25), (-3.75, 3.75), (-1.0, 1.0))
    e, vx, vy, vz = vacuum_ic(nx, ny, nz, lims, 7)
    g, o = pair((nx, ny, nz), lims, dict(etaS=0.12), 0.8, e, vx, vy, vz, nb=0.01 * e)
    tg = to = 0.8
    dt = 0.05
    for s in range(8):
        st, tg = g.step(tg, dt)
        rc, to = o.run(to, dt, 1)
        assert st == 0 and rc == 0
        close(g.get_dbeta(), o.get_dbeta(), f"step {s}")
    db = o.get_dbeta()
    assert np.any(db != 0.0) and np.max(np.abs(db)) <= 100.0


def sheet_state(o, shape, lims):
    """a sheet of matter at iz = 4 of 9 between two layers with Q_t < 0 (m = 1
    everywhere): the layers stay empty through the predictor (negative Q_t
    is cleared, src/cll.h:75-86), so no sheet cell forms row 3 and every one
    adds the Christoffel terms to [3][0], [3][3]"""
    nx, ny, nz = shape
    X, Y, H = setups.axes(nx, ny, nz, *lims)
    e = np.where(np.abs(H) < 0.1, 10.0 * np.exp(-(X ** 2 + Y ** 2) / 8.0), 0.0)
    z = np.zeros_like(e)
    o.set_prim(1.0, e, 0.2 * np.tanh(X), 0.2 * np.tanh(Y), z)
    st = o.get_state()
    iz = np.arange(nx * ny * nz) // (nx * ny)
    for k in ("Q", "Qh", "Qprev"):
        st[k][:, (iz == 3) | (iz == 5)] = 0.0
        st[k][0, (iz == 3) | (iz == 5)] = -50.0
    st["m"][:] = 1.0
    o.set_state(st)
    return st


@pytest.mark.parametrize("nx, ny", [(20, 17), (48, 45)])
def test_thin_sheet_running_sums(gpu, nx, ny):
    """one running sum over the whole grid in loop order (the resolver's
    chains of columns that only add); 48 x 45 columns span three scan tiles
    of 1024"""
    nz = 9
    lims = ((-0.125 * (nx - 1), 0.125 * (nx - 1)), (-0.125 * (ny - 1), 0.125 * (ny - 1)), (-1.0, 1.0))
    g = abi.Context(nx, ny, nz, *lims, trc=abi.make_trc(etaS=0.08))
    o = po.Solver(nx, ny, nz, *lims, trc=po.make_trc(etaS=0.08))
    g.set_vorticity(True)
    o.set_vorticity(True)
    g.upload(sheet_state(o, (nx, ny, nz), lims))
    tg = to = 1.0
    dt = 0.02
    for s in range(2):
        st, tg = g.step(tg, dt)
        rc, to = o.run(to, dt, 1)
        assert st == 0 and rc == 0
        got, want = g.get_dbeta(), o.get_dbeta()
        close(got, want, f"step {s}")
        if s == 0:  # the sums grew along the loop order and reached the clamp
            assert np.max(want[15]) == 100.0 and np.all(want[13:15] == 0.0)


def test_vacuum_fronts_over_several_scan_tiles(gpu):
    """64 x 40 columns (three scan tiles of 1024): stale cells and partial
    rows next to vacuum take the matrix of the cell before them across tile
    boundaries"""
    nx, ny, nz = 64, 40, 7
    lims = ((-7.875, 7.875), (-4.875, 4.875), (-0.75, 0.75))
    X, Y, H = setups.axes(nx, ny, nz, *lims)
    rng = np.random.default_rng(11)
    # a row of blobs along x: vacuum gaps between them in every tile
    e = np.zeros_like(X)
    for x0 in (-6.0, -2.0, 2.0, 6.0):
        r2 = (X - x0) ** 2 / 2.0 + Y ** 2 / 9.0 + H ** 2 / 0.5
        e += np.where(r2 < 1.0, 20.0 * (1.0 - r2) ** 2 + 0.3, 0.0)
    e *= 1 + 0.05 * rng.uniform(-1, 1, e.shape)
    vx = 0.2 * np.sin(X)
    vy = 0.3 * np.tanh(Y / 2)
    vz = 0.1 * np.tanh(H)
    g, o = pair((nx, ny, nz), lims, dict(etaS=0.1), 0.8, e, vx, vy, vz)
    tg = to = 0.8
    for s in range(3):
        st, tg = g.step(tg, 0.05)
        rc, to = o.run(to, 0.05, 1)
        assert st == 0 and rc == 0
        close(g.get_dbeta(), o.get_dbeta(), f"step {s}")
    op_stale = o.get_dbeta()
    assert np.any(op_stale != 0.0)


def test_graph_run_and_records(gpu):
    nx, ny, nz = 16, 14, 10
    lims = ((-3.75, 3.75), (-3.25, 3.25), (-1.125, 1.125))
    e, vx, vy, vz = vacuum_ic(nx, ny, nz, lims, 3)
    g, o = pair((nx, ny, nz), lims, dict(etaS=0.1, zetaS=0.04, zetaSparam=2), 0.7, e, vx, vy, vz)
    tg = to = 0.7
    dt = 0.04
    st, tg = g.run(tg, dt, 6)
    rc, to = o.run(to, dt, 6)
    assert st == 0 and rc == 0 and tg == to
    close(g.get_dbeta(), o.get_dbeta(), "graph replay")
    rows, beta = g.freezeout_beta(tg, dt, 2.0, extend=True)
    rows0, _ = g.freezeout(tg, dt, 2.0, extend=True)
    s = g.download()
    spacing = tuple((l[1] - l[0]) / (n - 1) for l, n in zip(lims, (nx, ny, nz)))
    ref, rbeta = po.freezeout((nx, ny, nz), spacing, tuple(l[0] for l in lims), s["Q"],
                              s["Qprev"], s["pi"], s["Pi"], tg, dt, 2.0, extend=True,
                              dbeta=o.get_dbeta())
    print(f"{len(rows)} elements", flush=True)
    assert len(rows) == len(ref) == len(beta) > 50
    assert np.array_equal(rows, rows0)
    assert np.array_equal(beta[:, :16], rows[:, :16]) and np.array_equal(beta[:, 32], rows[:, 27])
    # cosh/sinh of the device and host libm may differ by an ulp
    close(beta[:, 16:32], rbeta[:, 16:32], "beta d_i beta_j")


def test_ideal_fluid_keeps_zero_and_switch_off(gpu):
    n = 10
    lims = ((-2.25, 2.25), (-2.25, 2.25), (-0.5, 0.5))
    e, vx, vy, vz = vacuum_ic(n, n, 5, lims, 1)
    g = abi.Context(n, n, 5, *lims)
    g.set_vorticity(True)
    g.set_prim(1.0, e, vx, vy, vz)
    st, _ = g.run(1.0, 0.05, 3)
    assert st == 0
    assert np.all(g.get_dbeta() == 0.0)  # NSquant runs only in the viscous step
    g.set_vorticity(False)
    with pytest.raises(RuntimeError):
        g.get_dbeta()


@pytest.mark.parametrize("c2p", [abi.C2P_REFERENCE, abi.C2P_NEWTON])
def test_mixed_precision_close_to_fp64(gpu, c2p):
    """VH_PREC_MIXED forms the matrix in FP64 from its own states: within the
    mixed step's accuracy of the FP64 path's matrix, the entries inherited at
    the vacuum front included"""
    nx, ny, nz = 18, 16, 9
    lims = ((-4.25, 4.25), (-3.75, 3.75), (-1.0, 1.0))
    e, vx, vy, vz = vacuum_ic(nx, ny, nz, lims, 7)
    out = {}
    for prec in (abi.PREC_FP64, abi.PREC_MIXED):
        g = abi.Context(nx, ny, nz, *lims, trc=abi.make_trc(etaS=0.12), precision=prec)
        if prec == abi.PREC_MIXED and c2p != abi.C2P_REFERENCE:
            g.set_c2p(c2p)
        g.set_vorticity(True)
        g.set_prim(0.8, e, vx, vy, vz)
        st, tau = g.run(0.8, 0.05, 6)
        assert st == 0
        out[prec] = g.get_dbeta()
        g.close()
    a, b = out[abi.PREC_MIXED], out[abi.PREC_FP64]
    scale = np.max(np.abs(b))
    rel = np.max(np.abs(a - b)) / scale
    print(f"mixed vs FP64: max diff / scale = {rel:.3e}", flush=True)
    assert rel <= 1e-4
