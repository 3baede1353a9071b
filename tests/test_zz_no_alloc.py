"""vh_cells_noalloc (the entry point that allocates nothing) against
vh_cells, through the C-ABI: the same kernel over a fixed device scratch
array instead of buffers the context allocates, so the records must be equal
bit for bit -- on two calls in a row, the second with fewer cells, on two
contexts of different sizes and precisions sharing the GPU's array."""
import numpy as np
import pytest

from vhlle_amd import abi, setups

pytestmark = pytest.mark.gpu


def _ctx(nx, ny, nz, precision):
    xl, yl, zl = (-(nx - 1) * 0.25, (nx - 1) * 0.25), (-(ny - 1) * 0.25, (ny - 1) * 0.25), \
        (-1.0, 1.0)
    e, vx, vy, vz = setups.gaussian3d_ic(nx, ny, nz, xl, yl, zl, e0=20.0, sx=1.5, sy=1.2,
                                         eta_plateau=0.5, eta_sigma=0.5)
    X, Y, _ = setups.axes(nx, ny, nz, xl, yl, zl)
    c = abi.Context(nx, ny, nz, xl, yl, zl, trc=abi.make_trc(etaS=0.1), precision=precision)
    c.set_prim(0.6, e + 0.1, 0.2 * np.tanh(X), 0.2 * np.tanh(Y), vz)
    st, tau = c.run(0.6, 0.02, 4)
    assert st == 0
    return c, tau


def test_cells_noalloc_equals_cells(gpu):
    rng = np.random.default_rng(17)
    big, tb = _ctx(40, 36, 12, abi.PREC_FP64)
    small, ts = _ctx(14, 12, 6, abi.PREC_MIXED)
    calls = [(big, tb, rng.integers(0, big.N, abi.CELLS_NOALLOC_MAX)),  # the maximum
             (big, tb, rng.integers(0, big.N, 300)),                     # fewer cells
             (small, ts, rng.integers(0, small.N, 1000)),                # another context
             (small, ts, rng.integers(0, small.N, 7))]
    for c, tau, idx in calls:
        rc, a = c.cells_noalloc(tau, idx)
        assert rc == abi.VH_OK
        b = c.cells(tau, idx)
        assert a.shape == b.shape == (len(idx), abi.VH_CELL_COLS if hasattr(abi, "VH_CELL_COLS")
                                      else 21)
        np.testing.assert_array_equal(a, b)
        assert np.all(a[:, 0] > 0.0)
    # beyond the maximum: refused, nothing written
    idx = np.zeros(abi.CELLS_NOALLOC_MAX + 1, dtype=np.int64)
    rc, a = big.cells_noalloc(tb, idx)
    assert rc == abi.VH_ERR_ARG and not a.any()
    # an index outside the grid: refused
    rc, _ = small.cells_noalloc(ts, np.array([small.N]))
    assert rc == abi.VH_ERR_ARG
    rc, a = small.cells_noalloc(ts, np.array([], dtype=np.int64))
    assert rc == abi.VH_OK and a.shape == (0, 21)
    big.close()
    small.close()
