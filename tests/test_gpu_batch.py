"""Event-by-event batches on the GPU (vh_create_batch, vh_hotspot_ic) through
the C-ABI:

* the hot-spot deposition equals a numpy deposition of the returned spots,
  and the sampled spots follow Poisson(N), the 2-D Woods-Saxon radial law and
  Gamma(1) weights;
* every grid of a batch of uneven grids (2-D and 3-D) evolves bit-identically
  to the same grid run alone, for each stage kernel shape and con2prim;
* a slice of the benchmark batch (hot spots, eta/s = 0.12, T-peaked zeta/s)
  runs without error status.
"""
import numpy as np
import pytest

from vhlle_amd import abi, setups

pytestmark = pytest.mark.gpu


def axes(g):
    nx, ny, nz, xl, yl, zl = g
    x = np.linspace(xl[0], xl[1], nx)
    y = np.linspace(yl[0], yl[1], ny)
    z = np.linspace(zl[0], zl[1], nz) if nz > 1 else np.array([zl[0]])
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return X.ravel(), Y.ravel(), Z.ravel()


def test_hotspot_deposition_and_sampling(gpu):
    grids = [(60, 50, 1, (-3.0, 3.0), (-2.5, 2.5), (-0.5, 0.5)),
             (30, 30, 11, (-1.5, 1.5), (-1.5, 1.5), (-2.5, 2.5))]
    e, npart, sp = abi.hotspot_ic(grids, seed=7, amp=5.2, eta_plateau=1.0, eta_sigma=0.6,
                                  spots=True)
    off, k0 = 0, 0
    for g, n in zip(grids, npart):
        X, Y, Z = axes(g)
        s = sp[k0:k0 + n]
        ref = 5.2 * np.sum(s[:, 2][None, :] * np.exp(-((X[:, None] - s[None, :, 0]) ** 2 +
                                                        (Y[:, None] - s[None, :, 1]) ** 2) / 0.5),
                           axis=1)
        if g[2] > 1:
            a = np.abs(Z)
            ref = ref * np.where(a < 1.0, 1.0, np.exp(-(a - 1.0) ** 2 / (2 * 0.36)))
        got = e[off:off + len(X)]
        assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(ref)
        off += len(X)
        k0 += n
    # sampling statistics over many (tiny) events
    many = [(5, 5, 1, (-0.2, 0.2), (-0.2, 0.2), (-0.5, 0.5))] * 400
    _, npart, sp = abi.hotspot_ic(many, seed=11, spots=True)
    assert abs(npart.mean() - 150.0) < 4 * np.sqrt(150.0 / 400)
    assert abs(np.var(npart) / 150.0 - 1.0) < 0.25
    w = sp[:, 2]
    assert abs(w.mean() - 1.0) < 5 / np.sqrt(len(w)) and abs(w.var() - 1.0) < 0.05
    r = np.hypot(sp[:, 0], sp[:, 1])
    rr = np.linspace(0, 6.6 + 12 * 0.55, 20001)
    dens = rr / (1 + np.exp((rr - 6.6) / 0.55))
    mean_r = np.trapezoid(rr * dens, rr) / np.trapezoid(dens, rr)
    assert abs(r.mean() - mean_r) < 5 * r.std() / np.sqrt(len(r))


GRIDS = [(23, 23, 1, (-1.1, 1.1), (-1.1, 1.1), (-0.5, 0.5)),
         (31, 27, 1, (-1.5, 1.5), (-1.3, 1.3), (-0.5, 0.5)),
         (40, 40, 1, (-1.95, 1.95), (-1.95, 1.95), (-0.5, 0.5)),
         (17, 19, 1, (-0.8, 0.8), (-0.9, 0.9), (-0.5, 0.5)),
         (24, 20, 9, (-1.15, 1.15), (-0.95, 0.95), (-0.4, 0.4))]


@pytest.mark.parametrize("c2p,stage", [(abi.C2P_CLOSED, abi.STAGE_AUTO),
                                       (abi.C2P_CLOSED, abi.STAGE_ROWS),
                                       (abi.C2P_NEWTON, abi.STAGE_CELLS),
                                       (abi.C2P_NEWTON, abi.STAGE_BANDS)])
def test_batch_equals_single_grids(gpu, c2p, stage):
    trc = setups.batch_trc()
    e, _ = abi.hotspot_ic(GRIDS, seed=3, amp=5.2, eta_plateau=0.2, eta_sigma=0.3)
    z = np.zeros_like(e)
    b = abi.Context(grids=GRIDS, trc=trc, precision=abi.PREC_MIXED)
    b.set_c2p(c2p)
    if stage == abi.STAGE_BANDS:  # bands of 4 rows: ragged against ny = 23, 27, 19
        b.set_stage_bands(4, 3)
    elif stage != abi.STAGE_AUTO:
        b.set_stage(stage, 3)
    b.set_prim(0.5, e, z, z, z)
    st, _ = b.run(0.5, 0.02, 12)
    assert st == 0
    sb = b.download()
    b.close()
    P = len(e)
    off = 0
    for g in GRIDS:
        n = g[0] * g[1] * g[2]
        c = abi.Context(*g, trc=trc, precision=abi.PREC_MIXED)
        c.set_c2p(c2p)
        c.set_prim(0.5, e[off:off + n], z[:n], z[:n], z[:n])
        st, _ = c.run(0.5, 0.02, 12)
        assert st == 0
        s = c.download()
        c.close()
        for f, nc in (("Q", 7), ("pi", 10), ("Pi", 1)):
            got = sb[f].reshape(nc, P)[:, off:off + n]
            assert np.array_equal(got, s[f].reshape(nc, n)), (g, f)
        off += n


def test_batch_workload_slice(gpu):
    cfg = setups.BATCH_EBE
    grids = setups.batch_grids(events=12)
    e, npart = abi.hotspot_ic(grids, seed=cfg["seed"], npart_mean=cfg["npart_mean"], R=cfg["R"],
                              a=cfg["a"], sigma=cfg["sigma"], amp=cfg["amp"])
    assert np.all(e > 0)
    z = np.zeros_like(e)
    b = abi.Context(grids=grids, trc=setups.batch_trc(), precision=abi.PREC_MIXED)
    b.set_c2p(abi.C2P_CLOSED)
    b.set_prim(cfg["tau0"], e, z, z, z)
    st, tau = b.run(cfg["tau0"], cfg["dt"], 40)
    s = b.download()
    b.close()
    print(f"12 events, npart {npart.tolist()}, e_max {e.max():.1f}, status {st}", flush=True)
    assert st == 0 and np.all(np.isfinite(s["Q"])) and np.all(np.isfinite(s["pi"]))


def test_long_runs_multistep_graphs_and_retuning(gpu):
    """Runs long enough for the 8-step graphs and the second tuning of the
    stage shape (after 32 steps): a batch run of 45 steps in one vh_run, in
    uneven chunks, and a forced cell-kernel run all give the same states, bit
    for bit."""
    trc = setups.batch_trc()
    e, _ = abi.hotspot_ic(GRIDS, seed=5, amp=5.2, eta_plateau=0.2, eta_sigma=0.3)
    z = np.zeros_like(e)
    out = []
    for how in ("one_run", "chunks", "cells"):
        b = abi.Context(grids=GRIDS, trc=trc, precision=abi.PREC_MIXED)
        b.set_c2p(abi.C2P_CLOSED)
        if how == "cells":
            b.set_stage(abi.STAGE_CELLS)
        b.set_prim(0.5, e, z, z, z)
        if how == "chunks":
            tau = 0.5
            for n in (1, 9, 16, 3, 16):
                st, tau = b.run(tau, 0.02, n)
                assert st == 0
        else:
            st, tau = b.run(0.5, 0.02, 45)
            assert st == 0
        k, cz, ms = b.get_stage()
        print(f"{how}: stage kernel {k} CZ {cz} ({ms:.4f} ms per stage pair), tau {tau}",
              flush=True)
        assert k in (abi.STAGE_CELLS, abi.STAGE_ROWS)
        out.append(b.download())
        b.close()
    for f in ("Q", "Qh", "pi", "Pi"):
        assert np.array_equal(out[0][f], out[1][f]), f
        assert np.array_equal(out[0][f], out[2][f]), f
