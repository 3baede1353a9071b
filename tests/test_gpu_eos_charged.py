"""The charge-dependent EoSs on the GPU (FP64 path) against the oracle:
EoSChiral (bilinear, src/eoChiral.cpp:99-154) and EoSHadron (log-trilinear,
src/eoHadron.cpp:78-208), with the synthetic tables of charged_eos.py.

* the device lookups (vh_eos_eval on the GPU) equal the oracle's: bit for bit
  for the chiral tables (no transcendental outside the power law's T), to a
  few ulp for the hadron table (the device log);
* con2prim with the charge refresh (vh_c2p) equals the oracle's;
* viscous 3-D steps with a baryon-charged initial state, with eta/s from
  etaSparam 3 (which takes T and mu_B from the EoS), with and without the
  vorticity output, equal the oracle's steps to the FP64 parity bar;
* the freeze-out records with the hadron EoS as the surface EoS
  (vh_set_surface_eos: Fluid's eosH) equal the oracle's records of the same
  state with that EoS, including the T, mu_B, mu_Q, mu_S columns;
* the mixed-precision path refuses these EoSs.
"""
import numpy as np
import pytest

import charged_eos as ce
from oracle import pyoracle as po
from vhlle_amd import abi, setups

pytestmark = pytest.mark.gpu

TOL = 1e-10
FIELDS = ("Q", "Qh", "Qprev", "pi", "piH", "Pi", "PiH", "m", "viscCorrCut")


@pytest.fixture(scope="module")
def tables():
    return {"chiral": ce.chiral(), "hadron": ce.hadron()}


def _points(which, rng, n):
    if which == "chiral":
        e = np.concatenate([rng.uniform(-0.05, 1.6, n // 2), rng.uniform(0, 180.0, n // 2)])
        nb = np.concatenate([rng.uniform(-0.3, 0.35, n // 2), rng.uniform(-0.5, 7.0, n // 2)])
        return e, nb, np.zeros(n), np.zeros(n)
    e = np.exp(rng.uniform(np.log(1e-4), np.log(300.0), n))
    nb = rng.uniform(-1, 1, n) * np.exp(rng.uniform(np.log(1e-5), np.log(6.0), n))
    nq = rng.uniform(-1, 1, n) * np.exp(rng.uniform(np.log(1e-5), np.log(6.0), n))
    return e, nb, nq, rng.uniform(-0.1, 0.1, n)


@pytest.mark.parametrize("which", ["chiral", "hadron"])
def test_device_lookups_equal_oracle(gpu, tables, which):
    rng = np.random.default_rng(11)
    e, nb, nq, ns = _points(which, rng, 100_000)
    t = tables[which]
    dev = gpu.eos_eval(ce.to_abi(t), e, nb, nq, ns, on_gpu=True)
    host = gpu.eos_eval(ce.to_abi(t), e, nb, nq, ns, on_gpu=False)
    ref = po.eos_eval(t, e, nb, nq, ns)
    assert np.array_equal(host, ref)
    if which == "chiral":
        table = e < 146.0
        assert np.array_equal(dev[table & (nb < 6.0)], ref[table & (nb < 6.0)])
    rel = np.abs(dev - ref) / np.maximum(np.abs(ref), 1e-300)
    print(which, "max rel diff", rel.max(), "exact fraction", np.mean(dev == ref), flush=True)
    assert rel.max() <= 1e-13


@pytest.mark.parametrize("which", ["chiral", "hadron"])
def test_c2p_equals_oracle(gpu, tables, which):
    t = tables[which]
    rng = np.random.default_rng(12)
    n = 100_000
    e = 10 ** rng.uniform(-3, 1.5, n)
    nb = rng.uniform(-0.25, 0.28, n)
    nq = rng.uniform(-0.1, 0.1, n) if which == "hadron" else np.zeros(n)
    v = rng.uniform(0, 0.99, n) ** 0.5
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    prim = np.zeros((n, 8))
    prim[:, 0], prim[:, 1], prim[:, 2], prim[:, 3] = e, e / 3.0, nb, nq
    prim[:, 5:8] = v[:, None] * d
    Q = po.p2c(prim)
    Q[: n // 10, 1:4] = 0.0  # at rest
    og, itg, Qg = gpu.c2p(Q, eos=ce.to_abi(t))
    oo, ito, Qo = po.c2p(Q, t)
    agree = np.mean(itg == ito)
    rel = np.max(np.abs(og - oo) / np.maximum(np.abs(oo), 1e-300))
    print(which, "iterations agree", agree, "max rel diff", rel, flush=True)
    if which == "chiral":
        assert np.array_equal(og, oo) and np.array_equal(itg, ito)
    else:
        assert agree > 0.999 and rel <= 1e-9


def _charged_ic(nx, ny, nz, xl, yl, zl):
    e, vx, vy, vz = setups.gaussian3d_ic(nx, ny, nz, xl, yl, zl, e0=1.3, sx=1.0, sy=1.2,
                                         eta_plateau=0.3, eta_sigma=0.3)
    nb = 0.18 * e / (0.4 + e)
    nq = 0.4 * nb
    return e, vx, vy, vz, nb, nq


def _trc(param):
    if param == 3:
        return dict(etaS=0.0, etaSparam=3, T0=0.15, etaSMin=0.08, ah=0.3, al=-0.5,
                    etaSShiftMuB=0.4, etaSScaleMuB=0.6, zetaS=0.04, zetaSparam=5)
    return dict(etaS=0.1, zetaS=0.03, zetaSparam=5)


def _assert_close(a, b, fields, label):
    for f in fields:
        ref, got = np.ravel(b[f]), np.ravel(a[f])
        scale = np.max(np.abs(ref))
        diff = np.max(np.abs(got - ref)) if ref.size else 0.0
        assert np.all(np.isfinite(got)), f"{label} {f}"
        assert diff <= TOL * max(scale, 1e-300), f"{label} {f}: {diff:.3e} of {scale:.3e}"


@pytest.mark.parametrize("which,param,vort", [("chiral", 0, False), ("chiral", 3, True),
                                              ("hadron", 3, False), ("hadron", 0, True)])
def test_viscous_steps_equal_oracle(gpu, tables, which, param, vort):
    nx, ny, nz = 22, 20, 9
    xl, yl, zl = (-2.1, 2.1), (-1.9, 1.9), (-0.48, 0.48)
    e, vx, vy, vz, nb, nq = _charged_ic(nx, ny, nz, xl, yl, zl)
    if which == "chiral":
        nq = None
    t = tables[which]
    kw = _trc(param)
    ctx = gpu.Context(nx, ny, nz, xl, yl, zl, trc=gpu.make_trc(**kw), eos=ce.to_abi(t))
    orc = po.Solver(nx, ny, nz, xl, yl, zl, trc=po.make_trc(**kw), table=t)
    if vort:
        ctx.set_vorticity(True)
        orc.set_vorticity(True)
    ctx.set_prim(0.6, e, vx, vy, vz, nb=nb, nq=nq)
    orc.set_prim(0.6, e, vx, vy, vz, nb=nb, nq=nq)
    st, tau = ctx.run(0.6, 0.02, 6)
    rc, tau_o = orc.run(0.6, 0.02, 6)
    print(f"{which} etaSparam {param} vorticity {vort}: status {st}, oracle rc {rc}", flush=True)
    assert st == 0 and rc == 0 and tau == tau_o
    g = ctx.download()
    o = orc.get_state()
    assert np.max(np.abs(g["Q"].reshape(7, -1)[4])) > 0.05  # the baryon charge is there
    _assert_close(g, o, FIELDS, which)
    if vort:
        db_g, db_o = ctx.get_dbeta(), orc.get_dbeta()
        scale = np.max(np.abs(db_o))
        assert scale > 0 and np.max(np.abs(db_g - db_o)) <= 1e-9 * scale
    ctx.close()


def test_surface_eos_records_equal_oracle(gpu, tables):
    """chiral hydro, hadron EoS on the surface (the reference's eosH, swapped
    in for outputSurface): the records' con2prim, T and mu columns follow the
    surface EoS, as the oracle's records of the same state with it"""
    nx, ny, nz = 26, 24, 12
    xl, yl, zl = (-2.5, 2.5), (-2.3, 2.3), (-0.55, 0.55)
    e, vx, vy, vz, nb, _ = _charged_ic(nx, ny, nz, xl, yl, zl)
    e = e * 1.5
    ctx = gpu.Context(nx, ny, nz, xl, yl, zl, trc=gpu.make_trc(etaS=0.1),
                      eos=ce.to_abi(tables["chiral"]))
    ctx.set_prim(0.6, e, vx, vy, vz, nb=nb)
    st, tau = ctx.run(0.6, 0.02, 8)
    assert st == 0
    shape = (nx, ny, nz)
    spacing = ((xl[1] - xl[0]) / (nx - 1), (yl[1] - yl[0]) / (ny - 1), (zl[1] - zl[0]) / (nz - 1))
    mins = (xl[0], yl[0], zl[0])
    s = ctx.download()
    rows_h, _ = ctx.freezeout(tau, 0.02, 0.5, extend=True)
    ctx.set_surface_eos(ce.to_abi(tables["hadron"]))
    rows, stats = ctx.freezeout(tau, 0.02, 0.5, extend=True)
    ctx.close()
    ref = po.freezeout(shape, spacing, mins, s["Q"], s["Qprev"], s["pi"], s["Pi"], tau, 0.02, 0.5,
                       extend=True, table=tables["hadron"])
    print(f"{len(rows)} elements, T {rows[:, 12].min():.4f}..{rows[:, 12].max():.4f}, "
          f"mu_B {rows[:, 13].min():.4f}..{rows[:, 13].max():.4f}", flush=True)
    assert len(rows) == len(ref) and len(rows) > 100 and stats["err"] == 0
    assert np.array_equal(rows[:, :4], ref[:, :4])
    scale = np.maximum(np.abs(ref).max(axis=0), 1e-300)
    assert np.all(np.abs(rows - ref) <= 1e-11 * scale)
    assert np.abs(rows[:, 13]).max() > 1e-3 and np.abs(rows[:, 14]).max() == 0.0  # n_q = 0
    # the hydro EoS's surface differs in T (the two synthetic EoSs differ)
    assert len(rows_h) == len(rows) and not np.allclose(rows_h[:, 12], rows[:, 12])


def test_mixed_precision_refuses_charged_eos(gpu, tables):
    for t in tables.values():
        with pytest.raises(RuntimeError):
            gpu.Context(12, 12, 1, (-1, 1), (-1, 1), (-0.5, 0.5), precision=gpu.PREC_MIXED,
                        eos=ce.to_abi(t))
    ctx = gpu.Context(12, 12, 1, (-1, 1), (-1, 1), (-0.5, 0.5), precision=gpu.PREC_MIXED)
    with pytest.raises(RuntimeError):
        ctx.set_surface_eos(ce.to_abi(tables["hadron"]))
    ctx.close()
