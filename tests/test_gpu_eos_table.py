"""The tabulated EoS (VH_EOS_TABLE, the TABLE branch of EoSs, src/eos.cpp:27-120)
on the GPU against the oracle's table EoS (oracle/vh_oracle.cpp), with the
crossover table of vhlle_amd/setups.py (p/e down to 0.10 at the softest point).

* con2prim of the FP64 path: bit-identical to the oracle's transformPV with the
  same splines (the spline formula and the interval search are the same).
* FP64 viscous 3-D steps (cs2(e) = dpe(e) in the relaxation and the bulk
  coefficient, T(e) and s = (e+p)/T in the transport coefficients): 1e-10.
* Mixed precision with the reference con2prim and with the FP32 Newton from
  v = 0.5 (the closed form is for p = e/3 only and is refused).
* Evaluation outside the table raises VH_ST_EOS_RANGE (the reference aborts).
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from vhlle_amd import abi, setups

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.fixture(scope="module")
def table():
    return setups.crossover_eos()


def test_c2p_bitwise_with_table(gpu, table):
    rng = np.random.default_rng(11)
    n = 100_000
    prim = np.zeros((n, 8))
    prim[:, 0] = 10 ** rng.uniform(-3, 3, n)
    prim[:, 1] = table.spline(0, prim[:, 0])
    prim[:, 2] = rng.uniform(0, 1, n) * 0.1 * prim[:, 0]
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    prim[:, 5:8] = rng.uniform(0, 0.99, n)[:, None] * d
    Q = po.p2c(prim)
    og, itg, Qg = abi.c2p(Q, eos=table)
    oo, ito, Qo = po.c2p(Q, table)
    assert np.array_equal(itg, ito) and np.array_equal(Qg, Qo)
    assert np.array_equal(og, oo)
    # the roots satisfy the table's p: p = p(e)
    assert np.allclose(og[:, 1], table.spline(0, og[:, 0]), rtol=1e-14)


def viscous_ic(nx, ny, nz, xl, yl, zl, seed=5):
    rng = np.random.default_rng(seed)
    e, vx, vy, vz = setups.gaussian3d_ic(nx, ny, nz, xl, yl, zl, e0=20.0, sx=1.5, sy=1.2,
                                         eta_plateau=0.5, eta_sigma=0.5)
    X, Y, Z = setups.axes(nx, ny, nz, xl, yl, zl)
    vx = 0.3 * np.tanh(X / 2.0) + 0.03 * rng.uniform(-1, 1, e.shape)
    vy = 0.3 * np.tanh(Y / 2.0) + 0.03 * rng.uniform(-1, 1, e.shape)
    vz = 0.2 * np.tanh(Z)
    return e + 0.05, vx, vy, vz


@pytest.mark.parametrize("trc_kw", [dict(etaS=0.08),
                                    dict(etaS=0.12, zetaS=0.05, zetaSparam=5),
                                    dict(etaS=0.1, zetaS=0.04, zetaSparam=0)])
def test_fp64_viscous_steps_with_table(gpu, table, trc_kw):
    nx, ny, nz = 14, 12, 8
    xl, yl, zl = (-3.25, 3.25), (-2.75, 2.75), (-1.4, 1.4)
    e, vx, vy, vz = viscous_ic(nx, ny, nz, xl, yl, zl)
    ctx = abi.Context(nx, ny, nz, xl, yl, zl, trc=abi.make_trc(**trc_kw), eos=table)
    orc = po.Solver(nx, ny, nz, xl, yl, zl, trc=po.make_trc(**trc_kw), table=table)
    ctx.set_prim(0.6, e, vx, vy, vz)
    orc.set_prim(0.6, e, vx, vy, vz)
    tg = to = 0.6
    for s in range(10):
        st, tg = ctx.step(tg, 0.02)
        rc, to = orc.run(to, 0.02, 1)
        assert st == 0 and rc == 0, (s, st, rc)
    a, b = ctx.download(), orc.get_state()
    ctx.close()
    for f in ("Q", "Qh", "pi", "piH", "Pi", "PiH"):
        scale = max(np.max(np.abs(b[f])), 1e-300)
        d = np.max(np.abs(a[f] - b[f]))
        print(f"{trc_kw} {f}: {d / scale:.2e}", flush=True)
        assert d <= TOL * scale, f
    assert np.max(np.abs(b["pi"])) > 0.0


@pytest.mark.parametrize("c2p", [abi.C2P_REFERENCE, abi.C2P_NEWTON])
def test_mixed_with_table(gpu, table, c2p):
    nx, ny, nz = 40, 40, 16
    xl, yl, zl = (-5.0, 5.0), (-5.0, 5.0), (-1.5, 1.5)
    e, vx, vy, vz = viscous_ic(nx, ny, nz, xl, yl, zl)
    out = {}
    for name, prec, mode in (("fp64", abi.PREC_FP64, abi.C2P_REFERENCE),
                             ("mixed", abi.PREC_MIXED, c2p)):
        ctx = abi.Context(nx, ny, nz, xl, yl, zl, trc=abi.make_trc(etaS=0.08), precision=prec,
                          eos=table)
        if mode != abi.C2P_REFERENCE:
            ctx.set_c2p(mode)
        ctx.set_prim(0.6, e, vx, vy, vz)
        st, tau = ctx.run(0.6, 0.02, 30)
        assert st == 0
        out[name] = ctx.download()
        ctx.close()
    a, b = out["mixed"], out["fp64"]
    eT = np.max(np.abs(a["Q"][:4] - b["Q"][:4])) / np.max(np.abs(b["Q"][0]))
    print(f"c2p={c2p}: T^tau-mu rel err {eT:.2e}", flush=True)
    assert eT < 2e-6


def _probe_points(e_nodes, rng):
    """Every node and its 1-3 ulp neighbours, random log-uniform points across
    and beyond the table, values below the first positive node and negative
    values."""
    pos = e_nodes[e_nodes > 0]
    pts = [e_nodes]
    for k in (1, 2, 3):
        up = dn = pos.copy()
        for _ in range(k):
            up = np.nextafter(up, np.inf)
            dn = np.nextafter(dn, -np.inf)
        pts += [up, dn]
    lo, hi = np.log10(pos[0]), np.log10(e_nodes[-1])
    pts.append(10 ** rng.uniform(lo - 1, hi + 0.5, 200_000))
    pts.append(rng.uniform(0, pos[0], 1000))
    pts.append(-10 ** rng.uniform(lo - 1, hi, 1000))  # below the table: the first cubic
    return np.concatenate(pts)


@pytest.mark.parametrize("kind", ["log_uniform", "uniform_2000", "uniform_20000"])
def test_direct_lookup_matches_search(gpu, table, kind):
    """The device spline lookup by log2 bins (each bin holds at most one node,
    so the interval is the bin's first interval or the next) returns the
    values of the interval search bit for bit: the host evaluation bisects.
    Tables: the log-uniform crossover table; 2 000 uniformly spaced nodes (the
    bins must be finer than the top nodes' spacing: 32 768 bins); 20 000
    uniformly spaced nodes (no bin count up to 2^18 separates the top nodes:
    the device falls back to the hint search)."""
    rng = np.random.default_rng(41)
    if kind == "log_uniform":
        tab = table
    else:
        n = int(kind.split("_")[1])
        e_t = np.linspace(0.0, 50.0, n)
        tab = abi.EosTable(e_t, e_t / 3.0 - 0.01 * np.sin(e_t), 0.2 * e_t ** 0.25 + 0.001 * e_t)
    e = _probe_points(tab.e, rng)
    gpu_out = abi.eos_eval(tab, e, on_gpu=True)
    host_out = abi.eos_eval(tab, e, on_gpu=False)
    assert np.array_equal(gpu_out[:, :2], host_out[:, :2])
    assert np.array_equal(gpu_out[:, 5], host_out[:, 5])


def test_closed_form_refused_with_table(gpu, table):
    ctx = abi.Context(10, 10, 1, (-1, 1), (-1, 1), (-1, 1), precision=abi.PREC_MIXED, eos=table)
    assert abi.lib().vh_set_c2p_mode(ctx.h, abi.C2P_CLOSED) == abi.VH_ERR_ARG
    ctx.close()


def test_out_of_table_raises_status(gpu):
    """A table ending at e = 30 and a state reaching e = 40: the step's
    con2prim evaluates p(e) beyond the table (the reference's GSL call would
    abort); the kernels use the end interval and raise VH_ST_EOS_RANGE."""
    e_t = np.concatenate(([0.0], np.geomspace(1e-6, 30.0, 300)))
    tab = abi.EosTable(e_t, e_t / 3.0, 0.2 * e_t ** 0.25)
    n = 12
    X, Y, _ = setups.axes(n, n, 1, (-1, 1), (-1, 1), (-1, 1))
    e = 1.0 + 39.0 * np.exp(-(X * X + Y * Y) * 8)
    ctx = abi.Context(n, n, 1, (-1, 1), (-1, 1), (-1, 1), eos=tab)
    ctx.set_prim(1.0, e, 0 * e, 0 * e, 0 * e)
    st, _ = ctx.run(1.0, 0.02, 2)
    ctx.close()
    assert st & abi.ST_EOS_RANGE


@pytest.fixture(scope="module")
def table_model(table):
    """The con2prim MLP trained on states of the crossover EoS (Gubser-shaped
    runs with the table, half augmented), on the GPU."""
    Q, prim = abi.harvest(1 << 20, seed=21, steps=100, eos=table)
    p, hist = abi.mlp_train(abi.mlp_init(5), Q, prim, batch=512, epochs=6, lr=2e-3,
                            lr_final=1e-6, seed=3)
    return p, hist


def test_learned_c2p_on_table(gpu, table, table_model):
    """Without a closed form the Newton start matters: the MLP guess against
    v = 0.5 on held-out states of the table EoS, both polished to the
    1e-12 Q_t residual of the table's p(e)."""
    p, hist = table_model
    assert hist[-1] < 0.05 * hist[0]
    Q, _ = abi.harvest(200000, seed=77, steps=100, eos=table)
    res = {}
    for name, start in (("mlp", abi.START_MLP), ("half", abi.START_HALF)):
        out, it, Qa = abi.c2p_learned(p, Q, start, eos=table)
        m = Qa[:, 0] > 0
        M = np.linalg.norm(Qa[:, 1:4], axis=1)
        v = np.linalg.norm(out[:, 5:8], axis=1)
        pe = table.spline(0, np.maximum(out[:, 0], 0.0))
        resid = np.abs((Qa[:, 0] + pe) * v - M)
        assert np.all(resid[m] <= 2e-12 * Qa[m, 0]), name
        res[name] = (np.mean(np.where(it >= 0, it, 0)), np.mean(it < 0))
    print(f"table EoS: polish updates / fallback rate: MLP {res['mlp']}, v=0.5 {res['half']}",
          flush=True)
    assert res["mlp"][1] <= 1e-3
    assert res["mlp"][0] < 0.75 * res["half"][0]


def test_mixed_learned_with_table(gpu, table, table_model):
    p, _ = table_model
    nx, ny, nz = 40, 40, 16
    xl, yl, zl = (-5.0, 5.0), (-5.0, 5.0), (-1.5, 1.5)
    e, vx, vy, vz = viscous_ic(nx, ny, nz, xl, yl, zl)
    out = {}
    for name, prec, mode in (("fp64", abi.PREC_FP64, abi.C2P_REFERENCE),
                             ("learned", abi.PREC_MIXED, abi.C2P_LEARNED),
                             ("newton", abi.PREC_MIXED, abi.C2P_NEWTON)):
        ctx = abi.Context(nx, ny, nz, xl, yl, zl, trc=abi.make_trc(etaS=0.08), precision=prec,
                          eos=table)
        if mode != abi.C2P_REFERENCE:
            ctx.set_c2p(mode, p if mode == abi.C2P_LEARNED else None)
        ctx.set_prim(0.6, e, vx, vy, vz)
        st, _ = ctx.run(0.6, 0.02, 30)
        assert st == 0
        out[name] = ctx.download()
        ctx.close()
    qs = np.max(np.abs(out["fp64"]["Q"][0]))
    for k in ("learned", "newton"):
        eT = np.max(np.abs(out[k]["Q"][:4] - out["fp64"]["Q"][:4])) / qs
        print(f"table EoS {k}: T^tau-mu rel err {eT:.2e}", flush=True)
        assert eT < 2e-6
