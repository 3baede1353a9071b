"""ctypes wrapper of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

The oracle is a serial C++ restatement of vHLLE's Hydro::performStep
(src/hdo.cpp:1016-1134) in the reference's operation order.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module,
and only as the checker / timed CPU baseline, never as the product path.

Parity status of each oracle function (how it is pinned without the reference,
which cannot be built here: every hot-path file includes GSL headers that this
image lacks):
  con2prim         -- closed-form root of the conformal EoS, special cases,
                      round trip with prim2con, brute-force bisection
  prim2con         -- textbook T^{mu nu} of a perfect fluid
  ideal step       -- Gubser analytic flow, uniform-state (Bjorken) recurrence,
                      mirror symmetry, conservation in a closed box
  viscous step     -- Bjorken viscous ODE (Israel-Stewart with the second-order
                      coefficients of src/trancoeff.h:28-37), static fluid,
                      symmetry of pi
  ghost fill       -- copy identities
  transport coeffs -- closed forms of src/trancoeff.cpp
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class Trc(ctypes.Structure):
    """Mirror of or_trc_t (vh_oracle.h)."""
    _fields_ = [(n, ctypes.c_double) for n in (
        "etaS0", "zetaS0", "ah", "al", "aRho", "T0", "etaSMin", "etaSEpsilonMin",
        "etaSShiftMuB", "etaSScaleMuB", "zetaSPeakEpsilon", "zetaSScaleBeta",
        "zetaSSigmaMinus", "zetaSSigmaPlus")] + [
        ("etaSparam", ctypes.c_int), ("zetaSparam", ctypes.c_int),
        ("zetaSPeakT", ctypes.c_double), ("zetaSWidthMinus", ctypes.c_double),
        ("zetaSWidthPlus", ctypes.c_double)]


class State(ctypes.Structure):
    _fields_ = [(n, ctypes.POINTER(ctypes.c_double)) for n in (
        "Q", "Qh", "Qprev", "pi", "piH", "Pi", "PiH", "pi0", "piH0", "Pi0", "PiH0",
        "m", "viscCorrCut")]


# number of components of each state field
FIELDS = {"Q": 7, "Qh": 7, "Qprev": 7, "pi": 10, "piH": 10, "Pi": 1, "PiH": 1,
          "pi0": 10, "piH0": 10, "Pi0": 1, "PiH0": 1, "m": 3, "viscCorrCut": 1}


class FoGrid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int)] + [
        (n, ctypes.c_double) for n in ("dx", "dy", "dz", "xmin", "ymin", "zmin")]


def make_trc(etaS=0.0, zetaS=0.0, etaSparam=0, zetaSparam=0, **kw):
    """TransportCoeff with the reference driver's defaults (src/main.cpp:70-81)."""
    d = dict(etaS0=etaS, zetaS0=zetaS, ah=0.0, al=0.0, aRho=0.0, T0=0.15,
             etaSMin=0.08, etaSEpsilonMin=5.0, etaSShiftMuB=0.0, etaSScaleMuB=0.0,
             zetaSPeakEpsilon=5.0, zetaSScaleBeta=0.103, zetaSSigmaMinus=0.1,
             zetaSSigmaPlus=0.1, etaSparam=etaSparam, zetaSparam=zetaSparam,
             zetaSPeakT=0.170, zetaSWidthMinus=0.010, zetaSWidthPlus=0.030)
    d.update(kw)
    return Trc(**d)


class OrChiral(ctypes.Structure):
    _fields_ = [("ne", ctypes.c_int), ("nn", ctypes.c_int)] + [
        (k, ctypes.c_double) for k in ("emin", "emax", "nmin", "nmax")] + [
        (k, ctypes.POINTER(ctypes.c_double)) for k in ("p", "T", "mub", "mus")]


class OrHadron(ctypes.Structure):
    _fields_ = [("ne", ctypes.c_int), ("nnb", ctypes.c_int), ("nnq", ctypes.c_int)] + [
        (k, ctypes.c_double) for k in ("e0", "n0", "logemin", "logemax", "lognmax")] + [
        (k, ctypes.POINTER(ctypes.c_double)) for k in ("p", "T", "mub", "muq", "mus")] + [
        ("status", ctypes.POINTER(ctypes.c_int))]


class OrEos(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("n", ctypes.c_int)] + [
        (k, ctypes.POINTER(ctypes.c_double)) for k in ("e", "p", "T", "mu")] + [
        ("small", ctypes.POINTER(OrChiral)), ("big", ctypes.POINTER(OrChiral)),
        ("hadron", ctypes.POINTER(OrHadron))]


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _eos(table):
    """(pointer to an or_eos_t or None, arrays kept alive) of a table or None
    (the ideal gas).  A table is any object with arrays e, p, T [, mu] (the
    spline EoS), or with `small` and `big` (the chiral EoS: each with ne, nn,
    emin, emax, nmin, nmax and (ne, nn) arrays p, T, mub, mus), or with
    e0, n0, logemin, logemax, lognmax and (ne, nnb, nnq) arrays p, T, mub,
    muq, mus, status (the hadron EoS)."""
    if table is None:
        return None, []
    if hasattr(table, "small"):
        keep, tabs = [], []
        for t in (table.small, table.big):
            arrs = [_f64(getattr(t, k)) for k in ("p", "T", "mub", "mus")]
            ne, nn = arrs[0].shape
            c = OrChiral(ne, nn, t.emin, t.emax, t.nmin, t.nmax, *[_p(a) for a in arrs])
            keep += arrs + [c]
            tabs.append(c)
        d = OrEos(2, 0, None, None, None, None, ctypes.pointer(tabs[0]),
                  ctypes.pointer(tabs[1]), None)
        return ctypes.byref(d), [d] + keep
    if hasattr(table, "lognmax"):
        arrs = [_f64(getattr(table, k)) for k in ("p", "T", "mub", "muq", "mus")]
        st = np.ascontiguousarray(getattr(table, "status", np.zeros(arrs[0].shape)),
                                  dtype=np.int32)
        ne, nnb, nnq = arrs[0].shape
        h = OrHadron(ne, nnb, nnq, table.e0, table.n0, table.logemin, table.logemax,
                     table.lognmax, *[_p(a) for a in arrs],
                     st.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        d = OrEos(3, 0, None, None, None, None, None, None, ctypes.pointer(h))
        return ctypes.byref(d), [d, h, st] + arrs
    arrs = [_f64(getattr(table, k)) for k in ("e", "p", "T")]
    mu = getattr(table, "mu", None)
    arrs.append(None if mu is None else _f64(mu))
    d = OrEos(1, len(arrs[0]), *[None if a is None else _p(a) for a in arrs])
    return ctypes.byref(d), [d] + arrs


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run make -C oracle")
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        L.or_create.restype = ctypes.c_void_p
        L.or_create.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 6 + [
            ctypes.POINTER(OrEos), ctypes.POINTER(Trc)]
        L.or_destroy.argtypes = [ctypes.c_void_p]
        L.or_ncell.restype = ctypes.c_long
        L.or_ncell.argtypes = [ctypes.c_void_p]
        L.or_set_state.argtypes = [ctypes.c_void_p, ctypes.POINTER(State)]
        L.or_get_state.argtypes = [ctypes.c_void_p, ctypes.POINTER(State)]
        L.or_set_prim.argtypes = [ctypes.c_void_p, ctypes.c_double] + [dp] * 7
        L.or_run.restype = ctypes.c_int
        L.or_run.argtypes = [ctypes.c_void_p, dp, ctypes.c_double, ctypes.c_int]
        L.or_check_dtau.restype = ctypes.c_int
        L.or_branch_counts.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_long)]
        L.or_ghost.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_update_m.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        L.or_advect.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        L.or_hlle.restype = ctypes.c_int
        L.or_hlle.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 5 + [ctypes.c_double] * 2 + [dp]
        L.or_signal_speed.argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
        L.or_check_dtau.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.or_c2p.argtypes = [ctypes.POINTER(OrEos), dp, ctypes.c_long, dp,
                             ctypes.POINTER(ctypes.c_int)]
        L.or_p2c.argtypes = [dp, ctypes.c_long, dp]
        L.or_eos.argtypes = [ctypes.POINTER(OrEos), dp, ctypes.c_long, dp]
        L.or_eos_eval.argtypes = [ctypes.POINTER(OrEos), dp, ctypes.c_long, dp]
        L.or_eos_spline.argtypes = [ctypes.POINTER(OrEos), ctypes.c_int, dp, ctypes.c_long, dp]
        L.or_trcoeff.argtypes = [ctypes.POINTER(Trc), ctypes.POINTER(OrEos), dp, ctypes.c_long,
                                 dp]
        L.or_cornelius.argtypes = [dp, ctypes.c_long, dp, ctypes.c_double, dp, dp,
                                   ctypes.POINTER(ctypes.c_int)]
        L.or_freezeout.argtypes = [ctypes.POINTER(FoGrid), dp, dp, dp, dp, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_int, dp,
                                   ctypes.c_long, ctypes.POINTER(OrEos), dp, dp]
        L.or_corona.argtypes = [ctypes.POINTER(FoGrid), dp, dp, dp, ctypes.c_double,
                                ctypes.c_double, ctypes.c_int, dp, ctypes.c_long,
                                ctypes.POINTER(OrEos), dp]
        L.or_corona.restype = ctypes.c_long
        L.or_set_vorticity.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_get_dbeta.argtypes = [ctypes.c_void_p, dp]
        L.or_freezeout.restype = ctypes.c_long
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def c2p(Q, table=None):
    """transformPV over rows of Q (n x 7). Returns (prim n x 8, iters n, Q_after)."""
    Q = np.ascontiguousarray(Q, dtype=np.float64).copy()
    n = Q.shape[0]
    out = np.zeros((n, 8))
    it = np.zeros(n, dtype=np.int32)
    d, keep = _eos(table)
    lib().or_c2p(d, _p(Q), n, _p(out), it.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return out, it, Q


def p2c(prim):
    prim = np.ascontiguousarray(prim, dtype=np.float64)
    Q = np.zeros((prim.shape[0], 7))
    lib().or_p2c(_p(prim), prim.shape[0], _p(Q))
    return Q


def eos(e, table=None):
    """n x 4 {p, T, s, cs2} of the ideal gas or of a table"""
    e = np.ascontiguousarray(e, dtype=np.float64)
    out = np.zeros((e.shape[0], 4))
    d, keep = _eos(table)
    lib().or_eos(d, _p(e), e.shape[0], _p(out))
    return out


def eos_eval(table, e, nb=None, nq=None, ns=None):
    """EoS::eos and EoS::s: n x 6 {p, T, mub, muq, mus, s} at (e, nb, nq, ns)"""
    e = _f64(e)
    z = np.zeros_like(e)
    inp = _f64(np.stack([e, z if nb is None else nb, z if nq is None else nq,
                         z if ns is None else ns], axis=1))
    out = np.zeros((e.shape[0], 6))
    d, keep = _eos(table)
    lib().or_eos_eval(d, _p(inp), e.shape[0], _p(out))
    return out


def eos_spline(table, f, v):
    """the table's natural cubic spline of f (0 p, 1 T, 2 mu) at v"""
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.zeros(v.shape[0])
    d, keep = _eos(table)
    lib().or_eos_spline(d, f, _p(v), v.shape[0], _p(out))
    return out


def trcoeff(trc, inp, table=None):
    inp = np.ascontiguousarray(inp, dtype=np.float64)
    out = np.zeros((inp.shape[0], 4))
    d, keep = _eos(table)
    lib().or_trcoeff(ctypes.byref(trc), d, _p(inp), inp.shape[0], _p(out))
    return out


class Solver:
    """Oracle hydro grid + Hydro::performStep driver (CPU, serial)."""

    def __init__(self, nx, ny, nz, xlim, ylim, zlim, trc=None, table=None):
        self.trc = trc if trc is not None else make_trc()
        d, keep = _eos(table)
        self.h = lib().or_create(nx, ny, nz, xlim[0], xlim[1], ylim[0], ylim[1],
                                 zlim[0], zlim[1], d, ctypes.byref(self.trc))
        if not self.h:
            raise ValueError("bad grid")
        self.shape = (nx, ny, nz)
        self.N = nx * ny * nz

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_destroy(self.h)
            self.h = None

    def set_prim(self, tau, e, vx, vy, vz, nb=None, nq=None, ns=None):
        z = np.zeros(self.N)
        arrs = [np.ascontiguousarray(a if a is not None else z, dtype=np.float64).ravel()
                for a in (e, nb, nq, ns, vx, vy, vz)]
        lib().or_set_prim(self.h, tau, *[_p(a) for a in arrs])

    def get_state(self):
        st = State()
        out = {}
        for k, nc in FIELDS.items():
            a = np.zeros(nc * self.N)
            out[k] = a
            setattr(st, k, _p(a))
        lib().or_get_state(self.h, ctypes.byref(st))
        return {k: v.reshape(FIELDS[k], self.N) for k, v in out.items()}

    def set_state(self, d):
        st = State()
        keep = []
        for k, nc in FIELDS.items():
            if k in d:
                a = np.ascontiguousarray(d[k], dtype=np.float64).reshape(-1)
                assert a.size == nc * self.N, k
                keep.append(a)
                setattr(st, k, _p(a))
        lib().or_set_state(self.h, ctypes.byref(st))

    def run(self, tau, dt, nsteps):
        t = ctypes.c_double(tau)
        rc = lib().or_run(self.h, ctypes.byref(t), dt, nsteps)
        return rc, t.value

    def set_vorticity(self, on=True):
        lib().or_set_vorticity(self.h, int(on))

    def ghost(self, full=False):
        """the ghost fill alone (correctImagCells / correctImagCellsFull)"""
        lib().or_ghost(self.h, int(full))

    def update_m(self, tau, dt):
        """Fluid::updateM alone"""
        lib().or_update_m(self.h, tau, dt)

    def advect(self, tau, dt):
        """ISformal loop #2 alone (advection of pi0, piH0, Pi0, PiH0, limiter)"""
        lib().or_advect(self.h, tau, dt)

    def hlle(self, ix, iy, iz, d, mode, tau, dt):
        """(formed, flux[7]) of one interface's hlle_flux"""
        f = np.zeros(7)
        rc = lib().or_hlle(self.h, ix, iy, iz, d, mode, tau, dt, _p(f))
        return rc, f

    def branch_counts(self):
        """(negative-pressure re-evaluations of hlle_flux in the predictor, in
        the corrector, limiter rescales, Pi clamps) since creation"""
        out = (ctypes.c_long * 4)()
        lib().or_branch_counts(self.h, out)
        return tuple(int(v) for v in out)

    def signal_speed(self, reset=False):
        """9 values: the last step's predictor and corrector maxima of
        max(-bl, br) per axis (x, y, eta), then the per-axis maxima since the
        state was set (or the last reset)"""
        out = np.zeros(9)
        lib().or_signal_speed(self.h, int(reset), _p(out))
        return out

    def get_dbeta(self):
        """Cell::getDbeta of every cell: (16, N), row 4 i + j = d_i (u^j / T)"""
        out = np.zeros(16 * self.N)
        lib().or_get_dbeta(self.h, _p(out))
        return out.reshape(16, self.N)


def cornelius(h, dx, E0):
    """Cornelius::find_surface_4d over hypercubes h (n x 2 x 2 x 2 x 2):
    (n_elements per hypercube, centroids n x 10 x 4, normals n x 10 x 4)."""
    h = np.ascontiguousarray(h, dtype=np.float64).reshape(-1, 16)
    n = h.shape[0]
    cen = np.zeros((n, 10, 4))
    nrm = np.zeros((n, 10, 4))
    nel = np.zeros(n, dtype=np.int32)
    d = np.ascontiguousarray(dx, dtype=np.float64)
    lib().or_cornelius(_p(h), n, _p(d), E0, _p(cen), _p(nrm),
                       nel.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return nel, cen, nrm


def freezeout(shape, spacing, mins, Q, Qlow, pi, Pi, tau, dt, ecrit, extend=False,
              max_rows=1000000, table=None, dbeta=None):
    """Fluid::outputSurface's element records from SoA state arrays; with
    dbeta (16 x N), also the beta.dat records: (rows, beta)."""
    g = FoGrid(shape[0], shape[1], shape[2], spacing[0], spacing[1], spacing[2], mins[0],
               mins[1], mins[2])
    ncol = 29 if extend else 27
    rows = np.zeros((max_rows, ncol))
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (Q, Qlow, pi, Pi)]
    d, keep = _eos(table)
    db = beta = None
    if dbeta is not None:
        db = np.ascontiguousarray(dbeta, dtype=np.float64)
        beta = np.zeros((max_rows, 33))
    n = lib().or_freezeout(ctypes.byref(g), *[_p(a) for a in arrs], tau, dt, ecrit,
                           int(extend), _p(rows), max_rows, d,
                           None if db is None else _p(db), None if beta is None else _p(beta))
    if n < 0:
        raise RuntimeError("or_freezeout: construction failed")
    if beta is not None:
        return rows[:min(n, max_rows)], beta[:min(n, max_rows)]
    return rows[:min(n, max_rows)]


def corona(shape, spacing, mins, Q, pi, Pi, tau, ecrit, extend=False, max_rows=1000000,
           table=None):
    """Fluid::outputCorona's element records and stats (count, sum dV_eff,
    EtotSurf increment, sum n_b dV_eff)."""
    g = FoGrid(shape[0], shape[1], shape[2], spacing[0], spacing[1], spacing[2], mins[0],
               mins[1], mins[2])
    ncol = 29 if extend else 27
    rows = np.zeros((max_rows, ncol))
    st = np.zeros(4)
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (Q, pi, Pi)]
    d, keep = _eos(table)
    n = lib().or_corona(ctypes.byref(g), *[_p(a) for a in arrs], tau, ecrit, int(extend),
                        _p(rows), max_rows, d, _p(st))
    return rows[:min(n, max_rows)], st
