"""ctypes binding of the C-ABI in include/vhlle_gpu.h (libvhlle_gpu.so).

This is plumbing for tests and the benchmark: it loads the in-tree HIP library
and fails loudly if it is missing -- there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvhlle_gpu.so")

VH_OK, VH_ERR_ARG, VH_ERR_HIP, VH_ERR_NODEVICE, VH_ERR_DTAU = 0, -1, -2, -3, -4
ST_NAN_FLUX, ST_BIG_E, ST_NAN_SOURCE, ST_NAN_PI, ST_EOS_RANGE = 1, 2, 4, 8, 16
EOS_IDEAL, EOS_TABLE, EOS_CHIRAL, EOS_HADRON = 0, 1, 2, 3
PREC_FP64, PREC_MIXED = 0, 1
C2P_REFERENCE, C2P_LEARNED, C2P_CLOSED, C2P_NEWTON = 0, 1, 2, 3
START_MLP, START_HALF = 0, 1
FO_QPREV, FO_SNAPSHOT = 0, 1
BETA_COLS = 33  # VH_BETA_COLS: columns of a beta.dat record
STAGE_AUTO, STAGE_CELLS, STAGE_ROWS, STAGE_BANDS = 0, 1, 2, 3
CELLS_NOALLOC_MAX = 8192
MLP_NPARAM = 4674

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)


class Grid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int)] + [
        (n, ctypes.c_double) for n in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")]


class ChiralTab(ctypes.Structure):
    """vh_eos_chiral_tab_t: one EoSaux table of EoSChiral"""
    _fields_ = [("ne", ctypes.c_int), ("nn", ctypes.c_int)] + [
        (k, ctypes.c_double) for k in ("emin", "emax", "nmin", "nmax")] + [
        (k, ctypes.POINTER(ctypes.c_double)) for k in ("p", "T", "mub", "mus")]


class HadronTab(ctypes.Structure):
    """vh_eos_hadron_tab_t: the EoSHadron table"""
    _fields_ = [("ne", ctypes.c_int), ("nnb", ctypes.c_int), ("nnq", ctypes.c_int)] + [
        (k, ctypes.c_double) for k in ("e0", "n0", "logemin", "logemax", "lognmax")] + [
        (k, ctypes.POINTER(ctypes.c_double)) for k in ("p", "T", "mub", "muq", "mus")] + [
        ("status", ctypes.POINTER(ctypes.c_int))]


class Eos(ctypes.Structure):
    """vh_eos_t: kind (EOS_IDEAL, EOS_TABLE, EOS_CHIRAL, EOS_HADRON) and the
    kind's tables."""
    _fields_ = [("kind", ctypes.c_int), ("n", ctypes.c_int)] + [
        (k, ctypes.POINTER(ctypes.c_double)) for k in ("e", "p", "T", "mu")] + [
        ("small", ctypes.POINTER(ChiralTab)), ("big", ctypes.POINTER(ChiralTab)),
        ("hadron", ctypes.POINTER(HadronTab))]


class EosTable:
    """A tabulated EoS (the TABLE branch of EoSs): increasing e (GeV/fm^3),
    p (GeV/fm^3), T (GeV), mu (GeV, optional).  desc() is the vh_eos_t (the
    arrays stay alive with this object)."""

    def __init__(self, e, p, T, mu=None):
        self.e = np.ascontiguousarray(e, dtype=np.float64)
        self.p = np.ascontiguousarray(p, dtype=np.float64)
        self.T = np.ascontiguousarray(T, dtype=np.float64)
        self.mu = None if mu is None else np.ascontiguousarray(mu, dtype=np.float64)
        self._d = Eos(EOS_TABLE, len(self.e), _p(self.e), _p(self.p), _p(self.T),
                      None if self.mu is None else _p(self.mu))

    @classmethod
    def read(cls, path, ncols=3):
        """Columns e, p, T [, mu] of a whitespace table file (EoSs's format)."""
        a = np.loadtxt(path, ndmin=2)
        return cls(a[:, 0], a[:, 1], a[:, 2], a[:, 3] if ncols > 3 else None)

    def desc(self):
        return self._d

    def spline_table(self):
        """(13 n) spline table of vh_eos_table: nodes, then y, b, c, d of p, T, mu"""
        buf = np.zeros(13 * len(self.e))
        check(lib().vh_eos_table(ctypes.byref(self._d), _p(buf)), "vh_eos_table")
        return buf

    def spline(self, f, v):
        """the spline of f (0 p, 1 T, 2 mu) at v, evaluated in numpy from the
        library's coefficients (same formula as the device EoS)"""
        n = len(self.e)
        t = self.spline_table().reshape(13, n)
        x = t[0]
        y, b, c, d = t[1 + 4 * f:5 + 4 * f]
        v = np.asarray(v, dtype=np.float64)
        i = np.clip(np.searchsorted(x, v, side="right") - 1, 0, n - 2)
        dx = v - x[i]
        return y[i] + dx * (b[i] + dx * (c[i] + dx * d[i]))


class ChiralTable:
    """One EoSaux table of EoSChiral (src/eoChiral.cpp:38-84): fields p, T,
    mub, mus as (ne, nn) arrays over the uniform nodes e in [emin, emax], n_b
    in [nmin, nmax] (GeV/fm^3, 1/fm^3, GeV)."""

    def __init__(self, emin, emax, nmin, nmax, p, T, mub, mus):
        self.p, self.T, self.mub, self.mus = (np.ascontiguousarray(a, dtype=np.float64)
                                              for a in (p, T, mub, mus))
        self.ne, self.nn = self.p.shape
        self.emin, self.emax, self.nmin, self.nmax = emin, emax, nmin, nmax
        self._d = ChiralTab(self.ne, self.nn, emin, emax, nmin, nmax,
                            *(_p(a) for a in (self.p, self.T, self.mub, self.mus)))

    @classmethod
    def read(cls, path, ne, nn):
        """The reference's file format (EoSaux::EoSaux, src/eoChiral.cpp:38-84):
        nn blocks of ne rows of T [MeV], mu_q [MeV], e [e_0], p [e_0], n [n_0],
        s [n_0], mu_S [MeV], a placeholder; converted to GeV, GeV/fm^3, 1/fm^3
        with the reference's factors (mu_b = 3 mu_q, e_0 = 0.146 GeV/fm^3,
        n_0 = 0.15 /fm^3)."""
        a = np.loadtxt(path).reshape(nn, ne, 8)
        T = a[:, :, 0].T / 1000.0
        mub = a[:, :, 1].T * 3.0 / 1000.0
        p = a[:, :, 3].T * 0.146
        mus = a[:, :, 6].T / 1000.0
        e = a[0, :, 2] * 0.146
        n = a[:, 0, 4] * 0.15
        return cls(e[0], e[-1], n[0], n[-1], p, T, mub, mus)


class EosChiral:
    """EoSChiral (src/eoChiral.cpp:131-154): `small` where e < 1.46 and
    n_b < 0.3, `big` where e < 146 and n_b < 6, else p = 0.2964 e.  FP64
    contexts only."""

    def __init__(self, small, big):
        self.small, self.big = small, big
        self._d = Eos(EOS_CHIRAL, 0, None, None, None, None,
                      ctypes.pointer(small._d), ctypes.pointer(big._d), None)

    @classmethod
    def read(cls, big_path, small_path):
        """EoSChiral() reads eos/chiraleos.dat (2001 x 401) and
        eos/chiralsmall.dat (201 x 201)."""
        return cls(ChiralTable.read(small_path, 201, 201), ChiralTable.read(big_path, 2001, 401))

    def desc(self):
        return self._d


class EosHadron:
    """EoSHadron (src/eoHadron.cpp): fields p, T, mub, muq, mus and the node
    status as (ne, nnb, nnq) arrays; nodes uniform in log(e/e0) over [logemin,
    logemax] and in the log-mapped n_b, n_q over [-lognmax, lognmax].  FP64
    contexts only."""

    def __init__(self, e0, n0, logemin, logemax, lognmax, p, T, mub, muq, mus, status=None):
        self.p, self.T, self.mub, self.muq, self.mus = (
            np.ascontiguousarray(a, dtype=np.float64) for a in (p, T, mub, muq, mus))
        self.ne, self.nnb, self.nnq = self.p.shape
        self.status = np.ascontiguousarray(np.zeros(self.p.shape) if status is None else status,
                                           dtype=np.int32)
        self.e0, self.n0, self.logemin, self.logemax, self.lognmax = (
            e0, n0, logemin, logemax, lognmax)
        self._t = HadronTab(self.ne, self.nnb, self.nnq, e0, n0, logemin, logemax, lognmax,
                            *(_p(a) for a in (self.p, self.T, self.mub, self.muq, self.mus)),
                            self.status.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        self._d = Eos(EOS_HADRON, 0, None, None, None, None, None, None, ctypes.pointer(self._t))

    @classmethod
    def read(cls, path):
        """The reference's file format (EoSHadron::EoSHadron,
        src/eoHadron.cpp:13-65): a header ne nnb nnq e0 n0 logemin logemax
        lognmax, then rows e nb nq p T mub muq mus status in (ie, inb, inq)
        order."""
        with open(path) as f:
            head = f.readline().split()
            while len(head) < 8:
                head += f.readline().split()
            ne, nnb, nnq = (int(v) for v in head[:3])
            e0, n0, logemin, logemax, lognmax = (float(v) for v in head[3:8])
            a = np.loadtxt(f, ndmin=2).reshape(ne, nnb, nnq, 9)
        return cls(e0, n0, logemin, logemax, lognmax, *(a[..., k] for k in range(3, 8)),
                   status=a[..., 8].astype(np.int32))

    def desc(self):
        return self._d


def eos_desc(eos):
    """vh_eos_t of None (ideal gas), an EosTable, an EosChiral or an EosHadron"""
    return Eos(EOS_IDEAL, 0, None, None, None, None) if eos is None else eos.desc()


def eos_eval(eos, e, nb=None, nq=None, ns=None, on_gpu=True):
    """vh_eos_eval: n x 6 {p, T, mu_b, mu_q, mu_s, s} at (e, nb, nq, ns)"""
    e = np.asarray(e, dtype=np.float64)
    z = np.zeros_like(e)
    inp = np.ascontiguousarray(np.stack([e, z if nb is None else nb, z if nq is None else nq,
                                         z if ns is None else ns], axis=1), dtype=np.float64)
    out = np.zeros((len(e), 6))
    d = eos_desc(eos)
    check(lib().vh_eos_eval(ctypes.byref(d), _p(inp), len(e), _p(out), 1 if on_gpu else 0),
          "vh_eos_eval")
    return out


class Trc(ctypes.Structure):
    """vh_trcoeff_t: TransportCoeff fields (src/trancoeff.h:8-11) + zetaSparam 5."""
    _fields_ = [(n, ctypes.c_double) for n in (
        "etaS0", "zetaS0", "ah", "al", "aRho", "T0", "etaSMin", "etaSEpsilonMin",
        "etaSShiftMuB", "etaSScaleMuB", "zetaSPeakEpsilon", "zetaSScaleBeta",
        "zetaSSigmaMinus", "zetaSSigmaPlus")] + [
        ("etaSparam", ctypes.c_int), ("zetaSparam", ctypes.c_int),
        ("zetaSPeakT", ctypes.c_double), ("zetaSWidthMinus", ctypes.c_double),
        ("zetaSWidthPlus", ctypes.c_double)]


class MlpTrainCfg(ctypes.Structure):
    _fields_ = [("batch", ctypes.c_int), ("epochs", ctypes.c_int), ("lr", ctypes.c_double),
                ("lr_final", ctypes.c_double), ("seed", ctypes.c_ulonglong)]


class HarvestCfg(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_ulonglong), ("steps", ctypes.c_int), ("namp", ctypes.c_int),
                ("amp", ctypes.c_double * 4), ("augment", ctypes.c_double),
                ("vmax", ctypes.c_double), ("eos", ctypes.POINTER(Eos))]


class HotspotCfg(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_ulonglong)] + [(n, ctypes.c_double) for n in (
        "npart_mean", "R", "a", "sigma", "amp", "eta_plateau", "eta_sigma")]


class FoStats(ctypes.Structure):
    _fields_ = [("nelements", ctypes.c_long), ("nsusp", ctypes.c_long),
                ("dsigma", ctypes.c_double * 4), ("v_eff", ctypes.c_double),
                ("e_surf", ctypes.c_double), ("nb_surf", ctypes.c_double),
                ("err", ctypes.c_int)]


class State(ctypes.Structure):
    _fields_ = [(n, _dp) for n in ("Q", "Qh", "Qprev", "pi", "piH", "Pi", "PiH", "m",
                                   "viscCorrCut")]


STATE_FIELDS = {"Q": 7, "Qh": 7, "Qprev": 7, "pi": 10, "piH": 10, "Pi": 1, "PiH": 1, "m": 3,
                "viscCorrCut": 1}


def make_trc(etaS=0.0, zetaS=0.0, etaSparam=0, zetaSparam=0, **kw):
    """TransportCoeff with the reference driver's defaults (src/main.cpp:70-81)."""
    d = dict(etaS0=etaS, zetaS0=zetaS, ah=0.0, al=0.0, aRho=0.0, T0=0.15, etaSMin=0.08,
             etaSEpsilonMin=5.0, etaSShiftMuB=0.0, etaSScaleMuB=0.0, zetaSPeakEpsilon=5.0,
             zetaSScaleBeta=0.103, zetaSSigmaMinus=0.1, zetaSSigmaPlus=0.1,
             etaSparam=etaSparam, zetaSparam=zetaSparam, zetaSPeakT=0.170,
             zetaSWidthMinus=0.010, zetaSWidthPlus=0.030)
    d.update(kw)
    return Trc(**d)


_LIB = None


def lib():
    """Loads libvhlle_gpu.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} missing: run `python __graft_entry__.py build` first")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    ip = ctypes.POINTER(ctypes.c_int)
    sig = {
        "vh_create": (ctypes.c_int, [ctypes.POINTER(Grid), ctypes.POINTER(Eos),
                                     ctypes.POINTER(Trc), ctypes.c_int, ctypes.POINTER(vp)]),
        "vh_destroy": (ctypes.c_int, [vp]),
        "vh_ncell": (ctypes.c_long, [vp]),
        "vh_upload": (ctypes.c_int, [vp, ctypes.POINTER(State)]),
        "vh_download": (ctypes.c_int, [vp, ctypes.POINTER(State)]),
        "vh_set_prim": (ctypes.c_int, [vp, ctypes.c_double] + [_dp] * 7),
        "vh_step": (ctypes.c_int, [vp, _dp, ctypes.c_double, ip]),
        "vh_run": (ctypes.c_int, [vp, _dp, ctypes.c_double, ctypes.c_int, ip]),
        "vh_check_dtau": (ctypes.c_int, [vp, ctypes.c_double]),
        "vh_signal_speed": (ctypes.c_int, [vp, ctypes.c_int, _dp]),
        "vh_set_c2p_mode": (ctypes.c_int, [vp, ctypes.c_int]),
        "vh_enable_timing": (ctypes.c_int, [vp, ctypes.c_int]),
        "vh_kernel_times": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_void_p, ip, _dp]),
        "vh_c2p": (ctypes.c_int, [ctypes.POINTER(Eos), _dp, ctypes.c_long, _dp, ip]),
        "vh_p2c": (ctypes.c_int, [_dp, ctypes.c_long, _dp]),
        "vh_device_info": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int]),
        "vh_mlp_init": (ctypes.c_int, [ctypes.c_ulonglong, _fp]),
        "vh_mlp_forward": (ctypes.c_int, [_fp, _dp, ctypes.c_long, _fp]),
        "vh_mlp_grad": (ctypes.c_int, [_fp, _dp, _dp, ctypes.c_long, _fp, _dp]),
        "vh_mlp_train": (ctypes.c_int, [ctypes.POINTER(MlpTrainCfg), _dp, _dp, ctypes.c_long,
                                        _fp, _dp]),
        "vh_c2p_learned": (ctypes.c_int, [ctypes.POINTER(Eos), _fp, _dp, ctypes.c_long,
                                          ctypes.c_int, _dp, ip]),
        "vh_set_c2p_model": (ctypes.c_int, [vp, _fp]),
        "vh_mlp_harvest": (ctypes.c_int, [ctypes.POINTER(HarvestCfg), ctypes.c_long, _dp, _dp]),
        "vh_set_stage": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int]),
        "vh_get_stage": (ctypes.c_int, [vp, ip, ip, _dp]),
        "vh_set_stage_bands": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int]),
        "vh_get_stage_bands": (ctypes.c_int, [vp, ip]),
        "vh_hotspot_ic": (ctypes.c_int, [ctypes.POINTER(HotspotCfg), ctypes.c_int,
                                         ctypes.POINTER(Grid), _dp, ip, _dp, ctypes.c_long]),
        "vh_copy_bandwidth": (ctypes.c_int, [ctypes.c_long, ctypes.c_int, _dp]),
        "vh_fo_snapshot": (ctypes.c_int, [vp, ctypes.c_double]),
        "vh_freezeout": (ctypes.c_int, [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_int, ctypes.c_int, _dp, ctypes.c_long,
                                        ctypes.POINTER(ctypes.c_long), ctypes.POINTER(FoStats)]),
        "vh_cornelius": (ctypes.c_int, [_dp, ctypes.c_long, _dp, ctypes.c_double, _dp, _dp, ip]),
        "vh_corona": (ctypes.c_int, [vp, ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp,
                                     ctypes.c_long, ctypes.POINTER(ctypes.c_long),
                                     ctypes.POINTER(FoStats)]),
        "vh_prim_fields": (ctypes.c_int, [vp, ctypes.c_double, _dp]),
        "vh_fluid_stats": (ctypes.c_int, [vp, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                          _dp]),
        "vh_cells": (ctypes.c_int, [vp, ctypes.c_double, ctypes.POINTER(ctypes.c_long),
                                    ctypes.c_long, _dp]),
        "vh_cells_noalloc": (ctypes.c_int, [vp, ctypes.c_double, ctypes.POINTER(ctypes.c_long),
                                            ctypes.c_long, _dp]),
        "vh_set_vorticity": (ctypes.c_int, [vp, ctypes.c_int]),
        "vh_get_dbeta": (ctypes.c_int, [vp, _dp]),
        "vh_freezeout_beta": (ctypes.c_int, [vp, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp,
                                             _dp, ctypes.c_long, ctypes.POINTER(ctypes.c_long),
                                             ctypes.POINTER(FoStats)]),
        "vh_eos_table": (ctypes.c_int, [ctypes.POINTER(Eos), _dp]),
        "vh_eos_eval": (ctypes.c_int, [ctypes.POINTER(Eos), _dp, ctypes.c_long, _dp, ctypes.c_int]),
        "vh_set_surface_eos": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Eos)]),
        "vh_create_batch": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(Grid), ctypes.POINTER(Eos),
                                           ctypes.POINTER(Trc), ctypes.c_int, ctypes.POINTER(vp)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


def exported_symbols():
    return ["vh_create", "vh_destroy", "vh_ncell", "vh_upload", "vh_download", "vh_set_prim",
            "vh_step", "vh_run", "vh_check_dtau", "vh_set_c2p_mode", "vh_enable_timing",
            "vh_kernel_times", "vh_c2p", "vh_p2c", "vh_device_info", "vh_mlp_init",
            "vh_mlp_forward", "vh_mlp_grad", "vh_mlp_train", "vh_c2p_learned",
            "vh_set_c2p_model", "vh_mlp_harvest", "vh_set_stage", "vh_get_stage", "vh_create_batch", "vh_hotspot_ic",
            "vh_copy_bandwidth", "vh_fo_snapshot", "vh_freezeout", "vh_cornelius",
            "vh_eos_table", "vh_eos_eval", "vh_set_surface_eos", "vh_set_vorticity", "vh_get_dbeta", "vh_freezeout_beta",
            "vh_corona", "vh_prim_fields", "vh_fluid_stats", "vh_cells", "vh_signal_speed",
            "vh_cells_noalloc", "vh_set_stage_bands", "vh_get_stage_bands"]


def _p(a):
    return a.ctypes.data_as(_dp)


def check(rc, what="call"):
    if rc != VH_OK:
        raise RuntimeError(f"vhlle_gpu {what} failed with code {rc}")


def copy_bandwidth(nbytes=1 << 30, reps=10):
    """GB/s of a plain device copy (read + write bytes / time)."""
    g = ctypes.c_double(0.0)
    check(lib().vh_copy_bandwidth(nbytes, reps, ctypes.byref(g)), "vh_copy_bandwidth")
    return g.value


def device_info():
    buf = ctypes.create_string_buffer(256)
    check(lib().vh_device_info(buf, 256), "vh_device_info")
    return buf.value.decode()


def c2p(Q, eos=None):
    """transformPV on the GPU over rows of Q (n x 7): (prim n x 8, iters, Q_after)."""
    Q = np.ascontiguousarray(Q, dtype=np.float64).copy()
    n = Q.shape[0]
    out = np.zeros((n, 8))
    it = np.zeros(n, dtype=np.int32)
    e = eos_desc(eos)
    check(lib().vh_c2p(ctypes.byref(e), _p(Q), n, _p(out),
                       it.ctypes.data_as(ctypes.POINTER(ctypes.c_int))), "vh_c2p")
    return out, it, Q


def _f(a):
    return a.ctypes.data_as(_fp)


def mlp_init(seed):
    """He-uniform initial MLP parameters (float32 vector)."""
    p = np.zeros(MLP_NPARAM, dtype=np.float32)
    check(lib().vh_mlp_init(seed, _f(p)), "vh_mlp_init")
    return p


def mlp_forward(params, Q):
    """MFMA inference over states Q (n x 7, already / tau): (n x 2) float32."""
    params = np.ascontiguousarray(params, dtype=np.float32)
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    out = np.zeros((Q.shape[0], 2), dtype=np.float32)
    check(lib().vh_mlp_forward(_f(params), _p(Q), Q.shape[0], _f(out)), "vh_mlp_forward")
    return out


def mlp_grad(params, Q, prim):
    params = np.ascontiguousarray(params, dtype=np.float32)
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    prim = np.ascontiguousarray(prim, dtype=np.float64)
    g = np.zeros(MLP_NPARAM, dtype=np.float32)
    L = ctypes.c_double(0.0)
    check(lib().vh_mlp_grad(_f(params), _p(Q), _p(prim), Q.shape[0], _f(g), ctypes.byref(L)),
          "vh_mlp_grad")
    return g, L.value


def mlp_train(params, Q, prim, batch=4096, epochs=4, lr=2e-3, lr_final=1e-5, seed=1):
    """Adam training on the GPU: (trained params, per-epoch loss)."""
    p = np.ascontiguousarray(params, dtype=np.float32).copy()
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    prim = np.ascontiguousarray(prim, dtype=np.float64)
    hist = np.zeros(epochs)
    cfg = MlpTrainCfg(batch, epochs, lr, lr_final, seed)
    check(lib().vh_mlp_train(ctypes.byref(cfg), _p(Q), _p(prim), Q.shape[0], _f(p), _p(hist)),
          "vh_mlp_train")
    return p, hist


def c2p_learned(params, Q, start=START_MLP, eos=None):
    """con2prim from the MLP guess (or v = 0.5): (prim n x 8, iters, Q_after)."""
    params = np.ascontiguousarray(params, dtype=np.float32)
    Q = np.ascontiguousarray(Q, dtype=np.float64).copy()
    n = Q.shape[0]
    out = np.zeros((n, 8))
    it = np.zeros(n, dtype=np.int32)
    e = eos_desc(eos)
    check(lib().vh_c2p_learned(ctypes.byref(e), _f(params), _p(Q), n, start, _p(out),
                               it.ctypes.data_as(ctypes.POINTER(ctypes.c_int))),
          "vh_c2p_learned")
    return out, it, Q


def harvest(n, seed=1, steps=200, amps=(10.0, 1.0), augment=0.5, vmax=0.99, eos=None):
    """Training pairs (Q n x 7, prim n x 8) from Gubser runs + augmentation,
    with the ideal gas or an EosTable."""
    a = (ctypes.c_double * 4)(*(list(amps) + [0.0] * (4 - len(amps))))
    d = eos_desc(eos)
    cfg = HarvestCfg(seed, steps, len(amps), a, augment, vmax,
                     None if eos is None else ctypes.pointer(d))
    Q = np.zeros((n, 7))
    prim = np.zeros((n, 8))
    check(lib().vh_mlp_harvest(ctypes.byref(cfg), n, _p(Q), _p(prim)), "vh_mlp_harvest")
    return Q, prim


def hotspot_ic(grids, seed, npart_mean=150.0, R=6.6, a=0.55, sigma=0.5, amp=1.0,
               eta_plateau=2.0, eta_sigma=1.0, spots=False):
    """Hot-spot e of every cell of `grids` ((nx, ny, nz, xlim, ylim, zlim) list),
    concatenated in the batch layout: (e, npart per grid[, spots (n, 3)])."""
    arr = (Grid * len(grids))(*[Grid(a_, b, c, x[0], x[1], y[0], y[1], z[0], z[1])
                                for (a_, b, c, x, y, z) in grids])
    n = sum(a_ * b * c for (a_, b, c, _, _, _) in grids)
    e = np.zeros(n)
    npart = np.zeros(len(grids), dtype=np.int32)
    cfg = HotspotCfg(seed, npart_mean, R, a, sigma, amp, eta_plateau, eta_sigma)
    maxs = int(len(grids) * (npart_mean + 20 * np.sqrt(npart_mean) + 50)) if spots else 0
    sp = np.zeros((max(maxs, 1), 3))
    check(lib().vh_hotspot_ic(ctypes.byref(cfg), len(grids), arr, _p(e),
                              npart.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                              _p(sp) if spots else None, maxs), "vh_hotspot_ic")
    if spots:
        assert npart.sum() <= maxs
        return e, npart, sp[:npart.sum()]
    return e, npart


def cornelius(h, dx, E0):
    """The Cornelius construction on the GPU over hypercubes h (n x 2 x 2 x 2 x 2):
    (elements per hypercube, centroids n x 10 x 4, normals n x 10 x 4)."""
    h = np.ascontiguousarray(h, dtype=np.float64).reshape(-1, 16)
    n = h.shape[0]
    cen = np.zeros((n, 10, 4))
    nrm = np.zeros((n, 10, 4))
    nel = np.zeros(n, dtype=np.int32)
    d = np.ascontiguousarray(dx, dtype=np.float64)
    check(lib().vh_cornelius(_p(h), n, _p(d), E0, _p(cen), _p(nrm),
                             nel.ctypes.data_as(ctypes.POINTER(ctypes.c_int))), "vh_cornelius")
    return nel, cen, nrm


def p2c(prim):
    prim = np.ascontiguousarray(prim, dtype=np.float64)
    Q = np.zeros((prim.shape[0], 7))
    check(lib().vh_p2c(_p(prim), prim.shape[0], _p(Q)), "vh_p2c")
    return Q


class Context:
    """One Fluid + Hydro on the GPU through the C-ABI (or, with `grids`, a batch
    of independent grids: a list of (nx, ny, nz, xlim, ylim, zlim))."""

    def __init__(self, nx=None, ny=None, nz=None, xlim=None, ylim=None, zlim=None, trc=None,
                 precision=PREC_FP64, grids=None, eos=None):
        self.trc = trc if trc is not None else make_trc()
        self.eos = eos
        h = ctypes.c_void_p()
        if grids is None:
            g = Grid(nx, ny, nz, xlim[0], xlim[1], ylim[0], ylim[1], zlim[0], zlim[1])
            check(lib().vh_create(ctypes.byref(g), ctypes.byref(eos_desc(eos)),
                                  ctypes.byref(self.trc), precision, ctypes.byref(h)),
                  "vh_create")
            self.N = nx * ny * nz
            self.shape = (nx, ny, nz)
        else:
            arr = (Grid * len(grids))(*[Grid(a, b, c, x[0], x[1], y[0], y[1], z[0], z[1])
                                        for (a, b, c, x, y, z) in grids])
            check(lib().vh_create_batch(len(grids), arr, ctypes.byref(eos_desc(eos)),
                                        ctypes.byref(self.trc), precision, ctypes.byref(h)),
                  "vh_create_batch")
            self.N = sum(a * b * c for (a, b, c, _, _, _) in grids)
            self.shape = None
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().vh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_prim(self, tau, e, vx, vy, vz, nb=None, nq=None, ns=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64).ravel()
                for a in (e, nb, nq, ns, vx, vy, vz)]
        ptrs = [None if a is None else _p(a) for a in arrs]
        check(lib().vh_set_prim(self.h, tau, *ptrs), "vh_set_prim")

    def upload(self, d):
        st = State()
        keep = []
        for k, nc in STATE_FIELDS.items():
            if k in d:
                a = np.ascontiguousarray(d[k], dtype=np.float64).reshape(-1)
                assert a.size == nc * self.N, k
                keep.append(a)
                setattr(st, k, _p(a))
        check(lib().vh_upload(self.h, ctypes.byref(st)), "vh_upload")

    def download(self):
        st = State()
        out = {}
        for k, nc in STATE_FIELDS.items():
            a = np.zeros(nc * self.N)
            out[k] = a
            setattr(st, k, _p(a))
        check(lib().vh_download(self.h, ctypes.byref(st)), "vh_download")
        return {k: v.reshape(STATE_FIELDS[k], self.N) for k, v in out.items()}

    def step(self, tau, dt):
        t = ctypes.c_double(tau)
        s = ctypes.c_int(0)
        check(lib().vh_step(self.h, ctypes.byref(t), dt, ctypes.byref(s)), "vh_step")
        return s.value, t.value

    def run(self, tau, dt, nsteps):
        t = ctypes.c_double(tau)
        s = ctypes.c_int(0)
        check(lib().vh_run(self.h, ctypes.byref(t), dt, nsteps, ctypes.byref(s)), "vh_run")
        return s.value, t.value

    def set_c2p(self, mode, params=None):
        """con2prim mode (C2P_*); params: the MLP for C2P_LEARNED."""
        if params is not None:
            p = np.ascontiguousarray(params, dtype=np.float32)
            check(lib().vh_set_c2p_model(self.h, _f(p)), "vh_set_c2p_model")
        check(lib().vh_set_c2p_mode(self.h, mode), "vh_set_c2p_mode")

    def set_stage(self, kind=0, CZ=0):
        check(lib().vh_set_stage(self.h, kind, CZ), "vh_set_stage")

    def set_stage_bands(self, CY, CZ):
        check(lib().vh_set_stage_bands(self.h, CY, CZ), "vh_set_stage_bands")

    def get_stage(self):
        k, cz = ctypes.c_int(), ctypes.c_int()
        ms = ctypes.c_double()
        check(lib().vh_get_stage(self.h, ctypes.byref(k), ctypes.byref(cz), ctypes.byref(ms)),
              "vh_get_stage")
        return k.value, cz.value, ms.value

    def get_stage_bands(self):
        cy = ctypes.c_int()
        check(lib().vh_get_stage_bands(self.h, ctypes.byref(cy)), "vh_get_stage_bands")
        return cy.value

    def fo_snapshot(self, tau):
        check(lib().vh_fo_snapshot(self.h, tau), "vh_fo_snapshot")

    def freezeout(self, tau, dt, ecrit, extend=False, lower=FO_QPREV, max_rows=2000000):
        """Surface elements (rows of 27 or 29 columns) and the stats."""
        ncol = 29 if extend else 27
        n = ctypes.c_long(0)
        st = FoStats()
        check(lib().vh_freezeout(self.h, tau, dt, ecrit, int(extend), lower, None, 0,
                                 ctypes.byref(n), ctypes.byref(st)), "vh_freezeout")
        rows = np.zeros((max(min(n.value, max_rows), 1), ncol))
        if n.value and max_rows > 0:
            check(lib().vh_freezeout(self.h, tau, dt, ecrit, int(extend), lower, _p(rows),
                                     min(n.value, max_rows), ctypes.byref(n), ctypes.byref(st)),
                  "vh_freezeout")
        return rows[:min(n.value, max_rows)], dict(
            nelements=st.nelements, nsusp=st.nsusp, dsigma=list(st.dsigma), v_eff=st.v_eff,
            e_surf=st.e_surf, nb_surf=st.nb_surf, err=st.err)

    def corona(self, tau, ecrit, extend=False, max_rows=2000000):
        """Fluid::outputCorona's elements (rows of 27 or 29 columns) and stats"""
        ncol = 29 if extend else 27
        n = ctypes.c_long(0)
        st = FoStats()
        check(lib().vh_corona(self.h, tau, ecrit, int(extend), None, 0, ctypes.byref(n),
                              ctypes.byref(st)), "vh_corona")
        rows = np.zeros((max(min(n.value, max_rows), 1), ncol))
        if n.value and max_rows > 0:
            check(lib().vh_corona(self.h, tau, ecrit, int(extend), _p(rows),
                                  min(n.value, max_rows), ctypes.byref(n), ctypes.byref(st)),
                  "vh_corona")
        return rows[:min(n.value, max_rows)], dict(nelements=st.nelements, v_eff=st.v_eff,
                                                   e_surf=st.e_surf, nb_surf=st.nb_surf)

    def prim_fields(self, tau):
        """Cell::getPrimVar of every cell: (8, N) rows e, p, nb, nq, ns, vx, vy, vz"""
        out = np.zeros(8 * self.N)
        check(lib().vh_prim_fields(self.h, tau, _p(out)), "vh_prim_fields")
        return out.reshape(8, self.N)

    def fluid_stats(self, tau, ecrit, central=True):
        """the 16 grid totals of the reference's output routines"""
        out = np.zeros(16)
        check(lib().vh_fluid_stats(self.h, tau, ecrit, int(central), _p(out)), "vh_fluid_stats")
        return out

    def cells(self, tau, idx):
        """getPrimVar (8), pi (10), Pi, viscCorrCut, 0 of the listed cells"""
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        out = np.zeros((len(idx), 21))
        check(lib().vh_cells(self.h, tau, idx.ctypes.data_as(ctypes.POINTER(ctypes.c_long)),
                             len(idx), _p(out)), "vh_cells")
        return out

    def cells_noalloc(self, tau, idx):
        """cells() through vh_cells_noalloc (no device allocation; at most
        CELLS_NOALLOC_MAX cells); returns (status code, records)"""
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        out = np.zeros((len(idx), 21))
        rc = lib().vh_cells_noalloc(self.h, tau, idx.ctypes.data_as(ctypes.POINTER(ctypes.c_long)),
                                    len(idx), _p(out))
        return rc, out

    def set_vorticity(self, on=True):
        check(lib().vh_set_vorticity(self.h, int(on)), "vh_set_vorticity")

    def set_surface_eos(self, eos):
        """vh_set_surface_eos: the EoS of the freeze-out outputs (None: the
        hydro EoS)"""
        self._surface_eos = eos  # keeps the tables alive
        d = None if eos is None else ctypes.byref(eos.desc())
        check(lib().vh_set_surface_eos(self.h, d), "vh_set_surface_eos")

    def get_dbeta(self):
        """Cell::getDbeta of every cell: (16, N), row 4 i + j = d_i (u^j / T)"""
        out = np.zeros(16 * self.N)
        check(lib().vh_get_dbeta(self.h, _p(out)), "vh_get_dbeta")
        return out.reshape(16, self.N)

    def freezeout_beta(self, tau, dt, ecrit, extend=False, lower=FO_QPREV, max_rows=2000000):
        """freezeout() plus the beta.dat records (rows of BETA_COLS columns)"""
        ncol = 29 if extend else 27
        n = ctypes.c_long(0)
        st = FoStats()
        check(lib().vh_freezeout_beta(self.h, tau, dt, ecrit, int(extend), lower, None, None, 0,
                                      ctypes.byref(n), ctypes.byref(st)), "vh_freezeout_beta")
        m = max(min(n.value, max_rows), 1)
        rows, beta = np.zeros((m, ncol)), np.zeros((m, BETA_COLS))
        if n.value and max_rows > 0:
            check(lib().vh_freezeout_beta(self.h, tau, dt, ecrit, int(extend), lower, _p(rows),
                                          _p(beta), min(n.value, max_rows), ctypes.byref(n),
                                          ctypes.byref(st)), "vh_freezeout_beta")
        k = min(n.value, max_rows)
        return rows[:k], beta[:k]

    def enable_timing(self, on=True):
        check(lib().vh_enable_timing(self.h, 1 if on else 0), "vh_enable_timing")

    def kernel_times(self):
        mx = 64
        names = (ctypes.c_char * 32 * mx)()
        launches = np.zeros(mx, dtype=np.int32)
        ms = np.zeros(mx)
        n = lib().vh_kernel_times(self.h, mx, names,
                                  launches.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _p(ms))
        return {names[i].value.decode(): (int(launches[i]), float(ms[i])) for i in range(n)}

    def check_dtau(self, dt):
        return lib().vh_check_dtau(self.h, dt)

    def signal_speed(self, reset=False):
        """9 values: the last step's predictor and corrector maxima of the
        HLLE signal speed max(-b_l, b_r) per axis (x, y, eta), then the per-axis
        maxima since the state was set (or the last reset)"""
        out = np.zeros(9)
        check(lib().vh_signal_speed(self.h, int(reset), _p(out)), "vh_signal_speed")
        return out
