"""Synthetic tables of the charge-dependent EoSs for the tests (the
reference's own tables, eos/chiraleos.dat, eos/chiralsmall.dat and
eos/eosHadronLog.dat, live in a separate repository and are not available):

* EoSChiral (src/eoChiral.cpp): two uniform (e, n_b) tables of the reference's
  sizes and ranges -- small 201 x 201 over [0, 1.46] x [0, 0.3], big
  2001 x 401 over [0, 146] x [0, 6] (GeV/fm^3, 1/fm^3);
* EoSHadron (src/eoHadron.cpp): a table uniform in log(e/e0) and in the
  log-mapped n_b, n_q.

`smooth` fields are a plausible EoS (p near e/3, T near the massless gas,
mu_b growing with n_b) for hydro runs; `linear` fields are exact for the
interpolation (bilinear in (e, n_b), trilinear in the mapped coordinates), for
the pins.  `mapped_to_density` is the inverse of the hadron table's log
mapping, derived from its definition: a node at x (|x| >= dx) holds
n = sign(x) n0 exp(2|x| - lognmax), and for |x| < dx the mapping is linear,
n = x / dx * n0 exp(2 dx - lognmax).
"""
import numpy as np

from vhlle_amd import abi

CNST = (16 + 0.5 * 21.0 * 2.5) * 3.14159265358979312 ** 2 / 30.0 / 0.197326968 ** 3


class ChiralTable:
    def __init__(self, emin, emax, nmin, nmax, p, T, mub, mus):
        self.emin, self.emax, self.nmin, self.nmax = emin, emax, nmin, nmax
        self.p, self.T, self.mub, self.mus = (np.ascontiguousarray(a, dtype=np.float64)
                                              for a in (p, T, mub, mus))
        self.ne, self.nn = self.p.shape

    def nodes(self):
        e = np.linspace(self.emin, self.emax, self.ne)
        n = np.linspace(self.nmin, self.nmax, self.nn)
        return np.meshgrid(e, n, indexing="ij")


class Chiral:
    def __init__(self, small, big):
        self.small, self.big = small, big


class Hadron:
    def __init__(self, e0, n0, logemin, logemax, lognmax, p, T, mub, muq, mus, status=None):
        self.e0, self.n0, self.logemin, self.logemax, self.lognmax = (
            e0, n0, logemin, logemax, lognmax)
        self.p, self.T, self.mub, self.muq, self.mus = (
            np.ascontiguousarray(a, dtype=np.float64) for a in (p, T, mub, muq, mus))
        self.ne, self.nnb, self.nnq = self.p.shape
        self.status = np.ascontiguousarray(np.zeros(self.p.shape) if status is None else status,
                                           dtype=np.int32)

    def dx(self):
        return ((self.logemax - self.logemin) / (self.ne - 1),
                2.0 * self.lognmax / (self.nnb - 1), 2.0 * self.lognmax / (self.nnq - 1))

    def axes(self):
        """node coordinates: x_e = log(e/e0), mapped x_nb, x_nq"""
        dxe, dxb, dxq = self.dx()
        return (self.logemin + dxe * np.arange(self.ne), -self.lognmax + dxb * np.arange(self.nnb),
                -self.lognmax + dxq * np.arange(self.nnq))


def mapped_to_density(x, n0, lognmax, dx):
    x = np.asarray(x, dtype=np.float64)
    lin = x / dx * n0 * np.exp(2 * dx - lognmax)
    log = np.sign(x) * n0 * np.exp(2 * np.abs(x) - lognmax)
    return np.where(np.abs(x) >= dx, log, lin)


def _smooth_fields(e, nb, nq=0.0):
    p = e / 3.0 * (1.0 - 0.08 * np.tanh(np.abs(nb) / (e + 0.05)))
    T = (np.maximum(e, 0.0) / CNST) ** 0.25 * (1.0 - 0.05 * nb * nb / (1.0 + nb * nb))
    mub = 0.25 * nb / np.sqrt(e + 0.05)
    muq = 0.04 * nq / np.sqrt(e + 0.05)
    mus = 0.01 * nb
    return p, T, mub, muq, mus


def chiral(kind="smooth", coef=None):
    """EoSChiral tables of the reference's sizes.  kind "smooth": the
    plausible fields above; "linear": field k = a_k + b_k e + c_k nb +
    d_k e nb with coef[k] = (a, b, c, d) for k in p, T, mub, mus, the same
    in both tables unless coef is a pair (small, big)."""
    tabs = []
    for i, (emax, nmax, ne, nn) in enumerate(((1.46, 0.3, 201, 201), (146.0, 6.0, 2001, 401))):
        t = ChiralTable(0.0, emax, 0.0, nmax, *(np.zeros((ne, nn)),) * 4)
        E, N = t.nodes()
        if kind == "smooth":
            p, T, mub, _, mus = _smooth_fields(E, N)
        else:
            c = coef[i] if isinstance(coef, tuple) else coef
            p, T, mub, mus = (a + b * E + cc * N + d * E * N for (a, b, cc, d) in c)
        tabs.append(ChiralTable(0.0, emax, 0.0, nmax, p, T, mub, mus))
    return Chiral(tabs[0], tabs[1])


def hadron(kind="smooth", coef=None, ne=101, nnb=41, nnq=41, e0=1.0, n0=0.15,
           logemin=np.log(1e-3), logemax=np.log(200.0), lognmax=np.log(5.0 / 0.15)):
    """An EoSHadron table.  kind "smooth": the plausible fields at the nodes'
    (e, nb, nq); "linear": field k = a + b X + c Y + d Z + g X Y Z in the
    node coordinates X = log(e/e0), Y, Z = mapped nb, nq, coef[k] =
    (a, b, c, d, g) for k in p, T, mub, muq, mus."""
    h = Hadron(e0, n0, logemin, logemax, lognmax, *(np.zeros((ne, nnb, nnq)),) * 5)
    X, Y, Z = np.meshgrid(*h.axes(), indexing="ij")
    if kind == "smooth":
        _, dxb, dxq = h.dx()
        E = e0 * np.exp(X)
        NB = mapped_to_density(Y, n0, lognmax, dxb)
        NQ = mapped_to_density(Z, n0, lognmax, dxq)
        fields = _smooth_fields(E, NB, NQ)
    else:
        fields = [a + b * X + c * Y + d * Z + g * X * Y * Z for (a, b, c, d, g) in coef]
    return Hadron(e0, n0, logemin, logemax, lognmax, *fields)


def to_abi(t):
    """the library's descriptor object (abi.EosChiral / abi.EosHadron) of a
    table built here"""
    if isinstance(t, Chiral):
        return abi.EosChiral(*(abi.ChiralTable(x.emin, x.emax, x.nmin, x.nmax, x.p, x.T, x.mub,
                                               x.mus) for x in (t.small, t.big)))
    return abi.EosHadron(t.e0, t.n0, t.logemin, t.logemax, t.lognmax, t.p, t.T, t.mub, t.muq,
                         t.mus, t.status)


def write_chiral_files(eos_dir, t):
    """chiraleos.dat (big) and chiralsmall.dat (small) in the reference's
    format (src/eoChiral.cpp:60-72): nn blocks of ne rows of T [MeV],
    mu_q [MeV], e [e_0], p [e_0], n [n_0], s [n_0], mu_S [MeV], unused"""
    for name, x in (("chiraleos.dat", t.big), ("chiralsmall.dat", t.small)):
        E, N = x.nodes()
        cols = [x.T * 1000.0, x.mub * 1000.0 / 3.0, E / 0.146, x.p / 0.146, N / 0.15,
                np.ones_like(E), x.mus * 1000.0, np.zeros_like(E)]
        rows = np.stack([c.T.ravel() for c in cols], axis=1)  # in-major, ie-minor
        np.savetxt(f"{eos_dir}/{name}", rows, fmt="%.17g")


def write_hadron_file(path, h):
    """eosHadronLog.dat in the reference's format (src/eoHadron.cpp:19-40): a
    header ne nnb nnq e0 n0 logemin logemax lognmax, then rows e nb nq p T
    mub muq mus status in (ie, inb, inq) order"""
    X, Y, Z = np.meshgrid(*h.axes(), indexing="ij")
    _, dxb, dxq = h.dx()
    cols = [h.e0 * np.exp(X), mapped_to_density(Y, h.n0, h.lognmax, dxb),
            mapped_to_density(Z, h.n0, h.lognmax, dxq), h.p, h.T, h.mub, h.muq, h.mus]
    rows = np.column_stack([c.ravel() for c in cols] + [h.status.ravel()])
    head = f"{h.ne} {h.nnb} {h.nnq} " + " ".join(
        repr(float(v)) for v in (h.e0, h.n0, h.logemin, h.logemax, h.lognmax))
    np.savetxt(path, rows, fmt=["%.17g"] * 8 + ["%d"], header=head, comments="")
