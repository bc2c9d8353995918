"""Cluster-model definitions of the ED workloads used by the tests and the bench.

A model is ``Workload(name, n_sites, n_bath, operators, values, sector, gf_orbitals)``
where ``operators`` maps an operator name to ``(type, elements)`` with 0-based
spin-orbital indices exactly as ``ED::new_operator`` receives them
(src_ed/qcm_ED.cpp:74-90; the Python binding subtracts 1 from the 1-based
indices users type, src_python/qcm_ED_wrap.hpp:200-204).  Spin-orbital a < L is
spin up, a >= L spin down, L = n_sites + n_bath.

Operator conventions (the lattice layer of the reference builds the same):
  * hopping ``t`` on bond (i, j): elements (i, j, -1) per spin, so H gets -t(c+c + h.c.);
  * ``U``: interaction elements (i, i+L, 1), i.e. U n_up n_dn;
  * ``mu``: one-body (a, a, -1) on the cluster orbitals (src_qcm/lattice_model.cpp:286-294).
"""
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Workload:
    name: str
    n_sites: int
    n_bath: int
    operators: dict            # name -> (type, [(r, c, v), ...])
    values: dict               # name -> coefficient
    sector: str                # target sector string
    gf_orbitals: list = field(default_factory=list)

    @property
    def L(self):
        return self.n_sites + self.n_bath

    def terms(self):
        """(type, value, elements) list in operator-definition order (oracle input)."""
        return [(typ, float(self.values.get(name, 0.0)), el)
                for name, (typ, el) in self.operators.items() if name in self.values]


def _hop(L, bonds):
    el = []
    for (i, j) in bonds:
        a, b = min(i, j), max(i, j)
        el.append((a, b, -1.0))
        el.append((a + L, b + L, -1.0))
    return el


def _hubbard_U(L, sites):
    return [(i, i + L, 1.0) for i in sites]


def _mu(L, sites):
    return [(i, i, -1.0) for i in sites] + [(i + L, i + L, -1.0) for i in sites]


def _sector(N, S=0):
    return "R0:N%d:S%d" % (N, S)


def hubbard_chain(n=8, t=1.0, U=4.0, mu=None, periodic=False, N=None):
    """Open (or periodic) n-site Hubbard chain, half filling by default."""
    mu = U / 2 if mu is None else mu
    L = n
    bonds = [(i, i + 1) for i in range(n - 1)] + ([(n - 1, 0)] if periodic and n > 2 else [])
    ops = {"t": ("one-body", _hop(L, bonds)),
           "U": ("interaction", _hubbard_U(L, range(n))),
           "mu": ("one-body", _mu(L, range(n)))}
    return Workload("chain%d" % n, n, 0, ops, {"t": t, "U": U, "mu": mu},
                    _sector(n if N is None else N), list(range(n)))


def torus(nx=4, ny=4, t=1.0, U=8.0, mu=None, N=None):
    """nx x ny periodic Hubbard cluster (4x4: 32 bonds), half filling by default."""
    mu = U / 2 if mu is None else mu
    L = nx * ny
    bonds = set()
    for y in range(ny):
        for x in range(nx):
            i = x + nx * y
            for (dx, dy) in ((1, 0), (0, 1)):
                j = (x + dx) % nx + nx * ((y + dy) % ny)
                if i != j:
                    bonds.add((min(i, j), max(i, j)))
    bonds = sorted(bonds)
    ops = {"t": ("one-body", _hop(L, bonds)),
           "U": ("interaction", _hubbard_U(L, range(L))),
           "mu": ("one-body", _mu(L, range(L)))}
    return Workload("torus%dx%d" % (nx, ny), L, 0, ops, {"t": t, "U": U, "mu": mu},
                    _sector(L if N is None else N), list(range(L)))


def plaquette_bath(seed=2025, t=1.0, tp=-0.3, U=8.0, mu=None, n_bath=8):
    """2x2 plaquette (NN t, diagonal t') + n_bath bath orbitals.

    Bath energies eps_b ~ U[-2, 2] and hybridisations V_b ~ U[0.2, 0.8] are drawn
    from numpy default_rng(seed) in that order (eps first); bath b couples to
    cluster site b mod 4.  mu = U/2 acts on the cluster sites only.
    """
    mu = U / 2 if mu is None else mu
    ns = 4
    L = ns + n_bath
    rng = np.random.default_rng(seed)
    eps = rng.uniform(-2.0, 2.0, n_bath)
    hyb = rng.uniform(0.2, 0.8, n_bath)
    ops = {"t": ("one-body", _hop(L, [(0, 1), (1, 3), (3, 2), (2, 0)])),
           "tp": ("one-body", _hop(L, [(0, 3), (1, 2)])),
           "U": ("interaction", _hubbard_U(L, range(ns))),
           "mu": ("one-body", _mu(L, range(ns)))}
    values = {"t": t, "tp": tp, "U": U, "mu": mu}
    for b in range(n_bath):
        o = ns + b
        ops["eb%d" % (b + 1)] = ("one-body", [(o, o, 1.0), (o + L, o + L, 1.0)])
        s = b % ns
        ops["tb%d" % (b + 1)] = ("one-body", [(s, o, 1.0), (s + L, o + L, 1.0)])
        values["eb%d" % (b + 1)] = float(eps[b])
        values["tb%d" % (b + 1)] = float(hyb[b])
    return Workload("plaquette_%db" % n_bath, ns, n_bath, ops, values, _sector(L), list(range(ns)))


def sformat_cluster():
    """4 correlated sites + 8 bath orbitals of tests/test_diagonalization/S_format.ref:2-20,
    with the instance values of S_format.ref:113-131 (U=14, mu=9.68, eb*, tb*)."""
    ns, nb = 4, 8
    L = ns + nb
    signs = [(1, 1, 1, 1), (1, 1, 1, 1), (1, 1, -1, -1), (1, 1, -1, -1),
             (1, -1, 1, -1), (1, -1, 1, -1), (1, -1, -1, 1), (1, -1, -1, 1)]
    ops = {"U": ("interaction", _hubbard_U(L, range(ns))),
           "mu": ("one-body", _mu(L, range(ns)))}
    for b in range(nb):
        o = ns + b
        ops["eb%d" % (b + 1)] = ("one-body", [(o, o, 1.0), (o + L, o + L, 1.0)])
        el = []
        for s in range(ns):
            el.append((s, o, float(signs[b][s])))
        for s in range(ns):
            el.append((s + L, o + L, float(signs[b][s])))
        ops["tb%d" % (b + 1)] = ("one-body", el)
    eb = [-0.006655534, -3.3502828, -4.1133163, -0.10235806, -4.113222, -0.1023468, 0.14108001, -3.7537133]
    tb = [-0.11240159, -1.1061747, 1.7476608, 0.057892245, -1.7476438, -0.057886658, 0.22749261, -2.3061557]
    values = {"U": 14.0, "mu": 9.68}
    for b in range(nb):
        values["eb%d" % (b + 1)] = eb[b]
        values["tb%d" % (b + 1)] = tb[b]
    return Workload("clus_2x2_8b", ns, nb, ops, values, "R0:N12:S0", list(range(ns)))


def afm_cluster(nx, ny, U, M, mu, t=1.0):
    """nx x ny open Hubbard cluster with a Neel Weiss field M sum_i (-1)^(x+y) (n_up - n_dn)."""
    L = nx * ny
    bonds = []
    for y in range(ny):
        for x in range(nx):
            i = x + nx * y
            if x + 1 < nx:
                bonds.append((i, i + 1))
            if y + 1 < ny:
                bonds.append((i, i + nx))
    stag = []
    for y in range(ny):
        for x in range(nx):
            i = x + nx * y
            s = 1.0 if (x + y) % 2 == 0 else -1.0
            stag.append((i, i, s))
            stag.append((i + L, i + L, -s))
    ops = {"t": ("one-body", _hop(L, bonds)),
           "U": ("interaction", _hubbard_U(L, range(L))),
           "mu": ("one-body", _mu(L, range(L))),
           "M": ("one-body", stag)}
    return Workload("afm%dx%d" % (nx, ny), L, 0, ops, {"t": t, "U": U, "mu": mu, "M": M},
                    _sector(L), list(range(L)))


def sector_numbers(sector):
    """(N, S) of a sector string like 'R0:N8:S0' (src_ed/sector.hpp:54-95)."""
    N = int(sector.split("N")[1].split(":")[0])
    S = int(sector.split("S")[1].split(":")[0])
    return N, S


def matsubara(n, beta):
    """i omega_n = i (2n+1) pi / beta, n = 0..n-1 (pyqcm/cdmft.py:841-846)."""
    return 1j * (2 * np.arange(n) + 1) * np.pi / beta


def real_axis(n=512, wmin=-8.0, wmax=8.0, eta=0.1):
    return np.linspace(wmin, wmax, n) + 1j * eta


def vca_cluster(nx=2, ny=2):
    """nx x ny open Hubbard cluster with the Neel Weiss field of an
    antiferromagnetic VCA sweep: operators t (nearest neighbours), U, mu and
    M = sum_i (-1)^(x+y) (n_i,up - n_i,down); target sectors N = n-1 (S = +-1),
    n (S = 0), n+1 (S = +-1), n = nx * ny."""
    L = nx * ny
    bonds = []
    for y in range(ny):
        for x in range(nx):
            i = x + nx * y
            if x + 1 < nx:
                bonds.append((i, i + 1))
            if y + 1 < ny:
                bonds.append((i, i + nx))
    neel = []
    for y in range(ny):
        for x in range(nx):
            i = x + nx * y
            s = 1.0 if (x + y) % 2 == 0 else -1.0
            neel.append((i, i, s))
            neel.append((i + L, i + L, -s))
    ops = {"t": ("one-body", _hop(L, bonds)),
           "U": ("interaction", _hubbard_U(L, range(L))),
           "mu": ("one-body", _mu(L, range(L))),
           "M": ("one-body", neel)}
    sectors = "/".join([_sector(L - 1, 1), _sector(L - 1, -1), _sector(L, 0), _sector(L + 1, 1),
                        _sector(L + 1, -1)])
    return Workload("vca%dx%d" % (nx, ny), L, 0, ops, {"t": 1.0, "U": 4.0, "mu": 0.0, "M": 0.1}, sectors,
                    list(range(L)))


def vca_cluster_bath(nx=2, ny=3, n_bath=6, seed=7, N=None):
    """vca_cluster(nx, ny) with n_bath bath orbitals: energies eps_b ~ U[-2, 2]
    and hybridisations V_b ~ U[0.2, 0.8] from default_rng(seed) (eps first),
    bath b coupled to cluster site b mod (nx ny); one target sector N (default
    nx ny, i.e. quarter filling of the ns + n_bath orbitals), S = 0.  The 2x3
    cluster with 6 bath orbitals at N = 6 has 220 x 220 = 48 400 states."""
    base = vca_cluster(nx, ny)
    ns = nx * ny
    L = ns + n_bath
    N = ns if N is None else N
    rng = np.random.default_rng(seed)
    eps = rng.uniform(-2.0, 2.0, n_bath)
    hyb = rng.uniform(0.2, 0.8, n_bath)

    def widen(el):  # spin-down orbitals move from index ns + i to L + i
        return [((r if r < ns else r - ns + L), (c if c < ns else c - ns + L), v) for (r, c, v) in el]
    ops = {k: (typ, widen(el)) for k, (typ, el) in base.operators.items()}
    values = dict(base.values)
    for b in range(n_bath):
        o = ns + b
        ops["eb%d" % (b + 1)] = ("one-body", [(o, o, 1.0), (o + L, o + L, 1.0)])
        s = b % ns
        ops["tb%d" % (b + 1)] = ("one-body", [(s, o, 1.0), (s + L, o + L, 1.0)])
        values["eb%d" % (b + 1)] = float(eps[b])
        values["tb%d" % (b + 1)] = float(hyb[b])
    return Workload("vca%dx%d_%db" % (nx, ny, n_bath), ns, n_bath, ops, values, _sector(N), list(range(ns)))


def vca_points(n=512, seed=4):
    """n parameter points of the sweep: U ~ U[2, 10], M ~ U[0, 0.3], mu ~ U[-1, 1]
    (numpy default_rng(seed)), t = 1."""
    rng = np.random.default_rng(seed)
    U = rng.uniform(2.0, 10.0, n)
    M = rng.uniform(0.0, 0.3, n)
    mu = rng.uniform(-1.0, 1.0, n)
    return [{"t": 1.0, "U": float(U[k]), "M": float(M[k]), "mu": float(mu[k])} for k in range(n)]


def setup(wl, label=0, sectors=None, model_name=None):
    """Register ``wl`` through the reference-style API (1-based elements) and
    create instance ``label``; returns the model name."""
    from . import qcm
    name = model_name or wl.name
    qcm.new_model(name, wl.n_sites, wl.n_bath)
    for op, (typ, el) in wl.operators.items():
        qcm.new_operator(name, op, typ, [(r + 1, c + 1, v) for (r, c, v) in el])
    qcm.new_model_instanceC(name, dict(wl.values), sectors or wl.sector, label)
    return name
