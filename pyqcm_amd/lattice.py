"""Lattice-level quantities of a lattice tiled by one cluster, computed from the
cluster's Green function on the device: the CPT Green function averaged over
the superlattice Brillouin zone and the self-energy functional (Potthoff
functional) on a discrete (frequency, wavevector) grid.

A superlattice spec is a dict:
  V0        (d, d) complex: the k-independent part of V = t(k) - t' (the
            lattice one-body matrix inside the cluster minus the cluster's own,
            e.g. minus a Weiss field that exists only on the cluster),
  elements  list of (row, col, neighbor index, complex value): the hoppings to
            other clusters, V(k)[row, col] += value exp(2 pi i k . R[neighbor]),
  R         (nnbr, 3) superlattice neighbor vectors.
k is in reduced coordinates of the superlattice.  A wavevector grid is
(dim, (nkx, nky, nkz), offset): the points (i + offset) / nk per axis.
"""
import numpy as np

from . import qcm


def legendre_frequency_grid(w1, w2, n1, n2, n3):
    """pyqcm.legendre_frequency_grid (pyqcm/__init__.py:3040-3066): Gauss-Legendre
    points in [0, w1] (n1), [w1, w2] (n2) and, in the variable 1/w, [w2, inf)
    (n3); weights include the 1/pi of the imaginary-axis integral."""
    w = np.zeros(n1 + n2 + n3)
    p = np.zeros(n1 + n2 + n3)
    x, a = np.polynomial.legendre.leggauss(n1)
    w[:n1] = 0.5 * w1 * (x + 1.0)
    p[:n1] = 0.5 * w1 * a
    x, a = np.polynomial.legendre.leggauss(n2)
    w[n1:n1 + n2] = w1 + 0.5 * (w2 - w1) * (x + 1.0)
    p[n1:n1 + n2] = 0.5 * (w2 - w1) * a
    # u = 1 / w on (0, 1 / w2]: nodes in decreasing u so that w increases
    x, a = np.polynomial.legendre.leggauss(n3)
    u = 0.5 * (x[::-1] + 1.0) / w2
    w[n1 + n2:] = 1.0 / u
    p[n1 + n2:] = 0.5 / w2 * a[::-1] * w[n1 + n2:] ** 2
    return w, p / np.pi


def regular_frequency_grid(wc, n1, n2):
    """pyqcm.regular_frequency_grid (pyqcm/__init__.py:3069-3089): midpoint rule
    with n1 points on [0, wc] and n2 points uniform in 1/w on (0, 1/wc];
    weights include 1/pi."""
    w = np.zeros(n1 + n2)
    p = np.zeros(n1 + n2)
    dw = wc / n1
    w[:n1] = (np.arange(n1) + 0.5) * dw
    p[:n1] = dw / np.pi
    du = 1.0 / (n2 * wc)
    u = (np.arange(n2) + 0.5) * du
    w[n1:] = 1.0 / u[::-1]
    p[n1:] = w[n1:] ** 2 * du / np.pi
    return w, p


def potthoff_functional(label, w, weights, superlattice, kgrid, superlattice_down=None):
    """lattice_model_instance::Potthoff_functional, discrete-grid path
    (src_qcm/lattice_model_instance.cpp:503-584), for a lattice tiled by one
    cluster without Nambu doubling:
        Omega = ( -I - Omega_diag + Omega_cluster ) / n_sites,
    I the grid integral of log|det(V G - 1)| (ed_sef_grid), Omega_diag =
    -1/2 sum of the diagonals of V0 (both spin blocks; twice the up block
    without spin mixing) and Omega_cluster = ED::Potthoff_functional."""
    I, _ = qcm.SEF_grid(w, weights, superlattice, kgrid, superlattice_down, label)
    two = qcm.mixingC(label) & 4
    diag = -np.sum(np.real(np.diag(np.asarray(superlattice["V0"]))))
    if two:
        diag -= np.sum(np.real(np.diag(np.asarray(superlattice_down["V0"]))))
    else:
        diag *= 2.0
    diag *= 0.5
    n_sites = qcm.Green_function_dimensionC(label)
    return (-I - diag + qcm.Potthoff_functionalC(label)) / n_sites


def cluster_superlattice(n_sites, bonds, t, weiss=None):
    """Superlattice spec of a lattice tiled by one cluster of n_sites sites.
    bonds: inter-cluster bonds (i, j, R): site i of the cell at 0 to site j of
    the cell at R (R a 3-vector in superlattice units), hopping -t each way.
    weiss: optional diagonal of a field present on the cluster only (V0 gets
    minus it)."""
    V0 = np.zeros((n_sites, n_sites), dtype=np.complex128)
    if weiss is not None:
        V0 -= np.diag(np.asarray(weiss, dtype=np.float64))
    vecs, els = [], []

    def index(R):
        R = tuple(float(x) for x in R)
        if R not in vecs:
            vecs.append(R)
        return vecs.index(R)

    for i, j, R in bonds:
        R = np.asarray(list(R) + [0.0] * (3 - len(R)), dtype=np.float64)
        # H = -t (c_i^dagger c_j(R) + h.c.): V[i, j] ~ exp(2 pi i k.R), V[j, i] ~ exp(-2 pi i k.R)
        els.append((i, j, index(R), -t))
        els.append((j, i, index(-R), -t))
    return {"V0": V0, "elements": els, "R": np.array(vecs if vecs else np.zeros((0, 3)))}


def chain_superlattice(L, t=1.0, weiss=None):
    """1D chain tiled by open clusters of L sites: bond (L-1, 0) to the next cell."""
    return cluster_superlattice(L, [(L - 1, 0, (1.0,))], t, weiss)


def square_superlattice(nx, ny, t=1.0, weiss=None):
    """Square lattice tiled by open nx x ny clusters (site x + nx y)."""
    bonds = []
    for y in range(ny):
        bonds.append((nx - 1 + nx * y, nx * y, (1.0, 0.0)))
    for x in range(nx):
        bonds.append((x + nx * (ny - 1), x, (0.0, 1.0)))
    return cluster_superlattice(nx * ny, bonds, t, weiss)
