"""CPU-side checks of the C-ABI library: it loads, exports every symbol the
header declares, and the host-only logic (bases, sectors, registries, error
paths) behaves like the reference.  No GPU work is launched here."""
import os
import re

import numpy as np
import pytest

from pyqcm_amd import _lib, qcm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_symbols():
    txt = open(os.path.join(ROOT, "include", "pyqcm_ed.h")).read()
    return sorted(set(re.findall(r"\b(ed_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    syms = _header_symbols()
    assert len(syms) >= 40
    for s in syms:
        assert hasattr(L, s), s
    assert sorted(_lib.SIGNATURES) == syms


def test_half_basis_host_matches_oracle(oracle):
    for L, N in [(8, 4), (12, 6), (12, 7), (5, 0), (5, 5), (16, 8)]:
        assert np.array_equal(qcm.half_basis(L, N), oracle.half_basis(L, N))


def test_sector_parse_matches_reference_rules(oracle):
    for s in ["R0:N8:S0", "R1:N7:S-1", "R0:N7", "R0:S1", "R0", "R3:N12:S2"]:
        assert qcm.sector_parse(s) == oracle.sector_parse(s)


def test_error_paths_raise():
    qcm.new_model("capi_m", 2, 0)
    with pytest.raises(_lib.EDError):
        qcm.new_operator("capi_m", "bad", "one-body", [(2, 1, 1.0)])   # r > c
    with pytest.raises(_lib.EDError):
        qcm.new_operator("capi_m", "bad0", "one-body", [(0, 1, 1.0)])  # 0-based index
    with pytest.raises(_lib.EDError) as e:
        qcm.new_model_instanceC("no_such_model", {"t": 1.0}, "R0:N2:S0", 5)
    assert e.value.code == 7
    qcm.new_operator("capi_m", "t", "one-body", [(1, 2, -1.0), (3, 4, -1.0)])
    with pytest.raises(_lib.EDError) as e:
        qcm.new_model_instanceC("capi_m", {"t": 1.0}, "R0:N9:S0", 5)   # N > 2L
    assert e.value.code == 2
    with pytest.raises(_lib.EDError):
        qcm.new_model_instanceC("capi_m", {"t": 1.0}, "R0:N2:S1", 5)   # parity
    with pytest.raises(_lib.EDError):
        qcm.set_global_parameterC("no_such_parameter", 1)
    with pytest.raises(_lib.EDError):
        qcm.new_model("capi_big", 30, 4)                                # > 32 orbitals


def test_global_parameters_roundtrip():
    qcm.set_global_parameterC("accur_lanczos", 1e-11)
    import ctypes as C
    v = C.c_double()
    _lib.check(_lib.lib().ed_get_global_double(b"accur_lanczos", C.byref(v)))
    assert v.value == 1e-11
    qcm.set_global_parameterC("accur_lanczos", 1e-12)
    n = C.c_long()
    _lib.check(_lib.lib().ed_get_global_int(b"max_iter_lanczos", C.byref(n)))
    assert n.value == 1000


def _tri_dense(d, e):
    return np.diag(d) + np.diag(e, 1) + np.diag(e, -1)


def _tri_cases():
    rng = np.random.default_rng(11)
    cases = []
    for n in (1, 2, 3, 7, 50, 200, 400):
        cases.append((rng.normal(size=n), rng.uniform(0.1, 1.0, size=n - 1)))
    # decoupled blocks (zero couplings), lowest block in the middle
    d = rng.normal(size=30)
    e = rng.uniform(0.2, 1.0, size=29)
    e[[9, 19]] = 0.0
    d[10:20] -= 5.0
    cases.append((d, e))
    # tiny couplings: nearly degenerate lowest pair, separated from the rest
    d = np.concatenate([[-3.0, -3.0 + 1e-9], rng.uniform(0, 1, size=20)])
    e = np.concatenate([[1e-12], rng.uniform(0.1, 0.5, size=20)])
    cases.append((d, e))
    # Lanczos-like tridiagonal with large entries (ghost-free part of a chain spectrum)
    n = 120
    cases.append((np.full(n, -4.0) + 0.01 * rng.normal(size=n), 2.0 + 0.01 * rng.normal(size=n - 1)))
    return cases


@pytest.mark.parametrize("seed", ["none", "near", "second", "far"])
@pytest.mark.parametrize("case", range(10))
def test_tridiag_lowest_matches_dense_eigh(case, seed):
    """Lowest Ritz pair of the Lanczos driver (bisection, or Rayleigh-quotient
    iteration from a guess, + inverse iteration) against numpy.linalg.eigh of
    the dense tridiagonal matrix; a guess at the second eigenvalue or far away
    must still give the lowest one."""
    d, e = _tri_cases()[case]
    T = _tri_dense(d, e)
    w, V = np.linalg.eigh(T)
    guess = {"none": None, "near": w[0] + 1e-3 * (1 + abs(w[0])),
             "second": w[min(1, len(w) - 1)], "far": w[-1] + 1.0}[seed]
    ev, z = qcm.tridiag_lowest(d, e, guess)
    norm = np.abs(T).sum(axis=1).max()
    assert abs(ev - w[0]) <= 1e-13 * max(norm, 1.0)
    assert abs(np.linalg.norm(z) - 1.0) <= 1e-12
    assert np.linalg.norm(T @ z - ev * z) <= 1e-11 * max(norm, 1.0)
    if len(w) == 1 or w[1] - w[0] > 1e-6:
        assert abs(abs(z @ V[:, 0]) - 1.0) <= 1e-10
        # the component the Lanczos stop rule uses, |z_last|
        assert abs(abs(z[-1]) - abs(V[-1, 0])) <= 1e-12


def test_hdf5_library_found_at_run_time():
    """the solver opens the HDF5 C library at run time (solved instances to
    and from files); where no library exists the entry points report code 8"""
    from pyqcm_amd import qcm
    from pyqcm_amd._lib import EDError
    try:
        v = qcm.hdf5_version()
    except EDError as e:
        assert e.code == 8
        pytest.skip("no HDF5 C library on this machine")
    major, minor, _ = (int(x) for x in v.split("."))
    assert (major, minor) >= (1, 10)
