"""The surrogate's cluster family (2x2 plaquette + 8 bath orbitals, 12
parameters) on the device: its symmetry reduction and the continued-fraction
export / import that carry the surrogate's targets and predictions.

An instance rebuilt from the five independent spin-up pairs' fractions
(exported at full length, expanded by the diagonal mirrors and, for spin down,
by spin flip x sublattice exchange) has the Green function of the exact
instance (GF_method 'F') to 1e-8 on a Matsubara grid, in both spin blocks, and
the same Potthoff functional on the lattice to 1e-9.  (Equivalent fractions
computed apart agree level by level only at the start: without
reorthogonalisation the Lanczos coefficients of two runs drift apart deep in
the fraction, while the functions they represent stay equal.)"""
import numpy as np
import pytest

from pyqcm_amd import _lib, qcm, vca_surrogate as V
from pyqcm_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    _lib.require_device()
    qcm.set_global_parameterC("Hamiltonian_format", "G")
    yield
    qcm.set_global_parameterC("GF_method", "L")


def _point():
    p = V.sample(3, seed=11)[2].copy()
    p[4] = 0.12   # a clear Neel field: the two spin blocks differ
    return p


def test_symmetry_reduction_and_fraction_import():
    p = _point()
    qcm.set_global_parameterC("GF_method", "F")
    exact = V.new_instance(p, 7101)
    z = W.matsubara(16, 20.0)
    G_up = qcm.Green_function_grid(z, 0, exact)
    G_dn = qcm.Green_function_grid(z, 1, exact)
    assert np.max(np.abs(G_up - G_dn)) > 1e-3          # the Weiss field splits the spins
    # the five classes, as long as the instance's own (max_iter_CF = 400 levels)
    y, ln = V.exact_targets(exact, ncf=400)
    E0, sec = qcm.ground_state_solve(exact)
    assert sec.startswith(V.SECTOR) and abs(y[0] - E0) < 1e-12 * abs(E0)
    rebuilt = V.set_predicted(7102, p, y)
    R_up = qcm.Green_function_grid(z, 0, rebuilt)
    R_dn = qcm.Green_function_grid(z, 1, rebuilt)
    assert np.max(np.abs(R_up - G_up)) < 1e-8, np.max(np.abs(R_up - G_up))
    assert np.max(np.abs(R_dn - G_dn)) < 1e-8, np.max(np.abs(R_dn - G_dn))
    # the same functional on the lattice
    grid = V.OmegaGrid(nk=16, w=(2.0, 20.0, 16, 16, 16))
    om_exact = grid.omega(exact, p)
    om_rebuilt = grid.omega(rebuilt, p)
    assert abs(om_exact - om_rebuilt) < 1e-9, (om_exact, om_rebuilt)


def test_exported_fractions_match_the_full_set():
    """ed_continued_fractions (all ten pairs) and the pair list agree level by level."""
    import ctypes as C
    p = _point()
    qcm.set_global_parameterC("GF_method", "F")
    lab = V.new_instance(p, 7103)
    qcm.Green_function_solveC(lab)
    cap = 400   # max_iter_CF: the same stop as the instance's own fractions
    nf = C.c_int()
    a = np.zeros((20, cap))
    b = np.zeros((20, cap))
    ln = np.zeros(20, dtype=np.int32)
    _lib.check(_lib.lib().ed_continued_fractions(lab, 0, 0, cap, a.ctypes.data_as(C.POINTER(C.c_double)),
                                                 b.ctypes.data_as(C.POINTER(C.c_double)),
                                                 ln.ctypes.data_as(C.POINTER(C.c_int)), C.byref(nf)))
    assert nf.value == 20
    y, l5 = V.exact_targets(lab, ncf=cap)
    _, a5, b5 = V.unpack(y)
    for k, (o1, o2) in enumerate(V.PAIRS):
        q = o1 * (o1 + 1) // 2 + o2
        for eh in range(2):
            n = ln[2 * q + eh]
            assert l5[2 * k + eh] == n
            assert np.allclose(a5[2 * k + eh, :n], a[2 * q + eh, :n], rtol=0, atol=1e-10)
            assert np.allclose(b5[2 * k + eh, :n], b[2 * q + eh, :n], rtol=0, atol=1e-10)
    # mirror classes: (3,3) equals (0,0), (2,0) and (3,1) equal (1,0), (2,2)
    # equals (1,1): level by level over the first 12 levels (two runs on
    # equivalent start vectors drift apart beyond, ~1e-7 at level 15, by loss
    # of orthogonality; the functions stay equal, test above)
    for (q1, q2) in ((0, 9), (1, 3), (2, 5), (1, 7)):
        for eh in range(2):
            n = min(12, ln[2 * q1 + eh], ln[2 * q2 + eh])
            assert np.allclose(a[2 * q1 + eh, :n], a[2 * q2 + eh, :n], rtol=0, atol=1e-8)
            assert np.allclose(b[2 * q1 + eh, :n], b[2 * q2 + eh, :n], rtol=0, atol=1e-8)
