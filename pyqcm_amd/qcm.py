"""Python mirror of the reference's ED bindings (module ``qcm``,
src_python/qcm_ED_wrap.hpp:25-382), backed by the MI355X C-ABI.

Function names, argument meaning, index conventions (1-based matrix
elements from Python, 0-based in C) and error behaviour (an exception on any
failure; qcm.error in the reference, EDError here) follow the reference, so
code written against ``qcm`` ED functions runs unchanged:

    new_model, new_operator, new_model_instanceC, ground_state_solve,
    Green_function_solveC, Green_function, Green_function_dimensionC,
    self_energyC, hybridization_functionC, cluster_averages, qmatrix,
    mixingC, model_sizeC, set_global_parameterC, erase_model_instance,
    Green_function_average, Green_function_density, susceptibility,
    susceptibility_poles (and tr_sigma_inf, a C++-level function there)

Vectorized variants (``Green_function_grid``) take whole frequency grids, so
the device kernels evaluate every frequency in one launch.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import ED_ERR_INVALID, EDError, check, lib

__all__ = ["EDError", "new_model", "new_operator", "new_model_instanceC", "new_model_instance",
           "ground_state_solve", "Green_function_solveC", "Green_function", "Green_function_grid",
           "Green_function_dimensionC", "self_energyC", "hybridization_functionC", "cluster_averages",
           "qmatrix", "mixingC", "model_sizeC", "set_global_parameterC", "set_global_parameter",
           "erase_model_instance", "Hamiltonian", "half_basis", "sector_parse", "states", "solve_batch", "Green_function_batch",
           "Green_function_average", "Green_function_density", "tr_sigma_inf", "susceptibility", "susceptibility_poles",
           "interrupt", "write_instance_to_hdf5", "read_instance_from_hdf5", "hdf5_version", "ED_complex_HS",
           "hopping_matrix", "interactions", "hybridization", "matrix_elements", "fidelity", "density_matrix",
           "print_wavefunction", "parametersC", "new_operator_complex"]

_models = {}      # name -> (n_sites, n_bath, [operator names])
_instances = {}   # label -> (model name, {name: value})

_BOOL, _INT, _DOUBLE, _CHAR = "b", "i", "d", "c"
# global parameter types (src_util/global_parameter.cpp:44-90)
_PARAM_TYPES = {
    "check_lanczos_residual": _BOOL, "no_degenerate_BL": _BOOL, "nosym": _BOOL, "continued_fraction": _BOOL,
    "verb_ED": _BOOL, "verb_Hilbert": _BOOL, "verb_warning": _BOOL, "parallel_sectors": _BOOL,
    "merge_states": _BOOL, "combine_mcf": _BOOL, "block_Lanczos_QR": _BOOL, "store_lanczos_basis": _BOOL,
    "sell16": _BOOL, "lanczos_reorth": _BOOL, "host_random_start": _BOOL,
    "poll_reductions": _BOOL, "lanczos_graph": _BOOL,
    "accur_band_lanczos": _DOUBLE, "accur_continued_fraction": _DOUBLE, "accur_Davidson": _DOUBLE,
    "accur_deflation": _DOUBLE, "accur_lanczos": _DOUBLE, "accur_Q_matrix": _DOUBLE,
    "band_lanczos_minimum_gap": _DOUBLE, "minimum_weight": _DOUBLE, "Qmatrix_wtol": _DOUBLE,
    "Qmatrix_vtol": _DOUBLE, "temperature": _DOUBLE, "lanczos_basis_budget_GB": _DOUBLE,
    "n_states": _INT, "max_dim_full": _INT, "max_dim_print": _INT, "max_iter_BL": _INT, "max_iter_CF": _INT,
    "max_iter_lanczos": _INT, "seed": _INT, "GF_lookup_depth": _INT, "hv_variant": _INT,
    "sell_sort_window": _INT, "hv_block": _INT, "hv_row_order": _INT, "vector_block": _INT, "reduction_blocks_per_cu": _INT,
    "lanczos_block": _INT, "band_lanczos_block": _INT, "gf_streams": _INT,
    "Hamiltonian_format": _CHAR, "Ground_state_method": _CHAR, "GF_method": _CHAR,
}


def set_global_parameterC(name, value=None):
    """pyqcm.set_global_parameter (pyqcm/__init__.py:2259-2273): None sets a boolean to True."""
    t = _PARAM_TYPES.get(name)
    if t is None:
        raise EDError(1, "unknown global parameter " + name)
    L = lib()
    if value is None:
        if t != _BOOL:
            raise EDError(1, "global parameter %s needs a value" % name)
        check(L.ed_set_global_bool(name.encode(), 1))
    elif t == _BOOL:
        check(L.ed_set_global_bool(name.encode(), int(bool(value))))
    elif t == _INT:
        check(L.ed_set_global_int(name.encode(), int(value)))
    elif t == _DOUBLE:
        check(L.ed_set_global_double(name.encode(), float(value)))
    else:
        check(L.ed_set_global_char(name.encode(), str(value)[0].encode()))


set_global_parameter = set_global_parameterC


def new_model(name, n_sites, n_bath, generators=None, bath_irrep=False):
    """qcm.new_model (src_python/qcm_ED_wrap.hpp:145-163): generators are lists of
    n_sites + n_bath 1-based orbital images (with bath_irrep, the bath entries
    are phases and kept as they are), converted to 0-based as the binding does."""
    n_orb = int(n_sites) + int(n_bath)
    gens = [] if generators is None else [list(g) for g in generators]
    flat = []
    for j, g in enumerate(gens):
        if len(g) != n_orb:
            raise EDError(1, "generator %d should have %d elements" % (j + 1, n_orb))
        flat += [int(v) - 1 for v in g[:n_sites]]
        flat += [int(v) if bath_irrep else int(v) - 1 for v in g[n_sites:]]
    arr = (C.c_int * max(len(flat), 1))(*flat)
    check(lib().ed_new_model(name.encode(), int(n_sites), int(n_bath), arr, len(gens), int(bool(bath_irrep))))
    _models.setdefault(name, (int(n_sites), int(n_bath), []))


def new_operator(model, op, type, elements):
    """qcm.new_operator (src_python/qcm_ED_wrap.hpp:174-217): elements are 1-based (r, c, v)
    with r <= c; they are decremented before reaching ED::new_operator."""
    rr, cc, vv = [], [], []
    fac = 0.5 if type == "anomalous" else 1.0
    for i, e in enumerate(elements):
        if len(e) != 3:
            raise EDError(1, "element %d of argument 4 of 'new_operator' should be a 3-tuple" % i)
        r, c, v = int(e[0]), int(e[1]), float(e[2])
        if r > c and type != "general_interaction":
            raise EDError(1, "the first index of element %d of argument 4 of 'new_operator' cannot be bigger "
                             "than the second index" % i)
        if r == 0 or c == 0:
            raise EDError(1, "indices in matrix elements of operators are labelled starting at 1, not at 0.")
        rr.append(r - 1)
        cc.append(c - 1)
        vv.append(fac * v)
    n = len(rr)
    ra = (C.c_int * max(n, 1))(*rr)
    ca = (C.c_int * max(n, 1))(*cc)
    va = (C.c_double * max(n, 1))(*vv)
    check(lib().ed_new_operator(model.encode(), op.encode(), type.encode(), n, ra, ca, va))
    if model in _models and op not in _models[model][2]:
        _models[model][2].append(op)


def new_model_instanceC(name, values, sectors, label=0):
    """qcm.new_model_instanceC (src_python/qcm_ED_wrap.hpp:165-172)."""
    names = list(values.keys())
    vals = [float(values[k]) for k in names]
    na = (C.c_char_p * max(len(names), 1))(*[k.encode() for k in names])
    va = (C.c_double * max(len(vals), 1))(*vals)
    check(lib().ed_new_model_instance(name.encode(), len(names), na, va, sectors.encode(), int(label)))
    _instances[int(label)] = (name, dict(values))


new_model_instance = new_model_instanceC


def erase_model_instance(label=0):
    check(lib().ed_erase_model_instance(int(label)))
    _instances.pop(int(label), None)


def ground_state_solve(label=0):
    """qcm.ground_state_solve: returns (energy, sector string)."""
    E0 = C.c_double()
    buf = C.create_string_buffer(1024)
    check(lib().ed_ground_state_solve(int(label), C.byref(E0), buf, 1024))
    return E0.value, buf.value.decode()


def solve_batch(labels, green_function=True):
    """Ground states (and, with green_function, the Green functions) of many
    instances at once.  Instances whose operators factorize by spin and that
    use the Lanczos ground state run batched on the device, one workgroup per
    (instance, sector), and with GF_method 'F' one workgroup per continued
    fraction; the others are solved one after the other.  Afterwards every
    instance answers the usual queries (ground_state_solve, Green_function...)."""
    labels = [int(x) for x in labels]
    arr = (C.c_int * max(len(labels), 1))(*labels)
    check(lib().ed_solve_batch(arr, len(labels), 1 if green_function else 0))


def Green_function_solveC(label=0):
    check(lib().ed_green_function_solve(int(label)))


def Green_function_dimensionC(label=0):
    d = C.c_int()
    check(lib().ed_green_function_dimension(int(label), C.byref(d)))
    return d.value


def _grid(z):
    z = np.atleast_1d(np.asarray(z, dtype=np.complex128))
    return np.ascontiguousarray(z), z.ctypes.data_as(C.POINTER(C.c_double))


def _gf_call(fn, label, z, spin_down, *extra):
    zz, zp = _grid(z)
    d = Green_function_dimensionC(label)
    out = np.zeros(len(zz) * d * d, dtype=np.complex128)
    check(fn(int(label), len(zz), zp, int(spin_down), *extra, out.ctypes.data_as(C.POINTER(C.c_double))))
    # each block is column-major d x d, exposed with the reference's flat copy
    # into a (d, d) C-ordered array (src_python/qcm_nb.hpp:27-41)
    return out.reshape(len(zz), d, d)


def Green_function_grid(z, spin_down=0, label=0, blocks=0):
    """G(z) for a whole frequency grid: array (nz, d, d)."""
    return _gf_call(lib().ed_green_function, label, z, spin_down, int(blocks))


def Green_function_batch(z, labels, spin_down=0):
    """G(z) of many instances on one grid, evaluated together: list of arrays (nz, d, d)."""
    zz, zp = _grid(z)
    labels = [int(x) for x in labels]
    dims = [Green_function_dimensionC(l) for l in labels]
    out = np.zeros(len(zz) * sum(d * d for d in dims), dtype=np.complex128)
    arr = (C.c_int * max(len(labels), 1))(*labels)
    check(lib().ed_green_function_batch(arr, len(labels), len(zz), zp, int(spin_down),
                                        out.ctypes.data_as(C.POINTER(C.c_double))))
    res, at = [], 0
    for d in dims:
        res.append(out[at:at + len(zz) * d * d].reshape(len(zz), d, d))
        at += len(zz) * d * d
    return res


def Green_function(z, spin_down=0, label=0, blocks=0):
    """qcm.Green_function (src_python/qcm_ED_wrap.hpp:65-72): d x d matrix at one z."""
    return Green_function_grid([z], spin_down, label, blocks)[0]


def self_energyC(z, spin_down=0, label=0):
    return _gf_call(lib().ed_self_energy, label, [z], spin_down)[0]


def hybridization_functionC(z, spin_down=0, label=0):
    return _gf_call(lib().ed_hybridization_function, label, [z], spin_down)[0]


def _superlattice_struct(sl):
    """ctypes ed_superlattice from a dict with V0 ((d, d) complex, V0[a, b] the
    (a, b) element), elements (row, col, neighbor, complex value) and R
    ((nnbr, 3) neighbor vectors); the arrays are kept on the struct."""
    V0 = np.asarray(sl["V0"], dtype=np.complex128)
    d = V0.shape[0]
    els = list(sl.get("elements", []))
    R = np.ascontiguousarray(np.asarray(sl.get("R", np.zeros((0, 3))), dtype=np.float64).reshape(-1, 3))
    v0 = np.ascontiguousarray(V0.T).reshape(-1)  # column-major
    row = np.array([e[0] for e in els] or [0], dtype=np.int32)
    col = np.array([e[1] for e in els] or [0], dtype=np.int32)
    nbr = np.array([e[2] for e in els] or [0], dtype=np.int32)
    val = np.array([complex(e[3]) for e in els] or [0j], dtype=np.complex128)
    ip = C.POINTER(C.c_int)
    dp = C.POINTER(C.c_double)
    st = _lib.Superlattice(d, v0.ctypes.data_as(dp), len(els), row.ctypes.data_as(ip), col.ctypes.data_as(ip),
                           nbr.ctypes.data_as(ip), val.ctypes.data_as(dp), R.shape[0], R.ctypes.data_as(dp))
    st._keep = (v0, row, col, nbr, val, R)
    return st


def _kgrid_struct(kgrid):
    """ctypes ed_kgrid from (dim, (nkx, nky, nkz), offset)."""
    dim, nk, off = kgrid
    nk = list(nk) + [1] * (3 - len(nk))
    return _lib.KGridC(int(dim), (C.c_int * 3)(*[int(x) for x in nk[:3]]), float(off))


def projected_Green_function(z, superlattice, kgrid, spin_down=0, label=0):
    """lattice_model_instance::projected_Green_function (src_qcm/CPT.cpp:283-343)
    for a lattice tiled by one cluster: (1/nk) sum_k G (1 - V(k) G)^{-1} at every
    frequency of z; array (nz, d, d) laid out as Green_function_grid."""
    zz, zp = _grid(z)
    d = Green_function_dimensionC(label)
    out = np.zeros(len(zz) * d * d, dtype=np.complex128)
    st, kg = _superlattice_struct(superlattice), _kgrid_struct(kgrid)
    check(lib().ed_cpt_projected_gf(int(label), len(zz), zp, int(spin_down), C.byref(st), C.byref(kg),
                                    out.ctypes.data_as(C.POINTER(C.c_double))))
    return out.reshape(len(zz), d, d)


def SEF_grid(w, weights, superlattice, kgrid, superlattice_down=None, label=0):
    """The (frequency, k) grid integral of lattice_model_instance::Potthoff_functional
    (src_qcm/lattice_model_instance.cpp:503-584): sum_n weights[n] (1/nk) sum_k
    log|det(V(k) G(i w_n) - 1)|, spin blocks summed (SEF_integrand_k, :624-650).
    Returns (integral, per-frequency k-averages)."""
    w = np.ascontiguousarray(np.asarray(w, dtype=np.float64))
    p = np.ascontiguousarray(np.asarray(weights, dtype=np.float64))
    per = np.zeros(len(w))
    I = C.c_double()
    up = _superlattice_struct(superlattice)
    dn = _superlattice_struct(superlattice_down) if superlattice_down is not None else None
    kg = _kgrid_struct(kgrid)
    dp = C.POINTER(C.c_double)
    check(lib().ed_sef_grid(int(label), len(w), w.ctypes.data_as(dp), p.ctypes.data_as(dp), C.byref(up),
                            C.byref(dn) if dn is not None else None, C.byref(kg), C.byref(I),
                            per.ctypes.data_as(dp)))
    return I.value, per


def Potthoff_functionalC(label=0):
    """ED::Potthoff_functional (src_ed/qcm_ED.cpp:330-334): E0 plus the bath term."""
    x = C.c_double()
    check(lib().ed_potthoff_functional(int(label), C.byref(x)))
    return x.value


def parametersC(label=0):
    """qcm.parametersC (src_python/qcm_ED_wrap.hpp:263-273): (dict of the instance's
    parameters, model name)."""
    n = C.c_int()
    model = C.create_string_buffer(256)
    check(lib().ed_instance_parameters(int(label), 0, 0, None, None, C.byref(n), model, 256))
    cap, w = max(n.value, 1), 128
    names = C.create_string_buffer(cap * w)
    vals = np.zeros(cap)
    check(lib().ed_instance_parameters(int(label), n.value, w, names, vals.ctypes.data_as(C.POINTER(C.c_double)),
                                       C.byref(n), model, 256))
    raw = names.raw
    d = {raw[k * w:(k + 1) * w].split(b"\0", 1)[0].decode(): float(vals[k]) for k in range(n.value)}
    return d, model.value.decode()


def new_operator_complex(model, op, type, elements):
    """qcm.new_operator_complex (src_python/qcm_ED_wrap.hpp:219-261): elements
    are 1-based (r, c, v) with r <= c and v complex; the lower triangle holds
    their conjugates.  An instance with a nonzero value on the operator works
    in a complex Hilbert space."""
    rr, cc, vv = [], [], []
    fac = 0.5 if type == "anomalous" else 1.0
    for i, e in enumerate(elements):
        if len(e) != 3:
            raise EDError(1, "element %d of argument 4 of 'new_operator' should be a 3-tuple" % i)
        r, c, v = int(e[0]), int(e[1]), complex(e[2])
        if r > c:
            raise EDError(1, "the first index of element %d of argument 4 of 'new_operator' cannot be bigger "
                             "than the second index" % i)
        if r == 0 or c == 0:
            raise EDError(1, "indices in matrix elements of operators are labelled starting at 1, not at 0.")
        rr.append(r - 1)
        cc.append(c - 1)
        vv.extend([fac * v.real, fac * v.imag])
    n = len(rr)
    ra = (C.c_int * max(n, 1))(*rr)
    ca = (C.c_int * max(n, 1))(*cc)
    va = (C.c_double * max(2 * n, 1))(*vv)
    check(lib().ed_new_operator_complex(model.encode(), op.encode(), type.encode(), n, ra, ca, va))
    if model in _models and op not in _models[model][2]:
        _models[model][2].append(op)


def ED_complex_HS(label=0):
    """qcm.ED_complex_HS (src_python/qcm_ED_wrap.hpp:27-30): 1 in a complex
    Hilbert space, else 0."""
    x = C.c_int()
    check(lib().ed_complex_hs(int(label), C.byref(x)))
    return x.value


def _model_sizes(label):
    name = _instances.get(int(label), (None,))[0]
    if name is None or name not in _models:
        raise EDError(7, "unknown instance %d" % label)
    return _models[name][0], _models[name][1]


def hopping_matrix(spin_down=0, diag=0, label=0, full=0):
    """qcm.hopping_matrix (src_python/qcm_ED_wrap.hpp:94-106): the one-body matrix of
    the cluster (full=0) or of cluster and bath (full=1; diag=1: bath diagonalized),
    with the reference's flat copy of the column-major matrix."""
    ns, nb = _model_sizes(label)
    d = ns + nb if full else ns
    out = np.zeros(d * d, dtype=np.complex128)
    check(lib().ed_hopping_matrix(int(label), int(spin_down), int(diag), int(full),
                                  out.ctypes.data_as(C.POINTER(C.c_double))))
    return out.reshape(d, d)


def interactions(label=0):
    """qcm.interactions (src_python/qcm_ED_wrap.hpp:108-111): [(r, c, U v)]"""
    n = C.c_int()
    check(lib().ed_interactions(int(label), 0, None, None, None, C.byref(n)))
    m = max(n.value, 1)
    r, c, v = np.zeros(m, np.int32), np.zeros(m, np.int32), np.zeros(m)
    check(lib().ed_interactions(int(label), n.value, r.ctypes.data_as(C.POINTER(C.c_int)),
                                c.ctypes.data_as(C.POINTER(C.c_int)), v.ctypes.data_as(C.POINTER(C.c_double)),
                                C.byref(n)))
    return [(int(r[k]), int(c[k]), float(v[k])) for k in range(n.value)]


def hybridization(label=0):
    """qcm.hybridization (src_python/qcm_ED_wrap.hpp:351-362): (bath energies,
    cluster-bath matrix) with the reference's (L, M) flat copy of the column-major
    n_sites x n_bath block."""
    ns, nb = _model_sizes(label)
    e = np.zeros(max(nb, 1))
    t = np.zeros(max(ns * nb, 1), dtype=np.complex128)
    check(lib().ed_hybridization_lehmann(int(label), 0, e.ctypes.data_as(C.POINTER(C.c_double)),
                                         t.ctypes.data_as(C.POINTER(C.c_double))))
    return e[:nb], t[:ns * nb].reshape(ns, nb)


def matrix_elements(model, op):
    """qcm.matrix_elements (src_python/qcm_ED_wrap.hpp:121-130): (type, [(r, c, v)])"""
    n = C.c_int()
    tb = C.create_string_buffer(32)
    check(lib().ed_matrix_elements(model.encode(), op.encode(), tb, 32, 0, None, None, None, C.byref(n)))
    m = max(n.value, 1)
    r, c, v = np.zeros(m, np.int32), np.zeros(m, np.int32), np.zeros(m, np.complex128)
    check(lib().ed_matrix_elements(model.encode(), op.encode(), tb, 32, n.value, r.ctypes.data_as(C.POINTER(C.c_int)),
                                   c.ctypes.data_as(C.POINTER(C.c_int)), v.ctypes.data_as(C.POINTER(C.c_double)),
                                   C.byref(n)))
    return tb.value.decode(), [(int(r[k]), int(c[k]), complex(v[k])) for k in range(n.value)]


def density_matrix(sites, label=0):
    """qcm.density_matrix (src_python/qcm_ED_wrap.hpp:32-43): (rho, basis) for the
    orbitals `sites` (0-based) in the ground state; rho with the reference's flat
    copy of the column-major matrix, basis as uint64 words."""
    s = (C.c_int * max(len(sites), 1))(*[int(x) for x in sites])
    d = C.c_int()
    check(lib().ed_density_matrix(int(label), s, len(sites), 0, None, None, C.byref(d)))
    n = d.value
    rho = np.zeros(max(n * n, 1), dtype=np.complex128)
    basis = np.zeros(max(n, 1), dtype=np.uint64)
    check(lib().ed_density_matrix(int(label), s, len(sites), n, rho.ctypes.data_as(C.POINTER(C.c_double)),
                                  basis.ctypes.data_as(C.POINTER(C.c_uint64)), C.byref(d)))
    return rho[:n * n].reshape(n, n), basis[:n]


def print_wavefunction(label=0):
    """qcm.print_wavefunction (src_python/qcm_ED_wrap.hpp:335-337): the
    ground-state wavefunction(s) as a string."""
    n = C.c_int()
    check(lib().ed_print_wavefunction(int(label), None, 0, C.byref(n)))
    buf = C.create_string_buffer(n.value + 1)
    check(lib().ed_print_wavefunction(int(label), buf, n.value + 1, C.byref(n)))
    return buf.value.decode()


def fidelity(label1, label2):
    """qcm.fidelity (src_python/qcm_ED_wrap.hpp:378-381)"""
    x = C.c_double()
    check(lib().ed_fidelity(int(label1), int(label2), C.byref(x)))
    return x.value


def write_instance_to_hdf5(filename, group, label=0):
    """qcm.write_instance_to_hdf5 (src_python/qcm_ED_wrap.hpp:364-369): the solved
    instance into a group of an HDF5 file, in the reference's layout."""
    check(lib().ed_write_instance_hdf5(str(filename).encode(), int(label), str(group).encode()))


def read_instance_from_hdf5(filename, group, label=0):
    """qcm.read_instance_from_hdf5 (src_python/qcm_ED_wrap.hpp:371-376): the
    instance takes the states and Green functions stored in the group."""
    check(lib().ed_read_instance_hdf5(str(filename).encode(), int(label), str(group).encode()))


def hdf5_version():
    """version of the HDF5 C library the solver found (EDError code 8 if none)"""
    buf = C.create_string_buffer(64)
    check(lib().ed_hdf5_version(buf, 64))
    return buf.value.decode()


def Green_function_average(label=0, spin_down=0):
    """qcm.Green_function_average (src_python/qcm_ED_wrap.hpp:44-50): the
    frequency-integrated Green function, the matrix of <c^dagger_i c_j>, as a
    (d, d) complex array (the reference's flat copy of the column-major block)."""
    d = Green_function_dimensionC(label)
    out = np.zeros(d * d, dtype=np.complex128)
    check(lib().ed_green_function_average(int(label), int(spin_down), out.ctypes.data_as(C.POINTER(C.c_double))))
    return out.reshape(d, d)


def Green_function_density(label=0):
    """qcm.Green_function_density (src_python/qcm_ED_wrap.hpp:52-54): electrons per site."""
    x = C.c_double()
    check(lib().ed_green_function_density(int(label), C.byref(x)))
    return x.value


def tr_sigma_inf(label=0):
    """ED::tr_sigma_inf (src_ed/qcm_ED.cpp:345-348): trace of the self-energy at
    infinite frequency, both spins."""
    x = C.c_double()
    check(lib().ed_tr_sigma_inf(int(label), C.byref(x)))
    return x.value


def susceptibility(op, w, label=0):
    """qcm.susceptibility (src_python/qcm_ED_wrap.hpp:327-333): the dynamic
    susceptibility of operator op at the complex frequencies w."""
    ww, wp = _grid(w)
    out = np.zeros(len(ww), dtype=np.complex128)
    check(lib().ed_susceptibility(int(label), str(op).encode(), len(ww), wp, out.ctypes.data_as(C.POINTER(C.c_double))))
    return out


def susceptibility_poles(op, label=0):
    """qcm.susceptibility_poles (src_python/qcm_ED_wrap.hpp:318-325): array (n, 2)
    of (pole, residue)."""
    n = C.c_int()
    check(lib().ed_susceptibility_poles(int(label), str(op).encode(), 0, None, C.byref(n)))
    out = np.zeros(2 * max(n.value, 1))
    check(lib().ed_susceptibility_poles(int(label), str(op).encode(), n.value,
                                        out.ctypes.data_as(C.POINTER(C.c_double)), C.byref(n)))
    return out[:2 * n.value].reshape(n.value, 2)


def cluster_averages(label=0):
    """qcm.cluster_averages: dict name -> (average, variance)."""
    cap = 256
    ave = np.zeros(cap)
    var = np.zeros(cap)
    n = C.c_int()
    check(lib().ed_cluster_averages(int(label), cap, ave.ctypes.data_as(C.POINTER(C.c_double)),
                                    var.ctypes.data_as(C.POINTER(C.c_double)), C.byref(n)))
    names = list(parametersC(label)[0].keys())
    return {names[k] if k < len(names) else str(k): (ave[k], var[k]) for k in range(n.value)}


def qmatrix(label=0, spin_down=0):
    """Lehmann representation (poles e[M], residues Q[L, M]) of the first state
    (ED::qmatrix, src_ed/qcm_ED.cpp:406-416): G(r, c) = sum_i Q[c, i]
    conj(Q[r, i]) / (z - e[i]).  Q is real in a real Hilbert space and complex
    in a complex one."""
    L_, M_ = C.c_int(), C.c_int()
    check(lib().ed_qmatrix(int(label), int(spin_down), 0, None, None, C.byref(L_), C.byref(M_)))
    e = np.zeros(max(M_.value, 1))
    v = np.zeros(max(M_.value * L_.value, 1), dtype=np.complex128)
    check(lib().ed_qmatrix(int(label), int(spin_down), M_.value, e.ctypes.data_as(C.POINTER(C.c_double)),
                           v.ctypes.data_as(C.POINTER(C.c_double)), C.byref(L_), C.byref(M_)))
    Q = v[:M_.value * L_.value].reshape(M_.value, L_.value).T.copy()
    if not ED_complex_HS(label):
        Q = Q.real.copy()
    return e[:M_.value], Q


def combined_mcf(label=0, spin_down=0):
    """(W, [A_j], [B_j]) of the combined matrix continued fraction (pyqcm's
    combined_mcf, src_python/qcm_wrap.hpp:502-520; needs GF_method 'M' and
    combine_mcf = True); W is p x d."""
    p, nfl = C.c_int(), C.c_int()
    check(lib().ed_combined_mcf(int(label), int(spin_down), 0, None, None, None, C.byref(p), C.byref(nfl)))
    d = Green_function_dimensionC(label)
    W = np.zeros(p.value * d)
    A = np.zeros(nfl.value * p.value * p.value)
    B = np.zeros(nfl.value * p.value * p.value)
    check(lib().ed_combined_mcf(int(label), int(spin_down), nfl.value, _dptr(W), _dptr(A), _dptr(B), C.byref(p),
                                C.byref(nfl)))
    P = p.value
    W = W.reshape(d, P).T
    A = [A[j * P * P:(j + 1) * P * P].reshape(P, P).T for j in range(nfl.value)]
    B = [B[j * P * P:(j + 1) * P * P].reshape(P, P).T for j in range(nfl.value)]
    return W, A, B


def continued_fractions(label=0, spin_down=0, state=0):
    """continued fractions of low-energy state `state` of a solved instance
    (GF_method 'F'; continued_fraction_set, src_ed/continued_fraction_set.cpp:12-98):
    (a, b, lengths), fraction f = 2 q + eh for the orbital pair q = o1 (o1 + 1) / 2 + o2
    (o1 >= o2), eh 0 electron / 1 hole; a, b are (nfrac, max length) arrays as
    evaluated (a_n shifted by the state energy, b_n squared, b_0 the weighted
    norm), zero past each fraction's length."""
    args = (int(label), int(spin_down), int(state))
    nfr = C.c_int()
    check(lib().ed_continued_fractions(*args, 0, None, None, None, C.byref(nfr)))
    nf = nfr.value
    ln = np.zeros(nf, dtype=np.int32)
    one = np.zeros(1)
    check(lib().ed_continued_fractions(*args, 0, _dptr(one), _dptr(one), ln.ctypes.data_as(C.POINTER(C.c_int)),
                                       C.byref(nfr)))
    cap = max(1, int(ln.max(initial=0)))
    a = np.zeros(nf * cap)
    b = np.zeros(nf * cap)
    check(lib().ed_continued_fractions(*args, cap, _dptr(a), _dptr(b), ln.ctypes.data_as(C.POINTER(C.c_int)),
                                       C.byref(nfr)))
    return a.reshape(nf, cap), b.reshape(nf, cap), ln


def continued_fractions_pairs(label, pairs, ncf, spin_down=0):
    """the first ncf levels of the fractions of the orbital pairs (o1, o2)
    (0-based) from the ground state of instance `label`: (a, b, lengths) with
    fraction 2 k + eh for pair k (layout of continued_fractions)."""
    n = len(pairs)
    o1 = (C.c_int * n)(*[int(x) for (x, y) in pairs])
    o2 = (C.c_int * n)(*[int(y) for (x, y) in pairs])
    a = np.zeros((2 * n, ncf))
    b = np.zeros((2 * n, ncf))
    ln = np.zeros(2 * n, dtype=np.int32)
    check(lib().ed_continued_fractions_pairs(int(label), int(spin_down), n, o1, o2, int(ncf), _dptr(a), _dptr(b),
                                             ln.ctypes.data_as(C.POINTER(C.c_int))))
    return a, b, ln


def set_continued_fractions(label, E0, sector, spin_down, a, b, lengths):
    """instance `label` becomes solved with one state of energy E0 in `sector`
    whose spin block holds these fractions (layout of continued_fractions);
    call once per spin block."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    ln = np.ascontiguousarray(lengths, dtype=np.int32)
    check(lib().ed_set_continued_fractions(int(label), float(E0), sector.encode(), int(spin_down), a.shape[0],
                                           a.shape[1], _dptr(a), _dptr(b), ln.ctypes.data_as(C.POINTER(C.c_int))))


def states(label=0):
    cap = 64
    e = np.zeros(cap)
    w = np.zeros(cap)
    n = C.c_int()
    check(lib().ed_states(int(label), cap, e.ctypes.data_as(C.POINTER(C.c_double)),
                          w.ctypes.data_as(C.POINTER(C.c_double)), C.byref(n)))
    return e[:n.value], w[:n.value]


def interrupt():
    """Stop the running (or the next) solve at its next Lanczos / band-Lanczos /
    Davidson step or Green-function start vector: that call raises EDError
    with code 11.  The reference polls Python signals there (check_signals,
    src_python/common_Py.cpp:29-33); here any thread, or a C signal handler
    through ed_request_cancel, makes the request."""
    check(lib().ed_request_cancel())


def mixingC(label=0):
    m = C.c_int()
    check(lib().ed_mixing(int(label), C.byref(m)))
    return m.value


def model_sizeC(name):
    return _models[name][0], _models[name][1]


def half_basis(L, N):
    dim = C.c_int64()
    check(lib().ed_half_basis(L, N, None, 0, C.byref(dim)))
    out = np.zeros(max(dim.value, 1), dtype=np.uint32)
    check(lib().ed_half_basis(L, N, out.ctypes.data_as(C.POINTER(C.c_uint32)), dim.value, C.byref(dim)))
    return out[:dim.value]


def sector_parse(s):
    N, S, R = C.c_int(), C.c_int(), C.c_int()
    check(lib().ed_sector_parse(s.encode(), C.byref(N), C.byref(S), C.byref(R)))
    return N.value, S.value, R.value


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Hamiltonian:
    """Inner plug-in: a sector Hamiltonian of an instance (Hamiltonian<T>,
    src_ed/Hamiltonian/Hamiltonian_base.hpp:43-82) living on the device."""

    def __init__(self, label, sector, format="G"):
        h = C.c_int64()
        check(lib().ed_hamiltonian_create(int(label), sector.encode(), format.encode(), C.byref(h)))
        self.h = h.value
        d = C.c_int64()
        check(lib().ed_hamiltonian_dim(self.h, C.byref(d)))
        self.dim = d.value

    def __del__(self):
        try:
            lib().ed_hamiltonian_destroy(self.h)
        except Exception:
            pass

    def mult_add(self, x, y):
        """y += H x (host arrays); returns y."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        check(lib().ed_hamiltonian_mult_add(self.h, _dptr(x), _dptr(y)))
        return y

    def apply(self, x):
        return self.mult_add(x, np.zeros(self.dim))

    def apply_device(self, dx_ptr, dy_ptr, beta=0.0):
        check(lib().ed_hamiltonian_apply_device(self.h, C.c_void_p(dx_ptr), C.c_void_p(dy_ptr), float(beta)))

    def apply_dot_noalloc(self, dx_ptr, dy_ptr, ddot_ptr, beta=0.0):
        """y = H x + beta y and dot = <x, y> on device pointers; allocates nothing
        (the partials use the library's fixed device scratch)."""
        check(lib().ed_hamiltonian_apply_dot_noalloc(self.h, C.c_void_p(dx_ptr), C.c_void_p(dy_ptr), float(beta),
                                                     C.c_void_p(ddot_ptr)))

    def diag(self):
        d = np.zeros(self.dim)
        check(lib().ed_hamiltonian_diag(self.h, _dptr(d)))
        return d

    def csr_pattern(self):
        nnz = C.c_int64(0)
        check(lib().ed_hamiltonian_csr_pattern(self.h, C.byref(nnz), None, None, None))
        Iptr = np.zeros(self.dim + 1, dtype=np.int32)
        J = np.zeros(max(nnz.value, 1), dtype=np.int32)
        v = np.zeros(max(nnz.value, 1))
        check(lib().ed_hamiltonian_csr_pattern(self.h, C.byref(nnz), Iptr.ctypes.data_as(C.POINTER(C.c_int32)),
                                               J.ctypes.data_as(C.POINTER(C.c_int32)), _dptr(v)))
        return Iptr, J[:nnz.value], v[:nnz.value]

    def time(self, nrep=10, variant=0):
        ms = C.c_double()
        check(lib().ed_hamiltonian_time(self.h, int(nrep), C.byref(ms), int(variant)))
        return ms.value

    def autotune(self):
        """Time the candidate schedules; returns (ms per candidate, index chosen)."""
        ms = np.zeros(16)
        n, best = C.c_int(), C.c_int()
        check(lib().ed_hamiltonian_autotune(self.h, 16, _dptr(ms), C.byref(n), C.byref(best)))
        return ms[:n.value], best.value

    def set_variant(self, variant):
        check(lib().ed_hamiltonian_set_variant(self.h, int(variant)))

    def lanczos_gs(self, start):
        start = np.ascontiguousarray(start, dtype=np.float64)
        start = start / np.linalg.norm(start)
        gs = np.zeros(self.dim)
        E0 = C.c_double()
        nit = C.c_int()
        check(lib().ed_lanczos_gs(self.h, _dptr(start), C.byref(E0), _dptr(gs), C.byref(nit)))
        return E0.value, gs, nit.value

    def lanczos_green(self, psi, cap=2000):
        psi = np.ascontiguousarray(psi, dtype=np.float64)
        while True:
            a = np.zeros(cap)
            b = np.zeros(cap + 1)
            n = C.c_int()
            rc = lib().ed_lanczos_green(self.h, _dptr(psi), cap, _dptr(a), _dptr(b), C.byref(n))
            if rc == ED_ERR_INVALID and n.value > cap:
                cap = n.value  # the entry point reports the size it needs
                continue
            check(rc)
            return a[:n.value], b[:n.value + 1]

    def band_lanczos(self, phi, cap=4096):
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        p = phi.shape[0]
        while True:
            e = np.zeros(cap)
            Q = np.zeros(cap * p)
            M = C.c_int()
            rc = lib().ed_band_lanczos(self.h, p, _dptr(phi), cap, _dptr(e), _dptr(Q), C.byref(M))
            if rc == ED_ERR_INVALID and M.value > cap:
                cap = M.value
                continue
            check(rc)
            m = M.value
            return e[:m], Q[:p * m].reshape(m, p).T.copy()

    def davidson(self, K):
        ev = np.zeros(K)
        vec = np.zeros(K * self.dim)
        nit = C.c_int()
        check(lib().ed_davidson(self.h, int(K), _dptr(ev), _dptr(vec), C.byref(nit)))
        return ev, vec.reshape(K, self.dim), nit.value

    def block_lanczos(self, q0, M0, cap=2000):
        q0 = np.ascontiguousarray(q0, dtype=np.float64)
        p = q0.shape[0]
        while True:
            A = np.zeros(cap * p * p)
            B = np.zeros(cap * p * p)
            n = C.c_int()
            rc = lib().ed_block_lanczos(self.h, p, _dptr(q0), int(M0), cap, _dptr(A), _dptr(B), C.byref(n))
            if rc == ED_ERR_INVALID and n.value > cap:
                cap = n.value
                continue
            check(rc)
            break
        k = n.value
        return (A[:k * p * p].reshape(k, p, p).transpose(0, 2, 1).copy(),
                B[:k * p * p].reshape(k, p, p).transpose(0, 2, 1).copy())


def create_or_destroy(L, pm, a, Nup, Ndw, x, z=1.0):
    """y = z c_a x (pm=-1) or z c^dagger_a x (pm=+1) between factorized sectors (device)."""
    spin = 1 if a >= L else 0
    tNup, tNdw = Nup + (pm if spin == 0 else 0), Ndw + (pm if spin == 1 else 0)
    dim = len(half_basis(L, tNup)) * len(half_basis(L, tNdw))
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros(dim)
    check(lib().ed_create_or_destroy(L, pm, a, Nup, Ndw, _dptr(x), float(z), _dptr(y)))
    return y


def tridiag_lowest(d, e, guess=None):
    """Lowest eigenvalue and normalized eigenvector of the symmetric tridiagonal
    matrix with diagonal d and off-diagonal e (the Ritz pair of the Lanczos
    driver); guess, e.g. the previous Ritz value, seeds a Rayleigh-quotient
    iteration."""
    d = np.ascontiguousarray(d, dtype=np.float64)
    e = np.ascontiguousarray(e, dtype=np.float64)
    if len(e) != max(len(d) - 1, 0):
        raise EDError(1, "tridiag_lowest: e must have len(d) - 1 entries")
    ev = C.c_double()
    z = np.zeros(len(d))
    g = float("nan") if guess is None else float(guess)
    check(lib().ed_tridiag_lowest(len(d), _dptr(d), _dptr(e) if len(e) else _dptr(np.zeros(1)), g, C.byref(ev),
                                  _dptr(z)))
    return ev.value, z


def cf_eval(fractions, z):
    """Evaluate continued fractions [(a, b), ...] (already transformed) on grid z -> (nz, ncf)."""
    off = [0]
    A, B = [], []
    for (a, b) in fractions:
        A.extend(list(a))
        B.extend(list(b)[:len(a)])
        off.append(len(A))
    A = np.array(A + [0.0])
    B = np.array(B + [0.0])
    offa = (C.c_int * len(off))(*off)
    zz, zp = _grid(z)
    out = np.zeros(len(zz) * len(fractions), dtype=np.complex128)
    check(lib().ed_cf_eval(len(fractions), offa, _dptr(A), _dptr(B), len(zz), zp, _dptr(out.view(np.float64))))
    return out.reshape(len(zz), len(fractions))


def vdvh(e, V, z):
    e = np.ascontiguousarray(e, dtype=np.float64)
    L, M = V.shape
    Vf = np.asfortranarray(V, dtype=np.float64)
    zz, zp = _grid(z)
    G = np.zeros(len(zz) * L * L, dtype=np.complex128)
    check(lib().ed_vdvh(L, M, _dptr(e), Vf.ctypes.data_as(C.POINTER(C.c_double)), len(zz), zp,
                        _dptr(G.view(np.float64))))
    return G.reshape(len(zz), L, L).transpose(0, 2, 1)


def mcf_eval(A, B, W, z):
    A = np.asarray(A, dtype=np.float64)
    B = np.asarray(B, dtype=np.float64)
    nfl, p, _ = A.shape
    q = W.shape[1]
    Af = np.ascontiguousarray(A.transpose(0, 2, 1))
    Bf = np.ascontiguousarray(B.transpose(0, 2, 1))
    Wf = np.asfortranarray(W, dtype=np.float64)
    zz, zp = _grid(z)
    G = np.zeros(len(zz) * q * q, dtype=np.complex128)
    check(lib().ed_mcf_eval(p, nfl, _dptr(Af), _dptr(Bf), Wf.ctypes.data_as(C.POINTER(C.c_double)), q, len(zz),
                            zp, _dptr(G.view(np.float64))))
    return G.reshape(len(zz), q, q).transpose(0, 2, 1)
