"""ctypes binding of libpyqcm_ed.so (the C-ABI declared in include/pyqcm_ed.h).

The library is built in-tree (pyqcm_amd/libpyqcm_ed.so) by ``pyqcm_amd.build``;
importing a compute entry point without it raises immediately -- there is no
CPU fallback in the product path.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpyqcm_ed.so")

ED_OK = 0
ED_ERR_INVALID = 1
ED_ERR_CANCELLED = 11
_STATUS = {1: "invalid argument", 2: "invalid sector", 3: "not converged", 4: "Q-matrix norm violated",
           5: "out of memory", 6: "HIP error", 7: "not found", 8: "unsupported", 9: "internal error",
           10: "degenerate band Lanczos gap", 11: "cancelled"}


class EDError(RuntimeError):
    """Raised for a non-zero status of the C-ABI (qcm.error in the reference)."""

    def __init__(self, code, msg):
        super().__init__("[%s] %s" % (_STATUS.get(code, code), msg))
        self.code = code


_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_i64p = C.POINTER(C.c_int64)
_u32p = C.POINTER(C.c_uint32)
_i32p = C.POINTER(C.c_int32)


class Superlattice(C.Structure):
    """ed_superlattice: one spin block of V(k) = V0 + sum_e val_e exp(2 pi i k.R[nbr_e])."""
    _fields_ = [("d", C.c_int), ("V0", _dp), ("ne", C.c_int), ("row", _ip), ("col", _ip), ("nbr", _ip),
                ("val", _dp), ("nnbr", C.c_int), ("R", _dp)]


class KGridC(C.Structure):
    """ed_kgrid: dim axes of nk[a] points at (i + offset) / nk[a]."""
    _fields_ = [("dim", C.c_int), ("nk", C.c_int * 3), ("offset", C.c_double)]


_slp = C.POINTER(Superlattice)
_kgp = C.POINTER(KGridC)


SIGNATURES = {
    "ed_last_error": (C.c_char_p, []),
    "ed_request_cancel": (C.c_int, []),
    "ed_version": (C.c_int, [C.c_char_p, C.c_int]),
    "ed_device_count": (C.c_int, [_ip]),
    "ed_set_global_bool": (C.c_int, [C.c_char_p, C.c_int]),
    "ed_set_global_int": (C.c_int, [C.c_char_p, C.c_long]),
    "ed_set_global_double": (C.c_int, [C.c_char_p, C.c_double]),
    "ed_set_global_char": (C.c_int, [C.c_char_p, C.c_char]),
    "ed_get_global_double": (C.c_int, [C.c_char_p, _dp]),
    "ed_get_global_int": (C.c_int, [C.c_char_p, C.POINTER(C.c_long)]),
    "ed_new_model": (C.c_int, [C.c_char_p, C.c_int, C.c_int, _ip, C.c_int, C.c_int]),
    "ed_new_operator": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, _ip, _ip, _dp]),
    "ed_new_operator_complex": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, _ip, _ip, _dp]),
    "ed_new_model_instance": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_char_p), _dp, C.c_char_p, C.c_int]),
    "ed_erase_model_instance": (C.c_int, [C.c_int]),
    "ed_ground_state_solve": (C.c_int, [C.c_int, _dp, C.c_char_p, C.c_int]),
    "ed_ground_state_energy": (C.c_int, [C.c_int, _dp]),
    "ed_green_function_solve": (C.c_int, [C.c_int]),
    "ed_green_function_dimension": (C.c_int, [C.c_int, _ip]),
    "ed_green_function": (C.c_int, [C.c_int, C.c_int, _dp, C.c_int, C.c_int, _dp]),
    "ed_self_energy": (C.c_int, [C.c_int, C.c_int, _dp, C.c_int, _dp]),
    "ed_hybridization_function": (C.c_int, [C.c_int, C.c_int, _dp, C.c_int, _dp]),
    "ed_cluster_averages": (C.c_int, [C.c_int, C.c_int, _dp, _dp, _ip]),
    "ed_qmatrix": (C.c_int, [C.c_int, C.c_int, C.c_int, _dp, _dp, _ip, _ip]),
    "ed_states": (C.c_int, [C.c_int, C.c_int, _dp, _dp, _ip]),
    "ed_mixing": (C.c_int, [C.c_int, _ip]),
    "ed_solve_batch": (C.c_int, [_ip, C.c_int, C.c_int]),
    "ed_green_function_batch": (C.c_int, [_ip, C.c_int, C.c_int, _dp, C.c_int, _dp]),
    "ed_cpt_projected_gf": (C.c_int, [C.c_int, C.c_int, _dp, C.c_int, _slp, _kgp, _dp]),
    "ed_sef_grid": (C.c_int, [C.c_int, C.c_int, _dp, _dp, _slp, _slp, _kgp, _dp, _dp]),
    "ed_potthoff_functional": (C.c_int, [C.c_int, _dp]),
    "ed_complex_hs": (C.c_int, [C.c_int, _ip]),
    "ed_instance_parameters": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_char_p, _dp, _ip, C.c_char_p, C.c_int]),
    "ed_hopping_matrix": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, _dp]),
    "ed_interactions": (C.c_int, [C.c_int, C.c_int, _ip, _ip, _dp, _ip]),
    "ed_hybridization_lehmann": (C.c_int, [C.c_int, C.c_int, _dp, _dp]),
    "ed_matrix_elements": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, C.c_int, _ip, _ip, _dp, _ip]),
    "ed_fidelity": (C.c_int, [C.c_int, C.c_int, _dp]),
    "ed_print_wavefunction": (C.c_int, [C.c_int, C.c_char_p, C.c_int, _ip]),
    "ed_density_matrix": (C.c_int, [C.c_int, _ip, C.c_int, C.c_int, _dp, C.POINTER(C.c_uint64), _ip]),
    "ed_write_instance_hdf5": (C.c_int, [C.c_char_p, C.c_int, C.c_char_p]),
    "ed_read_instance_hdf5": (C.c_int, [C.c_char_p, C.c_int, C.c_char_p]),
    "ed_hdf5_version": (C.c_int, [C.c_char_p, C.c_int]),
    "ed_green_function_average": (C.c_int, [C.c_int, C.c_int, _dp]),
    "ed_green_function_density": (C.c_int, [C.c_int, _dp]),
    "ed_tr_sigma_inf": (C.c_int, [C.c_int, _dp]),
    "ed_susceptibility": (C.c_int, [C.c_int, C.c_char_p, C.c_int, _dp, _dp]),
    "ed_susceptibility_poles": (C.c_int, [C.c_int, C.c_char_p, C.c_int, _dp, _ip]),
    "ed_hamiltonian_create": (C.c_int, [C.c_int, C.c_char_p, C.c_char, _i64p]),
    "ed_hamiltonian_destroy": (C.c_int, [C.c_int64]),
    "ed_hamiltonian_dim": (C.c_int, [C.c_int64, _i64p]),
    "ed_hamiltonian_mult_add": (C.c_int, [C.c_int64, _dp, _dp]),
    "ed_hamiltonian_apply_device": (C.c_int, [C.c_int64, C.c_void_p, C.c_void_p, C.c_double]),
    "ed_hamiltonian_apply_dot_noalloc": (C.c_int, [C.c_int64, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p]),
    "ed_hamiltonian_diag": (C.c_int, [C.c_int64, _dp]),
    "ed_hamiltonian_csr_pattern": (C.c_int, [C.c_int64, _i64p, _i32p, _i32p, _dp]),
    "ed_hamiltonian_time": (C.c_int, [C.c_int64, C.c_int, _dp, C.c_int]),
    "ed_hamiltonian_set_variant": (C.c_int, [C.c_int64, C.c_int]),
    "ed_hamiltonian_autotune": (C.c_int, [C.c_int64, C.c_int, _dp, _ip, _ip]),
    "ed_half_basis": (C.c_int, [C.c_int, C.c_int, _u32p, C.c_int64, _i64p]),
    "ed_random_start": (C.c_int, [C.c_int64, C.c_int64, _dp]),
    "ed_sector_parse": (C.c_int, [C.c_char_p, _ip, _ip, _ip]),
    "ed_lanczos_gs": (C.c_int, [C.c_int64, _dp, _dp, _dp, _ip]),
    "ed_lanczos_green": (C.c_int, [C.c_int64, _dp, C.c_int, _dp, _dp, _ip]),
    "ed_band_lanczos": (C.c_int, [C.c_int64, C.c_int, _dp, C.c_int, _dp, _dp, _ip]),
    "ed_davidson": (C.c_int, [C.c_int64, C.c_int, _dp, _dp, _ip]),
    "ed_block_lanczos": (C.c_int, [C.c_int64, C.c_int, _dp, C.c_int, C.c_int, _dp, _dp, _ip]),
    "ed_create_or_destroy": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp, C.c_double, _dp]),
    "ed_tridiag_lowest": (C.c_int, [C.c_int, _dp, _dp, C.c_double, _dp, _dp]),
    "ed_cf_eval": (C.c_int, [C.c_int, _ip, _dp, _dp, C.c_int, _dp, _dp]),
    "ed_vdvh": (C.c_int, [C.c_int, C.c_int, _dp, _dp, C.c_int, _dp, _dp]),
    "ed_mcf_eval": (C.c_int, [C.c_int, C.c_int, _dp, _dp, _dp, C.c_int, C.c_int, _dp, _dp]),
    "ed_continued_fractions": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, _ip, _ip]),
    "ed_continued_fractions_pairs": (C.c_int, [C.c_int, C.c_int, C.c_int, _ip, _ip, C.c_int, _dp, _dp, _ip]),
    "ed_set_continued_fractions": (C.c_int, [C.c_int, C.c_double, C.c_char_p, C.c_int, C.c_int, C.c_int, _dp, _dp,
                                             _ip]),
    "ed_combined_mcf": (C.c_int, [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _ip, _ip]),
    "ed_surrogate_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int64, _i64p]),
    "ed_surrogate_destroy": (C.c_int, [C.c_int64]),
    "ed_surrogate_n_params": (C.c_int, [C.c_int64, _i64p]),
    "ed_surrogate_set_scaling": (C.c_int, [C.c_int64, _dp, _dp, _dp, _dp, _dp]),
    "ed_surrogate_set_optimizer": (C.c_int, [C.c_int64, C.c_double, C.c_double, C.c_double, C.c_double]),
    "ed_surrogate_get_weights": (C.c_int, [C.c_int64, _dp]),
    "ed_surrogate_set_weights": (C.c_int, [C.c_int64, _dp]),
    "ed_surrogate_predict": (C.c_int, [C.c_int64, C.c_int, _dp, _dp]),
    "ed_surrogate_loss_grad": (C.c_int, [C.c_int64, C.c_int, _dp, _dp, _dp, _dp]),
    "ed_surrogate_train": (C.c_int, [C.c_int64, C.c_int, _dp, _dp, C.c_int, C.c_int, C.c_int64, _dp, _dp]),
    "ed_surrogate_time": (C.c_int, [C.c_int64, C.c_int, C.c_int, _dp, _dp]),
    "ed_device_alloc": (C.c_int, [C.c_int64, C.POINTER(C.c_void_p)]),
    "ed_device_free": (C.c_int, [C.c_void_p, C.c_int64]),
    "ed_device_copy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int]),
    "ed_device_sync": (C.c_int, []),
    "ed_release_cache": (C.c_int, []),
}


def _declare(L):
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def lib():
    """Load libpyqcm_ed.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libpyqcm_ed.so not built (run __graft_entry__.build() or make -C pyqcm_amd/csrc)")
        L = C.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def check(rc):
    if rc != ED_OK:
        raise EDError(rc, lib().ed_last_error().decode(errors="replace"))


def device_count():
    n = C.c_int()
    check(lib().ed_device_count(C.byref(n)))
    return n.value


def require_device():
    """Compute entry points run on the GPU only: fail loudly without one."""
    if device_count() < 1:
        raise RuntimeError("no HIP device visible: the ED solver runs on MI355X (gfx950) only")
