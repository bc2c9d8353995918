"""ctypes front-end of the CPU oracle (oracle_ed.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, for reference values / the CPU baseline.  The product
package (pyqcm_amd) never imports this module.

A model is given as the list of terms the reference's ``ED::new_operator``
receives (src_ed/qcm_ED.cpp:74-90): ``(type, value, [(r, c, v), ...])`` with
0-based spin-orbital indices (a < L spin up, a >= L spin down).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_ed.so")
ONE_BODY, INTERACTION = 0, 1
_TYPES = {"one-body": ONE_BODY, "interaction": INTERACTION}


def build():
    """Compile liboracle_ed.so with the recipe in oracle/Makefile."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def _load():
    if not os.path.exists(_LIB_PATH):
        build()
    return C.CDLL(_LIB_PATH)


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
        _declare(_lib)
    return _lib


class _Model(C.Structure):
    _fields_ = [("L", C.c_int), ("n_terms", C.c_int), ("type", C.POINTER(C.c_int)),
                ("value", C.POINTER(C.c_double)), ("off", C.POINTER(C.c_int)),
                ("r", C.POINTER(C.c_int)), ("c", C.POINTER(C.c_int)), ("v", C.POINTER(C.c_double))]


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_up = C.POINTER(C.c_uint32)
_i32p = C.POINTER(C.c_int32)
_szp = C.POINTER(C.c_size_t)


def _declare(L):
    P = C.POINTER(_Model)
    L.orc_half_basis.restype = C.c_size_t
    L.orc_half_basis.argtypes = [C.c_int, C.c_int, _up]
    L.orc_sector_parse.argtypes = [C.c_char_p, _ip, _ip, _ip]
    L.orc_sector_name.argtypes = [C.c_int, C.c_int, C.c_int, C.c_char_p, C.c_int]
    L.orc_half_table.argtypes = [P, C.c_int, C.c_int, C.c_int, _szp, _szp, _up, _up, _dp, _ip, _up, _dp]
    L.orc_hv_factorized.argtypes = [P, C.c_int, C.c_int, _dp, _dp]
    L.orc_hv_factorized_omp.argtypes = [P, C.c_int, C.c_int, _dp, _dp, C.c_int]
    L.orc_diag.argtypes = [P, C.c_int, C.c_int, _dp]
    L.orc_csr_build.argtypes = [P, C.c_int, C.c_int, _szp, _i32p, _i32p, _dp, _dp]
    L.orc_hv_csr.argtypes = [C.c_size_t, _i32p, _i32p, _dp, _dp, _dp, _dp]
    L.orc_hv_csr_full_omp.argtypes = [C.c_size_t, _i32p, _i32p, _dp, _dp, _dp, _dp, C.c_int]
    L.orc_hv_factorized_omp_timed.argtypes = [P, C.c_int, C.c_int, _dp, _dp, C.c_int, _dp]
    L.orc_hv_csr_full_omp_timed.argtypes = [C.c_size_t, _i32p, _i32p, _dp, _dp, _dp, _dp, C.c_int, _dp]
    L.orc_hv_otf.argtypes = [P, C.c_int, C.c_int, _dp, _dp]
    L.orc_create_or_destroy.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp, C.c_double, _dp]
    L.orc_tridiag_eig.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
    L.orc_symeig.argtypes = [C.c_int, _dp, _dp, _dp]
    L.orc_lanczos_gs.argtypes = [P, C.c_int, C.c_int, _dp, C.c_int, C.c_double, _dp, _dp, _ip, _dp, _dp]
    L.orc_lanczos_green.argtypes = [P, C.c_int, C.c_int, _dp, C.c_double, C.c_int, _dp, _dp, _ip]
    L.orc_cf_eval.argtypes = [C.c_int, _dp, _dp, C.c_int, _dp, _dp]
    L.orc_band_lanczos.argtypes = [P, C.c_int, C.c_int, C.c_int, _dp, C.c_int, C.c_int, C.c_double,
                                   C.c_double, _ip, _dp, _dp]
    L.orc_qmatrix_streamline.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_double, C.c_double, _dp, _dp]
    L.orc_qmatrix_norm_error.restype = C.c_double
    L.orc_qmatrix_norm_error.argtypes = [C.c_int, C.c_int, _dp]
    L.orc_vdvh.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_int, _dp, _dp]
    L.orc_davidson.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, _dp, _dp, _ip]
    L.orc_block_lanczos.argtypes = [P, C.c_int, C.c_int, C.c_int, _dp, C.c_int, C.c_int, C.c_double,
                                    C.c_double, _ip, _dp, _dp]
    L.orc_block_lanczos_v.argtypes = [P, C.c_int, C.c_int, C.c_int, _dp, C.c_int, C.c_int, C.c_double,
                                      C.c_double, C.c_int, _ip, _dp, _dp]
    L.orc_mcf_eval.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, C.c_int, C.c_int, _dp, _dp]
    L.orc_openmp_threads.restype = C.c_int


def _d(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def half_basis(L, N):
    n = lib().orc_half_basis(L, N, None)
    out = np.zeros(max(n, 1), dtype=np.uint32)
    lib().orc_half_basis(L, N, out.ctypes.data_as(_up))
    return out[:n]


def sector_parse(s):
    N, S, R = C.c_int(), C.c_int(), C.c_int()
    if lib().orc_sector_parse(s.encode(), C.byref(N), C.byref(S), C.byref(R)):
        raise ValueError("sector string %s does not conform to standard!" % s)
    return N.value, S.value, R.value


def sector_name(N, S, R):
    buf = C.create_string_buffer(64)
    lib().orc_sector_name(N, S, R, buf, 64)
    return buf.value.decode()


class Model:
    """Flat term list for the oracle; keeps the numpy buffers alive."""

    def __init__(self, L, terms):
        self.L = L
        self.terms = [(_TYPES.get(t, t), float(val), list(el)) for (t, val, el) in terms]
        types, values, off, rr, cc, vv = [], [], [0], [], [], []
        for (t, val, el) in self.terms:
            types.append(t)
            values.append(val)
            for (r, c, v) in el:
                rr.append(r)
                cc.append(c)
                vv.append(v)
            off.append(len(rr))
        self._types = np.array(types, dtype=np.int32)
        self._values = np.array(values, dtype=np.float64)
        self._off = np.array(off, dtype=np.int32)
        self._r = np.array(rr + [0], dtype=np.int32)
        self._c = np.array(cc + [0], dtype=np.int32)
        self._v = np.array(vv + [0.0], dtype=np.float64)
        self.s = _Model(L, len(types), self._types.ctypes.data_as(_ip), self._values.ctypes.data_as(_dp),
                        self._off.ctypes.data_as(_ip), self._r.ctypes.data_as(_ip),
                        self._c.ctypes.data_as(_ip), self._v.ctypes.data_as(_dp))

    @property
    def ref(self):
        return C.byref(self.s)

    def dims(self, Nup, Ndw):
        return len(half_basis(self.L, Nup)), len(half_basis(self.L, Ndw))

    def dim(self, Nup, Ndw):
        du, dd = self.dims(Nup, Ndw)
        return du * dd

    # ---- H.x in the three reference formats ----
    def hv_factorized(self, Nup, Ndw, x, y=None):
        x = _f64(x)
        y = np.zeros_like(x) if y is None else _f64(y).copy()
        _chk(lib().orc_hv_factorized(self.ref, Nup, Ndw, _d(x), _d(y)))
        return y

    def hv_factorized_omp(self, Nup, Ndw, x, nrep=1):
        x = _f64(x)
        y = np.zeros_like(x)
        _chk(lib().orc_hv_factorized_omp(self.ref, Nup, Ndw, _d(x), _d(y), nrep))
        return y

    def hv_factorized_omp_times(self, Nup, Ndw, x, nrep):
        """ms of each of nrep OpenMP products (tables built once, untimed)"""
        x = _f64(x)
        y = np.zeros_like(x)
        ms = np.zeros(nrep)
        _chk(lib().orc_hv_factorized_omp_timed(self.ref, Nup, Ndw, _d(x), _d(y), nrep, _d(ms)))
        return ms

    def diag(self, Nup, Ndw):
        d = np.zeros(self.dim(Nup, Ndw))
        _chk(lib().orc_diag(self.ref, Nup, Ndw, _d(d)))
        return d

    def csr(self, Nup, Ndw):
        nnz = C.c_size_t()
        _chk(lib().orc_csr_build(self.ref, Nup, Ndw, C.byref(nnz), None, None, None, None))
        dim = self.dim(Nup, Ndw)
        Iptr = np.zeros(dim + 1, dtype=np.int32)
        J = np.zeros(max(nnz.value, 1), dtype=np.int32)
        v = np.zeros(max(nnz.value, 1))
        dg = np.zeros(dim)
        _chk(lib().orc_csr_build(self.ref, Nup, Ndw, C.byref(nnz), Iptr.ctypes.data_as(_i32p),
                                 J.ctypes.data_as(_i32p), _d(v), _d(dg)))
        return Iptr, J[:nnz.value], v[:nnz.value], dg

    def hv_otf(self, Nup, Ndw, x):
        x = _f64(x)
        y = np.zeros_like(x)
        _chk(lib().orc_hv_otf(self.ref, Nup, Ndw, _d(x), _d(y)))
        return y

    def half_table(self, t, spin, N):
        npair, ndiag = C.c_size_t(), C.c_size_t()
        _chk(lib().orc_half_table(self.ref, t, spin, N, C.byref(npair), C.byref(ndiag),
                                  None, None, None, None, None, None))
        n, m = npair.value, ndiag.value
        I = np.zeros(max(n, 1), np.uint32)
        J = np.zeros(max(n, 1), np.uint32)
        v = np.zeros(max(n, 1))
        g = np.zeros(max(n, 1), np.int32)
        dI = np.zeros(max(m, 1), np.uint32)
        dv = np.zeros(max(m, 1))
        _chk(lib().orc_half_table(self.ref, t, spin, N, C.byref(npair), C.byref(ndiag),
                                  I.ctypes.data_as(_up), J.ctypes.data_as(_up), _d(v), g.ctypes.data_as(_ip),
                                  dI.ctypes.data_as(_up), _d(dv)))
        return I[:n], J[:n], v[:n], g[:n], dI[:m], dv[:m]

    # ---- Krylov drivers ----
    def lanczos_gs(self, Nup, Ndw, start, max_iter=1000, accur=1e-12):
        start = _f64(start)
        start = start / np.linalg.norm(start)
        gs = np.zeros_like(start)
        E0 = C.c_double()
        nit = C.c_int()
        al = np.zeros(max_iter + 2)
        be = np.zeros(max_iter + 2)
        _chk(lib().orc_lanczos_gs(self.ref, Nup, Ndw, _d(start), max_iter, accur, C.byref(E0), _d(gs),
                                  C.byref(nit), _d(al), _d(be)))
        k = nit.value
        return E0.value, gs, al[:k], be[:k]

    def lanczos_green(self, Nup, Ndw, psi, accur_cf=0.01, max_iter_cf=400):
        psi = _f64(psi)
        a = np.zeros(max_iter_cf + 1)
        b = np.zeros(max_iter_cf + 2)
        n = C.c_int()
        _chk(lib().orc_lanczos_green(self.ref, Nup, Ndw, _d(psi), accur_cf, max_iter_cf, _d(a), _d(b),
                                     C.byref(n)))
        return a[:n.value], b[:n.value + 1]

    def band_lanczos(self, Nup, Ndw, phi, M=-1, max_iter_bl=600, accur_deflation=1e-7, accur_bl=1e-12):
        phi = _f64(phi)
        p = phi.shape[0]
        cap = max_iter_bl + 2 * p + 8
        ev = np.zeros(cap)
        Q = np.zeros(p * cap)
        Mo = C.c_int()
        _chk(lib().orc_band_lanczos(self.ref, Nup, Ndw, p, _d(phi), M, max_iter_bl, accur_deflation, accur_bl,
                                    C.byref(Mo), _d(ev), _d(Q)))
        Mv = Mo.value
        return ev[:Mv], Q[:p * Mv].reshape(Mv, p).T.copy()

    def davidson(self, Nup, Ndw, K, accuracy=1e-5, max_iter=1000):
        dim = self.dim(Nup, Ndw)
        ev = np.zeros(K)
        vec = np.zeros(K * dim)
        nit = C.c_int()
        _chk(lib().orc_davidson(self.ref, Nup, Ndw, K, accuracy, max_iter, _d(ev), _d(vec), C.byref(nit)))
        return ev, vec.reshape(K, dim), nit.value

    def block_lanczos(self, Nup, Ndw, q0, M0, max_iter_bl=600, accur_deflation=1e-7, accur_bl=1e-12, qr=True):
        q0 = _f64(q0)
        p = q0.shape[0]
        A = np.zeros(M0 * p * p)
        B = np.zeros(M0 * p * p)
        nf = C.c_int()
        _chk(lib().orc_block_lanczos_v(self.ref, Nup, Ndw, p, _d(q0), M0, max_iter_bl, accur_deflation, accur_bl,
                                       1 if qr else 0, C.byref(nf), _d(A), _d(B)))
        n = nf.value
        return (A[:n * p * p].reshape(n, p, p).transpose(0, 2, 1).copy(),
                B[:n * p * p].reshape(n, p, p).transpose(0, 2, 1).copy())


def _chk(rc):
    if rc:
        raise RuntimeError("oracle call failed with status %d" % rc)


def create_or_destroy(L, pm, a, Nup, Ndw, x, z=1.0):
    """y = z c_a x (pm=-1) or z c^dagger_a x (pm=+1); a is the spin-orbital."""
    spin = 1 if a >= L else 0
    tNup = Nup + (pm if spin == 0 else 0)
    tNdw = Ndw + (pm if spin == 1 else 0)
    dim = len(half_basis(L, tNup)) * len(half_basis(L, tNdw))
    x = _f64(x)
    y = np.zeros(dim)
    _chk(lib().orc_create_or_destroy(L, pm, a, Nup, Ndw, _d(x), z, _d(y)))
    return y


def tridiag_eig(d, e):
    d = _f64(d)
    n = len(d)
    e = _f64(np.append(np.asarray(e, dtype=np.float64)[:max(n - 1, 0)], 0.0))
    w = np.zeros(n)
    z = np.zeros(n * n)
    _chk(lib().orc_tridiag_eig(n, _d(d), _d(e), _d(w), _d(z)))
    return w, z.reshape(n, n).T.copy()


def symeig(A):
    A = np.asfortranarray(A, dtype=np.float64)
    n = A.shape[0]
    w = np.zeros(n)
    V = np.zeros(n * n)
    _chk(lib().orc_symeig(n, A.ctypes.data_as(_dp), _d(w), _d(V)))
    return w, V.reshape(n, n).T.copy()


def cf_eval(a, b, z):
    """a, b already transformed (continued_fraction ctor); z complex array."""
    a = _f64(a)
    b = _f64(b)
    z = np.ascontiguousarray(z, dtype=np.complex128)
    g = np.zeros(len(z), dtype=np.complex128)
    _chk(lib().orc_cf_eval(len(a), _d(a), _d(b), len(z), z.ctypes.data_as(_dp), g.ctypes.data_as(_dp)))
    return g


def cf_transform(a, b, e0, norm, create):
    """continued_fraction ctor (src_ed/continued_fraction.cpp:33-42)."""
    a = np.array(a, dtype=np.float64)
    b = np.array(b, dtype=np.float64) ** 2
    b[0] = norm
    a = a - e0 if create else e0 - a
    return a, b


def qmatrix_streamline(e, V, wtol=1e-5, vtol=1e-10):
    e = _f64(e)
    L, M = V.shape
    Vf = np.asfortranarray(V, dtype=np.float64)
    e2 = np.zeros(max(M, 1))
    V2 = np.zeros(max(M, 1) * L)
    M2 = lib().orc_qmatrix_streamline(L, M, _d(e), Vf.ctypes.data_as(_dp), wtol, vtol, _d(e2), _d(V2))
    return e2[:M2], V2[:M2 * L].reshape(M2, L).T.copy()


def qmatrix_norm_error(V):
    Vf = np.asfortranarray(V, dtype=np.float64)
    L, M = V.shape
    return lib().orc_qmatrix_norm_error(L, M, Vf.ctypes.data_as(_dp))


def vdvh(e, V, z):
    e = _f64(e)
    L, M = V.shape
    Vf = np.asfortranarray(V, dtype=np.float64)
    z = np.ascontiguousarray(z, dtype=np.complex128)
    G = np.zeros(len(z) * L * L, dtype=np.complex128)
    _chk(lib().orc_vdvh(L, M, _d(e), Vf.ctypes.data_as(_dp), len(z), z.ctypes.data_as(_dp),
                        G.ctypes.data_as(_dp)))
    return G.reshape(len(z), L, L).transpose(0, 2, 1)  # G[k][b, a] column-major -> [k, row, col]


def mcf_eval(A, B, W, z):
    A = np.asarray(A, dtype=np.float64)
    B = np.asarray(B, dtype=np.float64)
    nfl, p, _ = A.shape
    q = W.shape[1]
    Af = np.ascontiguousarray(A.transpose(0, 2, 1))
    Bf = np.ascontiguousarray(B.transpose(0, 2, 1))
    Wf = np.asfortranarray(W, dtype=np.float64)
    z = np.ascontiguousarray(z, dtype=np.complex128)
    G = np.zeros(len(z) * q * q, dtype=np.complex128)
    _chk(lib().orc_mcf_eval(p, nfl, _d(Af), _d(Bf), Wf.ctypes.data_as(_dp), q, len(z), z.ctypes.data_as(_dp),
                            G.ctypes.data_as(_dp)))
    return G.reshape(len(z), q, q).transpose(0, 2, 1)


def csr_hv(Iptr, J, v, dg, x):
    x = _f64(x)
    y = np.zeros_like(x)
    _chk(lib().orc_hv_csr(len(x), Iptr.ctypes.data_as(_i32p), J.ctypes.data_as(_i32p), _d(v), _d(dg), _d(x),
                          _d(y)))
    return y


def csr_hv_full_omp(Iptr, J, v, dg, x, nrep=1):
    x = _f64(x)
    y = np.zeros_like(x)
    _chk(lib().orc_hv_csr_full_omp(len(x), Iptr.ctypes.data_as(_i32p), J.ctypes.data_as(_i32p), _d(v), _d(dg),
                                   _d(x), _d(y), nrep))
    return y


def csr_hv_full_omp_times(Iptr, J, v, dg, x, nrep):
    """ms of each of nrep row-parallel full-CSR products (rows expanded once, untimed)"""
    x = _f64(x)
    y = np.zeros_like(x)
    ms = np.zeros(nrep)
    _chk(lib().orc_hv_csr_full_omp_timed(len(x), Iptr.ctypes.data_as(_i32p), J.ctypes.data_as(_i32p), _d(v),
                                         _d(dg), _d(x), _d(y), nrep, _d(ms)))
    return ms


def openmp_threads():
    return lib().orc_openmp_threads()
