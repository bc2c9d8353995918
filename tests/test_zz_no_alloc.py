"""The allocation-free product ed_hamiltonian_apply_dot_noalloc against
ed_hamiltonian_apply_device, the entry point it is a variant of: the same
product (bit-exact: same kernels, same schedule) and <x, y> equal to the host
dot product of the two vectors to 1e-12 relative, on two calls in a row, the
second on a smaller sector (the fixed scratch keeps the first call's partials,
which the second must not read)."""
import ctypes as C

import numpy as np
import pytest

from pyqcm_amd import _lib, qcm, workloads as W

pytestmark = pytest.mark.gpu


class _Dev:
    """A device array through the C-ABI's own allocator (test plumbing)."""

    def __init__(self, n):
        self.n = max(int(n), 1)
        p = C.c_void_p()
        _lib.check(_lib.lib().ed_device_alloc(8 * self.n, C.byref(p)))
        self.p = p.value

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        _lib.check(_lib.lib().ed_device_copy(C.c_void_p(self.p), a.ctypes.data_as(C.c_void_p), 8 * a.size, 0))

    def get(self, n=None):
        out = np.empty(self.n if n is None else n)
        _lib.check(_lib.lib().ed_device_copy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.p), 8 * out.size, 1))
        return out

    def free(self):
        _lib.check(_lib.lib().ed_device_free(C.c_void_p(self.p), 8 * self.n))


@pytest.fixture(scope="module", autouse=True)
def _device():
    _lib.require_device()
    qcm.set_global_parameterC("Hamiltonian_format", "G")
    yield


def _handle(wl, label, sector, fmt):
    W.setup(wl, label=label, sectors=sector, model_name="noalloc_%s_%d" % (wl.name, label))
    return qcm.Hamiltonian(label, sector, fmt)


@pytest.mark.parametrize("fmt", ["G", "S"])
def test_noalloc_product_equals_apply_device(fmt):
    cases = [(W.hubbard_chain(10, U=4.0, mu=2.0), "R0:N10:S0", 9501),   # 63 504 states
             (W.hubbard_chain(8, U=4.0, mu=2.0), "R0:N8:S0", 9502)]      # 4 900 states, second and smaller
    rng = np.random.default_rng(7)
    dims = []
    for wl, sector, label in cases:
        H = _handle(wl, label + (0 if fmt == "G" else 10), sector, fmt)
        n = H.dim
        dims.append(n)
        x = rng.normal(size=n)
        y0 = rng.normal(size=n)
        dx, dy_ref, dy, ddot = _Dev(n), _Dev(n), _Dev(n), _Dev(1)
        try:
            for beta in (0.0, -0.75):
                dx.put(x)
                dy_ref.put(y0)
                dy.put(y0)
                ddot.put(np.array([np.nan]))
                H.apply_device(dx.p, dy_ref.p, beta)
                H.apply_dot_noalloc(dx.p, dy.p, ddot.p, beta)
                y_ref = dy_ref.get(n)
                y = dy.get(n)
                dot = ddot.get(1)[0]
                assert np.array_equal(y, y_ref), (fmt, sector, beta, float(np.max(np.abs(y - y_ref))))
                want = float(np.dot(x, y_ref))
                assert abs(dot - want) <= 1e-12 * max(abs(want), np.linalg.norm(x) * np.linalg.norm(y_ref)), \
                    (fmt, sector, beta, dot, want)
        finally:
            for d in (dx, dy_ref, dy, ddot):
                d.free()
    assert dims[1] < dims[0]


def test_noalloc_rejects_null_dot():
    wl = W.hubbard_chain(6, U=4.0, mu=2.0)
    H = _handle(wl, 9530, "R0:N6:S0", "G")
    d = _Dev(H.dim)
    try:
        with pytest.raises(_lib.EDError):
            H.apply_dot_noalloc(d.p, d.p, 0, 0.0)
    finally:
        d.free()
