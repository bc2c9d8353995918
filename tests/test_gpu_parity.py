"""GPU parity of the HIP path against the CPU oracle (and, for tiny sectors, the
dense Jordan-Wigner Hamiltonian).  Bars: bit-exact for bases / patterns /
sign tables; 1e-12 relative for H.x; 1e-10 relative for E0; 1e-8 for G(z)."""
import numpy as np
import pytest

from pyqcm_amd import qcm, workloads as W, _lib
from tests.dense_ref import dense_hamiltonian

pytestmark = pytest.mark.gpu

_label = [1000]


def _new_label():
    _label[0] += 1
    return _label[0]


@pytest.fixture(scope="module", autouse=True)
def _device():
    _lib.require_device()
    qcm.set_global_parameterC("Hamiltonian_format", "G")
    qcm.set_global_parameterC("GF_method", "L")
    yield


def _instance(wl, sector=None, name=None):
    lab = _new_label()
    W.setup(wl, label=lab, sectors=sector, model_name=name or ("%s_%d" % (wl.name, lab)))
    return lab


def _random_workload(seed, L=6):
    rng = np.random.default_rng(seed)
    ops, vals = {}, {}
    for t in range(3):
        el = []
        for _ in range(2 * L):
            i, j = sorted(rng.choice(L, 2, replace=False))
            v = float(rng.normal())
            el.append((int(i), int(j), v))
            el.append((int(i) + L, int(j) + L, v if t != 1 else float(rng.normal())))
        for i in range(L):
            el.append((i, i, float(rng.normal())))
        ops["h%d" % t] = ("one-body", el)
        vals["h%d" % t] = float(rng.uniform(0.5, 1.5))
    ops["V"] = ("interaction", [(i, i + L, 1.0) for i in range(L)] + [(0, 1 + L, 0.3), (1, 2, 0.2)])
    vals["V"] = 2.0
    return W.Workload("rnd%d" % seed, L, 0, ops, vals, "R0:N6:S0", list(range(L)))


@pytest.mark.parametrize("fmt", ["G", "S", "N"])
@pytest.mark.parametrize("seed,Nup,Ndw", [(0, 3, 3), (1, 2, 4), (2, 4, 3)])
def test_hv_matches_dense_and_oracle(oracle, fmt, seed, Nup, Ndw):
    wl = _random_workload(seed)
    lab = _instance(wl, "R0:N%d:S%d" % (Nup + Ndw, Nup - Ndw))
    H = qcm.Hamiltonian(lab, "R0:N%d:S%d" % (Nup + Ndw, Nup - Ndw), fmt)
    Hd = dense_hamiltonian(wl.L, wl.terms(), Nup, Ndw)
    assert H.dim == Hd.shape[0]
    rng = np.random.default_rng(seed + 10)
    x = rng.normal(size=H.dim)
    y0 = rng.normal(size=H.dim)
    got = H.mult_add(x, y0.copy())
    ref = y0 + Hd @ x
    M = oracle.Model(wl.L, wl.terms())
    ref2 = M.hv_factorized(Nup, Ndw, x, y0)
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() <= 1e-12 * scale
    assert np.abs(got - ref2).max() <= 1e-12 * scale
    assert np.allclose(H.diag(), np.diag(Hd), atol=1e-13)


def test_csr_pattern_bit_exact(oracle):
    wl = W.plaquette_bath(n_bath=2)
    lab = _instance(wl, "R0:N6:S0")
    H = qcm.Hamiltonian(lab, "R0:N6:S0", "S")
    Ip, J, v = H.csr_pattern()
    M = oracle.Model(wl.L, wl.terms())
    Ip2, J2, v2, dg2 = M.csr(3, 3)
    assert np.array_equal(Ip, Ip2)
    assert np.array_equal(J, J2)
    assert np.allclose(v, v2, rtol=0, atol=1e-15)


@pytest.mark.parametrize("variant", [1, 2, 3])
def test_factorized_variants_agree(oracle, variant):
    wl = W.plaquette_bath(n_bath=4)     # L = 8, dim(4,4) = 4900
    lab = _instance(wl, "R0:N8:S0")
    H = qcm.Hamiltonian(lab, "R0:N8:S0", "G")
    H.set_variant(variant)
    x = np.random.default_rng(3).normal(size=H.dim)
    got = H.apply(x)
    ref = oracle.Model(wl.L, wl.terms()).hv_factorized(4, 4, x)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("row_order", [0, 1])
@pytest.mark.parametrize("block", [0, 1024])
@pytest.mark.parametrize("sell16", [True, False])
@pytest.mark.parametrize("variant", [1, 2, 3])
@pytest.mark.parametrize("kind", ["chain12", "plaquette8b"])
def test_factorized_sliced_ell_modes(oracle, kind, variant, sell16, block, row_order):
    """Spin-up rows of >= 512 states take the sliced-ELL row kernel; a single
    hopping magnitude (chain) allows the 16-bit sign-only table, several
    magnitudes (plaquette + bath) the packed 32-bit one; rows in label or
    breadth-first order."""
    if kind == "chain12":
        wl, sec, Nup, Ndw = W.hubbard_chain(12, U=4.0, mu=2.0), "R0:N11:S-1", 5, 6   # du = 792
    else:
        wl, sec, Nup, Ndw = W.plaquette_bath(seed=2025), "R0:N12:S0", 6, 6        # du = 924
    qcm.set_global_parameterC("sell16", sell16)
    qcm.set_global_parameterC("hv_block", block)
    qcm.set_global_parameterC("hv_row_order", row_order)
    try:
        lab = _instance(wl, sec)
        H = qcm.Hamiltonian(lab, sec, "G")
    finally:
        qcm.set_global_parameterC("sell16", True)
        qcm.set_global_parameterC("hv_block", 0)
        qcm.set_global_parameterC("hv_row_order", 1)
    H.set_variant(variant)
    x = np.random.default_rng(5).normal(size=H.dim)
    got = H.apply(x)
    ref = oracle.Model(wl.L, wl.terms()).hv_factorized(Nup, Ndw, x)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_torus4x4_full_sector_hv_matches_oracle(oracle):
    """The 4x4 periodic cluster at half filling (dim 165 636 900): every kernel
    schedule against the oracle's OpenMP product, and symmetry of H."""
    wl = W.torus(4, 4, U=8.0)
    sec = "R0:N16:S0"
    lab = _instance(wl, sec)
    H = qcm.Hamiltonian(lab, sec, "G")
    assert H.dim == 165636900
    rng = np.random.default_rng(7)
    x = rng.normal(size=H.dim)
    ref = oracle.Model(wl.L, wl.terms()).hv_factorized_omp(8, 8, x, nrep=1)
    scale = np.abs(ref).max()
    for variant in (1, 2, 0):
        H.set_variant(variant)
        got = H.apply(x)
        err = np.abs(got - ref).max()
        print("variant", variant, "max |err| / scale", err / scale, flush=True)
        assert err <= 1e-12 * scale
        del got
    w = rng.normal(size=H.dim)
    lhs = float(w @ H.apply(x))
    rhs = float(H.apply(w) @ x)
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)


@pytest.mark.parametrize("poll", [True, False])
def test_ground_state_chain8_known_answer(poll):
    """Dense eigh: -20.23580699912965; reductions returned by polling pinned
    memory or by stream synchronisation."""
    wl = W.hubbard_chain(8, U=4.0, mu=2.0)
    lab = _instance(wl)
    qcm.set_global_parameterC("poll_reductions", poll)
    try:
        E0, sec = qcm.ground_state_solve(lab)
    finally:
        qcm.set_global_parameterC("poll_reductions", True)
    assert abs(E0 - (-20.23580699912965)) < 1e-10 * 20.3
    assert sec.startswith("R0:N8:S0")


def test_ground_state_chain8_full_reorth():
    """Lanczos with full FP64 reorthogonalisation against the stored basis
    gives the same ground state (dense eigh: -20.23580699912965)."""
    wl = W.hubbard_chain(8, U=4.0, mu=2.0)
    lab = _instance(wl)
    qcm.set_global_parameterC("lanczos_reorth", True)
    try:
        E0, sec = qcm.ground_state_solve(lab)
    finally:
        qcm.set_global_parameterC("lanczos_reorth", False)
    assert abs(E0 - (-20.23580699912965)) <= 1e-10 * 20.2358


def test_ground_state_1D_4_known_answer():
    wl = W.hubbard_chain(4, U=4.0, mu=2.0)
    lab = _instance(wl)
    E0, sec = qcm.ground_state_solve(lab)
    assert round(E0, 7) == -9.9531453   # examples/1D_4/GS.tsv:2


def test_cluster_averages_1D_4_known_answer():
    """examples/1D_4/averages.tsv:2 (U=4, mu=2, t=1): averages and variances
    per site.  The lattice layer scales the cluster operators by their norms
    (1/4 for t and U, -1/4 for mu on the 4-site cluster); the ED-level values
    are <O> and <O^2> - <O>^2 of the operators as defined."""
    wl = W.hubbard_chain(4, U=4.0, mu=2.0)
    lab = _instance(wl)
    E0, _ = qcm.ground_state_solve(lab)
    ave = qcm.cluster_averages(lab)
    norm = {"t": 0.25, "U": 0.25, "mu": -0.25}
    ref = {"U": (0.084896413, 0.01648638), "mu": (1.0, 0.0), "t": (-0.82787198, 0.26378208)}
    for name, (a, v) in ref.items():
        got_a, got_v = ave[name]
        assert abs(got_a * norm[name] - a) <= 1e-8, name
        assert abs(got_v * norm[name] ** 2 - v) <= 1e-8, name
    # the energy is the sum of value * <O> over the instance operators
    assert abs(sum(wl.values[n] * ave[n][0] for n in ave) - E0) <= 1e-10 * abs(E0)


@pytest.mark.parametrize("fmt", ["G", "S", "N"])
def test_lanczos_gs_matches_oracle(oracle, fmt):
    wl = W.plaquette_bath(n_bath=4)
    lab = _instance(wl, "R0:N8:S0")
    H = qcm.Hamiltonian(lab, "R0:N8:S0", fmt)
    x = np.random.default_rng(4).normal(size=H.dim)
    E0, gs, nit = H.lanczos_gs(x)
    E0r, gsr, al, be = oracle.Model(wl.L, wl.terms()).lanczos_gs(4, 4, x)
    assert abs(E0 - E0r) <= 1e-10 * abs(E0r)
    g = gs / np.linalg.norm(gs)
    gr = gsr / np.linalg.norm(gsr)
    assert abs(abs(g @ gr) - 1.0) < 1e-9


def test_create_destroy_bit_exact(oracle):
    L, Nup, Ndw = 6, 3, 2
    x = np.random.default_rng(5).normal(size=len(qcm.half_basis(L, Nup)) * len(qcm.half_basis(L, Ndw)))
    for a in [0, 3, 5, 6, 9, 11]:
        for pm in (-1, 1):
            got = qcm.create_or_destroy(L, pm, a, Nup, Ndw, x)
            ref = oracle.create_or_destroy(L, pm, a, Nup, Ndw, x)
            assert np.array_equal(got, ref), (a, pm)


def test_gf_kernels_match_oracle(oracle):
    rng = np.random.default_rng(6)
    z = W.real_axis(64, -5, 5, 0.1)
    fr = []
    for n in [1, 5, 40]:
        a = rng.normal(size=n)
        b = np.abs(rng.normal(size=n + 1))
        fr.append((a, b))
    got = qcm.cf_eval(fr, z)
    for k, (a, b) in enumerate(fr):
        assert np.allclose(got[:, k], oracle.cf_eval(a, b, z), rtol=1e-13, atol=1e-14)
    e = rng.normal(size=50)
    V = rng.normal(size=(4, 50))
    assert np.allclose(qcm.vdvh(e, V, z), oracle.vdvh(e, V, z), rtol=1e-13, atol=1e-14)
    p, nfl = 3, 7
    A = rng.normal(size=(nfl, p, p))
    A = A + A.transpose(0, 2, 1)
    B = np.triu(rng.normal(size=(nfl, p, p)))
    Wm = np.triu(rng.normal(size=(p, p)))
    assert np.allclose(qcm.mcf_eval(A, B, Wm, z), oracle.mcf_eval(A, B, Wm, z), rtol=1e-12, atol=1e-13)


def _oracle_gf_band(oracle, wl, Nup, Ndw, z, start):
    """oracle pipeline: GS + band Lanczos per (pm) + streamline + VDVH."""
    M = oracle.Model(wl.L, wl.terms())
    E0, gs, _, _ = M.lanczos_gs(Nup, Ndw, start)
    gs = gs / np.linalg.norm(gs)
    L = wl.L
    orbs = list(range(wl.n_sites))
    es, Vs = [], []
    for pm in (-1, 1):
        phi = np.array([oracle.create_or_destroy(L, pm, a, Nup, Ndw, gs) for a in orbs])
        e, Q = M.band_lanczos(Nup + pm, Ndw, phi)
        e = e - E0
        if pm == -1:
            e = -e
        e, Q = oracle.qmatrix_streamline(e, Q)
        es.append(e)
        Vs.append(Q)
    e = np.concatenate(es)
    V = np.concatenate(Vs, axis=1)
    return E0, oracle.vdvh(e, V, z)


def test_green_function_band_lanczos_chain8(oracle):
    wl = W.hubbard_chain(8, U=4.0, mu=2.0)
    lab = _instance(wl)
    qcm.set_global_parameterC("GF_method", "L")
    E0, _ = qcm.ground_state_solve(lab)
    z = W.matsubara(64, 50.0)
    G = qcm.Green_function_grid(z, label=lab)
    x = np.random.default_rng(7).normal(size=70 * 70)
    E0r, Gr = _oracle_gf_band(oracle, wl, 4, 4, z, x)
    assert abs(E0 - E0r) < 1e-10 * abs(E0r)
    assert np.abs(G - Gr.transpose(0, 2, 1)).max() < 1e-8


def test_band_lanczos_deflation_matches_exact(oracle):
    """Band Lanczos run to exhaustion on a 56-state sector with a dependent
    start vector: initial and late deflations.  G(z) from (e, Q) equals the
    exact resolvent phi^T (z - H)^-1 phi and the oracle's."""
    wl = W.hubbard_chain(8, U=4.0, mu=2.0)
    lab = _instance(wl, "R0:N3:S3")
    H = qcm.Hamiltonian(lab, "R0:N3:S3", "G")
    assert H.dim == 56
    rng = np.random.default_rng(21)
    phi = rng.normal(size=(5, 56))
    phi[4] = phi[0] - 2.0 * phi[1]
    qcm.set_global_parameterC("accur_band_lanczos", 0.0)
    try:
        e, Q = H.band_lanczos(phi)
    finally:
        qcm.set_global_parameterC("accur_band_lanczos", 1e-12)
    eo, Qo = oracle.Model(wl.L, wl.terms()).band_lanczos(3, 0, phi, accur_bl=0.0)
    Hd = dense_hamiltonian(wl.L, wl.terms(), 3, 0)
    z = W.matsubara(16, 10.0)
    for zz in z:
        G = (Q / (zz - e)) @ Q.T
        Go = (Qo / (zz - eo)) @ Qo.T
        Gx = phi @ np.linalg.solve(zz * np.eye(56) - Hd, phi.T)
        assert np.abs(G - Gx).max() < 1e-10 * np.abs(Gx).max()
        assert np.abs(G - Go).max() < 1e-10 * np.abs(Gx).max()


def _cf_reference(oracle, wl, E0r, gs, z):
    """oracle pipeline of build_cf + continued_fraction_set::Green_function."""
    n, L = wl.n_sites, wl.L
    Nup = Ndw = L // 2
    M = oracle.Model(L, wl.terms())
    F = {}
    for o1 in range(n):
        for o2 in range(o1 + 1):
            tot = np.zeros(len(z), complex)
            for pm in (1, -1):
                psi = oracle.create_or_destroy(L, pm, o1, Nup, Ndw, gs)
                if o2 != o1:
                    psi = psi + oracle.create_or_destroy(L, pm, o2, Nup, Ndw, gs)
                nrm = psi @ psi
                a, b = M.lanczos_green(Nup + pm, Ndw, psi / np.sqrt(nrm))
                a, b = oracle.cf_transform(a, b, E0r, nrm, pm == 1)
                tot += oracle.cf_eval(a, b, z)
            F[(o1, o2)] = tot
    Gr = np.zeros((len(z), n, n), complex)
    for o in range(n):
        Gr[:, o, o] = F[(o, o)]
    for o1 in range(n):
        for o2 in range(o1):
            v = 0.5 * (F[(o1, o2)] - F[(o1, o1)] - F[(o2, o2)])
            Gr[:, o1, o2] = v
            Gr[:, o2, o1] = v
    return Gr


def _exact_g(wl, z):
    """Lehmann sum from dense eigendecompositions (sectors N-1, N+1 of spin up)."""
    from tests.dense_ref import dense_c
    L = wl.L
    Nup = Ndw = L // 2
    H0 = dense_hamiltonian(L, wl.terms(), Nup, Ndw)
    ev, U = np.linalg.eigh(H0)
    E0, gs = ev[0], U[:, 0]
    Hp = dense_hamiltonian(L, wl.terms(), Nup + 1, Ndw)
    Hm = dense_hamiltonian(L, wl.terms(), Nup - 1, Ndw)
    ep, Up = np.linalg.eigh(Hp)
    em, Um = np.linalg.eigh(Hm)
    n = wl.n_sites
    C = np.array([Up.T @ (dense_c(L, a, Nup, Ndw, True) @ gs) for a in range(n)])
    D = np.array([Um.T @ (dense_c(L, a, Nup, Ndw, False) @ gs) for a in range(n)])
    return (np.einsum("ai,bi,zi->zab", C, C, 1.0 / (z[:, None] - (ep - E0)[None, :])) +
            np.einsum("ai,bi,zi->zab", D, D, 1.0 / (z[:, None] + (em - E0)[None, :])))


def test_green_function_cf_chain8(oracle):
    """C1: continued-fraction G for all 8 orbitals.  On the Matsubara axis the
    recurrence converges and GPU == oracle to 1e-8.  On the real-axis grid
    (512 points in [-8, 8], eta = 0.1) the reference algorithm itself (no
    reorthogonalisation, 400 levels) is ~1e-7 away from the exact Lehmann sum,
    so GPU and oracle are each checked against the exact G at 1e-6."""
    wl = W.hubbard_chain(8, U=4.0, mu=2.0)
    qcm.set_global_parameterC("GF_method", "F")
    lab = _instance(wl)
    E0, _ = qcm.ground_state_solve(lab)
    zm = W.matsubara(256, 50.0)
    zr = W.real_axis(512, -8, 8, 0.1)
    Gm = qcm.Green_function_grid(zm, label=lab)
    Gre = qcm.Green_function_grid(zr, label=lab)
    qcm.set_global_parameterC("GF_method", "L")
    M = oracle.Model(wl.L, wl.terms())
    x = np.random.default_rng(8).normal(size=70 * 70)
    E0r, gs, _, _ = M.lanczos_gs(4, 4, x)
    gs = gs / np.linalg.norm(gs)
    assert abs(E0 - E0r) < 1e-10 * abs(E0r)
    assert np.abs(Gm - _cf_reference(oracle, wl, E0r, gs, zm)).max() < 1e-8
    Gx = _exact_g(wl, zr)
    assert np.abs(Gre - Gx).max() < 1e-6
    assert np.abs(_cf_reference(oracle, wl, E0r, gs, zr) - Gx).max() < 1e-6


def test_davidson_matches_dense():
    wl = W.hubbard_chain(6, U=4.0, mu=2.0)
    lab = _instance(wl)
    H = qcm.Hamiltonian(lab, "R0:N6:S0", "G")
    ev, vec, nit = H.davidson(4)
    Hd = dense_hamiltonian(wl.L, wl.terms(), 3, 3)
    ref = np.linalg.eigvalsh(Hd)[:4]
    assert np.allclose(ev, ref, atol=1e-8)


@pytest.mark.parametrize("qr", [True, False])
def test_block_lanczos_mcf_matches_oracle(oracle, qr):
    """blockLanczos (QR) and blockLanczosSVD (polar, the reference's default,
    block_Lanczos_QR = false) floor by floor against the oracle"""
    wl = W.hubbard_chain(6, U=4.0, mu=2.0)
    lab = _instance(wl, "R0:N7:S1")
    H = qcm.Hamiltonian(lab, "R0:N7:S1", "G")
    rng = np.random.default_rng(9)
    q = np.linalg.qr(rng.normal(size=(H.dim, 3)))[0].T.copy()
    qcm.set_global_parameterC("block_Lanczos_QR", qr)
    try:
        A, B = H.block_lanczos(q, 30)
    finally:
        qcm.set_global_parameterC("block_Lanczos_QR", False)
    Ar, Br = oracle.Model(wl.L, wl.terms()).block_lanczos(4, 3, q, 30, qr=qr)
    k = min(len(A), len(Ar), 8)
    assert np.allclose(A[:k], Ar[:k], atol=1e-9)
    if qr:
        assert np.allclose(np.abs(B[:k]), np.abs(Br[:k]), atol=1e-9)
    else:
        assert np.allclose(B[:k], Br[:k], atol=1e-9)
        assert np.allclose(B[:k], B[:k].transpose(0, 2, 1), atol=1e-12)


def test_krylov_outputs_longer_than_cap_fail_with_the_size_needed():
    """ed_lanczos_green / ed_band_lanczos / ed_block_lanczos write nothing and
    return ED_ERR_INVALID with the size needed when the result exceeds cap; the
    Python wrappers then retry with that capacity"""
    import ctypes as C
    wl = W.hubbard_chain(6, U=4.0, mu=2.0)
    lab = _instance(wl, "R0:N7:S1")
    H = qcm.Hamiltonian(lab, "R0:N7:S1", "G")
    rng = np.random.default_rng(10)
    psi = rng.normal(size=H.dim)
    psi /= np.linalg.norm(psi)
    a_full, b_full = H.lanczos_green(psi, cap=4)  # retried with the size needed
    assert len(a_full) > 4 and len(b_full) == len(a_full) + 1
    a = np.full(4, -7.0)
    b = np.full(5, -7.0)
    n = C.c_int()
    rc = _lib.lib().ed_lanczos_green(H.h, psi.ctypes.data_as(_lib._dp), 4, a.ctypes.data_as(_lib._dp),
                                     b.ctypes.data_as(_lib._dp), C.byref(n))
    assert rc == _lib.ED_ERR_INVALID and n.value == len(a_full)
    assert np.all(a == -7.0) and np.all(b == -7.0)
    phi = rng.normal(size=(2, H.dim))
    e_full, Q_full = H.band_lanczos(phi, cap=3)
    assert len(e_full) > 3 and Q_full.shape == (2, len(e_full))
    e = np.full(3, -7.0)
    Q = np.full(6, -7.0)
    M = C.c_int()
    rc = _lib.lib().ed_band_lanczos(H.h, 2, phi.ctypes.data_as(_lib._dp), 3, e.ctypes.data_as(_lib._dp),
                                    Q.ctypes.data_as(_lib._dp), C.byref(M))
    assert rc == _lib.ED_ERR_INVALID and M.value == len(e_full)
    assert np.all(e == -7.0) and np.all(Q == -7.0)
    q = np.linalg.qr(rng.normal(size=(H.dim, 2)))[0].T.copy()
    A_full, B_full = H.block_lanczos(q, 20, cap=2)
    assert len(A_full) > 2
    A = np.full(8, -7.0)
    B = np.full(8, -7.0)
    nfl = C.c_int()
    rc = _lib.lib().ed_block_lanczos(H.h, 2, q.ctypes.data_as(_lib._dp), 20, 2, A.ctypes.data_as(_lib._dp),
                                     B.ctypes.data_as(_lib._dp), C.byref(nfl))
    assert rc == _lib.ED_ERR_INVALID and nfl.value == len(A_full)
    assert np.all(A == -7.0) and np.all(B == -7.0)


def test_mcf_green_function_matches_band(oracle):
    wl = W.hubbard_chain(6, U=4.0, mu=2.0)
    z = W.matsubara(32, 30.0)
    qcm.set_global_parameterC("GF_method", "L")
    lab1 = _instance(wl)
    G1 = qcm.Green_function_grid(z, label=lab1)
    qcm.set_global_parameterC("GF_method", "M")
    qcm.set_global_parameterC("accur_band_lanczos", 0.0)
    lab2 = _instance(wl)
    G2 = qcm.Green_function_grid(z, label=lab2)
    qcm.set_global_parameterC("accur_band_lanczos", 1e-12)
    qcm.set_global_parameterC("GF_method", "L")
    assert np.abs(G1 - G2).max() < 1e-7


@pytest.mark.parametrize("vector_block,blocks_per_cu", [(256, 4), (512, 2), (1024, 1)])
def test_krylov_vector_kernel_shapes(vector_block, blocks_per_cu):
    """Ground state (dense E0) and band-Lanczos G(i w) of a plaquette + 4 bath
    sector with the Krylov vector kernels (multi-dot, multi-axpy, combine)
    at several workgroup sizes and grids: the same results as the default."""
    wl = W.plaquette_bath(n_bath=4)
    z = W.matsubara(8, 20.0)

    def solve():
        lab = _instance(wl, "R0:N8:S0")
        E0, _ = qcm.ground_state_solve(lab)
        G = qcm.Green_function_grid(z, label=lab)
        qcm.erase_model_instance(lab)
        return E0, G

    E_ref, G_ref = solve()
    qcm.set_global_parameterC("vector_block", vector_block)
    qcm.set_global_parameterC("reduction_blocks_per_cu", blocks_per_cu)
    try:
        E0, G = solve()
    finally:
        qcm.set_global_parameterC("vector_block", 1024)
        qcm.set_global_parameterC("reduction_blocks_per_cu", 1)
    Ed = np.linalg.eigvalsh(dense_hamiltonian(wl.L, wl.terms(), 4, 4))[0]
    assert abs(E0 - Ed) <= 1e-10 * abs(Ed)
    assert abs(E0 - E_ref) <= 1e-12 * abs(Ed)
    assert np.abs(G - G_ref).max() < 1e-10 * np.abs(G_ref).max()


def test_bath_cluster_1D_2_4b_known_answer():
    """examples/1D_2_4b/cdmft.tsv:3 (the converged CDMFT bath of the reference's
    own run): 2 cluster sites + 4 bath orbitals (model_1D_2_4b.py), U = 4,
    mu = 2, t = 1, eb1 = 1.1438241, eb2 = -1.1438239, tb1 = 0.60939833,
    tb2 = 0.60939828, sector N = 6, S = 0.  E0_1 and the averages / variances
    of every operator (lattice operators t, U, mu scaled by 1/2, 1/2, -1/2 as
    the lattice layer does on a 2-site cluster; bath operators as defined),
    printed to 8 digits."""
    ns, nb = 2, 4
    L = ns + nb
    one = {  # 0-based spin-up elements; the down ones are shifted by L
        "t": [(0, 1, -1.0)], "mu": [(0, 0, -1.0), (1, 1, -1.0)],
        "tb1": [(0, 2, -1.0), (1, 3, -1.0)], "tb2": [(0, 4, -1.0), (1, 5, -1.0)],
        "eb1": [(2, 2, 1.0), (3, 3, 1.0)], "eb2": [(4, 4, 1.0), (5, 5, 1.0)],
    }
    name = "L2_4b"
    qcm.new_model(name, ns, nb)
    for op, el in one.items():
        qcm.new_operator(name, op, "one-body", [(r + 1, c + 1, v) for (r, c, v) in el + [(r + L, c + L, v) for (r, c, v) in el]])
    qcm.new_operator(name, "U", "interaction", [(i + 1, i + L + 1, 1.0) for i in range(ns)])
    vals = {"U": 4.0, "mu": 2.0, "t": 1.0, "eb1": 1.1438241, "eb2": -1.1438239, "tb1": 0.60939833, "tb2": 0.60939828}
    lab = _new_label()
    qcm.new_model_instanceC(name, vals, "R0:N6:S0", lab)
    E0, sec = qcm.ground_state_solve(lab)
    assert abs(E0 - (-9.8979868)) < 2e-7, E0
    ave = qcm.cluster_averages(lab)
    norm = {"t": 0.5, "U": 0.5, "mu": -0.5}
    ref = {"U": (0.10367723, 0.0416604), "eb1": (0.099838804, 0.094703557), "eb2": (3.9001612, 0.094703556),
           "mu": (1.0, 0.038057271), "t": (-0.67470015, 0.41831011), "tb1": (-0.84600224, 1.7731342),
           "tb2": (-0.84600223, 1.7731342)}
    for op, (a, v) in ref.items():
        f = norm.get(op, 1.0)
        got_a, got_v = ave[op]
        assert abs(got_a * f - a) < 1e-6, (op, got_a * f, a)
        assert abs(got_v * f * f - v) < 1e-6, (op, got_v * f * f, v)
    qcm.erase_model_instance(lab)
