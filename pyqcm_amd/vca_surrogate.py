"""Surrogate-assisted variational cluster approximation (no counterpart in the
reference, whose VCA driver is pyqcm/vca.py).

The cluster-model family.  A 2x2 plaquette (sites 0 = (0, 0), 1 = (1, 0),
2 = (0, 1), 3 = (1, 1); sublattice A = {0, 3}, B = {1, 2}, eta = +1 / -1) with
8 bath orbitals in two shells: bath orbital 4 + 4 s + i couples to site i
(shell s = 0, 1).  Twelve parameters:

    U, mu      Hubbard U and chemical potential on the cluster sites
    t, tp      nearest-neighbour and diagonal hopping (-t, -tp)
    M          Neel Weiss field  M sum_i eta_i (n_i,up - n_i,dn)
    e0, e1     bath energies of the two shells
    m0, m1     staggered bath fields  m_s eta_i (n_b,up - n_b,dn)
    v0, v1     cluster-bath hybridization of each shell
    tb         hopping between the two bath orbitals of one site

The family keeps two symmetries, used to reduce what the surrogate learns:
the diagonal mirrors (sites 1 <-> 2, and 0 <-> 3, with their bath orbitals)
make the spin-up Green function depend on five orbital pairs only,
(0,0) = (3,3), (1,1) = (2,2), (1,0) = (2,0) = (3,1) = (3,2), (3,0) and (2,1); and
spin flip combined with the sublattice exchange 0 <-> 1, 2 <-> 3 gives the
spin-down block G_dn[i, j] = G_up[p(i), p(j)].

The learned map: the 12 parameters -> E0 and, for each of the five pairs, the
first NCF levels (a_n, b_n) of the electron and hole continued fractions of
the spin-up block (the instance's own fractions, continued_fraction.cpp:33-42),
1 + 5 x 2 x 2 x NCF outputs.

The VCA problem: the square-lattice Hubbard model (t, tp, U, mu) tiled by this
cluster; the Potthoff functional Omega(M) on a fixed (frequency, wavevector)
grid, with the bath parameters held fixed; a stationary point in M is found by
the reference's Newton-Raphson procedure (pyqcm/vca.py:241-430: 1 + 2n + n(n-1)/2
functional evaluations per iteration, steps redefined from the last move, stop
when every |dx| < accur, then one more evaluation at the solution).  Every
evaluation on the exact path is a full ED solve.  The surrogate path starts the
same procedure at the surrogate's own stationary point, located on the
surrogate's Omega (the predicted fractions evaluated through the same lattice
sums); exact ED performs every evaluation of the procedure, so each accepted
point is verified by the exact solver, and a failed proposal simply costs more
iterations.
"""
import ctypes as C

import numpy as np

from . import lattice as LT
from . import qcm
from ._lib import check, lib
from .surrogate import MLP

PARAMS = ("U", "mu", "t", "tp", "M", "e0", "e1", "m0", "m1", "v0", "v1", "tb")
NS, NB = 4, 8
L = NS + NB
SECTOR = "R0:N12:S0"
ETA = (1.0, -1.0, -1.0, 1.0)
# independent spin-up pairs (o1 >= o2) and the pair classes of all ten
PAIRS = ((0, 0), (1, 1), (1, 0), (3, 0), (2, 1))
_CLASS = {(0, 0): 0, (3, 3): 0, (1, 1): 1, (2, 2): 1, (1, 0): 2, (2, 0): 2, (3, 1): 2, (3, 2): 2, (3, 0): 3,
          (2, 1): 4}
_FLIP = (1, 0, 3, 2)  # spin-down block: G_dn[i, j] = G_up[p(i), p(j)]
NCF = 32
N_OUT = 1 + len(PAIRS) * 2 * 2 * NCF

# sampling box of the training corpus (default_rng(seed); mu = U/2 + dmu)
BOX = {"U": (4.0, 10.0), "dmu": (-1.0, 1.0), "t": (0.8, 1.2), "tp": (-0.5, 0.0), "M": (0.0, 0.3),
       "e0": (-2.0, 2.0), "e1": (-2.0, 2.0), "m0": (-0.3, 0.3), "m1": (-0.3, 0.3), "v0": (0.2, 0.8),
       "v1": (0.2, 0.8), "tb": (0.0, 0.3)}

_model_done = set()


def _bath(s, i):
    return NS + 4 * s + i


def model_operators():
    """name -> (type, 0-based elements) of the family (ED::new_operator input)."""
    hop = lambda bonds: [(a, b, -1.0) for (a, b) in bonds] + [(a + L, b + L, -1.0) for (a, b) in bonds]
    ops = {"t": ("one-body", hop([(0, 1), (1, 3), (2, 3), (0, 2)])),
           "tp": ("one-body", hop([(0, 3), (1, 2)])),
           "U": ("interaction", [(i, i + L, 1.0) for i in range(NS)]),
           "mu": ("one-body", [(i, i, -1.0) for i in range(NS)] + [(i + L, i + L, -1.0) for i in range(NS)]),
           "M": ("one-body", [(i, i, ETA[i]) for i in range(NS)] + [(i + L, i + L, -ETA[i]) for i in range(NS)])}
    for s in range(2):
        b = [_bath(s, i) for i in range(NS)]
        ops["e%d" % s] = ("one-body", [(o, o, 1.0) for o in b] + [(o + L, o + L, 1.0) for o in b])
        ops["m%d" % s] = ("one-body", [(o, o, ETA[i]) for i, o in enumerate(b)] +
                          [(o + L, o + L, -ETA[i]) for i, o in enumerate(b)])
        ops["v%d" % s] = ("one-body", [(i, o, 1.0) for i, o in enumerate(b)] +
                          [(i + L, o + L, 1.0) for i, o in enumerate(b)])
    ops["tb"] = ("one-body", [(_bath(0, i), _bath(1, i), 1.0) for i in range(NS)] +
                 [(_bath(0, i) + L, _bath(1, i) + L, 1.0) for i in range(NS)])
    return ops


def ensure_model(name="plaq8b"):
    if name in _model_done:
        return name
    qcm.new_model(name, NS, NB)
    for op, (typ, el) in model_operators().items():
        qcm.new_operator(name, op, typ, [(r + 1, c + 1, v) for (r, c, v) in el])
    _model_done.add(name)
    return name


def values(p):
    """instance values of a 12-vector of parameters (PARAMS order)."""
    return {k: float(v) for k, v in zip(PARAMS, p)}


def sample(n, seed=5):
    """n parameter vectors drawn uniformly from BOX (numpy default_rng(seed),
    one column at a time in PARAMS order; mu = U/2 + dmu)."""
    rng = np.random.default_rng(seed)
    X = np.zeros((n, len(PARAMS)))
    for j, k in enumerate(PARAMS):
        lo, hi = BOX["dmu" if k == "mu" else k]
        X[:, j] = rng.uniform(lo, hi, n)
    X[:, 1] += 0.5 * X[:, 0]
    return X


def new_instance(p, label, model="plaq8b"):
    ensure_model(model)
    qcm.new_model_instanceC(model, values(p), SECTOR, label)
    return label


def exact_targets(label, ncf=NCF):
    """E0 and the spin-up fractions of the five pairs (truncated to ncf levels)
    of a solved (or solvable) instance: the surrogate's training target."""
    E0 = C.c_double()
    sec = C.create_string_buffer(64)
    check(lib().ed_ground_state_solve(int(label), C.byref(E0), sec, 64))
    n = len(PAIRS)
    o1 = (C.c_int * n)(*[a for (a, b) in PAIRS])
    o2 = (C.c_int * n)(*[b for (a, b) in PAIRS])
    a = np.zeros((2 * n, ncf))
    b = np.zeros((2 * n, ncf))
    ln = np.zeros(2 * n, dtype=np.int32)
    check(lib().ed_continued_fractions_pairs(int(label), 0, n, o1, o2, int(ncf),
                                             a.ctypes.data_as(C.POINTER(C.c_double)),
                                             b.ctypes.data_as(C.POINTER(C.c_double)),
                                             ln.ctypes.data_as(C.POINTER(C.c_int))))
    return pack(E0.value, a, b), ln


def pack(E0, a, b):
    """output vector: E0, then per pair k and part eh: a_0..a_{NCF-1}, b_0..b_{NCF-1}."""
    n = len(PAIRS)
    y = [np.array([E0])]
    for f in range(2 * n):
        y += [a[f], b[f]]
    return np.concatenate(y)


def unpack(y):
    n = len(PAIRS)
    E0 = float(y[0])
    ncf = (len(y) - 1) // (4 * n)
    body = np.asarray(y[1:]).reshape(2 * n, 2, ncf)
    return E0, body[:, 0, :].copy(), body[:, 1, :].copy()


def _fraction_set(a5, b5, flip):
    """the ten pairs' fractions (layout of ed_set_continued_fractions) of one
    spin block, from the five classes; flip: the spin-down block"""
    npair = NS * (NS + 1) // 2
    cap = a5.shape[1]
    A = np.zeros((2 * npair, cap))
    B = np.zeros((2 * npair, cap))
    for o1 in range(NS):
        for o2 in range(o1 + 1):
            p1, p2 = (_FLIP[o1], _FLIP[o2]) if flip else (o1, o2)
            key = (max(p1, p2), min(p1, p2))
            k = _CLASS[key]
            q = o1 * (o1 + 1) // 2 + o2
            for eh in range(2):
                A[2 * q + eh] = a5[2 * k + eh]
                B[2 * q + eh] = b5[2 * k + eh]
    return A, B


def set_predicted(label, p, y, model="plaq8b"):
    """instance `label` at parameters p whose solution is the prediction y:
    E0 and both spin blocks' continued fractions (b_n clipped at 0; a fraction
    ends at its first vanishing b_n)."""
    new_instance(p, label, model)
    E0, a5, b5 = unpack(y)
    b5 = np.maximum(b5, 0.0)
    for spin in (0, 1):
        A, B = _fraction_set(a5, b5, spin == 1)
        ln = np.zeros(A.shape[0], dtype=np.int32)
        for f in range(A.shape[0]):
            z = np.nonzero(B[f, 1:] <= 1e-12)[0]
            ln[f] = (z[0] + 1) if z.size else A.shape[1]
            if B[f, 0] <= 1e-12:
                ln[f] = 0
        check(lib().ed_set_continued_fractions(int(label), float(E0), SECTOR.encode(), spin, A.shape[0], A.shape[1],
                                               np.ascontiguousarray(A).ctypes.data_as(C.POINTER(C.c_double)),
                                               np.ascontiguousarray(B).ctypes.data_as(C.POINTER(C.c_double)),
                                               ln.ctypes.data_as(C.POINTER(C.c_int))))
    return label


# ---------------------------------------------------------------------------
# the lattice: square lattice of 2x2 plaquettes, NN t and diagonal tp
# ---------------------------------------------------------------------------
_POS = ((0, 0), (1, 0), (0, 1), (1, 1))


def superlattice(p, spin):
    """Superlattice spec of the square lattice (t, tp of p) tiled by the cluster,
    for one spin: the inter-cluster NN and diagonal bonds, and V0 = minus the
    cluster-only Weiss field (M eta_i for spin up, -M eta_i for spin down)."""
    t, tp, M = p[2], p[3], p[4]
    site = {xy: i for i, xy in enumerate(_POS)}
    els, vecs = [], []

    def index(R):
        if R not in vecs:
            vecs.append(R)
        return vecs.index(R)

    for i, (x, y) in enumerate(_POS):
        for (dx, dy, amp) in ((1, 0, t), (0, 1, t), (-1, 0, t), (0, -1, t),
                              (1, 1, tp), (1, -1, tp), (-1, 1, tp), (-1, -1, tp)):
            X, Y = x + dx, y + dy
            R = (X // 2, Y // 2, 0)
            if R == (0, 0, 0) or amp == 0.0:
                continue   # inside the cluster: part of the cluster model
            j = site[(X % 2, Y % 2)]
            els.append((i, j, index((float(R[0]), float(R[1]), 0.0)), -amp))
    sgn = 1.0 if spin == 0 else -1.0
    V0 = -np.diag([sgn * M * e for e in ETA]).astype(np.complex128)
    return {"V0": V0, "elements": els, "R": np.array(vecs)}


class OmegaGrid:
    """The fixed (frequency, wavevector) grid of the Potthoff functional."""

    def __init__(self, nk=32, w=(2.0, 20.0, 32, 32, 32)):
        self.w, self.weights = LT.legendre_frequency_grid(*w)
        self.kgrid = (2, (nk, nk, 1), 0.5)

    def omega(self, label, p):
        return LT.potthoff_functional(label, self.w, self.weights, superlattice(p, 0), self.kgrid,
                                      superlattice(p, 1))


# ---------------------------------------------------------------------------
# the reference's Newton-Raphson procedure, batched over independent problems
# ---------------------------------------------------------------------------
def nr_points(x, step):
    """_NR_Hessian's 1 + 2n + n(n-1)/2 evaluation points (pyqcm/vca.py:241-279)."""
    n = len(x)
    pts = [np.array(x, dtype=float)]
    for i in range(n):
        for s in (1.0, -1.0):
            q = np.array(x, dtype=float)
            q[i] += s * step[i]
            pts.append(q)
    for i in range(n):
        for j in range(i):
            q = np.array(x, dtype=float)
            q[i] += 0.707 * (2 * (i % 2) - 1) * step[i]
            q[j] += 0.707 * (2 * (j % 2) - 1) * step[j]
            pts.append(q)
    return pts


def nr_fit(pts, y):
    """quadratic fit of _NR_Hessian (pyqcm/vca.py:281-313): value, gradient and
    Hessian at the base point pts[0]."""
    x = pts[0]
    n = len(x)
    N = len(pts)
    A = np.empty((N, N))
    for k, q in enumerate(pts):
        A[k, 0] = 1.0
        A[k, 1:n + 1] = q
        m = n + 1
        for i in range(n):
            for j in range(i + 1):
                A[k, m] = q[i] * q[j] * (0.5 if i == j else 1.0)
                m += 1
    c = np.linalg.solve(A, np.asarray(y, dtype=float))
    val = c[0] + np.dot(c[1:n + 1], x)
    grad = c[1:n + 1].copy()
    H = np.zeros((n, n))
    m = n + 1
    for i in range(n):
        for j in range(i + 1):
            H[i, j] = H[j, i] = c[m]
            val += x[i] * x[j] * c[m] * (0.5 if i == j else 1.0)
            m += 1
    grad += H @ x
    return val, grad, H


def newton_raphson(evaluate, starts, step=0.01, accur=1e-4, gtol=1e-6, max_iter=30):
    """The reference's Newton-Raphson VCA (pyqcm/vca.py:318-410) run for several
    independent problems at once: evaluate(list of (problem index, x)) returns
    the functional at each, so that one call can solve all pending points in a
    batch.  Returns per problem (solution, iterations, evaluations, omega at the
    solution, last gradient, Hessian)."""
    P = len(starts)
    x = [np.array(s, dtype=float) for s in starts]
    n = len(x[0])
    step0 = np.full(n, step) if np.isscalar(step) else np.array(step, dtype=float)
    acc = np.full(n, accur) if np.isscalar(accur) else np.array(accur, dtype=float)
    st = [step0.copy() for _ in range(P)]
    it = [0] * P
    nev = [0] * P
    done = [False] * P
    grad = [None] * P
    hess = [None] * P
    while not all(done):
        active = [k for k in range(P) if not done[k]]
        req, owner = [], []
        for k in active:
            pts = nr_points(x[k], st[k])
            owner.append((k, len(req), len(pts)))
            req += [(k, q) for q in pts]
        vals = evaluate(req)
        for (k, off, m) in owner:
            pts = [q for (_, q) in req[off:off + m]]
            nev[k] += m
            it[k] += 1
            F, g, H = nr_fit(pts, vals[off:off + m])
            grad[k], hess[k] = g, H
            if np.linalg.norm(g) < gtol:
                done[k] = True
                continue
            dx = -np.linalg.solve(H, g)
            x[k] = x[k] + dx
            for i in range(n):
                st[k][i] = abs(dx[i])
                if st[k][i] > step0[i]:
                    st[k][i] = 2.0 * step0[i]
                if st[k][i] < 2.0 * acc[i]:
                    st[k][i] = 2.0 * acc[i]
            if np.all(np.abs(dx) <= acc) or it[k] >= max_iter:
                done[k] = True
    # final, converged value (vca.py:766)
    fin = evaluate([(k, x[k]) for k in range(P)])
    return [{"x": x[k], "iterations": it[k], "evaluations": nev[k] + 1, "omega": fin[k], "grad": grad[k],
             "hessian": hess[k]} for k in range(P)]



# ---------------------------------------------------------------------------
# the VCA sweep: exact and surrogate-proposed starts of the same procedure
# ---------------------------------------------------------------------------
def with_M(p, M):
    q = np.array(p, dtype=float)
    q[4] = float(M)
    return q


def exact_evaluator(problems, grid, label=7301):
    """evaluate(req) for newton_raphson: one full ED solve (ground state and
    Green function, the instance's GF_method) and the lattice functional per
    point; counts the solves in .solves"""
    def evaluate(req):
        out = []
        for (k, x) in req:
            p = with_M(problems[k], x[0])
            new_instance(p, label)
            out.append(grid.omega(label, p))
            evaluate.solves += 1
        return out
    evaluate.solves = 0
    return evaluate


def surrogate_omega(net, problems, Ms, grid, label=7401):
    """the functional of every problem at every M of Ms from the surrogate's
    predictions (E0 and fractions evaluated through the same lattice sums);
    shape (len(problems), len(Ms))"""
    X = np.array([with_M(p, M) for p in problems for M in Ms])
    Y = net.predict(X)
    om = np.zeros(len(X))
    for i in range(len(X)):
        set_predicted(label, X[i], Y[i])
        om[i] = grid.omega(label, X[i])
    return om.reshape(len(problems), len(Ms))


def surrogate_starts(net, problems, grid, Ms=None):
    """for each problem, the stationary point of the surrogate's functional:
    its lowest sampled minimum over Ms, refined by the parabola through the
    neighbouring samples (an end point is kept as it is)"""
    Ms = np.linspace(0.0, 0.3, 31) if Ms is None else np.asarray(Ms)
    om = surrogate_omega(net, problems, Ms, grid)
    starts = []
    for k in range(len(problems)):
        j = int(np.argmin(om[k]))
        if 0 < j < len(Ms) - 1:
            y0, y1, y2 = om[k, j - 1], om[k, j], om[k, j + 1]
            h = Ms[1] - Ms[0]
            den = y0 - 2 * y1 + y2
            x = Ms[j] + (0.5 * h * (y0 - y2) / den if den > 0 else 0.0)
        else:
            x = Ms[j]
        starts.append([float(x)])
    return starts, om


def vca_compare(net, problems, grid, M0=0.1, step=0.01, accur=1e-4):
    """the same Newton-Raphson VCA from the default start M0 (the plain path)
    and from the surrogate's proposals; every evaluation of both is a full ED
    solve.  Returns both results and the solve counts."""
    ev_plain = exact_evaluator(problems, grid)
    plain = newton_raphson(ev_plain, [[M0] for _ in problems], step=step, accur=accur)
    starts, om = surrogate_starts(net, problems, grid)
    ev_sur = exact_evaluator(problems, grid)
    sur = newton_raphson(ev_sur, starts, step=step, accur=accur)
    return {"plain": plain, "surrogate": sur, "starts": starts, "solves_plain": ev_plain.solves,
            "solves_surrogate": ev_sur.solves, "surrogate_omega": om}


def sweep_problems(n=6, seed=7):
    """the VCA sweep: U from 6 to 9, half filling (mu = U/2), t = 1, tp = -0.2, a
    fixed bath (e0 = -1.2, e1 = 1.2, v0 = v1 = 0.5, tb = 0.1, no bath field); M
    is the variational parameter"""
    Us = np.linspace(6.0, 9.0, n)
    return [np.array([U, U / 2, 1.0, -0.2, 0.1, -1.2, 1.2, 0.0, 0.0, 0.5, 0.5, 0.1]) for U in Us]

# ---------------------------------------------------------------------------
# training corpus, training, persistence
# ---------------------------------------------------------------------------
def generate_corpus(n, seed=5, start=0, label=7001, ncf=NCF, progress=None):
    """Rows start .. start + n of the corpus drawn by sample(seed): inputs X
    (n x 12), exact targets Y (n x N_OUT) from the ED (ground state, then the
    five pairs' fractions truncated to ncf levels) and the fractions' full
    lengths (n x 10, capped at ncf)."""
    X = sample(start + n, seed)[start:]
    Y = np.zeros((n, N_OUT))
    lens = np.zeros((n, 2 * len(PAIRS)), dtype=np.int32)
    for k in range(n):
        new_instance(X[k], label)
        Y[k], lens[k] = exact_targets(label, ncf)
        if progress and (k + 1) % progress == 0:
            print("corpus: %d / %d" % (k + 1, n), flush=True)
    qcm.erase_model_instance(label)
    return X, Y, lens


def output_weights(w_E0=64.0):
    """loss weight per output: E0 first; every fraction coefficient 1"""
    w = np.ones(N_OUT)
    w[0] = w_E0
    return w


def train(X, Y, epochs=(200, 100, 100), lrs=(1e-3, 3e-4, 1e-4), batch=2048, seed=5, width=256, depth=4,
          w_E0=64.0):
    """A new network trained on (X, Y): inputs and outputs standardized per
    column, loss weights from output_weights, Adam at each learning rate of
    lrs for the matching number of epochs.  Returns (net, epoch losses, device
    ms per step, wall seconds)."""
    import time
    net = MLP(X.shape[1], Y.shape[1], width, depth, seed=seed)
    xs = np.maximum(X.std(0), 1e-12)
    ys = np.maximum(Y.std(0), 1e-9)
    net.set_scaling(X.mean(0), xs, Y.mean(0), ys, output_weights(w_E0))
    hist, msl = [], []
    t0 = time.time()
    for e, lr in zip(epochs, lrs):
        net.set_optimizer(lr=lr)
        h, ms = net.train(X, Y, e, batch, seed=seed + len(hist))
        hist += list(h)
        msl.append(ms)
    return net, np.array(hist), float(np.mean(msl)), time.time() - t0


def save(path, net, X, Y, extra=None):
    xs = np.maximum(X.std(0), 1e-12)
    ys = np.maximum(Y.std(0), 1e-9)
    meta = dict(extra or {})
    np.savez_compressed(path, weights=net.get_weights().astype(np.float32), x_mean=X.mean(0), x_std=xs,
                        y_mean=Y.mean(0), y_std=ys, shape=np.array([net.n_in, net.n_out, net.width, net.depth]),
                        meta=np.array(repr(meta)))


def load(path, seed=0):
    d = np.load(path, allow_pickle=False)
    n_in, n_out, width, depth = (int(v) for v in d["shape"])
    net = MLP(n_in, n_out, width, depth, seed=seed)
    net.set_scaling(d["x_mean"], d["x_std"], d["y_mean"], d["y_std"], output_weights())
    net.set_weights(d["weights"].astype(np.float64))
    return net, str(d["meta"])


def _save_chunk(path, X, Y, lens):
    np.savez(path, X=X, E0=Y[:, 0], cf=Y[:, 1:].astype(np.float32), lens=lens.astype(np.int16))


def load_corpus(paths):
    """X, Y (E0 in float64, the fractions from float32) of the chunk files, in order."""
    Xs, Ys = [], []
    for p in paths:
        d = np.load(p, allow_pickle=False)
        Xs.append(d["X"])
        Ys.append(np.concatenate([d["E0"][:, None], d["cf"].astype(np.float64)], axis=1))
    return np.concatenate(Xs), np.concatenate(Ys)


def main(argv=None):
    """corpus: python -m pyqcm_amd.vca_surrogate corpus N START OUT.npz
    train:  python -m pyqcm_amd.vca_surrogate train OUT_WEIGHTS.npz CHUNK.npz ...
            [--epochs E1,E2,E3] [--wE0 W] [--holdout N]
            (epochs at learning rates 1e-3, 3e-4, 1e-4; the E0 loss weight;
            the last N rows kept out of training and reported)
    time:   python -m pyqcm_amd.vca_surrogate time WEIGHTS.npz [BATCH]
            (forward and forward + backward of the trained network on BATCH
            resident rows, and 20 training steps on random rows, for kernel
            statistics)"""
    import sys
    import time
    a = list(sys.argv[1:] if argv is None else argv)
    qcm.set_global_parameterC("Hamiltonian_format", "G")
    if a[0] == "corpus":
        n, start, out = int(a[1]), int(a[2]), a[3]
        t0 = time.time()
        X, Y, lens = generate_corpus(n, start=start, progress=max(1, n // 20))
        dt = time.time() - t0
        _save_chunk(out, X, Y, lens)
        print("corpus rows %d..%d: %.1f s, %.2f ms per solve" % (start, start + n, dt, 1e3 * dt / n), flush=True)
    elif a[0] == "train":
        opts = {"--epochs": "200,100,100", "--wE0": "64", "--holdout": "0"}
        rest = []
        i = 1
        while i < len(a):
            if a[i] in opts:
                opts[a[i]] = a[i + 1]
                i += 2
            else:
                rest.append(a[i])
                i += 1
        out, chunks = rest[0], rest[1:]
        X, Y = load_corpus(chunks)
        nh = int(opts["--holdout"])
        Xh, Yh = X[len(X) - nh:], Y[len(Y) - nh:]
        X, Y = X[:len(X) - nh], Y[:len(Y) - nh]
        epochs = tuple(int(v) for v in opts["--epochs"].split(","))
        w_E0 = float(opts["--wE0"])
        net, hist, ms, wall = train(X, Y, epochs=epochs, w_E0=w_E0)
        meta = {"samples": int(X.shape[0]), "final_loss": float(hist[-1]), "ms_per_step": ms, "train_s": wall,
                "epochs": list(epochs), "w_E0": w_E0}
        if nh:
            err = np.abs(net.predict(Xh) - Yh)
            meta["holdout"] = nh
            meta["holdout_E0_mae"] = float(err[:, 0].mean())
            meta["holdout_cf_mae_level0"] = float(err[:, 1:].reshape(nh, -1, 2, NCF)[:, :, :, 0].mean())
        save(out, net, X, Y, meta)
        print("trained on %d samples: loss %.4g -> %.4g, %.3f ms per step, %.1f s; %s" %
              (X.shape[0], hist[0], hist[-1], ms, wall, meta), flush=True)
    elif a[0] == "time":
        net, _ = load(a[1])
        B = int(a[2]) if len(a) > 2 else 4096
        ms_f, ms_fb = net.time(B, nrep=20)
        ff, ffb = net.flops(B)
        rng = np.random.default_rng(0)
        X = rng.normal(size=(B * 20, net.n_in))
        Y = rng.normal(size=(B * 20, net.n_out))
        _, ms_step = net.train(X, Y, 1, B, seed=1)
        print("batch %d: forward %.4f ms (%.1f TFLOP/s), forward + backward %.4f ms (%.1f TFLOP/s), "
              "training step %.4f ms" % (B, ms_f, ff / ms_f / 1e9, ms_fb, ffb / ms_fb / 1e9, ms_step), flush=True)
    else:
        raise SystemExit(main.__doc__)


if __name__ == "__main__":
    main()
