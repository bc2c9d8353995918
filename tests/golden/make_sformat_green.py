"""Writes tests/golden/s_format_c2v_green.json: for the C2v cluster of
s_format_c2v.json (sector N = 12, S = 0, 853 776 states), the spin-up Green
function of the four symmetric d orbitals u = (1, +-1, +-1, +-1) / 2 on a
grid of complex frequencies away from the real axis, from the oracle:
ground state by Lanczos (accuracy 1e-12 on E0), then continued fractions of
u.c^dagger |gs> and u.c |gs> with 400 steps (on this grid the fraction of
the first 300 steps agrees to rounding; the script prints the difference).

About 6 minutes on 8 cores.  Run from the repository root:
    python tests/golden/make_sformat_green.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
IRREPS = [[1, 1, 1, 1], [1, -1, 1, -1], [1, 1, -1, -1], [1, -1, -1, 1]]
Z = np.concatenate([1j * np.array([0.25, 0.5, 1.0, 2.0, 4.0, 8.0]), np.linspace(-12.0, 12.0, 9) + 1.0j])


def model():
    fix = json.load(open(os.path.join(HERE, "s_format_c2v.json")))
    P = fix["parameters"]
    terms = [("one-body", P[op + "_1"], [(i - 1, j - 1, float(v)) for (i, j, v) in d["elements"]])
             for op, d in fix["operators"].items() if d["type"] == "one-body" and P.get(op + "_1", 0.0) != 0.0]
    terms.append(("interaction", P["U"], [(i, i + 12, 1.0) for i in range(4)]))
    terms.append(("one-body", P["mu"], [(i, i, -1.0) for i in range(4)] + [(i + 12, i + 12, -1.0) for i in range(4)]))
    return oracle.Model(12, terms)


def green(m, gs, E0, u, steps, short):
    """G_u(z) from `steps` Lanczos steps per part, and from the first `short`
    of them (the difference measures the convergence)"""
    g = np.zeros(len(Z), complex)
    g2 = np.zeros(len(Z), complex)
    for pm in (1, -1):
        phi = None
        for i in range(4):
            y = oracle.create_or_destroy(12, pm, i, 6, 6, gs, u[i] / 2.0)
            phi = y if phi is None else phi + y
        nrm = float(phi @ phi)
        a, b = m.lanczos_green(6 + pm, 6, phi / np.sqrt(nrm), accur_cf=1e-300, max_iter_cf=steps)
        a, b = oracle.cf_transform(a, b, E0, nrm, pm > 0)
        g += oracle.cf_eval(a, b, Z)
        g2 += oracle.cf_eval(a[:short], b[:short + 1], Z)
    return g, g2


def main():
    m = model()
    dim = m.dim(6, 6)
    E0, gs, _, _ = m.lanczos_gs(6, 6, np.random.default_rng(7).normal(size=dim), max_iter=1000, accur=1e-12)
    gs = gs / np.linalg.norm(gs)
    res = float(np.linalg.norm(m.hv_factorized_omp(6, 6, gs) - E0 * gs))
    out = {"z": [[float(z.real), float(z.imag)] for z in Z], "E0": E0, "residual": res, "G": []}
    for u in IRREPS:
        g, g2 = green(m, gs, E0, u, 400, 300)
        print(u, "change 300 -> 400 steps:", np.abs(g - g2).max(), flush=True)
        out["G"].append([[float(x.real), float(x.imag)] for x in g])
    out["irreps"] = IRREPS
    with open(os.path.join(HERE, "s_format_c2v_green.json"), "w") as f:
        json.dump(out, f, indent=0)
    print("E0", E0, "residual", res)


if __name__ == "__main__":
    main()
