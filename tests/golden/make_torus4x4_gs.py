"""Writes tests/golden/torus4x4_u8_gs.json: the ground-state energy of the 4x4
periodic Hubbard cluster (t = 1, U = 8, mu = 4, sector N = 16, S = 0, dim
165 636 900) from the oracle's Lanczos (C + OpenMP, two vectors, eigenvector
by a second pass), with the residual |H psi - E0 psi| / |psi| of the Ritz
vector measured by the oracle's OpenMP product.

About 8 GB of host memory and a few minutes on 8 cores.  Run from the
repository root:  python tests/golden/make_torus4x4_gs.py
"""
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))

from oracle import oracle  # noqa: E402
from pyqcm_amd import workloads as W  # noqa: E402


def main():
    wl = W.torus(4, 4, U=8.0)
    m = oracle.Model(wl.L, wl.terms())
    dim = m.dim(8, 8)
    t0 = time.time()
    start = np.random.default_rng(2024).normal(size=dim)
    E0, gs, al, be = m.lanczos_gs(8, 8, start, max_iter=1000, accur=1e-10)
    del start
    t1 = time.time()
    gs /= np.linalg.norm(gs)
    hg = m.hv_factorized_omp(8, 8, gs, nrep=1)
    rq = float(gs @ hg)
    hg -= E0 * gs
    res = float(np.linalg.norm(hg))
    out = {
        "workload": "4x4 periodic Hubbard cluster t=1 U=8 mu=4, sector R0:N16:S0",
        "dim": int(dim),
        "E0": E0,
        "rayleigh_quotient": rq,
        "residual": res,
        "lanczos_iterations": int(len(al)),
        "stop_rule": "Ritz residual |z_k beta_k| < 1e-10",
        "seconds": round(t1 - t0, 1),
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "torus4x4_u8_gs.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
