"""Writes tests/golden/s_format_complex.json from the reference's own test
output tests/test_diagonalization/S_format_complex.ref (read as text): an
8-site open chain ('clus', no bath) whose only active one-body term is the
imaginary hopping 'ti' (hopping_operator with tau = 2: element +i on each
bond (j, j + 1), src_qcm/lattice_model.cpp:56-59, :1003-1016), with U = 1e-9
and mu = 0 in sector R0:N8:S0; the fixture keeps the parameter values, the
ground-state energy and the printed Q-matrix (8 poles, 8 complex residues
each, rows of "w L M" in the reference's layout: pole, then residue per site).

Run from the repository root:  python tests/golden/make_sformat_complex_fixture.py
"""
import json
import os
import re

SRC = "/root/reference/tests/test_diagonalization/S_format_complex.ref"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "s_format_complex.json")


def _cplx(tok):
    re_, im_ = tok.strip("()").split(",")
    return [float(re_), float(im_)]


def main():
    lines = open(SRC).read().splitlines()
    params = {}
    for ln in lines:
        m = re.match(r'set_parameter\("(\w+)", ([-+0-9.eE]+)\)$', ln)
        if m:
            params[m.group(1)] = float(m.group(2))
    i0 = next(i for i, ln in enumerate(lines) if ln.startswith("GS_energy:"))
    E0 = float(lines[i0].split()[1])
    sector = lines[i0].split()[-1].rsplit(":", 1)[0]
    iw = next(i for i in range(i0, len(lines)) if lines[i].startswith("w\t"))
    L, M = (int(x) for x in lines[iw].split()[1:])
    poles, residues = [], []
    for k in range(M):
        tok = lines[iw + 1 + k].split("\t")
        assert len(tok) == L + 1
        poles.append(float(tok[0]))
        residues.append([_cplx(t) for t in tok[1:]])
    out = {
        "source": "tests/test_diagonalization/S_format_complex.ref (reference test output)",
        "model": {"n_sites": 8, "n_bath": 0, "chain": "open, bonds (j, j + 1), element +i of operator ti"},
        "parameters": params,
        "sector": sector,
        "E0": E0,
        "qmatrix": {"L": L, "M": M, "poles": poles, "residues": residues,
                    "layout": "residues[i][a] = (re, im) of Q(a, i); G(r, c) = sum_i Q(c, i) conj(Q(r, i)) / (z - e_i)"},
    }
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
