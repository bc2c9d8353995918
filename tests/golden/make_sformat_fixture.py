"""Writes tests/golden/s_format_c2v.json from the reference's own test output
tests/test_diagonalization/S_format.ref (read as text): the 2x2 cluster with
8 bath orbitals and C2v symmetry ('clus', 4 sites + 8 bath orbitals in the
irreducible-representation basis), its cluster operators (bath energies eb1-8,
hybridizations tb1-8), the parameter values of the run, the ground-state
energy of sector R0:N12:S0 and the four Lehmann blocks of the spin-up Green
function (one symmetric d orbital per irreducible representation: poles and
amplitudes, as printed).

Run from the repository root:  python tests/golden/make_sformat_fixture.py
"""
import ast
import json
import os
import re

SRC = "/root/reference/tests/test_diagonalization/S_format.ref"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "s_format_c2v.json")


def main():
    lines = open(SRC).read().splitlines()
    ops = {}
    params = {}
    for ln in lines:
        m = re.match(r"new_cluster_operator\('clus', '(\w+)', '([\w-]+)', (\[.*\])\)$", ln)
        if m:
            ops[m.group(1)] = {"type": m.group(2), "elements": ast.literal_eval(m.group(3))}
        m = re.match(r'set_parameter\("(\w+)", ([-+0-9.eE]+)\)$', ln)
        if m:
            params[m.group(1)] = float(m.group(2))
    # the first solution: its energy and Lehmann blocks
    i0 = next(i for i, ln in enumerate(lines) if ln.startswith("GS_energy:"))
    E0 = float(lines[i0].split()[1])
    sector = lines[i0].split()[-1].rsplit(":", 1)[0]
    blocks = []
    i = i0
    while i < len(lines) and not (lines[i].startswith("GS_energy:") and i > i0):
        if lines[i].startswith("w\t"):
            L, M = (int(x) for x in lines[i].split()[1:])
            rows = [[float(x) for x in lines[i + 1 + k].split()] for k in range(M)]
            assert all(len(r) == L + 1 for r in rows)
            blocks.append(rows)
            i += M
        i += 1
    out = {
        "source": "tests/test_diagonalization/S_format.ref (reference test output)",
        "model": {"n_sites": 4, "n_bath": 8},
        "operators": ops,
        "parameters": params,
        "sector": sector,
        "E0": E0,
        "blocks": blocks,
    }
    with open(OUT, "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", OUT, "blocks", [len(b) for b in blocks], "E0", E0, sector)


if __name__ == "__main__":
    main()
