"""Solved instances written to and read from HDF5 files in the reference's
layout (ED::write_instance_hdf5 / read_instance_hdf5):
  * round trip for the three Green-function methods, with up/down spin blocks,
    and for continued fractions held in the batched solver's device store: the
    instance read back gives the same ground-state energy, sector string,
    Green function, integrated Green function and Potthoff functional;
  * the file's structure (group, attribute and dataset names, their types) as
    the reference's writers lay it out, read with the HDF5 tools' h5dump;
  * errors: no wavefunction in a read instance, a missing group, an unknown
    label; a parameter that differs from the file's is reported, as there."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from pyqcm_amd import _lib, qcm
from pyqcm_amd._lib import EDError

pytestmark = pytest.mark.gpu

_label = [13000]
_models = set()


@pytest.fixture(scope="module", autouse=True)
def _device():
    _lib.require_device()
    try:
        qcm.hdf5_version()
    except EDError:
        pytest.skip("no HDF5 C library on this machine")
    qcm.set_global_parameterC("Hamiltonian_format", "G")
    qcm.set_global_parameterC("GF_method", "L")
    yield
    qcm.set_global_parameterC("GF_method", "L")


def _chain_ops(L):
    return {
        "t": ("one-body", [(i, i + 1, -1.0) for i in range(L - 1)] + [(i + L, i + 1 + L, -1.0) for i in range(L - 1)]),
        "U": ("interaction", [(i, i + L, 1.0) for i in range(L)]),
        "mu": ("one-body", [(i, i, -1.0) for i in range(L)] + [(i + L, i + L, -1.0) for i in range(L)]),
    }


def _instance(L, values, sector, method, solve=True):
    name = "h5_chain%d" % L
    if name not in _models:
        qcm.new_model(name, L, 0)
        for op, (typ, el) in _chain_ops(L).items():
            qcm.new_operator(name, op, typ, [(r + 1, c + 1, v) for (r, c, v) in el])
        _models.add(name)
    qcm.set_global_parameterC("GF_method", method)
    _label[0] += 1
    lab = _label[0]
    qcm.new_model_instanceC(name, values, sector, lab)
    if solve:
        qcm.ground_state_solve(lab)
    return lab


_Z = np.array([0.4 + 0.1j, -1.3 + 0.05j, 2.2j, 3.0 + 0.3j])


def _snapshot(lab, mixing):
    out = {"gs": qcm.ground_state_solve(lab), "omega": qcm.Potthoff_functionalC(lab)}
    for sd in ((0, 1) if mixing else (0,)):
        out["G%d" % sd] = qcm.Green_function_grid(_Z, spin_down=sd, label=lab)
        out["M%d" % sd] = qcm.Green_function_average(lab, sd)
    return out


def _same(a, b):
    assert a["gs"][1] == b["gs"][1]
    assert a["gs"][0] == b["gs"][0]
    assert a["omega"] == b["omega"]
    for k in a:
        if k[0] in "GM":
            assert np.abs(a[k] - b[k]).max() <= 1e-14 * max(1.0, np.abs(a[k]).max()), (k, np.abs(a[k] - b[k]).max())


@pytest.mark.parametrize("method", ["L", "F", "M"])
@pytest.mark.parametrize("case", ["normal", "updown"])
def test_round_trip(tmp_path, method, case):
    if case == "normal":
        L, vals, sec = 4, {"t": 1.0, "U": 4.0, "mu": 2.0}, "R0:N4:S0"
    else:
        L, vals, sec = 3, {"t": 1.0, "U": 4.0, "mu": 2.0}, "R0:N3:S1/R0:N3:S-1"
    lab = _instance(L, vals, sec, method)
    ref = _snapshot(lab, case == "updown")
    f = str(tmp_path / "sol.h5")
    qcm.write_instance_to_hdf5(f, "clus0", lab)
    qcm.write_instance_to_hdf5(f, "clus0", lab)  # an existing group is replaced
    other = _instance(L, vals, sec, "L", solve=False)
    qcm.read_instance_from_hdf5(f, "clus0", other)
    _same(ref, _snapshot(other, case == "updown"))
    # the read instance has no wavefunction
    with pytest.raises(EDError) as e:
        qcm.cluster_averages(other)
    assert e.value.code == 1
    # written again from the read instance, read into a third: the same
    qcm.write_instance_to_hdf5(f, "again", other)
    third = _instance(L, vals, sec, "L", solve=False)
    qcm.read_instance_from_hdf5(f, "again", third)
    _same(ref, _snapshot(third, case == "updown"))
    for x in (lab, other, third):
        qcm.erase_model_instance(x)


def test_round_trip_from_the_batched_store(tmp_path):
    pts = [(3.0, 1.5), (5.0, 2.5)]
    labs = [_instance(4, {"t": 1.0, "U": U, "mu": mu}, "R0:N4:S0", "F", solve=False) for (U, mu) in pts]
    qcm.set_global_parameterC("GF_method", "F")
    qcm.solve_batch(labs)
    f = str(tmp_path / "batch.h5")
    for k, lab in enumerate(labs):
        qcm.write_instance_to_hdf5(f, "point%d" % k, lab)
    for k, (lab, (U, mu)) in enumerate(zip(labs, pts)):
        ref = _snapshot(lab, False)
        other = _instance(4, {"t": 1.0, "U": U, "mu": mu}, "R0:N4:S0", "F", solve=False)
        qcm.read_instance_from_hdf5(f, "point%d" % k, other)
        _same(ref, _snapshot(other, False))
        qcm.erase_model_instance(other)
    for lab in labs:
        qcm.erase_model_instance(lab)


def _h5dump():
    for p in (shutil.which("h5dump"), "/opt/conda/bin/h5dump"):
        if p and os.path.exists(p):
            return p
    pytest.skip("h5dump not available")


@pytest.mark.parametrize("method", ["L", "F", "M"])
def test_file_layout(tmp_path, method):
    """names and types of the reference's writers: instance attributes
    GS_energy (double), GS_sectors, GF_format (variable-length strings),
    mixing, complex_HS, nstates (int); group params; state_i with sec,
    energy, weight, gf_format and group gf; per method the Q_matrix_set,
    continued_fraction_set or mcf_set layout"""
    lab = _instance(4, {"t": 1.0, "U": 4.0, "mu": 2.0}, "R0:N4:S0", method)
    f = str(tmp_path / "layout.h5")
    qcm.write_instance_to_hdf5(f, "inst", lab)
    txt = subprocess.run([_h5dump(), "-H", f], check=True, capture_output=True, text=True).stdout
    flat = " ".join(txt.split())
    for name, typ in (("GS_energy", "H5T_IEEE_F64LE"), ("mixing", "H5T_STD_I32LE"), ("complex_HS", "H5T_STD_I32LE"),
                      ("nstates", "H5T_STD_I32LE"), ("energy", "H5T_IEEE_F64LE"), ("weight", "H5T_IEEE_F64LE"),
                      ("t", "H5T_IEEE_F64LE"), ("U", "H5T_IEEE_F64LE"), ("mu", "H5T_IEEE_F64LE")):
        assert 'ATTRIBUTE "%s" { DATATYPE %s' % (name, typ) in flat, name
    for name in ("GS_sectors", "GF_format", "sec", "gf_format"):
        assert 'ATTRIBUTE "%s" { DATATYPE H5T_STRING { STRSIZE H5T_VARIABLE;' % name in flat, name
    for g in ('GROUP "inst"', 'GROUP "params"', 'GROUP "state_0"', 'GROUP "gf"'):
        assert g in flat, g
    if method == "L":
        for s in ('ATTRIBUTE "nblocks"', 'GROUP "block_0"', 'ATTRIBUTE "L"', 'ATTRIBUTE "M"',
                  'DATASET "e" { DATATYPE H5T_IEEE_F64LE', 'GROUP "v"', 'ATTRIBUTE "r"', 'ATTRIBUTE "c"',
                  'DATASET "data" { DATATYPE H5T_IEEE_F64LE'):
            assert s in flat, s
    elif method == "F":
        assert 'ATTRIBUTE "is_complex" { DATATYPE H5T_STD_I32LE' in flat
        for a in range(4):
            for b in range(a + 1):
                assert 'GROUP "r0_a%d_b%d"' % (a, b) in flat
        assert 'GROUP "r0_a0_b1"' not in flat
        for s in ('GROUP "e"', 'GROUP "h"', 'DATASET "a" { DATATYPE H5T_IEEE_F64LE',
                  'DATASET "b" { DATATYPE H5T_IEEE_F64LE'):
            assert s in flat, s
    else:
        for s in ('ATTRIBUTE "nblocks"', 'GROUP "block_0"', 'GROUP "e"', 'GROUP "h"', 'ATTRIBUTE "p"',
                  'ATTRIBUTE "floors"', 'GROUP "A_0"', 'GROUP "B_0"', 'GROUP "W"',
                  'DATASET "re" { DATATYPE H5T_IEEE_F64LE', 'DATASET "im" { DATATYPE H5T_IEEE_F64LE'):
            assert s in flat, s
    qcm.erase_model_instance(lab)


def test_errors_and_parameter_warning(tmp_path, capfd):
    lab = _instance(4, {"t": 1.0, "U": 4.0, "mu": 2.0}, "R0:N4:S0", "L")
    f = str(tmp_path / "e.h5")
    qcm.write_instance_to_hdf5(f, "inst", lab)
    with pytest.raises(EDError) as e:
        qcm.write_instance_to_hdf5(f, "x", 999999)
    assert e.value.code == 7
    other = _instance(4, {"t": 1.0, "U": 4.5, "mu": 2.0}, "R0:N4:S0", "L", solve=False)
    with pytest.raises(EDError) as e:
        qcm.read_instance_from_hdf5(f, "missing", other)
    assert e.value.code == 1
    with pytest.raises(EDError) as e:
        qcm.read_instance_from_hdf5(str(tmp_path / "nofile.h5"), "inst", other)
    assert e.value.code == 1
    capfd.readouterr()
    qcm.read_instance_from_hdf5(f, "inst", other)
    err = capfd.readouterr().err
    assert "WARNING : The value of U from the HDF5 solution" in err, err
    for x in (lab, other):
        qcm.erase_model_instance(x)
