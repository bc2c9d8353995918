"""MI355X-native exact-diagonalization impurity solver with pyqcm's ED API.

``pyqcm_amd.qcm`` mirrors the ED functions of pyqcm's compiled module ``qcm``;
the work runs in libpyqcm_ed.so (HIP kernels for gfx950 + C++ host driver).
"""
from . import qcm, workloads  # noqa: F401
from ._lib import EDError, LIB_PATH, device_count, lib  # noqa: F401

__version__ = "0.1.0"
