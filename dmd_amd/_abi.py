"""ctypes view of include/dmdx.h and loading of the HIP library.

The product path is ``dmd_amd/libdmdx.so`` (HIP, gfx950).  ``load()`` raises
when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdmdx.so")

DMDX_OK = 0
STATUS_NAMES = {
    0: "DMDX_OK", -1: "DMDX_E_NAN", -2: "DMDX_E_BOND_EXPLOSION", -3: "DMDX_E_MISMATCH",
    -4: "DMDX_E_PARM", -5: "DMDX_E_GPU", -6: "DMDX_E_OOM", -7: "DMDX_E_INVALID",
    -8: "DMDX_E_BUSY", -9: "DMDX_E_IO",
}

FAITHFUL, CORRECTED = 0, 1
FP64, MIXED = 0, 1
COULOMB_NONE, COULOMB_CUTOFF, COULOMB_PME = 0, 1, 2
THERMO_NONE, THERMO_BERENDSEN, THERMO_NOSE_HOOVER, THERMO_ANDERSEN = 0, 1, 2, 3
BARO_NONE, BARO_BERENDSEN, BARO_ANDERSEN = 0, 1, 2

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class SystemDesc(C.Structure):
    """Mirror of ``dmdx_system_desc`` (include/dmdx.h)."""

    _fields_ = [
        ("n_atoms", C.c_int64),
        ("box", C.c_double * 3),
        ("pos_x", _dp), ("pos_y", _dp), ("pos_z", _dp),
        ("vel_x", _dp), ("vel_y", _dp), ("vel_z", _dp),
        ("masses", _dp), ("charges", _dp), ("atom_types", _ip),
        ("use_lj", C.c_int32), ("lj_n_types", C.c_int32),
        ("lj_sigma", _dp), ("lj_epsilon", _dp), ("lj_cutoff", C.c_double),
        ("list_skin", C.c_double), ("rebuild_interval", C.c_int32),
        ("step_since_rebuild", C.c_int32),
        ("n_bonds", C.c_int64), ("bond_i", _ip), ("bond_j", _ip),
        ("bond_k", _dp), ("bond_r0", _dp),
        ("n_angles", C.c_int64), ("angle_i", _ip), ("angle_j", _ip), ("angle_k", _ip),
        ("angle_k_theta", _dp), ("angle_theta0", _dp),
        ("n_dihedrals", C.c_int64), ("dih_i", _ip), ("dih_j", _ip), ("dih_k", _ip),
        ("dih_l", _ip), ("dih_periodicity", _ip), ("dih_k_phi", _dp), ("dih_phi0", _dp),
        ("n_impropers", C.c_int64), ("imp_i", _ip), ("imp_j", _ip), ("imp_k", _ip),
        ("imp_l", _ip), ("imp_k_phi", _dp), ("imp_phi0", _dp),
        ("coulomb_type", C.c_int32), ("coulomb_cutoff", C.c_double),
        ("ewald_coeff", C.c_double), ("pme_grid", C.c_int32 * 3), ("pme_order", C.c_int32),
        ("n_excl", C.c_int64), ("excl_i", _ip), ("excl_j", _ip),
        ("dt", C.c_double),
        ("thermostat", C.c_int32), ("thermo_temperature", C.c_double),
        ("thermo_tau", C.c_double),
        ("barostat", C.c_int32), ("baro_pressure", C.c_double), ("baro_tau", C.c_double),
        ("baro_compressibility", C.c_double),
        ("physics", C.c_int32), ("precision", C.c_int32),
        ("n_constraints", C.c_int64), ("con_i", _ip), ("con_j", _ip),
        ("con_distance", _dp), ("con_tolerance", C.c_double),
        ("andersen_frequency", C.c_double), ("andersen_calls", C.c_int64),
    ]


class DmdxError(RuntimeError):
    """A non-zero status from the C-ABI (mirrors ForceException, src/core/error.h:14-21)."""

    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


_lib = None
_lock = threading.Lock()


def load(path: str | None = None):
    """Load libdmdx.so once.  Raises if the HIP library has not been built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        p = path or os.environ.get("DMDX_LIBRARY", LIB_PATH)
        if not os.path.exists(p):
            raise RuntimeError(
                f"HIP library not found at {p}; build it with `python -m dmd_amd.build` "
                "(there is no CPU fallback)")
        lib = C.CDLL(p, mode=C.RTLD_GLOBAL)
        lib.dmdx_last_error.restype = C.c_char_p
        _lib = lib
        return lib


def check(status: int) -> None:
    if status != DMDX_OK:
        msg = _lib.dmdx_last_error() if _lib is not None else b""
        raise DmdxError(status, (msg or b"").decode(errors="replace"))
