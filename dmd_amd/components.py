"""DMD's ForceComponent / ForceEngine interface over the C-ABI (dmdx_component_*).

Mirrors src/core/cell.h, src/core/system_data.h, src/force/force_component.h,
src/force/force_engine.h and the component headers under src/force/, with the
same names and argument meanings.  ``compute`` adds to the force arrays in
place and, since Python has no ``double&``, returns ``energy_out`` plus the
component's energy.  Non-zero C-ABI statuses raise ``ForceException``
(src/core/error.h:14-21).
"""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass, field

import numpy as np

from . import _abi

ForceException = _abi.DmdxError

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_cp = C.c_void_p


def _setup(lib):
    if getattr(lib, "_comp_ready", False):
        return
    P = C.POINTER(C.c_void_p)
    lib.dmdx_lj_create.argtypes = [C.c_int32, _dp, _dp, C.c_double, P]
    lib.dmdx_lj_set_atom_types.argtypes = [_cp, _ip, C.c_int64]
    lib.dmdx_lj_get_step_since_rebuild.argtypes = [_cp, _ip]
    lib.dmdx_lj_set_step_since_rebuild.argtypes = [_cp, C.c_int32]
    lib.dmdx_coulomb_direct_create.argtypes = [C.c_double, C.c_double, P]
    lib.dmdx_coulomb_exclusion_create.argtypes = [P]
    lib.dmdx_coulomb_exclusion_set_exclusions.argtypes = [_cp, _ip, _ip, C.c_int64]
    lib.dmdx_coulomb_exclusion_set_ewald_coefficient.argtypes = [_cp, C.c_double]
    lib.dmdx_pme_create.argtypes = [C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_int32,
                                    C.c_int32, P]
    lib.dmdx_pme_theta4.argtypes = [C.c_double, _dp]
    lib.dmdx_pme_dtheta4.argtypes = [C.c_double, _dp]
    lib.dmdx_pme_self_energy.argtypes = [C.c_double, _dp, C.c_int64, C.c_double, _dp]
    lib.dmdx_harmonic_bond_create.argtypes = [C.c_int64, _ip, _ip, _dp, _dp, P]
    lib.dmdx_harmonic_angle_create.argtypes = [C.c_int64, _ip, _ip, _ip, _dp, _dp, P]
    lib.dmdx_periodic_dihedral_create.argtypes = [C.c_int64, _ip, _ip, _ip, _ip, _ip, _dp, _dp, P]
    lib.dmdx_harmonic_improper_create.argtypes = [C.c_int64, _ip, _ip, _ip, _ip, _dp, _dp, P]
    lib.dmdx_component_set_charges.argtypes = [_cp, _dp, C.c_int64]
    lib.dmdx_component_rebuild_pair_list.argtypes = [_cp, _dp, _dp, _dp, C.c_int64, _dp]
    lib.dmdx_component_compute.argtypes = [_cp, _dp, _dp, _dp, C.c_int64, _dp, _dp, _dp, _dp, _dp]
    lib.dmdx_component_rebuild_pair_list_cell.argtypes = [_cp, _dp, _dp, _dp, C.c_int64, _dp,
                                                          C.c_int32]
    lib.dmdx_component_compute_cell.argtypes = [_cp, _dp, _dp, _dp, C.c_int64, _dp, C.c_int32,
                                                _dp, _dp, _dp, _dp]
    lib.dmdx_component_compute_noalloc.argtypes = [_cp, _cp, _cp, _cp, C.c_int64, _dp, _cp, _cp,
                                                   _cp, _cp]
    lib.dmdx_component_name.argtypes = [_cp, C.c_char_p, C.c_int64]
    lib.dmdx_component_set_mode.argtypes = [_cp, C.c_int32, C.c_int32]
    lib.dmdx_component_destroy.argtypes = [_cp]
    lib._comp_ready = True


def _lib():
    lib = _abi.load()
    _setup(lib)
    return lib


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _p(a):
    return a.ctypes.data_as(_dp if a.dtype == np.float64 else _ip)


# ------------------------------------------------------------------ core

class CellType(enum.IntEnum):
    """src/core/cell.h:3-8 (values of DMDX_CELL_*)."""
    orthorhombic = 0
    triclinic = 1
    hexagonal = 2
    rhombic_dodecahedron = 3


class Cell:
    """src/core/cell.h:10-23: Cell(a, b, c) is orthorhombic; Cell.from_matrix
    (the Cell(m[9], type) constructor) takes any row-major matrix and type.
    The minimum image is the identity unless orthorhombic
    (src/core/cell.cpp:26-38); the volume is the determinant."""

    def __init__(self, a: float = 1.0, b: float = 1.0, c: float = 1.0):
        self.matrix = np.zeros(9)
        self.matrix[0], self.matrix[4], self.matrix[8] = a, b, c
        self.type = CellType.orthorhombic

    @classmethod
    def from_matrix(cls, m, t=CellType.orthorhombic):
        cell = cls()
        cell.matrix = np.array(m, dtype=np.float64).reshape(9)
        cell.type = CellType(t)
        return cell

    def volume(self) -> float:
        m = self.matrix
        return (m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6])
                + m[2] * (m[3] * m[7] - m[4] * m[6]))

    def scale(self, sx, sy, sz):
        self.matrix[0] *= sx
        self.matrix[4] *= sy
        self.matrix[8] *= sz


class SystemData:
    """src/core/system_data.h:7-35 (numpy SoA)."""

    def __init__(self, n: int = 0):
        self.n_atoms = n
        for name in ("pos_x", "pos_y", "pos_z", "vel_x", "vel_y", "vel_z", "forces_x",
                     "forces_y", "forces_z", "masses", "charges"):
            setattr(self, name, np.zeros(n))
        self.atom_types = np.zeros(n, np.int32)
        self.potential_energy = 0.0
        self.step = 0
        self.time = 0.0
        self.excl_i = np.zeros(0, np.int32)
        self.excl_j = np.zeros(0, np.int32)

    def x(self):
        return self.pos_x

    def y(self):
        return self.pos_y

    def z(self):
        return self.pos_z

    def fx(self):
        return self.forces_x

    def fy(self):
        return self.forces_y

    def fz(self):
        return self.forces_z


# ------------------------------------------------------------ components

class ForceComponent:
    """src/force/force_component.h:8-22 over a dmdx_component handle."""

    _h = None

    def __init__(self, physics=_abi.FAITHFUL, precision=_abi.FP64):
        if physics != _abi.FAITHFUL or precision != _abi.FP64:
            self.set_mode(physics, precision)

    def __del__(self):
        if self._h:
            try:
                _lib().dmdx_component_destroy(self._h)
            except Exception:
                pass
            self._h = None

    def set_mode(self, physics, precision=_abi.FP64):
        _abi.check(_lib().dmdx_component_set_mode(self._h, int(physics), int(precision)))

    def name(self) -> str:
        buf = C.create_string_buffer(64)
        _abi.check(_lib().dmdx_component_name(self._h, buf, 64))
        return buf.value.decode()

    def compute(self, pos_x, pos_y, pos_z, cell: Cell, forces_x, forces_y, forces_z,
                energy_out: float = 0.0) -> float:
        """Adds this component's forces into forces_* (float64 arrays, in
        place) and returns energy_out + its energy."""
        x, y, z = _f64(pos_x), _f64(pos_y), _f64(pos_z)
        for f in (forces_x, forces_y, forces_z):
            if not (isinstance(f, np.ndarray) and f.dtype == np.float64 and f.flags.c_contiguous):
                raise TypeError("forces must be contiguous float64 numpy arrays (updated in place)")
        n = len(x)
        if not (len(y) == len(z) == len(forces_x) == len(forces_y) == len(forces_z) == n):
            raise ForceException(-3, "span sizes differ")
        e = C.c_double(energy_out)
        _abi.check(_lib().dmdx_component_compute_cell(
            self._h, _p(x), _p(y), _p(z), n, _p(_f64(cell.matrix)), int(cell.type), _p(forces_x),
            _p(forces_y), _p(forces_z), C.byref(e)))
        return e.value

    def compute_noalloc(self, n: int, x_ptr: int, y_ptr: int, z_ptr: int, cell: Cell,
                        fx_ptr: int, fy_ptr: int, fz_ptr: int, energy_ptr: int):
        """dmdx_component_compute_noalloc: the same compute on DEVICE arrays
        (integer device addresses of n doubles each, and of one double for
        the energy), with no device allocation by the library.  Orthorhombic
        cells only, as the C entry point."""
        if cell.type != CellType.orthorhombic:
            raise ForceException(-4, "compute_noalloc takes orthorhombic cells only")
        _abi.check(_lib().dmdx_component_compute_noalloc(
            self._h, x_ptr, y_ptr, z_ptr, int(n), _p(_f64(cell.matrix)), fx_ptr, fy_ptr, fz_ptr,
            energy_ptr))

    def _rebuild(self, pos_x, pos_y, pos_z, cell):
        x, y, z = _f64(pos_x), _f64(pos_y), _f64(pos_z)
        _abi.check(_lib().dmdx_component_rebuild_pair_list_cell(
            self._h, _p(x), _p(y), _p(z), len(x), _p(_f64(cell.matrix)), int(cell.type)))


@dataclass
class LennardJonesParams:
    """src/force/lennard_jones.h:7-11"""
    n_types: int = 0
    sigma: list = field(default_factory=list)
    epsilon: list = field(default_factory=list)


class LennardJones(ForceComponent):
    """src/force/lennard_jones.h:13-47"""

    def __init__(self, params: LennardJonesParams, cutoff: float = 1.2, **mode):
        sig, eps = _f64(params.sigma).ravel(), _f64(params.epsilon).ravel()
        h = C.c_void_p()
        _abi.check(_lib().dmdx_lj_create(int(params.n_types), _p(sig), _p(eps), float(cutoff),
                                         C.byref(h)))
        self._h = h
        super().__init__(**mode)

    def set_atom_types(self, types):
        t = _i32(types)
        _abi.check(_lib().dmdx_lj_set_atom_types(self._h, _p(t), len(t)))

    def rebuild_pair_list(self, pos_x, pos_y, pos_z, cell: Cell):
        self._rebuild(pos_x, pos_y, pos_z, cell)

    def step_since_rebuild(self) -> int:
        v = C.c_int32()
        _abi.check(_lib().dmdx_lj_get_step_since_rebuild(self._h, C.byref(v)))
        return v.value

    def set_step_since_rebuild(self, s: int):
        _abi.check(_lib().dmdx_lj_set_step_since_rebuild(self._h, int(s)))


@dataclass
class CoulombDirectParams:
    """src/force/coulomb_direct.h:6-9"""
    cutoff: float = 1.2
    ewald_coefficient: float = 3.5


class _Charged(ForceComponent):
    def set_charges(self, charges):
        q = _f64(charges)
        _abi.check(_lib().dmdx_component_set_charges(self._h, _p(q), len(q)))


class CoulombDirect(_Charged):
    """src/force/coulomb_direct.h:11-39"""

    def __init__(self, params: CoulombDirectParams, **mode):
        h = C.c_void_p()
        _abi.check(_lib().dmdx_coulomb_direct_create(float(params.cutoff),
                                                     float(params.ewald_coefficient), C.byref(h)))
        self._h = h
        super().__init__(**mode)

    def rebuild_pair_list(self, pos_x, pos_y, pos_z, cell: Cell):
        self._rebuild(pos_x, pos_y, pos_z, cell)


class CoulombExclusion(_Charged):
    """src/force/coulomb_exclusion.h:6-25"""

    def __init__(self, **mode):
        h = C.c_void_p()
        _abi.check(_lib().dmdx_coulomb_exclusion_create(C.byref(h)))
        self._h = h
        super().__init__(**mode)

    def set_exclusions(self, excl_i, excl_j):
        a, b = _i32(excl_i), _i32(excl_j)
        if len(a) != len(b):
            raise ForceException(-3, "excl_i and excl_j differ in length")
        _abi.check(_lib().dmdx_coulomb_exclusion_set_exclusions(self._h, _p(a), _p(b), len(a)))

    def set_ewald_coefficient(self, alpha: float):
        """Ewald coefficient of the CORRECTED (erf) exclusion term; no
        counterpart in the reference, whose term is the bare 1/r."""
        _abi.check(_lib().dmdx_coulomb_exclusion_set_ewald_coefficient(self._h, float(alpha)))


@dataclass
class PMEParams:
    """src/force/coulomb_pme.h:7-12"""
    cutoff: float = 1.2
    ewald_coefficient: float = 3.5
    nx: int = 32
    ny: int = 32
    nz: int = 32
    spline_order: int = 4


class CoulombPME(_Charged):
    """src/force/coulomb_pme.h:14-70"""

    def __init__(self, params: PMEParams, **mode):
        h = C.c_void_p()
        _abi.check(_lib().dmdx_pme_create(float(params.cutoff), float(params.ewald_coefficient),
                                          int(params.nx), int(params.ny), int(params.nz),
                                          int(params.spline_order), C.byref(h)))
        self._h = h
        super().__init__(**mode)

    @staticmethod
    def theta4(u: float) -> float:
        v = C.c_double()
        _abi.check(_lib().dmdx_pme_theta4(float(u), C.byref(v)))
        return v.value

    @staticmethod
    def dtheta4(u: float) -> float:
        v = C.c_double()
        _abi.check(_lib().dmdx_pme_dtheta4(float(u), C.byref(v)))
        return v.value

    @staticmethod
    def calc_self_energy(alpha: float, charges, k_e: float) -> float:
        q = _f64(charges)
        v = C.c_double()
        _abi.check(_lib().dmdx_pme_self_energy(float(alpha), _p(q), len(q), float(k_e),
                                               C.byref(v)))
        return v.value


@dataclass
class HarmonicBondParams:
    """src/force/harmonic_bond.h:7-11"""
    n: int = 0
    i: list = field(default_factory=list)
    j: list = field(default_factory=list)
    k: list = field(default_factory=list)
    r0: list = field(default_factory=list)


class HarmonicBond(ForceComponent):
    def __init__(self, p: HarmonicBondParams, **mode):
        h = C.c_void_p()
        a = [_i32(p.i), _i32(p.j), _f64(p.k), _f64(p.r0)]
        _abi.check(_lib().dmdx_harmonic_bond_create(int(p.n), *[_p(v) for v in a], C.byref(h)))
        self._h = h
        super().__init__(**mode)


@dataclass
class HarmonicAngleParams:
    """src/force/harmonic_angle.h:7-11"""
    n: int = 0
    i: list = field(default_factory=list)
    j: list = field(default_factory=list)
    k: list = field(default_factory=list)
    k_theta: list = field(default_factory=list)
    theta0: list = field(default_factory=list)


class HarmonicAngle(ForceComponent):
    def __init__(self, p: HarmonicAngleParams, **mode):
        h = C.c_void_p()
        a = [_i32(p.i), _i32(p.j), _i32(p.k), _f64(p.k_theta), _f64(p.theta0)]
        _abi.check(_lib().dmdx_harmonic_angle_create(int(p.n), *[_p(v) for v in a], C.byref(h)))
        self._h = h
        super().__init__(**mode)


@dataclass
class PeriodicDihedralParams:
    """src/force/periodic_dihedral.h:7-12"""
    n: int = 0
    i: list = field(default_factory=list)
    j: list = field(default_factory=list)
    k: list = field(default_factory=list)
    l: list = field(default_factory=list)  # noqa: E741 (reference field name)
    periodicity: list = field(default_factory=list)
    k_phi: list = field(default_factory=list)
    phi0: list = field(default_factory=list)


class PeriodicDihedral(ForceComponent):
    def __init__(self, p: PeriodicDihedralParams, **mode):
        h = C.c_void_p()
        a = [_i32(p.i), _i32(p.j), _i32(p.k), _i32(p.l), _i32(p.periodicity), _f64(p.k_phi),
             _f64(p.phi0)]
        _abi.check(_lib().dmdx_periodic_dihedral_create(int(p.n), *[_p(v) for v in a],
                                                        C.byref(h)))
        self._h = h
        super().__init__(**mode)


@dataclass
class ImproperParams:
    """src/force/harmonic_improper.h:7-11"""
    n: int = 0
    i: list = field(default_factory=list)
    j: list = field(default_factory=list)
    k: list = field(default_factory=list)
    l: list = field(default_factory=list)  # noqa: E741
    k_phi: list = field(default_factory=list)
    phi0: list = field(default_factory=list)


class HarmonicImproper(ForceComponent):
    def __init__(self, p: ImproperParams, **mode):
        h = C.c_void_p()
        a = [_i32(p.i), _i32(p.j), _i32(p.k), _i32(p.l), _f64(p.k_phi), _f64(p.phi0)]
        _abi.check(_lib().dmdx_harmonic_improper_create(int(p.n), *[_p(v) for v in a],
                                                        C.byref(h)))
        self._h = h
        super().__init__(**mode)


# --------------------------------------------------------------- engine

@dataclass
class EnergyReport:
    total: float = 0.0


class ForceEngine:
    """src/force/force_engine.h:14-26 / force_engine.cpp:10-25: sets the
    potential energy to 0 and lets every component add to the forces (which
    it does NOT zero) and to the energy, in registration order."""

    def __init__(self):
        self._components = []

    def add_component(self, component: ForceComponent):
        self._components.append(component)

    def components(self):
        return self._components

    def compute(self, sys: SystemData, cell: Cell) -> EnergyReport:
        sys.potential_energy = 0.0
        for comp in self._components:
            sys.potential_energy = comp.compute(sys.pos_x, sys.pos_y, sys.pos_z, cell,
                                                sys.forces_x, sys.forces_y, sys.forces_z,
                                                sys.potential_energy)
        return EnergyReport(total=sys.potential_energy)
