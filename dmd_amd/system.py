"""Host-side description of one system, mirroring DMD's SimulationConfig.

``SystemSpec`` holds numpy arrays for everything ``build_simulation``
(src/sim/simulation_engine.cpp:248-386) copies into SystemData and the force
components, and turns them into the POD ``dmdx_system_desc`` of
include/dmdx.h.  Velocities are always explicit (DMD's ``gen_vel`` uses an
implementation-defined std::normal_distribution, so seeded velocities are
generated here and handed to every path alike).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field, replace

import numpy as np

from . import _abi

KB = 0.008314462618  # src/core/constants.h:6
KE = 138.935456      # src/force/coulomb_direct.cpp:42


def _f64(a, n=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if n is not None and a.shape != (n,):
        raise ValueError(f"expected shape ({n},), got {a.shape}")
    return a


def _i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


@dataclass
class SystemSpec:
    pos: np.ndarray                      # (n, 3) float64, unwrapped positions (nm)
    masses: np.ndarray                   # (n,)
    box: tuple                           # (Lx, Ly, Lz)
    vel: np.ndarray | None = None        # (n, 3)
    charges: np.ndarray | None = None
    atom_types: np.ndarray | None = None
    use_lj: bool = False
    lj_sigma: np.ndarray | None = None   # (nt, nt)
    lj_epsilon: np.ndarray | None = None
    lj_cutoff: float = 1.2
    list_skin: float = 0.0
    rebuild_interval: int = 10
    step_since_rebuild: int = 10
    bonds: tuple | None = None           # (i, j, k, r0)
    angles: tuple | None = None          # (i, j, k, k_theta, theta0)
    dihedrals: tuple | None = None       # (i, j, k, l, periodicity, k_phi, phi0)
    impropers: tuple | None = None       # (i, j, k, l, k_phi, phi0)
    coulomb_type: int = _abi.COULOMB_NONE
    coulomb_cutoff: float = 1.2
    ewald_coeff: float = 3.5
    pme_grid: tuple = (32, 32, 32)
    pme_order: int = 4
    exclusions: tuple | None = None      # (i, j)
    dt: float = 0.002
    thermostat: int = _abi.THERMO_NONE
    thermo_temperature: float = 300.0
    thermo_tau: float = 0.1
    barostat: int = _abi.BARO_NONE
    baro_pressure: float = 1.0
    baro_tau: float = 1.0
    baro_compressibility: float = 4.5e-5
    physics: int = _abi.FAITHFUL
    precision: int = _abi.FP64
    constraints: tuple | None = None     # (i, j, distance) SHAKE/RATTLE pairs
    constraint_tolerance: float = 1e-6   # SimulationConfig::constraint_tolerance
    andersen_frequency: float = 10.0     # AndersenParams::frequency (1/ps)
    andersen_calls: int = 0              # apply() calls already made
    name: str = ""
    _keep: list = field(default_factory=list, repr=False)

    @property
    def n(self) -> int:
        return int(self.pos.shape[0])

    def with_(self, **kw) -> "SystemSpec":
        return replace(self, _keep=[], **kw)

    def to_desc(self) -> _abi.SystemDesc:
        """Build the POD descriptor; the arrays it points to stay alive with self."""
        n = self.n
        keep = []

        def dp(a):
            if a is None:
                return None
            a = _f64(a)
            keep.append(a)
            return a.ctypes.data_as(C.POINTER(C.c_double))

        def ip(a):
            if a is None:
                return None
            a = _i32(a)
            keep.append(a)
            return a.ctypes.data_as(C.POINTER(C.c_int32))

        d = _abi.SystemDesc()
        d.n_atoms = n
        d.box = (C.c_double * 3)(*[float(b) for b in self.box])
        pos = np.asarray(self.pos, dtype=np.float64)
        d.pos_x, d.pos_y, d.pos_z = dp(pos[:, 0]), dp(pos[:, 1]), dp(pos[:, 2])
        if self.vel is not None:
            vel = np.asarray(self.vel, dtype=np.float64)
            d.vel_x, d.vel_y, d.vel_z = dp(vel[:, 0]), dp(vel[:, 1]), dp(vel[:, 2])
        d.masses = dp(_f64(self.masses, n))
        d.charges = dp(None if self.charges is None else _f64(self.charges, n))
        d.atom_types = ip(self.atom_types)
        d.use_lj = int(bool(self.use_lj))
        if self.use_lj:
            sig = np.atleast_2d(np.asarray(self.lj_sigma, dtype=np.float64))
            eps = np.atleast_2d(np.asarray(self.lj_epsilon, dtype=np.float64))
            d.lj_n_types = sig.shape[0]
            d.lj_sigma, d.lj_epsilon = dp(sig.ravel()), dp(eps.ravel())
        d.lj_cutoff = float(self.lj_cutoff)
        d.list_skin = float(self.list_skin)
        d.rebuild_interval = int(self.rebuild_interval)
        d.step_since_rebuild = int(self.step_since_rebuild)
        if self.bonds is not None and len(self.bonds[0]):
            bi, bj, bk, br = self.bonds
            d.n_bonds = len(bi)
            d.bond_i, d.bond_j, d.bond_k, d.bond_r0 = ip(bi), ip(bj), dp(bk), dp(br)
        if self.angles is not None and len(self.angles[0]):
            ai, aj, ak, kt, t0 = self.angles
            d.n_angles = len(ai)
            d.angle_i, d.angle_j, d.angle_k = ip(ai), ip(aj), ip(ak)
            d.angle_k_theta, d.angle_theta0 = dp(kt), dp(t0)
        if self.dihedrals is not None and len(self.dihedrals[0]):
            a, b, c, e, per, kp, p0 = self.dihedrals
            d.n_dihedrals = len(a)
            d.dih_i, d.dih_j, d.dih_k, d.dih_l = ip(a), ip(b), ip(c), ip(e)
            d.dih_periodicity, d.dih_k_phi, d.dih_phi0 = ip(per), dp(kp), dp(p0)
        if self.impropers is not None and len(self.impropers[0]):
            a, b, c, e, kp, p0 = self.impropers
            d.n_impropers = len(a)
            d.imp_i, d.imp_j, d.imp_k, d.imp_l = ip(a), ip(b), ip(c), ip(e)
            d.imp_k_phi, d.imp_phi0 = dp(kp), dp(p0)
        d.coulomb_type = int(self.coulomb_type)
        d.coulomb_cutoff = float(self.coulomb_cutoff)
        d.ewald_coeff = float(self.ewald_coeff)
        d.pme_grid = (C.c_int32 * 3)(*[int(g) for g in self.pme_grid])
        d.pme_order = int(self.pme_order)
        if self.exclusions is not None and len(self.exclusions[0]):
            ei, ej = self.exclusions
            d.n_excl = len(ei)
            d.excl_i, d.excl_j = ip(ei), ip(ej)
        d.dt = float(self.dt)
        d.thermostat = int(self.thermostat)
        d.thermo_temperature = float(self.thermo_temperature)
        d.thermo_tau = float(self.thermo_tau)
        d.barostat = int(self.barostat)
        d.baro_pressure = float(self.baro_pressure)
        d.baro_tau = float(self.baro_tau)
        d.baro_compressibility = float(self.baro_compressibility)
        d.physics = int(self.physics)
        d.precision = int(self.precision)
        if self.constraints is not None and len(self.constraints[0]):
            ci, cj, cd = self.constraints
            d.n_constraints = len(ci)
            d.con_i, d.con_j, d.con_distance = ip(ci), ip(cj), dp(cd)
        d.con_tolerance = float(self.constraint_tolerance)
        d.andersen_frequency = float(self.andersen_frequency)
        d.andersen_calls = int(self.andersen_calls)
        self._keep = keep
        return d
