"""DMD's Python core API (module ``_dmd_core``) over the device engine.

DMD's runner does ``from _dmd_core import SimulationConfig, Engine``
(python/dmd/runner.py:3) and drives ``Engine.run_n`` (runner.py:46-58).
This module offers the same names with the same attributes -- the
``SimulationConfig`` fields bound in python/dmd/core.cpp:106-410 and the
``Engine`` surface of core.cpp:415-427 -- so the runner, the system builder
and the validation scripts can drive the MI355X path by importing from here
instead.  Differences, all explicit:

* ``gen_vel``: DMD draws with std::mt19937 + std::normal_distribution, whose
  output is implementation-defined (simulation_engine.cpp:339-383); here the
  velocities are drawn with a seeded Philox generator, COM removed and
  rescaled to ``init_temperature`` with 3N-3 degrees of freedom, as DMD does.
* ``constraint_type = "shake"`` (SHAKE after the drift, RATTLE after the
  closing kick) and the Andersen thermostat (DMD's own mt19937_64 stream)
  run on the device like the other couplings.
* trajectories: ``run_n`` writes an H5MD frame (DMD's layout,
  dmd_amd.trajectory) every ``nstxout`` steps, as EngineWrapper::run_n
  (core.cpp:29-44), and ``run`` also writes DMD's binary checkpoint every
  ``checkpoint_interval`` steps, as SimulationEngine::run
  (simulation_engine.cpp:88-106).  Frames come from asynchronous snapshots
  written while the next steps run; ``nstvout`` and the energy file are not
  written, as in DMD's engine.  ``run`` and ``run_n`` share one writer
  (DMD opens a second one on the same path, which HDF5 refuses).
* ``precision`` (``"fp64"`` or ``"mixed"``) and ``physics`` (``"faithful"``
  reproduces DMD's arithmetic including its quirks, ``"corrected"`` the
  textbook forms) are extra attributes; the defaults reproduce DMD.
"""
from __future__ import annotations

import math

import numpy as np

from . import _abi
from .system import KB, SystemSpec


class ThermostatConfig:
    """src/sim/simulation_config.h:13-18"""

    def __init__(self):
        self.type = "none"
        self.temperature = 300.0
        self.tau = 0.1
        self.frequency = 10.0


class BarostatConfig:
    """src/sim/simulation_config.h:20-25"""

    def __init__(self):
        self.type = "none"
        self.pressure = 1.0
        self.tau = 1.0
        self.compressibility = 4.5e-5


_ARRAYS_F = ("mass", "charges", "pos_x", "pos_y", "pos_z", "vel_x", "vel_y", "vel_z",
             "lj_sigma", "lj_epsilon", "bond_k", "bond_r0", "angle_k_theta", "angle_theta0",
             "dihedral_k_phi", "dihedral_phi0", "improper_k_phi", "improper_phi0",
             "constraint_dist")
_ARRAYS_I = ("atom_types", "bond_i", "bond_j", "angle_i", "angle_j", "angle_k",
             "dihedral_i", "dihedral_j", "dihedral_k", "dihedral_l", "dihedral_periodicity",
             "improper_i", "improper_j", "improper_k", "improper_l", "excl_i", "excl_j",
             "constraint_i", "constraint_j")


class SimulationConfig:
    """Field names and defaults of SimulationConfig (simulation_config.h:27-84)
    as bound by python/dmd/core.cpp; array fields are numpy arrays."""

    def __init__(self):
        self.dt = 0.002
        self.n_steps = 1000
        self.init_temperature = 300.0
        self.gen_vel = False
        self.seed = 42
        self.box_size = 3.0
        self.use_lj = False
        self.lj_cutoff = 1.2
        self.coulomb_type = "cutoff"
        self.coulomb_cutoff = 1.2
        self.pme_order = 4
        self.pme_grid_spacing = 0.12
        self.ewald_coeff = 0.34
        self.trajectory_path = "trajectory.h5"
        self.nstxout = 0
        self.nstvout = 0
        self.nstenergy = 10
        self.energy_path = "energy.h5"
        self.checkpoint_interval = 500
        self.checkpoint_path = "checkpoint.bin"
        self.thermostat = ThermostatConfig()
        self.barostat = BarostatConfig()
        self.constraint_type = "none"
        self.constraint_tolerance = 1e-6
        for a in _ARRAYS_F:
            setattr(self, a, np.zeros(0))
        for a in _ARRAYS_I:
            setattr(self, a, np.zeros(0, dtype=np.int32))
        # device-path options (not in DMD)
        self.precision = "fp64"
        self.physics = "faithful"

    def __setattr__(self, k, v):
        if k in _ARRAYS_F:
            v = np.asarray(v, dtype=np.float64).ravel()
        elif k in _ARRAYS_I:
            v = np.asarray(v, dtype=np.int32).ravel()
        object.__setattr__(self, k, v)


def _velocities(cfg, masses):
    """gen_vel: Maxwell-Boltzmann at init_temperature, COM removed, rescaled
    with 3N-3 degrees of freedom (as simulation_engine.cpp:339-383)."""
    n = len(masses)
    rng = np.random.Generator(np.random.Philox(key=int(cfg.seed)))
    sig = np.sqrt(KB * cfg.init_temperature / masses)
    v = rng.standard_normal((n, 3)) * sig[:, None]
    v -= (masses[:, None] * v).sum(0) / masses.sum()
    ke = 0.5 * (masses[:, None] * v * v).sum()
    if ke > 0 and n > 1:
        v *= math.sqrt(cfg.init_temperature / (2.0 * ke / ((3 * n - 3) * KB)))
    return v


def spec_from_config(cfg: SimulationConfig) -> SystemSpec:
    """build_simulation (simulation_engine.cpp:248-386) as a SystemSpec."""
    thermo = {"none": _abi.THERMO_NONE, "berendsen": _abi.THERMO_BERENDSEN,
              "nose-hoover": _abi.THERMO_NOSE_HOOVER, "andersen": _abi.THERMO_ANDERSEN}
    baro = {"none": _abi.BARO_NONE, "berendsen": _abi.BARO_BERENDSEN,
            "andersen": _abi.BARO_ANDERSEN}
    if cfg.thermostat.type not in thermo:
        raise ValueError(f"Unknown thermostat type: {cfg.thermostat.type}")
    if cfg.barostat.type not in baro:
        raise ValueError(f"Unknown barostat type: {cfg.barostat.type}")
    n = len(cfg.pos_x)
    pos = np.stack([cfg.pos_x, cfg.pos_y, cfg.pos_z], 1)
    masses = np.asarray(cfg.mass, dtype=np.float64)
    if cfg.gen_vel and n > 0:
        vel = _velocities(cfg, masses)
    elif len(cfg.vel_x) == n:
        vel = np.stack([cfg.vel_x, cfg.vel_y, cfg.vel_z], 1)
    else:
        vel = np.zeros((n, 3))
    nt = int(round(math.sqrt(len(cfg.lj_sigma)))) if len(cfg.lj_sigma) else 1
    types = cfg.atom_types if len(cfg.atom_types) else np.zeros(n, np.int32)
    charges = cfg.charges if len(cfg.charges) else np.zeros(n)
    L = float(cfg.box_size)
    kw = {}
    if len(cfg.bond_i):
        kw["bonds"] = (cfg.bond_i, cfg.bond_j, cfg.bond_k, cfg.bond_r0)
    if len(cfg.angle_i):
        kw["angles"] = (cfg.angle_i, cfg.angle_j, cfg.angle_k, cfg.angle_k_theta,
                        cfg.angle_theta0)
    if len(cfg.dihedral_i):
        kw["dihedrals"] = (cfg.dihedral_i, cfg.dihedral_j, cfg.dihedral_k, cfg.dihedral_l,
                           cfg.dihedral_periodicity, cfg.dihedral_k_phi, cfg.dihedral_phi0)
    if len(cfg.improper_i):
        kw["impropers"] = (cfg.improper_i, cfg.improper_j, cfg.improper_k, cfg.improper_l,
                           cfg.improper_k_phi, cfg.improper_phi0)
    if len(cfg.excl_i):
        kw["exclusions"] = (cfg.excl_i, cfg.excl_j)
    if cfg.constraint_type == "shake" and len(cfg.constraint_i):  # simulation_engine.cpp:293-301
        kw["constraints"] = (cfg.constraint_i, cfg.constraint_j, cfg.constraint_dist)
        kw["constraint_tolerance"] = float(cfg.constraint_tolerance)
    ct = cfg.coulomb_type
    coul = (_abi.COULOMB_PME if ct == "pme" else
            _abi.COULOMB_CUTOFF if ct in ("direct", "cutoff") else _abi.COULOMB_NONE)
    g = max(8, int(L / cfg.pme_grid_spacing + 0.5))  # simulation_engine.cpp:183-185
    return SystemSpec(
        pos=pos, vel=vel, masses=masses, box=(L, L, L), charges=charges, atom_types=types,
        use_lj=bool(cfg.use_lj),
        lj_sigma=np.asarray(cfg.lj_sigma).reshape(nt, nt) if len(cfg.lj_sigma) else None,
        lj_epsilon=np.asarray(cfg.lj_epsilon).reshape(nt, nt) if len(cfg.lj_epsilon) else None,
        lj_cutoff=float(cfg.lj_cutoff), list_skin=0.0, rebuild_interval=10,
        step_since_rebuild=10, coulomb_type=coul, coulomb_cutoff=float(cfg.coulomb_cutoff),
        ewald_coeff=float(cfg.ewald_coeff), pme_grid=(g, g, g), pme_order=int(cfg.pme_order),
        dt=float(cfg.dt), thermostat=thermo[cfg.thermostat.type],
        thermo_temperature=float(cfg.thermostat.temperature),
        thermo_tau=float(cfg.thermostat.tau),
        andersen_frequency=float(cfg.thermostat.frequency), barostat=baro[cfg.barostat.type],
        baro_pressure=float(cfg.barostat.pressure), baro_tau=float(cfg.barostat.tau),
        baro_compressibility=float(cfg.barostat.compressibility),
        physics=_abi.CORRECTED if cfg.physics == "corrected" else _abi.FAITHFUL,
        precision=_abi.MIXED if cfg.precision == "mixed" else _abi.FP64, **kw)


class Engine:
    """EngineWrapper of python/dmd/core.cpp:15-93 on the device."""

    def __init__(self, cfg: SimulationConfig):
        from .engine import Engine as _Engine
        from .trajectory import Clock
        self._spec = spec_from_config(cfg)
        self._e = _Engine(self._spec)
        self._cfg = cfg
        self._clock = Clock.start(float(cfg.dt))
        # one writer for run and run_n: DMD holds one in EngineWrapper and
        # one in SimulationEngine, and the second to open the same path
        # fails (HDF5 will not truncate an open file)
        self._writer = None

    def _steps(self, steps: int, every: int):
        from . import trajectory as T
        if every > 0:
            if self._writer is None:
                self._writer = T.H5MDWriter(self._cfg.trajectory_path, self._spec.n, True)
            T.run(self._e, steps, every, self._clock, self._writer)
            return
        self._e.step(steps)
        c = self._clock
        for _ in range(steps):  # SimulationEngine::step's clock (simulation_engine.cpp:54, 82-85)
            c.sys_time = c.time
            c.step += 1
            c.time += c.dt

    def run(self):
        """SimulationEngine::run: n_steps steps, a frame every nstxout steps,
        DMD's binary checkpoint every checkpoint_interval steps."""
        left = int(self._cfg.n_steps)
        every = int(self._cfg.checkpoint_interval)
        while left > 0:
            k = left if every <= 0 else min(left, every - self._clock.step % every)
            self._steps(k, int(self._cfg.nstxout))
            left -= k
            if every > 0 and self._clock.step % every == 0:
                self.save_checkpoint(self._cfg.checkpoint_path)

    def run_n(self, steps: int):
        """EngineWrapper::run_n: steps, with a frame every nstxout steps."""
        self._steps(int(steps), int(self._cfg.nstxout))

    def save_checkpoint(self, path: str):
        """SimulationEngine::save_checkpoint (simulation_engine.cpp:117-125):
        sys_ and the LJ counter in DMD's binary format."""
        from . import checkpoint as CK
        pos, vel, frc = self._e.positions(), self._e.velocities(), self._e.forces()
        sc = self._e.scalars()
        arrays = {"pos_x": pos[:, 0], "pos_y": pos[:, 1], "pos_z": pos[:, 2],
                  "vel_x": vel[:, 0], "vel_y": vel[:, 1], "vel_z": vel[:, 2],
                  "forces_x": frc[:, 0], "forces_y": frc[:, 1], "forces_z": frc[:, 2],
                  "masses": np.asarray(self._spec.masses, float),
                  "charges": (np.zeros(self._spec.n) if self._spec.charges is None
                              else np.asarray(self._spec.charges, float)),
                  "atom_types": (np.zeros(self._spec.n, np.int32)
                                 if self._spec.atom_types is None
                                 else np.asarray(self._spec.atom_types, np.int32))}
        counter = sc["counter"] if self._spec.use_lj else 0
        CK.save(path, CK.Checkpoint(int(self._clock.step), float(self._clock.sys_time),
                                    float(sc["pe"]), arrays, int(counter)))

    def load_checkpoint(self, path: str):
        """SimulationEngine::load_checkpoint (simulation_engine.cpp:127-137):
        positions, velocities, step, time and the LJ counter."""
        from . import checkpoint as CK
        ck = CK.load(path)
        if ck.n != self._spec.n:
            raise ValueError(f"{path} holds {ck.n} atoms, the engine {self._spec.n}")
        a = ck.arrays
        self._e.set_positions(np.stack([a["pos_x"], a["pos_y"], a["pos_z"]], 1))
        self._e.set_velocities(np.stack([a["vel_x"], a["vel_y"], a["vel_z"]], 1))
        if self._spec.use_lj:
            self._e.set_step_since_rebuild(ck.lj_step_since_rebuild)
        self._clock.step = ck.step
        self._clock.time = ck.time
        self._clock.sys_time = ck.time

    @property
    def current_step(self) -> int:
        return int(self._clock.step)

    @property
    def current_time(self) -> float:
        """SimulationEngine::time_: dt added once per step."""
        return float(self._clock.time)

    @property
    def positions(self):
        return self._e.positions()

    @property
    def velocities(self):
        return self._e.velocities()

    @property
    def forces(self):
        return self._e.forces()

    @property
    def box_size(self) -> float:
        # the configured edge, as EngineWrapper::box_size (core.cpp:22-24,46)
        return float(self._cfg.box_size)

    @property
    def n_atoms(self) -> int:
        return self._spec.n

    @property
    def potential_energy(self) -> float:
        return self._e.scalars()["pe"]
