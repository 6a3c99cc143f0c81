"""Seeded synthetic systems for the benchmark configurations.

Every generator is deterministic in its seed (SplitMix64 derivation of
per-system seeds, numpy's Philox bit generator for the draws), so the
oracle, the FP64 device path and the mixed device path all start from the
same arrays.  Velocities are explicit: Maxwell-Boltzmann draws with the
centre-of-mass momentum removed and a rescale to the target temperature with
3N-3 degrees of freedom, the steps of build_simulation's gen_vel
(src/sim/simulation_engine.cpp:339-383).

CORRECTED-physics workloads rebuild their lists by the Verlet criterion
(rebuild_interval -1: an atom moved more than half the skin since the last
build); argon_fcc keeps DMD's fixed cadence (every 10 force computes).

Workloads (sizes as named):
  argon_fcc      4000-atom LJ argon on an FCC lattice, rho* 0.8442, NVE.
  water_box      flexible 3-site water (O -0.82, H +0.41), SPME, Nose-Hoover.
  lj_fluid       LJ argon replica of arbitrary size (batched workloads).
  chain_melt     bead-spring chains with bonds, angles, dihedrals.
  replica_mix    256 uneven replicas, 60% LJ fluids / 40% chain melts.
"""
from __future__ import annotations

import math

import numpy as np

from . import _abi
from .system import KB, SystemSpec

MASK64 = (1 << 64) - 1

AR_SIGMA = 0.34
AR_EPS = 0.996
AR_MASS = 39.948


def splitmix64(x: int) -> int:
    """One SplitMix64 output for state x (Steele, Lea & Flood 2014)."""
    z = (x + 0x9E3779B97F4A7C15) & MASK64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return z ^ (z >> 31)


def derive_seed(master: int, *path: int) -> int:
    s = master & MASK64
    for p in path:
        s = splitmix64(s ^ ((p * 0x9E3779B97F4A7C15) & MASK64))
    return s


def rng_for(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(key=seed & MASK64))


def maxwell_boltzmann(rng, masses, temperature):
    n = len(masses)
    sig = np.sqrt(KB * temperature / masses)
    v = rng.standard_normal((n, 3)) * sig[:, None]
    com = (masses[:, None] * v).sum(0) / masses.sum()
    v -= com
    if temperature > 0 and n > 1:
        ekin = 0.5 * (masses[:, None] * v * v).sum()
        t_now = 2.0 * ekin / ((3.0 * n - 3.0) * KB)
        v *= math.sqrt(temperature / t_now)
    return v


def fcc_lattice(n_cells: int, box: float) -> np.ndarray:
    """Basis order of tests/integration/test_argon_nve.cpp:66-90."""
    a = box / n_cells
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    out = []
    for i in range(n_cells):
        for j in range(n_cells):
            for k in range(n_cells):
                for b in basis:
                    out.append([(i + b[0]) * a, (j + b[1]) * a, (k + b[2]) * a])
    return np.array(out, dtype=np.float64)


def ewald_alpha(cutoff: float, rtol: float = 1e-5) -> float:
    """alpha with erfc(alpha * cutoff) = rtol (bisection)."""
    lo, hi = 0.0, 50.0 / cutoff
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if math.erfc(mid * cutoff) > rtol:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def smooth_grid(n: int) -> int:
    """Smallest n' >= n whose prime factors are all in {2, 3, 5, 7}."""
    m = max(8, n)
    while True:
        k = m
        for p in (2, 3, 5, 7):
            while k % p == 0:
                k //= p
        if k == 1:
            return m
        m += 1


# ----------------------------------------------------------------- argon

def argon_fcc(n_cells: int = 10, rho_star: float = 0.8442, t_star: float = 0.72,
              cutoff: float = 0.85, skin: float = 0.0, seed: int = 42,
              dt: float = 0.001, **kw) -> SystemSpec:
    n = 4 * n_cells ** 3
    box = (n * AR_SIGMA ** 3 / rho_star) ** (1.0 / 3.0)
    pos = fcc_lattice(n_cells, box)
    masses = np.full(n, AR_MASS)
    temp = t_star * AR_EPS / KB
    vel = maxwell_boltzmann(rng_for(derive_seed(seed, 0)), masses, temp)
    spec = SystemSpec(
        pos=pos, vel=vel, masses=masses, box=(box, box, box), charges=np.zeros(n),
        atom_types=np.zeros(n, np.int32), use_lj=True,
        lj_sigma=np.array([[AR_SIGMA]]), lj_epsilon=np.array([[AR_EPS]]),
        lj_cutoff=cutoff, list_skin=skin, dt=dt, coulomb_type=_abi.COULOMB_NONE,
        name=f"argon_fcc_{n}")
    return spec.with_(**kw) if kw else spec


def lj_fluid(n: int, rho_star: float, t_star: float, seed: int, cutoff: float | None = None,
             skin: float = 0.1, dt: float = 0.001, thermostat=_abi.THERMO_BERENDSEN,
             **kw) -> SystemSpec:
    """LJ argon of n atoms: a seeded subset of an FCC lattice in the box of density rho*."""
    box = (n * AR_SIGMA ** 3 / rho_star) ** (1.0 / 3.0)
    n_cells = int(math.ceil((n / 4.0) ** (1.0 / 3.0)))
    sites = fcc_lattice(n_cells, box)
    rng = rng_for(seed)
    pick = np.sort(rng.permutation(len(sites))[:n])
    pos = sites[pick] + (box / n_cells) * 0.02 * (rng.random((n, 3)) - 0.5)
    masses = np.full(n, AR_MASS)
    temp = t_star * AR_EPS / KB
    vel = maxwell_boltzmann(rng, masses, temp)
    spec = SystemSpec(
        pos=pos, vel=vel, masses=masses, box=(box, box, box), charges=np.zeros(n),
        atom_types=np.zeros(n, np.int32), use_lj=True,
        lj_sigma=np.array([[AR_SIGMA]]), lj_epsilon=np.array([[AR_EPS]]),
        lj_cutoff=cutoff if cutoff is not None else 2.5 * AR_SIGMA, list_skin=skin, dt=dt,
        thermostat=thermostat, thermo_temperature=temp, thermo_tau=0.1,
        physics=_abi.CORRECTED, rebuild_interval=-1, name=f"lj_fluid_{n}")
    return spec.with_(**kw) if kw else spec


# ----------------------------------------------------------------- water

WATER_Q_O, WATER_Q_H = -0.82, 0.41
WATER_M_O, WATER_M_H = 15.999, 1.008
WATER_R0, WATER_KB = 0.1012, 443153.0
WATER_THETA0, WATER_KA = math.radians(113.24), 317.5
WATER_SIG_O, WATER_EPS_O = 0.316557, 0.650194
NA = 6.02214076e23


def _random_rotations(rng, n):
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1),
    ], 1)


def water_box(n_side: int = 16, cutoff: float = 0.9, rtol: float = 1e-5, grid: int | None = None,
              spacing: float = 0.1, seed: int = 42, dt: float = 0.0005, temperature: float = 300.0,
              thermostat=_abi.THERMO_NOSE_HOOVER, skin: float = 0.1, n_mol: int | None = None,
              rigid: bool = False, **kw) -> SystemSpec:
    """Flexible 3-site water at 1.0 g/cm^3: O on a jittered cubic lattice
    (+-0.02 nm), H at r0 with the HOH angle theta0 in a random orientation.
    rigid: no bond or angle terms; instead three SHAKE/RATTLE constraints per
    molecule (O-H1, O-H2, H1-H2) at the same geometry, as
    docs/adr/0004-use-shake-for-water.md lays them out."""
    nmol = n_mol if n_mol is not None else n_side ** 3
    side = int(math.ceil(nmol ** (1.0 / 3.0) - 1e-9))
    vol = nmol * (WATER_M_O + 2 * WATER_M_H) / (NA * 1.0) * 1e21  # nm^3
    box = vol ** (1.0 / 3.0)
    rng = rng_for(seed)
    a = box / side
    idx = np.array([(i, j, k) for i in range(side) for j in range(side) for k in range(side)])
    idx = idx[np.sort(rng.permutation(len(idx))[:nmol])] if len(idx) > nmol else idx
    o = (idx + 0.5) * a + rng.uniform(-0.02, 0.02, (nmol, 3))
    half = 0.5 * WATER_THETA0
    tmpl = np.array([[math.sin(half), 0.0, math.cos(half)],
                     [-math.sin(half), 0.0, math.cos(half)]]) * WATER_R0
    rot = _random_rotations(rng, nmol)
    h = o[:, None, :] + np.einsum("mij,hj->mhi", rot, tmpl)
    n = 3 * nmol
    pos = np.empty((n, 3))
    pos[0::3] = o
    pos[1::3] = h[:, 0]
    pos[2::3] = h[:, 1]
    masses = np.tile([WATER_M_O, WATER_M_H, WATER_M_H], nmol)
    charges = np.tile([WATER_Q_O, WATER_Q_H, WATER_Q_H], nmol)
    types = np.tile(np.array([0, 1, 1], np.int32), nmol)
    base = np.arange(nmol, dtype=np.int32) * 3
    bonds = (np.concatenate([base, base]), np.concatenate([base + 1, base + 2]),
             np.full(2 * nmol, WATER_KB), np.full(2 * nmol, WATER_R0))
    angles = (base + 1, base, base + 2, np.full(nmol, WATER_KA), np.full(nmol, WATER_THETA0))
    excl = (np.concatenate([base, base, base + 1]), np.concatenate([base + 1, base + 2, base + 2]))
    cons = None
    if rigid:
        d_hh = 2.0 * WATER_R0 * math.sin(half)
        cons = (np.stack([base, base, base + 1], 1).ravel(),
                np.stack([base + 1, base + 2, base + 2], 1).ravel(),
                np.tile([WATER_R0, WATER_R0, d_hh], nmol))
        bonds = angles = None
    sig = np.array([[WATER_SIG_O, 1.0], [1.0, 1.0]])
    eps = np.array([[WATER_EPS_O, 0.0], [0.0, 0.0]])
    vel = maxwell_boltzmann(rng, masses, temperature)
    g = grid if grid is not None else smooth_grid(int(round(box / spacing)))
    spec = SystemSpec(
        pos=pos, vel=vel, masses=masses, box=(box, box, box), charges=charges,
        atom_types=types, use_lj=True, lj_sigma=sig, lj_epsilon=eps, lj_cutoff=cutoff,
        list_skin=skin, bonds=bonds, angles=angles, exclusions=excl,
        coulomb_type=_abi.COULOMB_PME, coulomb_cutoff=cutoff,
        ewald_coeff=ewald_alpha(cutoff, rtol), pme_grid=(g, g, g), dt=dt,
        thermostat=thermostat, thermo_temperature=temperature, thermo_tau=0.1,
        physics=_abi.CORRECTED, rebuild_interval=-1, constraints=cons,
        name=f"{'rigid_' if rigid else ''}water_{n}")
    return spec.with_(**kw) if kw else spec


# ----------------------------------------------------------- chain melts

CH_SIGMA, CH_EPS, CH_MASS = 0.1, 1.0, 12.0
CH_R0, CH_KB = 0.1, 2.0e5
CH_THETA0, CH_KA = math.radians(120.0), 200.0
CH_KD, CH_NDIH = 2.0, 3


def chain_melt(n: int, seed: int, rho_star: float = 0.85, t_star: float = 1.0,
               skin: float = 0.05, dt: float = 0.001, thermostat=_abi.THERMO_BERENDSEN,
               **kw) -> SystemSpec:
    """Bead-spring melt: chains of 20-100 beads laid as all-trans zigzags
    (120 degree angles, bond r0) along rows of the box, rows on a square grid
    in the y-z plane, zigzag plane rotated 45 degrees and adjacent rows
    offset by half a step; the path turns row to row (boustrophedon) and is
    cut into chains."""
    rng = rng_for(seed)
    lens = []
    total = 0
    while total < n:
        L = int(rng.integers(20, 101))
        L = min(L, n - total) if n - total >= 20 else n - total
        lens.append(L)
        total += L
    if len(lens) > 1 and lens[-1] < 3:
        lens[-2] += lens[-1]
        lens.pop()
    box = (n * CH_SIGMA ** 3 / rho_star) ** (1.0 / 3.0)
    d = CH_R0 * math.sin(CH_THETA0 / 2.0)          # step along the row
    hgt = CH_R0 * math.cos(CH_THETA0 / 2.0)        # zigzag amplitude
    cap = max(2, int(box / d))
    rows_needed = int(math.ceil(n / cap))
    side = int(math.ceil(math.sqrt(rows_needed)))
    s = box / side
    pos = np.empty((n, 3))
    for b in range(n):
        r, i = divmod(b, cap)
        ry, rz = divmod(r, side)
        if rz % 2 == 1:
            pass
        xi = i if r % 2 == 0 else cap - 1 - i
        off = 0.5 * d if (ry + rz) % 2 else 0.0
        zz = (xi % 2) * hgt / math.sqrt(2.0)
        pos[b] = [xi * d + off, (ry + 0.25) * s + zz, (rz + 0.25) * s + zz]
    pos += rng.uniform(-0.002, 0.002, pos.shape)
    masses = np.full(n, CH_MASS)
    bi, bj, ai, aj, ak, di, dj, dk, dl = ([] for _ in range(9))
    start = 0
    for L in lens:
        idx = np.arange(start, start + L)
        bi += list(idx[:-1]); bj += list(idx[1:])
        ai += list(idx[:-2]); aj += list(idx[1:-1]); ak += list(idx[2:])
        di += list(idx[:-3]); dj += list(idx[1:-2]); dk += list(idx[2:-1]); dl += list(idx[3:])
        start += L
    nb, na, nd = len(bi), len(ai), len(di)
    bonds = (np.array(bi, np.int32), np.array(bj, np.int32), np.full(nb, CH_KB), np.full(nb, CH_R0))
    angles = (np.array(ai, np.int32), np.array(aj, np.int32), np.array(ak, np.int32),
              np.full(na, CH_KA), np.full(na, CH_THETA0))
    dihedrals = (np.array(di, np.int32), np.array(dj, np.int32), np.array(dk, np.int32),
                 np.array(dl, np.int32), np.full(nd, CH_NDIH, np.int32), np.full(nd, CH_KD),
                 np.zeros(nd))
    excl = (np.concatenate([bonds[0], angles[0]]), np.concatenate([bonds[1], angles[2]]))
    temp = t_star * CH_EPS / KB
    vel = maxwell_boltzmann(rng, masses, temp)
    spec = SystemSpec(
        pos=pos, vel=vel, masses=masses, box=(box, box, box), charges=np.zeros(n),
        atom_types=np.zeros(n, np.int32), use_lj=True, lj_sigma=np.array([[CH_SIGMA]]),
        lj_epsilon=np.array([[CH_EPS]]), lj_cutoff=2.5 * CH_SIGMA, list_skin=skin,
        bonds=bonds, angles=angles, dihedrals=dihedrals, exclusions=excl,
        coulomb_type=_abi.COULOMB_NONE, dt=dt, thermostat=thermostat,
        thermo_temperature=temp, thermo_tau=0.1, physics=_abi.CORRECTED,
        rebuild_interval=-1, name=f"chain_melt_{n}")
    return spec.with_(**kw) if kw else spec


def replica_mix(n_replicas: int = 256, n_min: int = 500, n_max: int = 8000, seed: int = 42,
                lj_fraction: float = 0.6, **kw) -> list:
    """Uneven replicas: sizes log-uniform in [n_min, n_max]; 60% LJ fluids
    (rho* U(0.70, 0.90), T* U(0.7, 1.5)), 40% bead-spring chain melts."""
    rng = rng_for(derive_seed(seed, 3))
    out = []
    for r in range(n_replicas):
        n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        s = derive_seed(seed, 3, r)
        if rng.random() < lj_fraction:
            out.append(lj_fluid(n, float(rng.uniform(0.70, 0.90)), float(rng.uniform(0.7, 1.5)),
                                s, **kw))
        else:
            out.append(chain_melt(n, s, **kw))
    return out


def serving_requests(n_requests: int, seed: int = 42, steps_scale: float = 1.0,
                     n_lj=(1000, 32000), n_water=(216, 8192), steps=(1000, 20000),
                     precision=_abi.MIXED) -> list:
    """Request stream for the serving loop: with equal probability an LJ argon
    fluid of N ~ U[1k, 32k] atoms (rho* 0.8442, T* 0.72, rc 0.85 nm, skin
    0.1 nm, NVE) or a box of U[216, 8192] flexible waters (SPME, grid
    round(L/0.1) made 7-smooth, Nose-Hoover 300 K); step count log-uniform
    in [1000, 20000], times `steps_scale`.  Returns [(spec, n_steps), ...]."""
    rng = rng_for(derive_seed(seed, 4))
    out = []
    for r in range(n_requests):
        s = derive_seed(seed, 4, r)
        if rng.random() < 0.5:
            n = int(rng.integers(n_lj[0], n_lj[1] + 1))
            spec = lj_fluid(n, 0.8442, 0.72, s, cutoff=0.85, skin=0.1,
                            thermostat=_abi.THERMO_NONE, precision=precision)
        else:
            m = int(rng.integers(n_water[0], n_water[1] + 1))
            spec = water_box(n_mol=m, seed=s, precision=precision)
        k = math.exp(rng.uniform(math.log(steps[0]), math.log(steps[1])))
        out.append((spec, max(1, int(round(k * steps_scale)))))
    return out
