"""ctypes wrapper of the FP64 CPU oracle (oracle/liboracle.so) and of the
reference build (oracle/_ref/libdmdref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product package dmd_amd
never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libdmdref.so")

FAITHFUL, CORRECTED = 0, 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


def build(ref: bool = True) -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if ref:
        subprocess.run(["bash", os.path.join(HERE, "build_ref.sh")], check=True)


def _load(path):
    if not os.path.exists(path):
        raise RuntimeError(f"{path} missing: run `make -C oracle`")
    return C.CDLL(path)


_o = None
_r = None


def lib():
    global _o
    if _o is None:
        _o = _load(LIB)
        for name in ("orc_cell_volume", "orc_theta4", "orc_dtheta4", "orc_pme_self_energy",
                     "orc_kinetic_energy", "orc_temperature", "orc_virial", "orc_pressure"):
            getattr(_o, name).restype = C.c_double
        for name in ("orc_pairs_n2", "orc_pairs_cells", "orc_pairs_filter_excl"):
            getattr(_o, name).restype = C.c_int64
        _o.orc_sim_create.restype = C.c_void_p
        _o.orc_sim_create.argtypes = [C.c_void_p]
        for name in ("orc_sim_destroy", "orc_sim_step", "orc_sim_compute", "orc_sim_forces_exact",
                     "orc_sim_get", "orc_sim_set_positions", "orc_sim_set_velocities",
                     "orc_sim_scalars"):
            getattr(_o, name).argtypes = None
    return _o


def ref_available() -> bool:
    return os.path.exists(REF_LIB)


def ref():
    global _r
    if _r is None:
        _r = _load(REF_LIB)
        for name in ("ref_kinetic_energy", "ref_temperature", "ref_virial", "ref_pressure"):
            getattr(_r, name).restype = C.c_double
        _r.ref_sim_create.restype = C.c_void_p
    return _r


# ------------------------------------------------------------- helpers

def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def P(a):
    if a is None:
        return None
    if a.dtype == np.float64:
        return a.ctypes.data_as(_dp)
    if a.dtype == np.int32:
        return a.ctypes.data_as(_ip)
    raise TypeError(a.dtype)


def cell9(box):
    """Row-major Cell::matrix from three edges (orthorhombic) or a full
    3x3 / 9-element matrix (any cell type)."""
    b = np.asarray(box, dtype=float).ravel()
    if b.size == 9:
        return b.copy()
    c = np.zeros(9)
    c[0], c[4], c[8] = b
    return c


def xyz(pos):
    pos = np.asarray(pos, dtype=np.float64)
    return f64(pos[:, 0]), f64(pos[:, 1]), f64(pos[:, 2])


def zeros3(n):
    return np.zeros(n), np.zeros(n), np.zeros(n)


# ------------------------------------------------------------- core

CELL_TYPES = ("orthorhombic", "triclinic", "hexagonal", "rhombic_dodecahedron")


class cell_type:
    """Context manager: every cell the oracle (and, with ref=True, the
    reference build) sees inside the block has this CellType
    (src/core/cell.h:3-8); orthorhombic again on exit."""

    def __init__(self, t, ref=False):
        self.t = CELL_TYPES.index(t) if isinstance(t, str) else int(t)
        self.ref = ref

    def __enter__(self):
        lib().orc_set_cell_type(C.c_int(self.t))
        if self.ref:
            ref().ref_set_cell_type(C.c_int(self.t))
        return self

    def __exit__(self, *exc):
        lib().orc_set_cell_type(C.c_int(0))
        if self.ref:
            ref().ref_set_cell_type(C.c_int(0))
        return False


def cell_volume(m):
    """Cell::volume of a 3x3 row-major matrix (src/core/cell.cpp:19-24)."""
    return lib().orc_cell_volume(P(f64(np.asarray(m, float).ravel())))


def ref_cell_volume(m, t):
    r = ref()
    r.ref_cell_volume.restype = C.c_double
    return r.ref_cell_volume(P(f64(np.asarray(m, float).ravel())), C.c_int(int(t)))


def min_image(box, d):
    c = cell9(box)
    dx, dy, dz = C.c_double(d[0]), C.c_double(d[1]), C.c_double(d[2])
    lib().orc_min_image(P(c), C.byref(dx), C.byref(dy), C.byref(dz))
    return np.array([dx.value, dy.value, dz.value])


def pairs(pos, box, rc, method="n2"):
    x, y, z = xyz(pos)
    n = len(x)
    fn = lib().orc_pairs_n2 if method == "n2" else lib().orc_pairs_cells
    cap = 0
    pi = np.zeros(1, np.int32)
    pj = np.zeros(1, np.int32)
    c = cell9(box)
    while True:
        cnt = fn(C.c_int64(n), P(x), P(y), P(z), P(c), C.c_double(rc), P(pi), P(pj),
                 C.c_int64(cap))
        if cnt <= cap:
            return pi[:cnt].copy(), pj[:cnt].copy()
        cap = int(cnt)
        pi = np.zeros(cap, np.int32)
        pj = np.zeros(cap, np.int32)


def filter_excl(pi, pj, ei, ej):
    pi, pj = i32(pi).copy(), i32(pj).copy()
    ei, ej = i32(ei), i32(ej)
    cnt = lib().orc_pairs_filter_excl(C.c_int64(len(pi)), P(pi), P(pj), C.c_int64(len(ei)),
                                      P(ei), P(ej))
    return pi[:cnt], pj[:cnt]


# ------------------------------------------------------------- components

def lj(pos, box, rc, sigma, epsilon, types=None, pi=None, pj=None):
    """LennardJones on the exact list (or the given one); returns (F (n,3), E)."""
    x, y, z = xyz(pos)
    n = len(x)
    if pi is None:
        pi, pj = pairs(pos, box, rc)
    pi, pj = i32(pi), i32(pj)
    sig = f64(np.atleast_2d(sigma)).ravel().copy()
    eps = f64(np.atleast_2d(epsilon)).ravel().copy()
    nt = int(round(np.sqrt(len(sig))))
    t = i32(types) if types is not None else None
    fx, fy, fz = zeros3(n)
    e = C.c_double(0.0)
    lib().orc_lj_pairs(C.c_int64(len(pi)), P(pi), P(pj), P(x), P(y), P(z), P(cell9(box)),
                       C.c_double(rc), P(t), C.c_int32(nt), P(sig), P(eps), P(fx), P(fy), P(fz),
                       C.byref(e))
    return np.stack([fx, fy, fz], 1), e.value


def coulomb(pos, box, rc, alpha, q, pi=None, pj=None):
    x, y, z = xyz(pos)
    n = len(x)
    if pi is None:
        pi, pj = pairs(pos, box, rc)
    pi, pj, q = i32(pi), i32(pj), f64(q)
    fx, fy, fz = zeros3(n)
    e = C.c_double(0.0)
    lib().orc_coulomb_pairs(C.c_int64(len(pi)), P(pi), P(pj), P(x), P(y), P(z), P(cell9(box)),
                            C.c_double(rc), C.c_double(alpha), P(q), P(fx), P(fy), P(fz),
                            C.byref(e))
    return np.stack([fx, fy, fz], 1), e.value


def exclusion(pos, box, alpha, q, ei, ej, mode):
    x, y, z = xyz(pos)
    n = len(x)
    ei, ej, q = i32(ei), i32(ej), f64(q)
    fx, fy, fz = zeros3(n)
    e = C.c_double(0.0)
    lib().orc_exclusion(C.c_int64(len(ei)), P(ei), P(ej), P(x), P(y), P(z), P(cell9(box)),
                        C.c_double(alpha), P(q), C.c_int(mode), P(fx), P(fy), P(fz), C.byref(e))
    return np.stack([fx, fy, fz], 1), e.value


def bonds(pos, box, bi, bj, k, r0):
    x, y, z = xyz(pos)
    n = len(x)
    fx, fy, fz = zeros3(n)
    e = C.c_double(0.0)
    bi, bj, k, r0 = i32(bi), i32(bj), f64(k), f64(r0)
    lib().orc_bonds(C.c_int64(len(bi)), P(bi), P(bj), P(k), P(r0), P(x), P(y), P(z),
                    P(cell9(box)), P(fx), P(fy), P(fz), C.byref(e))
    return np.stack([fx, fy, fz], 1), e.value


def angles(pos, box, ai, aj, ak, kt, t0, mode):
    x, y, z = xyz(pos)
    n = len(x)
    fx, fy, fz = zeros3(n)
    e = C.c_double(0.0)
    ai, aj, ak, kt, t0 = i32(ai), i32(aj), i32(ak), f64(kt), f64(t0)
    lib().orc_angles(C.c_int64(len(ai)), P(ai), P(aj), P(ak), P(kt), P(t0), P(x), P(y), P(z),
                     P(cell9(box)), C.c_int(mode), P(fx), P(fy), P(fz), C.byref(e))
    return np.stack([fx, fy, fz], 1), e.value


def dihedrals(pos, box, di, dj, dk, dl, per, kp, p0, mode):
    x, y, z = xyz(pos)
    n = len(x)
    fx, fy, fz = zeros3(n)
    e = C.c_double(0.0)
    a = [i32(v) for v in (di, dj, dk, dl, per)]
    kp, p0 = f64(kp), f64(p0)
    lib().orc_dihedrals(C.c_int64(len(a[0])), *[P(v) for v in a], P(kp), P(p0), P(x), P(y),
                        P(z), P(cell9(box)), C.c_int(mode), P(fx), P(fy), P(fz), C.byref(e))
    return np.stack([fx, fy, fz], 1), e.value


def impropers(pos, box, di, dj, dk, dl, kp, p0, mode):
    x, y, z = xyz(pos)
    n = len(x)
    fx, fy, fz = zeros3(n)
    e = C.c_double(0.0)
    a = [i32(v) for v in (di, dj, dk, dl)]
    kp, p0 = f64(kp), f64(p0)
    lib().orc_impropers(C.c_int64(len(a[0])), *[P(v) for v in a], P(kp), P(p0), P(x), P(y),
                        P(z), P(cell9(box)), C.c_int(mode), P(fx), P(fy), P(fz), C.byref(e))
    return np.stack([fx, fy, fz], 1), e.value


def theta4(u):
    return lib().orc_theta4(C.c_double(u))


def dtheta4(u):
    return lib().orc_dtheta4(C.c_double(u))


def self_energy(alpha, q, k_e=138.935456):
    q = f64(q)
    return lib().orc_pme_self_energy(C.c_double(alpha), C.c_int64(len(q)), P(q), C.c_double(k_e))


def fft3d(a, sign=-1):
    """Complex 3D DFT of array a with shape (nz, ny, nx)."""
    a = np.asarray(a, dtype=np.complex128)
    nz, ny, nx = a.shape
    re = f64(a.real).copy()
    im = f64(a.imag).copy()
    lib().orc_fft3d(C.c_int(nx), C.c_int(ny), C.c_int(nz), P(re), P(im), C.c_int(sign))
    return re + 1j * im


def pme(pos, box, q, grid, alpha, mode):
    x, y, z = xyz(pos)
    n = len(x)
    q = f64(q)
    fx, fy, fz = zeros3(n)
    e = C.c_double(0.0)
    lib().orc_pme(C.c_int64(n), P(x), P(y), P(z), P(q), P(cell9(box)), C.c_int(grid[0]),
                  C.c_int(grid[1]), C.c_int(grid[2]), C.c_double(alpha), C.c_int(mode), P(fx),
                  P(fy), P(fz), C.byref(e))
    return np.stack([fx, fy, fz], 1), e.value


def ewald(pos, box, q, alpha, rc_real, kmax):
    x, y, z = xyz(pos)
    n = len(x)
    q = f64(q)
    fx, fy, fz = zeros3(n)
    e = C.c_double(0.0)
    lib().orc_ewald(C.c_int64(n), P(x), P(y), P(z), P(q), P(cell9(box)), C.c_double(alpha),
                    C.c_double(rc_real), C.c_int(kmax), P(fx), P(fy), P(fz), C.byref(e))
    return np.stack([fx, fy, fz], 1), e.value


def kinetic_energy(m, v):
    m = f64(m)
    vx, vy, vz = xyz(v)
    return lib().orc_kinetic_energy(C.c_int64(len(m)), P(m), P(vx), P(vy), P(vz))


def temperature(ekin, n):
    return lib().orc_temperature(C.c_double(ekin), C.c_int64(n))


def virial(pos, f):
    x, y, z = xyz(pos)
    fx, fy, fz = xyz(f)
    return lib().orc_virial(C.c_int64(len(x)), P(x), P(y), P(z), P(fx), P(fy), P(fz))


def pressure(ekin, w, volume):
    return lib().orc_pressure(C.c_double(ekin), C.c_double(w), C.c_double(volume))


def integrate_op(kind, m, pos, vel, force, box, dt, target=0.0, tau=0.0, beta=0.0, state=0.0,
                 use_ref=False):
    """kind: 0 half_kick, 1 advance, 2 Berendsen, 3 Nose-Hoover (state=zeta),
    4 Berendsen barostat, 5 Andersen barostat (state=eps_dot).
    Returns (pos, vel, box, state)."""
    m = f64(m)
    n = len(m)
    x, y, z = [a.copy() for a in xyz(pos)]
    vx, vy, vz = [a.copy() for a in xyz(vel)]
    fx, fy, fz = xyz(force)
    b = f64(box).copy()
    st = C.c_double(state)
    if use_ref:
        s = np.array([state])
        ref().ref_apply(C.c_int(kind), C.c_int64(n), P(m), P(x), P(y), P(z), P(vx), P(vy),
                        P(vz), P(fx), P(fy), P(fz), P(b), C.c_double(dt), C.c_double(target),
                        C.c_double(tau), C.c_double(beta), P(s))
        return np.stack([x, y, z], 1), np.stack([vx, vy, vz], 1), b, float(s[0])
    o = lib()
    nn = C.c_int64(n)
    if kind == 0:
        o.orc_half_kick(nn, P(vx), P(vy), P(vz), P(fx), P(fy), P(fz), P(m), C.c_double(dt))
    elif kind == 1:
        o.orc_advance(nn, P(x), P(y), P(z), P(vx), P(vy), P(vz), C.c_double(dt))
    elif kind == 2:
        o.orc_berendsen(nn, P(m), P(vx), P(vy), P(vz), C.c_double(dt), C.c_double(target),
                        C.c_double(tau))
    elif kind == 3:
        o.orc_nose_hoover(nn, P(m), P(vx), P(vy), P(vz), C.c_double(dt), C.c_double(target),
                          C.c_double(tau), C.byref(st))
    elif kind in (4, 5):
        c = cell9(b)
        if kind == 4:
            o.orc_berendsen_barostat(nn, P(m), P(vx), P(vy), P(vz), P(x), P(y), P(z), P(fx),
                                     P(fy), P(fz), P(c), C.c_double(dt), C.c_double(target),
                                     C.c_double(tau), C.c_double(beta))
        else:
            o.orc_andersen_barostat(nn, P(m), P(vx), P(vy), P(vz), P(x), P(y), P(z), P(fx),
                                    P(fy), P(fz), P(c), C.c_double(dt), C.c_double(target),
                                    C.c_double(tau), C.byref(st))
        b = np.array([c[0], c[4], c[8]])
    return np.stack([x, y, z], 1), np.stack([vx, vy, vz], 1), b, st.value


def shake(pos, box, m, ci, cj, dist, tol, use_ref=False):
    """SHAKE positions (src/integrate/constraints.cpp:9-49); returns pos."""
    m = f64(m)
    n = len(m)
    x, y, z = [a.copy() for a in xyz(pos)]
    ci, cj, dist = i32(ci), i32(cj), f64(dist)
    if use_ref:
        v = [np.zeros(n) for _ in range(3)]
        ref().ref_constraints(C.c_int(0), C.c_int64(len(ci)), P(ci), P(cj), P(dist),
                              C.c_double(tol), C.c_int64(n), P(m), P(x), P(y), P(z),
                              *[P(a) for a in v], P(f64(box)))
    else:
        lib().orc_shake(C.c_int64(len(ci)), P(ci), P(cj), P(dist), C.c_double(tol),
                        P(cell9(box)), P(m), P(x), P(y), P(z))
    return np.stack([x, y, z], 1)


def rattle(pos, vel, box, m, ci, cj, tol, use_ref=False):
    """RATTLE velocities (src/integrate/constraints.cpp:51-91); returns vel."""
    m = f64(m)
    n = len(m)
    x, y, z = xyz(pos)
    vx, vy, vz = [a.copy() for a in xyz(vel)]
    ci, cj = i32(ci), i32(cj)
    if use_ref:
        xs = [a.copy() for a in (x, y, z)]
        d = np.ones(len(ci))
        ref().ref_constraints(C.c_int(1), C.c_int64(len(ci)), P(ci), P(cj), P(d),
                              C.c_double(tol), C.c_int64(n), P(m), *[P(a) for a in xs],
                              P(vx), P(vy), P(vz), P(f64(box)))
    else:
        lib().orc_rattle(C.c_int64(len(ci)), P(ci), P(cj), C.c_double(tol), P(cell9(box)),
                         P(m), P(x), P(y), P(z), P(vx), P(vy), P(vz))
    return np.stack([vx, vy, vz], 1)


class _Mt64(C.Structure):
    _fields_ = [("mt", C.c_uint64 * 312), ("idx", C.c_int)]


def mt64(seed, count):
    """The first `count` outputs of mt19937_64 seeded with `seed`."""
    g = _Mt64()
    o = lib()
    o.orc_mt64_next.restype = C.c_uint64
    o.orc_mt64_seed(C.byref(g), C.c_uint64(seed))
    return np.array([o.orc_mt64_next(C.byref(g)) for _ in range(count)], dtype=np.uint64)


def andersen(m, vel, dt, T0, frequency, calls=0, use_ref=False):
    """One Andersen apply() (src/thermostat/andersen.cpp:14-32) after `calls`
    earlier calls; returns vel."""
    m = f64(m)
    n = len(m)
    vx, vy, vz = [a.copy() for a in xyz(vel)]
    if use_ref:
        ref().ref_andersen(C.c_int64(n), P(m), P(vx), P(vy), P(vz), C.c_double(dt),
                           C.c_double(T0), C.c_double(frequency), C.c_int64(calls))
    else:
        lib().orc_andersen(C.c_int64(n), P(m), P(vx), P(vy), P(vz), C.c_double(dt),
                           C.c_double(T0), C.c_double(frequency),
                           C.c_uint32((42 + calls) & 0xFFFFFFFF))
    return np.stack([vx, vy, vz], 1)


# ------------------------------------------------------------- simulation

class Sim:
    """Step-level oracle over a dmd_amd.system.SystemSpec."""

    def __init__(self, spec):
        self.spec = spec
        self._desc = spec.to_desc()
        self.n = spec.n
        self.h = lib().orc_sim_create(C.byref(self._desc))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_sim_destroy(C.c_void_p(self.h))
            self.h = None

    def step(self, nsteps=1):
        lib().orc_sim_step(C.c_void_p(self.h), C.c_int64(nsteps))

    def compute(self):
        lib().orc_sim_compute(C.c_void_p(self.h))

    def forces_exact(self):
        fx, fy, fz = zeros3(self.n)
        pe = C.c_double(0.0)
        lib().orc_sim_forces_exact(C.c_void_p(self.h), P(fx), P(fy), P(fz), C.byref(pe))
        return np.stack([fx, fy, fz], 1), pe.value

    def state(self):
        a = [np.zeros(self.n) for _ in range(9)]
        lib().orc_sim_get(C.c_void_p(self.h), *[P(v) for v in a])
        s = np.zeros(11)
        lib().orc_sim_scalars(C.c_void_p(self.h), P(s))
        keys = ["pe", "ke", "step", "time", "zeta", "eps_dot", "box_x", "box_y", "box_z",
                "counter", "n_pairs"]
        out = dict(zip(keys, s.tolist()))
        out["pos"] = np.stack(a[0:3], 1)
        out["vel"] = np.stack(a[3:6], 1)
        out["force"] = np.stack(a[6:9], 1)
        return out

    def set_positions(self, pos):
        x, y, z = xyz(pos)
        lib().orc_sim_set_positions(C.c_void_p(self.h), P(x), P(y), P(z))

    def set_velocities(self, vel):
        x, y, z = xyz(vel)
        lib().orc_sim_set_velocities(C.c_void_p(self.h), P(x), P(y), P(z))


class RefSim:
    """DMD's own components (oracle/_ref) driven in SimulationEngine::step
    order; PME-free specs only."""

    def __init__(self, spec):
        if spec.coulomb_type == 2:
            raise ValueError("the reference PME translation unit is not buildable here")
        self.n = n = spec.n
        self._keep = []

        def k(a, f):
            if a is None:
                return None
            a = f(a)
            self._keep.append(a)
            return P(a)

        x, y, z = xyz(spec.pos)
        vx, vy, vz = xyz(spec.vel) if spec.vel is not None else zeros3(n)
        self._keep += [x, y, z, vx, vy, vz]
        sig = f64(np.atleast_2d(spec.lj_sigma if spec.use_lj else [[1.0]])).ravel().copy()
        eps = f64(np.atleast_2d(spec.lj_epsilon if spec.use_lj else [[0.0]])).ravel().copy()
        self._keep += [sig, eps]
        nt = int(round(np.sqrt(len(sig))))
        b = spec.bonds if spec.bonds is not None else ([], [], [], [])
        a = spec.angles if spec.angles is not None else ([], [], [], [], [])
        d = spec.dihedrals if spec.dihedrals is not None else ([], [], [], [], [], [], [])
        e = spec.exclusions if spec.exclusions is not None else ([], [])
        box = f64(spec.box)
        self._keep.append(box)
        r = ref()
        self.h = r.ref_sim_create(
            C.c_int64(n), P(x), P(y), P(z), P(vx), P(vy), P(vz), k(spec.masses, f64),
            k(spec.charges, f64), k(spec.atom_types, i32), P(box), C.c_int(int(spec.use_lj)),
            C.c_int32(nt), P(sig), P(eps), C.c_double(spec.lj_cutoff),
            C.c_int64(len(b[0])), k(b[0], i32), k(b[1], i32), k(b[2], f64), k(b[3], f64),
            C.c_int64(len(a[0])), k(a[0], i32), k(a[1], i32), k(a[2], i32), k(a[3], f64),
            k(a[4], f64),
            C.c_int64(len(d[0])), k(d[0], i32), k(d[1], i32), k(d[2], i32), k(d[3], i32),
            k(d[4], i32), k(d[5], f64), k(d[6], f64),
            C.c_int64(len(e[0])), k(e[0], i32), k(e[1], i32),
            C.c_int(1 if spec.coulomb_type == 1 else 0), C.c_double(spec.coulomb_cutoff),
            C.c_double(spec.ewald_coeff), C.c_double(spec.dt), C.c_int(spec.thermostat),
            C.c_double(spec.thermo_temperature), C.c_double(spec.thermo_tau),
            C.c_int(spec.barostat), C.c_double(spec.baro_pressure), C.c_double(spec.baro_tau),
            C.c_double(spec.baro_compressibility))
        if spec.constraints is not None and len(spec.constraints[0]):
            ci, cj, cd = i32(spec.constraints[0]), i32(spec.constraints[1]), f64(spec.constraints[2])
            r.ref_sim_set_constraints(C.c_void_p(self.h), C.c_int64(len(ci)), P(ci), P(cj), P(cd),
                                      C.c_double(spec.constraint_tolerance))
        if spec.thermostat == 3:
            if spec.andersen_calls:
                raise ValueError("the reference thermostat starts at seed 42")
            r.ref_sim_set_andersen(C.c_void_p(self.h), C.c_double(spec.thermo_temperature),
                                   C.c_double(spec.andersen_frequency))

    def __del__(self):
        if getattr(self, "h", None):
            ref().ref_sim_destroy(C.c_void_p(self.h))
            self.h = None

    def step(self, nsteps=1):
        ref().ref_sim_step(C.c_void_p(self.h), C.c_int64(nsteps))

    def state(self):
        a = [np.zeros(self.n) for _ in range(9)]
        s = np.zeros(4)
        ref().ref_sim_get(C.c_void_p(self.h), *[P(v) for v in a], P(s))
        return {"pos": np.stack(a[0:3], 1), "vel": np.stack(a[3:6], 1),
                "force": np.stack(a[6:9], 1), "pe": s[0], "ke": s[1], "box_x": s[2],
                "counter": s[3]}
