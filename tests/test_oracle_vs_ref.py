"""FAITHFUL oracle == DMD's own translation units, bit for bit.

oracle/_ref/libdmdref.so is built from the reference sources by
oracle/build_ref.sh (with the same -ffp-contract=off as the oracle).  It is
not present where the reference tree is absent; these tests then skip and
the oracle stays pinned by the known-answer and physics tests."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")

FA = 0


def rand_sys(seed, n=60, L=2.6):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-0.5 * L, 1.5 * L, (n, 3))  # unwrapped
    return rng, pos, (L, L * 1.05, L * 0.97)


def ref_call(fn, pos, box, *args):
    x, y, z = O.xyz(pos)
    n = len(x)
    fx, fy, fz = O.zeros3(n)
    e = np.zeros(1)
    b = O.f64(box)
    fn(C.c_int64(n), O.P(x), O.P(y), O.P(z), O.P(b), *args, O.P(fx), O.P(fy), O.P(fz), O.P(e))
    return np.stack([fx, fy, fz], 1), e[0]


def same(a, b):
    return np.array_equal(np.asarray(a), np.asarray(b))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_lj_bitexact(seed):
    rng, pos, box = rand_sys(seed, 300, 3.0)
    nt = 2
    sig = O.f64(rng.uniform(0.25, 0.35, (nt, nt))).ravel().copy()
    eps = O.f64(rng.uniform(0.2, 1.0, (nt, nt))).ravel().copy()
    types = O.i32(rng.integers(0, nt, len(pos)))
    F, E = O.lj(pos, box, 0.9, sig.reshape(nt, nt), eps.reshape(nt, nt), types)
    Fr, Er = ref_call(O.ref().ref_lj, pos, box, C.c_double(0.9), C.c_int32(nt), O.P(sig),
                      O.P(eps), O.P(types))
    assert same(F, Fr) and E == Er


def test_coulomb_direct_and_exclusion_bitexact():
    rng, pos, box = rand_sys(3, 200, 3.0)
    q = O.f64(rng.uniform(-1, 1, len(pos)))
    F, E = O.coulomb(pos, box, 1.0, 3.1, q)
    Fr, Er = ref_call(O.ref().ref_coulomb_direct, pos, box, C.c_double(1.0), C.c_double(3.1),
                      O.P(q), C.c_int(1))
    assert same(F, Fr) and E == Er
    # without rebuild_pair_list the reference list is empty (production path)
    Fr, Er = ref_call(O.ref().ref_coulomb_direct, pos, box, C.c_double(1.0), C.c_double(3.1),
                      O.P(q), C.c_int(0))
    assert not Fr.any() and Er == 0.0
    ei = O.i32(rng.integers(0, len(pos), 50))
    ej = O.i32((ei + 1 + rng.integers(0, 5, 50)) % len(pos))
    F, E = O.exclusion(pos, box, 3.1, q, ei, ej, FA)
    Fr, Er = ref_call(O.ref().ref_exclusion, pos, box, O.P(q), C.c_int64(len(ei)), O.P(ei),
                      O.P(ej))
    assert same(F, Fr) and E == Er


def test_bonded_bitexact():
    rng, pos, box = rand_sys(4, 40, 2.0)
    n = len(pos)
    nb = 60
    bi, bj = O.i32(rng.integers(0, n, nb)), O.i32(rng.integers(0, n, nb))
    bj = O.i32(np.where(bi == bj, (bj + 1) % n, bj))
    bk, br = O.f64(rng.uniform(100, 1000, nb)), O.f64(rng.uniform(0.1, 0.3, nb))
    F, E = O.bonds(pos, box, bi, bj, bk, br)
    Fr, Er = ref_call(O.ref().ref_bonds, pos, box, C.c_int64(nb), O.P(bi), O.P(bj), O.P(bk),
                      O.P(br))
    assert same(F, Fr) and E == Er
    trip = np.array([rng.permutation(n)[:3] for _ in range(nb)], np.int32)
    ai, aj, ak = (O.i32(trip[:, c]) for c in range(3))
    kt, t0 = O.f64(rng.uniform(50, 400, nb)), O.f64(rng.uniform(1.5, 2.5, nb))
    F, E = O.angles(pos, box, ai, aj, ak, kt, t0, FA)
    Fr, Er = ref_call(O.ref().ref_angles, pos, box, C.c_int64(nb), O.P(ai), O.P(aj), O.P(ak),
                      O.P(kt), O.P(t0))
    assert same(F, Fr) and E == Er
    quad = np.array([rng.permutation(n)[:4] for _ in range(nb)], np.int32)
    di, dj, dk, dl = (O.i32(quad[:, c]) for c in range(4))
    per = O.i32(rng.integers(1, 4, nb))
    kp, p0 = O.f64(rng.uniform(1, 10, nb)), O.f64(rng.uniform(-math.pi, math.pi, nb))
    F, E = O.dihedrals(pos, box, di, dj, dk, dl, per, kp, p0, FA)
    Fr, Er = ref_call(O.ref().ref_dihedrals, pos, box, C.c_int64(nb), O.P(di), O.P(dj),
                      O.P(dk), O.P(dl), O.P(per), O.P(kp), O.P(p0))
    assert same(F, Fr) and E == Er
    F, E = O.impropers(pos, box, di, dj, dk, dl, kp, p0, FA)
    Fr, Er = ref_call(O.ref().ref_impropers, pos, box, C.c_int64(nb), O.P(di), O.P(dj),
                      O.P(dk), O.P(dl), O.P(kp), O.P(p0))
    assert same(F, Fr) and E == Er


@pytest.mark.parametrize("kind", [0, 1, 2, 3, 4, 5])
def test_integration_ops_bitexact(kind):
    rng = np.random.default_rng(5 + kind)
    n = 77
    m = rng.uniform(1, 40, n)
    pos = rng.uniform(0, 3, (n, 3))
    v = rng.standard_normal((n, 3))
    f = rng.standard_normal((n, 3)) * 100
    args = (m, pos, v, f, (3.0, 3.1, 3.2), 0.002, 250.0, 0.1, 4.5e-5, 0.05)
    a = O.integrate_op(kind, *args)
    b = O.integrate_op(kind, *args, use_ref=True)
    assert same(a[0], b[0]) and same(a[1], b[1]) and same(a[2], b[2]) and a[3] == b[3]


def test_scalar_reductions_bitexact():
    rng = np.random.default_rng(12)
    n = 123
    m = O.f64(rng.uniform(1, 40, n))
    v = rng.standard_normal((n, 3))
    vx, vy, vz = O.xyz(v)
    ke = O.kinetic_energy(m, v)
    assert ke == O.ref().ref_kinetic_energy(C.c_int64(n), O.P(m), O.P(vx), O.P(vy), O.P(vz))
    assert O.temperature(ke, n) == O.ref().ref_temperature(C.c_double(ke), C.c_int64(n))
    pos = rng.uniform(0, 3, (n, 3))
    f = rng.standard_normal((n, 3))
    x, y, z = O.xyz(pos)
    fx, fy, fz = O.xyz(f)
    assert O.virial(pos, f) == O.ref().ref_virial(C.c_int64(n), O.P(x), O.P(y), O.P(z), O.P(fx),
                                                  O.P(fy), O.P(fz))
    assert O.pressure(1.5, 2.5, 27.0) == O.ref().ref_pressure(C.c_double(1.5), C.c_double(2.5),
                                                              C.c_double(27.0))


def _spec_lj_small(thermostat=0, barostat=0):
    from dmd_amd import workloads as W
    s = W.argon_fcc(n_cells=3, seed=7)
    return s.with_(thermostat=thermostat, thermo_temperature=90.0, thermo_tau=0.1,
                   barostat=barostat, baro_pressure=1.0, baro_tau=1.0, lj_cutoff=0.75)


@pytest.mark.parametrize("thermo,baro", [(0, 0), (1, 0), (2, 0), (0, 1), (2, 2)])
def test_step_sequence_bitexact(thermo, baro):
    """The oracle's SimulationEngine::step sequence equals DMD's components
    driven in the same order, for 60 steps (12 list rebuilds)."""
    spec = _spec_lj_small(thermo, baro)
    a, b = O.Sim(spec), O.RefSim(spec)
    a.step(60)
    b.step(60)
    sa, sb = a.state(), b.state()
    for k in ("pos", "vel", "force"):
        assert same(sa[k], sb[k]), k
    assert sa["pe"] == sb["pe"] and sa["box_x"] == sb["box_x"]
    assert sa["counter"] == sb["counter"]


def test_step_bonded_bitexact():
    """A faithful bonded mini-system (chain melt topology) through the step."""
    from dmd_amd import workloads as W
    spec = W.chain_melt(120, seed=3).with_(physics=FA, thermostat=1, list_skin=0.0)
    a, b = O.Sim(spec), O.RefSim(spec)
    a.step(20)
    b.step(20)
    sa, sb = a.state(), b.state()
    # faithful LJ ignores exclusions, as the reference does (Q2)
    for k in ("pos", "vel", "force"):
        assert same(sa[k], sb[k]), k
