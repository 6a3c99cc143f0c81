"""The FAITHFUL oracle against committed vectors of DMD's own translation
units (tests/golden/dmd_ref_vectors.npz, written by tests/golden/
make_golden.py from oracle/_ref, the reference sources compiled here).

The same comparisons run live in test_oracle_vs_ref.py and
test_oracle_constraints.py where oracle/_ref can be built; these keep the
oracle pinned bit for bit where the reference tree is absent."""
import os

import numpy as np
import pytest

import oracle as O
from dmd_amd import _abi
from dmd_amd import workloads as W

FA = _abi.FAITHFUL
G = np.load(os.path.join(os.path.dirname(__file__), "golden", "dmd_ref_vectors.npz"))


def same(a, b):
    return np.array_equal(np.asarray(a), np.asarray(b))


def test_lj():
    F, E = O.lj(G["lj_pos"], G["lj_box"], float(G["lj_rc"]), G["lj_sig"], G["lj_eps"],
                G["lj_types"])
    assert same(F, G["lj_F"]) and E == G["lj_E"]


def test_coulomb_direct_and_exclusion():
    pos, box, q = G["coul_pos"], G["coul_box"], G["coul_q"]
    F, E = O.coulomb(pos, box, float(G["coul_rc"]), float(G["coul_alpha"]), q)
    assert same(F, G["coul_F"]) and E == G["coul_E"]
    F, E = O.exclusion(pos, box, float(G["coul_alpha"]), q, G["excl_i"], G["excl_j"], FA)
    assert same(F, G["excl_F"]) and E == G["excl_E"]


def test_bonded():
    pos, box = G["bnd_pos"], G["bnd_box"]
    F, E = O.bonds(pos, box, G["bond_i"], G["bond_j"], G["bond_k"], G["bond_r0"])
    assert same(F, G["bond_F"]) and E == G["bond_E"]
    t = G["angle_ijk"]
    F, E = O.angles(pos, box, O.i32(t[:, 0]), O.i32(t[:, 1]), O.i32(t[:, 2]), G["angle_k"],
                    G["angle_t0"], FA)
    assert same(F, G["angle_F"]) and E == G["angle_E"]
    d = G["dih_ijkl"]
    di, dj, dk, dl = (O.i32(d[:, c]) for c in range(4))
    F, E = O.dihedrals(pos, box, di, dj, dk, dl, G["dih_n"], G["dih_k"], G["dih_p0"], FA)
    assert same(F, G["dih_F"]) and E == G["dih_E"]
    F, E = O.impropers(pos, box, di, dj, dk, dl, G["dih_k"], G["dih_p0"], FA)
    assert same(F, G["imp_F"]) and E == G["imp_E"]


@pytest.mark.parametrize("kind", range(6))
def test_integration_ops(kind):
    p = f"integ{kind}_"
    out = O.integrate_op(kind, G[p + "m"], G[p + "pos"], G[p + "vel"], G[p + "f"],
                         (3.0, 3.1, 3.2), 0.002, 250.0, 0.1, 4.5e-5, 0.05)
    for k in range(4):
        assert same(out[k], G[p + f"out{k}"]), k


def test_scalar_reductions():
    m, vel = G["red_m"], G["red_vel"]
    ke = O.kinetic_energy(m, vel)
    assert ke == G["red_ke"]
    assert O.temperature(ke, len(m)) == G["red_T"]
    assert O.virial(G["red_pos"], G["red_f"]) == G["red_virial"]
    assert O.pressure(1.5, 2.5, 27.0) == G["red_pressure"]


def _check_state(sim, p, keys=("pos", "vel", "force")):
    st = sim.state()
    for k in keys:
        assert same(st[k], G[p + k]), k
    return st


@pytest.mark.parametrize("thermo,baro", [(0, 0), (1, 0), (2, 0), (0, 1), (2, 2)])
def test_step_sequence(thermo, baro):
    p = f"step{thermo}{baro}_"
    spec = W.argon_fcc(n_cells=3, seed=7).with_(
        thermostat=thermo, thermo_temperature=90.0, thermo_tau=0.1, barostat=baro,
        baro_pressure=1.0, baro_tau=1.0, lj_cutoff=0.75, pos=G[p + "pos0"], vel=G[p + "vel0"])
    sim = O.Sim(spec)
    sim.step(60)
    st = _check_state(sim, p)
    assert st["pe"] == G[p + "pe"] and st["box_x"] == G[p + "box_x"]
    assert st["counter"] == G[p + "counter"]


def test_step_chain_faithful():
    spec = W.chain_melt(120, seed=3).with_(physics=FA, thermostat=1, list_skin=0.0)
    assert same(spec.pos, G["chain_pos0"]), "chain generator changed: regenerate the vectors"
    spec = spec.with_(vel=G["chain_vel0"])
    sim = O.Sim(spec)
    sim.step(20)
    _check_state(sim, "chain_")


@pytest.mark.parametrize("tol", [1e-4, 1e-14])
def test_shake_rattle_chain(tol):
    pos, m = G["shake_pos"], G["shake_m"]
    ci, cj, cd = G["shake_ci"], G["shake_cj"], G["shake_cd"]
    big = (1e3, 1e3, 1e3)
    sp = O.shake(pos, big, m, ci, cj, cd, tol)
    assert same(sp, G[f"shake_out_{tol:g}"])
    assert same(O.rattle(sp, G["shake_vel"], big, m, ci, cj, tol), G[f"rattle_out_{tol:g}"])


@pytest.mark.parametrize("freq,calls", [(500.0, 0), (25.0, 3)])
def test_andersen(freq, calls):
    p = f"and{int(freq)}_{calls}_"
    out = O.andersen(G[p + "m"], G[p + "vel"], 0.002, 300.0, freq, calls=calls)
    # the normals go through log(); the collided atoms are the same, their
    # velocities agree to the last bits of glibc's log
    assert np.array_equal(out == G[p + "vel"], G[p + "out"] == G[p + "vel"])
    assert np.allclose(out, G[p + "out"], rtol=1e-14, atol=0.0)


def test_step_rigid_water():
    spec = W.water_box(n_mol=64, seed=11, rigid=True, dt=0.002).with_(
        physics=FA, coulomb_type=_abi.COULOMB_NONE, charges=None, exclusions=None,
        list_skin=0.0, rebuild_interval=10, thermostat=_abi.THERMO_BERENDSEN)
    assert same(spec.pos, G["water_pos0"]), "water generator changed: regenerate the vectors"
    spec = spec.with_(vel=G["water_vel0"])
    sim = O.Sim(spec)
    sim.step(30)
    _check_state(sim, "water_")
