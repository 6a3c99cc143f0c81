"""Non-orthorhombic cells on the device component path.

DMD's Cell(matrix, type) (src/core/cell.h:10-23) wraps by the matrix diagonal
only for orthorhombic cells; for triclinic, hexagonal and rhombic-dodecahedral
cells its minimum image is the identity (src/core/cell.cpp:26-38), and its
volume is the determinant of the matrix for every type (:19-24).  PME takes the
box edges from the diagonal and the volume from Cell::volume
(src/force/coulomb_pme.cpp:86-88, :148-151).

Every component under a triclinic cell is compared with the oracle run under
the same cell type (the oracle is itself pinned bit for bit against DMD's own
translation units under a triclinic cell, tests/test_oracle_cell_types.py),
and PME's volume dependence is checked against what SPME fixes: with the
same diagonal, energy and forces scale as 1/V."""
import numpy as np
import pytest

import oracle as O
from dmd_amd import components as D

pytestmark = pytest.mark.gpu

# a matrix whose determinant (26.34) differs from the product of its diagonal
M_TRI = np.array([[3.0, 0.4, 0.0], [0.3, 3.1, 0.2], [0.0, 0.5, 2.9]])
DIAG = (3.0, 3.1, 2.9)


def say(*a):
    print(*a, flush=True)


def system(pos):
    pos = np.asarray(pos, dtype=float)
    s = D.SystemData(len(pos))
    s.pos_x[:], s.pos_y[:], s.pos_z[:] = pos[:, 0], pos[:, 1], pos[:, 2]
    return s


def run(comp, s, cell, e0=0.0):
    return comp.compute(s.pos_x, s.pos_y, s.pos_z, cell, s.forces_x, s.forces_y, s.forces_z, e0)


def forces(s):
    return np.stack([s.forces_x, s.forces_y, s.forces_z], 1)


def close(F, G, rel):
    return np.abs(F - G).max() <= rel * np.abs(G).max()


def positions(seed, n):
    # atoms also outside the box: DMD never wraps them, and with the identity
    # image a pair's distance is the plain difference of its positions
    rng = np.random.default_rng(seed)
    return rng, rng.uniform(-0.3, 1.3, (n, 3)) * np.asarray(DIAG)


@pytest.mark.parametrize("ctype", ["triclinic", "hexagonal", "rhombic_dodecahedron"])
def test_lj_identity_image_vs_oracle(ctype):
    say(f"case: LennardJones under a {ctype} cell, 700 atoms")
    rng, pos = positions(11, 700)
    types = rng.integers(0, 2, len(pos)).astype(np.int32)
    sig = np.array([[0.30, 0.32], [0.32, 0.34]])
    eps = np.array([[0.5, 0.7], [0.7, 0.9]])
    cell = D.Cell.from_matrix(M_TRI, D.CellType[ctype])
    lj = D.LennardJones(D.LennardJonesParams(2, sig, eps), 0.9)
    lj.set_atom_types(types)
    s = system(pos)
    lj.rebuild_pair_list(s.x(), s.y(), s.z(), cell)
    e = run(lj, s, cell)
    with O.cell_type(ctype):
        F, E = O.lj(pos, M_TRI, 0.9, sig, eps, types)
    assert close(forces(s), F, 1e-12) and abs(e - E) <= 1e-12 * abs(E)
    # and the image really is the identity: the orthorhombic cell of the same
    # diagonal gives other forces
    s2 = system(pos)
    orth = D.Cell(*DIAG)
    lj.rebuild_pair_list(s2.x(), s2.y(), s2.z(), orth)
    run(lj, s2, orth)
    assert not close(forces(s2), F, 1e-3)


def test_lj_identity_image_equals_isolated_system():
    """Size-independent property at 6000 atoms: under the identity image a
    system is isolated, so its forces equal those in an orthorhombic box so
    large that no image enters (positions shifted into it)."""
    say("case: LennardJones identity image vs isolated system, 6000 atoms")
    rng = np.random.default_rng(5)
    L = 4.0
    pos = rng.uniform(0.0, L, (6000, 3))
    sig, eps = np.array([[0.3]]), np.array([[0.5]])
    lj = D.LennardJones(D.LennardJonesParams(1, sig, eps), 1.0)
    tri = D.Cell.from_matrix([[L, 0.0, 0.0], [0.7, L, 0.0], [0.2, 0.4, L]], D.CellType.triclinic)
    s = system(pos)
    lj.rebuild_pair_list(s.x(), s.y(), s.z(), tri)
    e = run(lj, s, tri)
    big = D.Cell(60.0, 60.0, 60.0)
    lj2 = D.LennardJones(D.LennardJonesParams(1, sig, eps), 1.0)
    s2 = system(pos + 25.0)
    lj2.rebuild_pair_list(s2.x(), s2.y(), s2.z(), big)
    e2 = run(lj2, s2, big)
    assert close(forces(s), forces(s2), 1e-11) and abs(e - e2) <= 1e-11 * abs(e2)


def test_coulomb_and_exclusion_vs_oracle():
    say("case: CoulombDirect / CoulombExclusion under a triclinic cell, 400 atoms")
    rng, pos = positions(12, 400)
    q = rng.uniform(-1, 1, len(pos))
    cell = D.Cell.from_matrix(M_TRI, D.CellType.triclinic)
    cd = D.CoulombDirect(D.CoulombDirectParams(1.0, 3.1))
    cd.set_charges(q)
    s = system(pos)
    cd.rebuild_pair_list(s.x(), s.y(), s.z(), cell)
    e = run(cd, s, cell)
    with O.cell_type("triclinic"):
        F, E = O.coulomb(pos, M_TRI, 1.0, 3.1, q)
    assert close(forces(s), F, 1e-12) and abs(e - E) <= 1e-12 * abs(E)
    ei = np.arange(0, 300, 2, dtype=np.int32)
    ej = (ei * 7 + 3) % len(pos)
    ej = ej.astype(np.int32)
    for physics in (0, 1):
        ex = D.CoulombExclusion(physics=physics)
        ex.set_ewald_coefficient(3.1)
        ex.set_charges(q)
        ex.set_exclusions(ei, ej)
        s = system(pos)
        e = run(ex, s, cell)
        with O.cell_type("triclinic"):
            F, E = O.exclusion(pos, M_TRI, 3.1, q, ei, ej, physics)
        assert close(forces(s), F, 1e-12) and abs(e - E) <= 1e-12 * abs(E)


def test_bonded_vs_oracle():
    say("case: bond / angle / dihedral / improper under a triclinic cell, 600 atoms")
    rng, pos = positions(13, 600)
    n = 500
    quad = np.array([rng.choice(len(pos), 4, replace=False) for _ in range(n)], dtype=np.int32)
    bi, bj, bk, bl = (np.ascontiguousarray(quad[:, c]) for c in range(4))
    k1 = rng.uniform(50.0, 400.0, n)
    k2 = rng.uniform(0.2, 2.0, n)
    per = rng.integers(1, 4, n).astype(np.int32)
    cell = D.Cell.from_matrix(M_TRI, D.CellType.triclinic)
    for physics in (0, 1):
        comps = [
            (D.HarmonicBond(D.HarmonicBondParams(n, bi, bj, k1, k2)),
             lambda: O.bonds(pos, M_TRI, bi, bj, k1, k2)),
            (D.HarmonicAngle(D.HarmonicAngleParams(n, bi, bj, bk, k1, k2)),
             lambda: O.angles(pos, M_TRI, bi, bj, bk, k1, k2, physics)),
            (D.PeriodicDihedral(D.PeriodicDihedralParams(n, bi, bj, bk, bl, per, k1, k2)),
             lambda: O.dihedrals(pos, M_TRI, bi, bj, bk, bl, per, k1, k2, physics)),
            (D.HarmonicImproper(D.ImproperParams(n, bi, bj, bk, bl, k1, k2)),
             lambda: O.impropers(pos, M_TRI, bi, bj, bk, bl, k1, k2, physics)),
        ]
        for comp, ref in comps:
            comp.set_mode(physics)
            s = system(pos)
            e = run(comp, s, cell)
            with O.cell_type("triclinic"):
                F, E = ref()
            assert close(forces(s), F, 1e-11), comp.name()
            assert abs(e - E) <= 1e-11 * abs(E), comp.name()


@pytest.mark.parametrize("physics", [0, 1])
def test_pme_volume_is_the_determinant(physics):
    say(f"case: PME with an off-diagonal cell matrix, physics {physics}, 500 atoms")
    rng = np.random.default_rng(14)
    pos = rng.uniform(0.0, 1.0, (500, 3)) * np.asarray(DIAG)
    q = rng.uniform(-1, 1, len(pos))
    q -= q.mean()
    grid = (24, 25, 24)
    res = {}
    for ctype in ("orthorhombic", "triclinic"):
        pm = D.CoulombPME(D.PMEParams(1.0, 3.1, *grid, 4))
        pm.set_mode(physics)
        pm.set_charges(q)
        cell = D.Cell.from_matrix(M_TRI, D.CellType[ctype])
        s = system(pos)
        e = run(pm, s, cell)
        res[ctype] = (forces(s), e)
    with O.cell_type("triclinic"):
        F, E = O.pme(pos, M_TRI, q, grid, 3.1, physics)
    for ctype, (Fd, e) in res.items():
        assert close(Fd, F, 1e-10) and abs(e - E) <= 1e-10 * abs(E), ctype
    # SPME: the volume is a prefactor of the reciprocal sum, the edges come
    # from the diagonal -- the same system in the diagonal cell scales by V
    # (the self energy, -k_e alpha/sqrt(pi) sum q^2, does not depend on it)
    pm = D.CoulombPME(D.PMEParams(1.0, 3.1, *grid, 4))
    pm.set_mode(physics)
    pm.set_charges(q)
    s = system(pos)
    e_diag = run(pm, s, D.Cell(*DIAG))
    ratio = (DIAG[0] * DIAG[1] * DIAG[2]) / np.linalg.det(M_TRI)
    assert abs(ratio - 1.0) > 1e-3
    e_self = O.self_energy(3.1, q)
    assert abs((res["triclinic"][1] - e_self) - (e_diag - e_self) * ratio) <= 1e-11 * abs(e_diag)
    assert close(res["triclinic"][0], forces(s) * ratio, 1e-11)


def test_mixed_precision_and_noalloc_refuse_open_cells():
    say("case: MIXED and compute_noalloc refuse non-orthorhombic cells")
    cell = D.Cell.from_matrix(M_TRI, D.CellType.triclinic)
    b = D.HarmonicBond(D.HarmonicBondParams(1, [0], [1], [100.0], [0.3]))
    b.set_mode(0, 1)
    s = system([[0.1, 0.1, 0.1], [0.4, 0.1, 0.1]])
    with pytest.raises(D.ForceException):
        run(b, s, cell)
    with pytest.raises(D.ForceException):
        b.compute_noalloc(2, 0, 0, 0, cell, 0, 0, 0, 0)
    # the orthorhombic cell of the same matrix still runs in MIXED
    run(b, s, D.Cell.from_matrix(M_TRI, D.CellType.orthorhombic))
