"""Non-orthorhombic cells (CellType, src/core/cell.h:3-8).

DMD's minimum image is the identity for every cell type but orthorhombic
(src/core/cell.cpp:26-38) and Cell::volume is the 3x3 determinant (:19-24).
The oracle under a triclinic cell is checked bit for bit against vectors of
DMD's own translation units (tests/golden/make_golden.py, triclinic block),
and against what the identity image implies: the same forces as in a box so
large that no image ever enters."""
import os

import numpy as np

import oracle as O
from dmd_amd import _abi

FA = _abi.FAITHFUL
G = np.load(os.path.join(os.path.dirname(__file__), "golden", "dmd_ref_vectors.npz"))


def same(a, b):
    return np.array_equal(np.asarray(a), np.asarray(b))


def test_volume_is_the_determinant_for_every_type():
    for m, vols in zip(G["cellvol_m"], G["cellvol_V"]):
        for v in vols:
            assert O.cell_volume(m) == v
        assert abs(O.cell_volume(m) - np.linalg.det(np.asarray(m).reshape(3, 3))) <= 1e-12 * 40


def test_triclinic_identity_image_golden():
    pos, box, q = G["tric_pos"], G["tric_box"], G["tric_q"]
    with O.cell_type("triclinic"):
        assert same(O.min_image(box, [2.0, -1.7, 1.3]), [2.0, -1.7, 1.3])
        F, E = O.lj(pos, box, 0.9, [[0.3]], [[0.5]])
        assert same(F, G["tric_lj_F"]) and E == G["tric_lj_E"]
        F, E = O.coulomb(pos, box, 0.8, 3.3, q)
        assert same(F, G["tric_coul_F"]) and E == G["tric_coul_E"]
        F, E = O.exclusion(pos, box, 3.3, q, G["tric_excl_i"], G["tric_excl_j"], FA)
        assert same(F, G["tric_excl_F"]) and E == G["tric_excl_E"]
        d = G["tric_quad"]
        qi, qj, qk, ql = (O.i32(d[:, c]) for c in range(4))
        F, E = O.bonds(pos, box, qi, qj, G["tric_bk"], G["tric_br"])
        assert same(F, G["tric_bond_F"]) and E == G["tric_bond_E"]
        F, E = O.angles(pos, box, qi, qj, qk, G["tric_kt"], G["tric_t0"], FA)
        assert same(F, G["tric_angle_F"]) and E == G["tric_angle_E"]
        F, E = O.dihedrals(pos, box, qi, qj, qk, ql, G["tric_per"], G["tric_kp"], G["tric_p0"],
                           FA)
        assert same(F, G["tric_dih_F"]) and E == G["tric_dih_E"]
        F, E = O.impropers(pos, box, qi, qj, qk, ql, G["tric_kp"], G["tric_p0"], FA)
        assert same(F, G["tric_imp_F"]) and E == G["tric_imp_E"]


def test_identity_image_equals_an_isolated_system():
    """With the identity image, a system is isolated: its LJ forces equal the
    orthorhombic forces in a box too large for any image to matter, and they
    differ from the periodic forces in the small box (atoms reach past it)."""
    pos, box = G["tric_pos"], G["tric_box"]
    with O.cell_type("triclinic"):
        F_open, E_open = O.lj(pos, box, 0.9, [[0.3]], [[0.5]])
    big = (100.0, 100.0, 100.0)
    F_big, E_big = O.lj(pos + 40.0, big, 0.9, [[0.3]], [[0.5]])
    assert np.allclose(F_open, F_big, rtol=1e-12, atol=1e-9) and abs(E_open - E_big) <= 1e-9 * abs(E_big)
    F_per, _ = O.lj(pos, box, 0.9, [[0.3]], [[0.5]])
    assert not np.allclose(F_per, F_open)
