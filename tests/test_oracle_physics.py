"""Pins of the oracle by what the mathematics fixes: finite differences of
the energy, closed forms, brute force, a plain Ewald sum and numpy's FFT.

A plausible mistake (dropped term, wrong sign, transposed operand) in any
restated function fails at least one of these."""
import math

import numpy as np
import pytest

FA, CO = 0, 1


def fd_forces(energy_fn, pos, h=1e-6):
    """-dE/dx by central differences."""
    F = np.zeros_like(pos)
    for i in range(pos.shape[0]):
        for d in range(3):
            p = pos.copy()
            p[i, d] += h
            ep = energy_fn(p)
            p[i, d] -= 2 * h
            em = energy_fn(p)
            F[i, d] = -(ep - em) / (2 * h)
    return F


def cos_sim(a, b):
    return float((a * b).sum() / (np.linalg.norm(a) * np.linalg.norm(b)))


def rand_cluster(rng, n, box, spread=0.35):
    c = np.array(box) / 2
    return c + rng.uniform(-spread, spread, (n, 3))


def test_pairs_bruteforce_and_cells(orc):
    rng = np.random.default_rng(1)
    box = (3.1, 2.9, 3.3)
    pos = rng.uniform(-1.0, 4.0, (400, 3))  # unwrapped, as DMD never wraps
    pi, pj = orc.pairs(pos, box, 0.9)
    d = pos[None, :, :] - pos[:, None, :]
    L = np.array(box)
    d = d - np.round(d / L) * L  # numpy round is half-to-even; only ties differ
    r2 = (d * d).sum(-1)
    iu = np.triu_indices(len(pos), 1)
    mask = r2[iu] < 0.81
    assert set(zip(pi.tolist(), pj.tolist())) == set(zip(iu[0][mask], iu[1][mask]))
    assert np.all(np.diff(pi.astype(np.int64) * 10**6 + pj) > 0)  # lexicographic
    ci, cj = orc.pairs(pos, box, 0.9, "cells")
    assert np.array_equal(pi, ci) and np.array_equal(pj, cj)


def test_lj_and_coulomb_gradients(orc):
    rng = np.random.default_rng(2)
    box = (2.5, 2.5, 2.5)
    pos = rand_cluster(rng, 6, box)
    pos[:, 0] += 1.0  # straddle the boundary to exercise the minimum image
    sig, eps = [[0.3]], [[0.8]]
    F, _ = orc.lj(pos, box, 1.1, sig, eps)
    Ffd = fd_forces(lambda p: orc.lj(p, box, 1.1, sig, eps)[1], pos)
    assert np.allclose(F, Ffd, rtol=1e-5, atol=1e-4)
    q = rng.choice([-1.0, 1.0], 6)
    F, _ = orc.coulomb(pos, box, 1.1, 3.0, q)
    Ffd = fd_forces(lambda p: orc.coulomb(p, box, 1.1, 3.0, q)[1], pos)
    assert np.allclose(F, Ffd, rtol=1e-5, atol=1e-4)


def test_exclusion_gradient_and_faithful_sign(orc):
    rng = np.random.default_rng(3)
    box = (3.0, 3.0, 3.0)
    pos = rand_cluster(rng, 4, box, 0.15)
    q = np.array([-0.8, 0.4, 0.4, 0.3])
    ei, ej = [0, 0, 1], [1, 2, 2]
    F, _ = orc.exclusion(pos, box, 3.2, q, ei, ej, CO)
    Ffd = fd_forces(lambda p: orc.exclusion(p, box, 3.2, q, ei, ej, CO)[1], pos)
    assert np.allclose(F, Ffd, rtol=1e-5, atol=1e-4)
    # FAITHFUL: energy -k q q / r but force of +k q q / r (cos = -1 with -grad E)
    F, _ = orc.exclusion(pos, box, 3.2, q, ei, ej, FA)
    Ffd = fd_forces(lambda p: orc.exclusion(p, box, 3.2, q, ei, ej, FA)[1], pos)
    assert np.allclose(F, -Ffd, rtol=1e-5, atol=1e-4)


def test_angle_gradient(orc):
    rng = np.random.default_rng(4)
    box = (5.0, 5.0, 5.0)
    pos = rand_cluster(rng, 3, box, 0.2)
    args = ([0], [1], [2], [317.5], [math.radians(113.24)])
    F, _ = orc.angles(pos, box, *args, CO)
    Ffd = fd_forces(lambda p: orc.angles(p, box, *args, CO)[1], pos)
    assert np.allclose(F, Ffd, rtol=1e-5, atol=1e-4)
    F, _ = orc.angles(pos, box, *args, FA)
    assert cos_sim(F, Ffd) < -0.999999  # anti-restoring as written in DMD


@pytest.mark.parametrize("kind", ["dihedral", "improper"])
def test_torsion_gradient(orc, kind):
    rng = np.random.default_rng(5)
    box = (5.0, 5.0, 5.0)
    for trial in range(5):
        pos = rand_cluster(rng, 4, box, 0.25)
        if kind == "dihedral":
            fn = lambda p, m: orc.dihedrals(p, box, [0], [1], [2], [3], [3], [2.0], [0.3], m)
        else:
            fn = lambda p, m: orc.impropers(p, box, [0], [1], [2], [3], [40.0], [0.2], m)
        F, _ = fn(pos, CO)
        Ffd = fd_forces(lambda p: fn(p, CO)[1], pos)
        assert np.allclose(F, Ffd, rtol=1e-5, atol=1e-5)
        # FAITHFUL conserves momentum but is not -grad E
        Ff, _ = fn(pos, FA)
        assert np.all(np.abs(Ff.sum(0)) < 1e-9)
        assert abs(cos_sim(Ff, Ffd)) < 0.9


def test_fft_matches_numpy(orc):
    rng = np.random.default_rng(6)
    for shape in [(6, 10, 12), (7, 5, 9), (8, 8, 8), (3, 14, 15)]:
        a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        assert np.allclose(orc.fft3d(a, -1), np.fft.fftn(a), atol=1e-10)
        assert np.allclose(orc.fft3d(a, +1), np.fft.ifftn(a) * a.size, atol=1e-10)


def _charges(rng, n):
    q = rng.uniform(-1, 1, n)
    return q - q.mean()


def test_pme_corrected_vs_ewald(orc):
    rng = np.random.default_rng(7)
    box = (2.0, 2.2, 2.4)
    pos = rng.uniform(0, 2.0, (24, 3)) * np.array(box) / 2.0
    q = _charges(rng, 24)
    alpha = 3.0
    Fe, Ee = orc.ewald(pos, box, q, alpha, 0.0, 14)  # reciprocal + self only
    errs = []
    for K in [(20, 22, 24), (40, 44, 48)]:
        Fp, Ep = orc.pme(pos, box, q, K, alpha, CO)
        errs.append((abs(Ep - Ee), np.sqrt(((Fp - Fe) ** 2).sum() / (Fe ** 2).sum())))
    # order-4 interpolation: the error falls ~16x when the grid spacing halves
    assert errs[0][0] / errs[1][0] > 12 and errs[0][1] / errs[1][1] > 6
    assert errs[1][0] / abs(Ee) < 5e-5 and errs[1][1] < 5e-4


def test_ewald_total_forces_gradient(orc):
    rng = np.random.default_rng(8)
    box = (1.5, 1.5, 1.5)
    pos = rng.uniform(0, 1.5, (5, 3))
    q = _charges(rng, 5)
    F, _ = orc.ewald(pos, box, q, 2.5, 2.4, 8)
    Ffd = fd_forces(lambda p: orc.ewald(p, box, q, 2.5, 2.4, 8)[1], pos, h=1e-5)
    assert np.allclose(F, Ffd, rtol=1e-5, atol=1e-3)


def test_ewald_madelung_nacl(orc):
    # NaCl rock salt: E per ion pair = -M k_e / r0 with M = 1.747565
    a = 1.0
    pos, q = [], []
    for i in range(2):
        for j in range(2):
            for k in range(2):
                for b, s in [((0, 0, 0), 1), ((0.5, 0.5, 0), 1), ((0.5, 0, 0.5), 1),
                             ((0, 0.5, 0.5), 1), ((0.5, 0, 0), -1), ((0, 0.5, 0), -1),
                             ((0, 0, 0.5), -1), ((0.5, 0.5, 0.5), -1)]:
                    pos.append([(i + b[0]) * a, (j + b[1]) * a, (k + b[2]) * a])
                    q.append(s)
    pos, q = np.array(pos), np.array(q, float)
    box = (2 * a,) * 3
    _, E = orc.ewald(pos, box, q, 4.0, 0.99, 10)
    assert abs(E / 32 - (-1.747565 * 138.935456 / 0.5)) / (1.747565 * 138.935456 / 0.5) < 1e-5
    _, Ep = orc.pme(pos, box, q, (64, 64, 64), 4.0, CO)
    Fr, Er = orc.coulomb(pos, box, 0.99, 4.0, q)
    assert abs((Ep + Er) / 32 - (-1.747565 * 138.935456 / 0.5)) / (1.747565 * 138.935456 / 0.5) < 1e-4


def test_pme_corrected_gradient(orc):
    rng = np.random.default_rng(9)
    box = (2.0, 2.0, 2.0)
    pos = rng.uniform(0, 2.0, (6, 3))
    q = _charges(rng, 6)
    F, _ = orc.pme(pos, box, q, (20, 20, 20), 2.8, CO)
    Ffd = fd_forces(lambda p: orc.pme(p, box, q, (20, 20, 20), 2.8, CO)[1], pos, h=1e-6)
    assert np.allclose(F, Ffd, rtol=1e-4, atol=1e-3)


def test_pme_faithful_relations(orc):
    # FAITHFUL forces are -grad(E_faithful) / (nx ny nz) (quirk Q7) and the
    # faithful energy multiplies |b|^2-like factors instead of dividing (Q6)
    rng = np.random.default_rng(10)
    box = (2.0, 2.0, 2.0)
    pos = rng.uniform(0, 2.0, (5, 3))
    q = _charges(rng, 5)
    K = (12, 12, 12)
    F, Ef = orc.pme(pos, box, q, K, 2.8, FA)
    Ffd = fd_forces(lambda p: orc.pme(p, box, q, K, 2.8, FA)[1], pos, h=1e-6)
    assert np.allclose(F * np.prod(K), Ffd, rtol=1e-4, atol=1e-3)
    _, Ec = orc.pme(pos, box, q, K, 2.8, CO)
    _, Ee = orc.ewald(pos, box, q, 2.8, 0.0, 12)
    assert abs(Ec - Ee) < abs(Ef - Ee)


def test_thermostat_closed_forms(orc):
    rng = np.random.default_rng(11)
    n = 50
    m = rng.uniform(1, 40, n)
    v = rng.standard_normal((n, 3))
    z = np.zeros((n, 3))
    kb = 0.008314462618
    ke = 0.5 * (m[:, None] * v * v).sum()
    T = 2 * ke / ((3 * n - 3) * kb)
    lam = math.sqrt(1 + 0.002 / 0.1 * (300.0 / T - 1))
    _, v2, _, _ = orc.integrate_op(2, m, z, v, z, (3, 3, 3), 0.002, 300.0, 0.1)
    assert np.allclose(v2, v * lam, rtol=1e-14)
    Q = (3 * n - 3) * kb * 300.0 * 0.01
    zeta = 0.1 + 0.5 * 0.002 * kb * (T - 300.0) / Q
    _, v3, _, z3 = orc.integrate_op(3, m, z, v, z, (3, 3, 3), 0.002, 300.0, 0.1, state=0.1)
    assert abs(z3 - zeta) < 1e-14 * abs(zeta)
    assert np.allclose(v3, v * math.exp(-zeta * 0.001), rtol=1e-14)
