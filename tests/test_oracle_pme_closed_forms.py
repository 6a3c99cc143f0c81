"""Closed-form pins of the oracle's particle-mesh Ewald energy.

Charges placed exactly on grid points make the spread grid known in closed
form: the centred order-4 B-spline takes the values 1/6, 2/3, 1/6 at the
offsets -1, 0, +1, so along one axis the grid's DFT at mode m is the exact
structure factor times b(m) = 2/3 + cos(2 pi m / K) / 3.  The reciprocal
energy of DMD's loop (src/force/coulomb_pme.cpp:143-201) is then

    E_rec = 1/2 sum_{m != 0} G(m) B(m) prod_d b(m_d)^2 |S(m)|^2,
    G(m)  = 4 pi k_e / V exp(-k^2 / (4 alpha^2)) / k^2,  k = 2 pi |m / L|,

with DMD's B(m) = prod_d (sin t_d / t_d)^(2p), t_d = pi |m_d| / K_d, p = 4
(coulomb_pme.cpp:159-163; quirk Q6, FAITHFUL) or the textbook
B(m) = prod_d b(m_d)^-2 (CORRECTED), for which E_rec is the Ewald
reciprocal sum itself.  S(m) is summed directly over the charges here, so
neither the spreading nor the FFT of the oracle enters the expected value.

The neutralising-background term (CORRECTED only; DMD has none, quirk Q9)
is pinned by one charge in a cubic box: the Ewald energy of a simple cubic
lattice of charges q in a uniform compensating background is
-k_e q^2 xi / (2 L), xi = 2.837297479 (the Wigner constant of the simple
cubic lattice), for every splitting parameter alpha.
"""
import math

import numpy as np
import pytest

FA, CO = 0, 1
KE = 138.935456
XI_SC = 2.837297479480620


def _modes(K):
    """Representative frequency of each DFT bin, as the oracle and DMD pick it
    (k <= K/2 -> k, else k - K)."""
    k = np.arange(K)
    return np.where(k <= K // 2, k, k - K)


def _closed_form_rec(pos, q, box, K, alpha, mode, p_exp=8, t_scale=1.0):
    """Direct structure-factor sum of the reciprocal energy for charges on grid
    points.  p_exp and t_scale exist only to show that a mistaken B factor
    (exponent p instead of 2p, or the wrong argument t) moves the result."""
    L = np.asarray(box, float)
    V = float(np.prod(L))
    mx, my, mz = np.meshgrid(_modes(K[0]), _modes(K[1]), _modes(K[2]), indexing="ij")
    m = np.stack([mx, my, mz], -1).reshape(-1, 3).astype(float)
    m = m[np.any(m != 0, axis=1)]
    mL = m / L
    k2 = 4.0 * math.pi ** 2 * (mL ** 2).sum(1)
    green = 4.0 * math.pi * KE / V * np.exp(-k2 / (4.0 * alpha ** 2)) / k2
    bt = np.prod(2.0 / 3.0 + np.cos(2.0 * math.pi * m / np.asarray(K, float)) / 3.0, axis=1)
    if mode == FA:
        t = t_scale * math.pi * np.abs(m) / np.asarray(K, float)
        safe = np.where(t == 0, 1.0, t)
        B = np.prod(np.where(t == 0, 1.0, (np.sin(safe) / safe) ** p_exp), axis=1)
    else:
        B = 1.0 / bt ** 2
    phase = 2.0 * math.pi * (mL @ np.asarray(pos, float).T)  # (modes, atoms)
    S = (q[None, :] * np.exp(-1j * phase)).sum(1)
    return 0.5 * float((green * B * bt ** 2 * np.abs(S) ** 2).sum())


def _on_grid(rng, n, box, K):
    g = np.stack([rng.integers(0, K[d], n) for d in range(3)], 1)
    return g * (np.asarray(box) / np.asarray(K))


def test_pme_faithful_b_factor_closed_form(orc):
    """FAITHFUL E_rec equals the structure-factor sum with DMD's multiplied
    (sin t / t)^(2p) factor, on even and odd grids of unequal sides."""
    rng = np.random.default_rng(21)
    for box, K, alpha in [((2.0, 2.2, 1.8), (12, 16, 10), 2.6),
                          ((1.7, 1.7, 2.3), (15, 13, 18), 3.1)]:
        pos = _on_grid(rng, 7, box, K)
        q = rng.uniform(-1.0, 1.0, 7)  # not neutral: FAITHFUL has no background term
        _, E = orc.pme(pos, box, q, K, alpha, FA)
        e_rec = E - orc.self_energy(alpha, q)
        want = _closed_form_rec(pos, q, box, K, alpha, FA)
        assert abs(e_rec - want) <= 1e-11 * abs(want), (e_rec, want)
        # a plausible slip in the B factor is far outside that tolerance
        for p_exp, t_scale in [(4, 1.0), (8, 2.0), (6, 1.0)]:
            wrong = _closed_form_rec(pos, q, box, K, alpha, FA, p_exp, t_scale)
            assert abs(wrong - want) > 1e-4 * abs(want)


def test_pme_corrected_on_grid_is_ewald_reciprocal(orc):
    """CORRECTED divides by |b|^2, so for charges on grid points E_rec is the
    plain Ewald reciprocal sum over the grid's modes."""
    rng = np.random.default_rng(22)
    box, K, alpha = (2.0, 2.2, 1.8), (12, 16, 10), 2.6
    pos = _on_grid(rng, 6, box, K)
    q = rng.uniform(-1.0, 1.0, 6)
    q -= q.mean()  # neutral: no background term in the comparison
    _, E = orc.pme(pos, box, q, K, alpha, CO)
    e_rec = E - orc.self_energy(alpha, q)
    want = _closed_form_rec(pos, q, box, K, alpha, CO)
    assert abs(e_rec - want) <= 1e-11 * abs(want), (e_rec, want)


def test_pme_background_single_charge_wigner(orc):
    """One charge in a cubic box: CORRECTED PME (real space vanishes at
    alpha L >= 5) gives the simple-cubic Wigner energy for every alpha, which
    holds only with the -pi k_e Q^2 / (2 V alpha^2) background term."""
    L, K = 1.6, (32, 32, 32)
    for q0 in (1.0, -0.82):
        for aL in (5.0, 6.5):
            alpha = aL / L
            pos = np.array([[0.25 * L, 0.5 * L, 0.75 * L]])  # on grid points of 32^3
            _, E = orc.pme(pos, (L, L, L), np.array([q0]), K, alpha, CO)
            want = -KE * q0 * q0 * XI_SC / (2.0 * L)
            assert abs(E - want) <= 1e-9 * abs(want), (aL, E, want)
            background = -math.pi * KE * q0 * q0 / (2.0 * L ** 3 * alpha ** 2)
            assert abs(background) > 1e4 * 1e-9 * abs(want)  # the term is resolved
    # off the grid the same holds to the spline interpolation error
    pos = np.array([[0.123, 0.771, 1.402]])
    _, E = orc.pme(pos, (L, L, L), np.array([1.0]), K, 5.5 / L, CO)
    want = -KE * XI_SC / (2.0 * L)
    assert abs(E - want) <= 1e-4 * abs(want), (E, want)


def test_pme_faithful_has_no_background(orc):
    """DMD adds no background term (quirk Q9): a non-neutral FAITHFUL energy
    is the closed-form reciprocal sum plus the self energy, nothing else."""
    L, K = 1.6, (16, 16, 16)
    pos = np.array([[0.2, 0.4, 0.6], [1.0, 1.1, 0.3]])
    q = np.array([1.0, 0.5])
    for alpha in (3.0, 4.0):
        _, E = orc.pme(pos, (L, L, L), q, K, alpha, FA)
        want = _closed_form_rec(pos, q, (L, L, L), K, alpha, FA) + orc.self_energy(alpha, q)
        assert abs(E - want) <= 1e-11 * abs(want)


# ---------------------------------------------------------------- device

def _device_pme(pos, box, q, K, alpha, physics, precision=0):
    from dmd_amd import components as D
    pos = np.asarray(pos, float)
    s = D.SystemData(len(pos))
    s.pos_x[:], s.pos_y[:], s.pos_z[:] = pos[:, 0], pos[:, 1], pos[:, 2]
    pm = D.CoulombPME(D.PMEParams(1.0, alpha, K[0], K[1], K[2], 4))
    pm.set_mode(physics, precision)
    pm.set_charges(q)
    e = pm.compute(s.pos_x, s.pos_y, s.pos_z, D.Cell(*box), s.forces_x, s.forces_y,
                   s.forces_z, 0.0)
    return np.stack([s.forces_x, s.forces_y, s.forces_z], 1), e


@pytest.mark.gpu
def test_device_pme_closed_forms():
    """The HIP CoulombPME meets the same closed forms: DMD's multiplied B
    factor (FAITHFUL, non-neutral, on-grid charges) and the Wigner energy of
    one charge (CORRECTED, background term on the device)."""
    print("case: device PME closed forms (FAITHFUL B factor, CORRECTED background)", flush=True)
    rng = np.random.default_rng(23)
    box, K, alpha = (2.0, 2.2, 1.8), (12, 16, 10), 2.6
    pos = _on_grid(rng, 7, box, K)
    q = rng.uniform(-1.0, 1.0, 7)
    _, E = _device_pme(pos, box, q, K, alpha, FA)
    want = _closed_form_rec(pos, q, box, K, alpha, FA) + _self(alpha, q)
    assert abs(E - want) <= 1e-10 * abs(want), (E, want)
    L = 1.6
    for aL in (5.0, 6.5):
        _, E = _device_pme([[0.4, 0.8, 1.2]], (L, L, L), np.array([-0.82]), (32, 32, 32),
                           aL / L, CO)
        want = -KE * 0.82 ** 2 * XI_SC / (2.0 * L)
        assert abs(E - want) <= 1e-9 * abs(want), (aL, E, want)


@pytest.mark.gpu
def test_device_pme_non_neutral_vs_oracle(orc):
    """A non-neutral random charge set: device forces and energy against the
    oracle in both modes (FP64 within 1e-10; mixed within 1e-5)."""
    print("case: device PME non-neutral set vs oracle", flush=True)
    rng = np.random.default_rng(24)
    box, K, alpha = (2.4, 2.0, 2.2), (24, 20, 22), 3.1
    pos = rng.uniform(-0.5, 2.9, (200, 3))
    q = rng.uniform(-1.0, 1.0, 200) + 0.01  # net charge about +2 e
    for physics in (FA, CO):
        F0, E0 = orc.pme(pos, box, q, K, alpha, physics)
        F, E = _device_pme(pos, box, q, K, alpha, physics)
        assert np.abs(F - F0).max() <= 1e-10 * np.abs(F0).max()
        assert abs(E - E0) <= 1e-10 * abs(E0)
        Fm, Em = _device_pme(pos, box, q, K, alpha, physics, precision=1)
        rms = np.sqrt(((Fm - F0) ** 2).sum() / (F0 ** 2).sum())
        assert rms <= 1e-5, rms
        assert abs(Em - E0) <= 1e-5 * abs(E0)


def _self(alpha, q):
    return -KE * alpha * float((np.asarray(q) ** 2).sum()) / math.sqrt(math.pi)
