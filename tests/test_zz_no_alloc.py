"""dmdx_component_compute_noalloc against dmdx_component_compute.

The allocation-free entry point takes device arrays and lays the
component's batch out in the library's per-GPU scratch array.  It runs the
same kernels in the same order, so forces and energy must equal the
host-array entry point bit for bit -- on two calls in a row, the second with
a smaller system, with non-zero forces already in the output arrays
(accumulate semantics), and with other components' calls in between (the
scratch array keeps the previous call's contents)."""
import numpy as np
import pytest

from dmd_amd import _abi
from dmd_amd import components as D
from dmd_amd import workloads as W

pytestmark = pytest.mark.gpu

FA, CO, FP64, MIXED = _abi.FAITHFUL, _abi.CORRECTED, _abi.FP64, _abi.MIXED


def say(*a):
    print(*a, flush=True)


def _dev(a):
    import torch
    return torch.tensor(np.ascontiguousarray(a, dtype=np.float64), device="cuda:0",
                        dtype=torch.float64)


def _both(comp, pos, box, rng):
    """Run the host-array and the allocation-free entry points from the same
    starting forces and energy; return both results."""
    import torch
    n = len(pos)
    cell = D.Cell(*box)
    f0 = rng.normal(0.0, 10.0, (n, 3))
    e0 = float(rng.normal(0.0, 100.0))
    fx, fy, fz = (np.ascontiguousarray(f0[:, k]) for k in range(3))
    e_ref = comp.compute(pos[:, 0].copy(), pos[:, 1].copy(), pos[:, 2].copy(), cell, fx, fy, fz, e0)
    X, Y, Z = (_dev(pos[:, k]) for k in range(3))
    FX, FY, FZ = (_dev(f0[:, k]) for k in range(3))
    E = _dev([e0])
    torch.cuda.synchronize()
    comp.compute_noalloc(n, X.data_ptr(), Y.data_ptr(), Z.data_ptr(), cell, FX.data_ptr(),
                         FY.data_ptr(), FZ.data_ptr(), E.data_ptr())
    torch.cuda.synchronize()
    F_ref = np.stack([fx, fy, fz], 1)
    F = np.stack([FX.cpu().numpy(), FY.cpu().numpy(), FZ.cpu().numpy()], 1)
    return F, float(E.cpu().numpy()[0]), F_ref, e_ref


def _same(F, E, F_ref, E_ref, what):
    assert np.array_equal(F, F_ref), (what, float(np.abs(F - F_ref).max()))
    assert E == E_ref, (what, E, E_ref)


def _water(n_mol, seed):
    return W.water_box(n_mol=n_mol, seed=seed)


@pytest.mark.parametrize("physics,precision", [(CO, FP64), (FA, FP64), (CO, MIXED)])
def test_noalloc_pme_equals_compute(physics, precision):
    say(f"case: noalloc PME physics {physics} precision {precision}, 900 then 648 atoms")
    rng = np.random.default_rng(5)
    big, small = _water(300, 1), _water(216, 2)
    K = big.pme_grid
    pm = D.CoulombPME(D.PMEParams(big.coulomb_cutoff, big.ewald_coeff, K[0], K[1], K[2], 4))
    pm.set_mode(physics, precision)
    for spec in (big, small):  # two calls in a row, the second smaller
        pm.set_charges(spec.charges)
        F, E, F_ref, E_ref = _both(pm, spec.pos, spec.box, rng)
        _same(F, E, F_ref, E_ref, spec.name)


def test_noalloc_bonded_equals_compute():
    say("case: noalloc bonds, angles, dihedrals, impropers, exclusions; two sizes each")
    rng = np.random.default_rng(6)
    for spec in (W.chain_melt(1200, 3), W.chain_melt(700, 4)):
        b, a, d = spec.bonds, spec.angles, spec.dihedrals
        comps = [
            D.HarmonicBond(D.HarmonicBondParams(len(b[0]), b[0], b[1], b[2], b[3])),
            D.HarmonicAngle(D.HarmonicAngleParams(len(a[0]), a[0], a[1], a[2], a[3], a[4])),
            D.PeriodicDihedral(D.PeriodicDihedralParams(len(d[0]), d[0], d[1], d[2], d[3], d[4],
                                                        d[5], d[6])),
            D.HarmonicImproper(D.ImproperParams(len(d[0]), d[0], d[1], d[2], d[3], d[5] * 50.0,
                                                np.full(len(d[0]), 0.3))),
        ]
        q = rng.uniform(-1, 1, spec.n)
        ex = D.CoulombExclusion()
        ex.set_charges(q)
        ex.set_exclusions(*spec.exclusions)
        comps.append(ex)
        for physics in (FA, CO):
            for c in comps:
                c.set_mode(physics, FP64)
                F, E, F_ref, E_ref = _both(c, spec.pos, spec.box, rng)
                _same(F, E, F_ref, E_ref, (c.name(), spec.name, physics))


def test_noalloc_refusals():
    say("case: noalloc refuses list components and systems beyond the scratch array")
    import torch
    lj = D.LennardJones(D.LennardJonesParams(1, [0.3], [0.5]), 1.0)
    x = torch.zeros(8, dtype=torch.float64, device="cuda:0")
    with pytest.raises(D.ForceException):
        lj.compute_noalloc(8, x.data_ptr(), x.data_ptr(), x.data_ptr(), D.Cell(3, 3, 3),
                           x.data_ptr(), x.data_ptr(), x.data_ptr(), x.data_ptr())
    # a 256^3 PME grid needs more than the scratch array holds
    pm = D.CoulombPME(D.PMEParams(1.0, 3.0, 256, 256, 256, 4))
    pm.set_charges(np.array([1.0, -1.0] * 4))
    pos = torch.rand(8, dtype=torch.float64, device="cuda:0")
    f = torch.zeros(8, dtype=torch.float64, device="cuda:0")
    e = torch.zeros(1, dtype=torch.float64, device="cuda:0")
    with pytest.raises(D.ForceException) as err:
        pm.compute_noalloc(8, pos.data_ptr(), pos.data_ptr(), pos.data_ptr(), D.Cell(5, 5, 5),
                           f.data_ptr(), f.data_ptr(), f.data_ptr(), e.data_ptr())
    assert "scratch" in str(err.value)
    torch.cuda.synchronize()
    assert float(e.cpu().numpy()[0]) == 0.0
