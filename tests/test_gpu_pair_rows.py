"""Row-tile LJ pair kernel (k_pair_rows, tuning key 6): neighbour rows staged
in LDS, list blocks dealt evenly over the workgroup's lanes, partial sums
combined in a fixed order.  Checked against the per-atom kernel (key 6 = 0)
on the same lists, against the FP64 path after steps, alone against in a
batch, and on the tile kinds the kernel distinguishes: staged rows, rows too
large to stage (global reads), the FP32 minimum image of small boxes (fewer
than three cells along an edge), and typed LJ tables."""
import numpy as np
import pytest

from dmd_amd import _abi
from dmd_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def _systems():
    mix = W.replica_mix(n_replicas=6, n_min=500, n_max=4000, seed=11, precision=_abi.MIXED)
    # small box: 200 atoms, about 2 cells of rlist per edge (minimum image)
    small = W.lj_fluid(200, 0.8, 1.0, seed=3, precision=_abi.MIXED)
    # 20k atoms: nine rows of cells exceed the staging capacity
    large = W.lj_fluid(20000, 0.8442, 0.72, seed=5, cutoff=0.85, skin=0.1,
                       precision=_abi.MIXED)
    # two LJ types (water's O and H without charges)
    typed = W.water_box(n_mol=1000, seed=9, precision=_abi.MIXED,
                        coulomb_type=_abi.COULOMB_NONE, charges=np.zeros(3000))
    return mix + [small, large, typed]


def _rel(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / max((b ** 2).sum(), 1e-300)))


def _compute(specs, rows, steps=0):
    from dmd_amd.engine import Engine
    e = Engine(specs)
    e.set_tuning(pair_rows=rows)
    if steps:
        e.step(steps)
    e.compute()
    return e


def test_rows_match_per_atom_kernel():
    specs = _systems()
    say(f"case: {len(specs)} systems, row tiles 4/8 vs the per-atom kernel")
    ref = _compute(specs, 0)
    for rows in (4, 8):
        e = _compute(specs, rows)
        for k, s in enumerate(specs):
            fr, fe = ref.forces(k), e.forces(k)
            err = _rel(fe, fr)
            pr, pe = ref.scalars(k)["pe"], e.scalars(k)["pe"]
            say(f"rows {rows} system {k} ({s.name}, n={s.n}): force rel {err:.2e}, "
                f"pe {pe:.9g} vs {pr:.9g}")
            assert err <= 2e-6, (rows, k, err)
            assert abs(pe - pr) <= 1e-6 * max(1.0, abs(pr)), (rows, k, pe, pr)


def test_rows_vs_fp64_after_steps():
    from dmd_amd.engine import Engine
    specs = _systems()
    say("case: row tiles (8) after 20 steps, forces vs FP64 at the reached positions")
    e = _compute(specs, 8, steps=20)
    for k, s in enumerate(specs):
        f64 = Engine(s.with_(precision=_abi.FP64, pos=e.positions(k), vel=e.velocities(k)))
        f64.compute()
        err = _rel(e.forces(k), f64.forces(0))
        say(f"system {k} ({s.name}): force rel {err:.2e}")
        assert err <= 1e-5, (k, err)


def test_rows_alone_equals_in_mix():
    from dmd_amd.engine import Engine
    specs = _systems()
    say("case: row tiles, each system alone vs in the batch, 10 steps, bit for bit")
    mix = Engine(specs)
    mix.set_tuning(pair_rows=8)
    mix.step(10)
    for k in (0, 3, len(specs) - 3, len(specs) - 1):
        alone = Engine([specs[k]])
        alone.set_tuning(pair_rows=8)
        alone.step(10)
        assert np.array_equal(alone.positions(0), mix.positions(k)), k
        assert np.array_equal(alone.forces(0), mix.forces(k)), k
        assert alone.scalars(0)["pe"] == mix.scalars(k)["pe"], k
