"""Staged per-cell list build (k_nbr_build_staged, tuning key 7): each cell's
candidate stream copied into LDS once, scanned from there by 8 or 16 lanes
per atom, in passes of up to 1024 candidates.  Its rows must be the rows of
the other builds entry for entry: the pair set equals the oracle's, and the
forces are bit-identical to those of the per-cell register build and the
per-atom build.  Water (half-width cells, about 1500 candidates per cell)
takes two passes over its stream; the chain melt has excluded partners."""
import numpy as np
import pytest

import oracle as O
from dmd_amd import _abi
from dmd_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def _spec(kind):
    if kind == "argon":
        return W.lj_fluid(4000, 0.8442, 0.72, seed=3, cutoff=0.85, skin=0.1,
                          precision=_abi.MIXED)
    if kind == "water":
        return W.water_box(n_side=10, seed=2, precision=_abi.MIXED)
    return W.chain_melt(3000, seed=9, precision=_abi.MIXED)


@pytest.mark.parametrize("kind", ["argon", "water", "chain"])
def test_staged_build_rows_equal_other_builds(kind):
    from dmd_amd.engine import Engine
    spec = _spec(kind)
    say(f"case: staged list build, {kind} ({spec.n} atoms)")
    rl = max(spec.lj_cutoff, spec.coulomb_cutoff if spec.coulomb_type else 0.0) + spec.list_skin
    oi, oj = O.pairs(spec.pos, spec.box, rl, "cells")
    if spec.exclusions is not None:
        oi, oj = O.filter_excl(oi, oj, *spec.exclusions)
    forces = {}
    for name, kw in (("stage8", dict(build_cells=True, build_stage=8)),
                     ("stage16", dict(build_cells=True, build_stage=16)),
                     ("cell4", dict(build_cells=True, build_stage=0, build_tasks=4)),
                     ("atom16", dict(build_cells=False, build_lanes=16))):
        e = Engine(spec)
        e.set_tuning(**kw)
        e.compute()
        gi, gj = e.pairs()
        assert np.array_equal(gi, oi) and np.array_equal(gj, oj), name
        forces[name] = e.forces()
        say(f"  {name}: {len(gi)} pairs")
    for name, f in forces.items():
        assert np.array_equal(f, forces["atom16"]), name


def test_staged_build_along_a_run():
    """Twenty steps of a small replica mix with the staged build and with the
    per-atom build: partial rebuilds of some systems each step, the same
    trajectory bit for bit."""
    from dmd_amd.engine import Engine
    specs = W.replica_mix(n_replicas=8, n_min=500, n_max=3000, seed=5, precision=_abi.MIXED)
    say(f"case: staged vs per-atom build along 20 steps, {len(specs)} systems")
    a = Engine(specs)
    a.set_tuning(build_cells=True, build_stage=8)
    b = Engine(specs)
    b.set_tuning(build_cells=False, build_lanes=16)
    a.step(20)
    b.step(20)
    for k in range(len(specs)):
        assert np.array_equal(a.positions(k), b.positions(k)), k
        assert a.scalars(k)["list_builds"] == b.scalars(k)["list_builds"], k
