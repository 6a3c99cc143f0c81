"""HIP step engine (dmd_amd/libdmdx.so) against the FP64 oracle.

Bars: the FP64 plain path reproduces the oracle trajectory within 1e-10
relative forces over 100 steps (an LJ-only system is bit-identical); the
mixed path is within 1e-5 force RMS relative error of the CORRECTED FP64
oracle on identical positions."""
import numpy as np
import pytest

import oracle as O
from dmd_amd import _abi
from dmd_amd import workloads as W

pytestmark = pytest.mark.gpu

FA, CO = _abi.FAITHFUL, _abi.CORRECTED


def say(*a):
    print(*a, flush=True)


def engine(spec, **kw):
    from dmd_amd.engine import Engine
    return Engine(spec, **kw)


def rel_max(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def rms_rel(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def jittered_argon(n_cells=10, skin=0.0, seed=42, amp=0.01):
    s = W.argon_fcc(n_cells=n_cells, skin=skin, seed=seed)
    rng = np.random.default_rng(seed)
    return s.with_(pos=s.pos + rng.uniform(-amp, amp, s.pos.shape))


def test_c0_pair_set_and_forces_bitexact():
    say("case: argon 4000 FP64 faithful, pair set + forces")
    spec = jittered_argon()
    e = engine(spec)
    e.compute()
    gi, gj = e.pairs()
    oi, oj = O.pairs(spec.pos, spec.box, spec.lj_cutoff)
    assert np.array_equal(gi, oi) and np.array_equal(gj, oj)
    sim = O.Sim(spec)
    sim.compute()
    st = sim.state()
    F = e.forces()
    assert np.array_equal(F, st["force"]), rel_max(F, st["force"])
    assert abs(e.scalars()["pe"] - st["pe"]) <= 1e-12 * abs(st["pe"])


@pytest.mark.parametrize("skin", [0.0, 0.1])
def test_c0_trajectory_100_steps(skin):
    say(f"case: argon 4000 FP64 faithful, 100 steps, skin {skin}")
    spec = jittered_argon(skin=skin)
    e = engine(spec)
    sim = O.Sim(spec)
    worst = 0.0
    for block in range(4):
        e.step(25)
        sim.step(25)
        st = sim.state()
        worst = max(worst, rel_max(e.forces(), st["force"]))
        assert rel_max(e.positions(), st["pos"]) < 1e-12
    assert worst <= 1e-10, worst
    # LJ-only NVE: identical additions in identical order -> identical bits
    assert np.array_equal(e.positions(), st["pos"])
    assert np.array_equal(e.velocities(), st["vel"])
    sc = e.scalars()
    assert sc["counter"] == st["counter"]
    assert abs(sc["pe"] - st["pe"]) <= 1e-12 * abs(st["pe"])


def test_reuse_of_forces_changes_nothing():
    say("case: argon 864 FP64, force reuse on/off")
    spec = jittered_argon(n_cells=6, skin=0.05)
    a = engine(spec, reuse_forces=True)
    b = engine(spec, reuse_forces=False)
    a.step(37)
    b.step(37)
    assert np.array_equal(a.positions(), b.positions())
    assert np.array_equal(a.velocities(), b.velocities())
    assert a.scalars()["pe"] == b.scalars()["pe"]


@pytest.mark.parametrize("thermo", [_abi.THERMO_BERENDSEN, _abi.THERMO_NOSE_HOOVER])
def test_thermostats(thermo):
    say(f"case: argon 864 FP64 thermostat {thermo}")
    spec = jittered_argon(n_cells=6).with_(thermostat=thermo, thermo_temperature=120.0,
                                           thermo_tau=0.05)
    e = engine(spec)
    sim = O.Sim(spec)
    e.step(50)
    sim.step(50)
    st = sim.state()
    assert rel_max(e.forces(), st["force"]) <= 1e-10
    assert rel_max(e.velocities(), st["vel"]) <= 1e-10
    sc = e.scalars()
    assert abs(sc["zeta"] - st["zeta"]) <= 1e-10 * max(abs(st["zeta"]), 1e-30)
    assert abs(sc["ke"] - st["ke"]) <= 1e-10 * st["ke"]


@pytest.mark.parametrize("baro", [_abi.BARO_BERENDSEN, _abi.BARO_ANDERSEN])
def test_barostats(baro):
    say(f"case: argon 864 FP64 barostat {baro}")
    spec = jittered_argon(n_cells=6, skin=0.1).with_(
        barostat=baro, baro_pressure=100.0, baro_tau=0.5, lj_cutoff=0.8)
    e = engine(spec)
    sim = O.Sim(spec)
    e.step(30)
    sim.step(30)
    st = sim.state()
    assert rel_max(e.positions(), st["pos"]) <= 1e-10
    assert abs(e.scalars()["box"][0] - st["box_x"]) <= 1e-12 * st["box_x"]


@pytest.mark.parametrize("mode", [FA, CO])
def test_bonded_chain(mode):
    say(f"case: chain melt 600 FP64 physics {mode}")
    # FAITHFUL: DMD's list semantics (no skin, rebuild every 10 computes)
    spec = W.chain_melt(600, seed=5).with_(physics=mode, list_skin=0.05 if mode else 0.0,
                                           rebuild_interval=-1 if mode else 10)
    e = engine(spec)
    sim = O.Sim(spec)
    e.compute()
    sim.compute()
    st = sim.state()
    assert rel_max(e.forces(), st["force"]) <= 1e-12
    steps = 20 if mode == FA else 100
    e.step(steps)
    sim.step(steps)
    st = sim.state()
    assert rel_max(e.forces(), st["force"]) <= 1e-10
    assert rel_max(e.positions(), st["pos"]) <= 1e-10


def test_water_pme_fp64_corrected():
    say("case: water 648 FP64 corrected PME")
    spec = W.water_box(n_mol=216, seed=3)
    e = engine(spec)
    sim = O.Sim(spec)
    e.compute()
    F_ref, pe_ref = sim.forces_exact()
    assert rel_max(e.forces(), F_ref) <= 1e-10
    assert abs(e.scalars()["pe"] - pe_ref) <= 1e-10 * abs(pe_ref)
    e.step(20)
    sim.step(20)
    st = sim.state()
    assert rel_max(e.forces(), st["force"]) <= 1e-9


def test_water_pme_fp64_faithful_components():
    say("case: water 648 FP64 faithful (PME layout-consistent)")
    spec = W.water_box(n_mol=216, seed=4).with_(physics=FA, list_skin=0.0, rebuild_interval=10)
    e = engine(spec)
    sim = O.Sim(spec)
    e.compute()
    sim.compute()
    st = sim.state()
    assert rel_max(e.forces(), st["force"]) <= 1e-10
    assert abs(e.scalars()["pe"] - st["pe"]) <= 1e-10 * abs(st["pe"])


@pytest.mark.parametrize("kind", ["argon", "water", "chain"])
def test_mixed_force_rms(kind):
    say(f"case: mixed precision forces, {kind}")
    if kind == "argon":
        spec = jittered_argon(skin=0.1).with_(physics=CO)
    elif kind == "water":
        spec = W.water_box(n_side=10, seed=2)
    else:
        spec = W.chain_melt(2000, seed=9)
    e = engine(spec.with_(precision=_abi.MIXED))
    e.compute()
    # the mixed path takes its pair decisions in FP64: same set as the oracle
    gi, gj = e.pairs()
    rl = max(spec.lj_cutoff, spec.coulomb_cutoff if spec.coulomb_type else 0.0) + spec.list_skin
    oi, oj = O.pairs(spec.pos, spec.box, rl, "cells")
    if spec.exclusions is not None:
        oi, oj = O.filter_excl(oi, oj, *spec.exclusions)
    assert np.array_equal(gi, oi) and np.array_equal(gj, oj)
    F_ref, pe_ref = O.Sim(spec).forces_exact()
    err = rms_rel(e.forces(), F_ref)
    say(f"  force rms rel err {err:.3e}")
    assert err <= 1e-5
    # the potential energy is a sum of terms that cancel (water: large real-
    # and reciprocal-space Coulomb terms); its FP32 error scales with the
    # terms' magnitudes, not with the total
    sc = e.scalars()
    scale = max(abs(pe_ref), sum(abs(v) for v in sc["terms"].values()))
    say(f"  pe rel err {abs(sc['pe'] - pe_ref) / scale:.3e} (of the terms' magnitude)")
    assert abs(sc["pe"] - pe_ref) <= 1e-5 * scale


def test_batch_alone_equals_in_mix():
    say("case: batch of 3 vs alone, FP64")
    specs = [jittered_argon(n_cells=5, seed=1, skin=0.1).with_(physics=CO),
             W.chain_melt(700, seed=2),
             jittered_argon(n_cells=6, seed=3, skin=0.1).with_(physics=CO)]
    mix = engine(specs)
    mix.step(40)
    for k, s in enumerate(specs):
        alone = engine(s)
        alone.step(40)
        assert np.array_equal(alone.positions(), mix.positions(k)), k
        assert alone.scalars()["pe"] == mix.scalars(k)["pe"]


@pytest.mark.parametrize("precision", [_abi.FP64, _abi.MIXED])
def test_both_binning_paths_in_one_batch(precision):
    """A small system (binned by one workgroup) and one above the workgroup
    binning limit (the multi-kernel path) in one batch: both pair sets equal
    the oracle's, and each system's forces equal its forces alone."""
    small = jittered_argon(n_cells=5, seed=12, skin=0.1).with_(precision=precision)
    large = jittered_argon(n_cells=17, seed=13, skin=0.1).with_(precision=precision)
    say(f"case: argon {small.n} + {large.n} atoms, precision {precision}")
    e = engine([small, large])
    e.compute()
    for k, s in enumerate((small, large)):
        gi, gj = e.pairs(k)
        oi, oj = O.pairs(s.pos, s.box, s.lj_cutoff + s.list_skin, "cells")
        assert np.array_equal(gi, oi) and np.array_equal(gj, oj), k
        alone = engine(s)
        alone.compute()
        assert np.array_equal(alone.forces(), e.forces(k)), k


@pytest.mark.parametrize("kind", ["argon", "water", "chain", "small_box"])
def test_mixed_list_build_every_lane_count(kind):
    """The mixed list build at 8, 16 and 32 lanes per atom lists the
    oracle's pair set, in rows of the same order: forces are bit-identical
    across lane counts.  Water uses half-width cells searched +-2; the small
    box (fewer than three cells per edge) takes a minimum image per
    candidate."""
    if kind == "argon":
        spec = jittered_argon(skin=0.1).with_(physics=CO)
    elif kind == "water":
        spec = W.water_box(n_side=10, seed=2)
    elif kind == "chain":
        spec = W.chain_melt(2000, seed=9)
    else:
        spec = jittered_argon(n_cells=3, skin=0.1, seed=5).with_(physics=CO)
    spec = spec.with_(precision=_abi.MIXED)
    say(f"case: mixed list build lanes 8/16/32, {kind} ({spec.n} atoms)")
    rl = max(spec.lj_cutoff, spec.coulomb_cutoff if spec.coulomb_type else 0.0) + spec.list_skin
    oi, oj = O.pairs(spec.pos, spec.box, rl, "cells")
    if spec.exclusions is not None:
        oi, oj = O.filter_excl(oi, oj, *spec.exclusions)
    F_ref, _ = O.Sim(spec.with_(precision=_abi.FP64)).forces_exact()
    first = None
    for lanes in (8, 16, 32):
        e = engine(spec)
        e.set_tuning(build_lanes=lanes)
        e.compute()
        gi, gj = e.pairs()
        assert np.array_equal(gi, oi) and np.array_equal(gj, oj), lanes
        F = e.forces()
        if first is None:
            first = F
            err = rms_rel(F, F_ref)
            say(f"  force rms rel err vs FP64 {err:.3e}")
            assert err <= 1e-5
        else:
            assert np.array_equal(F, first), lanes


def test_mixed_batch_alone_equals_in_mix():
    """Mixed precision: an LJ system batched with a water system (real-space
    Coulomb pairs) and a chain melt gives the bits it gives alone -- each
    system is computed by the pair-kernel variant its own terms select."""
    say("case: batch of 3 vs alone, mixed precision, LJ + water + chain")
    M = _abi.MIXED
    specs = [W.lj_fluid(900, 0.8, 1.0, 4, precision=M),
             W.water_box(n_mol=216, seed=6, precision=M),
             W.chain_melt(800, seed=8, precision=M)]
    mix = engine(specs)
    mix.step(25)
    for k, s in enumerate(specs):
        alone = engine(s)
        alone.step(25)
        assert np.array_equal(alone.positions(), mix.positions(k)), k
        assert np.array_equal(alone.forces(), mix.forces(k)), k
        assert alone.scalars()["pe"] == mix.scalars(k)["pe"], k


def test_graph_replay_matches_eager():
    say("case: graph replay vs eager, argon 864 FP64")
    spec = jittered_argon(n_cells=6, skin=0.1).with_(thermostat=_abi.THERMO_BERENDSEN,
                                                      thermo_temperature=100.0)
    a = engine(spec, use_graph=True, graph_steps=5)
    b = engine(spec)
    for _ in range(4):
        a.step(10)
        b.step(10)
        a.sync()
        assert np.array_equal(a.positions(), b.positions())
        assert np.array_equal(a.velocities(), b.velocities())


def test_mixed_graph_replay_matches_eager():
    """Mixed batch with list rebuilds every few steps and bonded terms on
    the second stream beside the list build and pair kernel: captured in a
    graph or launched eagerly, the trajectories are the same bit for bit."""
    say("case: graph replay vs eager, mixed replica batch with rebuilds")
    specs = W.replica_mix(n_replicas=10, n_min=500, n_max=3000, seed=3,
                          precision=_abi.MIXED)
    a = engine(specs, use_graph=True, graph_steps=5)
    b = engine(specs)
    for _ in range(4):
        a.step(10)
        b.step(10)
        a.sync()
        for k in range(len(specs)):
            assert np.array_equal(a.positions(k), b.positions(k)), k
            assert np.array_equal(a.velocities(k), b.velocities(k)), k
    builds = [b.scalars(k)["list_builds"] for k in range(len(specs))]
    say(f"  list builds per system: {builds}")
    assert max(builds) > 1  # some computes rebuilt lists


PAIR_SHAPES = [(2, 8), (4, 4), (4, 8), (8, 2), (8, 4), (8, 8), (16, 2), (16, 4), (32, 2), (32, 4)]


@pytest.mark.parametrize("kind", ["argon", "water"])
def test_mixed_every_pair_shape(kind):
    """Every launch shape of the mixed pair kernel (lanes per atom x entries
    per lane pass) meets the FP64 force bar, for the LJ-only and the
    Coulomb variants; the shapes differ only in FP32 summation order."""
    say(f"case: mixed pair kernel shapes, {kind}")
    spec = (jittered_argon(skin=0.1).with_(physics=CO) if kind == "argon"
            else W.water_box(n_side=10, seed=2))
    spec = spec.with_(precision=_abi.MIXED)
    F_ref, _ = O.Sim(spec.with_(precision=_abi.FP64)).forces_exact()
    first = None
    for lanes, width in PAIR_SHAPES:
        e = engine(spec)
        e.set_tuning(pair_lanes=lanes, pair_width=width, coul_lanes=lanes, coul_width=width)
        e.compute()
        F = e.forces()
        err = rms_rel(F, F_ref)
        say(f"  {lanes}x{width}: force rms rel err {err:.3e}")
        assert err <= 1e-5, (lanes, width)
        if first is None:
            first = F
        else:
            assert rms_rel(F, first) <= 2e-6, (lanes, width)


@pytest.mark.parametrize("lj_shape,coul_shape", [((4, 8), (8, 4)), ((2, 8), (32, 2)),
                                                 ((16, 4), (4, 8))])
def test_mixed_split_pair_shapes(lj_shape, coul_shape):
    """LJ-only and Coulomb systems in one batch with different pair-kernel
    shapes (tuning keys 0/2 and 10/11): each system's forces and potential
    energy meet the FP64 bar -- the per-wave energy partials of the two
    variants have different layouts, and each system's are summed with its
    own -- and a Coulomb system's results do not depend on the LJ shape."""
    specs = [W.water_box(n_side=8, seed=3), jittered_argon(skin=0.1).with_(physics=CO),
             W.water_box(n_side=9, seed=4)]
    specs = [s.with_(precision=_abi.MIXED) for s in specs]
    say(f"case: split pair shapes LJ {lj_shape} Coulomb {coul_shape}, {len(specs)} systems")
    e = engine(specs)
    e.set_tuning(pair_lanes=lj_shape[0], pair_width=lj_shape[1], coul_lanes=coul_shape[0],
                 coul_width=coul_shape[1])
    e.step(5)
    e.compute()
    other = engine(specs)
    other.set_tuning(pair_lanes=coul_shape[0], pair_width=coul_shape[1],
                     coul_lanes=coul_shape[0], coul_width=coul_shape[1])
    other.step(5)
    other.compute()
    for k, s in enumerate(specs):
        f64 = engine(s.with_(precision=_abi.FP64, pos=e.positions(k), vel=e.velocities(k)))
        f64.compute()
        err = rms_rel(e.forces(k), f64.forces(0))
        sc = e.scalars(k)
        pe, pr = sc["pe"], f64.scalars(0)["pe"]
        # FP32 error of a PE whose terms cancel scales with the terms
        scale = max(abs(pr), sum(abs(v) for v in sc["terms"].values()))
        say(f"  system {k} ({s.name}): force rms rel {err:.2e}, pe {pe:.9g} vs {pr:.9g} "
            f"(terms {scale:.4g})")
        assert err <= 1e-5, k
        assert abs(pe - pr) <= 1e-5 * scale, (k, pe, pr, scale)
        if s.coulomb_type != 0:
            assert np.array_equal(e.positions(k), other.positions(k)), k
            assert e.scalars(k)["pe"] == other.scalars(k)["pe"], k


def test_mixed_chain_forces_after_steps():
    """After steps on the mixed path, a chain melt's forces are completed by
    the closing kick (pair forces plus the bonded forces gathered beside the
    pair kernel): they meet the FP64 bar at the reached positions."""
    spec = W.chain_melt(2000, seed=4).with_(precision=_abi.MIXED)
    say(f"case: mixed chain melt {spec.n} atoms, 30 steps; forces vs FP64 oracle")
    e = engine(spec)
    e.step(30)
    ref = O.Sim(spec.with_(precision=_abi.FP64))
    ref.set_positions(e.positions())
    F_ref, _ = ref.forces_exact()
    err = rms_rel(e.forces(), F_ref)
    say(f"  force rms rel err {err:.3e}")
    assert err <= 1e-5


@pytest.mark.parametrize("kind", ["replica_mix", "water"])
def test_list_build_per_cell_equals_per_atom(kind):
    """The per-cell list build (one wave per cell, candidates shared by the
    cell's atoms) writes the same rows, entry for entry, as the per-atom
    build: pair sets equal and trajectories identical bit for bit over
    steps with rebuilds."""
    if kind == "water":
        specs = [W.water_box(n_mol=500, seed=8, precision=_abi.MIXED)]
    else:
        specs = W.replica_mix(n_replicas=12, n_min=500, n_max=4000, seed=11,
                              precision=_abi.MIXED)
    say(f"case: list build per cell vs per atom, {kind} ({sum(s.n for s in specs)} atoms)")
    b = engine(specs)
    b.set_tuning(build_cells=False)
    b.compute()
    b_pairs = [b.pairs(k) for k in range(len(specs))]
    b_forces = [b.forces(k) for k in range(len(specs))]
    b.step(30)
    for tasks in (4, 1, 8):  # waves per cell of the per-cell build
        a = engine(specs)
        a.set_tuning(build_tasks=tasks)
        a.compute()
        for k in range(len(specs)):
            pa = a.pairs(k)
            assert np.array_equal(pa[0], b_pairs[k][0]), (tasks, k)
            assert np.array_equal(pa[1], b_pairs[k][1]), (tasks, k)
            assert np.array_equal(a.forces(k), b_forces[k]), (tasks, k)
        a.step(30)
        for k in range(len(specs)):
            assert np.array_equal(a.positions(k), b.positions(k)), (tasks, k)
            assert np.array_equal(a.forces(k), b.forces(k)), (tasks, k)
        builds = [a.scalars(k)["list_builds"] for k in range(len(specs))]
        say(f"  waves per cell {tasks}: list builds per system {builds}")
        assert max(builds) > 1
        a.close()
