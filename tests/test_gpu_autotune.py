"""Launch-shape autotuner (dmdx_engine_autotune): picks one of the measured
candidates for each kernel, leaves the engine's state alone, caches its
choice, and the tuned engine still meets the mixed-precision force bar."""
import numpy as np
import pytest

from dmd_amd import _abi
from dmd_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_autotune_choice_cache_and_state(tmp_path):
    from dmd_amd.engine import Engine, PAIR_SHAPES
    specs = W.replica_mix(n_replicas=12, n_min=500, n_max=3000, seed=7, precision=_abi.MIXED)
    e = Engine(specs)
    e.step(3)
    before = [e.positions(k) for k in range(len(specs))]
    cache = str(tmp_path / "tune.txt")
    r = e.autotune(steps=3, cache=cache)
    print(r, flush=True)
    assert f"{r['pair_lanes']}x{r['pair_width']}" in PAIR_SHAPES
    assert r["build_lanes"] in (8, 16, 32)
    assert r["build"] in ("atom8", "atom16", "atom32", "cell2", "cell4", "cell8", "stage8",
                          "stage16")
    pm = r["pair_ms"]
    assert all(v > 0 for v in pm.values())
    # the fastest, or the default 4x8 when no candidate beats it by 2%
    best = min(pm.values())
    chosen = pm[f"{r['pair_lanes']}x{r['pair_width']}"]
    assert chosen == best or ((r["pair_lanes"], r["pair_width"]) == (4, 8) and
                              best >= 0.98 * chosen)
    assert r["pair_rows"] in (0, 4, 8)
    rm = r["rows_ms"]
    assert all(v > 0 for v in rm.values())
    if r["pair_rows"]:
        assert rm[f"rows{r['pair_rows']}"] < chosen
    else:
        assert chosen <= min(rm.values())
    # no system has Coulomb pairs: that variant is not timed and keeps 8x2
    assert all(v == -1.0 for v in r["coul_ms"].values())
    assert (r["coul_lanes"], r["coul_width"]) == (8, 2)
    bm = r["build_ms"]
    assert bm[r["build"]] == min(bm.values())
    assert bm[f"atom{r['build_lanes']}"] == min(bm[k] for k in bm if k.startswith("atom"))
    # the engine's own systems did not move
    for k in range(len(specs)):
        assert np.array_equal(e.positions(k), before[k])
    # a second engine on the same systems hits the cache (skins left as they
    # are, as on the first engine, which had stepped)
    e2 = Engine(specs)
    e2.set_tuning(tune_skin=False)
    r2 = e2.autotune(steps=3, cache=cache)
    assert (r2["pair_lanes"], r2["pair_width"], r2["build_lanes"], r2["build"],
            r2["pair_rows"]) == \
        (r["pair_lanes"], r["pair_width"], r["build_lanes"], r["build"], r["pair_rows"])
    assert all(v == -1.0 for v in r2["pair_ms"].values())
    # forces of the tuned engine against the FP64 path
    e.compute()
    f64 = Engine([s.with_(precision=_abi.FP64) for s in specs])
    e3 = Engine(specs)
    e3.set_tuning(tune_skin=False)
    e3.autotune(steps=1, cache=cache)
    e3.compute()
    f64.compute()
    for k in range(len(specs)):
        a, b = f64.forces(k), e3.forces(k)
        err = float(np.sqrt(((b - a) ** 2).sum() / (a ** 2).sum()))
        assert err <= 1e-5, (k, err)


def test_autotune_skins(tmp_path):
    """On a fresh engine the autotuner may widen the skins of Verlet-criterion
    systems (whole groups of equal cutoffs and skin), caches the choice, never
    touches a stepped engine's skins, and the tuned engine's forces still meet
    the mixed-precision bar against the FP64 path along its run."""
    from dmd_amd.engine import Engine
    specs = W.replica_mix(n_replicas=12, n_min=500, n_max=3000, seed=7, precision=_abi.MIXED)
    cache = str(tmp_path / "tune.txt")
    e = Engine(specs)
    r = e.autotune(steps=3, cache=cache)
    print(r["skins"], flush=True)
    skins = [e.list_skin(k) for k in range(len(specs))]
    for s, now in zip(specs, skins):
        if s.rebuild_interval == -1 and s.list_skin > 0:
            ratio = now / s.list_skin
            assert min(abs(ratio - m) for m in (1.0, 1.3, 1.6, 2.0)) < 1e-9
        else:
            assert now == s.list_skin
    # one multiplier per group of equal (cutoffs, skin)
    groups = {}
    for s, now in zip(specs, skins):
        groups.setdefault((s.lj_cutoff, s.coulomb_cutoff, s.list_skin), set()).add(now)
    assert all(len(v) == 1 for v in groups.values())
    # the cache gives a second fresh engine the same skins without timing
    e2 = Engine(specs)
    r2 = e2.autotune(steps=3, cache=cache)
    assert [e2.list_skin(k) for k in range(len(specs))] == skins
    assert all(v == -1.0 for v in r2["pair_ms"].values())
    # a stepped engine keeps its skins
    e3 = Engine(specs)
    e3.step(1)
    e3.autotune(steps=1, cache=cache)
    assert [e3.list_skin(k) for k in range(len(specs))] == [s.list_skin for s in specs]
    # the tuned engine's forces along its run, against the FP64 path
    e.step(40)
    for k in range(len(specs)):
        st = specs[k].with_(pos=e.positions(k), vel=e.velocities(k), precision=_abi.FP64)
        f64 = Engine(st)
        f64.compute()
        a, b = f64.forces(), e.forces(k)
        err = float(np.sqrt(((b - a) ** 2).sum() / (a ** 2).sum()))
        assert err <= 1e-5, (k, err)


def test_cell_grid_forced():
    """Tuning key 9 sets the cell grid before the first step: water (half-width
    cells by the density rule) on cells of the list radius, an LJ fluid (cells
    of the list radius) on half-width cells.  The pair sets stay the oracle's
    and the forces meet the mixed-precision bar; after a step the key is
    refused."""
    import oracle as O
    from dmd_amd.engine import Engine
    cases = [(W.water_box(n_side=10, seed=2, precision=_abi.MIXED), 1),
             (W.lj_fluid(4000, 0.8442, 0.72, seed=3, cutoff=0.85, skin=0.1,
                         precision=_abi.MIXED), 2)]
    for spec, cd in cases:
        print(f"case: {spec.name} forced to {cd} cells per list radius", flush=True)
        e = Engine(spec)
        rule = e.cell_div()
        e.set_tuning(cell_div=cd)
        assert e.cell_div() == cd != rule
        e.compute()
        rl = max(spec.lj_cutoff, spec.coulomb_cutoff if spec.coulomb_type else 0.0) + spec.list_skin
        oi, oj = O.pairs(spec.pos, spec.box, rl, "cells")
        if spec.exclusions is not None:
            oi, oj = O.filter_excl(oi, oj, *spec.exclusions)
        gi, gj = e.pairs()
        assert np.array_equal(gi, oi) and np.array_equal(gj, oj)
        f64 = Engine(spec.with_(precision=_abi.FP64))
        f64.compute()
        a, b = f64.forces(), e.forces()
        err = float(np.sqrt(((b - a) ** 2).sum() / (a ** 2).sum()))
        assert err <= 1e-5, err
        e.step(1)
        with pytest.raises(RuntimeError):
            e.set_tuning(cell_div=rule)


def test_autotune_cells_and_fft(tmp_path):
    """On a fresh batch with PME the autotuner also picks each group's cells
    and the FFT tiling; a second engine replays both from the cache; the
    tuned forces meet the mixed-precision bar against the FP64 path."""
    from dmd_amd.engine import Engine
    specs = [W.water_box(n_mol=512, seed=4, precision=_abi.MIXED),
             W.lj_fluid(1500, 0.8, 1.0, seed=5, precision=_abi.MIXED),
             W.chain_melt(1200, seed=6, precision=_abi.MIXED)]
    cache = str(tmp_path / "tune.txt")
    e = Engine(specs)
    r = e.autotune(steps=3, cache=cache)
    print({k: r[k] for k in ("fft_lines", "cell_div", "skins")}, flush=True)
    assert r["fft_lines"] in (0, 4, 8, 16)
    assert sum(r["cell_div"].values()) == len(specs)
    # both pair-kernel variants timed, each with its own fastest shape
    cm, pm = r["coul_ms"], r["pair_ms"]
    print({"coul": cm, "pair": pm}, flush=True)
    assert all(v > 0 for v in cm.values()) and all(v > 0 for v in pm.values())
    for ms, shape, default in ((cm, (r["coul_lanes"], r["coul_width"]), (8, 2)),
                               (pm, (r["pair_lanes"], r["pair_width"]), (4, 8))):
        chosen, best = ms[f"{shape[0]}x{shape[1]}"], min(ms.values())
        assert chosen == best or (shape == default and best >= 0.98 * chosen)
    e2 = Engine(specs)
    r2 = e2.autotune(steps=3, cache=cache)
    assert r2["fft_lines"] == r["fft_lines"]
    assert (r2["coul_lanes"], r2["coul_width"], r2["pair_lanes"], r2["pair_width"]) == \
        (r["coul_lanes"], r["coul_width"], r["pair_lanes"], r["pair_width"])
    assert [e2.cell_div(k) for k in range(3)] == [e.cell_div(k) for k in range(3)]
    assert [e2.list_skin(k) for k in range(3)] == [e.list_skin(k) for k in range(3)]
    e.compute()
    for k, s in enumerate(specs):
        f64 = Engine(s.with_(precision=_abi.FP64))
        f64.compute()
        a, b = f64.forces(), e.forces(k)
        err = float(np.sqrt(((b - a) ** 2).sum() / (a ** 2).sum()))
        assert err <= 1e-5, (k, err)
