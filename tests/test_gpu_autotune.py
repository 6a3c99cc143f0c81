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
    assert pm[f"{r['pair_lanes']}x{r['pair_width']}"] == min(pm.values())
    assert r["pair_rows"] in (0, 4, 8)
    rm = r["rows_ms"]
    assert all(v > 0 for v in rm.values())
    if r["pair_rows"]:
        assert rm[f"rows{r['pair_rows']}"] == min(min(rm.values()), min(pm.values()))
    else:
        assert min(pm.values()) <= min(rm.values())
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
