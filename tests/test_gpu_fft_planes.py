"""xy planes of the mixed PME transforms (k_fft_plane, tuning keys 12/13):
the forward x and y passes of one z plane -- with the fixed-point grid
converted as it loads -- and the inverse y and x passes run in LDS by one
workgroup.  Every line goes through the same Stockham stages and twiddles as
on line tiles, so the grid, and with it the trajectory, is the line tiles'
bit for bit: a batch of water boxes with enough planes for the default rule,
one box forced onto planes (also against the FP64 forces), and a batch that
mixes plane systems with a grid too large for a plane (line tiles)."""
import numpy as np
import pytest

from dmd_amd import _abi
from dmd_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def _run(specs, planes, steps=10):
    from dmd_amd.engine import Engine
    e = Engine(specs)
    if planes is not None:
        e.set_tuning(fft_planes=planes)
    e.step(steps)
    e.compute()
    return e


def _same(a, b, n):
    for k in range(n):
        assert np.array_equal(a.positions(k), b.positions(k)), k
        assert np.array_equal(a.forces(k), b.forces(k)), k
        assert a.scalars(k)["pe"] == b.scalars(k)["pe"], k


def test_planes_by_rule_equal_line_tiles():
    # eight 32^3 grids: 256 planes, the default rule's threshold
    specs = [W.water_box(n_side=10, seed=k, precision=_abi.MIXED) for k in range(8)]
    say(f"case: {len(specs)} water boxes (32^3), planes by the rule vs line tiles, 10 steps")
    a = _run(specs, None)
    b = _run(specs, 0)
    assert a.get_tuning(13) == 1 and b.get_tuning(13) == 0
    _same(a, b, len(specs))


def test_one_box_forced_onto_planes():
    from dmd_amd.engine import Engine
    spec = W.water_box(n_side=10, seed=3, precision=_abi.MIXED)
    say("case: one water box (32^3): below the rule's threshold; forced onto planes")
    rule = _run([spec], None)
    assert rule.get_tuning(13) == 0
    a = _run([spec], 1)
    assert a.get_tuning(13) == 1
    _same(a, rule, 1)
    f64 = Engine(spec.with_(precision=_abi.FP64, pos=a.positions(0), vel=a.velocities(0)))
    f64.compute()
    fa, fr = a.forces(0), f64.forces(0)
    err = float(np.sqrt(((fa - fr) ** 2).sum() / (fr ** 2).sum()))
    say(f"  force rms rel vs FP64 {err:.2e}")
    assert err <= 1e-5


def test_planes_beside_line_tiles():
    # 72^3: two 72 x 72 complex buffers exceed 64 KiB, so that box keeps its
    # line tiles while the two 32^3 boxes run on planes
    specs = [W.water_box(n_side=10, seed=5, precision=_abi.MIXED),
             W.water_box(n_side=10, seed=6, grid=72, precision=_abi.MIXED),
             W.water_box(n_side=10, seed=7, precision=_abi.MIXED)]
    say("case: 32^3 boxes on planes beside a 72^3 box on line tiles")
    a = _run(specs, 1)
    b = _run(specs, 0)
    assert a.get_tuning(13) == 1
    _same(a, b, len(specs))
    from dmd_amd.engine import Engine
    e = Engine(specs)
    e.step(1)
    with pytest.raises(RuntimeError):
        e.set_tuning(fft_planes=0)
