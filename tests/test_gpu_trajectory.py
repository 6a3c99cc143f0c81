"""Frames of running engines: asynchronous snapshots and H5MD trajectories.

A snapshot taken after k steps must hold exactly the state that
dmdx_engine_get_arrays returns after k steps of an identical engine (the
velocities with any pending thermostat scale applied), must not disturb the
steps that follow it, and the frame loop must write DMD's frames: one after
every step whose number is a multiple of the interval, with
step = SimulationEngine::step_ and time = SystemData::time, which step() sets
at its start (src/sim/simulation_engine.cpp:53-55, 82-85, 88-106)."""
import os

import numpy as np
import pytest

import h5md_read as R
from dmd_amd import _abi
from dmd_amd import workloads as W
from dmd_amd.engine import Engine

pytestmark = pytest.mark.gpu

A = "/particles/atoms"


def say(*a):
    print(*a, flush=True)


def specs():
    """FP64 LJ with a pending Berendsen scale, the same in MIXED, and a
    Berendsen-barostat system whose box changes every step."""
    lj = W.lj_fluid(1500, 0.8, 1.0, seed=3)
    baro = W.lj_fluid(1200, 0.7, 1.2, seed=4).with_(
        barostat=_abi.BARO_BERENDSEN, baro_pressure=1.0, baro_tau=0.5)
    return {"lj_fp64": lj, "lj_mixed": lj.with_(precision=_abi.MIXED), "baro": baro}


@pytest.mark.parametrize("name", ["lj_fp64", "lj_mixed", "baro"])
def test_snapshot_equals_state_and_changes_nothing(name):
    say(f"case: snapshot vs get_arrays, {name}")
    spec = specs()[name]
    a, b = Engine(spec), Engine(spec)
    a.step(7)
    t = a.snapshot_begin()
    a.step(5)
    fr = a.snapshot_read(t)
    b.step(7)
    assert fr["steps_done"] == 7
    assert np.array_equal(fr["pos"], b.positions())
    assert np.array_equal(fr["vel"], b.velocities())
    assert np.array_equal(fr["box"], np.asarray(b.scalars()["box"]))
    b.step(5)
    assert np.array_equal(a.positions(), b.positions())
    assert np.array_equal(a.velocities(), b.velocities())
    a.close()
    b.close()


def test_snapshot_ring_bounds_and_order():
    say("case: snapshot ring, busy and out-of-order reads")
    spec = specs()["lj_fp64"]
    e = Engine(spec)
    tickets = []
    for _ in range(4):  # DMDX_SNAPSHOT_SLOTS
        e.step(1)
        tickets.append(e.snapshot_begin())
    with pytest.raises(_abi.DmdxError) as ex:
        e.snapshot_begin()
    assert ex.value.status == -8  # DMDX_E_BUSY
    frames = {t: e.snapshot_read(t) for t in reversed(tickets)}
    assert [frames[t]["steps_done"] for t in tickets] == [1, 2, 3, 4]
    with pytest.raises(_abi.DmdxError):
        e.snapshot_read(tickets[0])  # already read
    e.step(1)
    t = e.snapshot_begin()
    assert e.snapshot_read(t)["steps_done"] == 5
    e.close()


T = pytest.importorskip("dmd_amd.trajectory")
needs_h5 = pytest.mark.skipif(not (os.path.exists(T.LIB_PATH) and R.available()),
                              reason="libdmdx_h5md.so or the HDF5 library is absent")


def dmd_clock(dt, k):
    """SimulationEngine's time_ after k steps and SystemData::time (the
    time_ at the start of the k-th step): a running sum of dt."""
    t, sys_t = 0.0, 0.0
    for _ in range(k):
        sys_t = t
        t += dt
    return t, sys_t


@needs_h5
@pytest.mark.parametrize("name", ["lj_fp64", "baro"])
def test_run_writes_dmd_frames(tmp_path, name):
    say(f"case: frame loop, {name}, 23 steps, interval 5")
    spec = specs()[name]
    n, dt = spec.n, spec.dt
    path = str(tmp_path / "traj.h5")
    e = Engine(spec)
    clock = T.Clock.start(dt)
    with T.H5MDWriter(path, n, write_vel=True) as w:
        T.run(e, 13, 5, clock, w)  # split runs: the cadence follows the step number
        T.run(e, 10, 5, clock, w)
        assert w.frame_count() == 4
    assert clock.step == 23
    t23, s23 = dmd_clock(dt, 23)
    assert clock.time == t23 and clock.sys_time == s23
    ref = Engine(spec)
    with R.File(path) as f:
        pos = f.read(f"{A}/position/value")
        vel = f.read(f"{A}/velocity/value")
        box = f.read(f"{A}/box/edges/value")
        assert list(f.read(f"{A}/position/step")) == [5, 10, 15, 20]
        assert list(f.read(f"{A}/box/edges/step")) == [5, 10, 15, 20]
        times = f.read(f"{A}/velocity/time")
        for k, step in enumerate([5, 10, 15, 20]):
            ref.step(5)
            assert times[k] == dmd_clock(dt, step)[1]
            assert np.array_equal(pos[k], ref.positions())
            assert np.array_equal(vel[k], ref.velocities())
            L = np.asarray(ref.scalars()["box"])
            assert np.array_equal(box[k], np.diag(L))
    if name == "baro":
        assert not np.array_equal(box[0], box[-1])  # the box moved
    ref.step(3)
    assert np.array_equal(e.positions(), ref.positions())
    e.close()
    ref.close()


@needs_h5
def test_run_without_frames_and_water_mixed(tmp_path):
    """interval 0 only steps; a MIXED water box (PME, Nose-Hoover) frames
    equal the stepped state."""
    say("case: frame loop on MIXED water 648 atoms")
    spec = W.water_box(n_side=6, precision=_abi.MIXED)
    e, ref = Engine(spec), Engine(spec)
    clock = T.Clock.start(spec.dt)
    T.run(e, 4, 0, clock, None)
    assert clock.step == 4
    path = str(tmp_path / "w.h5")
    with T.H5MDWriter(path, spec.n, write_vel=False) as w:
        T.run(e, 8, 4, clock, w)
    ref.step(8)
    with R.File(path) as f:
        assert list(f.read(f"{A}/position/step")) == [8, 12]
        assert not f.exists(f"{A}/velocity")
        assert np.array_equal(f.read(f"{A}/position/value")[0], ref.positions())
    ref.step(4)
    assert np.array_equal(e.positions(), ref.positions())
    e.close()
    ref.close()


@needs_h5
def test_compat_run_frames_and_checkpoint_restart(tmp_path):
    """dmd_compat.Engine.run as SimulationEngine::run: frames every nstxout
    steps, DMD's binary checkpoint every checkpoint_interval steps, and a
    restart from that checkpoint that continues bit for bit."""
    from test_dmd_compat import argon_config

    from dmd_amd import checkpoint as CK
    from dmd_amd.dmd_compat import Engine as CompatEngine

    say("case: compat run, nstxout 5, checkpoint every 10, restart")
    c, s = argon_config()
    c.n_steps = 20
    c.nstxout = 5
    c.checkpoint_interval = 10
    c.trajectory_path = str(tmp_path / "traj.h5")
    c.checkpoint_path = str(tmp_path / "ck.bin")
    e = CompatEngine(c)
    e.run()
    assert e.current_step == 20 and e.current_time == dmd_clock(c.dt, 20)[0]
    ck = CK.load(c.checkpoint_path)  # the last one, at step 20
    assert ck.step == 20 and ck.time == dmd_clock(c.dt, 20)[1]
    assert np.array_equal(np.stack([ck.arrays["pos_x"], ck.arrays["pos_y"], ck.arrays["pos_z"]],
                                   1), e.positions)
    assert np.array_equal(np.stack([ck.arrays["forces_x"], ck.arrays["forces_y"],
                                    ck.arrays["forces_z"]], 1), e.forces)
    assert ck.lj_step_since_rebuild == e._e.scalars()["counter"]
    with R.File(c.trajectory_path) as f:
        assert list(f.read(f"{A}/position/step")) == [5, 10, 15, 20]
        assert np.array_equal(f.read(f"{A}/position/value")[3], e.positions)
    # restart: a fresh engine loads the checkpoint and runs 10 more steps
    c2, _ = argon_config()
    c2.nstxout = 0
    c2.checkpoint_interval = 0
    r = CompatEngine(c2)
    r.load_checkpoint(c.checkpoint_path)
    assert r.current_step == 20
    e.run_n(10)  # frames at 25 and 30 go on in the same file
    r.run_n(10)
    assert np.array_equal(r.positions, e.positions)
    assert np.array_equal(r.velocities, e.velocities)
    e._writer.close()
    with R.File(c.trajectory_path) as f:
        assert list(f.read(f"{A}/position/step")) == [5, 10, 15, 20, 25, 30]
