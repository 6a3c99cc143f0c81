"""Every benchmark configuration against the FP64 oracle, at its stated size
where the oracle finishes in seconds.

- Batched uneven replicas (256-replica mix, here 16 replicas in one batch):
  FP64 trajectories against oracle.Sim within 1e-10 over 50 steps; mixed
  forces at the reached positions within 1e-5 RMS of the oracle's.
- Served request stream (Poisson-style LJ / water requests): every retired
  request's final forces against the oracle at its final positions (FP64
  within 1e-10, mixed within 1e-5 RMS); FP64 LJ requests also against the
  oracle's trajectory over the same steps.
- Water at full size: 12,288 atoms on a 48^3 grid (rc 0.9 nm) and 98,304
  atoms on a 100^3 grid (rc 1.0 nm), one compute each, FP64 within 1e-10
  and mixed within 1e-5 RMS.
"""
import numpy as np
import pytest

import oracle as O
from dmd_amd import _abi
from dmd_amd import workloads as W

pytestmark = pytest.mark.gpu

FP64, MIXED = _abi.FP64, _abi.MIXED


def say(*a):
    print(*a, flush=True)


def rel_max(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def rms_rel(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def exact_at(spec, pos):
    ref = O.Sim(spec.with_(precision=FP64))
    ref.set_positions(pos)
    return ref.forces_exact()


def test_replica_mix_fp64_trajectories_vs_oracle():
    from dmd_amd.engine import Engine
    specs = W.replica_mix(n_replicas=16, seed=7, precision=FP64)
    say(f"case: replica mix, 16 replicas ({sum(s.n for s in specs)} atoms) FP64, 50 steps")
    e = Engine(specs)
    e.step(50)
    e.sync()
    worst = 0.0
    for k, s in enumerate(specs):
        sim = O.Sim(s)
        sim.step(50)
        st = sim.state()
        ef, ep = rel_max(e.forces(k), st["force"]), rel_max(e.positions(k), st["pos"])
        worst = max(worst, ef, ep)
        assert ef <= 1e-10 and ep <= 1e-10, (k, s.name, ef, ep)
        assert e.scalars(k)["steps_done"] == 50
    say(f"  worst relative force/position difference {worst:.2e}")
    e.close()


def test_replica_mix_mixed_forces_vs_oracle():
    from dmd_amd.engine import Engine
    specs = W.replica_mix(n_replicas=16, seed=7, precision=MIXED)
    say(f"case: replica mix, 16 replicas mixed precision, 50 steps; forces vs oracle")
    e = Engine(specs)
    e.step(50)
    e.sync()
    errs = []
    for k, s in enumerate(specs):
        F_ref, _ = exact_at(s, e.positions(k))
        err = rms_rel(e.forces(k), F_ref)
        errs.append(err)
        assert err <= 1e-5, (k, s.name, err)
        assert e.scalars(k)["flags"] == 0
    say(f"  force rms rel err: max {max(errs):.2e}, median {float(np.median(errs)):.2e}")
    e.close()


def _drain(srv, n_total, max_pumps=5000):
    done = 0
    for _ in range(max_pumps):
        if done >= n_total:
            return
        done += srv.pump()
    raise AssertionError("server did not retire every request")


@pytest.mark.parametrize("precision", [MIXED, FP64])
def test_served_stream_vs_oracle(precision):
    from dmd_amd.server import Server
    reqs = W.serving_requests(8, seed=42, steps_scale=0.005, precision=precision)
    say(f"case: served stream, precision {precision}: "
        + ", ".join(f"{s.name}x{n}" for s, n in reqs))
    srv = Server(chunk_steps=10, max_cohorts=3)
    # requests arrive while others run: half up front, the rest after a round
    ids = [srv.submit(s, n) for s, n in reqs[:4]]
    srv.pump()
    ids += [srv.submit(s, n) for s, n in reqs[4:]]
    _drain(srv, len(reqs))
    for rid, (s, n) in zip(ids, reqs):
        info = srv.info(rid)
        assert info["state"] == "done" and info["steps_done"] == n, (s.name, info)
        res = srv.result(rid)
        F_ref, _ = exact_at(s, res["pos"])
        if precision == MIXED:
            err = rms_rel(res["force"], F_ref)
            assert err <= 1e-5, (s.name, err)
        else:
            err = rel_max(res["force"], F_ref)
            assert err <= 1e-10, (s.name, err)
            if s.coulomb_type == _abi.COULOMB_NONE:
                sim = O.Sim(s)
                sim.step(n)
                st = sim.state()
                assert rel_max(res["pos"], st["pos"]) <= 1e-10, s.name
                assert rel_max(res["force"], st["force"]) <= 1e-10, s.name
        say(f"  {s.name} x {n} steps: force error {err:.2e}")
    st = srv.stats()
    assert st["retired"] == len(reqs) and st["failed"] == 0 and st["bytes_in_use"] == 0
    srv.close()


@pytest.mark.parametrize("size", ["12288 atoms, 48^3, rc 0.9", "98304 atoms, 100^3, rc 1.0"])
def test_water_full_size_vs_oracle(size):
    from dmd_amd.engine import Engine
    if size.startswith("12288"):
        spec = W.water_box(n_side=16, cutoff=0.9, grid=48, seed=1)
    else:
        spec = W.water_box(n_side=32, cutoff=1.0, grid=100, seed=2)
    say(f"case: water full size ({size}); one compute, FP64 and mixed vs oracle")
    F_ref, pe_ref = O.Sim(spec).forces_exact()
    e = Engine(spec.with_(precision=FP64))
    e.compute()
    err64 = rel_max(e.forces(), F_ref)
    say(f"  FP64 max rel force error {err64:.2e}")
    assert err64 <= 1e-10
    # the potential energy of a water box is a small sum of large terms of
    # both signs (real- and reciprocal-space Coulomb, self energy): its FP64
    # rounding scales with the terms, not with the total
    sc = e.scalars()
    scale = max(abs(pe_ref), sum(abs(v) for v in sc["terms"].values()))
    say(f"  FP64 pe error {abs(sc['pe'] - pe_ref):.2e} of terms {scale:.3e}")
    assert abs(sc["pe"] - pe_ref) <= 1e-10 * scale
    e.close()
    e = Engine(spec.with_(precision=MIXED))
    e.compute()
    err = rms_rel(e.forces(), F_ref)
    say(f"  mixed force rms rel error {err:.2e}")
    assert err <= 1e-5
    e.close()
