"""Serving loop (dmdx_server_*): many independent simulations on one device.

Bars: a request's final state (positions, velocities, forces, energies) is
bit-identical to the same system stepped alone by the engine, whatever it
was batched with and however the rounds cut its steps; a failing request
(non-finite energy) retires as FAILED without disturbing the requests it
shares a batch with; the Poisson driver retires every request and reports
finite latency percentiles."""
import numpy as np
import pytest

from dmd_amd import _abi
from dmd_amd import workloads as W

pytestmark = pytest.mark.gpu

M = _abi.MIXED


def say(*a):
    print(*a, flush=True)


def mix():
    return [
        (W.lj_fluid(700, 0.80, 1.0, 11, precision=M), 23),
        (W.chain_melt(600, 12, precision=M), 40),
        (W.water_box(n_mol=216, seed=13, precision=M), 17),
        (W.lj_fluid(1500, 0.85, 0.9, 14, precision=M), 31),
        (W.chain_melt(900, 15, precision=M), 9),
        (W.water_box(n_mol=300, seed=16, precision=M), 50),
    ]


def alone(spec, n):
    from dmd_amd.engine import Engine
    e = Engine(spec)
    e.step(n)
    out = {"pos": e.positions(), "vel": e.velocities(), "force": e.forces(),
           "potential_energy": e.scalars()["pe"], "steps_done": e.scalars()["steps_done"]}
    e.close()
    return out


def drain(srv, n_total, max_pumps=2000):
    done = 0
    for _ in range(max_pumps):
        if done >= n_total:
            return
        done += srv.pump()
    if done < n_total:
        raise AssertionError("server did not retire every request")


def same(a, b):
    for k in ("pos", "vel", "force"):
        assert np.array_equal(a[k], b[k]), (k, float(np.abs(a[k] - b[k]).max()))
    assert a["potential_energy"] == b["potential_energy"]
    assert a["steps_done"] == b["steps_done"]


@pytest.mark.parametrize("precision", [M, _abi.FP64])
def test_server_result_equals_alone(precision):
    from dmd_amd.server import Server
    reqs = [(s.with_(precision=precision), n) for s, n in mix()]
    say(f"case: server vs alone, precision {precision}, {len(reqs)} requests")
    ref = [alone(s, n) for s, n in reqs]
    srv = Server(chunk_steps=7, max_cohorts=2)
    ids = [srv.submit(s, n) for s, n in reqs[:3]]
    srv.pump()
    ids += [srv.submit(s, n) for s, n in reqs[3:]]
    drain(srv, len(reqs) - 0)
    st = srv.stats()
    assert st["retired"] == len(reqs) and st["failed"] == 0
    assert st["cohorts_made"] >= 2 and st["bytes_in_use"] == 0
    for rid, (s, n), r in zip(ids, reqs, ref):
        info = srv.info(rid)
        assert info["state"] == "done" and info["steps_done"] == n
        same(srv.result(rid), r)
    srv.close()


def test_server_failure_isolation():
    from dmd_amd.server import Server
    say("case: one request with a non-finite velocity among three")
    good = [(W.lj_fluid(800, 0.8, 1.0, 21, precision=M), 12),
            (W.chain_melt(700, 22, precision=M), 15)]
    bad = W.lj_fluid(600, 0.8, 1.0, 23, precision=M)
    vel = bad.vel.copy()
    vel[7, 1] = np.nan
    bad = bad.with_(vel=vel)
    ref = [alone(s, n) for s, n in good]
    srv = Server(chunk_steps=5)
    ids = [srv.submit(good[0][0], good[0][1]), srv.submit(bad, 20),
           srv.submit(good[1][0], good[1][1])]
    drain(srv, 3)
    ib = srv.info(ids[1])
    assert ib["state"] == "failed" and ib["flags"] & 2 and ib["steps_done"] == 1, ib
    assert "non-finite" in ib["error"]
    for rid, r in zip((ids[0], ids[2]), ref):
        assert srv.info(rid)["state"] == "done"
        same(srv.result(rid), r)
    assert srv.stats()["cohorts_made"] == 1


def test_server_budget_admission():
    """With a budget that holds about two of the requests, the server never
    exceeds it and still retires everything."""
    from dmd_amd.server import Server
    reqs = [(W.lj_fluid(2000, 0.8, 1.0, 30 + k, precision=M), 10 + k) for k in range(5)]
    srv = Server(hbm_budget_bytes=1, chunk_steps=4)
    rid = srv.submit(*reqs[0])
    srv.pump()
    one = srv.info(rid)["est_bytes"]
    assert srv.info(rid)["state"] == "failed"  # does not fit in 1 byte
    srv.close()
    srv = Server(hbm_budget_bytes=int(2.5 * one), chunk_steps=4)
    ids = [srv.submit(s, n) for s, n in reqs]
    done, peak = 0, 0
    for _ in range(1000):
        done += srv.pump()
        st = srv.stats()
        peak = max(peak, st["bytes_in_use"])
        assert st["bytes_in_use"] <= st["bytes_budget"]
        if done == len(reqs):
            break
    assert done == len(reqs)
    assert all(srv.info(i)["state"] == "done" for i in ids)
    assert srv.stats()["cohorts_made"] >= 2
    say(f"  budget {int(2.5 * one)} B, peak in use {peak} B")


def test_run_poisson_small():
    from dmd_amd.server import Server, run_poisson
    reqs = W.serving_requests(6, seed=5, steps_scale=0.005, n_lj=(1000, 3000),
                              n_water=(216, 400))
    say(f"case: poisson, {len(reqs)} requests, steps {[n for _, n in reqs]}")
    srv = Server(chunk_steps=10)
    rep = run_poisson(srv, reqs, rate_per_s=200.0, seed=3)
    say(f"  {rep}")
    assert rep["failed"] == 0 and rep["requests"] == 6
    assert rep["atom_steps"] == sum(s.n * n for s, n in reqs)
    for k in ("step_latency_ms", "completion_s"):
        assert np.isfinite(rep[k]["p50"]) and rep[k]["p99"] >= rep[k]["p50"]


@pytest.mark.parametrize("precision", [M, _abi.FP64])
def test_server_merges_keep_results(precision):
    """Requests arriving one per round: every round merges the newcomer's
    cohort into the running batch (systems move with their whole state, lists
    included); each result still equals the system run alone."""
    from dmd_amd.server import Server
    reqs = [(s.with_(precision=precision), n) for s, n in mix()]
    say(f"case: merging server, precision {precision}")
    ref = [alone(s, n) for s, n in reqs]
    srv = Server(chunk_steps=3)
    ids = []
    for s, n in reqs:
        ids.append(srv.submit(s, n))
        srv.pump()
    drain(srv, len(reqs) - srv.stats()["retired"])
    st = srv.stats()
    say(f"  merges {st['merges']}, cohorts made {st['cohorts_made']}")
    assert st["merges"] >= 3 and st["retired"] == len(reqs)
    for rid, r in zip(ids, ref):
        assert srv.info(rid)["state"] == "done"
        same(srv.result(rid), r)


def test_server_cancel_frees_memory_and_keeps_others():
    """A queued and a running request are cancelled: the running one keeps
    the state it reached, its device memory is returned at the next pump,
    and the other requests' results equal the systems run alone."""
    from dmd_amd.server import Server
    reqs = [(W.lj_fluid(1500, 0.80, 1.0, 41, precision=M), 30),
            (W.lj_fluid(1600, 0.85, 0.9, 42, precision=M), 200),
            (W.chain_melt(900, 43, precision=M), 25),
            (W.lj_fluid(900, 0.80, 1.0, 44, precision=M), 20)]
    say("case: server cancel (one running, one queued)")
    ref = {k: alone(*reqs[k]) for k in (0, 2)}
    srv = Server(chunk_steps=5, max_cohorts=1)
    ids = [srv.submit(s, n) for s, n in reqs[:3]]
    srv.pump()  # one cohort of three, 5 steps
    late = srv.submit(*reqs[3])
    assert srv.info(late)["state"] == "queued"
    srv.cancel(late)
    assert srv.info(late)["state"] == "cancelled"
    before = srv.stats()["bytes_in_use"]
    srv.cancel(ids[1])
    info = srv.info(ids[1])
    assert info["state"] == "cancelled" and info["steps_done"] == 5, info
    res = srv.result(ids[1])
    assert np.isfinite(res["pos"]).all() and res["steps_done"] == 5
    with pytest.raises(Exception):
        srv.cancel(ids[1])  # finished requests cannot be cancelled
    srv.pump()  # the cohort is compacted without the cancelled system
    after = srv.stats()["bytes_in_use"]
    say(f"  bytes in use {before} -> {after}")
    assert after < before
    drain(srv, 2 - srv.stats()["retired"])
    st = srv.stats()
    assert st["retired"] == 2 and st["cancelled"] == 2 and st["failed"] == 0
    for k in (0, 2):
        same(srv.result(ids[k]), ref[k])
    assert srv.info(ids[1])["steps_done"] == 5  # it never stepped again
    srv.close()


def test_engine_adopt_moves_state_and_refuses_other_layouts():
    """dmdx_engine_adopt: a system moved into another engine continues the
    same trajectory; a slot made from another description is refused."""
    from dmd_amd.engine import Engine
    say("case: engine adopt (same description moves, another layout refused)")
    a_spec = W.lj_fluid(1200, 0.80, 1.0, 51, precision=M)
    b_spec = W.chain_melt(800, 52, precision=M)
    src = Engine([a_spec, b_spec])
    src.step(12)
    dst = Engine([b_spec, a_spec])
    dst.adopt(1, src, 0)
    dst.adopt(0, src, 1)
    ref = Engine([a_spec, b_spec])
    ref.step(30)
    src.step(18)  # the sources continue too: the moved copies must follow them
    dst.step(18)
    for k_src, k_dst in ((0, 1), (1, 0)):
        assert np.array_equal(dst.positions(k_dst), ref.positions(k_src))
        assert np.array_equal(dst.positions(k_dst), src.positions(k_src))
    # same atom count, other box: other cells, refused
    other = a_spec.with_(box=tuple(1.3 * np.asarray(a_spec.box)))
    bad = Engine([other])
    with pytest.raises(Exception):
        bad.adopt(0, src, 0)
    with pytest.raises(Exception):
        dst.adopt(0, src, 0)  # different atoms
