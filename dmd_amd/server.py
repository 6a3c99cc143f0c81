"""Serving many simulations on one device, over the C-ABI (dmdx_server_*).

The reference advances one system per process: ``SimulationEngine::run``
(src/sim/simulation_engine.cpp:89-106) driven by ``dmd.run``
(python/dmd/runner.py:44-58).  ``Server`` takes requests -- a system and a
step count -- at any time, admits them when their device memory fits in the
budget, advances them in batched cohorts, and retires each exactly at its
step target.  ``run_poisson`` drives a server with Poisson arrivals and
reports throughput and latency percentiles.
"""
from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np

from . import _abi
from .engine import Scalars, _dp
from .system import SystemSpec

QUEUED, RUNNING, DONE, FAILED, CANCELLED = 0, 1, 2, 3, 4
STATE_NAMES = {QUEUED: "queued", RUNNING: "running", DONE: "done", FAILED: "failed",
               CANCELLED: "cancelled"}


class ServerOptions(C.Structure):
    _fields_ = [
        ("hbm_budget_bytes", C.c_int64), ("max_cohorts", C.c_int32),
        ("chunk_steps", C.c_int32), ("graph_steps", C.c_int32), ("min_admit", C.c_int32),
        ("max_wait_ms", C.c_double), ("max_cohort_systems", C.c_int32),
        ("keep_cohorts", C.c_int32),
    ]


class RequestInfo(C.Structure):
    _fields_ = [
        ("state", C.c_int32), ("flags", C.c_int32), ("n_atoms", C.c_int64),
        ("steps_target", C.c_int64), ("steps_done", C.c_int64), ("t_submit", C.c_double),
        ("t_admit", C.c_double), ("t_done", C.c_double), ("est_bytes", C.c_int64),
        ("error", C.c_char * 128),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("queued", C.c_int64), ("running", C.c_int64), ("retired", C.c_int64),
        ("failed", C.c_int64), ("cohorts", C.c_int64), ("cohorts_made", C.c_int64),
        ("rounds", C.c_int64), ("merges", C.c_int64), ("bytes_in_use", C.c_int64),
        ("bytes_budget", C.c_int64),
        ("estimate_scale", C.c_double), ("clock_s", C.c_double), ("atom_steps", C.c_int64),
        ("cancelled", C.c_int64),
    ]


def _setup(lib):
    if getattr(lib, "_server_ready", False):
        return
    vp = C.c_void_p
    lib.dmdx_server_create.argtypes = [C.POINTER(ServerOptions), C.POINTER(vp)]
    lib.dmdx_server_destroy.argtypes = [vp]
    lib.dmdx_server_submit.argtypes = [vp, C.c_void_p, C.c_int64, C.POINTER(C.c_int64)]
    lib.dmdx_server_pump.argtypes = [vp, C.POINTER(C.c_int32)]
    lib.dmdx_server_request_info.argtypes = [vp, C.c_int64, C.POINTER(RequestInfo)]
    lib.dmdx_server_result.argtypes = [vp, C.c_int64, C.c_int32, _dp, _dp, _dp]
    lib.dmdx_server_result_scalars.argtypes = [vp, C.c_int64, C.POINTER(Scalars)]
    lib.dmdx_server_release.argtypes = [vp, C.c_int64]
    lib.dmdx_server_cancel.argtypes = [vp, C.c_int64]
    lib.dmdx_server_get_stats.argtypes = [vp, C.POINTER(ServerStats)]
    lib.dmdx_server_step_latency.argtypes = [vp, _dp, C.POINTER(C.c_int64), C.c_int64,
                                             C.POINTER(C.c_int64)]
    lib._server_ready = True


class Server:
    """Device-side request server (include/dmdx.h, dmdx_server_*)."""

    def __init__(self, hbm_budget_bytes: int = 0, max_cohorts: int = 8, chunk_steps: int = 100,
                 graph_steps: int = 0, min_admit: int = 1, max_wait_ms: float = 0.0,
                 max_cohort_systems: int = 0, merge: bool = True):
        self.lib = _abi.load()
        _setup(self.lib)
        o = ServerOptions(int(hbm_budget_bytes), int(max_cohorts), int(chunk_steps),
                          int(graph_steps), int(min_admit), float(max_wait_ms),
                          int(max_cohort_systems), 0 if merge else 1)
        h = C.c_void_p()
        _abi.check(self.lib.dmdx_server_create(C.byref(o), C.byref(h)))
        self.h = h
        self._n = {}

    def close(self):
        if getattr(self, "h", None):
            self.lib.dmdx_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, spec: SystemSpec, n_steps: int) -> int:
        desc = spec.to_desc()
        rid = C.c_int64()
        _abi.check(self.lib.dmdx_server_submit(self.h, C.cast(C.byref(desc), C.c_void_p),
                                               int(n_steps), C.byref(rid)))
        self._n[rid.value] = spec.n
        return rid.value

    def pump(self) -> int:
        n = C.c_int32()
        _abi.check(self.lib.dmdx_server_pump(self.h, C.byref(n)))
        return n.value

    def info(self, rid: int) -> dict:
        i = RequestInfo()
        _abi.check(self.lib.dmdx_server_request_info(self.h, int(rid), C.byref(i)))
        return {"state": STATE_NAMES[i.state], "flags": i.flags, "n_atoms": i.n_atoms,
                "steps_target": i.steps_target, "steps_done": i.steps_done,
                "t_submit": i.t_submit, "t_admit": i.t_admit, "t_done": i.t_done,
                "est_bytes": i.est_bytes, "error": i.error.decode()}

    def result(self, rid: int) -> dict:
        n = self._n[rid]
        out = {}
        for what, name in ((0, "pos"), (1, "vel"), (2, "force")):
            a = [np.empty(n) for _ in range(3)]
            _abi.check(self.lib.dmdx_server_result(self.h, int(rid), what,
                                                   *[v.ctypes.data_as(_dp) for v in a]))
            out[name] = np.stack(a, 1)
        s = Scalars()
        _abi.check(self.lib.dmdx_server_result_scalars(self.h, int(rid), C.byref(s)))
        out["potential_energy"] = s.potential_energy
        out["kinetic_energy"] = s.kinetic_energy
        out["flags"] = s.flags
        out["steps_done"] = s.steps_done
        out["box"] = tuple(s.box)
        return out

    def cancel(self, rid: int):
        """Stop a queued or running request; a running one keeps the state
        it reached as its result."""
        _abi.check(self.lib.dmdx_server_cancel(self.h, int(rid)))

    def release(self, rid: int):
        _abi.check(self.lib.dmdx_server_release(self.h, int(rid)))
        self._n.pop(rid, None)

    def stats(self) -> dict:
        s = ServerStats()
        _abi.check(self.lib.dmdx_server_get_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def step_latency(self):
        n = C.c_int64()
        _abi.check(self.lib.dmdx_server_step_latency(self.h, None, None, 0, C.byref(n)))
        ms = np.empty(n.value)
        w = np.empty(n.value, dtype=np.int64)
        _abi.check(self.lib.dmdx_server_step_latency(
            self.h, ms.ctypes.data_as(_dp), w.ctypes.data_as(C.POINTER(C.c_int64)), n.value,
            C.byref(n)))
        return ms, w


def weighted_percentile(values, weights, p: float) -> float:
    """Smallest value v such that the weight of samples <= v is >= p% of
    the total."""
    values = np.asarray(values, dtype=np.float64)
    weights = np.asarray(weights, dtype=np.float64)
    if len(values) == 0:
        return float("nan")
    o = np.argsort(values, kind="stable")
    cw = np.cumsum(weights[o])
    k = int(np.searchsorted(cw, p / 100.0 * cw[-1], side="left"))
    return float(values[o][min(k, len(values) - 1)])


def run_poisson(server: Server, requests, rate_per_s: float, seed: int = 42,
                keep_results: bool = False, log=None) -> dict:
    """Submit `requests` ([(spec, n_steps), ...]) with exponential
    inter-arrival times of mean 1/rate_per_s (seeded), pump until every
    request has retired, and report aggregate atom-steps/s over the whole run
    (arrivals and drain), the loaded atom-steps/s (from the last arrival
    until the running requests fall below half their peak),
    per-step latency p50/p99 (weighted by steps), completion latency p50/p99
    (submit -> retire) and queueing delay p50/p99 (submit -> admit)."""
    rng = np.random.Generator(np.random.Philox(key=seed))
    gaps = rng.exponential(1.0 / rate_per_s, len(requests))
    arrive = np.cumsum(gaps) - gaps[0]
    ids = [None] * len(requests)
    t0 = time.perf_counter()
    nxt = 0
    retired = 0
    last_log = t0
    trace = []  # (seconds since start, atom-steps so far, requests running) after each pump
    while retired < len(requests):
        now = time.perf_counter() - t0
        while nxt < len(requests) and arrive[nxt] <= now:
            spec, steps = requests[nxt]
            ids[nxt] = server.submit(spec, steps)
            nxt += 1
        st = server.stats()
        if st["queued"] == 0 and st["running"] == 0:
            if nxt < len(requests):
                time.sleep(max(0.0, min(0.01, arrive[nxt] - now)))
            continue
        retired += server.pump()
        st2 = server.stats()
        trace.append((time.perf_counter() - t0, st2["atom_steps"], st2["running"]))
        if log and time.perf_counter() - last_log > 5.0:
            last_log = time.perf_counter()
            log(f"[serve] t={last_log - t0:.1f}s submitted {nxt}/{len(requests)} "
                f"retired {retired} cohorts {st['cohorts']} queued {st['queued']}")
    wall = time.perf_counter() - t0
    infos = [server.info(r) for r in ids]
    comp = np.array([i["t_done"] - i["t_submit"] for i in infos])
    wait = np.array([i["t_admit"] - i["t_submit"] for i in infos if i["t_admit"] >= 0])
    atom_steps = sum(i["n_atoms"] * i["steps_done"] for i in infos)
    ms, w = server.step_latency()
    st = server.stats()
    # loaded rate: from the last arrival until the running requests fall
    # below half their peak -- the server's throughput while it is full
    span = float(arrive[-1]) if len(arrive) else 0.0
    peak = max((r for _, _, r in trace), default=0)
    ta = span
    tb = next((t for t, _, r in trace if t > ta and r < 0.5 * peak), None)
    pts = [(t, a) for t, a, _ in trace if ta <= t <= (tb if tb is not None else t)]
    sustained = ((pts[-1][1] - pts[0][1]) / (pts[-1][0] - pts[0][0])
                 if len(pts) >= 2 and pts[-1][0] > pts[0][0] else None)
    tb = pts[-1][0] if pts else ta
    out = {
        "requests": len(requests), "rate_per_s": rate_per_s, "wall_s": wall,
        "atom_steps": int(atom_steps), "atom_steps_per_s": atom_steps / wall,
        "loaded_atom_steps_per_s": sustained,
        "loaded_window_s": [ta, tb], "peak_running": peak,
        "step_latency_ms": {"p50": weighted_percentile(ms, w, 50),
                            "p99": weighted_percentile(ms, w, 99)},
        "completion_s": {"p50": float(np.percentile(comp, 50)),
                         "p99": float(np.percentile(comp, 99))},
        "queue_wait_s": {"p50": float(np.percentile(wait, 50)) if len(wait) else math.nan,
                         "p99": float(np.percentile(wait, 99)) if len(wait) else math.nan},
        "failed": sum(1 for i in infos if i["state"] == "failed"),
        "cohorts_made": st["cohorts_made"], "rounds": st["rounds"], "merges": st["merges"],
        "bytes_budget": st["bytes_budget"], "estimate_scale": st["estimate_scale"],
    }
    if keep_results:
        out["ids"] = ids
    return out
