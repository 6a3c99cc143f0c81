#!/usr/bin/env python3
"""Throughput of DMD's step path on one MI355X: batched mixed precision vs
the plain one-system-at-a-time FP64 path vs the CPU (DMD's own compiled
components, 1 core).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Default workload ``replica_mix``: 256 independent replicas, sizes
log-uniform in 500-8000 atoms, 60% LJ fluids (rho* 0.70-0.90, T* 0.7-1.5)
and 40% bead-spring chain melts (harmonic bonds, angles, n=3 dihedrals,
LJ on non-excluded pairs), Berendsen NVT, dt 1 fs -- all replicas advanced
together in the same launches, pair arithmetic in FP32, integration and
reductions in FP64.  One "step" is one velocity-Verlet step of every
replica; ``value`` = sum(atoms) * K / wall seconds of the K timed steps
(state resident in HBM before the timed region starts).  The default K =
2000 is the whole run of the workload from its generated start (the chain
melts relax from a hot initial packing during it).  Per-step latency
percentiles come from separately synchronised steps after the timed region.

The default run also reports, beside ``value``: the plain path (each
replica alone, FP64, best of eager and hipGraph launches), the CPU
baseline -- DMD's own translation units built under oracle/_ref (the C
restatement where they are absent or the workload needs PME) -- on one
pinned core and on all host cores (one replica per worker), and a
served leg -- a bounded Poisson stream of LJ / SPME-water requests on
dmdx_server with its p50/p99 step and completion latencies and its speedup
over the plain path on the same requests (``serving``; --no-serving skips
it).

    python bench.py --workload serving [--requests N --rate R ...]

Serving alone: Poisson arrivals of independent LJ / water requests on
dmdx_server (see run_serving).

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "aggregate atom-steps/s (batched mixed vs plain GPU vs CPU); p50/p99 step latency"
PEAK_FP32_TFLOPS = 157.3   # MI355X vector FP32 (datasheet)
PEAK_FP64_TFLOPS = 78.6    # MI355X vector FP64 (datasheet)
PEAK_HBM_GBS = 8000.0      # MI355X HBM3E (datasheet)
# algorithmic work per in-cutoff pair (reference expressions,
# src/force/lennard_jones.cpp:62-88, src/force/coulomb_direct.cpp:60-83)
FLOP_LJ, FLOP_ERFC, FLOP_FUSED = 43, 49, 72


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(name, n_replicas, seed):
    from dmd_amd import workloads as W
    if name == "replica_mix":
        specs = W.replica_mix(n_replicas=n_replicas, seed=seed)
        desc = (f"replica_mix: {len(specs)} replicas, log-uniform 500-8000 atoms, "
                "60% LJ fluids / 40% bead-spring chain melts, Berendsen NVT, dt 1 fs")
    elif name == "water":
        specs = [W.water_box(n_side=16, grid=48, seed=seed)]
        desc = "water_box: 4096 flexible 3-site waters (12288 atoms), SPME 48^3, NH 300 K"
    elif name == "water_large":
        specs = [W.water_box(n_side=32, cutoff=1.0, grid=100, seed=seed,
                             thermostat=0)]
        desc = ("water_box: 32768 flexible 3-site waters (98304 atoms), SPME 100^3, rc 1.0 nm, "
                "NVE")
    elif name == "argon":
        specs = [W.argon_fcc(skin=0.1, seed=seed).with_(physics=1)]
        desc = "argon_fcc: 4000 LJ argon atoms, rho* 0.8442, NVE, dt 1 fs"
    else:
        raise SystemExit(f"unknown workload {name}")
    return specs, desc


def count_pairs(eng, specs):
    """In-cutoff unique pairs of every system (list pairs filtered by the
    cutoffs at the current positions) -> algorithmic FLOP of the pair
    forces of one step;
    also the unique pairs of each system's list (within the list radius)."""
    n_lj = n_c = n_both = 0
    list_pairs = []
    for k, s in enumerate(specs):
        pi, pj = eng.pairs(k)
        list_pairs.append(len(pi))
        if len(pi) == 0:
            continue
        pos = eng.positions(k)
        L = np.array(s.box)
        d = pos[pj] - pos[pi]
        d -= np.rint(d / L) * L
        r2 = (d * d).sum(1)
        lj = (r2 < s.lj_cutoff ** 2) if s.use_lj else np.zeros(len(r2), bool)
        c = (r2 < s.coulomb_cutoff ** 2) if s.coulomb_type != 0 else np.zeros(len(r2), bool)
        n_both += int((lj & c).sum())
        n_lj += int((lj & ~c).sum())
        n_c += int((c & ~lj).sum())
    return (n_lj * FLOP_LJ + n_c * FLOP_ERFC + n_both * FLOP_FUSED, n_lj + n_c + n_both,
            list_pairs)


def percentile(v, p):
    if len(v) == 0:
        return None
    v = np.sort(np.asarray(v))
    return float(v[min(len(v) - 1, int(math.ceil(p / 100.0 * len(v))) - 1)])


def run_batched(specs, steps, warmup, latency_steps):
    """Timed region: `steps` steps of every replica enqueued back to back,
    bracketed by device synchronisations, with HIP events around the pair
    kernel (the roofline kernel) on the engine's stream.  Then, outside the
    timed region: a pass with events around every kernel group for the
    per-group breakdown and the list-build roofline (events between all
    groups add gaps, so the timed region carries only the pair kernel's),
    and `latency_steps` single steps, each synchronised, for the per-step
    latency percentiles."""
    from dmd_amd import _abi
    from dmd_amd.engine import Engine
    mixed = [s.with_(precision=_abi.MIXED) for s in specs]
    eng = Engine(mixed, reuse_forces=True, use_graph=False)
    tune = eng.autotune(steps=10, cache=os.environ.get("DMDX_TUNE_CACHE"))
    log(f"[bench] autotune: {tune}")
    eng.step(warmup)
    eng.sync()
    eng.set_profiling(True, groups=("pair",))
    base = eng.kernel_times()
    t0 = time.perf_counter()
    eng.step(steps)
    eng.sync()
    wall = time.perf_counter() - t0
    kt = eng.kernel_times()
    kt = {k: (kt[k][0] - base[k][0], kt[k][1] - base[k][1]) for k in kt}
    # every group, outside the timed region: breakdown and list roofline
    nb = min(steps, 500)
    eng.set_profiling(True)
    builds0 = [eng.scalars(k)["list_builds"] for k in range(len(specs))]
    base = eng.kernel_times()
    eng.step(nb)
    eng.sync()
    kb = eng.kernel_times()
    builds = [eng.scalars(k)["list_builds"] - builds0[k] for k in range(len(specs))]
    kb = {k: (kb[k][0] - base[k][0], kb[k][1] - base[k][1]) for k in kb}
    breakdown = {"steps": nb, "ms_per_step": {k: round(v[0] / nb, 4) for k, v in kb.items()},
                 "list": kb["list"]}
    eng.set_profiling(False)
    lat = []
    for _ in range(latency_steps):
        a = time.perf_counter()
        eng.step(1)
        eng.sync()
        lat.append(time.perf_counter() - a)
    flags = [eng.scalars(k)["flags"] for k in range(len(specs))]
    return eng, wall, lat, kt, breakdown, flags, tune, builds


PLAIN_GRAPH_STEPS = 25


def run_plain(specs, steps, budget_s):
    """One system at a time, FP64: each system of the sample stepped alone,
    timed both with eager launches and with hipGraph-captured blocks of
    PLAIN_GRAPH_STEPS steps (the same trajectory); the faster of the two
    counts for that system (the plain path at its best launch mode).
    Returns (best atom-steps/s, systems used, {eager, graph} atom-steps/s)."""
    from dmd_amd.engine import Engine
    atom_steps, t_best, t_mode = 0, 0.0, {"eager": 0.0, "graph": 0.0}
    used = 0
    for s in specs:
        times = {}
        for mode in ("eager", "graph"):
            eng = Engine(s, reuse_forces=True, use_graph=(mode == "graph"),
                         graph_steps=PLAIN_GRAPH_STEPS)
            eng.step(PLAIN_GRAPH_STEPS)  # warm-up (captures the graph)
            eng.sync()
            a = time.perf_counter()
            eng.step(steps)
            eng.sync()
            times[mode] = time.perf_counter() - a
            eng.close()
        atom_steps += s.n * steps
        t_best += min(times.values())
        for mode in times:
            t_mode[mode] += times[mode]
        used += 1
        if t_best > budget_s:
            break
    return atom_steps / t_best, used, {m: atom_steps / t for m, t in t_mode.items()}


CPU_STEPS = 10


def _cpu_sim(spec, kind):
    """A CPU simulation of one replica: DMD's own translation units
    (oracle/_ref, "reference") or the C restatement of DMD's loops
    ("port"); both rebuild DMD's O(N^2) list every 10 computes and compute
    twice per step."""
    import oracle
    return oracle.RefSim(spec) if kind == "reference" else oracle.Sim(spec)


def cpu_kind(specs):
    """"reference" where DMD's translation units were compiled (oracle/_ref)
    and the workload has no PME (DMD's PME unit needs FFTW, which is not
    built), else "port"."""
    import oracle
    pme = any(s.coulomb_type == 2 for s in specs)
    return "reference" if oracle.ref_available() and not pme else "port"


def run_cpu(specs, budget_s, kind):
    """One CPU simulation kind on a bounded sample of the replicas, pinned to
    one host core."""
    atoms_steps, t, used = 0, 0.0, 0
    prev = os.sched_getaffinity(0)
    core = min(prev)
    os.sched_setaffinity(0, {core})
    try:
        for s in specs:
            sim = _cpu_sim(s, kind)
            a = time.perf_counter()
            sim.step(CPU_STEPS)
            t += time.perf_counter() - a
            atoms_steps += s.n * CPU_STEPS
            used += 1
            if t > budget_s:
                break
    finally:
        os.sched_setaffinity(0, prev)
    return atoms_steps / t, used, core


def _cpu_worker(args):
    spec, kind = args
    sim = _cpu_sim(spec, kind)
    a = time.perf_counter()
    sim.step(CPU_STEPS)
    return spec.n * CPU_STEPS, time.perf_counter() - a


def run_cpu_all_cores(specs, n_workers, kind):
    """The same CPU simulation with independent replicas spread one per
    worker process over `n_workers` host cores: aggregate atom-steps/s over
    the wall time of the sample (the replicas `specs`)."""
    import multiprocessing as mp
    ctx = mp.get_context("spawn")
    with ctx.Pool(n_workers) as pool:
        # warm the workers (imports, library load)
        pool.map(_cpu_worker, [(s, kind) for s in specs[:n_workers]])
        a = time.perf_counter()
        res = pool.map(_cpu_worker, [(s, kind) for s in specs], chunksize=1)
        wall = time.perf_counter() - a
    return sum(r[0] for r in res) / wall


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def accuracy(specs, n_check=4):
    """Force RMS relative error of the mixed path vs the FP64 device path on
    identical positions (same lists, same physics)."""
    from dmd_amd import _abi
    from dmd_amd.engine import Engine
    errs = []
    for s in specs[:n_check]:
        a = Engine(s.with_(precision=_abi.FP64))
        b = Engine(s.with_(precision=_abi.MIXED))
        a.compute()
        b.compute()
        fa, fb = a.forces(), b.forces()
        errs.append(float(np.sqrt(((fb - fa) ** 2).sum() / (fa ** 2).sum())))
    return max(errs)


def load_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` from profiles/traffic_<workload>.json
    (written by `python -m dmd_amd.profiling traffic` from two rocprofv3
    --pmc runs, FETCH_SIZE and WRITE_SIZE, calibrated on a copy kernel)."""
    p = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(p):
        return None
    try:
        rec = json.load(open(p)).get("kernels", {}).get(kernel)
        return rec.get("hbm_bytes_per_launch") if rec else None
    except (OSError, ValueError):
        return None


def run_serving(args):
    """Poisson arrivals of independent requests (LJ argon 1k-32k atoms or
    216-8192 waters with SPME; 1000-20000 steps log-uniform) served by
    dmdx_server: admission by device memory, batched cohorts, retirement at
    each request's step target."""
    rep, plain = serve(args)
    out = {
        "metric": "serving: aggregate atom-steps/s under Poisson arrivals; p50/p99 per-step "
                  "and completion latency",
        "value": rep["atom_steps_per_s"], "unit": "atom-steps/s", "n_gpus": 1,
        "steps": None, "warmup": 0, "ms_per_step": rep["step_latency_ms"]["p50"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 pair/PME arithmetic, fp64 accumulation and integration",
        "data": "synthetic (seeded generators; no datasets)",
        "config": {"workload": f"serving: {rep['requests']} requests, 50% LJ argon N~U[1k,32k], "
                               "50% water U[216,8192] molecules (SPME), steps log-uniform "
                               f"[1000,20000] x {args.steps_scale}, Poisson {args.rate}/s",
                   "chunk_steps": args.chunk, "max_cohorts": args.cohorts,
                   "graph_steps": args.graph_steps},
        "serving": rep,
    }
    if plain:
        out["plain_gpu"] = plain
        out["speedup_vs_plain"] = rep["atom_steps_per_s"] / plain["value"]
        if rep.get("loaded_atom_steps_per_s"):
            out["loaded_speedup_vs_plain"] = rep["loaded_atom_steps_per_s"] / plain["value"]
    print(json.dumps(out))


def serve(args):
    """The served request stream and, unless --no-plain, the plain path on
    the same requests.  Returns (run_poisson report, plain record or None)."""
    from dmd_amd import _abi
    from dmd_amd import workloads as W
    from dmd_amd.server import Server, run_poisson
    reqs = W.serving_requests(args.requests, seed=args.seed, steps_scale=args.steps_scale)
    log(f"[bench] serving: {len(reqs)} requests at {args.rate}/s, "
        f"{sum(s.n * n for s, n in reqs):.4g} atom-steps")
    srv = Server(chunk_steps=args.chunk, max_cohorts=args.cohorts, graph_steps=args.graph_steps)
    rep = run_poisson(srv, reqs, rate_per_s=args.rate, seed=args.seed, log=log)
    srv.close()
    log(f"[bench] serving: {rep['atom_steps_per_s']:.4g} atom-steps/s over {rep['wall_s']:.1f} s")
    if args.no_plain:
        return rep, None
    # the plain path on the same request mix: each request alone, FP64, at
    # its best launch mode, a bounded number of steps (throughput per
    # request is flat in the step count)
    log("[bench] serving: plain one-system-at-a-time FP64 GPU path on the same requests")
    value, used, modes = run_plain([s.with_(precision=_abi.FP64) for s, _ in reqs],
                                   args.plain_steps, args.plain_budget)
    return rep, {"value": value, "unit": "atom-steps/s", "by_launch_mode": modes,
                 "sample": f"first {used} requests, each alone, {args.plain_steps} steps, "
                           f"FP64, best of eager and {PLAIN_GRAPH_STEPS}-step hipGraph blocks "
                           "per request"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--latency-steps", type=int, default=200)
    ap.add_argument("--workload", default="replica_mix")
    ap.add_argument("--replicas", type=int, default=256)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--plain-steps", type=int, default=100)
    ap.add_argument("--plain-budget", type=float, default=10.0)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serving", action="store_true", help="skip the served-stream leg")
    ap.add_argument("--quick", action="store_true", help="batched leg only (profiling runs)")
    ap.add_argument("--requests", type=int, default=120, help="serving: number of requests")
    ap.add_argument("--rate", type=float, default=8.0, help="serving: arrivals per second")
    ap.add_argument("--steps-scale", type=float, default=1.0, help="serving: step-count factor")
    ap.add_argument("--chunk", type=int, default=100, help="serving: steps per round")
    ap.add_argument("--cohorts", type=int, default=8, help="serving: concurrent cohorts")
    ap.add_argument("--graph-steps", type=int, default=0, help="serving: captured block size")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: every system here is independent and one MI355X "
              "holds the whole workload; multi-GPU runs are not part of this build.",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "serving":
        run_serving(args)
        return
    from dmd_amd import engine as E
    log(f"[bench] building workload {args.workload}")
    specs, desc = make_workload(args.workload, args.replicas, args.seed)
    n_atoms = sum(s.n for s in specs)
    cpu_rec = None
    if not args.quick and not args.no_cpu:
        # before anything touches the GPU: the worker processes of the
        # all-cores leg start from a process without device state
        kind = cpu_kind(specs)
        what = {"reference": "DMD's own translation units (oracle/_ref: its ForceEngine, "
                             "LennardJones, bonded components, Integrator and thermostat "
                             "driven in SimulationEngine::step order)",
                "port": "the FP64 oracle (C restatement of DMD's loops)"}
        loops = "O(N^2) list rebuild every 10 computes, 2 computes per step"
        log(f"[bench] CPU baseline ({kind}): one pinned core, then one replica per core")
        cpu, used, core = run_cpu(specs, args.cpu_budget, kind)
        n_workers = max(1, min(16, len(os.sched_getaffinity(0))))
        cpu_all = run_cpu_all_cores(specs, n_workers, kind)
        cpu_rec = {
            "value": cpu, "unit": "atom-steps/s", "cores": 1, "kind": kind,
            "sample": f"first {used} replicas x {CPU_STEPS} steps of the same workload, "
                      f"{what[kind]}, FP64 ({loops}), pinned to host core {core}",
            "host": cpu_model(), "nproc": os.cpu_count(),
            "all_cores": {"value": cpu_all, "unit": "atom-steps/s", "cores": n_workers,
                          "sample": f"all {len(specs)} replicas x {CPU_STEPS} steps, one "
                                    f"replica at a time per worker process, {n_workers} "
                                    "workers"}}
        log(f"[bench] CPU ({kind}): {cpu:.4g} atom-steps/s on 1 core, {cpu_all:.4g} on "
            f"{n_workers}")
        if kind == "reference":
            # the C restatement beside it, on one core
            port, used_p, _ = run_cpu(specs, args.cpu_budget / 2, "port")
            cpu_rec["port"] = {
                "value": port, "unit": "atom-steps/s", "cores": 1, "kind": "port",
                "sample": f"first {used_p} replicas x {CPU_STEPS} steps, {what['port']}, "
                          f"FP64 ({loops}), pinned to host core {core}"}
            log(f"[bench] CPU (port): {port:.4g} atom-steps/s on 1 core")
    log(f"[bench] {len(specs)} systems, {n_atoms} atoms; batched mixed: warmup {args.warmup}, "
        f"timed {args.steps}")
    eng, wall, lat, kt, breakdown, flags, tune, builds = run_batched(
        specs, args.steps, args.warmup, args.latency_steps)
    value = n_atoms * args.steps / wall
    ms_per_step = wall / args.steps * 1e3
    log(f"[bench] batched mixed: {value:.4g} atom-steps/s, {ms_per_step:.3f} ms/step")

    # rooflines: the pair kernel (FP32 VALU) and the list build (HBM bytes
    # of the lists it writes)
    gms = breakdown["ms_per_step"]
    dom = "list" if gms["list"] > gms["pair"] else "pair"
    # achieved = FLOP of a step's pair forces / pair-kernel time per step
    # (the durations of its launches in the step)
    flop_per_step, n_pairs, list_pairs = count_pairs(eng, specs)
    pair_ms, pair_n = kt["pair"]
    pair_avg_s = pair_ms / max(pair_n, 1) * 1e-3
    pair_step_s = pair_ms / args.steps * 1e-3
    achieved = flop_per_step / pair_step_s / 1e12 if pair_n else 0.0
    roof_pair = {
        "kernel": "k_pair_mixed (LJ + real-space erfc pair forces)",
        "bound": "valu",
        "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
        "frac": round(achieved / PEAK_FP32_TFLOPS, 5),
        "traffic": load_traffic("k_pair_mixed", args.workload),
        "algorithmic_flop_per_step": flop_per_step, "pairs_in_cutoff": n_pairs,
        "ms_per_step": round(pair_step_s * 1e3, 4),
        "avg_launch_ms": round(pair_avg_s * 1e3, 4), "launches": pair_n,
    }
    # list build: per rebuilt system, 24 B read per atom and 4 B per row
    # entry (2 per unique pair within the list radius), over the breakdown
    # pass
    list_ms, list_n = breakdown["list"]
    list_bytes = sum(b * (24 * s.n + 8 * p) for b, s, p in zip(builds, specs, list_pairs))
    list_avg_s = list_ms / max(list_n, 1) * 1e-3
    gbs = list_bytes / max(list_n, 1) / list_avg_s / 1e9 if list_n else 0.0
    roof_list = {
        "kernel": "k_nbr_build_mixed + cell binning (list rebuilds of the systems that need one)",
        "bound": "hbm", "achieved": round(gbs, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
        "frac": round(gbs / PEAK_HBM_GBS, 5),
        "traffic": load_traffic("k_nbr_build_mixed", args.workload),
        "algorithmic_bytes_per_launch": list_bytes / max(list_n, 1),
        "system_rebuilds": int(sum(builds)), "avg_launch_ms": round(list_avg_s * 1e3, 4),
        "launches": list_n,
    }
    # the pair kernel is the kernel timed inside the timed region; the
    # breakdown pass names the largest group beside it
    roofline = dict(roof_pair)
    roofline["largest_group"] = dom
    # separate pass after the timed region, events around every group
    roofline["group_ms_per_step"] = gms
    roofline["group_pass_steps"] = breakdown["steps"]
    eng.close()

    out = {
        "metric": METRIC, "value": value, "unit": "atom-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 pair arithmetic, fp64 accumulation and integration",
        "data": "synthetic (seeded generators; no datasets)",
        "config": {"workload": desc, "systems": len(specs), "atoms": n_atoms,
                   "precision": "mixed", "parallelism": "batched systems, one GPU"},
        "latency_ms": {"p50": percentile(lat, 50) * 1e3 if lat else None,
                       "p99": percentile(lat, 99) * 1e3 if lat else None},
        "roofline": roofline,
        "roofline_list": roof_list,
        "failed_systems": int(sum(1 for f in flags if f)),
        "autotune": tune,
    }
    if not args.quick:
        out["force_rms_rel_err_mixed_vs_fp64"] = accuracy(specs)
        if not args.no_plain:
            log("[bench] plain one-system-at-a-time FP64 GPU path")
            plain, used, modes = run_plain(specs, args.plain_steps, args.plain_budget)
            out["plain_gpu"] = {"value": plain, "unit": "atom-steps/s", "by_launch_mode": modes,
                                "sample": f"first {used} replicas, each alone, "
                                          f"{args.plain_steps} steps, FP64, best of eager and "
                                          f"{PLAIN_GRAPH_STEPS}-step hipGraph blocks per replica"}
            out["speedup_vs_plain"] = value / plain
        if cpu_rec:
            out["cpu_baseline"] = cpu_rec
            out["speedup_vs_cpu"] = value / cpu_rec["value"]
            out["speedup_vs_cpu_all_cores"] = value / cpu_rec["all_cores"]["value"]
    if not args.quick and not args.no_serving:
        # the served form of the north-star target: a bounded Poisson stream
        # of LJ / SPME-water requests, and the plain path on the same requests
        rep, plain = serve(args)
        sv = {"workload": f"{rep['requests']} requests, 50% LJ argon N~U[1k,32k], 50% water "
                          "U[216,8192] molecules (SPME), steps log-uniform [1000,20000] x "
                          f"{args.steps_scale}, Poisson {args.rate}/s",
              "value": rep["atom_steps_per_s"], "unit": "atom-steps/s",
              "loaded_value": rep.get("loaded_atom_steps_per_s"),
              "wall_s": rep["wall_s"], "atom_steps": rep["atom_steps"],
              "step_latency_ms": rep["step_latency_ms"], "completion_s": rep["completion_s"],
              "queue_wait_s": rep.get("queue_wait_s"), "failed": rep.get("failed")}
        if plain:
            sv["plain_gpu"] = plain
            sv["speedup_vs_plain"] = rep["atom_steps_per_s"] / plain["value"]
        out["serving"] = sv
    # plain copy kernel of known size (256 MiB read + 256 MiB written per
    # launch): calibrates FETCH_SIZE / WRITE_SIZE counters in profiled runs
    out["copy_bandwidth_GBs"] = E.copy_bandwidth(1 << 28, 10)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
