#!/usr/bin/env python3
"""Throughput of DMD's step path on one MI355X: batched mixed precision vs
the plain one-system-at-a-time FP64 path vs the CPU (oracle, 1 core).

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

    python bench.py --workload serving [--requests N --rate R ...]

Serving: Poisson arrivals of independent LJ / water requests on
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


def run_plain(specs, steps, budget_s):
    """One system at a time, FP64, eager launches: each replica of the sample
    stepped alone."""
    from dmd_amd.engine import Engine
    atoms_steps, t = 0, 0.0
    used = 0
    for s in specs:
        eng = Engine(s, reuse_forces=True)
        eng.step(5)
        eng.sync()
        a = time.perf_counter()
        eng.step(steps)
        eng.sync()
        t += time.perf_counter() - a
        atoms_steps += s.n * steps
        used += 1
        eng.close()
        if t > budget_s:
            break
    return atoms_steps / t, used


def run_cpu(specs, budget_s):
    """The FP64 oracle (single-threaded C restatement of DMD's loops, O(N^2)
    list rebuild every 10 computes) on a bounded sample of the replicas."""
    import oracle
    atoms_steps, t, used, nsteps = 0, 0.0, 0, 10
    for s in specs:
        sim = oracle.Sim(s)
        a = time.perf_counter()
        sim.step(nsteps)
        t += time.perf_counter() - a
        atoms_steps += s.n * nsteps
        used += 1
        if t > budget_s:
            break
    return atoms_steps / t, used, nsteps


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
    from dmd_amd import _abi
    from dmd_amd import workloads as W
    from dmd_amd.server import Server, run_poisson
    reqs = W.serving_requests(args.requests, seed=args.seed, steps_scale=args.steps_scale)
    log(f"[bench] serving: {len(reqs)} requests at {args.rate}/s, "
        f"{sum(s.n * n for s, n in reqs):.4g} atom-steps")
    srv = Server(chunk_steps=args.chunk, max_cohorts=args.cohorts, graph_steps=args.graph_steps)
    rep = run_poisson(srv, reqs, rate_per_s=args.rate, seed=args.seed, log=log)
    srv.close()
    # the plain path on the same request mix: each request alone, FP64,
    # eager, a bounded number of steps (throughput per request is flat in
    # the step count)
    plain_value, used = None, 0
    if not args.no_plain:
        log("[bench] serving: plain one-system-at-a-time FP64 GPU path on the same requests")
        plain_value, used = run_plain([s.with_(precision=_abi.FP64) for s, _ in reqs],
                                      args.plain_steps, args.plain_budget)
    out = {
        "metric": "serving: aggregate atom-steps/s under Poisson arrivals; p50/p99 per-step "
                  "and completion latency",
        "value": rep["atom_steps_per_s"], "unit": "atom-steps/s", "n_gpus": 1,
        "steps": None, "warmup": 0, "ms_per_step": rep["step_latency_ms"]["p50"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 pair/PME arithmetic, fp64 accumulation and integration",
        "data": "synthetic (seeded generators; no datasets)",
        "config": {"workload": f"serving: {len(reqs)} requests, 50% LJ argon N~U[1k,32k], "
                               "50% water U[216,8192] molecules (SPME), steps log-uniform "
                               f"[1000,20000] x {args.steps_scale}, Poisson {args.rate}/s",
                   "chunk_steps": args.chunk, "max_cohorts": args.cohorts,
                   "graph_steps": args.graph_steps},
        "serving": rep,
    }
    if plain_value:
        out["plain_gpu"] = {"value": plain_value, "unit": "atom-steps/s",
                            "sample": f"first {used} requests, each alone, {args.plain_steps} "
                                      "steps, FP64, eager"}
        out["speedup_vs_plain"] = rep["atom_steps_per_s"] / plain_value
        if rep.get("loaded_atom_steps_per_s"):
            out["loaded_speedup_vs_plain"] = rep["loaded_atom_steps_per_s"] / plain_value
    print(json.dumps(out))


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
    ap.add_argument("--quick", action="store_true", help="batched leg only (profiling runs)")
    ap.add_argument("--requests", type=int, default=120, help="serving: number of requests")
    ap.add_argument("--rate", type=float, default=8.0, help="serving: arrivals per second")
    ap.add_argument("--steps-scale", type=float, default=1.0, help="serving: step-count factor")
    ap.add_argument("--chunk", type=int, default=50, help="serving: steps per round")
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
            plain, used = run_plain(specs, args.plain_steps, args.plain_budget)
            out["plain_gpu"] = {"value": plain, "unit": "atom-steps/s",
                                "sample": f"first {used} replicas, each alone, "
                                          f"{args.plain_steps} steps, FP64, eager"}
            out["speedup_vs_plain"] = value / plain
        if not args.no_cpu:
            log("[bench] CPU baseline (oracle, 1 core)")
            cpu, used, nst = run_cpu(specs, args.cpu_budget)
            out["cpu_baseline"] = {
                "value": cpu, "unit": "atom-steps/s", "cores": 1, "kind": "port",
                "sample": f"first {used} replicas x {nst} steps of the same workload, FP64 "
                          "oracle (DMD's loops: O(N^2) list rebuild every 10 computes, 2 "
                          "computes per step)",
                "host": platform.processor() or platform.machine(),
                "nproc": os.cpu_count()}
            out["speedup_vs_cpu"] = value / cpu
    # plain copy kernel of known size (256 MiB read + 256 MiB written per
    # launch): calibrates FETCH_SIZE / WRITE_SIZE counters in profiled runs
    out["copy_bandwidth_GBs"] = E.copy_bandwidth(1 << 28, 10)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
