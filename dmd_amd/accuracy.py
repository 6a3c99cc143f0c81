"""Accuracy of the mixed-precision path against the FP64 path.

    python -m dmd_amd.accuracy [--which c0,c1,c2] [--steps N] [--ensemble R]
                               [--power S] [--variant V] [--out FILE]

For each workload, from a configuration relaxed by 200 FP64 steps: force
RMS relative error and potential-energy error of the mixed path against
the FP64 path on identical positions, and the NVE energy drift (slope of a
least-squares line through E_pot + E_kin, in kJ/mol/ns/atom) of both paths
over the same run from the same start, plus the drift of a second FP64 run
from positions moved by 1e-9 nm (the statistical noise floor of comparing
drifts of decorrelating trajectories).  With --ensemble R, both precisions
also run from R starts moved by independent 1e-9 nm perturbations: the mean
and standard deviation of each precision's drift, and the difference of the
means with its standard error, separate a change of precision from the
spread of chaotic trajectories.  The
FP64 device path is the reference here because it reproduces the FP64 CPU
oracle (DMD's own arithmetic) to 1e-10 in forces; running the CPU oracle for
10^4 steps of 10^5 atoms is not practical.

--power S measures the rate at which the mixed path's force error does
work: P = sum_i (F_mixed(x) - F_fp64(x)) . v_i per atom, at S states spaced
--every steps apart along one FP64 NVE trajectory, each mixed force from a
fresh engine on the state (lists and cell frames built there); with
--along-mixed, along the mixed trajectory with the forces its own steps
computed (lists and cell-local frames as old as its last rebuild), each FP64
force from a fresh engine.  Its mean is the first-order energy injection of
the force error, with a standard error far below that of drift ensembles,
because both forces see the same positions and velocities.  --variant changes the electrostatics of every
workload first, to localize an effect: "no-recip" keeps the real-space erfc
term and drops PME (COULOMB_CUTOFF), "no-elec" drops all electrostatics,
"rebuild-every-step" rebuilds lists (and the mixed path's cell-local frames)
every step.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import _abi
from . import workloads as W


def relaxed(spec, steps: int = 200):
    """The spec moved on by `steps` FP64 steps (a liquid configuration; a
    perfect lattice has vanishing net forces, which makes relative force
    errors meaningless)."""
    from .engine import Engine
    e = Engine(spec.with_(precision=_abi.FP64))
    e.step(steps)
    out = spec.with_(pos=e.positions(), vel=e.velocities())
    e.close()
    return out


def force_error(spec):
    """RMS and max relative force error, relative PE error (mixed vs FP64)."""
    from .engine import Engine
    a = Engine(spec.with_(precision=_abi.FP64))
    b = Engine(spec.with_(precision=_abi.MIXED))
    a.compute()
    b.compute()
    fa, fb = a.forces(), b.forces()
    pa, pb = a.scalars()["pe"], b.scalars()["pe"]
    a.close()
    b.close()
    d = fb - fa
    return {"force_rms_rel": float(np.sqrt((d * d).sum() / (fa * fa).sum())),
            "force_max_rel": float(np.abs(d).max() / np.abs(fa).max()),
            "pe_rel": float(abs(pb - pa) / abs(pa))}


def nve_drift(spec, steps: int, every: int = 10, precision=_abi.MIXED):
    """Total energy every `every` steps over an NVE run; drift = slope of the
    least-squares line, kJ/mol/ns per atom."""
    from .engine import Engine
    s = spec.with_(precision=precision, thermostat=_abi.THERMO_NONE,
                   barostat=_abi.BARO_NONE)
    e = Engine(s)
    t, E = [], []
    done = 0
    t0 = time.perf_counter()
    while done < steps:
        k = min(every, steps - done)
        e.step(k)
        done += k
        sc = e.scalars()
        t.append(done * s.dt)
        E.append(sc["pe"] + sc["ke"])
    wall = time.perf_counter() - t0
    flags = sc["flags"]
    e.close()
    t = np.array(t)  # ps
    E = np.array(E)
    slope = float(np.polyfit(t, E, 1)[0]) if len(t) > 2 else float("nan")  # kJ/mol/ps
    return {"drift_kj_mol_ns_atom": slope * 1000.0 / s.n, "e0": float(E[0]),
            "e_span": float(E.max() - E.min()), "steps": steps, "dt_ps": s.dt,
            "flags": flags, "wall_s": wall}


def force_error_power(spec, samples: int, every: int, along_mixed: bool = False):
    """Mean and standard error of P = sum (F_mixed - F_fp64) . v / n, in
    kJ/mol/ns per atom (the unit of the drifts), at `samples` states `every`
    steps apart of one NVE trajectory.
    along_mixed=False: an FP64 trajectory; each mixed force from a fresh
    mixed engine on the state (list and cell-local frames built there).
    along_mixed=True: the mixed trajectory itself, with the forces its own
    step computed (lists and frames as old as its last rebuild); each FP64
    force from a fresh FP64 engine on that state."""
    from .engine import Engine
    s = spec.with_(precision=_abi.FP64, thermostat=_abi.THERMO_NONE, barostat=_abi.BARO_NONE)
    lead = _abi.MIXED if along_mixed else _abi.FP64
    other = _abi.FP64 if along_mixed else _abi.MIXED
    ref = Engine(s.with_(precision=lead))
    p = []
    for _ in range(samples):
        ref.step(every)
        x, v, f_lead = ref.positions(), ref.velocities(), ref.forces()
        m = Engine(s.with_(pos=x, vel=v, precision=other))
        m.compute()
        f_other = m.forces()
        m.close()
        df = (f_lead - f_other) if along_mixed else (f_other - f_lead)
        p.append(float((df * v).sum()) / s.n * 1000.0)  # kJ/mol/ps -> /ns
    ref.close()
    p = np.array(p)
    return {"samples": samples, "every": every, "mean": float(p.mean()),
            "stderr": float(p.std(ddof=1) / np.sqrt(len(p))), "rms": float(np.sqrt((p * p).mean()))}


def variant(spec, name: str):
    if name in ("", "as-is"):
        return spec
    if name == "no-recip":
        return spec.with_(coulomb_type=_abi.COULOMB_CUTOFF)
    if name == "no-elec":
        return spec.with_(coulomb_type=_abi.COULOMB_NONE)
    if name == "rebuild-every-step":
        return spec.with_(rebuild_interval=1)
    raise ValueError(name)


def workload(name: str, seed: int = 42):
    if name == "c0":
        return W.argon_fcc(skin=0.1, seed=seed).with_(physics=_abi.CORRECTED,
                                                      rebuild_interval=-1)
    if name == "c1":
        return W.water_box(n_side=16, grid=48, seed=seed)
    if name == "c2":
        return W.water_box(n_side=32, cutoff=1.0, grid=100, seed=seed,
                           thermostat=_abi.THERMO_NONE)
    raise ValueError(name)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--which", default="c0,c1,c2")
    ap.add_argument("--steps", type=int, default=0,
                    help="NVE steps (0: the workload's own run length)")
    ap.add_argument("--every", type=int, default=20)
    ap.add_argument("--ensemble", type=int, default=0,
                    help="perturbed starts per precision for drift statistics")
    ap.add_argument("--power", type=int, default=0,
                    help="states for the force-error power along an FP64 trajectory")
    ap.add_argument("--variant", default="as-is",
                    choices=["as-is", "no-recip", "no-elec", "rebuild-every-step"])
    ap.add_argument("--along-mixed", action="store_true",
                    help="--power along the mixed trajectory (its own forces)")
    ap.add_argument("--no-drift", action="store_true", help="skip the single NVE runs")
    ap.add_argument("--out", default="")
    args = ap.parse_args(argv)
    own = {"c0": 1000, "c1": 2000, "c2": 10000}
    rows = {}
    for name in args.which.split(","):
        spec = variant(workload(name), args.variant)
        steps = args.steps or own[name]
        print(f"[accuracy] {name} ({args.variant}): {spec.n} atoms, {steps} NVE steps",
              file=sys.stderr, flush=True)
        spec = relaxed(spec)
        r = {"atoms": spec.n, "variant": args.variant, **force_error(spec)}
        if args.power:
            r["force_error_power"] = force_error_power(spec, args.power, args.every,
                                                       args.along_mixed)
            r["force_error_power"]["along"] = "mixed" if args.along_mixed else "fp64"
            print(json.dumps({name: r["force_error_power"]}), file=sys.stderr, flush=True)
        if args.no_drift:
            rows[name] = r
            continue
        r["nve_fp64"] = nve_drift(spec, steps, args.every, _abi.FP64)
        r["nve_mixed"] = nve_drift(spec, steps, args.every, _abi.MIXED)
        # noise floor: the FP64 run again from positions moved by 1e-9 nm
        # (trajectories decorrelate, so drift estimates differ by this much
        # without any change of precision)
        rng = np.random.default_rng(1)
        jit = spec.with_(pos=spec.pos + rng.uniform(-1e-9, 1e-9, spec.pos.shape))
        r["nve_fp64_perturbed"] = nve_drift(jit, steps, args.every, _abi.FP64)
        d0 = r["nve_fp64"]["drift_kj_mol_ns_atom"]
        r["drift_difference"] = r["nve_mixed"]["drift_kj_mol_ns_atom"] - d0
        r["drift_noise_floor"] = r["nve_fp64_perturbed"]["drift_kj_mol_ns_atom"] - d0
        if args.ensemble > 1:
            ens = {}
            for prec, label in ((_abi.FP64, "fp64"), (_abi.MIXED, "mixed")):
                dr = []
                for k in range(args.ensemble):
                    g = np.random.default_rng(100 + k)
                    sk = spec.with_(pos=spec.pos + g.uniform(-1e-9, 1e-9, spec.pos.shape))
                    dr.append(nve_drift(sk, steps, args.every, prec)["drift_kj_mol_ns_atom"])
                ens[label] = {"drifts": dr, "mean": float(np.mean(dr)),
                              "std": float(np.std(dr, ddof=1))}
            n = args.ensemble
            se = float(np.sqrt(ens["fp64"]["std"] ** 2 / n + ens["mixed"]["std"] ** 2 / n))
            ens["mean_difference"] = ens["mixed"]["mean"] - ens["fp64"]["mean"]
            ens["difference_stderr"] = se
            r["drift_ensemble"] = ens
        rows[name] = r
        print(json.dumps({name: r}), file=sys.stderr, flush=True)
    out = json.dumps(rows, indent=1)
    if args.out:
        with open(args.out, "w") as f:
            f.write(out + "\n")
    print(out)


if __name__ == "__main__":
    main()
