"""SimulationEngine mirror over the C-ABI (include/dmdx.h, dmdx_engine_*).

``Engine`` replaces DMD's SimulationEngine (src/sim/simulation_engine.h:16-76)
for one system or for a batch of independent systems that step together in
the same device launches.  State lives in HBM; the getters copy to host.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from .system import SystemSpec


class Scalars(C.Structure):
    _fields_ = [
        ("potential_energy", C.c_double), ("energy_terms", C.c_double * 8),
        ("kinetic_energy", C.c_double), ("zeta", C.c_double), ("piston_velocity", C.c_double),
        ("virial", C.c_double), ("pressure", C.c_double), ("box", C.c_double * 3),
        ("step_since_rebuild", C.c_int32), ("flags", C.c_int32), ("steps_done", C.c_int64),
        ("list_builds", C.c_int64),
    ]


TERM_NAMES = ("lj", "coulomb", "bond", "angle", "dihedral", "improper", "pme", "exclusion")

_dp = C.POINTER(C.c_double)


def _setup(lib):
    if getattr(lib, "_engine_ready", False):
        return
    lib.dmdx_engine_create.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_void_p)]
    for name in ("dmdx_engine_destroy", "dmdx_engine_compute", "dmdx_engine_sync"):
        getattr(lib, name).argtypes = [C.c_void_p]
    lib.dmdx_engine_step.argtypes = [C.c_void_p, C.c_int64]
    lib.dmdx_engine_set_options.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]
    lib.dmdx_engine_get_arrays.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _dp, _dp, _dp]
    lib.dmdx_engine_set_arrays.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _dp, _dp, _dp]
    lib.dmdx_engine_get_scalars.argtypes = [C.c_void_p, C.c_int32, C.POINTER(Scalars)]
    lib.dmdx_engine_set_step_since_rebuild.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
    lib.dmdx_engine_set_active.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
    lib.dmdx_engine_adopt.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32]
    lib.dmdx_engine_get_pairs.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_int32),
                                          C.POINTER(C.c_int32), C.c_int64,
                                          C.POINTER(C.c_int64)]
    lib.dmdx_engine_n_systems.argtypes = [C.c_void_p, C.POINTER(C.c_int32)]
    lib.dmdx_engine_set_profiling.argtypes = [C.c_void_p, C.c_int32]
    lib.dmdx_engine_set_profiling_groups.argtypes = [C.c_void_p, C.c_uint32]
    lib.dmdx_engine_kernel_time.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_double),
                                            C.POINTER(C.c_int64)]
    lib.dmdx_copy_bandwidth.argtypes = [C.c_int64, C.c_int32, C.POINTER(C.c_double)]
    lib.dmdx_engine_set_tuning.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
    lib.dmdx_engine_get_tuning.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]
    lib.dmdx_engine_autotune.argtypes = [C.c_void_p, C.c_int32, C.c_char_p,
                                         C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                         C.POINTER(C.c_int32), C.POINTER(C.c_double)]
    lib.dmdx_engine_get_list_skin.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_double)]
    lib.dmdx_engine_get_cell_div.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]
    lib.dmdx_engine_snapshot_begin.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_int64)]
    lib.dmdx_engine_snapshot_read.argtypes = [C.c_void_p, C.c_int64, _dp, _dp, _dp,
                                              C.POINTER(C.c_int64)]
    lib._engine_ready = True


KERNEL_GROUPS = ("pair", "list", "bonded", "pme", "gather", "integrate")
# autotuner candidates in the order of dmdx_engine_autotune's times_ms:
# per-atom mixed pair kernel (lanes per atom x list entries per lane pass),
# row tiles for LJ-only systems (entries per block), list build (per atom
# with 8/16/32 lanes, per cell with 2/4/8 waves per cell, per cell staged in
# LDS with 8/16 lanes per atom)
PAIR_SHAPES = ("2x8", "4x8", "8x8", "4x4", "8x4", "16x4", "8x2", "16x2")
ROW_SHAPES = ("rows4", "rows8")
BUILD_SHAPES = ("atom8", "atom16", "atom32", "cell2", "cell4", "cell8", "stage8", "stage16")


def copy_bandwidth(nbytes: int = 1 << 28, repeats: int = 20) -> float:
    """GB/s of a plain device-to-device copy kernel (read + write bytes)."""
    lib = _abi.load()
    _setup(lib)
    g = C.c_double()
    _abi.check(lib.dmdx_copy_bandwidth(int(nbytes), int(repeats), C.byref(g)))
    return g.value


class Engine:
    """Device-resident SimulationEngine for one or more systems."""

    def __init__(self, specs, reuse_forces: bool = True, use_graph: bool = False,
                 graph_steps: int = 10):
        if isinstance(specs, SystemSpec):
            specs = [specs]
        self.specs = list(specs)
        self.lib = _abi.load()
        _setup(self.lib)
        descs = (_abi.SystemDesc * len(self.specs))()
        for k, s in enumerate(self.specs):
            descs[k] = s.to_desc()
        h = C.c_void_p()
        _abi.check(self.lib.dmdx_engine_create(C.cast(descs, C.c_void_p), len(self.specs),
                                               C.byref(h)))
        self.h = h
        self.set_options(reuse_forces, use_graph, graph_steps)

    def close(self):
        if getattr(self, "h", None):
            self.lib.dmdx_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_options(self, reuse_forces=True, use_graph=False, graph_steps=10):
        _abi.check(self.lib.dmdx_engine_set_options(self.h, int(reuse_forces), int(use_graph),
                                                    int(graph_steps)))

    def step(self, nsteps: int = 1):
        _abi.check(self.lib.dmdx_engine_step(self.h, int(nsteps)))

    def compute(self):
        _abi.check(self.lib.dmdx_engine_compute(self.h))

    def sync(self):
        _abi.check(self.lib.dmdx_engine_sync(self.h))

    def _get(self, sys, what):
        n = self.specs[sys].n
        a = [np.empty(n) for _ in range(3)]
        _abi.check(self.lib.dmdx_engine_get_arrays(self.h, sys, what,
                                                   *[v.ctypes.data_as(_dp) for v in a]))
        return np.stack(a, 1)

    def _set(self, sys, what, arr):
        arr = np.asarray(arr, dtype=np.float64)
        a = [np.ascontiguousarray(arr[:, k]) for k in range(3)]
        _abi.check(self.lib.dmdx_engine_set_arrays(self.h, sys, what,
                                                   *[v.ctypes.data_as(_dp) for v in a]))

    def positions(self, sys: int = 0):
        return self._get(sys, 0)

    def velocities(self, sys: int = 0):
        return self._get(sys, 1)

    def forces(self, sys: int = 0):
        return self._get(sys, 2)

    def set_positions(self, pos, sys: int = 0):
        self._set(sys, 0, pos)

    def set_velocities(self, vel, sys: int = 0):
        self._set(sys, 1, vel)

    def scalars(self, sys: int = 0) -> dict:
        s = Scalars()
        _abi.check(self.lib.dmdx_engine_get_scalars(self.h, sys, C.byref(s)))
        return {
            "pe": s.potential_energy, "terms": dict(zip(TERM_NAMES, list(s.energy_terms))),
            "ke": s.kinetic_energy, "zeta": s.zeta, "eps_dot": s.piston_velocity,
            "virial": s.virial, "pressure": s.pressure, "box": tuple(s.box),
            "counter": s.step_since_rebuild, "flags": s.flags, "steps_done": s.steps_done,
            "list_builds": s.list_builds,
        }

    def set_step_since_rebuild(self, counter: int, sys: int = 0):
        _abi.check(self.lib.dmdx_engine_set_step_since_rebuild(self.h, sys, int(counter)))

    def set_active(self, sys: int, active: bool):
        _abi.check(self.lib.dmdx_engine_set_active(self.h, sys, int(active)))

    def adopt(self, sys: int, src: "Engine", src_sys: int = 0):
        """Move system src_sys of `src`, with its whole state, into slot `sys`
        (both created from the same description; dmdx_engine_adopt)."""
        _abi.check(self.lib.dmdx_engine_adopt(self.h, int(sys), src.h, int(src_sys)))

    def set_tuning(self, pair_lanes: int | None = None, build_lanes: int | None = None,
                   pair_width: int | None = None, build_cells: bool | None = None,
                   build_tasks: int | None = None, tune_skin: bool | None = None,
                   pair_rows: int | None = None, build_stage: int | None = None,
                   cell_div: int | None = None, coul_lanes: int | None = None,
                   coul_width: int | None = None, fft_planes: int | None = None):
        """Launch shapes of the mixed kernels (dmdx_engine_set_tuning);
        build_cells: list built per cell (default) or per atom; build_tasks:
        waves per cell of the per-cell build (1, 2, 4, 8); pair_rows: LJ-only
        systems on row tiles staged in LDS with 4 or 8 list entries per lane
        block, or 0 (default) for the per-atom kernel; build_stage: the
        per-cell build with each cell's candidates staged in LDS for 8 or 16
        lanes per atom, or 0 (default) for candidates shared in registers by
        build_tasks waves; cell_div (before the first step): cells of the
        list radius (1) or half of it (2) for every mixed system with a list,
        0 for the density rule; coul_lanes, coul_width: the pair kernel's
        shape for the systems with real-space Coulomb pairs (pair_lanes and
        pair_width apply to the LJ-only systems); fft_planes (before the first
        step): the mixed PME x and y passes as one xy plane per workgroup, 0
        never, 1 where a plane fits in LDS, 2 (default) when the batch has at
        least 256 planes."""
        if cell_div is not None:  # before the first step: the batch is made again
            _abi.check(self.lib.dmdx_engine_set_tuning(self.h, 9, int(cell_div)))
        if fft_planes is not None:  # likewise
            _abi.check(self.lib.dmdx_engine_set_tuning(self.h, 12, int(fft_planes)))
        if build_stage is not None:
            _abi.check(self.lib.dmdx_engine_set_tuning(self.h, 7, int(build_stage)))
        if pair_rows is not None:
            _abi.check(self.lib.dmdx_engine_set_tuning(self.h, 6, int(pair_rows)))
        if build_cells is not None:
            _abi.check(self.lib.dmdx_engine_set_tuning(self.h, 3, int(bool(build_cells))))
        if build_tasks is not None:
            _abi.check(self.lib.dmdx_engine_set_tuning(self.h, 4, int(build_tasks)))
        if tune_skin is not None:
            _abi.check(self.lib.dmdx_engine_set_tuning(self.h, 5, int(bool(tune_skin))))
        if pair_lanes is not None:
            _abi.check(self.lib.dmdx_engine_set_tuning(self.h, 0, int(pair_lanes)))
        if build_lanes is not None:
            _abi.check(self.lib.dmdx_engine_set_tuning(self.h, 1, int(build_lanes)))
        if pair_width is not None:
            _abi.check(self.lib.dmdx_engine_set_tuning(self.h, 2, int(pair_width)))
        if coul_lanes is not None:
            _abi.check(self.lib.dmdx_engine_set_tuning(self.h, 10, int(coul_lanes)))
        if coul_width is not None:
            _abi.check(self.lib.dmdx_engine_set_tuning(self.h, 11, int(coul_width)))

    def autotune(self, steps: int = 5, cache: str | None = None) -> dict:
        """Time launch shapes of the mixed pair kernel (lanes per atom x
        entries per lane pass) and list build (8/16/32 lanes per atom, or per
        cell with 2/4/8 waves per cell) on a scratch copy of the systems and
        apply the fastest (dmdx_engine_autotune); on an engine that has not
        stepped, first the list skins of displacement-criterion systems.
        `cache`: text file of earlier results.  "build" names the chosen list
        build, "pair_rows" the row-tile setting of LJ-only systems (0: the
        per-atom kernel), "skins" the widened skins, "cell_div" how many
        systems have cells of the list radius (1) or half of it (2),
        "fft_lines" the FFT tiling (0: the grid-size rule), "coul_lanes" and
        "coul_width" the shape of the systems with Coulomb pairs ("coul_ms":
        their timings; -1 where no system has such pairs)."""
        p, w, b = C.c_int32(), C.c_int32(), C.c_int32()
        npr = len(PAIR_SHAPES) + len(ROW_SHAPES)
        nco = npr + len(BUILD_SHAPES)
        t = (C.c_double * (nco + len(PAIR_SHAPES)))()
        _abi.check(self.lib.dmdx_engine_autotune(
            self.h, int(steps), cache.encode() if cache else None, C.byref(p), C.byref(w),
            C.byref(b), t))
        cells, tasks, stage = self.get_tuning(3), self.get_tuning(4), self.get_tuning(7)
        # skins the autotuner widened: (original skin, chosen skin) -> systems
        widened = {}
        for k, s in enumerate(self.specs):
            now = self.list_skin(k)
            if now != s.list_skin:
                key = (round(s.list_skin, 6), round(now, 6))
                widened[key] = widened.get(key, 0) + 1
        return {"pair_lanes": p.value, "pair_width": w.value, "build_lanes": b.value,
                "build_cells": bool(cells), "build_tasks": tasks,
                "build": (f"stage{stage}" if stage else f"cell{tasks}") if cells
                else f"atom{b.value}",
                "skins": [{"skin": a, "tuned": t, "systems": n}
                          for (a, t), n in sorted(widened.items())],
                "pair_rows": self.get_tuning(6),
                "fft_lines": self.get_tuning(8),
                "cell_div": {c: sum(1 for k in range(len(self.specs)) if self.cell_div(k) == c)
                             for c in (1, 2)},
                "pair_ms": {k: t[i] for i, k in enumerate(PAIR_SHAPES)},
                "rows_ms": {k: t[len(PAIR_SHAPES) + i] for i, k in enumerate(ROW_SHAPES)},
                "build_ms": {k: t[npr + i] for i, k in enumerate(BUILD_SHAPES)},
                "coul_lanes": self.get_tuning(10), "coul_width": self.get_tuning(11),
                "coul_ms": {k: t[nco + i] for i, k in enumerate(PAIR_SHAPES)}}

    def cell_div(self, sys: int = 0) -> int:
        """Cells per list radius of system `sys`'s list grid (1 or 2;
        dmdx_engine_get_cell_div)."""
        v = C.c_int32()
        _abi.check(self.lib.dmdx_engine_get_cell_div(self.h, int(sys), C.byref(v)))
        return v.value

    def list_skin(self, sys: int = 0) -> float:
        """The list skin system `sys` is built with (dmdx_engine_get_list_skin)."""
        v = C.c_double()
        _abi.check(self.lib.dmdx_engine_get_list_skin(self.h, int(sys), C.byref(v)))
        return v.value

    def get_tuning(self, key: int) -> int:
        """Current value of a tuning key (dmdx_engine_get_tuning)."""
        v = C.c_int32()
        _abi.check(self.lib.dmdx_engine_get_tuning(self.h, int(key), C.byref(v)))
        return v.value

    def set_profiling(self, on: bool, groups=None):
        """HIP events around every kernel group, or only around `groups`
        (names from KERNEL_GROUPS)."""
        if groups is None or not on:
            _abi.check(self.lib.dmdx_engine_set_profiling(self.h, int(on)))
        else:
            mask = 0
            for g in groups:
                mask |= 1 << KERNEL_GROUPS.index(g)
            _abi.check(self.lib.dmdx_engine_set_profiling_groups(self.h, mask))

    def kernel_times(self) -> dict:
        """{group: (total_ms, launches)} of the HIP-event-timed kernel groups."""
        out = {}
        for k, name in enumerate(KERNEL_GROUPS):
            ms, n = C.c_double(), C.c_int64()
            _abi.check(self.lib.dmdx_engine_kernel_time(self.h, k, C.byref(ms), C.byref(n)))
            out[name] = (ms.value, n.value)
        return out

    def snapshot_begin(self, sys: int = 0) -> int:
        """Enqueue a frame of system `sys` after the work enqueued so far and
        return its ticket (dmdx_engine_snapshot_begin); later steps overlap
        its copy to host memory."""
        t = C.c_int64()
        _abi.check(self.lib.dmdx_engine_snapshot_begin(self.h, int(sys), C.byref(t)))
        return t.value

    def snapshot_read(self, ticket: int, sys: int = 0) -> dict:
        """Wait for a frame (dmdx_engine_snapshot_read): positions and
        velocities as (n, 3) arrays, box edges and steps_done."""
        n = self.specs[sys].n
        pos = np.empty((n, 3))
        vel = np.empty((n, 3))
        box = np.empty(3)
        done = C.c_int64()
        _abi.check(self.lib.dmdx_engine_snapshot_read(
            self.h, int(ticket), pos.ctypes.data_as(_dp), vel.ctypes.data_as(_dp),
            box.ctypes.data_as(_dp), C.byref(done)))
        return {"pos": pos, "vel": vel, "box": box, "steps_done": done.value}

    def pairs(self, sys: int = 0):
        cnt = C.c_int64(0)
        _abi.check(self.lib.dmdx_engine_get_pairs(self.h, sys, None, None, 0, C.byref(cnt)))
        pi = np.empty(cnt.value, np.int32)
        pj = np.empty(cnt.value, np.int32)
        _abi.check(self.lib.dmdx_engine_get_pairs(
            self.h, sys, pi.ctypes.data_as(C.POINTER(C.c_int32)),
            pj.ctypes.data_as(C.POINTER(C.c_int32)), cnt.value, C.byref(cnt)))
        return pi, pj
