"""H5MD trajectories from device state (libdmdx_h5md.so, include/dmdx_h5md.h).

``H5MDWriter`` mirrors DMD's writer (src/trajectory/h5md_writer.h:10-31):
``H5MDWriter(path, n_atoms, write_vel)``, ``write_frame(sys, cell)``,
``frame_count()``; the file has DMD's layout.  ``run`` is the loop of
SimulationEngine::run (src/sim/simulation_engine.cpp:88-106) over a device
engine: frames are taken with asynchronous snapshots and written while the
next steps run.  ``Clock`` is SimulationEngine's step_/time_ and
SystemData::time, which lags time_ by one step (step() sets it at its start).
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

from . import _abi

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libdmdx_h5md.so")

_dp = C.POINTER(C.c_double)
_lib = None
_lock = threading.Lock()


class Clock(C.Structure):
    """dmdx_clock: step_ / time_ of SimulationEngine and SystemData::time."""
    _fields_ = [("step", C.c_int64), ("time", C.c_double), ("sys_time", C.c_double),
                ("dt", C.c_double)]

    @classmethod
    def start(cls, dt: float, step: int = 0, time: float = 0.0):
        return cls(int(step), float(time), float(time), float(dt))


def load():
    """Load libdmdx_h5md.so (after libdmdx.so).  Raises if it was not built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        _abi.load()
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (python -m dmd_amd.build; needs HDF5)")
        lib = C.CDLL(LIB_PATH)
        lib.dmdx_h5md_open.argtypes = [C.c_char_p, C.c_int64, C.c_int32, C.POINTER(C.c_void_p)]
        lib.dmdx_h5md_write_frame.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_int64, C.c_double]
        lib.dmdx_h5md_frame_count.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
        lib.dmdx_h5md_close.argtypes = [C.c_void_p]
        lib.dmdx_h5md_last_error.restype = C.c_char_p
        lib.dmdx_h5md_run.argtypes = [C.c_void_p, C.c_int32, C.c_int64, C.c_int64,
                                      C.POINTER(Clock), C.c_void_p]
        _lib = lib
        return lib


def _check(rc: int):
    if rc != 0:
        raise _abi.DmdxError(rc, (load().dmdx_h5md_last_error() or b"").decode(errors="replace"))


def _rows(a, n):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.shape != (n, 3):
        raise ValueError(f"expected an ({n}, 3) array, got {a.shape}")
    return a


class H5MDWriter:
    """src/trajectory/h5md_writer.h:10-31."""

    def __init__(self, path: str, n_atoms: int, write_vel: bool = True):
        lib = load()
        h = C.c_void_p()
        _check(lib.dmdx_h5md_open(os.fsencode(path), int(n_atoms), int(bool(write_vel)),
                                  C.byref(h)))
        self.h = h
        self.n = int(n_atoms)
        self.write_vel = bool(write_vel)

    def write_rows(self, pos_xyz, vel_xyz, cell_matrix, step: int, time: float):
        """One frame from (n, 3) position / velocity rows and a 3x3 matrix."""
        pos = _rows(pos_xyz, self.n)
        vel = _rows(vel_xyz, self.n) if self.write_vel else None
        cell = np.ascontiguousarray(cell_matrix, dtype=np.float64).reshape(9)
        _check(load().dmdx_h5md_write_frame(
            self.h, pos.ctypes.data_as(_dp), vel.ctypes.data_as(_dp) if vel is not None else None,
            cell.ctypes.data_as(_dp), int(step), float(time)))

    def write_frame(self, sys, cell):
        """H5MDWriter::write_frame(sys, cell): sys has pos_/vel_ x, y, z, step
        and time (SystemData), cell has matrix (Cell)."""
        pos = np.stack([sys.pos_x, sys.pos_y, sys.pos_z], 1)
        vel = np.stack([sys.vel_x, sys.vel_y, sys.vel_z], 1) if self.write_vel else None
        self.write_rows(pos, vel, cell.matrix, int(sys.step), float(sys.time))

    def frame_count(self) -> int:
        v = C.c_int64()
        _check(load().dmdx_h5md_frame_count(self.h, C.byref(v)))
        return v.value

    def close(self):
        if getattr(self, "h", None):
            load().dmdx_h5md_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run(engine, nsteps: int, interval: int, clock: Clock, writer: H5MDWriter | None,
        sys: int = 0):
    """dmdx_h5md_run: nsteps steps of `engine` (dmd_amd.engine.Engine); after
    every step whose number is a multiple of interval, system sys's frame
    goes to writer with step = clock.step and time = clock.sys_time.  The
    clock advances in place."""
    if interval > 0 and writer is None:
        raise ValueError("frames requested without a writer")
    _check(load().dmdx_h5md_run(engine.h, int(sys), int(nsteps), int(interval), C.byref(clock),
                                writer.h if writer is not None else None))
    return clock


def _bench(argv=None):
    """python -m dmd_amd.trajectory [--steps N] [--side S]: ms per step of a
    MIXED water box (3 S^3 atoms) without frames, with asynchronous frames
    every 100, 10 and 1 steps, and with frames read by get_arrays and written
    synchronously every 10 steps."""
    import argparse
    import tempfile
    import time

    from . import workloads as W
    from .engine import Engine

    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--side", type=int, default=16)
    a = ap.parse_args(argv)
    spec = W.water_box(n_side=a.side, precision=_abi.MIXED)
    n_steps = a.steps
    tmp = tempfile.mkdtemp()
    e = Engine(spec)
    e.step(50)
    e.sync()

    def timed(fn):
        t0 = time.perf_counter()
        fn()
        return (time.perf_counter() - t0) * 1e3 / n_steps

    base = timed(lambda: (e.step(n_steps), e.sync()))
    print(f"water {spec.n} atoms, MIXED: no frames {base:.4f} ms/step", flush=True)
    for every in (100, 10, 1):
        clock = Clock.start(spec.dt)
        with H5MDWriter(os.path.join(tmp, f"a{every}.h5"), spec.n, True) as w:
            ms = timed(lambda: run(e, n_steps, every, clock, w))
            frames = w.frame_count()
        print(f"asynchronous frames every {every}: {ms:.4f} ms/step, {frames} frames, "
              f"+{(ms - base) * n_steps / max(frames, 1):.3f} ms/frame", flush=True)
    with H5MDWriter(os.path.join(tmp, "sync.h5"), spec.n, True) as w:
        def loop():
            for k in range(n_steps // 10):
                e.step(10)
                w.write_rows(e.positions(), e.velocities(), np.diag(e.scalars()["box"]),
                             10 * k, 0.0)
        ms = timed(loop)
    print(f"get_arrays + synchronous frames every 10: {ms:.4f} ms/step", flush=True)


if __name__ == "__main__":
    _bench()
