"""DMD's binary checkpoint file (src/sim/checkpoint.cpp:9-75), from device state.

Layout, little-endian as written by the reference on x86-64: magic u32
0x444D4450, version u32 1, n u64, step u64 (size_t), time f64, potential
energy f64, then n f64 each of pos_x, pos_y, pos_z, vel_x, vel_y, vel_z,
forces_x, forces_y, forces_z, masses, charges, n i32 atom types, and the
LennardJones step_since_rebuild counter as i32.  The device engine restores
positions, velocities and the list counter from it
(SimulationEngine::load_checkpoint, simulation_engine.cpp:127-137).
"""
from __future__ import annotations

import struct
from dataclasses import dataclass

import numpy as np

MAGIC = 0x444D4450
VERSION = 1
_HEAD = struct.Struct("<IIQQdd")
_VECTORS = ("pos_x", "pos_y", "pos_z", "vel_x", "vel_y", "vel_z", "forces_x", "forces_y",
            "forces_z", "masses", "charges")


@dataclass
class Checkpoint:
    step: int
    time: float
    potential_energy: float
    arrays: dict            # the eleven FP64 vectors by SystemData name, plus atom_types
    lj_step_since_rebuild: int

    @property
    def n(self) -> int:
        return int(len(self.arrays["pos_x"]))


def save(path: str, ck: Checkpoint) -> None:
    """checkpoint::save (checkpoint.cpp:9-38)."""
    n = ck.n
    with open(path, "wb") as f:
        f.write(_HEAD.pack(MAGIC, VERSION, n, int(ck.step), float(ck.time),
                           float(ck.potential_energy)))
        for name in _VECTORS:
            a = np.ascontiguousarray(ck.arrays[name], dtype="<f8")
            if a.shape != (n,):
                raise ValueError(f"{name} has shape {a.shape}, expected ({n},)")
            f.write(a.tobytes())
        f.write(np.ascontiguousarray(ck.arrays["atom_types"], dtype="<i4").tobytes())
        f.write(struct.pack("<i", int(ck.lj_step_since_rebuild)))


def load(path: str) -> Checkpoint:
    """checkpoint::load (checkpoint.cpp:40-75); the reference does not check
    the magic or version, this reader does."""
    with open(path, "rb") as f:
        data = f.read()
    magic, version, n, step, time, pe = _HEAD.unpack_from(data, 0)
    if magic != MAGIC or version != VERSION:
        raise ValueError(f"{path}: not a DMD checkpoint (magic {magic:#x}, version {version})")
    off = _HEAD.size
    need = off + n * (8 * len(_VECTORS) + 4) + 4
    if len(data) != need:
        raise ValueError(f"{path}: {len(data)} bytes, expected {need} for {n} atoms")
    arrays = {}
    for name in _VECTORS:
        arrays[name] = np.frombuffer(data, "<f8", n, off).astype(np.float64)
        off += 8 * n
    arrays["atom_types"] = np.frombuffer(data, "<i4", n, off).astype(np.int32)
    off += 4 * n
    (counter,) = struct.unpack_from("<i", data, off)
    return Checkpoint(int(step), float(time), float(pe), arrays, int(counter))
