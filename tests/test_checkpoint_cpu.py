"""DMD's binary checkpoint format (src/sim/checkpoint.cpp:9-75).

tests/golden/dmd_checkpoint_n5.bin was written by DMD's own checkpoint::save
(tests/golden/make_golden.py); dmd_amd.checkpoint must write the same bytes
from the same inputs and read that file back to them."""
import os

import numpy as np
import pytest

from dmd_amd import checkpoint as CK

HERE = os.path.dirname(__file__)
G = np.load(os.path.join(HERE, "golden", "dmd_ref_vectors.npz"))
GOLD = os.path.join(HERE, "golden", "dmd_checkpoint_n5.bin")
NAMES = ("pos_x", "pos_y", "pos_z", "vel_x", "vel_y", "vel_z", "forces_x", "forces_y",
         "forces_z", "masses", "charges")


def golden_checkpoint():
    v, types, sc = G["ckpt_v"], G["ckpt_types"], G["ckpt_scalars"]
    n = len(types)
    arrays = {name: v[k * n:(k + 1) * n] for k, name in enumerate(NAMES)}
    arrays["atom_types"] = types
    return CK.Checkpoint(int(sc[0]), float(sc[1]), float(sc[2]), arrays, int(sc[3]))


def test_save_writes_dmd_bytes(tmp_path):
    p = str(tmp_path / "c.bin")
    CK.save(p, golden_checkpoint())
    with open(p, "rb") as a, open(GOLD, "rb") as b:
        assert a.read() == b.read()


def test_load_reads_dmd_file():
    ck, ref = CK.load(GOLD), golden_checkpoint()
    assert (ck.step, ck.time, ck.potential_energy, ck.lj_step_since_rebuild) == \
        (ref.step, ref.time, ref.potential_energy, ref.lj_step_since_rebuild)
    for name in NAMES + ("atom_types",):
        assert np.array_equal(ck.arrays[name], ref.arrays[name])
    assert np.signbit(ck.arrays["pos_x"][3])  # -0.0 survives


def test_load_rejects_other_files(tmp_path):
    p = str(tmp_path / "bad.bin")
    with open(GOLD, "rb") as f:
        data = bytearray(f.read())
    data[0] ^= 1
    open(p, "wb").write(bytes(data))
    with pytest.raises(ValueError):
        CK.load(p)
    data[0] ^= 1
    open(p, "wb").write(bytes(data[:-1]))
    with pytest.raises(ValueError):
        CK.load(p)
