"""H5MD files of libdmdx_h5md.so on the CPU (no device calls).

The expectations are those of DMD's own writer tests
(tests/unit/trajectory/test_h5md_writer.cpp: CreateAndWriteOneFrame,
MultipleFrames, NoVelocityGroup) plus the layout of
src/trajectory/h5md_writer.cpp:20-58, 113-178 -- dataset shapes, unlimited
frame axis, chunking, element types -- that a reader of DMD's files relies
on.  Files are read back with the HDF5 C library."""
import os

import numpy as np
import pytest

import h5md_read as R
from dmd_amd import components as D

T = pytest.importorskip("dmd_amd.trajectory")

pytestmark = pytest.mark.skipif(
    not (os.path.exists(T.LIB_PATH) and R.available()),
    reason="libdmdx_h5md.so or the HDF5 library is absent (built only where HDF5 is)")

A = "/particles/atoms"


def test_create_and_write_one_frame(tmp_path):
    """CreateAndWriteOneFrame (test_h5md_writer.cpp:8-88)."""
    path = str(tmp_path / "test_one_frame.h5")
    n = 4
    s = D.SystemData(n)
    s.pos_x[:] = [1.0, 2.0, 3.0, 4.0]
    s.pos_y[:] = [5.0, 6.0, 7.0, 8.0]
    s.pos_z[:] = [9.0, 10.0, 11.0, 12.0]
    s.vel_x[:] = [0.1, 0.2, 0.3, 0.4]
    s.vel_y[:] = [0.5, 0.6, 0.7, 0.8]
    s.vel_z[:] = [0.9, 1.0, 1.1, 1.2]
    s.time, s.step = 0.0, 0
    cell = D.Cell.from_matrix([3.0, 0, 0, 0, 3.0, 0, 0, 0, 3.0], D.CellType.orthorhombic)
    with T.H5MDWriter(path, n, write_vel=True) as w:
        w.write_frame(s, cell)
        assert w.frame_count() == 1
    with R.File(path) as f:
        assert list(f.read("/h5md/version")) == [1, 1]
        pos = f.read(f"{A}/position/value")
        assert pos.shape == (1, 4, 3)
        flat = pos.reshape(-1)
        assert flat[0] == 1.0 and flat[1] == 5.0 and flat[2] == 9.0 and flat[9] == 4.0
        assert list(f.read(f"{A}/position/step")) == [0]
        assert f.exists(f"{A}/velocity/value")
        vel = f.read(f"{A}/velocity/value")
        assert np.array_equal(vel[0], np.stack([s.vel_x, s.vel_y, s.vel_z], 1))
        box = f.read(f"{A}/box/edges/value").reshape(-1)
        assert box[0] == 3.0 and box[4] == 3.0 and box[8] == 3.0


def test_multiple_frames_without_velocities(tmp_path):
    """MultipleFrames (test_h5md_writer.cpp:90-133)."""
    path = str(tmp_path / "test_multi_frame.h5")
    s = D.SystemData(2)
    cell = D.Cell(3.0, 3.0, 3.0)
    with T.H5MDWriter(path, 2, write_vel=False) as w:
        for step in range(5):
            s.pos_x[:] = [1.0 + step, 4.0]
            s.pos_y[:] = [2.0, 5.0]
            s.pos_z[:] = [3.0, 6.0]
            s.time = step * 0.002
            s.step = step
            w.write_frame(s, cell)
        assert w.frame_count() == 5
    with R.File(path) as f:
        pos = f.read(f"{A}/position/value")
        assert pos.shape == (5, 2, 3)
        assert list(pos[:, 0, 0]) == [1.0, 2.0, 3.0, 4.0, 5.0]
        assert list(f.read(f"{A}/position/step")) == [0, 1, 2, 3, 4]
        assert np.array_equal(f.read(f"{A}/position/time"), np.arange(5) * 0.002)
        assert not f.exists(f"{A}/velocity")


def test_no_velocity_group(tmp_path):
    """NoVelocityGroup (test_h5md_writer.cpp:135-151)."""
    path = str(tmp_path / "test_novel.h5")
    with T.H5MDWriter(path, 1, write_vel=False) as w:
        w.write_frame(D.SystemData(1), D.Cell(3.0, 3.0, 3.0))
    with R.File(path) as f:
        assert not f.exists(f"{A}/velocity")
        assert f.exists(f"{A}/position/value") and f.exists(f"{A}/box/edges/value")


def test_layout_matches_dmd(tmp_path):
    """Shapes, unlimited frame axis, chunks and element types of every
    dataset (h5md_writer.cpp:20-58 create the series; :132-177 the groups)."""
    path = str(tmp_path / "layout.h5")
    n = 7
    rng = np.random.default_rng(3)
    with T.H5MDWriter(path, n, write_vel=True) as w:
        for k in range(3):
            w.write_rows(rng.standard_normal((n, 3)), rng.standard_normal((n, 3)),
                         np.diag([2.0, 2.5, 3.0]) * (1 + k), step=10 * k, time=0.01 * k)
    with R.File(path) as f:
        v = f.info("/h5md/version")
        assert v["shape"] == (2,) and v["class"] == "int" and v["size"] == 4
        for g, inner in (("position", (n, 3)), ("velocity", (n, 3)), ("box/edges", (3, 3))):
            val = f.info(f"{A}/{g}/value")
            assert val["shape"] == (3, *inner) and val["maxshape"] == (-1, *inner)
            assert val["chunk"] == (1, *inner) and val["class"] == "float" and val["size"] == 8
            t = f.info(f"{A}/{g}/time")
            assert t["shape"] == (3,) and t["maxshape"] == (-1,) and t["chunk"] == (1024,)
            assert t["class"] == "float" and t["size"] == 8
            st = f.info(f"{A}/{g}/step")
            assert st["shape"] == (3,) and st["chunk"] == (1024,)
            assert st["class"] == "int" and st["size"] == 4
            assert list(f.read(f"{A}/{g}/step")) == [0, 10, 20]
        assert np.array_equal(f.read(f"{A}/box/edges/value")[2], np.diag([6.0, 7.5, 9.0]))


def test_writer_errors(tmp_path):
    with pytest.raises(Exception):
        T.H5MDWriter(str(tmp_path / "no_such_dir" / "x.h5"), 3)
    with T.H5MDWriter(str(tmp_path / "e.h5"), 3) as w:
        with pytest.raises(ValueError):
            w.write_rows(np.zeros((2, 3)), np.zeros((3, 3)), np.eye(3), 0, 0.0)
