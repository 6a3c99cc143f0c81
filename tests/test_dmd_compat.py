"""DMD's Python core API (SimulationConfig / Engine of python/dmd/core.cpp)
over the device engine: config translation on the CPU, a short run on the
GPU checked against the oracle."""
import numpy as np
import pytest

from dmd_amd import _abi
from dmd_amd import workloads as W
from dmd_amd.dmd_compat import Engine, SimulationConfig, spec_from_config


def argon_config(n_cells=4):
    s = W.argon_fcc(n_cells=n_cells, seed=3)
    c = SimulationConfig()
    c.dt = 0.001
    c.n_steps = 20
    c.box_size = s.box[0]
    c.mass = s.masses
    c.pos_x, c.pos_y, c.pos_z = s.pos[:, 0], s.pos[:, 1], s.pos[:, 2]
    c.vel_x, c.vel_y, c.vel_z = s.vel[:, 0], s.vel[:, 1], s.vel[:, 2]
    c.use_lj = True
    c.lj_cutoff = 0.85
    c.lj_sigma = [0.34]
    c.lj_epsilon = [0.996]
    return c, s


def test_config_translation():
    c, s = argon_config()
    spec = spec_from_config(c)
    assert spec.n == s.n and spec.box == (c.box_size,) * 3
    assert spec.rebuild_interval == 10 and spec.list_skin == 0.0  # DMD's list cadence
    assert spec.coulomb_type == _abi.COULOMB_CUTOFF  # default "cutoff"
    assert spec.physics == _abi.FAITHFUL and spec.precision == _abi.FP64
    c.coulomb_type = "pme"
    c.pme_grid_spacing = 0.12
    g = max(8, int(c.box_size / 0.12 + 0.5))  # simulation_engine.cpp:183-185
    assert spec_from_config(c).pme_grid == (g, g, g)
    c.coulomb_type = "none"
    assert spec_from_config(c).coulomb_type == _abi.COULOMB_NONE
    c.thermostat.type = "andersen"
    c.thermostat.frequency = 5.0
    sp = spec_from_config(c)
    assert sp.thermostat == _abi.THERMO_ANDERSEN and sp.andersen_frequency == 5.0
    c.thermostat.type = "bogus"
    with pytest.raises(ValueError):
        spec_from_config(c)
    c.thermostat.type = "nose-hoover"
    assert spec_from_config(c).thermostat == _abi.THERMO_NOSE_HOOVER
    c.constraint_type = "shake"
    c.constraint_i, c.constraint_j, c.constraint_dist = [0], [1], [0.1]
    c.constraint_tolerance = 1e-7
    sp = spec_from_config(c)
    assert [list(a) for a in sp.constraints] == [[0], [1], [0.1]]
    assert sp.constraint_tolerance == 1e-7
    c.constraint_type = "none"
    assert spec_from_config(c).constraints is None


def test_gen_vel():
    c, s = argon_config(6)
    c.gen_vel = True
    c.init_temperature = 120.0
    v = spec_from_config(c).vel
    m = s.masses
    assert np.allclose((m[:, None] * v).sum(0), 0.0, atol=1e-9)
    ke = 0.5 * (m[:, None] * v * v).sum()
    T = 2 * ke / ((3 * s.n - 3) * 0.008314462618)
    assert abs(T - 120.0) < 1e-9


@pytest.mark.gpu
def test_engine_run_matches_oracle():
    import oracle
    c, s = argon_config()
    rng = np.random.default_rng(2)
    pos = s.pos + rng.uniform(-0.01, 0.01, s.pos.shape)
    c.pos_x, c.pos_y, c.pos_z = pos[:, 0], pos[:, 1], pos[:, 2]
    e = Engine(c)
    e.run()
    assert e.current_step == 20 and abs(e.current_time - 0.02) < 1e-15
    sim = oracle.Sim(spec_from_config(c))
    sim.step(20)
    st = sim.state()
    assert np.array_equal(e.forces, st["force"])
    assert np.array_equal(e.positions, st["pos"])
    # energies are reduced in a different order than DMD's serial sum
    assert abs(e.potential_energy - st["pe"]) <= 1e-12 * abs(st["pe"])
    assert e.n_atoms == s.n and e.box_size == c.box_size
