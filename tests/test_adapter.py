"""The C++ adapters of include/dmdx_dmd.hpp, run inside DMD's own classes.

oracle/build_ref.sh compiles tests/adapter_check.cpp against DMD's headers
and translation units (src/force/*.cpp, src/integrate/integrator.cpp,
src/sim/checkpoint.cpp) and links libdmdx.so, so the GPU components are
driven by DMD's ForceEngine and Integrator exactly as SimulationEngine::step
drives its own (src/sim/simulation_engine.cpp:54-86).  The binary exists only
where the reference tree was present at build time; the tests skip without it.
"""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "adapter_check")

needs_exe = pytest.mark.skipif(not os.path.exists(EXE),
                               reason="oracle/_ref/adapter_check not built (no reference tree)")


def _run(*args, timeout):
    r = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, f"adapter_check failed ({r.returncode}):\n{r.stdout}\n{r.stderr}"
    return json.loads(r.stdout.strip().splitlines()[-1])


@needs_exe
def test_adapter_interface_cpu():
    """GpuLennardJones is found by dynamic_cast<LennardJones*> (the lookup of
    simulation_engine.cpp:108-137), names match, and C-ABI errors surface as
    ForceException with the matching ForceError (src/core/error.h:6-21)."""
    out = _run("--no-gpu", timeout=60)
    assert out["ok"]


@needs_exe
@pytest.mark.gpu
def test_adapter_in_dmd_loop_gpu():
    """30 velocity-Verlet steps of a 200-atom chain system: GPU components in
    DMD's loop (periodic and under a triclinic Cell) and DeviceSimulationEngine
    both match DMD's CPU components to 1e-10 relative; a restart from DMD's
    checkpoint format is bit-exact; run() writes H5MD frames that match DMD's
    loop at the frame steps and checkpoints on cadence."""
    out = _run(timeout=110)
    print(out)
    assert out["ok"]
    assert out["components_force_rel"] <= 1e-10
    assert out["device_force_rel"] <= 1e-10
    assert out["restart_rel"] == 0.0
