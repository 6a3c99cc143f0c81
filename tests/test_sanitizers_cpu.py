"""Host code under AddressSanitizer + UndefinedBehaviorSanitizer.

oracle/_ref/sanitize_check is the oracle (oracle/dmd_oracle.c) built with
-fsanitize=address,undefined and driven through every entry point on small
systems (oracle/Makefile, target sanitize); oracle/_ref/adapter_check_san is
tests/adapter_check.cpp -- DMD's own translation units with the adapters of
include/dmdx_dmd.hpp -- built the same way (oracle/build_ref.sh), run with
--no-gpu.  Any out-of-bounds access, use after free, leak or undefined
behaviour ends the program with a non-zero status."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")


def run(exe, *args):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, f"{exe} failed ({r.returncode}):\n{r.stdout}\n{r.stderr[-4000:]}"
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
    return r.stdout


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "sanitize_check")),
                    reason="oracle/_ref/sanitize_check not built (make -C oracle sanitize)")
def test_oracle_under_asan_ubsan():
    assert '"ok": true' in run(os.path.join(REF, "sanitize_check"))


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "adapter_check_san")),
                    reason="adapter_check_san not built (needs the reference tree)")
def test_adapter_check_no_gpu_under_asan_ubsan():
    assert '"ok": true' in run(os.path.join(REF, "adapter_check_san"), "--no-gpu")
