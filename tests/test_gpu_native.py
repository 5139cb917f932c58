"""Runs the native C++ checks of the reference-interface runtime classes (tests/native/*.cpp, built by
build()): the delay known answer and STDP trajectory, and the two-population RunStep parity run
against the CPU oracle."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("name", ["dense_runtime_smoke", "dense_runtime_parity"])
def test_native_executable(gpu, name):
    exe = os.path.join(HERE, "native", name)
    assert os.path.exists(exe), f"{exe} not built (run __graft_entry__.build())"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and "result=PASS" in r.stdout, r.stdout + r.stderr
