"""The C++ drop-in class (edgepredict::MPMWorkpieceSolver) driven through the
reference's IPhysicsSolver / IMetricsProvider / IThermalCoupling interfaces
by tests/cpp/solver_api_test.cpp, built by __graft_entry__.build()."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "edegpredict-engine-rotary_amd", "lib", "solver_api_test")


@pytest.mark.gpu
def test_cpp_drop_in_class_through_reference_interfaces():
    assert os.path.exists(BIN), "build() first"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout


def test_cpp_driver_is_built_and_linked():
    assert os.path.exists(BIN)
    out = subprocess.run(["ldd", BIN], capture_output=True, text=True).stdout
    assert "libmpmw_hip.so" in out and "not found" not in out.split("libmpmw_hip.so")[1].split("\n")[0]
