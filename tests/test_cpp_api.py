"""Runs the C++ restatement of the reference's unit tests
(tests/cpp/test_reference_api.cpp, built against include/raptr/*.hpp)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "raptr_amd", "lib", "test_reference_api")


def _run(*args):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} not built (run __graft_entry__.build())")
    r = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_cpp_mirror_host_geometry():
    assert "0 failure(s)" in _run("--cpu-only")


@pytest.mark.gpu
def test_cpp_mirror_reference_unit_tests():
    assert "0 failure(s)" in _run()
