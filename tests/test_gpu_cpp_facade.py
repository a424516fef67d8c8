"""Runs the C++ facade test (tests/cpp/test_facade.cpp, built by
__graft_entry__.build()): the reference-shaped C++ classes over the C ABI,
checked against the oracle inside the program."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_cpp_facade_matches_oracle():
    exe = os.path.join(HERE, "cpp", "test_facade")
    assert os.path.exists(exe), "build it with __graft_entry__.build()"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASSED" in r.stdout
