"""The C++ facade (include/kemhip.hpp) driven like KEMField host code (tests/cpp/facade_test.cpp,
built by __graft_entry__.build()): container, assembly, GMRES / Gauss-Jordan / Robin Hood
solves, capacitance, boundary field, error behaviour."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_facade_example():
    exe = os.path.join(ROOT, "kassiopeia_amd", "_build", "facade_test")
    assert os.path.exists(exe), "build the library first (__graft_entry__.build())"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "facade OK" in r.stdout
