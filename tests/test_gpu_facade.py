"""The C++ facade (include/kemhip.hpp) driven like KEMField host code (tests/cpp/facade_test.cpp,
built by __graft_entry__.build()): container, assembly, GMRES / Gauss-Jordan / Robin Hood
solves, capacitance, boundary field, Kassiopeia's field components, error behaviour.  The
step-level root field the example evaluates over two components (a triangle cube and a
rectangle cube, 3,000 sample points in one call) is checked against the oracle's fields of
both containers summed per point, at 1e-12 of the largest value."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def read_dump(path):
    with open(path, "rb") as f:
        m = int(np.fromfile(f, np.int64, 1)[0])
        P = np.fromfile(f, np.float64, 3 * m).reshape(m, 3)
        E = np.fromfile(f, np.float64, 3 * m).reshape(m, 3)
        phi = np.fromfile(f, np.float64, m)
        meshes = []
        for _ in range(2):
            n = int(np.fromfile(f, np.int64, 1)[0])
            shapes = np.fromfile(f, np.int32, n)
            recs = np.fromfile(f, np.float64, 11 * n).reshape(n, 11)
            sigma = np.fromfile(f, np.float64, n)
            meshes.append((shapes, recs, sigma))
    return P, E, phi, meshes


def test_cpp_facade_example(tmp_path, oracle):
    exe = os.path.join(ROOT, "kassiopeia_amd", "_build", "facade_test")
    assert os.path.exists(exe), "build the library first (__graft_entry__.build())"
    dump = str(tmp_path / "root_field.bin")
    r = subprocess.run([exe, dump], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "facade OK" in r.stdout
    P, E, phi, meshes = read_dump(dump)
    phr, Er = np.zeros(len(P)), np.zeros((len(P), 3))
    for shapes, recs, sigma in meshes:
        p, e = oracle.field(shapes, recs, sigma, P)
        phr += p
        Er += e
    assert np.abs(phi - phr).max() <= 1e-12 * np.abs(phr).max()
    assert np.linalg.norm(E - Er, axis=1).max() <= 1e-12 * np.linalg.norm(Er, axis=1).max()
