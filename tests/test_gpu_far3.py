"""The mixed far pass's 3-point rule: a cluster of elements whose every element lies beyond
80 sizes of every point of a wavefront is evaluated with the degree-2 3-point rule (interior
nodes, equal weights).  Points far from a 5,120-triangle sphere take that path, points near it
the 6-point and per-pair paths, points in between a mixture; every group stays within the
mixed-precision bar (1e-6) of the FP64 path, point by point where the field has no null, and a
pass over several charge-density sets agrees with each set's own pass."""
import numpy as np
import pytest

from kassiopeia_amd import geometry as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from kassiopeia_amd import kemfield as K

    c = K.Context(0)
    c.set_integrator("numeric")
    yield c
    c.close()


@pytest.fixture(scope="module")
def sphere(ctx):
    from kassiopeia_amd import kemfield as K

    m = G.sphere_mesh(4)  # 5,120 triangles, sizes about 0.06: 80 sizes about 5 radii
    cont = K.KSurfaceContainer(ctx, m)
    A = K.KBoundaryIntegralMatrix(cont)
    x, st, rc = A.solve(cont.rhs(), tolerance=1e-10)
    assert rc == 0
    return m, cont, x


@pytest.mark.parametrize("shell", [(1.5, 4.0), (5.0, 9.0), (10.0, 40.0)])
def test_far3_against_fp64(sphere, shell):
    from kassiopeia_amd import kemfield as K

    m, cont, x = sphere
    pts = G.points_shell(20000, shell[0], shell[1], seed=41)
    E64, p64 = K.KIntegratingElectrostaticFieldSolver(cont, x, "fp64").ElectricFieldAndPotential(pts)
    Em, pm = K.KIntegratingElectrostaticFieldSolver(cont, x, "mixed").ElectricFieldAndPotential(pts)
    # a charged sphere: no field null outside it, so the bar holds point by point
    rel = np.linalg.norm(Em - E64, axis=1) / np.linalg.norm(E64, axis=1)
    assert rel.max() < 1e-6, rel.max()
    assert (np.abs(pm - p64) / np.abs(p64)).max() < 1e-6
    # outside a uniformly charged sphere at potential 1: phi = 1/r, within the mesh's accuracy
    r = np.linalg.norm(pts, axis=1)
    assert np.abs(pm * r - 1.0).max() < 5e-3
    # the paths the pass took: the far shell runs the 3-point branch, the near shell never
    c = paths(cont, pts)
    assert c.sum() - c[9] == len(pts) * m.n
    if shell == (10.0, 40.0):
        assert c[0] > 0, c
    if shell == (1.5, 4.0):
        assert c[0] == 0 and c[1] > 0, c


def paths(cont, pts):
    import ctypes

    from kassiopeia_amd import kemfield as K

    ctx = cont.ctx
    X = [K.DeviceArray(ctx, np.ascontiguousarray(pts[:, k])) for k in range(3)]
    out = (ctypes.c_int64 * 10)()
    try:
        ctx.check(ctx.lib.kem_field_census_paths(ctx.h, cont.h, len(pts), X[0].ptr, X[1].ptr, X[2].ptr, out))
    finally:
        for a in X:
            a.close()
    return np.array(out[:], dtype=np.int64)


def test_far3_sets(ctx, sphere):
    from kassiopeia_amd import kemfield as K

    m, cont, x = sphere
    rng = np.random.default_rng(5)
    S = np.stack([x, -2.0 * x, x * rng.uniform(0.5, 1.5, size=x.size)], 1)
    pts = G.points_shell(30000, 8.0, 40.0, seed=43)
    sid = rng.integers(0, 3, size=len(pts)).astype(np.int32)
    Es, ps = K.field_sets(cont, S, sid, pts, "mixed")
    for k in range(3):
        sel = sid == k
        Ek, pk = K.KIntegratingElectrostaticFieldSolver(cont, S[:, k], "mixed").ElectricFieldAndPotential(pts[sel])
        assert np.linalg.norm(Es[sel] - Ek) / np.linalg.norm(Ek) < 1e-7, k
        E64, _ = K.KIntegratingElectrostaticFieldSolver(cont, S[:, k], "fp64").ElectricFieldAndPotential(pts[sel])
        assert (np.linalg.norm(Ek - E64, axis=1) / np.linalg.norm(E64, axis=1)).max() < 1e-6, k
