"""Edge cases of the C-ABI on the GPU: empty point sets, a one-element mesh, point counts on
either side of the wavefront (64), block (256) and Morton-sort (4,096) boundaries, which must
all give the same answers as one large call and as the oracle, and the rejection of degenerate
or unknown shapes."""
import ctypes

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


def test_empty_point_set(ctx):
    from kassiopeia_amd import kemfield as K

    m = G.cube_mesh(2)
    cont = K.KSurfaceContainer(ctx, m)
    sigma = np.ones(m.n) * 1e-10
    E, phi = K.KIntegratingElectrostaticFieldSolver(cont, sigma).ElectricFieldAndPotential(np.zeros((0, 3)))
    assert E.shape == (0, 3) and phi.shape == (0,)
    out = np.zeros(4)
    assert ctx.lib.kem_element_eval(ctx.h, cont.h, 0, None, None, K._ptr(out)) == 0


def test_one_element_mesh(ctx, oracle):
    from kassiopeia_amd import kemfield as K

    rec = G.records_from_points([0.0, 0.0, 0.0], [0.3, 0.0, 0.0], [0.1, 0.25, 0.0])
    m = G.Mesh(np.zeros(1, np.int32), rec, np.zeros(1, np.int32), np.ones(1), np.zeros(1), np.zeros(1, np.int32))
    cont = K.KSurfaceContainer(ctx, m)
    A = K.KBoundaryIntegralMatrix(cont)
    Aref = oracle.assemble(m.shapes, m.recs)
    assert abs(A.to_numpy()[0, 0] - Aref[0, 0]) <= 1e-12 * abs(Aref[0, 0])
    x, st, _ = A.solve(cont.rhs(), tolerance=1e-12)
    assert x[0] == pytest.approx(1.0 / Aref[0, 0], rel=1e-12)
    P = np.array([[0.1, 0.1, 0.05], [5.0, -2.0, 1.0]])
    E, phi = K.KIntegratingElectrostaticFieldSolver(cont, x).ElectricFieldAndPotential(P)
    pr, Er = oracle.field(m.shapes, m.recs, x, P)
    assert np.abs(phi - pr).max() <= 1e-12 * np.abs(pr).max()
    assert np.abs(E - Er).max() <= 1e-11 * np.abs(Er).max()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_ragged_point_counts(ctx, oracle, precision):
    from kassiopeia_amd import kemfield as K

    m = G.cube_mesh(4)
    cont = K.KSurfaceContainer(ctx, m)
    sigma = np.random.default_rng(2).uniform(0.5, 1.5, m.n) * 1e-10
    solver = K.KIntegratingElectrostaticFieldSolver(cont, sigma, precision)
    P = G.points_shell(5000, 0.05, 4.0, 31)
    E_all, phi_all = solver.ElectricFieldAndPotential(P)
    if precision == "fp64":
        pr, Er = oracle.field(m.shapes, m.recs, sigma, P[:300], threads=8)
        assert np.abs(phi_all[:300] - pr).max() <= 1e-12 * np.abs(pr).max()
    # the mixed pass groups points into wavefronts differently for each count, which changes
    # which pairs take the wavefront-uniform path: FP32 rounding level
    tol = 1e-13 if precision == "fp64" else 2e-7
    for n in (1, 2, 63, 64, 65, 255, 256, 257, 4095, 4096, 4097):
        E, phi = solver.ElectricFieldAndPotential(P[:n])
        assert np.abs(phi - phi_all[:n]).max() <= tol * np.abs(phi_all).max(), n
        assert np.abs(E - E_all[:n]).max() <= 10 * tol * np.abs(E_all).max(), n


def test_rejected_shapes(ctx):
    from kassiopeia_amd import _lib
    from kassiopeia_amd import kemfield as K

    def create(shape, rec):
        r = np.zeros((1, 11))
        r[0, : len(rec)] = rec
        h = ctypes.c_void_p()
        rc = ctx.lib.kem_mesh_create(ctx.h, 1, K._ptr(np.array([shape], np.int32)), K._ptr(r), None, None, None,
                                     ctypes.byref(h))
        if rc == 0:
            ctx.lib.kem_mesh_destroy(h)
        return rc

    assert create(G.RING, [0.0, 0.0, 1.0]) == _lib.KEM_E_GEOM  # zero radius
    assert create(G.CONICSECTION, [0.5, 0.0, 1.0, 0.5, 0.0, 1.0]) == _lib.KEM_E_GEOM  # zero length
    assert create(G.LINESEGMENT, [0, 0, 0, 1, 0, 0, 0.0]) == _lib.KEM_E_GEOM  # zero diameter
    assert create(7, [1.0, 1.0, 0, 0, 0, 1, 0, 0, 0, 1, 0]) == _lib.KEM_E_GEOM  # unknown shape type
    assert create(G.RING, [0.5, 0.0, 1.0]) == 0
