"""Symmetry groups on the GPU through the C-ABI (kem_mesh_create_groups): the grouped
collocation matrix (Dirichlet and Neumann rows) against the oracle's to 1e-12, the grouped
solve against the full mirrored solve, fields (FP64 to 1e-12, mixed to 1e-6) against the
oracle's group sums, element integrals of groups, and the derivative entry points' refusal."""
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
def mirrored():
    cube = G.cube_mesh(6)
    half = cube.subset(np.nonzero(cube.centroids()[:, 2] > 1e-9)[0])
    grouped = G.mirror_groups(half, [0.0, 0.0, 0.0], [0.0, 0.0, 1.0])
    full = G.Mesh(grouped.shapes, grouped.recs, np.zeros(2 * half.n, np.int32), np.ones(2 * half.n),
                  np.zeros(2 * half.n), np.zeros(2 * half.n, np.int32))
    return grouped, full


def test_grouped_matrix_and_solve(ctx, oracle, mirrored):
    from kassiopeia_amd import kemfield as K

    grouped, full = mirrored
    cont = K.KSurfaceContainer(ctx, grouped)
    assert cont.size() == grouped.n
    A = K.KBoundaryIntegralMatrix(cont, fp64=True, fp32=True)
    Aref = oracle.assemble_groups(grouped.group_start, grouped.shapes, grouped.recs)
    assert (np.abs(A.to_numpy() - Aref) / np.abs(Aref)).max() < 1e-12
    xg, st, _ = A.solve(cont.rhs(), tolerance=1e-12)
    fcont = K.KSurfaceContainer(ctx, full)
    Af = K.KBoundaryIntegralMatrix(fcont, fp64=True)
    xf, _, _ = Af.solve(fcont.rhs(), tolerance=1e-12)
    assert np.abs(xf[0::2] - xg).max() <= 1e-9 * np.abs(xf).max()
    assert np.abs(xf[1::2] - xg).max() <= 1e-9 * np.abs(xf).max()
    xm, stm, _ = A.solve(cont.rhs(), tolerance=1e-10, precision="mixed")
    assert np.linalg.norm(xm - xg) <= 1e-8 * np.linalg.norm(xg)


def test_grouped_neumann_rows(ctx, oracle, mirrored):
    from kassiopeia_amd import kemfield as K

    grouped, _ = mirrored
    bc = np.zeros(grouped.n, np.int32)
    bc[::4] = 1
    flux = np.where(bc == 1, 2.5, 0.0)
    g2 = G.Mesh(grouped.shapes, grouped.recs, bc, grouped.bc_value, flux, grouped.group, grouped.group_start)
    A = K.KBoundaryIntegralMatrix(K.KSurfaceContainer(ctx, g2), fp64=True).to_numpy()
    Aref = oracle.assemble_groups(g2.group_start, g2.shapes, g2.recs, bc, flux)
    err = np.abs(A - Aref) / np.maximum(np.abs(Aref), 1e-300)
    assert err[np.abs(Aref) > 1e-6 * np.abs(Aref).max()].max() < 1e-11


@pytest.mark.parametrize("npts", [500, 9000])
def test_grouped_field(ctx, oracle, mirrored, npts):
    from kassiopeia_amd import kemfield as K

    grouped, full = mirrored
    cont = K.KSurfaceContainer(ctx, grouped)
    x = np.random.default_rng(4).uniform(0.5, 1.5, grouped.n) * 1e-11
    P = G.points_shell(npts, 0.3, 3.0, npts)
    E, phi = K.KIntegratingElectrostaticFieldSolver(cont, x).ElectricFieldAndPotential(P)
    ref_p, ref_E = oracle.field_groups(grouped.group_start, grouped.shapes, grouped.recs, x, P[:400])
    assert np.abs(phi[:400] - ref_p).max() <= 1e-12 * np.abs(ref_p).max()
    assert (np.linalg.norm(E[:400] - ref_E, axis=1) / np.linalg.norm(ref_E, axis=1)).max() < 1e-11
    # the full mirrored mesh with each member carrying its group's density gives the same field
    Ef, pf = K.KIntegratingElectrostaticFieldSolver(K.KSurfaceContainer(ctx, full),
                                                      np.repeat(x, 2)).ElectricFieldAndPotential(P)
    assert np.abs(phi - pf).max() <= 1e-13 * np.abs(pf).max()
    Em, pm = K.KIntegratingElectrostaticFieldSolver(cont, x, "mixed").ElectricFieldAndPotential(P)
    assert np.abs(pm - phi).max() <= 1e-6 * np.abs(phi).max()


def test_grouped_element_integrals_and_refusals(ctx, oracle, mirrored):
    from kassiopeia_amd import _lib
    from kassiopeia_amd import kemfield as K

    grouped, full = mirrored
    cont = K.KSurfaceContainer(ctx, grouped)
    integ = K.KElectrostaticBoundaryIntegrator(cont, "numeric")
    P = np.array([[0.2, 0.1, 0.9], [1.5, -0.3, 0.0], [0.0, 0.0, 2.0]])
    E, phi = integ.ElectricFieldAndPotential([0, 5, grouped.n - 1], P)
    for k, g in enumerate([0, 5, grouped.n - 1]):
        ref = sum(oracle.element_eval(full.shapes[m], full.recs[m], P[k])[0] for m in (2 * g, 2 * g + 1))
        assert abs(phi[k] - ref[3]) <= 1e-12 * abs(ref[3])
        assert np.linalg.norm(E[k] - ref[:3]) <= 1e-11 * np.linalg.norm(ref[:3])
    import ctypes
    dsig, dv = np.zeros(grouped.n), np.zeros((2 * grouped.n, 9))
    g = np.ones(3)
    rc = ctx.lib.kem_field_vjp(ctx.h, cont.h, K._ptr(np.ones(grouped.n)), 3, K._ptr(np.ascontiguousarray(P)),
                               K._ptr(g), K._ptr(dsig), K._ptr(dv))
    assert rc == _lib.KEM_E_GEOM
    bad = np.array([0, 2, 2], np.int64)
    h = ctypes.c_void_p()
    rc = ctx.lib.kem_mesh_create_groups(ctx.h, 2, K._ptr(bad), K._ptr(np.zeros(2, np.int32)),
                                        K._ptr(np.ascontiguousarray(grouped.recs[:2])), None, None, None,
                                        ctypes.byref(h))
    assert rc == _lib.KEM_E_ARG
