"""Field evaluation captured as a HIP graph on fixed device buffers and replayed after the
points change in place (kem_field_graph_create / kem_graph_launch): every replay equals an
eager evaluation of the same points, with eager work and synchronisations between replays."""
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


@pytest.mark.parametrize("npts,precision", [(1000, 0), (9000, 0), (9000, 1)])
def test_field_graph_replays(ctx, npts, precision):
    from kassiopeia_amd import kemfield as K

    m = G.cube_mesh(6)
    cont = K.KSurfaceContainer(ctx, m)
    sigma = np.linalg.solve(K.KBoundaryIntegralMatrix(cont).to_numpy(), cont.rhs())
    S = K.DeviceArray(ctx, sigma)
    X = [K.DeviceArray.empty(ctx, npts) for _ in range(3)]
    O = [K.DeviceArray.empty(ctx, npts) for _ in range(4)]
    h = ctypes.c_void_p()
    ctx.check(ctx.lib.kem_field_graph_create(ctx.h, cont.h, S.ptr, npts, X[0].ptr, X[1].ptr, X[2].ptr, precision,
                                             O[0].ptr, O[1].ptr, O[2].ptr, O[3].ptr, ctypes.byref(h)))
    try:
        for step in range(4):
            pts = G.points_shell(npts, 0.9, 2.0, 100 + step)
            for k in range(3):
                X[k].upload(pts[:, k])
            ctx.check(ctx.lib.kem_graph_launch(h))
            ctx.check(ctx.lib.kem_ctx_sync(ctx.h))
            phi_g = O[0].get()
            E_g = np.stack([O[k].get() for k in (1, 2, 3)], 1)
            E, phi = K.KIntegratingElectrostaticFieldSolver(cont, sigma, ["fp64", "mixed"][precision]) \
                .ElectricFieldAndPotential(pts)
            assert np.array_equal(phi_g, phi) or np.abs(phi_g - phi).max() <= 1e-15 * np.abs(phi).max()
            assert np.abs(E_g - E).max() <= 1e-14 * np.abs(E).max()
    finally:
        ctx.lib.kem_graph_destroy(h)
        for a in X + O + [S]:
            a.close()


def test_field_graph_survives_context_growth(ctx):
    """The graph owns its workspace: eager calls between replays that grow the context's
    buffers (a larger point set through the per-XCD ranges, a solve's matrix-vector partials,
    several charge-density sets) leave the replays' results unchanged."""
    from kassiopeia_amd import kemfield as K

    npts = 5000
    m = G.cube_mesh(6)
    cont = K.KSurfaceContainer(ctx, m)
    Am = K.KBoundaryIntegralMatrix(cont)
    sigma = np.linalg.solve(Am.to_numpy(), cont.rhs())
    S = K.DeviceArray(ctx, sigma)
    X = [K.DeviceArray.empty(ctx, npts) for _ in range(3)]
    O = [K.DeviceArray.empty(ctx, npts) for _ in range(4)]
    h = ctypes.c_void_p()
    ctx.check(ctx.lib.kem_field_graph_create(ctx.h, cont.h, S.ptr, npts, X[0].ptr, X[1].ptr, X[2].ptr, 1,
                                             O[0].ptr, O[1].ptr, O[2].ptr, O[3].ptr, ctypes.byref(h)))
    solver = K.KIntegratingElectrostaticFieldSolver(cont, sigma, "mixed")
    try:
        for step in range(3):
            pts = G.points_shell(npts, 0.9, 2.0, 300 + step)
            for k in range(3):
                X[k].upload(pts[:, k])
            ctx.check(ctx.lib.kem_graph_launch(h))
            ctx.check(ctx.lib.kem_ctx_sync(ctx.h))
            phi_g = O[0].get()
            E_g = np.stack([O[k].get() for k in (1, 2, 3)], 1)
            E, phi = solver.ElectricFieldAndPotential(pts)
            assert np.abs(phi_g - phi).max() <= 1e-15 * np.abs(phi).max(), step
            assert np.abs(E_g - E).max() <= 1e-14 * np.abs(E).max(), step
            # eager work that grows the context's own workspaces
            big = G.points_shell(300_000 * (step + 1), 0.9, 2.0, 400 + step)
            solver.ElectricFieldAndPotential(big)
            Am.solve(cont.rhs(), tolerance=1e-8)
            sets = np.stack([sigma, 2 * sigma], 1)
            K.field_sets(cont, sets, np.arange(70_000) % 2, G.points_shell(70_000, 0.9, 2.0, 500 + step), "fp64")
    finally:
        ctx.lib.kem_graph_destroy(h)
        for a in X + O + [S]:
            a.close()
