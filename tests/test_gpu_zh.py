"""Zonal harmonics on the GPU (kem_zh_*) against the oracle's restatement and direct integration.

* The reference's test (UnitTest/KEMField/Fields/src/ZonalHarmonicsTest.cc:154-202): a ring at
  (7, 0, 0) of charge 1 and a conic section (7, 0, 0)-(6, 0, 1) of density 1e-8 at 1,000 points
  from glibc's rand() seeded 1337: the sums of relative deviations from direct integration
  within 1e-8, and the oracle's expansion to 1e-12 with the same central / remote / direct
  choices.
* A solved conic-section disk with rings beside it (tree of several levels, all three paths),
  and a mesh mixing triangles with rings (the triangles integrated directly at every point).
"""
import ctypes

import numpy as np
import pytest

from kassiopeia_amd import geometry as G

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def ctx():
    from kassiopeia_amd import kemfield as K

    c = K.Context(0)
    c.set_integrator("numeric")
    yield c
    c.close()


def mesh_of(shapes, recs, values=None):
    n = len(shapes)
    recs = np.asarray(recs, dtype=np.float64)
    full = np.zeros((n, 11))
    full[:, : recs.shape[1]] = recs
    return G.Mesh(np.asarray(shapes, np.int32), full, np.zeros(n, np.int32),
                  np.ones(n) if values is None else np.asarray(values, np.float64), np.zeros(n),
                  np.zeros(n, np.int32))


def glibc_points(n, seed):
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(seed)
    v = np.array([libc.rand() for _ in range(3 * n)], dtype=np.float64)
    return -4.0 + 8.0 * v.reshape(n, 3) / 2147483647.0


def device_zh(ctx, mesh, sigma, **params):
    from kassiopeia_amd import kemfield as K
    from kassiopeia_amd import zonal as Z

    cont = K.KSurfaceContainer(ctx, mesh)
    zh = Z.KZonalHarmonicContainer(cont, sigma)
    for k, v in params.items():
        setattr(zh.GetParameters(), k, v)
    zh.ComputeCoefficients()
    return cont, zh, Z.KZonalHarmonicFieldSolver(zh)


def assert_close(E, phi, Er, phir, tol):
    assert np.abs(phi - phir).max() <= tol * np.abs(phir).max()
    assert np.linalg.norm(E - Er, axis=1).max() <= tol * np.linalg.norm(Er, axis=1).max()


def test_reference_ring_and_conic(ctx, oracle):
    from oracle import zh as OZ

    mesh = mesh_of([4, 3], np.concatenate([G.ring_records(7.0, 0.0), G.conic_records(7.0, 0.0, 6.0, 1.0)]))
    sigma = np.array([1.0, 1e-8])
    cont, zh, fs = device_zh(ctx, mesh, sigma)
    pts = glibc_points(1000, 1337)
    E, phi = fs.ElectricFieldAndPotential(pts)
    ref = OZ.ZH(mesh.shapes, mesh.recs, sigma)
    Eo, phio, cnt = ref.field(pts)
    assert tuple(fs.counts) == tuple(cnt)
    assert_close(E, phi, Eo, phio, TOL)
    phd, Ed = oracle.field(mesh.shapes, mesh.recs, sigma, pts)

    def rel_sum(a, b):
        return float(np.sum((a - b) / np.where(np.abs(b) > 1e-14, b, 1.0)))

    assert abs(rel_sum(phi, phd)) < 1e-8
    for j in range(3):
        assert abs(rel_sum(E[:, j], Ed[:, j])) < 1e-8
    assert zh.info()[:3] == ref.info()
    zh.close()
    cont.close()


def test_solved_disk_with_rings(ctx, oracle):
    from kassiopeia_amd import kemfield as K
    from oracle import zh as OZ

    # the disk's charge densities from its own solve; six charged rings beside it
    disk = G.disk_mesh(64)
    cont = K.KSurfaceContainer(ctx, disk)
    A = K.KBoundaryIntegralMatrix(cont)
    sd, st, rc = A.solve(cont.rhs(), tolerance=1e-12)
    assert rc == 0
    A.close()
    cont.close()
    rings = mesh_of([4] * 6, G.ring_records(np.full(6, 1.4), np.linspace(-1.0, 1.0, 6)), values=np.zeros(6))
    mesh = G.concat(disk, rings)
    sigma = np.concatenate([sd, np.linspace(-1.0, 1.0, 6) * 1e-10])
    cont, zh, fs = device_zh(ctx, mesh, sigma)
    rng = np.random.default_rng(12)
    pts = np.concatenate([np.stack([rng.uniform(-0.5, 0.5, 4000), rng.uniform(-0.5, 0.5, 4000),
                                    rng.uniform(-1.5, 1.5, 4000)], 1),
                          G.points_shell(4000, 0.2, 6.0, 13)])
    E, phi = fs.ElectricFieldAndPotential(pts)
    cnt = fs.counts.copy()
    assert cnt[0] > 0 and cnt[1] > 0 and cnt[2] > 0
    ref = OZ.ZH(mesh.shapes, mesh.recs, sigma)
    Eo, phio, cnto = ref.field(pts)
    assert tuple(cnt) == tuple(cnto)
    assert_close(E, phi, Eo, phio, TOL)
    # against the device's direct field of the same charge densities
    Ed, phid = K.KIntegratingElectrostaticFieldSolver(cont, sigma).ElectricFieldAndPotential(pts)
    assert np.abs(phi - phid).max() <= 1e-9 * np.abs(phid).max()
    assert np.linalg.norm(E - Ed, axis=1).max() <= 1e-8 * np.linalg.norm(Ed, axis=1).max()
    # device points: the same results
    X = [K.DeviceArray(ctx, np.ascontiguousarray(pts[:, k])) for k in range(3)]
    O = [K.DeviceArray.empty(ctx, len(pts)) for _ in range(4)]
    c3 = np.zeros(3, np.int64)
    ctx.check(ctx.lib.kem_zh_field_eval_device(zh.h, len(pts), X[0].ptr, X[1].ptr, X[2].ptr, O[0].ptr, O[1].ptr,
                                               O[2].ptr, O[3].ptr, K._ptr(c3)))
    assert np.array_equal(O[0].get(), phi) and np.array_equal(O[3].get(), E[:, 2])
    assert tuple(c3) == tuple(cnt)
    for a in X + O:
        a.close()
    zh.close()
    cont.close()


def test_triangles_beside_rings(ctx, oracle):
    from oracle import zh as OZ

    cube = G.cube_mesh(3)
    rings = mesh_of([4] * 4, G.ring_records(np.full(4, 2.0), np.linspace(-0.6, 0.6, 4)))
    mesh = G.concat(cube, rings)
    sigma = np.concatenate([np.linspace(1.0, 2.0, cube.n) * 1e-10, [1e-9, -2e-9, 3e-9, -1e-9]])
    cont, zh, fs = device_zh(ctx, mesh, sigma, n_bifurcations=3)
    pts = np.concatenate([G.points_shell(500, 0.9, 1.6, 4), G.points_shell(500, 3.0, 9.0, 5)])
    E, phi = fs.ElectricFieldAndPotential(pts)
    assert fs.counts[2] >= len(pts)  # the cube's node at every point
    ref = OZ.ZH(mesh.shapes, mesh.recs, sigma, n_bifurcations=3)
    Eo, phio, cnto = ref.field(pts)
    assert tuple(fs.counts) == tuple(cnto)
    assert_close(E, phi, Eo, phio, TOL)
    zh.close()
    cont.close()


def test_rejected_inputs(ctx):
    from kassiopeia_amd import _lib
    from kassiopeia_amd import kemfield as K
    from kassiopeia_amd import zonal as Z

    mesh = mesh_of([4], G.ring_records(1.0, 0.0))
    cont = K.KSurfaceContainer(ctx, mesh)
    zh = Z.KZonalHarmonicContainer(cont, [1e-9])
    zh.GetParameters().n_central_coefficients = 1
    with pytest.raises(_lib.KemError):
        zh.ComputeCoefficients()
    grouped = G.mirror_groups(G.cube_mesh(2), [0.0, 0.0, 2.0], [0.0, 0.0, 1.0])
    gc = K.KSurfaceContainer(ctx, grouped)
    zg = Z.KZonalHarmonicContainer(gc, np.ones(gc.size()))
    with pytest.raises(_lib.KemError):
        zg.ComputeCoefficients()
    gc.close()
    cont.close()
