"""GPU parity of the axisymmetric elements (KConicSection, KRing) with the CPU oracle, through
the C-ABI: element integrals in every branch of the reference's conic-section integrator
(20 / 100 / 500 nodes and the split towards the generating line), the collocation matrix and
solve of the reference's disk test (C = 8 eps0 R, DiskCapacitanceTest.cc:103-113), and fields
of a mesh that mixes conic sections and rings with triangles and rectangles (FP64 to 1e-12,
mixed precision to 1e-6).  The device evaluates K and E by the arithmetic-geometric mean, the
oracle by Carlson's symmetric integrals: the two agree only if both are right."""
import numpy as np
import pytest

import axisym as X
from kassiopeia_amd import geometry as G

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def ctx():
    from kassiopeia_amd import kemfield as K

    c = K.Context(0)
    yield c
    c.close()


def mesh_of(shapes, recs, bc=None, flux=None):
    n = len(shapes)
    recs = np.asarray(recs, dtype=np.float64)
    full = np.zeros((n, 11))
    full[:, : recs.shape[1]] = recs
    return G.Mesh(np.asarray(shapes, np.int32), full, np.zeros(n, np.int32) if bc is None else np.asarray(bc, np.int32),
                  np.ones(n), np.zeros(n) if flux is None else np.asarray(flux, np.float64), np.zeros(n, np.int32))


def points_around(rec, shape, rng):
    """Points in every regime of the conic integrator: far, the 100/500-node bands, on and next to
    the generating line (inside and beyond its ends), on the axis; for rings near and on the ring.
    Returns the points and whether each lies within 1e-4 lengths of the surface, where the field
    integrand cancels to a principal value (the potential stays well conditioned)."""
    pts, close = [], []
    if shape == X.RING:
        R, Z = rec[0], rec[2]
        for rr, dz in ((0.0, 0.3), (0.5 * R, 0.0), (R * 1.01, 0.0), (R, 0.02), (3 * R, -2 * R), (R * 0.999, 1e-3)):
            t = rng.uniform(0, 2 * np.pi)
            pts.append([rr * np.cos(t), rr * np.sin(t), Z + dz])
            close.append(False)
        return np.array(pts), np.array(close)
    r0, z0, r1, z1 = rec[0], rec[2], rec[3], rec[5]
    L = np.hypot(r1 - r0, z1 - z0)
    u = np.array([r1 - r0, z1 - z0]) / L
    nrm = np.array([-u[1], u[0]])
    for s, d in ((0.5, 2.0 * L), (0.3, 0.2 * L), (0.7, 0.02 * L), (0.4, 3e-3 * L), (0.6, 2e-4 * L), (0.5, 0.0),
                 (0.25, 1e-7 * L), (-0.2, 1e-5 * L), (1.3, 0.0), (0.999, 1e-9 * L)):
        rz = np.array([r0, z0]) + s * L * u + d * nrm
        t = rng.uniform(0, 2 * np.pi)
        pts.append([rz[0] * np.cos(t), rz[0] * np.sin(t), rz[1]])
        # distance in the (r, z) half-plane from the point to the generating segment
        q = np.array([abs(rz[0]), rz[1]]) - np.array([r0, z0])
        foot = np.clip(q @ u, 0.0, L)
        close.append(np.linalg.norm(q - foot * u) < 1e-4 * L)
    pts.append([0.0, 0.0, 0.5 * (z0 + z1) + 0.1])
    close.append(False)
    return np.array(pts), np.array(close)


ELEMENTS = [(X.CONIC, X.conic(0.5, 0.1, 0.8, 0.6)), (X.CONIC, X.conic(0.4, -0.2, 0.4, 0.9)),
            (X.CONIC, X.conic(1.0, 0.0, 0.3, 0.0)), (X.CONIC, X.conic(0.9, 0.4, 0.0, 0.4)),
            (X.RING, X.ring(0.7, 0.3)), (X.RING, X.ring(0.05, -1.0))]


@pytest.mark.parametrize("policy", ["numeric", "rwg", "analytic", "reference"])
def test_element_integrals(ctx, oracle, policy):
    from kassiopeia_amd import kemfield as K

    rng = np.random.default_rng(11)
    mesh = mesh_of([s for s, _ in ELEMENTS], [np.pad(r, (0, 6 - len(r))) for _, r in ELEMENTS])
    cont = K.KSurfaceContainer(ctx, mesh)
    integ = K.KElectrostaticBoundaryIntegrator(cont, policy)
    elems, pts, refs, close = [], [], [], []
    for j, (shape, rec) in enumerate(ELEMENTS):
        Ps, cl = points_around(rec, shape, rng)
        for P, c in zip(Ps, cl):
            elems.append(j)
            pts.append(P)
            close.append(c)
            refs.append(oracle.element_eval(shape, rec, P)[0])
    E, phi = integ.ElectricFieldAndPotential(elems, pts)
    refs = np.array(refs)
    fin = np.isfinite(refs).all(axis=1)
    assert fin.sum() >= len(refs) - 2
    for k in np.nonzero(fin)[0]:
        assert abs(phi[k] - refs[k, 3]) <= TOL * abs(refs[k, 3]), (k, pts[k])
        en = np.linalg.norm(refs[k, :3])
        # on the surface the radial and axial integrands cancel to 1e-7 of their magnitude, so
        # the 1e-16 differences of K and E between the two methods grow to ~1e-9 there
        tol = 1e-7 if close[k] else 1e-11
        assert np.linalg.norm(E[k] - refs[k, :3]) <= tol * en + 1e-300, (k, pts[k])
    # where the reference's radial integrand takes the square root of a negative number, so does the device
    assert np.array_equal(np.isnan(E[~fin]), np.isnan(refs[~fin, :3]))


def test_disk_capacitance(ctx, oracle):
    from kassiopeia_amd import kemfield as K

    recs = X.disk(128)
    mesh = mesh_of(np.full(128, X.CONIC), recs)
    ctx.set_integrator("numeric")
    cont = K.KSurfaceContainer(ctx, mesh)
    A = K.KBoundaryIntegralMatrix(cont, fp64=True)
    Aref = oracle.assemble(mesh.shapes, recs)
    assert np.abs(A.to_numpy() - Aref).max() <= TOL * np.abs(Aref).max()
    assert (np.abs(A.to_numpy() - Aref) / np.abs(Aref)).max() < TOL
    x, st, _ = A.solve(np.ones(128), tolerance=1e-12)
    xref = np.linalg.solve(Aref, np.ones(128))
    assert np.linalg.norm(x - xref) <= 1e-9 * np.linalg.norm(xref)
    C = float((mesh.areas() * x).sum()) / oracle.EPS0()
    assert abs(C - 8.0) < 1e-2  # the reference test's tolerance
    assert abs(C - float((mesh.areas() * xref).sum()) / oracle.EPS0()) < 1e-8


def mixed_mesh():
    cube = G.cube_mesh(4)
    # a closed annular can: bottom, outer wall, top, inner wall
    cs = G.conic_records([1.5, 2.0, 2.0, 1.5], [-1.0, -1.0, 1.0, 1.0], [2.0, 2.0, 1.5, 1.5], [-1.0, 1.0, 1.0, -1.0])
    rg = G.ring_records([1.2, 1.3, 0.4], [1.5, -1.4, 2.0])
    rect = G.cube_rect_mesh(2, edge=0.4)
    n = cube.n + len(cs) + len(rg) + rect.n
    shapes = np.concatenate([cube.shapes, np.full(len(cs), X.CONIC, np.int32), np.full(len(rg), X.RING, np.int32),
                             rect.shapes])
    recs = np.concatenate([cube.recs, cs, rg, rect.recs + np.array([0, 0, 3.0, 0, 0, 0, 0, 0, 0, 0, 0])])
    return G.Mesh(shapes, recs, np.zeros(n, np.int32), np.ones(n), np.zeros(n), np.zeros(n, np.int32))


@pytest.mark.parametrize("npts", [700, 6000])
def test_field_of_a_mixed_mesh(ctx, oracle, npts):
    from kassiopeia_amd import kemfield as K

    mesh = mixed_mesh()
    ctx.set_integrator("numeric")
    cont = K.KSurfaceContainer(ctx, mesh)
    sigma = np.random.default_rng(5).uniform(0.5, 1.5, mesh.n) * 1e-10
    rng = np.random.default_rng(npts)
    P = rng.uniform(-2.5, 2.5, size=(npts, 3))
    phi_ref, E_ref = oracle.field(mesh.shapes, mesh.recs, sigma, P, threads=8)
    ok = np.isfinite(E_ref).all(axis=1) & np.isfinite(phi_ref)
    assert ok.mean() > 0.99
    E, phi = K.KIntegratingElectrostaticFieldSolver(cont, sigma).ElectricFieldAndPotential(P)
    assert np.abs(phi[ok] - phi_ref[ok]).max() <= TOL * np.abs(phi_ref[ok]).max()
    en = np.linalg.norm(E_ref[ok], axis=1)
    assert (np.linalg.norm(E[ok] - E_ref[ok], axis=1) / en).max() < 1e-11
    Em, phim = K.KIntegratingElectrostaticFieldSolver(cont, sigma, "mixed").ElectricFieldAndPotential(P)
    assert (np.linalg.norm(Em[ok] - E[ok], axis=1) / en).max() < 1e-5
    assert np.abs(phim[ok] - phi[ok]).max() <= 1e-6 * np.abs(phi[ok]).max()


def test_conic_neumann_rows(ctx, oracle):
    from kassiopeia_amd import kemfield as K

    recs = G.conic_records([0.5, 0.6, 0.8, 0.8], [0.0, 0.2, 0.5, 0.9], [0.6, 0.8, 0.8, 0.6], [0.2, 0.5, 0.9, 1.2])
    bc = np.array([0, 1, 0, 1], np.int32)
    flux = np.array([0.0, 0.25, 0.0, 3.0])
    mesh = mesh_of(np.full(4, X.CONIC), recs[:, :6], bc, flux)
    ctx.set_integrator("numeric")
    cont = K.KSurfaceContainer(ctx, mesh)
    A = K.KBoundaryIntegralMatrix(cont, fp64=True).to_numpy()
    Aref = oracle.assemble(mesh.shapes, mesh.recs, bc, flux)
    assert (np.abs(A - Aref) / np.abs(Aref)).max() < TOL
