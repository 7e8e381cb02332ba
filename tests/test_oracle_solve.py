"""Pins of the CPU oracle's collocation assembly, solvers and field sums (no GPU).

Known answers: unit-cube capacitance 0.6606785 m (UnitTest/KEMField/Fields/src/
CubeCapacitanceTest.cc:115-116, tolerance 1e-2 at :106), unit-sphere capacitance 1 m
(SphereCapacitanceTest.cc:121, tolerance 5e-2 at :112), and the iteration count / final
residual of the reference's GMRES on the 1,728-triangle cube measured with the compiled
reference (28 iterations, 6.6e-11).  Solvers are checked against LAPACK; the field sum
against the boundary values it must reproduce and against the monopole far field.
"""
import numpy as np
import pytest

from kassiopeia_amd import geometry as G


@pytest.fixture(scope="module")
def cube12(oracle):
    m = G.cube_mesh(12)
    A = oracle.assemble(m.shapes, m.recs)
    return m, A


@pytest.fixture(scope="module")
def cube6(oracle):
    m = G.cube_mesh(6)
    A = oracle.assemble(m.shapes, m.recs)
    return m, A


def test_cube_gmres_and_capacitance(oracle, cube12):
    m, A = cube12
    b = oracle.rhs(m.bc_value)
    x, it, res, hist = oracle.gmres(A, b, tol=1e-10)
    # compiled reference on this mesh: 28 iterations, final relative residual 6.6e-11
    assert it == 28
    assert 6.55e-11 < res < 6.7e-11
    assert np.all(np.diff(hist) <= 0)  # GMRES residual is monotone
    # true residual agrees with the Givens estimate
    assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) < 2e-10
    C = (m.areas() * x).sum() * oracle.K()
    assert abs(C - 0.6606785) < 1e-2
    # the 1,728-triangle discretisation sits 0.27 % below the converged value
    assert -0.004 < (C - 0.6606785) / 0.6606785 < -0.002


def test_cube_capacitance_converges(oracle):
    """Capacitance error shrinks with refinement towards the accepted 0.6606785 m."""
    errs = []
    for n in (4, 8):
        m = G.cube_mesh(n)
        A = oracle.assemble(m.shapes, m.recs)
        x = np.linalg.solve(A, oracle.rhs(m.bc_value))
        errs.append(abs((m.areas() * x).sum() * oracle.K() - 0.6606785))
    assert errs[1] < 0.75 * errs[0]


def test_sphere_capacitance(oracle):
    m = G.sphere_mesh(3)
    A = oracle.assemble(m.shapes, m.recs)
    x, it, res, _ = oracle.gmres(A, oracle.rhs(m.bc_value), tol=1e-10)
    C = (m.areas() * x).sum() * oracle.K()
    assert abs(C - 1.0) < 5e-2  # SphereCapacitanceTest.cc:112
    assert abs(C - 1.0) < 1e-2  # inscribed polyhedron: a little below 1


def test_gmres_matches_lapack(oracle, cube6):
    m, A = cube6
    b = oracle.rhs(m.bc_value)
    x, it, res, _ = oracle.gmres(A, b, tol=1e-12)
    xs = np.linalg.solve(A, b)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-10


def test_gmres_restarted(oracle, cube6):
    m, A = cube6
    rng = np.random.default_rng(2)
    b = rng.normal(size=m.n)
    x, it, res, _ = oracle.gmres(A, b, tol=1e-11, restart=5, maxit=2000)
    xs = np.linalg.solve(A, b)
    assert res <= 1e-11
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-8


def test_gmres_stops_at_maxit(oracle, cube6):
    m, A = cube6
    b = oracle.rhs(m.bc_value)
    x, it, res, _ = oracle.gmres(A, b, tol=1e-14, maxit=3)
    assert it == 3 and res > 1e-14


def test_robin_hood_matches_lapack(oracle, cube6):
    m, A = cube6
    b = oracle.rhs(m.bc_value)
    x, sub, res = oracle.robin_hood(A, b, tol=1e-10)
    assert res <= 1e-10
    assert sub % m.n == 0  # the norm is checked every N sub-iterations (KRobinHood.hh:76-80)
    xs = np.linalg.solve(A, b)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-8


def test_gaussian_elimination(oracle, cube6):
    m, A = cube6
    b = oracle.rhs(m.bc_value)
    assert np.allclose(oracle.gauss(A, b), np.linalg.solve(A, b), rtol=1e-11, atol=0)


def test_matrix_element_matches_assembly(oracle, cube6):
    m, A = cube6
    rng = np.random.default_rng(4)
    for i, j in rng.integers(0, m.n, size=(20, 2)):
        assert oracle.matrix_element(int(i), int(j), m.shapes, m.recs) == A[i, j]
    # diagonal: potential of a triangle at its own centroid, positive
    assert np.all(np.diag(A) > 0) and np.all(np.diag(A) > A.max(axis=1) * 0.999)


def test_neumann_rows(oracle):
    """Neumann target: E_j(c_i).n_i off the diagonal, (1+f)/(1-f)/(2 eps0) on it
    (KElectrostaticBoundaryIntegrator.hh:264-279, .cc:56-62); b_i = 0."""
    m = G.cube_mesh(2)
    bc = np.zeros(m.n, np.int32)
    bc[3] = 1
    flux = np.zeros(m.n)
    flux[3] = 0.6
    A = oracle.assemble(m.shapes, m.recs, bc=bc, flux=flux)
    assert A[3, 3] == ((1 + 0.6) / (1 - 0.6)) / (2 * oracle.EPS0())
    c3 = m.centroids()[3]
    n3 = oracle.normal(0, m.recs[3])
    out, _ = oracle.element_eval(0, m.recs[7], c3)
    assert A[3, 7] == out[0] * n3[0] + out[1] * n3[1] + out[2] * n3[2]
    b = oracle.rhs(np.ones(m.n), bc=bc)
    assert b[3] == 0.0 and b[0] == 1.0


def test_field_reproduces_boundary_values(oracle, cube6):
    m, A = cube6
    b = oracle.rhs(m.bc_value)
    x = np.linalg.solve(A, b)
    phi, E = oracle.field(m.shapes, m.recs, x, m.centroids()[:40])
    assert np.abs(phi - 1.0).max() < 1e-10


def test_field_far_monopole_and_linearity(oracle, cube6):
    m, A = cube6
    x = np.linalg.solve(A, oracle.rhs(m.bc_value))
    Q = (m.areas() * x).sum()
    P = np.array([[1e4, 0.0, 0.0], [0.0, -2e4, 3e3]])
    phi, E = oracle.field(m.shapes, m.recs, x, P)
    r = np.linalg.norm(P, axis=1)
    assert np.allclose(phi, Q * oracle.K() / r, rtol=1e-7)
    rng = np.random.default_rng(8)
    s1, s2 = rng.normal(size=m.n), rng.normal(size=m.n)
    pts = G.points_shell(5, 0.9, 2.0, 3)
    p1, e1 = oracle.field(m.shapes, m.recs, s1, pts)
    p2, e2 = oracle.field(m.shapes, m.recs, s2, pts)
    p3, e3 = oracle.field(m.shapes, m.recs, 2 * s1 - 3 * s2, pts)
    assert np.allclose(p3, 2 * p1 - 3 * p2, rtol=1e-12, atol=1e-12 * np.abs(p3).max())
    assert np.allclose(e3, 2 * e1 - 3 * e2, rtol=1e-12, atol=1e-12 * np.abs(e3).max())


def test_field_divergence_free(oracle, cube6):
    m, A = cube6
    x = np.linalg.solve(A, oracle.rhs(m.bc_value))
    P = np.array([0.7, -0.4, 0.9])
    h = 1e-5
    div = 0.0
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        _, Ep = oracle.field(m.shapes, m.recs, x, (P + e)[None])
        _, Em = oracle.field(m.shapes, m.recs, x, (P - e)[None])
        div += (Ep[0, k] - Em[0, k]) / (2 * h)
    _, E0 = oracle.field(m.shapes, m.recs, x, P[None])
    assert abs(div) < 1e-5 * np.linalg.norm(E0) / 0.1


def test_field_threads_equal_serial(oracle):
    """ko_field_mt shares points over OpenMP threads; every point's Kahan sum is unchanged."""
    from kassiopeia_amd import geometry as G

    m = G.cube_mesh(3)
    s = np.random.default_rng(4).normal(size=m.n)
    P = G.points_shell(300, 0.9, 2.0, 3)
    a = oracle.field(m.shapes, m.recs, s, P)
    b = oracle.field(m.shapes, m.recs, s, P, threads=4)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
