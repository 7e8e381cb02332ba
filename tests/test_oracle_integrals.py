"""Pins of the CPU oracle's element integrals (no GPU).

The reference build is not available here (its path needs a CMake-generated header), so the
oracle is pinned by the mathematics: polynomial exactness of each cubature table, an
independent 30-digit semi-analytic integral (tests/bruteforce.py), the Gauss jump, E = -grad
phi, far-field limits, vertex-relabelling invariance and closed-surface solid angles.  Cases
mirror UnitTest/KEMField/Fields/src/IntegratorsTest.cc:33-168.
"""
from math import factorial

import numpy as np
import pytest

import bruteforce as BF


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def test_constants(oracle):
    # Kommon/Base/Utility/KConst_2021.h:139-146
    assert oracle.EPS0() == 8.8541878128e-12
    assert abs(oracle.K() - 8.9875517923e9) / 8.9875517923e9 < 1e-10


def test_records_match_ktriangle_setvalues(oracle):
    from kassiopeia_amd import geometry as G

    rng = np.random.default_rng(0)
    p = rng.normal(size=(20, 3, 3))
    recs = G.records_from_points(p[:, 0], p[:, 1], p[:, 2])
    for k in range(20):
        assert np.array_equal(recs[k], oracle.set_from_points(p[k, 0], p[k, 1], p[k, 2]))


@pytest.mark.parametrize("n,deg", [(7, 5), (12, 7), (33, 12)])
def test_triangle_cubature_exactness(oracle, n, deg):
    """Each table integrates all monomials s^i t^j of total degree <= deg exactly on the unit
    right triangle (mean value 2 i! j!/(i+j+2)!) and fails at deg+1 (the table is the one
    claimed, not a lower-order one)."""
    rec = oracle.set_from_points([0, 0, 0], [1, 0, 0], [0, 1, 0])
    Q, w = oracle.tri_nodes(n, rec)
    assert len(w) == n
    s, t = Q[:, 0], Q[:, 1]
    for d in range(deg + 2):
        errs = [abs((w * s ** i * t ** (d - i)).sum() - 2 * factorial(i) * factorial(d - i) / factorial(d + 2))
                for i in range(d + 1)]
        if d <= deg:
            assert max(errs) < 2e-15, (d, errs)
        else:
            assert max(errs) > 1e-12


@pytest.mark.parametrize("n,deg", [(7, 5), (12, 7), (33, 13)])
def test_rectangle_cubature_exactness(oracle, n, deg):
    rec = oracle.set_from_points([0, 0, 0], [1, 0, 0], [0, 1, 0])
    Q, w = oracle.rect_nodes(n, rec)
    assert len(w) == n
    s, t = Q[:, 0], Q[:, 1]
    for d in range(deg + 1):
        for i in range(d + 1):
            assert abs((w * s ** i * t ** (d - i)).sum() - 1.0 / ((i + 1) * (d - i + 1))) < 2e-15


def test_cubature_nodes_on_general_triangle(oracle):
    """Nodes are affine images of the reference nodes (barycentric a*A + b*B + c*C)."""
    rng = np.random.default_rng(1)
    p = rng.normal(size=(3, 3))
    rec = oracle.set_from_points(*p)
    V = BF.triangle_vertices(rec)
    ref = oracle.set_from_points([0, 0, 0], [1, 0, 0], [0, 1, 0])
    for n in (7, 12, 33):
        Q, w = oracle.tri_nodes(n, rec)
        Qr, wr = oracle.tri_nodes(n, ref)
        expect = V[0] + Qr[:, 0:1] * (V[1] - V[0]) + Qr[:, 1:2] * (V[2] - V[0])
        assert np.allclose(Q, expect, atol=1e-14)
        assert np.array_equal(w, wr)


TRI_CASE = ([-1.75, 0.25, 0.0], [1.75, 0.25, 0.0], [0.0, 2.25, 0.0])  # IntegratorsTest.cc:102-121
RECT_CASE = ([-2.0, -2.0, 0.0], [2.0, -2.0, 0.0], [-2.0, 2.0, 0.0])  # IntegratorsTest.cc:37-48


@pytest.mark.parametrize("P,tol", [
    ([0.01, 0.5, 0.0], 1e-13),    # on the surface
    ([0.0, 0.25, 0.0], 1e-13),    # on a side line (Taylor branch)
    ([-1.75, 0.25, 0.0], 1e-7),   # on a vertex: reference averages +-1e-7 N3 shifted points
    ([10.0, 1.0, 0.0], 1e-13),    # in plane, far outside
    ([10.0, -20.5, 33.58], 1e-13),
    ([0.01, 0.5, 0.3], 1e-13),
    ([0.3, 0.7, -0.05], 1e-13),
    ([1.0, 1.0, 1e-6], 1e-13),    # just above an edge
])
def test_triangle_closed_form_potential(oracle, P, tol):
    rec = oracle.set_from_points(*TRI_CASE)
    ref = BF.polygon_field(BF.triangle_vertices(rec), P, want_field=False)
    out = oracle.tri_rwg(rec, P)
    assert abs(out[3] - ref[3]) / abs(ref[3]) < tol


@pytest.mark.parametrize("P", [[10.0, -20.5, 33.58], [0.01, 0.5, 0.3], [0.3, 0.7, -0.05], [-3.0, 4.0, 0.2],
                               [0.0, 1.0, 1e-3]])
def test_triangle_closed_form_field(oracle, P):
    rec = oracle.set_from_points(*TRI_CASE)
    ref = BF.polygon_field(BF.triangle_vertices(rec), P)
    out = oracle.tri_rwg(rec, P)
    assert rel(out[:3], ref[:3]) < 1e-12
    assert abs(out[3] - ref[3]) / abs(ref[3]) < 1e-13


@pytest.mark.parametrize("P,tol", [
    ([-2.0, -2.0, 0.0], 1e-7),    # corner P0
    ([0.1, -1.5, 0.0], 1e-13),    # on the surface
    ([0.0, 0.0, 0.0], 1e-13),     # centre
    ([10.0, 10.0, 0.0], 1e-13),   # in plane, far
    ([0.0, -2.0, 0.0], 1e-13),    # on a side line
    ([10.0, -20.5, 33.58], 1e-13),
    ([-1.0, -1.0, 0.0], 1e-13),   # on the diagonal
    ([0.5, 0.7, 0.4], 1e-13),
])
def test_rectangle_closed_form_potential(oracle, P, tol):
    rec = oracle.set_from_points(*RECT_CASE)
    ref = BF.polygon_field(BF.rectangle_vertices(rec), P, want_field=False)
    out = oracle.rect_rwg(rec, P)
    assert abs(out[3] - ref[3]) / abs(ref[3]) < tol


@pytest.mark.parametrize("P", [[10.0, -20.5, 33.58], [0.5, 0.7, 0.4], [-2.5, 1.0, -0.3]])
def test_rectangle_closed_form_field(oracle, P):
    rec = oracle.set_from_points(*RECT_CASE)
    ref = BF.polygon_field(BF.rectangle_vertices(rec), P)
    out = oracle.rect_rwg(rec, P)
    assert rel(out[:3], ref[:3]) < 1e-12


def test_rectangle_is_two_triangles(oracle):
    """Two independent closed forms: rectangle == triangle(P0,P1,P2') + triangle(P2',P3,P0)."""
    rng = np.random.default_rng(5)
    for _ in range(10):
        p0 = rng.normal(size=3)
        n1 = rng.normal(size=3)
        n1 /= np.linalg.norm(n1)
        n2 = np.cross(n1, rng.normal(size=3))
        n2 /= np.linalg.norm(n2)
        a, b = rng.uniform(0.2, 2.0, 2)
        rec = oracle.set_from_points(p0, p0 + a * n1, p0 + b * n2)
        V = BF.rectangle_vertices(rec)
        t1 = oracle.set_from_points(V[0], V[1], V[2])
        t2 = oracle.set_from_points(V[2], V[3], V[0])
        P = p0 + rng.normal(size=3) * 1.5
        r = oracle.rect_rwg(rec, P)
        s = oracle.tri_rwg(t1, P) + oracle.tri_rwg(t2, P)
        assert rel(r, s) < 1e-12


def _general_triangle(oracle):
    return oracle.set_from_points([0.1, 0.2, 0.3], [0.9, 0.1, 0.35], [0.3, 0.8, 0.2])


@pytest.mark.parametrize("dr,regime", [(0.5, 0), (2.0, 0), (3.43, 33), (10.0, 33), (30.5, 12), (100.0, 12),
                                       (131.5, 7), (1000.0, 7)])
def test_numeric_dispatch_vs_bruteforce(oracle, dr, regime):
    """The default (numeric) integrator picks 7/12/33-point cubature or the closed form by the
    distance ratio (KElectrostaticCubatureTriangleIntegrator.cc:595-665); every regime is
    accurate to rounding against the 30-digit integral."""
    rec = _general_triangle(oracle)
    V = BF.triangle_vertices(rec)
    size = (rec[0] + rec[1] + np.linalg.norm(V[2] - V[1])) / 3
    c = oracle.centroid(0, rec)
    rng = np.random.default_rng(int(dr * 100))
    for _ in range(3):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        P = c + d * dr * size
        out, reg = oracle.element_eval(0, rec, P)
        assert reg == regime
        ref = BF.polygon_field(V, P)
        assert abs(out[3] - ref[3]) / abs(ref[3]) < 1e-13
        assert rel(out[:3], ref[:3]) < 1e-13


@pytest.mark.parametrize("dr,regime", [(2.0, 0), (4.7, 33), (50.0, 12), (200.0, 7)])
def test_rectangle_numeric_dispatch(oracle, dr, regime):
    rec = oracle.set_from_points([0.1, 0.2, 0.3], [0.9, 0.1, 0.35], [0.1 - 0.05 * 0.6, 0.2 + 0.8 * 0.6, 0.3 + 0.1 * 0.6])
    # make N2 orthogonal to N1 (a rectangle)
    n1 = rec[5:8]
    n2 = rec[8:11] - np.dot(rec[8:11], n1) * n1
    n2 /= np.linalg.norm(n2)
    rec[8:11] = n2
    V = BF.rectangle_vertices(rec)
    c = oracle.centroid(1, rec)
    size = 0.5 * (rec[0] + rec[1])
    rng = np.random.default_rng(7)
    d = rng.normal(size=3)
    d /= np.linalg.norm(d)
    P = c + d * dr * size
    out, reg = oracle.element_eval(1, rec, P)
    assert reg == regime
    ref = BF.polygon_field(V, P)
    assert abs(out[3] - ref[3]) / abs(ref[3]) < 1e-13
    assert rel(out[:3], ref[:3]) < 1e-12


def test_gauss_jump(oracle):
    """Normal field jumps by sigma/eps0 across a charged sheet (one-sided limits)."""
    rec = _general_triangle(oracle)
    n = oracle.normal(0, rec)
    c = oracle.centroid(0, rec)
    P = c + 0.1 * (rec[2:5] - c)
    up, _ = oracle.element_eval(0, rec, P + 1e-9 * n)
    dn, _ = oracle.element_eval(0, rec, P - 1e-9 * n)
    jump = np.dot(up[:3] - dn[:3], n)
    assert abs(jump * oracle.EPS0() - 1.0) < 1e-6


@pytest.mark.parametrize("dr", [0.7, 2.5, 5.0, 50.0, 500.0])
def test_field_is_minus_grad_phi(oracle, dr):
    rec = _general_triangle(oracle)
    c = oracle.centroid(0, rec)
    size = 0.6
    d = np.array([0.3, -0.5, 0.81])
    d /= np.linalg.norm(d)
    P = c + d * dr * size
    E, _ = oracle.element_eval(0, rec, P)
    h = 1e-5 * dr * size
    g = np.zeros(3)
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        g[k] = (oracle.element_eval(0, rec, P + e)[0][3] - oracle.element_eval(0, rec, P - e)[0][3]) / (2 * h)
    assert rel(E[:3], -g) < 1e-7


def test_far_field_monopole(oracle):
    rec = _general_triangle(oracle)
    c = oracle.centroid(0, rec)
    r = 1.0e4
    P = c + np.array([0.0, 0.0, r])
    out, reg = oracle.element_eval(0, rec, P)
    assert reg == 7
    q = oracle.area(0, rec) * oracle.K()
    assert abs(out[3] - q / r) / (q / r) < 1e-8
    assert abs(out[2] - q / r ** 2) / (q / r ** 2) < 1e-7


def test_vertex_relabelling_invariance(oracle):
    rng = np.random.default_rng(11)
    p = rng.normal(size=(3, 3))
    r1 = oracle.set_from_points(p[0], p[1], p[2])
    r2 = oracle.set_from_points(p[1], p[2], p[0])
    for dr in (0.3, 1.0, 3.0, 8.0, 40.0, 200.0):
        P = p.mean(axis=0) + dr * np.array([0.2, 0.9, -0.4])
        a, ra = oracle.element_eval(0, r1, P)
        b, rb = oracle.element_eval(0, r2, P)
        assert rel(a, b) < 1e-13


def test_solid_angle_closed_surface(oracle):
    """A closed surface subtends 4 pi from inside and 0 from outside (KSolidAngle.hh:58-144)."""
    from kassiopeia_amd import geometry as G

    m = G.cube_mesh(2)
    for P, expect in (([0.1, -0.2, 0.15], 4 * np.pi), ([1.3, 0.2, -0.4], 0.0)):
        s = sum(oracle.solid_angle_tri(m.recs[k], P) for k in range(m.n))
        assert abs(abs(s) - expect) < 1e-12
    m = G.cube_rect_mesh(2)
    for P, expect in (([0.1, -0.2, 0.15], 4 * np.pi), ([1.3, 0.2, -0.4], 0.0)):
        s = sum(oracle.solid_angle_rect(m.recs[k], P) for k in range(m.n))
        assert abs(abs(s) - expect) < 1e-12


def _line_exact(rec, P):
    """(E, phi) of a uniformly charged segment (charge per length sigma*pi*d) by 30-digit
    quadrature of 1/r and (P-Q)/r^3 along the segment."""
    import mpmath as mp

    p0 = [mp.mpf(float(x)) for x in rec[0:3]]
    p1 = [mp.mpf(float(x)) for x in rec[3:6]]
    Pm = [mp.mpf(float(x)) for x in P]
    L = mp.sqrt(sum((p1[k] - p0[k]) ** 2 for k in range(3)))
    pre = L * mp.mpf(float(rec[6])) / (4 * BF.EPS0)

    def dv(s):
        return [Pm[k] - (p0[k] + s * (p1[k] - p0[k])) for k in range(3)]

    def r(s):
        return mp.sqrt(sum(x * x for x in dv(s)))

    phi = mp.quad(lambda s: 1 / r(s), [0, 1]) * pre
    E = [mp.quad(lambda s: dv(s)[k] / r(s) ** 3, [0, 1]) * pre for k in range(3)]
    return np.array([float(E[0]), float(E[1]), float(E[2]), float(phi)])


@pytest.mark.parametrize("dr,nodes", [(0.6, 0), (1.5, 0), (2.5, 16), (10.0, 16), (30.0, 4), (300.0, 4)])
def test_line_segment(oracle, dr, nodes):
    rec = np.zeros(11)
    rec[0:3] = [0.1, 0.2, -0.3]
    rec[3:6] = [0.9, 0.5, 0.1]
    rec[6] = 1e-3
    L = np.linalg.norm(rec[3:6] - rec[0:3])
    c = 0.5 * (rec[0:3] + rec[3:6])
    P = c + dr * L * np.array([0.48, -0.6, 0.64])
    out, reg = oracle.element_eval(2, rec, P)
    assert reg == nodes
    ref = _line_exact(rec, P)
    # analytic branch: exact to rounding; quadrature branches: the reference's 15-digit
    # Gauss-Legendre tables and cut-offs (dr > 2: 16 nodes, dr > 27.7: 4 nodes)
    tol = {0: 1e-12, 16: 1e-12, 4: 1e-11}[nodes]
    assert abs(out[3] - ref[3]) / abs(ref[3]) < tol
    assert rel(out[:3], ref[:3]) < tol
