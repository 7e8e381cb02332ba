"""Pins of the oracle's "analytic" and "reference" integrator policies (no GPU).

analytic: KElectrostaticAnalyticTriangleIntegrator / ...RectangleIntegrator (Analytic/src) --
closed forms derived independently of the RWG ones; both are checked against the 30-digit
semi-analytic integral of tests/bruteforce.py and against each other.
reference: KElectrostaticBiQuadrature{Triangle,Rectangle}Integrator and the 256-node line
rule (Reference/src) -- nested 32-node Gauss-Legendre; at distances of a few element sizes a
32 x 32 product rule converges to the exact integral, which pins node placement, weights,
the chord construction and the prefactors.
"""
import numpy as np
import pytest

import bruteforce as BF

TRI_CASE = ([-1.75, 0.25, 0.0], [1.75, 0.25, 0.0], [0.0, 2.25, 0.0])  # IntegratorsTest.cc:102-121
RECT_CASE = ([-2.0, -2.0, 0.0], [2.0, -2.0, 0.0], [-2.0, 2.0, 0.0])  # IntegratorsTest.cc:37-48
OFF_PLANE = [[10.0, -20.5, 33.58], [0.01, 0.5, 0.3], [0.3, 0.7, -0.05], [-3.0, 4.0, 0.2], [0.0, 1.0, 1e-3],
             [0.5, 0.7, 0.4], [-2.5, 1.0, -0.3], [0.2, 0.3, 5.0], [3.0, 3.0, 3.0]]


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.mark.parametrize("P", OFF_PLANE)
def test_analytic_triangle_vs_bruteforce(oracle, P):
    rec = oracle.set_from_points(*TRI_CASE)
    ref = BF.polygon_field(BF.triangle_vertices(rec), P)
    out = oracle.tri_analytic(rec, P)
    assert abs(out[3] - ref[3]) / abs(ref[3]) < 1e-13
    assert rel(out[:3], ref[:3]) < 1e-12


@pytest.mark.parametrize("P", OFF_PLANE)
def test_analytic_rectangle_vs_bruteforce(oracle, P):
    """The reference's rectangle primitive adds 1e-12 inside its logarithms, so the potential
    carries an absolute error of order 1e-12 * size / (4 pi eps0)."""
    rec = oracle.set_from_points(*RECT_CASE)
    ref = BF.polygon_field(BF.rectangle_vertices(rec), P)
    out = oracle.rect_analytic(rec, P)
    assert abs(out[3] - ref[3]) / abs(ref[3]) < 5e-12
    assert rel(out[:3], ref[:3]) < 1e-13


def test_analytic_general_triangles_match_rwg(oracle):
    """Two independent closed forms agree on random triangles and points."""
    rng = np.random.default_rng(11)
    for _ in range(200):
        p = rng.normal(size=(3, 3))
        rec = oracle.set_from_points(*p)
        P = p.mean(0) + rng.normal(size=3) * rng.choice([0.1, 1.0, 10.0])
        a, r = oracle.tri_analytic(rec, P), oracle.tri_rwg(rec, P)
        assert abs(a[3] - r[3]) / abs(r[3]) < 1e-11
        assert rel(a[:3], r[:3]) < 1e-9


def test_analytic_in_plane(oracle):
    """In the element plane (|z| <= 1e-14).  Potential: against the brute force.  Triangle
    field: the reference's in-plane branch feeds the unscaled local coordinates to primitives
    written for coordinates scaled by z, whose '+1' terms stand for z^2/z^2; its in-plane
    field is therefore the field at height z = 1 (in the element's length unit), not the
    in-plane limit.  The oracle keeps that behaviour; the pin is the identity
    E_inplane(P) == E(P + N3) (tangential components) with E(P + N3) against the brute force.
    Rectangle in-plane field: equals the RWG closed form.  Normal component: 1/(2 eps0) at the
    centroid (the reference's dist < 1e-12 branch) or inside the rectangle, 0 elsewhere."""
    rec = oracle.set_from_points(*TRI_CASE)
    n3 = oracle.normal(0, rec)
    for P in ([0.01, 0.5, 0.0], [10.0, 1.0, 0.0], [-4.0, -3.0, 0.0]):
        ref = BF.polygon_field(BF.triangle_vertices(rec), P, want_field=False)
        out = oracle.tri_analytic(rec, P)
        assert abs(out[3] - ref[3]) / abs(ref[3]) < 1e-13
    for P in ([10.0, 1.0, 0.0], [-4.0, -3.0, 0.0], [0.3, 0.9, 0.0]):
        out = oracle.tri_analytic(rec, P)
        up = np.asarray(P) + n3
        assert out[2] == 0.0
        assert rel(out[:2], oracle.tri_analytic(rec, up)[:2]) < 1e-14
        assert rel(out[:2], BF.polygon_field(BF.triangle_vertices(rec), up)[:2]) < 1e-12
    c = oracle.centroid(0, rec)
    out = oracle.tri_analytic(rec, c)
    assert out[2] == 1.0 / (2.0 * oracle.EPS0())
    rr = oracle.set_from_points(*RECT_CASE)
    for P in ([10.0, 1.0, 0.0], [-4.0, -3.0, 0.0], [0.3, 0.9, 0.0]):
        assert rel(oracle.rect_analytic(rr, P)[:2], oracle.rect_rwg(rr, P)[:2]) < 1e-12
    out = oracle.rect_analytic(rr, [0.3, -0.4, 0.0])
    assert out[2] == 1.0 / (2.0 * oracle.EPS0())
    out = oracle.rect_analytic(rr, [3.0, 0.0, 0.0])
    assert out[2] == 0.0


@pytest.mark.parametrize("P", [[10.0, -20.5, 33.58], [-3.0, 4.0, 0.2], [0.2, 0.3, 5.0], [3.0, 3.0, 3.0],
                               [-2.5, 1.0, -1.5]])
def test_biquadrature_converges_far(oracle, P):
    rt = oracle.set_from_points(*TRI_CASE)
    ref = BF.polygon_field(BF.triangle_vertices(rt), P)
    out = oracle.tri_biquad(rt, P)
    assert rel(out, ref) < 1e-13
    rr = oracle.set_from_points(*RECT_CASE)
    ref = BF.polygon_field(BF.rectangle_vertices(rr), P)
    out = oracle.rect_biquad(rr, P)
    assert rel(out, ref) < 1e-13


def test_biquadrature_general_triangles(oracle):
    """Random triangles (every vertex ordering the chord construction can meet) against the
    closed form at 3-10 sizes away (the closed form's own cancellation error grows as
    1e-14 * dr^2, dr = distance / mean side)."""
    rng = np.random.default_rng(12)
    for _ in range(100):
        p = rng.normal(size=(3, 3))
        rec = oracle.set_from_points(*p)
        size = np.linalg.norm(p[1] - p[0]) + np.linalg.norm(p[2] - p[0])
        mean_side = (size + np.linalg.norm(p[2] - p[1])) / 3.0
        d = rng.normal(size=3)
        dist = size * rng.uniform(3, 10)
        P = p.mean(0) + d / np.linalg.norm(d) * dist
        tol = max(1e-12, 1e-14 * (dist / mean_side) ** 2)
        assert rel(oracle.tri_biquad(rec, P), oracle.tri_rwg(rec, P)) < tol


def test_line_256_converges(oracle):
    """8 x 32-node rule against the exact wire integral (Analytic line integrator) at 0.5-20
    lengths away, and the symmetric split of nodes (reversing P0 and P1 changes nothing)."""
    rng = np.random.default_rng(13)
    for _ in range(50):
        p0, p1 = rng.normal(size=3), rng.normal(size=3)
        rec = np.zeros(11)
        rec[0:3], rec[3:6], rec[6] = p0, p1, 1e-3
        L = np.linalg.norm(p1 - p0)
        d = rng.normal(size=3)
        P = 0.5 * (p0 + p1) + d / np.linalg.norm(d) * L * rng.uniform(0.5, 20)
        ref = oracle.line_analytic(rec, P)
        out = oracle.line_256(rec, P)
        assert rel(out, ref) < 1e-12
        rev = rec.copy()
        rev[0:3], rev[3:6] = p1, p0
        assert rel(oracle.line_256(rev, P), out) < 1e-14


def test_policy_dispatch(oracle):
    """KElectrostaticBoundaryIntegratorFactory::MakeAnalytic / MakeReference (:42-77)."""
    rt = oracle.set_from_points(*TRI_CASE)
    rr = oracle.set_from_points(*RECT_CASE)
    line = np.zeros(11)
    line[0:7] = [0, 0, 0, 1, 0, 0, 1e-3]
    P = [0.4, 0.9, 0.7]
    A, R = oracle.ANALYTIC, oracle.REFERENCE
    assert np.array_equal(oracle.element_eval(0, rt, P, A)[0], oracle.tri_analytic(rt, P))
    assert np.array_equal(oracle.element_eval(1, rr, P, A)[0], oracle.rect_analytic(rr, P))
    assert np.array_equal(oracle.element_eval(2, line, P, A)[0], oracle.line_analytic(line, P))
    assert np.array_equal(oracle.element_eval(0, rt, P, R)[0], oracle.tri_biquad(rt, P))
    assert np.array_equal(oracle.element_eval(1, rr, P, R)[0], oracle.rect_biquad(rr, P))
    assert np.array_equal(oracle.element_eval(2, line, P, R)[0], oracle.line_256(line, P))
