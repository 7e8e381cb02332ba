"""The oracle's axisymmetric elements (KConicSection, KRing under every policy: the analytic
ring and conic-section integrators) against mathematics that fixes them: closed forms of the
complete elliptic integrals and the Legendre relation, the potential and field of a ring and of
cylinder and annulus bands on their axis, direct surface quadrature away from the surface,
E = -grad phi, and the reference's own disk-capacitance test (C = 8 eps0 R within 1e-2,
UnitTest/KEMField/Fields/src/DiskCapacitanceTest.cc:103-113)."""
import math

import numpy as np
import pytest

import axisym as X

K_HALF = 1.8540746773013719184  # K(1/sqrt 2) = Gamma(1/4)^2 / (4 sqrt(pi))
E_HALF = 1.3506438810476755025  # E(1/sqrt 2)


def test_elliptic_closed_forms(oracle):
    assert oracle.ell_K(0.0) == pytest.approx(math.pi / 2, rel=1e-15)
    assert oracle.ell_E(0.0) == pytest.approx(math.pi / 2, rel=1e-15)
    k = 1 / math.sqrt(2)
    assert oracle.ell_K(k) == pytest.approx(K_HALF, rel=1e-15)
    assert oracle.ell_E(k) == pytest.approx(E_HALF, rel=1e-15)
    assert oracle.ell_E(1.0) == 1.0


@pytest.mark.parametrize("k", [0.05, 0.2, 0.5, 0.8, 0.95, 0.999, 0.99999])
def test_legendre_relation(oracle, k):
    kp = math.sqrt(1 - k * k)
    E, K, Ep, Kp = oracle.ell_E(k), oracle.ell_K(k), oracle.ell_E(kp), oracle.ell_K(kp)
    assert E * Kp + Ep * K - K * Kp == pytest.approx(math.pi / 2, rel=2e-14)


def test_e_minus_k_over_k2(oracle):
    for k in (0.5, 0.7, 0.9, 0.99):
        direct = (oracle.ell_E(k) - oracle.ell_K(k)) / (k * k)
        assert oracle.ell_EK(k) == pytest.approx(direct, rel=1e-13)
    # small k: -pi/4 - 3 pi k^2 / 32 - 15 pi k^4 / 256 - ...
    for k in (1e-4, 1e-3):
        series = -math.pi / 4 - 3 * math.pi / 32 * k * k - 15 * math.pi / 256 * k**4
        assert oracle.ell_EK(k) == pytest.approx(series, rel=1e-14)


def test_k_logarithm_as_the_reference_regularises_it(oracle):
    # K ~ ln(4/k') for k' -> 0; the reference's fit takes log(eta + 1e-19), eta = 1 - k^2, and so
    # stays finite at k = 1 (Math/src/KEllipticIntegrals.cc:54)
    kp = 1e-5
    k = math.sqrt(1 - kp * kp)
    eta = 1 - k * k
    assert oracle.ell_K(k) == pytest.approx(math.log(4) - 0.5 * math.log(eta), rel=1e-9)
    assert oracle.ell_K(1.0) == pytest.approx(math.log(4) - 0.5 * math.log(1e-19), rel=1e-15)


@pytest.mark.parametrize("z", [-1.0, 0.0, 0.35, 2.5])
def test_ring_on_axis(oracle, z):
    # unit charge on a ring of radius R at height Z: phi = 1/(4 pi eps0 S), Ez = (z - Z)/(4 pi eps0 S^3)
    R, Z = 0.7, 0.3
    out = oracle.ring(X.ring(R, Z), np.array([0.0, 0.0, z]))
    S = math.hypot(R, z - Z)
    k = oracle.K()
    assert out[3] == pytest.approx(k / S, rel=1e-14)
    assert out[2] == pytest.approx(k * (z - Z) / S**3, rel=1e-13, abs=1e-12 * k / S**2)
    assert out[0] == 0.0 and out[1] == 0.0


def test_ring_field_is_minus_gradient(oracle):
    rec = X.ring(0.6, -0.2)
    rng = np.random.default_rng(3)
    for _ in range(6):
        P = rng.normal(size=3)
        out = oracle.ring(rec, P)
        h = 1e-5
        g = np.array([(oracle.ring(rec, P + h * e)[3] - oracle.ring(rec, P - h * e)[3]) / (2 * h) for e in np.eye(3)])
        assert np.linalg.norm(out[:3] + g) <= 1e-7 * np.linalg.norm(out[:3])


def test_ring_far_field_is_a_point_charge(oracle):
    out = oracle.ring(X.ring(0.5, 0.0), np.array([300.0, 100.0, 200.0]))
    d = math.sqrt(300.0**2 + 100.0**2 + 200.0**2)
    assert out[3] == pytest.approx(oracle.K() / d, rel=1e-6)


@pytest.mark.parametrize("P", [[0.3, 0.2, 1.5], [1.2, -0.4, 0.2], [0.1, 0.0, -0.5], [2.0, 1.0, 3.0]])
def test_conic_against_surface_quadrature(oracle, P):
    rec = X.conic(0.5, 0.1, 0.8, 0.6)
    P = np.array(P)
    o, b = oracle.conic(rec, P), X.brute_conic(rec, P, oracle.EPS0())
    assert abs(o[3] - b[3]) <= 1e-10 * abs(b[3])
    assert np.linalg.norm(o[:3] - b[:3]) <= 1e-9 * np.linalg.norm(b[:3])


@pytest.mark.parametrize("z", [-2.0, -0.3, 0.5, 1.1, 3.0])
def test_cylinder_band_on_axis(oracle, z):
    # within about a length of the band the reference integrator's own quadrature leaves up to
    # 3.3e-7 (direct surface quadrature agrees with the closed form to 1e-14 there)
    R, z0, z1 = 0.4, -0.2, 0.9
    out = oracle.conic(X.conic(R, z0, R, z1), np.array([0.0, 0.0, z]))
    tol = 1e-12 if (z < z0 - 1 or z > z1 + 1) else 1e-6
    assert out[3] == pytest.approx(X.cylinder_axis_potential(R, z0, z1, z, oracle.EPS0()), rel=tol)


@pytest.mark.parametrize("z", [0.02, 0.1, 0.5, 2.0])
def test_annulus_on_axis(oracle, z):
    # the 100- and 500-node branches and the 20-node one, on the axis of a flat annulus
    r0, r1 = 0.3, 1.0
    out = oracle.conic(X.conic(r1, 0.0, r0, 0.0), np.array([0.0, 0.0, z]))
    assert out[3] == pytest.approx(X.annulus_axis_potential(r0, r1, 0.0, z, oracle.EPS0()), rel=1e-8)


def test_conic_field_is_minus_gradient(oracle):
    rec = X.conic(0.5, 0.1, 0.8, 0.6)
    for P in ([0.3, 0.2, 1.5], [1.2, -0.4, 0.2], [0.05, 0.3, -0.6]):
        P = np.array(P)
        out = oracle.conic(rec, P)
        h = 1e-5
        g = np.array([(oracle.conic(rec, P + h * e)[3] - oracle.conic(rec, P - h * e)[3]) / (2 * h) for e in np.eye(3)])
        assert np.linalg.norm(out[:3] + g) <= 1e-6 * np.linalg.norm(out[:3])


def test_area_centroid_normal(oracle):
    rec = X.conic(0.5, 0.1, 0.8, 0.6)
    assert oracle.area(X.CONIC, rec) == pytest.approx(math.pi * 1.3 * math.hypot(0.3, 0.5), rel=1e-15)
    assert np.allclose(oracle.centroid(X.CONIC, rec), [0.65, 0.0, 0.35])
    n = oracle.normal(X.CONIC, rec)
    assert np.dot(n, [0.3, 0.0, 0.5]) == pytest.approx(0.0, abs=1e-15) and np.linalg.norm(n) == pytest.approx(1.0)
    assert oracle.area(X.RING, X.ring(0.4, 1.0)) == pytest.approx(2 * math.pi * 0.4)
    assert np.allclose(oracle.normal(X.RING, X.ring(0.4, 1.0)), [1.0, 0.0, 0.0])


def test_disk_capacitance(oracle):
    recs = X.disk(128)
    shapes = np.full(len(recs), X.CONIC, np.int32)
    A = oracle.assemble(shapes, recs)
    x = np.linalg.solve(A, np.ones(len(recs)))
    Q = sum(oracle.area(X.CONIC, recs[i]) * x[i] for i in range(len(recs)))
    C = Q / oracle.EPS0()
    assert abs(C - X.disk_capacitance_over_eps0()) < 1e-2  # the reference test's tolerance
    assert abs(C - X.disk_capacitance_over_eps0()) < 2e-3  # this discretisation: 7.99902


def test_geometry_records_match_the_oracle(oracle):
    from kassiopeia_amd import geometry as G

    m = G.disk_mesh(16)
    assert m.n == 16 and np.all(m.shapes == X.CONIC)
    assert m.recs[-1, 3] == 0.0  # the centre closes at r = 0 exactly
    rg = G.ring_records([0.3, 1.2], [0.5, -0.1])
    mesh = G.Mesh(np.array([X.CONIC, X.RING, X.RING], np.int32), np.vstack([m.recs[:1], rg]), np.zeros(3, np.int32),
                  np.ones(3), np.zeros(3), np.zeros(3, np.int32))
    for i in range(3):
        assert mesh.areas()[i] == pytest.approx(oracle.area(mesh.shapes[i], mesh.recs[i]), rel=1e-15)
        assert np.allclose(mesh.centroids()[i], oracle.centroid(mesh.shapes[i], mesh.recs[i]), rtol=0, atol=1e-15)
