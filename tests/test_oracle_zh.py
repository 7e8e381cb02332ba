"""The oracle's zonal-harmonic restatement (oracle/kem_oracle_zh.c) pinned without the
reference's binaries:

* the reference's own test (UnitTest/KEMField/Fields/src/ZonalHarmonicsTest.cc:154-202): a ring
  at (7, 0, 0) of charge 1 and a conic section (7, 0, 0)-(6, 0, 1) of density 1e-8, 1,000
  points uniform in [-4, 4]^3 from glibc's rand() seeded 1337, the sums of the relative
  deviations of phi and E from direct integration within 1e-8;
* a ring's potential on the axis in closed form, Q / (4 pi eps0 |P - Q|), and its Legendre
  coefficients against scipy's Legendre polynomials;
* the expansion off the axis against the elliptic-integral ring integrator and the conic-section
  quadrature (series and integrals computed by unrelated methods);
* the tree: bifurcation levels, source points, and the paths (central, remote, direct) taken.
"""
import ctypes
import os

import numpy as np
import pytest

from kassiopeia_amd import geometry as G

EPS0 = 8.8541878128e-12
K = 1.0 / (4.0 * np.pi * EPS0)


@pytest.fixture(scope="module")
def Z():
    from oracle import zh

    return zh


def direct(oracle, shapes, recs, sigma, pts):
    phi, E = oracle.field(np.asarray(shapes, np.int32), np.asarray(recs), np.asarray(sigma, np.float64), pts)
    return E, phi


def glibc_points(n, seed):
    """the reference test's points: -4 + 8 rand() / RAND_MAX per coordinate (glibc's rand)"""
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(seed)
    v = np.array([libc.rand() for _ in range(3 * n)], dtype=np.float64)
    return -4.0 + 8.0 * v.reshape(n, 3) / 2147483647.0


def rel_sum(a, b):
    """the reference test's sum of relative deviations (|b| <= 1e-14 counts absolutely)"""
    return float(np.sum((a - b) / np.where(np.abs(b) > 1e-14, b, 1.0)))


def test_reference_ring_and_conic(Z, oracle):
    shapes = [4, 3]
    recs = np.concatenate([G.ring_records(7.0, 0.0), G.conic_records(7.0, 0.0, 6.0, 1.0)])
    sigma = [1.0, 1e-8]
    zh = Z.ZH(shapes, recs, sigma)
    pts = glibc_points(1000, 1337)
    E, phi, cnt = zh.field(pts)
    Ed, phid = direct(oracle, shapes, recs, sigma, pts)
    assert abs(rel_sum(phi, phid)) < 1e-8
    for j in range(3):
        assert abs(rel_sum(E[:, j], Ed[:, j])) < 1e-8
    # both expansions and the direct fallback run; the head has two leaves, one per element
    assert cnt[0] > 0 and cnt[2] > 0 and cnt.sum() == 2 * len(pts)
    assert zh.info()[0] == 3


def test_ring_on_axis_closed_form(Z):
    R, zr, Q = 1.3, 0.4, 2.5e-9
    zh = Z.ZH([4], G.ring_records(R, zr), [Q])
    z = np.concatenate([np.linspace(-1.0, 1.8, 29), [zr + 50.0, zr - 80.0]])
    pts = np.stack([0 * z, 0 * z, z], 1)
    E, phi, cnt = zh.field(pts)
    rho = np.sqrt(R ** 2 + (z - zr) ** 2)
    assert np.abs(phi - K * Q / rho).max() <= 1e-13 * np.abs(K * Q / rho).max()
    Ez = K * Q * (z - zr) / rho ** 3
    assert np.abs(E[:, 2] - Ez).max() <= 1e-12 * np.abs(Ez).max()
    assert np.all(E[:, :2] == 0.0)
    assert cnt[0] > 0 and cnt[1] > 0  # near points central, the far two remote


def test_ring_coefficients_are_legendre(Z):
    from scipy.special import eval_legendre

    R, zr, Q = 0.8, 0.25, 1e-9
    zh = Z.ZH([4], G.ring_records(R, zr), [Q])
    z0, rho, c = zh.central_sp(0)
    rr = np.hypot(R, zr - z0)
    assert rho == pytest.approx(rr, rel=1e-15)
    n = np.arange(60)
    ref = K * Q / rr * (rho / rr) ** n * eval_legendre(n, (zr - z0) / rr)
    assert np.abs(c[:60] - ref).max() <= 1e-13 * np.abs(ref).max()


def test_ring_off_axis_against_elliptic_integrals(Z, oracle):
    R, zr, Q = 1.0, 0.0, 1e-9
    zh = Z.ZH([4], G.ring_records(R, zr), [Q])
    rng = np.random.default_rng(3)
    # inside the central sphere (r < 0.99 R about the source point) and beyond the remote one
    r = np.concatenate([rng.uniform(0.0, 0.6, 200), rng.uniform(1.6, 6.0, 200)])
    t = rng.uniform(0, 2 * np.pi, 400)
    zz = np.concatenate([rng.uniform(-0.3, 0.3, 200), rng.uniform(-6, 6, 200)])
    pts = np.stack([r * np.cos(t), r * np.sin(t), zz], 1)
    E, phi, cnt = zh.field(pts)
    Ed, phid = direct(oracle, [4], G.ring_records(R, zr), [Q], pts)
    assert cnt[2] == 0
    assert np.abs(phi - phid).max() <= 1e-12 * np.abs(phid).max()
    assert np.linalg.norm(E - Ed, axis=1).max() <= 1e-11 * np.linalg.norm(Ed, axis=1).max()


def test_conic_sections_against_quadrature(Z, oracle):
    # a cone of 12 sections (r 1 -> 0.4 over z 0 -> 1.2) with uneven charge densities
    r = np.linspace(1.0, 0.4, 13)
    z = np.linspace(0.0, 1.2, 13)
    recs = G.conic_records(r[:-1], z[:-1], r[1:], z[1:])
    sigma = np.linspace(1.0, 3.0, 12) * 1e-9
    shapes = np.full(12, 3)
    zh = Z.ZH(shapes, recs, sigma)
    rng = np.random.default_rng(7)
    pts = np.concatenate([np.stack([rng.uniform(-0.2, 0.2, 150), rng.uniform(-0.2, 0.2, 150),
                                    rng.uniform(0.0, 1.2, 150)], 1),
                          G.points_shell(150, 3.0, 8.0, 5)])
    E, phi, cnt = zh.field(pts)
    Ed, phid = direct(oracle, shapes, recs, sigma, pts)
    assert cnt[0] > 0 and cnt[1] > 0
    assert np.abs(phi - phid).max() <= 1e-9 * np.abs(phid).max()
    assert np.linalg.norm(E - Ed, axis=1).max() <= 1e-8 * np.linalg.norm(Ed, axis=1).max()


def test_tree_and_mixed_elements(Z, oracle):
    # 8 rings along z and a triangle: the head groups the rings apart from the triangle, the
    # rings bifurcate twice; the triangle's node integrates directly at every point
    zr = np.linspace(-1.0, 1.0, 8)
    rings = G.ring_records(np.full(8, 0.9), zr)
    tri = G.records_from_points(np.array([2.0, 0.0, 0.0]), np.array([2.2, 0.0, 0.0]), np.array([2.0, 0.3, 0.1]))
    shapes = np.array([4] * 8 + [0])
    recs = np.concatenate([rings, tri])
    sigma = np.concatenate([np.linspace(-1, 1, 8) * 1e-9, [2e-9]])
    zh = Z.ZH(shapes, recs, sigma, n_bifurcations=2)
    nodes, ncen, nrem = zh.info()
    # head, ring group, triangle, 2 + 4 bifurcation nodes
    assert nodes == 1 + 2 + 2 + 4
    pts = np.concatenate([np.stack([np.full(50, 0.1), np.zeros(50), np.linspace(-1.2, 1.2, 50)], 1),
                          G.points_shell(100, 4.0, 9.0, 2)])
    E, phi, cnt = zh.field(pts)
    assert cnt[2] >= len(pts)  # the triangle at every point
    Ed, phid = direct(oracle, shapes, recs, sigma, pts)
    assert np.abs(phi - phid).max() <= 1e-11 * np.abs(phid).max()
    assert np.linalg.norm(E - Ed, axis=1).max() <= 1e-10 * np.linalg.norm(Ed, axis=1).max()
