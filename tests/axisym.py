"""Axisymmetric test geometries (conic sections and rings) and closed forms shared by the
oracle and GPU tests.

Records follow the reference's shapes: a conic section is [P0x, P0y, P0z, P1x, P1y, P1z] with
radius = x and axial position = z of each end (KConicSection::SetValues(r0, z0, r1, z1) puts
y = 0); a ring is [Px, Py, Pz].
"""
import math

import numpy as np

CONIC, RING = 3, 4


def conic(r0, z0, r1, z1):
    return np.array([r0, 0.0, z0, r1, 0.0, z1])


def ring(r, z):
    return np.array([r, 0.0, z])


def discretize_interval(interval, n, power):
    """Segment lengths of the reference's disk test (UnitTest/KEMField/Fields/src/
    DiskCapacitanceTest.cc:121-145): n segments, finer towards both ends as power > 1."""
    seg = np.zeros(n)
    mid = interval * 0.5
    if n % 2 == 1:
        seg[n // 2] = interval / n
        mid -= interval / (2 * n)
    for i in range(n // 2):
        a = (i / (n / 2.0)) ** power
        b = ((i + 1) / (n / 2.0)) ** power
        seg[i] = seg[n - (i + 1)] = mid * (b - a)
    return seg


def disk(scale=128, power=1.5):
    """The unit disk of DiskCapacitanceTest.cc:63-86: `scale` conic sections from r = 1 inwards
    (the outer half of a symmetric 2*scale-segment split of the diameter), all at 1 V."""
    seg = discretize_interval(2.0, 2 * scale, power)
    recs, r0 = [], 1.0
    for i in range(scale):
        r1 = r0 - seg[i]
        recs.append(conic(r0, 0.0, r1, 0.0))
        r0 = r1
    return np.array(recs)


def disk_capacitance_over_eps0():
    """C = 8 eps0 R for a conducting disk of radius R = 1 (DiskCapacitanceTest.cc:103-113)."""
    return 8.0


def cylinder_axis_potential(R, z0, z1, z, eps0):
    """phi on the axis of a cylindrical shell (radius R, z0..z1) of unit surface charge density."""
    return R / (2 * eps0) * (math.asinh((z1 - z) / R) - math.asinh((z0 - z) / R))


def annulus_axis_potential(r0, r1, zc, z, eps0):
    """phi on the axis of a flat annulus (radii r0 < r1 at height zc) of unit surface charge density."""
    h = z - zc
    return (math.sqrt(r1 * r1 + h * h) - math.sqrt(r0 * r0 + h * h)) / (2 * eps0)


def brute_conic(rec, P, eps0, n_p=400, n_theta=1024):
    """(Ex, Ey, Ez, phi) of a conic section of unit surface charge density by direct quadrature
    of the surface: Gauss-Legendre along the generating line, the trapezoid rule (spectral for
    periodic integrands) around the axis."""
    r0, z0, r1, z1 = rec[0], rec[2], rec[3], rec[5]
    L = math.hypot(r1 - r0, z1 - z0)
    x, w = np.polynomial.legendre.leggauss(n_p)
    p = (x + 1) / 2 * L
    w = w * L / 2
    R = r0 + p / L * (r1 - r0)
    Z = z0 + p / L * (z1 - z0)
    th = np.arange(n_theta) * 2 * np.pi / n_theta
    dx = P[0] - R[:, None] * np.cos(th)
    dy = P[1] - R[:, None] * np.sin(th)
    dz = P[2] - Z[:, None]
    d = np.sqrt(dx * dx + dy * dy + dz * dz)
    W = (w * R)[:, None] * (2 * np.pi / n_theta) / (4 * np.pi * eps0)
    return np.array([(W * dx / d**3).sum(), (W * dy / d**3).sum(), (W * dz / d**3).sum(), (W / d).sum()])
