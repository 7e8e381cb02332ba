"""Spherical capacitor of the reference's ComputeSphericalCapacitor.cc: conductor spheres
r1 = 1 m (1 V) and r3 = 3 m (0 V), dielectrics eps1 = 2 for r1 < r < r2 = 2 m and eps2 = 3 for
r2 < r < r3 (the r2 sphere is a Neumann interface with flux eps1 / eps2), meshes from the
reference's sphericalCapacitorFiles (fixtures under tests/golden).  The closed forms follow
ComputeSphericalCapacitor.cc:996-1025 (charges) and :1528-1563 (Field_Analytic); the charges are
summed per sphere by centroid radius as at :1032-1041."""
import os

import numpy as np

from kassiopeia_amd import geometry as G

EPS0 = 8.8541878128e-12
R1, R2, R3, EPS1, EPS2, V1 = 1.0, 2.0, 3.0, 2.0, 3.0, 1.0
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def mesh(scale: int) -> G.Mesh:
    return G.read_triangle_file(os.path.join(GOLDEN, f"spherical_capacitor_{scale}.dat.gz"))


def free_charge() -> float:
    return V1 * 4.0 * np.pi / (-1.0 / (EPS0 * EPS2 * R3) + 1.0 / (EPS0 * EPS2 * R2) - 1.0 / (EPS0 * EPS1 * R2) +
                               1.0 / (EPS0 * EPS1 * R1))


def sphere_charges():
    """(Q_1, Q_2, Q_3): free plus bound charge on the r1, r2 and r3 spheres."""
    q = free_charge()
    qb1, qb2 = -(EPS1 - 1.0) * q / EPS1, (EPS1 - 1.0) * q / EPS1
    qb3, qb4 = -(EPS2 - 1.0) * q / EPS2, (EPS2 - 1.0) * q / EPS2
    return q + qb1, qb2 + qb3, qb4 - q


def measured_charges(m: G.Mesh, sigma):
    q = m.areas() * sigma
    r = np.linalg.norm(m.centroids(), axis=1)
    return (q[r < 0.5 * (R1 + R2)].sum(), q[(r >= 0.5 * (R1 + R2)) & (r < 0.5 * (R2 + R3))].sum(),
            q[r >= 0.5 * (R2 + R3)].sum())


def potential(P):
    """Field_Analytic's potential between the conductors (r1 < |P| < r3)."""
    r = np.linalg.norm(np.atleast_2d(P), axis=1)
    q = free_charge()
    inner = q / (4 * np.pi) * (-1 / (EPS0 * EPS2 * R3) + 1 / (EPS0 * EPS2 * R2) - 1 / (EPS0 * EPS1 * R2) +
                               1 / (EPS0 * EPS1 * r))
    outer = q / (4 * np.pi) * (-1 / (EPS0 * EPS2 * R3) + 1 / (EPS0 * EPS2 * r))
    return np.where(r < R2, inner, outer)


def field(P):
    """Field_Analytic's radial field between the conductors."""
    P = np.atleast_2d(P)
    r = np.linalg.norm(P, axis=1)
    eps = np.where(r < R2, EPS1, EPS2)
    return free_charge() / (4 * np.pi * EPS0 * eps * r**2)[:, None] * P / r[:, None]


def probe_points(n: int, seed: int):
    """Points between the conductors, away from the meshed spheres (|r - r_k| >= 0.2)."""
    rng = np.random.default_rng(seed)
    P = rng.normal(size=(n, 3))
    P /= np.linalg.norm(P, axis=1)[:, None]
    r = np.where(rng.uniform(size=n) < 0.5, rng.uniform(1.2, 1.8, n), rng.uniform(2.2, 2.8, n))
    return P * r[:, None]
