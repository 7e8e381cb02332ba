"""Symmetry groups (KSymmetryGroup) in the oracle, pinned by symmetry: a mesh of groups of an
element and its mirror image is the symmetric reduction of the full mirrored mesh, so its
solve gives every member the charge density the full solve gives it, its charge is half the
full charge, and its field equals the full mesh's (the reference's group integral sums the
members, Core/include/KElectrostaticElementIntegrator.hh:52-79, and collocates at the first)."""
import numpy as np
import pytest

from kassiopeia_amd import geometry as G


def upper_half_cube(n=4):
    cube = G.cube_mesh(n)
    keep = np.nonzero(cube.centroids()[:, 2] > 1e-9)[0]
    return cube.subset(keep)


@pytest.fixture(scope="module")
def mirrored(oracle):
    half = upper_half_cube()
    grouped = G.mirror_groups(half, [0.0, 0.0, 0.0], [0.0, 0.0, 1.0])
    full = G.Mesh(grouped.shapes, grouped.recs, np.zeros(2 * half.n, np.int32), np.ones(2 * half.n),
                  np.zeros(2 * half.n), np.zeros(2 * half.n, np.int32))
    return half, grouped, full


def test_members_are_mirror_images(mirrored):
    half, grouped, full = mirrored
    c = full.centroids()
    assert np.allclose(c[1::2] * [1, 1, -1], c[0::2], atol=1e-15)
    assert grouped.n == half.n and len(grouped.group_start) == half.n + 1
    assert np.allclose(grouped.centroids(), c[0::2]) and np.allclose(grouped.areas(), full.areas()[0::2])


def test_grouped_solve_is_the_symmetric_reduction(oracle, mirrored):
    half, grouped, full = mirrored
    Ag = oracle.assemble_groups(grouped.group_start, grouped.shapes, grouped.recs)
    Af = oracle.assemble(full.shapes, full.recs)
    # row i of the grouped matrix: the full rows of member 2i, columns of each group's members added
    assert np.allclose(Ag, Af[0::2, 0::2] + Af[0::2, 1::2], rtol=1e-14, atol=0)
    xg = np.linalg.solve(Ag, np.ones(grouped.n))
    xf = np.linalg.solve(Af, np.ones(full.n))
    assert np.abs(xf[0::2] - xg).max() <= 1e-10 * np.abs(xf).max()
    assert np.abs(xf[1::2] - xg).max() <= 1e-10 * np.abs(xf).max()
    # capacitance: a group's charge counts its first member's area (KSymmetryGroup::Area)
    assert (grouped.areas() * xg).sum() == pytest.approx(0.5 * (full.areas() * xf).sum(), rel=1e-10)


def test_grouped_field_equals_the_full_field(oracle, mirrored):
    half, grouped, full = mirrored
    xg = np.linalg.solve(oracle.assemble_groups(grouped.group_start, grouped.shapes, grouped.recs),
                         np.ones(grouped.n))
    xf = np.repeat(xg, 2)
    P = G.points_shell(200, 0.9, 3.0, 4)
    pg, Eg = oracle.field_groups(grouped.group_start, grouped.shapes, grouped.recs, xg, P)
    pf, Ef = oracle.field(full.shapes, full.recs, xf, P)
    assert np.abs(pg - pf).max() <= 1e-13 * np.abs(pf).max()
    assert np.abs(Eg - Ef).max() <= 1e-12 * np.abs(Ef).max()


def test_neumann_group_rows(oracle, mirrored):
    half, grouped, full = mirrored
    bc = np.zeros(grouped.n, np.int32)
    bc[::3] = 1
    flux = np.where(bc == 1, 0.3, 0.0)
    Ag = oracle.assemble_groups(grouped.group_start, grouped.shapes, grouped.recs, bc, flux)
    Af = oracle.assemble(full.shapes, full.recs, np.repeat(bc, 2), np.repeat(flux, 2))
    # off the diagonal a Neumann row is the normal field of both members; on it the reference's
    # self term alone (the mirror member's field at the first member is not added)
    both = Af[0::2, 0::2] + Af[0::2, 1::2]
    off = ~np.eye(grouped.n, dtype=bool)
    assert np.allclose(Ag[off], both[off], rtol=1e-13, atol=0)
    neu = bc == 1
    assert np.array_equal(np.diag(Ag)[neu], np.diag(Af[0::2, 0::2])[neu])
    assert np.allclose(np.diag(Ag)[~neu], np.diag(both)[~neu], rtol=1e-13, atol=0)
