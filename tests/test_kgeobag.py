"""KGeoBag mesh import (kassiopeia_amd/kgeobag.py) on the host: element areas and aspect ratios
against closed forms, frames, filters, the container's group-major order, axial meshes and
discrete rotational symmetry groups (KEMField/Source/Interface/KGeoBag/src/KGBEMConverter.cc)."""
import math

import numpy as np
import pytest

from kassiopeia_amd import geometry as G
from kassiopeia_amd import kgeobag as KG


def rot(axis, angle):
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(angle) * K + (1 - math.cos(angle)) * K @ K


def test_element_areas_and_aspects():
    t = KG.KGMeshTriangle([0, 0, 0], [3, 0, 0], [0, 4, 0])
    assert t.Area() == pytest.approx(6.0, rel=1e-15)
    assert t.Aspect() == pytest.approx(5.0 / (12.0 / 5.0), rel=1e-14)  # hypotenuse / its height
    e = KG.KGMeshTriangle([0, 0, 0], [2, 0, 0], [1, math.sqrt(3), 0])
    assert e.Aspect() == pytest.approx(2.0 / math.sqrt(3.0), rel=1e-14)
    # the longest side found whichever vertex is P0
    s = KG.KGMeshTriangle([3, 0, 0], [0, 4, 0], [0, 0, 0])
    assert s.Aspect() == pytest.approx(25.0 / 12.0, rel=1e-14)
    r = KG.KGMeshRectangle([0, 0, 0], [2, 0, 0], [2, 0.5, 0], [0, 0.5, 0])
    assert (r.Area(), r.Aspect()) == (pytest.approx(1.0), pytest.approx(4.0))
    w = KG.KGMeshWire([0, 0, 0], [0, 0, 3], 0.1)
    assert w.Aspect() == pytest.approx(30.0)
    assert w.Area() == pytest.approx((0.5 * math.pi * 0.1 + 3.0) * math.pi * 0.1)
    # a loop (z, r) = (0, 1) -> (1, 1) is the mantle of a unit cylinder of height 1
    lp = KG.KGAxialMeshLoop([0.0, 1.0], [1.0, 1.0])
    assert lp.Area() == pytest.approx(2 * math.pi, rel=1e-15) and lp.Aspect() == pytest.approx(1.0)
    ring = KG.KGAxialMeshRing(0.01, [0.0, 2.0])
    assert ring.Area() == pytest.approx(2 * math.pi ** 2 * 2.0 * 0.01)


def _cube_surfaces(n, rng):
    """The triangles of cube_mesh(n), each face a surface in a frame of its own."""
    m = G.cube_mesh(n)
    per = 2 * n * n
    surfaces, R_all = [], []
    for f in range(6):
        R = rot(rng.normal(size=3), rng.uniform(0, 2 * math.pi))
        o = rng.normal(size=3)
        recs = m.recs[f * per:(f + 1) * per]
        p0 = recs[:, 2:5]
        p1 = p0 + recs[:, 5:8] * recs[:, 0:1]
        p2 = p0 + recs[:, 8:11] * recs[:, 1:2]
        loc = [(P - o) @ R for P in (p0, p1, p2)]  # local coordinates: global = o + R local
        elems = [KG.KGMeshTriangle(loc[0][k], loc[1][k], loc[2][k]) for k in range(per)]
        surfaces.append(KG.KGSurface(mesh=elems, bem=KG.KGBEMData("dirichlet", 1.0), origin=o, x_axis=R[:, 0],
                                     y_axis=R[:, 1], z_axis=R[:, 2]))
        R_all.append(R)
    return m, surfaces


@pytest.mark.parametrize("system", ["global", "rotated"])
def test_mesh_converter_frames(system):
    rng = np.random.default_rng(5)
    m, surfaces = _cube_surfaces(3, rng)
    conv = KG.KGBEMMeshConverter()
    if system == "rotated":
        Q, o2 = rot([1, 2, 3], 0.7), np.array([0.3, -0.2, 0.5])
        conv.SetSystem(o2, Q[:, 0], Q[:, 1], Q[:, 2])
    for s in surfaces:
        assert conv.VisitSurface(s)
    mesh = conv.container.mesh()
    assert mesh.n == m.n and mesh.group_start is None
    p0 = m.recs[:, 2:5]
    p1 = p0 + m.recs[:, 5:8] * m.recs[:, 0:1]
    p2 = p0 + m.recs[:, 8:11] * m.recs[:, 1:2]
    if system == "rotated":
        p0, p1, p2 = ((P - o2) @ Q for P in (p0, p1, p2))
    ref = G.records_from_points(p0, p1, p2)
    assert np.abs(mesh.recs - ref).max() < 1e-14
    assert np.all(mesh.bc_type == 0) and np.all(mesh.bc_value == 1.0)


def test_filters_and_ignored():
    good = KG.KGMeshTriangle([0, 0, 0], [1, 0, 0], [0, 1, 0])  # area 0.5, aspect 2
    tiny = KG.KGMeshTriangle([0, 0, 0], [1e-4, 0, 0], [0, 1e-4, 0])  # area 5e-9
    sliver = KG.KGMeshTriangle([0, 0, 0], [1, 0, 0], [0.5, 1e-3, 0])  # aspect 1000
    wire = KG.KGMeshWire([0, 0, 0], [0, 0, 1], 0.5)  # aspect 2
    conv = KG.KGBEMMeshConverter()
    conv.SetMinimumArea(1e-6)
    conv.SetMaximumAspectRatio(100.0)
    conv.VisitSurface(KG.KGSurface(mesh=[good, tiny, sliver, wire, "not an element"]))
    mesh = conv.container.mesh()
    assert list(mesh.shapes) == [G.TRIANGLE, G.LINESEGMENT] and conv.ignored == 3
    # the bounds are strict: area > minimum, aspect < maximum
    conv = KG.KGBEMMeshConverter()
    conv.SetMinimumArea(0.5)
    conv.VisitSurface(KG.KGSurface(mesh=[good]))
    assert conv.container.size() == 0


def test_container_group_major_order():
    """KSurfaceContainer::push_back: (boundary, shape) groups in first-seen order."""
    tri = lambda z: KG.KGMeshTriangle([0, 0, z], [1, 0, z], [0, 1, z])
    rect = KG.KGMeshRectangle([0, 0, 5], [1, 0, 5], [1, 1, 5], [0, 1, 5])
    wire = KG.KGMeshWire([0, 0, 7], [0, 0, 8], 0.01)
    conv = KG.KGBEMMeshConverter()
    conv.VisitSurface(KG.KGSurface(mesh=[tri(1), rect, tri(2)], bem=KG.KGBEMData("dirichlet", 1.0)))
    conv.VisitSurface(KG.KGSurface(mesh=[tri(3)], bem=KG.KGBEMData("neumann", flux=2.0)))
    conv.VisitSurface(KG.KGSurface(mesh=[wire, tri(4)], bem=KG.KGBEMData("dirichlet", 3.0)))
    mesh = conv.container.mesh()
    # (D, tri): z = 1, 2, 4; (D, rect); (N, tri): z = 3; (D, line)
    assert list(mesh.shapes) == [G.TRIANGLE] * 3 + [G.RECTANGLE, G.TRIANGLE, G.LINESEGMENT]
    assert list(mesh.recs[:3, 4]) == [1.0, 2.0, 4.0] and mesh.recs[4, 4] == 3.0
    assert list(mesh.bc_type) == [0, 0, 0, 0, 1, 0]
    assert list(mesh.bc_value) == [1.0, 1.0, 3.0, 1.0, 0.0, 3.0] and mesh.bc_flux[4] == 2.0
    # the rectangle record: sides P1 - P0 and P3 - P0
    assert np.allclose(mesh.recs[3], [1, 1, 0, 0, 5, 1, 0, 0, 0, 1, 0])


def test_axial_mesh_converter():
    conv = KG.KGBEMAxialMeshConverter()
    loop = KG.KGAxialMeshLoop([0.5, 1.0], [1.0, 1.5])  # (z, r)
    ring = KG.KGAxialMeshRing(0.01, [0.25, 2.0])
    # a surface shifted along the axis: same axis; z measured from the converter's origin
    assert conv.VisitSurface(KG.KGSurface(axial_mesh=[loop, ring], origin=(0, 0, 2.0),
                                          bem=KG.KGBEMData("dirichlet", -5.0)))
    mesh = conv.container.mesh()
    assert list(mesh.shapes) == [G.CONICSECTION, G.RING]
    assert np.array_equal(mesh.recs[0, :6], [1.0, 0.0, 2.5, 1.5, 0.0, 3.0])
    assert np.array_equal(mesh.recs[1, :3], [2.0, 0.0, 2.25])
    assert np.allclose(mesh.recs[0], G.conic_records(1.0, 2.5, 1.5, 3.0)[0])
    # off-axis or tilted surfaces are not coaxial: not added
    n = conv.container.size()
    assert not conv.VisitSurface(KG.KGSurface(axial_mesh=[loop], origin=(0.1, 0, 0)))
    assert not conv.VisitSurface(KG.KGSurface(axial_mesh=[loop], z_axis=(0, 0.1, 1.0)))
    assert conv.container.size() == n
    # within the 1e-5 tolerance of KAxis::EqualTo: coaxial
    assert conv.VisitSurface(KG.KGSurface(axial_mesh=[loop], origin=(1e-6, 0, 0)))


def test_discrete_rotational_groups():
    tri = KG.KGMeshTriangle([1.0, 0.0, 0.0], [1.5, 0.1, 0.0], [1.2, 0.0, 0.4])
    conv = KG.KGBEMDiscreteRotationalMeshConverter()
    plain = KG.KGBEMMeshConverter(conv.container)
    plain.VisitSurface(KG.KGSurface(mesh=[KG.KGMeshTriangle([0, 0, 3], [1, 0, 3], [0, 1, 3])]))
    assert conv.VisitSurface(KG.KGSurface(discrete_rotational_mesh=[KG.KGDiscreteRotationalMeshElement(tri, 6)],
                                          bem=KG.KGBEMData("dirichlet", 2.0)))
    mesh = conv.container.mesh()
    # a plain triangle (a group of one) and a group of 6 rotated copies
    assert mesh.n == 2 and list(mesh.group_start) == [0, 1, 7]
    assert list(mesh.bc_value) == [0.0, 2.0]
    rec0 = mesh.recs[1]
    for k in range(6):
        R = rot([0, 0, 1], k * math.pi / 3)
        exp = np.concatenate([rec0[:2], R @ rec0[2:5], R @ rec0[5:8], R @ rec0[8:11]])
        assert np.abs(mesh.recs[1 + k] - exp).max() < 1e-14, k
    # the rotation is right-handed about +z: (1, 0, 0) goes to (cos 60, sin 60, 0)
    assert np.allclose(mesh.recs[2, 2:5], [0.5, math.sqrt(3) / 2, 0.0])
    # a wire group rotates both end points
    conv2 = KG.KGBEMDiscreteRotationalMeshConverter()
    conv2.VisitSurface(KG.KGSurface(discrete_rotational_mesh=[
        KG.KGDiscreteRotationalMeshElement(KG.KGMeshWire([1, 0, 0], [1, 0, 1], 0.01), 4)]))
    m2 = conv2.container.mesh()
    assert np.allclose(m2.recs[1, :7], [0, 1, 0, 0, 1, 1, 0.01], atol=1e-15)
