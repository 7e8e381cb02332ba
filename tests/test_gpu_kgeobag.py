"""Containers imported from KGeoBag-shaped meshes (kassiopeia_amd/kgeobag.py) solved on the GPU:
a cube of rectangles converted from six face surfaces in frames of their own has the
capacitance of the cube built directly (and the oracle's matrix), and a discrete rotational
mesh (a plate of 3 x 3 rectangles repeated six times about z) solved as symmetry groups matches
the explicit container of all six plates, every plate's charge density equal to its group's."""
import math

import numpy as np
import pytest

from kassiopeia_amd import geometry as G
from kassiopeia_amd import kgeobag as KG

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from kassiopeia_amd import kemfield as K

    c = K.Context(0)
    c.set_integrator("numeric")
    yield c
    c.close()


def rot(axis, angle):
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(angle) * K + (1 - math.cos(angle)) * K @ K


def solve(ctx, mesh):
    from kassiopeia_amd import kemfield as K

    cont = K.KSurfaceContainer(ctx, mesh)
    A = K.KBoundaryIntegralMatrix(cont, fp64=True)
    x, st, rc = A.solve(cont.rhs(), tolerance=1e-12)
    assert rc == 0
    Ah = A.to_numpy()
    A.close()
    cont.close()
    return x, Ah


def test_rectangle_cube_via_converter(ctx, oracle):
    m = G.cube_rect_mesh(6)
    per = m.n // 6
    rng = np.random.default_rng(11)
    conv = KG.KGBEMMeshConverter()
    for f in range(6):
        R, o = rot(rng.normal(size=3), rng.uniform(0, 2 * math.pi)), rng.normal(size=3)
        r = m.recs[f * per:(f + 1) * per]
        p0 = r[:, 2:5]
        p1 = p0 + r[:, 5:8] * r[:, 0:1]
        p3 = p0 + r[:, 8:11] * r[:, 1:2]
        p2 = p1 + (p3 - p0)
        loc = [(P - o) @ R for P in (p0, p1, p2, p3)]
        elems = [KG.KGMeshRectangle(loc[0][k], loc[1][k], loc[2][k], loc[3][k]) for k in range(per)]
        conv.VisitSurface(KG.KGSurface(mesh=elems, bem=KG.KGBEMData("dirichlet", 1.0), origin=o, x_axis=R[:, 0],
                                       y_axis=R[:, 1], z_axis=R[:, 2]))
    mc = conv.container.mesh()
    assert mc.n == m.n and np.abs(mc.recs - m.recs).max() < 1e-14
    x0, _ = solve(ctx, m)
    x1, A1 = solve(ctx, mc)
    Aref = oracle.assemble(mc.shapes, mc.recs)
    assert (np.abs(A1 - Aref) / np.abs(Aref)).max() < 1e-12
    C0, C1 = (m.areas() * x0).sum(), (mc.areas() * x1).sum()
    print(f"cube of rectangles: C/(4 pi eps0) {C1 / (4 * math.pi * 8.8541878128e-12):.9f}", flush=True)
    assert abs(C1 - C0) <= 1e-10 * abs(C0)


def test_discrete_rotational_plates(ctx):
    # a radial plate in the x-z plane: r in [1, 1.5], z in [-0.5, 0.5], 3 x 3 rectangles
    r = np.linspace(1.0, 1.5, 4)
    z = np.linspace(-0.5, 0.5, 4)
    cells = [KG.KGMeshRectangle([r[i], 0, z[j]], [r[i + 1], 0, z[j]], [r[i + 1], 0, z[j + 1]], [r[i], 0, z[j + 1]])
             for i in range(3) for j in range(3)]
    conv = KG.KGBEMDiscreteRotationalMeshConverter()
    assert conv.VisitSurface(KG.KGSurface(discrete_rotational_mesh=[KG.KGDiscreteRotationalMeshElement(c, 6)
                                                                    for c in cells],
                                          bem=KG.KGBEMData("dirichlet", 1.0)))
    mg = conv.container.mesh()
    assert mg.n == 9 and list(mg.group_start) == list(range(0, 55, 6))
    nm = mg.recs.shape[0]
    plain = G.Mesh(mg.shapes, mg.recs, np.zeros(nm, np.int32), np.ones(nm), np.zeros(nm), np.zeros(nm, np.int32))
    xg, _ = solve(ctx, mg)
    xp, _ = solve(ctx, plain)
    xp = xp.reshape(9, 6)
    print(f"plates: group sigma {xg[:3]}, spread over the six plates {np.ptp(xp, axis=1).max():.2e}", flush=True)
    assert np.abs(xp - xg[:, None]).max() <= 1e-9 * np.abs(xg).max()
