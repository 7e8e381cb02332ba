"""KGeoBag mesh import: the KGBEMConverter family of KEMField's KGeoBag interface
(KEMField/Source/Interface/KGeoBag/{include/KGBEMConverter.hh, src/KGBEMConverter.cc}) over plain
mesh-element records, filling a container (``geometry.Mesh``) in KSurfaceContainer order.

A KGeoBag surface (or space) carries its frame (origin and x / y / z axes), a mesh of triangles,
rectangles and wires (KGMesh), an axial mesh of loops and rings (KGAxialMesh) or a discrete
rotational mesh (KGDiscreteRotationalMesh), and its boundary data (KGBEM: a Dirichlet potential
or a Neumann normal flux).  The converters

* map every point from the surface frame into the converter's system (``SetSystem``; the global
  frame by default): ``LocalToInternal`` (KGBEMConverter.cc:188-203), for axial meshes from the
  (z, r) plane of the surface's axis to the position (r, 0, z) of the converter's;
* keep an element when its area is above the minimum and (triangles, rectangles, wires) its
  aspect ratio below the maximum (defaults 0 and 1e100), with the areas and aspect ratios of
  KGeoBag's mesh elements (KGeoBag/Source/Extensions/Mesh/Source/KGMeshTriangle.cc:34-106,
  KGMeshRectangle.cc:43-54, KGMeshWire.cc:14-21, AxialMesh/Source/KGAxialMeshLoop.cc:14-21,
  KGAxialMeshRing.cc:11-18);
* set the KEMField shape from the element's corner points as KGeoBag returns them
  (``GetP1() = P0 + N1 A`` etc.), so the triangle of a KGMeshTriangle built from three points
  is that of the points as KGeoBag stores them;
* push each surface's shapes, triangles first, then rectangles, line segments, conic sections,
  rings and the symmetry groups (KGBEMConverter.hh:214-266), into a container that groups them
  by (boundary, shape) type in first-seen order (KSurfaceContainer::push_back,
  Surfaces/src/KSurfaceContainer.cc:36-50), so the container's global order is group-major;
* refuse an axial or discrete rotational mesh whose surface axis differs from the converter's
  (KAxis::EqualTo: directions within 1e-5 of parallel, centres within 1e-5), returning False
  as the reference warns and skips the surface;
* turn a discrete rotational element repeated n times into a symmetry group of n members, member
  j the rotation of member j - 1 by 2 pi / n about the converter's axis
  (KSymmetryGroup::AddRotationsAboutAxis, Surfaces/include/KSymmetryGroup.hh:216-239, with the
  rotation of KThreeVector_::RotateAboutAxis, Math/include/KThreeVector_KEMField.hh:61-72).

Points are 3-vectors (x, y, z); axial points are 2-vectors (z, r) as KTwoVector stores them.
The C++ path needs none of this: KEMField's own KGBEMConverter fills a KSurfaceContainer that the
library's container upload flattens.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional, Sequence

import numpy as np

from . import geometry as G

PI = 3.141592653589793  # KEMConstants::Pi / KConst::Pi()


def _v(x) -> np.ndarray:
    return np.asarray(x, dtype=np.float64).reshape(3)


def _mag(v) -> float:
    return math.sqrt(float(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]))


def _dot(a, b) -> float:
    return float(a[0] * b[0] + a[1] * b[1] + a[2] * b[2])


def _cross(a, b) -> np.ndarray:
    return np.array([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]])


# ------------------------------------------------------------------ KGeoBag mesh elements
class KGMeshTriangle:
    """P0 with unit side directions N1, N2 and side lengths A, B (KGMeshTriangle.cc:14-23)."""

    def __init__(self, p0, p1, p2):
        self.p0 = _v(p0)
        n1, n2 = _v(p1) - self.p0, _v(p2) - self.p0
        self.a, self.b = _mag(n1), _mag(n2)
        self.n1, self.n2 = n1 / self.a, n2 / self.b

    def GetP0(self):
        return self.p0

    def GetP1(self):
        return self.p0 + self.n1 * self.a

    def GetP2(self):
        return self.p0 + self.n2 * self.b

    def Area(self) -> float:
        return 0.5 * self.a * self.b * _mag(_cross(self.n1, self.n2))

    def Aspect(self) -> float:
        """Longest side over the height onto it (KGMeshTriangle.cc:38-106)."""
        p1 = self.p0 + self.a * self.n1
        p2 = self.p0 + self.b * self.n2
        c = _mag(p1 - p2)
        if self.a > self.b:
            longest, pa, pb, pc = self.a, p2, self.p0, p1
        else:
            longest, pa, pb, pc = self.b, p1, p2, self.p0
        if c > longest:
            longest, pa, pb, pc = c, self.p0, p1, p2
        v = pc - pb
        y = v / _mag(v)
        t = (_dot(pa, v) - _dot(pb, v)) / _dot(v, v)
        x = (pb + t * v) - pa
        x = x - _dot(x, y) * y
        return longest / _mag(x)


class KGMeshRectangle:
    """P0 with sides P1 - P0 (A, N1) and P3 - P0 (B, N2); P2 unused (KGMeshRectangle.cc:20-30)."""

    def __init__(self, p0, p1, p2, p3):
        self.p0 = _v(p0)
        n1, n2 = _v(p1) - self.p0, _v(p3) - self.p0
        self.a, self.b = _mag(n1), _mag(n2)
        self.n1, self.n2 = n1 / self.a, n2 / self.b

    def GetP0(self):
        return self.p0

    def GetP1(self):
        return self.p0 + self.n1 * self.a

    def GetP2(self):
        return self.p0 + self.n1 * self.a + self.n2 * self.b

    def GetP3(self):
        return self.p0 + self.n2 * self.b

    def Area(self) -> float:
        return self.a * self.b

    def Aspect(self) -> float:
        return self.a / self.b if self.a > self.b else self.b / self.a


class KGMeshWire:
    """Wire P0-P1 of the given diameter (KGMeshWire.cc:14-21)."""

    def __init__(self, p0, p1, diameter: float):
        self.p0, self.p1, self.diameter = _v(p0), _v(p1), float(diameter)

    def GetP0(self):
        return self.p0

    def GetP1(self):
        return self.p1

    def Area(self) -> float:
        return (0.5 * math.pi * self.diameter + _mag(self.p1 - self.p0)) * math.pi * self.diameter

    def Aspect(self) -> float:
        return _mag(self.p1 - self.p0) / self.diameter


class KGAxialMeshLoop:
    """Loop from (z0, r0) to (z1, r1) about the surface's axis (KGAxialMeshLoop.cc:14-21)."""

    def __init__(self, p0, p1):
        self.p0 = np.asarray(p0, dtype=np.float64).reshape(2)
        self.p1 = np.asarray(p1, dtype=np.float64).reshape(2)

    def Area(self) -> float:
        d = self.p0 - self.p1
        return PI * math.sqrt(float(d[0] * d[0] + d[1] * d[1])) * float(self.p0[1] + self.p1[1])

    def Aspect(self) -> float:
        d = self.p0 - self.p1
        return 2.0 * math.sqrt(float(d[0] * d[0] + d[1] * d[1])) / float(self.p0[1] + self.p1[1])


class KGAxialMeshRing:
    """Ring at (z, r) of line width d (KGAxialMeshRing.cc:11-18)."""

    def __init__(self, d: float, p0):
        self.d = float(d)
        self.p0 = np.asarray(p0, dtype=np.float64).reshape(2)

    def Area(self) -> float:
        return 2.0 * PI ** 2 * float(self.p0[1]) * self.d

    def Aspect(self) -> float:
        return self.d / float(self.p0[1])


@dataclasses.dataclass
class KGDiscreteRotationalMeshElement:
    """A mesh triangle, rectangle or wire repeated n times about the axis
    (KGDiscreteRotationalMeshElement.hh:37-78)."""

    element: object
    n: int

    def Element(self):
        return self.element

    def NumberOfElements(self) -> int:
        return self.n

    def Area(self) -> float:
        return self.element.Area()

    def Aspect(self) -> float:
        return self.element.Aspect()


@dataclasses.dataclass
class KGBEMData:
    """A surface's boundary: Dirichlet (KDirichletBoundary, the potential) or Neumann
    (KNeumannBoundary, the normal boundary flux)."""

    boundary: str = "dirichlet"
    value: float = 0.0  # Dirichlet potential
    flux: float = 0.0  # Neumann normal flux

    def __post_init__(self):
        if self.boundary not in ("dirichlet", "neumann"):
            raise ValueError("boundary: 'dirichlet' or 'neumann'")


@dataclasses.dataclass
class KGSurface:
    """A meshed KGeoBag surface (or space) in its own frame, with its boundary data."""

    mesh: Sequence = ()
    axial_mesh: Sequence = ()
    discrete_rotational_mesh: Sequence = ()
    bem: KGBEMData = dataclasses.field(default_factory=KGBEMData)
    origin: Sequence[float] = (0.0, 0.0, 0.0)
    x_axis: Sequence[float] = (1.0, 0.0, 0.0)
    y_axis: Sequence[float] = (0.0, 1.0, 0.0)
    z_axis: Sequence[float] = (0.0, 0.0, 1.0)
    name: str = ""


# ------------------------------------------------------------------ axis (katrin::KAxis)
class KAxis:
    """An axis through two points: direction oriented towards +z (then +y, +x), centre the
    axis point closest to the origin (Kommon/Math/Geometry/src/KAxis.cc:12-62)."""

    def __init__(self):
        self.center = np.zeros(3)
        self.direction = np.array([0.0, 0.0, 1.0])

    def SetPoints(self, a, b):
        a, b = _v(a), _v(b)
        for k in (2, 1, 0):
            if a[k] > b[k]:
                d = a - b
                break
            if a[k] < b[k]:
                d = b - a
                break
        else:
            d = None
        self.direction = np.array([0.0, 0.0, 1.0]) if d is None else d / _mag(d)
        self.center = a - _dot(a, self.direction) * self.direction

    def EqualTo(self, other: "KAxis") -> bool:
        if abs(_dot(self.direction, other.direction)) < 1.0 - 1.0e-5:
            return False
        return _mag(self.center - other.center) <= 1.0e-5


def _rotate(p, center, axis, angle, displacement: bool):
    """KThreeVector_::RotateAboutAxis: p cos + d (d.p)(1 - cos) - (p x d) sin, about the axis
    through center for positions, through the origin for directions."""
    q = p - center if displacement else p.copy()
    c, s = math.cos(angle), math.sin(angle)
    q = q * c + axis * _dot(axis, q) * (1.0 - c) - _cross(q, axis) * s
    return q + center if displacement else q


# ------------------------------------------------------------------ container builder
_SHAPE_KEYS = {"triangle": G.TRIANGLE, "rectangle": G.RECTANGLE, "line": G.LINESEGMENT,
               "conic": G.CONICSECTION, "ring": G.RING}


class ContainerBuilder:
    """KSurfaceContainer::push_back: surfaces grouped by (boundary, shape) type in first-seen
    order; a symmetry group of a shape is a shape type of its own."""

    def __init__(self):
        self._groups = {}  # (bc, shape key, symmetric) -> list of (members records, data)
        self._order = []

    def push_back(self, shape: str, members: List[np.ndarray], bem: KGBEMData, symmetric: bool = False):
        key = (bem.boundary, shape, symmetric)
        if key not in self._groups:
            self._groups[key] = []
            self._order.append(key)
        self._groups[key].append((members, bem))

    def size(self) -> int:
        return sum(len(v) for v in self._groups.values())

    def mesh(self) -> G.Mesh:
        """The container in its global order; with any symmetry group present, every surface is
        a group (a plain surface a group of one member)."""
        shapes, recs, bct, val, flux, gstart = [], [], [], [], [], [0]
        grouped = any(k[2] for k in self._order)
        for key in self._order:
            for members, bem in self._groups[key]:
                for r in members:
                    shapes.append(_SHAPE_KEYS[key[1]])
                    recs.append(r)
                gstart.append(gstart[-1] + len(members))
                bct.append(0 if bem.boundary == "dirichlet" else 1)
                val.append(bem.value if bem.boundary == "dirichlet" else 0.0)
                flux.append(bem.flux if bem.boundary == "neumann" else 0.0)
        n = len(bct)
        full = np.zeros((len(recs), 11))
        if recs:
            full[:] = np.asarray(recs)
        return G.Mesh(np.asarray(shapes, np.int32), full, np.asarray(bct, np.int32), np.asarray(val, np.float64),
                      np.asarray(flux, np.float64), np.zeros(n, np.int32),
                      np.asarray(gstart, np.int64) if grouped else None)


def _triangle_record(p0, p1, p2) -> np.ndarray:
    # KTriangle::SetValues(p0, p1, p2) (Surfaces/src/KTriangle.cc:18-28)
    return G.records_from_points(p0[None], p1[None], p2[None])[0]


def _rectangle_record(p0, p1, p3) -> np.ndarray:
    # KRectangle::SetValues(p0, p1, p2, p3) (Surfaces/src/KRectangle.cc:18-28): sides p1 - p0, p3 - p0
    n1, n2 = p1 - p0, p3 - p0
    a, b = _mag(n1), _mag(n2)
    return np.concatenate([[a, b], p0, n1 / a, n2 / b])


def _line_record(p0, p1, d) -> np.ndarray:
    # KLineSegment::SetValues(p0, p1, diameter) (Surfaces/src/KLineSegment.cc:7-12)
    return np.concatenate([p0, p1, [d], np.zeros(4)])


class KGBEMConverter:
    """The converters' common part (KGBEMConverter.cc:26-203): the converter's system, the
    element filters, the current surface's frame and the container the shapes go to."""

    def __init__(self, container: Optional[ContainerBuilder] = None):
        self.container = container if container is not None else ContainerBuilder()
        self.min_area = 0.0
        self.max_aspect = 1e100
        self.origin = np.zeros(3)
        self.x_axis, self.y_axis, self.z_axis = np.eye(3)
        self.axis = KAxis()
        self._cur = None

    def SetMinimumArea(self, a: float):
        self.min_area = float(a)

    def SetMaximumAspectRatio(self, r: float):
        self.max_aspect = float(r)

    def SetSystem(self, origin, x_axis, y_axis, z_axis):
        self.origin, self.x_axis, self.y_axis, self.z_axis = _v(origin), _v(x_axis), _v(y_axis), _v(z_axis)
        self.axis = KAxis()
        self.axis.SetPoints(self.origin, self.origin + self.z_axis)

    def _set_current(self, s: KGSurface):
        self._cur = (_v(s.origin), _v(s.x_axis), _v(s.y_axis), _v(s.z_axis))
        self._cur_axis = KAxis()
        self._cur_axis.SetPoints(self._cur[0], self._cur[0] + self._cur[3])

    def LocalToInternal(self, p) -> np.ndarray:
        o, x, y, z = self._cur
        p = np.asarray(p, dtype=np.float64)
        if p.shape == (2,):  # (z, r) about the surface's axis -> (r, 0, z)
            g = o + z * p[0]
            return np.array([p[1], 0.0, _dot(g - self.origin, self.z_axis)])
        g = o + p[0] * x + p[1] * y + p[2] * z
        d = g - self.origin
        return np.array([_dot(d, self.x_axis), _dot(d, self.y_axis), _dot(d, self.z_axis)])

    def _keep(self, e, aspect=True) -> bool:
        return e.Area() > self.min_area and (not aspect or e.Aspect() < self.max_aspect)


class KGBEMMeshConverter(KGBEMConverter):
    """Triangles, rectangles and wires of a surface's mesh (KGBEMConverter.cc:229-300); ``ignored``
    counts the elements not added."""

    def VisitSurface(self, s: KGSurface) -> bool:
        self._set_current(s)
        tri, rect, line, ignored = [], [], [], 0
        for e in s.mesh:
            if isinstance(e, KGMeshTriangle) and self._keep(e):
                tri.append(_triangle_record(self.LocalToInternal(e.GetP0()), self.LocalToInternal(e.GetP1()),
                                            self.LocalToInternal(e.GetP2())))
            elif isinstance(e, KGMeshRectangle) and self._keep(e):
                rect.append(_rectangle_record(self.LocalToInternal(e.GetP0()), self.LocalToInternal(e.GetP1()),
                                              self.LocalToInternal(e.GetP3())))
            elif isinstance(e, KGMeshWire) and self._keep(e):
                line.append(_line_record(self.LocalToInternal(e.GetP0()), self.LocalToInternal(e.GetP1()),
                                         e.diameter))
            else:  # another element type, or one the filters drop (KGBEMConverter.cc:292-294)
                ignored += 1
        for shape, recs in (("triangle", tri), ("rectangle", rect), ("line", line)):
            for r in recs:
                self.container.push_back(shape, [r], s.bem)
        self.ignored = ignored
        return True


class KGBEMAxialMeshConverter(KGBEMConverter):
    """Loops (conic sections) and rings of a surface's axial mesh (KGBEMConverter.cc:322-383);
    a surface whose axis is not the converter's is not added (returns False)."""

    def VisitSurface(self, s: KGSurface) -> bool:
        self._set_current(s)
        if not self.axis.EqualTo(self._cur_axis):
            return False
        conic, ring = [], []
        for e in s.axial_mesh:
            if isinstance(e, KGAxialMeshLoop) and self._keep(e, aspect=False):
                conic.append(np.concatenate([self.LocalToInternal(e.p0), self.LocalToInternal(e.p1), np.zeros(5)]))
            elif isinstance(e, KGAxialMeshRing) and self._keep(e, aspect=False):
                ring.append(np.concatenate([self.LocalToInternal(e.p0), np.zeros(8)]))
        for shape, recs in (("conic", conic), ("ring", ring)):
            for r in recs:
                self.container.push_back(shape, [r], s.bem)
        return True


class KGBEMDiscreteRotationalMeshConverter(KGBEMConverter):
    """Discrete rotational triangles, rectangles and wires as symmetry groups of rotations
    about the converter's axis (KGBEMConverter.cc:433-512)."""

    def VisitSurface(self, s: KGSurface) -> bool:
        self._set_current(s)
        if not self.axis.EqualTo(self._cur_axis):
            return False
        center, direction = self.axis.center, self.axis.direction
        groups = {"triangle": [], "rectangle": [], "line": []}
        for de in s.discrete_rotational_mesh:
            e = de.Element()
            if not self._keep(de):
                continue
            if isinstance(e, KGMeshTriangle):
                pts = [self.LocalToInternal(e.GetP0()), self.LocalToInternal(e.GetP1()),
                       self.LocalToInternal(e.GetP2())]
                groups["triangle"].append(_rotations("triangle", _triangle_record(*pts), de.NumberOfElements(),
                                                     center, direction))
            elif isinstance(e, KGMeshRectangle):
                pts = [self.LocalToInternal(e.GetP0()), self.LocalToInternal(e.GetP1()),
                       self.LocalToInternal(e.GetP3())]
                groups["rectangle"].append(_rotations("rectangle", _rectangle_record(*pts), de.NumberOfElements(),
                                                      center, direction))
            elif isinstance(e, KGMeshWire):
                rec = _line_record(self.LocalToInternal(e.GetP0()), self.LocalToInternal(e.GetP1()), e.diameter)
                groups["line"].append(_rotations("line", rec, de.NumberOfElements(), center, direction))
        for shape in ("triangle", "rectangle", "line"):
            for members in groups[shape]:
                self.container.push_back(shape, members, s.bem, symmetric=True)
        return True


def _rotations(shape: str, rec: np.ndarray, n: int, center, direction) -> List[np.ndarray]:
    """KSymmetryGroup::AddRotationsAboutAxis: member j = member j - 1 rotated by 2 pi / n; the
    shape's positions rotate about the axis, its directions about the origin."""
    angle = (2.0 * PI) / n
    out = [rec]
    for _ in range(1, n):
        r = out[-1].copy()
        if shape == "line":
            r[0:3] = _rotate(r[0:3], center, direction, angle, True)
            r[3:6] = _rotate(r[3:6], center, direction, angle, True)
        else:  # [A, B, P0, N1, N2]
            r[2:5] = _rotate(r[2:5], center, direction, angle, True)
            r[5:8] = _rotate(r[5:8], center, direction, angle, False)
            r[8:11] = _rotate(r[8:11], center, direction, angle, False)
        out.append(r)
    return out
