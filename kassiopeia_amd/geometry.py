"""Element records and synthetic meshes for the electrostatic BEM path.

Element records follow KEMField's shape data (11 doubles per element):
triangles and rectangles ``[A, B, P0x, P0y, P0z, N1x, N1y, N1z, N2x, N2y, N2z]``
(``KEMField/Source/Surfaces/include/KTriangle.hh:9-131``, ``KRectangle.hh``), line
segments ``[P0x, P0y, P0z, P1x, P1y, P1z, diameter, 0, 0, 0, 0]``
(``KEMField/Source/Surfaces/include/KLineSegment.hh``).

The meshes are the workloads named by the benchmark configurations:

* ``cube_mesh``: unit cube, each face an n x n grid of squares split into 2 triangles;
* ``coax_mesh``: two coaxial cylinders (r = 0.5 m, 1.0 m, length 4 m) with annular end
  caps, tangentially jittered interior vertices;
* ``ring_stack_mesh``: an axisymmetric stack of annular ring electrodes;
* ``ring_bore_mesh``: coaxial ring electrodes plus an outer cylinder;
* ``sphere_mesh``: subdivided icosahedron on the unit sphere.
"""
from __future__ import annotations

import dataclasses

import numpy as np

TRIANGLE, RECTANGLE, LINESEGMENT, CONICSECTION, RING = 0, 1, 2, 3, 4
DIRICHLET, NEUMANN = 0, 1
REC = 11


@dataclasses.dataclass
class Mesh:
    """A surface container: elements in KSurfaceContainer global order.  With group_start the
    unknowns are symmetry groups (KSymmetryGroup): shapes / recs list their members, group g
    holding members group_start[g] .. group_start[g+1]-1, and the per-unknown arrays (bc_type,
    bc_value, bc_flux, group) have one entry per group."""

    shapes: np.ndarray  # int32[members]
    recs: np.ndarray  # float64[members, 11]
    bc_type: np.ndarray  # int32[n]  (0 Dirichlet, 1 Neumann)
    bc_value: np.ndarray  # float64[n]  (Dirichlet voltage)
    bc_flux: np.ndarray  # float64[n]  (Neumann normal boundary flux, eps ratio)
    group: np.ndarray  # int32[n]  (electrode id)
    group_start: np.ndarray = None  # int64[n+1] (symmetry groups) or None

    @property
    def n(self) -> int:
        return int(self.bc_type.shape[0])

    def _first(self, a):
        return a if self.group_start is None else a[self.group_start[:-1]]

    def areas(self) -> np.ndarray:
        a = np.empty(self.shapes.shape[0])
        t = self.shapes == TRIANGLE
        r = self.recs
        cr = np.cross(r[:, 5:8], r[:, 8:11])
        a[t] = 0.5 * r[t, 0] * r[t, 1] * np.sqrt(cr[t, 0] ** 2 + cr[t, 1] ** 2 + cr[t, 2] ** 2)
        q = self.shapes == RECTANGLE
        a[q] = r[q, 0] * r[q, 1]
        ln = self.shapes == LINESEGMENT
        if ln.any():
            d = r[ln, 0:3] - r[ln, 3:6]
            a[ln] = np.pi * r[ln, 6] * np.sqrt(d[:, 0] ** 2 + d[:, 1] ** 2 + d[:, 2] ** 2)
        # KConicSection::Area (KConicSection.cc:24-28), KRing::Area (KRing.cc:20-23)
        cs = self.shapes == CONICSECTION
        a[cs] = np.pi * (r[cs, 0] + r[cs, 3]) * np.sqrt((r[cs, 0] - r[cs, 3]) ** 2 + (r[cs, 2] - r[cs, 5]) ** 2)
        rg = self.shapes == RING
        a[rg] = 2.0 * np.pi * r[rg, 0]
        return self._first(a)  # a group's area is its first member's (KSymmetryGroup.hh:44-47)

    def centroids(self) -> np.ndarray:
        r = self.recs
        c = np.empty((self.shapes.shape[0], 3))
        t = self.shapes == TRIANGLE
        c[t] = r[t, 2:5] + (r[t, 0:1] * r[t, 5:8] + r[t, 1:2] * r[t, 8:11]) / 3.0
        q = self.shapes == RECTANGLE
        c[q] = r[q, 2:5] + r[q, 0:1] * r[q, 5:8] * 0.5 + r[q, 1:2] * r[q, 8:11] * 0.5
        ln = (self.shapes == LINESEGMENT) | (self.shapes == CONICSECTION)
        c[ln] = (r[ln, 0:3] + r[ln, 3:6]) * 0.5
        rg = self.shapes == RING
        c[rg] = r[rg, 0:3]
        return self._first(c)

    def subset(self, idx) -> "Mesh":
        if self.group_start is not None:
            raise ValueError("subset of a mesh with symmetry groups")
        idx = np.asarray(idx)
        return Mesh(self.shapes[idx], self.recs[idx], self.bc_type[idx], self.bc_value[idx],
                    self.bc_flux[idx], self.group[idx])

    def with_voltages(self, group_volts) -> "Mesh":
        """Copy with Dirichlet values set per electrode group."""
        v = np.asarray(group_volts, dtype=np.float64)[self.group]
        return dataclasses.replace(self, bc_value=np.ascontiguousarray(v))


def _norm(v):
    # evaluation order of KThreeVector::Magnitude: sqrt((x*x + y*y) + z*z)
    return np.sqrt(v[..., 0] * v[..., 0] + v[..., 1] * v[..., 1] + v[..., 2] * v[..., 2])


def records_from_points(p0, p1, p2) -> np.ndarray:
    """Vectorised KTriangle/KRectangle::SetValues(p0, p1, p2) (KTriangle.cc:18-28):
    N1 = unit(p1 - p0), A = |p1 - p0|, N2 = unit(p2 - p0), B = |p2 - p0|."""
    p0 = np.asarray(p0, dtype=np.float64).reshape(-1, 3)
    p1 = np.asarray(p1, dtype=np.float64).reshape(-1, 3)
    p2 = np.asarray(p2, dtype=np.float64).reshape(-1, 3)
    n1 = p1 - p0
    n2 = p2 - p0
    a = _norm(n1)
    b = _norm(n2)
    rec = np.empty((p0.shape[0], REC))
    rec[:, 0] = a
    rec[:, 1] = b
    rec[:, 2:5] = p0
    rec[:, 5:8] = n1 / a[:, None]
    rec[:, 8:11] = n2 / b[:, None]
    return rec


def line_records(p0, p1, diameter) -> np.ndarray:
    p0 = np.asarray(p0, dtype=np.float64).reshape(-1, 3)
    p1 = np.asarray(p1, dtype=np.float64).reshape(-1, 3)
    rec = np.zeros((p0.shape[0], REC))
    rec[:, 0:3] = p0
    rec[:, 3:6] = p1
    rec[:, 6] = diameter
    return rec


def reflected_records(recs, shapes, point, normal) -> np.ndarray:
    """Triangle / rectangle records reflected through the plane (point, unit normal): P0 and the
    side directions N1, N2 mirrored, the side lengths kept (the member a reflection adds to a
    symmetry group, KSymmetryGroup::AddReflectionThroughPlane)."""
    recs = np.array(recs, dtype=np.float64, copy=True)
    if np.any((np.asarray(shapes) != TRIANGLE) & (np.asarray(shapes) != RECTANGLE)):
        raise ValueError("reflected_records: triangles and rectangles only")
    o, nrm = np.asarray(point, dtype=np.float64), np.asarray(normal, dtype=np.float64)
    nrm = nrm / np.linalg.norm(nrm)
    recs[:, 2:5] -= 2.0 * ((recs[:, 2:5] - o) @ nrm)[:, None] * nrm
    for k in (5, 8):
        recs[:, k:k + 3] -= 2.0 * (recs[:, k:k + 3] @ nrm)[:, None] * nrm
    return recs


def mirror_groups(mesh: Mesh, point, normal) -> Mesh:
    """Each element of `mesh` grouped with its reflection through the plane: a mesh of
    mesh.n symmetry groups of two members (element, then its mirror image)."""
    if mesh.group_start is not None:
        raise ValueError("mirror_groups: the mesh is grouped already")
    refl = reflected_records(mesh.recs, mesh.shapes, point, normal)
    n = mesh.n
    recs = np.empty((2 * n, mesh.recs.shape[1]))
    recs[0::2], recs[1::2] = mesh.recs, refl
    return Mesh(np.repeat(mesh.shapes, 2).astype(np.int32), recs, mesh.bc_type.copy(), mesh.bc_value.copy(),
                mesh.bc_flux.copy(), mesh.group.copy(), np.arange(0, 2 * n + 1, 2, dtype=np.int64))


def conic_records(r0, z0, r1, z1) -> np.ndarray:
    """KConicSection::SetValues(r0, z0, r1, z1) (KConicSection.cc:13-22): P0 = (r0, 0, z0),
    P1 = (r1, 0, z1); the section is the surface of revolution of P0-P1 about the z axis."""
    r0, z0, r1, z1 = (np.atleast_1d(np.asarray(v, dtype=np.float64)) for v in (r0, z0, r1, z1))
    rec = np.zeros((r0.shape[0], REC))
    rec[:, 0], rec[:, 2], rec[:, 3], rec[:, 5] = r0, z0, r1, z1
    return rec


def ring_records(r, z) -> np.ndarray:
    """KRing::SetValues(r, z) (KRing.cc:14-18): P = (r, 0, z), a charged circle about the z axis."""
    r, z = (np.atleast_1d(np.asarray(v, dtype=np.float64)) for v in (r, z))
    rec = np.zeros((r.shape[0], REC))
    rec[:, 0], rec[:, 2] = r, z
    return rec


def disk_mesh(scale: int = 128, power: float = 1.5, volts: float = 1.0) -> Mesh:
    """Unit disk of `scale` conic sections from r = 1 inwards, graded towards the rim and the
    centre as the reference's disk-capacitance test builds it (UnitTest/KEMField/Fields/src/
    DiskCapacitanceTest.cc:63-86, 121-145)."""
    n = 2 * scale
    seg = np.zeros(n)
    mid = 1.0
    for i in range(n // 2):
        a, b = (i / (n / 2.0)) ** power, ((i + 1) / (n / 2.0)) ** power
        seg[i] = seg[n - (i + 1)] = mid * (b - a)
    r = [1.0]
    for i in range(scale):
        r.append(r[-1] - seg[i])
    r = np.array(r)
    # the running subtraction ends a few ulps from 0 (-2.8e-16 for 128 sections); a negative
    # radius makes the radial-field integrand of the reference's ring integrator the square root
    # of a negative number, so the centre is closed at r = 0
    r[-1] = 0.0
    return _mesh(conic_records(r[:-1], 0.0, r[1:], 0.0), CONICSECTION, volts, 0)


def _mesh(recs, shape, volts, group, bc_type=None, flux=None) -> Mesh:
    n = recs.shape[0]
    return Mesh(
        shapes=np.full(n, shape, dtype=np.int32),
        recs=np.ascontiguousarray(recs),
        bc_type=np.zeros(n, dtype=np.int32) if bc_type is None else np.asarray(bc_type, np.int32),
        bc_value=np.ascontiguousarray(np.asarray(volts, dtype=np.float64) * np.ones(n)),
        bc_flux=np.zeros(n) if flux is None else np.asarray(flux, np.float64),
        group=np.asarray(group, dtype=np.int32) * np.ones(n, dtype=np.int32),
    )


def read_triangle_file(path) -> Mesh:
    """KEMField's triangle input cards (the sphericalCapacitorFiles format), read as
    ReadInTriangles does (Applications/Calculation/ComputeSphericalCapacitor.cc:1403-1476):
    '#' starts a comment, a one-token line is the count, every other line holds 4 integers
    (index, group, unused, replicas) and 15 doubles d = P0 (3), N1 (3), N2 (3), A, B, potential,
    permittivity above, permittivity below.  A triangle is Dirichlet at the potential when
    |below / above| < 1e-10, else Neumann with normal boundary flux below / above.  The shape
    record is taken as written ([A, B, P0, N1, N2], no renormalisation).  A path ending in .gz is
    read through gzip."""
    import gzip

    opener = gzip.open if str(path).endswith(".gz") else open
    rows = []
    with opener(path, "rt") as f:
        for line in f:
            tok = line.split("#", 1)[0].split()
            if len(tok) < 19:
                continue
            rows.append([float(v) for v in tok[4:19]])
    d = np.array(rows, dtype=np.float64).reshape(-1, 15)
    recs = np.concatenate([d[:, 9:11], d[:, 0:9]], axis=1)
    neumann = np.abs(d[:, 13] / d[:, 12]) >= 1e-10
    return _mesh(recs, TRIANGLE, np.where(neumann, 0.0, d[:, 11]), 0, bc_type=neumann.astype(np.int32),
                 flux=np.where(neumann, d[:, 13] / d[:, 12], 0.0))


def concat(*meshes: Mesh) -> Mesh:
    """Containers one after another (meshes without symmetry groups)."""
    if any(m.group_start is not None for m in meshes):
        raise ValueError("concat of meshes with symmetry groups")
    return Mesh(*[np.concatenate([getattr(m, f.name) for m in meshes])
                  for f in dataclasses.fields(Mesh) if f.name != "group_start"])


def _quad_grid_triangles(P):
    """Split a (nu+1) x (nv+1) grid of vertices P[i, j] into 2 triangles per cell, with
    (p1 - p0) x (p2 - p0) along the grid's (u x v) orientation."""
    a = P[:-1, :-1].reshape(-1, 3)
    b = P[1:, :-1].reshape(-1, 3)
    c = P[1:, 1:].reshape(-1, 3)
    d = P[:-1, 1:].reshape(-1, 3)
    p0 = np.concatenate([a, c])
    p1 = np.concatenate([b, d])
    p2 = np.concatenate([c, a])
    # interleave so the two triangles of one cell are adjacent in container order
    n = a.shape[0]
    order = np.stack([np.arange(n), np.arange(n) + n], axis=1).reshape(-1)
    return p0[order], p1[order], p2[order]


def cube_mesh(n: int = 12, edge: float = 1.0, volts: float = 1.0) -> Mesh:
    """Cube centred at the origin; each face an n x n grid of squares, each square split
    along a diagonal into two triangles (6*2*n*n triangles), outward normals."""
    h = edge / 2.0
    s = np.linspace(-h, h, n + 1)
    U, V = np.meshgrid(s, s, indexing="ij")
    faces = []
    for axis in range(3):
        for sign in (1.0, -1.0):
            P = np.zeros((n + 1, n + 1, 3))
            u_ax, v_ax = (axis + 1) % 3, (axis + 2) % 3
            if sign < 0:
                u_ax, v_ax = v_ax, u_ax
            P[..., axis] = sign * h
            P[..., u_ax] = U
            P[..., v_ax] = V
            faces.append(records_from_points(*_quad_grid_triangles(P)))
    recs = np.concatenate(faces)
    return _mesh(recs, TRIANGLE, volts, 0)


def cube_rect_mesh(n: int = 12, edge: float = 1.0, volts: float = 1.0) -> Mesh:
    """Cube with each face an n x n grid of rectangles (KRectangle records)."""
    h = edge / 2.0
    s = np.linspace(-h, h, n + 1)
    recs = []
    for axis in range(3):
        for sign in (1.0, -1.0):
            u_ax, v_ax = (axis + 1) % 3, (axis + 2) % 3
            if sign < 0:
                u_ax, v_ax = v_ax, u_ax
            I, J = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
            p0 = np.zeros((n * n, 3))
            p1 = np.zeros((n * n, 3))
            p2 = np.zeros((n * n, 3))
            for p in (p0, p1, p2):
                p[:, axis] = sign * h
            p0[:, u_ax] = s[I.ravel()]
            p0[:, v_ax] = s[J.ravel()]
            p1[:, u_ax] = s[I.ravel() + 1]
            p1[:, v_ax] = s[J.ravel()]
            p2[:, u_ax] = s[I.ravel()]
            p2[:, v_ax] = s[J.ravel() + 1]
            recs.append(records_from_points(p0, p1, p2))
    return _mesh(np.concatenate(recs), RECTANGLE, volts, 0)


def _jitter(rng, shape, amp):
    return rng.uniform(-amp, amp, size=shape)


def _cylinder(radius, z0, z1, nphi, nz, rng, jitter, inward=False):
    """Cylinder side r = radius, z in [z0, z1]; vertices jittered tangentially (phi, z) by
    +-jitter x local edge length except on the two rim circles."""
    phi = np.linspace(0.0, 2.0 * np.pi, nphi + 1)
    z = np.linspace(z0, z1, nz + 1)
    PH, Z = np.meshgrid(phi, z, indexing="ij")
    dphi = 2.0 * np.pi / nphi
    dz = (z1 - z0) / nz
    if jitter > 0:
        jp = _jitter(rng, PH.shape, jitter * dphi)
        jz = _jitter(rng, Z.shape, jitter * dz)
        jp[:, 0] = jp[:, -1] = 0.0
        jz[:, 0] = jz[:, -1] = 0.0
        jp[-1] = jp[0]  # seam: phi = 2 pi copies phi = 0
        jz[-1] = jz[0]
        PH = PH + jp
        Z = Z + jz
    P = np.stack([radius * np.cos(PH), radius * np.sin(PH), Z], axis=-1)
    if inward:
        P = P[::-1]
    return records_from_points(*_quad_grid_triangles(P))


def _annulus(r0, r1, z, nr, nphi, rng, jitter, up=True):
    """Flat annulus r in [r0, r1] at height z; jittered tangentially (r, phi) except on rims."""
    r = np.linspace(r0, r1, nr + 1)
    phi = np.linspace(0.0, 2.0 * np.pi, nphi + 1)
    R, PH = np.meshgrid(r, phi, indexing="ij")
    if jitter > 0:
        jr = _jitter(rng, R.shape, jitter * (r1 - r0) / nr)
        jp = _jitter(rng, PH.shape, jitter * 2.0 * np.pi / nphi)
        jr[0] = jr[-1] = 0.0
        jp[0] = jp[-1] = 0.0
        jr[:, -1] = jr[:, 0]
        jp[:, -1] = jp[:, 0]
        R = R + jr
        PH = PH + jp
    P = np.stack([R * np.cos(PH), R * np.sin(PH), np.full_like(R, z)], axis=-1)
    if not up:
        P = P[::-1]
    return records_from_points(*_quad_grid_triangles(P))


def coax_mesh(seed: int = 1234, r_in: float = 0.5, r_out: float = 1.0, length: float = 4.0,
              v_in: float = -18600.0, v_out: float = 0.0, nphi_in: int = 48, nphi_out: int = 96,
              nz: int = 62, ncap_r: int = 8, ncap_phi: int = 67, jitter: float = 0.01) -> Mesh:
    """Two coaxial cylinders with annular end caps: 2*(48*62 + 96*62 + 2*8*67) = 20,000
    triangles.  Groups: 0 = inner cylinder (v_in); 1 = outer cylinder, 2 = end caps (both
    v_out).  Interior vertices of each surface are jittered tangentially by up to 1 % of the
    local edge length (seeded)."""
    rng = np.random.default_rng(seed)
    z0, z1 = -length / 2.0, length / 2.0
    inner = _cylinder(r_in, z0, z1, nphi_in, nz, rng, jitter)
    outer = _cylinder(r_out, z0, z1, nphi_out, nz, rng, jitter, inward=True)
    cap_lo = _annulus(r_in, r_out, z0, ncap_r, ncap_phi, rng, jitter, up=True)
    cap_hi = _annulus(r_in, r_out, z1, ncap_r, ncap_phi, rng, jitter, up=False)
    return concat(_mesh(inner, TRIANGLE, v_in, 0), _mesh(outer, TRIANGLE, v_out, 1),
                  _mesh(np.concatenate([cap_lo, cap_hi]), TRIANGLE, v_out, 2))


def coax_ring_groups(mesh: Mesh, ngroups: int = 16, length: float = 4.0) -> np.ndarray:
    """Split a coax mesh into `ngroups` axial ring groups by centroid z (for multi-RHS sets)."""
    z = mesh.centroids()[:, 2]
    g = np.floor((z + length / 2.0) / length * ngroups).astype(np.int32)
    return np.clip(g, 0, ngroups - 1)


def ring_stack_mesh(nrings: int = 8, r_min: float = 0.2, r_max: float = 1.0, dz: float = 0.25,
                    width: float = 0.1, n_target: int = 5000, volts=None) -> Mesh:
    """Axisymmetric stack of flat annular ring electrodes (inner radius r_i spaced linearly in
    [r_min, r_max], radial width `width`), one ring per z-level spaced by dz, centred on z = 0.
    Triangles are distributed over rings in proportion to area, ~n_target in total."""
    radii = np.linspace(r_min, r_max, nrings)
    areas = np.pi * ((radii + width) ** 2 - radii ** 2)
    per = np.maximum(2 * 8 * 2, np.round(n_target * areas / areas.sum())).astype(int)
    parts = []
    zs = (np.arange(nrings) - (nrings - 1) / 2.0) * dz
    for k in range(nrings):
        nr = 2
        nphi = max(8, int(round(per[k] / (2 * nr))))
        recs = _annulus(radii[k], radii[k] + width, zs[k], nr, nphi, None, 0.0)
        v = 0.0 if volts is None else volts[k]
        parts.append(_mesh(recs, TRIANGLE, v, k))
    return concat(*parts)


def ring_stack_indexed(nrings: int = 8, r_min: float = 0.2, r_max: float = 1.0, dz: float = 0.25,
                       width: float = 0.1, n_target: int = 5000):
    """The ring stack of ring_stack_mesh as a shared-vertex mesh: (X (nv, 3) vertex positions,
    tri (ne, 3) vertex indices in SetValues order (p0, p1, p2), electrode (ne,) ring index).
    Each annulus is an (nr+1) x nphi vertex grid wrapping around in phi, two triangles per cell
    with normals along +z."""
    radii = np.linspace(r_min, r_max, nrings)
    areas = np.pi * ((radii + width) ** 2 - radii ** 2)
    per = np.maximum(2 * 8 * 2, np.round(n_target * areas / areas.sum())).astype(int)
    zs = (np.arange(nrings) - (nrings - 1) / 2.0) * dz
    X, tris, elec = [], [], []
    base = 0
    for k in range(nrings):
        nr = 2
        nphi = max(8, int(round(per[k] / (2 * nr))))
        r = np.linspace(radii[k], radii[k] + width, nr + 1)
        ph = 2.0 * np.pi * np.arange(nphi) / nphi
        R, PH = np.meshgrid(r, ph, indexing="ij")
        X.append(np.stack([R * np.cos(PH), R * np.sin(PH), np.full_like(R, zs[k])], -1).reshape(-1, 3))
        ir, ip = np.meshgrid(np.arange(nr), np.arange(nphi), indexing="ij")
        ir, ip = ir.reshape(-1), ip.reshape(-1)
        a = base + ir * nphi + ip
        b = base + (ir + 1) * nphi + ip
        c = base + (ir + 1) * nphi + (ip + 1) % nphi
        d = base + ir * nphi + (ip + 1) % nphi
        t = np.stack([np.stack([a, b, c], 1), np.stack([c, d, a], 1)], 1).reshape(-1, 3)
        tris.append(t)
        elec.append(np.full(t.shape[0], k))
        base += (nr + 1) * nphi
    return np.concatenate(X), np.concatenate(tris), np.concatenate(elec)


def ring_bore_mesh(nrings: int = 32, r_ring: float = 1.9, ring_width: float = 0.05,
                   z_span: float = 8.0, r_outer: float = 2.0, n_target: int = 100000, volts=None,
                   seed: int = 99) -> Mesh:
    """Coaxial ring electrodes (short cylinders of radius r_ring and axial width ring_width,
    evenly spaced over z_span) inside a grounded outer cylinder of radius r_outer: exactly
    n_target triangles (even), split between rings and the outer cylinder by area.  The outer
    cylinder takes an nz x nphi grid; the rings share the remaining cells, one axial row each
    (rings this narrow get one), the first (remainder % nrings) rings one azimuthal cell more."""
    if n_target % 2:
        raise ValueError("ring_bore_mesh: n_target must be even (two triangles per grid cell)")
    rng = np.random.default_rng(seed)
    zc = (np.arange(nrings) - (nrings - 1) / 2.0) * (z_span / nrings)
    a_ring = 2 * np.pi * r_ring * ring_width
    a_out = 2 * np.pi * r_outer * (z_span + 0.5)
    total = nrings * a_ring + a_out
    cells = n_target // 2
    n_out = cells - nrings * (int(n_target * a_ring / total) // 2)
    nz = max(1, int(round(np.sqrt(n_out * (z_span + 0.5) / (2 * np.pi * r_outer)))))
    nphi_out = max(8, n_out // nz)
    left = cells - nz * nphi_out  # cells of the rings
    if left < 8 * nrings:
        raise ValueError("ring_bore_mesh: n_target too small for the rings")
    parts = []
    for k in range(nrings):
        nphi = left // nrings + (1 if k < left % nrings else 0)
        recs = _cylinder(r_ring, zc[k] - ring_width / 2, zc[k] + ring_width / 2, nphi, 1, rng, 0.0)
        v = 0.0 if volts is None else volts[k]
        parts.append(_mesh(recs, TRIANGLE, v, k))
    recs = _cylinder(r_outer, -(z_span + 0.5) / 2, (z_span + 0.5) / 2, nphi_out, nz, rng, 0.0, inward=True)
    parts.append(_mesh(recs, TRIANGLE, 0.0, nrings))
    mesh = concat(*parts)
    assert mesh.n == n_target
    return mesh


def sphere_mesh(subdiv: int = 3, radius: float = 1.0, volts: float = 1.0) -> Mesh:
    """Icosahedron subdivided `subdiv` times and projected to the sphere (20*4**subdiv
    triangles), outward normals."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    V = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    F = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in V]
    faces = list(F)
    for _ in range(subdiv):
        cache = {}
        nf = []

        def mid(i, j):
            key = (min(i, j), max(i, j))
            if key not in cache:
                m = verts[i] + verts[j]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    Vv = np.array(verts) * radius
    Fa = np.array(faces)
    return _mesh(records_from_points(Vv[Fa[:, 0]], Vv[Fa[:, 1]], Vv[Fa[:, 2]]), TRIANGLE, volts, 0)


# ------------------------------------------------------------------ field points
def points_shell(m: int, r_min: float, r_max: float, seed: int) -> np.ndarray:
    """m points uniform in volume in the spherical shell r_min <= r <= r_max."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(r_min ** 3, r_max ** 3, m) ** (1.0 / 3.0)
    v = rng.normal(size=(m, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    return np.ascontiguousarray(v * u[:, None])


def points_annulus(m: int, r_min: float, r_max: float, z_half: float, seed: int) -> np.ndarray:
    """m points uniform in volume in r_min < r < r_max, |z| < z_half."""
    rng = np.random.default_rng(seed)
    r = np.sqrt(rng.uniform(r_min ** 2, r_max ** 2, m))
    ph = rng.uniform(0.0, 2.0 * np.pi, m)
    z = rng.uniform(-z_half, z_half, m)
    return np.ascontiguousarray(np.stack([r * np.cos(ph), r * np.sin(ph), z], axis=1))


def points_axis(m: int, z_min: float, z_max: float) -> np.ndarray:
    z = np.linspace(z_min, z_max, m)
    return np.ascontiguousarray(np.stack([np.zeros(m), np.zeros(m), z], axis=1))
