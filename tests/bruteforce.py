"""Independent high-precision potential/field of a uniformly charged planar polygon.

Test helper (checker of the oracle).  The polygon is split into sub-triangles with a common
vertex at P', the projection of the field point P onto the plane.  In polar coordinates about
P' the radial integrals are exact:

    phi :  int_0^R rho drho / sqrt(rho^2+h^2)          = sqrt(R^2+h^2) - |h|
    E_n :  h int_0^R rho drho / (rho^2+h^2)^(3/2)       = h (1/|h| - 1/sqrt(R^2+h^2))
    E_t :  int_0^R rho^2 drho / (rho^2+h^2)^(3/2)       = asinh(R/|h|) - R/sqrt(R^2+h^2)

and the remaining angular integral is taken along each edge line (t = position along the edge,
d theta = (n . w) dt / (d^2 + t^2)) with mpmath's adaptive Gauss-Legendre at 30 digits.
This shares no formula with the solid-angle + edge-logarithm closed form it checks.
"""
import mpmath as mp
import numpy as np

mp.mp.dps = 30
EPS0 = mp.mpf("8.8541878128e-12")
K = 1 / (4 * mp.pi * EPS0)


def _v(a):
    return mp.matrix([mp.mpf(float(x)) for x in a])


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def _cross(a, b):
    return mp.matrix([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]])


def _norm(a):
    return mp.sqrt(_dot(a, a))


def polygon_field(verts, P, want_field=True):
    """(Ex, Ey, Ez, phi) at P of a planar polygon (vertices in order) carrying sigma = 1."""
    V = [_v(v) for v in verts]
    P = _v(P)
    n = _cross(V[1] - V[0], V[2] - V[0])
    n = n / _norm(n)
    h = _dot(n, P - V[0])
    Pp = P - h * n
    ah = abs(h)
    phi = mp.mpf(0)
    E = mp.matrix([0, 0, 0])
    m = len(V)
    for e in range(m):
        A, B = V[e], V[(e + 1) % m]
        L = _norm(B - A)
        u = (B - A) / L
        tA = _dot(A - Pp, u)
        F = A - tA * u  # foot of P' on the edge line
        w = _cross(F - Pp, u)
        nw = _dot(n, w)
        d2 = _dot(F - Pp, F - Pp)
        if abs(nw) < mp.mpf(10) ** (-25):
            continue  # P' on the edge line: degenerate sub-triangle
        tB = tA + L

        def dth(t):
            return nw / (d2 + t * t)

        def R(t):
            return mp.sqrt(d2 + t * t)

        pts = [tA, tB] if not (tA < 0 < tB) else [tA, 0, tB]
        phi += mp.quad(lambda t: dth(t) * (mp.sqrt(R(t) ** 2 + h * h) - ah), pts)
        if want_field:
            if ah > 0:
                En = mp.quad(lambda t: dth(t) * h * (1 / ah - 1 / mp.sqrt(R(t) ** 2 + h * h)), pts)
                for k in range(3):
                    Et = mp.quad(
                        lambda t: dth(t) * ((F[k] - Pp[k]) + t * u[k]) / R(t)
                        * (mp.asinh(R(t) / ah) - R(t) / mp.sqrt(R(t) ** 2 + h * h)), pts)
                    E[k] += n[k] * En - Et
    out = np.array([float(K * E[0]), float(K * E[1]), float(K * E[2]), float(K * phi)])
    return out


def triangle_vertices(rec):
    rec = np.asarray(rec, dtype=np.float64)
    p0 = rec[2:5]
    return [p0, p0 + rec[0] * rec[5:8], p0 + rec[1] * rec[8:11]]


def rectangle_vertices(rec):
    rec = np.asarray(rec, dtype=np.float64)
    p0 = rec[2:5]
    a, b = rec[0] * rec[5:8], rec[1] * rec[8:11]
    return [p0, p0 + a, p0 + a + b, p0 + b]
