"""ctypes loader for the CPU FP64 oracle (oracle/libkem_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` as the checker.  The product path
(``kassiopeia_amd``) never imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libkem_oracle.so")

TRIANGLE, RECTANGLE, LINESEGMENT, CONICSECTION, RING = 0, 1, 2, 3, 4
NUMERIC, RWG, ANALYTIC, REFERENCE = 0, 1, 2, 3
DIRICHLET, NEUMANN = 0, 1
REC = 11


def build():
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def _load():
    if not os.path.exists(_LIB):
        build()
    lib = ctypes.CDLL(_LIB)
    d, i32, i64 = ctypes.c_double, ctypes.c_int32, ctypes.c_int64
    P = ctypes.c_void_p
    sig = {
        "ko_pi": (d, []),
        "ko_eps0": (d, []),
        "ko_one_over_4pi_eps0": (d, []),
        "ko_set_from_points": (None, [P, P, P, P]),
        "ko_area": (d, [ctypes.c_int, P]),
        "ko_centroid": (None, [ctypes.c_int, P, P]),
        "ko_normal": (None, [ctypes.c_int, P, P]),
        "ko_tri_nodes": (ctypes.c_int, [ctypes.c_int, P, P, P]),
        "ko_rect_nodes": (ctypes.c_int, [ctypes.c_int, P, P, P]),
        "ko_tri_cubature": (None, [ctypes.c_int, P, P, P]),
        "ko_rect_cubature": (None, [ctypes.c_int, P, P, P]),
        "ko_solid_angle_tri": (d, [P, P]),
        "ko_solid_angle_rect": (d, [P, P]),
        "ko_tri_rwg": (None, [P, P, P]),
        "ko_rect_rwg": (None, [P, P, P]),
        "ko_line_quadrature": (None, [ctypes.c_int, P, P, P]),
        "ko_line_analytic": (None, [P, P, P]),
        "ko_tri_analytic": (None, [P, P, P]),
        "ko_rect_analytic": (None, [P, P, P]),
        "ko_tri_biquad": (None, [P, P, P]),
        "ko_rect_biquad": (None, [P, P, P]),
        "ko_line_256": (None, [P, P, P]),
        "ko_tri_numeric": (ctypes.c_int, [P, P, P]),
        "ko_rect_numeric": (ctypes.c_int, [P, P, P]),
        "ko_line_numeric": (ctypes.c_int, [P, P, P]),
        "ko_element_eval": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, P, P, P]),
        "ko_assemble": (None, [ctypes.c_int, i64, P, P, P, P, P]),
        "ko_matrix_element": (d, [ctypes.c_int, i64, i64, P, P, P, P]),
        "ko_rhs": (None, [i64, P, P, P]),
        "ko_field": (None, [ctypes.c_int, i64, P, P, P, i64, P, P, P]),
        "ko_field_mt": (None, [ctypes.c_int, ctypes.c_int, i64, P, P, P, i64, P, P, P]),
        "ko_gmres": (ctypes.c_int, [i64, P, P, P, d, ctypes.c_int, ctypes.c_int, P, P]),
        "ko_gmres_ortho": (ctypes.c_int, [i64, P, P, P, d, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P]),
        "ko_robin_hood": (i64, [i64, P, P, P, d, i64, i64, P]),
        "ko_gauss": (ctypes.c_int, [i64, P, P, P]),
        "ko_assemble_groups": (None, [ctypes.c_int, i64, P, P, P, P, P, P]),
        "ko_field_groups": (None, [ctypes.c_int, i64, P, P, P, P, i64, P, P, P]),
        "ko_ell_K": (d, [d]),
        "ko_ell_E": (d, [d]),
        "ko_ell_EK": (d, [d]),
        "ko_ring": (None, [P, P, P]),
        "ko_conic": (None, [P, P, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _ = i32
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def _recs(recs):
    recs = np.ascontiguousarray(recs, dtype=np.float64)
    if recs.ndim == 1:
        recs = recs.reshape(1, -1)
    if recs.shape[1] < REC:
        recs = np.concatenate([recs, np.zeros((recs.shape[0], REC - recs.shape[1]))], axis=1)
    return np.ascontiguousarray(recs)


def K():
    return lib().ko_one_over_4pi_eps0()


def EPS0():
    return lib().ko_eps0()


def set_from_points(p0, p1, p2):
    rec = np.zeros(REC)
    lib().ko_set_from_points(_p(_f64(p0)), _p(_f64(p1)), _p(_f64(p2)), _p(rec))
    return rec


def area(shape, rec):
    return lib().ko_area(shape, _p(_recs(rec)))


def centroid(shape, rec):
    c = np.zeros(3)
    lib().ko_centroid(shape, _p(_recs(rec)), _p(c))
    return c


def normal(shape, rec):
    c = np.zeros(3)
    lib().ko_normal(shape, _p(_recs(rec)), _p(c))
    return c


def tri_nodes(n, rec):
    Q, w = np.zeros(3 * 33), np.zeros(33)
    m = lib().ko_tri_nodes(n, _p(_recs(rec)), _p(Q), _p(w))
    return Q[: 3 * m].reshape(m, 3), w[:m]


def rect_nodes(n, rec):
    Q, w = np.zeros(3 * 33), np.zeros(33)
    m = lib().ko_rect_nodes(n, _p(_recs(rec)), _p(Q), _p(w))
    return Q[: 3 * m].reshape(m, 3), w[:m]


def _eval4(fn, *args):
    out = np.zeros(4)
    fn(*args, _p(out))
    return out


def tri_cubature(n, rec, P):
    return _eval4(lib().ko_tri_cubature, n, _p(_recs(rec)), _p(_f64(P)))


def rect_cubature(n, rec, P):
    return _eval4(lib().ko_rect_cubature, n, _p(_recs(rec)), _p(_f64(P)))


def tri_rwg(rec, P):
    return _eval4(lib().ko_tri_rwg, _p(_recs(rec)), _p(_f64(P)))


def rect_rwg(rec, P):
    return _eval4(lib().ko_rect_rwg, _p(_recs(rec)), _p(_f64(P)))


def solid_angle_tri(rec, P):
    return lib().ko_solid_angle_tri(_p(_recs(rec)), _p(_f64(P)))


def solid_angle_rect(rec, P):
    return lib().ko_solid_angle_rect(_p(_recs(rec)), _p(_f64(P)))


def line_quadrature(nn, rec, P):
    return _eval4(lib().ko_line_quadrature, nn, _p(_recs(rec)), _p(_f64(P)))


def line_analytic(rec, P):
    return _eval4(lib().ko_line_analytic, _p(_recs(rec)), _p(_f64(P)))


def tri_analytic(rec, P):
    return _eval4(lib().ko_tri_analytic, _p(_recs(rec)), _p(_f64(P)))


def rect_analytic(rec, P):
    return _eval4(lib().ko_rect_analytic, _p(_recs(rec)), _p(_f64(P)))


def tri_biquad(rec, P):
    return _eval4(lib().ko_tri_biquad, _p(_recs(rec)), _p(_f64(P)))


def rect_biquad(rec, P):
    return _eval4(lib().ko_rect_biquad, _p(_recs(rec)), _p(_f64(P)))


def line_256(rec, P):
    return _eval4(lib().ko_line_256, _p(_recs(rec)), _p(_f64(P)))


def ell_K(k):
    return lib().ko_ell_K(float(k))


def ell_E(k):
    return lib().ko_ell_E(float(k))


def ell_EK(k):
    return lib().ko_ell_EK(float(k))


def ring(rec, P):
    return _eval4(lib().ko_ring, _p(_recs(rec)), _p(_f64(P)))


def conic(rec, P):
    return _eval4(lib().ko_conic, _p(_recs(rec)), _p(_f64(P)))


def element_eval(shape, rec, P, policy=NUMERIC):
    """(E, phi) of one element at one point for unit charge density; also returns the regime."""
    out = np.zeros(4)
    reg = lib().ko_element_eval(policy, shape, _p(_recs(rec)), _p(_f64(P)), _p(out))
    return out, reg


def assemble(shapes, recs, bc=None, flux=None, policy=NUMERIC):
    shapes = np.ascontiguousarray(shapes, dtype=np.int32)
    recs = _recs(recs)
    n = len(shapes)
    A = np.zeros((n, n))
    bcp = None if bc is None else _p(np.ascontiguousarray(bc, dtype=np.int32))
    flp = None if flux is None else _p(_f64(flux))
    lib().ko_assemble(policy, n, _p(shapes), _p(recs), bcp, flp, _p(A))
    return A


def assemble_groups(group_start, shapes, recs, bc=None, flux=None, policy=NUMERIC):
    """Collocation matrix over symmetry groups (members group_start[g] .. group_start[g+1]-1)."""
    gs = np.ascontiguousarray(group_start, dtype=np.int64)
    shapes = np.ascontiguousarray(shapes, dtype=np.int32)
    recs = _recs(recs)
    n = len(gs) - 1
    A = np.zeros((n, n))
    bcp = None if bc is None else _p(np.ascontiguousarray(bc, dtype=np.int32))
    flp = None if flux is None else _p(_f64(flux))
    lib().ko_assemble_groups(policy, n, _p(gs), _p(shapes), _p(recs), bcp, flp, _p(A))
    return A


def field_groups(group_start, shapes, recs, sigma, points, policy=NUMERIC):
    gs = np.ascontiguousarray(group_start, dtype=np.int64)
    shapes = np.ascontiguousarray(shapes, dtype=np.int32)
    recs = _recs(recs)
    sigma = _f64(sigma)
    pts = _f64(points).reshape(-1, 3)
    m = pts.shape[0]
    phi = np.zeros(m)
    E = np.zeros((m, 3))
    lib().ko_field_groups(policy, len(gs) - 1, _p(gs), _p(shapes), _p(recs), _p(sigma), m, _p(pts), _p(phi), _p(E))
    return phi, E


def matrix_element(i, j, shapes, recs, bc=None, flux=None, policy=NUMERIC):
    shapes = np.ascontiguousarray(shapes, dtype=np.int32)
    bcp = None if bc is None else _p(np.ascontiguousarray(bc, dtype=np.int32))
    flp = None if flux is None else _p(_f64(flux))
    return lib().ko_matrix_element(policy, i, j, _p(shapes), _p(_recs(recs)), bcp, flp)


def rhs(values, bc=None):
    values = _f64(values)
    b = np.zeros(len(values))
    bcp = None if bc is None else _p(np.ascontiguousarray(bc, dtype=np.int32))
    lib().ko_rhs(len(values), bcp, _p(values), _p(b))
    return b


def field(shapes, recs, sigma, points, policy=NUMERIC, threads=1):
    """Kahan-summed (phi, E) at points (m x 3); threads > 1 shares the points out over OpenMP
    threads (identical results)."""
    shapes = np.ascontiguousarray(shapes, dtype=np.int32)
    recs = _recs(recs)
    sigma = _f64(sigma)
    pts = _f64(points).reshape(-1, 3)
    m = pts.shape[0]
    phi = np.zeros(m)
    E = np.zeros((m, 3))
    if threads > 1:
        lib().ko_field_mt(threads, policy, len(shapes), _p(shapes), _p(recs), _p(sigma), m, _p(pts), _p(phi), _p(E))
    else:
        lib().ko_field(policy, len(shapes), _p(shapes), _p(recs), _p(sigma), m, _p(pts), _p(phi), _p(E))
    return phi, E


def gmres(A, b, tol=1e-10, maxit=10000, restart=0, x0=None, ortho="mgs"):
    """ortho: "mgs" (the reference's modified Gram-Schmidt) or "cgs2" (the HIP path's)."""
    A = _f64(A)
    b = _f64(b)
    n = len(b)
    x = np.zeros(n) if x0 is None else _f64(x0).copy()
    resid = np.zeros(1)
    hist = np.zeros(maxit)
    it = lib().ko_gmres_ortho(n, _p(A), _p(b), _p(x), tol, maxit, restart, {"mgs": 0, "cgs2": 1}[ortho], _p(resid),
                              _p(hist))
    return x, it, resid[0], hist[:it]


def robin_hood(A, b, tol=1e-10, check_interval=0, max_sub=10**9, x0=None):
    A = _f64(A)
    b = _f64(b)
    n = len(b)
    x = np.zeros(n) if x0 is None else _f64(x0).copy()
    resid = np.zeros(1)
    sub = lib().ko_robin_hood(n, _p(A), _p(b), _p(x), tol, check_interval, max_sub, _p(resid))
    return x, sub, resid[0]


def gauss(A, b):
    A = _f64(A)
    b = _f64(b)
    x = np.zeros(len(b))
    lib().ko_gauss(len(b), _p(A), _p(b), _p(x))
    return x
