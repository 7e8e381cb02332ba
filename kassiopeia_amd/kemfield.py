"""KEMField-shaped host interface over the HIP C-ABI (include/kemhip.h).

Class and method names follow the reference so that calling code and tests read like
KEMField's (paths relative to KEMField/Source):

* ``KElectrostaticBoundaryIntegrator``  Electrostatic/Core/include/KElectrostaticBoundaryIntegrator.hh:57-149
* ``KBoundaryIntegralMatrix``           BoundaryIntegrals/Core/include/KBoundaryIntegralMatrix.hh:16-99
  (a KSquareMatrix: Dimension, operator(), Multiply, MultiplyTranspose; LinearAlgebra/Core/include/KMatrix.hh)
* ``KKrylovChargeDensitySolver``        Interface/ChargeDensitySolvers/Electric/src/KKrylovChargeDensitySolver.cc
* ``KRobinHoodChargeDensitySolver``     .../KRobinHoodChargeDensitySolver.cc
* ``KIntegratingElectrostaticFieldSolver`` Interface/FieldSolvers/Electric/src/KIntegratingElectrostaticFieldSolver.cc
* ``KElectrostaticBoundaryField``       Interface/Fields/Electric/src/KElectrostaticBoundaryField.cc

Errors: unknown integrator / solver names raise ``KEMSimpleException`` (the reference's
exception type for them); HIP failures raise ``KemError``.
"""
from __future__ import annotations

import ctypes
import weakref

import numpy as np

from . import _lib
from .geometry import Mesh

_P = ctypes.c_void_p


# precision modes of solves and block products by name (fields take "fp64" and "mixed")
PRECISIONS = {"fp64": _lib.KEM_FP64, "mixed": _lib.KEM_MIXED, "bf16": _lib.KEM_MIXED_BF16}


class KEMSimpleException(RuntimeError):
    """Raised for unknown integrator / solver names (KEMSimpleException in the reference)."""


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_P)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a if shape is None else a.reshape(shape)


class Context:
    """One HIP device and stream (kem_ctx)."""

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        h = _P()
        rc = self.lib.kem_ctx_create(device, ctypes.byref(h))
        if rc != 0:
            raise _lib.KemError(rc, "kem_ctx_create failed (no HIP device?)")
        self.h = h
        self.device = device
        # objects holding library handles that depend on this context; closed before it
        self._deps = weakref.WeakSet()

    def track(self, obj):
        self._deps.add(obj)
        return obj

    def check(self, rc):
        if rc < 0:
            msg = self.lib.kem_last_error(self.h).decode()
            if rc == _lib.KEM_E_NAME:
                raise KEMSimpleException(msg)
            raise _lib.KemError(rc, msg)
        return rc

    def set_stream(self, stream_handle):
        self.check(self.lib.kem_ctx_set_stream(self.h, _P(stream_handle) if stream_handle else None))

    def set_integrator(self, name: str):
        self.check(self.lib.kem_set_integrator(self.h, name.encode()))

    def sync(self):
        self.check(self.lib.kem_ctx_sync(self.h))

    def close(self):
        if getattr(self, "h", None):
            for d in list(self._deps):
                d.close()
            self.lib.kem_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class KSurfaceContainer:
    """Device-resident surface container (kem_mesh) built from a geometry.Mesh."""

    def __init__(self, ctx: Context, mesh: Mesh):
        self.ctx = ctx
        self.mesh = mesh
        h = _P()
        shapes = np.ascontiguousarray(mesh.shapes, np.int32)
        recs = _f64(mesh.recs)
        bc = np.ascontiguousarray(mesh.bc_type, np.int32)
        val = _f64(mesh.bc_value)
        flux = _f64(mesh.bc_flux)
        if mesh.group_start is None:
            ctx.check(ctx.lib.kem_mesh_create(ctx.h, mesh.n, _ptr(shapes), _ptr(recs), _ptr(bc), _ptr(val),
                                              _ptr(flux), ctypes.byref(h)))
        else:
            gs = np.ascontiguousarray(mesh.group_start, np.int64)
            ctx.check(ctx.lib.kem_mesh_create_groups(ctx.h, mesh.n, _ptr(gs), _ptr(shapes), _ptr(recs), _ptr(bc),
                                                     _ptr(val), _ptr(flux), ctypes.byref(h)))
        self.h = h
        ctx.track(self)

    def size(self):
        return int(self.ctx.lib.kem_mesh_size(self.h))

    def rhs(self) -> np.ndarray:
        b = np.zeros(self.mesh.n)
        self.ctx.check(self.ctx.lib.kem_mesh_rhs(self.h, _ptr(b)))
        return b

    def set_values(self, values):
        v = _f64(values)
        self.ctx.check(self.ctx.lib.kem_mesh_set_values(self.h, _ptr(v)))
        self.mesh = self.mesh.__class__(**{**self.mesh.__dict__, "bc_value": v})

    def close(self):
        if getattr(self, "h", None) and self.ctx.h:
            self.ctx.lib.kem_mesh_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class KElectrostaticBoundaryIntegrator:
    """Element-level integrals for unit charge density (numeric / rwg policy)."""

    def __init__(self, container: KSurfaceContainer, policy: str = "numeric"):
        self.container = container
        self.ctx = container.ctx
        self.ctx.set_integrator(policy)

    def ElectricFieldAndPotential(self, elements, points):
        """Batched: returns (E[n,3], phi[n]) of element elements[k] at points[k]."""
        el = np.ascontiguousarray(np.atleast_1d(elements), np.int64)
        P = _f64(points, (-1, 3))
        if P.shape[0] != el.shape[0]:
            raise ValueError("elements and points differ in length")
        out = np.zeros((el.shape[0], 4))
        self.ctx.check(self.ctx.lib.kem_element_eval(self.ctx.h, self.container.h, el.shape[0], _ptr(el), _ptr(P),
                                                     _ptr(out)))
        return out[:, :3], out[:, 3]

    def Potential(self, element, point):
        return float(self.ElectricFieldAndPotential([element], [point])[1][0])

    def ElectricField(self, element, point):
        return self.ElectricFieldAndPotential([element], [point])[0][0]


class KBoundaryIntegralMatrix:
    """Dense, device-resident collocation matrix (a KSquareMatrix<double>)."""

    def __init__(self, container: KSurfaceContainer, fp64: bool = True, fp32: bool = False, bf16: bool = False):
        self.ctx = container.ctx
        storage = (_lib.KEM_STORE_FP64 if fp64 else 0) | (_lib.KEM_STORE_FP32 if fp32 else 0) | \
            (_lib.KEM_STORE_BF16 if bf16 else 0)
        h = _P()
        self.ctx.check(self.ctx.lib.kem_matrix_assemble(self.ctx.h, container.h, storage, ctypes.byref(h)))
        self.h = h
        self.n = container.mesh.n
        self.ctx.track(self)

    def Dimension(self):
        return int(self.ctx.lib.kem_matrix_dimension(self.h))

    def __call__(self, i, j):
        v = ctypes.c_double()
        self.ctx.check(self.ctx.lib.kem_matrix_get(self.h, int(i), int(j), ctypes.byref(v)))
        return v.value

    def to_numpy(self):
        A = np.zeros((self.n, self.n))
        self.ctx.check(self.ctx.lib.kem_matrix_copy(self.h, _ptr(A)))
        return A

    def Multiply(self, x):
        x = _f64(x)
        y = np.zeros(self.n)
        self.ctx.check(self.ctx.lib.kem_matrix_multiply(self.h, 0, _ptr(x), _ptr(y)))
        return y

    def MultiplyBatch(self, X, precision="fp64"):
        """Y = A X for the columns of X (n x k); precision "mixed" uses the FP32 matrix (FP32
        MFMA), "bf16" the BF16 matrix (BF16 MFMA, X rounded to BF16)."""
        X = np.asfortranarray(np.asarray(X, dtype=np.float64).reshape(self.n, -1))
        Y = np.zeros_like(X, order="F")
        prec = PRECISIONS[precision]
        self.ctx.check(self.ctx.lib.kem_matrix_multiply_batch(self.h, prec, X.shape[1], _ptr(X), _ptr(Y)))
        return Y

    def solve_batch(self, B, X0=None, tolerance=1e-8, max_iterations=1 << 30, restart=0, precision="fp64",
                    time_limit_s=0.0, inner_tolerance=1e-5):
        """Solve A X = B for the columns of B (n x k) by lockstep GMRES; returns (X, residuals,
        stats, rc) with the per-column relative residuals ||b - A x|| / ||b||."""
        o = _lib.SolveOpts()
        self.ctx.lib.kem_solve_opts_default(ctypes.byref(o))
        o.method = _lib.KEM_SOLVER_GMRES
        o.precision = PRECISIONS[precision]
        o.tolerance = tolerance
        o.max_iterations = int(min(max_iterations, (1 << 31) - 1))
        o.restart = restart
        o.time_limit_s = time_limit_s
        o.inner_tolerance = inner_tolerance
        B = np.asfortranarray(np.asarray(B, dtype=np.float64).reshape(self.n, -1))
        X = np.zeros_like(B, order="F") if X0 is None else np.asfortranarray(np.array(X0, dtype=np.float64))
        res = np.zeros(B.shape[1])
        st = _lib.SolveStats()
        rc = self.ctx.check(self.ctx.lib.kem_solve_batch(self.h, ctypes.byref(o), B.shape[1], _ptr(B), _ptr(X),
                                                         _ptr(res), ctypes.byref(st)))
        return X, res, st, rc

    def MultiplyTranspose(self, x):
        x = _f64(x)
        y = np.zeros(self.n)
        self.ctx.check(self.ctx.lib.kem_matrix_multiply(self.h, 1, _ptr(x), _ptr(y)))
        return y

    def solve(self, b, x0=None, method="gmres", tolerance=1e-8, max_iterations=1 << 30, restart=0,
              check_interval=0, precision="fp64", time_limit_s=0.0, inner_tolerance=1e-5):
        o = _lib.SolveOpts()
        self.ctx.lib.kem_solve_opts_default(ctypes.byref(o))
        methods = {"gmres": _lib.KEM_SOLVER_GMRES, "robin_hood": _lib.KEM_SOLVER_ROBIN_HOOD,
                   "gauss": _lib.KEM_SOLVER_GAUSS}
        if method not in methods:
            raise KEMSimpleException(f"unknown solver '{method}'")
        o.method = methods[method]
        o.precision = PRECISIONS[precision]
        o.tolerance = tolerance
        o.max_iterations = int(min(max_iterations, (1 << 31) - 1))
        o.restart = restart
        o.check_interval = check_interval
        o.time_limit_s = time_limit_s
        o.inner_tolerance = inner_tolerance
        b = _f64(b)
        x = np.zeros(self.n) if x0 is None else _f64(x0).copy()
        st = _lib.SolveStats()
        rc = self.ctx.check(self.ctx.lib.kem_solve(self.h, ctypes.byref(o), _ptr(b), _ptr(x), ctypes.byref(st)))
        return x, st, rc

    def close(self):
        if getattr(self, "h", None) and self.ctx.h:
            self.ctx.lib.kem_matrix_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class KChargeDensitySolver:
    """Base of the charge-density solvers: the solution cache's hash properties and the
    thresholds a solver looks up and stores solutions with (KChargeDensitySolver.hh:17-44;
    FindSolution / SaveSolution in KChargeDensitySolver.cc:51-280, through
    kassiopeia_amd.cache)."""

    def __init__(self):
        self.hash_masked_bits = 20
        self.hash_threshold = 1e-14

    def SetHashProperties(self, masked_bits, hash_threshold):
        self.hash_masked_bits = int(masked_bits)
        self.hash_threshold = float(hash_threshold)

    # the tolerance a stored solution must meet, and the threshold a new one is stored with
    # (KKrylovChargeDensitySolver.cc:66-75: the tolerance / the final residual norm)
    def find_threshold(self):
        return self.kw["tolerance"]

    def save_threshold(self, stats):
        return stats.residual

    def FindSolution(self, directory, mesh):
        from . import cache

        return cache.find_solution(directory, mesh, self.find_threshold(), self.hash_masked_bits,
                                   self.hash_threshold)

    def SaveSolution(self, directory, file, mesh, sigma, stats):
        from . import cache

        cache.save_solution(directory, file, mesh, sigma, self.save_threshold(stats), self.hash_masked_bits,
                            self.hash_threshold)


class KKrylovChargeDensitySolver(KChargeDensitySolver):
    """GMRES charge-density solve (krylov_bem_solver: solver_name, tolerance, max_iterations,
    iterations_between_restarts, time_limit_in_seconds)."""

    def __init__(self, tolerance=1e-8, max_iterations=1 << 30, iterations_between_restarts=0, solver_name="gmres",
                 precision="fp64", time_limit_in_seconds=0.0):
        super().__init__()
        if solver_name != "gmres":
            raise KEMSimpleException(f"unknown Krylov solver '{solver_name}'")
        self.kw = dict(method="gmres", tolerance=tolerance, max_iterations=max_iterations,
                       restart=iterations_between_restarts, precision=precision, time_limit_s=time_limit_in_seconds)

    def solve(self, matrix, b):
        return matrix.solve(b, **self.kw)


class KRobinHoodChargeDensitySolver(KChargeDensitySolver):
    """Robin Hood charge-density solve (robin_hood_bem_solver: tolerance, check_sub_interval)."""

    def __init__(self, tolerance=1e-8, check_sub_interval=0, max_checks=1 << 30, time_limit_in_seconds=0.0):
        super().__init__()
        self.kw = dict(method="robin_hood", tolerance=tolerance, check_interval=check_sub_interval,
                       max_iterations=max_checks, time_limit_s=time_limit_in_seconds)

    def solve(self, matrix, b):
        return matrix.solve(b, **self.kw)

    # found and stored at the tolerance (KRobinHoodChargeDensitySolver.cc:89, 157, 212)
    def save_threshold(self, stats):
        return self.kw["tolerance"]


class DeviceArray:
    """A float64 / int32 array in the context's device memory (kem_dev_alloc), for the entry
    points that take device pointers."""

    def __init__(self, ctx: "Context", host: np.ndarray):
        self.ctx = ctx
        self.host = np.ascontiguousarray(host)
        self.ptr = _P()
        ctx.check(ctx.lib.kem_dev_alloc(ctx.h, self.host.nbytes, ctypes.byref(self.ptr)))
        ctx.track(self)
        ctx.check(ctx.lib.kem_dev_upload(ctx.h, self.ptr, self.host.ctypes.data_as(_P), self.host.nbytes))

    @classmethod
    def empty(cls, ctx: "Context", shape, dtype=np.float64):
        return cls(ctx, np.zeros(shape, dtype=dtype))

    def get(self) -> np.ndarray:
        out = np.empty_like(self.host)
        self.ctx.check(self.ctx.lib.kem_dev_download(self.ctx.h, out.ctypes.data_as(_P), self.ptr, out.nbytes))
        return out

    def upload(self, host: np.ndarray):
        host = np.ascontiguousarray(host, dtype=self.host.dtype)
        assert host.nbytes == self.host.nbytes
        self.ctx.check(self.ctx.lib.kem_dev_upload(self.ctx.h, self.ptr, host.ctypes.data_as(_P), host.nbytes))

    def close(self):
        if self.ptr and self.ctx.h:
            self.ctx.lib.kem_dev_free(self.ctx.h, self.ptr)
        self.ptr = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def field_sets(container: "KSurfaceContainer", sigma_sets, set_index, points, precision="fp64"):
    """Field of several charge-density sets in one pass (kem_field_eval_sets): point i uses
    column set_index[i] of sigma_sets (n x k)."""
    ctx = container.ctx
    S = DeviceArray(ctx, np.asfortranarray(np.asarray(sigma_sets, dtype=np.float64)).T)
    idx = DeviceArray(ctx, np.asarray(set_index, dtype=np.int32))
    P = _f64(points, (-1, 3))
    X = [DeviceArray(ctx, P[:, k]) for k in range(3)]
    m = P.shape[0]
    O = [DeviceArray.empty(ctx, m) for _ in range(4)]
    prec = {"fp64": _lib.KEM_FP64, "mixed": _lib.KEM_MIXED}[precision]
    try:
        ctx.check(ctx.lib.kem_field_eval_sets(ctx.h, container.h, S.host.shape[0], S.ptr, idx.ptr, m, X[0].ptr,
                                              X[1].ptr, X[2].ptr, prec, O[0].ptr, O[1].ptr, O[2].ptr, O[3].ptr))
        ctx.check(ctx.lib.kem_ctx_sync(ctx.h))
        E = np.stack([O[k].get() for k in (1, 2, 3)], 1)
        return E, O[0].get()
    finally:
        for a in [S, idx] + X + O:
            a.close()


class KGaussianEliminationChargeDensitySolver(KChargeDensitySolver):
    """Direct solve (gaussian_elimination_bem_solver; KGaussianEliminationChargeDensitySolver.cc:29-46)."""

    kw = {"method": "gauss"}

    def solve(self, matrix, b):
        return matrix.solve(b, method="gauss")

    # found and stored at threshold 0 (KGaussianEliminationChargeDensitySolver.cc:35, 44)
    def find_threshold(self):
        return 0.0

    def save_threshold(self, stats):
        return 0.0


class KIntegratingElectrostaticFieldSolver:
    """Direct Kahan-summed field of a solved container: fused, batched evaluation."""

    def __init__(self, container: KSurfaceContainer, sigma, precision="fp64"):
        self.container = container
        self.ctx = container.ctx
        self.sigma = _f64(sigma)
        self.precision = {"fp64": _lib.KEM_FP64, "mixed": _lib.KEM_MIXED}[precision]

    def ElectricFieldAndPotential(self, points):
        P = _f64(points, (-1, 3))
        m = P.shape[0]
        phi = np.zeros(m)
        E = np.zeros((m, 3))
        self.ctx.check(self.ctx.lib.kem_field_eval(self.ctx.h, self.container.h, _ptr(self.sigma), m, _ptr(P),
                                                   self.precision, _ptr(phi), _ptr(E)))
        return E, phi

    def Potential(self, points):
        return self.ElectricFieldAndPotential(points)[1]

    def ElectricField(self, points):
        return self.ElectricFieldAndPotential(points)[0]


class KElectrostaticBoundaryField:
    """Top-level field: Initialize() finds or solves the charge densities, then evaluates
    fields (KElectrostaticBoundaryField.cc:77-95).  With a directory set (SetDirectory), solved
    charge densities are stored in SetFile's .kbd file there and found again by later
    Initialize() calls on the same geometry (the reference's cache, with its hash properties:
    20 masked bits, threshold 1e-14); without one nothing is read or written."""

    def __init__(self, ctx: Context, mesh: Mesh, integrator="numeric", solver=None, precision="fp64",
                 directory=None, file=None):
        self.ctx = ctx
        self.mesh = mesh
        self.integrator = integrator
        self.solver = solver or KKrylovChargeDensitySolver(tolerance=1e-8)
        self.precision = precision
        self.container = None
        self.sigma = None
        self.stats = None
        self.directory = directory
        self.file = file
        self.hash_masked_bits = 20
        self.hash_threshold = 1e-14
        self.cache_result = 0  # 0 solved, 1 stored solution, 2 superposed stored solutions
        self.cache_file = ""

    def SetChargeDensitySolver(self, solver):
        self.solver = solver

    def SetDirectory(self, directory):
        self.directory = directory

    def SetFile(self, file):
        self.file = file

    def Initialize(self):
        self.ctx.set_integrator(self.integrator)
        self.container = KSurfaceContainer(self.ctx, self.mesh)
        cached = self.directory is not None and isinstance(self.solver, KChargeDensitySolver) and \
            self.mesh.group_start is None
        self.cache_result, self.cache_file, self.stats = 0, "", None
        if cached:
            self.solver.SetHashProperties(self.hash_masked_bits, self.hash_threshold)
            kind, sigma, fname = self.solver.FindSolution(self.directory, self.mesh)
            if kind:
                self.cache_result, self.cache_file, self.sigma = kind, fname, sigma
        if not self.cache_result:
            need32 = self.solver.kw.get("precision") == "mixed"
            A = KBoundaryIntegralMatrix(self.container, fp64=True, fp32=need32)
            self.sigma, self.stats, _ = self.solver.solve(A, self.container.rhs())
            A.close()
            if cached:
                from . import cache

                if self.file is None:
                    self.file = cache.default_file_name()
                self.solver.SaveSolution(self.directory, self.file, self.mesh, self.sigma, self.stats)
        self.field = KIntegratingElectrostaticFieldSolver(self.container, self.sigma, self.precision)

    def _ready(self):
        if self.sigma is None:
            raise RuntimeError("KElectrostaticBoundaryField: call Initialize() first")

    def ElectricFieldAndPotential(self, P):
        self._ready()
        P = np.asarray(P, dtype=np.float64)
        E, phi = self.field.ElectricFieldAndPotential(P)
        if P.ndim == 1:
            return E[0], float(phi[0])
        return E, phi

    def Potential(self, P):
        return self.ElectricFieldAndPotential(P)[1]

    def ElectricField(self, P):
        return self.ElectricFieldAndPotential(P)[0]

    def capacitance(self):
        """Q / V / (4 pi eps0) for a single electrode at uniform voltage (CubeCapacitanceTest.cc:108-113)."""
        self._ready()
        Q = float((self.mesh.areas() * self.sigma).sum())
        V = float(self.mesh.bc_value[0])
        return Q / V / (4.0 * np.pi * 8.8541878128e-12)
