"""Differentiable BEM potential: electrode voltages and mesh vertices -> potential at points.

A thin harness over the C-ABI adjoint entry points (kem_field_vjp, kem_solve_transposed,
kem_assemble_vjp).  The forward pass is the reference's path: collocation matrix A(X) (numeric
policy), A sigma = b(V) by GMRES, phi_k = sum_j G(P_k; T_j) sigma_j.  For a loss L(phi) with
g = dL/dphi the backward pass is

    c        = G^T g                          (kem_field_vjp: dL/dsigma)
    A^T lam  = c                              (kem_solve_transposed)
    dL/dV_e  = sum_{i in electrode e} lam_i   (b_i = V_{e(i)})
    dL/dX    = g^T (dG/dX) sigma - lam^T (dA/dX) sigma   (kem_field_vjp, kem_assemble_vjp)

with the per-element corner gradients scattered to the shared vertices.  `BEMPotential` wraps
this as a torch.autograd.Function (CPU tensors; torch is used for autograd only).  KEMField
itself has no derivatives.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from . import geometry as G
from . import kemfield as K


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class BEMProblem:
    """A triangle mesh with per-triangle electrode index, evaluation points and solver
    settings.  X: (nv, 3) vertex positions, tri: (ne, 3) vertex indices (p0, p1, p2 of each
    KTriangle::SetValues), electrode: (ne,) in [0, nelectrodes)."""

    def __init__(self, ctx: K.Context, tri, electrode, points, nelectrodes=None, tolerance=1e-12,
                 precision="fp64"):
        self.ctx = ctx
        self.tri = np.ascontiguousarray(tri, dtype=np.int64)
        self.electrode = np.ascontiguousarray(electrode, dtype=np.int64)
        self.points = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        self.ne = self.tri.shape[0]
        self.nelectrodes = int(nelectrodes if nelectrodes is not None else self.electrode.max() + 1)
        self.tolerance = tolerance
        self.precision = precision
        self._state = None

    def mesh(self, V, X) -> G.Mesh:
        X = np.asarray(X, dtype=np.float64)
        recs = G.records_from_points(X[self.tri[:, 0]], X[self.tri[:, 1]], X[self.tri[:, 2]])
        n = self.ne
        return G.Mesh(np.zeros(n, np.int32), recs, np.zeros(n, np.int32),
                      np.asarray(V, dtype=np.float64)[self.electrode], np.zeros(n),
                      self.electrode.astype(np.int32))

    def forward(self, V, X):
        """phi at the points; keeps what the backward pass needs."""
        self.ctx.set_integrator("numeric")
        m = self.mesh(V, X)
        cont = K.KSurfaceContainer(self.ctx, m)
        A = K.KBoundaryIntegralMatrix(cont, fp64=True, fp32=self.precision == "mixed")
        sigma, st, rc = A.solve(cont.rhs(), tolerance=self.tolerance, precision=self.precision)
        if rc != 0:
            raise RuntimeError(f"BEM solve did not converge (residual {st.residual:.2e})")
        E, phi = K.KIntegratingElectrostaticFieldSolver(cont, sigma).ElectricFieldAndPotential(self.points)
        self._state = (cont, A, sigma)
        return phi

    def vjp(self, gphi, X_shape):
        """(dL/dV, dL/dX) for gphi = dL/dphi at the points of the last forward pass."""
        cont, A, sigma = self._state
        lib, h = self.ctx.lib, self.ctx.h
        g = np.ascontiguousarray(gphi, dtype=np.float64)
        n = self.ne
        dsig = np.zeros(n)
        dvf = np.zeros((n, 9))
        self.ctx.check(lib.kem_field_vjp(h, cont.h, _ptr(sigma), self.points.shape[0], _ptr(self.points), _ptr(g),
                                         _ptr(dsig), _ptr(dvf)))
        o = _lib.SolveOpts()
        lib.kem_solve_opts_default(ctypes.byref(o))
        o.tolerance = self.tolerance
        o.precision = {"fp64": _lib.KEM_FP64, "mixed": _lib.KEM_MIXED}[self.precision]
        lam = np.zeros(n)
        st = _lib.SolveStats()
        rc = self.ctx.check(lib.kem_solve_transposed(A.h, ctypes.byref(o), _ptr(dsig), _ptr(lam), ctypes.byref(st)))
        if rc != 0:
            raise RuntimeError(f"adjoint solve did not converge (residual {st.residual:.2e})")
        dvm = np.zeros((n, 9))
        self.ctx.check(lib.kem_assemble_vjp(h, cont.h, _ptr(lam), _ptr(sigma), _ptr(dvm)))
        dV = np.bincount(self.electrode, weights=lam, minlength=self.nelectrodes)
        dcorner = (dvf + dvm).reshape(n, 3, 3)
        dX = np.zeros(X_shape)
        for m in range(3):
            np.add.at(dX, self.tri[:, m], dcorner[:, m, :])
        return dV, dX


def potential_fn():
    """torch.autograd.Function: phi = BEMPotential.apply(V, X, problem) for CPU float64 tensors."""
    import torch

    class BEMPotential(torch.autograd.Function):
        @staticmethod
        def forward(fctx, V, X, problem):
            phi = problem.forward(V.detach().numpy(), X.detach().numpy())
            fctx.problem = problem
            fctx.xshape = tuple(X.shape)
            return torch.from_numpy(phi)

        @staticmethod
        def backward(fctx, gphi):
            dV, dX = fctx.problem.vjp(gphi.detach().numpy(), fctx.xshape)
            return torch.from_numpy(dV), torch.from_numpy(dX), None

    return BEMPotential
