"""Zonal-harmonic field of axisymmetric elements through the C-ABI (kem_zh_*).

Mirrors KEMField's KZonalHarmonicContainer<KElectrostaticBasis> (ComputeCoefficients:
FieldSolvers/ZonalHarmonic/Generator/include/KZonalHarmonicContainer.hh:284-345) and
KZonalHarmonicFieldSolver<KElectrostaticBasis> (ElectricFieldAndPotential, Potential,
ElectricField: Solver/src/KElectrostaticZonalHarmonicFieldSolver.cc:36-124), with the
parameters of KZonalHarmonicParameters (Generator/include/KZonalHarmonicParameters.hh).
The expansion is generated on the host and evaluated on the device, one lane per point.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .kemfield import KSurfaceContainer, _f64, _ptr


class KZonalHarmonicParameters:
    """KZonalHarmonicParameters: the same names and defaults (Set*/Get* as attributes)."""

    FIELDS = ("n_bifurcations", "convergence_ratio", "proximity_to_source_point", "convergence_parameter",
              "n_central_coefficients", "central_fractional_spacing", "central_fractional_distance",
              "central_delta_z", "central_z1", "central_z2", "n_remote_coefficients", "n_remote_source_points",
              "remote_z1", "remote_z2")

    def __init__(self, lib=None):
        lib = lib or _lib.load()
        p = _lib.ZHParams()
        lib.kem_zh_params_default(ctypes.byref(p))
        for f in self.FIELDS:
            setattr(self, f, getattr(p, f))

    def SetNBifurcations(self, n):
        self.n_bifurcations = int(n)

    def SetNCentralCoefficients(self, n):
        self.n_central_coefficients = int(n)

    def SetNRemoteCoefficients(self, n):
        self.n_remote_coefficients = int(n)

    def SetNRemoteSourcePoints(self, n):
        self.n_remote_source_points = int(n)

    def SetConvergenceRatio(self, v):
        self.convergence_ratio = float(v)

    def SetConvergenceParameter(self, v):
        self.convergence_parameter = float(v)

    def SetCentralFractionalSpacing(self, b):
        self.central_fractional_spacing = 1 if b else 0

    def SetCentralFractionalDistance(self, v):
        self.central_fractional_distance = float(v)

    def SetCentralDeltaZ(self, v):
        self.central_delta_z = float(v)

    def struct(self):
        p = _lib.ZHParams()
        for f in self.FIELDS:
            setattr(p, f, getattr(self, f))
        return p


class KZonalHarmonicContainer:
    """The expansion of a surface container's axisymmetric elements with their charge
    densities (rings: total charge); ComputeCoefficients() builds it on the device."""

    def __init__(self, container: KSurfaceContainer, sigma):
        self.container = container
        self.ctx = container.ctx
        self.sigma = _f64(sigma)
        self.parameters = KZonalHarmonicParameters(self.ctx.lib)
        self.h = None

    def GetParameters(self):
        return self.parameters

    def ComputeCoefficients(self):
        self.close()
        h = ctypes.c_void_p()
        p = self.parameters.struct()
        self.ctx.check(self.ctx.lib.kem_zh_create(self.ctx.h, self.container.h, _ptr(self.sigma), ctypes.byref(p),
                                                  ctypes.byref(h)))
        self.h = h
        self.ctx.track(self)

    def info(self):
        """(tree nodes, central source points, remote source points, depth)"""
        out = np.zeros(4, np.int64)
        self.ctx.check(self.ctx.lib.kem_zh_info(self.h, _ptr(out)))
        return tuple(int(v) for v in out)

    def close(self):
        if getattr(self, "h", None) and self.ctx.h:
            self.ctx.lib.kem_zh_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class KZonalHarmonicFieldSolver:
    """ElectricFieldAndPotential / Potential / ElectricField of a computed container over
    batches of points; the last call's central / remote / direct execution counts in
    `counts` (KZonalHarmonicComputer::Get*ExecutionCount)."""

    def __init__(self, zh: KZonalHarmonicContainer):
        if zh.h is None:
            raise RuntimeError("KZonalHarmonicFieldSolver: call ComputeCoefficients() first")
        self.zh = zh
        self.counts = np.zeros(3, np.int64)

    def Initialize(self):
        pass

    def ElectricFieldAndPotential(self, points):
        P = _f64(points, (-1, 3))
        m = P.shape[0]
        phi = np.zeros(m)
        E = np.zeros((m, 3))
        self.zh.ctx.check(self.zh.ctx.lib.kem_zh_field_eval(self.zh.h, m, _ptr(P), _ptr(phi), _ptr(E),
                                                             _ptr(self.counts)))
        return E, phi

    def Potential(self, points):
        return self.ElectricFieldAndPotential(points)[1]

    def ElectricField(self, points):
        return self.ElectricFieldAndPotential(points)[0]

    def GetCentralExecutionCount(self):
        return int(self.counts[0])

    def GetRemoteExecutionCount(self):
        return int(self.counts[1])

    def GetDirectExecutionCount(self):
        return int(self.counts[2])
