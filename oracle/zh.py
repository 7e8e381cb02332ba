"""ctypes wrapper of the oracle's zonal-harmonic restatement (kem_oracle_zh.c).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of kassiopeia_amd.zonal.
"""
import ctypes

import numpy as np

from . import oracle as O

_P = ctypes.c_void_p
_DEFAULTS = [5, 0.99, 1.e-14, 1.e-15, 500, 1, 0.2, 0.01, 0., 0., 500, 3, 0., 0.]


def _lib():
    lib = O.lib()
    if not getattr(lib, "_zh_ready", False):
        lib.ko_zh_build.restype = _P
        lib.ko_zh_build.argtypes = [ctypes.c_int64, _P, _P, _P, _P]
        lib.ko_zh_eval.restype = None
        lib.ko_zh_eval.argtypes = [_P, ctypes.c_int64, _P, _P, _P]
        lib.ko_zh_info.restype = None
        lib.ko_zh_info.argtypes = [_P, _P]
        lib.ko_zh_central_sp.restype = ctypes.c_int
        lib.ko_zh_central_sp.argtypes = [_P, ctypes.c_int, _P]
        lib.ko_zh_free.restype = None
        lib.ko_zh_free.argtypes = [_P]
        lib._zh_ready = True
    return lib


class ZH:
    """The restated expansion of elements (shape codes, KO_REC-stride records, charge
    densities); params: overrides of the 14 parameters by index name."""

    NAMES = ["n_bifurcations", "convergence_ratio", "proximity_to_source_point", "convergence_parameter",
             "n_central_coefficients", "central_fractional_spacing", "central_fractional_distance",
             "central_delta_z", "central_z1", "central_z2", "n_remote_coefficients", "n_remote_source_points",
             "remote_z1", "remote_z2"]

    def __init__(self, shapes, recs, sigma, **params):
        self.lib = _lib()
        self.shapes = np.ascontiguousarray(shapes, np.int32)
        self.recs = np.ascontiguousarray(np.asarray(recs, np.float64).reshape(len(self.shapes), -1)[:, :O.REC])
        if self.recs.shape[1] < O.REC:
            self.recs = np.ascontiguousarray(np.pad(self.recs, ((0, 0), (0, O.REC - self.recs.shape[1]))))
        self.sigma = np.ascontiguousarray(sigma, np.float64)
        vals = list(_DEFAULTS)
        for k, v in params.items():
            vals[self.NAMES.index(k)] = float(v)
        self.params = np.array(vals, np.float64)
        self.h = self.lib.ko_zh_build(len(self.shapes), O._p(self.shapes), O._p(self.recs), O._p(self.sigma),
                                      O._p(self.params))

    def field(self, points):
        """(E (m, 3), phi (m,), counts (central, remote, direct))"""
        P = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
        out = np.zeros((P.shape[0], 4))
        cnt = np.zeros(3, np.int64)
        self.lib.ko_zh_eval(self.h, P.shape[0], O._p(P), O._p(out), O._p(cnt))
        return out[:, :3], out[:, 3], cnt

    def info(self):
        out = np.zeros(3, np.int64)
        self.lib.ko_zh_info(self.h, O._p(out))
        return tuple(int(v) for v in out)

    def central_sp(self, k):
        out = np.zeros(2 + int(self.params[4]))
        n = self.lib.ko_zh_central_sp(self.h, k, O._p(out))
        return out[0], out[1], out[2:2 + n]

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.ko_zh_free(self.h)
            self.h = None
