"""Charge-density solution cache over the C-ABI (host code; no device needed).

Mirrors the cache of KEMField's charge-density solvers (paths relative to KEMField/Source):

* ``FindSolution`` / ``SaveSolution``  Interface/ChargeDensitySolvers/Electric/src/KChargeDensitySolver.cc:51-280
* the directory of ``.kbd`` files       IO/FileManipulation/include/KEMFileInterface.hh:23-81, KEMFile.hh
* masked-bit MD5 hashes                 IO/HashGenerator/include/KMD5HashGenerator.hh

A solved container is stored with its residual threshold, labelled by the hash of its shapes
(charge densities and Dirichlet values left out) and of its shapes with boundary values; a later
solve of a container with the same hashes and a tolerance at least that threshold reads the
stored charge densities instead of solving.  Containers of the same shapes with other boundary
values are superposed when their boundary vectors span the new one.  The hashes round and mask
the ``masked_bits`` low bits of every double and leave out values below ``hash_threshold`` in
magnitude (the reference's field sets 20 bits and 1e-14).
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

from . import _lib
from .geometry import Mesh

OMIT_BASIS = 1  # KElectrostaticBasis: the charge densities
OMIT_DIRICHLET = 2  # KBoundaryType<KElectrostaticBasis, KDirichletBoundary>: the Dirichlet values
MASKED_BITS = 20  # KElectrostaticBoundaryField's fHashMaskedBits
HASH_THRESHOLD = 1e-14  # and fHashThreshold (KElectrostaticBoundaryField.cc:20-21)

NOT_FOUND, FOUND, SUPERPOSED = 0, 1, 2

_P = ctypes.c_void_p


class CacheError(RuntimeError):
    pass


def _check(lib, rc):
    if rc < 0:
        raise CacheError(lib.kem_cache_error().decode())
    return rc


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_P)


def _surfaces(mesh: Mesh):
    if mesh.group_start is not None:
        raise CacheError("the solution cache takes containers without symmetry groups")
    return (np.ascontiguousarray(mesh.shapes, np.int32), np.ascontiguousarray(mesh.recs, np.float64),
            np.ascontiguousarray(mesh.bc_type, np.int32), np.ascontiguousarray(mesh.bc_value, np.float64),
            np.ascontiguousarray(mesh.bc_flux, np.float64))


def md5_hex(data: bytes) -> str:
    """RFC 1321 MD5 of bytes as 32 lowercase hex digits (md5.hh's hex_digest)."""
    lib = _lib.load()
    buf = ctypes.create_string_buffer(bytes(data), len(data)) if data else None
    out = ctypes.create_string_buffer(33)
    _check(lib, lib.kem_md5_hex(buf, len(data), out))
    return out.value.decode()


def hash_doubles(x, masked_bits: int = 1, threshold: float = 1e-12) -> str:
    """KMD5HashGenerator's hash of doubles streamed in order (its defaults: 1 bit, 1e-12)."""
    lib = _lib.load()
    x = np.ascontiguousarray(x, np.float64).ravel()
    out = ctypes.create_string_buffer(33)
    _check(lib, lib.kem_hash_doubles(_ptr(x), x.size, masked_bits, threshold, out))
    return out.value.decode()


def hash_surfaces(mesh: Mesh, sigma=None, omit: int = 0, masked_bits: int = 1, threshold: float = 1e-12) -> str:
    """KMD5HashGenerator::GenerateHash of the container with charge densities sigma, leaving out
    the policies in omit (OMIT_BASIS, OMIT_DIRICHLET)."""
    lib = _lib.load()
    s = _surfaces(mesh)
    sg = None if sigma is None else np.ascontiguousarray(sigma, np.float64)
    out = ctypes.create_string_buffer(33)
    _check(lib, lib.kem_hash_surfaces(mesh.n, *(_ptr(a) for a in s), _ptr(sg), omit, masked_bits, threshold, out))
    return out.value.decode()


def default_file_name() -> str:
    """KEMFile's default name: KEM_<date>_<time>.kbd (IO/FileManipulation/src/KEMFile.cc:18-31)."""
    return time.strftime("KEM_%Y-%m-%d_%H-%M-%S.kbd")


def save_solution(directory: str, file: str, mesh: Mesh, sigma, residual_threshold: float,
                  masked_bits: int = MASKED_BITS, hash_threshold: float = HASH_THRESHOLD):
    """KChargeDensitySolver::SaveSolution: appends the threshold record and the solved
    container to directory/file."""
    lib = _lib.load()
    s = _surfaces(mesh)
    sg = np.ascontiguousarray(sigma, np.float64)
    if sg.shape != (mesh.n,):
        raise CacheError("sigma must have one entry per element")
    _check(lib, lib.kem_cache_save(str(directory).encode(), str(file).encode(), mesh.n, *(_ptr(a) for a in s),
                                   _ptr(sg), float(residual_threshold), masked_bits, hash_threshold))


def find_solution(directory: str, mesh: Mesh, tolerance: float, masked_bits: int = MASKED_BITS,
                  hash_threshold: float = HASH_THRESHOLD):
    """KChargeDensitySolver::FindSolution: (NOT_FOUND | FOUND | SUPERPOSED, sigma or None,
    file read)."""
    lib = _lib.load()
    s = _surfaces(mesh)
    sigma = np.zeros(mesh.n)
    found = ctypes.c_int()
    fname = ctypes.create_string_buffer(4096)
    _check(lib, lib.kem_cache_find(str(directory).encode(), mesh.n, *(_ptr(a) for a in s), float(tolerance),
                                   masked_bits, hash_threshold, _ptr(sigma), ctypes.byref(found), fname, 4096))
    if found.value == NOT_FOUND:
        return NOT_FOUND, None, ""
    return found.value, sigma, fname.value.decode()
