"""The C-ABI library builds for gfx950, loads on a machine without a GPU and exports every
entry point that include/kemhip.h declares (no compute calls here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "kemhip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(kem_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    from kassiopeia_amd import _lib

    if not os.path.exists(_lib.LIBPATH):
        from kassiopeia_amd import build

        build.build()
    return _lib.load()


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in ("kem_ctx_create", "kem_field_eval", "kem_matrix_assemble", "kem_solve", "kem_element_eval"):
        assert s in syms


def test_library_exports_every_declared_symbol(lib):
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_header():
    from kassiopeia_amd import _lib

    assert sorted(_lib.SIGNATURES) == declared_symbols()


def test_version_and_null_handling(lib):
    assert lib.kem_version() == 1
    # argument checks run before any device work
    assert lib.kem_ctx_create(0, None) == -1
    assert lib.kem_mesh_size(None) == 0
    assert lib.kem_last_error(None) == b"null context"


def test_gfx950_code_object_present():
    from kassiopeia_amd import _lib

    data = open(_lib.LIBPATH, "rb").read()
    assert b"gfx950" in data
