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


def test_cpp_facade_header_compiles(tmp_path):
    """include/kemhip.hpp is self-contained C++17 over the C header."""
    import shutil
    import subprocess

    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    src = tmp_path / "t.cpp"
    src.write_text('#include "kemhip.hpp"\nint main() { return 0; }\n')
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    subprocess.run([gxx, "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.join(root, "include"),
                    str(src)], check=True)
