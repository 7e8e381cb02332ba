"""Workspace arithmetic of the field passes (kem_field_plan: host code of the library, no device
access).  Every partial sum a plan writes must lie inside the workspace the library allocates,
for single-set passes (with and without the per-XCD element ranges), passes over several
charge-density sets and a workspace smaller than the preferred size, up to point counts past
2^24 where 8 doubles per point exceed the per-XCD partial-sum budget."""
import ctypes
import os

import numpy as np
import pytest

NUMERIC, RWG = 0, 1
FP64, MIXED = 0, 1


@pytest.fixture(scope="module")
def lib():
    from kassiopeia_amd import _lib

    if not os.path.exists(_lib.LIBPATH):
        from kassiopeia_amd import build

        build.build()
    return _lib.load()


def plan(lib, policy, clustered, precision, multi, nelem, nsets, npts, range_elems, ws=-1):
    out = (ctypes.c_int64 * 8)()
    rc = lib.kem_field_plan(policy, clustered, precision, multi, nelem, nsets, npts, range_elems, ws, out)
    assert rc == 0
    keys = ("ws", "weights", "partials", "ranges", "slices", "chunk", "xcd", "kind")
    return dict(zip(keys, list(out)))


CASES = [(p, cl, pr, mu, ne, ns, npts, rg)
         for p, cl in ((NUMERIC, 1), (NUMERIC, 0), (RWG, 0))
         for pr in (FP64, MIXED)
         for mu, ns in ((0, 1), (1, 1), (1, 64))
         for ne in (1, 31, 1728, 20000, 100000)
         for npts in (1, 255, 4096, 262143, 262144, 1 << 20, (1 << 24) + 1, 40_000_000)
         for rg in (0, 640, 2560)]


def test_partials_fit_workspace(lib):
    for p, cl, pr, mu, ne, ns, npts, rg in CASES:
        q = plan(lib, p, cl, pr, mu, ne, ns, npts, rg)
        case = (p, cl, pr, mu, ne, ns, npts, rg, q)
        assert q["weights"] + q["partials"] <= q["ws"], case
        assert q["ranges"] >= 1 and q["slices"] in (q["ranges"], q["ranges"] + 1), case
        if q["xcd"]:
            # chunks of points: the partials of one chunk (4 components x slices) are reused
            assert q["partials"] == 4 * q["slices"] * min(q["chunk"], npts), case
            assert q["partials"] <= 128 << 20, case
            assert mu == 0 and rg > 0 and p == NUMERIC and cl == 1, case
        else:
            assert q["chunk"] == npts, case
            if q["slices"] > 1:
                assert q["partials"] == 4 * q["slices"] * npts, case


def test_several_sets_and_range_zero_past_2_24(lib):
    """The culled near pass needs a slice beside the far pass's even with one element range:
    8 doubles per point, more than the 128 Mi-double per-XCD budget past 2^24 points."""
    npts = (1 << 24) + 1
    for multi, rg in ((1, 2560), (0, 0)):
        q = plan(lib, NUMERIC, 1, MIXED, multi, 20000, 1, npts, rg)
        assert not q["xcd"]
        assert q["slices"] >= 2 and q["partials"] >= 8 * npts
        assert q["weights"] + q["partials"] <= q["ws"]


def test_small_workspace_takes_fewer_ranges(lib):
    npts, ne = 100_000, 20000
    full = plan(lib, NUMERIC, 1, MIXED, 0, ne, 1, npts, 0)
    # room for the weights and two slices only: one element range plus the near slice
    small = full["weights"] + 8 * npts
    q = plan(lib, NUMERIC, 1, MIXED, 0, ne, 1, npts, 0, ws=small)
    assert full["ranges"] > 1 and q["ranges"] == 1 and q["slices"] == 2
    assert q["weights"] + q["partials"] <= small
    # too small for even that: the plan overflows and launch_field refuses it
    q = plan(lib, NUMERIC, 1, MIXED, 0, ne, 1, npts, 0, ws=small - 1)
    assert q["weights"] + q["partials"] > small - 1


def test_xcd_plan_covers_every_element(lib):
    for ne in (1000, 20000, 99668, 100000):
        for rg in (640, 2560, 10240):
            q = plan(lib, NUMERIC, 1, MIXED, 0, ne, 1, 1 << 20, rg)
            assert q["xcd"] and q["ranges"] % 8 == 0
            # ranges x elements per range cover the mesh (per range from the plan's own split)
            per = -(-ne // q["ranges"])
            per = -(-per // 32) * 32
            assert per * q["ranges"] >= ne
