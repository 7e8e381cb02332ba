"""The launch-shape autotuner (kassiopeia_amd/autotune.py) on the device: every knob's sweep
returns a time per candidate and leaves the fastest set, results do not depend on the knobs
beyond summation order, and tables round-trip through JSON."""
import numpy as np
import pytest

from kassiopeia_amd import geometry as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    from kassiopeia_amd import kemfield as K

    ctx = K.Context(0)
    m = G.coax_mesh(nphi_in=12, nphi_out=24, nz=16, ncap_r=2, ncap_phi=16)
    cont = K.KSurfaceContainer(ctx, m)
    A = K.KBoundaryIntegralMatrix(cont, fp64=True, fp32=True)
    x, st, rc = A.solve(cont.rhs(), tolerance=1e-10)
    yield ctx, m, cont, A, x
    A.close()
    cont.close()
    ctx.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_tune_field(setup, precision):
    from kassiopeia_amd import autotune as T
    from kassiopeia_amd import kemfield as K

    ctx, m, cont, A, x = setup
    pts = G.points_annulus(300_000, 0.5, 1.0, 2.0, 3)
    res = T.tune_field(cont, x, pts, precision, repeats=1, candidates=(0, 64, 2560))
    assert set(res["ms"]) == {0, 64, 2560} and all(t > 0 for t in res["ms"].values())
    assert res["best"] == min(res["ms"], key=res["ms"].get) == T.get(ctx, "far_range_" + precision)
    # the field does not depend on the split beyond summation order (and, mixed, the last FP32
    # bit of some pairs)
    out = {}
    for v in (0, 64, 2560):
        T.set_knob(ctx, "far_range_" + precision, v)
        out[v] = K.KIntegratingElectrostaticFieldSolver(cont, x, precision).ElectricFieldAndPotential(pts)
    tol = 1e-14 if precision == "fp64" else 1e-9
    for v in (64, 2560):
        assert np.linalg.norm(out[v][0] - out[0][0]) / np.linalg.norm(out[0][0]) < tol
        assert np.abs(out[v][1] - out[0][1]).max() / 18600.0 < 10 * tol
    T.set_knob(ctx, "far_range_" + precision, 2560)


def test_tune_products(setup):
    from kassiopeia_amd import autotune as T

    from kassiopeia_amd import kemfield as K

    ctx, m, cont, A, x = setup
    v = np.random.default_rng(2).normal(size=A.n)
    ref = A.Multiply(v)
    r = T.tune_matvec(A, repeats=2, candidates=(256, 2048))
    assert set(r["ms"]) == {256, 2048} and T.get(ctx, "gemv_blocks") == r["best"]
    for c in (1, 256, 4096):
        T.set_knob(ctx, "gemv_blocks", c)
        assert np.linalg.norm(A.Multiply(v) - ref) <= 1e-14 * np.linalg.norm(ref)
    X = np.random.default_rng(3).normal(size=(A.n, 40))
    ref = A.MultiplyBatch(X, "mixed")
    ref64 = A.MultiplyBatch(X[:, :20], "fp64")
    # on this FP32 matrix (n % 4 == 0) 40 right-hand sides stream through LDS on 32 x 32 tiles
    # and 7 on 16 x 16 tiles, as the FP64 matrix (n even) does: each sweeps its workgroup target
    r = T.tune_block_product(A, nrhs=40, repeats=1, candidates=(60, 480))
    assert r["knob"] == "bmm_stream_blocks" and T.get(ctx, "bmm_stream_blocks") == r["best"]
    assert set(r["ms"]) == {60, 480}
    r = T.tune_block_product(A, nrhs=7, repeats=1, candidates=(256, 2048))
    assert r["knob"] == "bmm_tile_blocks" and T.get(ctx, "bmm_tile_blocks") == r["best"]
    assert set(r["ms"]) == {256, 2048}
    r = T.tune_block_product(A, nrhs=5, precision="fp64", repeats=1, candidates=(256, 2048))
    assert r["knob"] == "bmm_tile_blocks" and set(r["ms"]) == {256, 2048}
    # every workgroup target: the product of the FP32 matrix to FP32 accuracy, FP64 to FP64's
    for c in (1, 64, 10240):
        T.set_knob(ctx, "bmm_stream_blocks", c)
        T.set_knob(ctx, "bmm_tile_blocks", c)
        for nrhs in (7, 20, 40):
            R = A.MultiplyBatch(X[:, :nrhs], "mixed")
            assert np.linalg.norm(R - ref[:, :nrhs]) <= 1e-6 * np.linalg.norm(ref[:, :nrhs]), (c, nrhs)
        R = A.MultiplyBatch(X[:, :20], "fp64")
        assert np.linalg.norm(R - ref64) <= 1e-14 * np.linalg.norm(ref64), c
    # the register-loaded kernels (an FP32 matrix with n % 4 != 0): split target and tile
    sub = G.Mesh(m.shapes[:1278].copy(), m.recs[:1278].copy(), m.bc_type[:1278].copy(), m.bc_value[:1278].copy(),
                 m.bc_flux[:1278].copy(), m.group[:1278].copy())
    cont2 = K.KSurfaceContainer(ctx, sub)
    A2 = K.KBoundaryIntegralMatrix(cont2, fp64=False, fp32=True)
    try:
        X2 = X[:1278, :7]
        ref2 = A2.MultiplyBatch(X2, "mixed")
        r = T.tune_block_product(A2, nrhs=7, repeats=1, candidates=(320, 2560))
        assert r["knob"] == "bmm_blocks" and T.get(ctx, "bmm_blocks") == r["best"]
        assert T.get(ctx, "bmm_rows") == r["rows"]["best"] and set(r["rows"]["ms"]) == {64}
        for c in (64, 10240):
            T.set_knob(ctx, "bmm_blocks", c)
            assert np.linalg.norm(A2.MultiplyBatch(X2, "mixed") - ref2) <= 1e-6 * np.linalg.norm(ref2), c
    finally:
        A2.close()
        cont2.close()
    T.apply(ctx, {"gemv_blocks": 2048, "bmm_blocks": 2560, "bmm_rows": 64, "bmm_stream_blocks": 480,
                  "bmm_tile_blocks": 4096})


def test_tuning_table_roundtrip_and_errors(setup, tmp_path):
    from kassiopeia_amd import _lib
    from kassiopeia_amd import autotune as T

    ctx = setup[0]
    table = {"far_range_fp64": 1280, "far_range_mixed": {"best": 5120, "ms": {}}, "gemv_blocks": 1024,
             "bmm_blocks": 2560, "near_blocks": 256, "bmm_rows": 128, "robin_hood_groups": 1,
             "bmm_stream_blocks": 240, "bmm_tile_blocks": 2048}
    T.save(table, tmp_path / "t.json")
    loaded = T.load(tmp_path / "t.json")
    assert loaded == {"far_range_fp64": 1280, "far_range_mixed": 5120, "gemv_blocks": 1024, "bmm_blocks": 2560,
                      "near_blocks": 256, "bmm_rows": 128, "robin_hood_groups": 1, "bmm_stream_blocks": 240,
                      "bmm_tile_blocks": 2048}
    T.apply(ctx, loaded)
    assert T.current(ctx) == loaded
    with pytest.raises(_lib.KemError):
        T.set_knob(ctx, "gemv_blocks", 0)
    with pytest.raises(_lib.KemError):
        ctx.check(ctx.lib.kem_ctx_set_tuning(ctx.h, 99, 1))
    with pytest.raises(_lib.KemError):
        T.set_knob(ctx, "near_blocks", -1)
    with pytest.raises(_lib.KemError):
        T.set_knob(ctx, "bmm_rows", 96)
    with pytest.raises(_lib.KemError):
        T.set_knob(ctx, "robin_hood_groups", 2)
    with pytest.raises(_lib.KemError):
        T.set_knob(ctx, "bmm_stream_blocks", 0)
    with pytest.raises(_lib.KemError):
        T.set_knob(ctx, "bmm_tile_blocks", 0)
    T.apply(ctx, {"far_range_fp64": 2560, "far_range_mixed": 2560, "gemv_blocks": 2048, "near_blocks": 384,
                  "bmm_rows": 64, "robin_hood_groups": 0, "bmm_stream_blocks": 480, "bmm_tile_blocks": 4096})


def test_tune_near(setup):
    """The near pass's grid beside the mixed far pass: a sweep over the whole field pass, and the
    field is bit-identical for every grid (each point's near sum runs in the same order)."""
    from kassiopeia_amd import autotune as T
    from kassiopeia_amd import kemfield as K

    ctx, m, cont, A, x = setup
    pts = G.points_annulus(300_000, 0.5, 1.0, 2.0, 5)
    r = T.tune_near(cont, x, pts, repeats=1, candidates=(0, 7, 384))
    assert set(r["ms"]) == {0, 7, 384} and T.get(ctx, "near_blocks") == r["best"]
    out = {}
    for v in (0, 7, 384):
        T.set_knob(ctx, "near_blocks", v)
        out[v] = K.KIntegratingElectrostaticFieldSolver(cont, x, "mixed").ElectricFieldAndPotential(pts)
    for v in (7, 384):
        assert np.array_equal(out[v][0], out[0][0]) and np.array_equal(out[v][1], out[0][1])
    T.set_knob(ctx, "near_blocks", 384)
