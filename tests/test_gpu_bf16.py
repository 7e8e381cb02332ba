"""BF16 matrix storage and the BF16 MFMA block product (v_mfma_f32_32x32x16_bf16) on the device:
the product equals the FP64 product of the BF16-rounded operands to FP32 accumulation accuracy
(and differs from the unrounded product by BF16 rounding, so the BF16 path is the one that
ran); lockstep GMRES on BF16 block products with FP64 refinement reaches the FP64 solution."""
import numpy as np
import pytest

from kassiopeia_amd import geometry as G

pytestmark = pytest.mark.gpu


def bf16_round(a):
    """Round to FP32, then to BF16 (nearest even on the FP32 bits), as float64."""
    f = np.asarray(a, dtype=np.float64).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


@pytest.fixture(scope="module")
def setup():
    from kassiopeia_amd import kemfield as K

    ctx = K.Context(0)
    m = G.coax_mesh(nphi_in=12, nphi_out=24, nz=16, ncap_r=2, ncap_phi=17)  # 1,288 triangles
    cont = K.KSurfaceContainer(ctx, m)
    A = K.KBoundaryIntegralMatrix(cont, fp64=True, fp32=True, bf16=True)
    yield ctx, m, cont, A
    A.close()
    cont.close()
    ctx.close()


@pytest.mark.parametrize("k", [1, 7, 32, 40, 64])
def test_bf16_block_product(setup, k):
    ctx, m, cont, A = setup
    print(f"bf16 product: n {A.n} k {k}", flush=True)
    A64 = A.to_numpy()
    X = np.random.default_rng(k).normal(size=(A.n, k))
    Y = A.MultiplyBatch(X, "bf16")
    Ab, Xb = bf16_round(A64), bf16_round(X)
    ref = Ab @ Xb
    scale = np.abs(Ab) @ np.abs(Xb)  # the sum of |terms| per entry bounds the accumulation error
    err = np.abs(Y - ref) / scale
    print(f"bf16 product: max error / sum|terms| {err.max():.2e}", flush=True)
    assert err.max() < 2e-6
    # the BF16 rounding is visible against the exact product
    exact = A64 @ X
    assert np.linalg.norm(Y - exact) / np.linalg.norm(exact) > 1e-5


def test_bf16_storage_combinations(setup):
    """BF16 storage beside FP64 only, and alone (packed from an FP32 matrix assembled for it and
    dropped): the same BF16 matrix (FP64 -> FP32 -> BF16 and FP32 -> BF16 round alike, the FP32
    matrix being the FP64 one rounded)."""
    from kassiopeia_amd import kemfield as K

    ctx, m, cont, A = setup
    X = np.random.default_rng(3).normal(size=(A.n, 8))
    ref = A.MultiplyBatch(X, "bf16")
    for fp64 in (True, False):
        B = K.KBoundaryIntegralMatrix(cont, fp64=fp64, fp32=False, bf16=True)
        assert np.array_equal(B.MultiplyBatch(X, "bf16"), ref)
        B.close()


def test_bf16_lockstep_solve(setup):
    ctx, m, cont, A = setup
    groups = G.coax_ring_groups(m, 16)
    V = np.random.default_rng(16).uniform(-1000.0, 0.0, size=(16, 16))
    B = V[groups, :]
    X64, res64, st64, _ = A.solve_batch(B, tolerance=1e-10, precision="fp64")
    Xb, resb, stb, rc = A.solve_batch(B, tolerance=1e-10, precision="bf16")
    print(f"bf16 lockstep: {stb.iterations} it, {stb.refinements} ref, worst residual {resb.max():.2e}; "
          f"fp64 {st64.iterations} it", flush=True)
    assert rc == 0 and resb.max() <= 1e-10
    assert np.linalg.norm(Xb - X64) / np.linalg.norm(X64) < 1e-8


def test_bf16_single_solve(setup):
    ctx, m, cont, A = setup
    b = cont.rhs() + np.linspace(0.0, 1.0, A.n)
    x64, s64, _ = A.solve(b, tolerance=1e-10)
    xb, sb, rc = A.solve(b, tolerance=1e-10, precision="bf16")
    print(f"bf16 single: {sb.iterations} it, {sb.refinements} ref, true residual {sb.true_residual:.2e}", flush=True)
    assert rc == 0 and sb.true_residual <= 1e-10
    assert np.linalg.norm(xb - x64) / np.linalg.norm(x64) < 1e-8
