"""Multi-right-hand-side path on the GPU: block matrix products (FP64 vector ALU and FP32
MFMA) and the lockstep GMRES solve with FP64 refinement, through the C-ABI.

The reference solves one right-hand side per charge-density solve
(KKrylovChargeDensitySolver.cc:42-78); the batched solve is checked against single-RHS
solves of the same matrix, against a dense LU solve, and by the recomputed FP64 residual.
"""
import numpy as np
import pytest

from kassiopeia_amd import geometry as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from kassiopeia_amd import kemfield as K

    c = K.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def coax(ctx):
    """1,280-triangle coax split into 16 axial ring groups, both matrix storages."""
    from kassiopeia_amd import kemfield as K

    m = G.coax_mesh(nphi_in=12, nphi_out=24, nz=16, ncap_r=2, ncap_phi=16)
    ctx.set_integrator("numeric")
    cont = K.KSurfaceContainer(ctx, m)
    A = K.KBoundaryIntegralMatrix(cont, fp64=True, fp32=True)
    return m, cont, A, A.to_numpy()


@pytest.mark.parametrize("k", [1, 5, 16, 17, 33, 64, 70])
def test_block_product_fp64(coax, k):
    m, cont, A, Ah = coax
    X = np.random.default_rng(k).normal(size=(m.n, k))
    Y = A.MultiplyBatch(X, "fp64")
    ref = Ah @ X
    assert np.abs(Y - ref).max() / np.abs(ref).max() < 1e-13
    if k == 1:
        assert np.abs(Y[:, 0] - A.Multiply(X[:, 0])).max() / np.abs(ref).max() < 1e-13


def _first_elements(m, n):
    return G.Mesh(m.shapes[:n].copy(), m.recs[:n].copy(), m.bc_type[:n].copy(), m.bc_value[:n].copy(),
                  m.bc_flux[:n].copy(), m.group[:n].copy())


@pytest.mark.parametrize("n", [1352, 1349])
def test_block_product_streamed_shapes(ctx, n):
    """33-64 right-hand sides: FP32 matrices with n % 4 == 0 and BF16 matrices stream through
    LDS (256-row workgroups; chunks of 16 FP32 columns or 4 BF16 column groups), FP32 with
    n % 4 != 0 loads into registers.  n = 1,352 leaves 72 rows past the last full workgroup and
    partial last chunks; the workgroup target 1 gives one column split, 480 five, 4,096 the
    most (one per 256 columns).  Exact check: unit right-hand sides e_j pick columns j of the
    stored matrix bit for bit (one nonzero term, products by 1 and sums with 0 exact) at 64
    columns spread over every chunk and split, first and last included.  Random right-hand
    sides against the FP64 product of the rounded operands, within the recursive-summation
    bound (n + 2) 2^-24 of the sum of |terms|."""
    from kassiopeia_amd import autotune as T
    from kassiopeia_amd import kemfield as K

    m = _first_elements(G.coax_mesh(nphi_in=12, nphi_out=24, nz=16, ncap_r=2, ncap_phi=25), n)
    cont = K.KSurfaceContainer(ctx, m)
    A = K.KBoundaryIntegralMatrix(cont, fp64=True, fp32=True, bf16=True)
    Ah = A.to_numpy()

    def rnd16(a):
        u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
        u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
        return u.astype(np.uint32).view(np.float32).astype(np.float64)

    rounded = {"mixed": Ah.astype(np.float32).astype(np.float64), "bf16": rnd16(Ah.astype(np.float32))}
    cols = np.linspace(0, n - 1, 64).astype(int)
    E = np.zeros((n, 64))
    E[cols, np.arange(64)] = 1.0
    X = np.random.default_rng(n).normal(size=(n, 64))
    Xr = {"mixed": X.astype(np.float32).astype(np.float64), "bf16": rnd16(X)}
    bound = (n + 2) * 2.0 ** -24
    try:
        for blocks in (1, 480, 4096):
            T.set_knob(ctx, "bmm_stream_blocks", blocks)
            for k in (33, 64):
                for prec in ("mixed", "bf16"):
                    print(f"streamed product: n {n} blocks {blocks} k {k} {prec}", flush=True)
                    Y = A.MultiplyBatch(E[:, :k], prec)
                    assert np.array_equal(Y, rounded[prec][:, cols[:k]]), (prec, blocks, k)
                    Y = A.MultiplyBatch(X[:, :k], prec)
                    Ar, Xk = rounded[prec], Xr[prec][:, :k]
                    assert np.all(np.abs(Y - Ar @ Xk) <= bound * (np.abs(Ar) @ np.abs(Xk))), (prec, blocks, k)
    finally:
        T.set_knob(ctx, "bmm_stream_blocks", 480)
        A.close()
        cont.close()


@pytest.mark.parametrize("n", [1352, 1350, 1349, 260, 8, 6, 4])
def test_block_product_tile_shapes(ctx, n):
    """Block products on 16 x 16 MFMA tiles streamed through LDS: FP32 matrices with n % 4 == 0
    and up to 32 right-hand sides (16-column panels up to 16, 32 columns above), FP64 matrices
    with n even (16 or 64 columns); the other shapes (n = 1,350 FP32, n = 1,349 both) take the
    register-loaded kernels.  Chunks of 16 FP32 or 8 FP64 columns; the workgroup target 1 gives
    one column split, 1,024 and 4,096 more.  Exact check: unit right-hand sides pick columns of
    the stored matrix bit for bit.  Random right-hand sides against the FP64 product of the
    rounded operands within the recursive-summation bound (n + 2) u of the sum of |terms|
    (u = 2^-24 for FP32, 2^-53 for FP64)."""
    from kassiopeia_amd import autotune as T
    from kassiopeia_amd import kemfield as K

    m = _first_elements(G.coax_mesh(nphi_in=12, nphi_out=24, nz=16, ncap_r=2, ncap_phi=25), n)
    cont = K.KSurfaceContainer(ctx, m)
    A = K.KBoundaryIntegralMatrix(cont, fp64=True, fp32=True)
    Ah = A.to_numpy()
    rounded = {"mixed": Ah.astype(np.float32).astype(np.float64), "fp64": Ah}
    unit = {"mixed": 2.0 ** -24, "fp64": 2.0 ** -53}
    cols = np.linspace(0, n - 1, 64).astype(int)
    E = np.zeros((n, 64))
    E[cols, np.arange(64)] = 1.0
    X = np.random.default_rng(n + 7).normal(size=(n, 64))
    Xr = {"mixed": X.astype(np.float32).astype(np.float64), "fp64": X}
    cases = [("mixed", k) for k in (2, 5, 16, 17, 32)] + [("fp64", k) for k in (1, 5, 16, 17, 64)]
    # matrices smaller than one workgroup's 256 rows and one chunk's columns: clamped loads
    try:
        for blocks in (1, 1024, 4096):
            T.set_knob(ctx, "bmm_tile_blocks", blocks)
            for prec, k in cases:
                print(f"tile product: n {n} blocks {blocks} k {k} {prec}", flush=True)
                Y = A.MultiplyBatch(E[:, :k], prec)
                assert np.array_equal(Y, rounded[prec][:, cols[:k]]), (prec, blocks, k)
                Y = A.MultiplyBatch(X[:, :k], prec)
                Ar, Xk = rounded[prec], Xr[prec][:, :k]
                # FP64: the reference product's own rounding doubles the bound
                bound = (n + 2) * unit[prec] * (2 if prec == "fp64" else 1)
                assert np.all(np.abs(Y - Ar @ Xk) <= bound * (np.abs(Ar) @ np.abs(Xk))), (prec, blocks, k)
    finally:
        T.set_knob(ctx, "bmm_tile_blocks", 4096)
        A.close()
        cont.close()


@pytest.mark.parametrize("k", [1, 7, 32, 33, 64, 70])
def test_block_product_fp32_mfma(coax, k):
    """FP32 MFMA product against the FP64 product of the FP32-rounded operands: FP32 products
    summed in FP32 over column splits, then in FP64 (tolerance 2e-6 of the column scale)."""
    m, cont, A, Ah = coax
    X = np.random.default_rng(100 + k).normal(size=(m.n, k))
    Y = A.MultiplyBatch(X, "mixed")
    A32 = Ah.astype(np.float32).astype(np.float64)
    ref = A32 @ X.astype(np.float32).astype(np.float64)
    scale = np.abs(A32) @ np.abs(X)
    assert (np.abs(Y - ref) / scale).max() < 2e-6
    # every column and row is covered (no stale or missing tile)
    assert np.all(np.abs(Y - ref) <= 2e-6 * scale)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_solve_batch_voltage_sets(coax, precision):
    """16 ring groups x 12 voltage sets (U[-1000, 0] V per group): every column converges to
    1e-10, matches the dense solution and the single-RHS GMRES solve of the same column."""
    m, cont, A, Ah = coax
    groups = G.coax_ring_groups(m, 16)
    V = np.random.default_rng(7).uniform(-1000.0, 0.0, size=(16, 12))
    B = V[groups, :]
    X, res, st, rc = A.solve_batch(B, tolerance=1e-10, precision=precision)
    assert rc == 0
    assert res.max() <= 1e-10
    assert np.all(np.linalg.norm(Ah @ X - B, axis=0) / np.linalg.norm(B, axis=0) <= 1e-10)
    Xd = np.linalg.solve(Ah, B)
    assert np.abs(X - Xd).max() / np.abs(Xd).max() < 1e-8
    x1, st1, rc1 = A.solve(B[:, 3], tolerance=1e-10, precision=precision)
    assert np.abs(X[:, 3] - x1).max() / np.abs(x1).max() < 1e-8
    if precision == "mixed":
        assert st.refinements >= 2


def test_solve_batch_more_than_64(coax):
    """Groups of 64 right-hand sides; zero columns stay zero; initial guesses are used."""
    m, cont, A, Ah = coax
    rng = np.random.default_rng(9)
    B = rng.normal(size=(m.n, 70))
    B[:, 5] = 0.0
    X0 = np.linalg.solve(Ah, B) * (1 + 1e-6)
    X, res, st, rc = A.solve_batch(B, X0=X0, tolerance=1e-10, precision="mixed")
    assert rc == 0
    assert np.all(X[:, 5] == 0.0)
    nz = np.linalg.norm(B, axis=0) > 0
    assert np.all(np.linalg.norm(Ah @ X - B, axis=0)[nz] / np.linalg.norm(B, axis=0)[nz] <= 1e-10)


def test_solve_batch_errors(coax):
    from kassiopeia_amd import _lib

    m, cont, A, Ah = coax
    with pytest.raises(_lib.KemError):
        A.solve_batch(np.ones((m.n, 2)), tolerance=-1.0)


@pytest.mark.parametrize("sizes", [[700, 1500, 123, 900], [1000, 5000, 123, 7000, 2500, 64]])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_field_sets_uneven_batches(coax, sizes, precision):
    """Uneven point batches, each with its own charge-density set, in one pass (unsorted below
    4096 points, Morton-sorted by (set, position) above): every batch equals the single-set
    evaluation of its points."""
    from kassiopeia_amd import kemfield as K

    m, cont, A, Ah = coax
    groups = G.coax_ring_groups(m, 16)
    V = np.random.default_rng(3).uniform(-1000.0, 0.0, size=(16, 3))
    S = np.linalg.solve(Ah, V[groups, :])
    pts = G.points_annulus(sum(sizes), 0.5, 1.0, 2.0, 77)
    bset = np.arange(len(sizes)) % 3
    set_index = np.repeat(bset, sizes).astype(np.int32)
    E, phi = K.field_sets(cont, S, set_index, pts, precision)
    off = np.concatenate([[0], np.cumsum(sizes)])
    for b in range(len(sizes)):
        sl = slice(off[b], off[b + 1])
        Eb, pb = K.KIntegratingElectrostaticFieldSolver(cont, S[:, bset[b]], precision).ElectricFieldAndPotential(
            pts[sl])
        tol = 1e-13 if precision == "fp64" else 1e-7
        assert np.abs(phi[sl] - pb).max() <= tol * np.abs(pb).max()
        assert np.abs(E[sl] - Eb).max() <= tol * np.abs(Eb).max()


def test_field_sets_rejects_bad_index(coax):
    from kassiopeia_amd import _lib
    from kassiopeia_amd import kemfield as K

    m, cont, A, Ah = coax
    S = np.ones((m.n, 2)) * 1e-9
    with pytest.raises(_lib.KemError):
        K.field_sets(cont, S, np.array([0, 1, 2], np.int32), np.ones((3, 3)) * 0.7)
