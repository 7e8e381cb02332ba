"""Matrix storage handed back to the context on kem_matrix_destroy and reused by the next
assembly of the same size: a matrix assembled into reused storage equals, bit for bit, the same
matrix assembled into fresh storage on another context (every entry rewritten, nothing of the
previous matrix left), in FP64, FP32 and BF16, with a matrix of another size in between."""
import numpy as np
import pytest

from kassiopeia_amd import geometry as G

pytestmark = pytest.mark.gpu


def _scaled(m, f):
    recs = m.recs.copy()
    recs[:, :2] *= f  # side lengths
    recs[:, 2:5] *= f  # first vertex
    return G.Mesh(m.shapes.copy(), recs, m.bc_type.copy(), m.bc_value.copy(), m.bc_flux.copy(), m.group.copy())


def test_matrix_storage_reuse():
    from kassiopeia_amd import kemfield as K

    m1 = G.cube_mesh(8)
    m2 = _scaled(m1, 1.5)
    mo = G.cube_mesh(6)  # another size
    ctx, fresh = K.Context(0), K.Context(0)
    try:
        X = np.random.default_rng(4).normal(size=(m1.n, 5))
        c1, c2, co = (K.KSurfaceContainer(ctx, m) for m in (m1, m2, mo))
        A1 = K.KBoundaryIntegralMatrix(c1, fp64=True, fp32=True, bf16=True)
        A1.to_numpy()
        A1.close()
        Ao = K.KBoundaryIntegralMatrix(co, fp64=True, fp32=True)
        A2 = K.KBoundaryIntegralMatrix(c2, fp64=True, fp32=True, bf16=True)  # the storage of A1
        Ao.close()
        f2 = K.KSurfaceContainer(fresh, m2)
        R2 = K.KBoundaryIntegralMatrix(f2, fp64=True, fp32=True, bf16=True)
        assert np.array_equal(A2.to_numpy(), R2.to_numpy())
        for prec in ("mixed", "bf16", "fp64"):
            assert np.array_equal(A2.MultiplyBatch(X, prec), R2.MultiplyBatch(X, prec)), prec
        A2.close()
        R2.close()
        for c in (c1, c2, co, f2):
            c.close()
    finally:
        ctx.close()
        fresh.close()
