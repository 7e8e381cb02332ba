"""The no-allocation field entry point (kem_field_eval_device_noalloc) against the eager one
(kem_field_eval_device), through the C-ABI: the same kernels and sums, so bit-equal outputs,
on two calls in a row, the second with a smaller input (the static scratch keeps the first
call's contents), in both precisions and on a mesh with symmetry groups; sizes beyond the
scratch are refused before any device work."""
import ctypes

import numpy as np
import pytest

from kassiopeia_amd import geometry as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from kassiopeia_amd import kemfield as K

    c = K.Context(0)
    c.set_integrator("numeric")
    yield c
    c.close()


def run(ctx, cont, S, pts, precision, noalloc):
    from kassiopeia_amd import kemfield as K

    m = pts.shape[0]
    X = [K.DeviceArray(ctx, np.ascontiguousarray(pts[:, k])) for k in range(3)]
    O = [K.DeviceArray(ctx, np.full(m, np.nan)) for _ in range(4)]
    fn = ctx.lib.kem_field_eval_device_noalloc if noalloc else ctx.lib.kem_field_eval_device
    try:
        ctx.check(fn(ctx.h, cont.h, S.ptr, m, X[0].ptr, X[1].ptr, X[2].ptr, precision, O[0].ptr, O[1].ptr, O[2].ptr,
                     O[3].ptr))
        ctx.sync()
        return np.stack([o.get() for o in O], 1)
    finally:
        for a in X + O:
            a.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_matches_eager_twice(ctx, precision):
    from kassiopeia_amd import kemfield as K

    mesh = G.coax_mesh(nphi_in=12, nphi_out=24, nz=16, ncap_r=2, ncap_phi=16)
    cont = K.KSurfaceContainer(ctx, mesh)
    sigma = np.random.default_rng(1).normal(size=mesh.n) * 1e-9
    S = K.DeviceArray(ctx, sigma)
    for npts, seed in ((65536, 3), (5000, 4), (1000, 5)):
        pts = G.points_annulus(npts, 0.5, 1.0, 2.0, seed=seed)
        print(f"noalloc: precision {precision}, {npts} points", flush=True)
        ref = run(ctx, cont, S, pts, precision, False)
        got = run(ctx, cont, S, pts, precision, True)
        assert np.all(np.isfinite(got))
        assert np.array_equal(got, ref)
    S.close()
    cont.close()


def test_noalloc_groups(ctx):
    """symmetry groups: the members' charge densities are expanded into the scratch"""
    from kassiopeia_amd import kemfield as K

    m = G.mirror_groups(G.cube_mesh(4), [0.0, 0.0, 2.0], [0.0, 0.0, 1.0])
    cont = K.KSurfaceContainer(ctx, m)
    S = K.DeviceArray(ctx, np.linspace(1.0, 2.0, cont.size()) * 1e-10)
    for npts, seed in ((20000, 8), (300, 9)):
        pts = G.points_shell(npts, 0.9, 3.0, seed)
        ref = run(ctx, cont, S, pts, 0, False)
        got = run(ctx, cont, S, pts, 0, True)
        assert np.array_equal(got, ref)
    S.close()
    cont.close()


def test_noalloc_rejects_beyond_scratch(ctx):
    from kassiopeia_amd import kemfield as K

    cont = K.KSurfaceContainer(ctx, G.cube_mesh(2))
    # argument checks only: the call returns before touching these (never dereferenced)
    fake = ctypes.c_void_p(256)
    rc = ctx.lib.kem_field_eval_device_noalloc(ctx.h, cont.h, fake, 65537, fake, fake, fake, 0, fake, fake, fake,
                                               fake)
    assert rc == -1
    assert b"65536" in ctx.lib.kem_last_error(ctx.h)
    cont.close()
