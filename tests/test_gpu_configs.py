"""The benchmark configurations at their full sizes, on the GPU.

* 20,000-triangle coax (-18.6 kV / 0 V, 1 % jitter): sampled matrix entries and the FP64 field
  against the oracle at 1e-12, GMRES to 1e-10, mixed vs FP64 over 1,000,000 points at 1e-6.
* The same mesh in 16 axial ring groups with 64 voltage sets (U[-1000, 0] V, seeded): lockstep
  GMRES on FP32 and BF16 MFMA block products with FP64 refinement, every column to 1e-10; the
  field of the 64 sets over 4,000,000 points in uneven batches (sizes log-uniform in 1k-256k),
  one pass, against each batch's own single-set pass.
* The 5,000-triangle stack of 8 annular rings (radii 0.2-1.0 m, dz 0.25 m, voltages U[-1, 0] kV)
  with a mean-squared potential loss on 10,000 on-axis points: voltage and vertex gradients
  of the adjoint against central finite differences (h = 1e-6 relative) and the FP64 gradient.
* 100,000 triangles (32 rings + a grounded r = 2 m cylinder), FP64 (80 GB) and FP32 (40 GB)
  matrices on the device, 16 voltage sets, 10,000,000 bore points: property checks (per-set
  FP64 residual, mixed vs FP64, oracle sample, voltage gradient against forward sensitivities).
"""
import numpy as np
import pytest

from kassiopeia_amd import geometry as G

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def ctx():
    from kassiopeia_amd import kemfield as K

    c = K.Context(0)
    c.set_integrator("numeric")
    yield c
    c.close()


def field_errors(E, phi, Em, phim, vscale):
    """Norm-wise field error, largest field deviation over the RMS field, potential over the
    voltage scale."""
    dE = np.linalg.norm(Em - E, axis=1)
    return (float(np.linalg.norm(Em - E) / np.linalg.norm(E)),
            float(dE.max() / np.sqrt((E ** 2).sum(1).mean())),
            float(np.abs(phim - phi).max() / vscale))


def check_oracle_sample(oracle, mesh, sigma, pts, E, phi):
    """FP64 field at a point sample against the oracle: |dphi| and |dE| within 1e-12 of the
    sample's largest |phi| and |E|."""
    import os

    phr, Er = oracle.field(mesh.shapes, mesh.recs, sigma, pts, threads=max(1, min(16, os.cpu_count() or 1)))
    assert np.abs(phi - phr).max() <= TOL * np.abs(phr).max()
    assert np.linalg.norm(E - Er, axis=1).max() <= TOL * np.linalg.norm(Er, axis=1).max()


def residuals(A, X, B):
    """Per-column relative residual ||b - A x|| / ||b|| with the FP64 matrix on the device."""
    AX = A.MultiplyBatch(X, "fp64")
    return np.linalg.norm(B - AX, axis=0) / np.linalg.norm(B, axis=0)


# ------------------------------------------------------------------ 20,000-triangle coax
@pytest.fixture(scope="module")
def coax20k(ctx):
    from kassiopeia_amd import kemfield as K

    m = G.coax_mesh()
    assert m.n == 20000
    cont = K.KSurfaceContainer(ctx, m)
    A = K.KBoundaryIntegralMatrix(cont, fp64=True, fp32=True, bf16=True)
    yield m, cont, A
    A.close()
    cont.close()


def test_coax20k_matrix_entries(coax20k, oracle):
    m, cont, A = coax20k
    rng = np.random.default_rng(11)
    idx = np.concatenate([np.stack([np.arange(0, m.n, 97)] * 2, 1), rng.integers(0, m.n, size=(1500, 2))])
    ref = np.array([oracle.matrix_element(int(i), int(j), m.shapes, m.recs) for i, j in idx])
    got = np.array([A(int(i), int(j)) for i, j in idx])
    assert np.abs(got - ref).max() <= TOL * np.abs(ref).max()
    assert (np.abs(got - ref) / np.abs(ref)).max() <= 1e-11


def test_coax20k_solve_and_field(ctx, coax20k, oracle):
    from kassiopeia_amd import _lib
    from kassiopeia_amd import kemfield as K

    m, cont, A = coax20k
    b = cont.rhs()
    x64, st64, rc = A.solve(b, tolerance=1e-10)
    assert rc == 0 and st64.true_residual <= 1e-10
    assert residuals(A, x64[:, None], b[:, None])[0] <= 1e-10
    xmx, stmx, rc = A.solve(b, tolerance=1e-10, precision="mixed")
    assert rc == 0 and residuals(A, xmx[:, None], b[:, None])[0] <= 1e-10
    assert np.linalg.norm(xmx - x64) / np.linalg.norm(x64) < 1e-8
    pts = G.points_annulus(1_000_000, 0.5, 1.0, 2.0, seed=2024)
    M = pts.shape[0]
    X = [K.DeviceArray(ctx, pts[:, k]) for k in range(3)]
    S = K.DeviceArray(ctx, x64)
    outs = {}
    for prec in (_lib.KEM_FP64, _lib.KEM_MIXED):
        O = [K.DeviceArray.empty(ctx, M) for _ in range(4)]
        ctx.check(ctx.lib.kem_field_eval_device(ctx.h, cont.h, S.ptr, M, X[0].ptr, X[1].ptr, X[2].ptr, prec,
                                                O[0].ptr, O[1].ptr, O[2].ptr, O[3].ptr))
        ctx.sync()
        outs[prec] = (np.stack([O[k].get() for k in (1, 2, 3)], 1), O[0].get())
        for a in O:
            a.close()
    for a in X + [S]:
        a.close()
    E, phi = outs[_lib.KEM_FP64]
    Em, phim = outs[_lib.KEM_MIXED]
    l2, mx, ph = field_errors(E, phi, Em, phim, 18600.0)
    print(f"coax20k 1M points: mixed vs FP64 l2 {l2:.2e} max/rms {mx:.2e} phi/V {ph:.2e}", flush=True)
    assert l2 <= 1e-6 and mx <= 1e-6 and ph <= 1e-6
    sel = np.random.default_rng(5).choice(M, 256, replace=False)
    check_oracle_sample(oracle, m, x64, pts[sel], E[sel], phi[sel])


@pytest.mark.parametrize("precision", ["mixed", "bf16"])
def test_coax20k_64_voltage_sets(coax20k, precision):
    m, cont, A = coax20k
    groups = G.coax_ring_groups(m, 16)
    V = np.random.default_rng(64).uniform(-1000.0, 0.0, size=(16, 64))
    B = V[groups, :]
    X, res, st, rc = A.solve_batch(B, tolerance=1e-10, precision=precision)
    print(f"64 sets {precision}: {st.iterations} it, {st.refinements} ref, {st.seconds * 1e3:.1f} ms", flush=True)
    assert rc == 0 and res.max() <= 1e-10
    assert residuals(A, X, B).max() <= 1e-10
    # spot check: one column against the single-RHS FP64 solve
    x, _, _ = A.solve(B[:, 17], tolerance=1e-12)
    assert np.linalg.norm(X[:, 17] - x) / np.linalg.norm(x) < 1e-8


def test_coax20k_uneven_batches_4M(ctx, coax20k):
    from kassiopeia_amd import _lib
    from kassiopeia_amd import kemfield as K

    m, cont, A = coax20k
    groups = G.coax_ring_groups(m, 16)
    V = np.random.default_rng(64).uniform(-1000.0, 0.0, size=(16, 64))
    S, res, _, rc = A.solve_batch(V[groups, :], tolerance=1e-10, precision="mixed")
    assert rc == 0
    npb = 4_000_000
    rng = np.random.default_rng(4096)
    sizes = []
    while sum(sizes) < npb:
        sizes.append(int(np.exp(rng.uniform(np.log(1000), np.log(256000)))))
    sizes[-1] -= sum(sizes) - npb
    assert min(sizes) > 0 and sum(sizes) == npb
    pts = G.points_annulus(npb, 0.5, 1.0, 2.0, seed=77)
    sidx = (np.repeat(np.arange(len(sizes)) % 64, sizes)).astype(np.int32)
    off = np.concatenate([[0], np.cumsum(sizes)])
    Sd = K.DeviceArray(ctx, np.ascontiguousarray(S.T))
    Id = K.DeviceArray(ctx, sidx)
    Xd = [K.DeviceArray(ctx, pts[:, k]) for k in range(3)]
    # FP64: the same sums in another order; mixed: the several-set kernel and the single-set
    # cluster pass (3-point clusters, element ranges per XCD) each carry FP32 rounding of
    # ~1e-7 against FP64, so they agree to the mixed-precision bar
    for prec, tol in ((_lib.KEM_FP64, 1e-13), (_lib.KEM_MIXED, 1e-6)):
        Od = [K.DeviceArray.empty(ctx, npb) for _ in range(4)]
        ctx.check(ctx.lib.kem_field_eval_sets(ctx.h, cont.h, 64, Sd.ptr, Id.ptr, npb, Xd[0].ptr, Xd[1].ptr,
                                              Xd[2].ptr, prec, Od[0].ptr, Od[1].ptr, Od[2].ptr, Od[3].ptr))
        ctx.sync()
        phi = Od[0].get()
        E = np.stack([Od[k].get() for k in (1, 2, 3)], 1)
        for a in Od:
            a.close()
        pname = "fp64" if prec == _lib.KEM_FP64 else "mixed"
        worst = 0.0
        for b in range(len(sizes)):
            sl = slice(off[b], off[b + 1])
            Eb, pb = K.KIntegratingElectrostaticFieldSolver(cont, S[:, b % 64], pname).ElectricFieldAndPotential(
                pts[sl])
            worst = max(worst, np.abs(phi[sl] - pb).max() / np.abs(pb).max(),
                        np.abs(E[sl] - Eb).max() / np.abs(Eb).max())
        print(f"4M points in {len(sizes)} batches, {pname}: worst batch deviation {worst:.2e}", flush=True)
        assert worst <= tol
    for a in Xd + [Sd, Id]:
        a.close()


# ------------------------------------------------------------------ 5,000-triangle ring stack, adjoint
def test_ring_stack_adjoint_10k_axis(ctx):
    from kassiopeia_amd import adjoint as AD

    X, tri, el = G.ring_stack_indexed(n_target=5000)
    assert abs(tri.shape[0] - 5000) <= 100
    z = np.linspace(-1.5, 1.5, 10_000)
    pts = np.stack([np.zeros_like(z), np.zeros_like(z), z], 1)
    V = np.random.default_rng(44).uniform(-1000.0, 0.0, 8)
    target = -600.0 * np.exp(-(z / 0.8) ** 2) + 50.0 * np.sin(2.0 * z)

    def grad(precision):
        prob = AD.BEMProblem(ctx, tri, el, pts, nelectrodes=8, tolerance=1e-13, precision=precision)
        phi = prob.forward(V, X)
        r = phi - target
        dV, dX = prob.vjp(2 * r / r.size, X.shape)
        return float(np.mean(r ** 2)), dV, dX, prob

    L, dV, dX, prob = grad("fp64")

    def loss(Vv, Xv):
        return float(np.mean((prob.forward(Vv, Xv) - target) ** 2))

    for e in range(8):
        h = 1e-6 * abs(V[e])
        Vp, Vm = V.copy(), V.copy()
        Vp[e] += h
        Vm[e] -= h
        fd = (loss(Vp, X) - loss(Vm, X)) / (2 * h)
        assert abs(dV[e] - fd) <= 1e-6 * np.abs(dV).max(), (e, dV[e], fd)
    rng = np.random.default_rng(12)
    scale = np.abs(X).max()
    for v in rng.choice(X.shape[0], 4, replace=False):
        for k in range(3):
            h = 1e-6 * scale
            Xp, Xm = X.copy(), X.copy()
            Xp[v, k] += h
            Xm[v, k] -= h
            fd = (loss(V, Xp) - loss(V, Xm)) / (2 * h)
            assert abs(dX[v, k] - fd) <= 1e-5 * np.abs(dX).max(), (v, k, dX[v, k], fd)
    Lm, dVm, dXm, _ = grad("mixed")
    assert abs(Lm - L) <= 1e-9 * L
    assert np.abs(dVm - dV).max() <= 1e-6 * np.abs(dV).max()
    assert np.abs(dXm - dX).max() <= 1e-6 * np.abs(dX).max()


# ------------------------------------------------------------------ 100,000 triangles, 10 M points
def test_bore100k_pipeline(ctx, oracle):
    import ctypes

    from kassiopeia_amd import _lib
    from kassiopeia_amd import kemfield as K

    rng = np.random.default_rng(20)
    volts = rng.uniform(-20000.0, 0.0, 32)
    m = G.ring_bore_mesh(volts=volts, n_target=100_000)
    assert m.n == 100_000
    cont = K.KSurfaceContainer(ctx, m)
    A = K.KBoundaryIntegralMatrix(cont, fp64=True, fp32=True)
    b = cont.rhs()
    x64, st64, rc = A.solve(b, tolerance=1e-10)
    assert rc == 0 and st64.true_residual <= 1e-10
    xmx, stmx, rc = A.solve(b, tolerance=1e-10, precision="mixed")
    assert rc == 0 and residuals(A, xmx[:, None], b[:, None])[0] <= 1e-10
    # 16 voltage sets in lockstep (FP32 MFMA block products, FP64 refinement)
    V16 = rng.uniform(-20000.0, 0.0, size=(33, 16))
    V16[32] = 0.0
    B16 = V16[m.group, :]
    X16, res16, st16, rc = A.solve_batch(B16, tolerance=1e-10, precision="mixed")
    assert rc == 0 and residuals(A, X16, B16).max() <= 1e-10
    print(f"bore100k: FP64 GMRES {st64.iterations} it, mixed {stmx.iterations} it / {stmx.refinements} ref, "
          f"16 sets {st16.iterations} it", flush=True)
    # 10 M points in the bore, both precisions
    pts = G.points_annulus(10_000_000, 0.0, 1.8, 3.5, seed=55)
    M = pts.shape[0]
    Xd = [K.DeviceArray(ctx, pts[:, k]) for k in range(3)]
    S = K.DeviceArray(ctx, x64)
    outs = {}
    for prec in (_lib.KEM_FP64, _lib.KEM_MIXED):
        O = [K.DeviceArray.empty(ctx, M) for _ in range(4)]
        ctx.check(ctx.lib.kem_field_eval_device(ctx.h, cont.h, S.ptr, M, Xd[0].ptr, Xd[1].ptr, Xd[2].ptr, prec,
                                                O[0].ptr, O[1].ptr, O[2].ptr, O[3].ptr))
        ctx.sync()
        outs[prec] = (np.stack([O[k].get() for k in (1, 2, 3)], 1), O[0].get())
        for a in O:
            a.close()
    for a in Xd + [S]:
        a.close()
    E, phi = outs[_lib.KEM_FP64]
    Em, phim = outs[_lib.KEM_MIXED]
    l2, mx, ph = field_errors(E, phi, Em, phim, 20000.0)
    print(f"bore100k 10M points: mixed vs FP64 l2 {l2:.2e} max/rms {mx:.2e} phi/V {ph:.2e}", flush=True)
    assert l2 <= 1e-6 and mx <= 1e-6 and ph <= 1e-6
    sel = np.random.default_rng(6).choice(M, 48, replace=False)
    check_oracle_sample(oracle, m, x64, pts[sel], E[sel], phi[sel])
    # voltage gradient of an axial loss: the adjoint (field VJP + transposed solve) against the
    # forward sensitivities dphi/dV_e = G A^-1 b_e of two electrodes (the potential is linear
    # in the voltages, so these are exact up to the solve tolerance)
    zax = np.linspace(-3.5, 3.5, 10_000)
    pax = np.ascontiguousarray(np.stack([0 * zax, 0 * zax, zax], 1))
    _, phax = K.KIntegratingElectrostaticFieldSolver(cont, x64).ElectricFieldAndPotential(pax)
    g = 2.0 * (phax + 10000.0) / phax.size
    n = m.n
    dsig = np.zeros(n)
    dvf = np.zeros((n, 9))
    ctx.check(ctx.lib.kem_field_vjp(ctx.h, cont.h, K._ptr(x64), pax.shape[0], K._ptr(pax), K._ptr(g),
                                    K._ptr(dsig), K._ptr(dvf)))
    o = _lib.SolveOpts()
    ctx.lib.kem_solve_opts_default(ctypes.byref(o))
    o.tolerance = 1e-12
    lam = np.zeros(n)
    sta = _lib.SolveStats()
    ctx.check(ctx.lib.kem_solve_transposed(A.h, ctypes.byref(o), K._ptr(dsig), K._ptr(lam), ctypes.byref(sta)))
    dV = np.bincount(m.group, weights=lam, minlength=33)
    for e in (3, 20):
        be = (m.group == e).astype(np.float64)
        se, _, rc = A.solve(be, tolerance=1e-12)
        assert rc == 0
        _, dphi = K.KIntegratingElectrostaticFieldSolver(cont, se).ElectricFieldAndPotential(pax)
        ref = float(g @ dphi)
        assert abs(dV[e] - ref) <= 1e-8 * np.abs(dV).max(), (e, dV[e], ref)
    A.close()
    cont.close()
