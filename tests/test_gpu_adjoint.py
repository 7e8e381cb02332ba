"""Adjoint of the BEM path on the GPU (no reference counterpart): vertex and charge-density
derivatives of the potential at points and of the collocation matrix, the transposed solve,
and the full voltage / vertex gradient of a potential loss, checked against central finite
differences of the forward pass (h = 1e-6 relative, quadrature regimes frozen) and between
the FP64 and mixed-precision adjoints.
"""
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


def one_triangle(p):
    recs = G.records_from_points(p[0], p[1], p[2])
    return G.Mesh(np.zeros(1, np.int32), recs, np.zeros(1, np.int32), np.zeros(1), np.zeros(1),
                  np.zeros(1, np.int32))


def element_phi(ctx, p, P):
    from kassiopeia_amd import kemfield as K

    cont = K.KSurfaceContainer(ctx, one_triangle(p))
    E, phi = K.KElectrostaticBoundaryIntegrator(cont, "numeric").ElectricFieldAndPotential([0], [P])
    return phi[0]


@pytest.mark.parametrize("dr", [0.2, 1.0, 2.5, 5.0, 50.0, 200.0])
def test_field_vjp_single_pair(ctx, dr):
    """dsigma = G(P; T) and dG/dvertices against central differences of the element integral,
    in the closed-form (dr <= 3.42, dual numbers) and cubature (33 / 12 / 7-point) regimes."""
    from kassiopeia_amd import kemfield as K

    rng = np.random.default_rng(int(dr * 10))
    p = rng.normal(size=(3, 3))
    cont = K.KSurfaceContainer(ctx, one_triangle(p))
    c = p.mean(0)
    size = (np.linalg.norm(p[1] - p[0]) + np.linalg.norm(p[2] - p[1]) + np.linalg.norm(p[0] - p[2])) / 3
    d = rng.normal(size=3)
    P = c + d / np.linalg.norm(d) * dr * size
    sigma = np.array([1.0])
    ds, dv = np.zeros(1), np.zeros(9)
    ctx.check(ctx.lib.kem_field_vjp(ctx.h, cont.h, K._ptr(sigma), 1, K._ptr(np.ascontiguousarray(P)),
                                    K._ptr(np.ones(1)), K._ptr(ds), K._ptr(dv)))
    G0 = element_phi(ctx, p, P)
    assert abs(ds[0] - G0) <= 1e-12 * abs(G0)
    h = 1e-6 * size
    fd = np.zeros(9)
    for m in range(3):
        for k in range(3):
            pp, pm = p.copy(), p.copy()
            pp[m, k] += h
            pm[m, k] -= h
            fd[3 * m + k] = (element_phi(ctx, pp, P) - element_phi(ctx, pm, P)) / (2 * h)
    assert np.abs(dv - fd).max() <= 1e-6 * np.abs(fd).max(), (dv, fd)
    # translation invariance: moving every vertex by t equals moving P by -t, so the vertex
    # gradients sum to -grad_P G = E(P) of the element
    E, _ = K.KElectrostaticBoundaryIntegrator(cont, "numeric").ElectricFieldAndPotential([0], [P])
    assert np.linalg.norm(dv.reshape(3, 3).sum(0) - E[0]) <= 1e-9 * np.linalg.norm(E[0])


@pytest.mark.parametrize("where", ["before", "beyond", "inside"])
def test_field_vjp_point_over_edge_line(ctx, where):
    """Points whose projection lies on an edge line (t_e = 0: the edge term vanishes but its
    derivative does not), e.g. axis points over the radial edges of a ring."""
    from kassiopeia_amd import kemfield as K

    p = np.array([[0.2, 0.0, 0.0], [0.3, 0.0, 0.0], [0.25, 0.15, 0.0]])
    cont = K.KSurfaceContainer(ctx, one_triangle(p))
    u = p[1] - p[0]
    P = {"before": p[0] - 0.8 * u, "beyond": p[1] + 0.8 * u, "inside": p[0] + 0.5 * u}[where] + [0, 0, 0.07]
    ds, dv = np.zeros(1), np.zeros(9)
    ctx.check(ctx.lib.kem_field_vjp(ctx.h, cont.h, K._ptr(np.ones(1)), 1, K._ptr(np.ascontiguousarray(P)),
                                    K._ptr(np.ones(1)), K._ptr(ds), K._ptr(dv)))
    h = 1e-7
    fd = np.zeros(9)
    for m in range(3):
        for k in range(3):
            pp, pm = p.copy(), p.copy()
            pp[m, k] += h
            pm[m, k] -= h
            fd[3 * m + k] = (element_phi(ctx, pp, P) - element_phi(ctx, pm, P)) / (2 * h)
    assert np.abs(dv - fd).max() <= 1e-6 * np.abs(fd).max(), (dv, fd)


def test_matrix_vjp_fd(ctx):
    """-lambda^T (dA/dX) sigma against central differences of lambda^T A(X) sigma, on a small
    patch (self, edge-adjacent and far pairs; centroids move with their triangles)."""
    from kassiopeia_amd import kemfield as K

    X, tri, el = G.ring_stack_indexed(nrings=2, r_min=0.3, r_max=0.5, dz=0.2, width=0.1, n_target=64)
    n = tri.shape[0]
    rng = np.random.default_rng(5)
    lam, sig = rng.normal(size=n), rng.normal(size=n)

    def quad(Xv):
        m = G.Mesh(np.zeros(n, np.int32), G.records_from_points(Xv[tri[:, 0]], Xv[tri[:, 1]], Xv[tri[:, 2]]),
                   np.zeros(n, np.int32), np.zeros(n), np.zeros(n), np.zeros(n, np.int32))
        cont = K.KSurfaceContainer(ctx, m)
        return cont, lam @ K.KBoundaryIntegralMatrix(cont).to_numpy() @ sig

    cont, _ = quad(X)
    dvm = np.zeros((n, 9))
    ctx.check(ctx.lib.kem_assemble_vjp(ctx.h, cont.h, K._ptr(lam), K._ptr(sig), K._ptr(dvm)))
    dX = np.zeros_like(X)
    for m in range(3):
        np.add.at(dX, tri[:, m], dvm.reshape(n, 3, 3)[:, m, :])
    dX = -dX  # the entry point returns -lambda^T dA sigma
    h = 1e-6
    for v in rng.choice(X.shape[0], 8, replace=False):
        for k in range(3):
            Xp, Xm = X.copy(), X.copy()
            Xp[v, k] += h
            Xm[v, k] -= h
            fd = (quad(Xp)[1] - quad(Xm)[1]) / (2 * h)
            assert abs(dX[v, k] - fd) <= 1e-6 * np.abs(dX).max(), (v, k, dX[v, k], fd)


def test_transposed_solve(ctx):
    from kassiopeia_amd import _lib
    from kassiopeia_amd import kemfield as K
    import ctypes

    m = G.cube_mesh(4)
    cont = K.KSurfaceContainer(ctx, m)
    A = K.KBoundaryIntegralMatrix(cont, fp64=True, fp32=True)
    Ah = A.to_numpy()
    c = np.random.default_rng(1).normal(size=m.n)
    for prec in (_lib.KEM_FP64, _lib.KEM_MIXED):
        o = _lib.SolveOpts()
        ctx.lib.kem_solve_opts_default(ctypes.byref(o))
        o.tolerance, o.precision = 1e-11, prec
        lam = np.zeros(m.n)
        st = _lib.SolveStats()
        assert ctx.check(ctx.lib.kem_solve_transposed(A.h, ctypes.byref(o), K._ptr(c), K._ptr(lam),
                                                      ctypes.byref(st))) == 0
        assert np.linalg.norm(Ah.T @ lam - c) / np.linalg.norm(c) < 1e-11


@pytest.fixture(scope="module")
def stack_problem(ctx):
    from kassiopeia_amd import adjoint as AD

    X, tri, el = G.ring_stack_indexed(n_target=600)
    z = np.linspace(-1.2, 1.2, 100)
    pts = np.stack([np.zeros_like(z), np.zeros_like(z), z], 1)
    V = np.random.default_rng(8).uniform(-1000.0, 0.0, 8)
    target = -500.0 * np.exp(-z ** 2)
    return AD, X, tri, el, pts, V, target


def loss_and_grad(AD, ctx, X, tri, el, pts, V, target, precision="fp64"):
    prob = AD.BEMProblem(ctx, tri, el, pts, nelectrodes=8, tolerance=1e-13, precision=precision)
    phi = prob.forward(V, X)
    r = phi - target
    L = float(np.mean(r ** 2))
    dV, dX = prob.vjp(2 * r / r.size, X.shape)
    return L, dV, dX, prob


def test_adjoint_gradient_vs_fd(ctx, stack_problem):
    AD, X, tri, el, pts, V, target = stack_problem
    L, dV, dX, prob = loss_and_grad(AD, ctx, X, tri, el, pts, V, target)

    def loss(Vv, Xv):
        phi = prob.forward(Vv, Xv)
        return float(np.mean((phi - target) ** 2))

    for e in range(8):
        h = 1e-6 * abs(V[e])
        Vp, Vm = V.copy(), V.copy()
        Vp[e] += h
        Vm[e] -= h
        fd = (loss(Vp, X) - loss(Vm, X)) / (2 * h)
        assert abs(dV[e] - fd) <= 1e-6 * np.abs(dV).max(), (e, dV[e], fd)
    rng = np.random.default_rng(3)
    for v in rng.choice(X.shape[0], 6, replace=False):
        for k in range(3):
            h = 1e-6
            Xp, Xm = X.copy(), X.copy()
            Xp[v, k] += h
            Xm[v, k] -= h
            fd = (loss(V, Xp) - loss(V, Xm)) / (2 * h)
            assert abs(dX[v, k] - fd) <= 1e-5 * np.abs(dX).max(), (v, k, dX[v, k], fd)


def test_adjoint_mixed_matches_fp64(ctx, stack_problem):
    AD, X, tri, el, pts, V, target = stack_problem
    L, dV, dX, _ = loss_and_grad(AD, ctx, X, tri, el, pts, V, target)
    Lm, dVm, dXm, _ = loss_and_grad(AD, ctx, X, tri, el, pts, V, target, precision="mixed")
    assert abs(Lm - L) <= 1e-9 * L
    assert np.abs(dVm - dV).max() <= 1e-6 * np.abs(dV).max()
    assert np.abs(dXm - dX).max() <= 1e-6 * np.abs(dX).max()


def test_torch_autograd(ctx, stack_problem):
    AD, X, tri, el, pts, V, target = stack_problem
    import torch

    Vt = torch.tensor(V, dtype=torch.float64, requires_grad=True)
    Xt = torch.tensor(X, dtype=torch.float64, requires_grad=True)
    prob = AD.BEMProblem(ctx, tri, el, pts, nelectrodes=8, tolerance=1e-13)
    phi = AD.potential_fn().apply(Vt, Xt, prob)
    loss = ((phi - torch.from_numpy(target)) ** 2).mean()
    loss.backward()
    L, dV, dX, _ = loss_and_grad(AD, ctx, X, tri, el, pts, V, target)
    # the same computation up to the rounding of dL/dphi (torch forms 2 (phi - t) / m its own way)
    assert np.abs(Vt.grad.numpy() - dV).max() <= 1e-12 * np.abs(dV).max()
    assert np.abs(Xt.grad.numpy() - dX).max() <= 1e-12 * np.abs(dX).max()
