"""The oracle on the reference's spherical-capacitor meshes (Dirichlet conductors and a Neumann
dielectric interface; fixtures copied from KEMField/data/input_files/sphericalCapacitorFiles):
the sphere charges and the potential converge to the closed forms of
ComputeSphericalCapacitor.cc with refinement.  The flat triangles inscribe the spheres, so the
charges approach the analytic values from below, halving the error per refinement step."""
import numpy as np
import pytest

import spherical as S


@pytest.fixture(scope="module")
def solved(oracle):
    out = {}
    for scale in (1, 2):
        m = S.mesh(scale)
        A = oracle.assemble(m.shapes, m.recs, m.bc_type, m.bc_flux)
        x = np.linalg.solve(A, oracle.rhs(m.bc_value, m.bc_type))
        out[scale] = (m, A, x)
    return out


def test_fixture_contents():
    m = S.mesh(1)
    assert m.n == 156 and int(m.bc_type.sum()) == 32
    assert set(np.unique(m.bc_value)) == {0.0, 1.0}
    assert np.allclose(m.bc_flux[m.bc_type == 1], S.EPS1 / S.EPS2)
    r = np.linalg.norm(m.centroids(), axis=1)
    assert np.all((r > 0.8) & (r < 3.0))
    assert S.mesh(2).n == 1030 and S.mesh(3).n == 2034


def test_charges_converge(solved):
    exact = np.array(S.sphere_charges())
    err = {s: np.array(S.measured_charges(m, x)) / exact - 1 for s, (m, A, x) in solved.items()}
    assert abs(err[1][0]) < 0.14 and abs(err[1][2]) < 0.21
    assert abs(err[2][0]) < 0.075 and abs(err[2][2]) < 0.09
    assert abs(err[2][0]) < 0.6 * abs(err[1][0]) and abs(err[2][2]) < 0.6 * abs(err[1][2])
    assert np.all(err[1][[0, 2]] < 0) and np.all(err[2][[0, 2]] < 0)
    # total charge: free charges cancel, bound charges cancel in pairs
    for s, (m, A, x) in solved.items():
        assert abs(sum(S.measured_charges(m, x))) < 0.1 * exact[0]


def test_potential_converges(oracle, solved):
    P = S.probe_points(40, 3)
    errs = []
    for s, (m, A, x) in solved.items():
        phi, E = oracle.field(m.shapes, m.recs, x, P)
        errs.append(np.abs(phi - S.potential(P)).max() / S.V1)
    assert errs[0] < 0.2 and errs[1] < 0.07 and errs[1] < 0.6 * errs[0]


def test_neumann_rows_hold(oracle, solved):
    """The solution satisfies every row: Dirichlet values and the Neumann interface condition."""
    for s, (m, A, x) in solved.items():
        b = oracle.rhs(m.bc_value, m.bc_type)
        assert np.linalg.norm(A @ x - b) <= 1e-10 * np.linalg.norm(b)
