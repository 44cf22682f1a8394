"""Pin the oracle's assembly, Dirichlet enforcement, solve, compliance and
energies: brute-force dense assembly on tiny meshes, the uniaxial-bar closed
form, C = f.u, sum_e U_e = C/2, heat J = T^T K T, and CG == SuperLU."""
import numpy as np
import pytest

from oracle import elements as el
from oracle import fem as ofem
from oracle import mesh as om
from oracle import problems as op
from oracle import transforms as otr


def _dense_assembly(mesh, E, nu):
    K0 = el.hex8_elasticity_K0(nu, mesh.h)
    n = 3 * mesh.nnodes
    K = np.zeros((n, n))
    for e in range(mesh.nelements):
        d = np.array([3 * mesh.t[a, e] + c for a in range(8) for c in range(3)])
        K[np.ix_(d, d)] += E[e] * K0
    return K


def test_assembly_matches_dense_loop():
    mesh = om.structured_hex(2, 1, 1)
    E = np.array([1.0, 0.25])
    op_ = ofem.ElementOperator(mesh, "elastic", 0.3)
    np.testing.assert_allclose(op_.assemble(E).toarray(), _dense_assembly(mesh, E, 0.3),
                               rtol=1e-14, atol=1e-15)


def test_enforce_zeroes_rows_and_matches_elimination():
    task = op.cantilever(3, 2, 2)
    op_ = task.element_operator()
    rng = np.random.default_rng(0)
    E = rng.uniform(0.1, 1.0, task.n_elements)
    K = op_.assemble(E)
    f = task.forces[0]
    Ke, fe = ofem.enforce(K, f, task.dirichlet_dofs)
    D = task.dirichlet_dofs
    A = Ke.toarray()
    assert np.all(A[D][:, np.setdiff1d(np.arange(task.ndof), D)] == 0)
    np.testing.assert_array_equal(np.diag(A)[D], 1.0)
    u = np.linalg.solve(A, fe)
    free = np.setdiff1d(np.arange(task.ndof), D)
    Kd = K.toarray()
    uf = np.linalg.solve(Kd[np.ix_(free, free)], f[free])
    np.testing.assert_allclose(u[free], uf, rtol=1e-10)
    assert np.all(u[D] == 0)


def test_uniaxial_bar_closed_form():
    """nu = 0 bar pulled at x = L, rollers on x = 0: u_x = F L / (E A) exactly."""
    nx, ny, nz = 6, 2, 2
    mesh = om.structured_hex(nx, ny, nz)
    n = np.arange(mesh.nnodes)
    i, j, k = n % (nx + 1), (n // (nx + 1)) % (ny + 1), n // ((nx + 1) * (ny + 1))
    D = list(3 * n[i == 0])                       # u_x = 0 on x = 0
    D += [3 * n[(i == 0) & (j == 0) & (k == 0)][0] + 1, 3 * n[(i == 0) & (j == 0) & (k == 0)][0] + 2]
    D += [3 * n[(i == 0) & (j == ny) & (k == 0)][0] + 2]
    D += [3 * n[(i == 0) & (j == 0) & (k == nz)][0] + 1]
    D = np.array(sorted(set(D)))
    # consistent nodal loads of a uniform traction on the x = L face
    f = np.zeros(3 * mesh.nnodes)
    face = (i == nx)
    w = np.where((j[face] == 0) | (j[face] == ny), 0.5, 1.0) * \
        np.where((k[face] == 0) | (k[face] == nz), 0.5, 1.0)
    F = 3.0
    f[3 * n[face]] = F * w / w.sum()
    op_ = ofem.ElementOperator(mesh, "elastic", 0.0)
    K = op_.assemble(np.full(mesh.nelements, 2.0))
    Ke, fe = ofem.enforce(K, f, D)
    u = ofem.solve(Ke, fe)
    ux = u[0::3]
    np.testing.assert_allclose(ux, F * i / (2.0 * ny * nz), rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("kind", ["hex8", "tet4"])
def test_compliance_and_energy_identities(kind):
    task = op.cantilever(5, 3, 3, kind=kind, jitter=0.2, seed=2)
    fem = ofem.ElasticFEM(task, 1e-9)
    rng = np.random.default_rng(1)
    rho = rng.uniform(0.01, 1.0, task.n_elements)
    U = np.zeros((task.ndof, 1))
    C = fem.objectives_multi_load(rho, 3.0, U)
    en = fem.energy_multi_load(rho, 3.0, U)
    np.testing.assert_allclose(en.sum(), 0.5 * C[0], rtol=1e-9)
    K = fem.op.assemble(otr.simp_interpolation(rho, 1.0, 1e-9, 3.0))
    np.testing.assert_allclose(U[:, 0] @ K @ U[:, 0], C[0], rtol=1e-9)


def test_heat_compliance_is_TKT():
    task = op.heat_patch(5, patch=2)
    fem = ofem.HeatFEM(task, 1e-3)
    rho = np.random.default_rng(3).uniform(0.01, 1.0, task.n_elements)
    U = np.zeros((task.ndof, 1))
    J = fem.objectives_multi_load(rho, 3.0, U)
    en = fem.energy_multi_load(rho, 3.0, U)
    np.testing.assert_allclose(J[0], 2 * en.sum(), rtol=1e-10)
    np.testing.assert_allclose(J[0], task.forces[0] @ U[:, 0], rtol=1e-9)


def test_cg_jacobi_agrees_with_superlu():
    task = op.multiload(6, 3, 3, n_loads=3, seed=4)
    rho = np.random.default_rng(5).uniform(0.01, 1.0, task.n_elements)
    U1 = np.zeros((task.ndof, 3))
    U2 = np.zeros((task.ndof, 3))
    c1 = ofem.ElasticFEM(task, 1e-9, solver="spsolve").objectives_multi_load(rho, 3.0, U1)
    c2 = ofem.ElasticFEM(task, 1e-9, solver="cg_jacobi", rtol=1e-13).objectives_multi_load(
        rho, 3.0, U2)
    np.testing.assert_allclose(c2, c1, rtol=1e-11)


def test_hex_volumes_and_centres():
    mesh = om.structured_hex(3, 2, 2, 1.0, 0.5, 2.0)
    np.testing.assert_allclose(om.elements_volume(mesh), 1.0, rtol=1e-14)
    c = om.element_centers(mesh)
    np.testing.assert_allclose(c[:, 0], [0.5, 0.25, 1.0])


def test_tet_mesh_conforming_and_volume():
    mesh = om.jittered_tet(4, 3, 2, jitter=0.2, seed=0)
    v = om.elements_volume(mesh)
    assert np.all(v > 0)
    np.testing.assert_allclose(v.sum(), 4 * 3 * 2, rtol=1e-12)
    # every interior face is shared by exactly two tets
    faces = np.sort(np.concatenate([mesh.t[[0, 1, 2]], mesh.t[[0, 1, 3]], mesh.t[[0, 2, 3]],
                                    mesh.t[[1, 2, 3]]], axis=1), axis=0)
    _, cnt = np.unique(faces, axis=1, return_counts=True)
    assert cnt.max() == 2
