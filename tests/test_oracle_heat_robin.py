"""Pins of the Robin heat oracle (fea/solver_heat.py:306-980,
mesh/task_heat.py:60-139) against what the mathematics fixes."""
import numpy as np
import pytest

from oracle import heat_robin as hr
from oracle import helmholtz as oh
from oracle import mesh as om
from oracle import problems as oprob
from oracle.fem import HeatFEM


def test_nodal_average_pairing_equals_the_reference_statement():
    """k-th term of np.add.at(rn, t.ravel(), np.repeat(rho, 8)) is rho[k // 8]
    at node t.ravel()[k] -- written out as a plain loop."""
    m = om.structured_hex(4, 3, 2)
    rho = np.random.default_rng(0).uniform(size=m.nelements)
    rn, cnt = hr.nodal_average_reference(m, rho)
    flat = m.t.ravel()
    acc = np.zeros(m.nnodes)
    c = np.zeros(m.nnodes)
    for k in range(flat.size):
        acc[flat[k]] += rho[k // 8]
        c[flat[k]] += 1.0
    np.testing.assert_array_equal(cnt, c)
    np.testing.assert_array_equal(rn, acc / np.maximum(c, 1.0))
    # constant fields survive the pairing
    rn1, _ = hr.nodal_average_reference(m, np.full(m.nelements, 0.3))
    np.testing.assert_allclose(rn1, 0.3, rtol=1e-15)


def test_virtual_robin_reduces_to_a_mass_matrix():
    """p = q = 0 and a linear nodal field: h a |grad rho| = h |beta| constant,
    so B = h |beta| M and L = h |beta| T_env M 1."""
    m = om.structured_hex(3, 2, 2, 0.5, 1.0, 2.0)
    beta = 0.37
    rn = 0.2 + beta * m.p[0]
    B, L = hr.virtual_robin_nodal(m, rn, 0.0, 0.0, 2.5, 300.0, npts=2)
    _, M = oh.helmholtz_matrices(m, 0.0)
    np.testing.assert_allclose(B.toarray(), 2.5 * beta * M.toarray(), rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(L, 2.5 * beta * 300.0 * (M @ np.ones(m.nnodes)), rtol=1e-13)
    # p = 1: the integrand is cubic along x, exact with 2 points per axis
    B2, _ = hr.virtual_robin_nodal(m, rn, 1.0, 0.0, 1.0, 1.0, npts=2)
    B4, _ = hr.virtual_robin_nodal(m, rn, 1.0, 0.0, 1.0, 1.0, npts=4)
    np.testing.assert_allclose(B2.toarray(), B4.toarray(), rtol=1e-12, atol=1e-15)


def test_facet_robin_integrals():
    m = om.structured_hex(2, 2, 1)
    # the z = 1 face: 4 unit squares in cyclic order
    nid = lambda i, j, k: i + 3 * (j + 3 * k)  # noqa: E731
    fc = np.array([[nid(i, j, 1), nid(i + 1, j, 1), nid(i + 1, j + 1, 1), nid(i, j + 1, 1)]
                   for j in range(2) for i in range(2)]).T
    B, L = hr.facet_robin(m.p, fc, 2.0, 5.0, m.nnodes)
    np.testing.assert_allclose(B.sum(), 2.0 * 4.0, rtol=1e-14)
    np.testing.assert_allclose(L.sum(), 2.0 * 5.0 * 4.0, rtol=1e-14)
    one, _ = hr.facet_robin(m.p, fc[:, :1], 36.0, 0.0, m.nnodes)
    loc = one.toarray()[np.ix_(fc[:, 0], fc[:, 0])]
    ref = np.array([[4, 2, 1, 2], [2, 4, 2, 1], [1, 2, 4, 2], [2, 1, 2, 4]], dtype=float)
    np.testing.assert_allclose(loc, ref, rtol=1e-14)


def test_explicit_gradient_is_the_nodal_derivative():
    """The nodal form of _robin_compliance_explicit_grad_ is the derivative of
    int h a(rho) |grad rho| (2 T_env T - T^2) in the nodal densities."""
    m = om.structured_hex(3, 2, 2)
    rng = np.random.default_rng(4)
    rn = 0.3 + 0.4 * rng.uniform(size=m.nnodes)
    T = 300.0 + 50.0 * rng.uniform(size=m.nnodes)
    g = hr.robin_explicit_grad_nodal(m, rn, T, 3.0, 4.0, 0.7, 310.0)
    eps = 1e-6
    for n in rng.choice(m.nnodes, size=6, replace=False):
        a, b = rn.copy(), rn.copy()
        a[n] += eps
        b[n] -= eps
        fd = (hr.robin_functional(m, a, T, 3.0, 4.0, 0.7, 310.0) -
              hr.robin_functional(m, b, T, 3.0, 4.0, 0.7, 310.0)) / (2 * eps)
        np.testing.assert_allclose(g[n], fd, rtol=1e-6, atol=1e-6 * np.abs(g).max())


def test_without_robin_equals_the_plain_heat_oracle():
    task = oprob.heat_patch(6, patch=3, source=1e-2)
    fem0 = HeatFEM(task, 1e-3)
    fem1 = hr.HeatRobinFEM(task.mesh, task.dirichlet_dofs, task.dirichlet_values, task.E0,
                           1e-3, [], 0.0, 0.0, emit=task.forces[0])
    rho = np.random.default_rng(1).uniform(0.1, 1.0, task.n_elements)
    U0 = np.zeros((task.ndof, 1))
    U1 = np.zeros((task.ndof, 1))
    c0 = fem0.objectives_multi_load(rho, 3.0, U0)
    c1 = fem1.objectives_multi_load(rho, 3.0, U1)
    np.testing.assert_allclose(c1, c0, rtol=1e-12)
    s0 = fem0.compliance_sensitivity_multi_load(rho, 3.0, U0)
    np.testing.assert_allclose(fem1.compliance_sensitivity_multi_load(rho, 3.0, U1), s0,
                               rtol=1e-12, atol=1e-12 * np.abs(s0).max())


def test_averaged_temp_adjoint():
    """J = sum(T) and lam solves K lam = 1 (enforced with the state's values):
    for a zero Dirichlet value, sum(T) = lam . F on the free dofs."""
    task = oprob.heat_patch(5, patch=3, source=1e-2)
    fem = hr.HeatRobinFEM(task.mesh, task.dirichlet_dofs, [0.0], task.E0, 1e-3, [], 0.0,
                          0.0, emit=task.forces[0], objective="averaged_temp")
    rho = np.random.default_rng(2).uniform(0.1, 1.0, task.n_elements)
    U = np.zeros((task.ndof, 1))
    J = fem.objectives_multi_load(rho, 3.0, U)
    free = np.setdiff1d(np.arange(task.ndof), task.dirichlet_dofs)
    np.testing.assert_allclose(J[0], fem.lam[free, 0] @ task.forces[0][free], rtol=1e-10)
    e = fem.energy_multi_load(rho, 3.0, U)
    assert e.shape == (task.n_elements, 1)


def test_heat_exchange_terms_closed_forms():
    """Linear nodal fields pin the heat-exchange functionals: a trilinear
    interpolant reproduces a linear field, so |grad rho_h| = |g| everywhere,
    J_den = |g| V and J_num = -T_env h |g| (V T(centroid) - T_env V) for a
    linear T; the adjoint load sums (partition of unity) to
    -T_env |g| sum_e V_e mean_corners(h_eff), a trilinear field's element
    mean being its corner mean.  All integrands are polynomial, so the 2-
    and 4-point rules agree."""
    m = om.structured_hex(4, 3, 2, 0.5, 0.7, 1.1)
    X = m.p
    g = np.array([0.05, -0.03, 0.02])
    rn = 0.4 + g @ X
    c = np.array([[3.0, 1.0, -2.0], [-1.0, 0.5, 4.0]])
    T = np.column_stack([310.0 + c[0] @ X, 280.0 + c[1] @ X])
    V = (4 * 0.5) * (3 * 0.7) * (2 * 1.1)
    ctr = np.array([1.0, 1.05, 1.1])
    gn = np.linalg.norm(g)
    h, Te, p, q = 0.04, 300.0, 3.0, 4.0
    heff = h * rn ** p * (1.0 - rn) ** q
    rhs_sum = -Te * gn * (0.5 * 0.7 * 1.1) * heff[m.t].mean(axis=0).sum()
    for npts in (2, 4):
        den, num, rhs = hr.heat_exchange_terms(m, rn, T, p, q, h, Te, npts)
        np.testing.assert_allclose(den, gn * V, rtol=1e-13)
        Tc = np.array([310.0 + c[0] @ ctr, 280.0 + c[1] @ ctr])
        np.testing.assert_allclose(num, -Te * h * gn * (V * Tc - Te * V), rtol=1e-12)
        np.testing.assert_allclose(rhs.sum(), rhs_sum, rtol=1e-12)
        assert np.all(rhs <= 0.0)


def _hx_fem(task, objective, w=0.0, values=(0.0,)):
    return hr.HeatRobinFEM(task.mesh, task.dirichlet_dofs, list(values), task.E0, 1e-3, [],
                           0.04, 300.0, emit=task.forces[0], objective=objective,
                           avg_temp_weight=w)


def test_heat_exchange_blend_and_flat_density():
    """``avg_temp_weight`` (solver_heat.py:845-869): one load case adds w sum(T)
    to J and w times the averaged-temperature adjoint to lam; several load
    cases first divide each part by its mean magnitude.  A density without
    interface (J_den = 0) gives J = 0 and lam = 0."""
    task = oprob.heat_patch(5, patch=3, source=1e-2)
    rho = np.random.default_rng(4).uniform(0.1, 1.0, task.n_elements)
    for values in ((0.0,), (0.0, 5.0)):
        k = len(values)
        res = {}
        for name, obj, w in (("hx", "heat_exchange", 0.0), ("mix", "heat_exchange", 0.3),
                             ("avg", "averaged_temp", 0.0)):
            fem = _hx_fem(task, obj, w, values)
            U = np.zeros((task.ndof, k))
            res[name] = (fem.objectives_multi_load(rho, 3.0, U), fem.lam.copy(), U)
        J0, l0, U = res["hx"]
        Ja, la, _ = res["avg"]
        Jw, lw, _ = res["mix"]
        np.testing.assert_allclose(Ja, U.sum(axis=0), rtol=1e-12)
        if k == 1:
            np.testing.assert_allclose(Jw, J0 + 0.3 * Ja, rtol=1e-12)
            np.testing.assert_allclose(lw, l0 + 0.3 * la, rtol=1e-10, atol=1e-12)
        else:
            hs, avs = np.abs(J0).mean(), np.abs(Ja).mean()
            np.testing.assert_allclose(Jw, J0 / hs + 0.3 * Ja / avs, rtol=1e-12)
            np.testing.assert_allclose(lw, l0 / hs + (0.3 / avs) * la, rtol=1e-10, atol=1e-12)
        assert np.all(J0 != 0.0)
    fem = _hx_fem(task, "heat_exchange", 0.3)
    U = np.zeros((task.ndof, 1))
    # a zero field: its gradient is zero to the last bit (a flat non-zero
    # field leaves rounding-level gradients, J_den ~ 1e-15, as in the
    # reference's arithmetic)
    J = fem.objectives_multi_load(np.zeros(task.n_elements), 3.0, U)
    assert J[0] == 0.0 and not fem.lam.any()


# --------------------------------------------------------------- tet4 meshes

def _tet_mesh(n=3, seed=5):
    return om.jittered_tet(n, n, n, 0.2, seed)


def _p1_mass(mesh):
    """Assembled P1 mass matrix from the closed form int N_a N_b = V (1 + d_ab) / 20."""
    import scipy.sparse as sp
    vol = om.tet_signed_volumes(mesh.p, mesh.t)
    loc = (np.ones((4, 4)) + np.eye(4)) / 20.0
    rows = np.repeat(mesh.t.T, 4, axis=1).ravel()
    cols = np.tile(mesh.t.T, (1, 4)).ravel()
    vals = (vol[:, None, None] * loc[None]).ravel()
    return sp.csr_matrix((vals, (rows, cols)), shape=(mesh.nnodes, mesh.nnodes))


def test_tet_rule_degree():
    """The 4-point rule integrates every barycentric monomial of degree <= 2
    exactly (int l_a l_b = V (1 + d_ab) / 20, int l_a = V / 4) and a cubic
    one not (int l_0^3 = V / 20); the centroid rule only degree 1."""
    lam, w = hr.tet_rule(2)
    np.testing.assert_allclose(w.sum(), 1.0, rtol=1e-15)
    np.testing.assert_allclose(w @ lam, np.full(4, 0.25), rtol=1e-15)
    np.testing.assert_allclose(np.einsum("q,qa,qb->ab", w, lam, lam),
                               (np.ones((4, 4)) + np.eye(4)) / 20.0, rtol=1e-14)
    assert abs(w @ lam[:, 0] ** 3 - 1.0 / 20.0) > 1e-3
    lam1, w1 = hr.tet_rule(1)
    np.testing.assert_allclose(w1 @ lam1, np.full(4, 0.25), rtol=1e-15)
    assert abs(np.einsum("q,qa,qb->ab", w1, lam1, lam1)[0, 0] - 0.1) > 1e-3


def test_tet_virtual_robin_reduces_to_a_mass_matrix():
    """p = q = 0 and a linear nodal field on a jittered tet mesh: the P1
    interpolant reproduces it, h a |grad rho| = h |g|, so B = h |g| M and
    L = h |g| T_env M 1 with the closed-form P1 mass matrix."""
    m = _tet_mesh()
    g = np.array([0.05, -0.03, 0.02])
    rn = 0.3 + g @ m.p
    B, L = hr.virtual_robin_nodal(m, rn, 0.0, 0.0, 2.5, 300.0, npts=2)
    M = _p1_mass(m)
    gn = np.linalg.norm(g)
    np.testing.assert_allclose(B.toarray(), 2.5 * gn * M.toarray(), rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(L, 2.5 * gn * 300.0 * (M @ np.ones(m.nnodes)), rtol=1e-12)


def test_tet_facet_robin_integrals():
    """Triangle facets: the edge-midpoint rule gives the closed-form facet
    mass h A / 12 (1 + d_ab) and the load h T_env A / 3 per node."""
    P = np.array([[0.0, 2.0, 0.3], [0.1, 0.0, 1.5], [0.0, 0.2, 0.4]])
    fc = np.array([[0], [1], [2]])
    B, L = hr.facet_robin(P, fc, 3.0, 7.0, 3)
    A = 0.5 * np.linalg.norm(np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]))
    np.testing.assert_allclose(B.toarray(), 3.0 * A / 12.0 * (np.ones((3, 3)) + np.eye(3)),
                               rtol=1e-14)
    np.testing.assert_allclose(L, np.full(3, 3.0 * 7.0 * A / 3.0), rtol=1e-14)


@pytest.mark.parametrize("order", [1, 2])
def test_tet_explicit_gradient_is_the_nodal_derivative(order):
    m = _tet_mesh()
    rng = np.random.default_rng(7)
    rn = 0.3 + 0.4 * rng.uniform(size=m.nnodes)
    T = 300.0 + 50.0 * rng.uniform(size=m.nnodes)
    g = hr.robin_explicit_grad_nodal(m, rn, T, 3.0, 4.0, 0.7, 310.0, order)
    eps = 1e-6
    for n in rng.choice(m.nnodes, size=6, replace=False):
        a, b = rn.copy(), rn.copy()
        a[n] += eps
        b[n] -= eps
        fd = (hr.robin_functional(m, a, T, 3.0, 4.0, 0.7, 310.0, order) -
              hr.robin_functional(m, b, T, 3.0, 4.0, 0.7, 310.0, order)) / (2 * eps)
        np.testing.assert_allclose(g[n], fd, rtol=1e-6, atol=1e-6 * np.abs(g).max())


def test_tet_heat_exchange_terms_closed_forms():
    """Per tet, |grad rho_h| is constant and T_h linear: J_den = sum V |g_e|,
    J_num = -T_env h sum V |g_e| (mean of the corner T - T_env), and the
    adjoint load is -T_env |g_e| times the P1 mass applied to the nodal
    h_eff, for any nodal fields."""
    m = _tet_mesh()
    rng = np.random.default_rng(8)
    rn = 0.2 + 0.6 * rng.uniform(size=m.nnodes)
    T = 290.0 + 30.0 * rng.uniform(size=(m.nnodes, 2))
    h, Te, p, q = 0.04, 300.0, 3.0, 4.0
    from oracle import elements as oel
    gb, vol = oel.tet4_grads(m.p, m.t)
    gn = np.linalg.norm(np.einsum("ead,ae->ed", gb, rn[m.t]), axis=1)
    den, num, rhs = hr.heat_exchange_terms(m, rn, T, p, q, h, Te)
    np.testing.assert_allclose(den, np.sum(vol * gn), rtol=1e-13)
    Tm = T[m.t].mean(axis=0)                                   # (n_el, 2)
    np.testing.assert_allclose(num, -Te * h * np.einsum("e,ek->k", vol * gn, Tm - Te),
                               rtol=1e-12)
    heff = h * rn ** p * (1.0 - rn) ** q
    loc = (np.ones((4, 4)) + np.eye(4)) / 20.0
    Le = -Te * (vol * gn)[:, None] * np.einsum("ab,be->ea", loc, heff[m.t])
    ref = np.zeros(m.nnodes)
    np.add.at(ref, m.t.T.ravel(), Le.ravel())
    np.testing.assert_allclose(rhs, ref, rtol=1e-12, atol=1e-14 * np.abs(ref).max())


def test_tet_without_robin_equals_the_plain_heat_oracle():
    task = oprob.heat_patch(4, patch=3, source=1e-2, kind="tet4", jitter=0.2)
    fem0 = HeatFEM(task, 1e-3)
    fem1 = hr.HeatRobinFEM(task.mesh, task.dirichlet_dofs, task.dirichlet_values, task.E0,
                           1e-3, [], 0.0, 0.0, emit=task.forces[0])
    rho = np.random.default_rng(1).uniform(0.1, 1.0, task.n_elements)
    U0 = np.zeros((task.ndof, 1))
    U1 = np.zeros((task.ndof, 1))
    c0 = fem0.objectives_multi_load(rho, 3.0, U0)
    c1 = fem1.objectives_multi_load(rho, 3.0, U1)
    np.testing.assert_allclose(c1, c0, rtol=1e-12)
    s0 = fem0.compliance_sensitivity_multi_load(rho, 3.0, U0)
    np.testing.assert_allclose(fem1.compliance_sensitivity_multi_load(rho, 3.0, U1), s0,
                               rtol=1e-12, atol=1e-12 * np.abs(s0).max())


def test_tet_averaged_temp_adjoint():
    task = oprob.heat_patch(4, patch=3, source=1e-2, kind="tet4", jitter=0.2)
    fem = hr.HeatRobinFEM(task.mesh, task.dirichlet_dofs, [0.0], task.E0, 1e-3, [], 0.04,
                          300.0, emit=task.forces[0], objective="averaged_temp")
    rho = np.random.default_rng(2).uniform(0.1, 1.0, task.n_elements)
    U = np.zeros((task.ndof, 1))
    J = fem.objectives_multi_load(rho, 3.0, U)
    # with Robin data the right-hand side is emit + the virtual Robin load
    _, Lv = hr.virtual_robin(task.mesh, rho, 3.0, 4.0, 0.04, 300.0, 2)
    F = task.forces[0] + Lv
    free = np.setdiff1d(np.arange(task.ndof), task.dirichlet_dofs)
    np.testing.assert_allclose(J[0], fem.lam[free, 0] @ F[free], rtol=1e-10)
