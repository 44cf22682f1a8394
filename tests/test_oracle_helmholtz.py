"""Pin the oracle's nodal Helmholtz filter (filters/helmholtz_filter_nodal.py
restated with scipy): closed-form element matrices, the reference's
element-by-element averaging loop, constant-field and zero-radius
invariants, and the condensed solve against a dense solve."""
import numpy as np
import pytest

from oracle import helmholtz as oh
from oracle import mesh as om
from oracle import problems as oprob


def test_hex_mass_closed_form():
    h = (1.0, 2.0, 0.5)
    M = oh.hex8_mass_M0(h)
    c = om.HEX_CORNERS
    ref = np.ones((8, 8))
    for ax in range(3):
        same = c[:, ax][:, None] == c[:, ax][None, :]
        ref *= h[ax] / 6.0 * np.where(same, 2.0, 1.0)
    np.testing.assert_allclose(M, ref, rtol=1e-14)
    assert abs(M.sum() - np.prod(h)) <= 1e-14


def test_tet_mass_sums_to_volume():
    m = om.jittered_tet(3, 2, 2, 0.2, 1)
    A, M = oh.helmholtz_matrices(m, 0.0)
    np.testing.assert_allclose(M.sum(), om.elements_volume(m).sum(), rtol=1e-13)
    np.testing.assert_allclose((A - M).toarray(), 0.0, atol=1e-14)


def _reference_loop(mesh, vol, rho, mask, weighted, fixed):
    """helmholtz_filter_nodal.py:31-59 as written (Python loop)."""
    rho_node = np.zeros(mesh.nnodes)
    wsum = np.zeros(mesh.nnodes)
    for e, nodes in enumerate(mesh.t.T):
        v = rho[e] if mask[e] else fixed
        w = vol[e] if weighted else 1.0
        rho_node[nodes] += w * v
        wsum[nodes] += w
    wsum[wsum == 0.0] = 1.0
    rho_node /= wsum
    return rho_node


@pytest.mark.parametrize("kind", ["hex8", "tet4"])
def test_averaging_equals_the_reference_loop_bitwise(kind):
    task = oprob.cantilever(6, 3, 3, kind=kind, jitter=0.2, seed=2)
    rng = np.random.default_rng(0)
    rho = rng.uniform(0, 1, task.n_elements)
    for fixed in (1.0, 0.0):
        a = oh.element_to_node_density_averaging(task.mesh, task.elements_volume, rho,
                                                 task.design_mask, True, fixed)
        b = _reference_loop(task.mesh, task.elements_volume, rho, task.design_mask, True, fixed)
        np.testing.assert_array_equal(a, b)


def test_constant_field_is_preserved():
    m = om.structured_hex(5, 4, 3)
    f = oh.HelmholtzFilterNodal(m, om.elements_volume(m), 1.5, None)
    np.testing.assert_allclose(f.forward(np.full(m.nelements, 0.37)), 0.37, rtol=1e-12)


def test_ones_stay_ones_with_fixed_nodes():
    task = oprob.cantilever(6, 3, 3)
    f = oh.HelmholtzFilterNodal(task.mesh, task.elements_volume, 2.0, task.design_mask)
    np.testing.assert_allclose(f.forward(np.ones(task.n_elements)), 1.0, rtol=1e-12)


def test_zero_radius_is_average_then_mean():
    m = om.structured_hex(4, 3, 3)
    vol = om.elements_volume(m)
    rho = np.random.default_rng(1).uniform(0, 1, m.nelements)
    f = oh.HelmholtzFilterNodal(m, vol, 0.0, None)
    rn = oh.element_to_node_density_averaging(m, vol, rho, None, True, 1.0)
    np.testing.assert_allclose(f.forward(rho), oh.node_to_element_density(m, rn), rtol=1e-12)


def test_condensed_solve_matches_dense_and_gradient_is_clamped():
    task = oprob.cantilever(4, 2, 2)
    m = task.mesh
    rng = np.random.default_rng(3)
    rho = rng.uniform(0, 1, m.nelements)
    f = oh.HelmholtzFilterNodal(m, task.elements_volume, 1.2, task.design_mask)
    A, M = oh.helmholtz_matrices(m, 1.2)
    rn = oh.element_to_node_density_averaging(m, task.elements_volume, rho, task.design_mask,
                                              True, 1.0)
    fixed = np.unique(m.t[:, ~task.design_mask].ravel())
    Ad = A.toarray()
    b = M @ rn
    Ad[fixed, :] = 0.0
    Ad[fixed, fixed] = 1.0
    b[fixed] = 1.0
    x = np.linalg.solve(Ad, b)
    np.testing.assert_allclose(f.forward(rho), oh.node_to_element_density(m, x), rtol=1e-12)
    # sensitivities are negative: the result is too; positive input is clamped
    g = f.gradient(-rng.uniform(0.1, 1.0, m.nelements))
    assert np.all(g < 0.0)
    np.testing.assert_array_equal(f.gradient(np.abs(rng.standard_normal(m.nelements))), 0.0)
