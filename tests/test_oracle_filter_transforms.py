"""Pin the oracle's filter (brute-force neighbour search, row sums, adjoint,
distance ties) and design-variable transforms (closed forms and their
quirks, central differences)."""
import numpy as np
import pytest

from oracle import filter as of
from oracle import mesh as om
from oracle import problems as op
from oracle import transforms as tr


def _brute_filter(c, r, mask):
    ids = np.nonzero(mask)[0]
    s = r / 3.0
    sup = 3.0 * s
    n = len(ids)
    W = np.zeros((n, n))
    for a, i in enumerate(ids):
        for b, j in enumerate(ids):
            d2 = ((c[:, j] - c[:, i]) ** 2).sum()
            if d2 <= sup * sup:
                W[a, b] = np.exp(-0.5 * (np.sqrt(d2) / s) ** 2)
    rs = W.sum(1)
    rs[rs == 0] = 1.0
    return W / rs[:, None]


@pytest.mark.parametrize("r", [1.0, 1.5, 2.0, 2.5])
def test_filter_matches_brute_force(r):
    mesh = om.structured_hex(6, 5, 4)
    c = om.element_centers(mesh)
    mask = np.ones(mesh.nelements, bool)
    mask[::7] = False
    W = of.build_filter_matrix(c, r, mask).toarray()
    np.testing.assert_allclose(W, _brute_filter(c, r, mask), rtol=1e-14, atol=1e-16)


def test_filter_ties_at_the_radius_are_included():
    """r = 2.0 on a unit grid: the 6 axis neighbours at distance exactly 2
    belong to the stencil (33 entries per interior row)."""
    mesh = om.structured_hex(9, 9, 9)
    c = om.element_centers(mesh)
    W = of.build_filter_matrix(c, 2.0, np.ones(mesh.nelements, bool))
    centre = 4 + 9 * (4 + 9 * 4)
    assert W.indptr[centre + 1] - W.indptr[centre] == 33
    for r, n in [(1.5, 19), (2.5, 81)]:
        W = of.build_filter_matrix(c, r, np.ones(mesh.nelements, bool))
        assert W.indptr[centre + 1] - W.indptr[centre] == n


def test_filter_rows_sum_to_one_and_adjoint():
    task = op.cantilever(8, 4, 4, kind="tet4", jitter=0.2, seed=0)
    f = of.SpacialFilter(task.centers, 1.5, task.design_mask)
    np.testing.assert_allclose(np.asarray(f.W.sum(axis=1)).ravel(), 1.0, rtol=1e-14)
    rng = np.random.default_rng(0)
    x, y = rng.standard_normal(task.n_elements), rng.standard_normal(task.n_elements)
    xd = np.where(task.design_mask, x, 0.0)
    lhs = (f.forward(xd) * task.design_mask) @ y
    rhs = xd @ f.gradient(y)
    np.testing.assert_allclose(lhs, rhs, rtol=1e-12)
    out = f.forward(x)
    np.testing.assert_array_equal(out[~task.design_mask], x[~task.design_mask])
    assert np.all(f.gradient(y)[~task.design_mask] == 0)


def test_gaussian_support_equals_radius():
    for r in (1.5, 2.0, 2.5):
        assert of.gaussian_support(r)[1] == r


def test_heaviside_closed_forms():
    beta, eta = 4.0, 0.5
    d = np.tanh(beta * eta) + np.tanh(beta * (1 - eta)) + 1e-12
    x = np.array([0.0, eta, 1.0])
    np.testing.assert_allclose(tr.heaviside_projection(x, beta, eta),
                               [0.0, np.tanh(beta * eta) / d, (2 * np.tanh(2.0)) / d],
                               rtol=1e-15)
    # the +1e-12 keeps H(1) just below 1
    assert tr.heaviside_projection(np.array([1.0]), 1.0)[0] < 1.0


@pytest.mark.parametrize("beta,eta", [(1.0, 0.5), (8.0, 0.5), (3.0, 0.3)])
def test_heaviside_derivative_central_difference(beta, eta):
    x = np.linspace(0.01, 0.99, 97)
    h = 1e-6
    fd = (tr.heaviside_projection(x + h, beta, eta) - tr.heaviside_projection(x - h, beta, eta)) / (2 * h)
    np.testing.assert_allclose(tr.heaviside_projection_derivative(x, beta, eta), fd, rtol=1e-7)


def test_heaviside_derivative_overflow():
    assert tr.heaviside_projection_derivative(np.array([1e3]), 10.0)[0] == 0.0


def test_interpolation_closed_forms():
    rho = np.array([0.0, 0.5, 1.0])
    np.testing.assert_allclose(tr.simp_interpolation(rho, 2.0, 1e-3, 3.0),
                               [1e-3, 1e-3 + (2 - 1e-3) / 8, 2.0])
    np.testing.assert_allclose(tr.ramp_interpolation(rho, 2.0, 1e-3, 3.0),
                               [1e-3, 1e-3 + (2 - 1e-3) * 0.5 / 2.5, 2.0])


def test_dC_simp_is_minus_energy_derivative():
    """dC/drho = -2 U dE/E with U = E q / 2 equals -q dE/drho; dE by FD."""
    rho = np.linspace(0.05, 0.95, 19)
    q = np.linspace(1.0, 2.0, 19)
    E = tr.simp_interpolation(rho, 1.0, 1e-9, 3.0)
    U = 0.5 * E * q
    h = 1e-7
    dE = (tr.simp_interpolation(rho + h, 1.0, 1e-9, 3.0)
          - tr.simp_interpolation(rho - h, 1.0, 1e-9, 3.0)) / (2 * h)
    np.testing.assert_allclose(tr.dC_drho_simp(rho, U, 1.0, 1e-9, 3.0), -q * dE, rtol=1e-6)


def test_dC_ramp_reproduces_reference_quirk():
    """The reference's RAMP derivative uses (den - p rho)/den^2 with
    den = 1 + p(1 - rho) (core/derivatives.py:56-61); the exact derivative of
    rho/den is (1 + p)/den^2.  Hand values at p = 3: rho = 0.5 -> den = 2.5,
    quirk factor 1.0/6.25 = 0.16, exact 4/6.25 = 0.64."""
    rho = np.array([0.5])
    E = tr.ramp_interpolation(rho, 1.0, 0.0, 3.0)          # 0.5/2.5 = 0.2
    np.testing.assert_allclose(E, [0.2])
    U = 0.5 * E * 1.0                                        # q = 1
    np.testing.assert_allclose(tr.dC_drho_ramp(rho, U, 1.0, 0.0, 3.0), [-0.16], rtol=1e-14)


def test_dC_simp_clamps():
    """rho clamped at 1e-6 inside both E and dE (derivatives.py:25-52)."""
    U = np.array([1.0])
    a = tr.dC_drho_simp(np.array([0.0]), U, 1.0, 0.0, 3.0)
    b = tr.dC_drho_simp(np.array([1e-6]), U, 1.0, 0.0, 3.0)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_allclose(b, -2 * 3 * 1e-12 / max(1e-18, 1e-12))
