"""Pin the oracle's element matrices independently of the reference:
exact symbolic integration, rigid-body null space, constant-strain patch
test and closed-form bending energies."""
import numpy as np
import pytest
import sympy as sp

from oracle import elements as el
from oracle.mesh import HEX_CORNERS


def _sympy_hex_K(nu_val, h=(1, 1, 1), physics="elastic"):
    x, y, z = sp.symbols("x y z")
    hx, hy, hz = [sp.Rational(v) if isinstance(v, int) else sp.nsimplify(v) for v in h]
    N = []
    for cx, cy, cz in HEX_CORNERS:
        fx = x / hx if cx else 1 - x / hx
        fy = y / hy if cy else 1 - y / hy
        fz = z / hz if cz else 1 - z / hz
        N.append(fx * fy * fz)
    grads = [[sp.diff(n, v) for v in (x, y, z)] for n in N]

    def integ(expr):
        # exact term-wise integration of a polynomial over the box
        P = sp.Poly(sp.expand(expr), x, y, z)
        tot = sp.Integer(0)
        for (a, b, c), coef in P.terms():
            tot += coef * hx ** (a + 1) / (a + 1) * hy ** (b + 1) / (b + 1) * hz ** (c + 1) / (c + 1)
        return tot
    if physics == "heat":
        K = sp.zeros(8, 8)
        for a in range(8):
            for b in range(a, 8):
                K[a, b] = K[b, a] = integ(sum(grads[a][d] * grads[b][d] for d in range(3)))
        return np.array(K.evalf(), dtype=float)
    nu = sp.nsimplify(nu_val)
    lam = nu / ((1 + nu) * (1 - 2 * nu))
    mu = 1 / (2 * (1 + nu))
    K = sp.zeros(24, 24)
    for a in range(8):
        for b in range(8):
            for i in range(3):
                for j in range(3):
                    # lam * div u div v + mu (grad u : grad v + grad u : grad v^T)
                    ga, gb = grads[a], grads[b]
                    e = lam * ga[i] * gb[j] + mu * ga[j] * gb[i]
                    if i == j:
                        e += mu * sum(ga[d] * gb[d] for d in range(3))
                    K[3 * a + i, 3 * b + j] = integ(e)
    return np.array(K.evalf(), dtype=float)


@pytest.mark.parametrize("nu,h", [(0.3, (1, 1, 1)), (0.25, (2, 1, 0.5))])
def test_hex8_elasticity_matches_exact_integration(nu, h):
    K = el.hex8_elasticity_K0(nu, h)
    np.testing.assert_allclose(K, _sympy_hex_K(nu, h), rtol=1e-12, atol=1e-13)


def test_hex8_conduction_matches_exact_integration():
    np.testing.assert_allclose(el.hex8_conduction_K0((1, 1, 1)), _sympy_hex_K(0, physics="heat"),
                               rtol=1e-12, atol=1e-14)


def test_hex8_rigid_body_null_space_and_rank():
    K = el.hex8_elasticity_K0(0.3)
    X = HEX_CORNERS.astype(float)
    modes = []
    for c in range(3):
        t = np.zeros(24)
        t[c::3] = 1.0
        modes.append(t)
    for (i, j) in [(0, 1), (1, 2), (0, 2)]:
        r = np.zeros(24)
        r[i::3] = -X[:, j]
        r[j::3] = X[:, i]
        modes.append(r)
    for m in modes:
        assert np.abs(K @ m).max() < 1e-13
    w = np.linalg.eigvalsh(K)
    assert np.sum(w > 1e-10) == 18 and w.min() > -1e-12


@pytest.mark.parametrize("nu", [0.0, 0.3, 0.45])
def test_hex8_constant_strain_patch(nu):
    """u = G x gives energy 1/2 V (lam tr(eps)^2 + 2 mu eps:eps) exactly."""
    h = (1.5, 1.0, 0.75)
    K = el.hex8_elasticity_K0(nu, h)
    rng = np.random.default_rng(0)
    G = rng.standard_normal((3, 3))
    X = HEX_CORNERS * np.array(h)
    u = (X @ G.T).ravel()
    eps = 0.5 * (G + G.T)
    lam, mu = el.lame(1.0, nu)
    V = np.prod(h)
    ref = 0.5 * V * (lam * np.trace(eps) ** 2 + 2 * mu * np.sum(eps * eps))
    np.testing.assert_allclose(0.5 * u @ K @ u, ref, rtol=1e-12)


def test_hex8_bending_mode_energy_closed_form():
    """u_x = x*y on the unit cube: eps_xx = y, eps_xy = x/2, so the exact energy
    is 1/2 int[(lam+2mu) y^2 + mu x^2] = (lam+2mu)/6 + mu/6."""
    K = el.hex8_elasticity_K0(0.3)
    X = HEX_CORNERS.astype(float)
    u = np.zeros(24)
    u[0::3] = X[:, 0] * X[:, 1]
    lam, mu = el.lame(1.0, 0.3)
    np.testing.assert_allclose(0.5 * u @ K @ u, (lam + 2 * mu) / 6 + mu / 6, rtol=1e-12)


def test_tet4_patch_and_rigid_modes():
    rng = np.random.default_rng(1)
    p = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float).T + \
        0.1 * rng.standard_normal((3, 4))
    t = np.arange(4)[:, None]
    K = el.tet4_elasticity_K(p, t, 0.3)[0]
    g, vol = el.tet4_grads(p, t)
    G = rng.standard_normal((3, 3))
    u = (p.T @ G.T).ravel()
    eps = 0.5 * (G + G.T)
    lam, mu = el.lame(1.0, 0.3)
    ref = 0.5 * abs(vol[0]) * (lam * np.trace(eps) ** 2 + 2 * mu * np.sum(eps * eps))
    np.testing.assert_allclose(0.5 * u @ K @ u, ref, rtol=1e-12)
    for c in range(3):
        m = np.zeros(12)
        m[c::3] = 1.0
        assert np.abs(K @ m).max() < 1e-12
    assert np.sum(np.linalg.eigvalsh(K) > 1e-10) == 6


def test_tet4_conduction_linear_field():
    p = np.array([[0, 0, 0], [2, 0, 0], [0, 1, 0], [0, 0, 3]], float).T
    t = np.arange(4)[:, None]
    K = el.tet4_conduction_K(p, t)[0]
    a = np.array([0.3, -1.2, 0.7])
    T = p.T @ a
    np.testing.assert_allclose(T @ K @ T, 1.0 * a @ a, rtol=1e-12)   # V = 1
