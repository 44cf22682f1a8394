"""Pin the oracle's compliance sensitivity chain (filter -> Heaviside ->
SIMP/RAMP -> compliance, ``core/optimizers/common_density.py:1073-1108``)
against central finite differences of the oracle's own compliance, with
exact (SuperLU) solves on small meshes."""
import numpy as np
import pytest

from oracle import fem as ofem
from oracle import filter as ofilt
from oracle import problems as oprob
from oracle import transforms as tr


def _chain(task, rho, beta, r, p, interp, dcfn):
    f = ofilt.SpacialFilter(task.centers, r, task.design_mask)
    fem = ofem.ElasticFEM(task, 1e-9, interpolation=interp)

    def compliance(x):
        rf = f.forward(x)
        rp = tr.heaviside_projection(rf, beta, 0.5)
        U = np.zeros((task.ndof, task.n_tasks))
        return fem.objectives_multi_load(rp, p, U).mean(), rf, rp, U

    C, rf, rp, U = compliance(rho)
    en = fem.energy_multi_load(rp, p, U)
    dH = tr.heaviside_projection_derivative(rf, beta, 0.5)
    g = np.zeros(task.n_elements)
    for k in range(task.n_tasks):
        g += f.gradient(dcfn(rp, en[:, k], task.E0, task.E0 * 1e-9, p) * dH)
    g /= task.n_tasks
    return compliance, g


@pytest.mark.parametrize("interp,dcfn", [("SIMP", tr.dC_drho_simp),
                                         ("RAMP", tr.dC_drho_ramp_exact)])
@pytest.mark.parametrize("mk", [lambda: oprob.cantilever(8, 4, 4),
                                lambda: oprob.multiload(6, 3, 3, n_loads=2, seed=1)])
def test_chain_matches_central_differences(interp, dcfn, mk):
    task = mk()
    rng = np.random.default_rng(3)
    rho = np.ones(task.n_elements)
    d = np.nonzero(task.design_mask)[0]
    rho[d] = rng.uniform(0.2, 0.8, d.size)
    compliance, g = _chain(task, rho, 8.0, 1.5, 3.0, interp, dcfn)
    h = 1e-6
    idx = rng.choice(d, 10, replace=False)
    fd = []
    for e in idx:
        a = rho.copy(); a[e] += h
        b = rho.copy(); b[e] -= h
        fd.append((compliance(a)[0] - compliance(b)[0]) / (2 * h))
    fd = np.array(fd)
    assert np.abs(fd - g[idx]).max() <= 1e-6 * np.abs(g[idx]).max()
    # non-design elements carry no sensitivity (the filter adjoint zeroes them)
    assert np.all(g[~task.design_mask] == 0.0)


def test_reference_ramp_expression_is_not_the_derivative():
    """The reference's RAMP sensitivity (core/derivatives.py:56-61) differs
    from the derivative of ramp_interpolation by the factor
    (1 + p (1 - 2 rho)) / (1 + p) -- equal only at rho = 1/2."""
    rho = np.array([0.2, 0.5, 0.9])
    U = np.ones(3)
    q = tr.dC_drho_ramp(rho, U, 1.0, 1e-9, 3.0) / tr.dC_drho_ramp_exact(rho, U, 1.0, 1e-9, 3.0)
    np.testing.assert_allclose(q, (1 + 3 * (1 - 2 * rho)) / 4.0, rtol=1e-12)
