"""Differentiable compliance (sktopt_amd.core.Compliance): the backward
against the oracle's sensitivity chain and against central finite
differences, up to the 200x100x100 (6.2M DOF) cantilever with a radius-2.5
filter and beta = 8 projection."""
import numpy as np
import pytest

from oracle import driver as odrv
from oracle import fem as ofem
from oracle import filter as ofilt
from oracle import problems as oprob
from oracle import transforms as tr

pytestmark = pytest.mark.gpu


def _problem(task, r, beta, precision="fp64", rtol=1e-13, interp="SIMP"):
    from sktopt_amd import core, fea
    from sktopt_amd.filters import SpacialFilter
    from sktopt_amd.mesh import structured_hex, LinearElasticity
    pm = task.mesh
    m = structured_hex(*pm.grid, *pm.h)
    t = LinearElasticity(m, task.dirichlet_nodes, task.dirichlet_dofs, task.forces,
                         task.neumann_nodes, None, task.E0, task.nu)
    cfg = fea.LinearSolverConfig(rtol=rtol, precision=precision, maxiter=20000)
    di = fea.ramp_interpolation if interp == "RAMP" else fea.simp_interpolation
    fem = fea.FEM_SimpLinearElasticity(t, 1e-9, density_interpolation=di, solver_config=cfg)
    filt = SpacialFilter(None, task.elements_volume, r, task.design_mask,
                         element_centers=task.centers)
    return core.ComplianceProblem(fem, filt, beta=beta, eta=0.5, p=3.0)


def _rho(task, seed, lo=0.3, hi=0.7):
    rng = np.random.default_rng(seed)
    rho = np.ones(task.n_elements)
    d = np.nonzero(task.design_mask)[0]
    rho[d] = rng.uniform(lo, hi, d.size)
    return rho, d, rng


def _grad(prob, rho):
    import torch
    x = torch.tensor(rho, device="cuda", requires_grad=True)
    C = prob.compliance(x)
    C.backward()
    return C.item(), x.grad.detach().cpu().numpy()


def _fd(prob, rho, idx, h=1e-6):
    import torch
    U0 = prob.U.clone()
    out = []
    with torch.no_grad():
        for e in idx:
            v = []
            for s in (1.0, -1.0):
                x = torch.tensor(rho, device="cuda")
                x[e] += s * h
                prob.U.copy_(U0)
                v.append(prob.compliance(x).item())
            out.append((v[0] - v[1]) / (2 * h))
    prob.U.copy_(U0)
    return np.array(out)


def test_backward_matches_reference_sensitivities_c0(gpu):
    """40x20x20 cantilever: the gradient equals the reference chain's dC_full
    (common_density.py:1073-1108) at the first OC iteration."""
    task = oprob.cantilever(40, 20, 20)
    S = odrv.Sched
    cfg = odrv.Config(method="OC", max_iters=10, p=S("Constant", 3.0, 3.0),
                      vol_frac=S("Constant", 0.3, 0.3), beta=S("Constant", 8.0, 8.0),
                      filter_radius=S("Constant", 1.5, 1.5), move_limit=S("Constant", 0.2, 0.2),
                      E_min_coeff=1e-9, solver="spsolve")
    loop = odrv.DensityLoop(cfg, task)
    rho0 = loop.rho.copy()
    rec = loop.step(1)
    prob = _problem(oprob.cantilever(40, 20, 20), 1.5, 8.0)
    C, g = _grad(prob, rho0)
    # beta = 8 projects rho = 0.3 to a soft design (C ~ 3e4): 1e-9 relative is
    # the float64 bar, and about what SuperLU itself holds here
    assert abs(C - rec["compliance"]) <= 1e-9 * rec["compliance"]
    ref = rec["dC_full"]
    err = np.abs(g - ref).max() / np.abs(ref).max()
    assert err <= 1e-8, err


def test_backward_ramp_exact_matches_oracle(gpu):
    task = oprob.cantilever(12, 6, 6)
    rho, d, _ = _rho(task, 4)
    f = ofilt.SpacialFilter(task.centers, 2.0, task.design_mask)
    fem = ofem.ElasticFEM(task, 1e-9, interpolation="RAMP")
    rf = f.forward(rho)
    rp = tr.heaviside_projection(rf, 4.0, 0.5)
    U = np.zeros((task.ndof, 1))
    C = fem.objectives_multi_load(rp, 3.0, U).mean()
    en = fem.energy_multi_load(rp, 3.0, U)[:, 0]
    dH = tr.heaviside_projection_derivative(rf, 4.0, 0.5)
    ref = f.gradient(tr.dC_drho_ramp_exact(rp, en, 1.0, 1e-9, 3.0) * dH)
    prob = _problem(oprob.cantilever(12, 6, 6), 2.0, 4.0, interp="RAMP")
    Cg, g = _grad(prob, rho)
    assert abs(Cg - C) <= 1e-10 * C
    assert np.abs(g - ref).max() <= 1e-9 * np.abs(ref).max()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-5), ("mixed", 1e-4)])
def test_backward_matches_central_differences(gpu, precision, tol):
    """24x12x12, radius 2.5, beta 8: 32 seeded design elements, step 1e-6;
    |fd - grad| <= tol * max |grad| over the sample (the finite difference
    resolves compliance changes of ~1e-9 relative, so its own error sets the
    bar)."""
    task = oprob.cantilever(24, 12, 12)
    rho, d, rng = _rho(task, 7)
    prob = _problem(task, 2.5, 8.0, precision=precision,
                    rtol=1e-13 if precision == "fp64" else 1e-12)
    _, g = _grad(prob, rho)
    idx = rng.choice(d, 32, replace=False)
    fd = _fd(prob, rho, idx)
    assert np.abs(fd - g[idx]).max() <= tol * np.abs(g[idx]).max()


def test_backward_matches_central_differences_at_scale(gpu):
    """200x100x100 (2M elements, 6,151,203 DOF), single load, radius 2.5,
    beta 8, float64: 64 seeded design elements, step 1e-6.  Compliance
    differences here are ~1e-11 relative, so the float64 solves (rtol 1e-13)
    bound the agreement at a few 1e-4 of max |grad|."""
    task = oprob.cantilever(200, 100, 100)
    rho, d, rng = _rho(task, 0)
    prob = _problem(task, 2.5, 8.0)
    _, g = _grad(prob, rho)
    assert prob.fem.last_solve["preconditioner"] == "mg"
    idx = rng.choice(d, 64, replace=False)
    fd = _fd(prob, rho, idx)
    assert np.abs(fd - g[idx]).max() <= 5e-4 * np.abs(g[idx]).max()


def test_two_forwards_then_backward_of_the_first(gpu):
    """Each graph keeps its own displacement: a second forward (which
    overwrites the problem's warm-start block) leaves the first graph's
    gradient equal to that of a lone forward + backward."""
    import torch
    task = oprob.cantilever(12, 6, 6)
    rho_a, _, _ = _rho(task, 11)
    rho_b, _, _ = _rho(task, 12, lo=0.5, hi=0.9)
    ref = _problem(oprob.cantilever(12, 6, 6), 2.0, 4.0)
    _, g_ref = _grad(ref, rho_a)
    prob = _problem(oprob.cantilever(12, 6, 6), 2.0, 4.0)
    xa = torch.tensor(rho_a, device="cuda", requires_grad=True)
    xb = torch.tensor(rho_b, device="cuda", requires_grad=True)
    Ca = prob.compliance(xa)
    Cb = prob.compliance(xb)
    Ca.backward()
    Cb.backward()
    np.testing.assert_allclose(xa.grad.cpu().numpy(), g_ref, rtol=1e-9,
                               atol=1e-12 * np.abs(g_ref).max())
    _, g_b = _grad(_problem(oprob.cantilever(12, 6, 6), 2.0, 4.0), rho_b)
    np.testing.assert_allclose(xb.grad.cpu().numpy(), g_b, rtol=1e-9,
                               atol=1e-12 * np.abs(g_b).max())


def test_heat_objectives_without_self_adjoint_backward_are_rejected(gpu):
    """The backward is the self-adjoint compliance gradient: a heat FEM with
    the averaged-temperature objective has none, so the problem refuses it."""
    from sktopt_amd import core, fea
    from sktopt_amd.filters import SpacialFilter
    from sktopt_amd.mesh import structured_hex, LinearHeatConduction
    task = oprob.heat_patch(6, patch=2, source=1e-2)
    pm = task.mesh
    t = LinearHeatConduction(structured_hex(*pm.grid, *pm.h), task.dirichlet_nodes,
                             task.dirichlet_values, task.forces[0], None, task.E0,
                             objective="averaged_temp")
    fem = fea.FEM_SimpLinearHeatConduction(t, 1e-3)
    filt = SpacialFilter(None, task.elements_volume, 1.5, task.design_mask,
                         element_centers=task.centers)
    with pytest.raises(NotImplementedError):
        core.ComplianceProblem(fem, filt, beta=1.0)


def _heat_problem(n, patch, r, beta, rtol=1e-13, mask=None):
    from sktopt_amd import core, fea
    from sktopt_amd.filters import SpacialFilter
    from sktopt_amd.mesh import toy_problem
    task = toy_problem.heat_patch(n, patch=patch, source=1e-2)
    cfg = fea.LinearSolverConfig(rtol=rtol, precision="fp64", maxiter=20000)
    fem = fea.FEM_SimpLinearHeatConduction(task, 1e-3, solver_config=cfg)
    centers = task.mesh.element_centers()
    filt = SpacialFilter(None, task.elements_volume, r,
                         task.design_mask if mask is None else mask, element_centers=centers)
    return core.ComplianceProblem(fem, filt, beta=beta, eta=0.5, p=3.0), task


def test_heat_backward_matches_oracle_chain(gpu):
    """Heat compliance (conductivity k0 = 1, kmin = 1e-3, uniform source,
    T = 0 on a centred patch of the z = 0 face): the gradient equals the
    reference's chain (solver_heat.py:928-950 energies, common_density.py:
    1073-1108) computed by the oracle."""
    to = oprob.heat_patch(12, patch=4, source=1e-2)
    prob, task = _heat_problem(12, 4, 2.0, 4.0, mask=to.design_mask)
    rng = np.random.default_rng(3)
    rho = np.ones(to.n_elements)
    d = np.nonzero(to.design_mask)[0]
    rho[d] = rng.uniform(0.2, 0.9, d.size)
    f = ofilt.SpacialFilter(to.centers, 2.0, to.design_mask)
    fem = ofem.HeatFEM(to, 1e-3)
    rf = f.forward(rho)
    rp = tr.heaviside_projection(rf, 4.0, 0.5)
    U = np.zeros((to.ndof, 1))
    J = fem.objectives_multi_load(rp, 3.0, U).mean()
    dC = fem.compliance_sensitivity_multi_load(rp, 3.0, U)[:, 0]
    ref = f.gradient(dC * tr.heaviside_projection_derivative(rf, 4.0, 0.5))
    Jg, g = _grad(prob, rho)
    assert abs(Jg - J) <= 1e-10 * abs(J)
    assert np.abs(g - ref).max() <= 1e-9 * np.abs(ref).max()


def test_heat_backward_central_differences_at_160(gpu):
    """The heat configuration at size (160^3 elements, 4.2M DOF, r = 2.0,
    multigrid solves): 16 seeded design elements, central differences with
    step 1e-6 against the backward pass."""
    prob, task = _heat_problem(160, 16, 2.0, 1.0)
    rng = np.random.default_rng(5)
    rho = np.ones(task.mesh.nelements)
    d = np.asarray(task.design_elements)
    rho[d] = rng.uniform(0.2, 0.9, d.size)
    _, g = _grad(prob, rho)
    assert prob.fem.last_solve["preconditioner"] == "mg"
    idx = rng.choice(d, 16, replace=False)
    fd = _fd(prob, rho, idx)
    err = np.abs(fd - g[idx]).max() / np.abs(g[idx]).max()
    print(f"[heat 160^3 autograd] max |fd - grad| / max|grad| = {err:.2e}", flush=True)
    assert err <= 5e-4, err


def _tet_problem(task, r, beta, precision="fp64", rtol=1e-13):
    from sktopt_amd import core, fea
    from sktopt_amd.filters import SpacialFilter
    from sktopt_amd.mesh import LinearElasticity
    from sktopt_amd.mesh.meshes import MeshTet
    pm = task.mesh
    m = MeshTet(pm.p, pm.t, cell_grid=pm.grid)
    t = LinearElasticity(m, task.dirichlet_nodes, task.dirichlet_dofs, task.forces,
                         task.neumann_nodes, None, task.E0, task.nu)
    cfg = fea.LinearSolverConfig(rtol=rtol, precision=precision, maxiter=20000)
    fem = fea.FEM_SimpLinearElasticity(t, 1e-9, solver_config=cfg)
    filt = SpacialFilter(None, task.elements_volume, r, task.design_mask,
                         element_centers=task.centers)
    return core.ComplianceProblem(fem, filt, beta=beta, eta=0.5, p=3.0)


def test_tet_backward_matches_reference_chain_and_differences(gpu):
    """Jittered tet cantilever (16x8x8 cells, 6144 tets, multigrid solve): the
    gradient equals the oracle's dC_full at the first OC iteration, and central
    differences on 24 design elements."""
    mk = lambda: oprob.cantilever(16, 8, 8, kind="tet4", jitter=0.2, seed=4)  # noqa: E731
    S = odrv.Sched
    cfg = odrv.Config(method="OC", max_iters=10, p=S("Constant", 3.0, 3.0),
                      vol_frac=S("Constant", 0.3, 0.3), beta=S("Constant", 4.0, 4.0),
                      filter_radius=S("Constant", 1.5, 1.5), move_limit=S("Constant", 0.2, 0.2),
                      E_min_coeff=1e-9, solver="spsolve")
    loop = odrv.DensityLoop(cfg, mk())
    rho0 = loop.rho.copy()
    rec = loop.step(1)
    task = mk()
    prob = _tet_problem(task, 1.5, 4.0)
    C, g = _grad(prob, rho0)
    assert prob.fem.last_solve["preconditioner"] == "mg"
    assert abs(C - rec["compliance"]) <= 1e-9 * rec["compliance"]
    ref = rec["dC_full"]
    assert np.abs(g - ref).max() <= 1e-8 * np.abs(ref).max()
    rho, d, rng = _rho(task, 9)
    _, g = _grad(prob, rho)
    idx = rng.choice(d, 24, replace=False)
    fd = _fd(prob, rho, idx)
    assert np.abs(fd - g[idx]).max() <= 1e-5 * np.abs(g[idx]).max()
