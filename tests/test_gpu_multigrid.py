"""Multigrid-preconditioned solves against the oracle's direct solve.

The multigrid is only a preconditioner: the solution is fixed by the
operator and the tolerance, so these tests check the solution against
scipy's SuperLU on the oracle's assembled matrix, that hierarchies with
factor-2, 3 and 5 coarsening are taken, and that the preconditioner cuts the
iteration count far below Jacobi's."""
import numpy as np
import pytest

from oracle import fem as ofem
from oracle import problems as oprob

pytestmark = pytest.mark.gpu


def _fem(task, physics, precision, rtol, precond, max_coarse=2000, degree=2):
    from sktopt_amd import fea, mesh as pmesh
    pm = task.mesh
    m = pmesh.structured_hex(*pm.grid, *pm.h)
    cfg = fea.LinearSolverConfig(rtol=rtol, precision=precision, preconditioner=precond,
                                 mg_max_coarse=max_coarse, mg_smooth_degree=degree)
    if physics == "elastic":
        t = pmesh.LinearElasticity(m, task.dirichlet_nodes, task.dirichlet_dofs, task.forces,
                                   task.neumann_nodes, None, task.E0, task.nu)
        return fea.FEM_SimpLinearElasticity(t, 1e-9, solver_config=cfg)
    t = pmesh.LinearHeatConduction(m, task.dirichlet_nodes, task.dirichlet_values,
                                   task.forces[0], None, task.E0)
    return fea.FEM_SimpLinearHeatConduction(t, 1e-3, solver_config=cfg)


MG_CASES = [
    # name, task factory, physics, max_coarse, expected levels
    ("hex-16x8x8", lambda: oprob.cantilever(16, 8, 8), "elastic", 2000, 2),
    ("hex-16x8x8-deep", lambda: oprob.cantilever(16, 8, 8), "elastic", 40, 4),
    ("hex-15x6x6-f3f5", lambda: oprob.cantilever(15, 6, 6), "elastic", 40, 3),
    ("heat-12", lambda: oprob.heat_patch(12, patch=4), "heat", 30, 4),
]


@pytest.mark.parametrize("precision,rtol,ctol", [("fp64", 1e-12, 1e-9), ("mixed", 1e-10, 1e-8)])
@pytest.mark.parametrize("name,mk,physics,max_coarse,levels", MG_CASES)
def test_mg_solve_matches_direct(gpu, precision, rtol, ctol, name, mk, physics, max_coarse,
                                 levels):
    task = mk()
    fem = _fem(task, physics, precision, rtol, "mg", max_coarse)
    rng = np.random.default_rng(5)
    rho = rng.uniform(0.0, 1.0, task.n_elements)
    p = 3.0
    ofe = ofem.ElasticFEM(task, 1e-9) if physics == "elastic" else ofem.HeatFEM(task, 1e-3)
    Uo = np.zeros((task.ndof, task.n_tasks))
    co = ofe.objectives_multi_load(rho, p, Uo)
    Ug = np.zeros((task.ndof, task.n_tasks))
    cg = fem.objectives_multi_load(rho, p, Ug)
    assert fem.last_solve["preconditioner"] == "mg"
    assert fem.last_solve["levels"] == levels
    np.testing.assert_allclose(cg, co, rtol=ctol)
    assert np.abs(Ug - Uo).max() <= 1e-6 * np.abs(Uo).max()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_mg_cuts_iterations(gpu, precision):
    task = oprob.cantilever(24, 12, 12)
    rng = np.random.default_rng(9)
    rho = rng.uniform(0.05, 1.0, task.n_elements)
    its = {}
    for pc in ("jacobi", "mg"):
        fem = _fem(task, "elastic", precision, 1e-8, pc)
        fem.objectives_multi_load(rho, 3.0, np.zeros((task.ndof, 1)))
        its[pc] = fem.last_solve["iters"]
    assert its["mg"] * 8 <= its["jacobi"], its


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_mg_smoothing_degrees(gpu, precision):
    """Chebyshev degrees 1..4 (the fine level's first steps take a different
    kernel path for degree 1, 2 and >= 3): every degree solves to the direct
    solution, and more smoothing never needs more PCG iterations."""
    task = oprob.cantilever(24, 12, 12)
    rng = np.random.default_rng(11)
    rho = np.where(rng.uniform(size=task.n_elements) < 0.5, 1.0, 0.02)
    ofe = ofem.ElasticFEM(task, 1e-9)
    Uo = np.zeros((task.ndof, 1))
    co = ofe.objectives_multi_load(rho, 3.0, Uo)
    its = {}
    for degree in (1, 2, 3, 4):
        fem = _fem(task, "elastic", precision, 1e-10, "mg", degree=degree)
        U = np.zeros((task.ndof, 1))
        c = fem.objectives_multi_load(rho, 3.0, U)
        assert fem.last_solve["preconditioner"] == "mg" and fem.last_solve["converged"]
        np.testing.assert_allclose(c, co, rtol=1e-8, err_msg=f"degree {degree}")
        assert np.abs(U - Uo).max() <= 1e-6 * np.abs(Uo).max(), degree
        its[degree] = fem.last_solve["iters"]
    print(f"[mg degrees {precision}] iterations {its}", flush=True)
    for d in (2, 3, 4):
        assert its[d] <= its[d - 1], its


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_mg_lagged_coarsest_inverse(gpu, precision, monkeypatch):
    """Solves for rho1, rho2, rho2 on one solver.  With the lagged coarsest
    inverse the second solve preconditions with the inverse of rho1's coarse
    matrix (computed while nothing waits for it) and still reaches the direct
    solution; the third has rho2's own inverse again and is bit-identical to
    the same sequence without the lag."""
    task = oprob.cantilever(24, 12, 12)
    rng = np.random.default_rng(21)
    rho1 = np.where(rng.uniform(size=task.n_elements) < 0.5, 1.0, 0.05)
    rho2 = np.clip(rho1 + rng.uniform(-0.2, 0.2, task.n_elements), 0.05, 1.0)
    ofe = ofem.ElasticFEM(task, 1e-9)
    ref = {}
    for name, rho in (("1", rho1), ("2", rho2)):
        Uo = np.zeros((task.ndof, 1))
        ref[name] = (ofe.objectives_multi_load(rho, 3.0, Uo), Uo)
    runs = {}
    for lag in ("1", "0"):
        monkeypatch.setenv("SKT_MG_LAG", lag)
        fem = _fem(task, "elastic", precision, 1e-10, "mg")
        out = []
        for name, rho in (("1", rho1), ("2", rho2), ("2", rho2)):
            U = np.zeros((task.ndof, 1))
            c = fem.objectives_multi_load(rho, 3.0, U)
            assert fem.last_solve["preconditioner"] == "mg" and fem.last_solve["converged"]
            np.testing.assert_allclose(c, ref[name][0], rtol=1e-8)
            assert np.abs(U - ref[name][1]).max() <= 1e-6 * np.abs(ref[name][1]).max()
            out.append((U, fem.last_solve["iters"]))
        runs[lag] = out
    print(f"[mg lag {precision}] iterations lagged {[o[1] for o in runs['1']]}, "
          f"exact {[o[1] for o in runs['0']]}", flush=True)
    for i in (0, 2):
        np.testing.assert_array_equal(runs["1"][i][0], runs["0"][i][0], err_msg=f"solve {i}")
        assert runs["1"][i][1] == runs["0"][i][1]


def test_mg_multiload_columns_independent(gpu):
    """Each column of a blocked multi-load MG solve equals its own solve."""
    task = oprob.multiload(16, 8, 8, n_loads=4, seed=3)
    rng = np.random.default_rng(4)
    rho = rng.uniform(0.05, 1.0, task.n_elements)
    fem = _fem(task, "elastic", "fp64", 1e-12, "mg")
    U = np.zeros((task.ndof, 4))
    fem.objectives_multi_load(rho, 3.0, U)
    ofe = ofem.ElasticFEM(task, 1e-9)
    Uo = np.zeros((task.ndof, 4))
    ofe.objectives_multi_load(rho, 3.0, Uo)
    assert np.abs(U - Uo).max() <= 1e-8 * np.abs(Uo).max()


@pytest.mark.parametrize("precision,rtol", [("fp64", 1e-12), ("mixed", 1e-10)])
def test_mg_stopped_columns_are_skipped_safely(gpu, precision, rtol):
    """Columns 0 and 1 start from their exact solution, so they stop at the
    first check while columns 2 and 3 iterate; the operator and V-cycle
    then skip the workgroups of the stopped pair.  Every column must still
    match the direct solve."""
    task = oprob.multiload(16, 8, 8, n_loads=4, seed=3)
    rng = np.random.default_rng(4)
    rho = rng.uniform(0.05, 1.0, task.n_elements)
    ofe = ofem.ElasticFEM(task, 1e-9)
    Uo = np.zeros((task.ndof, 4))
    ofe.objectives_multi_load(rho, 3.0, Uo)
    fem = _fem(task, "elastic", precision, rtol, "mg")
    U = np.zeros((task.ndof, 4))
    U[:, :2] = Uo[:, :2]
    fem.objectives_multi_load(rho, 3.0, U)
    assert fem.last_solve["iters"] > 0
    assert np.abs(U - Uo).max() <= 100 * rtol * np.abs(Uo).max()


def test_mg_request_on_unstructured_mesh_is_rejected(gpu):
    from sktopt_amd import fea, mesh as pmesh
    task = oprob.cantilever(4, 3, 3, kind="tet4", jitter=0.2, seed=1)
    m = pmesh.MeshTet(task.mesh.p, task.mesh.t)
    t = pmesh.LinearElasticity(m, task.dirichlet_nodes, task.dirichlet_dofs, task.forces,
                               task.neumann_nodes, None, task.E0, task.nu)
    cfg = fea.LinearSolverConfig(rtol=1e-10, precision="fp64", preconditioner="mg")
    fem = fea.FEM_SimpLinearElasticity(t, 1e-9, solver_config=cfg)
    with pytest.raises(ValueError):
        fem.objectives_multi_load(np.full(task.n_elements, 0.5), 3.0,
                                  np.zeros((task.ndof, 1)))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_mg_stencil_layouts_match(gpu, precision, monkeypatch):
    """Coarse-level stencils are stored tile-major (big tiles on large
    levels, small tiles below) and applied by the LDS-tiled kernel of that
    tile shape, or node-major with one thread per row.  All three sum the 27
    blocks in the same order, so forcing one layout on every level
    (SKT_MG_STENCIL_LAYOUT) must reproduce the same solve exactly, including
    the iteration count."""
    task = oprob.multiload(32, 16, 16, n_loads=2, seed=7)
    rng = np.random.default_rng(11)
    rho = rng.uniform(0.05, 1.0, task.n_elements)
    out = {}
    for mode in ("node", "big", "small"):
        monkeypatch.setenv("SKT_MG_STENCIL_LAYOUT", mode)
        fem = _fem(task, "elastic", precision, 1e-10, "mg", max_coarse=100)
        U = np.zeros((task.ndof, 2))
        fem.objectives_multi_load(rho, 3.0, U)
        assert fem.last_solve["levels"] >= 3
        out[mode] = (U, fem.last_solve["iters"])
    for mode in ("big", "small"):
        assert out[mode][1] == out["node"][1], mode
        np.testing.assert_array_equal(out[mode][0], out["node"][0])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("physics", ["elastic", "heat"])
def test_mg_two_step_smoothing_pass_matches_separate_steps(gpu, precision, physics,
                                                            monkeypatch):
    """At degree 2 the fine level's two Chebyshev steps from the zero iterate
    run as one operator pass, the first iterate s0 D^-1 b formed while the
    node planes are staged.  The separate form (first step, then one fused
    step, forced by SKT_MG_CHEB0=0) does the same arithmetic, so the solves
    agree bit for bit, iteration count included."""
    if physics == "elastic":
        task, max_coarse = oprob.multiload(32, 16, 16, n_loads=2, seed=7), 100
    else:
        task, max_coarse = oprob.heat_patch(16, patch=4), 40
    rng = np.random.default_rng(5)
    rho = rng.uniform(0.05, 1.0, task.n_elements)
    out = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("SKT_MG_CHEB0", flag)
        fem = _fem(task, physics, precision, 1e-10, "mg", max_coarse=max_coarse)
        U = np.zeros((task.ndof, task.n_tasks))
        fem.objectives_multi_load(rho, 3.0, U)
        assert fem.last_solve["levels"] >= 2
        out[flag] = (U, fem.last_solve["iters"])
    assert out["0"][1] == out["1"][1]
    np.testing.assert_array_equal(out["0"][0], out["1"][0])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("physics", ["elastic", "heat"])
def test_mg_factor2_restriction_matches_row_grid_kernels(gpu, precision, physics, monkeypatch):
    """Restriction into a level coarsened by 2 on every axis runs as one
    LDS-staged pass that also forms the coarse level's first smoothing
    iterate; it sums in the order of the row-grid kernels
    (SKT_MG_TRANSFERS=rowgrid: two separable passes, then the separate first
    step), so the solves agree bit for bit."""
    if physics == "elastic":
        task, max_coarse = oprob.multiload(32, 16, 16, n_loads=2, seed=7), 100
    else:
        task, max_coarse = oprob.heat_patch(16, patch=4), 40
    rng = np.random.default_rng(9)
    rho = rng.uniform(0.05, 1.0, task.n_elements)
    out = {}
    for mode in ("fused", "rowgrid"):
        monkeypatch.setenv("SKT_MG_TRANSFERS", mode)
        fem = _fem(task, physics, precision, 1e-10, "mg", max_coarse=max_coarse)
        U = np.zeros((task.ndof, task.n_tasks))
        fem.objectives_multi_load(rho, 3.0, U)
        assert fem.last_solve["levels"] >= 3
        out[mode] = (U, fem.last_solve["iters"])
    assert out["fused"][1] == out["rowgrid"][1]
    np.testing.assert_array_equal(out["fused"][0], out["rowgrid"][0])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("case", ["elastic-8", "elastic-4", "elastic-3", "heat-1"])
def test_mg_row_blocked_coarsest_apply_matches_wave_per_row(gpu, precision, case, monkeypatch):
    """The coarsest level's dense product reads the inverse once for every
    column group (b staged in LDS, two rows per wavefront); per row and column
    it sums in the order of the wavefront-per-row kernel
    (SKT_MG_DENSE_APPLY=wave), so the solves agree bit for bit, for 1, 2 and 4
    column groups and the 3-column block that takes the other kernel."""
    physics, n = case.split("-")
    n = int(n)
    if physics == "elastic":
        task, max_coarse = oprob.multiload(32, 16, 16, n_loads=n, seed=7), 700
    else:
        task, max_coarse = oprob.heat_patch(16, patch=4), 800
    rng = np.random.default_rng(11)
    rho = rng.uniform(0.05, 1.0, task.n_elements)
    out = {}
    for mode in ("rows", "wave"):
        monkeypatch.setenv("SKT_MG_DENSE_APPLY", mode)
        fem = _fem(task, physics, precision, 1e-10, "mg", max_coarse=max_coarse)
        U = np.zeros((task.ndof, task.n_tasks))
        fem.objectives_multi_load(rho, 3.0, U)
        assert fem.last_solve["levels"] >= 2
        out[mode] = (U, fem.last_solve["iters"])
    assert out["rows"][1] == out["wave"][1]
    np.testing.assert_array_equal(out["rows"][0], out["wave"][0])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("precond", ["mg", "jacobi"])
def test_pcg_solution_update_in_direction_update_matches(gpu, precision, precond, monkeypatch):
    """x += alpha p runs in the direction update of the same iteration (one
    read of p less); SKT_PCG_X_IN_UPDATE=1 keeps it in the r update.  With 8
    load columns stopping at different iterations (a stopping column still
    needs its last alpha p) and the mixed path's reliable updates, the solves
    agree bit for bit, iteration counts included."""
    task = oprob.multiload(24, 12, 12, n_loads=8, seed=5)
    rng = np.random.default_rng(17)
    rho = rng.uniform(0.05, 1.0, task.n_elements)
    out = {}
    for flag in ("0", "1"):
        monkeypatch.setenv("SKT_PCG_X_IN_UPDATE", flag)
        fem = _fem(task, "elastic", precision, 1e-9, precond)
        U = np.zeros((task.ndof, task.n_tasks))
        c = fem.objectives_multi_load(rho, 3.0, U)
        assert fem.last_solve["converged"]
        out[flag] = (U, c, fem.last_solve["iters"])
    assert out["0"][2] == out["1"][2]
    np.testing.assert_array_equal(out["0"][0], out["1"][0])
    np.testing.assert_array_equal(out["0"][1], out["1"][1])


@pytest.mark.parametrize("precision,rtol,ctol", [("fp64", 1e-12, 1e-9), ("mixed", 1e-10, 1e-8)])
def test_mg_reused_coarsest_inverse_reaches_direct_solutions(gpu, precision, rtol, ctol,
                                                             monkeypatch):
    """Once the solves are steady (the last one took no more iterations than
    the one before), a set-up keeps the previous coarsest inverse, for at most
    two set-ups per inverse.  That is still an SPD coarse solve: every solve
    of a sequence with slowly changing densities reaches the direct solution,
    and the sequence differs from one with a fresh inverse at every set-up
    (SKT_MG_COARSE_REFRESH=1) only within the tolerance."""
    task = oprob.multiload(32, 16, 16, n_loads=2, seed=7)
    rng = np.random.default_rng(3)
    rho0 = rng.uniform(0.05, 1.0, task.n_elements)
    rhos = [np.clip(rho0 * (1.0 + 0.02 * k), 0.0, 1.0) for k in range(6)]
    runs = {}
    for refresh in ("1", "2"):
        monkeypatch.setenv("SKT_MG_COARSE_REFRESH", refresh)
        fem = _fem(task, "elastic", precision, rtol, "mg", max_coarse=100)
        U = np.zeros((task.ndof, task.n_tasks))
        out = []
        for rho in rhos:  # each solve starts from the previous solution
            c = fem.objectives_multi_load(rho, 3.0, U)
            assert fem.last_solve["levels"] >= 3
            out.append((np.array(c), U.copy()))
        runs[refresh] = out
    ofe = ofem.ElasticFEM(task, 1e-9)
    for k, rho in enumerate(rhos):
        Uo = np.zeros((task.ndof, task.n_tasks))
        co = ofe.objectives_multi_load(rho, 3.0, Uo)
        for refresh in runs:
            c, U = runs[refresh][k]
            np.testing.assert_allclose(c, co, rtol=ctol)
            assert np.abs(U - Uo).max() <= 1e-6 * np.abs(Uo).max(), (refresh, k)
    # the reuse took effect (a different coarse solve changes the iterates)
    assert any(not np.array_equal(runs["1"][k][1], runs["2"][k][1]) for k in range(6))
