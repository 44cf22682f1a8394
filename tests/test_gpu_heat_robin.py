"""Robin heat conduction on the GPU against the oracle
(fea/solver_heat.py:552-980, mesh/task_heat.py:60-139)."""
import numpy as np
import pytest

from oracle import driver as odrv
from oracle import heat_robin as hr
from oracle import mesh as om
from oracle import problems as oprob

pytestmark = pytest.mark.gpu


def _t(x):
    import torch
    return torch.tensor(np.ascontiguousarray(x), dtype=torch.float64, device="cuda")


def _s():
    import torch
    return torch.cuda.current_stream().cuda_stream


def _task(nx=10, ny=10, nz=4, objective="compliance", h=4e-2, design_robin=True, w=0.0,
          values=600.0):
    from sktopt_amd.mesh import structured_hex, toy_problem as tp
    from sktopt_amd.mesh.task import LinearHeatConduction
    m = structured_hex(nx, ny, nz)
    d = m.facets_satisfying(tp.get_points_in_range((nx, nx), (0, 2), (0, nz)))
    r0 = m.facets_satisfying(tp.get_points_in_range((0, 0), (0, ny), (0, nz)))
    r1 = m.facets_satisfying(tp.get_points_in_range((0, nx), (ny, ny), (0, nz)))
    return LinearHeatConduction.from_facets(m, d, values, [r0, r1], h, 300.0, design_robin,
                                            np.arange(m.nelements), 10.0, objective=objective,
                                            avg_temp_weight=w)


def _oracle_fem(t, objective):
    om_ = om.structured_hex(*t.mesh.grid, *t.mesh.h)
    facets = [t.mesh.facets[:, np.asarray(f)] for f in t.robin_facets_ids]
    return om_, hr.HeatRobinFEM(om_, t.dirichlet_dofs, t.dirichlet_values, t.k, 1e-3, facets,
                                t.robin_coefficient, t.robin_bc_value, emit=t.emit,
                                objective=objective, avg_temp_weight=t.avg_temp_weight)


def test_robin_kernels_match_oracle(gpu):
    import torch
    from sktopt_amd import _lib
    t = _task(7, 5, 3)
    m = om.structured_hex(*t.mesh.grid)
    nx, ny, nz = m.grid
    rng = np.random.default_rng(3)
    rho = rng.uniform(0.05, 1.0, m.nelements)
    rn = torch.empty(m.nnodes, dtype=torch.float64, device="cuda")
    cnt = torch.empty_like(rn)
    _lib.call("skt_hex8_nodal_average_pairing", nx, ny, nz, _t(rho).data_ptr(), rn.data_ptr(),
              cnt.data_ptr(), _s())
    rn_ref, cnt_ref = hr.nodal_average_reference(m, rho)
    np.testing.assert_array_equal(rn.cpu().numpy(), rn_ref)
    np.testing.assert_array_equal(cnt.cpu().numpy(), cnt_ref)
    for npts in (1, 2, 3):
        st = torch.empty((27, m.nnodes), dtype=torch.float64, device="cuda")
        L = torch.empty(m.nnodes, dtype=torch.float64, device="cuda")
        _lib.call("skt_hex8_robin_virtual", nx, ny, nz, 1.0, 1.0, 1.0, npts, rn.data_ptr(),
                  3.0, 4.0, 0.04, 300.0, st.data_ptr(), L.data_ptr(), _s())
        B, Lr = hr.virtual_robin_nodal(m, rn_ref, 3.0, 4.0, 0.04, 300.0, npts)
        x = rng.standard_normal((m.nnodes, 2))
        y = torch.empty((m.nnodes, 2), dtype=torch.float64, device="cuda")
        _lib.call("skt_stencil27_apply", nx, ny, nz, 2, st.data_ptr(), _t(x).data_ptr(),
                  y.data_ptr(), _s())
        ref = B @ x
        np.testing.assert_allclose(y.cpu().numpy(), ref, rtol=1e-12,
                                   atol=1e-13 * np.abs(ref).max())
        np.testing.assert_allclose(L.cpu().numpy(), Lr, rtol=1e-12, atol=1e-13 * np.abs(Lr).max())
    T = 300.0 + 100.0 * rng.uniform(size=(m.nnodes, 2))
    out = torch.empty((m.nelements, 2), dtype=torch.float64, device="cuda")
    work = torch.empty(m.nnodes, dtype=torch.float64, device="cuda")
    _lib.call("skt_hex8_robin_grad", nx, ny, nz, 1.0, 1.0, 1.0, 2, rn.data_ptr(), cnt.data_ptr(),
              2, _t(T).data_ptr(), 3.0, 4.0, 0.04, 300.0, work.data_ptr(), out.data_ptr(), _s())
    for c in range(2):
        ref = hr.robin_explicit_grad(m, rho, T[:, c], 3.0, 4.0, 0.04, 300.0)
        np.testing.assert_allclose(out[:, c].cpu().numpy(), ref, rtol=1e-11,
                                   atol=1e-12 * np.abs(ref).max())
    for npts in (2, 4):
        den = torch.empty(m.nelements, dtype=torch.float64, device="cuda")
        num = torch.empty((m.nelements, 2), dtype=torch.float64, device="cuda")
        rhs = torch.empty(m.nnodes, dtype=torch.float64, device="cuda")
        Td = _t(T)
        _lib.call("skt_hex8_heat_exchange", nx, ny, nz, 1.0, 1.0, 1.0, npts, rn.data_ptr(), 2,
                  Td.data_ptr(), 3.0, 4.0, 0.04, 300.0, den.data_ptr(), num.data_ptr(),
                  rhs.data_ptr(), _s())
        d_ref, n_ref, r_ref = hr.heat_exchange_terms(m, rn_ref, T, 3.0, 4.0, 0.04, 300.0, npts)
        np.testing.assert_allclose(den.sum().item(), d_ref, rtol=1e-12)
        np.testing.assert_allclose(-300.0 * 0.04 * num.sum(0).cpu().numpy(), n_ref, rtol=1e-11)
        r = rhs.cpu().numpy()
        np.testing.assert_allclose(r, r_ref, rtol=1e-12, atol=1e-13 * np.abs(r_ref).max())


@pytest.mark.parametrize("objective,precision,w,values",
                         [("compliance", "fp64", 0.0, 600.0),
                          ("averaged_temp", "fp64", 0.0, 600.0),
                          ("compliance", "mixed", 0.0, 600.0),
                          ("heat_exchange", "fp64", 0.0, 600.0),
                          ("heat_exchange", "fp64", 0.3, 600.0),
                          ("heat_exchange", "fp64", 0.3, [600.0, 450.0])])
def test_robin_objectives_match_oracle(gpu, objective, precision, w, values):
    from sktopt_amd import fea
    t = _task(objective=objective, w=w, values=values)
    _, ofem = _oracle_fem(t, objective)
    cfg = fea.LinearSolverConfig(rtol=1e-12 if precision == "fp64" else 1e-10,
                                 precision=precision)
    fem = fea.FEM_SimpLinearHeatConduction(t, 1e-3, solver_config=cfg)
    rho = np.random.default_rng(5).uniform(0.1, 1.0, t.mesh.nelements)
    k = len(t.dirichlet_values)
    U = np.zeros((t.ndof, k))
    Uo = np.zeros((t.ndof, k))
    J = fem.objectives_multi_load(rho, 3.0, U)
    Jo = ofem.objectives_multi_load(rho, 3.0, Uo)
    tol = 1e-9 if precision == "fp64" else 1e-7
    np.testing.assert_allclose(J, Jo, rtol=tol)
    np.testing.assert_allclose(U, Uo, rtol=tol, atol=tol * np.abs(Uo).max())
    e = fem.energy_multi_load(rho, 3.0, U)
    eo = ofem.energy_multi_load(rho, 3.0, Uo)
    np.testing.assert_allclose(e, eo, rtol=1e-6, atol=1e-8 * np.abs(eo).max())
    s = fem.compliance_sensitivity_multi_load(rho, 3.0, U)
    so = ofem.compliance_sensitivity_multi_load(rho, 3.0, Uo)
    np.testing.assert_allclose(s, so, rtol=1e-6, atol=1e-8 * np.abs(so).max())


@pytest.mark.parametrize("method,objective", [("OC", "compliance"), ("LogMOC", "compliance"),
                                              ("OC", "averaged_temp"),
                                              ("OC", "heat_exchange")])
def test_robin_heat_trajectory_matches_oracle(gpu, method, objective, tmp_path):
    from sktopt_amd import core
    from sktopt_amd.tools import SchedulerConfig as SC
    t = _task(objective=objective)
    om_, ofem = _oracle_fem(t, objective)
    otask = oprob.Task(om_, "heat", 10.0, 0.0, t.dirichlet_nodes, t.dirichlet_dofs, [t.emit],
                       None, dirichlet_values=list(t.dirichlet_values))
    S = odrv.Sched
    kw = dict(lambda_lower=-1e7, lambda_upper=1e7) if method == "LogMOC" else {}
    ocfg = odrv.Config(method=method, max_iters=50, p=S("Constant", 3.0, 3.0),
                       vol_frac=S("Constant", 0.3, 0.3), beta=S("Constant", 1.0, 1.0),
                       filter_radius=S("Constant", 1.5, 1.5), move_limit=S("Constant", 0.2, 0.2),
                       eta=S("Constant", 0.5, 0.5), E_min_coeff=1e-3, **kw)
    loop = odrv.DensityLoop(ocfg, otask, fem=ofem)
    common = dict(dst_path=str(tmp_path), max_iters=50, p=SC.constant(target_value=3.0),
                  vol_frac=SC.constant(target_value=0.3), beta=SC.constant(target_value=1.0),
                  filter_type="spacial", filter_radius=SC.constant(target_value=1.5),
                  move_limit=SC.constant(target_value=0.2), E_min_coeff=1e-3,
                  precision="fp64", solver_rtol=1e-13,
                  eta=SC.constant(target_value=0.5))
    if method == "OC":
        opt = core.OC_Optimizer(core.OC_Config(**common), t)
    else:
        opt = core.LogMOC_Optimizer(core.LogMOC_Config(**common, **kw), t)
    opt.parameterize()
    # the Robin compliance sensitivity is a difference of the conduction and
    # Robin terms (phi = 2 T_env T - T^2 ~ -1e5 here), so solver-level
    # differences (1e-9 in T) reach the densities amplified to ~1e-7
    for it in range(1, 4):
        rec = loop.step(it)
        opt.optimize_steps(1)
        c = opt.recorder.latest(objective)
        assert abs(c - rec["compliance"]) <= 1e-8 * abs(rec["compliance"]), (it, c, rec)
        assert np.abs(opt.rho_host() - loop.rho).max() <= 2e-6, it


# --------------------------------------------------------------- tet4 meshes

def _tet_task(n=(6, 6, 3), objective="compliance", h=4e-2, design_robin=True, w=0.0,
              values=600.0, seed=2):
    """The heat task above on a jittered six-tet split of the same box (box
    faces stay planar: only interior nodes move)."""
    from sktopt_amd.mesh import toy_problem as tp
    from sktopt_amd.mesh.meshes import MeshTet
    from sktopt_amd.mesh.task import LinearHeatConduction
    nx, ny, nz = n
    o = om.jittered_tet(nx, ny, nz, 0.2, seed)
    m = MeshTet(o.p, o.t, cell_grid=(nx, ny, nz))
    d = m.facets_satisfying(tp.get_points_in_range((nx, nx), (0, 2), (0, nz)), True)
    r0 = m.facets_satisfying(tp.get_points_in_range((0, 0), (0, ny), (0, nz)), True)
    r1 = m.facets_satisfying(tp.get_points_in_range((0, nx), (ny, ny), (0, nz)), True)
    return o, LinearHeatConduction.from_facets(m, d, values, [r0, r1], h, 300.0, design_robin,
                                               np.arange(m.nelements), 10.0,
                                               objective=objective, avg_temp_weight=w)


def _tet_oracle_fem(o, t, objective):
    facets = [t.mesh.facets[:, np.asarray(f)] for f in t.robin_facets_ids]
    return hr.HeatRobinFEM(o, t.dirichlet_dofs, t.dirichlet_values, t.k, 1e-3, facets,
                           t.robin_coefficient, t.robin_bc_value, emit=t.emit,
                           objective=objective, avg_temp_weight=t.avg_temp_weight)


def test_tet_robin_kernels_match_oracle(gpu):
    """Nodal pairing (bit-equal), virtual Robin element matrices and loads on
    both tetrahedral rules, the explicit gradient and the heat-exchange
    terms, through the C ABI on the task's scalar tet operator."""
    import torch
    from sktopt_amd import _lib, fea
    from test_host_robin import _elem_matrix_assemble
    o, t = _tet_task((4, 3, 3))
    fem = fea.FEM_SimpLinearHeatConduction(t, 1e-3)
    op = fem._op
    ne, nn = o.nelements, o.nnodes
    rng = np.random.default_rng(3)
    rho = rng.uniform(0.05, 1.0, ne)
    rn = torch.empty(nn, dtype=torch.float64, device="cuda")
    cnt = torch.empty_like(rn)
    _lib.call("skt_tet4_nodal_average_pairing", op, _t(rho).data_ptr(), rn.data_ptr(),
              cnt.data_ptr(), _s())
    rn_ref, cnt_ref = hr.nodal_average_reference(o, rho)
    np.testing.assert_array_equal(rn.cpu().numpy(), rn_ref)
    np.testing.assert_array_equal(cnt.cpu().numpy(), cnt_ref)
    work = torch.empty(4 * ne + nn, dtype=torch.float64, device="cuda")
    for order in (1, 2):
        R = torch.empty((10, ne), dtype=torch.float64, device="cuda")
        L = torch.empty(nn, dtype=torch.float64, device="cuda")
        _lib.call("skt_tet4_robin_virtual", op, order, rn.data_ptr(), 3.0, 4.0, 0.04, 300.0,
                  R.data_ptr(), L.data_ptr(), work.data_ptr(), _s())
        B, Lr = hr.virtual_robin_nodal(o, rn_ref, 3.0, 4.0, 0.04, 300.0, order)
        Bg = _elem_matrix_assemble(t.mesh, R.cpu().numpy()).toarray()
        np.testing.assert_allclose(Bg, B.toarray(), rtol=1e-12, atol=1e-13 * np.abs(Bg).max())
        np.testing.assert_allclose(L.cpu().numpy(), Lr, rtol=1e-12, atol=1e-13 * np.abs(Lr).max())
    T = 300.0 + 100.0 * rng.uniform(size=(nn, 2))
    out = torch.empty((ne, 2), dtype=torch.float64, device="cuda")
    for order in (1, 2):
        _lib.call("skt_tet4_robin_grad", op, order, rn.data_ptr(), cnt.data_ptr(), 2,
                  _t(T).data_ptr(), 3.0, 4.0, 0.04, 300.0, work.data_ptr(), out.data_ptr(), _s())
        for c in range(2):
            ref = hr.robin_explicit_grad(o, rho, T[:, c], 3.0, 4.0, 0.04, 300.0, order)
            np.testing.assert_allclose(out[:, c].cpu().numpy(), ref, rtol=1e-11,
                                       atol=1e-12 * np.abs(ref).max())
    den = torch.empty(ne, dtype=torch.float64, device="cuda")
    num = torch.empty((ne, 2), dtype=torch.float64, device="cuda")
    rhs = torch.empty(nn, dtype=torch.float64, device="cuda")
    _lib.call("skt_tet4_heat_exchange", op, rn.data_ptr(), 2, _t(T).data_ptr(), 3.0, 4.0, 0.04,
              300.0, den.data_ptr(), num.data_ptr(), rhs.data_ptr(), work.data_ptr(), _s())
    d_ref, n_ref, r_ref = hr.heat_exchange_terms(o, rn_ref, T, 3.0, 4.0, 0.04, 300.0)
    np.testing.assert_allclose(den.sum().item(), d_ref, rtol=1e-12)
    np.testing.assert_allclose(-300.0 * 0.04 * num.sum(0).cpu().numpy(), n_ref, rtol=1e-11)
    r = rhs.cpu().numpy()
    np.testing.assert_allclose(r, r_ref, rtol=1e-12, atol=1e-13 * np.abs(r_ref).max())


@pytest.mark.parametrize("objective,precision,w,values",
                         [("compliance", "fp64", 0.0, 600.0),
                          ("averaged_temp", "fp64", 0.0, 600.0),
                          ("compliance", "mixed", 0.0, 600.0),
                          ("heat_exchange", "fp64", 0.0, 600.0),
                          ("heat_exchange", "fp64", 0.3, [600.0, 450.0])])
def test_tet_robin_objectives_match_oracle(gpu, objective, precision, w, values):
    """Objectives, states, energies and sensitivities on a jittered tet mesh
    (the multigrid preconditions the conduction part; the operator carries
    the Robin element matrices)."""
    from sktopt_amd import fea
    o, t = _tet_task(objective=objective, w=w, values=values)
    ofem = _tet_oracle_fem(o, t, objective)
    cfg = fea.LinearSolverConfig(rtol=1e-12 if precision == "fp64" else 1e-10,
                                 precision=precision)
    fem = fea.FEM_SimpLinearHeatConduction(t, 1e-3, solver_config=cfg)
    rho = np.random.default_rng(5).uniform(0.1, 1.0, t.mesh.nelements)
    k = len(t.dirichlet_values)
    U = np.zeros((t.ndof, k))
    Uo = np.zeros((t.ndof, k))
    J = fem.objectives_multi_load(rho, 3.0, U)
    assert fem.last_solve["converged"]
    Jo = ofem.objectives_multi_load(rho, 3.0, Uo)
    tol = 1e-9 if precision == "fp64" else 1e-7
    np.testing.assert_allclose(J, Jo, rtol=tol)
    np.testing.assert_allclose(U, Uo, rtol=tol, atol=tol * np.abs(Uo).max())
    e = fem.energy_multi_load(rho, 3.0, U)
    eo = ofem.energy_multi_load(rho, 3.0, Uo)
    np.testing.assert_allclose(e, eo, rtol=1e-6, atol=1e-8 * np.abs(eo).max())
    s = fem.compliance_sensitivity_multi_load(rho, 3.0, U)
    so = ofem.compliance_sensitivity_multi_load(rho, 3.0, Uo)
    np.testing.assert_allclose(s, so, rtol=1e-6, atol=1e-8 * np.abs(so).max())


@pytest.mark.parametrize("objective", ["compliance", "heat_exchange"])
def test_tet_robin_heat_trajectory_matches_oracle(gpu, objective, tmp_path):
    from sktopt_amd import core
    from sktopt_amd.tools import SchedulerConfig as SC
    o, t = _tet_task(objective=objective)
    ofem = _tet_oracle_fem(o, t, objective)
    otask = oprob.Task(o, "heat", 10.0, 0.0, t.dirichlet_nodes, t.dirichlet_dofs, [t.emit],
                       None, dirichlet_values=list(t.dirichlet_values))
    S = odrv.Sched
    ocfg = odrv.Config(method="OC", max_iters=50, p=S("Constant", 3.0, 3.0),
                       vol_frac=S("Constant", 0.3, 0.3), beta=S("Constant", 1.0, 1.0),
                       filter_radius=S("Constant", 1.5, 1.5), move_limit=S("Constant", 0.2, 0.2),
                       eta=S("Constant", 0.5, 0.5), E_min_coeff=1e-3)
    loop = odrv.DensityLoop(ocfg, otask, fem=ofem)
    cfg = core.OC_Config(dst_path=str(tmp_path), max_iters=50, p=SC.constant(target_value=3.0),
                         vol_frac=SC.constant(target_value=0.3),
                         beta=SC.constant(target_value=1.0), filter_type="spacial",
                         filter_radius=SC.constant(target_value=1.5),
                         move_limit=SC.constant(target_value=0.2), E_min_coeff=1e-3,
                         precision="fp64", solver_rtol=1e-13,
                         eta=SC.constant(target_value=0.5))
    opt = core.OC_Optimizer(cfg, t)
    opt.parameterize()
    for it in range(1, 4):
        rec = loop.step(it)
        opt.optimize_steps(1)
        c = opt.recorder.latest(objective)
        assert abs(c - rec["compliance"]) <= 1e-8 * abs(rec["compliance"]), (it, c, rec)
        assert np.abs(opt.rho_host() - loop.rho).max() <= 2e-6, it
