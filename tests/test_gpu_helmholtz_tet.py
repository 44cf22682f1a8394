"""Nodal Helmholtz filter on tet4 meshes (``ElementTetP1``) on the GPU
against the oracle (``filters/helmholtz_filter_nodal.py`` restated with
scipy): the scalar a M + b K operator against the assembled matrices, the
element <-> node transfers bit for bit, forward and gradient, and OC
trajectories with the filter on a jittered tet cantilever."""
import numpy as np
import pytest

from oracle import driver as odrv
from oracle import helmholtz as oh
from oracle import problems as oprob

pytestmark = pytest.mark.gpu


def _task(n=(8, 4, 4), seed=3):
    return oprob.cantilever(*n, kind="tet4", jitter=0.2, seed=seed)


def _mask(task):
    m = np.zeros(task.n_elements, dtype=bool)
    m[task.design_elements] = True
    return m


def _gpu_mesh(task):
    from sktopt_amd.mesh import MeshTet
    return MeshTet(task.mesh.p, task.mesh.t)


def _scalar_op(task, a, b, fixed):
    import ctypes as C
    import torch
    from sktopt_amd import _lib
    h = C.c_void_p()
    p = torch.tensor(task.mesh.p, device="cuda")
    t = torch.tensor(task.mesh.t, dtype=torch.int32, device="cuda")
    f = torch.tensor(fixed.astype(np.uint8), device="cuda")
    _lib.call("skt_tet4_create_scalar", task.mesh.nnodes, task.n_elements, p.data_ptr(),
              t.data_ptr(), a, b, f.data_ptr(), C.byref(h))
    return h


@pytest.mark.parametrize("a,b", [(1.0, 2.25), (1.0, 0.0), (0.0, 1.0), (0.5, 3.0)])
def test_tet_scalar_operator_matches_assembled(gpu, a, b):
    import torch
    from sktopt_amd import _lib
    task = _task()
    n = task.mesh.nnodes
    rng = np.random.default_rng(1)
    fixed = np.zeros(n, dtype=bool)
    fixed[rng.choice(n, 7, replace=False)] = True
    A, M = oh.helmholtz_matrices(task.mesh, 1.0)
    K = A - M
    ref_mat = (a * M + b * K).tocsr()
    x = rng.standard_normal((n, 2))
    ref = ref_mat @ x
    ref[fixed] = x[fixed]
    h = _scalar_op(task, a, b, fixed)
    try:
        s = torch.cuda.current_stream().cuda_stream
        ones = torch.ones(task.n_elements, dtype=torch.float64, device="cuda")
        X = torch.tensor(x, device="cuda")
        Y = torch.empty_like(X)
        _lib.call("skt_operator_apply", h, _lib.F64, 2, ones.data_ptr(), X.data_ptr(),
                  Y.data_ptr(), s)
        np.testing.assert_allclose(Y.cpu().numpy(), ref, rtol=1e-12,
                                   atol=1e-13 * np.abs(ref).max())
        d = torch.empty(n, dtype=torch.float64, device="cuda")
        _lib.call("skt_operator_diag", h, _lib.F64, ones.data_ptr(), d.data_ptr(), s)
        dref = ref_mat.diagonal().copy()
        dref[fixed] = 1.0
        np.testing.assert_allclose(d.cpu().numpy(), dref, rtol=1e-13)
    finally:
        _lib.load().skt_operator_destroy(h)


def test_tet_element_node_transfers_are_bitwise(gpu):
    import torch
    from sktopt_amd import _lib
    task = _task((7, 5, 3), seed=11)
    mask = _mask(task)
    h = _scalar_op(task, 1.0, 0.0, np.zeros(task.mesh.nnodes, dtype=bool))
    try:
        s = torch.cuda.current_stream().cuda_stream
        rng = np.random.default_rng(0)
        k = 3
        R = rng.uniform(0, 1, (task.n_elements, k))
        vol = torch.tensor(task.elements_volume, device="cuda")
        des = torch.tensor(mask.astype(np.uint8), device="cuda")
        x = torch.tensor(R, device="cuda")
        rn = torch.empty((task.mesh.nnodes, k), dtype=torch.float64, device="cuda")
        for fixed in (1.0, 0.0):
            _lib.call("skt_tet4_elem_to_node", h, k, vol.data_ptr(), des.data_ptr(), fixed,
                      x.data_ptr(), rn.data_ptr(), s)
            for j in range(k):
                ref = oh.element_to_node_density_averaging(task.mesh, task.elements_volume,
                                                           R[:, j], mask, True, fixed)
                np.testing.assert_array_equal(rn[:, j].cpu().numpy(), ref)
        V = rng.standard_normal((task.mesh.nnodes, k))
        vn = torch.tensor(V, device="cuda")
        out = torch.empty((task.n_elements, k), dtype=torch.float64, device="cuda")
        for clamp in (0, 1):
            _lib.call("skt_tet4_node_to_elem", h, k, vn.data_ptr(), clamp, 0, out.data_ptr(), s)
            for j in range(k):
                ref = oh.node_to_element_density(task.mesh, V[:, j])
                if clamp:
                    ref = np.minimum(ref, 0.0)
                np.testing.assert_array_equal(out[:, j].cpu().numpy(), ref)
        # folded mean over the columns: acc += g_k from 0.0, then / k
        m = torch.empty(task.n_elements, dtype=torch.float64, device="cuda")
        _lib.call("skt_tet4_node_to_elem", h, k, vn.data_ptr(), 1, 1, m.data_ptr(), s)
        acc = np.zeros(task.n_elements)
        for j in range(k):
            acc += np.minimum(oh.node_to_element_density(task.mesh, V[:, j]), 0.0)
        np.testing.assert_array_equal(m.cpu().numpy(), acc / k)
    finally:
        _lib.load().skt_operator_destroy(h)


@pytest.mark.parametrize("r", [0.6, 1.5])
def test_tet_helmholtz_forward_gradient_match_oracle(gpu, r):
    import torch
    from sktopt_amd.filters import HelmholtzFilterNodal
    task = _task((10, 5, 5), seed=4)
    mask = _mask(task)
    f = HelmholtzFilterNodal(_gpu_mesh(task), task.elements_volume, r, mask)
    of = oh.HelmholtzFilterNodal(task.mesh, task.elements_volume, r, mask)
    rng = np.random.default_rng(2)
    rho = np.where(mask, rng.uniform(0.05, 1.0, task.n_elements), 1.0)
    got = f.forward(torch.tensor(rho, device="cuda")).cpu().numpy()
    np.testing.assert_allclose(got, of.forward(rho), rtol=1e-10, atol=1e-12)
    v = -rng.uniform(0, 1, task.n_elements)
    got = f.gradient(torch.tensor(v, device="cuda")).cpu().numpy()
    ref = of.gradient(v)
    assert (got <= 0.0).all()
    np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())
    assert f.last_solve["iters"] < 200


def test_tet_helmholtz_oc_trajectory_matches_oracle(gpu, tmp_path):
    """OC with the nodal Helmholtz filter (the reference's default filter
    type) on a jittered tet cantilever, 3 iterations, float64 solves."""
    from test_gpu_driver import _gpu_opt, _oracle_cfg
    mk = lambda: _task((8, 4, 4), seed=3)  # noqa: E731
    loop = odrv.DensityLoop(_oracle_cfg(r=1.5, filter_type="helmholtz"), mk())
    opt = _gpu_opt(mk(), r=1.5, precision="fp64", rtol=1e-13, tmp=str(tmp_path),
                   filter_type="helmholtz")
    for it in range(1, 4):
        rec = loop.step(it)
        opt.optimize_steps(1)
        c = opt.recorder.latest("compliance")
        assert abs(c - rec["compliance"]) <= 1e-9 * abs(rec["compliance"]), (it, c, rec)
        assert opt.recorder.latest("bisection_evals") == rec["n_evals"]
        assert np.abs(opt.rho_host() - loop.rho).max() <= 1e-8, it
