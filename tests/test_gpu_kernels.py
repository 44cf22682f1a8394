"""GPU parity of the individual HIP entry points against the FP64 oracle."""
import ctypes as C

import numpy as np
import pytest

from oracle import elements as oel
from oracle import fem as ofem
from oracle import filter as ofilt
from oracle import mesh as omesh
from oracle import oc as ooc
from oracle import problems as oprob
from oracle import transforms as otr

pytestmark = pytest.mark.gpu


def _t(x, dtype=None):
    import torch
    dtype = dtype or torch.float64
    return torch.as_tensor(np.ascontiguousarray(x), dtype=dtype, device="cuda")


_KEEP = []


def _p(t):
    """Device pointer of ``t``; the tensor is kept alive until the test ends
    (a temporary would be freed and its memory reused by the next argument
    before the kernel runs)."""
    _KEEP.append(t)
    return t.data_ptr()


@pytest.fixture(autouse=True)
def _release():
    yield
    _KEEP.clear()


def _s():
    import torch
    return torch.cuda.current_stream().cuda_stream


# ---------------------------------------------------------------- transforms

@pytest.mark.parametrize("beta,eta", [(1.0, 0.5), (8.0, 0.5), (3.0, 0.3), (64.0, 0.5)])
def test_heaviside_matches_oracle(gpu, beta, eta):
    from sktopt_amd.core import projection
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-0.5, 1.5, 5000), [0.0, 0.5, 1.0, eta]])
    y = projection.heaviside_projection_inplace(_t(x), beta, eta).cpu().numpy()
    d = projection.heaviside_projection_derivative_inplace(_t(x), beta, eta).cpu().numpy()
    np.testing.assert_allclose(y, otr.heaviside_projection(x, beta, eta), rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(d, otr.heaviside_projection_derivative(x, beta, eta),
                               rtol=1e-13, atol=1e-300)


def test_heaviside_host_arrays_round_trip(gpu):
    from sktopt_amd.core import projection
    x = np.linspace(0, 1, 101)
    out = np.empty_like(x)
    projection.heaviside_projection_inplace(x, 2.0, 0.5, out=out)
    np.testing.assert_allclose(out, otr.heaviside_projection(x, 2.0, 0.5), rtol=1e-13)


def test_heaviside_derivative_overflow_is_zero(gpu):
    from sktopt_amd.core import projection
    x = np.array([-1e3, 1e3, 0.5])
    d = projection.heaviside_projection_derivative_inplace(_t(x), 10.0, 0.5).cpu().numpy()
    assert d[0] == 0.0 and d[1] == 0.0 and d[2] > 0


@pytest.mark.parametrize("kind", ["SIMP", "RAMP"])
def test_interpolation_and_dC(gpu, kind):
    from sktopt_amd.fea import composer
    from sktopt_amd.core import derivatives
    rng = np.random.default_rng(1)
    rho = np.concatenate([rng.uniform(0, 1, 4000), [0.0, 1e-8, 1e-6, 1.0]])
    U = rng.uniform(0, 2, rho.size)
    f = composer.simp_interpolation if kind == "SIMP" else composer.ramp_interpolation
    of = otr.simp_interpolation if kind == "SIMP" else otr.ramp_interpolation
    np.testing.assert_allclose(f(_t(rho), 1.0, 1e-9, 3.0).cpu().numpy(),
                               of(rho, 1.0, 1e-9, 3.0), rtol=1e-14)
    g = derivatives.dC_drho_simp if kind == "SIMP" else derivatives.dC_drho_ramp
    og = otr.dC_drho_simp if kind == "SIMP" else otr.dC_drho_ramp
    np.testing.assert_allclose(g(_t(rho), _t(U), 1.0, 1e-9, 3.0).cpu().numpy(),
                               og(rho, U, 1.0, 1e-9, 3.0), rtol=1e-12)


# ---------------------------------------------------------------- reductions

@pytest.mark.parametrize("n", [1, 2, 7, 1000, 65537])
@pytest.mark.parametrize("q", [0.0, 3.3, 50.0, 95.0, 100.0])
def test_percentile_abs_is_numpy_exact(gpu, n, q):
    from sktopt_amd import _lib
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) * 10.0 ** rng.integers(-5, 5, n)
    if n > 10:
        x[: n // 3] = np.round(x[: n // 3], 1)          # ties
    out = _t(np.zeros(1))
    _lib.call("skt_percentile_abs", n, _p(_t(x)), q, _p(out), _s())
    assert out.item() == np.percentile(np.abs(x), q)


def test_reductions(gpu):
    from sktopt_amd import _lib
    rng = np.random.default_rng(3)
    x, y = rng.standard_normal(100001), rng.standard_normal(100001)
    out = _t(np.zeros(1))
    for op, ref in [(_lib.RED_SUM, x.sum()), (_lib.RED_DOT, x @ y),
                    (_lib.RED_MAXABS, np.abs(x).max()),
                    (_lib.RED_MAXABSDIFF, np.abs(x - y).max())]:
        _lib.call("skt_reduce", op, x.size, _p(_t(x)), _p(_t(y)), _p(out), _s())
        np.testing.assert_allclose(out.item(), ref, rtol=1e-12)
    X, Y = rng.standard_normal((5001, 5)), rng.standard_normal((5001, 5))
    o = _t(np.zeros(5))
    _lib.call("skt_reduce_cols", _lib.RED_DOT, 5001, 5, _p(_t(X)), _p(_t(Y)), _p(o), _s())
    np.testing.assert_allclose(o.cpu().numpy(), np.einsum("ij,ij->j", X, Y), rtol=1e-12)
    _lib.call("skt_reduce_cols", _lib.RED_SUM, 5001, 5, _p(_t(X)), None, _p(o), _s())
    np.testing.assert_allclose(o.cpu().numpy(), X.sum(0), rtol=1e-12)
    _lib.call("skt_reduce_cols", _lib.RED_MAXABS, 5001, 5, _p(_t(X)), None, _p(o), _s())
    np.testing.assert_array_equal(o.cpu().numpy(), np.abs(X).max(0))


@pytest.mark.parametrize("n", [1, 2, 1023, 1024, 1025, 100001, 432000, 1_500_001])
def test_stats4_matches_numpy(gpu, n):
    """[min, mean, max, std] as numpy computes them; the offset case keeps
    a large mean over a small spread (a one-pass sum of squares would lose
    the std there), the constant case has std exactly 0."""
    from sktopt_amd import _lib
    rng = np.random.default_rng(n)
    out = _t(np.zeros(4))
    for x in (rng.standard_normal(n), 1e6 + 1e-3 * rng.uniform(size=n), np.full(n, 0.3)):
        _lib.call("skt_stats4", n, _p(_t(x)), _p(out), _s())
        got = out.cpu().numpy()
        assert got[0] == x.min() and got[2] == x.max()
        np.testing.assert_allclose(got[1], x.mean(), rtol=1e-13)
        np.testing.assert_allclose(got[3], x.std(), rtol=1e-11, atol=1e-14 * abs(x.mean()))
    with pytest.raises(ValueError):
        _lib.call("skt_stats4", 0, _p(_t(np.zeros(1))), _p(out), _s())


def test_reductions_and_stats_propagate_nan(gpu):
    """numpy's max / min / mean / std of an array holding a NaN are all NaN;
    so are the device reductions' (a diverged iteration shows up in every
    statistic of its record)."""
    from sktopt_amd import _lib
    rng = np.random.default_rng(5)
    for n, pos in ((100001, 77777), (5000, 0), (3, 2)):
        x = rng.standard_normal(n)
        x[pos] = np.nan
        out = _t(np.zeros(4))
        _lib.call("skt_stats4", n, _p(_t(x)), _p(out), _s())
        assert np.isnan(out.cpu().numpy()).all(), (n, pos, out)
        o1 = _t(np.zeros(1))
        _lib.call("skt_reduce", _lib.RED_MAXABS, n, _p(_t(x)), None, _p(o1), _s())
        assert np.isnan(o1.item()), (n, pos)


def test_reductions_on_two_streams_use_separate_scratch(gpu):
    """Reductions issued on two streams at once give each stream's own
    result: the partials buffer belongs to the (device, stream) pair."""
    import torch
    from sktopt_amd import _lib
    rng = np.random.default_rng(9)
    n = 2_000_003
    xs = [_t(rng.standard_normal(n)), _t(rng.standard_normal(n) + 3.0)]
    refs = [float(x.sum().item()) for x in xs]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    outs = [_t(np.zeros(64)), _t(np.zeros(64))]
    torch.cuda.synchronize()
    for rep in range(64):
        for i in (0, 1):
            _lib.call("skt_reduce", _lib.RED_SUM, n, _p(xs[i]), None,
                      _p(outs[i][rep:rep + 1]), streams[i].cuda_stream)
    torch.cuda.synchronize()
    for i in (0, 1):
        got = outs[i].cpu().numpy()
        assert (got == got[0]).all(), i
        np.testing.assert_allclose(got[0], refs[i], rtol=1e-12)


# ---------------------------------------------------------------- filter

def _filter_case(kind, n, r):
    if kind == "hex":
        task = oprob.cantilever(*n)
    else:
        task = oprob.cantilever(*n, kind="tet4", jitter=0.2, seed=5)
    mask = task.design_mask.copy()
    mask[task.dirichlet_elements] = False
    return task, mask


@pytest.mark.parametrize("kind,n,r", [("hex", (40, 20, 20), 1.5), ("hex", (24, 12, 12), 2.0),
                                      ("hex", (20, 10, 10), 2.5), ("tet", (8, 4, 4), 1.5)])
def test_filter_connectivity_bit_exact(gpu, kind, n, r):
    from sktopt_amd.filters import SpacialFilter
    task, mask = _filter_case(kind, n, r)
    W = ofilt.build_filter_matrix(task.centers, r, mask)
    f = SpacialFilter(None, task.elements_volume, r, mask,
                      element_centers=task.centers)
    indptr, indices, data, ids = f.csr()
    np.testing.assert_array_equal(ids, np.nonzero(mask)[0])
    np.testing.assert_array_equal(indptr, W.indptr)
    np.testing.assert_array_equal(indices, W.indices)
    np.testing.assert_allclose(data, W.data, rtol=1e-14)
    rng = np.random.default_rng(7)
    x = rng.uniform(0, 1, task.n_elements)
    of = ofilt.SpacialFilter(task.centers, r, mask)
    np.testing.assert_allclose(f.forward(_t(x)).cpu().numpy(), of.forward(x), rtol=1e-13)
    np.testing.assert_allclose(f.gradient(_t(x)).cpu().numpy(), of.gradient(x), rtol=1e-13,
                               atol=1e-15)


# ---------------------------------------------------------------- operators

def _gpu_fem(task, physics, precision="fp64", rtol=1e-12):
    from sktopt_amd import fea, mesh as pmesh
    pm = task.mesh
    if pm.kind == "hex8":
        m = pmesh.structured_hex(*pm.grid, *pm.h)
    else:
        m = pmesh.MeshTet(pm.p, pm.t)
    cfg = fea.LinearSolverConfig(rtol=rtol, precision=precision)
    if physics == "elastic":
        t = pmesh.LinearElasticity(m, task.dirichlet_nodes, task.dirichlet_dofs, task.forces,
                                   task.neumann_nodes, None, task.E0, task.nu)
        return fea.FEM_SimpLinearElasticity(t, 1e-9, solver_config=cfg)
    t = pmesh.LinearHeatConduction(m, task.dirichlet_nodes, task.dirichlet_values,
                                   task.forces[0], None, task.E0)
    return fea.FEM_SimpLinearHeatConduction(t, 1e-3, solver_config=cfg)


CASES = [
    ("hex-elastic", lambda: oprob.cantilever(6, 4, 3), "elastic"),
    ("tet-elastic", lambda: oprob.cantilever(4, 3, 3, kind="tet4", jitter=0.2, seed=1), "elastic"),
    ("hex-heat", lambda: oprob.heat_patch(6, patch=3), "heat"),
]


@pytest.mark.parametrize("name,mk,physics", CASES)
def test_operator_apply_diag_energy(gpu, name, mk, physics):
    import torch
    from sktopt_amd import _lib
    task = mk()
    fem = _gpu_fem(task, physics)
    op = task.element_operator()
    rng = np.random.default_rng(11)
    E = rng.uniform(1e-3, 1.0, task.n_elements)
    K = op.assemble(E)
    Ke, _ = ofem.enforce(K, np.zeros(task.ndof), task.dirichlet_dofs)
    for k in (1, 3, 8):
        X = rng.standard_normal((task.ndof, k))
        Y = torch.empty((task.ndof, k), dtype=torch.float64, device="cuda")
        _lib.call("skt_operator_apply", fem._op, _lib.F64, k, _p(_t(E)), _p(_t(X)), _p(Y), _s())
        ref = Ke @ X
        assert np.abs(Y.cpu().numpy() - ref).max() <= 1e-12 * np.abs(ref).max()
        Y32 = torch.empty((task.ndof, k), dtype=torch.float32, device="cuda")
        _lib.call("skt_operator_apply", fem._op, _lib.F32, k, _p(_t(E, torch.float32)),
                  _p(_t(X, torch.float32)), _p(Y32), _s())
        assert np.abs(Y32.cpu().numpy() - ref).max() <= 2e-5 * np.abs(ref).max()
        en = torch.empty((task.n_elements, k), dtype=torch.float64, device="cuda")
        _lib.call("skt_operator_energy", fem._op, _lib.F64, k, _p(_t(E)), _p(_t(X)), _p(en), _s())
        np.testing.assert_allclose(en.cpu().numpy(), 0.5 * E[:, None] * op.element_quadratic(X),
                                   rtol=1e-11, atol=1e-13)
    d = torch.empty(task.ndof, dtype=torch.float64, device="cuda")
    _lib.call("skt_operator_diag", fem._op, _lib.F64, _p(_t(E)), _p(d), _s())
    np.testing.assert_allclose(d.cpu().numpy(), Ke.diagonal(), rtol=1e-13)


def test_hex8_element_matrix_matches_oracle(gpu):
    from sktopt_amd import _lib
    task = oprob.cantilever(3, 2, 2)
    fem = _gpu_fem(task, "elastic")
    K = (C.c_double * 576)()
    _lib.call("skt_operator_element_matrix", fem._op, K)
    np.testing.assert_allclose(np.array(K).reshape(24, 24), oel.hex8_elasticity_K0(0.3),
                               rtol=1e-13, atol=1e-15)


# ---------------------------------------------------------------- solver

@pytest.mark.parametrize("precision,rtol,ctol", [("fp64", 1e-12, 1e-9), ("mixed", 1e-10, 1e-8)])
@pytest.mark.parametrize("name,mk,physics", CASES)
def test_solver_matches_direct_solve(gpu, precision, rtol, ctol, name, mk, physics):
    task = mk()
    fem = _gpu_fem(task, physics, precision, rtol)
    rng = np.random.default_rng(2)
    rho = rng.uniform(0.0, 1.0, task.n_elements)
    p = 3.0
    if physics == "elastic":
        ofe = ofem.ElasticFEM(task, 1e-9)
    else:
        ofe = ofem.HeatFEM(task, 1e-3)
    Uo = np.zeros((task.ndof, task.n_tasks))
    co = ofe.objectives_multi_load(rho, p, Uo)
    Ug = np.zeros((task.ndof, task.n_tasks))
    cg = fem.objectives_multi_load(rho, p, Ug)
    np.testing.assert_allclose(cg, co, rtol=ctol)
    assert np.abs(Ug - Uo).max() <= 1e-6 * np.abs(Uo).max()
    eo = ofe.energy_multi_load(rho, p, Uo)
    eg = fem.energy_multi_load(rho, p, Ug)
    assert np.abs(eg - eo).max() <= 1e-6 * np.abs(eo).max()


def test_solver_reports_non_convergence(gpu):
    from sktopt_amd import _lib, fea
    task = oprob.cantilever(6, 3, 3)
    fem = _gpu_fem(task, "elastic", "fp64", 1e-12)
    fem.solver_config = fea.LinearSolverConfig(rtol=1e-12, maxiter=1, precision="fp64")
    with pytest.raises(_lib.ConvergenceError):
        fem.objectives_multi_load(np.full(task.n_elements, 0.5), 3.0,
                                  np.zeros((task.ndof, 1)))


# ---------------------------------------------------------------- OC bisection

@pytest.mark.parametrize("it_seed,kind", [(0, "hex8"), (1, "hex8"), (2, "hex8"), (3, "tet4")])
def test_oc_bisection_matches_oracle(gpu, it_seed, kind):
    """The device bisection on the CSR form (three tree levels per read of
    the matrix) against the oracle's loop: evaluation count, multiplier,
    volume error and densities."""
    import torch
    from sktopt_amd import _lib
    from sktopt_amd.filters import SpacialFilter
    task = oprob.cantilever(12, 6, 6) if kind == "hex8" else \
        oprob.cantilever(8, 4, 4, kind="tet4", jitter=0.2, seed=it_seed)
    mask = task.design_mask.copy()
    mask[task.dirichlet_elements] = False
    design = np.nonzero(mask)[0]
    rng = np.random.default_rng(it_seed)
    rho = np.ones(task.n_elements)
    rho[design] = rng.uniform(0.05, 0.9, design.size)
    dC = -rng.uniform(0.0, 1.0, design.size) ** 3
    vol = task.elements_volume[design]
    of = ofilt.SpacialFilter(task.centers, 1.5, mask)
    ref = ooc.bisection_with_physical_volume(
        dC, rho[design], rho, design, of, 1e-2, 1.0, 0.2, 0.5, 1e-12, 0.3, 2.0, 0.5,
        vol, vol.sum(), 0.7, 1.3, max_iter=1000, tolerance=1e-5, l1=1e-7, l2=1e7)
    f = SpacialFilter(None, task.elements_volume, 1.5, mask, element_centers=task.centers)
    out = torch.empty(design.size, dtype=torch.float64, device="cuda")
    res = torch.empty(4, dtype=torch.float64, device="cuda")
    _lib.call("skt_oc_bisect", f.handle, _p(_t(rho)), _p(_t(rho[design])), _p(_t(dC)),
              _p(_t(vol)), float(vol.sum()), 0.3, 0.2, 0.5, 1e-2, 1.0, 0.7, 1.3, 2.0, 0.5,
              1e-7, 1e7, 1e-5, 1e-4, 1000, _p(out), _p(res), _s())
    r = res.cpu().numpy()
    assert int(r[2]) == ref[3]
    np.testing.assert_allclose(r[0], ref[0], rtol=1e-12)
    np.testing.assert_allclose(r[1], ref[1], rtol=1e-8, atol=1e-12)
    np.testing.assert_allclose(out.cpu().numpy(), ref[2], rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("grid,physics,ks", [((31, 17, 9), "elastic", (1, 2, 3, 4, 8)),
                                              ((16, 16, 16), "elastic", (8,)),
                                              ((30, 15, 7), "elastic", (1, 2, 8)),
                                              ((15, 45, 4), "elastic", (2,)),
                                              ((33, 20, 7), "heat", (1, 2, 4, 5)),
                                              ((30, 16, 6), "heat", (1, 4))])
def test_hex8_tiled_apply_matches_assembled(gpu, grid, physics, ks):
    """Multi-tile / multi-z-chunk grids for the modal kernel (tiles of 15x15
    nodes; 16 along an axis whose node count is 1 mod 15, e.g. 31 and 16
    nodes), every column blocking, float64 and float32."""
    import torch
    from sktopt_amd import _lib
    task = oprob.cantilever(*grid) if physics == "elastic" else oprob.heat_patch(grid[0], 3)
    if physics == "heat":
        from oracle import mesh as om2
        task.mesh = om2.structured_hex(*grid)
        task.dirichlet_dofs = np.arange(0, (grid[0] + 1) * (grid[1] + 1), 7)
        task.dirichlet_nodes = task.dirichlet_dofs
        task.forces = [np.ones(task.mesh.nnodes)]
    fem = _gpu_fem(task, physics)
    op = task.element_operator()
    n_el = task.mesh.nelements
    rng = np.random.default_rng(5)
    E = rng.uniform(1e-9, 1.0, n_el)
    K = op.assemble(E)
    Ke, _ = ofem.enforce(K, np.zeros(K.shape[0]), task.dirichlet_dofs)
    for k in ks:
        X = rng.standard_normal((K.shape[0], k))
        ref = Ke @ X
        Y = torch.empty((K.shape[0], k), dtype=torch.float64, device="cuda")
        _lib.call("skt_operator_apply", fem._op, _lib.F64, k, _p(_t(E)), _p(_t(X)), _p(Y), _s())
        assert np.abs(Y.cpu().numpy() - ref).max() <= 1e-12 * np.abs(ref).max(), k
        Y32 = torch.empty((K.shape[0], k), dtype=torch.float32, device="cuda")
        _lib.call("skt_operator_apply", fem._op, _lib.F32, k, _p(_t(E, torch.float32)),
                  _p(_t(X, torch.float32)), _p(Y32), _s())
        assert np.abs(Y32.cpu().numpy() - ref).max() <= 2e-5 * np.abs(ref).max(), k


# ------------------------------------------------- filter, structured stencil

@pytest.mark.parametrize("n,r,h", [((40, 20, 20), 1.5, (1.0, 1.0, 1.0)),
                                   ((24, 12, 12), 2.0, (1.0, 1.0, 1.0)),
                                   ((20, 10, 10), 2.5, (1.0, 1.0, 1.0)),
                                   ((16, 8, 8), 1.0, (0.5, 0.5, 0.5))])
def test_filter_stencil_form_matches_csr_and_oracle(gpu, n, r, h):
    """On structured grids the filter switches to its stencil form only
    after an entry-by-entry check against the CSR; forward, gradient and the
    device bisection then agree with the CSR form and the oracle (ties at
    d = r included: r = 2.0 and 1.0 at h = 0.5)."""
    import torch
    from oracle import mesh as om2
    from sktopt_amd import _lib
    from sktopt_amd.filters import SpacialFilter
    from sktopt_amd.mesh import structured_hex
    task = oprob.cantilever(*n)
    task.mesh = om2.structured_hex(*n, *h)
    task.centers = om2.element_centers(task.mesh)
    mask = task.design_mask.copy()
    mask[task.dirichlet_elements] = False
    m = structured_hex(*n, *h)
    fs = SpacialFilter(m, task.elements_volume, r, mask)
    fc = SpacialFilter(m, task.elements_volume, r, mask)
    fc.use_stencil = False
    fs.handle, fc.handle
    assert fs.stencil_active and not fc.stencil_active
    of = ofilt.SpacialFilter(task.centers, r, mask)
    rng = np.random.default_rng(11)
    x = rng.uniform(0, 1, task.n_elements)
    a, b = fs.forward(_t(x)).cpu().numpy(), fc.forward(_t(x)).cpu().numpy()
    np.testing.assert_allclose(a, b, rtol=1e-14)
    np.testing.assert_allclose(a, of.forward(x), rtol=1e-13)
    g = rng.standard_normal(task.n_elements)
    a, b = fs.gradient(_t(g)).cpu().numpy(), fc.gradient(_t(g)).cpu().numpy()
    np.testing.assert_allclose(a, b, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(a, of.gradient(g), rtol=1e-12, atol=1e-14)
    # the device bisection on the stencil form
    design = np.nonzero(mask)[0]
    rho = np.ones(task.n_elements)
    rho[design] = rng.uniform(0.05, 0.9, design.size)
    dC = -rng.uniform(0.0, 1.0, design.size) ** 3
    vol = task.elements_volume[design]
    ref = ooc.bisection_with_physical_volume(
        dC, rho[design], rho, design, of, 1e-2, 1.0, 0.2, 0.5, 1e-12, 0.3, 2.0, 0.5,
        vol, vol.sum(), 0.7, 1.3, max_iter=1000, tolerance=1e-5, l1=1e-7, l2=1e7)
    out = torch.empty(design.size, dtype=torch.float64, device="cuda")
    res = torch.empty(4, dtype=torch.float64, device="cuda")
    _lib.call("skt_oc_bisect", fs.handle, _p(_t(rho)), _p(_t(rho[design])), _p(_t(dC)),
              _p(_t(vol)), float(vol.sum()), 0.3, 0.2, 0.5, 1e-2, 1.0, 0.7, 1.3, 2.0, 0.5,
              1e-7, 1e7, 1e-5, 1e-4, 1000, _p(out), _p(res), _s())
    rr = res.cpu().numpy()
    assert int(rr[2]) == ref[3]
    np.testing.assert_allclose(rr[0], ref[0], rtol=1e-12)
    np.testing.assert_allclose(out.cpu().numpy(), ref[2], rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("kind,seed", [("hex8", 0), ("tet4", 1), ("tet4", 2)])
def test_oc_bisection_tree_equals_one_midpoint_at_a_time(gpu, kind, seed, monkeypatch):
    """The CSR bisection evaluated three tree levels per pass
    (k_oc_volume_tree) and one midpoint at a time (SKT_OC_TREE=0) take the
    same decisions: evaluation count, multiplier, volume error, bracket and
    densities bit for bit, over several volume targets (terminations at
    every level of a pass)."""
    import torch
    from sktopt_amd import _lib
    from sktopt_amd.filters import SpacialFilter
    task = oprob.cantilever(14, 7, 7) if kind == "hex8" else \
        oprob.cantilever(10, 5, 5, kind="tet4", jitter=0.2, seed=seed)
    mask = task.design_mask.copy()
    mask[task.dirichlet_elements] = False
    design = np.nonzero(mask)[0]
    rng = np.random.default_rng(seed)
    rho = np.ones(task.n_elements)
    rho[design] = rng.uniform(0.05, 0.9, design.size)
    dC = -rng.uniform(0.0, 1.0, design.size) ** 3
    vol = task.elements_volume[design]
    f = SpacialFilter(None, task.elements_volume, 1.5, mask, element_centers=task.centers)
    counts = set()
    for vf, tol in ((0.3, 1e-5), (0.45, 1e-5), (0.2, 1e-3), (0.3, 1e-9), (0.25, 1e-5),
                    (0.35, 1e-4), (0.4, 1e-6), (0.5, 1e-7), (0.15, 1e-2)):
        out = {}
        for tree in ("1", "0"):
            monkeypatch.setenv("SKT_OC_TREE", tree)
            o = torch.empty(design.size, dtype=torch.float64, device="cuda")
            res = torch.empty(4, dtype=torch.float64, device="cuda")
            _lib.call("skt_oc_bisect", f.handle, _p(_t(rho)), _p(_t(rho[design])), _p(_t(dC)),
                      _p(_t(vol)), float(vol.sum()), vf, 0.2, 0.5, 1e-2, 1.0, 0.7, 1.3, 2.0,
                      0.5, 1e-7, 1e7, tol, 1e-4, 1000, _p(o), _p(res), _s())
            out[tree] = (res.cpu().numpy(), o.cpu().numpy())
        np.testing.assert_array_equal(out["1"][0], out["0"][0], err_msg=f"vf={vf} tol={tol}")
        np.testing.assert_array_equal(out["1"][1], out["0"][1], err_msg=f"vf={vf} tol={tol}")
        counts.add(int(out["1"][0][2]) % 3)
    print(f"[oc tree {kind}] evaluation counts mod 3: {sorted(counts)}", flush=True)
    assert counts == {0, 1, 2}


def test_filter_stencil_form_declines_other_grids(gpu):
    """A grid whose centre differences are not the stencil's offsets bit for
    bit (h = 8/27) keeps the CSR form; results are the CSR's either way."""
    from sktopt_amd.filters import SpacialFilter
    from sktopt_amd.mesh import toy_problem as tp
    t = tp.toy_base(0.3)
    f = SpacialFilter(t.mesh, t.elements_volume, 0.9, t.design_mask)
    f.handle
    x = np.random.default_rng(0).uniform(size=t.mesh.nelements)
    of = ofilt.SpacialFilter(t.mesh.element_centers(), 0.9, t.design_mask)
    np.testing.assert_allclose(f.forward(_t(x)).cpu().numpy(), of.forward(x), rtol=1e-13)


@pytest.mark.parametrize("physics", ["elastic", "heat"])
def test_tet_grid_apply_matches_assembled(gpu, physics, monkeypatch):
    """Tet4 meshes on a grid of cells apply through per-cell corner forces
    and a node gather over the 8 cells of a node (no per-tet staging); the
    result equals the assembled matrix, and the per-tet staging form
    (SKT_TET_APPLY=slots) to rounding."""
    import torch
    from sktopt_amd import _lib, fea, mesh as pmesh
    if physics == "elastic":
        task = oprob.cantilever(6, 4, 3, kind="tet4", jitter=0.2, seed=1)
    else:
        task = oprob.heat_patch(6, patch=3, kind="tet4", jitter=0.2, seed=1)
    pm = task.mesh
    m = pmesh.MeshTet(pm.p, pm.t, cell_grid=pm.grid)
    cfg = fea.LinearSolverConfig(rtol=1e-12, precision="fp64")
    if physics == "elastic":
        t = pmesh.LinearElasticity(m, task.dirichlet_nodes, task.dirichlet_dofs, task.forces,
                                   task.neumann_nodes, None, task.E0, task.nu)
        fem = fea.FEM_SimpLinearElasticity(t, 1e-9, solver_config=cfg)
    else:
        t = pmesh.LinearHeatConduction(m, task.dirichlet_nodes, task.dirichlet_values,
                                       task.forces[0], None, task.E0)
        fem = fea.FEM_SimpLinearHeatConduction(t, 1e-3, solver_config=cfg)
    op = task.element_operator()
    rng = np.random.default_rng(13)
    E = rng.uniform(1e-3, 1.0, task.n_elements)
    K = op.assemble(E)
    Ke, _ = ofem.enforce(K, np.zeros(task.ndof), task.dirichlet_dofs)
    for k in (1, 3, 8):
        X = rng.standard_normal((task.ndof, k))
        ref = Ke @ X
        out = {}
        for mode in ("grid", "slots"):
            monkeypatch.setenv("SKT_TET_APPLY", mode)
            Y = torch.empty((task.ndof, k), dtype=torch.float64, device="cuda")
            _lib.call("skt_operator_apply", fem._op, _lib.F64, k, _p(_t(E)), _p(_t(X)), _p(Y),
                      _s())
            Y32 = torch.empty((task.ndof, k), dtype=torch.float32, device="cuda")
            _lib.call("skt_operator_apply", fem._op, _lib.F32, k, _p(_t(E, torch.float32)),
                      _p(_t(X, torch.float32)), _p(Y32), _s())
            out[mode] = (Y.cpu().numpy(), Y32.cpu().numpy().astype(np.float64))
            assert np.abs(out[mode][0] - ref).max() <= 1e-12 * np.abs(ref).max(), (mode, k)
            assert np.abs(out[mode][1] - ref).max() <= 2e-5 * np.abs(ref).max(), (mode, k)
        assert np.abs(out["grid"][0] - out["slots"][0]).max() <= 1e-13 * np.abs(ref).max()
