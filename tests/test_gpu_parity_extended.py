"""GPU parity beyond the small trajectory cases:

* filter connectivity at the configurations' own sizes (120x60x60 hex at
  r = 2.0, where the axis neighbours at distance exactly 2 sit on the
  support; 100x50x50x6 jittered tets at r = 1.5), integer arrays equal to
  cKDTree's (``filters/spacial.py:51-103``);
* the OC driver with the reference's default continuation schedules
  (p Step 1 -> 3, beta StepAccelerating 1 -> 2, move limit SawtoothDecay
  0.3 -> 0.1), RAMP, a float percentile scale and the sensitivity filter
  (``core/optimizers/common_density.py:207-389,1102-1106``,
  ``oc.py:187-189``);
* LogMOC history records belong to the iteration that produced them
  (``logmoc.py:238-245``), whether the run is one ``optimize()`` call or
  one ``optimize_steps(1)`` call per iteration.
"""
import numpy as np
import pytest
from scipy.spatial import cKDTree

from oracle import driver as odrv
from oracle import filter as ofilt
from oracle import problems as oprob

pytestmark = pytest.mark.gpu


def _design_case(task):
    mask = task.design_mask.copy()
    mask[task.dirichlet_elements] = False
    return mask


def _ckdtree_rows(centers, radius, mask, row_lo, row_hi, design_ids, tree):
    """Rows [row_lo, row_hi) of the filter's connectivity as cKDTree gives it:
    (row lengths, column indices in the compressed design numbering, sorted
    ascending in each row as ``searchsorted`` on sorted ids leaves them)."""
    _, support = ofilt.gaussian_support(radius)
    rows = design_ids[row_lo:row_hi]
    lists = tree.query_ball_point(centers[:, rows].T, support, workers=-1,
                                  return_sorted=True)
    lens = np.fromiter((len(x) for x in lists), dtype=np.int64, count=len(lists))
    flat = np.fromiter((j for x in lists for j in x), dtype=np.int64, count=int(lens.sum()))
    keep = mask[flat]
    row_of = np.repeat(np.arange(len(rows)), lens)
    counts = np.bincount(row_of[keep], minlength=len(rows))
    cols = np.searchsorted(design_ids, flat[keep])
    return counts, cols


@pytest.mark.parametrize("case", ["hex8-120x60x60-r2.0", "tet4-100x50x50-r1.5"])
def test_filter_connectivity_bit_exact_at_config_size(gpu, case):
    import torch
    from sktopt_amd.filters import SpacialFilter
    if case.startswith("hex8"):
        task, r = oprob.cantilever(120, 60, 60), 2.0
    else:
        task, r = oprob.cantilever(100, 50, 50, kind="tet4", jitter=0.2, seed=0), 1.5
    mask = _design_case(task)
    print(f"[{case}] n_el={task.n_elements} design={int(mask.sum())}", flush=True)
    f = SpacialFilter(None, task.elements_volume, r, mask, element_centers=task.centers)
    indptr, indices, data, ids = f.csr()
    del data
    torch.cuda.empty_cache()
    design_ids = np.nonzero(mask)[0]
    np.testing.assert_array_equal(ids, design_ids)
    tree = cKDTree(task.centers.T)
    chunk = 100_000
    for a in range(0, len(design_ids), chunk):
        b = min(a + chunk, len(design_ids))
        counts, cols = _ckdtree_rows(task.centers, r, mask, a, b, design_ids, tree)
        np.testing.assert_array_equal(np.diff(indptr[a:b + 1]), counts, err_msg=f"rows {a}:{b}")
        np.testing.assert_array_equal(indices[indptr[a]:indptr[b]], cols,
                                      err_msg=f"rows {a}:{b}")
    assert indptr[-1] == len(indices)
    print(f"[{case}] nnz={len(indices)} equal to cKDTree", flush=True)


def _default_cfgs(interp, percentile, sens_filter, iters, tmp):
    """The reference's default schedules, with vol 0.3, r = 1.5 and the
    chosen interpolation / percentile / sensitivity filter."""
    from sktopt_amd import core
    from sktopt_amd.tools import SchedulerConfig as SC
    S = odrv.Sched
    ocfg = odrv.Config(method="OC", interpolation=interp, max_iters=iters,
                       vol_frac=S("Constant", 0.3, 0.3), filter_radius=S("Constant", 1.5, 1.5),
                       percentile=S("Constant", percentile, percentile) if percentile
                       else S("None"), sensitivity_filter=sens_filter, solver="spsolve")
    gcfg = core.OC_Config(dst_path=tmp, max_iters=iters, interpolation=interp,
                          vol_frac=SC.constant(target_value=0.3), filter_type="spacial",
                          filter_radius=SC.constant(target_value=1.5),
                          percentile=SC.constant(target_value=percentile) if percentile
                          else SC.none(),
                          sensitivity_filter=sens_filter, precision="fp64", solver_rtol=1e-13)
    return ocfg, gcfg


@pytest.mark.parametrize("interp,percentile,sens_filter", [
    ("SIMP", None, False), ("RAMP", None, False), ("SIMP", 90.0, False),
    ("SIMP", None, True), ("RAMP", 75.0, True)])
def test_reference_default_schedules_match_oracle(gpu, interp, percentile, sens_filter,
                                                  tmp_path):
    """12 iterations: p steps 1, 2, 3 at iterations 1, 5, 9, beta climbs to
    2 and the move limit follows the sawtooth; every iteration is compared.

    The reference's RAMP derivative (``core/derivatives.py:56-61``) changes
    sign above rho = (1 + p) / (2 p), which makes the OC rate
    ``(-dC / lmid)^eta`` NaN (``oc.py:51-57``, computed under
    ``np.errstate(invalid="ignore")``): from then on the reference's densities
    hold NaN.  Both sides must produce NaN in the same elements; the run
    stops at that iteration, since the next solve would take NaN moduli."""
    from sktopt_amd import core
    from sktopt_amd.mesh import toy_problem
    iters = 12
    ocfg, gcfg = _default_cfgs(interp, percentile, sens_filter, iters, str(tmp_path))
    loop = odrv.DensityLoop(ocfg, oprob.cantilever(16, 8, 8))
    opt = core.OC_Optimizer(gcfg, toy_problem.cantilever(16, 8, 8))
    opt.parameterize()
    seen_p = set()
    for it in range(1, iters + 1):
        rec = loop.step(it)
        opt.optimize_steps(1)
        seen_p.add(rec["p"])
        c = opt.recorder.latest("compliance")
        assert abs(c - rec["compliance"]) <= 1e-9 * abs(rec["compliance"]), (it, c, rec)
        assert opt.recorder.latest("bisection_evals") == rec["n_evals"], it
        np.testing.assert_allclose(opt.recorder.latest("lmid"), rec["lmid"], rtol=1e-6)
        np.testing.assert_allclose(opt.recorder.latest("running_scale"), rec["running_scale"],
                                   rtol=1e-9)
        rho_g = opt.rho_host()
        nan_o = np.isnan(loop.rho)
        np.testing.assert_array_equal(np.isnan(rho_g), nan_o, err_msg=f"NaN pattern, it {it}")
        assert np.abs(rho_g[~nan_o] - loop.rho[~nan_o]).max() <= 1e-8, it
        if nan_o.any():
            assert interp == "RAMP", it
            break
    assert seen_p == ({1.0, 2.0, 3.0} if interp == "SIMP" else seen_p | {1.0, 2.0}), seen_p


def test_logmoc_records_belong_to_their_iteration(gpu, tmp_path):
    """One optimize() call, four optimize_steps(1) calls and the oracle give
    the same lambda_v / vol_error / constraint_coeff histories."""
    from test_gpu_driver import _gpu_opt, _oracle_cfg
    n = 4
    mk = dict(method="LogMOC", vol=0.3, r=2.0, emin=1e-3, iters=n, precision="fp64", rtol=1e-13)
    a = _gpu_opt(oprob.heat_patch(10, patch=4, source=1e-2), tmp=str(tmp_path / "a"), **mk)
    a.optimize()
    b = _gpu_opt(oprob.heat_patch(10, patch=4, source=1e-2), tmp=str(tmp_path / "b"), **mk)
    for _ in range(n):
        b.optimize_steps(1)
    cfg = _oracle_cfg(method="LogMOC", vol=0.3, r=2.0, emin=1e-3, iters=n,
                      eta=odrv.Sched("Constant", 0.6, 0.6), lambda_lower=-1e7, lambda_upper=1e7)
    loop = odrv.DensityLoop(cfg, oprob.heat_patch(10, patch=4, source=1e-2))
    recs = [loop.step(it) for it in range(1, n + 1)]
    ha, hb = a.recorder.as_object(), b.recorder.as_object()
    for name, oname in (("lambda_v", "lambda_v"), ("vol_error", "vol_error"),
                        ("constraint_coeff", "coeff")):
        va, vb = np.array(getattr(ha, name)), np.array(getattr(hb, name))
        assert va.shape == (n,), name
        np.testing.assert_array_equal(va, vb, err_msg=name)
        np.testing.assert_allclose(va, [r[oname] for r in recs], rtol=1e-7, atol=1e-12,
                                   err_msg=name)


def test_scaling_option_matches_scaled_oracle_problem(gpu, tmp_path):
    """``scaling=True`` (``core/optimizers/common_density.py:539-540,626-641``;
    ``mesh/task_common.py:403-420``): lengths divided by the largest
    bounding-box side (16 here), loads by 1e5, filter radius read in the
    scaled lengths; the oracle runs the same scaled problem directly.  The
    task is restored when the run finishes."""
    from oracle import mesh as omesh
    from sktopt_amd import core
    from sktopt_amd.mesh import toy_problem
    from sktopt_amd.tools import SchedulerConfig as SC
    n = 4
    base = oprob.cantilever(16, 8, 8)
    L = 16.0
    ms = omesh.structured_hex(16, 8, 8, 1.0 / L, 1.0 / L, 1.0 / L)
    scaled = oprob.Task(ms, "elastic", 1.0, 0.3, base.dirichlet_nodes, base.dirichlet_dofs,
                        [base.forces[0] * (1.0 / 1e5)], base.neumann_nodes)
    S = odrv.Sched
    ocfg = odrv.Config(max_iters=n, p=S("Constant", 3.0, 3.0), vol_frac=S("Constant", 0.3, 0.3),
                       beta=S("Constant", 1.0, 1.0), filter_radius=S("Constant", 0.1, 0.1),
                       move_limit=S("Constant", 0.2, 0.2), E_min_coeff=1e-9, solver="spsolve")
    loop = odrv.DensityLoop(ocfg, scaled)
    task = toy_problem.cantilever(16, 8, 8)
    p0 = task.mesh.p.copy()
    f0 = task.neumann_linear[0].copy()
    cfg = core.OC_Config(dst_path=str(tmp_path), max_iters=n, p=SC.constant(target_value=3.0),
                         vol_frac=SC.constant(target_value=0.3),
                         beta=SC.constant(target_value=1.0), filter_type="spacial",
                         filter_radius=SC.constant(target_value=0.1),
                         move_limit=SC.constant(target_value=0.2), E_min_coeff=1e-9,
                         precision="fp64", solver_rtol=1e-13, scaling=True)
    opt = core.OC_Optimizer(cfg, task)
    assert opt.L_scale == L and np.ptp(task.mesh.p, axis=1).max() == 1.0
    opt.parameterize()
    for it in range(1, n + 1):
        rec = loop.step(it)
        opt.optimize_steps(1)
        c = opt.recorder.latest("compliance")
        assert abs(c - rec["compliance"]) <= 1e-9 * abs(rec["compliance"]), (it, c, rec)
        assert opt.recorder.latest("bisection_evals") == rec["n_evals"], it
        assert np.abs(opt.rho_host() - loop.rho).max() <= 1e-8, it
    # the run ended: the task is back at its own lengths and loads
    np.testing.assert_array_equal(task.mesh.p, p0)
    np.testing.assert_allclose(task.neumann_linear[0], f0, rtol=1e-15)


@pytest.mark.parametrize("method", ["OC", "LogMOC"])
def test_design_dirichlet_true_matches_oracle(gpu, method, tmp_path):
    """``design_dirichlet=True`` (``common_density.py:736-744``): the elements
    touching Dirichlet nodes stay design variables (only the Neumann ones
    are pinned to 1), six iterations against the oracle's loop."""
    from sktopt_amd import core
    from sktopt_amd.mesh import toy_problem
    from sktopt_amd.tools import SchedulerConfig as SC
    n = 6
    S = odrv.Sched
    kw = dict(lambda_lower=-1e7, lambda_upper=1e7) if method == "LogMOC" else {}
    ocfg = odrv.Config(method=method, max_iters=n, p=S("Constant", 3.0, 3.0),
                       vol_frac=S("Constant", 0.3, 0.3), beta=S("Constant", 1.0, 1.0),
                       filter_radius=S("Constant", 1.5, 1.5), move_limit=S("Constant", 0.2, 0.2),
                       eta=S("Constant", 0.5, 0.5), E_min_coeff=1e-9, solver="spsolve",
                       design_dirichlet=True, **kw)
    loop = odrv.DensityLoop(ocfg, oprob.cantilever(16, 8, 8))
    common = dict(dst_path=str(tmp_path), max_iters=n, p=SC.constant(target_value=3.0),
                  vol_frac=SC.constant(target_value=0.3), beta=SC.constant(target_value=1.0),
                  filter_type="spacial", filter_radius=SC.constant(target_value=1.5),
                  move_limit=SC.constant(target_value=0.2), E_min_coeff=1e-9,
                  eta=SC.constant(target_value=0.5), precision="fp64", solver_rtol=1e-13,
                  design_dirichlet=True)
    task = toy_problem.cantilever(16, 8, 8)
    if method == "OC":
        opt = core.OC_Optimizer(core.OC_Config(**common), task)
    else:
        opt = core.LogMOC_Optimizer(core.LogMOC_Config(**common, **kw), task)
    opt.parameterize()
    # the design set keeps the clamped face's elements
    assert np.intersect1d(task.design_elements, task.dirichlet_elements).size > 0
    np.testing.assert_array_equal(np.sort(task.design_elements), np.sort(loop.design))
    for it in range(1, n + 1):
        rec = loop.step(it)
        opt.optimize_steps(1)
        c = opt.recorder.latest("compliance")
        assert abs(c - rec["compliance"]) <= 1e-9 * abs(rec["compliance"]), (it, c, rec)
        assert np.abs(opt.rho_host() - loop.rho).max() <= 1e-8, it
