"""The reference's end-to-end test (tests/test_global_flow.py:53-157) on the
GPU path: OC on toy_test(), LogMOC on toy2() and OC on a Robin heat task,
one iteration each with the default configuration (so the default nodal
Helmholtz filter), asserting a finite objective as the reference does."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _heat_task():
    """tests/test_global_flow.py:53-103: 8 x 8 x 1 box at mesh size 0.2,
    Robin (h = 4e-5, T_env = 300) on the x = 0 and y = 8 boundaries,
    T = 600 on the x >= 7.6, y <= 0.4 corner, k = 10, Robin boundary
    elements designable."""
    from sktopt_amd.mesh import toy_problem as tp
    from sktopt_amd.mesh.task import LinearHeatConduction
    x_len, y_len, z_len, ms = 8.0, 8.0, 1.0, 0.2
    m = tp.create_box_hex(x_len, y_len, z_len, ms)
    sel = lambda *r: m.facets_satisfying(tp.get_points_in_range(*r),  # noqa: E731
                                         boundaries_only=True)
    r0 = sel((0.0, 0.0), (0.0, y_len), (0.0, z_len))
    r1 = sel((0.0, x_len), (y_len, y_len), (0.0, z_len))
    d = sel((x_len - x_len / 20, x_len), (0.0, y_len / 20), (0.0, z_len))
    return LinearHeatConduction.from_facets(m, d, 600.0, [r0, r1], 4.0e-5, 300.0, True,
                                            np.arange(m.nelements), 10.0, "compliance")


def test_reference_global_flow(gpu, tmp_path):
    from sktopt_amd import core
    from sktopt_amd.mesh import toy_problem as tp
    from sktopt_amd.tools import SchedulerConfig
    # OC on toy_test
    cfg = core.OC_Config()
    cfg.max_iters = 1
    cfg.record_times = cfg.max_iters
    cfg.dst_path = str(tmp_path / "oc")
    opt = core.OC_Optimizer(cfg, tp.toy_test())
    opt.parameterize()
    opt.optimize()
    assert np.isfinite(opt.recorder.latest("compliance"))
    assert (tmp_path / "oc" / "histories.npz").exists()
    # LogMOC on toy2 (p Step 1 -> 3 over 3 steps, vol_frac target 0.6)
    cfg = core.LogMOC_Config(
        p=SchedulerConfig(init_value=1.0, target_value=3.0, num_steps=3,
                          scheduler_type="Step"),
        vol_frac=SchedulerConfig(target_value=0.6, scheduler_type="Step"))
    cfg.max_iters = 1
    cfg.record_times = cfg.max_iters
    cfg.dst_path = str(tmp_path / "logmoc")
    opt = core.LogMOC_Optimizer(cfg, tp.toy2())
    opt.parameterize()
    opt.optimize()
    assert np.isfinite(opt.recorder.latest("compliance"))
    # OC on the Robin heat task
    cfg = core.OC_Config()
    cfg.max_iters = 1
    cfg.record_times = cfg.max_iters
    cfg.dst_path = str(tmp_path / "heat")
    opt = core.OC_Optimizer(cfg, _heat_task())
    opt.parameterize()
    opt.optimize()
    assert np.isfinite(opt.recorder.latest("compliance"))
