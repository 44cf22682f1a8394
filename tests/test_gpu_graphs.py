"""The solver's PCG chunks replayed from captured hipGraphs give the same
bits as the same kernels launched one by one (SKT_GRAPH=0): every OC
iteration of several-iteration runs is compared, so each graph is replayed
many times with the driver's own eager work and host synchronisations
between the replays."""
import os

import numpy as np
import pytest

from oracle import problems as oprob

pytestmark = pytest.mark.gpu


def _run(task_fn, graphs, precision, iters, tmp, **kw):
    from test_gpu_driver import _gpu_opt
    old = os.environ.get("SKT_GRAPH")
    os.environ["SKT_GRAPH"] = "1" if graphs else "0"
    try:
        opt = _gpu_opt(task_fn(), iters=iters, precision=precision, rtol=1e-8, tmp=tmp, **kw)
    finally:
        if old is None:
            del os.environ["SKT_GRAPH"]
        else:
            os.environ["SKT_GRAPH"] = old
    comps = []
    for _ in range(iters):
        opt.optimize_steps(1)
        comps.append(np.array(opt.recorder.latest("compliance_each")))
    return opt, np.array(comps), opt.rho_host()


@pytest.mark.parametrize("case", ["hex-8load-mixed", "hex-8load-fp64", "hex-1load-mixed",
                                  "tet-jacobi-mixed"])
def test_graph_replay_bit_identical_to_direct_launches(gpu, case, tmp_path):
    if case.startswith("hex-8load"):
        fn, kw = (lambda: oprob.multiload(24, 12, 12, n_loads=8, seed=0)), dict(vol=0.2, r=2.0)
    elif case.startswith("hex-1load"):
        fn, kw = (lambda: oprob.cantilever(40, 20, 20)), dict()
    else:
        fn, kw = (lambda: oprob.cantilever(8, 4, 4, kind="tet4", jitter=0.2, seed=3)), dict()
    precision = case.rsplit("-", 1)[1]
    iters = 6
    a, ca, ra = _run(fn, True, precision, iters, str(tmp_path / "g"), **kw)
    b, cb, rb = _run(fn, False, precision, iters, str(tmp_path / "d"), **kw)
    ga, gb = a.fem.graph_stats(), b.fem.graph_stats()
    assert ga["launches"] > 0 and ga["captures"] >= 1, ga
    assert gb["launches"] == 0 and gb["captures"] == 0, gb
    # captures happen once per variant, not once per replay
    assert ga["captures"] <= 4 < ga["launches"], ga
    np.testing.assert_array_equal(ca, cb)
    np.testing.assert_array_equal(ra, rb)
    assert [s["iters"] for s in a.fem.solve_history] == [s["iters"] for s in b.fem.solve_history]
