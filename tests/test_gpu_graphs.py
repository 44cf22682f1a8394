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


@pytest.mark.parametrize("case", ["hex-8load-mixed", "hex-1load-fp64-sensfilter"])
def test_step_graph_replay_bit_identical_to_eager(gpu, case, tmp_path):
    """The iteration's device phases (filter and projection; energies to
    records) replayed from CUDA graphs give the same bits as running them
    eagerly (SKT_STEP_GRAPH=0): every iteration's compliance, record values
    and densities are compared, so each graph is replayed several times with
    the solve, the record copies and host synchronisations between replays."""
    from test_gpu_driver import _gpu_opt
    if case.startswith("hex-8load"):
        fn, kw = (lambda: oprob.multiload(24, 12, 12, n_loads=8, seed=0)), dict(vol=0.2, r=2.0)
        precision = "mixed"
    else:
        fn, kw = (lambda: oprob.cantilever(40, 20, 20)), dict(sensitivity_filter=True)
        precision = "fp64"
    runs = {}
    for mode in ("1", "0"):
        old = os.environ.get("SKT_STEP_GRAPH")
        os.environ["SKT_STEP_GRAPH"] = mode
        try:
            opt = _gpu_opt(fn(), iters=7, precision=precision, rtol=1e-8,
                           tmp=str(tmp_path / mode), **kw)
        finally:
            if old is None:
                del os.environ["SKT_STEP_GRAPH"]
            else:
                os.environ["SKT_STEP_GRAPH"] = old
        recs = []
        for _ in range(7):
            opt.optimize_steps(1)
            h = opt.recorder
            recs.append([np.array(h.latest(n)) for n in ("compliance_each", "lmid", "kkt_residual",
                                                         "rho_change_max", "rho_projected",
                                                         "energy", "-dC", "bisection_evals")])
        runs[mode] = (opt, recs, opt.rho_host())
    g, e = runs["1"], runs["0"]
    assert len(g[0]._phases) == 2 and len(e[0]._phases) == 0
    for it, (a, b) in enumerate(zip(g[1], e[1])):
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y, err_msg=f"iteration {it + 1}")
    np.testing.assert_array_equal(g[2], e[2])
