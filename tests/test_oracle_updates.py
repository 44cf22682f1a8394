"""Pin the oracle's OC bisection, LogMOC step and schedules independently:
a brute-force root of the volume constraint, bound/monotonicity invariants,
hand-derived multiplicative move limits, the reference's own scheduler test
properties and hand-computed schedule values."""
import numpy as np
import pytest
from scipy.optimize import brentq

from oracle import filter as of
from oracle import logmoc as olm
from oracle import oc as ooc
from oracle import problems as op
from oracle import scheduler as osch
from oracle import transforms as tr


def _bisect_case(seed):
    task = op.cantilever(8, 4, 4)
    mask = task.design_mask
    design = np.nonzero(mask)[0]
    rng = np.random.default_rng(seed)
    rho = np.ones(task.n_elements)
    rho[design] = rng.uniform(0.05, 0.9, design.size)
    dC = -rng.uniform(0.0, 1.0, design.size) ** 2
    f = of.SpacialFilter(task.centers, 1.5, mask)
    vol = task.elements_volume[design]
    return task, design, rho, dC, f, vol


def _volume(lam, dC, rho, design, f, vol, beta=2.0, eta=0.5, m=0.2):
    rate = np.clip((-dC / (lam + 1e-12)) ** 0.5, 0.7, 1.3)
    cand = np.clip(rho[design] * rate, np.maximum(rho[design] - m, 1e-2),
                   np.minimum(rho[design] + m, 1.0))
    full = rho.copy()
    full[design] = cand
    proj = tr.heaviside_projection(f.forward(full), beta, eta)
    return np.sum(proj[design] * vol) / vol.sum()


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_bisection_finds_the_volume_root(seed):
    task, design, rho, dC, f, vol = _bisect_case(seed)
    vf = 0.35
    lmid, err, cand, n = ooc.bisection_with_physical_volume(
        dC, rho[design], rho, design, f, 1e-2, 1.0, 0.2, 0.5, 1e-12, vf, 2.0, 0.5,
        vol, vol.sum(), 0.7, 1.3, max_iter=1000, tolerance=1e-5, l1=1e-7, l2=1e7)
    # invariants: exit condition, bounds, reported error is the candidate's
    assert abs(err) < 1e-4 or n == 41
    lo = np.maximum(rho[design] - 0.2, 1e-2)
    hi = np.minimum(rho[design] + 0.2, 1.0)
    assert np.all(cand >= lo - 1e-15) and np.all(cand <= hi + 1e-15)
    np.testing.assert_allclose(_volume(lmid, dC, rho, design, f, vol) - vf, err, rtol=1e-12)
    # brute force: the volume is non-increasing in lambda; bracket the root
    lams = np.logspace(-7, 7, 57)
    v = np.array([_volume(l_, dC, rho, design, f, vol) for l_ in lams])
    assert np.all(np.diff(v) <= 1e-12)
    if v[0] > vf > v[-1]:
        root = brentq(lambda l_: _volume(l_, dC, rho, design, f, vol) - vf, 1e-7, 1e7,
                      xtol=1e-12)
        # |vol error| < 1e-4 at the returned lambda, and the exact root is close
        assert abs(err) < 1e-4
        vdiff = abs(_volume(root, dC, rho, design, f, vol) - _volume(lmid, dC, rho, design, f, vol))
        assert vdiff < 1e-4


def test_bisection_evaluation_count_is_bounded():
    """Bracket [1e-7, 1e7] halved to 1e-5 allows at most 41 evaluations."""
    task, design, rho, dC, f, vol = _bisect_case(0)
    out = ooc.bisection_with_physical_volume(
        dC, rho[design], rho, design, f, 1e-2, 1.0, 0.2, 0.5, 1e-12, 2.0, 2.0, 0.5,
        vol, vol.sum(), 0.7, 1.3, max_iter=1000, tolerance=1e-5, l1=1e-7, l2=1e7)
    assert out[3] == 41          # vol_frac = 2 is unreachable: exit on the bracket


def test_logmoc_step_multiplicative_move_limits():
    rho = np.array([0.5, 0.5, 0.5, 0.005])
    dL = np.array([0.2, 1.5, -3.0, 0.0])
    out = olm.lagrangian_log_update(rho.copy(), dL, 0.6, 0.2, 0.01, 1.0, 1.0)
    # free step exp(log 0.5 - 0.6*0.2); clipped steps hit rho/(1+m/rho), rho(1+m/rho)
    np.testing.assert_allclose(out, [0.5 * np.exp(-0.12), 0.5 / 1.4, 0.5 * 1.4, 0.01],
                               rtol=1e-14)


def test_sched_hand_values():
    # Step 1 -> 3 in 3 steps over 10 iterations: step length 10/3
    got = [osch.value("Step", it, 10, 1.0, 3.0, 3) for it in range(1, 11)]
    assert got == [1.0, 1.0, 1.0, 2.0, 2.0, 2.0, 3.0, 3.0, 3.0, 3.0]
    # sawtooth 0.3 -> 0.1, 6 teeth over 12 iterations: period 2, uses it - 1
    got = [osch.value("SawtoothDecay", it, 12, 0.3, 0.1, 6) for it in range(1, 6)]
    np.testing.assert_allclose(got, [0.3, 0.2, 0.3, 0.2, 0.3])
    assert osch.value("Constant", 5, 10, None, 0.7) == 0.7
    assert osch.value("None", 5, 10) is None
    assert osch.value("Step", 5, 10, 1.0, 3.0, None) == 3.0      # num_steps None -> target
    assert osch.value("Step", 12, 10, 1.0, 3.0, 3) == 3.0        # it >= iters_max -> target


def _products_schedulers():
    from sktopt_amd.tools import scheduler as ps
    return ps


def test_reference_scheduler_properties_oracle_and_product():
    """tests/test_scheduler.py:16-114 of the reference, on both restatements."""
    ps = _products_schedulers()
    for mod in (osch, ps):
        dec = mod.schedule_step_decelerating
        acc = mod.schedule_step_accelerating
        lin = mod.schedule_step
        kw = dict(total=10, initial_value=1.0, target_value=5.0, num_steps=5, curvature=2.0)
        vals = [dec(it=i, **kw) for i in (1, 3, 5, 7, 9)]
        assert np.isclose(vals[0], 1.0) and np.isclose(vals[-1], 5.0)
        assert all(a <= b for a, b in zip(vals, vals[1:]))
        assert acc(it=3, **kw) < lin(it=3, **kw) < dec(it=3, **kw)
        assert acc(it=7, **kw) < lin(it=7, **kw) < dec(it=7, **kw)
        kw1 = dict(total=10, initial_value=0.5, target_value=1.5, num_steps=4, curvature=1.0)
        for i in (1, 4, 7, 9):
            assert np.isclose(lin(it=i, **kw1), dec(it=i, **kw1))
            assert np.isclose(lin(it=i, **kw1), acc(it=i, **kw1))
    cfg = ps.SchedulerConfig.step_to_one(name="x", num_steps=4, iters_max=8)
    s = ps.Scheduler.from_config(cfg)
    assert np.isclose(cfg.init_value, 0.25) and np.isclose(s.value(1), 0.25)
    assert np.isclose(s.value(8), 1.0)
    with pytest.raises(ValueError):
        ps.SchedulerConfig.from_defaults(name="x", num_steps=3, target_value=0.5,
                                         scheduler_type="StepToOne")
    for f in (ps.SchedulerConfig.step_accelerating_to_one,
              ps.SchedulerConfig.step_decelerating_to_one):
        cfg = f(name="x", num_steps=5, iters_max=10, curvature=2.0)
        s = ps.Scheduler.from_config(cfg)
        assert np.isclose(cfg.init_value, 0.2) and np.isclose(s.value(1), 0.2)
        assert s.value(10) <= 1.0


def test_product_scheduler_matches_oracle():
    ps = _products_schedulers()
    cases = [("Step", 1.0, 3.0, 3, None), ("StepAccelerating", 1.0, 8.0, 6, 3.0),
             ("StepDecelerating", 2.0, 0.5, 4, 2.0), ("SawtoothDecay", 0.3, 0.1, 6, None)]
    for kind, a, b, n, c in cases:
        if kind == "SawtoothDecay":
            cfg = ps.SchedulerConfig.sawtooth_decay("m", a, b, 3, n)
        else:
            cfg = ps.SchedulerConfig.from_defaults("m", a, b, n, None, c, kind)
        s = ps.Schedulers()
        s.add_object_from_config(cfg, "m")
        s.set_iters_max(40)
        for it in range(1, 45):
            assert s.value_on_a_scheduler("m", it) == osch.value(kind, it, 40, a, b, n, c)


def test_default_move_limit_binds_iters_max_3():
    """sawtooth_decay('move_limit', 0.3, 0.1, 3) puts 3 in iters_max, not
    num_steps (tools/scheduler.py:658-665); set_iters_max later overrides it."""
    ps = _products_schedulers()
    cfg = ps.SchedulerConfig.sawtooth_decay("move_limit", 0.3, 0.1, 3)
    assert cfg.iters_max == 3 and cfg.num_steps == 6
