"""Continuation schedules (oracle), restating ``tools/scheduler.py``.

Schedule functions: ``tools/scheduler.py:14-261``; ``Scheduler.value``
quirks (target when ``num_steps`` is None or < 0, or ``it >= iters_max``):
``tools/scheduler.py:809-830``; sawtooth uses ``it - 1``.
"""
from __future__ import annotations


def schedule_step(it, total, initial_value, target_value, num_steps, **_):
    if total <= 0:
        raise ValueError("total must be positive")
    if num_steps <= 1:
        return target_value
    step_length = total / num_steps
    step_index = min(int(it // step_length), num_steps - 1)
    alpha = step_index / (num_steps - 1)
    return (1 - alpha) * initial_value + alpha * target_value


def schedule_step_accelerating(it, total, initial_value, target_value,
                               num_steps, curvature=3.0, **_):
    if total <= 0:
        raise ValueError("total must be positive")
    if num_steps <= 1:
        return target_value
    step_length = total / num_steps
    step_index = min(int(it // step_length), num_steps - 1)
    a = (step_index / (num_steps - 1)) ** curvature
    return (1 - a) * initial_value + a * target_value


def schedule_step_decelerating(it, total, initial_value, target_value,
                               num_steps, curvature=3.0, **_):
    if total <= 0:
        raise ValueError("total must be positive")
    if num_steps <= 1:
        return target_value
    step_length = total / num_steps
    step_index = min(int(it // step_length), num_steps - 1)
    a = 1 - (1 - step_index / (num_steps - 1)) ** curvature
    return (1 - a) * initial_value + a * target_value


def schedule_sawtooth_decay(it, total, initial_value, target_value,
                            num_steps, **_):
    if total <= 0 or num_steps <= 0:
        raise ValueError("total and num_steps must be positive")
    it0 = it - 1
    step_size = total / num_steps
    step_index = int(it0 // step_size)
    local_index = it0 - step_index * step_size
    alpha = min(local_index / step_size, 1.0)
    return (1 - alpha) * initial_value + alpha * target_value


def schedule_constant(it, total, target_value, **_):
    return target_value


FUNCS = {
    "Constant": schedule_constant, "ConstantOne": schedule_constant,
    "Step": schedule_step, "StepToOne": schedule_step,
    "StepAccelerating": schedule_step_accelerating,
    "StepAcceleratingToOne": schedule_step_accelerating,
    "StepDecelerating": schedule_step_decelerating,
    "StepDeceleratingToOne": schedule_step_decelerating,
    "SawtoothDecay": schedule_sawtooth_decay,
    "None": None,
}


def value(kind, it, iters_max, init=None, target=None, num_steps=None,
          curvature=None):
    """``Scheduler.value`` (``tools/scheduler.py:809-830``) after
    ``Scheduler.from_config`` normalisation (``:746-808``)."""
    if kind in ("Constant",):
        init, num_steps = target, None
    if kind == "ConstantOne":
        init = target = 1.0
        num_steps = None
    if kind.endswith("ToOne") and kind != "ConstantOne":
        target = 1.0
        if num_steps:
            init = 1.0 / num_steps
    if target is None:
        return None
    if isinstance(num_steps, (int, float)):
        if num_steps < 0:
            return target
    elif num_steps is None:
        return target
    if it >= iters_max:
        return target
    return FUNCS[kind](it=it, total=iters_max, initial_value=init,
                       target_value=target, num_steps=num_steps,
                       curvature=curvature)
