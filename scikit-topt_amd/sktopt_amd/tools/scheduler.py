"""Continuation schedules (host scalars per iteration).

Restates ``tools/scheduler.py``: the schedule functions (``:14-261``),
``SchedulerConfig`` with its factories and validation (``:272-682``),
``Scheduler.from_config`` / ``value`` (``:746-830``) and ``Schedulers``
(``:1138-1189``).  Iterations are 1-based.  Quirks kept: ``value`` returns the
target when ``num_steps`` is None or negative or ``it >= iters_max``; the
sawtooth uses ``it - 1``; ``sawtooth_decay(name, init, target, iters_max,
num_steps)`` takes ``iters_max`` before ``num_steps``.  Plotting/export of
the reference's schedule figures is not reproduced.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable, Optional


def schedule_constant(it, total, target_value=0.4, **args):
    return target_value


def _step_index(it, total, num_steps):
    step_length = total / num_steps
    return min(int(it // step_length), num_steps - 1)


def schedule_step(it, total, initial_value=1.0, target_value=0.4, num_steps=10, **args):
    if total <= 0:
        raise ValueError("total must be positive")
    if num_steps <= 1:
        return target_value
    alpha = _step_index(it, total, num_steps) / (num_steps - 1)
    return (1 - alpha) * initial_value + alpha * target_value


def schedule_step_accelerating(it, total, initial_value=1.0, target_value=0.4,
                               num_steps=10, curvature=3.0, **args):
    if total <= 0:
        raise ValueError("total must be positive")
    if num_steps <= 1:
        return target_value
    a = (_step_index(it, total, num_steps) / (num_steps - 1)) ** curvature
    return (1 - a) * initial_value + a * target_value


def schedule_step_decelerating(it, total, initial_value=1.0, target_value=0.4,
                               num_steps=10, curvature=3.0, **args):
    if total <= 0:
        raise ValueError("total must be positive")
    if num_steps <= 1:
        return target_value
    a = 1 - (1 - _step_index(it, total, num_steps) / (num_steps - 1)) ** curvature
    return (1 - a) * initial_value + a * target_value


def schedule_sawtooth_decay(it, total, initial_value=0.1, target_value=0.05,
                            num_steps=6, **args):
    if total <= 0 or num_steps <= 0:
        raise ValueError("total and num_steps must be positive")
    it0 = it - 1
    step_size = total / num_steps
    step_index = int(it0 // step_size)
    local_index = it0 - step_index * step_size
    alpha = min(local_index / step_size, 1.0)
    return (1 - alpha) * initial_value + alpha * target_value


_TYPES = ("Constant", "ConstantOne", "Step", "StepAccelerating", "StepDecelerating",
          "SawtoothDecay", "StepToOne", "StepAcceleratingToOne", "StepDeceleratingToOne",
          "None")


@dataclass
class SchedulerConfig:
    name: Optional[str] = None
    init_value: Optional[float] = None
    target_value: Optional[float] = None
    num_steps: Optional[int] = None
    iters_max: Optional[int] = None
    curvature: Optional[float] = None
    scheduler_type: str = "Constant"

    @classmethod
    def from_defaults(cls, name=None, init_value=None, target_value=None, num_steps=None,
                      iters_max=None, curvature=None, scheduler_type="Constant"):
        st = scheduler_type
        if st == "Constant":
            if init_value is None:
                init_value = target_value
            if target_value is None:
                raise ValueError("Should set target_value")
        elif st == "ConstantOne":
            init_value = target_value = 1.0
        elif st in ("Step", "StepAccelerating", "StepDecelerating", "SawtoothDecay"):
            if init_value is None:
                raise ValueError("Should set init_value")
            if target_value is None:
                raise ValueError("Should set target_value")
            if num_steps is None:
                raise ValueError("Should set num_steps")
            if st in ("StepAccelerating", "StepDecelerating") and curvature is None:
                raise ValueError("Should set curvature")
        elif st in ("StepToOne", "StepAcceleratingToOne", "StepDeceleratingToOne"):
            if num_steps is None:
                raise ValueError("Should set num_steps")
            if num_steps <= 0:
                raise ValueError(f"num_steps must be positive for {st}")
            if st != "StepToOne" and curvature is None:
                raise ValueError("Should set curvature")
            if target_value is not None and not math.isclose(target_value, 1.0):
                raise ValueError(f"{st} fixes target_value to 1.0")
            target_value = 1.0
            init_value = 1.0 / num_steps
        elif st == "None":
            pass
        else:
            raise ValueError("")
        return cls(name=name, init_value=init_value, target_value=target_value,
                   num_steps=num_steps, iters_max=iters_max, curvature=curvature,
                   scheduler_type=st)

    @classmethod
    def none(cls):
        return cls.from_defaults(scheduler_type="None")

    @classmethod
    def constant(cls, name=None, target_value=1.0):
        return cls.from_defaults(name=name, init_value=target_value, target_value=target_value,
                                 scheduler_type="Constant")

    @classmethod
    def constant_one(cls, name=None):
        return cls.from_defaults(name=name, init_value=1.0, target_value=1.0,
                                 scheduler_type="ConstantOne")

    @classmethod
    def step(cls, name=None, init_value=None, target_value=None, num_steps=None, iters_max=None):
        return cls.from_defaults(name, init_value, target_value, num_steps, iters_max, None, "Step")

    @classmethod
    def step_to_one(cls, name=None, num_steps=None, iters_max=None):
        return cls.from_defaults(name, None, 1.0, num_steps, iters_max, None, "StepToOne")

    @classmethod
    def step_accelerating(cls, name=None, init_value=None, target_value=None, num_steps=None,
                          iters_max=None, curvature=None):
        return cls.from_defaults(name, init_value, target_value, num_steps, iters_max,
                                 curvature, "StepAccelerating")

    @classmethod
    def step_decelerating(cls, name=None, init_value=None, target_value=None, num_steps=None,
                          iters_max=None, curvature=None):
        return cls.from_defaults(name, init_value, target_value, num_steps, iters_max,
                                 curvature, "StepDecelerating")

    @classmethod
    def step_accelerating_to_one(cls, name=None, num_steps=None, iters_max=None, curvature=None):
        return cls.from_defaults(name, None, 1.0, num_steps, iters_max, curvature,
                                 "StepAcceleratingToOne")

    @classmethod
    def step_decelerating_to_one(cls, name=None, num_steps=None, iters_max=None, curvature=None):
        return cls.from_defaults(name, None, 1.0, num_steps, iters_max, curvature,
                                 "StepDeceleratingToOne")

    @classmethod
    def sawtooth_decay(cls, name=None, init_value=0.1, target_value=0.05, iters_max=100,
                       num_steps=6):
        return cls(name=name, init_value=init_value, target_value=target_value,
                   num_steps=num_steps, iters_max=iters_max, curvature=None,
                   scheduler_type="SawtoothDecay")


class Scheduler:
    def __init__(self, name, init_value, target_value, num_steps=None, iters_max=None,
                 curvature=None, func: Callable | None = schedule_step):
        self.name = name
        self.init_value = init_value
        self.target_value = target_value
        self.iters_max = iters_max
        self.num_steps = num_steps
        self.curvature = curvature
        self.func = func

    @classmethod
    def from_config(cls, cfg: SchedulerConfig):
        st = cfg.scheduler_type
        if st == "Constant":
            func = schedule_constant
            cfg.init_value = cfg.target_value
            cfg.iters_max = cfg.num_steps = cfg.curvature = None
        elif st == "ConstantOne":
            func = schedule_constant
            cfg.init_value = cfg.target_value = 1.0
            cfg.iters_max = cfg.num_steps = cfg.curvature = None
        elif st == "Step":
            func = schedule_step
        elif st in ("StepToOne", "StepAcceleratingToOne", "StepDeceleratingToOne"):
            func = {"StepToOne": schedule_step,
                    "StepAcceleratingToOne": schedule_step_accelerating,
                    "StepDeceleratingToOne": schedule_step_decelerating}[st]
            cfg.target_value = 1.0
            if cfg.num_steps:
                cfg.init_value = 1.0 / cfg.num_steps
        elif st == "StepAccelerating":
            func = schedule_step_accelerating
        elif st == "StepDecelerating":
            func = schedule_step_decelerating
        elif st == "SawtoothDecay":
            func = schedule_sawtooth_decay
        elif st == "None":
            func = None
        else:
            raise ValueError(f"{st} not in {list(_TYPES)}")
        return cls(cfg.name, cfg.init_value, cfg.target_value, cfg.num_steps,
                   iters_max=cfg.iters_max, curvature=cfg.curvature, func=func)

    def value(self, iter):
        if self.target_value is None:
            return None
        if isinstance(self.num_steps, (int, float)):
            if self.num_steps < 0:
                return self.target_value
        elif self.num_steps is None:
            return self.target_value
        if iter >= self.iters_max:
            return self.target_value
        return self.func(it=iter, total=self.iters_max, initial_value=self.init_value,
                         target_value=self.target_value, num_steps=self.num_steps,
                         curvature=self.curvature)


class SchedulerStepAcceleratingToOne(Scheduler):
    def __init__(self, name, num_steps, iters_max=None, curvature=None):
        if num_steps is None or num_steps <= 0:
            raise ValueError("num_steps must be positive for SchedulerStepAcceleratingToOne")
        if curvature is None:
            raise ValueError("curvature is required for SchedulerStepAcceleratingToOne")
        super().__init__(name, 1.0 / num_steps, 1.0, num_steps, iters_max, curvature,
                         schedule_step_accelerating)


class SchedulerStepDeceleratingToOne(Scheduler):
    def __init__(self, name, num_steps, iters_max=None, curvature=None):
        if num_steps is None or num_steps <= 0:
            raise ValueError("num_steps must be positive for SchedulerStepDeceleratingToOne")
        if curvature is None:
            raise ValueError("curvature is required for SchedulerStepDeceleratingToOne")
        super().__init__(name, 1.0 / num_steps, 1.0, num_steps, iters_max, curvature,
                         schedule_step_decelerating)


class Schedulers:
    def __init__(self, dst_path: str = "."):
        self.scheduler_list = []
        self.dst_path = dst_path

    def set_iters_max(self, iters_max: int):
        for s in self.scheduler_list:
            s.iters_max = iters_max

    def value_on_a_scheduler(self, key, iter):
        return self.values_as_list(iter, [key], export_log=False)[0]

    def values_as_dict(self, iter):
        return {s.name: s.value(iter) for s in self.scheduler_list}

    def values_as_list(self, iter, order, export_log=False, precision=4):
        d = self.values_as_dict(iter)
        return [d[k] for k in order]

    def add_object(self, s):
        self.scheduler_list.append(s)

    def add_object_from_config(self, cfg, rewrite_name=None):
        if isinstance(rewrite_name, str):
            cfg.name = rewrite_name
        self.add_object(Scheduler.from_config(cfg))

    def add(self, name, init_value, target_value, num_steps, iters_max=None, curvature=None,
            func=schedule_step):
        self.scheduler_list.append(
            Scheduler(name, init_value, target_value, num_steps, iters_max, curvature, func))
