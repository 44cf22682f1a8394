"""Host-side tools: continuation schedules (``tools/scheduler.py``)."""
from .scheduler import (SchedulerConfig, Scheduler, Schedulers,  # noqa: F401
                        SchedulerStepAcceleratingToOne, SchedulerStepDeceleratingToOne)
