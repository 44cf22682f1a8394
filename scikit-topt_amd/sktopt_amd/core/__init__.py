"""Optimizers and design-variable transforms (``sktopt/core``)."""
from . import projection, derivatives  # noqa: F401
from .optimizers import (OC_Config, OC_Optimizer, LogMOC_Config,  # noqa: F401
                         LogMOC_Optimizer, DensityMethodConfig)
from . import optimizers  # noqa: F401
from .autograd import Compliance, ComplianceProblem  # noqa: F401
