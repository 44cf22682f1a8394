"""Density-method optimizers (``core/optimizers/__init__.py``)."""
from .common_density import (DensityMethodConfig, DensityMethod_OC_Config,  # noqa: F401
                             DensityMethod, HistoryCollection)
from .oc import OC_Config, OC_Optimizer  # noqa: F401
from .logmoc import LogMOC_Config, LogMOC_Optimizer  # noqa: F401
