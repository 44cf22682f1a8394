"""GPU state solve with the reference's FEA call surface
(``fea/solver_elastic.py:540-676``, ``fea/solver_heat.py:628-980``)."""
from .composer import simp_interpolation, ramp_interpolation  # noqa: F401
from .solver import (LinearSolverConfig, FEM_SimpLinearElasticity,  # noqa: F401
                     FEM_SimpLinearHeatConduction, normalize_linear_solver_config)
