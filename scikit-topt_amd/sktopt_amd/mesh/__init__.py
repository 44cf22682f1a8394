"""Problem definitions: meshes, boundary conditions, loads and design sets.

Host-side setup (numpy).  The reference builds these with scikit-fem
(``mesh/task_common.py``, ``mesh/task_elastic.py``, ``mesh/task_heat.py``);
here meshes are this package's own structured hex grids and tet meshes, and
boundary conditions are given as node sets with explicit nodal load vectors.

Design-set rule (``mesh/task_common.py:211-227,398-401``): design = the
requested elements minus elements touching a load node, minus elements
touching a Dirichlet node (``design_dirichlet=False``); both excluded sets are
pinned to rho = 1 (``core/optimizers/common_density.py:736-744``).
"""
from .meshes import MeshHex, MeshTet, structured_hex, jittered_tet  # noqa: F401
from .task import FEMDomain, LinearElasticity, LinearHeatConduction  # noqa: F401
from . import toy_problem  # noqa: F401
