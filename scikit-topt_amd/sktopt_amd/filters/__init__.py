"""Density filters (``filters/__init__.py``): the neighbour-list Gaussian
filter ``SpacialFilter`` and the nodal Helmholtz filter
``HelmholtzFilterNodal`` on the GPU."""
from .spacial import SpacialFilter, make_kernel  # noqa: F401
from .helmholtz import HelmholtzFilterNodal  # noqa: F401
