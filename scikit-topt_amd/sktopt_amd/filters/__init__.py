"""Density filters (``filters/__init__.py``): the neighbour-list Gaussian
filter ``SpacialFilter`` on the GPU."""
from .spacial import SpacialFilter, make_kernel  # noqa: F401
