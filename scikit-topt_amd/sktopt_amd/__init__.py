"""sktopt_amd -- MI355X-native density-method hot path with scikit-topt's
optimizer, FEA and filter call surface.

Layout mirrors ``sktopt``: ``core`` (OC / LogMOC drivers, projection,
derivatives), ``fea`` (matrix-free GPU state solve), ``filters`` (neighbour-
list density filter), ``mesh`` (problem definitions), ``tools`` (schedules).
All numerics run in the HIP library ``_lib/libsktopt_amd.so``; importing a
compute entry point without it raises.
"""
from . import _lib  # noqa: F401
from . import tools, mesh  # noqa: F401
from . import fea, filters, core  # noqa: F401

__version__ = "0.1.0"
