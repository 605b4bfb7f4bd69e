"""gfsph: MI355X-native Geoflow-SPH propagation step (host side).

The compute path is the HIP library libgfsph.so (csrc/gfsph.hip) behind the
C ABI declared in include/gfsph.h; this package builds it, mirrors the
reference's Init_sph / time_integration_sph verbs over it (api.SPHModel),
and generates the synthetic DEMs and particle releases of the workloads
(dem, cases).
"""
from . import cases, dem  # noqa: F401
from ._build import LIB, build  # noqa: F401
from .api import AdjointModel, BatchAdjointModel, BatchModel, GfsphError, SPHModel, load_library, neighbour_pairs, write_netcdf  # noqa: F401,E501
from .inputs import Inputs  # noqa: F401
