"""MI355X-native SSE sweep (host side).

Host-side mirror of StochasticSeriesExpansion.jl's model/table layer and of
Carlo's AbstractMC interface, above the C ABI of libssegpu.so (gfx950 HIP
kernels).  The directory name contains a dot, so load it with
``__graft_entry__.load_package()`` (registered as module ``sse_amd``).
"""
from . import util, lattice, vertex_data, sse_data, models, capi, h5lite, carlo, carlo_mc, host_estimators  # noqa: F401
from .lattice import Lattice, UnitCell, UnitCells, UCBond, UCSite, neel_vector  # noqa: F401
from .vertex_data import VertexData  # noqa: F401
from .sse_data import SSEBond, SSEData, FlatModel  # noqa: F401
from .models import (MagnetModel, ClusterModel, ClusterBases, ClusterBasis,  # noqa: F401
                     MagnetizationEstimator, all_magnetization_estimators, spin_operators,
                     bosonic_a_operator, lift_twobody_operator)
from .capi import Batch, SSE_MODE_WAVE, SSE_MODE_LANE, walker_desc  # noqa: F401
