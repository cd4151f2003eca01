"""sie_mom_amd — MI355X-native SIE-MoM hot path (assembly, dense LU / IR,
GMRES, adjoint gradient) over the C ABI of libsie_amd.so."""
from ._lib import SieError, SieInvalidArgument, SieOverflow  # noqa: F401
from .physics import (  # noqa: F401
    Medium, GstcOptions, PecOptions, DielectricOptions, physical_parameters,
    calderon_params, plane_wave, frequency_for_k0, fibonacci_directions, fibonacci_plane_waves,
)
from .mesh import Mesh  # noqa: F401
from . import ops, linalg  # noqa: F401
