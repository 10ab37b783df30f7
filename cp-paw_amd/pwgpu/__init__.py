"""pwgpu: MI355X-native Car-Parrinello wavefunction step of CP-PAW.

Host-side mirror of the reference's PLANEWAVE and WAVES objects over the C ABI
of libpwgpu.so (include/pwgpu.h).  The compute runs in hand-written HIP
kernels for gfx950; this package only marshals arguments.
"""
from ._lib import PW_PREC_FP64, PW_PREC_MIXED, PwError, load  # noqa: F401
from .planewave import DeviceArray, PlaneWave  # noqa: F401
from .waves import FP64, MIXED, Waves, ortho_x  # noqa: F401
from .potential import Potential  # noqa: F401
from . import control, model, restart  # noqa: F401
