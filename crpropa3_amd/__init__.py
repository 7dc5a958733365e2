"""crpropa3_amd: MI355X-native batched charged-particle propagator.

Drop-in for CRPropa3's PropagationBP / PropagationCK hot path over Grid3f,
PlaneWaveTurbulence and uniform fields:

    from crpropa3_amd import *
    sim = ModuleList()
    sim.add(PropagationCK(MagneticFieldGrid(grid), 1e-4, 1 * kpc, 1 * Mpc))
    sim.add(MaximumTrajectoryLength(100 * Mpc))
    sim.run(candidates)

The classes come from the C++ host layer (crpropa3_amd/host, pybind11 module
_crpropa); the compute path is the HIP library libcrpropa3_amd.so (C ABI in
include/crpropa3_amd.h, Python binding in crpropa3_amd.capi).  There is no
CPU fallback: importing fails if the libraries are not built.
"""
# PyTorch ships its own copy of the HIP runtime (libamdhip64, same soname as ROCm's).  Two copies
# in one process cannot both use the GPU: whichever initialises second sees no device.  Importing
# torch before this package's HIP library is loaded makes the library bind to torch's copy, so
# torch tensors (crpropa3_amd.adjoint) and the propagator share one runtime.
try:
    import torch as _torch  # noqa: F401
except ImportError:
    pass

from . import capi  # noqa: F401,E402
from ._crpropa import *  # noqa: F401,F403,E402
from ._crpropa import __doc__ as _native_doc  # noqa: F401
