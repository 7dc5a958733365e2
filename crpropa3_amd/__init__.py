"""crpropa3_amd: MI355X-native batched charged-particle propagator.

Drop-in for CRPropa3's PropagationBP / PropagationCK hot path over Grid3f,
PlaneWaveTurbulence and uniform fields.  The compute path is the HIP library
libcrpropa3_amd.so (C ABI in include/crpropa3_amd.h), reached through
crpropa3_amd.capi.
"""
from . import capi  # noqa: F401

__all__ = ["capi"]
