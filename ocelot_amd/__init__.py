"""MI355X-native 3D space-charge kick, a drop-in for Ocelot's SpaceCharge
(and its 1D longitudinal counterpart LSC).

    from ocelot_amd import SpaceCharge, LSC, ParticleArray, DeviceParticleArray

The compute path is libocsc.so (HIP kernels for gfx950, C ABI in
include/ocsc.h); importing this package does not load it, the first call
does, and raises if it is missing.
"""
from .physproc import PhysProc
from .beam import ParticleArray, DeviceParticleArray
from .sc import SpaceCharge, apply_batched, drift, track_drift
from .lsc import LSC
from .io import load_particle_array, save_particle_array
from .envelope import Twiss, get_envelope
from .autotune import tune_tracking

__all__ = ["PhysProc", "ParticleArray", "DeviceParticleArray", "SpaceCharge", "LSC", "apply_batched", "drift",
           "track_drift", "Twiss", "get_envelope", "tune_tracking", "load_particle_array", "save_particle_array"]
