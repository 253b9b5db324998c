"""MI355X (gfx950) particle-tracing core of radiativeTransfer.

Drop-in for the path ``obs = radiativeTransfer(geometry, source, material,
observation)`` of cottereau/RadiativeTransferMonteCarlo: the particle kernels
run in librtmc.so (hand-written HIP, C ABI in include/rtmc.h); this package
is the host side that mirrors the reference's MATLAB interface.
"""
from .material import MaterialClass, InverseCDF, zoeppritz_fluid
from .observation import initializeObservation, normalize, Observation
from .solver import (radiativeTransfer, Particles, Draws, DeviceMaterial, DeviceBins, initializeParticle,
                     propagateParticleSmallDt, propagateParticle, observeTime, forward_histogram)

__all__ = ["MaterialClass", "InverseCDF", "zoeppritz_fluid", "initializeObservation", "normalize",
           "Observation", "radiativeTransfer", "Particles", "Draws", "DeviceMaterial", "DeviceBins",
           "initializeParticle", "propagateParticleSmallDt", "propagateParticle", "observeTime",
           "forward_histogram"]
