"""sfparticles_amd -- the sfparticles per-particle step on AMD MI355X (gfx950).

Same public API as sfparticles (sfparticles/__init__.py:1-5): ``Particles``,
``SpinParticles``, ``RadiationReactionType``, ``Simulation`` and ``Fields``
(plus the field builders in ``sfparticles_amd.fields``).  The compute runs in
hand-written HIP kernels in ``libsfp_hip.so``; importing the package does not
touch the GPU, the first compute call does (and fails loudly without one).
"""
from .particles import Particles, SpinParticles, RadiationReactionType
from .simulation import Simulation
from .fields import Fields

__all__ = ["Particles", "SpinParticles", "Simulation", "Fields", "RadiationReactionType"]
