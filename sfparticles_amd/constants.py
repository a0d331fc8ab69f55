"""Physical constants used by the particle step.

Pinned to CODATA 2018 -- the values ``scipy.constants`` held from scipy 1.4 to
1.14, which the reference (sfparticles/particles.py:5,
sfparticles/qed/*_tables.py) picks up at import time.  Pinning keeps every
result independent of the installed scipy (1.15 moved to CODATA 2022, a
~1e-9 relative shift in m_e, epsilon_0 and alpha).  The same literals are
compiled into the HIP library (csrc/sfp_device.h).
"""
import math

c = 299792458.0
m_e = 9.1093837015e-31
e = 1.602176634e-19
hbar = 1.0545718176461565e-34
epsilon_0 = 8.8541878128e-12
alpha = 0.0072973525693
pi = math.pi

__all__ = ["c", "m_e", "e", "hbar", "epsilon_0", "alpha", "pi"]
