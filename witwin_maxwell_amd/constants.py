"""Physical constants (CODATA 2018), matching witwin/maxwell/constants.py:20-34.

The reference imports ``EPSILON_0`` from ``witwin.core``; its docstring fixes it to
CODATA 2018 (witwin/maxwell/constants.py:1-4), which is the value used here.
"""

C_0 = 299792458.0
MU_0 = 1.25663706212e-6
EPSILON_0 = 8.8541878128e-12
# CODATA 2018 recommended value (witwin/maxwell/constants.py:32-34), not MU_0*C_0.
ETA_0 = 376.730313668

__all__ = ["C_0", "MU_0", "EPSILON_0", "ETA_0"]
