"""Physical constants (same definitions as the reference's src/fdtdx/constants.py:13-22)."""

import math

c: float = 299792458.0
mu0: float = 4e-7 * math.pi
eps0: float = 1.0 / (mu0 * c**2)
eta0: float = mu0 * c

relative_permittivity_air: float = 1.0
relative_permittivity_substrate: float = 2.1025
relative_permittivity_polymer: float = 2.368521
relative_permittivity_silicon: float = 12.25
relative_permittivity_silica: float = 2.25
