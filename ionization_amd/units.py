"""
SI unit table used by every host-side computation of this package.

The reference expresses all quantities in SI through ``simulacra.units``
(imported throughout ``ionization/mesh/*.py`` and ``ionization/potentials/*.py``).
That package is not available here, so this module fixes ONE table of values,
taken from ``scipy.constants`` (CODATA 2022), and defines the same names the
reference uses. Both the device path and the CPU oracle read these values, so
they agree with each other exactly; their agreement with simulacra's own
numeric constants is not pinned by any reference test.
"""

import math

import scipy.constants as _c

_pc = _c.physical_constants

pi = math.pi
twopi = 2 * math.pi

# fundamental
hbar = _pc["reduced Planck constant"][0]
c = _c.c
epsilon_0 = _pc["vacuum electric permittivity"][0]
coulomb_constant = 1 / (4 * pi * epsilon_0)
proton_charge = _c.e
electron_charge = -_c.e
electron_mass = _pc["electron mass"][0]
proton_mass = _pc["proton mass"][0]
electron_mass_reduced = electron_mass * proton_mass / (electron_mass + proton_mass)

# lengths / times / energies
m = 1.0
nm = 1e-9
bohr_radius = _pc["Bohr radius"][0]
s = 1.0
fsec = 1e-15
asec = 1e-18
atomic_time = _pc["atomic unit of time"][0]
J = 1.0
eV = _c.e
rydberg = _pc["Rydberg constant times hc in J"][0]
hartree = 2 * rydberg
Hz = 1.0
THz = 1e12
N = 1.0

# fields / fluence
V_per_m = 1.0
atomic_electric_field = _pc["atomic unit of electric field"][0]
Jcm2 = 1e4  # J / cm^2 expressed in J / m^2
