"""
Quantum states needed by the mesh path: hydrogen bound and Coulomb states
(used as keys and analytic initial states) and the numeric eigenstates that
replace them when ``use_numeric_eigenstates`` is on.

Equality, hashing and ordering go through ``tuple`` exactly as in
``ionization/states/state.py:155-175``; a numeric state compares equal to
nothing but another numeric state of the same class with the same tuple.
"""

import enum

import numpy as np
import scipy.special as special

from .. import units as u
from .. import exceptions


class Binding(str, enum.Enum):
    BOUND = "bound"
    FREE = "free"


class Derivation(str, enum.Enum):
    ANALYTIC = "analytic"
    NUMERIC = "numeric"


class QuantumState:
    binding = None
    derivation = Derivation.ANALYTIC

    def __init__(self, amplitude=1):
        self.amplitude = amplitude

    @property
    def numeric(self):
        return self.derivation == Derivation.NUMERIC

    @property
    def analytic(self):
        return self.derivation == Derivation.ANALYTIC

    @property
    def bound(self):
        return self.binding == Binding.BOUND

    @property
    def free(self):
        return self.binding == Binding.FREE

    @property
    def tuple(self):
        raise NotImplementedError

    def __iter__(self):
        yield self

    def __hash__(self):
        return hash((self.__class__.__name__,) + self.tuple)

    def __eq__(self, other):
        return isinstance(other, self.__class__) and self.tuple == other.tuple

    def __lt__(self, other):
        return isinstance(other, self.__class__) and self.tuple < other.tuple

    def __gt__(self, other):
        return isinstance(other, self.__class__) and self.tuple > other.tuple


class SphericalHarmonicState(QuantumState):
    def __init__(self, l=0, m=0, amplitude=1):
        super().__init__(amplitude=amplitude)
        self.l = l
        self.m = m

    @property
    def spherical_harmonic(self):
        return (self.l, self.m)


class HydrogenBoundState(SphericalHarmonicState):
    """Hydrogen bound state |n, l, m> (``ionization/states/three_d.py:298-391``)."""

    binding = Binding.BOUND

    def __init__(self, n=1, l=0, m=0, amplitude=1):
        if int(n) != n or n < 0:
            raise exceptions.IllegalQuantumState(f"n ({n}) must be an integer greater than zero")
        if int(l) != l or not 0 <= l < n:
            raise exceptions.IllegalQuantumState(
                f"l ({l}) must be an integer greater than or equal to zero and less than n ({n})"
            )
        self.n = int(n)
        super().__init__(l=int(l), m=m, amplitude=amplitude)

    @property
    def energy(self):
        return -u.rydberg / (self.n ** 2)

    @property
    def tuple(self):
        return self.n, self.l, self.m

    def radial_function(self, r):
        """R_nl(r) with Griffiths' normalisation (``three_d.py:375-391``)."""
        n, l = self.n, self.l
        normalization = np.sqrt(
            ((2 / (n * u.bohr_radius)) ** 3)
            * (special.factorial(n - l - 1) / (2 * n * special.factorial(n + l)))
        )
        x = 2 * r / (n * u.bohr_radius)
        r_dep = np.exp(-r / (n * u.bohr_radius)) * (x ** l)
        lag_poly = special.eval_genlaguerre(n - l - 1, (2 * l) + 1, x)
        return self.amplitude * normalization * r_dep * lag_poly

    def __call__(self, r, theta, phi):
        """psi(r, theta, phi) = R_nl(r) Y_l^m(theta, phi) (three_d.py:132-205, 298-391)."""
        return self.radial_function(r) * special.sph_harm_y(self.l, self.m, theta, phi)

    def __repr__(self):
        return f"HydrogenBoundState(n={self.n}, l={self.l}, m={self.m})"


class HydrogenCoulombState(SphericalHarmonicState):
    """Free Coulomb state keyed by energy and l (``three_d.py:403-470``); only its key is used here."""

    binding = Binding.FREE

    def __init__(self, energy=1 * u.eV, l=0, m=0, amplitude=1):
        self.energy = energy
        super().__init__(l=l, m=m, amplitude=amplitude)

    @property
    def wavenumber(self):
        return np.sqrt(2 * u.electron_mass * self.energy) / u.hbar

    @property
    def tuple(self):
        return self.wavenumber, self.l, self.m, self.energy

    def __repr__(self):
        return f"HydrogenCoulombState(energy={self.energy / u.eV:.6g} eV, l={self.l})"


class NumericSphericalHarmonicState(SphericalHarmonicState):
    """
    A numeric radial eigenstate on a spherical-harmonic mesh (``three_d.py:543-633``).
    ``g`` holds R(r) (the eigenvector divided by r); ``radial_function`` returns it.
    """

    derivation = Derivation.NUMERIC

    def __init__(self, *, g, l=0, m=0, energy, corresponding_analytic_state, binding, amplitude=1):
        self.g = g
        self.energy = energy
        self.corresponding_analytic_state = corresponding_analytic_state
        self.binding = binding
        super().__init__(l=l, m=m, amplitude=amplitude)

    @property
    def n(self):
        return self.corresponding_analytic_state.n

    @property
    def tuple(self):
        return self.corresponding_analytic_state.tuple

    def radial_function(self, r):
        return self.g

    def __repr__(self):
        return repr(self.corresponding_analytic_state) + "_n"


class NumericOneDState(QuantumState):
    """A numeric eigenstate on a LineMesh (``ionization/states/one_d.py:788-873``)."""

    derivation = Derivation.NUMERIC

    def __init__(self, *, g, energy, corresponding_analytic_state=None, binding=Binding.BOUND, amplitude=1):
        self.g = g
        self.energy = energy
        self.corresponding_analytic_state = corresponding_analytic_state
        self.binding = binding
        super().__init__(amplitude=amplitude)

    @property
    def tuple(self):
        if self.corresponding_analytic_state is not None:
            return self.corresponding_analytic_state.tuple
        return (self.energy,)

    def __call__(self, z):
        return self.g

    def __repr__(self):
        return f"NumericOneDState(energy={self.energy / u.eV:.6g} eV)"


class QHOState(QuantumState):
    """
    Quantum harmonic oscillator eigenstate n of spring constant k and mass m
    (``ionization/states/one_d.py:130-300``): psi_n(x) =
    (m w / (pi hbar))^(1/4) / sqrt(2^n n!) H_n(sqrt(m w / hbar) x) exp(-m w x^2 / (2 hbar)),
    w = sqrt(k / m).
    """

    def __init__(self, spring_constant: float, mass: float = u.electron_mass, n: int = 0, amplitude=1):
        self.n = int(n)
        self.spring_constant = spring_constant
        self.mass = mass
        super().__init__(amplitude=amplitude)

    @classmethod
    def from_omega_and_mass(cls, omega: float, mass: float = u.electron_mass, n: int = 0, amplitude=1):
        return cls(spring_constant=mass * (omega ** 2), mass=mass, n=n, amplitude=amplitude)

    @classmethod
    def from_potential(cls, potential, mass: float, n: int = 0, amplitude=1):
        return cls(spring_constant=potential.spring_constant, mass=mass, n=n, amplitude=amplitude)

    @property
    def omega(self):
        return np.sqrt(self.spring_constant / self.mass)

    @property
    def energy(self):
        return u.hbar * self.omega * (self.n + 0.5)

    @property
    def tuple(self):
        return self.n, self.mass, self.omega

    def __call__(self, x):
        mw = self.mass * self.omega / u.hbar
        norm = (mw / u.pi) ** 0.25 / (np.float64(2.0 ** (self.n / 2)) * np.sqrt(np.float64(special.factorial(self.n))))
        x = np.asarray(x, dtype=np.float64)
        return (norm * np.exp(-mw * x ** 2 / 2) * special.eval_hermite(self.n, np.sqrt(mw) * x)).astype(np.complex128)

    def __repr__(self):
        return f"QHOState(n={self.n}, energy={self.energy / u.eV:.6g} eV)"


class GroundStateLabel(QuantumState):
    """Key for the lowest eigenstate of a line potential with no analytic counterpart."""

    binding = Binding.BOUND

    def __init__(self, index=0):
        super().__init__()
        self.index = index

    @property
    def tuple(self):
        return (self.index,)

    def __repr__(self):
        return f"GroundStateLabel({self.index})"
