"""
Operator strategies of the mesh path (``ionization/mesh/mesh_operators.py``).

In the reference these classes build SciPy sparse ``dia`` matrices every
step. Here they hold the same coefficient formulas but produce what the
device needs: per-simulation coefficient tables built once at compile time
(the device regenerates the time-dependent parts from a per-step field
table). The small helpers (alpha, beta, c_l, single-l Hamiltonians) keep the
reference's names and are also used for the numeric eigenbasis.
"""

import enum

import numpy as np

from .. import core
from .. import units as u


class KineticEnergyDerivation(str, enum.Enum):
    HAMILTONIAN = "hamiltonian"
    LAGRANGIAN = "lagrangian"


class MeshOperators:
    gauge = None

    def __repr__(self):
        return f"{self.__class__.__name__}()"

    def info(self):
        return {"operators": self.__class__.__name__}


class SphericalHarmonicLengthGaugeOperators(MeshOperators):
    """
    Length-gauge operators of a SphericalHarmonicMesh
    (mesh_operators.py:815-1127). Only the Lagrangian kinetic derivation
    (the reference default, :822) is available on the device: the
    Hamiltonian derivation (:869-887) couples neighbouring l-lines through
    the r off-diagonal, which the line-decoupled device solve does not model.
    """

    gauge = core.Gauge.LENGTH

    def __init__(self, kinetic_energy_derivation=KineticEnergyDerivation.LAGRANGIAN,
                 hydrogen_zero_angular_momentum_correction: bool = True):
        kinetic_energy_derivation = KineticEnergyDerivation(kinetic_energy_derivation)
        if kinetic_energy_derivation != KineticEnergyDerivation.LAGRANGIAN:
            raise NotImplementedError(
                "only the Lagrangian kinetic energy derivation is implemented on the device"
            )
        self.kinetic_energy_derivation = kinetic_energy_derivation
        self.hydrogen_zero_angular_momentum_correction = hydrogen_zero_angular_momentum_correction

    def __repr__(self):
        return (f"{self.__class__.__name__}(kinetic_energy_derivation={self.kinetic_energy_derivation.value}, "
                f"hydrogen_zero_angular_momentum_correction={self.hydrogen_zero_angular_momentum_correction})")

    @staticmethod
    def alpha(j):
        """(mesh_operators.py:841-843)"""
        x = (j ** 2) + (2 * j)
        return (x + 1) / (x + 0.75)

    @staticmethod
    def beta(j):
        """(mesh_operators.py:845-847)"""
        x = 2 * (j ** 2) + (2 * j)
        return (x + 1) / (x + 0.5)

    @staticmethod
    def c_l(l):
        """(mesh_operators.py:853-855)"""
        return (l + 1) / np.sqrt(((2 * l) + 1) * ((2 * l) + 3))

    def internal_hamiltonian_for_single_l(self, mesh, l: int):
        """
        Real tridiagonal H_l = K_l + V(r) as (diagonal, off-diagonal) arrays
        (mesh_operators.py:930-986): -2 P beta_j (+ dr (1 + dr)/8 at j = 0 for
        l = 0) + hbar^2 l (l + 1) / (2 mu r^2) + V(r); off-diagonal P alpha_j.
        """
        r = mesh.r
        mu = u.electron_mass_reduced
        pr = -(u.hbar ** 2) / (2 * mu * (mesh.delta_r ** 2))
        eff = ((u.hbar ** 2) / (2 * mu)) * l * (l + 1) / (r ** 2)
        jj = np.arange(len(r), dtype=np.float64)
        r_beta = self.beta(jj)
        if l == 0 and self.hydrogen_zero_angular_momentum_correction:
            dr = mesh.delta_r / u.bohr_radius
            r_beta[0] += dr * (1 + dr) / 8
        diag = (-2 * pr * r_beta) + eff
        off = pr * self.alpha(jj[:-1])
        v = np.asarray(mesh.spec.internal_potential(r=r, test_charge=mesh.spec.test_charge), dtype=np.float64)
        return diag + v * np.ones_like(r), off


class SphericalHarmonicVelocityGaugeOperators(SphericalHarmonicLengthGaugeOperators):
    """
    Velocity-gauge operators of a SphericalHarmonicMesh
    (mesh_operators.py:1130-1408), for SplitInteractionOperator. H0 is the
    length gauge's; the interaction A(t) (h1 + h2) couples neighbouring l at
    fixed r (h1: i hbar (q / m) c_l (l + 1) / r) and neighbouring (l, r)
    together (h2: i hbar q / (2 m dr) kron(C, Alpha)), with the vector
    potential integrated from the field over the time grid so far
    (``engine.vector_potential_table``) and the simulation's test_mass.
    """

    gauge = core.Gauge.VELOCITY

    def __init__(self, hydrogen_zero_angular_momentum_correction: bool = True):
        super().__init__(hydrogen_zero_angular_momentum_correction=hydrogen_zero_angular_momentum_correction)

    def __repr__(self):
        return (f"{self.__class__.__name__}("
                f"hydrogen_zero_angular_momentum_correction={self.hydrogen_zero_angular_momentum_correction})")


class LineLengthGaugeOperators(MeshOperators):
    """Length-gauge operators of a LineMesh (mesh_operators.py:304-349)."""

    gauge = core.Gauge.LENGTH

    def internal_hamiltonian(self, mesh):
        """(diagonal, off-diagonal) of K + V on the z grid (mesh_operators.py:309-316, 244-269)."""
        prefactor = -(u.hbar ** 2) / (2 * mesh.spec.test_mass * (mesh.delta_z ** 2))
        n = len(mesh.z_mesh)
        v = np.asarray(mesh.spec.internal_potential(r=mesh.z_mesh, z=mesh.z_mesh, test_charge=mesh.spec.test_charge),
                       dtype=np.float64)
        return (-2 * prefactor) * np.ones(n) + v * np.ones(n), prefactor * np.ones(n - 1)


class LineVelocityGaugeOperators(LineLengthGaugeOperators):
    """
    Velocity-gauge interaction of a LineMesh (mesh_operators.py:352-427), for
    SplitInteractionOperator: A(t) i hbar (q / m) / (2 dz) above the
    diagonal and its negative below, split into rotations of the even and odd
    z-pairs. The reference labels this class with the length gauge
    (mesh_operators.py:355); the attribute is kept as it is there and the
    device path recognises the class itself.
    """

    gauge = core.Gauge.LENGTH


class CylindricalSliceLengthGaugeOperators(MeshOperators):
    """Length-gauge operators of a CylindricalSliceMesh (mesh_operators.py:550-631)."""

    gauge = core.Gauge.LENGTH

    @staticmethod
    def rho_c(j):
        """rho off-diagonal factor j / sqrt(j^2 - 1/4) (mesh_operators.py:578-579)."""
        return j / np.sqrt((j ** 2) - 0.25)
