"""
Meshes (``ionization/mesh/meshes.py``): grids, wavefunction storage and
observables.

``mesh.g`` is a host (numpy) copy of the wavefunction in the reference's
layout. During a run the authoritative copy lives on the device; the run
loop refreshes ``mesh.g`` whenever the wavefunction is materialised (per-step
callbacks, checkpoints, the end of the run). Host-side observables
(``norm``, ``inner_product``, ``norm_by_l``) follow the reference formulas on
that copy; the datastores of a run are filled from device reductions.
"""

import enum
import logging
from typing import Dict, Optional

import numpy as np
import scipy.linalg as sl

from .. import exceptions, states
from .. import units as u

logger = logging.getLogger(__name__)


class WrappingDirection(str, enum.Enum):
    Z = "z"
    RHO = "rho"
    R = "r"
    L = "l"


class QuantumMesh:
    """Common mesh behaviour (meshes.py:55-261)."""

    def __init__(self, sim):
        self.sim = sim
        self.spec = sim.spec
        self.operators = self.spec.operators
        self.g = None
        self.inner_product_multiplier = None

    def __str__(self):
        return f"{self.__class__.__name__} for {self.sim}"

    def flatten_mesh(self, mesh, flatten_along):
        order = self.wrapping_direction_to_order(flatten_along)
        return mesh if order is None else mesh.flatten(order)

    def wrap_vector(self, vector, wrap_along):
        order = self.wrapping_direction_to_order(wrap_along)
        return vector if order is None else np.reshape(vector, self.mesh_shape, order)

    def wrapping_direction_to_order(self, wrapping_direction):
        return None

    def state_to_g(self, state_or_mesh):
        if state_or_mesh is None:
            return self.g
        if isinstance(state_or_mesh, states.QuantumState):
            try:
                state_or_mesh = self.analytic_to_numeric[state_or_mesh]
            except (AttributeError, KeyError):
                pass
            return self.get_g_for_state(state_or_mesh)
        return state_or_mesh

    def inner_product(self, a=None, b=None) -> complex:
        """sum conj(a) b * weight (meshes.py:195-200)."""
        return np.sum(np.conj(self.state_to_g(a)) * self.state_to_g(b)) * self.inner_product_multiplier

    def state_overlap(self, a=None, b=None) -> float:
        return np.abs(self.inner_product(a, b)) ** 2

    def norm(self, state=None) -> float:
        """inner_product(g, g).real (meshes.py:206-217)."""
        g = self.state_to_g(state)
        return self.inner_product(g, g).real

    @property
    def g2(self):
        return np.abs(self.g) ** 2

    def evolve(self, time_step):
        """One step by the evolution method, then the mask (meshes.py:251-257)."""
        self.g = self.spec.evolution_method.evolve(self, self.g, time_step)
        self.g *= self.spec.mask(r=self.r_mesh)


class LineMesh(QuantumMesh):
    """Wavefunction on a line (meshes.py:263-341); r and r_mesh are the signed z."""

    def __init__(self, sim):
        super().__init__(sim)
        self.z_mesh = np.linspace(-self.spec.z_bound, self.spec.z_bound, self.spec.z_points)
        self.delta_z = np.abs(self.z_mesh[1] - self.z_mesh[0])
        self.z_center_index = int(np.argmin(np.abs(self.z_mesh)))
        self.mesh_points = len(self.z_mesh)
        self.mesh_shape = (self.mesh_points,)
        self.inner_product_multiplier = self.delta_z
        self.g_factor = 1
        if self.spec.use_numeric_eigenstates:
            self.analytic_to_numeric = self._get_numeric_eigenstate_basis(self.spec.number_of_numeric_eigenstates)
            self.spec.test_states = sorted(self.analytic_to_numeric.values(), key=lambda x: x.energy)
            if not self.spec.initial_state.numeric:
                self.spec.initial_state = self.analytic_to_numeric[self.spec.initial_state]
        self.g = self.get_g_for_state(self.spec.initial_state)

    @property
    def z(self):
        return self.z_mesh

    @property
    def r(self):
        return self.z_mesh

    @property
    def r_mesh(self):
        return self.z_mesh

    def get_g_for_state(self, state):
        """state(z) normalised on the mesh, times the amplitude (meshes.py:327-341)."""
        g = np.array(state(self.z_mesh), dtype=np.complex128)
        g /= np.sqrt(self.norm(g))
        g *= state.amplitude
        return g

    def _get_numeric_eigenstate_basis(self, number_of_eigenstates: int):
        """Lowest eigenstates of K + V (meshes.py:343-397), keyed by GroundStateLabel(i)."""
        d, e = self.operators.internal_hamiltonian(self)
        w, v = sl.eigh_tridiagonal(d, e, select="i", select_range=(0, number_of_eigenstates - 1))
        out = {}
        for i, (val, vec) in enumerate(zip(w, v.T)):
            vec = vec / np.sqrt(self.inner_product_multiplier * np.sum(np.abs(vec) ** 2))
            key = states.GroundStateLabel(i)
            out[key] = states.NumericOneDState(g=vec, energy=val, corresponding_analytic_state=key,
                                               binding=states.Binding.BOUND if val < 0 else states.Binding.FREE)
        return out


class SphericalHarmonicMesh(QuantumMesh):
    """Radial distance x orbital angular momentum mesh (meshes.py:985-1136)."""

    def __init__(self, sim):
        super().__init__(sim)
        self.r = np.linspace(0, self.spec.r_bound, self.spec.r_points)
        self.delta_r = self.r[1] - self.r[0]
        self.r += self.delta_r / 2
        self.r_max = np.max(self.r)
        self.inner_product_multiplier = self.delta_r
        self.l = np.array(range(self.spec.l_bound), dtype=int)
        self.mesh_points = len(self.r) * len(self.l)
        self.mesh_shape = (len(self.l), len(self.r))
        self._radial_g_cache: Dict = {}
        if self.spec.use_numeric_eigenstates:
            self.analytic_to_numeric = self.get_numeric_eigenstate_basis(
                self.spec.numeric_eigenstate_max_energy, self.spec.numeric_eigenstate_max_angular_momentum)
            self.spec.test_states = sorted(self.analytic_to_numeric.values(), key=lambda x: x.energy)
            if not self.spec.initial_state.numeric:
                self.spec.initial_state = self.analytic_to_numeric[self.spec.initial_state]
        self.g = self.get_g_for_state(self.spec.initial_state)

    @property
    def r_mesh(self):
        return np.meshgrid(self.l, self.r, indexing="ij")[1]

    @property
    def l_mesh(self):
        return np.meshgrid(self.l, self.r, indexing="ij")[0]

    @property
    def g_factor(self):
        return self.r

    def wrapping_direction_to_order(self, wrapping_direction):
        if wrapping_direction is None:
            return None
        if wrapping_direction == WrappingDirection.L:
            return "F"
        if wrapping_direction == WrappingDirection.R:
            return "C"
        raise ValueError(f"{wrapping_direction} is not a valid specifier for flatten_mesh (valid specifiers: 'l', 'r')")

    def get_g_for_state(self, state):
        """(meshes.py:1066-1088)"""
        if isinstance(state, states.QuantumState) and all(hasattr(s, "spherical_harmonic") for s in state):
            if state.analytic and self.spec.use_numeric_eigenstates:
                state = self.analytic_to_numeric.get(state, state)
            g = np.zeros(self.mesh_shape, dtype=np.complex128)
            for s in state:
                g[s.l, :] += self.get_radial_g_for_state(s)
            return g
        raise NotImplementedError(
            "States with non-definite angular momentum components are not currently supported by SphericalHarmonicMesh")

    def get_radial_g_for_state(self, state):
        """R(r) r normalised on the radial grid (meshes.py:1090-1097)."""
        key = (state.__class__.__name__,) + tuple(state.tuple)
        if key not in self._radial_g_cache:
            g = state.radial_function(self.r) * self.g_factor
            g = g / np.sqrt(self.norm(g))
            g = g * state.amplitude
            self._radial_g_cache[key] = g
        return self._radial_g_cache[key]

    def inner_product(self, a=None, b=None) -> complex:
        """Shortcut for a spherical-harmonic state against the current g (meshes.py:1099-1131)."""
        if isinstance(a, states.QuantumState) and all(hasattr(s, "spherical_harmonic") for s in a) and b is None:
            ip = 0
            for s in a:
                ip += np.sum(np.conj(self.get_radial_g_for_state(s)) * self.g[s.l, :])
            return ip * self.inner_product_multiplier
        return super().inner_product(a, b)

    def norm_by_l(self, state=None):
        """Norm in each l (meshes.py:1133-1136)."""
        g = self.state_to_g(state)
        return np.abs(np.sum(np.conj(g) * g, axis=1) * self.delta_r)

    def get_numeric_eigenstate_basis(self, max_energy: float, max_angular_momentum: int):
        """
        Numeric eigenstates of every H_l, l <= max_angular_momentum, with energy
        <= max_energy (meshes.py:1281-1356). The reference grows an ARPACK
        ``eigsh(k, which="SA")`` request until the largest eigenvalue exceeds
        max_energy and keeps those below it (at most n - 2); the same set is
        taken here directly with LAPACK's ``eigh_tridiagonal`` on the real
        tridiagonal H_l. Keys: HydrogenBoundState(n_guess, l) for E <= 0 with
        n_guess = max(1, int(sqrt(Ry / |E|))) (a later collision replaces an
        earlier one), HydrogenCoulombState(E, l) for E > 0.
        """
        analytic_to_numeric = {}
        for l in range(max_angular_momentum + 1):
            d, e = self.operators.internal_hamiltonian_for_single_l(self, l)
            n = len(d)
            w, v = sl.eigh_tridiagonal(d, e, select="v", select_range=(-np.inf, max_energy))
            if len(w) > n - 2:
                w, v = w[: n - 2], v[:, : n - 2]
            for eigenvalue, eigenvector in zip(w, v.T):
                eigenvector = eigenvector / np.sqrt(self.inner_product_multiplier * np.sum(np.abs(eigenvector) ** 2))
                eigenvector = eigenvector / self.g_factor
                if eigenvalue > max_energy:
                    continue
                if eigenvalue > 0:
                    analytic_state = states.HydrogenCoulombState(energy=eigenvalue, l=l)
                    binding = states.Binding.FREE
                else:
                    n_guess = int(np.sqrt(u.rydberg / np.abs(eigenvalue)))
                    if n_guess == 0:
                        n_guess = 1
                    analytic_state = states.HydrogenBoundState(n=n_guess, l=l)
                    binding = states.Binding.BOUND
                analytic_to_numeric[analytic_state] = states.NumericSphericalHarmonicState(
                    g=eigenvector, l=l, m=0, energy=eigenvalue, corresponding_analytic_state=analytic_state,
                    binding=binding)
        return analytic_to_numeric


class CylindricalSliceMesh(QuantumMesh):
    """(z, rho) slice of an azimuthally symmetric cylinder (meshes.py:509-599)."""

    def __init__(self, sim):
        super().__init__(sim)
        self.z = np.linspace(-self.spec.z_bound, self.spec.z_bound, self.spec.z_points)
        self.rho = np.delete(np.linspace(0, self.spec.rho_bound, self.spec.rho_points + 1), 0)
        self.delta_z = self.z[1] - self.z[0]
        self.delta_rho = self.rho[1] - self.rho[0]
        self.inner_product_multiplier = self.delta_z * self.delta_rho
        self.rho -= self.delta_rho / 2
        self.z_center_index = int(self.spec.z_points // 2)
        self.mesh_points = len(self.z) * len(self.rho)
        self.mesh_shape = (len(self.z), len(self.rho))
        self.g = self.get_g_for_state(self.spec.initial_state)

    @property
    def z_mesh(self):
        return np.meshgrid(self.z, self.rho, indexing="ij")[0]

    @property
    def rho_mesh(self):
        return np.meshgrid(self.z, self.rho, indexing="ij")[1]

    @property
    def r_mesh(self):
        return np.sqrt((self.z_mesh ** 2) + (self.rho_mesh ** 2))

    @property
    def theta_mesh(self):
        return np.arccos(self.z_mesh / self.r_mesh)

    @property
    def g_factor(self):
        return np.sqrt(u.twopi * self.rho_mesh)

    def wrapping_direction_to_order(self, wrapping_direction):
        if wrapping_direction is None:
            return None
        if wrapping_direction == WrappingDirection.Z:
            return "F"
        if wrapping_direction == WrappingDirection.RHO:
            return "C"
        raise ValueError(f"{wrapping_direction} is not a valid specifier for flatten_mesh (valid specifiers: 'z', 'rho')")

    def get_g_for_state(self, state):
        """sqrt(2 pi rho) psi(r, theta) normalised (meshes.py:601-607)."""
        if isinstance(state, np.ndarray):
            return state
        g = self.g_factor * state(self.r_mesh, self.theta_mesh, 0)
        g = g / np.sqrt(self.norm(g))
        g = g * state.amplitude
        return g.astype(np.complex128)
