"""
Evolution methods (``ionization/mesh/evolution_methods.py``) backed by the
HIP kernels.

Each method keeps the reference's strategy seam ``evolve(mesh, g, time_step)
-> g`` and adds ``compile(sim)``, which turns a simulation into the device
table form the engine runs. ``MeshSimulation.run`` uses ``compile`` and runs
the whole time grid on the device; a bare ``evolve`` call does one reference
step through the same kernels.
"""

import numpy as np

from .. import core, engine, potentials
from .. import units as u
from . import data, meshes, mesh_operators


class EvolutionMethod:
    """Strategy for advancing a mesh wavefunction (evolution_methods.py:16-43)."""

    def evolve(self, mesh, g, time_step):
        """One step of ``g`` by ``time_step`` on the device; returns the new host g."""
        sim = mesh.sim
        mesh.g = np.asarray(g, dtype=np.complex128)
        sim._device_single_step(time_step)
        return mesh.g

    def compile(self, sim, g0=None):
        raise NotImplementedError

    def __repr__(self):
        return f"{self.__class__.__name__}()"

    def info(self):
        return {"evolution_method": self.__class__.__name__}


def _field(sim):
    pot = sim.spec.electric_potential
    return lambda t: np.asarray(pot.get_electric_field_amplitude(t), dtype=np.float64)


def _mask(sim):
    mask = sim.spec.mask
    return lambda r: np.asarray(mask(r=r), dtype=np.float64)


def _potential(sim):
    pot = sim.spec.internal_potential
    return lambda r: np.asarray(pot(r=r, test_charge=sim.spec.test_charge), dtype=np.float64)


class SplitInteractionOperator(EvolutionMethod):
    """
    Split-operator step for a SphericalHarmonicMesh (evolution_methods.py:80-123):
    the split interaction factors, CN explicit, CN implicit, the factors in
    reverse, then the mask. Length gauge: E, O, the 2x2 rotations on the
    l-flattened pairs. Velocity gauge: the h1 rotations on the same pairs, then
    the four h2 similarity factors (mesh_operators.py:1188-1408).
    """

    def compile(self, sim, g0=None):
        mesh = sim.mesh
        if isinstance(mesh, meshes.LineMesh):
            # LineMesh: the interaction split around a CN step in K + V (mesh_operators.py:329-342, 352-427)
            velocity = isinstance(sim.spec.operators, mesh_operators.LineVelocityGaugeOperators)
            return _compile_line(sim, g0, method="sio", gauge="velocity" if velocity else "length")
        if not isinstance(mesh, meshes.SphericalHarmonicMesh):
            raise NotImplementedError("SplitInteractionOperator runs on SphericalHarmonicMesh and LineMesh")
        gauge = sim.spec.operators.gauge
        if gauge not in (core.Gauge.LENGTH, core.Gauge.VELOCITY):
            raise NotImplementedError(f"unknown gauge {gauge}")
        return _compile_sh(sim, g0, gauge="velocity" if gauge == core.Gauge.VELOCITY else "length", method="sio")


def _compile_sh(sim, g0, *, gauge, method):
    mesh = sim.mesh
    basis = []
    for state in sim.spec.test_states:
        comps = list(state)
        if len(comps) != 1:
            raise NotImplementedError("test states must have one spherical-harmonic component")
        s = comps[0]
        basis.append((np.real(mesh.get_radial_g_for_state(s)), s.l))
    return engine.compile_sh(
        r_bound=sim.spec.r_bound, r_points=sim.spec.r_points, l_bound=sim.spec.l_bound,
        potential=_potential(sim), field=_field(sim), mask=_mask(sim), times=sim.times,
        time_step=sim.spec.time_step, g0=mesh.g if g0 is None else g0, basis=basis,
        store_every=sim.spec.store_data_every, test_charge=sim.spec.test_charge,
        hydrogen_correction=sim.spec.operators.hydrogen_zero_angular_momentum_correction,
        rows=getattr(sim.spec, "device_rows", None),
        norm_by_l=data.NormBySphericalHarmonic in getattr(sim, "datastores_by_type", {}),
        gauge=gauge, test_mass=sim.spec.test_mass, method=method,
    )


class AlternatingDirectionImplicit(EvolutionMethod):
    """
    Crank-Nicolson / ADI step (evolution_methods.py:46-77). On a LineMesh the
    total Hamiltonian is one operator, so a step is (1 + i tau H)^-1 (1 - i tau H)
    followed by the mask; on a CylindricalSliceMesh it is explicit rho, implicit
    z, explicit z, implicit rho, then the mask; on a SphericalHarmonicMesh
    explicit H0 along r, implicit then explicit E H_int along l, implicit H0,
    then the mask (mesh_operators.py:1020-1035).
    """

    def compile(self, sim, g0=None):
        mesh = sim.mesh
        if isinstance(mesh, meshes.SphericalHarmonicMesh):
            if sim.spec.operators.gauge != core.Gauge.LENGTH:
                raise NotImplementedError("the device SphericalHarmonicMesh ADI step implements the length gauge")
            return _compile_sh(sim, g0, gauge="length", method="adi")
        if isinstance(mesh, meshes.LineMesh):
            if sim.spec.operators.gauge != core.Gauge.LENGTH or isinstance(
                    sim.spec.operators, mesh_operators.LineVelocityGaugeOperators):
                raise NotImplementedError("the device LineMesh ADI step implements the length gauge "
                                          "(the velocity gauge runs with SplitInteractionOperator)")
            return _compile_line(sim, g0, method="adi", gauge="length")
        if isinstance(mesh, meshes.CylindricalSliceMesh):
            if sim.spec.operators.gauge != core.Gauge.LENGTH:
                raise NotImplementedError("the device CylindricalSliceMesh step implements the length gauge")
            basis = [np.asarray(mesh.get_g_for_state(s), dtype=np.complex128) for s in sim.spec.test_states]
            return engine.compile_cyl(
                z_bound=sim.spec.z_bound, rho_bound=sim.spec.rho_bound, z_points=sim.spec.z_points,
                rho_points=sim.spec.rho_points, field=_field(sim), times=sim.times,
                g0=mesh.g if g0 is None else g0, basis=basis, store_every=sim.spec.store_data_every,
                charge=_coulomb_charge(sim.spec.internal_potential), test_mass=sim.spec.test_mass,
                test_charge=sim.spec.test_charge, mask=_radial_mask(sim.spec.mask),
            )
        raise NotImplementedError(f"device ADI is not implemented for {mesh.__class__.__name__}")


def _compile_line(sim, g0, *, method, gauge):
    mesh = sim.mesh
    basis = [np.asarray(mesh.get_g_for_state(s), dtype=np.complex128) for s in sim.spec.test_states]
    return engine.compile_line(
        z_bound=sim.spec.z_bound, z_points=sim.spec.z_points,
        potential=lambda z: np.asarray(sim.spec.internal_potential(r=z, z=z, test_charge=sim.spec.test_charge),
                                       dtype=np.float64),
        field=_field(sim), mask=_mask(sim), times=sim.times, g0=mesh.g if g0 is None else g0, basis=basis,
        store_every=sim.spec.store_data_every, test_mass=sim.spec.test_mass,
        test_charge=sim.spec.test_charge, rows=getattr(sim.spec, "device_rows", None), method=method, gauge=gauge,
    )


def _coulomb_charge(pot):
    """The point charge of a Coulomb internal potential: the cylindrical kernels evaluate
    k Q q / r on the fly (potentials/static.py:12-38)."""
    if isinstance(pot, potentials.PotentialEnergySum) and len(pot._container) == 1:
        pot = pot._container[0]
    if isinstance(pot, potentials.CoulombPotential):
        return pot.charge
    raise NotImplementedError("the device CylindricalSliceMesh step supports a CoulombPotential internal potential")


def _radial_mask(mask):
    """(inner, outer, smoothness) of a RadialCosineMask, None for NoMask."""
    if isinstance(mask, potentials.NoMask):
        return None
    if isinstance(mask, potentials.RadialCosineMask):
        return (mask.inner_radius, mask.outer_radius, mask.smoothness)
    raise NotImplementedError("the device CylindricalSliceMesh step supports NoMask and RadialCosineMask")
