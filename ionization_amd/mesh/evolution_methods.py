"""
Evolution methods (``ionization/mesh/evolution_methods.py``) backed by the
HIP kernels.

Each method keeps the reference's strategy seam ``evolve(mesh, g, time_step)
-> g`` and adds ``compile(sim)``, which turns a simulation into the device
table form the engine runs. ``MeshSimulation.run`` uses ``compile`` and runs
the whole time grid on the device. ``evolve`` is the per-step drop-in: called
from the reference's own ``QuantumMesh.evolve`` (meshes.py:251-257), it takes
one step on the device and returns g WITHOUT the mask, which that caller
applies afterwards.

The seam reads only what the reference's objects carry (``mesh.sim``,
``sim.spec``, ``sim.times``, ``sim.time_index``; the mesh and operator kinds
by class name), so it runs under the reference's loop as well as under this
package's mirror of it.
"""

import weakref

import numpy as np

from .. import engine
from . import data


def _class_names(obj):
    return {c.__name__ for c in type(obj).__mro__}


def _mesh_kind(mesh):
    names = _class_names(mesh)
    for name, kind in (("SphericalHarmonicMesh", "sh"), ("LineMesh", "line"), ("CylindricalSliceMesh", "cyl")):
        if name in names:
            return kind
    return None


def _gauge(operators):
    """'LEN' or 'VEL' (core.Gauge values, ionization/core.py:127-129)."""
    g = getattr(operators, "gauge", None)
    return str(getattr(g, "value", g))


class _SeamSession:
    """One simulation stepped through ``evolve``: compiled once without the mask (the caller applies it),
    its tables and step tables resident in a private device context."""

    def __init__(self, method, sim, g):
        self.compiled = method.compile(sim, g0=g, seam=True)
        c = self.compiled
        self.kind = (engine.KIND_SH if isinstance(c, engine.CompiledSH)
                     else engine.KIND_CYL if isinstance(c, engine.CompiledCyl) else engine.KIND_LINE)
        self.ctx = engine.DeviceContext(device=getattr(sim.spec, "device", 0),
                                        pool_bytes=int(c.footprint_bytes * 1.25) + (64 << 20))
        add = {engine.KIND_SH: self.ctx.add_sh, engine.KIND_LINE: self.ctx.add_line,
               engine.KIND_CYL: self.ctx.add_cyl}
        self.slot = add[self.kind](c)
        # every step ends at the exact reference g (cylindrical meshes; a no-op for the others)
        self.ctx.set_hold(self.kind, self.slot, True)

    def step(self, g, n):
        """The reference g after n + 1 steps (unmasked) from the g after n steps."""
        self.ctx.upload_state(self.kind, self.slot, g, n)
        self.ctx.run(1, use_graph=False)
        self.ctx.materialize(self.kind, self.slot)
        if self.kind == engine.KIND_SH:
            return self.ctx.download_sh(self.slot, with_g=True).g
        if self.kind == engine.KIND_CYL:
            return self.ctx.download_cyl(self.slot)[0]
        return self.ctx.download_line(self.slot)[0]

    def close(self):
        self.ctx.close()


class EvolutionMethod:
    """Strategy for advancing a mesh wavefunction (evolution_methods.py:16-43)."""

    def evolve(self, mesh, g, time_step):
        """
        One step of ``g`` by ``time_step`` (evolution_methods.py:19-24), on the device; returns the new g
        without the mask. ``mesh.sim.time_index`` is the index of the time being stepped to (the reference's
        loop advances it before calling, sims.py:317-321), and ``time_step`` must be that grid step, which the
        compiled step tables hold. The first call compiles the simulation; each call uploads ``g`` (the caller
        owns it and masks it between calls) and downloads the result; the device memory is released after the
        last step of the grid.
        """
        sim = mesh.sim
        times = np.asarray(sim.times)
        n = int(sim.time_index) - 1
        if not 0 <= n < len(times) - 1:
            raise ValueError(f"time index {sim.time_index} has no step before it on the time grid")
        grid_step = times[n + 1] - times[n]
        if not np.isclose(time_step, grid_step, rtol=1e-12, atol=0):
            raise ValueError(f"time_step {time_step} differs from the time grid's step {grid_step} at index "
                             f"{n + 1}: the device step takes the grid's own steps")
        sess = self._seam_session(sim, g)
        out = sess.step(np.ascontiguousarray(g, dtype=np.complex128), n)
        if n + 1 == len(times) - 1:
            self._seam_close(sim)
        return out

    def _seam_session(self, sim, g):
        sessions = self.__dict__.setdefault("_seam", {})
        entry = sessions.get(id(sim))
        if entry is not None and entry[0]() is sim:
            return entry[1]
        if entry is not None:
            entry[1].close()
        sess = _SeamSession(self, sim, np.ascontiguousarray(g, dtype=np.complex128))
        try:
            ref = weakref.ref(sim)
        except TypeError:
            ref = lambda sim=sim: sim  # noqa: E731  (objects without weak references are held strongly)
        sessions[id(sim)] = (ref, sess)
        return sess

    def _seam_close(self, sim):
        entry = self.__dict__.get("_seam", {}).pop(id(sim), None)
        if entry is not None:
            entry[1].close()

    def __getstate__(self):
        state = dict(self.__dict__)
        state.pop("_seam", None)
        return state

    def __deepcopy__(self, memo):
        new = self.__class__.__new__(self.__class__)
        new.__dict__.update({k: v for k, v in self.__dict__.items() if k != "_seam"})
        return new

    def compile(self, sim, g0=None, seam=False):
        raise NotImplementedError

    def __repr__(self):
        return f"{self.__class__.__name__}()"

    def info(self):
        return {"evolution_method": self.__class__.__name__}


def _field(sim):
    pot = sim.spec.electric_potential
    return lambda t: np.asarray(pot.get_electric_field_amplitude(t), dtype=np.float64)


def _mask(sim, seam=False):
    if seam:
        return lambda r: np.ones_like(np.asarray(r, dtype=np.float64))
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

    def compile(self, sim, g0=None, seam=False):
        kind = _mesh_kind(sim.mesh)
        if kind == "line":
            # LineMesh: the interaction split around a CN step in K + V (mesh_operators.py:329-342, 352-427)
            velocity = "LineVelocityGaugeOperators" in _class_names(sim.spec.operators)
            return _compile_line(sim, g0, method="sio", gauge="velocity" if velocity else "length", seam=seam)
        if kind != "sh":
            raise NotImplementedError("SplitInteractionOperator runs on SphericalHarmonicMesh and LineMesh")
        gauge = _gauge(sim.spec.operators)
        if gauge not in ("LEN", "VEL"):
            raise NotImplementedError(f"unknown gauge {gauge}")
        return _compile_sh(sim, g0, gauge="velocity" if gauge == "VEL" else "length", method="sio", seam=seam)


def _compile_sh(sim, g0, *, gauge, method, seam=False):
    mesh = sim.mesh
    basis = []
    for state in ([] if seam else sim.spec.test_states):
        comps = list(state)
        if len(comps) != 1:
            raise NotImplementedError("test states must have one spherical-harmonic component")
        s = comps[0]
        basis.append((np.real(mesh.get_radial_g_for_state(s)), s.l))
    return engine.compile_sh(
        r_bound=sim.spec.r_bound, r_points=sim.spec.r_points, l_bound=sim.spec.l_bound,
        potential=_potential(sim), field=_field(sim), mask=_mask(sim, seam), times=sim.times,
        time_step=sim.spec.time_step, g0=mesh.g if g0 is None else g0, basis=basis,
        store_every=-1 if seam else sim.spec.store_data_every, test_charge=sim.spec.test_charge,
        hydrogen_correction=sim.spec.operators.hydrogen_zero_angular_momentum_correction,
        rows=getattr(sim.spec, "device_rows", None),
        norm_by_l=(not seam) and data.NormBySphericalHarmonic in getattr(sim, "datastores_by_type", {}),
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

    def compile(self, sim, g0=None, seam=False):
        kind = _mesh_kind(sim.mesh)
        gauge = _gauge(sim.spec.operators)
        if kind == "sh":
            if gauge != "LEN":
                raise NotImplementedError("the device SphericalHarmonicMesh ADI step implements the length gauge")
            return _compile_sh(sim, g0, gauge="length", method="adi", seam=seam)
        if kind == "line":
            if gauge != "LEN" or "LineVelocityGaugeOperators" in _class_names(sim.spec.operators):
                raise NotImplementedError("the device LineMesh ADI step implements the length gauge "
                                          "(the velocity gauge runs with SplitInteractionOperator)")
            return _compile_line(sim, g0, method="adi", gauge="length", seam=seam)
        if kind == "cyl":
            if gauge != "LEN":
                raise NotImplementedError("the device CylindricalSliceMesh step implements the length gauge")
            mesh = sim.mesh
            basis = [] if seam else [np.asarray(mesh.get_g_for_state(s), dtype=np.complex128)
                                     for s in sim.spec.test_states]
            return engine.compile_cyl(
                z_bound=sim.spec.z_bound, rho_bound=sim.spec.rho_bound, z_points=sim.spec.z_points,
                rho_points=sim.spec.rho_points, field=_field(sim), times=sim.times,
                g0=mesh.g if g0 is None else g0, basis=basis, store_every=-1 if seam else sim.spec.store_data_every,
                charge=_coulomb_charge(sim.spec.internal_potential), test_mass=sim.spec.test_mass,
                test_charge=sim.spec.test_charge, mask=None if seam else _radial_mask(sim.spec.mask),
            )
        raise NotImplementedError(f"device ADI is not implemented for {sim.mesh.__class__.__name__}")


def _compile_line(sim, g0, *, method, gauge, seam=False):
    mesh = sim.mesh
    basis = [] if seam else [np.asarray(mesh.get_g_for_state(s), dtype=np.complex128) for s in sim.spec.test_states]
    return engine.compile_line(
        z_bound=sim.spec.z_bound, z_points=sim.spec.z_points,
        potential=lambda z: np.asarray(sim.spec.internal_potential(r=z, z=z, test_charge=sim.spec.test_charge),
                                       dtype=np.float64),
        field=_field(sim), mask=_mask(sim, seam), times=sim.times, g0=mesh.g if g0 is None else g0, basis=basis,
        store_every=-1 if seam else sim.spec.store_data_every, test_mass=sim.spec.test_mass,
        test_charge=sim.spec.test_charge, rows=getattr(sim.spec, "device_rows", None), method=method, gauge=gauge,
    )


def _coulomb_charge(pot):
    """The point charge of a Coulomb internal potential: the cylindrical kernels evaluate
    k Q q / r on the fly (potentials/static.py:12-38)."""
    container = getattr(pot, "_container", None)
    if "PotentialEnergySum" in _class_names(pot) and container is not None and len(container) == 1:
        pot = container[0]
    if "CoulombPotential" in _class_names(pot) and "SoftCoulombPotential" not in _class_names(pot):
        return pot.charge
    raise NotImplementedError("the device CylindricalSliceMesh step supports a CoulombPotential internal potential")


def _radial_mask(mask):
    """(inner, outer, smoothness) of a RadialCosineMask, None for NoMask."""
    names = _class_names(mask)
    if "NoMask" in names:
        return None
    if "RadialCosineMask" in names:
        return (mask.inner_radius, mask.outer_radius, mask.smoothness)
    raise NotImplementedError("the device CylindricalSliceMesh step supports NoMask and RadialCosineMask")
