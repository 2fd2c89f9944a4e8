"""
Specifications and simulations (``ionization/mesh/sims.py``).

``MeshSpecification(...).to_sim().run()`` keeps the reference's call shape.
``run`` compiles the simulation into device tables, runs the whole time grid
on the GPU (one captured hipGraph replay per step), and fills the
datastores from device reductions. With a per-step ``callback`` the loop
materialises ``mesh.g`` before every step, as the reference's loop exposes it
(sims.py:289-321). Checkpoints download the wavefunction and pickle the
simulation; ``MeshSimulation.load`` + ``run`` resumes on the device.
"""

import collections
import datetime
import gzip
import logging
import os
import pickle
from copy import deepcopy
from typing import Callable, Optional

import numpy as np

from .. import engine, exceptions, potentials, states
from .. import units as u
from ..core import SphericalHarmonic, WarningRecord
from . import data, evolution_methods, mesh_operators, meshes

logger = logging.getLogger(__name__)


class _DeviceSession:
    """A simulation's slot in a private device context."""

    def __init__(self, sim, device=0):
        method = sim.spec.evolution_method
        self.compiled = method.compile(sim)
        if isinstance(self.compiled, engine.CompiledSH):
            self.kind = engine.KIND_SH
        elif isinstance(self.compiled, engine.CompiledCyl):
            self.kind = engine.KIND_CYL
        else:
            self.kind = engine.KIND_LINE
        pool = int(self.compiled.footprint_bytes * 1.25) + (64 << 20)
        self.ctx = engine.DeviceContext(device=device, pool_bytes=pool)
        add = {engine.KIND_SH: self.ctx.add_sh, engine.KIND_LINE: self.ctx.add_line, engine.KIND_CYL: self.ctx.add_cyl}
        self.slot = add[self.kind](self.compiled)
        if sim.time_index > 0:
            self.ctx.upload_state(self.kind, self.slot, sim.mesh.g, sim.time_index)

    def hold(self, on: bool):
        """End later steps at exact states (cylindrical meshes; a no-op for the others)."""
        self.ctx.set_hold(self.kind, self.slot, on)

    def step_index(self):
        return self.ctx.state(self.kind, self.slot)[0]

    def g(self):
        self.ctx.materialize(self.kind, self.slot)
        if self.kind == engine.KIND_SH:
            return self.ctx.download_sh(self.slot, with_g=True).g
        if self.kind == engine.KIND_CYL:
            return self.ctx.download_cyl(self.slot)[0]
        return self.ctx.download_line(self.slot)[0]

    def flags(self):
        return int(self.ctx.flags(self.kind, self.slot)[0])

    def data(self):
        if self.kind == engine.KIND_SH:
            res = self.ctx.download_sh(self.slot, with_g=False)
            return {"norms": res.norms, "lnorms": res.lnorms, "ips": res.ips, "norm_by_l": res.norm_by_l}
        if self.kind == engine.KIND_CYL:
            _, norms, ips = self.ctx.download_cyl(self.slot, with_g=False)
            return {"norms": norms, "lnorms": None, "ips": ips}
        _, norms, ips = self.ctx.download_line(self.slot)
        return {"norms": norms, "lnorms": None, "ips": ips}

    def close(self):
        self.ctx.close()


class MeshSimulation:
    """A TDSE simulation on a mesh (sims.py:36-438)."""

    def __init__(self, spec):
        self.spec = spec
        self.name = spec.name
        self.latest_checkpoint_time = datetime.datetime.utcnow()
        self.times = self.get_times()
        self.time_index = 0
        self.data_time_index = 0
        self.time_steps = len(self.times)
        self.mesh = self.spec.mesh_type(self)
        time_indices = np.arange(self.time_steps)
        self.data_mask = (time_indices == 0) | (time_indices == self.time_steps - 1)
        if self.spec.store_data_every >= 1:
            self.data_mask |= time_indices % self.spec.store_data_every == 0
        self.data_times = self.times[self.data_mask]
        self.data_indices = time_indices[self.data_mask]
        self.data_time_steps = len(self.data_times)
        self.spacetime_points = self.time_steps * int(np.prod(self.mesh.g.shape))
        self.data = data.Data(self)
        self.datastores_by_type = data.clone(self.spec.datastores)
        for ds in self.datastores_by_type.values():
            ds.init(self)
        self.warnings = collections.defaultdict(list)
        self.lnorm = self.get_blank_data()
        self._session = None
        self.status = "initialized"

    def __str__(self):
        return f"{self.__class__.__name__}({self.name})"

    def __repr__(self):
        return str(self)

    def get_blank_data(self, dtype=np.float64):
        a = np.empty(self.data_time_steps, dtype=dtype)
        a.fill(np.nan)
        return a

    @property
    def time(self):
        return self.times[self.time_index]

    @property
    def times_to_current(self):
        return self.times[: self.time_index + 1]

    def get_times(self):
        """linspace(t0, tf, int((tf - t0) / dt) + 1) (sims.py:198-205)."""
        total_time = self.spec.time_final - self.spec.time_initial
        return np.linspace(self.spec.time_initial, self.spec.time_final, int(total_time / self.spec.time_step) + 1)

    @property
    def bound_states(self):
        yield from (s for s in self.spec.test_states if s.bound)

    @property
    def free_states(self):
        yield from (s for s in self.spec.test_states if not s.bound)

    # ---------------------------------------------------------- device plumbing

    def _session_open(self):
        if self._session is None:
            self._session = _DeviceSession(self, device=self.spec.device)
        return self._session

    def _session_close(self):
        if self._session is not None:
            self._session.close()
            self._session = None

    def _collect(self, first_time_index=0):
        sess = self._session
        dev = sess.data()
        slots = np.nonzero(self.data_indices >= first_time_index)[0]
        for ds in self.datastores_by_type.values():
            ds.collect(slots, dev)
        if dev["lnorms"] is not None:
            self.lnorm[slots] = dev["lnorms"][slots]

    def _store_host_data(self):
        """The datastores' values at the current data index from the host copy of g (sims.py:222-226)."""
        for ds in self.datastores_by_type.values():
            ds.store_host()
        if self.mesh.g.ndim == 2 and isinstance(self.mesh, meshes.SphericalHarmonicMesh):
            self.lnorm[self.data_time_index] = self.mesh.norm(self.mesh.g[-1])

    def _check_data_times(self, first_time_index=0):
        """check() at every stored data time from first_time_index on, with time_index and data_time_index
        set as the reference's loop has them when it calls check (sims.py:290-294)."""
        time_index, data_time_index = self.time_index, self.data_time_index
        try:
            for i in np.nonzero(self.data_indices >= first_time_index)[0]:
                self.data_time_index = int(i)
                self.time_index = int(self.data_indices[i])
                self.check()
        finally:
            self.time_index, self.data_time_index = time_index, data_time_index

    def _check_device_health(self, flags):
        """Record the device's health flags (a non-finite norm: the simulation diverged)."""
        if flags & engine.FLAG_NONFINITE and not self.warnings.get("non_finite"):
            msg = f"Wavefunction norm is not finite for {self}: the evolution diverged"
            logger.error(msg)
            self.warnings["non_finite"].append(WarningRecord(int(self.time_index), msg))

    # ---------------------------------------------------------- run loop

    def check(self):
        """Norm-growth warnings at the current data index (sims.py:228-240): above the initial norm or
        the previous stored norm by more than 0.1 %. At data index 0 the previous norm is the last,
        still blank (NaN) entry, which compares false, as in the reference."""
        if data.Norm not in self.datastores_by_type:
            return
        norms = self.data.norm
        i = self.data_time_index
        norm = norms[i]
        prev = norms[i - 1] if i > 0 else np.nan  # the reference reads norms[-1], still blank at index 0
        if norm > 1.001 * norms[0]:
            msg = f"Wavefunction norm ({norm}) has exceeded initial norm ({norms[0]}) by more than .1% for {self}"
            logger.warning(msg)
            self.warnings["norm_growth"].append(WarningRecord(int(self.time_index), msg))
        if norm > 1.001 * prev:
            msg = (f"Wavefunction norm ({norm}) at time_index = {i} has exceeded norm from previous time step "
                   f"({prev}) by more than .1% for {self}")
            logger.warning(msg)
            self.warnings["norm_growth"].append(WarningRecord(int(self.time_index), msg))

    def run(self, progress_bar: bool = False, callback: Optional[Callable] = None,
            checkpoint_callback: Optional[Callable] = None, chunk_steps: int = 2000):
        """Evolve over the whole time grid on the device and fill the datastores."""
        start_index = self.time_index
        sess = self._session_open()
        self.status = "running"
        try:
            if callback is not None:
                self._run_with_callback(sess, callback)
            else:
                first_collected = start_index
                if start_index > 0 and sess.kind != engine.KIND_SH and start_index in self.data_indices:
                    # Line and cylinder kernels observe after each step: the resumed state's own data
                    # time is stored from the host copy, as the reference's loop does (sims.py:290-294)
                    self.data_time_index = int(np.searchsorted(self.data_indices, start_index))
                    self._store_host_data()
                    first_collected = start_index + 1
                total = self.time_steps  # N + 1 replays from step 0
                done = sess.step_index()
                while done < total:
                    k = min(chunk_steps, total - done)
                    if self.spec.checkpoints and done + k < total:
                        # end the chunk at an exact state for the checkpoint
                        sess.ctx.run(k - 1)
                        sess.hold(True)
                        sess.ctx.run(1)
                        sess.hold(False)
                    else:
                        sess.ctx.run(k)
                    done += k
                    self._check_device_health(sess.flags())
                    if self.spec.checkpoints and done < total:
                        now = datetime.datetime.utcnow()
                        if (now - self.latest_checkpoint_time) > self.spec.checkpoint_every:
                            self.time_index = sess.step_index()
                            self.mesh.g = sess.g()
                            self._collect(first_collected)
                            self.do_checkpoint(now, checkpoint_callback)
                self.mesh.g = sess.g()
                self._collect(first_collected)
                self._check_data_times(start_index)
            self.time_index = self.time_steps - 1
            self.data_time_index = self.data_time_steps
            self.status = "finished"
        finally:
            self._session_close()

    def _run_with_callback(self, sess, callback):
        """The reference's per-step loop (sims.py:289-321): at every data time the datastores are stored
        from g and check() runs before the callback; then one device step, g materialised."""
        sess.hold(True)
        self.data_time_index = int(np.searchsorted(self.data_indices, self.time_index))
        while True:
            is_data_time = self.time_index in self.data_indices
            if is_data_time:
                self._store_host_data()
                self.check()
            callback(self)
            if is_data_time:
                self.data_time_index += 1
            if self.time_index == self.time_steps - 1:
                break
            sess.ctx.run(1)
            self.time_index += 1
            self.mesh.g = sess.g()
        self._check_device_health(sess.flags())

    def do_checkpoint(self, now, callback=None):
        """Pickle the simulation with its mesh (sims.py:348-356)."""
        path = self.save(target_dir=self.spec.checkpoint_dir, save_mesh=True)
        if callback is not None:
            callback(path)
        self.latest_checkpoint_time = now
        return path

    def __getstate__(self):
        state = dict(self.__dict__)
        state["_session"] = None
        return state

    def save(self, target_dir: Optional[str] = None, save_mesh: bool = True):
        """Atomically write ``{target_dir}/{name}.sim`` (gzip pickle of this package's own objects)."""
        target_dir = target_dir or os.getcwd()
        os.makedirs(target_dir, exist_ok=True)
        path = os.path.join(target_dir, f"{self.name}.sim")
        mesh = self.mesh
        if not save_mesh:
            self.mesh = None
        try:
            tmp = path + ".tmp"
            with gzip.open(tmp, "wb") as f:
                pickle.dump(self, f, protocol=pickle.HIGHEST_PROTOCOL)
            os.replace(tmp, path)
        finally:
            self.mesh = mesh
        return path

    @classmethod
    def load(cls, path):
        with gzip.open(path, "rb") as f:
            return pickle.load(f)


class SphericalHarmonicSimulation(MeshSimulation):
    """SH simulation with the largest-l norm check (sims.py:973-990)."""

    def check(self):
        """check() at the current data index plus the largest-l warning (sims.py:979-990); the norm in
        the largest l at that data time is reduced on the device (or taken from g per step)."""
        super().check()
        lnorm = self.lnorm[self.data_time_index]
        # state_overlap(g_L, g_L) = |<g_L|g_L>|^2 in the reference
        norm_in_largest_l = lnorm ** 2
        if norm_in_largest_l > 1e-6:
            fraction = norm_in_largest_l / self.data.norm[self.data_time_index] if data.Norm in \
                self.datastores_by_type else float("nan")
            msg = (f"Wavefunction norm in largest angular momentum state is large at time index {self.time_index} "
                   f"(norm at bound = {norm_in_largest_l}, fraction of norm = {fraction}), "
                   f"consider increasing l bound")
            logger.warning(msg)
            self.warnings["norm_in_largest_l"].append(WarningRecord(int(self.time_index), msg))


class MeshSpecification:
    """Common simulation parameters (sims.py:441-583)."""

    simulation_type = MeshSimulation
    mesh_type = meshes.QuantumMesh

    def __init__(self, name: str, test_mass: float = u.electron_mass_reduced, test_charge: float = u.electron_charge,
                 initial_state=None, test_states=(), internal_potential=None, electric_potential=None,
                 electric_potential_dc_correction: bool = False, electric_potential_fluence_correction: bool = False,
                 mask=None, operators=None, evolution_method=None, time_initial: float = 0 * u.asec,
                 time_final: float = 200 * u.asec, time_step: float = 1 * u.asec, checkpoints: bool = False,
                 checkpoint_every: datetime.timedelta = datetime.timedelta(hours=1),
                 checkpoint_dir: Optional[str] = None, store_data_every: int = 1, datastores=None,
                 device: int = 0, device_rows: Optional[int] = None, **kwargs):
        if electric_potential_dc_correction or electric_potential_fluence_correction:
            raise NotImplementedError("field corrections are not part of the device path")
        self.name = name
        self.test_mass = test_mass
        self.test_charge = test_charge
        self.initial_state = states.HydrogenBoundState(1, 0) if initial_state is None else initial_state
        self.test_states = sorted(test_states)
        if len(self.test_states) == 0:
            self.test_states = [self.initial_state]
        self.internal_potential = potentials.CoulombPotential(charge=u.proton_charge) if internal_potential is None else internal_potential
        self.electric_potential = potentials.NoElectricPotential() if electric_potential is None else electric_potential
        self.electric_potential_dc_correction = electric_potential_dc_correction
        self.electric_potential_fluence_correction = electric_potential_fluence_correction
        self.mask = potentials.NoMask() if mask is None else mask
        self.operators = operators
        self.evolution_method = evolution_method
        self.time_initial = time_initial
        self.time_final = time_final
        self.time_step = time_step
        self.checkpoints = checkpoints
        self.checkpoint_every = checkpoint_every
        self.checkpoint_dir = checkpoint_dir
        self.store_data_every = int(store_data_every)
        if datastores is None:
            datastores = [t() for t in data.DEFAULT_DATASTORE_TYPES]
        self.datastores = datastores
        if len({ds.__class__ for ds in datastores}) != len(datastores):
            raise exceptions.DuplicateDatastores("Cannot duplicate datastores")
        self.device = device
        self.device_rows = device_rows
        self.extra = kwargs

    def to_sim(self):
        return self.simulation_type(self)

    def __repr__(self):
        return f"{self.__class__.__name__}({self.name})"


class LineSpecification(MeshSpecification):
    """LineMesh simulation (sims.py:683-762)."""

    mesh_type = meshes.LineMesh

    def __init__(self, name: str, internal_potential=None, initial_state=None, z_bound: float = 10 * u.nm,
                 z_points: int = 2 ** 9, use_numeric_eigenstates: bool = False,
                 number_of_numeric_eigenstates: int = 100, analytic_eigenstate_type=None, operators=None,
                 evolution_method=None, **kwargs):
        if initial_state is None:
            raise ValueError("LineSpecification needs an initial_state on the device path")
        super().__init__(
            name, internal_potential=internal_potential, initial_state=initial_state,
            operators=mesh_operators.LineLengthGaugeOperators() if operators is None else operators,
            evolution_method=evolution_methods.AlternatingDirectionImplicit() if evolution_method is None else evolution_method,
            **kwargs)
        self.z_bound = z_bound
        self.z_points = int(z_points)
        self.analytic_eigenstate_type = analytic_eigenstate_type
        self.use_numeric_eigenstates = use_numeric_eigenstates
        self.number_of_numeric_eigenstates = number_of_numeric_eigenstates


class CylindricalSliceSpecification(MeshSpecification):
    """CylindricalSliceMesh simulation (sims.py:841-904)."""

    mesh_type = meshes.CylindricalSliceMesh

    def __init__(self, name: str, z_bound: float = 20 * u.bohr_radius, rho_bound: float = 20 * u.bohr_radius,
                 z_points: int = 2 ** 9, rho_points: int = 2 ** 8, operators=None, evolution_method=None, **kwargs):
        super().__init__(
            name,
            operators=mesh_operators.CylindricalSliceLengthGaugeOperators() if operators is None else operators,
            evolution_method=evolution_methods.AlternatingDirectionImplicit() if evolution_method is None else evolution_method,
            **kwargs)
        self.z_bound = z_bound
        self.rho_bound = rho_bound
        self.z_points = int(z_points)
        self.rho_points = int(rho_points)


class SphericalHarmonicSpecification(MeshSpecification):
    """SphericalHarmonicMesh simulation (sims.py:993-1094)."""

    simulation_type = SphericalHarmonicSimulation
    mesh_type = meshes.SphericalHarmonicMesh

    def __init__(self, name: str, r_bound: float = 100 * u.bohr_radius, r_points: int = 1000, l_bound: int = 300,
                 theta_points: int = 180, operators=None, evolution_method=None,
                 use_numeric_eigenstates: bool = True, numeric_eigenstate_max_energy: float = 20 * u.eV,
                 numeric_eigenstate_max_angular_momentum: int = 5, **kwargs):
        super().__init__(
            name,
            operators=mesh_operators.SphericalHarmonicLengthGaugeOperators() if operators is None else operators,
            evolution_method=evolution_methods.SplitInteractionOperator() if evolution_method is None else evolution_method,
            **kwargs)
        self.r_bound = r_bound
        self.r_points = int(r_points)
        self.l_bound = l_bound
        self.theta_points = theta_points
        self.use_numeric_eigenstates = use_numeric_eigenstates
        self.numeric_eigenstate_max_angular_momentum = min(self.l_bound - 1, numeric_eigenstate_max_angular_momentum)
        self.spherical_harmonics = tuple(SphericalHarmonic(l, 0) for l in range(self.l_bound))
        self.numeric_eigenstate_max_energy = numeric_eigenstate_max_energy
