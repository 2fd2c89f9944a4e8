"""
Time-indexed datastores (``ionization/mesh/data.py``).

The names, the ``sim.data.<name>`` access and the error types follow the
reference (data.py:14-100, 152-287). Values come from the device: the
evolve kernels reduce the norm, the norm in the largest l and the inner
products with the test states at every data time, in a fixed order per
simulation; ``collect`` copies them into the arrays after (or during) a run.
``Fields`` is host-side: the field and its time integral are known functions
of time.
"""

import copy

import numpy as np

from .. import exceptions, potentials


class Data:
    """``sim.data``: attribute access to every attached datastore's arrays (data.py:14-100)."""

    def __init__(self, sim):
        self.sim = sim
        self.times = sim.data_times

    def __getattr__(self, item):
        if item.startswith("__") and item.endswith("__"):
            raise AttributeError(item)
        datastore = DATA_NAME_TO_DATASTORE_TYPE.get(item)
        if datastore is None:
            raise exceptions.UnknownData(
                f"Couldn't find any data named '{item}' on {self.sim}. "
                "Ensure that the corresponding datastore is correctly implemented.")
        raise exceptions.MissingDatastore(
            f"Couldn't get data '{item}' for {self.sim} because it does not include a {datastore.__name__} datastore.")


DATA_NAME_TO_DATASTORE_TYPE = {}


class Datastore:
    def init(self, sim):
        self.sim = sim
        self.spec = sim.spec
        self.attach()

    def attach(self):
        raise NotImplementedError

    def collect(self, slots, device):
        """Copy device-reduced values for data slots ``slots`` (indices into data_times)."""

    def store_host(self):
        """Evaluate on the host copy of g at the current data index (per-step mode)."""

    def __repr__(self):
        return self.__class__.__name__


def _link(datastore_type, name):
    def getter(data):
        try:
            ds = data.sim.datastores_by_type[datastore_type]
        except KeyError:
            raise exceptions.MissingDatastore(
                f"Couldn't get data {name} for {data.sim} because it does not include a "
                f"{datastore_type.__name__} datastore.") from None
        return getattr(ds, name)()

    return property(getter)


class Fields(Datastore):
    """Electric field and vector potential at the data times (data.py:152-186)."""

    def init(self, sim):
        times = sim.data_times
        self.electric_field_amplitude = np.asarray(
            sim.spec.electric_potential.get_electric_field_amplitude(times), dtype=np.float64) * np.ones(len(times))
        # vector potential: -integral of E over times_to_current at each data time (data.py:161-170)
        all_t = sim.times
        e_all = np.asarray(sim.spec.electric_potential.get_electric_field_amplitude(all_t), dtype=np.float64)
        cum = potentials.simps_cumulative(e_all * np.ones(len(all_t)), all_t)
        self.vector_potential_amplitude = -cum[sim.data_indices]
        super().init(sim)

    def attach(self):
        self.sim.data.electric_field_amplitude = self.electric_field_amplitude
        self.sim.data.vector_potential_amplitude = self.vector_potential_amplitude


DATA_NAME_TO_DATASTORE_TYPE.update({"electric_field_amplitude": Fields, "vector_potential_amplitude": Fields})


class Norm(Datastore):
    """Wavefunction norm vs time (data.py:189-207)."""

    def init(self, sim):
        self.norm = sim.get_blank_data()
        super().init(sim)

    def attach(self):
        self.sim.data.norm = self.norm

    def collect(self, slots, device):
        self.norm[slots] = device["norms"][slots]

    def store_host(self):
        self.norm[self.sim.data_time_index] = self.sim.mesh.norm()


DATA_NAME_TO_DATASTORE_TYPE.update({"norm": Norm})


class InnerProducts(Datastore):
    """Inner products with the test states (data.py:210-287)."""

    def init(self, sim):
        self.inner_products = {state: sim.get_blank_data(dtype=np.complex128) for state in sim.spec.test_states}
        super().init(sim)

    def attach(self):
        self.sim.data.inner_products = self.inner_products
        self.sim.data.initial_state_inner_product = self.inner_products[self.spec.initial_state]

    def collect(self, slots, device):
        for i, state in enumerate(self.spec.test_states):
            self.inner_products[state][slots] = device["ips"][slots, i]

    def store_host(self):
        for state in self.spec.test_states:
            self.inner_products[state][self.sim.data_time_index] = self.sim.mesh.inner_product(state)

    def state_overlaps(self):
        return {state: np.abs(ip) ** 2 for state, ip in self.inner_products.items()}

    def initial_state_overlap(self):
        return np.abs(self.inner_products[self.spec.initial_state]) ** 2

    def bound_state_overlap(self):
        return sum(ov for state, ov in self.state_overlaps().items() if state.bound)

    def free_state_overlap(self):
        return sum(ov for state, ov in self.state_overlaps().items() if state.free)

    def total_state_overlap(self):
        return sum(self.state_overlaps().values())


for _name in ("state_overlaps", "initial_state_overlap", "bound_state_overlap", "free_state_overlap",
              "total_state_overlap"):
    setattr(Data, _name, _link(InnerProducts, _name))
DATA_NAME_TO_DATASTORE_TYPE.update({
    "inner_products": InnerProducts,
    "initial_state_inner_product": InnerProducts,
    "state_overlaps": InnerProducts,
    "initial_state_overlap": InnerProducts,
    "bound_state_overlap": InnerProducts,
    "free_state_overlap": InnerProducts,
    "total_state_overlap": InnerProducts,
})

class NormBySphericalHarmonic(Datastore):
    """
    Norm in every spherical harmonic vs time (data.py:438-464): sum over r of
    |g_l(r)|^2 times delta_r (meshes.py:1130-1133), keyed by the spec's
    ``spherical_harmonics``. On the device the L-phase of the step at a data
    time reduces it per l while the state streams through, so it costs no
    extra pass over the mesh.
    """

    def init(self, sim):
        self.norm_by_l = {sph_harm: sim.get_blank_data() for sph_harm in sim.spec.spherical_harmonics}
        super().init(sim)

    def attach(self):
        self.sim.data.norm_by_sph_harm = self.norm_by_l

    def collect(self, slots, device):
        by_l = device.get("norm_by_l")
        if by_l is None:
            raise exceptions.MissingDatastore(
                f"the device did not reduce the norm by spherical harmonic for {self.sim}")
        for sph_harm, col in zip(self.spec.spherical_harmonics, by_l.T):
            self.norm_by_l[sph_harm][slots] = col[slots]

    def store_host(self):
        for sph_harm, l_norm in zip(self.spec.spherical_harmonics, self.sim.mesh.norm_by_l()):
            self.norm_by_l[sph_harm][self.sim.data_time_index] = l_norm


DATA_NAME_TO_DATASTORE_TYPE.update({"norm_by_sph_harm": NormBySphericalHarmonic})

DEFAULT_DATASTORE_TYPES = (Fields, Norm, InnerProducts)


def clone(datastores):
    return {ds.__class__: copy.deepcopy(ds) for ds in datastores}
