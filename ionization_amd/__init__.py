"""
ionization_amd — MI355X-native evolve step for JoshKarpel/ionization's mesh
TDSE simulations.

The host side mirrors the reference's ``ionization.mesh`` API (specifications,
meshes, operators, evolution methods, datastores); the evolve step, the
absorbing mask and the observables run as hand-written HIP kernels for gfx950
in ``libtdse_hip.so`` (C ABI in include/tdse_hip.h), loaded through ctypes.
"""

from . import units, exceptions, core, potentials, states  # noqa: F401

__version__ = "0.1.0"
