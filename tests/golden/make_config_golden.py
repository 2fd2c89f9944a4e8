"""
Writes tests/golden/config_cases.npz: the CPU oracle's full runs of the two
single-simulation workloads at their real sizes, so the device can be checked
against them at those sizes without running the oracle on the test machine.

    python tests/golden/make_config_golden.py      (about two minutes on one core)

  line_cn    LineMesh CN, z in [-200, 200] a0 with 8192 points, soft-core -1/sqrt(z^2 + 2),
             sine field E0 = 0.05 au, w = 0.057 au with a 2-cycle sin^2 ramp, dt = 0.0413 au,
             4000 steps, mask on z in [180, 200] a0 (one-sided: the mask sees the signed z),
             start = the lowest eigenvector of the discrete K + V (LAPACK), data every 100 steps.
  sh_single  SphericalHarmonicMesh 200 l x 1000 r, r <= 100 a0, hydrogen, numeric 1s start,
             sinc pulse 200 as / 1 J/cm^2 / phase 0, t in [-4000, 4000] as, dt = 1 as (8000
             steps), mask on r in [90, 100] a0, data every 100 steps, test states = the numeric
             eigenbasis (l <= 5, E <= 20 eV).

Both starts are stored, so the device runs from the same vectors (eigenvector signs are
arbitrary, meshes.py:1303-1305).
"""

import os
import sys
import time

import numpy as np
import scipy.linalg as sl

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from ionization_amd import units as u  # noqa: E402
from ionization_amd import workloads  # noqa: E402
from oracle import mesh_cpu, physics  # noqa: E402

STORE_EVERY = 100


def line_oracle():
    omega = 0.057 / u.atomic_time
    period = u.twopi / omega
    field = lambda t: physics.sine_wave_field(t, omega=omega, amplitude=0.05 * u.atomic_electric_field,
                                              time_front=2 * period, time_plateau=10 * period)
    return mesh_cpu.LineOracle(
        z_bound=200 * u.bohr_radius, z_points=8192,
        potential=lambda z: physics.soft_coulomb(z, np.sqrt(2.0) * u.bohr_radius), field=field,
        mask=lambda z: physics.radial_cosine_mask(z, 180 * u.bohr_radius, 200 * u.bohr_radius),
        test_mass=u.electron_mass)


def line_ground_state(orc):
    w, v = sl.eigh_tridiagonal(orc.h_internal.diagonal().real, orc.h_internal.diagonal(1).real, select="i",
                               select_range=(0, 0))
    g = v[:, 0] / np.sqrt(orc.delta_z * np.sum(v[:, 0] ** 2))
    return (g * np.sign(g[np.argmax(np.abs(g))])).astype(np.complex128)


def sh_oracle(spec):
    pulse = spec.electric_potential
    field = lambda t: physics.sinc_pulse_field(t, pulse_width=pulse.pulse_width, fluence=pulse.fluence,
                                               phase=pulse.phase)
    return mesh_cpu.SHOracle(r_bound=spec.r_bound, r_points=spec.r_points, l_bound=spec.l_bound,
                             potential=physics.coulomb, field=field,
                             mask=lambda r: physics.radial_cosine_mask(r, 0.9 * spec.r_bound, spec.r_bound))


def build():
    out = {}
    t0 = time.time()
    orc = line_oracle()
    g0 = line_ground_state(orc)
    spec = workloads.line_cn_spec("line_cn", ground_state=g0, store_data_every=STORE_EVERY)
    sim = spec.to_sim()
    assert len(sim.times) == 4001
    g, norms, ips = orc.run(g0, sim.times, STORE_EVERY, basis=[g0])
    out.update(line_g0=g0, line_g=g, line_norms=norms, line_ips=ips)
    print(f"line_cn: {time.time() - t0:.1f} s", flush=True)

    t0 = time.time()
    spec = workloads.sh_single_spec("sh_single", store_data_every=STORE_EVERY)
    sim = spec.to_sim()
    assert len(sim.times) == 8001
    orc = sh_oracle(spec)
    basis = [(np.real(sim.mesh.get_radial_g_for_state(s)), s.l) for s in spec.test_states]
    g0 = sim.mesh.g.copy()
    g, norms, ips, lnorms = orc.run(g0, sim.times, spec.time_step, STORE_EVERY, basis=basis)
    bound = np.array([s.bound for s in spec.test_states])
    out.update(sh_g0=g0, sh_g=g, sh_norms=norms, sh_ips=ips, sh_lnorms=lnorms, sh_bound=bound,
               sh_basis_l=np.array([s.l for s in spec.test_states]))
    print(f"sh_single: {time.time() - t0:.1f} s, {len(basis)} test states", flush=True)
    return out


if __name__ == "__main__":
    path = os.path.join(HERE, "config_cases.npz")
    np.savez_compressed(path, **build())
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")
