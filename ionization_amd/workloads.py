"""
The benchmark workloads, built through the ``ionization_amd.mesh`` API.

  line_cn_8192        LineMesh CN: z in [-200, 200] a0, 8192 points, soft-core -1/sqrt(z^2 + 2),
                      ground state from imaginary-time CN, sine field E0 = 0.05 au, w = 0.057 au
                      with a 2-cycle sin^2 ramp, dt = 0.0413 au, 4000 steps, mask on z in [180, 200] a0.
  sh_single           SphericalHarmonicMesh 200 l x 1000 r, r <= 100 a0, hydrogen, numeric 1s,
                      sinc pulse 200 as / 1 J/cm^2 / phase 0, t in [-20 pw, 20 pw], dt = 1 as,
                      mask on r in [90, 100] a0.
  sh_scan             256 SphericalHarmonicMesh simulations drawn from a seed: r_points ~ U{500..4000},
                      r_bound = r_points / 10 a0, l_bound ~ U{50..300}, pw ~ U[50, 800] as,
                      F ~ logU[0.01, 10] J/cm^2, phase ~ U[0, pi]; t in [-20 pw, 20 pw], dt = 1 as,
                      mask on r in [0.9, 1] r_bound.
  served_stream       Poisson arrivals mixing 70 % sh_single-like, 20 % sh_scan draws, 10 % line_cn.

Choices these workloads fix where the workload description leaves them open: the scan's time
window and mask (as above), the line's test mass (1 m_e), its ramp plateau (irrelevant: the run
ends inside the ramp), and the scan's test states (bound numeric states only, i.e.
numeric_eigenstate_max_energy = 0: the observables compared are the norm and the bound-state
overlap).
"""

import numpy as np

from . import engine, potentials, states
from . import units as u
from .mesh import sims as _sims


AU_FIELD = u.atomic_electric_field


def sh_single_spec(name="sh_single", *, r_bound=100 * u.bohr_radius, r_points=1000, l_bound=200,
                   pulse_width=200 * u.asec, fluence=1 * u.Jcm2, phase=0.0, store_data_every=1,
                   numeric_eigenstate_max_energy=20 * u.eV, window=20.0, **kwargs):
    pulse = potentials.SincPulse(pulse_width=pulse_width, fluence=fluence, phase=phase)
    return _sims.SphericalHarmonicSpecification(
        name, r_bound=r_bound, r_points=r_points, l_bound=l_bound, electric_potential=pulse,
        mask=potentials.RadialCosineMask(0.9 * r_bound, r_bound),
        time_initial=-window * pulse_width, time_final=window * pulse_width, time_step=1 * u.asec,
        store_data_every=store_data_every, numeric_eigenstate_max_energy=numeric_eigenstate_max_energy, **kwargs)


def sh_scan_params(n=256, seed=0):
    """Seeded draws of the scan, in a fixed order per simulation."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        r_points = int(rng.integers(500, 4001))
        l_bound = int(rng.integers(50, 301))
        pw = float(rng.uniform(50, 800)) * u.asec
        fluence = float(np.exp(rng.uniform(np.log(0.01), np.log(10)))) * u.Jcm2
        phase = float(rng.uniform(0, np.pi))
        out.append(dict(index=i, r_points=r_points, r_bound=(r_points / 10) * u.bohr_radius, l_bound=l_bound,
                        pulse_width=pw, fluence=fluence, phase=phase))
    return out


def sh_scan_spec(p, store_data_every=100, **kwargs):
    return sh_single_spec(
        f"scan_{p['index']:04d}", r_bound=p["r_bound"], r_points=p["r_points"], l_bound=p["l_bound"],
        pulse_width=p["pulse_width"], fluence=p["fluence"], phase=p["phase"], store_data_every=store_data_every,
        numeric_eigenstate_max_energy=0.0, **kwargs)


def line_ground_state(*, z_bound, z_points, potential, test_mass, imaginary_dt=0.5 * u.atomic_time, n_iter=400,
                      ctx=None):
    """
    Lowest eigenstate of K + V by imaginary-time Crank-Nicolson on the device:
    time steps -i dt, renormalised after every step (the LineMesh kernel with a
    complex tau).
    """
    times = np.arange(n_iter + 1) * imaginary_dt
    comp = engine.compile_line(
        z_bound=z_bound, z_points=z_points, potential=potential, field=lambda t: 0.0 * t, mask=None, times=times,
        g0=np.exp(-(np.linspace(-z_bound, z_bound, z_points) / (4 * u.bohr_radius)) ** 2), basis=(),
        store_every=-1, test_mass=test_mass, time_steps=-1j * np.diff(times), renormalize=True)
    own = ctx is None
    ctx = engine.DeviceContext(pool_bytes=64 << 20) if own else ctx
    try:
        slot = ctx.add_line(comp)
        ctx.line_run(comp.n_steps)
        g, _, _ = ctx.download_line(slot)
        ctx.remove(engine.KIND_LINE, slot)
    finally:
        if own:
            ctx.close()
    # fix the global phase so the vector is real and positive at its peak
    k = int(np.argmax(np.abs(g)))
    g = g * (np.abs(g[k]) / g[k])
    return g


def line_cn_spec(name="line_cn_8192", *, z_points=8192, n_steps=4000, ground_state=None, store_data_every=1,
                 ctx=None, **kwargs):
    z_bound = 200 * u.bohr_radius
    soft = potentials.SoftCoulombPotential(softening_distance=np.sqrt(2.0) * u.bohr_radius)
    mass = u.electron_mass
    if ground_state is None:
        ground_state = line_ground_state(
            z_bound=z_bound, z_points=z_points, test_mass=mass, ctx=ctx,
            potential=lambda z: soft(r=z, test_charge=u.electron_charge))
    omega = 0.057 / u.atomic_time
    period = u.twopi / omega
    field = potentials.SineWave(omega=omega, amplitude=0.05 * AU_FIELD,
                                window=potentials.SmoothedTrapezoidalWindow(time_front=2 * period,
                                                                            time_plateau=10 * period))
    dt = 0.0413 * u.atomic_time
    init = states.NumericOneDState(g=ground_state, energy=0.0, corresponding_analytic_state=states.GroundStateLabel(0))
    spec = _sims.LineSpecification(
        name, internal_potential=soft, initial_state=init, z_bound=z_bound, z_points=z_points, test_mass=mass,
        electric_potential=field, mask=potentials.RadialCosineMask(180 * u.bohr_radius, 200 * u.bohr_radius),
        time_initial=0.0, time_final=n_steps * dt, time_step=dt, store_data_every=store_data_every, **kwargs)
    return fixed_step_count(spec, n_steps)


def fixed_step_count(spec, n_steps):
    """Nudge time_final so that int((tf - t0) / dt) == n_steps exactly (sims.py:204 truncates)."""
    dt = spec.time_step
    tf = spec.time_initial + n_steps * dt
    for _ in range(8):
        if int((tf - spec.time_initial) / dt) == n_steps:
            break
        tf = np.nextafter(tf, np.inf)
    spec.time_final = tf
    return spec


def cyl_scan_params(n=32, seed=0):
    """
    Seeded draws for the batched CylindricalSliceMesh ADI workload: per
    simulation, in this order, pulse width sigma ~ U[100, 500] as and peak field
    E0 ~ logU[0.005, 0.1] atomic units (GaussianPulse.from_amplitude with the
    default carrier, pulses.py:1151-1175). Mesh z in [-100, 100] bohr with 2000
    points, rho in (0, 50] bohr with 1000 points; time window [-5 sigma, 5 sigma],
    dt = 1 as; RadialCosineMask over the outer 10 % of rho_bound (45 -> 50 bohr).
    """
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        pw = float(rng.uniform(100, 500)) * u.asec
        e0 = float(np.exp(rng.uniform(np.log(0.005), np.log(0.1)))) * u.atomic_electric_field
        out.append(dict(index=i, pulse_width=pw, amplitude=e0))
    return out


def cyl_scan_spec(p, *, z_points=2000, rho_points=1000, store_data_every=100):
    from .mesh import CylindricalSliceSpecification

    pw = p["pulse_width"]
    return CylindricalSliceSpecification(
        f"cyl_{p['index']}", z_bound=100 * u.bohr_radius, rho_bound=50 * u.bohr_radius, z_points=z_points,
        rho_points=rho_points,
        electric_potential=potentials.GaussianPulse.from_amplitude(pulse_width=pw, amplitude=p["amplitude"]),
        mask=potentials.RadialCosineMask(45 * u.bohr_radius, 50 * u.bohr_radius),
        time_initial=-5 * pw, time_final=5 * pw, time_step=1 * u.asec, store_data_every=store_data_every)


def cyl_scan_compiled(params, **kw):
    """
    Compile every simulation of the cylindrical scan. All share one mesh, so the
    hydrogen 1s start (and test state) is evaluated once and reused.
    """
    from .mesh import evolution_methods

    compiled = []
    g0 = None
    for p in params:
        spec = cyl_scan_spec(p, **kw)
        if g0 is None:
            sim = spec.to_sim()
            g0 = sim.mesh.g
            times = sim.times
        else:
            total = spec.time_final - spec.time_initial
            times = np.linspace(spec.time_initial, spec.time_final, int(total / spec.time_step) + 1)
        pot = spec.electric_potential
        compiled.append(engine.compile_cyl(
            z_bound=spec.z_bound, rho_bound=spec.rho_bound, z_points=spec.z_points, rho_points=spec.rho_points,
            field=lambda t, pot=pot: np.asarray(pot.get_electric_field_amplitude(t), dtype=np.float64),
            times=times, g0=g0, basis=[g0], store_every=spec.store_data_every,
            charge=evolution_methods._coulomb_charge(spec.internal_potential), test_mass=spec.test_mass,
            test_charge=spec.test_charge, mask=evolution_methods._radial_mask(spec.mask)))
    return compiled
