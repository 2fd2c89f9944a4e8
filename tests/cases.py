"""
Small simulation cases built twice from the same parameters: once through
the device compiler (ionization_amd.engine) and once through the CPU oracle.
"""

import numpy as np
from scipy import linalg
import scipy.linalg as sl

from ionization_amd import engine
from ionization_amd import units as u
from oracle import mesh_cpu, physics


def sh_case(*, r_bound_bohr=30.0, r_points=200, l_bound=20, pulse_width_as=100.0, fluence_jcm2=1.0,
            phase=0.0, t_half_as=300.0, dt_as=1.0, n_basis_l=3, store_every=1, mask=True, rows=None,
            norm_by_l=False, gauge="length", layout=None, method="sio"):
    """SH + split operator (length or velocity gauge) with a sinc pulse, an absorber and a numeric 1s start."""
    r_bound = r_bound_bohr * u.bohr_radius
    pw = pulse_width_as * u.asec
    F = fluence_jcm2 * u.Jcm2
    field = lambda t: physics.sinc_pulse_field(t, pulse_width=pw, fluence=F, phase=phase)
    maskf = (lambda r: physics.radial_cosine_mask(r, 0.8 * r_bound, r_bound)) if mask else None
    oracle_type = mesh_cpu.SHVelocityOracle if gauge == "velocity" else mesh_cpu.SHOracle
    orc = oracle_type(r_bound=r_bound, r_points=r_points, l_bound=l_bound, potential=physics.coulomb,
                      field=field, mask=maskf)
    # numeric eigenstates of each low l block of H0 (dense, small); 1s is the l = 0 ground state
    h0 = orc.h0.tocsr()
    n = r_points
    basis = []
    for l in range(min(n_basis_l, l_bound)):
        blk = h0[l * n:(l + 1) * n, l * n:(l + 1) * n]  # real symmetric tridiagonal
        k = min(3, n)
        if n <= 2048:  # dense, as the committed golden fixtures were made
            _, v = np.linalg.eigh(blk.toarray().real)
        else:
            _, v = linalg.eigh_tridiagonal(blk.diagonal().real, blk.diagonal(1).real, select="i",
                                           select_range=(0, k - 1))
        for i in range(k):
            vec = v[:, i] / np.sqrt(orc.delta_r * np.sum(v[:, i] ** 2))
            basis.append((vec, l))
    g0 = np.zeros((l_bound, n), dtype=np.complex128)
    g0[0] = basis[0][0]
    dt = dt_as * u.asec
    times = mesh_cpu.time_grid(-t_half_as * u.asec, t_half_as * u.asec, dt)
    comp = engine.compile_sh(r_bound=r_bound, r_points=r_points, l_bound=l_bound, potential=physics.coulomb,
                             field=field, mask=maskf, times=times, time_step=dt, g0=g0, basis=basis,
                             store_every=store_every, rows=rows, norm_by_l=norm_by_l, gauge=gauge, layout=layout,
                             method=method)
    return dict(oracle=orc, compiled=comp, g0=g0, times=times, dt=dt, basis=basis, store_every=store_every,
                adi=method == "adi")


def line_case(*, z_bound_bohr=50.0, z_points=500, n_steps=300, e0_au=0.05, omega_au=0.057, dt_au=0.0413,
              store_every=1, mask=True, rows=None, method="adi", gauge="length", ramp=True):
    """Soft-Coulomb LineMesh (ADI, or the split operator in either gauge) with a sine field and a
    one-sided absorber (mask at signed z)."""
    z_bound = z_bound_bohr * u.bohr_radius
    soft = np.sqrt(2.0) * u.bohr_radius
    pot = lambda z: physics.soft_coulomb(z, soft)
    omega = omega_au / u.atomic_time
    period = u.twopi / omega
    window = dict(time_front=2 * period, time_plateau=10 * period) if ramp else {}
    field = lambda t: physics.sine_wave_field(t, omega=omega, amplitude=e0_au * u.atomic_electric_field, **window)
    maskf = (lambda z: physics.radial_cosine_mask(z, 0.8 * z_bound, z_bound)) if mask else None
    mass = u.electron_mass
    if method == "sio":
        orc = mesh_cpu.LineSIOOracle(z_bound=z_bound, z_points=z_points, potential=pot, field=field, mask=maskf,
                                     test_mass=mass, gauge=gauge)
    else:
        orc = mesh_cpu.LineOracle(z_bound=z_bound, z_points=z_points, potential=pot, field=field, mask=maskf,
                                  test_mass=mass)
    hd = orc.h_internal.diagonal().real
    ho = orc.h_internal.diagonal(1).real
    w, v = sl.eigh_tridiagonal(hd, ho, select="i", select_range=(0, 1))
    g0 = v[:, 0] / np.sqrt(orc.delta_z * np.sum(v[:, 0] ** 2))
    dt = dt_au * u.atomic_time
    times = np.linspace(0.0, n_steps * dt, n_steps + 1)
    basis = [g0.astype(np.complex128), (v[:, 1] / np.sqrt(orc.delta_z * np.sum(v[:, 1] ** 2))).astype(np.complex128)]
    comp = engine.compile_line(z_bound=z_bound, z_points=z_points, potential=pot, field=field, mask=maskf,
                               times=times, g0=g0, basis=basis, store_every=store_every, test_mass=mass, rows=rows,
                               method=method, gauge=gauge)
    return dict(oracle=orc, compiled=comp, g0=g0.astype(np.complex128), times=times, basis=basis,
                store_every=store_every)


def cyl_case(*, z_bound_bohr=20.0, rho_bound_bohr=10.0, z_points=120, rho_points=60, pulse_width_as=50.0,
             fluence_jcm2=5.0, t_half_as=40.0, dt_as=1.0, store_every=1, mask=True, z_rows=None, rho_shape=None,
             mask_smoothness=8.0):
    """CylindricalSliceMesh ADI with a sinc pulse, a radial absorber and a hydrogen 1s start."""
    zb = z_bound_bohr * u.bohr_radius
    rb = rho_bound_bohr * u.bohr_radius
    pw = pulse_width_as * u.asec
    F = fluence_jcm2 * u.Jcm2
    field = lambda t: physics.sinc_pulse_field(t, pulse_width=pw, fluence=F, phase=0.0)
    mask_params = (0.7 * zb, zb, mask_smoothness) if mask else None
    maskf = (lambda r: physics.radial_cosine_mask(r, mask_params[0], mask_params[1], mask_smoothness)) if mask else None
    orc = mesh_cpu.CylOracle(z_bound=zb, rho_bound=rb, z_points=z_points, rho_points=rho_points,
                             potential=physics.coulomb, field=field, mask=maskf)
    w = orc.delta_z * orc.delta_rho
    gf = np.sqrt(u.twopi * orc.rho_mesh)

    def normed(g):
        g = g.astype(np.complex128)
        return g / np.sqrt(np.sum(np.abs(g) ** 2) * w)

    a0 = u.bohr_radius
    g0 = normed(gf * np.exp(-orc.r_mesh / a0))                                   # 1s
    basis = [g0, normed(gf * orc.z_mesh * np.exp(-orc.r_mesh / (2 * a0))),        # 2p0-like
             normed(gf * (2 - orc.r_mesh / a0) * np.exp(-orc.r_mesh / (2 * a0)))]  # 2s-like
    dt = dt_as * u.asec
    times = mesh_cpu.time_grid(-t_half_as * u.asec, t_half_as * u.asec, dt)
    comp = engine.compile_cyl(z_bound=zb, rho_bound=rb, z_points=z_points, rho_points=rho_points, field=field,
                              times=times, g0=g0, basis=basis, store_every=store_every, mask=mask_params,
                              z_rows=z_rows, rho_shape=rho_shape)
    return dict(oracle=orc, compiled=comp, g0=g0, times=times, basis=basis, store_every=store_every)
