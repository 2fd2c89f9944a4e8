"""
Velocity-gauge SphericalHarmonicMesh (SphericalHarmonicVelocityGaugeOperators
+ SplitInteractionOperator, mesh_operators.py:1130-1408) on the device
against the CPU oracle, with the same tolerances as the length gauge: final
wavefunction relative L2 <= 1e-10; norm, bound-state overlaps and norm by l
<= 1e-9 absolute.
"""

import numpy as np
import pytest

from ionization_amd import mesh, potentials, states
from ionization_amd import units as u
from oracle import mesh_cpu, physics
from tests.cases import sh_case
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

WF_RTOL = 1e-10
OBS_ATOL = 1e-9


def _run(ctx, comp):
    slot = ctx.add_sh(comp)
    ctx.run(comp.n_steps + 1)
    assert ctx.finished(0, slot)
    res = ctx.download_sh(slot)
    ctx.remove(0, slot)
    return res


def _check(res, case, by_l=False):
    out = case["oracle"].run(case["g0"], case["times"], case["dt"], case["store_every"], basis=case["basis"],
                             by_l=by_l)
    g, norms, ips, lnorm = out[:4]
    assert rel_l2(res.g, g) <= WF_RTOL
    assert np.max(np.abs(res.norms - norms)) <= OBS_ATOL
    assert np.max(np.abs(res.lnorms - lnorm)) <= OBS_ATOL
    assert np.max(np.abs(res.ips - ips)) <= OBS_ATOL
    if by_l:
        assert np.max(np.abs(res.norm_by_l - out[4])) <= OBS_ATOL


@pytest.mark.parametrize(
    "r_points, l_bound",
    # odd / even r_points and l_bound (the pairing parities), a 3-point line, and
    # l_bound > 64 (several L-phase l-segments)
    [(200, 20), (201, 21), (7, 5), (3, 2), (150, 7), (33, 4), (96, 150)],
)
def test_velocity_gauge_matches_oracle(device_ctx, r_points, l_bound):
    case = sh_case(r_points=r_points, l_bound=l_bound, t_half_as=120.0, gauge="velocity")
    comp = case["compiled"]
    print(f"velocity SH r_points={r_points} l_bound={l_bound} P={comp.P} m={comp.m} steps={comp.n_steps}",
          flush=True)
    _check(_run(device_ctx, comp), case)


@pytest.mark.parametrize("rows", [2, 4, 6, 8, 10, 12, 14, 16])
def test_velocity_gauge_every_rows_per_thread(device_ctx, rows):
    """Every h2 kernel instance (rows per thread m; the odd radial pairs cross chunk
    edges every m rows) on a 700-point line: P = 512, 256, 128, 128, 128, 64, 64, 64."""
    case = sh_case(r_points=700, l_bound=6, t_half_as=20.0, rows=rows, gauge="velocity")
    comp = case["compiled"]
    print(f"velocity SH rows={rows} P={comp.P} m={comp.m}", flush=True)
    assert comp.m == rows
    _check(_run(device_ctx, comp), case)


@pytest.mark.parametrize("r_points", [4097, 6000, 8192])
def test_velocity_gauge_radial_lines_past_4096_points(device_ctx, r_points):
    """512-thread workgroups (P = 512, m = 10, 12, 16): the h2 kernel with one thread per chunk,
    and the odd l-pairs' factors as launches of their own around single-line R-phases (two
    lines of 512 chunks do not fit one workgroup)."""
    case = sh_case(r_points=r_points, l_bound=5, t_half_as=12.0, n_basis_l=1, r_bound_bohr=r_points / 20,
                   gauge="velocity")
    comp = case["compiled"]
    print(f"velocity SH long line r_points={r_points} P={comp.P} m={comp.m}", flush=True)
    assert comp.P == 512 and comp.m >= 10
    _check(_run(device_ctx, comp), case)


def test_velocity_gauge_norm_by_l_and_sparse_data(device_ctx):
    case = sh_case(r_points=120, l_bound=10, t_half_as=60.0, store_every=7, norm_by_l=True, gauge="velocity")
    _check(_run(device_ctx, case["compiled"]), case, by_l=True)


def test_velocity_and_length_in_one_batch(device_ctx):
    """Both gauges in one context: each is bit-identical to its run alone and matches its oracle."""
    cases = [sh_case(r_points=200, l_bound=20, t_half_as=60.0, gauge="velocity"),
             sh_case(r_points=200, l_bound=20, t_half_as=60.0, gauge="length"),
             sh_case(r_points=333, l_bound=15, t_half_as=30.0, fluence_jcm2=3.0, gauge="velocity"),
             sh_case(r_points=90, l_bound=9, t_half_as=80.0, gauge="length")]
    alone = [_run(device_ctx, c["compiled"]) for c in cases]
    slots = [device_ctx.add_sh(c["compiled"]) for c in cases]
    device_ctx.run(max(c["compiled"].n_steps for c in cases) + 1)
    mixed = [device_ctx.download_sh(s) for s in slots]
    for s in slots:
        device_ctx.remove(0, s)
    for c, a, m in zip(cases, alone, mixed):
        assert np.array_equal(m.g, a.g)
        assert np.array_equal(m.norms, a.norms)
        assert np.array_equal(m.ips, a.ips)
        _check(m, c)


def test_velocity_gauge_graph_and_direct_launch_agree(device_ctx):
    comp = sh_case(r_points=130, l_bound=11, t_half_as=30.0, gauge="velocity")["compiled"]
    s = device_ctx.add_sh(comp)
    device_ctx.run(comp.n_steps + 1, use_graph=False)
    direct = device_ctx.download_sh(s)
    device_ctx.remove(0, s)
    graph = _run(device_ctx, comp)
    assert np.array_equal(direct.g, graph.g)


def test_velocity_gauge_field_free_keeps_eigenstates():
    """tests/mesh/test_sims.py:30-82 of the reference, velocity-gauge row: with no field the
    norm and the numeric-eigenstate overlaps are unchanged after 100 steps (atol 1e-14)."""
    for initial_state in (states.HydrogenBoundState(1, 0), states.HydrogenBoundState(2, 1)):
        sim = mesh.SphericalHarmonicSpecification(
            "test", initial_state=initial_state,
            operators=mesh.SphericalHarmonicVelocityGaugeOperators(),
            evolution_method=mesh.SplitInteractionOperator(),
            time_initial=0, time_final=100 * u.asec, time_step=1 * u.asec,
            r_bound=50 * u.bohr_radius, r_points=250, l_bound=30, use_numeric_eigenstates=True,
            numeric_eigenstate_max_energy=10 * u.eV, numeric_eigenstate_max_angular_momentum=10,
        ).to_sim()
        sim.run()
        assert np.allclose(sim.data.norm[0], sim.data.norm[-1], atol=1e-14)
        first = np.array([v[0] for v in sim.data.state_overlaps.values()])
        last = np.array([v[-1] for v in sim.data.state_overlaps.values()])
        assert np.allclose(first, last, atol=1e-14)


def test_velocity_gauge_api_run_matches_oracle():
    """The drop-in API path (spec -> sim.run) with a pulse against the oracle on the same inputs."""
    pw, fl, rb = 100 * u.asec, 1.0 * u.Jcm2, 30 * u.bohr_radius
    spec = mesh.SphericalHarmonicSpecification(
        "vel", operators=mesh.SphericalHarmonicVelocityGaugeOperators(),
        evolution_method=mesh.SplitInteractionOperator(),
        electric_potential=potentials.SincPulse(pulse_width=pw, fluence=fl),
        mask=potentials.RadialCosineMask(25 * u.bohr_radius, rb),
        time_initial=-100 * u.asec, time_final=100 * u.asec, time_step=1 * u.asec,
        r_bound=rb, r_points=150, l_bound=12)
    sim = spec.to_sim()
    g0 = sim.mesh.g.copy()
    sim.run()
    orc = mesh_cpu.SHVelocityOracle(
        r_bound=rb, r_points=spec.r_points, l_bound=spec.l_bound, potential=physics.coulomb,
        field=lambda t: physics.sinc_pulse_field(t, pulse_width=pw, fluence=fl, phase=0.0),
        mask=lambda r: physics.radial_cosine_mask(r, 25 * u.bohr_radius, rb), test_mass=spec.test_mass)
    g, norms, _, _ = orc.run(g0, sim.times, spec.time_step, spec.store_data_every)
    assert rel_l2(sim.mesh.g, g) <= WF_RTOL
    assert np.max(np.abs(sim.data.norm - norms)) <= OBS_ATOL
