"""
LineMesh with SplitInteractionOperator on the device (length gauge:
LineLengthGaugeOperators.split_interaction_operators, mesh_operators.py:329-342;
velocity gauge: LineVelocityGaugeOperators, mesh_operators.py:352-427;
evolution_methods.py:89-123) against the CPU oracle: final wavefunction
relative L2 <= 1e-10; norm and overlaps <= 1e-9 absolute.
"""

import numpy as np
import pytest

from ionization_amd import mesh, potentials, states
from ionization_amd import units as u
from oracle import mesh_cpu, physics
from tests.cases import line_case
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

WF_RTOL = 1e-10
OBS_ATOL = 1e-9
GAUGES = ["length", "velocity"]
# an unramped strong field: split angles O(0.1 - 1) within the short runs
STRONG = dict(e0_au=2.0, omega_au=0.5, ramp=False)


def _run(ctx, comp):
    slot = ctx.add_line(comp)
    ctx.line_run(comp.n_steps)
    assert ctx.finished(1, slot)
    out = ctx.download_line(slot)
    ctx.remove(1, slot)
    return out


def _check(out, case):
    g_dev, norms_dev, ips_dev = out
    g, norms, ips = case["oracle"].run(case["g0"], case["times"], case["store_every"], basis=case["basis"])
    assert rel_l2(g_dev, g) <= WF_RTOL
    assert np.max(np.abs(norms_dev - norms)) <= OBS_ATOL
    assert np.max(np.abs(ips_dev - ips)) <= OBS_ATOL


@pytest.mark.parametrize("gauge", GAUGES)
@pytest.mark.parametrize("z_points", [500, 257, 8192])
def test_line_split_operator_matches_oracle(device_ctx, gauge, z_points):
    case = line_case(z_points=z_points, n_steps=200, method="sio", gauge=gauge, **STRONG)
    comp = case["compiled"]
    print(f"Line SIO {gauge} z_points={z_points} P={comp.P} m={comp.m}", flush=True)
    _check(_run(device_ctx, comp), case)


@pytest.mark.parametrize("gauge", GAUGES)
@pytest.mark.parametrize("rows", [2, 4, 6, 8, 10, 12, 14, 16])
def test_line_split_operator_every_rows_per_thread(device_ctx, gauge, rows):
    """Every kernel instance (the odd z-pairs cross chunk edges every m rows) on a 900-point line."""
    case = line_case(z_points=900, n_steps=60, rows=rows, method="sio", gauge=gauge, **STRONG)
    comp = case["compiled"]
    print(f"Line SIO {gauge} rows={rows} P={comp.P}", flush=True)
    assert comp.m == rows
    _check(_run(device_ctx, comp), case)


def test_line_methods_in_one_batch(device_ctx):
    """ADI and both split-operator gauges in one context: each bit-identical to its run alone."""
    cases = [line_case(z_points=600, n_steps=80, method="sio", gauge="velocity", **STRONG),
             line_case(z_points=600, n_steps=80),
             line_case(z_points=600, n_steps=80, method="sio", gauge="length", **STRONG),
             line_case(z_points=333, n_steps=50, method="sio", gauge="velocity", **STRONG)]
    alone = [_run(device_ctx, c["compiled"]) for c in cases]
    slots = [device_ctx.add_line(c["compiled"]) for c in cases]
    device_ctx.line_run(max(c["compiled"].n_steps for c in cases))
    for c, a, s in zip(cases, alone, slots):
        m = device_ctx.download_line(s)
        device_ctx.remove(1, s)
        assert np.array_equal(m[0], a[0]) and np.array_equal(m[1], a[1])
        _check(m, c)


@pytest.mark.parametrize("gauge", GAUGES)
def test_line_split_operator_stepwise_equals_single_launch(device_ctx, gauge):
    comp = line_case(z_points=300, n_steps=50, method="sio", gauge=gauge, **STRONG)["compiled"]
    s1 = device_ctx.add_line(comp)
    device_ctx.line_run(comp.n_steps)
    g1, n1, _ = device_ctx.download_line(s1)
    device_ctx.remove(1, s1)
    s2 = device_ctx.add_line(comp)
    device_ctx.run(comp.n_steps)  # one step per graph replay
    g2, n2, _ = device_ctx.download_line(s2)
    device_ctx.remove(1, s2)
    assert np.array_equal(g1, g2)
    assert np.array_equal(n1, n2)


@pytest.mark.parametrize("velocity", [False, True])
def test_line_split_operator_api_matches_oracle(velocity):
    """The drop-in path (LineSpecification + SplitInteractionOperator, the benchmark's
    Line_SO_LEN / Line_SO_VEL specs) against the oracle on the same initial g."""
    zb, omega, e0 = 50 * u.bohr_radius, 0.057 / u.atomic_time, 0.1 * u.atomic_electric_field
    soft = np.sqrt(2) * u.bohr_radius
    ops = mesh.LineVelocityGaugeOperators() if velocity else mesh.LineLengthGaugeOperators()
    z = np.linspace(-zb, zb, 400)
    start = states.NumericOneDState(g=np.exp(-(z / (3 * u.bohr_radius)) ** 2).astype(np.complex128), energy=0.0)
    spec = mesh.LineSpecification(
        "line_so", initial_state=start, z_bound=zb, z_points=400, operators=ops,
        evolution_method=mesh.SplitInteractionOperator(),
        internal_potential=potentials.SoftCoulombPotential(softening_distance=soft),
        electric_potential=potentials.SineWave(omega=omega, amplitude=e0),
        mask=potentials.RadialCosineMask(40 * u.bohr_radius, zb),
        time_initial=0.0, time_final=100 * u.atomic_time, time_step=0.05 * u.atomic_time)
    sim = spec.to_sim()
    g0 = sim.mesh.g.copy()
    sim.run()
    orc = mesh_cpu.LineSIOOracle(
        z_bound=zb, z_points=400, potential=lambda z: physics.soft_coulomb(z, soft),
        field=lambda t: physics.sine_wave_field(t, omega=omega, amplitude=e0),
        mask=lambda z: physics.radial_cosine_mask(z, 40 * u.bohr_radius, zb), test_mass=spec.test_mass,
        gauge="velocity" if velocity else "length")
    g, norms, _ = orc.run(g0, sim.times, spec.store_data_every)
    assert rel_l2(sim.mesh.g, g) <= WF_RTOL
    assert np.max(np.abs(sim.data.norm - norms)) <= OBS_ATOL
