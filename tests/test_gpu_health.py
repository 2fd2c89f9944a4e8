"""
Failure detection on the device: the health flags set where a data time's norm is stored
(tdse_sim_flags), the run loop's warnings, and the serving error path. A diverged simulation is
made by uploading a host wavefunction that holds a NaN: ordinary arithmetic then carries the NaN
into the norm, and every kernel runs to its normal end.
"""

import numpy as np
import pytest

from ionization_amd import engine, mesh, potentials, serve
from ionization_amd import units as u
from tests.cases import cyl_case, line_case, sh_case

pytestmark = pytest.mark.gpu


def _with_nan(g):
    g = np.array(g, dtype=np.complex128)
    g.reshape(-1)[g.size // 3] = np.nan
    return g


@pytest.mark.parametrize("kind", ["sh", "line", "cyl"])
def test_nan_state_sets_the_nonfinite_flag_of_that_simulation_only(kind):
    if kind == "sh":
        cases = [sh_case(r_points=120, l_bound=10, t_half_as=20.0, store_every=5) for _ in range(2)]
        k = engine.KIND_SH
    elif kind == "line":
        cases = [line_case(z_points=300, n_steps=40, store_every=5) for _ in range(2)]
        k = engine.KIND_LINE
    else:
        cases = [cyl_case(z_points=64, rho_points=33, t_half_as=10.0, store_every=5) for _ in range(2)]
        k = engine.KIND_CYL
    print(f"health flags: {kind}", flush=True)
    with engine.DeviceContext(pool_bytes=256 << 20) as ctx:
        add = {engine.KIND_SH: ctx.add_sh, engine.KIND_LINE: ctx.add_line, engine.KIND_CYL: ctx.add_cyl}[k]
        good, bad = add(cases[0]["compiled"]), add(cases[1]["compiled"])
        ctx.upload_state(k, bad, _with_nan(cases[1]["g0"]), 0)
        n = cases[0]["compiled"].n_steps
        if k == engine.KIND_LINE:
            ctx.line_run(n)
        else:
            ctx.run(n + 1 if k == engine.KIND_SH else n)
        flags = ctx.flags(k, 0, max(good, bad) + 1)
        assert flags[bad] & engine.FLAG_NONFINITE
        assert flags[good] == 0
        # a fresh upload clears the flags
        ctx.upload_state(k, bad, cases[1]["g0"], 0)
        assert ctx.flags(k, bad)[0] == 0


def test_norm_growth_sets_its_flag():
    case = sh_case(r_points=120, l_bound=10, t_half_as=20.0, store_every=5, mask=False)
    with engine.DeviceContext(pool_bytes=256 << 20) as ctx:
        s = ctx.add_sh(case["compiled"])
        ctx.run(11)
        ctx.materialize(engine.KIND_SH, s)
        st = ctx.state(engine.KIND_SH, s)[0]
        g = ctx.download_sh(s).g
        assert ctx.flags(engine.KIND_SH, s)[0] == 0  # unitary without a mask
        ctx.upload_state(engine.KIND_SH, s, 1.01 * g, st)  # 2 % more norm than the first data time
        ctx.run(10)
        assert ctx.flags(engine.KIND_SH, s)[0] == engine.FLAG_NORM_GROWTH


def test_run_records_a_diverged_simulation():
    spec = mesh.SphericalHarmonicSpecification(
        "nan", r_bound=20 * u.bohr_radius, r_points=100, l_bound=8,
        electric_potential=potentials.SincPulse(pulse_width=50 * u.asec, fluence=1 * u.Jcm2),
        time_initial=-10 * u.asec, time_final=10 * u.asec, time_step=1 * u.asec, store_data_every=5)
    sim = spec.to_sim()
    sim.mesh.g = _with_nan(sim.mesh.g)
    sim.run()
    assert sim.warnings["non_finite"]
    assert np.isnan(sim.data.norm[-1])


def _alone(c):
    with engine.DeviceContext(pool_bytes=int(c.footprint_bytes * 2) + (64 << 20)) as ctx:
        s = ctx.add_sh(c)
        ctx.run(c.n_steps + 1)
        r = ctx.download_sh(s)
        return r.g, r.norms, r.ips


def test_served_batch_retires_a_diverged_simulation_and_finishes_the_rest():
    comps = [sh_case(r_points=150, l_bound=12, t_half_as=25.0)["compiled"] for _ in range(3)]
    comps[1].g0 = np.array(comps[1].g0)
    comps[1].g0[0, 0] = np.nan  # l = 0, radial index 0 (device position order)
    reqs = [serve.Request(i, "sh", 0.0, (lambda c=c: c)) for i, c in enumerate(comps)]
    with serve.Server(pool_bytes=256 << 20, chunk_steps=10) as srv:
        rep = srv.run(reqs, max_seconds=120, wait_compiled=True)
    by_id = {r.rid: r for r in rep.requests}
    assert by_id[1].error is not None
    for i in (0, 2):
        assert by_id[i].error is None
        g, norms, ips = _alone(comps[i])
        assert np.array_equal(by_id[i].g, g) and np.array_equal(by_id[i].norms, norms)
