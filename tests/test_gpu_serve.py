"""
Serving on the device: a stream of SH, Line and cylindrical simulations with
staggered arrivals, admitted and retired while others run, gives every
simulation exactly the results it gets alone.
"""

import numpy as np
import pytest

from ionization_amd import engine, serve
from tests.cases import cyl_case, line_case, sh_case

pytestmark = pytest.mark.gpu


def _alone(c):
    with engine.DeviceContext(pool_bytes=int(c.footprint_bytes * 2) + (64 << 20)) as ctx:
        if isinstance(c, engine.CompiledSH):
            s = ctx.add_sh(c)
            ctx.run(c.n_steps + 1)
            r = ctx.download_sh(s)
            return r.g, r.norms, r.ips
        if isinstance(c, engine.CompiledCyl):
            s = ctx.add_cyl(c)
            ctx.run(c.n_steps)
            return ctx.download_cyl(s)
        s = ctx.add_line(c)
        ctx.run(c.n_steps)
        return ctx.download_line(s)


def test_served_results_equal_alone_results():
    comps = [
        sh_case(r_points=200, l_bound=20, t_half_as=30.0)["compiled"],
        line_case(z_points=500, n_steps=70)["compiled"],
        cyl_case(t_half_as=20.0)["compiled"],
        sh_case(r_points=333, l_bound=15, t_half_as=12.0, fluence_jcm2=3.0)["compiled"],
        sh_case(r_points=90, l_bound=9, t_half_as=45.0, fluence_jcm2=0.1)["compiled"],
        cyl_case(z_points=90, rho_points=31, t_half_as=9.0, fluence_jcm2=1.0)["compiled"],
        line_case(z_points=257, n_steps=33)["compiled"],
    ]
    arrivals = [0.0, 0.0, 0.01, 0.02, 0.05, 0.05, 0.08]
    reqs = [serve.Request(i, type(c).__name__, a, (lambda c=c: c)) for i, (c, a) in enumerate(zip(comps, arrivals))]
    # a pool that holds only some of them at once, so admissions wait for retirements
    pool = sum(serve.Server._reserve(c) for c in comps) * 2 // 3
    pool = max(pool, max(serve.Server._reserve(c) for c in comps))
    print(f"serving {len(reqs)} simulations in a {pool / 2**20:.1f} MiB pool", flush=True)
    with serve.Server(pool_bytes=pool, chunk_steps=9) as srv:
        rep = srv.run(reqs, max_seconds=120)
    assert rep.p99 >= rep.p50 > 0
    for r, c in zip(sorted(rep.requests, key=lambda r: r.rid), comps):
        g, norms, ips = _alone(c)
        assert np.array_equal(r.g, g), r.kind
        assert np.array_equal(r.norms, norms), r.kind
        assert np.array_equal(r.ips, ips), r.kind


def test_served_velocity_gauge_adi_and_line_split_operator_equal_alone():
    """The velocity-gauge SH path (its own L-phase instance and h2 chain), SH ADI and the Line
    split operator, served beside length-gauge simulations, give their alone results bit for bit."""
    comps = [
        sh_case(r_points=200, l_bound=20, t_half_as=30.0, gauge="velocity")["compiled"],
        sh_case(r_points=150, l_bound=11, t_half_as=20.0)["compiled"],
        line_case(z_points=500, n_steps=70, method="sio", gauge="velocity", e0_au=2.0, omega_au=0.5,
                  ramp=False)["compiled"],
        sh_case(r_points=333, l_bound=15, t_half_as=12.0, fluence_jcm2=3.0, gauge="velocity")["compiled"],
        line_case(z_points=257, n_steps=33, method="sio", gauge="length", e0_au=2.0, omega_au=0.5,
                  ramp=False)["compiled"],
        sh_case(r_points=180, l_bound=13, t_half_as=25.0, method="adi")["compiled"],
    ]
    arrivals = [0.0, 0.0, 0.01, 0.03, 0.05, 0.06]
    reqs = [serve.Request(i, type(c).__name__, a, (lambda c=c: c)) for i, (c, a) in enumerate(zip(comps, arrivals))]
    pool = max(sum(serve.Server._reserve(c) for c in comps) * 2 // 3, max(serve.Server._reserve(c) for c in comps))
    with serve.Server(pool_bytes=pool, chunk_steps=9) as srv:
        rep = srv.run(reqs, max_seconds=120)
    for r, c in zip(sorted(rep.requests, key=lambda r: r.rid), comps):
        g, norms, ips = _alone(c)
        assert np.array_equal(r.g, g), r.kind
        assert np.array_equal(r.norms, norms), r.kind
        assert np.array_equal(r.ips, ips), r.kind
