"""
SphericalHarmonicMesh with AlternatingDirectionImplicit on the device
(evolution_methods.py:46-77 with the total Hamiltonian of
mesh_operators.py:1020-1035: explicit H0 along r, implicit and explicit E H_int
along l, implicit H0, then the mask) against the CPU oracle: final wavefunction
relative L2 <= 1e-10; norm, largest-l norm, overlaps and norm by l <= 1e-9
absolute.
"""

import numpy as np
import pytest

from ionization_amd import mesh, potentials
from ionization_amd import units as u
from oracle import mesh_cpu, physics
from tests.cases import sh_case
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

WF_RTOL = 1e-10
OBS_ATOL = 1e-9


def _run(ctx, comp, use_graph=True):
    slot = ctx.add_sh(comp)
    ctx.run(comp.n_steps + 1, use_graph=use_graph)
    assert ctx.finished(0, slot)
    res = ctx.download_sh(slot)
    ctx.remove(0, slot)
    return res


def _check(res, case, by_l=False):
    out = case["oracle"].run(case["g0"], case["times"], case["dt"], case["store_every"], basis=case["basis"],
                             by_l=by_l, adi=True)
    g, norms, ips, lnorm = out[:4]
    assert rel_l2(res.g, g) <= WF_RTOL
    assert np.max(np.abs(res.norms - norms)) <= OBS_ATOL
    assert np.max(np.abs(res.lnorms - lnorm)) <= OBS_ATOL
    assert np.max(np.abs(res.ips - ips)) <= OBS_ATOL
    if by_l:
        assert np.max(np.abs(res.norm_by_l - out[4])) <= OBS_ATOL


@pytest.mark.parametrize(
    "r_points, l_bound",
    # odd / even sizes, a single l, a 3-point line, l_bound > 64; the l-direction partition solve
    # up to l_bound 512 (16 rows per thread) and the Thomas sweep beyond it
    [(200, 20), (201, 21), (7, 5), (3, 2), (40, 1), (150, 7), (96, 150), (12, 512), (12, 513), (20, 600)],
)
def test_sh_adi_matches_oracle(device_ctx, r_points, l_bound):
    case = sh_case(r_points=r_points, l_bound=l_bound, t_half_as=100.0, method="adi")
    comp = case["compiled"]
    print(f"SH ADI r_points={r_points} l_bound={l_bound} P={comp.P} m={comp.m} steps={comp.n_steps}", flush=True)
    _check(_run(device_ctx, comp), case)


@pytest.mark.parametrize("rows", [2, 4, 6, 8, 10, 12, 14, 16])
def test_sh_adi_every_rows_per_thread(device_ctx, rows):
    """Every r-sweep instance (explicit and implicit) on a 700-point line."""
    case = sh_case(r_points=700, l_bound=6, t_half_as=20.0, rows=rows, method="adi")
    comp = case["compiled"]
    print(f"SH ADI rows={rows} P={comp.P}", flush=True)
    assert comp.m == rows
    _check(_run(device_ctx, comp), case)


@pytest.mark.parametrize("r_points", [4097, 6000, 8192])
def test_sh_adi_radial_lines_past_4096_points(device_ctx, r_points):
    """512-thread r-sweep workgroups (P = 512, m = 10, 12, 16), explicit and implicit with the
    run-ahead, against the oracle."""
    case = sh_case(r_points=r_points, l_bound=5, t_half_as=12.0, n_basis_l=1, r_bound_bohr=r_points / 20,
                   method="adi")
    comp = case["compiled"]
    print(f"SH ADI long line r_points={r_points} P={comp.P} m={comp.m}", flush=True)
    assert comp.P == 512 and comp.m >= 10
    _check(_run(device_ctx, comp), case)


def test_sh_adi_norm_by_l_sparse_data_and_direct_launches(device_ctx):
    case = sh_case(r_points=120, l_bound=10, t_half_as=60.0, store_every=7, norm_by_l=True, method="adi")
    graph = _run(device_ctx, case["compiled"])
    _check(graph, case, by_l=True)
    direct = _run(device_ctx, case["compiled"], use_graph=False)
    assert np.array_equal(direct.g, graph.g)


@pytest.mark.parametrize("store_every", [1, 9, 1000])
def test_sh_adi_run_ahead_matches_single_steps(device_ctx, store_every):
    """On every replay of a run but the last, the implicit r-sweep may run ahead into the next
    step (the mask, the next explicit half) and the next replay skips its L-phase and explicit
    sweep. One run, runs in uneven chunks with a materialised state read in between, and one
    replay per run (never ahead) give the same bits; every store cadence against the oracle."""
    from ionization_amd import engine

    case = sh_case(r_points=201, l_bound=21, t_half_as=50.0, store_every=store_every, method="adi")
    comp = case["compiled"]
    total = comp.n_steps + 1
    one = _run(device_ctx, comp)
    _check(one, case)

    def chunked(chunks, mid):
        slot = device_ctx.add_sh(comp)
        done, i, g_mid = 0, 0, None
        while done < total:
            k = min(chunks[i % len(chunks)], total - done)
            if done < mid < done + k:
                k = mid - done
            device_ctx.run(k)
            done += k
            i += 1
            if done == mid:
                device_ctx.materialize(engine.KIND_SH, slot)
                assert device_ctx.state(engine.KIND_SH, slot)[0] == mid
                g_mid = device_ctx.download_sh(slot).g
        assert device_ctx.finished(engine.KIND_SH, slot)
        res = device_ctx.download_sh(slot)
        device_ctx.remove(engine.KIND_SH, slot)
        return res, g_mid

    mid = total // 2 + 1
    single, g_single = chunked([1], mid)
    uneven, g_uneven = chunked([3, 8, 2, 17], mid)
    print(f"SH ADI run-ahead, store_every={store_every}: {comp.n_steps} steps, state read at {mid}", flush=True)
    for r in (single, uneven):
        assert np.array_equal(r.g, one.g) and np.array_equal(r.norms, one.norms)
        assert np.array_equal(r.lnorms, one.lnorms) and np.array_equal(r.ips, one.ips)
    assert np.array_equal(g_uneven, g_single)


def test_sh_methods_in_one_batch(device_ctx):
    """ADI, split operator and velocity gauge in one context: each bit-identical to its run alone."""
    cases = [sh_case(r_points=200, l_bound=20, t_half_as=50.0, method="adi"),
             sh_case(r_points=200, l_bound=20, t_half_as=50.0),
             sh_case(r_points=333, l_bound=15, t_half_as=30.0, fluence_jcm2=3.0, method="adi"),
             sh_case(r_points=150, l_bound=9, t_half_as=40.0, gauge="velocity")]
    alone = [_run(device_ctx, c["compiled"]) for c in cases]
    slots = [device_ctx.add_sh(c["compiled"]) for c in cases]
    device_ctx.run(max(c["compiled"].n_steps for c in cases) + 1)
    for c, a, s in zip(cases, alone, slots):
        m = device_ctx.download_sh(s)
        device_ctx.remove(0, s)
        assert np.array_equal(m.g, a.g) and np.array_equal(m.norms, a.norms)


def test_sh_adi_api_matches_oracle():
    """The drop-in path (SphericalHarmonicSpecification + AlternatingDirectionImplicit: the
    reference benchmark's SphericalHarmonic_LAG_CN_LEN) with a pulse, against the oracle."""
    pw, fl, rb = 80 * u.asec, 1.0 * u.Jcm2, 30 * u.bohr_radius
    spec = mesh.SphericalHarmonicSpecification(
        "adi", evolution_method=mesh.AlternatingDirectionImplicit(),
        electric_potential=potentials.SincPulse(pulse_width=pw, fluence=fl, phase=0.3),
        mask=potentials.RadialCosineMask(25 * u.bohr_radius, rb),
        time_initial=-60 * u.asec, time_final=60 * u.asec, time_step=1 * u.asec,
        r_bound=rb, r_points=200, l_bound=16)
    sim = spec.to_sim()
    g0 = sim.mesh.g.copy()
    sim.run()
    orc = mesh_cpu.SHOracle(
        r_bound=rb, r_points=spec.r_points, l_bound=spec.l_bound, potential=physics.coulomb,
        field=lambda t: physics.sinc_pulse_field(t, pulse_width=pw, fluence=fl, phase=0.3),
        mask=lambda r: physics.radial_cosine_mask(r, 25 * u.bohr_radius, rb))
    g, norms, _, _ = orc.run(g0, sim.times, spec.time_step, spec.store_data_every, adi=True)
    assert rel_l2(sim.mesh.g, g) <= WF_RTOL
    assert np.max(np.abs(sim.data.norm - norms)) <= OBS_ATOL
