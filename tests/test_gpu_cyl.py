"""
CylindricalSliceMesh ADI on the device against the CPU oracle
(oracle/mesh_cpu.py CylOracle, a restatement of meshes.py:509-607,
mesh_operators.py:550-631 and evolution_methods.py:45-77).
"""

import numpy as np
import pytest

from tests.cases import cyl_case
from tests.conftest import rel_l2
from ionization_amd.engine import KIND_CYL

pytestmark = pytest.mark.gpu

WF_RTOL = 1e-10  # final wavefunction, relative L2
OBS_ATOL = 1e-9  # norms and inner products, absolute


def _run_device(ctx, comp, hold=False, per_step=False):
    slot = ctx.add_cyl(comp)
    ctx.set_hold(KIND_CYL, slot, hold)
    if per_step:
        for _ in range(comp.n_steps):
            ctx.run(1)
    else:
        ctx.run(comp.n_steps)
    assert ctx.finished(KIND_CYL, slot)
    ctx.materialize(KIND_CYL, slot)
    out = ctx.download_cyl(slot)
    ctx.remove(KIND_CYL, slot)
    return out


def _check(case, g_dev, norms_dev, ips_dev):
    g, norms, ips = case["oracle"].run(case["g0"], case["times"], case["store_every"], basis=case["basis"])
    assert len(norms_dev) == len(norms)
    assert rel_l2(g_dev, g) <= WF_RTOL
    assert np.max(np.abs(norms_dev - norms)) <= OBS_ATOL
    assert np.max(np.abs(ips_dev - ips)) <= OBS_ATOL
    assert norms[-1] < norms[0]  # the absorber acted
    return g, norms, ips


@pytest.mark.parametrize("rho_shape", [(4, 4), (8, 4), (4, 8), (16, 4), (8, 8)])
def test_cyl_adi_matches_oracle_every_rho_shape(device_ctx, rho_shape):
    case = cyl_case(rho_shape=rho_shape)
    c = case["compiled"]
    print(f"Cyl rho_shape={rho_shape} Cr={c.Cr} Mr={c.Mr} lanes={c.lanes} Pz={c.Pz} Mz={c.Mz}", flush=True)
    _check(case, *_run_device(device_ctx, c))


@pytest.mark.parametrize("smoothness", [1.0, 2.0, 3.0, 8.0, 16.0, 5.5])
def test_cyl_mask_smoothness(device_ctx, smoothness):
    """|cos|^(1/s): square roots for s = 2^k, pow otherwise (potentials/masks.py:76-89)."""
    case = cyl_case(z_points=96, rho_points=40, t_half_as=30.0, store_every=3, mask_smoothness=smoothness)
    print(f"Cyl mask smoothness={smoothness}", flush=True)
    _check(case, *_run_device(device_ctx, case["compiled"]))


@pytest.mark.parametrize("z_rows", [2, 8, 10, 16])
def test_cyl_adi_matches_oracle_z_rows(device_ctx, z_rows):
    case = cyl_case(z_points=150, rho_points=44, z_rows=z_rows, t_half_as=15.0, store_every=4)
    c = case["compiled"]
    print(f"Cyl z_rows={z_rows} Pz={c.Pz} Mz={c.Mz}", flush=True)
    assert c.Mz == z_rows
    _check(case, *_run_device(device_ctx, c))


@pytest.mark.parametrize("z_points, rho_points, z_rows", [(150, 41, None), (120, 45, 16), (4000, 21, 8)])
def test_cyl_odd_rho_rows_and_widest_z_sweep(device_ctx, z_points, rho_points, z_rows):
    """The rho rows are stored in interleaved pairs: an odd rho count ends in a pair with a
    padding row (solved by the z-sweep as zeros); 4000 z points with 8 rows per thread give
    512 z chunks, the z-sweep's widest workgroup (two lines, 1024 threads)."""
    case = cyl_case(z_bound_bohr=40.0 if z_points > 1000 else 20.0, z_points=z_points, rho_points=rho_points,
                    z_rows=z_rows, t_half_as=6.0 if z_points > 1000 else 15.0, store_every=3)
    c = case["compiled"]
    print(f"Cyl {z_points} x {rho_points}: Pz={c.Pz} Mz={c.Mz} Cr={c.Cr} Mr={c.Mr} nrpad={c.nrpad}", flush=True)
    if z_points > 1000:
        assert (c.Pz, c.Mz) == (512, 8)
    _check(case, *_run_device(device_ctx, c))


def test_cyl_hold_and_per_step_match_fused(device_ctx):
    """Exact-state stepping (one step per run, rho explicit half deferred) agrees with the fused run."""
    comp = cyl_case(t_half_as=10.0)["compiled"]
    g1, n1, i1 = _run_device(device_ctx, comp)
    g2, n2, i2 = _run_device(device_ctx, comp, hold=True, per_step=True)
    assert rel_l2(g2, g1) <= 1e-13
    assert np.max(np.abs(n2 - n1)) <= 1e-13
    assert np.max(np.abs(i2 - i1)) <= 1e-13


def test_cyl_materialize_mid_run_requires_hold(device_ctx):
    from ionization_amd.exceptions import DeviceError

    comp = cyl_case(t_half_as=10.0)["compiled"]
    slot = device_ctx.add_cyl(comp)
    device_ctx.run(3)
    with pytest.raises(DeviceError):
        device_ctx.materialize(KIND_CYL, slot)
    device_ctx.remove(KIND_CYL, slot)


def test_cyl_alone_equals_in_batch(device_ctx):
    a = cyl_case(t_half_as=12.0)["compiled"]
    alone = _run_device(device_ctx, a)
    others = [cyl_case(z_points=zp, rho_points=rp, t_half_as=th, fluence_jcm2=fl)["compiled"]
              for zp, rp, th, fl in [(90, 31, 20.0, 1.0), (200, 70, 6.0, 8.0)]]
    slots = [device_ctx.add_cyl(o) for o in others[:1]]
    sa = device_ctx.add_cyl(a)
    slots.append(device_ctx.add_cyl(others[1]))
    device_ctx.run(max(c.n_steps for c in [a] + others))
    mixed = device_ctx.download_cyl(sa)
    for s in slots + [sa]:
        device_ctx.remove(KIND_CYL, s)
    assert np.array_equal(mixed[0], alone[0])
    assert np.array_equal(mixed[1], alone[1])
    assert np.array_equal(mixed[2], alone[2])


def test_cyl_checkpoint_resume(device_ctx):
    """Upload an exact mid-run state and finish: same result as the uninterrupted run."""
    case = cyl_case(t_half_as=10.0)
    comp = case["compiled"]
    full = _run_device(device_ctx, comp)
    slot = device_ctx.add_cyl(comp)
    device_ctx.set_hold(KIND_CYL, slot, True)
    device_ctx.run(7)
    device_ctx.materialize(KIND_CYL, slot)
    g_mid = device_ctx.download_cyl(slot)[0]
    device_ctx.remove(KIND_CYL, slot)
    slot = device_ctx.add_cyl(comp)
    device_ctx.upload_state(KIND_CYL, slot, g_mid, 7)
    device_ctx.run(comp.n_steps)
    g_end = device_ctx.download_cyl(slot)[0]
    device_ctx.remove(KIND_CYL, slot)
    assert rel_l2(g_end, full[0]) <= 1e-13
