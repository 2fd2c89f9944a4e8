"""
Edge sizes on the device against the oracle: single-l and two-l SH meshes,
radial lines shorter than a wave or one point past it, three-point Line
meshes, cylinders of a few points per direction (mostly padding in every
workgroup). Tolerances as in the parity tests.
"""

import numpy as np
import pytest

from ionization_amd import engine
from tests.cases import cyl_case, line_case, sh_case
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

WF_RTOL = 1e-10
OBS_ATOL = 1e-9


@pytest.mark.parametrize("r_points, l_bound", [(10, 1), (3, 4), (33, 2), (64, 3), (65, 1), (130, 1), (129, 5)])
def test_sh_edge_sizes(device_ctx, r_points, l_bound):
    case = sh_case(r_points=r_points, l_bound=l_bound, t_half_as=15.0, n_basis_l=1,
                   r_bound_bohr=max(3.0, r_points / 6))
    comp = case["compiled"]
    print(f"SH edge r_points={r_points} l_bound={l_bound} P={comp.P} m={comp.m}", flush=True)
    s = device_ctx.add_sh(comp)
    device_ctx.run(comp.n_steps + 1)
    res = device_ctx.download_sh(s)
    device_ctx.remove(engine.KIND_SH, s)
    g, norms, ips, lnorm = case["oracle"].run(case["g0"], case["times"], case["dt"], 1, basis=case["basis"])
    assert rel_l2(res.g, g) <= WF_RTOL
    assert np.max(np.abs(res.norms - norms)) <= OBS_ATOL
    assert np.max(np.abs(res.lnorms - lnorm)) <= OBS_ATOL
    assert np.max(np.abs(res.ips - ips)) <= OBS_ATOL


@pytest.mark.parametrize("z_points", [3, 5, 64, 65])
def test_line_edge_sizes(device_ctx, z_points):
    case = line_case(z_points=z_points, n_steps=40, z_bound_bohr=max(5.0, z_points / 4))
    comp = case["compiled"]
    print(f"Line edge z_points={z_points} P={comp.P} m={comp.m}", flush=True)
    s = device_ctx.add_line(comp)
    device_ctx.line_run(comp.n_steps)
    g_dev, norms_dev, ips_dev = device_ctx.download_line(s)
    device_ctx.remove(engine.KIND_LINE, s)
    g, norms, ips = case["oracle"].run(case["g0"], case["times"], case["store_every"], basis=case["basis"])
    assert rel_l2(g_dev, g) <= WF_RTOL
    assert np.max(np.abs(norms_dev - norms)) <= OBS_ATOL
    assert np.max(np.abs(ips_dev - ips)) <= OBS_ATOL


@pytest.mark.parametrize("z_points, rho_points", [(4, 2), (10, 3), (64, 17), (65, 2)])
def test_cyl_edge_sizes(device_ctx, z_points, rho_points):
    case = cyl_case(z_points=z_points, rho_points=rho_points, t_half_as=6.0, store_every=2)
    comp = case["compiled"]
    print(f"Cyl edge {z_points} x {rho_points}: Pz={comp.Pz} Mz={comp.Mz} Cr={comp.Cr} Mr={comp.Mr} "
          f"lanes={comp.lanes}", flush=True)
    s = device_ctx.add_cyl(comp)
    device_ctx.run(comp.n_steps)
    device_ctx.materialize(engine.KIND_CYL, s)
    g_dev, norms_dev, ips_dev = device_ctx.download_cyl(s)
    device_ctx.remove(engine.KIND_CYL, s)
    g, norms, ips = case["oracle"].run(case["g0"], case["times"], case["store_every"], basis=case["basis"])
    assert rel_l2(g_dev, g) <= WF_RTOL
    assert np.max(np.abs(norms_dev - norms)) <= OBS_ATOL
    assert np.max(np.abs(ips_dev - ips)) <= OBS_ATOL


@pytest.mark.parametrize("r_points", [4097, 6000, 7000, 8192])
def test_sh_radial_lines_past_4096_points(device_ctx, r_points):
    """512-thread R-phase workgroups of 10-16 rows per thread (m = 10, 12, 14, 16)."""
    case = sh_case(r_points=r_points, l_bound=5, t_half_as=12.0, n_basis_l=1, r_bound_bohr=r_points / 20)
    comp = case["compiled"]
    print(f"SH long line r_points={r_points} P={comp.P} m={comp.m}", flush=True)
    assert comp.P == 512 and comp.m >= 10
    s = device_ctx.add_sh(comp)
    device_ctx.run(comp.n_steps + 1)
    res = device_ctx.download_sh(s)
    device_ctx.remove(engine.KIND_SH, s)
    g, norms, ips, _ = case["oracle"].run(case["g0"], case["times"], case["dt"], 1, basis=case["basis"])
    assert rel_l2(res.g, g) <= WF_RTOL
    assert np.max(np.abs(res.norms - norms)) <= OBS_ATOL
    assert np.max(np.abs(res.ips - ips)) <= OBS_ATOL
