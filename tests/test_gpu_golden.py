"""
Device results through the C ABI against the committed golden vectors
(tests/golden/oracle_cases.npz): final wavefunction rel L2 <= 1e-10,
norms and inner products <= 1e-9 absolute.
"""

import json
import os

import numpy as np
import pytest

from ionization_amd.engine import KIND_CYL, KIND_LINE
from tests.cases import cyl_case, line_case, sh_case
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

WF_RTOL = 1e-10
OBS_ATOL = 1e-9
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "oracle_cases.npz")


@pytest.fixture(scope="module")
def gold():
    with np.load(GOLDEN) as f:
        d = {k: f[k] for k in f.files}
    d["params"] = json.loads(str(d["params"]))
    return d


def test_tdma_matches_golden(device_ctx, gold):
    x = device_ctx.tdma(gold["tdma_sub"], gold["tdma_diag"], gold["tdma_sup"], gold["tdma_d"])
    assert rel_l2(x, gold["tdma_x"]) <= 1e-13


def test_sh_matches_golden(device_ctx, gold):
    p = gold["params"]["sh"]
    c = sh_case(r_points=p["r_points"], l_bound=p["l_bound"], t_half_as=p["t_half_as"],
                store_every=p["store_every"])["compiled"]
    s = device_ctx.add_sh(c)
    device_ctx.run(c.n_steps + 1)
    r = device_ctx.download_sh(s)
    device_ctx.remove(0, s)
    assert rel_l2(r.g, gold["sh_g"]) <= WF_RTOL
    assert np.max(np.abs(r.norms - gold["sh_norms"])) <= OBS_ATOL
    assert np.max(np.abs(r.lnorms - gold["sh_lnorms"])) <= OBS_ATOL
    assert np.max(np.abs(r.ips - gold["sh_ips"])) <= OBS_ATOL


def test_line_matches_golden(device_ctx, gold):
    p = gold["params"]["line"]
    c = line_case(z_points=p["z_points"], n_steps=p["n_steps"], store_every=p["store_every"])["compiled"]
    s = device_ctx.add_line(c)
    device_ctx.line_run(c.n_steps)
    g, norms, ips = device_ctx.download_line(s)
    device_ctx.remove(KIND_LINE, s)
    assert rel_l2(g, gold["line_g"]) <= WF_RTOL
    assert np.max(np.abs(norms - gold["line_norms"])) <= OBS_ATOL
    assert np.max(np.abs(ips - gold["line_ips"])) <= OBS_ATOL


def test_cyl_matches_golden(device_ctx, gold):
    p = gold["params"]["cyl"]
    c = cyl_case(z_points=p["z_points"], rho_points=p["rho_points"], t_half_as=p["t_half_as"],
                 store_every=p["store_every"])["compiled"]
    s = device_ctx.add_cyl(c)
    device_ctx.run(c.n_steps)
    g, norms, ips = device_ctx.download_cyl(s)
    device_ctx.remove(KIND_CYL, s)
    assert rel_l2(g, gold["cyl_g"]) <= WF_RTOL
    assert np.max(np.abs(norms - gold["cyl_norms"])) <= OBS_ATOL
    assert np.max(np.abs(ips - gold["cyl_ips"])) <= OBS_ATOL


def test_sh_velocity_matches_golden(device_ctx, gold):
    p = gold["params"]["shv"]
    c = sh_case(r_points=p["r_points"], l_bound=p["l_bound"], t_half_as=p["t_half_as"],
                store_every=p["store_every"], gauge="velocity")["compiled"]
    s = device_ctx.add_sh(c)
    device_ctx.run(c.n_steps + 1)
    r = device_ctx.download_sh(s)
    device_ctx.remove(0, s)
    assert rel_l2(r.g, gold["shv_g"]) <= WF_RTOL
    assert np.max(np.abs(r.norms - gold["shv_norms"])) <= OBS_ATOL
    assert np.max(np.abs(r.lnorms - gold["shv_lnorms"])) <= OBS_ATOL
    assert np.max(np.abs(r.ips - gold["shv_ips"])) <= OBS_ATOL


@pytest.mark.parametrize("key, gauge", [("lsl", "length"), ("lsv", "velocity")])
def test_line_split_operator_matches_golden(device_ctx, gold, key, gauge):
    p = gold["params"][key]
    c = line_case(z_points=p["z_points"], n_steps=p["n_steps"], store_every=p["store_every"], method="sio",
                  gauge=gauge, e0_au=p["e0_au"], omega_au=p["omega_au"], ramp=False)["compiled"]
    s = device_ctx.add_line(c)
    device_ctx.line_run(c.n_steps)
    g, norms, ips = device_ctx.download_line(s)
    device_ctx.remove(KIND_LINE, s)
    assert rel_l2(g, gold[f"{key}_g"]) <= WF_RTOL
    assert np.max(np.abs(norms - gold[f"{key}_norms"])) <= OBS_ATOL
    assert np.max(np.abs(ips - gold[f"{key}_ips"])) <= OBS_ATOL


def test_sh_adi_matches_golden(device_ctx, gold):
    p = gold["params"]["sha"]
    c = sh_case(r_points=p["r_points"], l_bound=p["l_bound"], t_half_as=p["t_half_as"],
                store_every=p["store_every"], method="adi")["compiled"]
    s = device_ctx.add_sh(c)
    device_ctx.run(c.n_steps + 1)
    r = device_ctx.download_sh(s)
    device_ctx.remove(0, s)
    assert rel_l2(r.g, gold["sha_g"]) <= WF_RTOL
    assert np.max(np.abs(r.norms - gold["sha_norms"])) <= OBS_ATOL
    assert np.max(np.abs(r.lnorms - gold["sha_lnorms"])) <= OBS_ATOL
    assert np.max(np.abs(r.ips - gold["sha_ips"])) <= OBS_ATOL
