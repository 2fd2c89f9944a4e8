"""
The committed golden vectors (tests/golden/oracle_cases.npz, written by
tests/golden/make_golden.py) against the oracle and the case builders on the
CPU: a change to the oracle or to the builders that moves their results shows
up here.
"""

import json
import os

import numpy as np
import pytest

from oracle import mesh_cpu
from tests.cases import cyl_case, line_case, sh_case
from tests.conftest import rel_l2
from tests.golden.make_golden import tdma_system

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "oracle_cases.npz")


@pytest.fixture(scope="module")
def gold():
    with np.load(GOLDEN) as f:
        d = {k: f[k] for k in f.files}
    d["params"] = json.loads(str(d["params"]))
    return d


def test_tdma_fixture(gold):
    sub, diag, sup, d = tdma_system(**gold["params"]["tdma"])
    assert np.array_equal(sub, gold["tdma_sub"]) and np.array_equal(d, gold["tdma_d"])
    x = mesh_cpu.thomas(sub, diag, sup, d)
    assert rel_l2(x, gold["tdma_x"]) <= 1e-14
    # and the fixture solves its system
    r = diag * x
    r[1:] += sub[1:] * x[:-1]
    r[:-1] += sup[:-1] * x[1:]
    assert rel_l2(r, d) <= 1e-13


def test_sh_fixture(gold):
    p = gold["params"]["sh"]
    case = sh_case(r_points=p["r_points"], l_bound=p["l_bound"], t_half_as=p["t_half_as"], store_every=p["store_every"])
    assert np.array_equal(case["g0"], gold["sh_g0"])
    g, norms, ips, lnorm = case["oracle"].run(case["g0"], case["times"], case["dt"], p["store_every"],
                                             basis=case["basis"])
    assert rel_l2(g, gold["sh_g"]) <= 1e-13
    assert np.max(np.abs(norms - gold["sh_norms"])) <= 1e-13
    assert np.max(np.abs(ips - gold["sh_ips"])) <= 1e-13


def test_line_fixture(gold):
    p = gold["params"]["line"]
    case = line_case(z_points=p["z_points"], n_steps=p["n_steps"], store_every=p["store_every"])
    assert rel_l2(case["g0"], gold["line_g0"]) <= 1e-15
    g, norms, ips = case["oracle"].run(case["g0"], case["times"], p["store_every"], basis=case["basis"])
    assert rel_l2(g, gold["line_g"]) <= 1e-13
    assert np.max(np.abs(norms - gold["line_norms"])) <= 1e-13


def test_cyl_fixture(gold):
    p = gold["params"]["cyl"]
    case = cyl_case(z_points=p["z_points"], rho_points=p["rho_points"], t_half_as=p["t_half_as"],
                    store_every=p["store_every"])
    assert np.array_equal(case["g0"], gold["cyl_g0"])
    g, norms, ips = case["oracle"].run(case["g0"], case["times"], p["store_every"], basis=case["basis"])
    assert rel_l2(g, gold["cyl_g"]) <= 1e-13
    assert np.max(np.abs(ips - gold["cyl_ips"])) <= 1e-13


def test_sh_velocity_fixture(gold):
    p = gold["params"]["shv"]
    case = sh_case(r_points=p["r_points"], l_bound=p["l_bound"], t_half_as=p["t_half_as"], store_every=p["store_every"],
                   gauge="velocity")
    assert np.array_equal(case["g0"], gold["shv_g0"])
    g, norms, ips, lnorm = case["oracle"].run(case["g0"], case["times"], case["dt"], p["store_every"],
                                             basis=case["basis"])
    assert rel_l2(g, gold["shv_g"]) <= 1e-13
    assert np.max(np.abs(norms - gold["shv_norms"])) <= 1e-13
    assert np.max(np.abs(ips - gold["shv_ips"])) <= 1e-13


@pytest.mark.parametrize("key, gauge", [("lsl", "length"), ("lsv", "velocity")])
def test_line_split_operator_fixture(gold, key, gauge):
    p = gold["params"][key]
    case = line_case(z_points=p["z_points"], n_steps=p["n_steps"], store_every=p["store_every"], method="sio",
                     gauge=gauge, e0_au=p["e0_au"], omega_au=p["omega_au"], ramp=False)
    assert rel_l2(case["g0"], gold[f"{key}_g0"]) <= 1e-15
    g, norms, ips = case["oracle"].run(case["g0"], case["times"], p["store_every"], basis=case["basis"])
    assert rel_l2(g, gold[f"{key}_g"]) <= 1e-13
    assert np.max(np.abs(norms - gold[f"{key}_norms"])) <= 1e-13


def test_sh_adi_fixture(gold):
    p = gold["params"]["sha"]
    case = sh_case(r_points=p["r_points"], l_bound=p["l_bound"], t_half_as=p["t_half_as"], store_every=p["store_every"],
                   method="adi")
    assert np.array_equal(case["g0"], gold["sha_g0"])
    g, norms, ips, lnorm = case["oracle"].run(case["g0"], case["times"], case["dt"], p["store_every"],
                                             basis=case["basis"], adi=True)
    assert rel_l2(g, gold["sha_g"]) <= 1e-13
    assert np.max(np.abs(norms - gold["sha_norms"])) <= 1e-13
    assert np.max(np.abs(ips - gold["sha_ips"])) <= 1e-13
