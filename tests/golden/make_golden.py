"""
Writes tests/golden/oracle_cases.npz: outputs of the CPU oracle (a restatement
of the reference's evolve step, pinned by tests/test_oracle.py) for small
seeded cases, so GPU tests can check the device against fixed vectors without
running the oracle, and the CPU suite can check that the oracle still
reproduces them.

    python tests/golden/make_golden.py

Cases (parameters are stored in the file; the builders are tests/cases.py):
  tdma  complex diagonally dominant tridiagonal system, n = 37, seed 7 (cy.pyx:9-50)
  sh    SphericalHarmonicMesh split operator, 120 r x 12 l, 80 steps, data every 5
  line  LineMesh CN, 300 z, 80 steps, data every 10
  cyl   CylindricalSliceMesh ADI, 64 z x 33 rho, 20 steps, data every 4
  shv   SphericalHarmonicMesh velocity gauge, 61 r x 9 l, 60 steps, data every 5
  lsl   LineMesh split operator, length gauge, 301 z, 60 steps, strong unramped field
  lsv   LineMesh split operator, velocity gauge, 301 z, 60 steps, strong unramped field
  sha   SphericalHarmonicMesh AlternatingDirectionImplicit, 61 r x 9 l, 60 steps, data every 5
"""

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import mesh_cpu  # noqa: E402
from tests.cases import cyl_case, line_case, sh_case  # noqa: E402

PARAMS = {
    "tdma": dict(n=37, seed=7),
    "sh": dict(r_points=120, l_bound=12, t_half_as=40.0, store_every=5),
    "line": dict(z_points=300, n_steps=80, store_every=10),
    "cyl": dict(z_points=64, rho_points=33, t_half_as=10.0, store_every=4),
    "shv": dict(r_points=61, l_bound=9, t_half_as=30.0, store_every=5),
    "lsl": dict(z_points=301, n_steps=60, store_every=10, e0_au=2.0, omega_au=0.5),
    "lsv": dict(z_points=301, n_steps=60, store_every=10, e0_au=2.0, omega_au=0.5),
    "sha": dict(r_points=61, l_bound=9, t_half_as=30.0, store_every=5),
}


def tdma_system(n, seed):
    rng = np.random.default_rng(seed)
    c = lambda: rng.standard_normal(n) + 1j * rng.standard_normal(n)
    sub, sup, d = c(), c(), c()
    diag = 4.0 + c()
    sub[0] = 0.0
    sup[-1] = 0.0
    return sub, diag, sup, d


def build():
    out = {"params": np.array(json.dumps(PARAMS))}
    sub, diag, sup, d = tdma_system(**PARAMS["tdma"])
    out.update(tdma_sub=sub, tdma_diag=diag, tdma_sup=sup, tdma_d=d, tdma_x=mesh_cpu.thomas(sub, diag, sup, d))

    p = PARAMS["sh"]
    case = sh_case(r_points=p["r_points"], l_bound=p["l_bound"], t_half_as=p["t_half_as"], store_every=p["store_every"])
    g, norms, ips, lnorm = case["oracle"].run(case["g0"], case["times"], case["dt"], p["store_every"],
                                             basis=case["basis"])
    out.update(sh_g0=case["g0"], sh_g=g, sh_norms=norms, sh_ips=ips, sh_lnorms=lnorm)

    p = PARAMS["line"]
    case = line_case(z_points=p["z_points"], n_steps=p["n_steps"], store_every=p["store_every"])
    g, norms, ips = case["oracle"].run(case["g0"], case["times"], p["store_every"], basis=case["basis"])
    out.update(line_g0=case["g0"], line_g=g, line_norms=norms, line_ips=ips)

    p = PARAMS["cyl"]
    case = cyl_case(z_points=p["z_points"], rho_points=p["rho_points"], t_half_as=p["t_half_as"],
                    store_every=p["store_every"])
    g, norms, ips = case["oracle"].run(case["g0"], case["times"], p["store_every"], basis=case["basis"])
    out.update(cyl_g0=case["g0"], cyl_g=g, cyl_norms=norms, cyl_ips=ips)

    p = PARAMS["shv"]
    case = sh_case(r_points=p["r_points"], l_bound=p["l_bound"], t_half_as=p["t_half_as"], store_every=p["store_every"],
                   gauge="velocity")
    g, norms, ips, lnorm = case["oracle"].run(case["g0"], case["times"], case["dt"], p["store_every"],
                                             basis=case["basis"])
    out.update(shv_g0=case["g0"], shv_g=g, shv_norms=norms, shv_ips=ips, shv_lnorms=lnorm)

    for key, gauge in (("lsl", "length"), ("lsv", "velocity")):
        p = PARAMS[key]
        case = line_case(z_points=p["z_points"], n_steps=p["n_steps"], store_every=p["store_every"], method="sio",
                         gauge=gauge, e0_au=p["e0_au"], omega_au=p["omega_au"], ramp=False)
        g, norms, ips = case["oracle"].run(case["g0"], case["times"], p["store_every"], basis=case["basis"])
        out.update({f"{key}_g0": case["g0"], f"{key}_g": g, f"{key}_norms": norms, f"{key}_ips": ips})

    p = PARAMS["sha"]
    case = sh_case(r_points=p["r_points"], l_bound=p["l_bound"], t_half_as=p["t_half_as"], store_every=p["store_every"],
                   method="adi")
    g, norms, ips, lnorm = case["oracle"].run(case["g0"], case["times"], case["dt"], p["store_every"],
                                             basis=case["basis"], adi=True)
    out.update(sha_g0=case["g0"], sha_g=g, sha_norms=norms, sha_ips=ips, sha_lnorms=lnorm)
    return out


if __name__ == "__main__":
    path = os.path.join(HERE, "oracle_cases.npz")
    np.savez_compressed(path, **build())
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")
