"""The layout autotuner on the device: it picks a measured candidate, and simulations compiled
under the tuned table (or re-laid out) still match the oracle."""

import numpy as np
import pytest

from ionization_amd import autotune, engine
from tests.cases import sh_case
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def test_tune_picks_a_measured_candidate_and_keeps_parity(device_ctx):
    sizes = [300, 700]
    print("tuning size classes", sorted({engine.sh_size_class(n) for n in sizes}), flush=True)
    table, costs = autotune.tune_sh(sizes, l_rows=16, lines=256, steps=2, limit=3, log=print)
    assert set(table) == {384, 768}
    for c, lay in table.items():
        assert lay in autotune.sh_candidates(c, limit=3, pow2=True)
        assert costs[c][lay] == min(costs[c].values()) and costs[c][lay] > 0
    try:
        engine.set_sh_layout_table(table)
        case = sh_case(r_points=700, l_bound=12, t_half_as=30.0)
        comp = case["compiled"]
        assert (comp.P, comp.m) == table[768]
        s = device_ctx.add_sh(comp)
        device_ctx.run(comp.n_steps + 1)
        res = device_ctx.download_sh(s)
        device_ctx.remove(engine.KIND_SH, s)
    finally:
        engine.set_sh_layout_table({})
    g, norms, ips, _ = case["oracle"].run(case["g0"], case["times"], case["dt"], case["store_every"],
                                          basis=case["basis"])
    assert rel_l2(res.g, g) <= 1e-10
    assert np.max(np.abs(res.norms - norms)) <= 1e-9
    assert np.max(np.abs(res.ips - ips)) <= 1e-9


def test_relayout_keeps_results(device_ctx):
    case = sh_case(r_points=450, l_bound=10, t_half_as=25.0)
    base = case["compiled"]
    outs = []
    for lay in [(base.P, base.m)] + [x for x in autotune.sh_candidates(450, max_extra=0.5, limit=4)
                                     if x != (base.P, base.m)][:2]:
        c = engine.relayout_sh(base, lay)
        print(f"relayout {lay}", flush=True)
        s = device_ctx.add_sh(c)
        device_ctx.run(c.n_steps + 1)
        outs.append(device_ctx.download_sh(s))
        device_ctx.remove(engine.KIND_SH, s)
    g, norms, _, _ = case["oracle"].run(case["g0"], case["times"], case["dt"], case["store_every"],
                                        basis=case["basis"])
    for r in outs:
        assert rel_l2(r.g, g) <= 1e-10
        assert np.max(np.abs(r.norms - norms)) <= 1e-9


def test_tune_batch_keeps_the_faster_of_rule_and_tuned():
    comps = [sh_case(r_points=rp, l_bound=lb, t_half_as=20.0)["compiled"] for rp, lb in [(300, 30), (700, 20), (450, 40)]]
    table, ms = autotune.tune_batch(comps, l_rows=16, lines=256, steps=3, limit=3, log=print)
    assert set(ms) == {"rule", "tuned"} and min(ms.values()) > 0
    assert table == {} or set(table) == {engine.sh_size_class(c.r_points) for c in comps}
    assert engine.sh_layout_table() == {}
