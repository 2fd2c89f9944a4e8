"""Host side of the layout autotuner: candidates, the tuned table, relayout, JSON round trip."""

import numpy as np
import pytest

from ionization_amd import autotune, engine
from ionization_amd.exceptions import DeviceError
from tests.cases import sh_case


@pytest.mark.parametrize("n", [96, 500, 1000, 2047, 3478, 4000, 4096])
def test_candidates_are_supported_and_least_padded_first(n):
    cands = autotune.sh_candidates(n)
    assert 1 <= len(cands) <= 8
    for P, m in cands:
        engine.check_layout(n, P, m, "sh")
    pads = [P * m - n for P, m in cands]
    assert pads == sorted(pads)
    assert pads[0] == min(P * m - n for P, m in autotune.sh_candidates(n, max_extra=10.0, limit=1000))
    assert all(P * m <= (n + pads[0]) * 1.10 + 1e-9 for P, m in cands)


def test_check_layout_rejects_unsupported():
    for P, m in [(64, 3), (96, 4), (576, 10), (1024, 16), (64, 2), (64, 18)]:
        with pytest.raises(DeviceError):
            engine.check_layout(700, P, m, "sh")
    engine.check_layout(700, 64, 12, "sh")


def test_tuned_table_drives_choose_layout_and_compile():
    rule_1000 = engine.choose_layout(1000, "sh")
    try:
        engine.set_sh_layout_table({768: (384, 2)})
        assert engine.choose_layout(700, "sh") == (384, 2)
        assert engine.choose_layout(700, "sh", rows=8) == (128, 8)  # forced rows win
        assert engine.choose_layout(1000, "sh") == rule_1000  # classes without an entry: the rule
        c = sh_case(r_points=700, l_bound=6, t_half_as=3.0)["compiled"]
        assert (c.P, c.m) == (384, 2)
    finally:
        engine.set_sh_layout_table({})
    assert engine.choose_layout(700, "sh") != (384, 2)


def test_relayout_is_a_permutation():
    c = sh_case(r_points=301, l_bound=9, t_half_as=3.0)["compiled"]
    for lay in autotune.sh_candidates(301, max_extra=2.0, limit=20):
        d = engine.relayout_sh(c, lay)
        assert (d.P, d.m) == lay
        n = c.r_points
        for a, b in [(c.tables, d.tables), (c.g0, d.g0), (c.basis, d.basis)]:
            assert np.array_equal(engine.from_positions(a, n, c.P, c.m), engine.from_positions(b, n, *lay))
            assert np.count_nonzero(b) == np.count_nonzero(a)  # padding stays zero
        back = engine.relayout_sh(d, (c.P, c.m))
        assert np.array_equal(back.tables, c.tables) and np.array_equal(back.g0, c.g0)


def test_table_json_round_trip(tmp_path):
    table = {512: (128, 4), 4096: (256, 16)}
    path = tmp_path / "layouts.json"
    autotune.save(table, str(path), meta={"device": "test"})
    assert autotune.load(str(path)) == table
