"""
Host logic of the serving scheduler (ionization_amd/serve.py), on the CPU
with a stand-in for the device context that counts replays per slot.
"""

import numpy as np
import pytest

from ionization_amd import engine, serve
from ionization_amd import units as u
from ionization_amd.exceptions import DeviceError


class FakeContext:
    """Records what the scheduler does; every slot advances one step per replay."""

    def __init__(self):
        self.slots = {}
        self.next_slot = 0
        self.log = []

    def _add(self, c):
        s = self.next_slot
        self.next_slot += 1
        self.slots[s] = {"c": c, "replays": 0}
        self.log.append(("add", s))
        return s

    add_sh = add_line = add_cyl = _add

    def run(self, k, use_graph=True):
        for v in self.slots.values():
            v["replays"] += k
        self.log.append(("run", k))
        return 0.0

    def materialize(self, kind, slot):
        v = self.slots[slot]
        need = v["c"].n_steps + (0 if kind == engine.KIND_CYL else 1)
        assert v["replays"] >= need, "retired before its last step"

    def download_line(self, slot):
        c = self.slots[slot]["c"]
        return np.zeros(c.z_points, np.complex128), np.zeros(2), np.zeros((2, 0), np.complex128)

    def remove(self, kind, slot):
        self.log.append(("remove", slot))
        del self.slots[slot]

    def flags(self, kind, first, n):
        return np.zeros(n, np.int32)

    def close(self):
        pass


def _line(n_steps, z_points=128):
    times = np.arange(n_steps + 1) * 1.0 * u.asec
    return engine.compile_line(z_bound=10 * u.bohr_radius, z_points=z_points, potential=lambda z: 0 * z,
                               field=lambda t: 0 * t, mask=None, times=times, g0=np.ones(z_points), basis=(),
                               store_every=-1)


def _requests(steps, arrival=0.0):
    return [serve.Request(i, "line", arrival, (lambda c=_line(n): c)) for i, n in enumerate(steps)]


def test_every_request_finishes_after_its_own_step_count():
    ctx = FakeContext()
    with serve.Server(pool_bytes=1 << 30, ctx=ctx, chunk_steps=7) as srv:
        rep = srv.run(_requests([5, 30, 12, 1]), wait_compiled=True)
    assert all(r.t_done is not None and r.latency >= 0 for r in rep.requests)
    assert not ctx.slots  # everything retired
    # all admitted at replay 0: the longest simulation (30 steps) needs 31 replays, none wasted
    assert rep.replays == 31


def test_admission_respects_the_pool_and_arrival_order():
    reqs = _requests([10, 10, 10, 10])
    one = serve.Server._reserve(reqs[0].make())
    ctx = FakeContext()
    with serve.Server(pool_bytes=2 * one + one // 2, ctx=ctx, chunk_steps=100) as srv:
        rep = srv.run(reqs, wait_compiled=True)
    adds = [e for e in ctx.log if e[0] == "add"]
    assert [s for _, s in adds] == [0, 1, 2, 3]
    # at most two resident at any time
    live, peak = 0, 0
    for e in ctx.log:
        live += {"add": 1, "remove": -1}.get(e[0], 0)
        peak = max(peak, live)
    assert peak == 2
    assert [r.admitted_replay for r in rep.requests] == [0, 0, 11, 11]


def test_request_larger_than_the_pool_is_an_error():
    reqs = _requests([3])
    with serve.Server(pool_bytes=1024, ctx=FakeContext()) as srv:
        with pytest.raises(DeviceError):
            srv.run(reqs)


def test_served_mix_is_seeded_and_mixes_the_three_classes():
    a = serve.served_mix(200, 10.0, seed=3, line_ground_state=np.ones(8192))
    b = serve.served_mix(200, 10.0, seed=3, line_ground_state=np.ones(8192))
    assert [(r.kind, r.arrival) for r in a] == [(r.kind, r.arrival) for r in b]
    kinds = [r.kind for r in a]
    frac = {k: kinds.count(k) / len(kinds) for k in ("sh_single", "sh_scan", "line_cn")}
    assert abs(frac["sh_single"] - 0.7) < 0.1 and abs(frac["sh_scan"] - 0.2) < 0.08
    assert abs(frac["line_cn"] - 0.1) < 0.06
    gaps = np.diff([0.0] + [r.arrival for r in a])
    assert np.all(gaps > 0) and abs(np.mean(gaps) - 0.1) < 0.03


class FlaggingContext(FakeContext):
    """A stand-in whose device reports a non-finite norm for chosen slots after a given replay."""

    def __init__(self, bad_after):
        super().__init__()
        self.bad_after = bad_after  # slot -> replays after which its norm is non-finite

    def flags(self, kind, first, n):
        out = np.zeros(n, np.int32)
        for s, v in self.slots.items():
            if first <= s < first + n and s in self.bad_after and v["replays"] >= self.bad_after[s]:
                out[s - first] = engine.FLAG_NONFINITE
        return out

    def materialize(self, kind, slot):
        assert slot not in self.bad_after, "a diverged simulation is not materialised"
        super().materialize(kind, slot)


def test_diverged_simulation_is_retired_with_an_error_and_the_others_finish():
    ctx = FlaggingContext({1: 8})
    with serve.Server(pool_bytes=1 << 30, ctx=ctx, chunk_steps=5) as srv:
        rep = srv.run(_requests([20, 40, 12]), wait_compiled=True)
    by_id = {r.rid: r for r in rep.requests}
    assert by_id[1].error is not None and "diverged" in by_id[1].error
    assert by_id[0].error is None and by_id[2].error is None
    assert not ctx.slots
    # the diverged one left after the first chunk that saw its flag, long before its 41 replays
    removes = [i for i, e in enumerate(ctx.log) if e == ("remove", 1)]
    runs_before = sum(e[1] for e in ctx.log[: removes[0]] if e[0] == "run")
    assert runs_before == 10 and rep.replays == 21
