"""
Layout autotuner for the SphericalHarmonicMesh step.

The device layout of a radial line of n points is (P, m): P chunk-threads
(the R-phase workgroup, a multiple of 64) of m rows each (the depth of the
per-thread tile that the partition solve keeps in registers; P m >= n, the
rest is zero padding). ``engine.choose_layout`` picks one by rule; this module
measures the candidates instead. For every radial size class (128 points wide)
it builds a synthetic batch of SH simulations of that size under each
candidate layout, times whole graph-captured steps (L-phase and R-phase
together) on the device and keeps the fastest per real mesh point. The result
is a table {size class: (P, m)} that ``engine.set_sh_layout_table`` installs,
so every later ``compile_sh`` without forced rows uses the measured layout;
``save`` / ``load`` keep it as JSON beside the caller's data.

The layout changes only the order of floating-point operations inside the
partition solve, so results stay within the parity tolerance whatever the
tuner picks (the every-rows parity tests cover each kernel instance).
"""

import json
import time
from typing import Callable, Dict, Iterable, List, Optional, Tuple

import numpy as np

from . import engine
from . import units as u

Layout = Tuple[int, int]


def sh_candidates(n: int, max_extra: float = 0.10, limit: int = 8, pow2: bool = False) -> List[Layout]:
    """Supported (P, m) for a line of n points whose padding is within max_extra of the
    least-padded one, least padding first (at most `limit`); pow2 keeps power-of-two P only
    (whole waves per SIMD: in a mixed batch 3-wave workgroups measured slow)."""
    out = []
    for m in range(2, 17, 2):
        cap = 512
        P = 64 * -(-n // (64 * m))
        while P <= cap:
            if not pow2 or (P & (P - 1)) == 0:
                out.append((P * m, -m, P, m))
            P += 64
    if not out:
        raise engine.DeviceError(f"no device layout for a line of {n} points")
    best = min(x[0] for x in out)
    ok = sorted(x for x in out if x[0] <= best * (1 + max_extra))
    return [(P, m) for _, _, P, m in ok[:limit]]


def _synthetic(n: int, l_rows: int, layout: Layout, n_steps: int, seed: int) -> engine.CompiledSH:
    """An SH simulation of n radial points and l_rows angular rows under `layout`: hydrogen
    Coulomb potential, a weak constant field, random normalised g (only timing matters)."""
    rng = np.random.default_rng(seed)
    g0 = rng.standard_normal((l_rows, n)) + 1j * rng.standard_normal((l_rows, n))
    dt = 1 * u.asec
    r_bound = (n / 10) * u.bohr_radius
    g0 /= np.sqrt(np.sum(np.abs(g0) ** 2) * (r_bound / n))
    kq = u.coulomb_constant * u.proton_charge * u.electron_charge
    return engine.compile_sh(
        r_bound=r_bound, r_points=n, l_bound=l_rows, potential=lambda r: kq / r,
        field=lambda t: 1e-3 * u.atomic_electric_field + 0.0 * t, mask=None,
        times=np.arange(n_steps + 1) * dt, time_step=dt, g0=g0, store_every=-1, layout=layout)


def time_layout(n: int, layout: Layout, *, l_rows: int = 64, lines: int = 2048, steps: int = 6,
                warmup: int = 2, device: int = 0, repeats: int = 2) -> float:
    """Device milliseconds per real mesh-point-step (x 1e9) of a batch of synthetic simulations
    of n radial points under `layout`, enough of them for `lines` R-phase workgroups."""
    copies = max(1, -(-lines // l_rows))
    comp = _synthetic(n, l_rows, layout, warmup + steps * repeats + 1, seed=n)
    pool = int(copies * comp.footprint_bytes * 1.1) + (64 << 20)
    with engine.DeviceContext(device=device, pool_bytes=pool) as ctx:
        for _ in range(copies):
            ctx.add_sh(comp)
        ctx.run(warmup)
        best = min(ctx.run(steps) for _ in range(repeats))
    return best * 1e9 / (steps * copies * l_rows * n)


def tune_sh(sizes: Iterable[int], *, l_rows: int = 64, lines: int = 2048, steps: int = 6,
            max_extra: float = 0.10, limit: int = 8, pow2: bool = True, device: int = 0,
            log: Optional[Callable[[str], None]] = None) -> Tuple[Dict[int, Layout], dict]:
    """
    Measure every candidate layout for each radial size class among `sizes` and return
    ({class: best (P, m)}, {class: {(P, m): cost}}). Costs are device ms per 1e9
    mesh-point-steps of whole steps; the class is timed at its upper end.
    """
    classes = sorted({engine.sh_size_class(n) for n in sizes})
    table, costs = {}, {}
    for c in classes:
        costs[c] = {}
        for lay in sh_candidates(c, max_extra, limit, pow2):
            costs[c][lay] = time_layout(c, lay, l_rows=l_rows, lines=lines, steps=steps, device=device)
        table[c] = min(costs[c], key=costs[c].get)
        if log:
            rule = engine.choose_layout(c, "sh") if not engine.sh_layout_table() else None
            log(f"size class {c}: best {table[c]} ({costs[c][table[c]]:.3f} ms/G), rule {rule}, "
                f"{len(costs[c])} candidates")
    return table, costs


def save(table: Dict[int, Layout], path: str, meta: Optional[dict] = None) -> None:
    with open(path, "w") as f:
        json.dump({"sh_layouts": {str(k): list(v) for k, v in sorted(table.items())},
                   "meta": dict(meta or {}, saved=time.strftime("%Y-%m-%dT%H:%M:%S"))}, f, indent=1)


def load(path: str) -> Dict[int, Layout]:
    with open(path) as f:
        data = json.load(f)
    return {int(k): (int(v[0]), int(v[1])) for k, v in data["sh_layouts"].items()}


def time_batch(compiled: List[engine.CompiledSH], table: Dict[int, Layout], *, steps: int = 20, warmup: int = 5,
               device: int = 0) -> float:
    """Device ms per step of the whole batch with every simulation laid out by `table`
    (an empty table: the rule)."""
    saved = engine.sh_layout_table()
    try:
        engine.set_sh_layout_table(table)
        relaid = [engine.relayout_sh(c) for c in compiled]
    finally:
        engine.set_sh_layout_table(saved)
    pool = int(sum(c.footprint_bytes for c in relaid) * 1.1) + (256 << 20)
    with engine.DeviceContext(device=device, pool_bytes=pool) as ctx:
        for c in relaid:
            ctx.add_sh(c)
        ctx.run(warmup)
        return min(ctx.run(steps) for _ in range(2)) / steps


def tune_batch(compiled: List[engine.CompiledSH], *, steps: int = 20, device: int = 0,
               log: Optional[Callable[[str], None]] = None, **kw) -> Tuple[Dict[int, Layout], dict]:
    """
    Tune per size class (tune_sh), then time the actual batch under the measured table and
    under the rule and keep the faster: concurrency between the batch's layout buckets is
    only seen in the whole batch. Returns (table, {"rule": ms, "tuned": ms}); the table is
    empty when the rule won.
    """
    table, _ = tune_sh([c.r_points for c in compiled], device=device, log=log, **kw)
    ms = {"rule": time_batch(compiled, {}, steps=steps, device=device),
          "tuned": time_batch(compiled, table, steps=steps, device=device)}
    if log:
        log(f"batch: rule {ms['rule']:.4f} ms/step, tuned {ms['tuned']:.4f} ms/step")
    return (table if ms["tuned"] < ms["rule"] else {}), ms
