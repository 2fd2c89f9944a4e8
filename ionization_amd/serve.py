"""
Serving many mesh simulations from one device context.

Requests arrive over time (``Request.arrival``, seconds from the start of the
run). A pool of host threads compiles each request's specification into
device table form as soon as it arrives (``spec.to_sim()`` and the evolution
method's ``compile``: the reference's setup, sims.py:46-118). The scheduler
admits compiled requests in arrival order while their device footprint fits
the context's memory pool, advances every resident simulation together with
one captured graph replay per step (in chunks of ``chunk_steps`` replays), and
retires each simulation as soon as its last step has run: its wavefunction and
stored observables are copied to the host and its slot and memory are freed
for the next admission. A simulation's results do not depend on which other
simulations share the device (every kernel works per simulation in a fixed
order), so a served result equals the result of running it alone.

Time-to-result is measured from arrival to the observables being on the host.
After every chunk the device health flags of the resident simulations are
read; one whose norm became non-finite (it diverged) is retired at once with
``Request.error`` set, and the others carry on.
"""

import concurrent.futures
import time
from dataclasses import dataclass, field
from typing import Any, List, Optional

import numpy as np

from . import engine
from . import units as u
from .exceptions import DeviceError

# Pool bytes a simulation needs beyond its footprint: every device block is
# rounded up to 256 B and a simulation holds at most a dozen blocks.
_ALLOC_SLACK = 16 << 10


@dataclass
class Request:
    rid: int
    kind: str          # label of the request class (for reporting)
    arrival: float     # seconds from the start of the run
    make: Any          # callable returning a compiled simulation (engine.Compiled*)
    compiled: Any = None
    slot: Optional[int] = None
    sim_kind: Optional[int] = None
    admitted_replay: Optional[int] = None
    t_admit: Optional[float] = None
    t_done: Optional[float] = None
    g: Optional[np.ndarray] = None
    norms: Optional[np.ndarray] = None
    ips: Optional[np.ndarray] = None
    error: Optional[str] = None  # set when the simulation was retired early (it diverged)

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.arrival

    @property
    def work(self):
        """Mesh-point-steps of the request."""
        c = self.compiled
        return c.mesh_points * c.n_steps

    def replays_to_finish(self):
        """Graph replays from admission until the last step has run."""
        c = self.compiled
        return c.n_steps if self.sim_kind == engine.KIND_CYL else c.n_steps + 1


def _kind_of(c):
    if isinstance(c, engine.CompiledSH):
        return engine.KIND_SH
    if isinstance(c, engine.CompiledCyl):
        return engine.KIND_CYL
    return engine.KIND_LINE


@dataclass
class ServeReport:
    requests: List[Request]
    makespan_s: float
    replays: int
    rebuilds: int
    latencies_s: np.ndarray = field(default=None)

    @property
    def p50(self):
        return float(np.percentile(self.latencies_s, 50))

    @property
    def p99(self):
        return float(np.percentile(self.latencies_s, 99))

    @property
    def throughput(self):
        """Aggregate mesh-point-steps per second over the makespan."""
        return sum(r.work for r in self.requests) / self.makespan_s


class Server:
    """
    One device context serving a stream of simulations.

    pool_bytes: device memory the context reserves; admission keeps the sum of
    resident footprints within it. chunk_steps: graph replays between
    scheduling points (admissions, retirements).
    """

    def __init__(self, pool_bytes: int, device: int = 0, chunk_steps: int = 25, compile_workers: int = 4,
                 keep_g: bool = True, ctx=None):
        # ctx: an existing engine.DeviceContext (or an object with its methods) to serve on
        self.ctx = engine.DeviceContext(device=device, pool_bytes=pool_bytes) if ctx is None else ctx
        self.pool_bytes = int(pool_bytes)
        self.chunk_steps = int(chunk_steps)
        self.keep_g = keep_g
        self._pool = concurrent.futures.ThreadPoolExecutor(max_workers=compile_workers)

    def close(self):
        self._pool.shutdown(wait=True)
        self.ctx.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @staticmethod
    def _reserve(c):
        return int(c.footprint_bytes) + _ALLOC_SLACK

    def _add(self, r):
        c = r.compiled
        r.sim_kind = _kind_of(c)
        add = {engine.KIND_SH: self.ctx.add_sh, engine.KIND_LINE: self.ctx.add_line,
               engine.KIND_CYL: self.ctx.add_cyl}[r.sim_kind]
        r.slot = add(c)

    def _diverged(self, resident):
        """Resident requests whose device health flags report a non-finite norm (one read per kind)."""
        out = []
        for kind in (engine.KIND_SH, engine.KIND_LINE, engine.KIND_CYL):
            rs = [r for r in resident if r.sim_kind == kind]
            if not rs:
                continue
            flags = self.ctx.flags(kind, 0, max(r.slot for r in rs) + 1)
            out += [r for r in rs if flags[r.slot] & engine.FLAG_NONFINITE]
        return out

    def _retire(self, r, now, error=None):
        ctx = self.ctx
        r.error = error
        if error is not None:
            # a diverged simulation is stopped where it is: its stored data and state go back as they are
            if r.sim_kind == engine.KIND_SH:
                res = ctx.download_sh(r.slot, with_g=self.keep_g)
                r.g, r.norms, r.ips = res.g, res.norms, res.ips
            elif r.sim_kind == engine.KIND_CYL:
                r.g, r.norms, r.ips = ctx.download_cyl(r.slot, with_g=self.keep_g)
            else:
                r.g, r.norms, r.ips = ctx.download_line(r.slot)
            ctx.remove(r.sim_kind, r.slot)
            r.t_done = now()
            return
        ctx.materialize(r.sim_kind, r.slot)
        if r.sim_kind == engine.KIND_SH:
            res = ctx.download_sh(r.slot, with_g=self.keep_g)
            r.g, r.norms, r.ips = res.g, res.norms, res.ips
        elif r.sim_kind == engine.KIND_CYL:
            r.g, r.norms, r.ips = ctx.download_cyl(r.slot, with_g=self.keep_g)
        else:
            r.g, r.norms, r.ips = ctx.download_line(r.slot)
        ctx.remove(r.sim_kind, r.slot)
        r.t_done = now()

    def run(self, requests: List[Request], max_seconds: float = 3600.0, wait_compiled: bool = False) -> ServeReport:
        """
        Serve `requests` (arrival times relative to now) to completion.
        wait_compiled: finish compiling every arrived request before each
        admission round (admission then follows arrival order exactly).
        """
        reqs = sorted(requests, key=lambda r: (r.arrival, r.rid))
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        futures = {}
        next_arrival = 0
        waiting = []      # compiled, not admitted (arrival order)
        resident = []
        reserved = 0
        replays = 0
        rebuilds = 0
        retired = False
        done = 0
        while done < len(reqs):
            t = now()
            if t > max_seconds:
                raise DeviceError(f"serving did not finish within {max_seconds} s")
            # arrivals: start compiling
            while next_arrival < len(reqs) and reqs[next_arrival].arrival <= t:
                r = reqs[next_arrival]
                futures[r.rid] = (r, self._pool.submit(r.make))
                next_arrival += 1
            if wait_compiled and futures:
                concurrent.futures.wait([f for _, f in futures.values()])
            # compiled requests join the admission queue in arrival order
            for rid in sorted(futures, key=lambda i: (futures[i][0].arrival, i)):
                r, fut = futures[rid]
                if fut.done():
                    r.compiled = fut.result()
                    waiting.append(r)
                    del futures[rid]
            waiting.sort(key=lambda r: (r.arrival, r.rid))
            # admission by footprint, first come first served
            changed = False
            while waiting and reserved + self._reserve(waiting[0].compiled) <= self.pool_bytes:
                r = waiting.pop(0)
                try:
                    self._add(r)
                except DeviceError:
                    waiting.insert(0, r)  # fragmentation: retry after the next retirement
                    break
                reserved += self._reserve(r.compiled)
                r.admitted_replay = replays
                r.t_admit = now()
                resident.append(r)
                changed = True
            if not resident:
                if waiting and not changed:
                    raise DeviceError(f"request {waiting[0].rid} does not fit the device pool")
                time.sleep(0.001)
                continue
            if changed or retired:
                rebuilds += 1  # the resident set changed: the next replay captures a new step graph
            # run until the next simulation finishes or the chunk ends
            left = min(r.admitted_replay + r.replays_to_finish() - replays for r in resident)
            k = max(1, min(self.chunk_steps, left))
            self.ctx.run(k)
            replays += k
            # retire finished simulations, and diverged ones with an error (the others go on)
            retired = False
            still = []
            diverged = {id(r) for r in self._diverged(resident)}
            for r in resident:
                if id(r) in diverged:
                    self._retire(r, now, error="non-finite norm: the simulation diverged")
                    reserved -= self._reserve(r.compiled)
                    done += 1
                    retired = True
                elif replays - r.admitted_replay >= r.replays_to_finish():
                    self._retire(r, now)
                    reserved -= self._reserve(r.compiled)
                    done += 1
                    retired = True
                else:
                    still.append(r)
            resident = still
        makespan = max(r.t_done for r in reqs) - min(r.arrival for r in reqs)
        lat = np.array([r.latency for r in reqs])
        return ServeReport(requests=reqs, makespan_s=makespan, replays=replays, rebuilds=rebuilds, latencies_s=lat)


def _line_ground_state(z_points=8192):
    """Ground state of the line_cn workload (imaginary-time CN on the device), computed once."""
    from . import potentials, workloads

    soft = potentials.SoftCoulombPotential(softening_distance=np.sqrt(2.0) * u.bohr_radius)
    return workloads.line_ground_state(z_bound=200 * u.bohr_radius, z_points=z_points, test_mass=u.electron_mass,
                                       potential=lambda z: soft(r=z, test_charge=u.electron_charge))


def served_mix(n: int, rate_per_s: float, seed: int = 0, *, single_steps: Optional[int] = None,
               scan_steps: Optional[int] = None, line_steps: Optional[int] = None,
               line_ground_state=None) -> List[Request]:
    """
    A seeded request stream: exponential inter-arrival times at `rate_per_s`,
    each request drawn as 70 % sh_single-like (200 l x 1000 r with a drawn
    fluence and phase), 20 % an sh_scan draw and 10 % line_cn (8192 x 4000).
    `single_steps`, `scan_steps` and `line_steps` shorten the runs of each
    class to that many steps at the same mesh sizes (tests); the Line ground
    state is computed once and shared.
    """
    from . import workloads

    rng = np.random.default_rng(seed)
    scan = workloads.sh_scan_params(n, seed + 1)
    kinds = []
    for _ in range(n):
        kinds.append((float(rng.exponential(1.0 / rate_per_s)), float(rng.uniform())))
    if line_ground_state is None and any(c >= 0.9 for _, c in kinds):
        line_ground_state = _line_ground_state()
    t = 0.0
    out = []
    for i in range(n):
        dt, c = kinds[i]
        t += dt
        if c < 0.7:
            fl = float(np.exp(rng.uniform(np.log(0.1), np.log(10)))) * u.Jcm2
            ph = float(rng.uniform(0, np.pi))

            def make(fl=fl, ph=ph, i=i):
                spec = workloads.sh_single_spec(f"sh_single_{i}", fluence=fl, phase=ph, store_data_every=100,
                                                numeric_eigenstate_max_energy=0.0)
                if single_steps is not None:
                    workloads.fixed_step_count(spec, single_steps)
                sim = spec.to_sim()
                return spec.evolution_method.compile(sim)

            out.append(Request(i, "sh_single", t, make))
        elif c < 0.9:
            p = scan[i]

            def make(p=p):
                spec = workloads.sh_scan_spec(p)
                if scan_steps is not None:
                    workloads.fixed_step_count(spec, scan_steps)
                sim = spec.to_sim()
                return sim.spec.evolution_method.compile(sim)

            out.append(Request(i, "sh_scan", t, make))
        else:
            def make(i=i):
                spec = workloads.line_cn_spec(f"line_{i}", z_points=8192, n_steps=line_steps or 4000,
                                              store_data_every=100, ground_state=line_ground_state)
                sim = spec.to_sim()
                return spec.evolution_method.compile(sim)

            out.append(Request(i, "line_cn", t, make))
    return out
