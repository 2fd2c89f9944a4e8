#!/usr/bin/env python
"""
Benchmark: mesh-point-steps per second of the batched SphericalHarmonicMesh
scan (256 simulations, one hipGraph replay advances every simulation by one
step) versus the plain path (one simulation, direct kernel launches, no
graph), complex128, on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
    python bench.py --calibrate-copy ...      (profiling runs: adds a copy of known size)
    python bench.py --summarize-profiles DIR  (CPU: rocprofv3 databases -> profiles/)

Prints one JSON line. ``value`` is the batched throughput with every
wavefunction resident in HBM, timed over a window in which every simulation
is at its pulse peak and takes one data step: each simulation starts (from its
initial state, uploaded at that step of its own time grid) so that the timed
steps are centred on its data time nearest its pulse peak. Beside it: the same
at store_data_every = 1 (the reference default) centred on the peak, the first
steps of every run (weak fields, no data step), and the whole scan run to its
end. ``plain_value`` times the 200 x 1000 simulation alone over the same
window relative to its peak, at the same cadence. ``roofline`` is for the kernel with the largest
share of device time, from HIP events recorded on the library's stream
around every launch of a profiled pass over K more steps; ``cpu_baseline``
is the CPU oracle (a restatement of the reference's evolve step) timed on
one host core for a bounded sample of the same workload; ``cpu_baseline_parallel``
runs it in --cpu-procs processes at once, one scan simulation each (the
reference runs one simulation per process), before this process touches the GPU.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
COPY_CAL_BYTES = 1 << 30  # read and written once per calibration copy launch
COPY_CAL_LAUNCHES = 3
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "sh_scan_hbm_traffic.txt")
BYTES_PER_POINT_STEP = 64  # two complex128 read+write sweeps (L-phase, R-phase)
BYTES_PER_POINT_PHASE = 32


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_scan(n_sims, seed, store_every):
    from ionization_amd import workloads

    params = workloads.sh_scan_params(n_sims, seed)
    compiled = []
    for p in params:
        sim = workloads.sh_scan_spec(p, store_data_every=store_every).to_sim()
        compiled.append(sim.spec.evolution_method.compile(sim))
    return params, compiled


def peak_starts(compiled, steps, warmup, store_every):
    """
    First step of each simulation for a timed window centred on its pulse peak (t = 0, the middle
    of its symmetric time grid): on the data time nearest the peak when data are stored every
    `store_every` > 1 steps (so the window holds exactly one data step, at most 50 steps from the
    peak for a cadence of 100), on the peak itself at a cadence of 1. The warm-up steps come first.
    """
    out = []
    for c in compiled:
        peak = c.n_steps // 2
        centre = int(round(peak / store_every)) * store_every if store_every > 1 else peak
        out.append(max(0, centre - steps // 2 - warmup))
    return out


def time_batched(compiled, steps, warmup, starts=None):
    """Graph-replayed steps of every simulation at once; `starts` puts each simulation's initial
    state at that step of its time grid first (see peak_starts)."""
    from ionization_amd.engine import DeviceContext, KIND_SH, from_positions

    pool = int(sum(c.footprint_bytes for c in compiled) * 1.1) + (256 << 20)
    ctx = DeviceContext(pool_bytes=pool)
    slots = [ctx.add_sh(c) for c in compiled]
    if starts is not None:
        for c, s, st in zip(compiled, slots, starts):
            if st > 0:
                ctx.upload_state(KIND_SH, s, from_positions(c.g0, c.r_points, c.P, c.m), st)
    points = sum(c.mesh_points for c in compiled)
    ctx.run(warmup)
    ctx.sync()
    t0 = time.perf_counter()
    dev_ms = ctx.run(steps, use_graph=True)
    ctx.sync()
    wall = time.perf_counter() - t0
    # profiled pass: events around every L-phase / R-phase launch
    ctx.profile(True)
    ctx.run(steps, use_graph=False)
    prof = ctx.profile_read()
    prof["vphase_ms"], prof["vphase_launches"] = ctx.profile_kernel(4)
    prof["rpair_ms"], prof["rpair_launches"] = ctx.profile_kernel(5)
    ctx.profile(False)
    used, total = ctx.pool_usage()
    ctx.close()
    return dict(points=points, wall_s=wall, dev_ms=dev_ms, prof=prof, pool_used=used)


def time_velocity(params, store_every, steps, warmup):
    """
    The same scan draws in the velocity gauge (SphericalHarmonicVelocityGaugeOperators +
    SplitInteractionOperator): per step the L-phase (h1 rotations, mask), the h2
    similarity factors of the even l-pairs, the R-phase (the odd l-pairs solved together
    with their h2 factors where two lines fit one workgroup), and the even pairs' factors
    again: four read+write sweeps of g (128 B per mesh point and step; six where the odd
    pairs' factors run as kernels of their own).
    """
    from ionization_amd import mesh, workloads

    t = time.perf_counter()
    compiled = []
    for p in params:
        sim = workloads.sh_scan_spec(p, store_data_every=store_every,
                                     operators=mesh.SphericalHarmonicVelocityGaugeOperators()).to_sim()
        compiled.append(sim.spec.evolution_method.compile(sim))
    build_s = time.perf_counter() - t
    b = time_batched(compiled, steps, warmup, peak_starts(compiled, steps, warmup, store_every))
    rate = b["points"] * steps / b["wall_s"]
    prof = b["prof"]
    # h2 sweeps of k_sh_vphase read and write the rows inside l-pairs, before and after the solve:
    # the even pairs always, the odd pairs only where they are not solved with the R-phase
    fused = lambda c: 2 * c.P <= 512
    h2_rows = lambda c: 2 * (c.l_bound // 2) + (0 if fused(c) else 2 * ((c.l_bound - 1) // 2))
    h2_bytes = sum(2 * 32 * c.r_points * h2_rows(c) for c in compiled)
    pts = sum(c.r_points * c.l_bound for c in compiled)
    bytes_step = 64 * pts + h2_bytes  # L-phase and R-phase sweeps plus the separate h2 sweeps
    h2_ms = prof["vphase_ms"] / steps
    out = {
        "workload": f"sh_scan draws ({len(compiled)} simulations), velocity gauge, one hipGraph replay per step, "
                    "the window on each simulation's data time nearest its pulse peak",
        "value": rate, "unit": "mesh-point-steps/s", "ms_per_step": b["wall_s"] * 1e3 / steps,
        "bytes_per_point_step": bytes_step / pts, "step_hbm_gbs": rate * (bytes_step / pts) / 1e9, "build_s": build_s,
        "kernels_ms_per_step": {"k_sh_lphase": prof["lphase_ms"] / steps, "k_sh_rphase": prof["rphase_ms"] / steps,
                                "k_sh_rphase_vpair": prof["rpair_ms"] / steps, "k_sh_vphase": h2_ms},
        "k_sh_vphase_gbs": h2_bytes / (h2_ms * 1e-3) / 1e9 if h2_ms > 0 else None,
    }
    log(f"velocity gauge: {out['ms_per_step']:.3f} ms/step, {rate / 1e9:.2f} G mesh-point-steps/s; "
        f"h2 factors {h2_ms:.3f} ms/step at {out['k_sh_vphase_gbs'] or 0:.0f} GB/s")
    return out


def time_sh_adi(params, store_every, steps, warmup):
    """
    The same scan draws with AlternatingDirectionImplicit (the reference benchmark's
    SphericalHarmonic_LAG_CN_LEN): per step the L-phase (mask, observables), explicit H0
    along r, the l-direction Cayley factor (a Thomas sweep per position that keeps one real
    pivot factor per row in the idle state buffer), implicit H0 along r.
    """
    from ionization_amd import mesh, workloads

    t = time.perf_counter()
    compiled = []
    for p in params:
        sim = workloads.sh_scan_spec(p, store_data_every=store_every,
                                     evolution_method=mesh.AlternatingDirectionImplicit()).to_sim()
        compiled.append(sim.spec.evolution_method.compile(sim))
    build_s = time.perf_counter() - t
    b = time_batched(compiled, steps, warmup, peak_starts(compiled, steps, warmup, store_every))
    rate = b["points"] * steps / b["wall_s"]
    out = {"workload": f"sh_scan draws ({len(compiled)} simulations), AlternatingDirectionImplicit, one hipGraph "
                       "replay per step, the window on each simulation's data time nearest its pulse peak",
           "value": rate, "unit": "mesh-point-steps/s",
           "ms_per_step": b["wall_s"] * 1e3 / steps, "build_s": build_s}
    log(f"SH ADI: {out['ms_per_step']:.3f} ms/step, {rate / 1e9:.2f} G mesh-point-steps/s")
    return out


def time_reference_specs(cpu_seconds=0.0):
    """
    The specifications of the reference's own benchmark (benchmarks/benchmark_mesh_evolution.py:
    43-166) that are on the device path, built and run one at a time through the drop-in API
    and counted the way it counts: spacetime points = len(times) x mesh points (:22-23) over the
    process CPU time of spec.to_sim() + sim.run() (:14-19, 36), with the wall time beside it.
    """
    from ionization_amd import mesh, potentials, states
    from ionization_amd import units as u

    energy_spacing, test_mass = 0.1 * u.eV, u.electron_mass
    qho = potentials.HarmonicOscillator.from_energy_spacing_and_mass(energy_spacing=energy_spacing, mass=test_mass)
    sine = potentials.SineWave.from_photon_energy(photon_energy=energy_spacing, amplitude=0.0001 * u.atomic_electric_field)
    line_kw = dict(time_initial=0, time_final=1 * sine.period, time_step=0.001 * sine.period, internal_potential=qho,
                   initial_state=states.QHOState.from_potential(qho, n=0, mass=test_mass), z_bound=100 * u.nm,
                   z_points=2 ** 12, store_data_every=-1, test_mass=test_mass, electric_potential=sine)
    three_d = dict(time_initial=0, time_final=1 * u.fsec, time_step=1 * u.asec, store_data_every=-1)
    sh_kw = dict(three_d, r_bound=50 * u.bohr_radius, r_points=250, l_bound=100, use_numeric_eigenstates=False)
    cyl_kw = dict(three_d, rho_bound=50 * u.bohr_radius, rho_points=100, z_bound=50 * u.bohr_radius, z_points=250)
    specs = [
        mesh.LineSpecification("Line_CN_LEN", operators=mesh.LineLengthGaugeOperators(),
                               evolution_method=mesh.AlternatingDirectionImplicit(), **line_kw),
        mesh.LineSpecification("Line_SO_LEN", operators=mesh.LineLengthGaugeOperators(),
                               evolution_method=mesh.SplitInteractionOperator(), **line_kw),
        mesh.LineSpecification("Line_SO_VEL", operators=mesh.LineVelocityGaugeOperators(),
                               evolution_method=mesh.SplitInteractionOperator(), **line_kw),
        mesh.CylindricalSliceSpecification("CylindricalSlice_CN_LEN", evolution_method=mesh.AlternatingDirectionImplicit(),
                                           **cyl_kw),
        mesh.SphericalHarmonicSpecification("SphericalHarmonic_LAG_CN_LEN",
                                            operators=mesh.SphericalHarmonicLengthGaugeOperators(),
                                            evolution_method=mesh.AlternatingDirectionImplicit(), **sh_kw),
        mesh.SphericalHarmonicSpecification("SphericalHarmonic_LAG_SO_LEN",
                                            operators=mesh.SphericalHarmonicLengthGaugeOperators(),
                                            evolution_method=mesh.SplitInteractionOperator(), **sh_kw),
        mesh.SphericalHarmonicSpecification("SphericalHarmonic_LAG_SO_VEL",
                                            operators=mesh.SphericalHarmonicVelocityGaugeOperators(),
                                            evolution_method=mesh.SplitInteractionOperator(), **sh_kw),
    ]
    out = {}
    if cpu_seconds > 0:
        # the oracle (a restatement of the reference's per-step sparse builds, dia products and Thomas
        # solves) on the SH split-operator spec for a bounded number of steps: the CPU side of the table
        from oracle import mesh_cpu

        sp = specs[5]
        sim0 = sp.to_sim()
        orc = mesh_cpu.SHOracle(r_bound=sp.r_bound, r_points=sp.r_points, l_bound=sp.l_bound,
                                potential=lambda r: sp.internal_potential(r=r, test_charge=sp.test_charge),
                                field=lambda t: sp.electric_potential.get_electric_field_amplitude(t))
        g, steps, t0 = np.array(sim0.mesh.g), 0, time.perf_counter()
        while time.perf_counter() - t0 < cpu_seconds and steps < len(sim0.times) - 1:
            g = orc.evolve(g, sim0.times[steps + 1] - sim0.times[steps], sim0.times[steps + 1], sp.time_step)
            steps += 1
        out["cpu_port"] = {"spec": sp.name, "mesh_point_steps_per_s": g.size * steps / (time.perf_counter() - t0),
                           "sample": f"{steps} steps, one core, evolve only"}
        log(f"reference benchmark spec {sp.name} on the CPU oracle: "
            f"{out['cpu_port']['mesh_point_steps_per_s'] / 1e6:.2f} M mesh-point-steps/s")
    for spec in specs:
        w0, c0 = time.perf_counter(), time.process_time()
        sim = spec.to_sim()
        sim.run()
        wall, cpu = time.perf_counter() - w0, time.process_time() - c0
        pts = len(sim.times) * int(np.prod(sim.mesh.g.shape))
        out[spec.name] = {"spacetime_points": pts, "cpu_s": cpu, "wall_s": wall,
                          "pts_per_cpu_s": pts / cpu, "pts_per_wall_s": pts / wall}
        log(f"reference benchmark spec {spec.name}: {pts / wall / 1e6:.1f} M spacetime points per wall second "
            f"({pts / cpu / 1e6:.1f} M per CPU second)")
    return out


def time_full_run(compiled, build_cpu_s):
    """
    The whole scan: every simulation from its first to its last step in one batch (the batch thins
    out as the shorter runs finish). Also counted as the reference benchmark counts
    (benchmarks/benchmark_mesh_evolution.py:14-23, 36): len(times) x mesh points of every
    simulation over the process CPU time of building and running them.
    """
    from ionization_amd.engine import DeviceContext

    pool = int(sum(c.footprint_bytes for c in compiled) * 1.1) + (256 << 20)
    c0 = time.process_time()
    ctx = DeviceContext(pool_bytes=pool)
    for c in compiled:
        ctx.add_sh(c)
    total = max(c.n_steps for c in compiled) + 1
    work = sum(c.mesh_points * c.n_steps for c in compiled)
    ctx.sync()
    t0 = time.perf_counter()
    done = 0
    while done < total:
        k = min(4000, total - done)
        ctx.run(k)
        done += k
    ctx.sync()
    wall = time.perf_counter() - t0
    cpu = time.process_time() - c0
    ctx.close()
    spacetime = sum(c.mesh_points * (c.n_steps + 1) for c in compiled)
    out = {"workload": f"the {len(compiled)} scan simulations, each from its first to its last step in one batch",
           "value": work / wall, "unit": "mesh-point-steps/s", "wall_s": wall, "replays": total,
           "reference_convention": {"spacetime_points": spacetime, "cpu_s": build_cpu_s + cpu,
                                    "pts_per_cpu_s": spacetime / (build_cpu_s + cpu)}}
    log(f"full run: {total} replays in {wall:.1f} s, {work / wall / 1e9:.2f} G mesh-point-steps/s; "
        f"{spacetime / (build_cpu_s + cpu) / 1e9:.2f} G spacetime points per CPU second with setup")
    return out


def time_plain(steps, warmup, store_every, repeats=5):
    """sh_single (200 l x 1000 r) alone with direct launches, over the window peak_starts gives it at
    the same cadence as the batch; the median of `repeats` passes over that window."""
    import dataclasses

    from ionization_amd import workloads
    from ionization_amd.engine import DeviceContext, KIND_SH, from_positions

    sim = workloads.sh_single_spec(store_data_every=store_every).to_sim()
    comp = dataclasses.replace(sim.spec.evolution_method.compile(sim), store_every=store_every)
    start = peak_starts([comp], steps, warmup, store_every)[0]
    g0 = from_positions(comp.g0, comp.r_points, comp.P, comp.m)
    ctx = DeviceContext(pool_bytes=comp.footprint_bytes * 2 + (64 << 20))
    slot = ctx.add_sh(comp)
    walls = []
    for _ in range(repeats):
        ctx.upload_state(KIND_SH, slot, g0, start)
        ctx.run(warmup, use_graph=False)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.run(steps, use_graph=False)
        ctx.sync()
        walls.append(time.perf_counter() - t0)
    ctx.close()
    return dict(points=comp.mesh_points, wall_s=float(np.median(walls)), start=start)


def time_cyl(n_sims, seed, steps, warmup):
    """The batched CylindricalSliceMesh ADI workload: one graph replay advances all simulations."""
    from ionization_amd import workloads
    from ionization_amd.engine import DeviceContext

    compiled = workloads.cyl_scan_compiled(workloads.cyl_scan_params(n_sims, seed))
    pool = int(sum(c.footprint_bytes for c in compiled) * 1.1) + (256 << 20)
    ctx = DeviceContext(pool_bytes=pool)
    for c in compiled:
        ctx.add_cyl(c)
    points = sum(c.mesh_points for c in compiled)
    padded = sum(c.npz * c.nrpad for c in compiled)
    ctx.run(warmup)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.run(steps, use_graph=True)
    ctx.sync()
    wall = time.perf_counter() - t0
    ctx.profile(True)
    ctx.run(steps, use_graph=False)
    zk = ctx.profile_kernel(2)
    rk = ctx.profile_kernel(3)
    ctx.profile(False)
    ctx.close()
    rate = points * steps / wall
    out = {"workload": f"cyl_scan: {n_sims} CylindricalSliceMesh simulations (z 2000 x rho 1000), seed {seed}, "
                       "ADI length gauge, dt = 1 as, one hipGraph replay per step",
           "value": rate, "unit": "mesh-point-steps/s", "ms_per_step": wall * 1e3 / steps,
           "mesh_points": points, "step_hbm_gbs": rate * BYTES_PER_POINT_STEP / 1e9, "kernels": {}}
    for name, (ms, n) in (("k_cyl_z", zk), ("k_cyl_rho", rk)):
        if n:
            avg = ms / n
            by = BYTES_PER_POINT_PHASE * points * steps / n
            out["kernels"][name] = {"avg_ms": avg, "algorithmic_bytes": by, "achieved_gbs": by / (avg * 1e-3) / 1e9,
                                    "padded_points_frac": padded / points}
    log(f"cyl: {out['ms_per_step']:.3f} ms/step, {rate / 1e9:.2f} G mesh-point-steps/s")
    return out


def time_line(n_batch, cpu_seconds):
    """
    LineMesh CN, 8192 points x 4000 steps (soft Coulomb, sine field, one-sided
    mask): the single simulation as one launch (latency), a batch of n_batch
    copies in one launch (throughput), and the oracle on one core for a bounded
    number of steps.
    """
    from ionization_amd.engine import DeviceContext, KIND_LINE

    ctx = DeviceContext(pool_bytes=(n_batch + 2) * (2 << 20) + (64 << 20))
    from ionization_amd import workloads

    spec = workloads.line_cn_spec(z_points=8192, n_steps=4000, store_data_every=100, ctx=ctx)
    sim = spec.to_sim()
    comp = spec.evolution_method.compile(sim)
    slot = ctx.add_line(comp)
    ms_single = ctx.line_run(comp.n_steps)
    ctx.remove(KIND_LINE, slot)
    slots = [ctx.add_line(comp) for _ in range(n_batch)]
    ms_batch = ctx.line_run(comp.n_steps)
    for s in slots:
        ctx.remove(KIND_LINE, s)
    # the same workload with SplitInteractionOperator (the reference benchmark's Line_SO_LEN / Line_SO_VEL)
    from ionization_amd import mesh

    so = {}
    for gauge, ops in (("length", mesh.LineLengthGaugeOperators()), ("velocity", mesh.LineVelocityGaugeOperators())):
        sspec = workloads.line_cn_spec(z_points=8192, n_steps=4000, store_data_every=100,
                                       ground_state=np.asarray(spec.initial_state.g), operators=ops,
                                       evolution_method=mesh.SplitInteractionOperator())
        ssim = sspec.to_sim()
        sc = sspec.evolution_method.compile(ssim)
        s1 = ctx.add_line(sc)
        ms = ctx.line_run(sc.n_steps)
        ctx.remove(KIND_LINE, s1)
        so[gauge] = {"single_ms": ms, "single_value": sc.z_points * sc.n_steps / (ms * 1e-3)}
    ctx.close()
    work = comp.z_points * comp.n_steps
    out = {"workload": "line_cn: LineMesh 8192 points x 4000 steps, soft Coulomb, sine field, CN (one launch)",
           "single_ms": ms_single, "single_value": work / (ms_single * 1e-3),
           "batch": n_batch, "batch_ms": ms_batch, "batch_value": n_batch * work / (ms_batch * 1e-3),
           "unit": "mesh-point-steps/s", "split_operator": so}
    if cpu_seconds > 0:
        from oracle import mesh_cpu

        orc = mesh_cpu.LineOracle(
            z_bound=spec.z_bound, z_points=spec.z_points,
            potential=lambda z: spec.internal_potential(r=z, z=z, test_charge=spec.test_charge),
            field=lambda t: spec.electric_potential.get_electric_field_amplitude(t),
            mask=lambda z: spec.mask(r=z), test_mass=spec.test_mass)
        g = np.array(sim.mesh.g)
        steps, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < cpu_seconds and steps < comp.n_steps:
            g = orc.evolve(g, sim.times[steps + 1] - sim.times[steps], sim.times[steps + 1])
            steps += 1
        out["cpu_port_value"] = comp.z_points * steps / (time.perf_counter() - t0)
    log(f"line: single {ms_single:.1f} ms ({out['single_value'] / 1e9:.3f} G/s), "
        f"batch of {n_batch} {ms_batch:.1f} ms ({out['batch_value'] / 1e9:.2f} G/s); split operator single "
        f"{so['length']['single_ms']:.1f} ms (length), {so['velocity']['single_ms']:.1f} ms (velocity)")
    return out


def time_serve(n, rate, seed, pool_gb, fallback_pool_gb=96.0):
    """
    The served stream: Poisson arrivals (seeded) of 70 % sh_single-like, 20 %
    sh_scan draws and 10 % line_cn requests; compile on arrival, admission by
    device footprint, retirement as soon as the last step has run. The admission
    pool is `pool_gb` GiB of the GPU's HBM when it can be reserved, else the fallback.
    """
    from ionization_amd import serve
    from ionization_amd.exceptions import DeviceError

    gs = serve._line_ground_state()
    reqs = serve.served_mix(n, rate, seed, line_ground_state=gs)
    try:
        srv = serve.Server(pool_bytes=int(pool_gb * (1 << 30)), chunk_steps=25, compile_workers=8)
    except DeviceError as e:
        log(f"serve: a {pool_gb} GiB pool could not be reserved ({e}); using {fallback_pool_gb} GiB")
        pool_gb = fallback_pool_gb
        srv = serve.Server(pool_bytes=int(pool_gb * (1 << 30)), chunk_steps=25, compile_workers=8)
    with srv:
        rep = srv.run(reqs)
    kinds = {}
    for r in rep.requests:
        kinds.setdefault(r.kind, []).append(r.latency)
    out = {"workload": f"served_stream: {n} requests, Poisson {rate}/s (seed {seed}), 70 % sh_single-like "
                       "(200 x 1000, 8000 steps), 20 % sh_scan draws, 10 % line_cn (8192 x 4000); "
                       f"{pool_gb} GiB admission pool",
           "value": rep.throughput, "unit": "mesh-point-steps/s", "p50_s": rep.p50, "p99_s": rep.p99,
           "makespan_s": rep.makespan_s, "replays": rep.replays, "graph_rebuilds": rep.rebuilds,
           "pool_gib": pool_gb, "errors": sum(r.error is not None for r in rep.requests),
           "by_kind": {k: {"n": len(v), "p50_s": float(np.percentile(v, 50)), "max_s": float(np.max(v))}
                       for k, v in kinds.items()}}
    log(f"serve: {n} requests, p50 {rep.p50:.2f} s, p99 {rep.p99:.2f} s, {rep.throughput / 1e9:.2f} G/s, "
        f"{rep.rebuilds} graph rebuilds over {rep.replays} replays, {pool_gb} GiB pool")
    return out


def host_info():
    """The host the CPU baselines ran on: its logical CPUs, the ones this process may use, the model."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"nproc": os.cpu_count(), "usable_cpus": usable, "cpu_model": model}


def cpu_baseline(params, seconds):
    """The oracle on one core: one simulation of the scan, as many steps as fit in about `seconds`."""
    from ionization_amd import units as u
    from ionization_amd import workloads
    from oracle import mesh_cpu, physics

    p = params[0]
    sim = workloads.sh_scan_spec(p).to_sim()
    field = lambda t: physics.sinc_pulse_field(t, pulse_width=p["pulse_width"], fluence=p["fluence"], phase=p["phase"])
    rb = p["r_bound"]
    orc = mesh_cpu.SHOracle(r_bound=rb, r_points=p["r_points"], l_bound=p["l_bound"], potential=physics.coulomb,
                            field=field, mask=lambda r: physics.radial_cosine_mask(r, 0.9 * rb, rb))
    g = np.array(sim.mesh.g)
    times = sim.times
    dt = sim.spec.time_step
    steps = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds and steps < len(times) - 1:
        g = orc.evolve(g, times[steps + 1] - times[steps], times[steps + 1], dt)
        steps += 1
    el = time.perf_counter() - t0
    return dict(value=orc.mesh_points * steps / el, steps=steps, seconds=el,
                sample=f"scan simulation 0 ({p['l_bound']} l x {p['r_points']} r), {steps} steps")


def _cpu_worker(args):
    params, i, seconds = args
    c = cpu_baseline([params[i]], seconds)
    return c["value"]


def cpu_baseline_parallel(params, seconds, procs):
    """The oracle in `procs` forked processes at once, each on its own scan simulation (the
    reference runs one simulation per process); run before this process touches the GPU."""
    import multiprocessing as mp

    from ionization_amd import _lib

    if _lib._LIB is not None:
        raise RuntimeError("the parallel CPU baseline must run before the device library is loaded")
    with mp.get_context("fork").Pool(procs) as pool:
        rates = pool.map(_cpu_worker, [(params, i % len(params), seconds) for i in range(procs)])
    return dict(value=float(sum(rates)), procs=procs,
                sample=f"scan simulations 0..{procs - 1}, one per process, ~{seconds:.0f} s each")


def calibrate_copy():
    """The library's streaming copy of COPY_CAL_BYTES (k_stream_copy): the
    known-size kernel that scales the FETCH_SIZE / WRITE_SIZE counters to bytes;
    also the attainable copy bandwidth."""
    from ionization_amd.engine import DeviceContext

    ctx = DeviceContext(pool_bytes=2 * COPY_CAL_BYTES + (16 << 20))
    ms = ctx.copy_probe(COPY_CAL_BYTES, COPY_CAL_LAUNCHES)
    ctx.close()
    gbs = 2 * COPY_CAL_BYTES / (ms * 1e-3) / 1e9
    log(f"copy calibration: {COPY_CAL_LAUNCHES} x ({COPY_CAL_BYTES} B read + {COPY_CAL_BYTES} B written), "
        f"{ms:.3f} ms each = {gbs:.0f} GB/s")
    return gbs


def load_traffic():
    """HBM bytes per launch measured by the committed counter summary:
    lines 'traffic <kernel> <fetch_bytes> <write_bytes>' (per launch, calibrated)."""
    out = {}
    if not os.path.exists(TRAFFIC_FILE):
        return out
    with open(TRAFFIC_FILE) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 4 and parts[0] == "traffic":
                out[parts[1]] = (float(parts[2]), float(parts[3]))
    return out


def _kernel_family(name):
    base = name.replace("(anonymous namespace)::", "").split("(")[0].replace("void ", "").strip()
    return base.split("<")[0]


def summarize_profiles(d, command_trace, command_fetch, command_write, trace_steps=20, trace_warmup=5,
                       trace_cyl_steps=20):
    """Reads rocprofv3 databases under d/{trace,fetch,write} and writes the
    kernel-time and HBM-traffic summaries under profiles/."""
    import glob
    import sqlite3

    def db(sub):
        files = sorted(glob.glob(os.path.join(d, sub, "**", "*_results.db"), recursive=True))
        if not files:
            raise SystemExit(f"no rocprofv3 database under {os.path.join(d, sub)}")
        return sqlite3.connect(files[0])

    prof_dir = os.path.join(ROOT, "profiles")
    os.makedirs(prof_dir, exist_ok=True)

    con = db("trace")
    rows = con.execute("select name, count(*), sum(duration), avg(duration), min(duration), max(duration) "
                       "from kernels group by name order by sum(duration) desc").fetchall()
    fam = {}
    for name, cnt, tot, avg, mn, mx in rows:
        f = _kernel_family(name)
        c, t = fam.get(f, (0, 0.0))
        fam[f] = (c + cnt, t + tot)
    lines = ["# rocprofv3 kernel trace of bench.py: the 256-simulation SphericalHarmonicMesh scan over its",
             "# headline window (each simulation's steps around its data time nearest its pulse peak) and the",
             "# 32-simulation CylindricalSliceMesh workload", "",
             "command:", "    " + command_trace, "",
             f"{'kernel':<70} {'calls':>7} {'total ms':>10} {'avg us':>9} {'min us':>9} {'max us':>9}"]
    for name, cnt, tot, avg, mn, mx in rows:
        if tot < 1e5:
            continue
        lines.append(f"{name[:70]:<70} {cnt:>7d} {tot / 1e6:>10.3f} {avg / 1e3:>9.1f} {mn / 1e3:>9.1f} {mx / 1e3:>9.1f}")
    # bench.py times every kernel in a profiled pass of direct launches, one kernel at a time,
    # after the graph-replayed steps: those are the last dispatches of each family
    starts = {}
    for name, dur, st in con.execute("select name, duration, start from kernels order by start"):
        starts.setdefault(_kernel_family(name), []).append(dur)
    steps_of = {"k_sh_lphase": (trace_warmup + 2 * trace_steps, trace_steps),
                "k_sh_rphase": (trace_warmup + 2 * trace_steps, trace_steps),
                "k_cyl_z": (10 + 2 * trace_cyl_steps, trace_cyl_steps),
                "k_cyl_rho": (10 + 2 * trace_cyl_steps, trace_cyl_steps)}
    lines += ["", "per kernel family (all template instances); 'profiled pass' = the direct launches that",
              "bench.py times with HIP events (its kernels.*.avg_ms and roofline.achieved):",
              f"{'family':<30} {'calls':>7} {'total ms':>10} {'avg us':>9} {'profiled pass avg us':>22}"]
    for f, (c, t) in sorted(fam.items(), key=lambda kv: -kv[1][1]):
        if t < 1e5:
            continue
        extra = ""
        if f in steps_of and f in starts:
            total_steps, prof_steps = steps_of[f]
            per_step = len(starts[f]) // total_steps
            last = starts[f][-per_step * prof_steps:] if per_step else []
            if last:
                extra = f"{np.mean(last) / 1e3:>22.1f}"
        lines.append(f"{f:<30} {c:>7d} {t / 1e6:>10.3f} {t / c / 1e3:>9.1f}{extra}")
    with open(os.path.join(prof_dir, "sh_scan_kernel_trace.txt"), "w") as fh:
        fh.write("\n".join(lines) + "\n")

    def counter(sub, cname):
        con = db(sub)
        q = ("select kernel_name, count(*), sum(value) from counters_collection "
             "where counter_name = ? group by kernel_name")
        return {k: (c, v * 1024.0) for k, c, v in con.execute(q, (cname,)).fetchall()}  # KiB -> B

    fetch = counter("fetch", "FETCH_SIZE")
    write = counter("write", "WRITE_SIZE")

    def cal(table):
        cands = [(k, c, v) for k, (c, v) in table.items() if _kernel_family(k) == "k_stream_copy"]
        if not cands:
            raise SystemExit("calibration copy kernel not found (run the profiled command with --calibrate-copy)")
        k, c, v = cands[0]
        return COPY_CAL_BYTES * c / v, k

    f_scale, f_kernel = cal(fetch)
    w_scale, w_kernel = cal(write)
    fams = {}
    for table, idx in ((fetch, 0), (write, 1)):
        for k, (c, v) in table.items():
            f = _kernel_family(k)
            if not f.startswith("k_"):
                continue
            e = fams.setdefault(f, [0, 0.0, 0, 0.0])
            e[2 * idx] += c
            e[2 * idx + 1] += v
    lines = ["# HBM traffic per launch from rocprofv3 FETCH_SIZE / WRITE_SIZE (KiB per dispatch),",
             "# scaled by what the same counters read for a float64 copy of 1 GiB (read 1 GiB, write 1 GiB).", "",
             "commands:", "    " + command_fetch, "    " + command_write, "",
             f"calibration kernel: {f_kernel[:90]}",
             f"FETCH_SIZE scale = {f_scale:.4f} (1 GiB / counter reading), WRITE_SIZE scale = {w_scale:.4f}", "",
             f"{'kernel family':<20} {'launches':>8} {'fetch MB/launch':>16} {'write MB/launch':>16}"]
    traffic = []
    for f, (cf, vf, cw, vw) in sorted(fams.items()):
        if cf == 0 or cw == 0:
            continue
        fb = vf / cf * f_scale
        wb = vw / cw * w_scale
        lines.append(f"{f:<20} {cf:>8d} {fb / 1e6:>16.2f} {wb / 1e6:>16.2f}")
        traffic.append(f"traffic {f} {fb:.6e} {wb:.6e}")
    lines += ["", "machine-readable (bytes per launch, calibrated):"] + traffic
    with open(TRAFFIC_FILE, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    print(open(os.path.join(prof_dir, "sh_scan_kernel_trace.txt")).read())
    print(open(TRAFFIC_FILE).read())


def summarize_velocity_profiles(d, command_trace, command_fetch, command_write, sims=256, seed=0,
                                trace_steps=20, trace_warmup=5, counter_steps=3, counter_warmup=1):
    """Velocity-gauge scan: kernel trace by full kernel name (the velocity L-phase is the
    instance k_sh_lphase<*, true>) from d/vtrace, and the h2 kernel's calibrated HBM traffic per
    step from d/vfetch and d/vwrite against its algorithmic bytes."""
    import glob
    import sqlite3

    from ionization_amd import workloads

    def db(sub):
        files = sorted(glob.glob(os.path.join(d, sub, "**", "*_results.db"), recursive=True))
        if not files:
            raise SystemExit(f"no rocprofv3 database under {os.path.join(d, sub)}")
        return sqlite3.connect(files[0])

    prof_dir = os.path.join(ROOT, "profiles")
    con = db("vtrace")
    rows = con.execute("select name, count(*), sum(duration), avg(duration), min(duration), max(duration) "
                       "from kernels group by name order by sum(duration) desc").fetchall()
    steps = trace_warmup + 2 * trace_steps  # warm-up, graph replays, profiled direct launches
    lines = ["# rocprofv3 kernel trace of bench.py with the velocity-gauge scan (the same 256 draws,",
             "# SphericalHarmonicVelocityGaugeOperators); the length-gauge scan runs first in the same process",
             "", "command:", "    " + command_trace, "",
             f"{'kernel':<72} {'calls':>7} {'total ms':>10} {'avg us':>9} {'min us':>9} {'max us':>9}"]
    vph = [0, 0.0]
    for name, cnt, tot, avg, mn, mx in rows:
        if _kernel_family(name) == "k_sh_vphase":
            vph[0] += cnt
            vph[1] += tot
        if tot < 1e5:
            continue
        lines.append(f"{name[:72]:<72} {cnt:>7d} {tot / 1e6:>10.3f} {avg / 1e3:>9.1f} {mn / 1e3:>9.1f} {mx / 1e3:>9.1f}")
    lines += ["", f"k_sh_vphase, all instances: {vph[0]} launches, {vph[1] / 1e6:.3f} ms over {steps} steps = "
              f"{vph[1] / 1e6 / steps:.3f} ms per step"]
    with open(os.path.join(prof_dir, "sh_velocity_scan_kernel_trace.txt"), "w") as fh:
        fh.write("\n".join(lines) + "\n")

    def counter(sub, cname):
        q = ("select kernel_name, count(*), sum(value) from counters_collection "
             "where counter_name = ? group by kernel_name")
        return {k: (c, v * 1024.0) for k, c, v in db(sub).execute(q, (cname,)).fetchall()}

    fetch, write = counter("vfetch", "FETCH_SIZE"), counter("vwrite", "WRITE_SIZE")

    def cal(table):
        for k, (c, v) in table.items():
            if _kernel_family(k) == "k_stream_copy":
                return COPY_CAL_BYTES * c / v
        raise SystemExit("calibration copy kernel not found (run the profiled command with --calibrate-copy)")

    fs, ws = cal(fetch), cal(write)
    csteps = counter_warmup + 2 * counter_steps
    from ionization_amd import engine

    def h2_rows(p):  # rows of k_sh_vphase sweeps: even l-pairs, and odd ones unless fused with the R-phase
        P, _ = engine.choose_layout(p["r_points"], "sh")
        return 2 * (p["l_bound"] // 2) + (0 if 2 * P <= 512 else 2 * ((p["l_bound"] - 1) // 2))

    alg = sum(2 * 32 * p["r_points"] * h2_rows(p) for p in workloads.sh_scan_params(sims, seed))
    fb = sum(v for k, (c, v) in fetch.items() if _kernel_family(k) == "k_sh_vphase") * fs / csteps
    wb = sum(v for k, (c, v) in write.items() if _kernel_family(k) == "k_sh_vphase") * ws / csteps
    lv = [(k, c, v) for k, (c, v) in fetch.items() if k.startswith("void k_sh_lphase<false, true>")]
    lw = [(k, c, v) for k, (c, v) in write.items() if k.startswith("void k_sh_lphase<false, true>")]
    lines = ["# HBM traffic of the velocity-gauge scan from rocprofv3 FETCH_SIZE / WRITE_SIZE (KiB per",
             "# dispatch), scaled by the same counters on the library's 1 GiB copy kernel", "",
             "commands:", "    " + command_fetch, "    " + command_write, "",
             f"FETCH_SIZE scale = {fs:.4f}, WRITE_SIZE scale = {ws:.4f}; {csteps} steps counted", "",
             f"k_sh_vphase (h2 factors, all instances) per step: fetch {fb / 1e6:.1f} MB, write {wb / 1e6:.1f} MB; "
             f"algorithmic {alg / 2e6:.1f} MB each way ({fb / (alg / 2):.2f}x / {wb / (alg / 2):.2f}x)"]
    if lv and lw:
        lines.append(f"k_sh_lphase<false, true> (velocity L-phase) per launch: fetch "
                     f"{lv[0][2] / lv[0][1] * fs / 1e6:.1f} MB, write {lw[0][2] / lw[0][1] * ws / 1e6:.1f} MB")
    pf = sum(v for k, (c, v) in fetch.items() if _kernel_family(k) == "k_sh_rphase_vpair") * fs / csteps
    pw = sum(v for k, (c, v) in write.items() if _kernel_family(k) == "k_sh_rphase_vpair") * ws / csteps
    pair_alg = sum(16 * p["r_points"] * 2 * ((p["l_bound"] - 1) // 2) for p in workloads.sh_scan_params(sims, seed)
                   if 2 * engine.choose_layout(p["r_points"], "sh")[0] <= 512)
    if pair_alg:
        lines.append(f"k_sh_rphase_vpair (odd l-pairs: R-phase with their h2 factors) per step: fetch {pf / 1e6:.1f} MB, "
                     f"write {pw / 1e6:.1f} MB; algorithmic {pair_alg / 1e6:.1f} MB each way "
                     f"({pf / pair_alg:.2f}x / {pw / pair_alg:.2f}x)")
    with open(os.path.join(prof_dir, "sh_velocity_scan_hbm_traffic.txt"), "w") as fh:
        fh.write("\n".join(lines) + "\n")
    print(open(os.path.join(prof_dir, "sh_velocity_scan_kernel_trace.txt")).read())
    print(open(os.path.join(prof_dir, "sh_velocity_scan_hbm_traffic.txt")).read())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--sims", type=int, default=256)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--store-every", type=int, default=100)
    ap.add_argument("--plain-steps", type=int, default=1,
                    help="0 skips the plain path (one simulation alone, timed over the same window)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-procs", type=int, default=16,
                    help="processes for the parallel CPU baseline (the host share of one GPU)")
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--autotune", action="store_true",
                    help="also measure every candidate SH layout per radial size class and time the scan "
                         "under the measured layouts (reported as 'autotuned')")
    ap.add_argument("--velocity", type=int, default=1, help="0 skips the velocity-gauge scan")
    ap.add_argument("--sh-adi", type=int, default=1, help="0 skips the SphericalHarmonicMesh ADI scan")
    ap.add_argument("--reference-specs", type=int, default=1,
                    help="0 skips the reference benchmark's own specifications (run one at a time)")
    ap.add_argument("--cyl-sims", type=int, default=32, help="0 skips the cylindrical ADI workload")
    ap.add_argument("--cyl-steps", type=int, default=100)
    ap.add_argument("--line-batch", type=int, default=256, help="0 skips the LineMesh workload")
    ap.add_argument("--serve-requests", type=int, default=96, help="0 skips the served stream")
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-pool-gb", type=float, default=256.0,
                    help="admission pool (GiB of HBM); 96 GiB if that much cannot be reserved")
    ap.add_argument("--full-run", type=int, default=1, help="0 skips running the whole scan to its end")
    ap.add_argument("--windows", choices=("all", "peak"), default="all",
                    help="peak: only the headline window (profiling runs)")
    ap.add_argument("--summarize-profiles", metavar="DIR")
    ap.add_argument("--summarize-velocity-profiles", metavar="DIR")
    ap.add_argument("--trace-command", default="")
    ap.add_argument("--fetch-command", default="")
    ap.add_argument("--write-command", default="")
    args = ap.parse_args()
    if args.summarize_velocity_profiles:
        summarize_velocity_profiles(args.summarize_velocity_profiles, args.trace_command, args.fetch_command,
                                    args.write_command, args.sims, args.seed, args.steps, args.warmup)
        return
    if args.summarize_profiles:
        summarize_profiles(args.summarize_profiles, args.trace_command, args.fetch_command, args.write_command,
                           args.steps, args.warmup, args.cyl_steps)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported: simulations are "
              "independent, so several GPUs would each run their own context over a share of the "
              "simulations with no exchange, and the whole-job rate is the sum", file=sys.stderr)
        sys.exit(2)

    t, c0 = time.perf_counter(), time.process_time()
    params, compiled = build_scan(args.sims, args.seed, args.store_every)
    build_cpu_s = time.process_time() - c0
    log(f"built {len(compiled)} simulations in {time.perf_counter() - t:.1f} s; "
        f"{sum(c.mesh_points for c in compiled) / 1e6:.1f} M mesh points")
    cpu_par = None
    if args.cpu_seconds > 0 and args.cpu_procs > 1:  # before any device work: the workers are forked
        cp = cpu_baseline_parallel(params, min(args.cpu_seconds, 10.0), args.cpu_procs)
        cpu_par = {"value": cp["value"], "unit": "mesh-point-steps/s", "cores": cp["procs"], "kind": "port",
                   "sample": cp["sample"]}
        log(f"cpu oracle, {cp['procs']} processes: {cp['value'] / 1e6:.2f} M mesh-point-steps/s")
    import dataclasses

    starts = peak_starts(compiled, args.steps, args.warmup, args.store_every)
    b = time_batched(compiled, args.steps, args.warmup, starts)
    rate = b["points"] * args.steps / b["wall_s"]
    log(f"batched, window on each data time nearest the pulse peak: {b['wall_s'] * 1e3 / args.steps:.3f} ms/step, "
        f"{rate / 1e9:.2f} G mesh-point-steps/s")
    every1_out = first_out = None
    if args.windows == "all":  # "peak" (profiling runs) launches the headline window's kernels only
        every1 = [dataclasses.replace(c, store_every=1) for c in compiled]
        b1 = time_batched(every1, args.steps, args.warmup, peak_starts(every1, args.steps, args.warmup, 1))
        rate1 = b1["points"] * args.steps / b1["wall_s"]
        log(f"batched, store_data_every = 1, window on the pulse peak: {b1['wall_s'] * 1e3 / args.steps:.3f} "
            f"ms/step, {rate1 / 1e9:.2f} G mesh-point-steps/s")
        del every1
        every1_out = {"value": rate1, "ms_per_step": b1["wall_s"] * 1e3 / args.steps,
                      "window": "centred on each simulation's pulse peak; every step is a data step"}
        b0 = time_batched(compiled, args.steps, args.warmup)
        rate0 = b0["points"] * args.steps / b0["wall_s"]
        log(f"batched, first steps of every run: {b0['wall_s'] * 1e3 / args.steps:.3f} ms/step, "
            f"{rate0 / 1e9:.2f} G mesh-point-steps/s")
        first_out = {"value": rate0, "ms_per_step": b0["wall_s"] * 1e3 / args.steps,
                     "window": f"steps {args.warmup}..{args.warmup + args.steps - 1} of every run "
                               "(weak fields, no data step)"}
    plain_rate = plain_rate1 = None
    if args.plain_steps > 0:  # 0 skips the plain pass (profiling runs)
        pl = time_plain(args.steps, args.warmup, args.store_every)
        plain_rate = pl["points"] * args.steps / pl["wall_s"]
        pl1 = time_plain(args.steps, args.warmup, 1)
        plain_rate1 = pl1["points"] * args.steps / pl1["wall_s"]
        log(f"plain, same window: {pl['wall_s'] * 1e3 / args.steps:.3f} ms/step, {plain_rate / 1e9:.3f} G "
            f"mesh-point-steps/s; at store_data_every = 1 {plain_rate1 / 1e9:.3f} G")
    if every1_out is not None:
        every1_out["plain_value"] = plain_rate1
        every1_out["batched_over_plain"] = None if plain_rate1 is None else every1_out["value"] / plain_rate1
    full =time_full_run(compiled, build_cpu_s) if args.full_run else None

    tuned = None
    if args.autotune:
        from ionization_amd import autotune, engine

        t = time.perf_counter()
        table, trial = autotune.tune_batch(compiled, log=log)
        tune_s = time.perf_counter() - t
        engine.set_sh_layout_table(table)
        relaid = [engine.relayout_sh(c) for c in compiled]
        engine.set_sh_layout_table({})
        bt = time_batched(relaid, args.steps, args.warmup)
        rate_t = bt["points"] * args.steps / bt["wall_s"]
        tuned = {"value": rate_t, "ms_per_step": bt["wall_s"] * 1e3 / args.steps, "tune_s": tune_s,
                 "size_classes": len(table), "trial_ms_per_step": trial,
                 "layouts_changed": sum((c.P, c.m) != (d.P, d.m) for c, d in zip(compiled, relaid)),
                 "table": {str(k): list(v) for k, v in sorted(table.items())}}
        log(f"autotuned: {tuned['ms_per_step']:.3f} ms/step, {rate_t / 1e9:.2f} G mesh-point-steps/s "
            f"({tuned['layouts_changed']} of {len(compiled)} layouts changed, tuning {tune_s:.1f} s)")

    prof = b["prof"]
    l_avg = prof["lphase_ms"] / max(1, prof["lphase_launches"])
    r_avg = prof["rphase_ms"] / max(1, prof["rphase_launches"])
    bytes_phase = BYTES_PER_POINT_PHASE * b["points"]
    l_bytes = bytes_phase  # one L-phase launch per step covers every simulation
    r_bytes = bytes_phase * args.steps / max(1, prof["rphase_launches"])  # R-phase launches split by layout bucket
    kern = {
        "k_sh_lphase": dict(avg_ms=l_avg, bytes=l_bytes, total_ms=prof["lphase_ms"]),
        "k_sh_rphase": dict(avg_ms=r_avg, bytes=r_bytes, total_ms=prof["rphase_ms"]),
    }
    dom_name = max(kern, key=lambda k: kern[k]["total_ms"])
    dom = kern[dom_name]
    achieved = dom["bytes"] / (dom["avg_ms"] * 1e-3) / 1e9
    measured = load_traffic().get(dom_name)
    traffic = None if measured is None else measured[0] + measured[1]
    step_gbs = rate * BYTES_PER_POINT_STEP / 1e9

    vel = time_velocity(params, args.store_every, args.steps, args.warmup) if args.velocity else None
    sh_adi = time_sh_adi(params, args.store_every, args.steps, args.warmup) if args.sh_adi else None
    ref_specs = time_reference_specs(min(args.cpu_seconds, 5.0)) if args.reference_specs else None
    cyl = time_cyl(args.cyl_sims, args.seed, args.cyl_steps, 10) if args.cyl_sims > 0 else None
    line = time_line(args.line_batch, min(args.cpu_seconds, 5.0)) if args.line_batch > 0 else None
    served = (time_serve(args.serve_requests, args.serve_rate, args.seed, args.serve_pool_gb)
              if args.serve_requests > 0 else None)

    copy_gbs = calibrate_copy()  # attainable streaming rate, beside the published peak

    cpu = None
    host = host_info()
    if cpu_par is not None:
        cpu_par["host"] = host
    if args.cpu_seconds > 0:
        c = cpu_baseline(params, args.cpu_seconds)
        cpu = {"value": c["value"], "unit": "mesh-point-steps/s", "cores": 1, "kind": "port", "sample": c["sample"],
               "host": host}
        log(f"cpu oracle: {c['value'] / 1e6:.2f} M mesh-point-steps/s ({c['sample']})")

    out = {
        "metric": "mesh-point-steps/sec batched vs plain on 1 GPU at <=1e-10 rel L2; p50/p99 latency",
        "value": rate,
        "unit": "mesh-point-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": b["wall_s"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex128",
        "data": "synthetic (seeded parameter draws; numeric 1s initial states)",
        "config": {
            "workload": f"sh_scan: {args.sims} SphericalHarmonicMesh simulations, seed {args.seed}, "
                        "split-operator length gauge, dt = 1 as, one hipGraph replay per step; each simulation "
                        "timed over the steps around its data time nearest its pulse peak",
            "simulations": args.sims,
            "mesh_points": b["points"],
            "store_data_every": args.store_every,
            "window": "each simulation starts from its initial state at step (d - steps // 2 - warmup) of its own "
                      "time grid, d = its data time nearest the pulse peak",
            "plain_workload": "sh_single: 200 l x 1000 r, one simulation, direct launches, no graph, the same "
                              "window relative to its peak and the same cadence",
        },
        "plain_value": plain_rate,
        "batched_over_plain": None if plain_rate is None else rate / plain_rate,
        "store_data_every_1": every1_out,
        "first_steps_of_each_run": first_out,
        "full_run": full,
        "device_ms_per_step": b["dev_ms"] / args.steps,
        "step_hbm_gbs": step_gbs,
        "step_hbm_frac": step_gbs / HBM_PEAK_GBS,
        "kernels": {k: {"avg_ms": v["avg_ms"], "algorithmic_bytes": v["bytes"], "ms_per_step": v["total_ms"] / args.steps}
                    for k, v in kern.items()},
        "roofline": {
            "bound": "hbm",
            "kernel": dom_name,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "copy_gbs": copy_gbs,
        },
        "cpu_baseline": cpu,
        "cpu_baseline_parallel": cpu_par,
        "velocity_gauge": vel,
        "sh_adi": sh_adi,
        "reference_benchmark_specs": ref_specs,
        "cyl_adi": cyl,
        "line_cn": line,
        "served_stream": served,
    }
    if tuned is not None:
        out["autotuned"] = tuned
    print(json.dumps(out), flush=True)
    if args.calibrate_copy:  # after the JSON line: a known-size kernel for the counter passes
        calibrate_copy()


if __name__ == "__main__":
    main()
