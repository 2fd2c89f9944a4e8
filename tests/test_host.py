"""CPU tests of the host side: layouts, tables, the mesh API, the C-ABI library surface."""

import ctypes
import os
import re

import numpy as np
import pytest

from ionization_amd import _lib, engine, exceptions, mesh, potentials, states, workloads
from ionization_amd import units as u
from oracle import mesh_cpu, physics

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "tdse_hip.h")).read()
    return sorted(set(re.findall(r"\b(tdse_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = header_symbols()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(_lib.SIGNATURES), "ctypes signatures and header disagree"
    assert _lib.load().tdse_abi_version() == 1


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_LIB", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "nope.so"))
    with pytest.raises(exceptions.DeviceError):
        _lib.load()


def test_error_path_without_device():
    """Creating a context with no visible GPU reports an error string instead of crashing."""
    lib = _lib.load()
    handle = ctypes.c_void_p()
    rc = lib.tdse_ctx_create(0, 1 << 20, ctypes.byref(handle))
    if rc == 0:  # a GPU is visible: nothing to check here
        lib.tdse_ctx_destroy(handle)
        pytest.skip("a HIP device is visible")
    assert rc < 0 and lib.tdse_last_error(None)


@pytest.mark.parametrize("n", [2, 63, 64, 200, 201, 500, 1000, 2049, 4000, 4096])
def test_layout_roundtrip(n):
    P, m = engine.choose_layout(n, "sh")
    assert P % 64 == 0 and m % 2 == 0 and 2 <= m <= 16 and P * m >= n
    assert P <= (512 if m <= 8 else 256)
    j = engine.position_index(n, P, m)
    assert np.array_equal(np.sort(j), np.arange(P * m))
    a = np.random.default_rng(n).standard_normal((3, n)) + 1j
    b = engine.from_positions(engine.to_positions(a, P, m), n, P, m)
    assert np.array_equal(a, b)
    # chunk k holds rows k m .. k m + m - 1: position i P + k
    pos = engine.to_positions(np.arange(n, dtype=float)[None], P, m)[0]
    for k in (0, P - 1):
        for i in range(m):
            jj = k * m + i
            assert pos[i * P + k] == (jj if jj < n else 0)


def test_line_layout_supports_8192():
    assert engine.choose_layout(8192, "line") in [(1024, 8), (512, 16)]


@pytest.mark.parametrize("N", [1, 2, 7, 100, 101])
@pytest.mark.parametrize("every", [1, 3, 7, 100, -1])
def test_data_cadence(N, every):
    idx = engine.data_indices(N, every)
    assert len(idx) == engine.data_count(N, every)
    assert idx[0] == 0 and idx[-1] == N
    # the device slot formula
    for slot, n in enumerate(idx):
        if every >= 1:
            dev = N // every + 1 if (n == N and N % every) else n // every
        else:
            dev = 0 if n == 0 else 1
        assert dev == slot


def _small_sh_spec(**kw):
    return mesh.SphericalHarmonicSpecification(
        "t", r_bound=30 * u.bohr_radius, r_points=150, l_bound=12,
        electric_potential=potentials.SincPulse(pulse_width=80 * u.asec, fluence=0.5 * u.Jcm2),
        mask=potentials.RadialCosineMask(25 * u.bohr_radius, 30 * u.bohr_radius),
        time_initial=-40 * u.asec, time_final=40 * u.asec, time_step=1 * u.asec, **kw)


@pytest.mark.parametrize("n, l", [(1, 0), (2, 0), (2, 1)])
def test_initial_wavefunction_is_normalized_with_numeric_eigenstates(n, l):
    """tests/mesh/test_sims.py:13-28 of the reference."""
    sim = mesh.SphericalHarmonicSpecification(
        "test", initial_state=states.HydrogenBoundState(n, l), r_bound=50 * u.bohr_radius, r_points=250, l_bound=30,
        use_numeric_eigenstates=True, numeric_eigenstate_max_energy=20 * u.eV,
        numeric_eigenstate_max_angular_momentum=10).to_sim()
    assert np.allclose(sim.mesh.norm(), 1)
    assert sim.spec.initial_state.numeric and sim.spec.initial_state.l == l


def test_numeric_basis_matches_arpack_set():
    """Same eigenvalue set as the reference's eigsh loop (meshes.py:1290-1305) for one l."""
    import scipy.sparse.linalg as sla

    sim = _small_sh_spec().to_sim()
    st = [s for s in sim.spec.test_states if s.l == 1]
    d, e = sim.spec.operators.internal_hamiltonian_for_single_l(sim.mesh, 1)
    import scipy.sparse as sparse

    h = sparse.diags((e, d, e), offsets=(-1, 0, 1))
    w, _ = sla.eigsh(h, k=len(st) + 3, which="SA")
    w = np.sort(w)
    kept = w[w <= sim.spec.numeric_eigenstate_max_energy]
    assert np.allclose(sorted(s.energy for s in st), kept[: len(st)], rtol=1e-10)


def test_device_tables_reproduce_oracle_hamiltonian():
    """The compiled tables encode the oracle's sparse H0 and interaction couplings exactly (to rounding)."""
    sim = _small_sh_spec().to_sim()
    comp = sim.spec.evolution_method.compile(sim)
    tau0 = sim.spec.time_step / (2 * u.hbar)
    orc = mesh_cpu.SHOracle(r_bound=sim.spec.r_bound, r_points=150, l_bound=12, potential=physics.coulomb,
                            field=lambda t: 0.0 * t)
    n, L = 150, 12
    tab = engine.from_positions(comp.tables, n, comp.P, comp.m)
    T, U, AP, AN, RQ, MK = tab
    h0 = orc.h0.tocsr()
    for l in (0, 1, 5, 11):
        blk = h0[l * n:(l + 1) * n, l * n:(l + 1) * n]
        diag = (T + l * (l + 1) * U) / tau0
        if l == 0:
            diag = diag.copy()
            diag[0] += comp.corr0 / tau0
        assert np.allclose(diag, blk.diagonal().real, rtol=1e-13, atol=0)
        assert np.allclose(AN[:-1] / tau0, blk.diagonal(1).real, rtol=1e-13, atol=0)
        assert np.allclose(AP[1:] / tau0, blk.diagonal(-1).real, rtol=1e-13, atol=0)
    assert AN[-1] == 0 and AP[0] == 0
    # interaction coupling of pair (l, l+1) at r_j: RQ_j c_l / tau0 == oracle's l_interaction entry
    li = orc.l_interaction.tocsr()
    for j in (0, 7, 149):
        for l in (0, 3, 10):
            k = j * L + l
            cl = (l + 1) / np.sqrt((2 * l + 1) * (2 * l + 3))
            assert np.isclose(RQ[j] * cl / tau0, li[k + 1, k].real, rtol=1e-13)
    assert np.allclose(MK, sim.spec.mask(r=sim.mesh.r))


def test_field_table_samples_t_next_plus_half_step():
    sim = _small_sh_spec().to_sim()
    comp = sim.spec.evolution_method.compile(sim)
    t = sim.times
    expect = sim.spec.electric_potential.get_electric_field_amplitude(t[1:] + sim.spec.time_step / 2)
    assert np.allclose(comp.fs, expect * np.diff(t) / sim.spec.time_step, rtol=1e-15, atol=0)


def test_data_access_errors():
    sim = _small_sh_spec(datastores=[mesh.Norm()]).to_sim()
    with pytest.raises(exceptions.MissingDatastore):
        sim.data.bound_state_overlap
    with pytest.raises(exceptions.UnknownData):
        sim.data.definitely_not_data
    with pytest.raises(exceptions.DuplicateDatastores):
        _small_sh_spec(datastores=[mesh.Norm(), mesh.Norm()])
    with pytest.raises(exceptions.MissingDatastore):
        sim.data.norm_by_sph_harm


def test_norm_by_spherical_harmonic_compiles_the_device_reduction():
    plain = _small_sh_spec().to_sim()
    assert not plain.spec.evolution_method.compile(plain).norm_by_l
    sim = _small_sh_spec(datastores=[mesh.Norm(), mesh.NormBySphericalHarmonic()]).to_sim()
    assert sim.spec.evolution_method.compile(sim).norm_by_l
    harmonics = sim.spec.spherical_harmonics
    assert len(harmonics) == sim.spec.l_bound and len(set(harmonics)) == len(harmonics)
    assert set(sim.data.norm_by_sph_harm) == set(harmonics)
    assert all(np.isnan(v).all() and len(v) == sim.data_time_steps for v in sim.data.norm_by_sph_harm.values())


def test_scan_draws_are_seeded_and_in_range():
    a = workloads.sh_scan_params(256, 0)
    b = workloads.sh_scan_params(256, 0)
    assert a == b
    for p in a:
        assert 500 <= p["r_points"] <= 4000 and 50 <= p["l_bound"] <= 300
        assert 50 * u.asec <= p["pulse_width"] <= 800 * u.asec
        assert 0.01 * u.Jcm2 <= p["fluence"] <= 10 * u.Jcm2 and 0 <= p["phase"] <= np.pi


def test_fixed_step_count():
    spec = mesh.LineSpecification("l", initial_state=states.NumericOneDState(g=np.ones(64), energy=0.0),
                                  z_bound=10 * u.bohr_radius, z_points=64, time_step=0.0413 * u.atomic_time,
                                  time_final=4000 * 0.0413 * u.atomic_time)
    workloads.fixed_step_count(spec, 4000)
    assert int((spec.time_final - spec.time_initial) / spec.time_step) == 4000


def test_hamiltonian_derivation_is_rejected():
    with pytest.raises(NotImplementedError):
        mesh.SphericalHarmonicLengthGaugeOperators(kinetic_energy_derivation="hamiltonian")


# ------------------------------------------------------------ cylindrical layout


@pytest.mark.parametrize("nz, nr", [(2000, 1000), (120, 60), (64, 33), (301, 129), (150, 44)])
def test_cyl_layout_covers_mesh_and_round_trips(nz, nr):
    Pz, Mz, Cr, Mr, lanes = engine.choose_cyl_layout(nz, nr)
    assert Pz % 64 == 0 and Pz * Mz >= nz
    assert (Mr, lanes) in engine.CYL_RHO_SHAPES and Cr <= engine.CYL_RHO_SHAPES[(Mr, lanes)]
    assert Cr * Mr >= nr and (Cr * lanes) % 64 == 0 and (Pz * Mz) % lanes == 0
    rng = np.random.default_rng(nz + nr)
    g = rng.standard_normal((nz, nr)) + 1j * rng.standard_normal((nz, nr))
    d = engine.to_cyl_positions(g, Pz, Mz, Cr * Mr)
    assert d.shape == (Cr * Mr, Pz * Mz)
    assert np.count_nonzero(d) == nz * nr  # padding is zero
    assert np.array_equal(engine.from_cyl_positions(d, nz, nr, Pz, Mz), g)
    # the device order the kernels index (cyl_at): rho rows in interleaved pairs
    flat, npz = d.reshape(-1), Pz * Mz
    for j, r in [(0, 0), (nz - 1, nr - 1), (nz // 2, 1), (1, nr // 2), (nz - 2, nr - 2)]:
        p = (j % Mz) * Pz + j // Mz  # position of z index j
        assert flat[((r // 2) * npz + p) * 2 + r % 2] == g[j, r]


def test_cyl_layout_forced_shapes_and_rejects_unknown():
    for shape in engine.CYL_RHO_SHAPES:
        _, _, Cr, Mr, lanes = engine.choose_cyl_layout(120, 60, rho_shape=shape)
        assert (Mr, lanes) == shape
    with pytest.raises(exceptions.DeviceError):
        engine.choose_cyl_layout(120, 60, rho_shape=(6, 4))
    with pytest.raises(exceptions.DeviceError):
        engine.choose_cyl_layout(120, 5000)


def test_cyl_compile_tables_follow_reference_grid():
    """z / rho grids, the rho coupling c(j) = j / sqrt(j^2 - 1/4) and tau_n (meshes.py:533-544,
    mesh_operators.py:578-600) in the device tables."""
    zb, rb = 20 * u.bohr_radius, 10 * u.bohr_radius
    times = np.linspace(-10 * u.asec, 10 * u.asec, 21)
    nz, nr = 90, 31
    g0 = np.ones((nz, nr), np.complex128)
    c = engine.compile_cyl(z_bound=zb, rho_bound=rb, z_points=nz, rho_points=nr, field=lambda t: 0 * t,
                           times=times, g0=g0)
    z = np.linspace(-zb, zb, nz)
    rho = np.delete(np.linspace(0, rb, nr + 1), 0)
    drho = rho[1] - rho[0]
    jz = engine.position_index(nz, c.Pz, c.Mz)
    assert np.array_equal(c.zt[jz < nz], z[jz[jz < nz]])
    assert np.array_equal(c.rt[:nr], rho - drho / 2) and not c.rt[nr:].any()
    rp = -(u.hbar ** 2) / (2 * u.electron_mass_reduced * drho ** 2)
    j = np.arange(1, nr)
    assert np.allclose(c.roff[: nr - 1], rp * j / np.sqrt(j ** 2 - 0.25), rtol=1e-15, atol=0)
    assert not c.roff[nr - 1:].any()
    assert np.allclose(c.ts, np.diff(times) / (2 * u.hbar), rtol=1e-15, atol=0)


# ------------------------------------------------------------------ velocity gauge (host side)


def test_vector_potential_table_matches_oracle_quadrature():
    """A_n = -simps(E, times[: n + 2]) per step (mesh_operators.py:1182-1186), from the O(n)
    running sums, equals the oracle's per-call restatement of old SciPy simps."""
    fld = lambda t: physics.sinc_pulse_field(t, pulse_width=80 * u.asec, fluence=0.5 * u.Jcm2, phase=0.4)
    times = mesh_cpu.time_grid(-40 * u.asec, 40 * u.asec, 1 * u.asec)
    table = engine.vector_potential_table(fld, times)
    assert len(table) == len(times) - 1
    ref = np.array([-mesh_cpu.simps_old(fld(times[: n + 2]), times[: n + 2]) for n in range(len(times) - 1)])
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(table - ref)) <= 1e-13 * scale


def test_fields_datastore_vector_potential_uses_the_same_rule():
    sim = _small_sh_spec().to_sim()
    fld = sim.spec.electric_potential.get_electric_field_amplitude
    ref = np.array([-mesh_cpu.simps_old(fld(sim.times[: i + 1]), sim.times[: i + 1]) for i in sim.data_indices])
    got = sim.datastores_by_type[mesh.Fields].vector_potential_amplitude
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))
    t = sim.times[:30]
    assert np.isclose(sim.spec.electric_potential.get_vector_potential_amplitude_numeric(t),
                      -mesh_cpu.simps_old(fld(t), t), rtol=1e-13)


def test_velocity_tables_reproduce_oracle_couplings():
    """The velocity rows and the field table encode the oracle's h1 / h2 entries and A(t)."""
    sim = _small_sh_spec(operators=mesh.SphericalHarmonicVelocityGaugeOperators()).to_sim()
    comp = sim.spec.evolution_method.compile(sim)
    assert comp.gauge == "velocity" and comp.tables.shape[0] == engine.SH_NTAB + 1
    n, L = 150, 12
    tau0 = sim.spec.time_step / (2 * u.hbar)
    orc = mesh_cpu.SHVelocityOracle(r_bound=sim.spec.r_bound, r_points=n, l_bound=L, potential=physics.coulomb,
                                    field=lambda t: 0.0 * t, test_mass=sim.spec.test_mass)
    tab = engine.from_positions(comp.tables, n, comp.P, comp.m)
    RQ, AV = tab[engine.SH_RQ], tab[engine.SH_NTAB]
    h1, h2 = orc.h1.tocsr(), orc.h2.tocsr()
    for j in (0, 7, 149):
        for l in (0, 3, 10):
            k = j * L + l  # l-flattening
            cl = (l + 1) / np.sqrt((2 * l + 1) * (2 * l + 3))
            assert np.isclose(RQ[j] * cl * (l + 1) / tau0, h1[k, k + 1].imag, rtol=1e-13)
            if j < n - 1:
                kr = l * n + j  # r-flattening: (l, j) -> (l + 1, j + 1)
                assert np.isclose(AV[j] * cl / tau0, h2[kr, kr + n + 1].imag, rtol=1e-13)
    assert AV[n - 1] == 0.0
    fld = sim.spec.electric_potential.get_electric_field_amplitude
    a = np.array([-mesh_cpu.simps_old(fld(sim.times[: i + 2]), sim.times[: i + 2]) for i in range(comp.n_steps)])
    assert np.allclose(comp.fs, a * np.diff(sim.times) / sim.spec.time_step, rtol=1e-12, atol=1e-15 * np.abs(a).max())


def test_velocity_gauge_descriptor_and_limits():
    assert "gauge" in [f[0] for f in _lib.ShDesc._fields_]
    with pytest.raises(exceptions.DeviceError):
        engine.compile_sh(r_bound=10 * u.bohr_radius, r_points=2, l_bound=4, potential=physics.coulomb,
                          field=lambda t: 0.0 * t, mask=None, times=np.arange(3) * u.asec, time_step=u.asec,
                          g0=np.zeros((4, 2)), gauge="velocity")
    with pytest.raises(exceptions.DeviceError):
        engine.compile_sh(r_bound=10 * u.bohr_radius, r_points=20, l_bound=4, potential=physics.coulomb,
                          field=lambda t: 0.0 * t, mask=None, times=np.arange(3) * u.asec, time_step=u.asec,
                          g0=np.zeros((4, 20)), gauge="magnetic")


def test_harmonic_oscillator_and_qho_states():
    """HarmonicOscillator (potentials/static.py:113-233) and QHOState (states/one_d.py:130-300):
    k = m w^2 from the energy spacing; the eigenstates are orthonormal and are eigenvectors of
    the discretised K + V with energies hbar w (n + 1/2) to the grid's accuracy."""
    m = u.electron_mass
    qho = potentials.HarmonicOscillator.from_energy_spacing_and_mass(energy_spacing=0.1 * u.eV, mass=m)
    assert np.isclose(qho.omega(m) * u.hbar, 0.1 * u.eV, rtol=1e-13)
    z = np.linspace(-30 * u.nm, 30 * u.nm, 6001)
    dz = z[1] - z[0]
    vecs = [states.QHOState.from_potential(qho, n=n, mass=m)(z) for n in range(4)]
    gram = np.array([[np.sum(np.conj(a) * b) * dz for b in vecs] for a in vecs])
    assert np.allclose(gram, np.eye(4), atol=1e-9)
    pref = -(u.hbar ** 2) / (2 * m * dz ** 2)
    for n, v in enumerate(vecs):
        hv = (-2 * pref + qho(r=z)) * v
        hv[1:] += pref * v[:-1]
        hv[:-1] += pref * v[1:]
        e = (np.sum(np.conj(v) * hv) * dz).real
        assert np.isclose(e, states.QHOState.from_potential(qho, n=n, mass=m).energy, rtol=1e-4)
    sine = potentials.SineWave.from_photon_energy(photon_energy=0.1 * u.eV)
    assert np.isclose(sine.omega * u.hbar, 0.1 * u.eV, rtol=1e-13)


def test_simps_cumulative_small_prefixes():
    """One sample integrates to 0, two samples are the trapezoid (old SciPy simps on N = 1, 2)."""
    t = np.array([0.0, 0.5, 1.25])
    y = np.array([1.0, 3.0, 2.0])
    c = potentials.simps_cumulative(y, t)
    assert c[0] == 0.0
    assert np.isclose(c[1], 0.5 * 0.5 * (1.0 + 3.0), rtol=1e-15)
    assert np.isclose(c[2], mesh_cpu.simps_old(y, t), rtol=1e-14)


def test_unsupported_method_combinations_raise():
    """The device path refuses combinations it does not implement instead of running elsewhere."""
    kw = dict(r_bound=10 * u.bohr_radius, r_points=20, l_bound=4, potential=physics.coulomb, field=lambda t: 0.0 * t,
              mask=None, times=np.arange(3) * u.asec, time_step=u.asec, g0=np.zeros((4, 20)))
    with pytest.raises(exceptions.DeviceError):
        engine.compile_sh(method="adi", gauge="velocity", **kw)
    with pytest.raises(exceptions.DeviceError):
        engine.compile_sh(method="crank", **kw)
    lkw = dict(z_bound=10 * u.bohr_radius, z_points=64, potential=lambda z: 0.0 * z, field=lambda t: 0.0 * t,
               mask=None, times=np.arange(3) * u.asec, g0=np.zeros(64))
    with pytest.raises(exceptions.DeviceError):
        engine.compile_line(method="sio", time_steps=np.array([1j, 1j]) * u.asec, **lkw)
    with pytest.raises(exceptions.DeviceError):
        engine.compile_line(method="adi", gauge="velocity", **lkw)
    spec = mesh.LineSpecification("l", initial_state=states.NumericOneDState(g=np.ones(64), energy=0.0),
                                  z_bound=10 * u.bohr_radius, z_points=64, operators=mesh.LineVelocityGaugeOperators(),
                                  evolution_method=mesh.AlternatingDirectionImplicit())
    with pytest.raises(NotImplementedError):
        spec.evolution_method.compile(spec.to_sim())
    sh = _small_sh_spec(operators=mesh.SphericalHarmonicVelocityGaugeOperators(),
                        evolution_method=mesh.AlternatingDirectionImplicit())
    with pytest.raises(NotImplementedError):
        sh.evolution_method.compile(sh.to_sim())


def test_field_integrals_follow_the_reference_rules():
    """pulses.py:59-98: rule='simps' (default) or 'trapz' for the integral and the vector potential,
    cumtrapz with initial 0 for the cumulative forms, eps0 c int E^2 for the fluence."""
    from scipy import integrate

    p = potentials.SincPulse(pulse_width=80 * u.asec, fluence=2 * u.Jcm2, phase=0.3)
    t = np.linspace(-300, 300, 601) * u.asec
    e = p.get_electric_field_amplitude(t)
    assert p.get_vector_potential_amplitude_numeric(t, rule="trapz") == pytest.approx(
        -integrate.trapezoid(e, x=t), rel=1e-12)
    assert p.get_vector_potential_amplitude_numeric(t) == pytest.approx(-integrate.simpson(e, x=t), rel=1e-12)
    cum = p.get_vector_potential_amplitude_numeric_cumulative(t)
    assert np.allclose(cum, -integrate.cumulative_trapezoid(e, x=t, initial=0), rtol=1e-12,
                       atol=1e-12 * np.max(np.abs(cum)))
    assert cum[0] == 0.0
    assert p.get_fluence_numeric(t, rule="trapz") == pytest.approx(
        u.epsilon_0 * u.c * integrate.trapezoid(e ** 2, x=t), rel=1e-12)
    with pytest.raises(ValueError):
        p.get_vector_potential_amplitude_numeric(t, rule="romb")
