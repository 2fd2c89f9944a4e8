"""
Pinning of the CPU oracle against things other than itself: dense linear
algebra, closed forms, matrix exponentials, and the reference's own
invariant tests (tests/test_tdma.py, tests/mesh/test_sims.py,
tests/potentials/test_masks.py, tests/potentials/test_pulses.py).
"""

import hypothesis as hyp
import hypothesis.strategies as st
import numpy as np
import pytest
import scipy.linalg as sl
import scipy.sparse as sparse
import scipy.sparse.linalg as sla
from scipy import integrate

from ionization_amd import potentials
from ionization_amd import units as u
from oracle import mesh_cpu, physics
from tests.conftest import complex_random_sample, rel_l2


# ------------------------------------------------------------------ Thomas (cy.tdma)


@hyp.settings(deadline=None, max_examples=40)
@hyp.given(n=st.integers(min_value=2, max_value=400))
def test_tdma_agrees_with_solution_via_inverse(n):
    """tests/test_tdma.py:12-26 of the reference, against the oracle's C Thomas."""
    a = complex_random_sample(n - 1)
    b = complex_random_sample(n)
    c = complex_random_sample(n - 1)
    d = complex_random_sample(n)
    dia = sparse.diags([a, b, c], offsets=[-1, 0, 1])
    inv_x = np.linalg.inv(dia.toarray()).dot(d)
    tdma_x = mesh_cpu.tdma(dia.todia(), d)
    assert np.allclose(dia.dot(inv_x), d)
    assert np.allclose(tdma_x, inv_x)
    assert np.allclose(dia.dot(tdma_x), d)


def test_tdma_matches_solve_banded_on_dominant_system():
    rng = np.random.default_rng(3)
    n = 500
    a, c = complex_random_sample(n - 1, rng), complex_random_sample(n - 1, rng)
    b = complex_random_sample(n, rng) + 3.0
    d = complex_random_sample(n, rng)
    ab = np.zeros((3, n), dtype=np.complex128)
    ab[0, 1:] = c
    ab[1] = b
    ab[2, :-1] = a
    ref = sl.solve_banded((1, 1), ab, d)
    assert rel_l2(mesh_cpu.tdma(sparse.diags([a, b, c], [-1, 0, 1]).todia(), d), ref) < 1e-13


def test_tdma_rejects_length_one():
    with pytest.raises(ValueError):
        mesh_cpu.thomas(np.zeros(1), np.ones(1), np.zeros(1), np.ones(1))


# ------------------------------------------------------------------ SH operators


def _tiny_sh(r_points, l_bound, field=lambda t: 0.0 * t, mask=None):
    return mesh_cpu.SHOracle(r_bound=8 * u.bohr_radius, r_points=r_points, l_bound=l_bound,
                             potential=physics.coulomb, field=field, mask=mask)


def _dense_h0_closed_form(o):
    """H0 built independently from the closed forms of mesh_operators.py:841-928 (R-flattened)."""
    n, L = o.r_points, o.l_bound
    mu = u.electron_mass_reduced
    pr = -(u.hbar ** 2) / (2 * mu * o.delta_r ** 2)
    H = np.zeros((n * L, n * L))
    for l in range(L):
        for j in range(n):
            beta = (2 * j * j + 2 * j + 1) / (2 * j * j + 2 * j + 0.5)
            if l == 0 and j == 0:
                dr = o.delta_r / u.bohr_radius
                beta += dr * (1 + dr) / 8
            k = l * n + j
            H[k, k] = -2 * pr * beta + u.hbar ** 2 * l * (l + 1) / (2 * mu * o.r[j] ** 2) + physics.coulomb(o.r[j])
            if j + 1 < n:
                alpha = (j * j + 2 * j + 1) / (j * j + 2 * j + 0.75)
                H[k, k + 1] = H[k + 1, k] = pr * alpha
    return H


@pytest.mark.parametrize("r_points, l_bound", [(6, 4), (5, 3), (7, 4), (6, 5)])
def test_sh_internal_hamiltonian_matches_closed_form(r_points, l_bound):
    o = _tiny_sh(r_points, l_bound)
    H = _dense_h0_closed_form(o)
    assert np.allclose(o.h0.toarray().real, H, rtol=1e-12, atol=0)
    assert np.all(o.h0.toarray().imag == 0)


def _perm_R_to_L(o):
    """Permutation matrix Q with v_L = Q v_R (meshes.py:1052-1064)."""
    n, L = o.r_points, o.l_bound
    N = n * L
    Q = np.zeros((N, N))
    for l in range(L):
        for j in range(n):
            Q[j * L + l, l * n + j] = 1
    return Q


@pytest.mark.parametrize("r_points, l_bound", [(6, 4), (5, 3), (5, 4), (7, 5)])
def test_split_operators_are_exponentials_of_even_odd_couplings(r_points, l_bound):
    """E = expm(-i tau H_even), O = expm(-i tau H_odd) on the L-flattened pairs (k, k+1)."""
    o = _tiny_sh(r_points, l_bound)
    efield = 0.7 * u.atomic_electric_field
    tau = 3 * u.asec / (2 * u.hbar)
    E, O = o.split_operators(efield, tau)
    N = r_points * l_bound
    # dense interaction coupling: (k, k+1) -> E (-q) r_{k//L} c_l(k % L), zero across l-line ends
    hint = np.zeros((N, N))
    for k in range(N - 1):
        l, j = k % l_bound, k // l_bound
        if (k + 1) % l_bound != 0:
            cl = (l + 1) / np.sqrt((2 * l + 1) * (2 * l + 3))
            hint[k, k + 1] = hint[k + 1, k] = efield * (-u.electron_charge) * o.r[j] * cl
    even = np.zeros_like(hint)
    odd = np.zeros_like(hint)
    for k in range(N - 1):
        tgt = even if k % 2 == 0 else odd
        tgt[k, k + 1] = hint[k, k + 1]
        tgt[k + 1, k] = hint[k + 1, k]
    assert np.allclose(E.matrix.toarray(), sl.expm(-1j * tau * even), atol=1e-13)
    assert np.allclose(O.matrix.toarray(), sl.expm(-1j * tau * odd), atol=1e-13)
    # both are unitary
    for op in (E, O):
        M = op.matrix.toarray()
        assert np.allclose(M.conj().T @ M, np.eye(N), atol=1e-13)


@pytest.mark.parametrize("r_points, l_bound", [(6, 4), (5, 3)])
def test_sh_step_matches_dense_composition(r_points, l_bound):
    """One SIO step = M E O CN O E with dense pieces (evolution_methods.py:113-121, Q4)."""
    rb = 8 * u.bohr_radius
    maskf = lambda r: physics.radial_cosine_mask(r, 0.5 * rb, rb)
    fld = lambda t: 0.3 * u.atomic_electric_field * np.cos(t / (20 * u.asec))
    o = _tiny_sh(r_points, l_bound, field=fld, mask=maskf)
    rng = np.random.default_rng(1)
    g = (rng.standard_normal((l_bound, r_points)) + 1j * rng.standard_normal((l_bound, r_points)))
    dt = 2 * u.asec
    t_next = 7 * u.asec
    out = o.evolve(g, dt, t_next, dt)
    tau = dt / (2 * u.hbar)
    H0 = _dense_h0_closed_form(o)
    N = r_points * l_bound
    cn = np.linalg.solve(np.eye(N) + 1j * tau * H0, np.eye(N) - 1j * tau * H0)
    E, O = o.split_operators(fld(t_next + dt / 2), tau)
    Ed, Od = E.matrix.toarray(), O.matrix.toarray()
    Q = _perm_R_to_L(o)
    vL = g.flatten("F")
    vL = Od @ (Ed @ vL)
    vR = Q.T @ vL
    vR = cn @ vR
    vL = Q @ vR
    vL = Ed @ (Od @ vL)
    ref = np.reshape(vL, (l_bound, r_points), "F") * maskf(o.r_mesh)
    assert rel_l2(out, ref) < 1e-12


def test_sh_field_free_eigenstate_is_stationary():
    """tests/mesh/test_sims.py:30-82: norm and overlaps conserved to 1e-14 over 100 as."""
    o = mesh_cpu.SHOracle(r_bound=50 * u.bohr_radius, r_points=250, l_bound=30, potential=physics.coulomb,
                          field=lambda t: 0.0 * t)
    n = o.r_points
    basis = []
    for l in range(3):
        blk = o.h0.tocsr()[l * n:(l + 1) * n, l * n:(l + 1) * n].toarray().real
        w, v = np.linalg.eigh(blk)
        for i in range(2):
            basis.append((v[:, i] / np.sqrt(o.delta_r * np.sum(v[:, i] ** 2)), l))
    for vec, l in basis[:3]:
        g = np.zeros((o.l_bound, n), dtype=np.complex128)
        g[l] = vec
        times = mesh_cpu.time_grid(0, 100 * u.asec, 1 * u.asec)
        _, norms, ips, _ = o.run(g, times, 1 * u.asec, basis=basis)
        assert np.allclose(norms[0], norms[-1], atol=1e-14)
        assert np.allclose(np.abs(ips[0]) ** 2, np.abs(ips[-1]) ** 2, atol=1e-14)


def test_sh_evolution_without_mask_is_unitary():
    fld = lambda t: physics.sinc_pulse_field(t, pulse_width=50 * u.asec, fluence=1 * u.Jcm2)
    o = mesh_cpu.SHOracle(r_bound=20 * u.bohr_radius, r_points=60, l_bound=8, potential=physics.coulomb, field=fld)
    rng = np.random.default_rng(5)
    g = rng.standard_normal((8, 60)) + 1j * rng.standard_normal((8, 60))
    times = mesh_cpu.time_grid(-40 * u.asec, 40 * u.asec, 1 * u.asec)
    gf, norms, _, _ = o.run(g, times, 1 * u.asec)
    assert abs(norms[-1] - norms[0]) < 1e-12 * norms[0]


def test_sh_norm_by_l_pinned_by_total_norm_and_pure_l_states():
    """
    NormBySphericalHarmonic (meshes.py:1130-1133): the per-l norms sum to the
    norm, the entry at l = L-1 is the largest-l norm the run stores separately
    (sims.py:245-256), and a state in one l keeps its whole norm there while no
    field couples it to its neighbours.
    """
    fld = lambda t: physics.sinc_pulse_field(t, pulse_width=50 * u.asec, fluence=1 * u.Jcm2)
    o = mesh_cpu.SHOracle(r_bound=20 * u.bohr_radius, r_points=60, l_bound=8, potential=physics.coulomb, field=fld)
    rng = np.random.default_rng(9)
    g = rng.standard_normal((8, 60)) + 1j * rng.standard_normal((8, 60))
    times = mesh_cpu.time_grid(-30 * u.asec, 30 * u.asec, 1 * u.asec)
    _, norms, _, lnorm, nbl = o.run(g, times, 1 * u.asec, store_every=7, by_l=True)
    assert nbl.shape == (len(norms), 8)
    assert np.all(nbl >= 0)
    assert np.allclose(nbl.sum(axis=1), norms, rtol=1e-13, atol=0)
    assert np.allclose(nbl[:, -1], lnorm, rtol=1e-13, atol=0)
    # the first data point is the input itself, row by row
    ref0 = np.array([np.vdot(g[l], g[l]).real for l in range(8)]) * o.delta_r
    assert np.allclose(nbl[0], ref0, rtol=1e-13, atol=0)

    free = mesh_cpu.SHOracle(r_bound=20 * u.bohr_radius, r_points=60, l_bound=8, potential=physics.coulomb,
                             field=lambda t: 0.0 * t)
    g1 = np.zeros((8, 60), dtype=np.complex128)
    g1[3] = rng.standard_normal(60)
    _, n1, _, _, nbl1 = free.run(g1, times, 1 * u.asec, store_every=10, by_l=True)
    assert np.allclose(nbl1[:, 3], n1, rtol=1e-13, atol=0)
    assert np.all(np.delete(nbl1, 3, axis=1) == 0.0)


# ------------------------------------------------------------------ Line and Cyl


def test_line_step_matches_dense_cn():
    zb = 10 * u.bohr_radius
    fld = lambda t: 0.05 * u.atomic_electric_field * np.sin(t / u.atomic_time)
    maskf = lambda z: physics.radial_cosine_mask(z, 0.5 * zb, zb)
    o = mesh_cpu.LineOracle(z_bound=zb, z_points=40, potential=lambda z: physics.soft_coulomb(z, u.bohr_radius),
                            field=fld, mask=maskf, test_mass=u.electron_mass)
    rng = np.random.default_rng(2)
    g = rng.standard_normal(40) + 1j * rng.standard_normal(40)
    dt = 0.05 * u.atomic_time
    t_next = 3 * u.atomic_time
    out = o.evolve(g, dt, t_next)
    # independent dense H: kinetic stencil + V + V_int (mesh_operators.py:309-327)
    z = o.z
    dz = o.delta_z
    pz = -(u.hbar ** 2) / (2 * u.electron_mass * dz ** 2)
    H = np.diag(-2 * pz + physics.soft_coulomb(z, u.bohr_radius) - z * u.electron_charge * fld(t_next))
    H += np.diag(np.full(39, pz), 1) + np.diag(np.full(39, pz), -1)
    tau = dt / (2 * u.hbar)
    ref = np.linalg.solve(np.eye(40) + 1j * tau * H, (np.eye(40) - 1j * tau * H) @ g) * maskf(z)
    assert rel_l2(out, ref) < 1e-12


def test_cyl_adi_step_matches_dense_composition():
    """exp(H_rho) -> TDMA(H_z) -> exp(H_z) -> TDMA(H_rho), V split half/half (Q5)."""
    zb, rb = 6 * u.bohr_radius, 4 * u.bohr_radius
    nz, nr = 7, 5
    fld = lambda t: 0.2 * u.atomic_electric_field
    o = mesh_cpu.CylOracle(z_bound=zb, rho_bound=rb, z_points=nz, rho_points=nr, potential=physics.coulomb,
                           field=fld, test_mass=u.electron_mass)
    rng = np.random.default_rng(9)
    g = rng.standard_normal((nz, nr)) + 1j * rng.standard_normal((nz, nr))
    dt = 1 * u.asec
    out = o.evolve(g, dt, 1 * u.asec)
    m = u.electron_mass
    pz = -(u.hbar ** 2) / (2 * m * o.delta_z ** 2)
    pr = -(u.hbar ** 2) / (2 * m * o.delta_rho ** 2)
    V = physics.coulomb(o.r_mesh) - o.z_mesh * u.electron_charge * fld(0)
    Kz = np.diag(np.full(nz, -2 * pz)) + np.diag(np.full(nz - 1, pz), 1) + np.diag(np.full(nz - 1, pz), -1)
    cj = lambda j: j / np.sqrt(j * j - 0.25)
    Kr = np.diag(np.full(nr, -2 * pr))
    for j in range(nr - 1):
        Kr[j, j + 1] = Kr[j + 1, j] = pr * cj(j + 1)
    # operators on the (z, rho) C-ordered vector
    Hz = np.kron(Kz, np.eye(nr)) + np.diag(0.5 * V.flatten())
    Hr = np.kron(np.eye(nz), Kr) + np.diag(0.5 * V.flatten())
    tau = dt / (2 * u.hbar)
    I = np.eye(nz * nr)
    v = g.flatten()
    v = (I - 1j * tau * Hr) @ v
    v = np.linalg.solve(I + 1j * tau * Hz, v)
    v = (I - 1j * tau * Hz) @ v
    v = np.linalg.solve(I + 1j * tau * Hr, v)
    assert rel_l2(out, v.reshape(nz, nr)) < 1e-12


# ------------------------------------------------------------------ inputs


def test_radial_cosine_mask_identities():
    """tests/potentials/test_masks.py:82-108."""
    for outer in [1e-10, 3e-9, 5.0]:
        assert np.allclose(physics.radial_cosine_mask(0.0, 0.0, outer), 1)
        assert np.allclose(physics.radial_cosine_mask(outer, 0.0, outer), 0, atol=1e-14)
        m = potentials.RadialCosineMask(inner_radius=0.0, outer_radius=outer)
        assert np.allclose(m(r=0.0), 1)
        assert np.allclose(m(r=outer), 0, atol=1e-14)
    r = np.linspace(0, 12, 97)
    m = potentials.RadialCosineMask(inner_radius=4.0, outer_radius=10.0)
    assert np.array_equal(m(r=r), physics.radial_cosine_mask(r, 4.0, 10.0))


@pytest.mark.parametrize("pw", [50 * u.asec, 200 * u.asec, 800 * u.asec])
def test_pulse_identities(pw):
    """tests/potentials/test_pulses.py:169-219: envelope 1 at centre, E = A (phase 0), ~0 (phase pi/2)."""
    for cls in (potentials.SincPulse, potentials.GaussianPulse):
        p0 = cls(pulse_width=pw, phase=0)
        assert np.isclose(p0.get_electric_field_envelope(0.0), 1)
        assert np.isclose(p0.get_electric_field_amplitude(0.0), p0.amplitude)
        p1 = cls(pulse_width=pw, phase=np.pi / 2)
        assert abs(p1.get_electric_field_amplitude(0.0)) < 1e-9 * p1.amplitude


def test_pulse_restatements_agree():
    t = np.linspace(-2000, 2000, 801) * u.asec
    for pw, F, ph in [(200 * u.asec, 1 * u.Jcm2, 0.0), (73 * u.asec, 0.3 * u.Jcm2, 2.1)]:
        a = potentials.SincPulse(pulse_width=pw, fluence=F, phase=ph).get_electric_field_amplitude(t)
        b = physics.sinc_pulse_field(t, pulse_width=pw, fluence=F, phase=ph)
        assert np.allclose(a, b, rtol=1e-14, atol=1e-14 * np.max(np.abs(b)))
    g = potentials.GaussianPulse.from_amplitude(pulse_width=300 * u.asec, amplitude=0.02 * u.atomic_electric_field,
                                                phase=0.4)
    b = physics.gaussian_pulse_field(t, pulse_width=300 * u.asec, amplitude=0.02 * u.atomic_electric_field, phase=0.4)
    assert np.allclose(g.get_electric_field_amplitude(t), b, rtol=1e-14, atol=0)
    w = potentials.SmoothedTrapezoidalWindow(time_front=100 * u.asec, time_plateau=50 * u.asec)
    s = potentials.SineWave(omega=1e16, amplitude=3.0, window=w)
    assert np.allclose(s.get_electric_field_amplitude(t),
                       physics.sine_wave_field(t, omega=1e16, amplitude=3.0, time_front=100 * u.asec,
                                               time_plateau=50 * u.asec), rtol=1e-15, atol=0)


def test_sinc_pulse_fluence_matches_definition():
    """Numerical fluence eps0 c int E^2 dt of a long sinc window approaches the nominal fluence."""
    pw = 100 * u.asec
    p = potentials.SincPulse(pulse_width=pw, fluence=1 * u.Jcm2)
    t = np.linspace(-400 * pw, 400 * pw, 400001)
    fl = u.epsilon_0 * u.c * integrate.simpson(p.get_electric_field_amplitude(t) ** 2, x=t)
    assert abs(fl / p.fluence - 1) < 5e-3


@pytest.mark.parametrize("n, l", [(1, 0), (2, 0), (2, 1), (3, 2)])
def test_hydrogen_radial_normalised(n, l):
    f = lambda r: (physics.hydrogen_radial(n, l, r * u.bohr_radius) ** 2) * (r * u.bohr_radius) ** 2 * u.bohr_radius
    val, _ = integrate.quad(f, 0, 200)
    assert abs(val - 1) < 1e-8


# ------------------------------------------------------------------ closed forms of the host inputs
#
# Each check below is a value or an invariant that the mathematics fixes independently of the
# oracle's own formula, chosen so that a plausible slip (a power inverted, a width or carrier
# frequency off by a factor, a ramp edge swapped) fails it.


@pytest.mark.parametrize("s", [1.0, 2.0, 3.0, 8.0, 16.0])
def test_radial_cosine_mask_interior_closed_forms(s):
    """masks.py:76-89: m = |cos(pi x / 2)|^(1/s) on x = (r - r_in) / (r_out - r_in) in [0, 1)."""
    r_in, r_out = 3.0, 11.0
    at = lambda x: float(physics.radial_cosine_mask(r_in + x * (r_out - r_in), r_in, r_out, s))
    # midpoint: cos(pi / 4) = 2^(-1/2), so m = 2^(-1/(2 s))
    assert at(0.5) == pytest.approx(2.0 ** (-1.0 / (2 * s)), rel=1e-14)
    # quarter point: cos(pi / 8) = sqrt(2 + sqrt 2) / 2
    assert at(0.25) == pytest.approx((np.sqrt(2 + np.sqrt(2)) / 2) ** (1 / s), rel=1e-14)
    # complementary points: m(x)^(2s) + m(1 - x)^(2s) = cos^2 + sin^2 = 1
    for x in [0.1, 0.3, 0.45, 0.7]:
        assert at(x) ** (2 * s) + at(1 - x) ** (2 * s) == pytest.approx(1.0, abs=1e-13)
    # the mask is 1 up to and including r_in, 0 from r_out on, and decreasing between
    r = np.linspace(0, 14, 281)
    m = physics.radial_cosine_mask(r, r_in, r_out, s)
    assert np.all(m[r <= r_in] == 1.0) and np.all(m[r >= r_out] == 0.0)
    assert np.all(np.diff(m) <= 0)


def _fourier(t, e, omegas):
    """int e(t) exp(-i w t) dt by the trapezoid rule on a uniform grid (spectrally accurate for smooth,
    decaying integrands)."""
    dt = t[1] - t[0]
    return np.array([np.sum(e * np.exp(-1j * w * t)) * dt for w in omegas])


@pytest.mark.parametrize("pw, amp, phase", [(100 * u.asec, 0.01, 0.0), (300 * u.asec, 0.05, 0.4),
                                            (500 * u.asec, 0.1, 2.5)])
def test_gaussian_pulse_closed_forms(pw, amp, phase):
    """pulses.py:1016-1175, 1361-1372: E = A exp(-tau^2 / (2 pw^2)) cos(w_c tau + phi),
    w_c = w_min + pi / pw (the carrier of a sinc pulse of the same width), fluence
    F = eps0 c A^2 pw sqrt(pi) / 2 (1 + exp(-w_c^2 pw^2) cos 2 phi)."""
    A = amp * u.atomic_electric_field
    f = lambda t: physics.gaussian_pulse_field(t, pulse_width=pw, amplitude=A, phase=phase)
    w_min = u.twopi * 30 * u.THz
    wc = w_min + u.pi / pw
    # envelope: exp(-1/2) one width from the centre, exp(-2) at two, where the carrier is known
    for k in (1.0, 2.0):
        tau = k * pw
        assert f(tau) == pytest.approx(A * np.exp(-k * k / 2) * np.cos(wc * tau + phase), rel=1e-12, abs=1e-15 * A)
    # fluence of the amplitude-specified pulse, against the Gaussian integral
    t = np.linspace(-14 * pw, 14 * pw, 20001)
    fl = u.epsilon_0 * u.c * np.sum(f(t) ** 2) * (t[1] - t[0])
    closed = u.epsilon_0 * u.c * A ** 2 * pw * np.sqrt(u.pi) / 2 * (1 + np.exp(-(wc * pw) ** 2) * np.cos(2 * phase))
    assert fl == pytest.approx(closed, rel=1e-10)
    g = potentials.GaussianPulse.from_amplitude(pulse_width=pw, amplitude=A, phase=phase)
    assert u.epsilon_0 * u.c * A ** 2 * pw * np.sqrt(u.pi) / 2 == pytest.approx(g.fluence, rel=1e-12)
    # spectrum: (A / 2) pw sqrt(2 pi) [e^{i phi} G(w - w_c) + e^{-i phi} G(w + w_c)], G(v) = exp(-v^2 pw^2 / 2);
    # the peak is at w_c
    omegas = wc + np.array([-1.5, -0.5, 0.0, 0.7, 2.0]) / pw
    got = _fourier(t, f(t), omegas)
    G = lambda v: np.exp(-(v * pw) ** 2 / 2)
    want = (A / 2) * pw * np.sqrt(u.twopi) * (np.exp(1j * phase) * G(omegas - wc) + np.exp(-1j * phase) * G(omegas + wc))
    assert np.allclose(got, want, rtol=0, atol=1e-10 * abs(want[2]))


@pytest.mark.parametrize("pw, phase", [(50 * u.asec, 0.0), (200 * u.asec, 1.1), (800 * u.asec, 3.0)])
def test_sinc_pulse_spectrum_is_flat_on_its_band(pw, phase):
    """pulses.py:646-695, 929-940: the sinc pulse's spectrum is flat, (A / 2)(2 pi / dw) e^{i phi}, on
    [w_min, w_min + 2 pi / pw] and zero above it, with A^2 = F dw / (pi eps0 c).
    A wide Gaussian window W makes the numerical transform converge; it blurs the band edges by only 1 / W."""
    F = 1 * u.Jcm2
    f = lambda t: physics.sinc_pulse_field(t, pulse_width=pw, fluence=F, phase=phase)
    dw = u.twopi / pw
    w_min = u.twopi * 30 * u.THz
    A = np.sqrt(F * dw / (u.pi * u.epsilon_0 * u.c))
    W = 40 * pw
    t = np.arange(-12 * W, 12 * W, pw / 24)
    e = f(t) * np.exp(-0.5 * (t / W) ** 2)
    inside = w_min + np.array([0.2, 0.35, 0.5, 0.65, 0.8]) * dw
    outside = w_min + np.array([1.2, 1.6, 3.0]) * dw
    flat = (A / 2) * (u.twopi / dw) * np.exp(1j * phase)
    assert np.allclose(_fourier(t, e, inside), flat, rtol=0, atol=1e-9 * abs(flat))
    assert np.all(np.abs(_fourier(t, e, outside)) < 1e-9 * abs(flat))
    # the mirror band at negative frequencies carries the conjugate phase
    assert np.allclose(_fourier(t, e, -inside), np.conj(flat), rtol=0, atol=1e-9 * abs(flat))
    # envelope zeros one and two pulse widths from the centre, 2 / pi at half a width
    tau = np.array([pw, 2 * pw])
    wc = w_min + dw / 2
    assert np.all(np.abs(f(tau)) < 1e-15 * A)
    assert f(pw / 2) == pytest.approx(A * (2 / u.pi) * np.cos(wc * pw / 2 + phase), rel=1e-12, abs=1e-15 * A)


def test_smoothed_trapezoidal_window_closed_forms():
    """windows.py:203-225: sin^2 rise over time_front, plateau, cos^2 fall, then 0."""
    tf, tp, omega, amp = 120 * u.asec, 70 * u.asec, 1e16, 2.0
    w = lambda t: physics.sine_wave_field(t, omega=omega, amplitude=amp, time_front=tf, time_plateau=tp) / (
        np.sin(omega * np.asarray(t)) * amp)
    end = 2 * tf + tp
    assert w(tf / 2) == pytest.approx(0.5, rel=1e-13)
    assert w(tf + tp + tf / 2) == pytest.approx(0.5, rel=1e-13)
    assert w(tf / 4) == pytest.approx((2 - np.sqrt(2)) / 4, rel=1e-12)  # sin^2(pi / 8)
    assert w(tf) == 1.0 and w(tf + tp) == 1.0 and w(tf + 0.5 * tp) == 1.0
    assert w(end + 3 * u.asec) == 0.0
    # mirror symmetry about the plateau's centre on [0, end]
    t = np.linspace(0.01 * end, 0.99 * end, 173)
    t = t[np.abs(np.sin(omega * t)) > 1e-3]
    t = t[np.abs(np.sin(omega * (end - t))) > 1e-3]
    assert np.allclose(w(t), w(end - t), rtol=1e-12, atol=1e-14)
    # without a window the field is A sin(w t + phi): zero at t = (k pi - phi) / w, A at a quarter period
    phi = 0.3
    f = lambda t: physics.sine_wave_field(t, omega=omega, amplitude=amp, phase=phi)
    assert abs(f((u.pi - phi) / omega)) < 1e-14 * amp
    assert f((u.pi / 2 - phi) / omega) == pytest.approx(amp, rel=1e-14)


def test_derived_atomic_units_follow_their_definitions():
    """The unit table's derived atomic units against their definitions from the primary constants
    (the primary constants themselves are parity-unpinned: simulacra.units is absent)."""
    a0 = 4 * u.pi * u.epsilon_0 * u.hbar ** 2 / (u.electron_mass * u.proton_charge ** 2)
    assert u.bohr_radius == pytest.approx(a0, rel=1e-9)
    e_h = u.hbar ** 2 / (u.electron_mass * u.bohr_radius ** 2)
    assert u.hartree == pytest.approx(e_h, rel=1e-9)
    assert u.hartree == 2 * u.rydberg
    assert u.atomic_time == pytest.approx(u.hbar / e_h, rel=1e-9)
    assert u.atomic_electric_field == pytest.approx(e_h / (u.proton_charge * u.bohr_radius), rel=1e-9)
    assert u.coulomb_constant * u.proton_charge ** 2 / u.bohr_radius == pytest.approx(e_h, rel=1e-9)
    assert u.electron_mass_reduced == pytest.approx(1 / (1 / u.electron_mass + 1 / u.proton_mass), rel=1e-15)
    assert u.Jcm2 == 1 / (1e-2) ** 2 and u.asec == 1e-18 and u.THz == 1e12 and u.eV == u.proton_charge


def test_coulomb_and_soft_coulomb_closed_forms():
    """static.py:26-38, 75-91: -1 hartree one bohr from a proton; -1/sqrt(z^2 + s^2) hartree softened."""
    assert physics.coulomb(u.bohr_radius) == pytest.approx(-u.hartree, rel=1e-9)
    assert physics.coulomb(3 * u.bohr_radius, charge=2 * u.proton_charge) == pytest.approx(-2 * u.hartree / 3,
                                                                                            rel=1e-9)
    s = np.sqrt(2) * u.bohr_radius
    for z in [0.0, 1.0, 5.0]:
        want = -u.hartree / np.sqrt(z * z + 2)
        assert physics.soft_coulomb(z * u.bohr_radius, s) == pytest.approx(want, rel=1e-9)


def test_hydrogen_radial_values_and_orthogonality():
    """three_d.py:375-391 against the textbook R_10(0) = 2 a0^-3/2, R_20(0) = a0^-3/2 / sqrt 2,
    R_21(r) = r e^{-r/2} / (2 sqrt 6) (a0 units), and <R_10|R_20> = 0."""
    a = u.bohr_radius
    assert physics.hydrogen_radial(1, 0, 0.0) == pytest.approx(2 * a ** -1.5, rel=1e-13)
    assert physics.hydrogen_radial(2, 0, 0.0) == pytest.approx(a ** -1.5 / np.sqrt(2), rel=1e-13)
    for x in [0.5, 2.0, 7.0]:
        assert physics.hydrogen_radial(2, 1, x * a) == pytest.approx(
            a ** -1.5 * x * np.exp(-x / 2) / (2 * np.sqrt(6)), rel=1e-13)
    f = lambda x: physics.hydrogen_radial(1, 0, x * a) * physics.hydrogen_radial(2, 0, x * a) * (x * a) ** 2 * a
    val, _ = integrate.quad(f, 0, 200)
    assert abs(val) < 1e-10


def test_time_grid_truncates_like_int():
    """sims.py:198-205: int() truncation of (tf - t0) / dt, so 0.3 / 0.1 = 2.999... gives 3 points."""
    assert len(mesh_cpu.time_grid(0.0, 0.3, 0.1)) == 3
    assert np.array_equal(mesh_cpu.time_grid(0.0, 0.3, 0.1), np.array([0.0, 0.15, 0.3]))
    g = mesh_cpu.time_grid(-4000 * u.asec, 4000 * u.asec, 1 * u.asec)
    assert len(g) == int(8000e-18 / 1e-18) + 1 and g[0] == -4000 * u.asec and g[-1] == 4000 * u.asec


def test_data_indices_known_answers():
    """sims.py:83-90: t0, every store_data_every-th index, and the last index; only the ends for < 1."""
    assert list(mesh_cpu.data_indices(11, 4)) == [0, 4, 8, 10]
    assert list(mesh_cpu.data_indices(12, 4)) == [0, 4, 8, 11]
    assert list(mesh_cpu.data_indices(5, 1)) == [0, 1, 2, 3, 4]
    assert list(mesh_cpu.data_indices(9, -1)) == [0, 8]
    assert list(mesh_cpu.data_indices(9, 0)) == [0, 8]


def test_arpack_and_lapack_eigenbases_agree():
    """The reference's eigsh (meshes.py:1303-1305) vs eigh_tridiagonal on one H_l."""
    o = mesh_cpu.SHOracle(r_bound=40 * u.bohr_radius, r_points=300, l_bound=3, potential=physics.coulomb,
                          field=lambda t: 0.0 * t)
    n = o.r_points
    blk = o.h0.tocsr()[n:2 * n, n:2 * n]  # l = 1
    w1, v1 = sla.eigsh(blk, k=6, which="SA")
    d = blk.diagonal().real
    e = blk.diagonal(1).real
    w2, v2 = sl.eigh_tridiagonal(d, e, select="i", select_range=(0, 5))
    assert np.allclose(w1, w2, rtol=1e-9)
    for i in range(6):
        assert abs(abs(np.vdot(v1[:, i], v2[:, i])) - 1) < 1e-8
    # hydrogen 2p within discretisation error of -Ry/4
    assert abs(w2[0] / (-u.rydberg / 4) - 1) < 5e-3


# ------------------------------------------------------------------ velocity gauge


def _tiny_vel(r_points, l_bound, **kw):
    args = dict(r_bound=10 * u.bohr_radius, r_points=r_points, l_bound=l_bound, potential=physics.coulomb,
                field=lambda t: 0.02 * u.atomic_electric_field * np.cos(t / (50 * u.asec)))
    args.update(kw)
    return mesh_cpu.SHVelocityOracle(**args)


def _masked_part(H, keep):
    out = np.zeros_like(H)
    rows, cols = np.nonzero(H)
    for a, b in zip(rows, cols):
        if keep(a, b):
            out[a, b] = H[a, b]
    return out


@pytest.mark.parametrize("r_points, l_bound", [(7, 5), (6, 4), (5, 6), (4, 3), (3, 2)])
def test_velocity_split_factors_are_exponentials_of_their_couplings(r_points, l_bound):
    """
    Each velocity-gauge split factor is exp(-i tau part) of its part of the interaction,
    the parts built from the reference's definitions (mesh_operators.py:1143-1180) as dense
    matrices: h1 (l-flattening) split by the parity of the pair's first index; h2 =
    i hbar q / (2 m dr) kron(C, Alpha) (r-flattening) split by the parities of the lower l
    and the lower j of each coupled pair. The similarity transform, the sign of the d row,
    the angles and the pairing are all pinned by this (split_h1 / split_h2,
    mesh_operators.py:1201-1408, SimilarityOperator :150-204).
    """
    L, n = l_bound, r_points
    o = _tiny_vel(n, L)
    rng = np.random.default_rng(n * 10 + L)
    g = rng.standard_normal((L, n)) + 1j * rng.standard_normal((L, n))
    tau = 3 * u.asec / (2 * u.hbar)
    A = 0.7 * u.atomic_electric_field * u.atomic_time
    # dense h2 from the definition, independently of the oracle's sparse build
    cvec = (np.arange(L - 1) + 1) / np.sqrt((2 * np.arange(L - 1) + 1) * (2 * np.arange(L - 1) + 3))
    jj = np.arange(n - 1)
    avec = (jj ** 2 + 2 * jj + 1) / (jj ** 2 + 2 * jj + 0.75)
    C = np.diag(cvec, 1) + np.diag(cvec, -1)
    Al = np.diag(avec, 1) - np.diag(avec, -1)
    H2 = A * 1j * u.hbar * u.electron_charge / (u.electron_mass_reduced * 2 * o.delta_r) * np.kron(C, Al)
    assert np.allclose((A * o.h2).toarray(), H2, rtol=1e-14, atol=0)
    for op, (lp, rp) in zip(o.split_h2(A, tau), [(0, 0), (0, 1), (1, 0), (1, 1)]):
        part = _masked_part(H2, lambda a, b: min(a // n, b // n) % 2 == lp and min(a % n, b % n) % 2 == rp)
        ref = (sl.expm(-1j * tau * part) @ g.reshape(-1)).reshape(L, n)
        got = mesh_cpu.apply_sequentially(o.flatten, o.wrap, g, [op])
        assert rel_l2(got, ref) < 1e-13
    # dense h1: (k, k+1) of the l-flattening k = j L + l, i hbar (q / m) c_l (l + 1) / r_j above the diagonal
    H1 = np.zeros((L * n, L * n), dtype=np.complex128)
    for k in range(L * n - 1):
        l, j = k % L, k // L
        if (k + 1) % L:
            v = A * 1j * u.hbar * (u.electron_charge / u.electron_mass_reduced) * cvec[l] * (l + 1) / o.r[j]
            H1[k, k + 1], H1[k + 1, k] = v, -v
    for op, par in zip(o.split_h1(A, tau), (0, 1)):
        part = _masked_part(H1, lambda a, b: min(a, b) % 2 == par)
        ref = sl.expm(-1j * tau * part) @ g.reshape(-1, order="F")
        got = mesh_cpu.apply_sequentially(o.flatten, o.wrap, g, [op])
        assert rel_l2(got.reshape(-1, order="F"), ref) < 1e-13


def test_velocity_split_product_is_third_order():
    """The symmetric product of the six factors approximates exp(-2 i tau (h1 + h2) A) with a
    local error that falls 8x per halving of tau (a dropped factor or a wrong sign would not)."""
    L, n = 5, 7
    o = _tiny_vel(n, L)
    rng = np.random.default_rng(2)
    g = rng.standard_normal((L, n)) + 1j * rng.standard_normal((L, n))
    A = 0.7 * u.atomic_electric_field * u.atomic_time
    Q = np.zeros((L * n, L * n))  # r-order index l n + j from l-order index j L + l
    for l in range(L):
        for j in range(n):
            Q[l * n + j, j * L + l] = 1
    H = (A * o.h2).toarray() + Q @ (A * o.h1).toarray() @ Q.T
    errs = []
    for tau in np.array([3.0, 1.5, 0.75]) * u.asec / (2 * u.hbar):
        split = (*o.split_h1(A, tau), *o.split_h2(A, tau))
        got = mesh_cpu.apply_sequentially(o.flatten, o.wrap, g, (*split, *reversed(split)))
        ref = (sl.expm(-2j * tau * H) @ g.reshape(-1)).reshape(L, n)
        errs.append(np.abs(got - ref).max())
    assert 7.0 < errs[0] / errs[1] < 9.0 and 7.0 < errs[1] / errs[2] < 9.0


def test_velocity_gauge_without_field_is_the_length_gauge_step():
    """A = 0: every split factor is the identity, so both gauges take the same CN step."""
    kw = dict(r_bound=10 * u.bohr_radius, r_points=9, l_bound=4, potential=physics.coulomb, field=lambda t: 0.0 * t,
              mask=lambda r: physics.radial_cosine_mask(r, 6 * u.bohr_radius, 10 * u.bohr_radius))
    rng = np.random.default_rng(3)
    g = rng.standard_normal((4, 9)) + 1j * rng.standard_normal((4, 9))
    times = np.arange(5) * 2 * u.asec
    gv = mesh_cpu.SHVelocityOracle(**kw).run(g, times)[0]
    gl = mesh_cpu.SHOracle(**kw).run(g, times, 2 * u.asec)[0]
    assert rel_l2(gv, gl) < 1e-14


def test_velocity_gauge_evolution_without_mask_is_unitary():
    fld = lambda t: physics.sinc_pulse_field(t, pulse_width=50 * u.asec, fluence=1 * u.Jcm2)
    o = mesh_cpu.SHVelocityOracle(r_bound=20 * u.bohr_radius, r_points=60, l_bound=8, potential=physics.coulomb,
                                  field=fld)
    rng = np.random.default_rng(5)
    g = rng.standard_normal((8, 60)) + 1j * rng.standard_normal((8, 60))
    times = mesh_cpu.time_grid(-40 * u.asec, 40 * u.asec, 1 * u.asec)
    gf, norms, _, _ = o.run(g, times)
    assert abs(norms[-1] - norms[0]) < 1e-12 * norms[0]


def test_velocity_oracle_resumed_run_equals_the_uninterrupted_one():
    """run(start=k) continues from the state at times[k] with the vector potential still
    integrated from times[0]: stopping after k steps and resuming gives the same g, and a
    resume that integrated A from times[k] instead would not."""
    fld = lambda t: physics.sinc_pulse_field(t, pulse_width=50 * u.asec, fluence=5 * u.Jcm2)
    o = mesh_cpu.SHVelocityOracle(r_bound=20 * u.bohr_radius, r_points=40, l_bound=6, potential=physics.coulomb,
                                  field=fld, mask=lambda r: physics.radial_cosine_mask(r, 15 * u.bohr_radius,
                                                                                       20 * u.bohr_radius))
    rng = np.random.default_rng(11)
    g = rng.standard_normal((6, 40)) + 1j * rng.standard_normal((6, 40))
    times = mesh_cpu.time_grid(-10 * u.asec, 10 * u.asec, 1 * u.asec)
    k = 9
    g_full = o.run(g, times)[0]
    g_mid = o.run(g, times[:k + 1])[0]
    assert np.array_equal(o.run(g_mid, times, start=k)[0], g_full)
    g_wrong = o.run(g_mid, times[k:])[0]
    assert np.linalg.norm(g_wrong - g_full) > 1e-6 * np.linalg.norm(g_full)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 8, 11, 30])
def test_simps_old_matches_scipy_simpson_on_odd_counts_and_avg_rule_on_even(n):
    """
    Odd sample counts: SciPy's composite Simpson (the rule is the same in every SciPy
    release, so the local scipy.integrate.simpson pins it). Even counts: the mean of the
    two Simpson + trapezoid splits, each Simpson part from scipy.integrate.simpson on an
    odd count. Parity with the reference's SciPy (< 1.11, even='avg') rests on this rule.
    """
    rng = np.random.default_rng(n)
    x = np.sort(rng.uniform(-3, 4, n)) if n > 1 else np.zeros(1)
    y = np.sin(x) + 0.1 * x ** 3
    got = mesh_cpu.simps_old(y, x)
    if n == 1:
        assert got == 0.0
        return
    if n % 2:
        assert np.isclose(got, integrate.simpson(y, x=x), rtol=1e-13, atol=1e-15)
        return
    simp = lambda a, b: integrate.simpson(y[a:b], x=x[a:b]) if b - a >= 3 else 0.0
    trap = lambda a: 0.5 * (x[a + 1] - x[a]) * (y[a + 1] + y[a])
    ref = 0.5 * ((simp(0, n - 1) + trap(n - 2)) + (trap(0) + simp(1, n)))
    assert np.isclose(got, ref, rtol=1e-13, atol=1e-15)


def test_simps_old_exact_for_cubics_on_uniform_odd_grids():
    x = np.linspace(-1.0, 2.0, 9)
    y = 2 * x ** 3 - x ** 2 + 3
    exact = (2 * 2 ** 4 / 4 - 2 ** 3 / 3 + 6) - (2 / 4 + 1 / 3 - 3)
    assert np.isclose(mesh_cpu.simps_old(y, x), exact, rtol=1e-14)


# ------------------------------------------------------------------ LineMesh split operator


def _line_sio(n, gauge, e0=0.3):
    return mesh_cpu.LineSIOOracle(
        z_bound=20 * u.bohr_radius, z_points=n, potential=lambda z: physics.soft_coulomb(z, np.sqrt(2) * u.bohr_radius),
        field=lambda t: e0 * u.atomic_electric_field * np.sin(t / (30 * u.asec)), gauge=gauge,
        test_mass=u.electron_mass)


@pytest.mark.parametrize("n", [9, 10])
def test_line_velocity_split_factors_are_exponentials(n):
    """LineVelocityGaugeOperators (mesh_operators.py:352-427): the even / odd factors are
    exp(-i tau h_part) of h = A i hbar (q / m) / (2 dz) (super - sub diagonal) restricted to the
    pairs (k, k + 1) of even / odd k. A strong field makes the angles O(1)."""
    o = _line_sio(n, "velocity", e0=40.0)
    ts = np.arange(8) * 5 * u.asec
    tau = 5 * u.asec / (2 * u.hbar)
    A = -mesh_cpu.simps_old(o.field(ts), ts)
    pref = A * 1j * u.hbar * (u.electron_charge / u.electron_mass) / (2 * o.delta_z)
    H = np.diag(np.full(n - 1, pref), 1) - np.diag(np.full(n - 1, pref), -1)
    ops = o.split_operators(tau, ts[-1], ts)
    assert 0.1 < abs(tau * pref) < 3.0
    rng = np.random.default_rng(n)
    g = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for op, par in zip(ops, (0, 1)):
        part = np.zeros_like(H)
        for k in range(par, n - 1, 2):
            part[k, k + 1], part[k + 1, k] = H[k, k + 1], H[k + 1, k]
        assert rel_l2(op.matrix.dot(g), sl.expm(-1j * tau * part) @ g) < 1e-13


def test_line_length_split_step_matches_dense_composition():
    """Line SIO, length gauge: D . CN(K + V) . D . mask with D = exp(-i tau V_int(t_{n+1}))
    (mesh_operators.py:329-342, evolution_methods.py:113-121)."""
    n = 12
    maskf = lambda z: physics.radial_cosine_mask(z, 10 * u.bohr_radius, 20 * u.bohr_radius)
    o = mesh_cpu.LineSIOOracle(z_bound=20 * u.bohr_radius, z_points=n,
                               potential=lambda z: physics.soft_coulomb(z, np.sqrt(2) * u.bohr_radius),
                               field=lambda t: 2.0 * u.atomic_electric_field * np.cos(t / (40 * u.asec)),
                               mask=maskf, test_mass=u.electron_mass, gauge="length")
    rng = np.random.default_rng(4)
    g = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    dt, t1 = 3 * u.asec, 9 * u.asec
    out = o.evolve_sio(g, dt, t1, np.array([t1 - dt, t1]))
    tau = dt / (2 * u.hbar)
    H0 = o.h_internal.toarray()
    D = np.diag(np.exp(-1j * tau * (-o.z * u.electron_charge * o.field(t1))))
    cn = np.linalg.solve(np.eye(n) + 1j * tau * H0, np.eye(n) - 1j * tau * H0)
    ref = (D @ cn @ D @ g) * maskf(o.z)
    assert rel_l2(out, ref) < 1e-12


def test_line_split_operator_without_field_is_the_adi_step():
    """E = 0: the split factors are the identity and both gauges take the ADI step in K + V."""
    kw = dict(z_bound=20 * u.bohr_radius, z_points=11, field=lambda t: 0.0 * t, test_mass=u.electron_mass,
              potential=lambda z: physics.soft_coulomb(z, np.sqrt(2) * u.bohr_radius))
    rng = np.random.default_rng(6)
    g = rng.standard_normal(11) + 1j * rng.standard_normal(11)
    times = np.arange(4) * 2 * u.asec
    adi = mesh_cpu.LineOracle(**kw).run(g, times)[0]
    for gauge in ("length", "velocity"):
        assert rel_l2(mesh_cpu.LineSIOOracle(gauge=gauge, **kw).run(g, times)[0], adi) < 1e-14


@pytest.mark.parametrize("gauge", ["length", "velocity"])
def test_line_split_operator_is_unitary_without_mask(gauge):
    o = _line_sio(40, gauge)
    rng = np.random.default_rng(7)
    g = rng.standard_normal(40) + 1j * rng.standard_normal(40)
    gf, norms, _ = o.run(g, np.arange(30) * 2 * u.asec)
    assert abs(norms[-1] - norms[0]) < 1e-12 * norms[0]


# ------------------------------------------------------------------ SphericalHarmonicMesh ADI


@pytest.mark.parametrize("r_points, l_bound", [(6, 4), (5, 3)])
def test_sh_adi_step_matches_dense_composition(r_points, l_bound):
    """AlternatingDirectionImplicit on the SH mesh: (1 + i tau H0)^-1 (1 - i tau E H_int)
    (1 + i tau E H_int)^-1 (1 - i tau H0), then the mask, with H0 along r and H_int along l
    (evolution_methods.py:46-77, mesh_operators.py:1020-1035), built densely with explicit
    permutation matrices between the flattenings."""
    rb = 8 * u.bohr_radius
    maskf = lambda r: physics.radial_cosine_mask(r, 0.5 * rb, rb)
    fld = lambda t: 0.3 * u.atomic_electric_field * np.cos(t / (20 * u.asec))
    o = _tiny_sh(r_points, l_bound, field=fld, mask=maskf)
    rng = np.random.default_rng(1)
    g = rng.standard_normal((l_bound, r_points)) + 1j * rng.standard_normal((l_bound, r_points))
    dt, t1 = 2 * u.asec, 7 * u.asec
    out = o.evolve_adi(g, dt, t1, dt)
    tau = dt / (2 * u.hbar)
    N = r_points * l_bound
    H0 = _dense_h0_closed_form(o)
    hint = np.zeros((N, N))  # l-flattening: (k, k+1) -> E (-q) r_{k//L} c_l(k % L)
    for k in range(N - 1):
        l, j = k % l_bound, k // l_bound
        if (k + 1) % l_bound != 0:
            cl = (l + 1) / np.sqrt((2 * l + 1) * (2 * l + 3))
            hint[k, k + 1] = hint[k + 1, k] = fld(t1 + dt / 2) * (-u.electron_charge) * o.r[j] * cl
    Q = _perm_R_to_L(o)
    eye = np.eye(N)
    vR = (eye - 1j * tau * H0) @ g.reshape(-1)
    vL = Q @ vR
    vL = (eye - 1j * tau * hint) @ np.linalg.solve(eye + 1j * tau * hint, vL)
    vR = np.linalg.solve(eye + 1j * tau * H0, Q.T @ vL)
    ref = vR.reshape(l_bound, r_points) * maskf(o.r_mesh)
    assert rel_l2(out, ref) < 1e-12


def test_sh_adi_field_free_stationary_and_its_invariant():
    o = mesh_cpu.SHOracle(r_bound=30 * u.bohr_radius, r_points=120, l_bound=6, potential=physics.coulomb,
                          field=lambda t: 0.0 * t)
    n = o.r_points
    blk = o.h0.tocsr()[:n, :n].toarray().real
    _, v = np.linalg.eigh(blk)
    phi = v[:, 0] / np.sqrt(o.delta_r * np.sum(v[:, 0] ** 2))
    g = np.zeros((6, n), dtype=np.complex128)
    g[0] = phi
    times = mesh_cpu.time_grid(0, 50 * u.asec, 1 * u.asec)
    _, norms, ips, _ = o.run(g, times, 1 * u.asec, basis=[(phi, 0)], adi=True)
    assert np.allclose(norms[0], norms[-1], atol=1e-14)
    assert np.allclose(np.abs(ips[0]) ** 2, np.abs(ips[-1]) ** 2, atol=1e-14)
    # with a field the step is S^-1 (C_int C_0) S, S = 1 + i tau H0 (C_0 = (1 - i tau H0) S^-1): the
    # norm of S g is conserved exactly (unit-modulus Cayley factors), not the norm of g
    fld = lambda t: physics.sinc_pulse_field(t, pulse_width=50 * u.asec, fluence=1 * u.Jcm2)
    o2 = mesh_cpu.SHOracle(r_bound=20 * u.bohr_radius, r_points=60, l_bound=8, potential=physics.coulomb, field=fld)
    rng = np.random.default_rng(5)
    g2 = rng.standard_normal((8, 60)) + 1j * rng.standard_normal((8, 60))
    times = mesh_cpu.time_grid(-40 * u.asec, 40 * u.asec, 1 * u.asec)
    gN, norms2, _, _ = o2.run(g2, times, 1 * u.asec, adi=True)
    S = mesh_cpu.add_to_diagonal((1j * (1 * u.asec) / (2 * u.hbar)) * o2.h0, value=1)
    s0, sN = np.linalg.norm(S.dot(g2.reshape(-1))), np.linalg.norm(S.dot(gN.reshape(-1)))
    assert abs(sN - s0) < 1e-12 * s0
    assert abs(norms2[-1] - norms2[0]) > 1e-6 * norms2[0]  # while the plain norm is not conserved


def test_line_velocity_split_step_matches_dense_composition():
    """Line SIO, velocity gauge: O E . CN(K + V) . E O order reversed around the solve and the mask,
    with E / O = expm(-i tau A h_even / h_odd) built densely from the definition (mesh_operators.py:
    352-427, evolution_methods.py:113-121) and A = -simps(E, t_0 .. t_{n+1})."""
    n = 11
    maskf = lambda z: physics.radial_cosine_mask(z, 10 * u.bohr_radius, 20 * u.bohr_radius)
    fld = lambda t: 30.0 * u.atomic_electric_field * np.cos(t / (40 * u.asec))
    o = mesh_cpu.LineSIOOracle(z_bound=20 * u.bohr_radius, z_points=n,
                               potential=lambda z: physics.soft_coulomb(z, np.sqrt(2) * u.bohr_radius),
                               field=fld, mask=maskf, test_mass=u.electron_mass, gauge="velocity")
    rng = np.random.default_rng(8)
    g = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ts = np.arange(5) * 3 * u.asec
    dt = ts[-1] - ts[-2]
    out = o.evolve_sio(g, dt, ts[-1], ts)
    tau = dt / (2 * u.hbar)
    A = -mesh_cpu.simps_old(fld(ts), ts)
    pref = A * 1j * u.hbar * (u.electron_charge / u.electron_mass) / (2 * o.delta_z)
    H = np.diag(np.full(n - 1, pref), 1) - np.diag(np.full(n - 1, pref), -1)
    parts = []
    for par in (0, 1):
        Pm = np.zeros_like(H)
        for k in range(par, n - 1, 2):
            Pm[k, k + 1], Pm[k + 1, k] = H[k, k + 1], H[k + 1, k]
        parts.append(sl.expm(-1j * tau * Pm))
    Ev, Od = parts
    H0 = o.h_internal.toarray()
    cn = np.linalg.solve(np.eye(n) + 1j * tau * H0, np.eye(n) - 1j * tau * H0)
    assert abs(tau * pref) > 0.05  # the factors are far from the identity
    ref = (Ev @ Od @ cn @ Od @ Ev @ g) * maskf(o.z)
    assert rel_l2(out, ref) < 1e-12


def test_velocity_sh_step_matches_dense_composition():
    """One velocity-gauge SIO step = M . h1E h1O . EE EO OE OO (reversed) . CN . OO OE EO EE . h1O h1E
    in application order (evolution_methods.py:113-121), every factor expm(-i tau part) of the
    dense h1 / h2 parts and A = -simps(E, t_0 .. t_{n+1})."""
    L, n = 4, 5
    rb = 10 * u.bohr_radius
    maskf = lambda r: physics.radial_cosine_mask(r, 0.5 * rb, rb)
    fld = lambda t: 20.0 * u.atomic_electric_field * np.cos(t / (30 * u.asec))
    o = _tiny_vel(n, L, field=fld, mask=maskf)
    rng = np.random.default_rng(9)
    g = rng.standard_normal((L, n)) + 1j * rng.standard_normal((L, n))
    ts = np.arange(4) * 2 * u.asec
    dt = ts[-1] - ts[-2]
    out = o.evolve_velocity(g, dt, ts)
    tau = dt / (2 * u.hbar)
    A = -mesh_cpu.simps_old(fld(ts), ts)
    Q = np.zeros((L * n, L * n))  # r-order index l n + j from l-order index j L + l
    for l in range(L):
        for j in range(n):
            Q[l * n + j, j * L + l] = 1
    H1 = Q @ (A * o.h1).toarray() @ Q.T  # both in r-order
    H2 = (A * o.h2).toarray()
    ex = lambda P_: sl.expm(-1j * tau * P_)
    h1 = [ex(_masked_part((A * o.h1).toarray(), lambda a, b, p=p: min(a, b) % 2 == p)) for p in (0, 1)]
    h1 = [Q @ M @ Q.T for M in h1]
    h2 = [ex(_masked_part(H2, lambda a, b, lp=lp, rp=rp: min(a // n, b // n) % 2 == lp and min(a % n, b % n) % 2 == rp))
          for lp, rp in [(0, 0), (0, 1), (1, 0), (1, 1)]]
    split = [h1[0], h1[1], *h2]  # application order
    H0 = _dense_h0_closed_form(o)
    N = L * n
    cn = np.linalg.solve(np.eye(N) + 1j * tau * H0, np.eye(N) - 1j * tau * H0)
    v = g.reshape(-1)
    for M in split:
        v = M @ v
    v = cn @ v
    for M in reversed(split):
        v = M @ v
    ref = v.reshape(L, n) * maskf(o.r_mesh)
    assert np.abs(tau * A * u.hbar * u.electron_charge / (u.electron_mass * o.delta_r)) > 0.05
    assert rel_l2(out, ref) < 1e-12
    assert np.allclose(H1 + H2, (H1 + H2).conj().T)  # the interaction is Hermitian
