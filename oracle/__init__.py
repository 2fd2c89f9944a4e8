"""
CPU ORACLE — TEST INFRASTRUCTURE ONLY.

This package restates, on the CPU, the reference's mesh-TDSE evolve step
(JoshKarpel/ionization 0.2.0): per-step construction of the sparse ``dia``
operators, the flatten/wrap layout changes between them, SciPy ``dia.dot``
for explicit applies, a C Thomas solver with the recurrence of
``ionization/cy.pyx:9-50`` for implicit solves, the absorbing mask, and the
norm / inner-product observables. Each function cites the reference lines it
follows.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker or as the timed CPU
baseline. The product package ``ionization_amd`` never imports it and has no
CPU fallback.

Pinning (what checks each part against something other than itself):
  * ``thomas``            — dense ``numpy.linalg.solve`` and ``solve_banded``
                            on random systems (mirrors tests/test_tdma.py:12-26).
  * split E/O operators   — closed form ``expm(-i a sigma_x)`` per pair via
                            ``scipy.linalg.expm`` of the even/odd parts of the
                            dense interaction matrix; pairing parity for odd and
                            even mesh sizes.
  * CN / ADI steps        — dense ``(1 + i tau H)^-1 (1 - i tau H)`` compositions
                            built from dense matrices on tiny meshes; unitarity
                            of the field-free step.
  * SH / Line evolution   — field-free eigenstate stationarity at 1e-14
                            (tests/mesh/test_sims.py:30-82).
  * mask, pulses          — identities of tests/potentials/test_masks.py:82-108
                            and tests/potentials/test_pulses.py:169-219, plus
                            closed forms: the mask's interior values (midpoint
                            2^(-1/(2s)), complementary points); the Gaussian
                            pulse's envelope, fluence integral and Fourier
                            transform; the sinc pulse's flat band on
                            [w_min, w_min + 2 pi / pw]; the trapezoidal
                            window's half points and mirror symmetry.
  * potentials, states    — Coulomb = -1 hartree at one bohr, the softened form
                            in atomic units; hydrogen R_nl values at the origin,
                            R_21 in closed form, orthogonality, normalisation.
  * time grid, data times — known answers of the int() truncation and cadence.
  * A(t) quadrature       — simps_old against scipy.integrate.simpson on odd
                            sample counts and the even='avg' composition on
                            even ones (the rule of SciPy < 1.11, which the
                            reference's Python 3.7 environment runs); exact on
                            cubics.
  * eigenbasis            — ARPACK ``eigsh`` (the reference's routine) against
                            LAPACK ``eigh_tridiagonal``; hydrogen energies.
  * unit constants        — the derived atomic units (bohr, hartree, atomic
                            time and field) against their definitions; the
                            primary constants are parity unpinned (no reference
                            test fixes simulacra's numeric values; one CODATA
                            table is shared by oracle and device path).
"""
