/*
 * tdse_hip — C ABI of the MI355X (gfx950) mesh-TDSE evolution library.
 *
 * Plain pointers and sizes only; every function returns 0 on success and a
 * negative code on failure, with the message available from
 * tdse_last_error(ctx) (ctx may be NULL for errors raised before a context
 * exists). No C++ exception crosses this boundary. complex128 arrays are
 * interleaved (re, im) doubles.
 *
 * Reference interfaces each entry point replaces are cited per declaration
 * (paths relative to the JoshKarpel/ionization source tree).
 */
#ifndef TDSE_HIP_H
#define TDSE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tdse_ctx tdse_ctx;

enum {
    TDSE_OK = 0,
    TDSE_ERR_ARG = -1,     /* invalid argument or layout */
    TDSE_ERR_HIP = -2,     /* HIP runtime error */
    TDSE_ERR_NOMEM = -3,   /* device pool exhausted */
    TDSE_ERR_STATE = -4    /* call not valid in the current state */
};

/* Library ABI version (increments when a signature changes). */
int tdse_abi_version(void);

/* Last error message of a context (or of the library when ctx is NULL). */
const char *tdse_last_error(const tdse_ctx *ctx);

/*
 * Create / destroy a context on one device: one HIP stream, one device memory
 * pool of pool_bytes reserved up front (all simulation buffers are carved
 * from it; nothing is allocated per step). Replaces the per-process state of
 * a reference run (one MeshSimulation per process, ionization/mesh/sims.py:36-118).
 */
int tdse_ctx_create(int device, size_t pool_bytes, tdse_ctx **out);
int tdse_ctx_destroy(tdse_ctx *ctx);
int tdse_sync(tdse_ctx *ctx);
int tdse_pool_usage(tdse_ctx *ctx, size_t *used_bytes, size_t *total_bytes);

/*
 * Streaming copy of `bytes` (a multiple of 16) between two blocks carved from
 * the context pool, `launches` times; *avg_ms = average kernel time from HIP
 * events on the context stream. No counterpart in the reference: it gives the
 * attainable copy bandwidth and the known-size kernel that scales HBM byte
 * counters (FETCH_SIZE / WRITE_SIZE) to bytes.
 */
int tdse_copy_probe(tdse_ctx *ctx, size_t bytes, int launches, double *avg_ms);

/*
 * Tridiagonal solve x = T^-1 d for one complex128 system of length n >= 1,
 * sub[i] multiplying x[i-1] (sub[0] ignored) and sup[i] multiplying x[i+1]
 * (sup[n-1] ignored). Host arrays in, host array out; inputs unchanged.
 * Replaces cy.tdma(matrix, d) (ionization/cy.pyx:9-50) as called by
 * TDMAOperator._apply (ionization/mesh/mesh_operators.py:109-111); the
 * caller passes the dia matrix rows data[0], data[1], data[2] re-aligned as
 * cy.pyx:19-22 does.
 */
int tdse_tdma(tdse_ctx *ctx, const double *sub, const double *diag, const double *sup, const double *d,
              double *x, int64_t n);

/*
 * Many independent systems packed back to back: system i occupies
 * [offsets[i], offsets[i] + lengths[i]) of every array. Same semantics as
 * tdse_tdma per system. Replaces a loop of cy.tdma calls
 * (ionization/cy.pyx:9-50), e.g. the decoupled lines of one block-diagonal
 * TDMAOperator (ionization/mesh/mesh_operators.py:109-111).
 */
int tdse_tdma_batched(tdse_ctx *ctx, const double *sub, const double *diag, const double *sup, const double *d,
                      double *x, const int64_t *offsets, const int64_t *lengths, int64_t n_systems);

/*
 * Allocation-free form of tdse_tdma_batched. No counterpart in the reference:
 * it is for callers that cannot allocate device memory (tight serving loops,
 * code that forbids hipMalloc). sub, diag, sup, d and x are DEVICE arrays of
 * `total` complex128 each, owned by the caller (x is written); offsets and
 * lengths are host arrays of n_systems, every system inside [0, total). The
 * Thomas scratch (primed super-diagonal) and the staged offsets / lengths live
 * in a static device array of the library, one per GPU, so the call allocates
 * nothing: total <= 2^21 unknowns and n_systems <= 2^16, else TDSE_ERR_ARG.
 * Synchronous: returns after x is written. Calls that overlap on one GPU
 * (several host threads or contexts) race on that array.
 */
int tdse_tdma_batched_noalloc(tdse_ctx *ctx, const double *sub, const double *diag, const double *sup,
                              const double *d, double *x, const int64_t *offsets, const int64_t *lengths,
                              int64_t n_systems, int64_t total);

/*
 * SphericalHarmonicMesh simulation, length gauge, SplitInteractionOperator
 * (ionization/mesh/meshes.py:985-1136, mesh_operators.py:815-1080,
 * evolution_methods.py:80-123). All arrays are in device position order:
 * position p = i * chunks + k holds radial index j = k * rows + i,
 * npad = chunks * rows.
 */
typedef struct {
    int32_t r_points;      /* n */
    int32_t l_bound;       /* L */
    int32_t chunks;        /* P: threads per l-line, multiple of 64, <= 512 */
    int32_t rows;          /* m: rows per thread, even, 2 <= m <= 16 */
    int32_t n_steps;       /* N = len(times) - 1 */
    int32_t store_every;   /* data cadence (sims.py:84-95); -1 = first and last only */
    int32_t n_basis;       /* test states for inner products */
    int32_t lb;            /* rows l < lb are captured at data steps (>= 1 + max basis l) */
    double delta_r;        /* inner-product weight (meshes.py:1013); r_j = (j + 1/2) delta_r */
    double corr0;          /* tau0 (-2 P_r) dr (1 + dr) / 8 (mesh_operators.py:901-903) */
    double centrifugal;    /* tau0 hbar^2 / (2 mu): the U table is centrifugal / r_j^2 */
    int32_t norm_by_l;     /* 1: also reduce the norm in every l at data times (NormBySphericalHarmonic) */
    int32_t gauge;         /* 0: length gauge; 1: velocity gauge (SphericalHarmonicVelocityGaugeOperators,
                              mesh_operators.py:1130-1408): r_points >= 3, l_bound >= 2 */
    int32_t method;        /* 0: SplitInteractionOperator; 1: AlternatingDirectionImplicit (length gauge;
                              evolution_methods.py:46-77 with the total Hamiltonian of mesh_operators.py:1020-1035) */
    int32_t reserved;
} tdse_sh_desc;

/*
 * Add one SH simulation; it starts at step 0 on the next tdse_run. tables:
 * 6 * npad doubles (T, U, A_{j-1}, A_j, tau0 (-q) r, mask), fs / ts: N
 * doubles (field at t_{n+1} + dt/2 times tau_n / tau0, and tau_n / tau0),
 * g0: 2 * L * npad doubles, basis: n_basis * npad doubles, basis_l: n_basis.
 * Velocity gauge (gauge = 1): row 4 holds tau0 hbar (q / m) / r_j (the h1
 * factor; the library multiplies by c_l (l + 1)), a seventh row holds
 * tau0 hbar q / (2 m dr) alpha_j for j < n - 1 and 0 beyond (the h2 factor;
 * times c_l), and fs holds the vector potential A(t_{n+1}) times tau_n / tau0
 * (mesh_operators.py:1143-1199).
 * Replaces MeshSpecification.to_sim() + the first call of
 * SplitInteractionOperator.evolve (sims.py:46-118, evolution_methods.py:89-123).
 */
int tdse_sh_add(tdse_ctx *ctx, const tdse_sh_desc *desc, const double *tables, const double *fs, const double *ts,
                const double *g0, const double *basis, const int32_t *basis_l, int32_t *slot);

/*
 * LineMesh simulation with the length-gauge Crank-Nicolson step
 * (meshes.py:263-325, mesh_operators.py:304-327, evolution_methods.py:46-77),
 * or with SplitInteractionOperator (method 1: length gauge, the phase
 * exp(-i tau V_int) around a CN step in K + V, mesh_operators.py:329-342;
 * method 2: velocity gauge, z-pair rotations around it, mesh_operators.py:
 * 352-427; evolution_methods.py:89-123).
 * tables: 5 * npad doubles (H0 diagonal tau0-scaled, off-diagonal to j-1,
 * off-diagonal to j+1, tau0 (-q) z, mask); wt: N complex (i tau_n / tau0);
 * fs: N doubles (field at t_{n+1}; method 2: the vector potential
 * -integral of E up to t_{n+1}); basis: n_basis * npad complex.
 */
typedef struct {
    int32_t z_points;      /* n */
    int32_t chunks;        /* P, multiple of 64, <= 512 */
    int32_t rows;          /* m, 2 <= m <= 16 */
    int32_t n_steps;
    int32_t store_every;
    int32_t n_basis;
    int32_t renormalize;   /* 1: renormalise after every step (imaginary-time relaxation) */
    int32_t method;        /* 0: ADI (CN in the total H); 1, 2: split operator, length / velocity gauge
                              (real time steps, no renormalisation) */
    double delta_z;
    double offdiag;        /* tau0 P_z, the off-diagonal inside the line (tables 1 and 2 hold it too) */
    double zq_first;       /* tau0 (-q) z_0: the z coupling of row j is zq_first + j zq_step */
    double zq_step;
    double vel_factor;     /* method 2: tau0 hbar (q / m) / (2 dz), the z-pair angle per unit fs * tau_n / tau0 */
} tdse_line_desc;

int tdse_line_add(tdse_ctx *ctx, const tdse_line_desc *desc, const double *tables, const double *wt,
                  const double *fs, const double *g0, const double *basis, int32_t *slot);

/*
 * CylindricalSliceMesh simulation, length gauge, AlternatingDirectionImplicit
 * (ionization/mesh/meshes.py:509-607, mesh_operators.py:550-631,
 * evolution_methods.py:45-77): per step the explicit rho half, implicit z,
 * explicit z, implicit rho, then the mask, with the Coulomb potential and the
 * interaction split half-and-half between the directions
 * (mesh_operators.py:244-298). Device layout: rho rows r < nrpad =
 * rho_chunks * rho_rows (even) by positions p < npz = z_chunks * z_rows,
 * position p holding z index (p % z_chunks) * z_rows + p / z_chunks; the two
 * rows of each pair (2 q, 2 q + 1) are interleaved position by position:
 * element (r, p) at ((r / 2) * npz + p) * 2 + r % 2 (g0, basis, downloads and
 * uploads alike).
 * The rho-sweep shape (rho_rows, lanes, rho_chunks) is one of (4, 4, <= 256),
 * (8, 4, <= 128), (4, 8, <= 128), (8, 8, <= 128), (16, 4, <= 128), with
 * rho_chunks * lanes a multiple of 64.
 */
typedef struct {
    int32_t z_points, rho_points;
    int32_t z_chunks, z_rows;      /* z-sweep: threads per z-line (multiple of 64), rows per thread */
    int32_t rho_chunks, rho_rows;  /* rho-sweep: threads per rho-line, rows per thread */
    int32_t lanes;                 /* rho-lines per rho-sweep workgroup */
    int32_t n_steps, store_every, n_basis;
    int32_t has_mask, reserved;
    double z_prefactor;            /* -hbar^2 / (2 mu dz^2)   (mesh_operators.py:561) */
    double rho_prefactor;          /* -hbar^2 / (2 mu drho^2) (mesh_operators.py:578-580) */
    double coulomb;                /* k Q q: V = coulomb / r  (potentials/static.py:26-38) */
    double test_charge;
    double delta;                  /* inner-product weight dz drho (meshes.py:541) */
    double mask_inner, mask_outer, mask_smoothness;  /* RadialCosineMask (potentials/masks.py:38-89) */
    double z_first, z_step;        /* z_j = z_first + j z_step (the kernels recompute the grid) */
    double rho_first, rho_step;    /* rho_j = rho_first + j rho_step */
} tdse_cyl_desc;

/*
 * Add one cylindrical simulation. zt: npz doubles (z of every position, 0 for
 * padding); rt: nrpad doubles (rho, 0 for padding); roff: nrpad doubles
 * (rho_prefactor * c(j + 1) between rho j and j + 1, 0 past the line,
 * mesh_operators.py:582-592); fs: N doubles E(t_{n+1}); ts: N doubles
 * tau_n = (t_{n+1} - t_n) / (2 hbar); g0 and basis: nrpad * npz complex each.
 * Replaces MeshSpecification.to_sim() + the first
 * AlternatingDirectionImplicit.evolve (sims.py:46-118, evolution_methods.py:48-77).
 */
int tdse_cyl_add(tdse_ctx *ctx, const tdse_cyl_desc *desc, const double *zt, const double *rt, const double *roff,
                 const double *fs, const double *ts, const double *g0, const double *basis, int32_t *slot);

/*
 * hold != 0: every later step of this simulation ends at the exact reference
 * g (the rho-sweep leaves the next step's explicit half to the next step), so
 * tdse_download after any run sees g after a whole number of steps; needed by
 * per-step callers of QuantumMesh.evolve, callbacks and checkpoints
 * (meshes.py:251-257, sims.py:296-330). hold = 0 fuses that half into the
 * previous sweep. No-op for SH and Line, whose states can always be made exact.
 */
int tdse_set_hold(tdse_ctx *ctx, int32_t kind, int32_t slot, int32_t hold);

/*
 * Advance every active simulation by up to n_steps steps (MeshSimulation.run,
 * sims.py:289-321). use_graph != 0 replays one captured hipGraph per step.
 * Finished simulations stay resident until removed. Inside one call, SH ADI
 * simulations may carry a step's mask and the next step's explicit half into
 * the previous replay; the call's last replay never does, so between calls
 * every simulation is at an ordinary step boundary.
 */
int tdse_run(tdse_ctx *ctx, int32_t n_steps, int32_t use_graph);

/* Advance only the LineMesh simulations, up to n_steps inside one launch. */
int tdse_line_run(tdse_ctx *ctx, int32_t n_steps);

/* kind: 0 = SH, 1 = Line, 2 = Cyl. step: steps taken (SH and Line: n_steps + 1
 * once finished; Cyl: n_steps once finished). */
int tdse_sim_state(tdse_ctx *ctx, int32_t kind, int32_t slot, int32_t *step, int32_t *n_steps, int32_t *n_data);

/*
 * Copy a simulation's wavefunction (position order) and its stored data back
 * to the host. Any pointer may be NULL to skip it. For SH: norms, lnorms
 * (norm in l = L-1), ips (n_data * n_basis complex). For Line: lnorms is
 * ignored. Replaces reading sim.mesh.g and the Norm / InnerProducts
 * datastores (data.py:189-287).
 */
int tdse_download(tdse_ctx *ctx, int32_t kind, int32_t slot, double *g, double *norms, double *lnorms, double *ips);

/*
 * Make the device copy of an SH simulation's wavefunction equal to the
 * reference g after its current step count (the step pipeline otherwise keeps
 * the last half of step n-1 fused with the first half of step n). Needed
 * before tdse_download when a caller inspects g mid-run: the per-step
 * QuantumMesh.evolve / callback / checkpoint points of MeshSimulation.run
 * (ionization/mesh/meshes.py:251-257, sims.py:296-330). No-op for LineMesh.
 * Cyl: succeeds when the state is exact (finished, or run with tdse_set_hold),
 * else returns TDSE_ERR_STATE.
 */
int tdse_materialize(tdse_ctx *ctx, int32_t kind, int32_t slot);

/*
 * The norm in every spherical harmonic l at every data time, n_data * L
 * doubles (data index major), for an SH simulation added with norm_by_l = 1.
 * Replaces SphericalHarmonicMesh.norm_by_l as stored by the
 * NormBySphericalHarmonic datastore (ionization/mesh/meshes.py:1130-1133,
 * data.py:438-458).
 */
int tdse_download_norm_by_l(tdse_ctx *ctx, int32_t slot, double *out);

/*
 * Overwrite a simulation's wavefunction (position order; the reference g after
 * `step` steps) and its step counter: checkpoint resume (sims.py:277-279, 348-356).
 */
int tdse_upload_state(tdse_ctx *ctx, int32_t kind, int32_t slot, const double *g, int32_t step);

/*
 * Health flags of simulations first_slot .. first_slot + n_slots - 1 of one
 * kind (0 SH, 1 Line, 2 Cyl), one int32 each: a bitwise OR of TDSE_FLAG_*,
 * set on the device where a data time's norm is stored and cleared by
 * tdse_upload_state. TDSE_FLAG_NONFINITE: the norm was NaN or infinite (a
 * diverged simulation); TDSE_FLAG_NORM_GROWTH: it exceeded the first or the
 * previous stored norm by more than 0.1 %, the two conditions
 * MeshSimulation.check warns about (ionization/mesh/sims.py:228-240). Slots
 * that are not live read as whatever they last held. Waits for the context's
 * stream.
 */
enum { TDSE_FLAG_NONFINITE = 1, TDSE_FLAG_NORM_GROWTH = 2 };
int tdse_sim_flags(tdse_ctx *ctx, int32_t kind, int32_t first_slot, int32_t n_slots, int32_t *flags);

/* Release a simulation's slot and device memory (serving retirement). */
int tdse_remove(tdse_ctx *ctx, int32_t kind, int32_t slot);

/*
 * Timing of the most recent tdse_run: total milliseconds between HIP events
 * recorded on the context's stream around the whole run.
 */
int tdse_last_run_ms(tdse_ctx *ctx, double *ms);

/*
 * Per-kernel timing over the next tdse_run: when enabled, the run launches
 * kernels directly (no graph) with HIP events around each L-phase and R-phase
 * launch on the context's stream, and accumulates their durations.
 */
int tdse_profile(tdse_ctx *ctx, int32_t enable);
int tdse_profile_read(tdse_ctx *ctx, double *lphase_ms, int64_t *lphase_launches, double *rphase_ms,
                      int64_t *rphase_launches);
/* Same accumulators by kernel: 0 = SH L-phase, 1 = SH R-phase, 2 = Cyl z-sweep, 3 = Cyl rho-sweep,
 * 4 = SH velocity-gauge h2 factors, 5 = SH velocity-gauge odd l-pairs (R-phase with their h2 factors). */
int tdse_profile_kernel(tdse_ctx *ctx, int32_t kernel, double *ms, int64_t *launches);

#ifdef __cplusplus
}
#endif

#endif /* TDSE_HIP_H */
