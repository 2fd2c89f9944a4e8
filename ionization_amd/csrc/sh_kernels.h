// SphericalHarmonicMesh, length gauge, split-interaction evolution on gfx950.
//
// One reference step n (ionization/mesh/evolution_methods.py:89-123, then the
// mask of meshes.py:256-257) is
//     g <- M . E_n O_n . CN . O_n E_n . g
// where E_n / O_n are the 2x2 rotations exp(-i a_k sigma_x) on the even / odd
// pairs (k, k+1) of the l-fastest flattening (mesh_operators.py:1037-1080),
// a_k = tau * E(t_{n+1} + dt/2) * (-q) * r_j * c_l (mesh_operators.py:988-1018),
// and CN = (1 + i tau H0)^-1 (1 - i tau H0) along r for every l
// (mesh_operators.py:889-928, cy.pyx:9-50).
//
// The device splits the step stream into two HBM sweeps per step:
//   L-phase: O_{n+1} E_{n+1} . M . E_n O_n   (one lane per radial point, streaming over l)
//   R-phase: CN                              (one workgroup per l-line, partition solve)
// so each step reads and writes g exactly twice (64 B per mesh point).
//
// Layout of one simulation in HBM: g[l][p], l in [0, L), p in [0, npad),
// npad = P * m with P = 64 w chunk-threads of m rows. Radial index j and
// position p are related by  j = (p % P) * m + p / P  (chunk k = p % P holds
// rows j = k m .. k m + m - 1; its row i sits at p = i P + k). With this
// permutation the R-phase thread k reading its row i and the L-phase lane p
// both issue unit-stride 16 B loads across the wave. Positions with j >= n
// are padding: zero in g, zero in every table.
#pragma once

#include "common.h"
#include "line_solver.h"

// Per-simulation device descriptor.
struct ShSim {
    double2* g;            // [L][npad] state before even steps (step counter n even)
    double2* g2;           // [L][npad] state before odd steps
    const double* tab;     // [SH_NTAB][npad] radial tables in position order
    const double* fs;      // [N] field times tau ratio per step: E(t_{n+1}+dt/2) * tau_n / tau0
    const double* ts;      // [N] tau_n / tau0
    double2* snap;         // [LB][npad] rows l < LB captured at data steps
    double* part;          // [2][nwaves] per-wave partial |g|^2 sums (all l, largest l); nwaves counts L-phase waves
    double* out_norm;      // [n_data]
    double* out_lnorm;     // [n_data]
    double2* out_ip;       // [n_data][n_basis]
    double* part_l;        // [nwaves][kShSegRows] per-wave |g|^2 sums of each row (norm by l), or null
    double* out_nl;        // [n_data][L] norm in each l, or null
    const double* basis;   // [n_basis][npad] radial g of each test state, position order
    const int* basis_l;    // [n_basis]
    const double2* ta;     // [npad] (T_j, A_j) interleaved: the R-phase's only table read per row
    double corr0;          // tau0 * (-2 P_r) * dr (1 + dr) / 8, added at l = 0, j = 0
    double delta_r;
    double cu;             // tau0 hbar^2 / (2 mu): U_j = cu / r_j^2, r_j = (j + 1/2) delta_r
    int n, L, P, m, npad, N, store_every, LB, n_basis, nwaves;
    int step;              // steps already advanced; device-updated
    int active;
    int phase;             // 0: g is mid-step (after the R-phase); 1: g is the reference g after `step` steps
    int norm_by_l;         // 1: also reduce the norm of every l at data times
    int adi;               // 1: AlternatingDirectionImplicit (fs is then zero: the L-phase only masks and observes)
    const double* fa;      // ADI: [N] field times tau ratio per step, E(t_{n+1} + dt/2) tau_n / tau0
    int flags;             // health flags (norm_health), device-updated at data times
    int ahead;             // ADI: 1 after a replay whose implicit r-sweep also applied the mask of its step
                           // and the explicit H0 half of the next one (device-updated; 0 between runs)
};

constexpr int kShSegRows = 64;  // most rows of one L-phase wave (its l-segment)

enum { SH_T = 0, SH_U = 1, SH_AP = 2, SH_AN = 3, SH_RQ = 4, SH_MASK = 5, SH_NTAB = 6 };
// velocity-gauge simulations carry one more row: tau0 hbar q / (2 m dr) alpha_j, zero for j >= n - 1
enum { SH_AV = 6 };

// The wavefunction ping-pongs between two buffers: the state with step
// counter n lives in sh_state(S, n). The L-phase of a step reads
// sh_state(S, n) and writes sh_state(S, n + 1) (its waves read input rows
// outside the l-segment they store, so it must not update in place); the
// R-phase then works in place on sh_state(S, n + 1).
__device__ __host__ __forceinline__ double2* sh_state(const ShSim& S, int n) { return (n & 1) ? S.g2 : S.g; }

// Time index n is a data time (sims.py:84-95): first, last and every store_every.
__device__ __forceinline__ bool sh_is_data_step(int n, int N, int every) {
    return n == 0 || n == N || (every >= 1 && (n % every) == 0);
}
// Slot of time index n among the stored data times.
__device__ __forceinline__ int sh_data_slot(int n, int N, int every) {
    if (every >= 1) return (n == N && (N % every) != 0) ? N / every + 1 : n / every;
    return n == 0 ? 0 : 1;
}
// AlternatingDirectionImplicit: may the implicit r-sweep of step n run ahead into step n + 1
// (apply the mask that ends step n and the explicit H0 half of step n + 1)? Only when the
// next replay of the same run follows (ok), step n + 1 starts there (n + 1 < N) and the g
// after n + 1 steps is not observed.
__device__ __forceinline__ bool sh_adi_ahead(const ShSim& S, int n, int ok) {
    return ok && n + 1 < S.N && !sh_is_data_step(n + 1, S.N, S.store_every);
}

// Length gauge (VEL = false): exp(-i a sigma_x) applied to (x0, x1),
//     y0 = cos a x0 - i sin a x1,  y1 = -i sin a x0 + cos a x1.
// Velocity gauge (VEL = true): the real rotation of split_h1 / split_h2
// (mesh_operators.py:1201-1250, 1252-1408), diagonal cos a, super-diagonal
// sin a, sub-diagonal -sin a:
//     y0 = cos a x0 + sin a x1,    y1 = -sin a x0 + cos a x1.
// When every lane's angle is small (a weak field: most steps away from a
// pulse's peak) the whole wave takes the reduction-free series.
template <bool VEL = false>
__device__ __forceinline__ void sh_rotate(double a, double2& x0, double2& x1) {
    double s, c;
    if (__all(fabs(a) <= kSmallAngle)) tdse_sincos_small(a, &s, &c);
    else tdse_sincos(a, &s, &c);
    if constexpr (VEL) {
        const double2 y0 = cmk(fma(c, x0.x, s * x1.x), fma(c, x0.y, s * x1.y));
        const double2 y1 = cmk(fma(c, x1.x, -s * x0.x), fma(c, x1.y, -s * x0.y));
        x0 = y0;
        x1 = y1;
    } else {
        const double2 y0 = cmk(fma(c, x0.x, s * x1.y), fma(c, x0.y, -s * x1.x));
        const double2 y1 = cmk(fma(c, x1.x, s * x0.y), fma(c, x1.y, -s * x0.x));
        x0 = y0;
        x1 = y1;
    }
}

// ------------------------------------------------------------------ L-phase
//
// A wave covers 64 consecutive positions of one simulation and one segment
// [lo, hi) of its l range (wave table entry: sim, first position, lo, hi).
// Each lane streams over l carrying a three-stage pipeline: virtual row
// v = l + q (q = parity of j*L) puts the E-pairs on even v and the O-pairs on
// odd v for every lane, so iteration it consumes rows (2it, 2it+1) and
// finishes rows (2it-3, 2it-2). A finished row depends on input rows within
// distance 3 (O, E, O pair chains), so a segment starts its pipeline 4 rows
// below lo and stores only rows in [lo, hi); segments let short and long l
// ranges spread evenly over the GPU. Row loads are issued PF iterations ahead.
//
// Mode per simulation, from its step counter n (N = total steps) and phase:
//   start      E_n, O_n                          (n < N)
//   end        O_{n-1}, E_{n-1}, mask            (n > 0 and g is mid-step)
// At data times the reference g after n steps (between "end" and "start")
// is observed: per-wave |g|^2 partials and the rows l < LB. Elsewhere the two
// E rotations of a pair, which share the generator sigma_x and commute with
// the mask (a function of r only), are applied as one rotation by the sum of
// their angles. end_only != 0 materialises the reference g after n steps for
// the listed waves without starting step n and without observing; it too
// writes sh_state(S, n + 1) (the host then swaps the two buffers).
struct ShWave {
    int sim, p0, lo, hi, part;  // part: index of this wave's norm partial within the simulation
};

// BY_L: the instance that can also reduce the norm in every l (launched only
// while some simulation asks for it: the reductions cost registers).
// VEL: velocity-gauge simulations. Their h1 coupling has the same pairs and
// the same place in the step as the length gauge's E / O (split_h1,
// mesh_operators.py:1201-1250): the host supplies the angle factors (RQ row:
// tau0 hbar (q / m) / r_j; fs: the vector potential; cl: c_l (l + 1)) and the
// rotation is the real one.
template <bool BY_L, bool VEL>
__global__ void __launch_bounds__(256) k_sh_lphase(ShSim* __restrict__ sims, const ShWave* __restrict__ waves,
                                                   int n_waves, const double* __restrict__ cl, int end_only) {
    constexpr int PF = 4;  // prefetch depth in iterations (2 rows each)
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / TDSE_WAVE;
    const int lane = threadIdx.x % TDSE_WAVE;
    if (wave >= n_waves) return;
    const ShWave wt = waves[wave];
    ShSim& S = sims[wt.sim];
    const int n = S.step;
    const int N = S.N;
    if (!S.active || n > N) return;
    if (S.ahead) return;  // ADI: the previous replay already masked g and started this step in place of it
    const bool do_end = n > 0 && S.phase == 0;
    const bool do_start = n < N && !end_only;
    const bool observe = !end_only && sh_is_data_step(n, N, S.store_every);
    // Every wave writes its rows to the other buffer, even with nothing to apply (a copy).
    const bool both = do_end && do_start;

    const int L = S.L, P = S.P, m = S.m, npad = S.npad;
    const int lo = wt.lo, hi = wt.hi;
    // read once: inside the row loop the snapshot stores could otherwise make the compiler
    // reload these fields every iteration
    const int LB = S.LB;
    double2* __restrict__ const snap = S.snap;
    double* __restrict__ const part_l = S.part_l;
    const int p = wt.p0 + lane;
    const int j = (p % P) * m + p / P;
    const bool live = j < S.n;
    const int q = (j & 1) & (L & 1);  // parity of j*L
    const double rq = ld1(S.tab, SH_RQ * npad + p);
    const double mk = ld1(S.tab, SH_MASK * npad + p);
    const double f_end = do_end ? ld1(S.fs, n - 1) * rq : 0.0;
    const double f_start = do_start ? ld1(S.fs, n) * rq : 0.0;
    const double f_sum = f_end + f_start;
    const double2* __restrict__ gin = sh_state(S, n) + p;
    double2* __restrict__ g = sh_state(S, n + 1) + p;
    // Output row r depends on input rows up to r + 3, so rows from hi + 3 on
    // (which the prefetch ring would otherwise read up to 2 PF rows ahead) are
    // not loaded: the next segment's wave reads them.
    const int l_stop = min(L, hi + 3);
    auto load_row = [gin, live, l_stop, npad](int l) -> double2 {
        double2 v = make_double2(0.0, 0.0);
        if (live && l >= 0 && l < l_stop) v = ld2(gin, (size_t)l * npad);
        return v;
    };
    // coupling of the pair starting at virtual row v (0 when the pair does not exist:
    // a rotation by 0 is exactly the identity)
    auto load_c = [cl, q, L](int v) -> double {
        const int l = v - q;
        return (l >= 0 && l + 1 < L) ? ld1(cl, l) : 0.0;
    };

    // first iteration: rows below lo - 4 do not influence rows >= lo
    // Rows of one lane's iteration depend on its parity q, so with odd L the
    // lanes of a wave may have different bounds. Reducing the norm of every
    // row across the wave (norm by l) needs the same bounds for all lanes: the
    // union is used then (an earlier start or later end only adds pipeline
    // warm-up / run-out on rows that are not stored).
    const bool by_l = BY_L && observe && S.norm_by_l;
    const int it0 = (lo <= 4) ? 0 : ((lo + (by_l ? 0 : q) - 4) >> 1);
    const int it_end = (hi - 1 + (by_l ? 1 : q) + 3) >> 1;  // last iteration stores row hi-1
    const double2 zero = make_double2(0.0, 0.0);
    double2 carry_u = zero, carry_s1 = zero, carry_s3 = zero;
    double c_m3 = load_c(2 * it0 - 3), c_m2 = load_c(2 * it0 - 2), c_m1 = load_c(2 * it0 - 1);
    double acc_norm = 0.0, acc_lnorm = 0.0;
    double pend = 0.0;  // norm by l: the q = 0 lanes' share of odd row v0 - 1, completed next iteration
    double2 ring0[PF], ring1[PF];
    double ringc0[PF], ringc1[PF];
#pragma unroll
    for (int d = 0; d < PF; ++d) {
        ring0[d] = load_row(2 * (it0 + d) - q);
        ring1[d] = load_row(2 * (it0 + d) + 1 - q);
        ringc0[d] = load_c(2 * (it0 + d));
        ringc1[d] = load_c(2 * (it0 + d) + 1);
    }

    for (int it = it0; it <= it_end; ++it) {
        const int v0 = 2 * it;
        const double2 u0 = ring0[0], u1 = ring1[0];
        const double c0 = ringc0[0], c1 = ringc1[0];
#pragma unroll
        for (int d = 0; d + 1 < PF; ++d) {
            ring0[d] = ring0[d + 1];
            ring1[d] = ring1[d + 1];
            ringc0[d] = ringc0[d + 1];
            ringc1[d] = ringc1[d + 1];
        }
        ring0[PF - 1] = load_row(v0 + 2 * PF - q);
        ring1[PF - 1] = load_row(v0 + 2 * PF + 1 - q);
        ringc0[PF - 1] = load_c(v0 + 2 * PF);
        ringc1[PF - 1] = load_c(v0 + 2 * PF + 1);

        // stage 1: O_end on (v0-1, v0)
        double2 s1m = carry_u, s10 = u0;
        if (do_end) sh_rotate<VEL>(f_end * c_m1, s1m, s10);
        // stage 2: E_end on (v0-2, v0-1), mask; stage 3: E_start on the same pair
        double2 s2a = carry_s1, s2b = s1m;
        const int le = v0 - 2 - q;
        if (both && !observe) {
            // E_end and E_start as one rotation (the mask, a function of r only, commutes with both)
            sh_rotate<VEL>(f_sum * c_m2, s2a, s2b);
            s2a = cscale(mk, s2a);
            s2b = cscale(mk, s2b);
        } else {
            // At a data time the stored rows are the same as elsewhere (one rotation when both halves
            // apply), and the pair's norm is that of the rotated rows (E_start keeps it). Rows observed
            // on their own -- below LB (snapshots for the inner products), the largest l, rows outside
            // [lo, hi) (counted by the neighbouring segment), every row for the norm by l -- take E_end
            // and the mask once more by themselves.
            // decided for the whole wave (lanes differ only by the parity q): no divergent rotations
            const bool per_row = __any(!both || le < LB || le < lo || le + 1 >= hi || le + 1 >= L - 1);
            double2 oa = s2a, ob = s2b;  // the observed rows: g after step n
            if (both) {
                if (by_l || per_row) {
                    sh_rotate<VEL>(f_end * c_m2, oa, ob);
                    oa = cscale(mk, oa);
                    ob = cscale(mk, ob);
                }
                sh_rotate<VEL>(f_sum * c_m2, s2a, s2b);
                s2a = cscale(mk, s2a);
                s2b = cscale(mk, s2b);
            } else if (do_end) {
                sh_rotate<VEL>(f_end * c_m2, s2a, s2b);
                s2a = cscale(mk, s2a);
                s2b = cscale(mk, s2b);
                oa = s2a;
                ob = s2b;
            }
            if (observe && live) {
                if (!per_row) {
                    acc_norm += cabs2(s2a) + cabs2(s2b);
                } else {
                    if (le >= lo && le < hi) {
                        acc_norm += cabs2(oa);
                        if (le == L - 1) acc_lnorm += cabs2(oa);
                        if (le < LB) st2(snap, (size_t)le * npad + p, oa);
                    }
                    if (le + 1 >= lo && le + 1 < hi) {
                        acc_norm += cabs2(ob);
                        if (le + 1 == L - 1) acc_lnorm += cabs2(ob);
                        if (le + 1 < LB) st2(snap, (size_t)(le + 1) * npad + p, ob);
                    }
                }
            }
            if (by_l) {
                // this iteration finishes rows v0-2 (all lanes), v0-1 (q = 0 lanes)
                // and v0-3 (q = 1 lanes); sums in a fixed order per wave
                const bool ok_a = live && le >= lo && le < hi, ok_b = live && le + 1 >= lo && le + 1 < hi;
                const double na = ok_a ? cabs2(oa) : 0.0, nb = ok_b ? cabs2(ob) : 0.0;
                const double r3 = wave_sum(q ? na : 0.0);  // row v0-3
                const double r2 = wave_sum(q ? nb : na);   // row v0-2
                const double r1 = wave_sum(q ? 0.0 : nb);  // row v0-1
                if (lane == 0) {
                    double* pl = part_l + (size_t)wt.part * kShSegRows;
                    if (v0 - 3 >= lo && v0 - 3 < hi) pl[v0 - 3 - lo] = pend + r3;
                    if (v0 - 2 >= lo && v0 - 2 < hi) pl[v0 - 2 - lo] = r2;
                }
                pend = r1;
            }
            if (!both && do_start) sh_rotate<VEL>(f_start * c_m2, s2a, s2b);
        }
        // stage 4: O_start on (v0-3, v0-2)
        double2 o_a = carry_s3, o_b = s2a;
        if (do_start) sh_rotate<VEL>(f_start * c_m3, o_a, o_b);
        const int lo_row = v0 - 3 - q;
        if (live) {
            if (lo_row >= lo && lo_row < hi) st2s(g, (size_t)lo_row * npad, o_a);
            if (lo_row + 1 >= lo && lo_row + 1 < hi) st2s(g, (size_t)(lo_row + 1) * npad, o_b);
        }
        carry_u = u1;
        carry_s1 = s10;
        carry_s3 = s2b;
        c_m3 = c_m1;
        c_m2 = c0;
        c_m1 = c1;
    }

    if (observe) {
        acc_norm = wave_sum(acc_norm);
        acc_lnorm = wave_sum(acc_lnorm);
        if (lane == 0) {
            S.part[wt.part] = acc_norm;
            S.part[S.nwaves + wt.part] = acc_lnorm;
        }
    }
}

// ------------------------------------------------------------------ observables
//
// Job table entry (sim, b): b == -1 reduces the norm partials, b >= 0 forms
// the inner product of test state b with the captured row l_b:
//     <phi_b | g> = sum_j phi_b(j) g(l_b, j) dr        (meshes.py:1117-1129)
// Sums run in a fixed order per simulation, independent of the batch.
__global__ void __launch_bounds__(256) k_sh_observe(ShSim* __restrict__ sims, const int2* __restrict__ jobs,
                                                    int n_jobs) {
    const int job = blockIdx.x;
    if (job >= n_jobs) return;
    const int2 jb = jobs[job];
    ShSim& S = sims[jb.x];
    const int n = S.step;
    if (!S.active || n > S.N || !sh_is_data_step(n, S.N, S.store_every)) return;
    const int slot = sh_data_slot(n, S.N, S.store_every);
    __shared__ double red_re[256];
    __shared__ double red_im[256];
    const int t = threadIdx.x;
    double re = 0.0, im = 0.0;
    if (jb.y == -2) {
        // norm in every l: rows of l-segment sg sit in the partials of its waves
        // sg * wpl .. sg * wpl + wpl - 1 (wave table order), summed in that order
        const int L = S.L;
        const int nseg = S.nwaves / (S.npad / TDSE_WAVE);
        const int seg = (L + nseg - 1) / nseg;
        const int wpl = S.npad / TDSE_WAVE;
        for (int l = t; l < L; l += blockDim.x) {
            const int sg = l / seg, row = l - sg * seg;
            double acc = 0.0;
            for (int w = 0; w < wpl; ++w) acc += ld1(S.part_l, (size_t)(sg * wpl + w) * kShSegRows + row);
            S.out_nl[(size_t)slot * L + l] = acc * S.delta_r;
        }
        return;
    }
    if (jb.y < 0) {
        for (int w = t; w < S.nwaves; w += blockDim.x) {
            re += ld1(S.part, w);
            im += ld1(S.part, S.nwaves + w);
        }
    } else {
        const double* phi = S.basis + (size_t)jb.y * S.npad;
        const double2* row = S.snap + (size_t)S.basis_l[jb.y] * S.npad;
        for (int p = t; p < S.npad; p += blockDim.x) {
            const double f = ld1(phi, p);
            const double2 z = ld2(row, p);
            re = fma(f, z.x, re);
            im = fma(f, z.y, im);
        }
    }
    red_re[t] = re;
    red_im[t] = im;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (t < s) {
            red_re[t] += red_re[t + s];
            red_im[t] += red_im[t + s];
        }
        __syncthreads();
    }
    if (t == 0) {
        if (jb.y < 0) {
            const double nrm = red_re[0] * S.delta_r;
            const int health = norm_health(S.out_norm, slot, nrm);
            S.out_norm[slot] = nrm;
            S.out_lnorm[slot] = red_im[0] * S.delta_r;
            if (health) S.flags |= health;
        } else {
            S.out_ip[(size_t)slot * S.n_basis + jb.y] = cmk(red_re[0] * S.delta_r, red_im[0] * S.delta_r);
        }
    }
}

// ------------------------------------------------------------------ R-phase

enum { CN_CAYLEY_MEM = 0, CN_CAYLEY_LDS = 1, CN_EXPLICIT = 2 };

// Shared tail of the CN line kernels: x holds the thread's M rows of the line
// at gl (row i at gl[i * stride]; stride P by default); solves and stores
// y = (1 + i H~)^-1 (1 - i H~) x. Ratio / ib: optional pass-2 pivot ratios
// (cn_implicit_lowreg). LANES / ln: LANES lines per workgroup, interleaved in LDS.
// store = false: the thread solves but writes nothing (rows that are all padding).
template <int M, int FORM, class Coef, class Coef2, class Ratio = NoRatio, int LANES = 1>
__device__ __forceinline__ void cn_line_finish(double2 (&x)[M], double2* __restrict__ gl, const int k, const int P,
                                               Coef coef, Coef2 coef2, double2* __restrict__ lds,
                                               Ratio ratio = Ratio{}, const int ib = -1, const int ln = 0,
                                               const int stride = -1, const bool store = true) {
    const size_t st = stride < 0 ? (size_t)P : (size_t)stride;
    if constexpr (FORM == CN_EXPLICIT) {
        double2 xl, xr;
        chunk_neighbours<M, LANES>(x, k, P, lds, xl, xr, false, ln);
        cn_implicit_lowreg<M, Coef, Coef2, true, LANES, Ratio>(x, k, P, coef, coef2, lds, xl, xr, ln, ratio, ib);
        if (store) {
#pragma unroll
            for (int i = 0; i < M; ++i) st2s(gl, (size_t)i * st, x[i]);
        }
    } else {
        double2* keep = lds + 3 * P * LANES;  // CN_CAYLEY_LDS: this thread's own rows only
        if constexpr (FORM == CN_CAYLEY_LDS) {
#pragma unroll
            for (int i = 0; i < M; ++i) keep[(i * P + k) * LANES + ln] = x[i];
        }
        cn_implicit_lowreg<M, Coef, Coef2, false, LANES, Ratio>(x, k, P, coef, coef2, lds, cmk(0.0, 0.0),
                                                                cmk(0.0, 0.0), ln, ratio, ib);
        if (!store) return;
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const double2 x0 = (FORM == CN_CAYLEY_LDS) ? keep[(i * P + k) * LANES + ln] : ld2(gl, (size_t)i * st);
            st2s(gl, (size_t)i * st, cmk(fma(2.0, x[i].x, -x0.x), fma(2.0, x[i].y, -x0.y)));
        }
    }
}
//
// (1 + i H~) y = (1 - i H~) x along r for one l-line, H~ = tau_n H0 real
// symmetric tridiagonal: diagonal s_n (T_j + l(l+1) U_j [+ corr0 at l = 0,
// j = 0]), off-diagonals s_n A_{j-1}, s_n A_j (mesh_operators.py:889-928,
// 244-269). One workgroup of P threads per line, M (= m) rows per thread,
// solved by cn_implicit_lowreg, which keeps two complex values per row between
// its passes and re-reads the (L2-resident) coefficient tables instead of
// holding them, so WPE waves fit per SIMD; that occupancy hides the HBM
// latency of the line loads. Two algebraically equal forms:
//  * EXPLICIT: the right-hand side (1 - i H~) x is formed inside the solver's
//    first pass from x and the neighbouring chunks' edge rows; every line is
//    read once and written once (faster for M >= 10);
//  * Cayley: z = (1 + i H~)^-1 x, y = 2 z - x (fewer live registers; faster
//    for M <= 8), with the original x kept in LDS (CN_CAYLEY_LDS: M P more
//    double2 of LDS; each thread reads back only what it wrote) or re-read
//    from memory at the end (CN_CAYLEY_MEM).

template <int M, int MAXT, int WPE, int FORM>
__global__ void __launch_bounds__(MAXT) __attribute__((amdgpu_waves_per_eu(WPE, 8)))
k_sh_rphase(ShSim* __restrict__ sims, const int2* __restrict__ lines, int n_lines) {
    extern __shared__ double2 lds[];  // 3 * P, then M * P for CN_CAYLEY_LDS
    // consecutive l-lines of a simulation (which share its coefficient rows) on one XCD
    const int line = xcd_remap(blockIdx.x, gridDim.x);
    if (line >= n_lines) return;
    const int2 lt = lines[line];
    ShSim& S = sims[lt.x];
    const int n = S.step;
    if (!S.active || n >= S.N) return;  // uniform across the block
    const int l = lt.y;
    const int P = S.P, npad = S.npad;
    const int k = threadIdx.x;
    const double sn = ld1(S.ts, n);
    const double ll1 = (double)l * (double)(l + 1);
    const double c0 = (l == 0 && k == 0) ? S.corr0 : 0.0;
    double2* __restrict__ gl = sh_state(S, n + 1) + (size_t)l * npad + k;
    // One 16 B read per row and pass: (T_j, A_j). U_j = cu / r_j^2 is recomputed
    // and the sub-diagonal A_{j-1} is carried from the previous row (the
    // chunk's first row takes it from the neighbouring chunk's last row).
    const double2* __restrict__ TA = S.ta + k;
    const double dr = S.delta_r, cu = S.cu;
    const int nr = S.n;
    const double a_left = (k > 0) ? ld1(S.tab, SH_AN * npad + (M - 1) * P + k - 1) : 0.0;

    // A thread whose rows all lie past the last radial point (j >= n) holds padding: zero in
    // memory, identity rows in the tables, decoupled from row n - 1 (A_{n-1} = 0). It solves
    // with zeros and neither reads nor writes its rows.
    const bool pad = k * M >= nr;
    double2 x[M];
    if (pad) {
#pragma unroll
        for (int i = 0; i < M; ++i) x[i] = cmk(0.0, 0.0);
    } else {
#pragma unroll
        for (int i = 0; i < M; ++i) x[i] = (FORM == CN_CAYLEY_MEM) ? ld2(gl, (size_t)i * P) : ld2s(gl, (size_t)i * P);
    }
    auto diag = [&](int i, double T) {
        const int j = k * M + i;
        const double r = fma((double)j, dr, 0.5 * dr);
        const double U = (j < nr) ? cu * tdse_rcp(r * r) : 0.0;
        return sn * (fma(ll1, U, T) + (i == 0 ? c0 : 0.0));
    };
    double carry = a_left;  // A_{j-1} of the next row (pass 1 visits rows in order)
    auto coef = [&](int i, double& hm, double& h, double& hp) {
        const double2 tv = ld2(TA, (size_t)i * P);
        h = diag(i, tv.x);
        hm = sn * carry;
        hp = sn * tv.y;
        carry = tv.y;
    };
    // Second read of the same coefficients (pass 2, rows in order). The empty
    // asm makes the pointer opaque so the compiler reloads instead of keeping
    // the first pass's values live across the solver's barriers.
    double carry2 = a_left;
    auto coef2 = [&](int i, double& hm, double& h) {
        const double2* TA2 = TA;
        asm volatile("" : "+v"(TA2));
        const double2 tv = ld2(TA2, (size_t)i * P);
        h = diag(i, tv.x);
        hm = sn * carry2;
        carry2 = tv.y;
    };
    cn_line_finish<M, FORM>(x, gl, k, P, coef, coef2, lds, NoRatio{}, -1, 0, -1, !pad);
}

// ------------------------------------------------------------------ velocity gauge: h2
//
// The radial part h2 of the velocity-gauge interaction couples (l, j) with
// (l +- 1, j +- 1): A(t) i hbar q / (2 m dr) kron(C, Alpha) with C the
// symmetric l-coupling (c_l off the diagonal) and Alpha the antisymmetric
// r-coupling (alpha_j above, -alpha_j below the diagonal)
// (mesh_operators.py:1143-1180). split_h2 (mesh_operators.py:1252-1408)
// factors exp(-i tau h2) into four SimilarityOperators (mesh_operators.py:
// 150-194): for the l-pairs (l, l + 1) of one parity (even l first, then odd
// l) it maps the pair to s = (g_l + g_{l+1}) / sqrt 2, d = (g_l - g_{l+1}) /
// sqrt 2, which diagonalises C within the pair, rotates s by +theta_j and d
// by -theta_j on the radial pairs (j, j + 1) of even j, then on those of odd
// j, and maps back. theta_j = tau_n A_n hbar q / (2 m dr) c_l alpha_j (zero
// for j >= n - 1: the pair (n - 1, n) does not exist). The reversed half of
// the step (after the CN solve) applies the odd-j pairs first.
//
// One workgroup per l-pair, the R-phase layout: thread k holds radial rows
// j = k M .. k M + M - 1 (position i P + k) of both l-rows in registers and
// updates them in place. With M even the even-j pairs lie inside a thread;
// of the odd-j pairs only the one across a chunk edge (k M + M - 1, (k + 1) M)
// needs the neighbouring thread's row, exchanged through LDS (both threads
// compute that pair's rotation from the same values and keep their half).
// Rotate (s0, s1) by +th and (d0, d1) by -th (one sine and cosine for both).
__device__ __forceinline__ void sh_vel_pair(double th, double2& s0, double2& s1, double2& d0, double2& d1) {
    double sn, cs;
    if (__all(fabs(th) <= kSmallAngle)) tdse_sincos_small(th, &sn, &cs);
    else tdse_sincos(th, &sn, &cs);
    const double2 a0 = cmk(fma(cs, s0.x, sn * s1.x), fma(cs, s0.y, sn * s1.y));
    const double2 a1 = cmk(fma(cs, s1.x, -sn * s0.x), fma(cs, s1.y, -sn * s0.y));
    const double2 b0 = cmk(fma(cs, d0.x, -sn * d1.x), fma(cs, d0.y, -sn * d1.y));
    const double2 b1 = cmk(fma(cs, d1.x, sn * d0.x), fma(cs, d1.y, sn * d0.y));
    s0 = a0;
    s1 = a1;
    d0 = b0;
    d1 = b1;
    // the rotated rows are produced here, pair by pair: left to itself the compiler
    // evaluates every pair's sine and cosine first and spills at 14 and 16 rows
    asm("" : "+v"(s0.x), "+v"(s0.y), "+v"(s1.x), "+v"(s1.y), "+v"(d0.x), "+v"(d0.y), "+v"(d1.x), "+v"(d1.y));
}

// Position of row i of "virtual chunk" c when SPLIT threads share each chunk of M rows:
// thread c holds MH = M / SPLIT consecutive radial rows j = c MH + i, rows h MH + i of
// chunk k = c / SPLIT (h = c % SPLIT), at position (h MH + i) P + k.
template <int M, int SPLIT>
__device__ __forceinline__ size_t sh_vel_pos(const int c, const int i, const int P) {
    constexpr int MH = M / SPLIT;
    return (size_t)((c % SPLIT) * MH + i) * P + c / SPLIT;
}

// The radial pairs of one parity for thread c's MH rows (MH even: the even-j pairs lie inside
// the thread; of the odd-j pairs the ones across the thread's edges, (c MH - 1, c MH) and
// (c MH + MH - 1, (c + 1) MH), go through LDS). nc = SPLIT P threads.
template <int M, int SPLIT>
__device__ __forceinline__ void sh_vel_rpairs(double2 (&s)[M / SPLIT], double2 (&d)[M / SPLIT],
                                              const double* __restrict__ av, const double f, const int P,
                                              const int c, const bool odd, double2 (*ex)[512]) {
    constexpr int MH = M / SPLIT;
    static_assert(MH % 2 == 0, "the even-j pairs must lie inside a thread");
    const int nc = SPLIT * P;
    auto pos = [&](int cc, int i) { return sh_vel_pos<M, SPLIT>(cc, i, P); };
    if (!odd) {
#pragma unroll
        for (int i = 0; i < MH; i += 2) sh_vel_pair(f * ld1(av, pos(c, i)), s[i], s[i + 1], d[i], d[i + 1]);
        return;
    }
#pragma unroll
    for (int i = 1; i + 1 < MH; i += 2) sh_vel_pair(f * ld1(av, pos(c, i)), s[i], s[i + 1], d[i], d[i + 1]);
    ex[0][c] = s[MH - 1];
    ex[1][c] = d[MH - 1];
    ex[2][c] = s[0];
    ex[3][c] = d[0];
    __syncthreads();
    // the edge pair to the right: (my row MH - 1, the next thread's row 0)
    double2 sr = (c + 1 < nc) ? ex[2][c + 1] : cmk(0.0, 0.0), dr = (c + 1 < nc) ? ex[3][c + 1] : cmk(0.0, 0.0);
    const double thr = (c + 1 < nc) ? f * ld1(av, pos(c, MH - 1)) : 0.0;
    sh_vel_pair(thr, s[MH - 1], sr, d[MH - 1], dr);
    // the edge pair to the left: (the previous thread's row MH - 1, my row 0)
    double2 sl = (c > 0) ? ex[0][c - 1] : cmk(0.0, 0.0), dl = (c > 0) ? ex[1][c - 1] : cmk(0.0, 0.0);
    const double thl = (c > 0) ? f * ld1(av, pos(c > 0 ? c - 1 : 0, MH - 1)) : 0.0;
    sh_vel_pair(thl, sl, s[0], dl, d[0]);
}

// REV = false: the half before the CN solve (even-j pairs, then odd-j);
// REV = true: the half after it (odd-j, then even-j). The launch covers the
// l-pairs of one parity; the host orders the launches (even l-pairs, odd
// l-pairs) before the solve and (odd, even) after it. SPLIT threads per chunk
// (SPLIT P per workgroup), each with M / SPLIT rows: fewer registers, more waves.
template <int M, bool REV, int SPLIT>
__global__ void __launch_bounds__(512) k_sh_vphase(ShSim* __restrict__ sims, const int2* __restrict__ pairs,
                                                   int n_pairs, const double* __restrict__ cl) {
    constexpr int MH = M / SPLIT;
    __shared__ double2 ex[4][512];
    const int pair = xcd_remap(blockIdx.x, gridDim.x);
    if (pair >= n_pairs) return;
    const int2 pt = pairs[pair];
    ShSim& S = sims[pt.x];
    const int n = S.step;
    if (!S.active || n >= S.N) return;  // uniform across the block
    const double fn = ld1(S.fs, n);
    if (fn == 0.0) return;  // every angle is zero: the factor is the identity
    const int l = pt.y;
    const int P = S.P, npad = S.npad;
    const int c = threadIdx.x;
    const double f = fn * ld1(cl, l);
    double2* __restrict__ g0 = sh_state(S, n + 1) + (size_t)l * npad;
    double2* __restrict__ g1 = g0 + npad;
    const double* __restrict__ av = S.tab + (size_t)SH_AV * npad;
    constexpr double kRsqrt2 = 0.70710678118654752440;
    // all rows padding (j >= n): zeros with zero angles (alpha_j = 0 from j = n - 1 on), neither
    // read nor written
    const bool pad = c * MH >= S.n;
    double2 s[MH], d[MH];
    if (pad) {
#pragma unroll
        for (int i = 0; i < MH; ++i) s[i] = d[i] = cmk(0.0, 0.0);
    } else {
#pragma unroll
        for (int i = 0; i < MH; ++i) {
            const size_t q = sh_vel_pos<M, SPLIT>(c, i, P);
            const double2 a = ld2s(g0, q), b = ld2s(g1, q);
            s[i] = cmk((a.x + b.x) * kRsqrt2, (a.y + b.y) * kRsqrt2);
            d[i] = cmk((a.x - b.x) * kRsqrt2, (a.y - b.y) * kRsqrt2);
        }
    }
    sh_vel_rpairs<M, SPLIT>(s, d, av, f, P, c, REV, ex);
    sh_vel_rpairs<M, SPLIT>(s, d, av, f, P, c, !REV, ex);
    if (pad) return;
#pragma unroll
    for (int i = 0; i < MH; ++i) {
        const size_t q = sh_vel_pos<M, SPLIT>(c, i, P);
        st2s(g0, q, cmk((s[i].x + d[i].x) * kRsqrt2, (s[i].y + d[i].y) * kRsqrt2));
        st2s(g1, q, cmk((s[i].x - d[i].x) * kRsqrt2, (s[i].y - d[i].y) * kRsqrt2));
    }
}

// Velocity gauge: the R-phase of one odd l-pair (l, l + 1) together with the
// pair's h2 similarity factors (OE, OO before the CN solve; OO, OE after it), so
// the odd pairs' factors cost no sweep of their own. The two rows are
// interleaved over the workgroup: thread t works on row l + (t & 1) and chunk
// k = t >> 1, so the partner value of the pair transform sits in the
// neighbouring lane (one lane swap, no LDS); the CN solve is the R-phase's,
// with two lines per workgroup (LANES = 2 in LDS). Both rows have l >= 1: no
// l = 0 correction.
// The two lanes of a chunk (rows l and l + 1 of the pair) turn by +theta and
// -theta on the same radial pairs: each lane computes the sine and cosine of
// every other pair's angle and takes the rest from its neighbour lane.
template <int M>
__device__ __forceinline__ void sh_vpair_factors(double2 (&x)[M], const int k, const int P, const int ln,
                                                 double2* __restrict__ lds, const double* __restrict__ av,
                                                 const double f, const bool rev) {
    constexpr double kRsqrt2 = 0.70710678118654752440;
    const double sg = ln ? -1.0 : 1.0;
    // to the pair basis and back: row l keeps (own + partner) / sqrt 2, row l + 1 (partner - own) / sqrt 2
    auto hadamard = [&]() {
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const double2 o = cmk(__shfl_xor(x[i].x, 1), __shfl_xor(x[i].y, 1));
            x[i] = ln ? cmk((o.x - x[i].x) * kRsqrt2, (o.y - x[i].y) * kRsqrt2)
                      : cmk((x[i].x + o.x) * kRsqrt2, (x[i].y + o.y) * kRsqrt2);
        }
    };
    // the real rotation by sg * theta, from the sine and cosine of theta
    auto rot = [&](double sn, const double cs, double2& u, double2& v) {
        sn *= sg;
        const double2 y0 = cmk(fma(cs, u.x, sn * v.x), fma(cs, u.y, sn * v.y));
        const double2 y1 = cmk(fma(cs, v.x, -sn * u.x), fma(cs, v.y, -sn * u.y));
        u = y0;
        v = y1;
    };
    // sine and cosine of lane 0's angle ta and lane 1's angle tb, on both lanes
    auto sincos2 = [&](const double ta, const double tb, double& sa, double& ca, double& sb, double& cb) {
        const double th = ln ? tb : ta;
        double sm, cm;
        if (__all(fabs(th) <= kSmallAngle)) tdse_sincos_small(th, &sm, &cm);
        else tdse_sincos(th, &sm, &cm);
        const double so = __shfl_xor(sm, 1), co = __shfl_xor(cm, 1);
        sa = ln ? so : sm;
        ca = ln ? co : cm;
        sb = ln ? sm : so;
        cb = ln ? cm : co;
    };
    auto th = [&](int i) { return f * ld1(av, (size_t)i * P + k); };
    // radial pairs (i, i + 1), i = first, first + 2, ... < M - 1, two per sine/cosine evaluation
    auto pairs = [&](const int first) {
#pragma unroll
        for (int i = first; i + 1 < M; i += 4) {
            const int i2 = i + 2;
            const double ta = th(i);
            const double tb = (i2 + 1 < M) ? th(i2) : ta;
            double sa, ca, sb, cb;
            sincos2(ta, tb, sa, ca, sb, cb);
            rot(sa, ca, x[i], x[i + 1]);
            if (i2 + 1 < M) rot(sb, cb, x[i2], x[i2 + 1]);
        }
    };
    auto odd = [&]() {
        pairs(1);
        __syncthreads();  // the LDS rows may still be read by the solver
        lds[k * 2 + ln] = x[M - 1];
        lds[(P + k) * 2 + ln] = x[0];
        __syncthreads();
        // the chunk-edge pairs: to the right (my row M - 1, the next chunk's row 0) and to the left
        double2 r = (k + 1 < P) ? lds[(P + k + 1) * 2 + ln] : cmk(0.0, 0.0);
        double2 lf = (k > 0) ? lds[(k - 1) * 2 + ln] : cmk(0.0, 0.0);
        const double thr = (k + 1 < P) ? th(M - 1) : 0.0;
        const double thl = (k > 0) ? f * ld1(av, (size_t)(M - 1) * P + k - 1) : 0.0;
        double sr, cr, sl, cl;
        sincos2(thr, thl, sr, cr, sl, cl);
        rot(sr, cr, x[M - 1], r);
        rot(sl, cl, lf, x[0]);
    };
    hadamard();
    if (rev) {
        odd();
        pairs(0);
    } else {
        pairs(0);
        odd();
    }
    hadamard();
}

template <int M, int MAXT, int WPE>
__global__ void __launch_bounds__(MAXT) __attribute__((amdgpu_waves_per_eu(WPE, 8)))
k_sh_rphase_vpair(ShSim* __restrict__ sims, const int2* __restrict__ pairs, int n_pairs,
                  const double* __restrict__ cl) {
    extern __shared__ double2 lds[];  // (3 + M - 2) * P * 2: the solver's rows and alpha_i of both lines, interleaved
    const int job = xcd_remap(blockIdx.x, gridDim.x);
    if (job >= n_pairs) return;
    const int2 pt = pairs[job];
    ShSim& S = sims[pt.x];
    const int n = S.step;
    if (!S.active || n >= S.N) return;  // uniform across the block
    const int P = S.P, npad = S.npad;
    const int ln = threadIdx.x & 1, k = threadIdx.x >> 1;
    const int l = pt.y + ln;
    const double sn = ld1(S.ts, n);
    const double ll1 = (double)l * (double)(l + 1);
    double2* __restrict__ gl = sh_state(S, n + 1) + (size_t)l * npad + k;
    const double2* __restrict__ TA = S.ta + k;
    const double dr = S.delta_r, cu = S.cu;
    const int nr = S.n;
    const double a_left = (k > 0) ? ld1(S.tab, SH_AN * npad + (M - 1) * P + k - 1) : 0.0;
    const double fh = ld1(S.fs, n);
    const double fp = fh * ld1(cl, pt.y);  // the pair's angle factor (row l's combination +theta, row l + 1's -theta)
    const double* __restrict__ av = S.tab + (size_t)SH_AV * npad;

    const bool pad = k * M >= nr;  // all rows padding: zeros, neither read nor written (as in the R-phase)
    double2 x[M];
    if (pad) {
#pragma unroll
        for (int i = 0; i < M; ++i) x[i] = cmk(0.0, 0.0);
    } else {
#pragma unroll
        for (int i = 0; i < M; ++i) x[i] = ld2s(gl, (size_t)i * P);
    }
    if (fh != 0.0) sh_vpair_factors<M>(x, k, P, ln, lds, av, fp, false);

    auto diag = [&](int i, double T) {
        const int j = k * M + i;
        const double r = fma((double)j, dr, 0.5 * dr);
        // computed for padding rows too (r > 0): without the branch the two-line
        // kernel's schedule fits its registers
        const double Ur = cu * tdse_rcp(r * r);
        const double U = (j < nr) ? Ur : 0.0;
        return sn * fma(ll1, U, T);
    };
    double carry = a_left;
    auto coef = [&](int i, double& hm, double& h, double& hp) {
        const double2 tv = ld2(TA, (size_t)i * P);
        h = diag(i, tv.x);
        hm = sn * carry;
        hp = sn * tv.y;
        carry = tv.y;
    };
    double carry2 = a_left;
    auto coef2 = [&](int i, double& hm, double& h) {
        const double2* TA2 = TA;
        asm volatile("" : "+v"(TA2));
        const double2 tv = ld2(TA2, (size_t)i * P);
        h = diag(i, tv.x);
        hm = sn * carry2;
        carry2 = tv.y;
    };
    double2 xl, xr;
    chunk_neighbours<M, 2>(x, k, P, lds, xl, xr, true, ln);
    cn_implicit_lowreg<M, decltype(coef), decltype(coef2), true, 2, NoRatio, true>(x, k, P, coef, coef2, lds, xl, xr, ln);
    if (fh != 0.0) sh_vpair_factors<M>(x, k, P, ln, lds, av, fp, true);
    if (pad) return;
#pragma unroll
    for (int i = 0; i < M; ++i) st2s(gl, (size_t)i * P, x[i]);
}

// ------------------------------------------------------------------ AlternatingDirectionImplicit
//
// SphericalHarmonicMesh with AlternatingDirectionImplicit
// (evolution_methods.py:46-77): the total Hamiltonian is [H0 along r,
// E H_int along l] (mesh_operators.py:1020-1035), so one reference step is
//     g <- M . (1 + i tau H0)^-1 . (1 - i tau H_int)(1 + i tau H_int)^-1 . (1 - i tau H0) g
// with the field at t_{n+1} + dt/2 (mesh_operators.py:1008-1018). On the device
// the L-phase (with zero rotation angles) applies the mask and observes at data
// times, then three in-place sweeps of sh_state(n + 1) follow: k_sh_adi_r
// (explicit H0), k_sh_adi_l (the l-direction Cayley factor), k_sh_adi_r
// (implicit H0). Within a run, where the next step's g is not observed, the
// implicit sweep also applies the mask and the next explicit half, and the next
// replay runs only the l-direction factor and the implicit sweep: two sweeps per
// step instead of four.

// H0 along r for one l-line (the R-phase's layout and coefficients):
// EXPL: y = (1 - i H~) x; otherwise y = (1 + i H~)^-1 x. Run ahead (implicit sweep
// only, sh_adi_ahead): then also y <- (1 - i H~_{n+1}) M y, the mask that ends step n
// and the explicit half of step n + 1, written to the state buffer of step n + 2; the
// next replay skips its L-phase and explicit sweep for the simulation (S.ahead). The
// products are the ones those two sweeps would form.
template <int M, int MAXT, int WPE, bool EXPL>
__global__ void __launch_bounds__(MAXT) __attribute__((amdgpu_waves_per_eu(WPE, 8)))
k_sh_adi_r(ShSim* __restrict__ sims, const int2* __restrict__ lines, int n_lines, const int* __restrict__ ctl) {
    extern __shared__ double2 lds[];  // 3 * P
    const int line = xcd_remap(blockIdx.x, gridDim.x);
    if (line >= n_lines) return;
    const int2 lt = lines[line];
    ShSim& S = sims[lt.x];
    const int n = S.step;
    if (!S.active || n >= S.N) return;  // uniform across the block
    if (EXPL && S.ahead) return;         // applied by the previous replay's implicit sweep
    const int l = lt.y;
    const int P = S.P, npad = S.npad;
    const int k = threadIdx.x;
    const double sn = ld1(S.ts, n);
    const double ll1 = (double)l * (double)(l + 1);
    const double c0 = (l == 0 && k == 0) ? S.corr0 : 0.0;
    double2* __restrict__ gl = sh_state(S, n + 1) + (size_t)l * npad + k;
    const double2* __restrict__ TA = S.ta + k;
    const double dr = S.delta_r, cu = S.cu;
    const int nr = S.n;
    const double a_left = (k > 0) ? ld1(S.tab, SH_AN * npad + (M - 1) * P + k - 1) : 0.0;

    const bool pad = k * M >= nr;  // all rows padding: zeros, neither read nor written (as in the R-phase)
    double2 x[M];
    if (pad) {
#pragma unroll
        for (int i = 0; i < M; ++i) x[i] = cmk(0.0, 0.0);
    } else {
#pragma unroll
        for (int i = 0; i < M; ++i) x[i] = ld2s(gl, (size_t)i * P);
    }
    // s (T_j + l(l+1) U_j [+ corr0]): the diagonal of s H in tau0 units (thread kk, grid step rd, factor cc)
    auto diag = [&](int i, double T, double s, int kk, double rd, double cc) {
        const int j = kk * M + i;
        const double r = fma((double)j, rd, 0.5 * rd);
        const double U = (j < nr) ? cc * tdse_rcp(r * r) : 0.0;
        return s * (fma(ll1, U, T) + (i == 0 ? c0 : 0.0));
    };
    // x <- (1 - i s H) x with the chunk neighbours' edge rows. Opaque copies of the table
    // pointer and the grid constants keep the compiler from holding the solver's
    // first-pass coefficients (the same loads and values) live across the solve for a
    // run-ahead; the values are the same.
    auto explicit_h0 = [&](double s, bool sync_first) {
        double2 xl, xr;
        chunk_neighbours<M>(x, k, P, lds, xl, xr, sync_first);
        double2 prev = xl;
        const double2* TA3 = TA;
        double dr3 = dr, cu3 = cu;
        int k3 = k;
        asm volatile("" : "+v"(TA3), "+v"(dr3), "+v"(cu3), "+v"(k3));
        double carry = (k3 > 0) ? ld1(S.tab, SH_AN * npad + (M - 1) * P + k3 - 1) : 0.0;  // a_left, reloaded
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const double2 tv = ld2(TA3, (size_t)i * P);
            const double h = diag(i, tv.x, s, k3, dr3, cu3), hm = s * carry, hp = s * tv.y;
            carry = tv.y;
            const double2 xp = (i + 1 < M) ? x[i + 1 < M ? i + 1 : i] : xr;
            const double2 hx = cmk(fma(hm, prev.x, fma(h, x[i].x, hp * xp.x)), fma(hm, prev.y, fma(h, x[i].y, hp * xp.y)));
            prev = x[i];
            x[i] = cmk(x[i].x + hx.y, x[i].y - hx.x);  // x - i H~ x
        }
    };
    if constexpr (EXPL) {
        explicit_h0(sn, false);
    } else {
        double carry = a_left;
        auto coef = [&](int i, double& hm, double& h, double& hp) {
            const double2 tv = ld2(TA, (size_t)i * P);
            h = diag(i, tv.x, sn, k, dr, cu);
            hm = sn * carry;
            hp = sn * tv.y;
            carry = tv.y;
        };
        double carry2 = a_left;
        auto coef2 = [&](int i, double& hm, double& h) {
            const double2* TA2 = TA;
            asm volatile("" : "+v"(TA2));
            const double2 tv = ld2(TA2, (size_t)i * P);
            h = diag(i, tv.x, sn, k, dr, cu);
            hm = sn * carry2;
            carry2 = tv.y;
        };
        cn_implicit_lowreg<M, decltype(coef), decltype(coef2), false, 1>(x, k, P, coef, coef2, lds);
        if (sh_adi_ahead(S, n, *ctl)) {  // uniform across the block
            const double* mk = S.tab + (size_t)SH_MASK * npad + k;
            asm volatile("" : "+v"(mk));  // loaded after the solve, not hoisted into it
#pragma unroll
            for (int i = 0; i < M; ++i) {
                x[i] = cscale(ld1(mk, (size_t)i * P), x[i]);
                // the rounded product, as the L-phase stores it: not contracted into the next sums
                asm volatile("" : "+v"(x[i].x), "+v"(x[i].y));
            }
            explicit_h0(ld1(S.ts, n + 1), true);  // the solver's LDS rows may still be read
            gl = sh_state(S, n + 2) + (size_t)l * npad + k;
        }
    }
    if (pad) return;
#pragma unroll
    for (int i = 0; i < M; ++i) st2s(gl, (size_t)i * P, x[i]);
}

// The l-direction factor (1 - i A)(1 + i A)^-1 at every position: A real
// symmetric tridiagonal along l with zero diagonal and off-diagonals
// a_l = tau_n E (-q) r_j c_l (zero from l = L - 1 on). One lane per position
// (64 consecutive positions per wave, unit-stride rows), the Thomas recurrence
// along l on the right-hand side (1 - i A) x: the pivots of 1 + i A are real,
// den_l = 1 + a_{l-1} q_{l-1} >= 1 with q_l = a_l / den_l, so each row keeps one
// real q_l. The forward pass stores d'_l in place and q_l in the other state
// buffer (free between the L-phase and the next step's), the backward pass
// y_l = d'_l - i q_l y_{l+1} reads them back. Padding positions are left alone.
__global__ void __launch_bounds__(256) k_sh_adi_l(ShSim* __restrict__ sims, const int2* __restrict__ waves,
                                                  int n_waves, const double* __restrict__ cl) {
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / TDSE_WAVE;
    const int lane = threadIdx.x % TDSE_WAVE;
    if (wave >= n_waves) return;
    const int2 wt = waves[wave];
    ShSim& S = sims[wt.x];
    const int n = S.step;
    if (!S.active || n >= S.N) return;
    const int L = S.L, P = S.P, m = S.m, npad = S.npad;
    const int p = wt.y + lane;
    const int j = (p % P) * m + p / P;
    if (j >= S.n) return;  // no barriers below: lanes may leave
    const double f = ld1(S.fa, n) * ld1(S.tab, SH_RQ * npad + p);
    double2* __restrict__ g = sh_state(S, n + 1) + p;
    double* __restrict__ q = reinterpret_cast<double*>(sh_state(S, n)) + p;
    double2 xm = cmk(0.0, 0.0), x0 = ld2(g, 0), dprev = cmk(0.0, 0.0);
    double am = 0.0, qprev = 0.0;
    for (int l = 0; l < L; ++l) {
        const bool more = l + 1 < L;
        const double a = more ? f * ld1(cl, l) : 0.0;
        const double2 xp = more ? ld2(g, (size_t)(l + 1) * npad) : cmk(0.0, 0.0);
        // right-hand side x - i (a_{l-1} x_{l-1} + a_l x_{l+1})
        const double hx = fma(am, xm.x, a * xp.x), hy = fma(am, xm.y, a * xp.y);
        const double2 rhs = cmk(x0.x + hy, x0.y - hx);
        const double rden = tdse_rcp(fma(am, qprev, 1.0));
        // d'_l = (rhs - i a_{l-1} d'_{l-1}) / den_l
        const double2 d = cmk(fma(am, dprev.y, rhs.x) * rden, fma(-am, dprev.x, rhs.y) * rden);
        const double ql = a * rden;
        st2(g, (size_t)l * npad, d);
        q[(size_t)l * npad] = ql;
        xm = x0;
        x0 = xp;
        am = a;
        dprev = d;
        qprev = ql;
    }
    double2 y = dprev;  // y_{L-1} = d'_{L-1}, already in place
    for (int l = L - 2; l >= 0; --l) {
        const double2 d = ld2(g, (size_t)l * npad);
        const double ql = ld1(q, (size_t)l * npad);
        y = cmk(fma(ql, y.y, d.x), fma(-ql, y.x, d.y));  // d'_l - i q_l y_{l+1}
        st2(g, (size_t)l * npad, y);
    }
}

// The same l-direction factor by the partition solver, for l_bound <= 32 M:
// one workgroup of kAdiLanes x kAdiChunks threads per group of kAdiLanes
// consecutive positions; thread t works on position lane t % kAdiLanes and
// rows l = c M .. c M + M - 1 of chunk c = t / kAdiLanes (a wave reads full
// 128 B runs of each row). The explicit right-hand side (1 - i A) x is formed
// in the solver's first pass, so every element is read and written once.
constexpr int kAdiLanes = 8, kAdiChunks = 32;

template <int M>
__global__ void __launch_bounds__(kAdiLanes * kAdiChunks) k_sh_adi_lp(ShSim* __restrict__ sims,
                                                                      const int2* __restrict__ groups, int n_groups,
                                                                      const double* __restrict__ cl) {
    extern __shared__ double2 lds[];  // 3 * kAdiChunks * kAdiLanes
    constexpr int LANES = kAdiLanes;
    const int job = xcd_remap(blockIdx.x, gridDim.x);
    if (job >= n_groups) return;
    const int2 gt = groups[job];
    ShSim& S = sims[gt.x];
    const int n = S.step;
    if (!S.active || n >= S.N) return;  // uniform across the block
    const int t = threadIdx.x;
    const int ln = t % LANES, c = t / LANES;
    const int L = S.L, npad = S.npad;
    const int p = gt.y * LANES + ln;
    const double f = ld1(S.fa, n) * ld1(S.tab, SH_RQ * npad + p);  // zero at padding positions
    double2* __restrict__ g = sh_state(S, n + 1) + p;
    double2 x[M];
#pragma unroll
    for (int i = 0; i < M; ++i) {
        const int l = c * M + i;
        x[i] = (l < L) ? ld2s(g, (size_t)l * npad) : cmk(0.0, 0.0);
    }
    // rows in order: sub-diagonal a_{l-1}, zero diagonal, super-diagonal a_l (zero from l = L - 1 on)
    auto coef = [&](int i, double& hm, double& h, double& hp) {
        const int l = c * M + i;
        hm = (l >= 1 && l < L) ? f * ld1(cl, l - 1) : 0.0;
        h = 0.0;
        hp = (l + 1 < L) ? f * ld1(cl, l) : 0.0;
    };
    auto coef2 = [&](int i, double& hm, double& h) {
        const int l = c * M + i;
        hm = (l >= 1 && l < L) ? f * ld1(cl, l - 1) : 0.0;
        h = 0.0;
    };
    double2 xl, xr;
    chunk_neighbours<M, LANES>(x, c, kAdiChunks, lds, xl, xr, false, ln);
    cn_implicit_lowreg<M, decltype(coef), decltype(coef2), true, LANES>(x, c, kAdiChunks, coef, coef2, lds, xl, xr,
                                                                        ln);
#pragma unroll
    for (int i = 0; i < M; ++i) {
        const int l = c * M + i;
        if (l < L) st2s(g, (size_t)l * npad, x[i]);
    }
}

// Advance every active simulation's step counter after a full graph step.
__global__ void k_sh_advance(ShSim* sims, int n_sims, const int* __restrict__ ctl) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n_sims && sims[s].active && sims[s].step <= sims[s].N) {
        ShSim& S = sims[s];
        // the same decision as this replay's implicit ADI r-sweep
        S.ahead = (S.adi && S.step < S.N && sh_adi_ahead(S, S.step, *ctl)) ? 1 : 0;
        S.step += 1;
        S.phase = 0;
    }
}
