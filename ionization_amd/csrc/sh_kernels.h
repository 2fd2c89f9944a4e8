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
    double2* g;            // [L][npad]
    const double* tab;     // [SH_NTAB][npad] radial tables in position order
    const double* fs;      // [N] field times tau ratio per step: E(t_{n+1}+dt/2) * tau_n / tau0
    const double* ts;      // [N] tau_n / tau0
    double2* snap;         // [LB][npad] rows l < LB captured at data steps
    double* part;          // [2][nwaves] per-wave partial |g|^2 sums (all l, largest l)
    double* out_norm;      // [n_data]
    double* out_lnorm;     // [n_data]
    double2* out_ip;       // [n_data][n_basis]
    const double* basis;   // [n_basis][npad] radial g of each test state, position order
    const int* basis_l;    // [n_basis]
    double corr0;          // tau0 * (-2 P_r) * dr (1 + dr) / 8, added at l = 0, j = 0
    double delta_r;
    int n, L, P, m, npad, N, store_every, LB, n_basis, nwaves;
    int step;              // steps already advanced; device-updated
    int active;
    int phase;             // 0: g is mid-step (after the R-phase); 1: g is the reference g after `step` steps
};

enum { SH_T = 0, SH_U = 1, SH_AP = 2, SH_AN = 3, SH_RQ = 4, SH_MASK = 5, SH_NTAB = 6 };

// Time index n is a data time (sims.py:84-95): first, last and every store_every.
__device__ __forceinline__ bool sh_is_data_step(int n, int N, int every) {
    return n == 0 || n == N || (every >= 1 && (n % every) == 0);
}
// Slot of time index n among the stored data times.
__device__ __forceinline__ int sh_data_slot(int n, int N, int every) {
    if (every >= 1) return (n == N && (N % every) != 0) ? N / every + 1 : n / every;
    return n == 0 ? 0 : 1;
}

// exp(-i a sigma_x) applied to (x0, x1):  y0 = cos a x0 - i sin a x1,  y1 = -i sin a x0 + cos a x1
__device__ __forceinline__ void sh_rotate(double a, double2& x0, double2& x1) {
    double s, c;
    sincos(a, &s, &c);
    double2 y0 = cmk(fma(c, x0.x, s * x1.y), fma(c, x0.y, -s * x1.x));
    double2 y1 = cmk(fma(c, x1.x, s * x0.y), fma(c, x1.y, -s * x0.x));
    x0 = y0;
    x1 = y1;
}

// ------------------------------------------------------------------ L-phase
//
// One wave covers 64 consecutive positions of one simulation (wave table
// entry: sim index, first position). Each lane streams over l carrying a
// three-stage pipeline: virtual row v = l + q (q = parity of j*L) puts the
// E-pairs on even v and the O-pairs on odd v for every lane, so iteration it
// loads rows (2it, 2it+1) and finishes rows (2it-3, 2it-2).
//
// Mode per simulation, from its step counter n (N = total steps):
//   n == 0      : start only     E_0, O_0
//   0 < n < N   : end + start    O_{n-1}, E_{n-1}, mask, E_n, O_n
//   n == N      : end only       O_{N-1}, E_{N-1}, mask
// At data times the wavefunction between "end" and "start" (the reference's
// g after n steps) is observed: per-wave |g|^2 partials and the rows l < LB.
// end_only != 0 materialises the reference g after n steps (O_{n-1}, E_{n-1},
// mask) for the listed waves without starting step n and without observing.
__global__ void __launch_bounds__(256) k_sh_lphase(ShSim* __restrict__ sims, const int2* __restrict__ waves,
                                                   int n_waves, const double* __restrict__ cl, int end_only) {
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / TDSE_WAVE;
    const int lane = threadIdx.x % TDSE_WAVE;
    if (wave >= n_waves) return;
    const int2 wt = waves[wave];
    ShSim& S = sims[wt.x];
    const int n = S.step;
    const int N = S.N;
    if (!S.active || n > N) return;
    const bool do_end = n > 0 && S.phase == 0;
    const bool do_start = n < N && !end_only;
    const bool observe = !end_only && sh_is_data_step(n, N, S.store_every);
    if (!do_end && !do_start && !observe) return;

    const int L = S.L, P = S.P, m = S.m, npad = S.npad;
    const int p = wt.y + lane;
    const int j = (p % P) * m + p / P;
    const bool live = j < S.n;
    const int q = (j & 1) & (L & 1);  // parity of j*L
    const double rq = S.tab[SH_RQ * npad + p];
    const double mk = S.tab[SH_MASK * npad + p];
    const double f_end = do_end ? S.fs[n - 1] * rq : 0.0;
    const double f_start = do_start ? S.fs[n] * rq : 0.0;
    double2* __restrict__ g = S.g + p;
    const int vmax = L - 1 + q;

    double2 zero = cmk(0.0, 0.0);
    double2 carry_u = zero, carry_s1 = zero, carry_s3 = zero;
    double acc_norm = 0.0, acc_lnorm = 0.0;

    for (int it = 0; 2 * it - 3 <= vmax; ++it) {
        const int v0 = 2 * it;
        const int l0 = v0 - q, l1 = v0 + 1 - q;
        double2 u0 = (live && l0 >= 0 && l0 < L) ? g[(size_t)l0 * npad] : zero;
        double2 u1 = (live && l1 >= 0 && l1 < L) ? g[(size_t)l1 * npad] : zero;

        // stage 1: O_end on (v0-1, v0), real pair start l = v0-1-q
        double2 s1m = carry_u, s10 = u0;
        {
            const int lp = v0 - 1 - q;
            if (do_end && lp >= 0 && lp + 1 < L) sh_rotate(f_end * cl[lp], s1m, s10);
        }
        // stage 2: E_end on (v0-2, v0-1), then mask
        double2 s2a = carry_s1, s2b = s1m;
        const int le = v0 - 2 - q;
        if (do_end) {
            if (le >= 0 && le + 1 < L) sh_rotate(f_end * cl[le], s2a, s2b);
            s2a = cscale(mk, s2a);
            s2b = cscale(mk, s2b);
        }
        if (observe && live) {
            if (le >= 0 && le < L) {
                acc_norm += cabs2(s2a);
                if (le == L - 1) acc_lnorm += cabs2(s2a);
                if (le < S.LB) S.snap[(size_t)le * npad + p] = s2a;
            }
            if (le + 1 >= 0 && le + 1 < L) {
                acc_norm += cabs2(s2b);
                if (le + 1 == L - 1) acc_lnorm += cabs2(s2b);
                if (le + 1 < S.LB) S.snap[(size_t)(le + 1) * npad + p] = s2b;
            }
        }
        // stage 3: E_start on the same pair
        if (do_start && le >= 0 && le + 1 < L) sh_rotate(f_start * cl[le], s2a, s2b);
        // stage 4: O_start on (v0-3, v0-2)
        double2 o_a = carry_s3, o_b = s2a;
        const int lo = v0 - 3 - q;
        if (do_start && lo >= 0 && lo + 1 < L) sh_rotate(f_start * cl[lo], o_a, o_b);
        if (live) {
            if (lo >= 0 && lo < L) g[(size_t)lo * npad] = o_a;
            if (lo + 1 >= 0 && lo + 1 < L) g[(size_t)(lo + 1) * npad] = o_b;
        }
        carry_u = u1;
        carry_s1 = s10;
        carry_s3 = s2b;
    }

    if (observe) {
        acc_norm = wave_sum(acc_norm);
        acc_lnorm = wave_sum(acc_lnorm);
        if (lane == 0) {
            const int w_in_sim = wt.y / TDSE_WAVE;
            S.part[w_in_sim] = acc_norm;
            S.part[S.nwaves + w_in_sim] = acc_lnorm;
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
    if (jb.y < 0) {
        for (int w = t; w < S.nwaves; w += blockDim.x) {
            re += S.part[w];
            im += S.part[S.nwaves + w];
        }
    } else {
        const double* phi = S.basis + (size_t)jb.y * S.npad;
        const double2* row = S.snap + (size_t)S.basis_l[jb.y] * S.npad;
        for (int p = t; p < S.npad; p += blockDim.x) {
            const double f = phi[p];
            const double2 z = row[p];
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
            S.out_norm[slot] = red_re[0] * S.delta_r;
            S.out_lnorm[slot] = red_im[0] * S.delta_r;
        } else {
            S.out_ip[(size_t)slot * S.n_basis + jb.y] = cmk(red_re[0] * S.delta_r, red_im[0] * S.delta_r);
        }
    }
}

// ------------------------------------------------------------------ R-phase
//
// (1 + i H~) y = (1 - i H~) x along r for one l-line, H~ = tau_n H0 real
// symmetric tridiagonal: diagonal s_n (T_j + l(l+1) U_j [+ corr0 at l = 0,
// j = 0]), off-diagonals s_n A_{j-1}, s_n A_j (mesh_operators.py:889-928,
// 244-269). One workgroup of P threads per line; the solve is cn_partition.
template <int M, int MAXT>
__global__ void __launch_bounds__(MAXT) k_sh_rphase(ShSim* __restrict__ sims, const int2* __restrict__ lines,
                                                   int n_lines) {
    extern __shared__ double2 lds[];  // 3 * P
    const int line = blockIdx.x;
    if (line >= n_lines) return;
    const int2 lt = lines[line];
    ShSim& S = sims[lt.x];
    const int n = S.step;
    if (!S.active || n >= S.N) return;  // uniform across the block
    const int l = lt.y;
    const int P = S.P, m = S.m, npad = S.npad;
    const int k = threadIdx.x;
    const double sn = S.ts[n];
    const double ll1 = (double)l * (double)(l + 1);
    const double c0 = (l == 0 && k == 0) ? S.corr0 : 0.0;
    double2* __restrict__ gl = S.g + (size_t)l * npad + k;
    const double* __restrict__ T = S.tab + SH_T * npad + k;
    const double* __restrict__ U = S.tab + SH_U * npad + k;
    const double* __restrict__ AP = S.tab + SH_AP * npad + k;
    const double* __restrict__ AN = S.tab + SH_AN * npad + k;

    double2 x[M];
#pragma unroll
    for (int i = 0; i < M; ++i)
        if (i < m) x[i] = gl[(size_t)i * P];
    const double2 xlast = gl[(size_t)(m - 1) * P];
    double2 xl, xr;
    chunk_neighbours<M>(x, xlast, k, P, lds, xl, xr, false);
    auto coef = [&](int i, double& hm, double& h, double& hp) {
        h = sn * (fma(ll1, U[i * P], T[i * P]) + (i == 0 ? c0 : 0.0));
        hm = sn * AP[i * P];
        hp = sn * AN[i * P];
    };
    cn_partition<M, true>(x, m, k, P, xl, xr, cmk(0.0, 1.0), coef, lds, true);
#pragma unroll
    for (int i = 0; i < M; ++i)
        if (i < m) gl[(size_t)i * P] = x[i];
}

// Advance every active simulation's step counter after a full graph step.
__global__ void k_sh_advance(ShSim* sims, int n_sims) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n_sims && sims[s].active && sims[s].step <= sims[s].N) {
        sims[s].step += 1;
        sims[s].phase = 0;
    }
}
