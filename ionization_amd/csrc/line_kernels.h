// LineMesh Crank-Nicolson on gfx950.
//
// One reference step (AlternatingDirectionImplicit with the single total
// Hamiltonian of a LineMesh, evolution_methods.py:46-77 and
// mesh_operators.py:271-327; then the mask, meshes.py:251-257):
//     g <- M . (1 + i tau H)^-1 (1 - i tau H) g,   H = K + V(z) - z q E(t_{n+1})
// with K the 3-point kinetic stencil (diagonal -2 P_z, off-diagonal P_z),
// tau = (t_{n+1} - t_n) / 2 hbar (complex tau for imaginary-time relaxation)
// and M the mask evaluated at the signed z (meshes.py:318-320).
//
// SplitInteractionOperator on a LineMesh (evolution_methods.py:89-123): the
// split interaction factors, CN with the internal H0 = K + V, the factors in
// reverse, then the mask. Length gauge (mesh_operators.py:329-342): the factor
// is the diagonal phase exp(-i tau V_int), V_int = -z q E(t_{n+1}). Velocity
// gauge (mesh_operators.py:352-427): real rotations by one angle
// a = tau A(t) hbar (q / m) / (2 dz) on the z-pairs (k, k + 1) of even k,
// then of odd k, A the vector potential -integral of E up to t_{n+1}.
//
// One workgroup owns one simulation's line for the whole launch: every thread
// holds its chunk of m rows in registers and the kernel advances up to
// n_steps steps before writing g back, so a whole run can be one launch.
// Layout: position p = i P + k holds z-index j = k m + i (as for SH lines).
#pragma once

#include "common.h"
#include "line_solver.h"

struct LineSim {
    double2* g;            // [npad]
    const double* tab;     // [LINE_NTAB][npad]: H0 diagonal, off-diag to j-1, off-diag to j+1, z coupling, mask
    const double2* wt;     // [N] w_n = i tau_n / tau0 (complex for imaginary time)
    const double* fs;      // [N] field E(t_{n+1})
    double* out_norm;      // [n_data]
    double2* out_ip;       // [n_data][n_basis]
    const double2* basis;  // [n_basis][npad] test states' g, position order
    double delta_z;
    double off, zq0, dzq;  // tau0 P_z; z coupling of row j = zq0 + j dzq (recomputed, not read)
    double vq;             // velocity gauge: tau0 hbar (q / m) / (2 dz); the step's angle is (tau_n / tau0) A_n vq
    int n, P, m, npad, N, store_every, n_basis, renormalize;
    int method;            // LINE_ADI, LINE_SIO_LEN or LINE_SIO_VEL (real time steps only for the split operator)
    int step;
    int active;
    int flags;             // health flags (norm_health), device-updated at data times
};

enum { LN_H = 0, LN_AP = 1, LN_AN = 2, LN_ZQ = 3, LN_MASK = 4, LN_NTAB = 5 };
enum { LINE_ADI = 0, LINE_SIO_LEN = 1, LINE_SIO_VEL = 2 };

// Length-gauge split factor: x_j *= exp(-i th_j), th_j = s (zq0 + j dzq) for the
// thread's rows j = k M + i. th_j is linear in j, so the phase advances by one
// complex multiplication per row from one sine and cosine at the chunk's first
// row and one of the step th_{j+1} - th_j.
template <int M>
__device__ __forceinline__ void line_phase(double2 (&x)[M], const int k, const double s, const double zq0,
                                           const double dzq) {
    double s0, c0, sd, cd;
    tdse_sincos(s * fma((double)(k * M), dzq, zq0), &s0, &c0);
    tdse_sincos(s * dzq, &sd, &cd);
    double2 ph = cmk(c0, -s0);
    const double2 stp = cmk(cd, -sd);
#pragma unroll
    for (int i = 0; i < M; ++i) {
        x[i] = cmul(ph, x[i]);
        ph = cmul(ph, stp);
    }
}

// Velocity-gauge split factors: the rotation [[cos a, sin a], [-sin a, cos a]]
// on the z-pairs (j, j + 1), j + 1 < nz, of even j then odd j (rev: odd, then
// even). With M even the even pairs lie inside a thread; of the odd pairs the
// one across a chunk edge uses the neighbouring thread's row through LDS.
// Forward (odd pairs last), each thread's edge rotations also produce the
// neighbouring chunks' final edge rows: xl / xr, when given, receive them (what
// chunk_neighbours would read), so the CN solve needs no exchange of its own.
template <int M>
__device__ __forceinline__ void line_vel_pairs(double2 (&x)[M], const int k, const int P,
                                               double2* __restrict__ lds, const double a, const int nz,
                                               const bool rev, double2* xl = nullptr, double2* xr = nullptr) {
    double sn, cs;  // one angle for the whole line (uniform across the block)
    if (fabs(a) <= kSmallAngle) tdse_sincos_small(a, &sn, &cs);
    else tdse_sincos(a, &sn, &cs);
    auto rot = [&](double2& u, double2& v) {
        const double2 y0 = cmk(fma(cs, u.x, sn * v.x), fma(cs, u.y, sn * v.y));
        const double2 y1 = cmk(fma(cs, v.x, -sn * u.x), fma(cs, v.y, -sn * u.y));
        u = y0;
        v = y1;
    };
    auto even = [&]() {
#pragma unroll
        for (int i = 0; i < M; i += 2)
            if (k * M + i + 1 < nz) rot(x[i], x[i + 1]);
    };
    auto odd = [&]() {
#pragma unroll
        for (int i = 1; i + 1 < M; i += 2)
            if (k * M + i + 1 < nz) rot(x[i], x[i + 1]);
        __syncthreads();  // the LDS rows may still be read by the previous exchange
        lds[k] = x[M - 1];
        lds[P + k] = x[0];
        __syncthreads();
        double2 r = (k + 1 < P) ? lds[P + k + 1] : cmk(0.0, 0.0);
        if (k + 1 < P && (k + 1) * M < nz) rot(x[M - 1], r);
        double2 l = (k > 0) ? lds[k - 1] : cmk(0.0, 0.0);
        if (k > 0 && k * M < nz) rot(l, x[0]);
        if (xl) *xl = l;
        if (xr) *xr = r;
    };
    if (rev) {
        odd();
        even();
    } else {
        even();
        odd();
    }
}

__device__ __forceinline__ bool line_is_data_step(int n, int N, int every) {
    return n == 0 || n == N || (every >= 1 && (n % every) == 0);
}
__device__ __forceinline__ int line_data_slot(int n, int N, int every) {
    if (every >= 1) return (n == N && (N % every) != 0) ? N / every + 1 : n / every;
    return n == 0 ? 0 : 1;
}

// Block-wide sum in a fixed order (deterministic for a given block size).
__device__ __forceinline__ double block_sum(double v, double* red) {
    const int t = threadIdx.x;
    v = wave_sum(v);
    __syncthreads();
    if ((t % TDSE_WAVE) == 0) red[t / TDSE_WAVE] = v;
    __syncthreads();
    double tot = 0.0;
    const int nw = (blockDim.x + TDSE_WAVE - 1) / TDSE_WAVE;
    for (int w = 0; w < nw; ++w) tot += red[w];
    return tot;
}

template <int M>
__device__ __forceinline__ void line_observe(LineSim& S, const double2 (&x)[M], int k, int P, int slot,
                                             double* red) {
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < M; ++i) acc += cabs2(x[i]);
    const double nrm = block_sum(acc, red);
    if (threadIdx.x == 0) {
        const int health = norm_health(S.out_norm, slot, nrm * S.delta_z);
        S.out_norm[slot] = nrm * S.delta_z;
        if (health) S.flags |= health;
    }
    for (int b = 0; b < S.n_basis; ++b) {
        const double2* phi = S.basis + (size_t)b * S.npad + k;
        double re = 0.0, im = 0.0;
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const double2 f = ld2(phi, i * P);
            // conj(f) * x
            re += f.x * x[i].x + f.y * x[i].y;
            im += f.x * x[i].y - f.y * x[i].x;
        }
        const double sre = block_sum(re, red);
        const double sim = block_sum(im, red);
        if (threadIdx.x == 0) S.out_ip[(size_t)slot * S.n_basis + b] = cmk(sre * S.delta_z, sim * S.delta_z);
    }
}

// From 10 rows per thread the real-time solve keeps alpha_i in LDS (cn_implicit_lowreg's
// ALPHA_LDS form): with pass 2 gone the deep tiles no longer spill registers.
constexpr bool line_alpha_lds(int M) { return M >= 10; }
// LDS rows of P double2 that the solver uses (the reduction scratch follows them)
constexpr int line_lds_rows(int M, bool IMAG) { return (IMAG && line_alpha_lds(M)) ? 3 + M - 2 : 3; }

// blockDim.x == P and m == M for every simulation launched together; one block per simulation.
// IMAG: every w_n is purely imaginary (real time steps), which halves the
// coefficient arithmetic; imaginary-time relaxation uses the general path.
// METHOD: the step of every simulation of the launch (LINE_ADI, LINE_SIO_LEN,
// LINE_SIO_VEL; the split operator only with IMAG).
template <int M, int MAXT, bool IMAG, int METHOD>
__global__ void __launch_bounds__(MAXT) k_line_steps(LineSim* __restrict__ sims, const int2* __restrict__ ids,
                                                     int n_ids, int n_steps) {
    extern __shared__ double2 lds[];  // the solver's line_lds_rows(M, IMAG) * P double2, then reduction scratch
    if ((int)blockIdx.x >= n_ids) return;
    LineSim& S = sims[ids[blockIdx.x].x];
    if (!S.active) return;
    int n = S.step;
    const int N = S.N;
    if (n > N) return;
    const int P = S.P, npad = S.npad;
    const int k = threadIdx.x;
    double* red = reinterpret_cast<double*>(lds + line_lds_rows(M, IMAG) * P);
    double2* __restrict__ g = S.g + k;
    const double* __restrict__ H = S.tab + LN_H * npad + k;
    const double* __restrict__ AP = S.tab + LN_AP * npad + k;
    const double* __restrict__ AN = S.tab + LN_AN * npad + k;
    const double* __restrict__ ZQ = S.tab + LN_ZQ * npad + k;
    const double* __restrict__ MK = S.tab + LN_MASK * npad + k;

    double2 x[M];
#pragma unroll
    for (int i = 0; i < M; ++i) x[i] = ld2(g, (size_t)i * P);

    if (n == 0 && line_is_data_step(0, N, S.store_every)) line_observe<M>(S, x, k, P, 0, red);

    static_assert(IMAG || METHOD == LINE_ADI, "the split operator takes real time steps");
    for (int it = 0; it < n_steps && n < N; ++it, ++n) {
        const double2 w = ld2(S.wt, n);
        const double fs = ld1(S.fs, n);
        // split operator: the interaction factors before the CN half
        if constexpr (METHOD == LINE_SIO_LEN) line_phase<M>(x, k, w.y * fs, S.zq0, S.dzq);
        double2 xl, xr;
        if constexpr (METHOD == LINE_SIO_VEL)  // the odd pairs' edge rotations give the neighbours' edge rows
            line_vel_pairs<M>(x, k, P, lds, w.y * fs * S.vq, S.n, false, &xl, &xr);
        else
            chunk_neighbours<M>(x, k, P, lds, xl, xr, true);
        // the interaction sits in the CN diagonal only for the ADI step
        const double f = (METHOD == LINE_ADI) ? fs : 0.0;
        if constexpr (IMAG) {
            // w = i s with s real: solve with H~ = s H by the low-register solver,
            // which re-reads the (L2-resident) tables instead of holding them
            const double sw = w.y;
            // One table read per row and pass (the H0 diagonal); the off-diagonals
            // (constant inside the line) and the z coupling are recomputed. The
            // pointer is made opaque once per step: the table values are invariant
            // across steps, and without this the compiler would hoist every load
            // out of the step loop and hold M doubles per thread.
            const double* H1 = H;
            double zq0 = S.zq0, dzq = S.dzq;
            asm volatile("" : "+v"(H1), "+v"(zq0), "+v"(dzq));  // and the z coupling per row
            const double so = sw * S.off;
            const int nz = S.n;
            auto coef = [&](int i, double& hm, double& h, double& hp) {
                const int j = k * M + i;
                const double zq = (j < nz) ? fma((double)j, dzq, zq0) : 0.0;
                h = sw * fma(f, zq, ld1(H1, i * P));
                hm = (j >= 1 && j < nz) ? so : 0.0;
                hp = (j + 1 < nz) ? so : 0.0;
            };
            auto coef2 = [&](int i, double& hm, double& h) {
                const double* H2 = H;
                asm volatile("" : "+v"(H2));
                const int j = k * M + i;
                const double zq = (j < nz) ? fma((double)j, dzq, zq0) : 0.0;
                h = sw * fma(f, zq, ld1(H2, i * P));
                hm = (j >= 1 && j < nz) ? so : 0.0;
            };
            // equal off-diagonals inside the line: pass 2 reuses the stored pivots
            // (hm / hp = 1 from row 1 to nz - 2); row nz - 1 (hp = 0) is rebuilt
            auto ratio = [&](int i) {
                const int j = k * M + i;
                return (j >= 1 && j + 1 < nz) ? 1.0 : 0.0;
            };
            const int ib = ((nz - 1) / M == k) ? (nz - 1) % M : -1;
            cn_implicit_lowreg<M, decltype(coef), decltype(coef2), true, 1, decltype(ratio), line_alpha_lds(M)>(
                x, k, P, coef, coef2, lds, xl, xr, 0, ratio, ib);
            if constexpr (METHOD == LINE_SIO_LEN) line_phase<M>(x, k, sw * fs, S.zq0, S.dzq);
            if constexpr (METHOD == LINE_SIO_VEL) line_vel_pairs<M>(x, k, P, lds, sw * fs * S.vq, nz, true);
        } else {
            auto coef = [&](int i, double& hm, double& h, double& hp) {
                h = fma(f, ld1(ZQ, i * P), ld1(H, i * P));
                hm = ld1(AP, i * P);
                hp = ld1(AN, i * P);
            };
            cn_partition<M, IMAG>(x, k, P, xl, xr, w, coef, lds, true);
        }
        const double* MK1 = MK;
        asm volatile("" : "+v"(MK1));
#pragma unroll
        for (int i = 0; i < M; ++i) x[i] = cscale(ld1(MK1, i * P), x[i]);
        if (S.renormalize) {
            // imaginary-time relaxation: restore unit norm after every step
            double acc = 0.0;
#pragma unroll
            for (int i = 0; i < M; ++i) acc += cabs2(x[i]);
            const double inv = 1.0 / sqrt(block_sum(acc, red) * S.delta_z);
#pragma unroll
            for (int i = 0; i < M; ++i) x[i] = cscale(inv, x[i]);
        }
        if (line_is_data_step(n + 1, N, S.store_every))
            line_observe<M>(S, x, k, P, line_data_slot(n + 1, N, S.store_every), red);
    }
    // the store addresses from an opaque copy of the base: not the load addresses kept live
    // (in registers or spilled) across the whole step loop
    double2* gs = g;
    asm volatile("" : "+v"(gs));
#pragma unroll
    for (int i = 0; i < M; ++i) st2(gs, (size_t)i * P, x[i]);
    __syncthreads();
    if (k == 0) S.step = (n == N) ? N + 1 : n;
}
