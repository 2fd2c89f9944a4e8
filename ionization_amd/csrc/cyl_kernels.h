// CylindricalSliceMesh, length gauge, alternating-direction implicit step on gfx950.
//
// One reference step n (evolution_methods.py:52-77 with the total Hamiltonian
// of mesh_operators.py:282-298, then the mask of meshes.py:256-257) is
//     g <- M . I_rho . E_z . I_z . E_rho . g
// with E_d = 1 - i tau_n H_d (DotOperator) and I_d = (1 + i tau_n H_d)^-1
// (TDMAOperator) along direction d. H_z and H_rho each carry the kinetic term
// of their direction (mesh_operators.py:561-600) plus half of the Coulomb
// potential and half of the length-gauge interaction -z q E(t_{n+1})
// (mesh_operators.py:244-298, 602-628); tau_n = (t_{n+1} - t_n) / (2 hbar).
//
// Two HBM sweeps per step (64 B per mesh point):
//   z-sweep   E_z I_z = 2 (1 + i tau H_z)^-1 - 1 along every z-line (the two
//             halves share one matrix, so they collapse to the Cayley form);
//   rho-sweep E_rho(n+1) . [observe] . M . I_rho(n) along every rho-line: the
//             implicit half of step n, the mask, the observables of the exact
//             g after n+1 steps, and the explicit half of step n+1.
// A simulation held at an exact state (phase 1: its last rho-sweep skipped the
// explicit half, at the end of the run or when the host asked for exact states)
// gets that half from the pre-sweep at the start of its next step.
//
// Layout of one simulation in HBM: g[rho][p], rho in [0, nrpad), p in [0, npz),
// npz = Pz * Mz. The z index of position p is jz = (p % Pz) * Mz + p / Pz, as
// for the SphericalHarmonicMesh radial layout, so the z-sweep thread k holds
// rows jz = k Mz .. k Mz + Mz - 1 of one rho-line at unit stride across the
// wave. The rho-sweep puts LANES consecutive positions (LANES z-lines) side by
// side in every wave: a wave reads 64 / LANES runs of LANES * 16 B. Padding
// (jz >= nz, rho >= nr) is zero in g and gets zero coefficients.
#pragma once

#include "common.h"
#include "line_solver.h"
#include "sh_kernels.h"

struct CylSim {
    double2* g;             // [nrpad][npz]
    const double* zt;       // [npz] z of every position (0 for padding)
    const double* rt;       // [nrpad] rho of every rho index (0 for padding)
    const double* roff;     // [nrpad] rho off-diagonal between j and j + 1: rp c(j + 1); 0 past the line
    const double* fs;       // [N] E(t_{n+1})
    const double* ts;       // [N] tau_n = (t_{n+1} - t_n) / (2 hbar)
    double* part;           // [n_groups][1 + 2 n_basis] observation partials of the rho-sweep blocks
    double* out_norm;       // [n_data]
    double2* out_ip;        // [n_data][n_basis]
    const double2* basis;   // [n_basis][nrpad][npz] test states in device layout
    double zp, rp;          // kinetic prefactors -hbar^2 / (2 mu dz^2), -hbar^2 / (2 mu drho^2)
    double kc;              // Coulomb numerator k Q q: V = kc / r
    double q;               // test charge
    double delta;           // inner-product weight dz drho
    double m_inner, m_outer, m_inv_s;  // RadialCosineMask (inner, outer, 1 / smoothness)
    double z0, dz, rho0, drho;         // grids: z_j = z0 + j dz, rho_j = rho0 + j drho
    int has_mask;
    int m_roots;  // smoothness = 2^m_roots: the root is m_roots square roots; -1 otherwise
    int nz, nr, Pz, Mz, npz, Cr, Mr, nrpad, lanes, n_groups;
    int N, store_every, n_basis;
    int step;   // steps already advanced; device-updated
    int active;
    int phase;  // 1: g is the reference g after `step` steps; 0: the explicit rho half of step `step` is applied
    int hold;   // 1: end every step at an exact state (phase 1)
    int flags;  // health flags (norm_health), device-updated at data times
};

// 1 / sqrt(x): the hardware estimate y refined by the series
// x^(-1/2) = y (1 - e)^(-1/2) = y (1 + e/2 + 3 e^2 / 8 + ...), e = 1 - x y^2,
// to third order (5 operations instead of 8 for two Newton steps; the
// truncation error ~e^3 is far below the final rounding).
__device__ __forceinline__ double cyl_rsqrt(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma(-(x * y), y, 1.0);
    return fma(y, e * fma(e, 0.375, 0.5), y);
}

// Half the Coulomb potential plus half the interaction at (z, rho):
// internal_hamiltonian adds V / 2, total_hamiltonian adds (-z q E) / 2.
__device__ __forceinline__ double cyl_half_potential(const CylSim& S, double z, double rho, double E) {
    return 0.5 * (S.kc * cyl_rsqrt(z * z + rho * rho)) + 0.5 * (-z * S.q * E);
}

// Off-diagonal of H_rho between rho indices j and j + 1: rp c(j + 1),
// c(m) = m / sqrt(m^2 - 1/4) (mesh_operators.py:582-592); 0 past the line.
// Device order of a cylindrical g (and of its basis states): rho rows in pairs
// (2 q, 2 q + 1) interleaved position by position, element (r, p) at
// ((r / 2) npz + p) 2 + r % 2 (nrpad is even). The z-sweep's threads then read
// whole lines in pairs and the rho-sweep's four positions whole 128 B lines.
__device__ __forceinline__ size_t cyl_at(int r, int p, int npz) {
    return ((size_t)(r >> 1) * npz + p) * 2 + (r & 1);
}

__device__ __forceinline__ double cyl_rho_off(const CylSim& S, int j) {
    if (j < 0 || j + 1 >= S.nr) return 0.0;
    const double m = (double)(j + 1);
    return (m * cyl_rsqrt(m * m - 0.25)) * S.rp;
}

// RadialCosineMask.__call__ (potentials/masks.py:76-89) at r.
__device__ __forceinline__ double cyl_mask_r(const CylSim& S, double r) {
    if (r < S.m_inner) return 1.0;
    if (r >= S.m_outer) return 0.0;
    double sn, cs;
    tdse_sincos(0.5 * 3.141592653589793 * (r - S.m_inner) / fabs(S.m_outer - S.m_inner), &sn, &cs);
    // |cos|^(1/s): for the usual s = 2^k (default 8) k square roots, a few
    // times cheaper than pow (1.49 -> 1.45 ms per cylinder step)
    if (S.m_roots >= 0) {
        double c = fabs(cs);
        for (int i = 0; i < S.m_roots; ++i) c = sqrt(c);
        return c;
    }
    return pow(fabs(cs), S.m_inv_s);
}

// ------------------------------------------------------------------ z-sweep
//
// One workgroup of 2 Pz threads per (simulation, pair of rho-lines 2 q, 2 q + 1);
// Mz rows per thread. Thread t works on line ln = t & 1 and chunk k = t >> 1:
// the two lines of a pair are interleaved in memory (cyl_at), so neighbouring
// threads read neighbouring 16 B and a wave reads 1 KiB contiguous.
// (1 + i tau H_z) y = (1 - i tau H_z) x with H_z tridiagonal: diagonal
// -2 zp + V/2 - z q E / 2, off-diagonals zp inside the line
// (mesh_operators.py:561-575); the launch uses the explicit form. A pair may end in a
// padding line (r = nr for odd nr): it holds zeros and stays zero.
template <int M, int MAXT, int WPE, int FORM>
__global__ void __launch_bounds__(MAXT) __attribute__((amdgpu_waves_per_eu(WPE, 8)))
k_cyl_z(CylSim* __restrict__ sims, const int2* __restrict__ pairs, int n_pairs) {
    extern __shared__ double2 lds[];  // 3 * Pz * 2, then M * Pz * 2 for CN_CAYLEY_LDS
    const int job = blockIdx.x;
    if (job >= n_pairs) return;
    const int2 lt = pairs[job];
    CylSim& S = sims[lt.x];
    const int n = S.step;
    if (!S.active || n >= S.N) return;  // uniform across the block
    const int P = S.Pz, nz = S.nz;
    const int ln = threadIdx.x & 1, k = threadIdx.x >> 1;
    const double tau = ld1(S.ts, n);
    const double E = ld1(S.fs, n);
    const double rho = fma((double)(2 * lt.y + ln), S.drho, S.rho0);
    const double tzp = tau * S.zp;
    const double tdiag = tau * (S.zp * -2.0);
    double2* __restrict__ gl = S.g + cyl_at(2 * lt.y + ln, k, S.npz);
    const double z0 = S.z0, dz = S.dz;

    double2 x[M];
#pragma unroll
    for (int i = 0; i < M; ++i) x[i] = ld2(gl, (size_t)i * P * 2);
    // coefficients from the grid, no table reads
    auto coef = [&](int i, double& hm, double& h, double& hp) {
        const int j = k * M + i;
        const bool live = j < nz;
        const double z = fma((double)j, dz, z0);
        h = live ? fma(tau, cyl_half_potential(S, z, rho, E), tdiag) : 0.0;
        hm = (live && j >= 1) ? tzp : 0.0;
        hp = (j + 1 < nz) ? tzp : 0.0;
    };
    auto coef2 = [&](int i, double& hm, double& h) {
        double hp;
        coef(i, hm, h, hp);
    };
    // equal off-diagonals inside the line: hm / hp = 1 from row 1 to nz - 2, so
    // pass 2 reuses the stored pivots; row nz - 1 (hp = 0) is rebuilt
    auto ratio = [&](int i) {
        const int j = k * M + i;
        return (j >= 1 && j + 1 < nz) ? 1.0 : 0.0;
    };
    const int ib = ((nz - 1) / M == k) ? (nz - 1) % M : -1;
    cn_line_finish<M, FORM, decltype(coef), decltype(coef2), decltype(ratio), 2>(x, gl, k, P, coef, coef2, lds, ratio,
                                                                               ib, ln, 2 * P);
}

// ---------------------------------------------------------------- rho-sweep
//
// One workgroup of LANES * Cr threads per (simulation, group of LANES
// consecutive positions): thread t works on line ln = t % LANES (position
// p = LANES * group + ln) and chunk c = t / LANES (rho rows c Mr .. c Mr + Mr - 1).
// H_rho: diagonal -2 rp + V/2 - z q E / 2, off-diagonals rp c(j + 1)
// (mesh_operators.py:578-600).
//
// MAIN (every step n < N):  implicit half of step n, mask, observables of the
//     exact g after n + 1 steps at data times, explicit half of step n + 1
//     unless n + 1 == N or the simulation holds exact states.
// PRE (phase 1 only): observables of g at step 0, explicit half of step n.
enum { CYL_PRE = 0, CYL_MAIN = 1 };

template <int LANES>
__device__ __forceinline__ double cyl_block_sum(double v, double* red, int t, int nthreads) {
    v = wave_sum(v);
    __syncthreads();
    if ((t & (TDSE_WAVE - 1)) == 0) red[t / TDSE_WAVE] = v;
    __syncthreads();
    double s = 0.0;
    const int nw = (nthreads + TDSE_WAVE - 1) / TDSE_WAVE;
    for (int w = 0; w < nw; ++w) s += red[w];  // fixed order
    return s;
}

template <int M, int LANES, int MODE, int MAXT, int WPE>
__global__ void __launch_bounds__(MAXT) __attribute__((amdgpu_waves_per_eu(WPE, 8))) k_cyl_rho(CylSim* __restrict__ sims, const int2* __restrict__ groups,
                                                  int n_groups) {
    extern __shared__ double2 lds[];  // 3 * Cr * LANES double2, then 16 doubles for reductions
    if ((int)blockIdx.x >= n_groups) return;
    // neighbouring position groups share 128 B lines: keep them on one XCD
    const int2 gt = groups[xcd_remap(blockIdx.x, n_groups)];
    CylSim& S = sims[gt.x];
    const int n = S.step;
    if (!S.active || n >= S.N) return;
    if (MODE == CYL_PRE && S.phase != 1) return;
    const int t = threadIdx.x;
    const int ln = t % LANES, c = t / LANES;
    const int P = S.Cr, nr = S.nr, npz = S.npz;
    const int p = gt.y * LANES + ln;
    const int jz = (p % S.Pz) * S.Mz + p / S.Pz;
    const bool line_live = jz < S.nz;
    const double z = ld1(S.zt, p);
    // rows c M .. c M + M - 1 (M even): row c M + i at gl[cyl_at(i, 0, npz)]; with the two
    // rows of a pair adjacent, LANES = 4 neighbouring positions read whole 128 B lines
    double2* __restrict__ gl = S.g + cyl_at(c * M, p, npz);
    double* red = reinterpret_cast<double*>(lds + 3 * P * LANES);

    double2 x[M];
#pragma unroll
    for (int i = 0; i < M; ++i) x[i] = ld2(gl, cyl_at(i, 0, npz));

    // rho-direction coefficients of row i (rho index j = c M + i) for (tau, E),
    // from the grid: rho_j recomputed, the sub-diagonal carried from the
    // previous row (callers visit rows in order, starting from `carry` =
    // cyl_rho_off(j0 - 1)).
    const double a_left = line_live ? cyl_rho_off(S, c * M - 1) : 0.0;
    // With few rows per thread, half the Coulomb potential (0.5 k Q q / r) and r
    // (for the mask) are computed once per row, from one refined reciprocal
    // square root, and kept; with more rows they are recomputed where needed.
    constexpr bool KEEP = M <= 4;
    double vhalf[KEEP ? M : 1], rad[KEEP ? M : 1];
    auto rho_of = [&](int i) { return fma((double)(c * M + i), S.drho, S.rho0); };
    if constexpr (KEEP) {
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const int j = c * M + i;
            const double rho = rho_of(i);
            const double r2 = z * z + rho * rho;
            const double ri = cyl_rsqrt(r2);
            vhalf[i] = (line_live && j < nr) ? 0.5 * (S.kc * ri) : 0.0;
            rad[i] = r2 * ri;
        }
    }
    auto half_v = [&](int i) {
        if constexpr (KEEP) return vhalf[i < M ? i : 0];
        const double rho = rho_of(i);
        return 0.5 * (S.kc * cyl_rsqrt(z * z + rho * rho));
    };
    auto radius = [&](int i) {
        if constexpr (KEEP) return rad[i < M ? i : 0];
        const double rho = rho_of(i);
        return sqrt(z * z + rho * rho);
    };
    auto rho_coef = [&](double& carry, double tau, double E, int i, double& hm, double& h, double& hp) {
        const int j = c * M + i;
        const bool live = line_live && j < nr;
        const double off = live ? cyl_rho_off(S, j) : 0.0;
        h = live ? fma(tau, half_v(i) + 0.5 * (-z * S.q * E), tau * (S.rp * -2.0)) : 0.0;
        hm = (live && j >= 1) ? tau * carry : 0.0;
        hp = tau * off;
        carry = off;
    };

    if constexpr (MODE == CYL_MAIN) {
        const double tau = ld1(S.ts, n), E = ld1(S.fs, n);
        double carry1 = a_left, carry3 = a_left;
        auto coef = [&](int i, double& hm, double& h, double& hp) { rho_coef(carry1, tau, E, i, hm, h, hp); };
        // row i's coefficients on their own (the one row whose pass-2 pivot is rebuilt)
        auto coef2 = [&](int i, double& hm, double& h) {
            double hp, left = cyl_rho_off(S, c * M + i - 1);
            rho_coef(left, tau, E, i, hm, h, hp);
        };
        // pass-2 pivot ratios hm / hp = c(j) / c(j + 1) (tau cancels), rows in order
        auto ratio = [&](int i) {
            const int j = c * M + i;
            const bool live = line_live && j < nr;
            const double off = live ? cyl_rho_off(S, j) : 0.0;
            const double q = (live && j >= 1 && off != 0.0) ? carry3 * tdse_rcp(off) : 0.0;
            carry3 = off;
            return q;
        };
        const int ib = (line_live && (nr - 1) / M == c) ? (nr - 1) % M : -1;
        cn_implicit_lowreg<M, decltype(coef), decltype(coef2), false, LANES, decltype(ratio)>(
            x, c, P, coef, coef2, lds, cmk(0.0, 0.0), cmk(0.0, 0.0), ln, ratio, ib);
        if (S.has_mask) {
            // inside the inner radius (most points) the mask is 1: decided on r^2, without the root
            const double in2 = S.m_inner * S.m_inner;
#pragma unroll
            for (int i = 0; i < M; ++i) {
                const int j = c * M + i;
                double mk = 0.0;
                if (line_live && j < nr) {
                    const double rho = rho_of(i);
                    mk = (!KEEP && z * z + rho * rho < in2) ? 1.0 : cyl_mask_r(S, radius(i));
                }
                x[i] = cscale(mk, x[i]);
            }
        }
    }

    // observables of the exact g after n + 1 (MAIN) or 0 (PRE) steps
    const int obs = (MODE == CYL_MAIN) ? n + 1 : 0;
    if ((MODE == CYL_MAIN || n == 0) && sh_is_data_step(obs, S.N, S.store_every)) {
        const int nb = S.n_basis;
        double* part = S.part + (size_t)gt.y * (1 + 2 * nb);
        double acc = 0.0;
#pragma unroll
        for (int i = 0; i < M; ++i) acc += cabs2(x[i]);
        const double tot = cyl_block_sum<LANES>(acc, red, t, blockDim.x);
        if (t == 0) part[0] = tot;
        for (int b = 0; b < nb; ++b) {
            const double2* phi = S.basis + (size_t)b * S.nrpad * npz + cyl_at(c * M, p, npz);
            double re = 0.0, im = 0.0;
#pragma unroll
            for (int i = 0; i < M; ++i) {
                const double2 f = ld2(phi, cyl_at(i, 0, npz));  // conj(f) x
                re += fma(f.x, x[i].x, f.y * x[i].y);
                im += fma(f.x, x[i].y, -f.y * x[i].x);
            }
            const double sre = cyl_block_sum<LANES>(re, red, t, blockDim.x);
            const double sim = cyl_block_sum<LANES>(im, red, t, blockDim.x);
            if (t == 0) {
                part[1 + 2 * b] = sre;
                part[2 + 2 * b] = sim;
            }
        }
    }

    // explicit half of the next step (MAIN) or of step n (PRE)
    const int ne = (MODE == CYL_MAIN) ? n + 1 : n;
    const bool expl = (MODE == CYL_PRE) || (ne < S.N && !S.hold);
    if (expl) {
        const double tau = ld1(S.ts, ne), E = ld1(S.fs, ne);
        double2 xl, xr;
        chunk_neighbours<M, LANES>(x, c, P, lds, xl, xr, true, ln);
        double2 prev = xl;
        double carry3 = a_left;
#pragma unroll
        for (int i = 0; i < M; ++i) {
            double hm, h, hp;
            rho_coef(carry3, tau, E, i, hm, h, hp);
            const double2 xp = (i + 1 < M) ? x[i + 1 < M ? i + 1 : i] : xr;
            const double2 hx = cmk(fma(hm, prev.x, fma(h, x[i].x, hp * xp.x)), fma(hm, prev.y, fma(h, x[i].y, hp * xp.y)));
            prev = x[i];
            x[i] = cmk(x[i].x + hx.y, x[i].y - hx.x);  // x - i hx
        }
    }
#pragma unroll
    for (int i = 0; i < M; ++i) st2(gl, cyl_at(i, 0, npz), x[i]);
}

// Reduce the rho-sweep partials of every simulation at a data time: job
// (sim, -1) gives the norm, (sim, b) the inner product with test state b.
// PRE: after the pre-sweep (data time 0); MAIN: after the main sweep (n + 1).
__global__ void __launch_bounds__(256) k_cyl_observe(CylSim* __restrict__ sims, const int2* __restrict__ jobs,
                                                     int n_jobs, int after_main) {
    const int job = blockIdx.x;
    if (job >= n_jobs) return;
    const int2 jb = jobs[job];
    CylSim& S = sims[jb.x];
    const int n = S.step;
    if (!S.active || n >= S.N) return;
    int obs;
    if (after_main) {
        obs = n + 1;
    } else {
        if (n != 0 || S.phase != 1) return;
        obs = 0;
    }
    if (!sh_is_data_step(obs, S.N, S.store_every)) return;
    const int slot = sh_data_slot(obs, S.N, S.store_every);
    const int stride = 1 + 2 * S.n_basis;
    const int off = jb.y < 0 ? 0 : 1 + 2 * jb.y;
    __shared__ double red_re[256];
    __shared__ double red_im[256];
    const int t = threadIdx.x;
    double re = 0.0, im = 0.0;
    for (int g = t; g < S.n_groups; g += blockDim.x) {
        re += ld1(S.part, (size_t)g * stride + off);
        if (jb.y >= 0) im += ld1(S.part, (size_t)g * stride + off + 1);
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
            const double nrm = red_re[0] * S.delta;
            const int health = norm_health(S.out_norm, slot, nrm);
            S.out_norm[slot] = nrm;
            if (health) S.flags |= health;
        } else {
            S.out_ip[(size_t)slot * S.n_basis + jb.y] = cmk(red_re[0] * S.delta, red_im[0] * S.delta);
        }
    }
}

// After a full step: advance the step counter and record whether g is exact.
__global__ void k_cyl_advance(CylSim* sims, int n_sims) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n_sims && sims[s].active && sims[s].step < sims[s].N) {
        const int next = sims[s].step + 1;
        sims[s].phase = (next < sims[s].N && !sims[s].hold) ? 0 : 1;
        sims[s].step = next;
    }
}
