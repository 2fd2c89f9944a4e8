// Crank-Nicolson line solve by the partition method, complex128, one chunk per thread.
//
// Solves   (1 + w H) y = (1 - w H) x   for one line, where H is tridiagonal and
// w = i tau (tau = dt / 2 hbar; complex tau gives imaginary-time relaxation).
// This is the pair DotOperator(1 - i tau H) then TDMAOperator(1 + i tau H)
// of evolution_methods.py:56-77 / 99-110; the reference solves the implicit
// half with the sequential Thomas recurrence of cy.pyx:31-48, which here is
// replaced by an algebraically equivalent parallel elimination:
//
//  * The line is cut into P chunks of m rows; the calling thread owns chunk k
//    (rows j = k m + i, i < m) in registers.
//  * Rows 0..m-2 of the chunk are eliminated downward in terms of the left
//    unknown L_k = y_{k m - 1}:  y_i = delta_i - alpha_i L_k - gamma_i y_{i+1}.
//  * A running backward pass expresses row 0 through (L_k, Y_k),
//    Y_k = y_{k m + m - 1}; the chunk's last row then gives one equation in
//    (Y_{k-1}, Y_k, Y_{k+1}).
//  * The P-unknown reduced system (every equation normalised to a unit
//    diagonal) is solved in LDS by cyclic reduction (reduced_cr) or parallel
//    cyclic reduction (reduced_pcr).
//  * Each thread back-substitutes its rows.
//
// M is the exact number of rows per chunk (a compile-time constant, 2 <= M).
// Requirements: all P threads of the line call this together
// (block-wide barriers), lds holds 3 P double2 for this line. Rows beyond the
// physical line are padding rows with H = 0 and x = 0 (they stay 0).
#pragma once

#include <type_traits>

#include "common.h"

template <bool IMAG>
__device__ __forceinline__ double2 wmul(double2 w, double2 v) {
    if constexpr (IMAG) return cimul(w.y, v);
    else return cmul(w, v);
}

// coef(i, hm, h, hp): H_{j,j-1}, H_{j,j}, H_{j,j+1} of the thread's row i.
// xl / xr: x_{km-1} and x_{km+m} (zero at the line ends).
// The reduced system  a_k Y_{k-1} + Y_k + c_k Y_{k+1} = d_k,  k < P  (one equation
// per thread, unit diagonal, a_0 = c_{P-1} = 0). Both solvers return Y_k and Y_{k-1}.
//
// Parallel cyclic reduction: ceil(log2 P) levels, every equation updated at every level.
__device__ __forceinline__ void reduced_pcr(double2 A, double2 C, double2 D, const int k, const int P,
                                            double2* __restrict__ lds, double2& Yk, double2& Ykm1) {
    double2* pa = lds;
    double2* pc = lds + P;
    double2* pd = lds + 2 * P;
    for (int s = 1; s < P; s <<= 1) {
        __syncthreads();
        pa[k] = A;
        pc[k] = C;
        pd[k] = D;
        __syncthreads();
        double2 am = cmk(0.0, 0.0), cm = cmk(0.0, 0.0), dm = cmk(0.0, 0.0);
        double2 apl = cmk(0.0, 0.0), cpl = cmk(0.0, 0.0), dpl = cmk(0.0, 0.0);
        if (k - s >= 0) {
            am = pa[k - s];
            cm = pc[k - s];
            dm = pd[k - s];
        }
        if (k + s < P) {
            apl = pa[k + s];
            cpl = pc[k + s];
            dpl = pd[k + s];
        }
        const double2 bn = csub(csub(cmk(1.0, 0.0), cmul(A, cm)), cmul(C, apl));
        const double2 an2 = cneg(cmul(A, am));
        const double2 cn2 = cneg(cmul(C, cpl));
        const double2 dn2 = csub(csub(D, cmul(A, dm)), cmul(C, dpl));
        const double2 ib = crecip(bn);
        A = cmul(an2, ib);
        C = cmul(cn2, ib);
        D = cmul(dn2, ib);
    }
    __syncthreads();
    pd[k] = D;
    __syncthreads();
    Yk = D;
    Ykm1 = (k > 0) ? pd[k - 1] : cmk(0.0, 0.0);
}

// Cyclic reduction with compacted threads: at stride s the surviving equations
// kk = 2 s (t + 1) - 1 are updated by threads t = 0, 1, ... (so whole waves drop
// out as the system shrinks), eliminating kk - s and kk + s; the back
// substitution then solves the eliminated equations s (2 t + 1) - 1 from their
// already solved neighbours. About 2 P equation updates in total instead of
// P log2 P, at the same number of barriers.
//
// LANES > 1: LANES independent lines share the block, equation k of line ln at
// LDS index k * LANES + ln (each array of 3 holds P * LANES entries); all lines
// have the same P and step through the levels together.
template <int LANES = 1>
__device__ __forceinline__ void reduced_cr(double2 A, double2 C, double2 D, const int k, const int P,
                                           double2* __restrict__ lds, double2& Yk, double2& Ykm1,
                                           const int ln = 0) {
    double2* pa = lds + ln;
    double2* pc = lds + P * LANES + ln;
    double2* pd = lds + 2 * P * LANES + ln;
    __syncthreads();
    pa[k * LANES] = A;
    pc[k * LANES] = C;
    pd[k * LANES] = D;
    int s = 1;
    for (; s < P; s <<= 1) {
        __syncthreads();
        const int kk = 2 * s * (k + 1) - 1;
        if (kk < P) {
            const int e = kk * LANES, em = (kk - s) * LANES, ep = (kk + s) * LANES;
            const double2 a0 = pa[e], c0 = pc[e], d0 = pd[e];
            const double2 am = pa[em], cm = pc[em], dm = pd[em];
            double2 apl = cmk(0.0, 0.0), cpl = cmk(0.0, 0.0), dpl = cmk(0.0, 0.0);
            if (kk + s < P) {
                apl = pa[ep];
                cpl = pc[ep];
                dpl = pd[ep];
            }
            const double2 bn = csub(csub(cmk(1.0, 0.0), cmul(a0, cm)), cmul(c0, apl));
            const double2 an2 = cneg(cmul(a0, am));
            const double2 cn2 = cneg(cmul(c0, cpl));
            const double2 dn2 = csub(csub(d0, cmul(a0, dm)), cmul(c0, dpl));
            const double2 ib = crecip(bn);
            pa[e] = cmul(an2, ib);
            pc[e] = cmul(cn2, ib);
            pd[e] = cmul(dn2, ib);
        }
    }
    // equations still standing have no neighbours in range: their d is their solution
    for (s >>= 1; s >= 1; s >>= 1) {
        __syncthreads();
        const int kk = s * (2 * k + 1) - 1;
        if (kk < P) {
            const double2 ym = (kk - s >= 0) ? pd[(kk - s) * LANES] : cmk(0.0, 0.0);
            const double2 yp = (kk + s < P) ? pd[(kk + s) * LANES] : cmk(0.0, 0.0);
            pd[kk * LANES] = csub(csub(pd[kk * LANES], cmul(pa[kk * LANES], ym)), cmul(pc[kk * LANES], yp));
        }
    }
    __syncthreads();
    Yk = pd[k * LANES];
    Ykm1 = (k > 0) ? pd[(k - 1) * LANES] : cmk(0.0, 0.0);
}

// Off-diagonal coefficients w*hm, w*hp: complex in general; for IMAG (w = i s)
// they are i times a real number and are kept as that real number.
template <bool IMAG>
struct OffDiag {
    double2 v;
    __device__ __forceinline__ static OffDiag make(double2 w, double h) {
        OffDiag o;
        o.v = cmk(w.x * h, w.y * h);
        return o;
    }
    __device__ __forceinline__ double2 mul(double2 a) const { return cmul(v, a); }
};
template <>
struct OffDiag<true> {
    double s;  // the coefficient is i * s
    __device__ __forceinline__ static OffDiag make(double2 w, double h) {
        OffDiag o;
        o.s = w.y * h;
        return o;
    }
    __device__ __forceinline__ double2 mul(double2 a) const { return cimul(s, a); }
};

// RHS_IS_X: the right-hand side is x itself (implicit solve only, no explicit
// half); xl / xr are then unused.
template <int M, bool IMAG, class Coef, bool RHS_IS_X = false, bool REDUCED_CR = true>
__device__ __forceinline__ void cn_partition(double2 (&x)[M], const int k, const int P,
                                             const double2 xl, const double2 xr, const double2 w,
                                             Coef coef, double2* __restrict__ lds, const bool sync_first) {
    using Off = OffDiag<IMAG>;
    double2 al[M - 1], ga[M - 1], de[M - 1];
    double2 alm = cmk(0.0, 0.0), gam = cmk(0.0, 0.0), dem = cmk(0.0, 0.0);  // row m-2
    double2 dlast = cmk(0.0, 0.0), diaglast = cmk(1.0, 0.0);
    Off sublast = Off::make(w, 0.0), suplast = Off::make(w, 0.0);
#pragma unroll
    for (int i = 0; i < M; ++i) {
        {
            double hm, h, hp;
            coef(i, hm, h, hp);
            double2 d = x[i];
            if constexpr (!RHS_IS_X) {
                const double2 xm = (i == 0) ? xl : x[i > 0 ? i - 1 : 0];
                const double2 xp = (i + 1 < M) ? x[i + 1 < M ? i + 1 : i] : xr;
                // explicit half: d = x - w (hm x_{j-1} + h x_j + hp x_{j+1})
                const double2 hx =
                    cmk(fma(hm, xm.x, fma(h, x[i].x, hp * xp.x)), fma(hm, xm.y, fma(h, x[i].y, hp * xp.y)));
                d = csub(x[i], wmul<IMAG>(w, hx));
            }
            // implicit half coefficients: sub = w hm, diag = 1 + w h, sup = w hp
            const Off sub = Off::make(w, hm);
            const Off sup = Off::make(w, hp);
            const double2 diag = cadd(cmk(1.0, 0.0), wmul<IMAG>(w, cmk(h, 0.0)));
            if (i < M - 1) {
                const double2 den = (i == 0) ? diag : csub(diag, sub.mul(ga[i > 0 ? i - 1 : 0]));
                const double2 r = crecip(den);
                const double2 rsub = sub.mul(r);
                if (i == 0) {
                    al[0] = rsub;
                    de[0] = cmul(r, d);
                } else {
                    al[i] = cneg(cmul(rsub, al[i - 1]));
                    de[i] = csub(cmul(r, d), cmul(rsub, de[i - 1]));
                }
                ga[i] = sup.mul(r);
                alm = al[i];
                gam = ga[i];
                dem = de[i];
            } else {
                dlast = d;
                diaglast = diag;
                sublast = sub;
                suplast = sup;
            }
        }
    }
    // running backward pass: row 0 in terms of (L_k, Y_k), starting from row m-2
    double2 ap0 = alm, gp0 = gam, dp0 = dem;
#pragma unroll
    for (int i = M - 3; i >= 0; --i) {
        ap0 = csub(al[i], cmul(ga[i], ap0));
        gp0 = cneg(cmul(ga[i], gp0));
        dp0 = csub(de[i], cmul(ga[i], dp0));
    }
    if (sync_first) __syncthreads();
    lds[k] = ap0;
    lds[P + k] = gp0;
    lds[2 * P + k] = dp0;
    __syncthreads();
    double2 na = cmk(0.0, 0.0), ng = cmk(0.0, 0.0), nd = cmk(0.0, 0.0);
    if (k + 1 < P) {
        na = lds[k + 1];
        ng = lds[P + k + 1];
        nd = lds[2 * P + k + 1];
    }
    // reduced equation of the chunk's last row: A Y_{k-1} + B Y_k + C Y_{k+1} = D
    double2 A = cneg(sublast.mul(alm));
    const double2 B = csub(csub(diaglast, sublast.mul(gam)), suplast.mul(na));
    double2 C = cneg(suplast.mul(ng));
    double2 D = csub(csub(dlast, sublast.mul(dem)), suplast.mul(nd));
    {
        const double2 ib = crecip(B);
        A = cmul(A, ib);
        C = cmul(C, ib);
        D = cmul(D, ib);
    }
    double2 Lk, Yk;
    if constexpr (REDUCED_CR) reduced_cr(A, C, D, k, P, lds, Yk, Lk);
    else reduced_pcr(A, C, D, k, P, lds, Yk, Lk);
    D = Yk;
    double2 y = D;
#pragma unroll
    for (int i = M - 1; i >= 0; --i) {
        if (i < M - 1) y = csub(csub(de[i], cmul(al[i], Lk)), cmul(ga[i], y));
        x[i] = y;
    }
}

// Real-time form through the Cayley identity
//     (1 + i H)^-1 (1 - i H) x = 2 (1 + i H)^-1 x - x,
// which needs no explicit stencil (and no neighbour exchange): solve
// (1 + i H) z = x in place with cn_partition's elimination (d = x); the
// caller forms y = 2 z - x, re-reading x from memory (a cache hit) rather than
// keeping a second copy live in registers.
template <int M, class Coef>
__device__ __forceinline__ void cn_cayley_solve(double2 (&x)[M], const int k, const int P, Coef coef,
                                                double2* __restrict__ lds) {
    cn_partition<M, true, Coef, true>(x, k, P, cmk(0.0, 0.0), cmk(0.0, 0.0), cmk(0.0, 1.0), coef, lds, false);
}

// Implicit solve (1 + i H) z = x for real H with a smaller register footprint
// than cn_partition: per row only gamma and delta are kept (4 doubles).
//
//  pass 1 (forward): pivots r_i, gamma_i = r_i sup_i, delta_i = r_i (x_i - sub_i delta_{i-1})
//          and the running alpha_i = -r_i sub_i alpha_{i-1}; the row-0 coefficients
//          through (L_k, Y_k) that the previous chunk needs are accumulated forward,
//          alpha'_0 = sum_i pi_i alpha_i, delta'_0 = sum_i pi_i delta_i,
//          gamma'_0 = pi_{m-2} gamma_{m-2}, with pi_i = prod_{k<i} (-gamma_k);
//  reduced system (one equation per chunk, unit diagonal) by cyclic reduction;
//  pass 2 (forward): the left-boundary term mu_i = alpha_i L_k = -r_i sub_i mu_{i-1},
//          pivots recomputed from the stored gamma, delta_i <- delta_i - mu_i;
//  pass 3 (backward): z_{m-1} = Y_k, z_i = delta_i - gamma_i z_{i+1}.
// coef(i, hm, h, hp) as for cn_partition; sub = i hm, diag = 1 + i h, sup = i hp.
// coef2 is called in pass 2 and must re-read the coefficients (not reuse
// pass-1 values held in registers across the reduced solve).
// EXPLICIT: the right-hand side is the explicit half (1 - i H~) x, formed
// row by row in pass 1 from x and the chunk neighbours xl = x_{km-1},
// xr = x_{km+m} (chunk_neighbours); otherwise it is x itself.
// LANES / ln: several interleaved lines per block, as for reduced_cr.
//
// Ratio: when given (not NoRatio), pass 2 takes the sub-diagonal pivot factor
// rsub_i = i hm_i / den_i from the stored super-diagonal one, g_i = i hp_i / den_i,
// as rsub_i = g_i (hm_i / hp_i), with ratio(i) = hm_i / hp_i (0 where hm_i = 0):
// a product instead of rebuilding den_i and its reciprocal. Row ib of the thread
// (the one row of the line with hp = 0 and hm != 0, or -1) is rebuilt as before.
//
// ALPHA_LDS (fewer instructions and registers; used where the two-line kernel
// would otherwise spill):
//  * pass 1 forms the pivots from the chunk's leading principal minors (no
//    division in the recurrence) and stores alpha_i, i < m - 2, in LDS after
//    the reduced system's 3 P LANES entries, at (3 P + i P + k) LANES + ln
//    (M - 2 more P LANES double2; alpha_{m-2} stays in its register);
//  * the row-0 coefficients through (L_k, Y_k) come from a backward sweep over
//    the stored alpha_i instead of the running forward sums;
//  * pass 2 disappears: pass 3 reads alpha_i back and forms
//    z_i = delta_i - alpha_i L_k - gamma_i z_{i+1} (coef2 is not called).
// Every thread reads back only what it wrote, so no barrier is needed.
struct NoRatio {};

template <int M, class Coef, class Coef2, bool EXPLICIT = false, int LANES = 1, class Ratio = NoRatio,
          bool ALPHA_LDS = false>
__device__ __forceinline__ void cn_implicit_lowreg(double2 (&x)[M], const int k, const int P, Coef coef,
                                                   Coef2 coef2, double2* __restrict__ lds,
                                                   const double2 xl = make_double2(0.0, 0.0),
                                                   const double2 xr = make_double2(0.0, 0.0), const int ln = 0,
                                                   Ratio ratio = Ratio{}, const int ib = -1) {
    double2* __restrict__ alds = lds + (3 * P + k) * LANES + ln;  // alpha_i at alds[i P LANES]
    double2 ga[M - 1], de[M - 1];
    double2 al = cmk(0.0, 0.0), pi = cmk(1.0, 0.0), sa = cmk(0.0, 0.0), sd = cmk(0.0, 0.0);
    double2 gprev = cmk(0.0, 0.0), dprev = cmk(0.0, 0.0);
    double2 dlast = cmk(0.0, 0.0), diaglast = cmk(1.0, 0.0);
    double sublast = 0.0, suplast = 0.0;
    double2 Dm1 = cmk(1.0, 0.0), Dm2 = cmk(0.0, 0.0);  // ALPHA_LDS: D_{i-1}, D_{i-2}
    double hpp = 0.0;                                  // ALPHA_LDS: hp_{i-1}
#pragma unroll
    for (int i = 0; i < M; ++i) {
        double hm, h, hp;
        coef(i, hm, h, hp);
        double2 rhs = x[i];
        if constexpr (EXPLICIT) {
            // (1 - i H~) x on row i: x_i - i (hm x_{i-1} + h x_i + hp x_{i+1})
            const double2 xm = (i == 0) ? xl : x[i > 0 ? i - 1 : 0];
            const double2 xp = (i + 1 < M) ? x[i + 1 < M ? i + 1 : i] : xr;
            const double2 hx =
                cmk(fma(hm, xm.x, fma(h, x[i].x, hp * xp.x)), fma(hm, xm.y, fma(h, x[i].y, hp * xp.y)));
            rhs = cmk(x[i].x + hx.y, x[i].y - hx.x);
        }
        if (i < M - 1) {
            double2 r;
            if constexpr (ALPHA_LDS) {
                // The pivots den_i = (1 + i h) - (i hm) gamma_{i-1} are ratios D_i / D_{i-1} of the
                // chunk's leading principal minors, D_i = (1 + i h_i) D_{i-1} + hm_i hp_{i-1} D_{i-2}
                // (D_{-1} = 1, D_{-2} = 0): the recurrence carries no division, and
                // r_i = D_{i-1} / D_i is formed off it. |D_i| >= |D_{i-1}|: every pivot has
                // real part >= 1 (the Hermitian part of 1 + i H, H real symmetric, is the
                // identity, and Schur complements keep that bound), so nothing underflows.
                const double c = hm * hpp;
                const double2 Di = cmk(fma(c, Dm2.x, fma(-h, Dm1.y, Dm1.x)), fma(c, Dm2.y, fma(h, Dm1.x, Dm1.y)));
                const double inv = tdse_rcp(fma(Di.x, Di.x, Di.y * Di.y));
                r = cmk(fma(Dm1.x, Di.x, Dm1.y * Di.y) * inv, fma(Dm1.y, Di.x, -Dm1.x * Di.y) * inv);
                Dm2 = Dm1;
                Dm1 = Di;
                hpp = hp;
            } else {
                // den = (1 + i h) - (i hm) gamma_{i-1}
                const double2 den = (i == 0) ? cmk(1.0, h) : cmk(fma(hm, gprev.y, 1.0), fma(-hm, gprev.x, h));
                r = crecip(den);
            }
            const double2 rsub = cimul(hm, r);
            al = (i == 0) ? rsub : cneg(cmul(rsub, al));
            if constexpr (ALPHA_LDS) {
                if (i < M - 2) alds[(size_t)i * P * LANES] = al;
            }
            const double2 d = (i == 0) ? cmul(r, rhs) : csub(cmul(r, rhs), cmul(rsub, dprev));
            const double2 g = cimul(hp, r);
            if constexpr (!ALPHA_LDS) {
                sa = cadd(sa, cmul(pi, al));
                sd = cadd(sd, cmul(pi, d));
                if (i < M - 2) pi = cneg(cmul(pi, g));
            }
            ga[i] = g;
            de[i] = d;
            gprev = g;
            dprev = d;
        } else {
            dlast = rhs;
            diaglast = cmk(1.0, h);
            sublast = hm;
            suplast = hp;
        }
    }
    double2 gp0;
    if constexpr (ALPHA_LDS) {
        // row 0 through (L_k, Y_k) by a backward sweep from row m-2 over the stored alpha_i
        double2 ap0 = al, dp0 = dprev;
        gp0 = gprev;
#pragma unroll
        for (int i = M - 3; i >= 0; --i) {
            int off = i * P * LANES;
            asm("" : "+v"(off));  // an opaque offset: read back from LDS, not forwarded from the store
            const double2 a = alds[off];
            ap0 = csub(a, cmul(ga[i], ap0));
            gp0 = cneg(cmul(ga[i], gp0));
            dp0 = csub(de[i], cmul(ga[i], dp0));
        }
        sa = ap0;
        sd = dp0;
    } else {
        gp0 = cmul(pi, gprev);  // pi_{m-2} gamma_{m-2}
    }
    __syncthreads();
    lds[k * LANES + ln] = sa;
    lds[(P + k) * LANES + ln] = gp0;
    lds[(2 * P + k) * LANES + ln] = sd;
    __syncthreads();
    double2 na = cmk(0.0, 0.0), ng = cmk(0.0, 0.0), nd = cmk(0.0, 0.0);
    if (k + 1 < P) {
        na = lds[(k + 1) * LANES + ln];
        ng = lds[(P + k + 1) * LANES + ln];
        nd = lds[(2 * P + k + 1) * LANES + ln];
    }
    // reduced equation of the chunk's last row: A Y_{k-1} + B Y_k + C Y_{k+1} = D
    double2 A = cneg(cimul(sublast, al));
    const double2 B = csub(csub(diaglast, cimul(sublast, gprev)), cimul(suplast, na));
    double2 C = cneg(cimul(suplast, ng));
    double2 D = csub(csub(dlast, cimul(sublast, dprev)), cimul(suplast, nd));
    {
        const double2 ib = crecip(B);
        A = cmul(A, ib);
        C = cmul(C, ib);
        D = cmul(D, ib);
    }
    double2 Lk, Yk;
    reduced_cr<LANES>(A, C, D, k, P, lds, Yk, Lk, ln);
    if constexpr (ALPHA_LDS) {
        // passes 2 and 3 in one: z_i = delta_i - alpha_i L_k - gamma_i z_{i+1}
        double2 z = Yk;
        x[M - 1] = z;
#pragma unroll
        for (int i = M - 2; i >= 0; --i) {
            const double2 a = (i == M - 2) ? al : alds[(size_t)i * P * LANES];
            z = csub(csub(de[i], cmul(a, Lk)), cmul(ga[i], z));
            x[i] = z;
        }
        return;
    }
    // pass 2: subtract alpha_i L_k
    double2 mu = cmk(0.0, 0.0);
#pragma unroll
    for (int i = 0; i < M - 1; ++i) {
        double2 rsub;
        auto direct = [&]() {
            double hm, h;
            coef2(i, hm, h);
            const double2 den = (i == 0) ? cmk(1.0, h) : cmk(fma(hm, ga[i - 1 > 0 ? i - 1 : 0].y, 1.0),
                                                              fma(-hm, ga[i - 1 > 0 ? i - 1 : 0].x, h));
            return cimul(hm, crecip(den));
        };
        if constexpr (std::is_same_v<Ratio, NoRatio>) {
            rsub = direct();
        } else {
            const double q = ratio(i);
            rsub = cmk(ga[i].x * q, ga[i].y * q);
            if (i == ib) rsub = direct();
        }
        mu = (i == 0) ? cmul(rsub, Lk) : cneg(cmul(rsub, mu));
        de[i] = csub(de[i], mu);
    }
    // pass 3: back substitution
    double2 z = Yk;
    x[M - 1] = z;
#pragma unroll
    for (int i = M - 2; i >= 0; --i) {
        z = csub(de[i], cmul(ga[i], z));
        x[i] = z;
    }
}

// Exchange the first / last rows of every chunk through LDS so each thread
// knows x_{km-1} and x_{km+m}. lds: 2 P double2.
template <int M, int LANES = 1>
__device__ __forceinline__ void chunk_neighbours(const double2 (&x)[M], const int k, const int P,
                                                 double2* __restrict__ lds, double2& xl, double2& xr,
                                                 const bool sync_first, const int ln = 0) {
    if (sync_first) __syncthreads();
    lds[k * LANES + ln] = x[0];
    lds[(P + k) * LANES + ln] = x[M - 1];
    __syncthreads();
    xl = (k > 0) ? lds[(P + k - 1) * LANES + ln] : cmk(0.0, 0.0);
    xr = (k + 1 < P) ? lds[(k + 1) * LANES + ln] : cmk(0.0, 0.0);
}
