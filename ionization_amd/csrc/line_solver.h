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
//  * The P-unknown reduced system is solved by parallel cyclic reduction in
//    LDS with every equation normalised to a unit diagonal.
//  * Each thread back-substitutes its rows.
//
// M is the exact number of rows per chunk (a compile-time constant, 2 <= M).
// Requirements: all P threads of the line call this together
// (block-wide barriers), lds holds 3 P double2 for this line. Rows beyond the
// physical line are padding rows with H = 0 and x = 0 (they stay 0).
#pragma once

#include "common.h"

template <bool IMAG>
__device__ __forceinline__ double2 wmul(double2 w, double2 v) {
    if constexpr (IMAG) return cimul(w.y, v);
    else return cmul(w, v);
}

// coef(i, hm, h, hp): H_{j,j-1}, H_{j,j}, H_{j,j+1} of the thread's row i.
// xl / xr: x_{km-1} and x_{km+m} (zero at the line ends).
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
template <int M, bool IMAG, class Coef, bool RHS_IS_X = false>
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
    pd[k] = D;  // Y_k
    __syncthreads();
    const double2 Lk = (k > 0) ? pd[k - 1] : cmk(0.0, 0.0);
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

// Exchange the first / last rows of every chunk through LDS so each thread
// knows x_{km-1} and x_{km+m}. lds: 2 P double2.
template <int M>
__device__ __forceinline__ void chunk_neighbours(const double2 (&x)[M], const int k, const int P,
                                                 double2* __restrict__ lds, double2& xl, double2& xr,
                                                 const bool sync_first) {
    if (sync_first) __syncthreads();
    lds[k] = x[0];
    lds[P + k] = x[M - 1];
    __syncthreads();
    xl = (k > 0) ? lds[P + k - 1] : cmk(0.0, 0.0);
    xr = (k + 1 < P) ? lds[k + 1] : cmk(0.0, 0.0);
}
