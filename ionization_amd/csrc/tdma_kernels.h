// Op-level Thomas solve: the device counterpart of cy.tdma (ionization/cy.pyx:9-50).
#pragma once

#include "common.h"

// One thread per line; line i occupies [off[i], off[i] + len[i]) of the
// packed arrays. a[t] multiplies x[t-1], c[t] multiplies x[t+1] (a[first] and
// c[last] are ignored). cp holds the primed super-diagonal; x holds the
// primed right-hand side during the forward sweep and the solution after the
// back substitution. Recurrence of cy.pyx:31-48 with one reciprocal of the
// pivot instead of two divisions.
__global__ void k_tdma_lines(const double2* __restrict__ a, const double2* __restrict__ b,
                             const double2* __restrict__ c, const double2* __restrict__ d,
                             double2* __restrict__ x, double2* __restrict__ cp,
                             const int64_t* __restrict__ off, const int64_t* __restrict__ len, int64_t n_lines) {
    const int64_t line = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (line >= n_lines) return;
    const int64_t o = off[line];
    const int64_t n = len[line];
    if (n < 1) return;
    double2 r = crecip(b[o]);
    double2 cprev = cmul(c[o], r);
    double2 dprev = cmul(d[o], r);
    if (n == 1) {
        x[o] = dprev;
        return;
    }
    cp[o] = cprev;
    x[o] = dprev;
    for (int64_t t = 1; t < n; ++t) {
        const double2 at = a[o + t];
        r = crecip(csub(b[o + t], cmul(at, cprev)));
        dprev = cmul(csub(d[o + t], cmul(at, dprev)), r);
        x[o + t] = dprev;
        if (t < n - 1) {
            cprev = cmul(c[o + t], r);
            cp[o + t] = cprev;
        }
    }
    double2 xn = dprev;
    for (int64_t t = n - 2; t >= 0; --t) {
        xn = csub(x[o + t], cmul(cp[o + t], xn));
        x[o + t] = xn;
    }
}
