/*
 * CPU oracle (test infrastructure only; never linked into the product).
 *
 * Thomas algorithm for one complex128 tridiagonal system, restating the
 * recurrence of ionization/cy.pyx:9-50 step for step:
 *   c'_0 = c_0 / b_0, d'_0 = d_0 / b_0                           (cy.pyx:31-32)
 *   den_i = b_i - a_i c'_{i-1}; c'_i = c_i / den_i;
 *   d'_i = (d_i - a_i d'_{i-1}) / den_i                           (cy.pyx:33-41)
 *   d'_{n-1} = (d_{n-1} - a_{n-1} d'_{n-2}) / (b_{n-1} - a_{n-1} c'_{n-2})  (cy.pyx:43)
 *   x_{n-1} = d'_{n-1}; x_i = d'_i - c'_i x_{i+1}                 (cy.pyx:45-48)
 * a is indexed so that a[i] multiplies x[i-1] (a[0] unused), c[i] multiplies
 * x[i+1] (c[n-1] unused): the same alignment cy.tdma builds from a dia matrix.
 * Divisions are C99 complex divisions, as Cython emits for complex128 with
 * cdivision enabled. n >= 2 is required (the reference leaves n < 2 undefined).
 */
#include <complex.h>
#include <stdint.h>
#include <stdlib.h>

int oracle_thomas(const double complex *a, const double complex *b, const double complex *c,
                  const double complex *d, double complex *x, int64_t n)
{
    if (n < 2) return -1;
    double complex *cp = (double complex *)malloc(sizeof(double complex) * (size_t)n);
    double complex *dp = (double complex *)malloc(sizeof(double complex) * (size_t)n);
    if (!cp || !dp) { free(cp); free(dp); return -2; }
    cp[0] = c[0] / b[0];
    dp[0] = d[0] / b[0];
    for (int64_t i = 1; i < n - 1; ++i) {
        double complex den = b[i] - a[i] * cp[i - 1];
        cp[i] = c[i] / den;
        dp[i] = (d[i] - a[i] * dp[i - 1]) / den;
    }
    dp[n - 1] = (d[n - 1] - a[n - 1] * dp[n - 2]) / (b[n - 1] - a[n - 1] * cp[n - 2]);
    x[n - 1] = dp[n - 1];
    for (int64_t i = n - 2; i >= 0; --i) x[i] = dp[i] - cp[i] * x[i + 1];
    free(cp);
    free(dp);
    return 0;
}
