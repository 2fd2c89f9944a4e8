// Shared device helpers for the complex128 mesh kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define TDSE_WAVE 64

// complex128 as double2 {re, im}
__device__ __forceinline__ double2 cmk(double re, double im) { return make_double2(re, im); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return cmk(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return cmk(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cneg(double2 a) { return cmk(-a.x, -a.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return cmk(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 cscale(double s, double2 a) { return cmk(s * a.x, s * a.y); }
// (i * s) * a for real s: i s (x + iy) = -s y + i s x
__device__ __forceinline__ double2 cimul(double s, double2 a) { return cmk(-s * a.y, s * a.x); }
__device__ __forceinline__ double cabs2(double2 a) { return fma(a.x, a.x, a.y * a.y); }

// 1 / z for complex z: conj(z) / |z|^2
__device__ __forceinline__ double2 crecip(double2 z) {
    double inv = 1.0 / fma(z.x, z.x, z.y * z.y);
    return cmk(z.x * inv, -z.y * inv);
}
__device__ __forceinline__ double2 cdiv(double2 a, double2 b) { return cmul(a, crecip(b)); }

// Sum over the 64 lanes of a wave; every lane gets the total.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, TDSE_WAVE);
    return v;
}
