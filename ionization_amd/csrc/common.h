// Shared device helpers for the complex128 mesh kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fastmath.h"

#define TDSE_WAVE 64

// Pointers read from descriptor structs are generic (flat) to the compiler;
// these helpers access them through the global address space so the loads
// and stores compile to global_load / global_store.
typedef double tdse_d2v __attribute__((ext_vector_type(2)));
#define TDSE_GLOBAL __attribute__((address_space(1)))

__device__ __forceinline__ double2 ld2(const double2* p, size_t i) {
    const tdse_d2v v = ((const TDSE_GLOBAL tdse_d2v*)p)[i];
    return make_double2(v.x, v.y);
}
__device__ __forceinline__ void st2(double2* p, size_t i, double2 v) {
    tdse_d2v w;
    w.x = v.x;
    w.y = v.y;
    ((TDSE_GLOBAL tdse_d2v*)p)[i] = w;
}
__device__ __forceinline__ double ld1(const double* p, size_t i) { return ((const TDSE_GLOBAL double*)p)[i]; }

// Streaming forms for data touched once per sweep: the nontemporal hint marks
// the lines for early eviction instead of keeping them in L2 / MALL. On the
// 256-simulation SH scan, streaming the L-phase stores and the R-phase loads
// and stores measured 1.76 -> 1.67 ms/step; streaming the L-phase loads (whose
// halo rows are read again by the neighbouring l-segment) or any access of the
// cylinder rho-sweep measured slower.
__device__ __forceinline__ double2 ld2s(const double2* p, size_t i) {
    const tdse_d2v v = __builtin_nontemporal_load(((const TDSE_GLOBAL tdse_d2v*)p) + i);
    return make_double2(v.x, v.y);
}
__device__ __forceinline__ void st2s(double2* p, size_t i, double2 v) {
    tdse_d2v w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, ((TDSE_GLOBAL tdse_d2v*)p) + i);
}

// Workgroups are dispatched round robin over the 8 XCDs (workgroup b runs on
// XCD b % 8), each with its own L2. This bijection on [0, n) gives each XCD
// one contiguous range of logical indices, so neighbouring logical work items
// (which share cache lines) run on the same XCD at about the same time.
#define TDSE_XCDS 8
__device__ __forceinline__ int xcd_remap(int b, int n) {
    const int q = n / TDSE_XCDS, r = n % TDSE_XCDS;
    const int xcd = b % TDSE_XCDS, slot = b / TDSE_XCDS;
    return (xcd < r) ? xcd * (q + 1) + slot : r * (q + 1) + (xcd - r) * q + slot;
}

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
    const double inv = tdse_rcp(fma(z.x, z.x, z.y * z.y));
    return cmk(z.x * inv, -z.y * inv);
}
__device__ __forceinline__ double2 cdiv(double2 a, double2 b) { return cmul(a, crecip(b)); }

// Health of a simulation, judged where its norm is stored at a data time:
// TDSE_FLAG_NONFINITE for a norm that is NaN or infinite, TDSE_FLAG_NORM_GROWTH
// for one more than 0.1 % above the first or the previous stored norm (the
// checks of MeshSimulation.check, ionization/mesh/sims.py:228-240). Stored
// norms that are zero (not recorded, e.g. before a resumed state) are skipped.
enum { TDSE_HEALTH_NONFINITE = 1, TDSE_HEALTH_NORM_GROWTH = 2 };
__device__ __forceinline__ int norm_health(const double* out_norm, int slot, double nrm) {
    if (!(fabs(nrm) <= 1.7976931348623157e308)) return TDSE_HEALTH_NONFINITE;
    if (slot == 0) return 0;
    const double first = out_norm[0], prev = out_norm[slot - 1];
    return ((first > 0.0 && nrm > 1.001 * first) || (prev > 0.0 && nrm > 1.001 * prev)) ? TDSE_HEALTH_NORM_GROWTH
                                                                                          : 0;
}

// Sum over the 64 lanes of a wave; every lane gets the total.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, TDSE_WAVE);
    return v;
}
