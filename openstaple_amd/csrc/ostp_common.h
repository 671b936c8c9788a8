// Shared device/host helpers of the OpenStaPLE-on-MI355X library.
//
// Field layout is the reference's structure-of-arrays (OpenAcc/struct_c_def.h:16-58):
// a vec3_soa is three planes of `sizeh` complex numbers (re,im interleaved), an
// su3_soa three rows of vec3_soa (+ one double plane K under parallel tempering),
// a gauge field su3_soa[8] indexed 2*mu+parity, phases double_soa[8].  On the
// device a complex double is a `double2` (one 128-bit load), a complex float a
// `float2` (one 64-bit load).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "ostp_phase.h"

namespace ostp {

// ------------------------------------------------------------- lattice shape
struct Lattice {
    int nd[4];          // nd0..nd3, d0 fastest
    long sizeh;         // nd0*nd1*nd2*nd3/2 half-sites per parity
    int par_temp;       // su3_soa carries the K plane
    int map[4];         // xmap, ymap, zmap, tmap (geometry.h geom_parameters)
    size_t vec3_bytes() const { return (size_t)48 * sizeh; }
    size_t vec3f_bytes() const { return (size_t)24 * sizeh; }
    size_t su3_bytes() const { return (size_t)(144 + (par_temp ? 8 : 0)) * sizeh; }
    size_t su3f_bytes() const { return (size_t)(72 + (par_temp ? 4 : 0)) * sizeh; }
    size_t dsoa_bytes() const { return (size_t)8 * sizeh; }
    size_t fsoa_bytes() const { return (size_t)4 * sizeh; }
};

// ------------------------------------------------------------ complex helpers
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
// acc + a*b
__device__ __forceinline__ double2 cfma(double2 a, double2 b, double2 acc) {
    return make_double2(fma(a.x, b.x, fma(-a.y, b.y, acc.x)), fma(a.x, b.y, fma(a.y, b.x, acc.y)));
}
// acc + conj(a)*b
__device__ __forceinline__ double2 cfmac(double2 a, double2 b, double2 acc) {
    return make_double2(fma(a.x, b.x, fma(a.y, b.y, acc.x)), fma(a.x, b.y, fma(-a.y, b.x, acc.y)));
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cfma(float2 a, float2 b, float2 acc) {
    return make_float2(fmaf(a.x, b.x, fmaf(-a.y, b.y, acc.x)), fmaf(a.x, b.y, fmaf(a.y, b.x, acc.y)));
}
__device__ __forceinline__ float2 cfmac(float2 a, float2 b, float2 acc) {
    return make_float2(fmaf(a.x, b.x, fmaf(a.y, b.y, acc.x)), fmaf(a.x, b.y, fmaf(-a.y, b.x, acc.y)));
}

// Non-temporal loads for data read exactly once per kernel (links, phases):
// they stream past the L2, which then serves the neighbour spinors that are
// re-read by adjacent sites.
typedef double ostp_dvec2 __attribute__((ext_vector_type(2)));
typedef float ostp_fvec2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 ld_nt(const double2* p) {
    const ostp_dvec2 v = __builtin_nontemporal_load(reinterpret_cast<const ostp_dvec2*>(p));
    return make_double2(v.x, v.y);
}
__device__ __forceinline__ float2 ld_nt(const float2* p) {
    const ostp_fvec2 v = __builtin_nontemporal_load(reinterpret_cast<const ostp_fvec2*>(p));
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ double ld_nt(const double* p) { return __builtin_nontemporal_load(p); }
// Non-temporal store of a result no kernel re-reads before the next kernel
// boundary (FP64 Dslash output: 78.4 vs 80.6 us per 32^4 launch).
__device__ __forceinline__ void st_nt(double2* p, const double2& v) {
    ostp_dvec2 w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, reinterpret_cast<ostp_dvec2*>(p));
}
__device__ __forceinline__ float ld_nt(const float* p) { return __builtin_nontemporal_load(p); }

// ------------------------------------------------------- deterministic reduce
// Block-wide sum with a fixed combination tree (wave shuffles, then LDS);
// result valid in thread 0.  Block size must be a multiple of 64, <= 1024.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* lds /* >= NT/64 */) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int w = 0; w < NT / 64; w++) r += lds[w];
    }
    return r;
}

// The same, with the total returned to every thread (each thread adds the
// wave sums in the same order, so all threads hold identical values).
template <int NT>
__device__ __forceinline__ double block_sum_all(double v, double* lds /* >= NT/64 */) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double r = 0.0;
#pragma unroll
    for (int w = 0; w < NT / 64; w++) r += lds[w];
    return r;
}

}  // namespace ostp
