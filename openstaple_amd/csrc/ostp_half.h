// Half-precision (FP16) storage records of the staggered operator, shared by
// the batched FP16-inner solver (batch.hip) and the learned preconditioner
// (ptc.hip): colour vectors and links stored in FP16, arithmetic in FP32.
#pragma once

#include "ostp_dslash.h"

namespace ostp {

typedef _Float16 h2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 h2f(h2 v) { return make_float2((float)v.x, (float)v.y); }
__device__ __forceinline__ h2 f2h(float x, float y) {
    h2 r;
    r.x = (_Float16)x;
    r.y = (_Float16)y;
    return r;
}
// FP16 storage records (one per half-site; each is one wide load):
//   h6: three complex numbers, 12 B (a colour vector; one dwordx3 access)
//   h8: four complex numbers, 16 B
//   h4: two complex numbers, 8 B
// A link of the FP16 copy is an h8 (row 0, then row 1 entry 0) plus an h6
// (row 1 entries 1, 2, then the phase as (cos, sin)): 28 B per link.  When
// every phase of the batch is +-1 (no background field, no imaginary
// chemical potential) the phase is carried instead by bit 14 of row 0 entry
// 0's real part (set: -1) and the second record is an h4 (row 1 entries 1, 2):
// 24 B.  Bit 14 is the top exponent bit, zero for every FP16 value of
// magnitude <= 1 (link entries), so the entry itself is kept exactly; the
// check kernel falls back to the 28 B records if that entry exceeds 1.
struct __attribute__((aligned(4))) h6 {
    h2 v[3];
};
struct __attribute__((aligned(16))) h8 {
    h2 v[4];
};
struct __attribute__((aligned(8))) h4 {
    h2 v[2];
};
// Record loads and stores through global-address-space pointers (the record
// pointers live in the system descriptor, where the compiler cannot infer the
// address space; element-wise accesses merge into one dwordx3 / dwordx4)
typedef __attribute__((address_space(1))) const h2 gch2;
typedef __attribute__((address_space(1))) h2 gh2;
__device__ __forceinline__ h6 ldg(const h6* p) {
    gch2* g = (gch2*)p;
    h6 r;
    r.v[0] = g[0]; r.v[1] = g[1]; r.v[2] = g[2];
    return r;
}
__device__ __forceinline__ h8 ldg(const h8* p) {
    gch2* g = (gch2*)p;
    h8 r;
    r.v[0] = g[0]; r.v[1] = g[1]; r.v[2] = g[2]; r.v[3] = g[3];
    return r;
}
__device__ __forceinline__ h4 ldg(const h4* p) {
    gch2* g = (gch2*)p;
    h4 r;
    r.v[0] = g[0]; r.v[1] = g[1];
    return r;
}
__device__ __forceinline__ void stg(h4* p, const h4& v) {
    gh2* g = (gh2*)p;
    g[0] = v.v[0]; g[1] = v.v[1];
}
__device__ __forceinline__ void stg(h6* p, const h6& v) {
    gh2* g = (gh2*)p;
    g[0] = v.v[0]; g[1] = v.v[1]; g[2] = v.v[2];
}
__device__ __forceinline__ void stg(h8* p, const h8& v) {
    gh2* g = (gh2*)p;
    g[0] = v.v[0]; g[1] = v.v[1]; g[2] = v.v[2]; g[3] = v.v[3];
}

// One hop in FP32 arithmetic on rows 0, 1 of the link and an input colour
// vector already multiplied by the phase: o += U v  or  o -= U^dag v.
template <bool DAG>
__device__ __forceinline__ void hop_rows(float2 a0, float2 a1, float2 a2, float2 b0, float2 b1, float2 b2, float2 v0,
                                         float2 v1, float2 v2, float2& o0, float2& o1, float2& o2) {
    const float2 x0 = csub(cmul(a1, b2), cmul(a2, b1));
    const float2 x1 = csub(cmul(a2, b0), cmul(a0, b2));
    const float2 x2 = csub(cmul(a0, b1), cmul(a1, b0));
    if (!DAG) {
        float2 r0 = cmul(a0, v0); r0 = cfma(a1, v1, r0); r0 = cfma(a2, v2, r0);
        float2 r1 = cmul(b0, v0); r1 = cfma(b1, v1, r1); r1 = cfma(b2, v2, r1);
        float2 r2 = cmul(cconj(x0), v0); r2 = cfmac(x1, v1, r2); r2 = cfmac(x2, v2, r2);
        o0 = cadd(o0, r0); o1 = cadd(o1, r1); o2 = cadd(o2, r2);
    } else {
        float2 r0 = cmul(cconj(a0), v0); r0 = cfmac(b0, v1, r0); r0 = cfma(x0, v2, r0);
        float2 r1 = cmul(cconj(a1), v0); r1 = cfmac(b1, v1, r1); r1 = cfma(x1, v2, r1);
        float2 r2 = cmul(cconj(a2), v0); r2 = cfmac(b2, v1, r2); r2 = cfma(x2, v2, r2);
        o0 = csub(o0, r0); o1 = csub(o1, r1); o2 = csub(o2, r2);
    }
}

// One hop with FP16 storage, link at iu: o += U e^{i th} v  or  o -= U^dag e^{-i th} v.
// SGN: 24 B link with the sign-carrying entry; otherwise 28 B with (cos, sin).
template <bool DAG, bool SGN>
__device__ __forceinline__ void hop_hv(const h8* __restrict__ LA, const h6* __restrict__ LB,
                                       const h4* __restrict__ LC, int iu, const h6& vin, float2& o0, float2& o1,
                                       float2& o2) {
    const h8 la = ldg(LA + iu);
    float2 a0 = h2f(la.v[0]);
    const float2 a1 = h2f(la.v[1]), a2 = h2f(la.v[2]), b0 = h2f(la.v[3]);
    float2 b1, b2, v0, v1, v2;
    if constexpr (SGN) {
        const h4 lc = ldg(LC + iu);
        b1 = h2f(lc.v[0]);
        b2 = h2f(lc.v[1]);
        const unsigned short bits = __builtin_bit_cast(unsigned short, la.v[0].x);
        const bool neg = (bits & 0x4000u) != 0;
        a0.x = (float)__builtin_bit_cast(_Float16, (unsigned short)(bits & 0xbfffu));
        const float2 i0 = h2f(vin.v[0]), i1 = h2f(vin.v[1]), i2 = h2f(vin.v[2]);
        v0 = neg ? make_float2(-i0.x, -i0.y) : i0;
        v1 = neg ? make_float2(-i1.x, -i1.y) : i1;
        v2 = neg ? make_float2(-i2.x, -i2.y) : i2;
    } else {
        const h6 lb = ldg(LB + iu);
        b1 = h2f(lb.v[0]);
        b2 = h2f(lb.v[1]);
        const float2 cs = h2f(lb.v[2]);
        const float2 ph = DAG ? make_float2(cs.x, -cs.y) : cs;
        v0 = cmul(h2f(vin.v[0]), ph);
        v1 = cmul(h2f(vin.v[1]), ph);
        v2 = cmul(h2f(vin.v[2]), ph);
    }
    hop_rows<DAG>(a0, a1, a2, b0, b1, b2, v0, v1, v2, o0, o1, o2);
}

}  // namespace ostp
