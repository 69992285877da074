// rb_philox.hh -- device-side counter-based random streams for gfx950.
//
// Philox4x32-10 (the RandBLAS DefaultRNG, RandBLAS/base.hh:53) and the maps
// RandBLAS applies to its words: uneg11 / u01 and the Box-Muller transform
// (RandBLAS/random_gen.hh:60-137, Random123 uniform.hpp / boxmuller.hpp).
//
// Box-Muller needs sin, cos and log of float arguments.  They are evaluated
// in double precision and rounded once to float:
//   sin/cos(t): Cody-Waite reduction by pi/2 with a two-part constant
//     (exact for |t| <= pi_f because |k| <= 2), then Taylor polynomials of
//     degree 17 (sin) and 16 (cos) in Horner form;
//   log(u):    u = 2^e * m, m in [1/sqrt2, sqrt2), log m = 2*atanh(f/(2+f))
//     with f = m-1, series through s^21.
// Every step is an IEEE double operation with an explicit fma where fused,
// and the library is compiled with -ffp-contract=off, so results do not
// depend on the compiler's contraction choices.  The resulting floats are
// correctly rounded except where the exact value lies within ~2^-50 of a
// rounding boundary.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rbx {

constexpr uint32_t PH_M0 = 0xD2511F53u;
constexpr uint32_t PH_M1 = 0xCD9E8D57u;
constexpr uint32_t PH_W0 = 0x9E3779B9u;
constexpr uint32_t PH_W1 = 0xBB67AE85u;

// 128-bit counter held as two 64-bit halves (word 0 = low 32 bits of lo).
struct Ctr128 {
    uint64_t lo, hi;
};

__host__ __device__ inline Ctr128 ctr_from_words(const uint32_t c[4]) {
    return Ctr128{(uint64_t)c[0] | ((uint64_t)c[1] << 32), (uint64_t)c[2] | ((uint64_t)c[3] << 32)};
}

__host__ __device__ inline Ctr128 ctr_add(Ctr128 c, uint64_t n) {
    uint64_t lo = c.lo + n;
    return Ctr128{lo, c.hi + (lo < n ? 1u : 0u)};
}

__host__ __device__ inline void ctr_to_words(Ctr128 c, uint32_t w[4]) {
    w[0] = (uint32_t)c.lo;
    w[1] = (uint32_t)(c.lo >> 32);
    w[2] = (uint32_t)c.hi;
    w[3] = (uint32_t)(c.hi >> 32);
}

// Philox4x32-10 on one counter.  The key words are wave-uniform in every
// caller, so the Weyl bumps stay on the scalar unit.
__device__ __forceinline__ uint4 philox4x32_10(Ctr128 c, uint32_t k0, uint32_t k1) {
    uint32_t x0 = (uint32_t)c.lo, x1 = (uint32_t)(c.lo >> 32);
    uint32_t x2 = (uint32_t)c.hi, x3 = (uint32_t)(c.hi >> 32);
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += PH_W0;
            k1 += PH_W1;
        }
        uint64_t p0 = (uint64_t)PH_M0 * x0;
        uint64_t p1 = (uint64_t)PH_M1 * x2;
        uint32_t y0 = (uint32_t)(p1 >> 32) ^ x1 ^ k0;
        uint32_t y2 = (uint32_t)(p0 >> 32) ^ x3 ^ k1;
        x1 = (uint32_t)p1;
        x3 = (uint32_t)p0;
        x0 = y0;
        x2 = y2;
    }
    return make_uint4(x0, x1, x2, x3);
}

// (float)(int32)w * 2^-31 + 2^-32; the product is exact.
__device__ __forceinline__ float uneg11(uint32_t w) {
    return __fadd_rn(__fmul_rn((float)(int32_t)w, 0x1p-31f), 0x1p-32f);
}
// (float)w * 2^-32 + 2^-33
__device__ __forceinline__ float u01(uint32_t w) {
    return __fadd_rn(__fmul_rn((float)w, 0x1p-32f), 0x1p-33f);
}

// sin and cos of a float argument with |t| <= pi_f.
__device__ __forceinline__ void rbm_sincos(float t, float &s_out, float &c_out) {
    const double TWO_OVER_PI = 0x1.45f306dc9c883p-1;
    const double PIO2_HI = 0x1.921fb54442d18p+0;
    const double PIO2_LO = 0x1.1a62633145c07p-54;
    double x = (double)t;
    double kd = __builtin_rint(__dmul_rn(x, TWO_OVER_PI));
    double r = __fma_rn(-kd, PIO2_HI, x);
    r = __fma_rn(-kd, PIO2_LO, r);
    double r2 = __dmul_rn(r, r);
    double ps = 0x1.952c77030ad4ap-49;
    ps = __fma_rn(ps, r2, -0x1.ae7f3e733b81fp-41);
    ps = __fma_rn(ps, r2, 0x1.6124613a86d09p-33);
    ps = __fma_rn(ps, r2, -0x1.ae64567f544e4p-26);
    ps = __fma_rn(ps, r2, 0x1.71de3a556c734p-19);
    ps = __fma_rn(ps, r2, -0x1.a01a01a01a01ap-13);
    ps = __fma_rn(ps, r2, 0x1.1111111111111p-7);
    ps = __fma_rn(ps, r2, -0x1.5555555555555p-3);
    double sv = __fma_rn(__dmul_rn(r, r2), ps, r);
    double pc = 0x1.ae7f3e733b81fp-45;
    pc = __fma_rn(pc, r2, -0x1.93974a8c07c9dp-37);
    pc = __fma_rn(pc, r2, 0x1.1eed8eff8d898p-29);
    pc = __fma_rn(pc, r2, -0x1.27e4fb7789f5cp-22);
    pc = __fma_rn(pc, r2, 0x1.a01a01a01a01ap-16);
    pc = __fma_rn(pc, r2, -0x1.6c16c16c16c17p-10);
    pc = __fma_rn(pc, r2, 0x1.5555555555555p-5);
    pc = __fma_rn(pc, r2, -0x1.0000000000000p-1);
    double cv = __fma_rn(r2, pc, 1.0);
    int q = ((int)kd) & 3;
    double S = (q & 1) ? cv : sv;
    double Cc = (q & 1) ? sv : cv;
    if (q == 1 || q == 2) Cc = -Cc;
    if (q >= 2) S = -S;
    s_out = (float)S;
    c_out = (float)Cc;
}

// log of a positive normal float.
__device__ __forceinline__ float rbm_logf(float u) {
    uint32_t bits = __float_as_uint(u);
    int e = (int)(bits >> 23) - 127;
    float mf = __uint_as_float((bits & 0x7fffffu) | 0x3f800000u);
    double m = (double)mf;
    if (m > 0x1.6a09e667f3bcdp+0) {
        m = __dmul_rn(m, 0.5);
        e += 1;
    }
    double f = __dsub_rn(m, 1.0);
    double s = __ddiv_rn(f, __dadd_rn(2.0, f));
    double s2 = __dmul_rn(s, s);
    double p = 0x1.8618618618618p-4;
    p = __fma_rn(p, s2, 0x1.af286bca1af28p-4);
    p = __fma_rn(p, s2, 0x1.e1e1e1e1e1e1ep-4);
    p = __fma_rn(p, s2, 0x1.1111111111111p-3);
    p = __fma_rn(p, s2, 0x1.3b13b13b13b14p-3);
    p = __fma_rn(p, s2, 0x1.745d1745d1746p-3);
    p = __fma_rn(p, s2, 0x1.c71c71c71c71cp-3);
    p = __fma_rn(p, s2, 0x1.2492492492492p-2);
    p = __fma_rn(p, s2, 0x1.999999999999ap-2);
    p = __fma_rn(p, s2, 0x1.5555555555555p-1);
    double lm = __fma_rn(__dmul_rn(s, s2), p, __dmul_rn(2.0, s));
    double ed = (double)e;
    double res = __fma_rn(ed, 0x1.62e42fee00000p-1, __fma_rn(ed, 0x1.a39ef35793c76p-33, lm));
    return (float)res;
}

// r123::boxmuller(u0, u1): (r*sin(pi_f*uneg11(u0)), r*cos(...)), r = sqrtf(-2 logf(u01(u1))).
__device__ __forceinline__ float2 boxmuller(uint32_t w0, uint32_t w1) {
    const float PIf = 3.1415926535897932f;
    float s, c;
    rbm_sincos(__fmul_rn(PIf, uneg11(w0)), s, c);
    // __builtin_sqrtf lowers to the correctly rounded sequence (HIP's default
    // -fhip-fp32-correctly-rounded-divide-sqrt); __fsqrt_rn would not be.
    float r = __builtin_sqrtf(__fmul_rn(-2.0f, rbm_logf(u01(w1))));
    return make_float2(__fmul_rn(s, r), __fmul_rn(c, r));
}

// One Philox call -> four operator entries (random_gen.hh:85-137).
__device__ __forceinline__ float4 gen4(bool gauss, Ctr128 c, uint32_t k0, uint32_t k1) {
    uint4 w = philox4x32_10(c, k0, k1);
    if (gauss) {
        float2 a = boxmuller(w.x, w.y);
        float2 b = boxmuller(w.z, w.w);
        return make_float4(a.x, a.y, b.x, b.y);
    }
    return make_float4(uneg11(w.x), uneg11(w.y), uneg11(w.z), uneg11(w.w));
}

// gen4 with the two Box-Muller pairs kept apart by scheduling fences, for
// kernels whose register budget cannot hold both pairs' temporaries at once.
__device__ __forceinline__ float4 gen4_lean(bool gauss, Ctr128 c, uint32_t k0, uint32_t k1) {
    uint4 w = philox4x32_10(c, k0, k1);
    if (gauss) {
        __builtin_amdgcn_sched_barrier(0);
        float2 a = boxmuller(w.x, w.y);
        __builtin_amdgcn_sched_barrier(0);
        float2 b = boxmuller(w.z, w.w);
        __builtin_amdgcn_sched_barrier(0);
        return make_float4(a.x, a.y, b.x, b.y);
    }
    return make_float4(uneg11(w.x), uneg11(w.y), uneg11(w.z), uneg11(w.w));
}

__device__ __forceinline__ float f4_get(const float4 &v, int t) {
    return t == 0 ? v.x : (t == 1 ? v.y : (t == 2 ? v.z : v.w));
}

// Entry value in precision T: Gaussians are promoted; uniforms are promoted
// and then scaled by (T)sqrt(3) as in dense_skops.hh:587-590.
template <typename T>
__device__ __forceinline__ T entry_value(float v, bool gauss);
template <>
__device__ __forceinline__ double entry_value<double>(float v, bool gauss) {
    return gauss ? (double)v : __dmul_rn((double)v, 0x1.bb67ae8584caap+0);
}
template <>
__device__ __forceinline__ float entry_value<float>(float v, bool gauss) {
    return gauss ? v : __fmul_rn(v, 0x1.bb67aep+0f);  // (float)sqrt(3.0)
}

// bfloat16 (as its 16 bits) nearest to x, ties to even; x not NaN.
__host__ __device__ __forceinline__ uint16_t bf16_rne(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    uint32_t u = __float_as_uint(x);
#else
    uint32_t u;
    __builtin_memcpy(&u, &x, 4);
#endif
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Operator entry of the BF16 operator: the FP32 entry rounded to bfloat16.
template <>
__device__ __forceinline__ uint16_t entry_value<uint16_t>(float v, bool gauss) {
    return bf16_rne(entry_value<float>(v, gauss));
}

}  // namespace rbx
