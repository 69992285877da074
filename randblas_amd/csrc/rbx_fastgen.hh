// rbx_fastgen.hh -- the fast bfloat16 Gaussian generator shared by the BF16
// kernels (rbx_bf16.hip) and the bfloat16 operator fill (rbx_core.hip):
// hardware transcendentals, exact by construction (each value is kept only
// when it provably rounds like the portable path's, else recomputed there).
#pragma once
#include "rbx_internal.hh"

namespace rbx {

typedef float fg_f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 fg_bf16x2 __attribute__((ext_vector_type(2)));

// ---------------------------------------------------------------- fast generator
// Gaussian operator entries rounded to bfloat16 with the hardware
// transcendentals (v_sin_f32, v_cos_f32, v_log_f32, v_sqrt_f32), exact by
// construction: the portable Box-Muller path (rb_philox.hh) defines each
// entry g = RN(s * r); the fast path computes g' = RN(s' * r') from hardware
// factors whose errors are bounded over every possible input word,
//   |s' - s|, |c' - c| <= BM_ES   (absolute, the sin / cos factor of w0)
//   |r' - r| <= BM_ER * r         (relative, the radius factor of w1),
// so |g' - g| <= D + 1.5 ulp(g') with D = r' (|s'| BM_ER + BM_ES) / (1 - BM_ER)
// (two half-ulp product roundings; ulp(g) <= 2 ulp(g')).  RNE to bfloat16
// cannot differ between g and g' when no rounding midpoint lies within that
// distance of g'.  The nearest midpoint is |low16(g') - 0x8000| ulps away
// (a midpoint across a binade edge is at least 0x4000 ulps away, far beyond
// any D accepted here).  Otherwise the lane recomputes on the exact path.
// The bounds are twice the maxima measured on gfx950 over all 2^32 words:
// max |s' - s| = 2.53e-7, max |c' - c| = 2.38e-7, max |r' - r| / r =
// 1.68e-7, and r' = 0 exactly when r = 0.
constexpr float BM_ES = 5.07e-7f;   // sin / cos factor, absolute
constexpr float BM_ER = 3.36e-7f;   // radius factor, relative

// true when RNE to bfloat16 of every value within D + 1.5 ulp of g equals RNE(g):
// mid is the rounding midpoint inside g's bfloat16 interval (same sign and
// exponent as g), so g - mid = (low16(g) - 0x8000) ulp(g) exactly (Sterbenz)
// and |g - mid| - 2 ulp(g) is exact too; the nearest other midpoint is a
// whole bfloat16 ulp further.  A zero or subnormal g (never a fast value worth
// keeping) fails: |g - mid| is then below any D the generator forms.
// The test is margin > 0 with margin = (|g - mid| - 2 ulp(g)) - D: rounding
// keeps the sign of that difference, so the minimum of several margins decides
// for all of them with one comparison.
__device__ __forceinline__ float bf16_round_margin(float g, float D) {
    const uint32_t b = __float_as_uint(g);
    const float mid = __uint_as_float((b & 0xffff0000u) | 0x8000u);
    const float ulp2 = __uint_as_float(b & 0x7f800000u) * 0x1p-22f;  // 2 ulp(g)
    return (fabsf(g - mid) - ulp2) - D;
}

// Fast bf16 pair of one Box-Muller draw on words (w0, w1), packed low =
// first; `ok` is cleared when either value is too close to a rounding
// midpoint for the bound above.
// The hardware factors of a Box-Muller draw, shared by the generator and by
// rbx_bf16_fastgen_bounds (which measures them against the portable ones).
// sin / cos of pi * uneg11(w0): v_sin_f32 / v_cos_f32 take revolutions.
// x = uneg11(w0) / 2 in one fma (the product by a power of two is exact, so
// this equals the portable mul-then-add halved).
__device__ __forceinline__ void fast_sincos_factor(uint32_t w0, float &sf, float &cf) {
    const float x = __fmaf_rn((float)(int32_t)w0, 0x1p-32f, 0x1p-33f);  // revolutions: sin(2 pi x) = sin(pi u0)
    sf = __builtin_amdgcn_sinf(x);
    cf = __builtin_amdgcn_cosf(x);
}
// sqrt(-2 log u01(w1)) with v_log_f32 (log2) scaled by -2 ln 2; u01 in one
// fma (equal to the portable u01, the product being exact).
__device__ __forceinline__ float fast_radius_factor(uint32_t w1) {
    const float u = __fmaf_rn((float)w1, 0x1p-32f, 0x1p-33f);
    return __builtin_amdgcn_sqrtf(__builtin_amdgcn_logf(u) * -0x1.62e430p+0f);
}

__device__ __forceinline__ uint32_t boxmuller_bf16_fast(uint32_t w0, uint32_t w1, float &margin) {
    float sf, cf;
    fast_sincos_factor(w0, sf, cf);
    const float rf = fast_radius_factor(w1);
    const float g0 = sf * rf, g1 = cf * rf;
    // D = r' (|s'| BM_ER + BM_ES) (1 + 2^-18): the margin covers the
    // roundings of these few operations
    const float kr = rf * (BM_ER * (1.0f + 0x1p-18f)), ks = rf * (BM_ES * (1.0f + 0x1p-18f));
    margin = fminf(bf16_round_margin(g0, __fmaf_rn(kr, fabsf(sf), ks)),
                   bf16_round_margin(g1, __fmaf_rn(kr, fabsf(cf), ks)));
    // both values rounded by one hardware RNE conversion (v_cvt_pk_bf16_f32)
    const fg_f32x2 gg = {g0, g1};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(gg, fg_bf16x2));
}

// One Philox counter -> four operator entries as bf16 (two per 32-bit word).
// Branch-free part: the four entries of words w, and (Gaussian) whether each
// pair's fast values are provably the exact path's.
template <bool GAUSS>
__device__ __forceinline__ uint2 gen4_bf16_fast(const uint4 &w, bool &ok) {
    uint2 pk;
    ok = true;
    if (GAUSS) {
        float m0, m1;
        pk.x = boxmuller_bf16_fast(w.x, w.y, m0);
        pk.y = boxmuller_bf16_fast(w.z, w.w, m1);
        ok = fminf(m0, m1) > 0.0f;
    } else {
        pk.x = (uint32_t)entry_value<uint16_t>(uneg11(w.x), false) |
               ((uint32_t)entry_value<uint16_t>(uneg11(w.y), false) << 16);
        pk.y = (uint32_t)entry_value<uint16_t>(uneg11(w.z), false) |
               ((uint32_t)entry_value<uint16_t>(uneg11(w.w), false) << 16);
    }
    return pk;
}

// A lane whose fast values are not provably exact recomputes all four on the
// exact (portable) path: one copy of that code, taken by a wave when any of
// its lanes needs it.  (Separate checks per pair need more live registers
// than the kernels have to spare.)
template <bool GAUSS>
__device__ __forceinline__ uint2 gen4_bf16(Ctr128 c, uint32_t k0, uint32_t k1) {
    const uint4 w = philox4x32_10(c, k0, k1);
    bool ok;
    uint2 pk = gen4_bf16_fast<GAUSS>(w, ok);
    if (GAUSS && !ok) {
        __builtin_amdgcn_sched_barrier(0);
        const float2 a = boxmuller(w.x, w.y);
        __builtin_amdgcn_sched_barrier(0);
        const float2 b = boxmuller(w.z, w.w);
        pk.x = (uint32_t)bf16_rne(a.x) | ((uint32_t)bf16_rne(a.y) << 16);
        pk.y = (uint32_t)bf16_rne(b.x) | ((uint32_t)bf16_rne(b.y) << 16);
    }
    return pk;
}


}  // namespace rbx
