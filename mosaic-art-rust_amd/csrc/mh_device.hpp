// mh_device.hpp -- device-side arithmetic shared by the gfx950 kernels.
//
// Bit-level contract: IEEE f32/f64 with the reference's operation order and
// no FMA contraction (the whole library is built with -ffp-contract=off; the
// pragma below keeps it so for anyone including this header).  f32 divide and
// sqrt are HIP's correctly rounded defaults.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mh_kernels.hpp"

#pragma clang fp contract(off)

namespace mh {

// ------------------------------------------------------- global pointers
// Pointers read out of job descriptors are generic to the compiler, which
// then emits flat loads; on gfx9 a flat load also counts against lgkmcnt, so
// every LDS wait would wait for it too.  These views pin the global address
// space (global_load / global_store).
typedef float v4f_t __attribute__((ext_vector_type(4)));
#define MH_GAS __attribute__((address_space(1)))

__device__ __forceinline__ float4 gload4(const float4* p, int64_t i) {
    const v4f_t v = ((const MH_GAS v4f_t*)p)[i];
    return make_float4(v.x, v.y, v.z, v.w);
}
template <typename T>
__device__ __forceinline__ T gload(const T* p, int64_t i) {
    return ((const MH_GAS T*)p)[i];
}
// A load at a 32-bit byte offset from a uniform base: one address register
// for any number of arrays indexed alike (the base stays scalar).
template <typename T>
__device__ __forceinline__ T gload_at(const T* p, uint32_t byte_off) {
    const MH_GAS char* base = (const MH_GAS char*)p;
    return *(const MH_GAS T*)(base + byte_off);
}
template <typename T>
__device__ __forceinline__ void gstore(T* p, int64_t i, T v) {
    ((MH_GAS T*)p)[i] = v;
}

// Lane mask of a predicate.  The builtin takes the compare's lane mask
// directly (HIP's __ballot goes through an int select and a second compare).
__device__ __forceinline__ uint64_t wave_ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// This lane's bit of a lane mask, as a condition: the mask itself becomes the
// predicate (no per-lane shift and compare).
__device__ __forceinline__ bool lane_in(uint64_t mask) {
    return __builtin_amdgcn_inverse_ballot_w64(mask);
}
// Set bits of a lane mask below this lane (v_mbcnt_lo / v_mbcnt_hi).
__device__ __forceinline__ uint32_t lanes_before(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Lane masks of one compare each (bit l set when the compare holds in lane l;
// inactive lanes read 0).  Each is a single v_cmp into a scalar register
// pair, so conditions combine with scalar mask logic instead of being
// rebuilt per lane for a ballot.  Predicates are LLVM's icmp/fcmp codes.
__device__ __forceinline__ uint64_t lanes_eq(int32_t a, int32_t b) {
    return __builtin_amdgcn_uicmp((uint32_t)a, (uint32_t)b, 32);
}
__device__ __forceinline__ uint64_t lanes_ne(int32_t a, int32_t b) {
    return __builtin_amdgcn_uicmp((uint32_t)a, (uint32_t)b, 33);
}
__device__ __forceinline__ uint64_t lanes_ult(uint32_t a, uint32_t b) {
    return __builtin_amdgcn_uicmp(a, b, 36);
}
__device__ __forceinline__ uint64_t lanes_ule(uint32_t a, uint32_t b) {
    return __builtin_amdgcn_uicmp(a, b, 37);
}
__device__ __forceinline__ uint64_t lanes_slt(int32_t a, int32_t b) {
    return __builtin_amdgcn_sicmp(a, b, 40);
}
__device__ __forceinline__ uint64_t lanes_nonneg(int32_t a) {
    return __builtin_amdgcn_sicmp(a, -1, 38);  // a > -1
}
__device__ __forceinline__ uint64_t lanes_fle(float a, float b) {
    return __builtin_amdgcn_fcmpf(a, b, 5);  // ordered <=: false for NaN, as in C
}

// A job descriptor read from memory through a generic pointer lands in
// VGPRs, and every branch on its fields then compiles as a divergent
// (exec-mask) branch.  The descriptor is uniform across the wave: copying it
// word by word through readfirstlane puts it in SGPRs.
template <typename T>
__device__ __forceinline__ T uniform_copy(const T& v) {
    static_assert(sizeof(T) % 4 == 0, "descriptor size must be a multiple of 4");
    T r;
    const int32_t* src = reinterpret_cast<const int32_t*>(&v);
    int32_t* dst = reinterpret_cast<int32_t*>(&r);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) dst[i] = __builtin_amdgcn_readfirstlane(src[i]);
    return r;
}

// Segment of a batched launch that owns block b: the j in [0, n_segs) with
// starts[j] <= b < starts[j+1] (starts ascending, starts[0] = 0, empty
// segments allowed).  Uniform per block, so the loads stay scalar.
__device__ __forceinline__ int32_t find_segment(const int32_t* __restrict__ starts, int32_t n_segs,
                                                int32_t b) {
    int32_t lo = 0, hi = n_segs;
    while (hi - lo > 1) {
        const int32_t mid = (lo + hi) >> 1;
        if (starts[mid] <= b)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// ----------------------------------------------------------------- helpers
__device__ __forceinline__ float clampf(float v, float lo, float hi) {
    // Rust f32::clamp: NaN passes through.
    if (v < lo) return lo;
    if (v > hi) return hi;
    return v;
}

// Cube root for Lab's f(t) (t in (216/24389, ~1.09]): a single-precision
// estimate from the hardware log2 / exp2 (about 2^-21 relative), then two
// Newton steps for t^3 = x in double, t <- t - (t^3 - x) * r (fused:
// t^3 - x = fma(t t, t, -x), then fma(-(t^3 - x), r, t)), with
// r ~ 1/(3 t0^2) from the single-precision reciprocal (the correction is so
// small that r's own error ends below double rounding), rounded once to f32.
// The host restatement reaches the same correctly rounded value with four
// division steps from an exponent seed; the two agree on every input the
// conversion can produce (all 2^24 colours are checked).
__device__ __forceinline__ float cbrt_f32(float xf) {
    const float e = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(xf) * (1.0f / 3.0f));
    const double x = (double)xf;
    const double r = (double)__builtin_amdgcn_rcpf(3.0f * e * e);
    double t = (double)e;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const double f = __builtin_fma(t * t, t, -x);  // t^3 - x, one rounding
        t = __builtin_fma(-f, r, t);
    }
    return (float)t;
}

// exp for f32 arguments: Cody-Waite reduction with fdlibm's ln2 split,
// degree-13 Taylor polynomial in double (Horner), exact 2^n scaling.
__device__ __forceinline__ float exp_f32(float xf) {
    if (xf != xf) return xf;
    if (xf > 88.8f) return __int_as_float(0x7f800000);
    if (xf < -104.0f) return 0.0f;
    double x = (double)xf;
    double n = floor(x * 1.4426950408889634 + 0.5);
    double r = (x - n * 6.93147180369123816490e-01) - n * 1.90821492927058770002e-10;
    double p = 1.0 / 6227020800.0;
    p = p * r + 1.0 / 479001600.0;
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    int ni = (int)n;
    double s = __longlong_as_double((long long)((uint64_t)(int64_t)(ni + 1023) << 52));
    return (float)(p * s);
}

// ------------------------------------------------------------ sRGB -> Lab
__device__ __forceinline__ float lab_f(float t) {
    const float eps = (float)(216.0 / 24389.0);
    const float kappa = (float)(841.0 / 108.0);
    const float delta = (float)(4.0 / 29.0);
    if (t > eps) return cbrt_f32(t);
    return kappa * t + delta;
}

// palette Lab::from_color(Srgb<f32>) on a u8 triple (src/lib.rs:102-107):
// lut[] is the linearised u8 ramp; IEC 61966-2-1 matrix, D65 white point,
// CIE f(t), FromColor clamp of L/a/b.
__device__ __forceinline__ void rgb8_to_lab(const float* lut, uint32_t r8, uint32_t g8, uint32_t b8,
                                            float& L, float& A, float& B) {
    float r = lut[r8], g = lut[g8], b = lut[b8];
    float x = 0.4124564f * r + 0.3575761f * g + 0.1804375f * b;
    float y = 0.2126729f * r + 0.7151522f * g + 0.0721750f * b;
    float z = 0.0193339f * r + 0.1191920f * g + 0.9503041f * b;
    // x / 0.95047f and z / 1.08883f as products with the rounded reciprocal,
    // corrected once by the fused residual: equal to the divisions for the
    // x and z of all 2^24 u8 colours (an exhaustive host check); y / 1 is y
    const float rx = 1.0f / 0.95047f, rz = 1.0f / 1.08883f;
    const float qx = x * rx, qz = z * rz;
    x = __builtin_fmaf(__builtin_fmaf(-qx, 0.95047f, x), rx, qx);
    z = __builtin_fmaf(__builtin_fmaf(-qz, 1.08883f, z), rz, qz);
    float fx = lab_f(x), fy = lab_f(y), fz = lab_f(z);
    // x, y, z >= +0 (non-negative ramp and coefficients), so f(.) is finite
    // and positive and none of the three values below is NaN or -0 (a zero
    // difference or sum rounds to +0): each clamp is a median
    L = __builtin_amdgcn_fmed3f(fy * 116.0f - 16.0f, 0.0f, 100.0f);
    A = __builtin_amdgcn_fmed3f((fx - fy) * 500.0f, -128.0f, 127.0f);
    B = __builtin_amdgcn_fmed3f((fy - fz) * 200.0f, -128.0f, 127.0f);
}

// ------------------------------------------------------------- distances
// calculate_lab_distance (src/similarity.rs:141-146).  (x-y)^2 == (y-x)^2
// exactly, so dist(i,j) and dist(j,i) have the same bits.
__device__ __forceinline__ float lab_distance(float4 p, float4 q) {
    float dl = p.x - q.x, da = p.y - q.y, db = p.z - q.z;
    return sqrtf(dl * dl + da * da + db * db);
}

// kiddo SquaredEuclidean: ((0 + dl^2) + da^2) + db^2; 0 + x is exact.
__device__ __forceinline__ float sq_distance(float ql, float qa, float qb, float ml, float ma,
                                             float mb) {
    float d0 = ql - ml, d1 = qa - ma, d2 = qb - mb;
    return d0 * d0 + d1 * d1 + d2 * d2;
}

// One adjacency term 1/(sim+1) (src/adjacency.rs:82, :110, :174).
// get_similarity(i,i) = 0.0, which the distance formula also gives.
__device__ __forceinline__ float adj_term(float4 p, float4 q) {
    return 1.0f / (lab_distance(p, q) + 1.0f);
}

// ----------------------------------------------------------- counter RNG
// SplitMix64 finaliser (Steele, Lea, Flood 2014) of
// seed + golden * (8*it + k + 1); high 32 bits.
__device__ __forceinline__ uint32_t draw(uint64_t seed, uint64_t it, uint32_t k) {
    uint64_t z = seed + 0x9E3779B97F4A7C15ULL * (it * 8u + k + 1u);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z = z ^ (z >> 31);
    return (uint32_t)(z >> 32);
}

__device__ __forceinline__ int32_t draw_range(uint64_t seed, uint64_t it, uint32_t k,
                                              int32_t bound) {
    return (int32_t)(((uint64_t)draw(seed, it, k) * (uint64_t)(uint32_t)bound) >> 32);
}

__device__ __forceinline__ float draw_unit(uint64_t seed, uint64_t it) {
    return (float)(draw(seed, it, 4) >> 8) * (1.0f / 16777216.0f);
}

// ------------------------------------------------------------ HSV helpers
// clampf(v, 0, 1) for a v that is neither NaN nor -0: the median of
// (v, 0, 1) is then the same value (+0 below, 1 above, v between).
__device__ __forceinline__ float clamp_unit(float v) { return __builtin_amdgcn_fmed3f(v, 0.0f, 1.0f); }

// palette Hsv::from_color(Srgb<f32>) (src/color_adjustment.rs:70, :122).
// kUnit: r, g, b lie in [0, 1] and none is -0 (see adjust_pixel), so the
// max(., 0) are identities and s, v are clamps of non-negative values.
template <bool kUnit = false>
__device__ __forceinline__ void rgb_to_hsv(float r, float g, float b, float& h, float& s,
                                           float& v) {
    if (!kUnit) {
        r = r > 0.0f ? r : 0.0f;
        g = g > 0.0f ? g : 0.0f;
        b = b > 0.0f ? b : 0.0f;
    }
    float mx, mn, sep, coeff;
    if (kUnit) {
        // the same choices as the branches below, as selects
        const bool c1 = r > g;
        const float mrg = c1 ? r : g, nrg = c1 ? g : r;
        const bool c2 = b > mrg;
        const float sep1 = c1 ? g - b : b - r;
        sep = c2 ? r - g : sep1;
        coeff = c2 ? 4.0f : (c1 ? 0.0f : 2.0f);
        mx = c2 ? b : mrg;
        mn = !c2 && b < nrg ? b : nrg;
    } else {
        if (r > g) {
            mx = r; mn = g; sep = g - b; coeff = 0.0f;
        } else {
            mx = g; mn = r; sep = b - r; coeff = 2.0f;
        }
        if (b > mx) {
            mx = b; sep = r - g; coeff = 4.0f;
        } else if (b < mn) {
            mn = b;
        }
    }
    float hh = 0.0f, ss = 0.0f;
    if (mx != mn) {
        float d = mx - mn;
        hh = (sep / d + coeff) * 60.0f;
        ss = d / mx;
    }
    h = hh;
    s = kUnit ? clamp_unit(ss) : clampf(ss, 0.0f, 1.0f);
    v = kUnit ? clamp_unit(mx) : clampf(mx, 0.0f, 1.0f);
}

// x / 60.0f for x == 0 and for 2^-124 <= x < 360: the product with
// RN(1/60), corrected once by its fused residual.  Equal to the division
// for every f32 in that range (an exhaustive host check); below 2^-124 the
// quotient nears the subnormal range and the shortcut can be off by one ulp.
__device__ __forceinline__ float div60_fast(float x) {
    const float c = 1.0f / 60.0f;
    const float q = x * c;
    const float r = __builtin_fmaf(-q, 60.0f, x);
    return __builtin_fmaf(r, c, q);
}

// palette Srgb::from_color(Hsv) (src/color_adjustment.rs:76).
// kUnit: v in [0, 1] and not -0, s in [0, 1], hdeg in [0, 360] and not -0
// (see adjust_pixel).
template <bool kUnit = false>
__device__ __forceinline__ void hsv_to_rgb(float hdeg, float s, float v, float& r, float& g,
                                           float& b) {
    float c = v * s;
    // hdeg - floor(hdeg/360)*360: for 0 <= hdeg < 360 the quotient rounds
    // below 1, so the floor is +0 (-0 for hdeg = -0) and the result is
    // hdeg + 0.0f (which turns -0 into +0, as the subtraction does)
    float hp = hdeg + 0.0f;
    if (kUnit) {
        // hdeg in [0, 360] and not -0: only 360 leaves the range, and
        // 360 - floor(360 / 360) * 360 is +0
        hp = hdeg < 360.0f ? hdeg : 0.0f;
    } else if (__builtin_expect(!(hdeg >= 0.0f && hdeg < 360.0f), 0)) {
        hp = hdeg - floorf(hdeg / 360.0f) * 360.0f;
    }
    float h = div60_fast(hp);
    if (__builtin_expect(!(hp == 0.0f || (hp >= 0x1p-124f && hp < 360.0f)), 0)) h = hp / 60.0f;
    float hm2 = h - 2.0f * floorf(h * 0.5f);
    float x = c * (1.0f - fabsf(hm2 - 1.0f));
    float m = v - c;
    // the sextant of h: [0,1) .. [4,5) are disjoint, anything else
    // (h >= 5, NaN) is the last case; each channel as selects.  kUnit: h is
    // not NaN and not below 0 (hp is in [0, 360)), so five thresholds fix it.
    bool s0, s1, s2, s3, s4;
    if (kUnit) {
        const bool a1 = h >= 1.0f, a2 = h >= 2.0f, a3 = h >= 3.0f, a4 = h >= 4.0f;
        const bool a5 = h >= 5.0f;
        s0 = !a1;
        s1 = a1 && !a2;
        s2 = a2 && !a3;
        s3 = a3 && !a4;
        s4 = a4 && !a5;
    } else {
        s0 = h >= 0.0f && h < 1.0f;
        s1 = h >= 1.0f && h < 2.0f;
        s2 = h >= 2.0f && h < 3.0f;
        s3 = h >= 3.0f && h < 4.0f;
        s4 = h >= 4.0f && h < 5.0f;
    }
    //        s0  s1  s2  s3  s4  else
    // rr:     c   x   0   0   x   c
    // gg:     x   c   c   x   0   0
    // bb:     0   0   x   c   c   x
    const float rr = (s1 || s4) ? x : ((s2 || s3) ? 0.0f : c);
    const float gg = (s0 || s3) ? x : ((s1 || s2) ? c : 0.0f);
    const float bb = (s0 || s1) ? 0.0f : ((s3 || s4) ? c : x);
    if (kUnit) {
        r = clamp_unit(rr + m);
        g = clamp_unit(gg + m);
        b = clamp_unit(bb + m);
    } else {
        r = clampf(rr + m, 0.0f, 1.0f);
        g = clampf(gg + m, 0.0f, 1.0f);
        b = clampf(bb + m, 0.0f, 1.0f);
    }
}

template <bool kUnit = false>
__device__ __forceinline__ float bright_contrast(float v, float brightness, float contrast) {
    if (kUnit) return clamp_unit(clamp_unit((v - 0.5f) * contrast + 0.5f) + brightness);
    float c = clampf((v - 0.5f) * contrast + 0.5f, 0.0f, 1.0f);
    return clampf(c + brightness, 0.0f, 1.0f);
}

// (v * 255).clamp(0, 255) as u8; kUnit: v in [0, 1] (not NaN), where the
// product already lies in [0, 255].
template <bool kUnit = false>
__device__ __forceinline__ uint32_t to_u8(float v) {
    if (kUnit) return (uint32_t)(v * 255.0f);
    float t = clampf(v * 255.0f, 0.0f, 255.0f);
    if (t != t) return 0;
    return (uint32_t)t;
}

// x / 255.0f for an integer 0 <= x <= 255, correctly rounded: the product
// with RN(1/255), corrected once by its exact fused residual.  The result
// equals the division for all 256 inputs (an exhaustive host check).
__device__ __forceinline__ float u8_unit(uint32_t x) {
    const float a = (float)x;
    const float c = 1.0f / 255.0f;
    const float q = a * c;
    const float r = __builtin_fmaf(-q, 255.0f, a);
    return __builtin_fmaf(r, c, q);
}

// fmodf(a, 360.0f).  For 360 <= |a| < 720 the remainder is a -/+ 360,
// exact by Sterbenz's lemma; a zero remainder keeps the sign of a.
// kUnit: |a| < 720 is known (see adjust_pixel), so no other case is tested.
template <bool kUnit = false>
__device__ __forceinline__ float fmod_360(float a) {
    const float m = __builtin_fabsf(a);
    if (!kUnit && !(m < 720.0f)) return fmodf(a, 360.0f);  // NaN and infinities too
    const float r = m >= 360.0f ? m - 360.0f : m;
    return __builtin_copysignf(r, a);
}

// ColorAdjustment::adjust_pixel (src/color_adjustment.rs:53-87).
// adj = {brightness, contrast, hue_shift, saturation}.
//
// kUnit: adj lies in ColorAdjustment::new's ranges (brightness in [-1, 1],
// contrast in [0, 2], hue_shift in [-180, 180], saturation in [0, 2], all
// finite), as the composite's per-cell adjustments do.  Then every value
// clamped to [0, 1] below is finite and not -0, so each clamp is a median:
//   * (v - 0.5) * contrast + 0.5 is a sum with +0.5 (a zero sum is +0);
//   * c + brightness adds a finite value to c in [0, 1] that is not -0 (a
//     zero sum of non-(-0) operands is +0);
//   * rgb_to_hsv's inputs are those results, so its max(., 0) are
//     identities, s = d / mx and v = mx are >= +0;
//   * hsv_to_rgb's rr + m has m = v - c >= +0 (c = v s <= v), and a zero
//     sum with a +0 operand is +0;
//   * the hue: rgb_to_hsv's h = (sep / d + coeff) * 60 has |sep| <= d (sep
//     is the difference of the two channels other than the maximum, d the
//     maximum minus the minimum, and rounding keeps the order), so h lies
//     in [-60, 300]; it is never -0 (sep is a difference of distinct values
//     or +0, and h starts at +0).  With hue_shift in [-180, 180],
//     h + hue_shift lies in [-240, 480] and is not -0 (a -0 sum needs two
//     -0 operands), so fmod_360 needs no range test, and rem_euclid gives
//     a value in [0, 360] that is not -0 (360 when a tiny negative
//     remainder rounds up);
// and every channel that reaches to_u8 lies in [0, 1].  The saturation
// product s * adj.w may be -0 (adj.w = -0) and keeps the general clamp.
//
// adjust_tail is everything after brightness/contrast: the HSV hue and
// saturation step and the conversion back to u8.
template <bool kUnit>
__device__ __forceinline__ void adjust_tail(float4 adj, float br, float bg, float bb,
                                            uint32_t& ro, uint32_t& go, uint32_t& bo) {
    if (adj.z != 0.0f || adj.w != 1.0f) {
        float h, s, v;
        rgb_to_hsv<kUnit>(br, bg, bb, h, s, v);
        float hr = fmod_360<kUnit>(h + adj.z);  // rem_euclid(360.0)
        if (hr < 0.0f) hr = hr + 360.0f;
        float s2 = clampf(s * adj.w, 0.0f, 1.0f);
        hsv_to_rgb<kUnit>(hr, s2, v, br, bg, bb);
    }
    ro = to_u8<kUnit>(br);
    go = to_u8<kUnit>(bg);
    bo = to_u8<kUnit>(bb);
}

template <bool kUnit = false>
__device__ __forceinline__ void adjust_pixel(float4 adj, uint32_t r8, uint32_t g8, uint32_t b8,
                                             uint32_t& ro, uint32_t& go, uint32_t& bo) {
    float r = u8_unit(r8), g = u8_unit(g8), b = u8_unit(b8);
    adjust_tail<kUnit>(adj, bright_contrast<kUnit>(r, adj.x, adj.y),
                       bright_contrast<kUnit>(g, adj.x, adj.y),
                       bright_contrast<kUnit>(b, adj.x, adj.y), ro, go, bo);
}

// The contrast stage at contrast 1 for channel value x in 0..255:
// clamp((x / 255 - 0.5) * 1 + 0.5, 0, 1), where the product by 1 is exact,
// so the value is RN(RN(x / 255 - 0.5) + 0.5) (not x / 255 itself below
// 64 / 255, where the subtraction rounds).  A 256-entry table of it,
// evaluated in f32 at compile time (round to nearest even, as at run
// time), stands in for bright_contrast's first clamp whenever a cell's
// contrast is exactly 1.
constexpr float contrast1_entry(int x) {
    const float v = (float)x / 255.0f;
    const float t = (v - 0.5f) + 0.5f;
    return t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
}

struct Contrast1Table {
    float v[256];
};

constexpr Contrast1Table make_contrast1_table() {
    Contrast1Table t{};
    for (int x = 0; x < 256; ++x) t.v[x] = contrast1_entry(x);
    return t;
}

// calculate_optimal_adjustment (src/color_adjustment.rs:99-154), followed
// by ColorAdjustment::new's clamps (:30-37).
__device__ __forceinline__ float4 optimal_adjustment(uint32_t t0, uint32_t t1, uint32_t t2,
                                                     uint32_t g0, uint32_t g1, uint32_t g2,
                                                     float k) {
    float tr = (float)t0 / 255.0f, tg = (float)t1 / 255.0f, tb = (float)t2 / 255.0f;
    float gr = (float)g0 / 255.0f, gg = (float)g1 / 255.0f, gb = (float)g2 / 255.0f;
    float tile_luma = 0.299f * tr + 0.587f * tg + 0.114f * tb;
    float target_luma = 0.299f * gr + 0.587f * gg + 0.114f * gb;
    float brightness = (target_luma - tile_luma) * k;
    float th, ts, tv, gh, gs, gv;
    rgb_to_hsv(tr, tg, tb, th, ts, tv);
    rgb_to_hsv(gr, gg, gb, gh, gs, gv);
    float hue = 0.0f;
    if (ts > 0.1f && gs > 0.1f) {
        float diff = gh - th;
        float wrapped = diff > 180.0f ? diff - 360.0f : (diff < -180.0f ? diff + 360.0f : diff);
        hue = wrapped * k * 0.5f;
    }
    float sat = 1.0f;
    if (ts > 0.01f) {
        float ratio = gs / ts;
        sat = 1.0f + (ratio - 1.0f) * k * 0.7f;
    }
    return make_float4(clampf(brightness, -1.0f, 1.0f), clampf(1.0f, 0.0f, 2.0f),
                       clampf(hue, -180.0f, 180.0f), clampf(sat, 0.0f, 2.0f));
}

// ------------------------------------------------------ wave reductions
// Minimum of a u32 over the 64 lanes with DPP lane moves (quad swaps, half-row
// and row mirrors, row broadcasts 15 and 31); the total lands in lane 63 and
// is returned uniformly.  All 64 lanes must be active.
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
    uint32_t t;
    t = (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0xB1, 0xF, 0xF, false);
    v = v < t ? v : t;
    t = (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x4E, 0xF, 0xF, false);
    v = v < t ? v : t;
    t = (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x141, 0xF, 0xF, false);
    v = v < t ? v : t;
    t = (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x140, 0xF, 0xF, false);
    v = v < t ? v : t;
    t = (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x142, 0xA, 0xF, false);
    v = v < t ? v : t;
    t = (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x143, 0xC, 0xF, false);
    v = v < t ? v : t;
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

__device__ __forceinline__ float readlane_f(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        uint64_t o = __shfl_xor(v, off, 64);
        v = o < v ? o : v;
    }
    return v;
}

}  // namespace mh
