// Float32 arithmetic of raw-native's AO path, written once for the host and for gfx950.
//
// Every expression keeps the reference's operation order and is compiled with
// -ffp-contract=off (no fused multiply-add), correctly rounded '/' and sqrt, and
// denormals kept, so the device computes the same bits as the reference's C++:
//   - normalize is v * (1/l), not v / l                        (raw/math/vec.hpp:16)
//   - std::min(a,b) is `b < a ? b : a`, std::max(a,b) `a < b ? b : a`; signed zeros
//     and NaNs make the operand order visible, so they are spelled out (fmin/fmax
//     would differ)
//   - std::lround rounds halves away from zero; floor(v + 0.5f) is wrong at
//     0.49999997f, so it is computed from the truncation
#pragma once
#include <cstdint>
#include <cmath>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define AO_HD __host__ __device__ __forceinline__
#else
#define AO_HD inline
#endif

namespace raw_ao {

struct V3 { float x, y, z; };
AO_HD V3 v3(float x, float y, float z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
AO_HD V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
AO_HD V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
AO_HD V3 scl(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
AO_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
AO_HD V3 cross(V3 a, V3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
AO_HD float stdmin(float a, float b) { return b < a ? b : a; }
AO_HD float stdmax(float a, float b) { return a < b ? b : a; }
AO_HD float length(V3 v) { return sqrtf(dot(v, v)); }
AO_HD V3 normalize(V3 v) { const float l = length(v); return l > 0.0f ? scl(v, 1.0f / l) : v; }

// The per-pixel counter hash (src/renderer/ray_ao.cpp:8-12): 24 bits, then / 2^24.
AO_HD uint32_t hash24(uint32_t x, uint32_t y, uint32_t s) {
    uint32_t h = x * 374761393u + y * 668265263u + s * 2246822519u;
    h = (h ^ (h >> 13)) * 1274126177u;
    h ^= h >> 16;
    return h & 0xFFFFFFu;
}
AO_HD float hash01(uint32_t x, uint32_t y, uint32_t s) { return float(hash24(x, y, s)) / float(0x1000000); }

// std::lround of a float (halves away from zero), exact for |v| < 2^23.
AO_HD int lround_f(float v) {
    const float t = truncf(v);
    const float f = v - t;   // exact
    return int(t) + (f >= 0.5f ? 1 : (f <= -0.5f ? -1 : 0));
}

// ---- sin and cos of phi = 6.2831853f * k / 2^24 -----------------------------------
// The reference calls the host libm (std::cos/std::sin of a float). Its angle has only
// 2^24 values, so the device reproduces the host's results exactly in two steps:
//  1. sincos_base(): sin/cos in float64 by Cody-Waite reduction with the three-part
//     pi/2 split of Cody and Waite ("Software Manual for the Elementary Functions",
//     1980), then Taylor polynomials to degree 17/18 on |r| <= pi/4 (truncation error
//     below 1e-19), rounded once to float. Plain IEEE adds and multiplies only, so the
//     host and the device compute the same bits.
//  2. a 2-bit-per-angle table of the host libm's difference from step 1 in units in the
//     last place (always -1, 0 or +1 for glibc), built on the host from the host libm
//     (trig_table.cpp) and applied with trig_adjust().
AO_HD void sincos_base(float phi, float& s, float& c) {
    const double x = double(phi);
    const int q = int(x * 0.63661977236758134308 + 0.5);   // nearest quadrant, x >= 0
    const double P1 = 1.57079632673412561417e+00;            // pi/2 to 33 bits: q * P1 is exact for q <= 4
    const double P2 = 6.07710050650619224932e-11;            // pi/2 - P1 to double precision
    const double dq = double(q);
    const double r = (x - dq * P1) - dq * P2;
    const double r2 = r * r;
    // sin r = r (1 - r^2/3! + r^4/5! - ... ), cos r = 1 - r^2/2! + r^4/4! - ...
    double ps = -1.0 / 355687428096000.0;                    // -1/17!
    ps = ps * r2 + 1.0 / 1307674368000.0;                    //  1/15!
    ps = ps * r2 - 1.0 / 6227020800.0;                       // -1/13!
    ps = ps * r2 + 1.0 / 39916800.0;                         //  1/11!
    ps = ps * r2 - 1.0 / 362880.0;                           // -1/9!
    ps = ps * r2 + 1.0 / 5040.0;                             //  1/7!
    ps = ps * r2 - 1.0 / 120.0;                              // -1/5!
    ps = ps * r2 + 1.0 / 6.0;                                //  1/3!
    const double sr = r - (r * r2) * ps;
    double pc = 1.0 / 6402373705728000.0;                    //  1/18!
    pc = pc * r2 - 1.0 / 20922789888000.0;                   // -1/16!
    pc = pc * r2 + 1.0 / 87178291200.0;                      //  1/14!
    pc = pc * r2 - 1.0 / 479001600.0;                        // -1/12!
    pc = pc * r2 + 1.0 / 3628800.0;                          //  1/10!
    pc = pc * r2 - 1.0 / 40320.0;                            // -1/8!
    pc = pc * r2 + 1.0 / 720.0;                              //  1/6!
    pc = pc * r2 - 1.0 / 24.0;                               // -1/4!
    pc = pc * r2 + 1.0 / 2.0;                                //  1/2!
    const double cr = 1.0 - r2 * pc;
    // quadrant q: (sin, cos) = (sr, cr), (cr, -sr), (-sr, -cr), (-cr, sr), by selects and sign
    // flips (rounding to float commutes with negation)
    const float fs = float(sr), fc = float(cr);
    const bool odd = (q & 1) != 0;
    const float as = odd ? fc : fs, ac = odd ? fs : fc;
    s = (q & 2) ? -as : as;
    c = ((q + 1) & 2) ? -ac : ac;
}

AO_HD float bits_step(float v, int d) {
    uint32_t u;
    __builtin_memcpy(&u, &v, 4);
    u = uint32_t(int32_t(u) + d);
    float r;
    __builtin_memcpy(&r, &u, 4);
    return r;
}
// Apply one table nibble: bits 0-1 the cos step, bits 2-3 the sin step (0: none,
// 1: +1 ulp of magnitude, 3: -1 ulp).
AO_HD void trig_adjust(uint32_t nib, float& s, float& c) {
    const int dc = (nib & 3u) == 1u ? 1 : ((nib & 3u) == 3u ? -1 : 0);
    const int ds = ((nib >> 2) & 3u) == 1u ? 1 : (((nib >> 2) & 3u) == 3u ? -1 : 0);
    c = bits_step(c, dc);
    s = bits_step(s, ds);
}
// The host-libm correction table: a bitmap of the 2^24 angle indices whose host sin or
// cos differs from the correctly rounded value (2 MB: it stays in L2), then one nibble an
// index (8 MB), read only where the bit is set (2.6% of the angles with glibc 2.35).
constexpr uint32_t kTrigBitmapBytes = 1u << 21;
AO_HD void trig_fix(const uint8_t* t, uint32_t k, float& s, float& c) {
    const uint32_t word = reinterpret_cast<const uint32_t*>(t)[k >> 5];
    if ((word >> (k & 31u)) & 1u) trig_adjust((t[kTrigBitmapBytes + (k >> 1)] >> ((k & 1u) * 4)) & 15u, s, c);
}

// A float32 sin/cos of the same angles for callers that only round a product of them to an
// integer (the SSAO taps): Cody-Waite reduction to |r| <= pi/4 with pi/2 in two float parts,
// then Taylor polynomials (sin to r^9, cos to r^10), fused multiply-adds written out. Checked
// against sincos_base and the host libm over all 2^24 angles (tests/cpp/fast_trig_check.cpp):
// at most 6e-8 from either, inside kFastTrigErr.
constexpr float kFastTrigErr = 4e-7f;
AO_HD void sincos_fast(float phi, float& s, float& c) {
    const int q = int(phi * 0.636619772f + 0.5f);   // nearest quadrant, phi >= 0
    const float qf = float(q);
    float r = fmaf(-qf, 1.57079637f, phi);
    r = fmaf(-qf, -4.37113883e-8f, r);
    const float r2 = r * r;
    float ps = fmaf(r2, 2.75573192e-6f, -1.98412698e-4f);
    ps = fmaf(r2, ps, 8.33333333e-3f);
    ps = fmaf(r2, ps, -1.66666667e-1f);
    const float sr = fmaf(r * r2, ps, r);
    float pc = fmaf(r2, -2.75573192e-7f, 2.48015873e-5f);
    pc = fmaf(r2, pc, -1.38888889e-3f);
    pc = fmaf(r2, pc, 4.16666667e-2f);
    pc = fmaf(r2, pc, -0.5f);
    const float cr = fmaf(r2, pc, 1.0f);
    switch (q & 3) {
        case 0: s = sr; c = cr; break;
        case 1: s = cr; c = -sr; break;
        case 2: s = -sr; c = -cr; break;
        default: s = -cr; c = sr; break;
    }
}
// An SSAO tap offset (cos or sin) x rad, rad <= 6, computed from the fast pair differs from
// the exact one by at most 6 x kFastTrigErr plus half an ulp of each product (2^-22 each
// below 8): 2.9e-6. Where an offset is within this margin (over three times that) of a
// half-integer, the tap takes the exact pair.
constexpr float kSsaoTapMargin = 1e-5f;
// True when v lies within m of a half-integer, where std::lround's result changes.
AO_HD bool near_half(float v, float m) { return fabsf(v - (floorf(v) + 0.5f)) <= m; }

// ---- watertight ray/triangle test (raw/math/primitives.hpp:17-64) ------------------
// The part that depends only on the ray direction is computed once per ray.
struct RayFrame {
    int kz;
    bool sw;
    float sx, sy, sz;
    int ix, iy;   // the components of v that wframe(v, kz, sw) puts first and second (its third is kz)
};
AO_HD V3 wframe(V3 v, int kz, bool swap) {
    const V3 p = kz == 0 ? v3(v.y, v.z, v.x) : (kz == 1 ? v3(v.z, v.x, v.y) : v);
    return swap ? v3(p.y, p.x, p.z) : p;
}
AO_HD RayFrame ray_frame(V3 rd) {
    RayFrame f;
    const float adx = fabsf(rd.x), ady = fabsf(rd.y), adz = fabsf(rd.z);
    f.kz = adx > ady ? (adx > adz ? 0 : 2) : (ady > adz ? 1 : 2);
    V3 d = wframe(rd, f.kz, false);
    f.sw = d.z < 0.0f;
    if (f.sw) d = v3(d.y, d.x, d.z);
    f.sx = d.x / d.z;
    f.sy = d.y / d.z;
    f.sz = 1.0f / d.z;
    const int i0 = f.kz == 2 ? 0 : f.kz + 1, i1 = f.kz == 0 ? 2 : f.kz - 1;   // (kz + 1) % 3, (kz + 2) % 3
    f.ix = f.sw ? i1 : i0;
    f.iy = f.sw ? i0 : i1;
    return f;
}
// Component i of v (0, 1, 2: x, y, z) by selects: wframe by per-ray indices, without the
// per-lane branches a rotation of the whole vector compiles to.
AO_HD float comp(V3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
AO_HD V3 wframe_f(V3 v, const RayFrame& f) { return v3(comp(v, f.ix), comp(v, f.iy), comp(v, f.kz)); }
// intersectTri's hit decision and t: true when the ray hits with t > 1e-4 (t written).
// SEL: the vertices' frame change by per-component selects (wframe_f), else by rotating and
// swapping whole vectors (wframe); the same values either way. Which compiles faster depends on
// the calling kernel (how coherent its rays' dominant axes are, and its register pressure).
template <bool SEL = true>
AO_HD bool intersect_t(V3 ro, V3 rd, const RayFrame& f, V3 ta, V3 tb, V3 tc, float& t) {
    const V3 A = SEL ? wframe_f(sub(ta, ro), f) : wframe(sub(ta, ro), f.kz, f.sw);
    const V3 B = SEL ? wframe_f(sub(tb, ro), f) : wframe(sub(tb, ro), f.kz, f.sw);
    const V3 C = SEL ? wframe_f(sub(tc, ro), f) : wframe(sub(tc, ro), f.kz, f.sw);
    const float Ax = A.x - f.sx * A.z, Ay = A.y - f.sy * A.z;
    const float Bx = B.x - f.sx * B.z, By = B.y - f.sy * B.z;
    const float Cx = C.x - f.sx * C.z, Cy = C.y - f.sy * C.z;
    const float U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const V3 e1 = sub(tb, ta), e2 = sub(tc, ta), p = cross(rd, e2);
    const float dm = dot(e1, p);
    if (dm > -1e-7f && dm < 1e-7f) {
        const float T = U * (f.sz * A.z) + V * (f.sz * B.z) + W * (f.sz * C.z), inv = 1.0f / det;
        t = T * inv;
        return t > 1e-4f;
    }
    const float inv = 1.0f / dm;
    const V3 tv = sub(ro, ta), q = cross(tv, e1);
    t = dot(e2, q) * inv;
    return t > 1e-4f;
}

// The whole of intersectTri: the hit decision, t and the barycentric weights u, v of b and c.
AO_HD bool intersect_tuv(V3 ro, V3 rd, const RayFrame& f, V3 ta, V3 tb, V3 tc, float& t, float& u, float& v) {
    const V3 A = wframe_f(sub(ta, ro), f), B = wframe_f(sub(tb, ro), f), C = wframe_f(sub(tc, ro), f);
    const float Ax = A.x - f.sx * A.z, Ay = A.y - f.sy * A.z;
    const float Bx = B.x - f.sx * B.z, By = B.y - f.sy * B.z;
    const float Cx = C.x - f.sx * C.z, Cy = C.y - f.sy * C.z;
    const float U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const V3 e1 = sub(tb, ta), e2 = sub(tc, ta), p = cross(rd, e2);
    const float dm = dot(e1, p);
    if (dm > -1e-7f && dm < 1e-7f) {
        const float T = U * (f.sz * A.z) + V * (f.sz * B.z) + W * (f.sz * C.z), inv = 1.0f / det;
        t = T * inv;
        u = V * inv;
        v = W * inv;
        return t > 1e-4f;
    }
    const float inv = 1.0f / dm;
    const V3 tv = sub(ro, ta), q = cross(tv, e1);
    u = stdmax(0.0f, dot(tv, p) * inv);
    v = stdmax(0.0f, dot(rd, q) * inv);
    t = dot(e2, q) * inv;
    const float s = u + v;
    if (s > 1.0f) {
        u /= s;
        v /= s;
    }
    return t > 1e-4f;
}

// basis (src/renderer/ray_ao.cpp:14-17).
AO_HD void basis(V3 n, V3& t, V3& b) {
    const V3 a = fabsf(n.x) > 0.9f ? v3(0, 1, 0) : v3(1, 0, 0);
    t = normalize(cross(a, n));
    b = cross(n, t);
}
// One AO sample direction (src/renderer/ray_ao.cpp:32-38) from u1, u2 and the angle's
// sin and cos.
AO_HD V3 ao_dir(float u1, float sphi, float cphi, V3 n, V3 t, V3 b) {
    const float r = sqrtf(u1);
    const float lx = r * cphi, ly = r * sphi;
    const float lz = sqrtf(stdmax(0.0f, 1.0f - u1));
    return normalize(add(add(scl(t, lx), scl(b, ly)), scl(n, lz)));
}

}  // namespace raw_ao
