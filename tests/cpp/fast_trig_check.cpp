// Exhaustive check of sincos_fast (raw-native_amd/csrc/ao_math.hpp) over the 2^24 angles
// 6.2831853f * k / 2^24 of the AO and SSAO samplers, against sincos_base and the host libm,
// and of the SSAO tap decision built on it (float32 pair unless an offset is near a
// half-integer) against the exact offsets on pseudo-random (angle, radius) pairs.
// Prints: max error vs sincos_base, max error vs libm, kFastTrigErr, kSsaoTapMargin,
// decisions checked, mismatches, exact fallbacks.
#include "ao_math.hpp"

#include <cmath>
#include <cstdio>

using namespace raw_ao;

static double worst(double e, float a, float b) {
    const double d = std::fabs(double(a) - double(b));
    return d > e ? d : e;
}

int main() {
    double e_base = 0.0, e_libm = 0.0;
    for (uint32_t k = 0; k < (1u << 24); ++k) {
        const float a = (float(k) / float(0x1000000)) * 6.2831853f;
        float fs, fc, bs, bc;
        sincos_fast(a, fs, fc);
        sincos_base(a, bs, bc);
        e_base = worst(worst(e_base, fs, bs), fc, bc);
        e_libm = worst(worst(e_libm, fs, std::sin(a)), fc, std::cos(a));
    }
    // the tap decision, as the SSAO kernel makes it, on hash-driven (angle, radius) pairs
    unsigned long long checked = 0, bad = 0, exact = 0;
    for (uint32_t i = 0; i < 4000000u; ++i) {
        const uint32_t ka = hash24(i, 7u, 11u), kr = hash24(i, 13u, 97u);
        const float a = (float(ka) / float(0x1000000)) * 6.2831853f;
        const float rad = (0.3f + 0.7f * (float(kr) / float(0x1000000))) * 6.0f;
        float fs, fc;
        sincos_fast(a, fs, fc);
        float ox = fc * rad, oy = fs * rad;
        const bool fallback = near_half(ox, kSsaoTapMargin) || near_half(oy, kSsaoTapMargin);
        exact += fallback;
        for (int which = 0; which < 2; ++which) {
            float es, ec;
            if (which == 0) sincos_base(a, es, ec);
            else { es = std::sin(a); ec = std::cos(a); }
            float gx = ox, gy = oy;
            if (fallback) { gx = ec * rad; gy = es * rad; }
            ++checked;
            bad += lround_f(gx) != lround_f(ec * rad) || lround_f(gy) != lround_f(es * rad);
        }
    }
    std::printf("%.9g %.9g %.9g %.9g %llu %llu %llu\n", e_base, e_libm, double(kFastTrigErr), double(kSsaoTapMargin), checked,
                bad, exact);
    return 0;
}
