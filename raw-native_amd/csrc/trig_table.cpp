// The host-libm correction table for the AO and SSAO sample angles.
//
// The reference computes std::cos/std::sin of phi = 6.2831853f * k / 2^24 with the
// host's libm (src/renderer/ray_ao.cpp:35-36, src/renderer/ssao.cpp:21-24). k has 2^24
// values, so this file evaluates the host libm on all of them, once per process, and
// records its difference from sincos_base() (ao_math.hpp), which the device evaluates
// bit-identically, as a 2-bit step per function: nibble k = cos step | sin step << 2
// with 0 none, 1 +1 ulp of magnitude, 3 -1 ulp. A difference beyond one ulp is an
// error (the table cannot express it), never a silent mismatch. Layout: a bitmap of the
// indices with a nonzero nibble (2^21 bytes, which stays in L2), then the nibbles, so the
// kernels touch the 8 MB nibble table only for the angles that need a step (2.6% with
// glibc 2.35).
#include "ao_math.hpp"
#include "ctx.hpp"
#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>
#include <thread>

namespace raw_ao {
namespace {
int32_t fbits(float f) { int32_t i; std::memcpy(&i, &f, 4); return i; }
int step_code(float libm, float mine, bool& bad) {
    if (std::signbit(libm) != std::signbit(mine)) { bad = true; return 0; }
    const int d = fbits(libm) - fbits(mine);
    if (d == 0) return 0;
    if (d == 1) return 1;
    if (d == -1) return 3;
    bad = true;
    return 0;
}
std::mutex g_mu;
std::vector<uint8_t> g_table;
bool g_built{false};
std::string g_err;
}  // namespace

bool build_trig_table(std::vector<uint8_t>& out, std::string& err) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_built) {
        const uint32_t N = 1u << 24;
        std::vector<uint8_t> t(N / 2, 0);
        std::atomic<bool> bad{false};
        unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> pool;
        const uint32_t per = (N / 2 + nt - 1) / nt;
        for (unsigned w = 0; w < nt; ++w)
            pool.emplace_back([&, w] {
                bool b = false;
                const uint32_t p0 = w * per, p1 = std::min(N / 2, p0 + per);
                for (uint32_t p = p0; p < p1; ++p) {
                    uint8_t byte = 0;
                    for (uint32_t h = 0; h < 2; ++h) {
                        const uint32_t k = 2 * p + h;
                        const float phi = 6.2831853f * (float(k) / float(0x1000000));
                        float s, c;
                        sincos_base(phi, s, c);
                        const volatile float vphi = phi;   // the libm call as the reference makes it
                        const float lc = std::cos(float(vphi)), ls = std::sin(float(vphi));
                        const int code = step_code(lc, c, b) | (step_code(ls, s, b) << 2);
                        byte = uint8_t(byte | (code << (4 * h)));
                    }
                    t[p] = byte;
                }
                if (b) bad = true;
            });
        for (auto& th : pool) th.join();
        if (bad) g_err = "host libm sin/cos differ from the correctly rounded value by more than one ulp";
        // the bitmap of indices with a nonzero nibble, then the nibbles
        std::vector<uint8_t> full(kTrigBitmapBytes + N / 2, 0);
        for (uint32_t k = 0; k < N; ++k)
            if ((t[k >> 1] >> ((k & 1u) * 4)) & 15u) full[k >> 3] = uint8_t(full[k >> 3] | (1u << (k & 7u)));
        std::copy(t.begin(), t.end(), full.begin() + kTrigBitmapBytes);
        g_table.swap(full);
        g_built = true;
    }
    if (!g_err.empty()) { err = g_err; return false; }
    out = g_table;
    return true;
}

}  // namespace raw_ao
