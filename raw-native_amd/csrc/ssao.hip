// Screen-space AO (computeSSAO, src/renderer/ssao.cpp:11-39) and the per-pixel part of
// the reconcile (src/cert/reconcile.cpp:11-16) on gfx950. Both are HBM-bound per-pixel
// passes: one thread a pixel, 16 x 16 pixel tiles so the SSAO taps (at most 6 px away)
// stay in the tile's cache lines.
#include "ao_math.hpp"
#include "ctx.hpp"
#include <algorithm>

namespace raw_ao {
namespace {

constexpr int TX = 16, TY = 16;

__global__ __launch_bounds__(TX* TY) void k_ssao(const float* __restrict__ pos, const float* __restrict__ nrm,
                                                 const uint8_t* __restrict__ mask, const uint8_t* __restrict__ trig,
                                                 int w, int h, int samples, float radius, float* __restrict__ ao) {
    const int x = int(blockIdx.x) * TX + int(threadIdx.x), y = int(blockIdx.y) * TY + int(threadIdx.y);
    if (x >= w || y >= h) return;
    const size_t i = size_t(y) * size_t(w) + size_t(x);
    if (!mask[i]) { ao[i] = 1.0f; return; }
    const V3 p = v3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    const V3 n = v3(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
    const int R = 6;
    int occ = 0, used = 0;
    // Taps in groups of U: the group's coordinates first, then all its loads (issued
    // together, from a clamped index where a tap leaves the image), then the tests. Each
    // tap's float operations are the ones of the one-tap-at-a-time loop.
    constexpr int U = 8;
    for (int s0 = 0; s0 < samples; s0 += U) {
        size_t jj[U];
        bool inb[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int s = s0 + u;
            const uint32_t ka = hash24(uint32_t(x), uint32_t(y), uint32_t(s));
            const float a = (float(ka) / float(0x1000000)) * 6.2831853f;
            const float rad = (0.3f + 0.7f * hash01(uint32_t(x), uint32_t(y), uint32_t(s + 97))) * float(R);
            // Only the rounded offsets are used, so a float32 sin/cos decides them unless an
            // offset lies within kSsaoTapMargin of a half-integer; there the exact pair is computed.
            float sa, ca;
            sincos_fast(a, sa, ca);
            float ox = ca * rad, oy = sa * rad;
            if (near_half(ox, kSsaoTapMargin) || near_half(oy, kSsaoTapMargin)) {
                sincos_base(a, sa, ca);
                if (trig) trig_fix(trig, ka, sa, ca);
                ox = ca * rad;
                oy = sa * rad;
            }
            const int sx = x + lround_f(ox);
            const int sy = y + lround_f(oy);
            inb[u] = s < samples && sx >= 0 && sy >= 0 && sx < w && sy < h;
            jj[u] = inb[u] ? size_t(sy) * size_t(w) + size_t(sx) : i;
        }
        uint8_t mk[U];
        float qx[U], qy[U], qz[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            mk[u] = mask[jj[u]];
            qx[u] = pos[3 * jj[u]];
            qy[u] = pos[3 * jj[u] + 1];
            qz[u] = pos[3 * jj[u] + 2];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (!inb[u] || !mk[u]) continue;
            ++used;
            const V3 d = sub(v3(qx[u], qy[u], qz[u]), p);
            const float dist = length(d);
            if (dist < 1e-4f || dist > radius) continue;
            const float ndl = dot(n, normalize(d));
            if (ndl > 0.15f) ++occ;
        }
    }
    const float a = used > 0 ? 1.0f - float(occ) / float(used) : 1.0f;
    ao[i] = a < 0.0f ? 0.0f : a;
}

// |approx - truth| where covered, 0 elsewhere; the maximum through an unsigned atomic on
// the bits (every value is >= 0, and a NaN never beats the running maximum, as
// `e > maxError` in the reference), one atomic a block: grid-stride over the pixels, the
// block's maximum reduced in LDS first.
__global__ __launch_bounds__(256) void k_error_map(const float* __restrict__ a, const float* __restrict__ t,
                                                   const uint8_t* __restrict__ mask, uint32_t npix, float* __restrict__ err,
                                                   uint32_t* __restrict__ maxbits) {
    __shared__ float wmax[4];
    float m = 0.0f;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < npix; i += gridDim.x * 256) {
        const float e = mask[i] ? fabsf(a[i] - t[i]) : 0.0f;
        err[i] = e;
        if (e > m) m = e;
    }
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        const float b = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
        if (b > 0.0f) atomicMax(maxbits, __float_as_uint(b));
    }
}

}  // namespace

void ssao_enqueue(const uint8_t* trig, const float* d_pos, const float* d_nrm, const uint8_t* d_mask, int w, int h, int samples,
                  float radius, float* d_ao, hipStream_t s) {
    const dim3 grid((w + TX - 1) / TX, (h + TY - 1) / TY), block(TX, TY);
    k_ssao<<<grid, block, 0, s>>>(d_pos, d_nrm, d_mask, trig, w, h, samples, radius, d_ao);
}

int ssao(Ctx& c, const float* d_pos, const float* d_nrm, const uint8_t* d_mask, int w, int h, int samples, float radius,
         float* d_ao, hipStream_t s) {
    (void)hipEventRecord(c.ev[4], s);
    ssao_enqueue(c.host_trig ? c.d_trig : nullptr, d_pos, d_nrm, d_mask, w, h, samples, radius, d_ao, s);
    (void)hipEventRecord(c.ev[5], s);
    return hip_ok(c, hipGetLastError(), "ssao launch") ? 0 : -2;
}

int error_map(Ctx& c, const float* d_a, const float* d_t, const uint8_t* d_mask, int w, int h, float* d_err, float* d_max,
              hipStream_t s) {
    const uint32_t npix = uint32_t(w) * uint32_t(h);
    if (!hip_ok(c, hipMemsetAsync(d_max, 0, 4, s), "max clear")) return -2;
    (void)hipEventRecord(c.ev[6], s);
    k_error_map<<<std::min(2048u, (npix + 255) / 256), 256, 0, s>>>(d_a, d_t, d_mask, npix, d_err,
                                                                    reinterpret_cast<uint32_t*>(d_max));
    (void)hipEventRecord(c.ev[7], s);
    return hip_ok(c, hipGetLastError(), "error map launch") ? 0 : -2;
}

}  // namespace raw_ao
