// The G-buffer on the GPU: raw::rasterize (src/renderer/raster.cpp:92-126, default
// RasterOptions) of one indexed mesh, bit for bit.
//
// The reference draws triangles in order and keeps a pixel's first nearest fragment
// (`depth >= stored` skips), so the winner of a pixel is the fragment with the smallest depth,
// ties to the earliest drawn. Here:
//   setup    one thread a triangle: clip-space vertices, the near-plane clip of triangles
//            with a vertex at w <= 1e-6 (raster.cpp:17-27, a fan of up to 2), each drawn
//            triangle's screen vertices, area and pixel box (raster.cpp:30-50); the draw
//            order id of fan triangle k of triangle t is 2t + k.
//   scan     the boxes' pixel counts, prefix-summed, spread the candidate pixels evenly
//            over the threads whatever the triangle sizes.
//   cover    each candidate pixel inside its triangle (the reference's edge weights) does a
//            64-bit atomicMin of (depth bits, draw id): depths are positive, so their bits
//            order as the floats do.
//   resolve  each pixel recomputes its winner's weights, depth, position and normal with the
//            reference's float operations.
// Depth NaNs (which the reference lets any later fragment overwrite) do not occur for
// triangles in front of the near plane.
#include "../../include/raw_ao_hip.h"
#include "ao_math.hpp"
#include "ctx.hpp"
#include <cmath>
#include <limits>
#include <vector>

namespace raw_ao {
namespace {

constexpr int RBT = 256;

struct Cam {   // the view-projection matrix, row-major, passed by value
    float m[16];
};

struct RTri {
    float sx[3], sy[3], sz[3], invw[3];
    float area;
    int minx, miny, bw;
    float wp[9], wn[9];
};

struct CV {
    float4 cs;
    V3 wp, wn;
};

__device__ CV lerp_cv(const CV& a, const CV& b, float t) {
    CV r;
    r.cs = make_float4(a.cs.x + (b.cs.x - a.cs.x) * t, a.cs.y + (b.cs.y - a.cs.y) * t, a.cs.z + (b.cs.z - a.cs.z) * t,
                       a.cs.w + (b.cs.w - a.cs.w) * t);
    r.wp = v3(a.wp.x + (b.wp.x - a.wp.x) * t, a.wp.y + (b.wp.y - a.wp.y) * t, a.wp.z + (b.wp.z - a.wp.z) * t);
    r.wn = v3(a.wn.x + (b.wn.x - a.wn.x) * t, a.wn.y + (b.wn.y - a.wn.y) * t, a.wn.z + (b.wn.z - a.wn.z) * t);
    return r;
}

// (int) of a float as x86-64 converts it (cvttss2si): INT_MIN when out of range or NaN, where
// the GPU's conversion saturates. A vertex far off screen then empties the box, as on the host.
__device__ int x86_int(float f) { return (f >= -2147483648.0f && f < 2147483648.0f) ? int(f) : int(0x80000000u); }

// The screen-space setup of drawTri (raster.cpp:30-50); count = box pixels (0: culled).
__device__ uint32_t setup_tri(const CV v[3], int w, int h, RTri& t) {
    for (int k = 0; k < 3; ++k) {
        t.invw[k] = 1.0f / v[k].cs.w;
        const float nx = v[k].cs.x * t.invw[k], ny = v[k].cs.y * t.invw[k];
        t.sx[k] = (nx * 0.5f + 0.5f) * float(w) + 0.0f;
        t.sy[k] = (1.0f - (ny * 0.5f + 0.5f)) * float(h) + 0.0f;
        t.sz[k] = v[k].cs.w;
        t.wp[3 * k] = v[k].wp.x;
        t.wp[3 * k + 1] = v[k].wp.y;
        t.wp[3 * k + 2] = v[k].wp.z;
        t.wn[3 * k] = v[k].wn.x;
        t.wn[3 * k + 1] = v[k].wn.y;
        t.wn[3 * k + 2] = v[k].wn.z;
    }
    const float mnx = stdmin(stdmin(t.sx[0], t.sx[1]), t.sx[2]), mxx = stdmax(stdmax(t.sx[0], t.sx[1]), t.sx[2]);
    const float mny = stdmin(stdmin(t.sy[0], t.sy[1]), t.sy[2]), mxy = stdmax(stdmax(t.sy[0], t.sy[1]), t.sy[2]);
    const int minx = max(0, x86_int(floorf(mnx))), maxx = min(w - 1, x86_int(ceilf(mxx)));
    const int miny = max(0, x86_int(floorf(mny))), maxy = min(h - 1, x86_int(ceilf(mxy)));
    t.area = (t.sx[1] - t.sx[0]) * (t.sy[2] - t.sy[0]) - (t.sy[1] - t.sy[0]) * (t.sx[2] - t.sx[0]);
    t.minx = minx;
    t.miny = miny;
    t.bw = maxx - minx + 1;
    if (fabsf(t.area) < 1e-9f || maxx < minx || maxy < miny) return 0u;
    return uint32_t(t.bw) * uint32_t(maxy - miny + 1);
}

__global__ __launch_bounds__(RBT) void k_rast_setup(const float* __restrict__ vpos, const float* __restrict__ vnrm,
                                                    const int32_t* __restrict__ idx, uint32_t ntri, const Cam vp, float nearz, int w, int h, RTri* __restrict__ tris,
                                                    uint64_t* __restrict__ counts) {
    const uint32_t i = blockIdx.x * RBT + threadIdx.x;
    if (i >= ntri) return;
    CV v[3];
    bool behind = false;
    for (int k = 0; k < 3; ++k) {
        const size_t vi = size_t(idx[3 * size_t(i) + k]);
        const V3 p = v3(vpos[3 * vi], vpos[3 * vi + 1], vpos[3 * vi + 2]);
        // mul(Mat4, Vec4) (raw/math/mat.hpp), left to right
        const float* m = vp.m;
        v[k].cs = make_float4(m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3] * 1.0f, m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7] * 1.0f,
                              m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11] * 1.0f,
                              m[12] * p.x + m[13] * p.y + m[14] * p.z + m[15] * 1.0f);
        v[k].wp = p;
        v[k].wn = v3(vnrm[3 * vi], vnrm[3 * vi + 1], vnrm[3 * vi + 2]);
        behind = behind || v[k].cs.w <= 1e-6f;
    }
    uint64_t c0 = 0, c1 = 0;
    if (!behind) {
        c0 = setup_tri(v, w, h, tris[2 * size_t(i)]);
    } else {   // clipNear (raster.cpp:17-27), then the fan (raster.cpp:117-121)
        CV c[4];
        int n = 0;
        for (int k = 0; k < 3; ++k) {
            const CV& a = v[k];
            const CV& b = v[(k + 1) % 3];
            const bool ain = a.cs.w >= nearz, bin = b.cs.w >= nearz;
            if (ain) c[n++] = a;
            if (ain != bin) c[n++] = lerp_cv(a, b, (nearz - a.cs.w) / (b.cs.w - a.cs.w));
        }
        for (int k = 1; k + 1 < n; ++k) {
            const CV t3[3] = {c[0], c[k], c[k + 1]};
            const uint32_t cnt = setup_tri(t3, w, h, tris[2 * size_t(i) + size_t(k - 1)]);
            if (k == 1) c0 = cnt;
            else c1 = cnt;
        }
    }
    counts[2 * size_t(i)] = c0;
    counts[2 * size_t(i) + 1] = c1;
}

// The edge weights of pixel (x, y) (raster.cpp:51-56); false outside.
__device__ bool weights(const RTri& t, int x, int y, float& w0, float& w1, float& w2) {
    const float px = float(x) + 0.5f, py = float(y) + 0.5f;
    w0 = ((t.sx[1] - px) * (t.sy[2] - py) - (t.sy[1] - py) * (t.sx[2] - px)) / t.area;
    w1 = ((t.sx[2] - px) * (t.sy[0] - py) - (t.sy[2] - py) * (t.sx[0] - px)) / t.area;
    w2 = 1.0f - w0 - w1;
    return !(w0 < 0.0f || w1 < 0.0f || w2 < 0.0f);
}

constexpr int PER = 4;   // candidates a thread

__global__ __launch_bounds__(RBT) void k_rast_cover(const RTri* __restrict__ tris, const uint64_t* __restrict__ offs, uint32_t nsub,
                                                    const uint64_t* __restrict__ total_p, int w,
                                                    unsigned long long* __restrict__ keys) {
    const uint64_t total = *total_p;
    const uint64_t stride = uint64_t(gridDim.x) * RBT * PER;
    for (uint64_t c0 = (uint64_t(blockIdx.x) * RBT + threadIdx.x) * PER; c0 < total; c0 += stride) {
        // the triangle of candidate c0: the last s with offs[s] <= c0
        uint32_t lo = 0, hi = nsub;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (offs[mid] <= c0) lo = mid;
            else hi = mid;
        }
        uint32_t s = lo;
        for (uint64_t c = c0; c < c0 + PER && c < total; ++c) {
            while (offs[s + 1] <= c) ++s;
            const RTri& t = tris[s];
            const uint32_t li = uint32_t(c - offs[s]);
            const int x = t.minx + int(li % uint32_t(t.bw)), y = t.miny + int(li / uint32_t(t.bw));
            float w0, w1, w2;
            if (!weights(t, x, y, w0, w1, w2)) continue;
            const float depth = w0 * t.sz[0] + w1 * t.sz[1] + w2 * t.sz[2];
            if (!(depth < __uint_as_float(0x7f800000u))) continue;   // never below the cleared +inf
            const unsigned long long key = (static_cast<unsigned long long>(__float_as_uint(depth)) << 32) | s;
            atomicMin(&keys[size_t(y) * size_t(w) + size_t(x)], key);
        }
    }
}

// alb_out (optional): the albedo of the winning triangle's mesh, mesh m holding the triangles
// below mesh_end[m] (and at or above mesh_end[m-1]).
__global__ __launch_bounds__(RBT) void k_rast_resolve(const RTri* __restrict__ tris, const unsigned long long* __restrict__ keys,
                                                      uint32_t npix, int w, float* __restrict__ depth, float* __restrict__ pos,
                                                      float* __restrict__ nrm, uint8_t* __restrict__ mask,
                                                      const uint32_t* __restrict__ mesh_end, const float* __restrict__ albedo,
                                                      uint32_t nmesh, float* __restrict__ alb_out) {
    const uint32_t i = blockIdx.x * RBT + threadIdx.x;
    if (i >= npix) return;
    const unsigned long long key = keys[i];
    if (key == ~0ull) {
        depth[i] = __uint_as_float(0x7f800000u);
        for (int k = 0; k < 3; ++k) {
            pos[3 * size_t(i) + k] = 0.0f;
            nrm[3 * size_t(i) + k] = 0.0f;
            if (alb_out) alb_out[3 * size_t(i) + k] = 0.0f;
        }
        mask[i] = 0;
        return;
    }
    const uint32_t sub = uint32_t(key & 0xffffffffull);
    if (alb_out) {
        const uint32_t tri = sub >> 1;
        uint32_t lo = 0, hi = nmesh - 1;   // the first mesh whose end is past tri
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (mesh_end[mid] > tri) hi = mid;
            else lo = mid + 1;
        }
        for (int k = 0; k < 3; ++k) alb_out[3 * size_t(i) + k] = albedo[3 * size_t(lo) + k];
    }
    const RTri& t = tris[sub];
    float w0, w1, w2;
    weights(t, int(i % uint32_t(w)), int(i / uint32_t(w)), w0, w1, w2);
    const float iw = w0 * t.invw[0] + w1 * t.invw[1] + w2 * t.invw[2];
    depth[i] = w0 * t.sz[0] + w1 * t.sz[1] + w2 * t.sz[2];
    const auto pc = [&](const float* a) {
        return v3((w0 * a[0] * t.invw[0] + w1 * a[3] * t.invw[1] + w2 * a[6] * t.invw[2]) / iw,
                  (w0 * a[1] * t.invw[0] + w1 * a[4] * t.invw[1] + w2 * a[7] * t.invw[2]) / iw,
                  (w0 * a[2] * t.invw[0] + w1 * a[5] * t.invw[1] + w2 * a[8] * t.invw[2]) / iw);
    };
    const V3 p = pc(t.wp), n = normalize(pc(t.wn));
    pos[3 * size_t(i)] = p.x;
    pos[3 * size_t(i) + 1] = p.y;
    pos[3 * size_t(i) + 2] = p.z;
    nrm[3 * size_t(i)] = n.x;
    nrm[3 * size_t(i) + 1] = n.y;
    nrm[3 * size_t(i) + 2] = n.z;
    mask[i] = 1;
}

__global__ void k_fill_keys(unsigned long long* keys, uint32_t n) {
    for (uint32_t i = blockIdx.x * RBT + threadIdx.x; i < n; i += gridDim.x * RBT) keys[i] = ~0ull;
}
// ---- exclusive prefix sum of the boxes' 64-bit pixel counts ---------------------------
// Three passes: each workgroup scans a tile of SCAN_TILE counts (loaded and stored coalesced
// through LDS, a thread summing SCAN_PER consecutive ones, then wave and workgroup scans) and
// leaves its total; one workgroup scans the totals and writes the grand total at out[n] (the
// cover pass's sentinel); then every tile adds its offset. Integer sums, so the order of
// additions does not change the result.
constexpr int SCAN_T = 256, SCAN_PER = 16, SCAN_TILE = SCAN_T * SCAN_PER;
__global__ __launch_bounds__(SCAN_T) void k_scan_tiles(const uint64_t* __restrict__ in, uint64_t* __restrict__ out, uint32_t n,
                                                       uint64_t* __restrict__ part) {
    __shared__ uint64_t tile[SCAN_TILE];
    __shared__ uint64_t wsum[SCAN_T / 64];
    const size_t t0 = size_t(blockIdx.x) * SCAN_TILE;
    for (int k = threadIdx.x; k < SCAN_TILE; k += SCAN_T) tile[k] = t0 + k < n ? in[t0 + k] : 0ull;
    __syncthreads();
    uint64_t v[SCAN_PER], sum = 0;
    for (int k = 0; k < SCAN_PER; ++k) {
        v[k] = tile[threadIdx.x * SCAN_PER + k];
        sum += v[k];
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint64_t x = sum;   // inclusive scan across the wave
    for (int off = 1; off < 64; off <<= 1) {
        const uint64_t y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    uint64_t run = x - sum;
    for (int k = 0; k < w; ++k) run += wsum[k];
    for (int k = 0; k < SCAN_PER; ++k) {
        tile[threadIdx.x * SCAN_PER + k] = run;
        run += v[k];
    }
    if (threadIdx.x == SCAN_T - 1) part[blockIdx.x] = run;   // the tile's total
    __syncthreads();
    for (int k = threadIdx.x; k < SCAN_TILE; k += SCAN_T)
        if (t0 + k < n) out[t0 + k] = tile[k];
}
__global__ __launch_bounds__(1024) void k_scan_parts(uint64_t* __restrict__ part, uint32_t nparts, uint64_t* __restrict__ total) {
    __shared__ uint64_t s[1024];
    const uint32_t per = (nparts + 1023) / 1024;
    const uint32_t a = threadIdx.x * per, b = min(nparts, a + per);
    uint64_t sum = 0;
    for (uint32_t i = a; i < b; ++i) sum += part[i];
    s[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const uint64_t v = threadIdx.x >= uint32_t(off) ? s[threadIdx.x - off] : 0ull;
        __syncthreads();
        s[threadIdx.x] += v;
        __syncthreads();
    }
    uint64_t run = s[threadIdx.x] - sum;
    for (uint32_t i = a; i < b; ++i) {
        const uint64_t v = part[i];
        part[i] = run;
        run += v;
    }
    if (threadIdx.x == 1023) *total = s[1023];
}
__global__ __launch_bounds__(SCAN_T) void k_scan_add(uint64_t* __restrict__ out, uint32_t n, const uint64_t* __restrict__ part) {
    const uint64_t add = part[blockIdx.x];
    if (add == 0) return;
    const size_t t0 = size_t(blockIdx.x) * SCAN_TILE;
    for (int k = threadIdx.x; k < SCAN_TILE; k += SCAN_T)
        if (t0 + k < n) out[t0 + k] += add;
}

// The camera matrices of raw/math/mat.hpp:7-38 on the host, as the reference computes them.
struct M4 {
    float m[16]{};
};
M4 mul(const M4& a, const M4& b) {
    M4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            float s = 0;
            for (int k = 0; k < 4; ++k) s += a.m[i * 4 + k] * b.m[k * 4 + j];
            r.m[i * 4 + j] = s;
        }
    return r;
}
M4 perspective(float fovy, float aspect, float n, float f) {
    const float t = 1.0f / std::tan(fovy * 0.5f);
    M4 r;
    r.m[0] = t / aspect;
    r.m[5] = t;
    r.m[10] = (f + n) / (n - f);
    r.m[11] = (2 * f * n) / (n - f);
    r.m[14] = -1;
    return r;
}
M4 look_at(V3 eye, V3 center, V3 up) {
    const V3 f = normalize(sub(center, eye)), s = normalize(cross(f, up)), u = cross(s, f);
    M4 r;
    r.m[0] = s.x;
    r.m[1] = s.y;
    r.m[2] = s.z;
    r.m[3] = -dot(s, eye);
    r.m[4] = u.x;
    r.m[5] = u.y;
    r.m[6] = u.z;
    r.m[7] = -dot(u, eye);
    r.m[8] = -f.x;
    r.m[9] = -f.y;
    r.m[10] = -f.z;
    r.m[11] = dot(f, eye);
    r.m[15] = 1;
    return r;
}

}  // namespace

void view_proj(const float eye[3], const float center[3], const float up[3], float fovy, float aspect, float nearz, float farz,
               float out[16]) {
    const M4 vp = mul(perspective(fovy, aspect, nearz, farz),
                      look_at(v3(eye[0], eye[1], eye[2]), v3(center[0], center[1], center[2]), v3(up[0], up[1], up[2])));
    for (int k = 0; k < 16; ++k) out[k] = vp.m[k];
}

int rasterize(Ctx& c, const float* d_vpos, const float* d_vnrm, const int32_t* d_idx, uint32_t ntri, const float eye[3],
              const float center[3], const float up[3], float fovy, float aspect, float nearz, float farz, int w, int h,
              float* d_depth, float* d_pos, float* d_nrm, uint8_t* d_mask, hipStream_t s, const uint32_t* d_mesh_end,
              const float* d_albedo, uint32_t nmesh, float* d_albedo_out) {
    const uint32_t npix = uint32_t(w) * uint32_t(h);
    M4 vp;
    view_proj(eye, center, up, fovy, aspect, nearz, farz, vp.m);
    const uint32_t nsub = 2 * ntri;
    if (!ensure(c, c.r_keys, 8 * size_t(npix)) || !ensure(c, c.r_tris, sizeof(RTri) * std::max<size_t>(1, nsub)) ||
        !ensure(c, c.r_counts, 8 * (size_t(nsub) + 1)) || !ensure(c, c.r_offs, 8 * (size_t(nsub) + 1)))
        return -2;
    Cam cam;
    for (int k = 0; k < 16; ++k) cam.m[k] = vp.m[k];
    k_fill_keys<<<std::min<uint32_t>(4096, (npix + RBT - 1) / RBT), RBT, 0, s>>>(c.r_keys.as<unsigned long long>(), npix);
    if (ntri) {
        k_rast_setup<<<(ntri + RBT - 1) / RBT, RBT, 0, s>>>(d_vpos, d_vnrm, d_idx, ntri, cam, nearz, w, h,
                                                           c.r_tris.as<RTri>(), c.r_counts.as<uint64_t>());
        // offs = the exclusive prefix sum of the counts; offs[nsub] = the candidate total, also the
        // search's sentinel
        const uint32_t ntiles = uint32_t((size_t(nsub) + SCAN_TILE - 1) / SCAN_TILE);
        if (!ensure(c, c.r_scan_tmp, 8 * size_t(ntiles))) return -2;
        k_scan_tiles<<<ntiles, SCAN_T, 0, s>>>(c.r_counts.as<uint64_t>(), c.r_offs.as<uint64_t>(), nsub, c.r_scan_tmp.as<uint64_t>());
        k_scan_parts<<<1, 1024, 0, s>>>(c.r_scan_tmp.as<uint64_t>(), ntiles, c.r_offs.as<uint64_t>() + nsub);
        k_scan_add<<<ntiles, SCAN_T, 0, s>>>(c.r_offs.as<uint64_t>(), nsub, c.r_scan_tmp.as<uint64_t>());
        k_rast_cover<<<8192, RBT, 0, s>>>(c.r_tris.as<RTri>(), c.r_offs.as<uint64_t>(), nsub, c.r_offs.as<uint64_t>() + nsub, w,
                                         c.r_keys.as<unsigned long long>());
    }
    k_rast_resolve<<<(npix + RBT - 1) / RBT, RBT, 0, s>>>(c.r_tris.as<RTri>(), c.r_keys.as<unsigned long long>(), npix, w, d_depth,
                                                         d_pos, d_nrm, d_mask, d_mesh_end, d_albedo, nmesh,
                                                         nmesh ? d_albedo_out : nullptr);
    return hip_ok(c, hipGetLastError(), "rasterize launch") ? 0 : -2;
}

}  // namespace raw_ao
