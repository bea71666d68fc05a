// The default frame on the GPU: the passes of renderWithParams (src/renderer/render.cpp:71-105)
// in its order, each bit for bit the reference's:
//   raster      rasterize() with albedo (raster.hip)
//   rtao        LinearAccel::build's triangle list (src/renderer/accel.cpp:13-25), the PLOC
//               build, then computeRTVisibility (lighting model 2) or computeRTAO (model 1)
//   ssao        computeSSAO
//   reconcile   the error map and maximum on the GPU; the float64 row-major sum on the host
//   shade       shadeV2 / shade and shadeHDRV2 / shadeHDR (src/renderer/composite.cpp:8-65),
//               with maxRadiance (composite.cpp:82-88) reduced on the way
//   motion      computeMotion (src/renderer/motion.cpp:7-31) and the covered-pixel count
#include "../../include/raw_ao_hip.h"
#include "ao_math.hpp"
#include "ctx.hpp"
#include <cmath>
#include <cstring>

namespace raw_ao {
namespace {

constexpr int FBT = 256;

struct Mat {   // a view-projection, row-major, passed by value
    float m[16];
};

__global__ __launch_bounds__(FBT) void k_gather_tris(const float* __restrict__ vpos, const int32_t* __restrict__ idx, uint32_t ntri,
                                                     float* __restrict__ tris9) {
    const uint32_t i = blockIdx.x * FBT + threadIdx.x;
    if (i >= ntri) return;
    for (int k = 0; k < 3; ++k) {
        const size_t v = size_t(idx[3 * size_t(i) + k]);
        for (int d = 0; d < 3; ++d) tris9[9 * size_t(i) + 3 * k + d] = vpos[3 * v + d];
    }
}

// projectToUV (src/renderer/motion.cpp:7-17)
__device__ bool project_uv(const Mat& vp, V3 p, float& u, float& v) {
    const float* m = vp.m;
    const float cx = m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3] * 1.0f;
    const float cy = m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7] * 1.0f;
    const float cw = m[12] * p.x + m[13] * p.y + m[14] * p.z + m[15] * 1.0f;
    if (cw <= 1e-6f) return false;
    const float invw = 1.0f / cw;
    const float nx = cx * invw, ny = cy * invw;
    u = nx * 0.5f + 0.5f;
    v = 1.0f - (ny * 0.5f + 0.5f);
    return true;
}

__device__ bool finite(float x) { return (__float_as_uint(x) & 0x7f800000u) != 0x7f800000u; }

// counts[0] += valid motion vectors, counts[1] += covered pixels (one atomic a wave).
__global__ __launch_bounds__(FBT) void k_motion(const float* __restrict__ pos, const uint8_t* __restrict__ mask, uint32_t npix,
                                                const Mat cur, const Mat prev, float* __restrict__ motion, uint32_t* __restrict__ counts) {
    __shared__ uint32_t red[2][FBT / 64];
    uint32_t nv = 0, nc = 0;
    for (size_t i = size_t(blockIdx.x) * FBT + threadIdx.x; i < npix; i += size_t(gridDim.x) * FBT) {
        float mx = 0.0f, my = 0.0f;
        const bool covered = mask[i] != 0;
        if (covered) {
            const V3 p = v3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
            float cu, cv, pu, pv;
            if (project_uv(cur, p, cu, cv) && project_uv(prev, p, pu, pv)) {
                const float dx = cu - pu, dy = cv - pv;
                if (finite(dx) && finite(dy)) {
                    mx = dx;
                    my = dy;
                    ++nv;
                }
            }
            ++nc;
        }
        motion[2 * i] = mx;
        motion[2 * i + 1] = my;
    }
    for (int o = 32; o > 0; o >>= 1) {
        nv += __shfl_xor(nv, o);
        nc += __shfl_xor(nc, o);
    }
    if ((threadIdx.x & 63u) == 0u) {
        red[0][threadIdx.x >> 6] = nv;
        red[1][threadIdx.x >> 6] = nc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t v = 0, c = 0;
        for (int k = 0; k < FBT / 64; ++k) {
            v += red[0][k];
            c += red[1][k];
        }
        if (v) atomicAdd(&counts[0], v);
        if (c) atomicAdd(&counts[1], c);
    }
}

// Model 2 (radianceV2, composite.cpp:18-25): lit = 0.2 sky + max(0, n.-l) li sun.
// Model 1 (radianceAt, composite.cpp:8-17): lit = (0.2 + max(0, n.-l) li) ao.
// frame = the radiance clamped to [0,1] per channel, hdr = unclamped; 0 off the mask.
// maxbits: the bits of maxRadiance(hdr), which starts at 0 (non-negative floats order as their bits).
__global__ __launch_bounds__(FBT) void k_shade(const float* __restrict__ nrm, const float* __restrict__ alb,
                                               const uint8_t* __restrict__ mask, uint32_t npix, int model,
                                               const float* __restrict__ plane_a, const float* __restrict__ plane_sun,
                                               float lx, float ly, float lz, float li, float* __restrict__ frame,
                                               float* __restrict__ hdr, uint32_t* __restrict__ maxbits) {
    __shared__ uint32_t red[FBT / 64];
    uint32_t mb = 0;
    for (size_t i = size_t(blockIdx.x) * FBT + threadIdx.x; i < npix; i += size_t(gridDim.x) * FBT) {
        float r[3] = {0.0f, 0.0f, 0.0f};
        if (mask[i]) {
            const float nx = nrm[3 * size_t(i)], ny = nrm[3 * size_t(i) + 1], nz = nrm[3 * size_t(i) + 2];
            const float d = nx * (lx * -1.0f) + ny * (ly * -1.0f) + nz * (lz * -1.0f);
            const float ndl = (0.0f < d ? d : 0.0f) * li;
            float lit;
            if (model == 2) lit = 0.2f * plane_a[i] + ndl * plane_sun[i];
            else lit = (0.2f + ndl) * plane_a[i];
            for (int k = 0; k < 3; ++k) r[k] = alb[3 * size_t(i) + k] * lit;
            // std::max(c.x, std::max(c.y, c.z)), then std::max(m, that) from m = 0
            const float yz = r[1] < r[2] ? r[2] : r[1];
            const float xyz = r[0] < yz ? yz : r[0];
            if (0.0f < xyz && __float_as_uint(xyz) > mb) mb = __float_as_uint(xyz);
        }
        for (int k = 0; k < 3; ++k) {
            hdr[3 * size_t(i) + k] = r[k];
            const float v = r[k];
            frame[3 * size_t(i) + k] = v < 0.0f ? 0.0f : (1.0f < v ? 1.0f : v);   // std::clamp(v, 0, 1)
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t other = __shfl_xor(mb, o);
        mb = other > mb ? other : mb;
    }
    if ((threadIdx.x & 63u) == 0u) red[threadIdx.x >> 6] = mb;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t m = red[0];
        for (int k = 1; k < FBT / 64; ++k) m = red[k] > m ? red[k] : m;
        if (m) atomicMax(maxbits, m);
    }
}

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
    for (size_t i = size_t(blockIdx.x) * FBT + threadIdx.x; i < n; i += size_t(gridDim.x) * FBT) p[i] = v;
}

uint32_t blocks(size_t n) { return uint32_t(std::min<size_t>((n + FBT - 1) / FBT, 1u << 20)); }
// The per-pixel passes that add counts or a maximum (k_shade, k_motion) run at most this many
// blocks, each striding over the pixels and adding its result once: one atomic a wave on one
// address over a 1080p frame (32k waves) serializes to ~0.6 ms.
uint32_t reduce_blocks(size_t n) { return std::min<uint32_t>(blocks(n), 1024u); }

}  // namespace

int render_frame(Ctx& c, const raw_ao_scene& sc, const raw_ao_render_opts& o, int w, int h, const raw_ao_frame& d,
                 raw_ao_frame_stats* stats, hipStream_t s) {
    const uint32_t npix = uint32_t(w) * uint32_t(h);
    const uint32_t nmesh = sc.nmesh;
    // the mesh table (host arrays), then the frame's device scalars: motion counts, the
    // radiance maximum's bits, the error maximum
    if (!ensure(c, c.fr_mesh, 16 * size_t(nmesh) + 16) || !ensure(c, c.fr_scalars, 64)) return -2;
    uint32_t* d_mesh_end = c.fr_mesh.as<uint32_t>();
    float* d_albedo = reinterpret_cast<float*>(d_mesh_end + nmesh);
    uint32_t* sc_dev = c.fr_scalars.as<uint32_t>();
    (void)hipEventRecord(c.ev[8], s);
    if (nmesh && (!hip_ok(c, hipMemcpyAsync(d_mesh_end, sc.mesh_end, 4 * size_t(nmesh), hipMemcpyHostToDevice, s), "mesh table") ||
                  !hip_ok(c, hipMemcpyAsync(d_albedo, sc.albedo, 12 * size_t(nmesh), hipMemcpyHostToDevice, s), "mesh table")))
        return -2;
    k_fill_u32<<<1, FBT, 0, s>>>(sc_dev, 16, 0u);

    // raster and ssao, on the context's second stream beside the build (they share no buffer):
    // forked from s after the mesh table's copies, joined back into s before the first pass that
    // reads the G-buffer, and before any return, so all of the frame stays ordered on s for the
    // caller (the default frame at 1080p, 1M triangles: 9.14 -> 8.79 ms with the raster beside
    // the build)
    if (!c.aux && (!hip_ok(c, hipStreamCreate(&c.aux), "hipStreamCreate") ||
                   !hip_ok(c, hipEventCreateWithFlags(&c.aux_fork, hipEventDisableTiming), "hipEventCreate") ||
                   !hip_ok(c, hipEventCreateWithFlags(&c.aux_join, hipEventDisableTiming), "hipEventCreate")))
        return -2;
    if (!hip_ok(c, hipEventRecord(c.aux_fork, s), "fork") || !hip_ok(c, hipStreamWaitEvent(c.aux, c.aux_fork, 0), "fork"))
        return -2;
    int r = rasterize(c, sc.vpos, sc.vnrm, sc.idx, sc.ntri, sc.eye, sc.center, sc.up, sc.fovy, sc.aspect, sc.near_z, sc.far_z, w, h,
                      d.depth, d.position, d.normal, d.mask, c.aux, d_mesh_end, d_albedo, nmesh, d.albedo);
    // ssao, which reads only the G-buffer, behind the raster on the same stream
    if (r == 0) r = ssao(c, d.position, d.normal, d.mask, w, h, o.ss_samples, o.radius, d.ao_ss, c.aux);
    bool joined = false;
    const auto join = [&]() -> bool {   // s waits for the raster (once)
        if (!joined) joined = hip_ok(c, hipEventRecord(c.aux_join, c.aux), "join") && hip_ok(c, hipStreamWaitEvent(s, c.aux_join, 0), "join");
        return joined;
    };
    if (r) {
        join();
        return r;
    }

    // rtao: the accelerator over every mesh's triangles in scene order, then the rays
    const float ldir[3] = {sc.has_light ? sc.light_dir[0] : 0.0f, sc.has_light ? sc.light_dir[1] : -1.0f,
                           sc.has_light ? sc.light_dir[2] : 0.0f};
    const float li = sc.has_light ? sc.light_intensity : 1.0f;
    const int model = o.rtao ? o.lighting : 1;
    if (o.rtao) {
        if (sc.ntri) {
            if (!ensure(c, c.tris_in, 36 * size_t(sc.ntri))) {
                join();
                return -2;
            }
            k_gather_tris<<<blocks(sc.ntri), FBT, 0, s>>>(sc.vpos, sc.idx, sc.ntri, c.tris_in.as<float>());
        }
        r = ploc_build(c, c.tris_in.as<float>(), sc.ntri, o.bvh_radius, s);
        if (r == 0) r = bvh_layout(c, s);
        if (r == 0 && c.trace_kernel != RAW_AO_TRACE_BINARY) r = wide_build(c, s);
        c.built_kernel = c.trace_kernel;
        if (r == 0) r = build_finish(c, s);
        if (r) {
            c.root = -1;
            join();
            return r;
        }
        if (!join()) return -2;
        if (model == 2) {
            const float to_light[3] = {ldir[0] * -1.0f, ldir[1] * -1.0f, ldir[2] * -1.0f};
            r = trace_vis4(c, d.position, d.normal, d.mask, w, h, o.rt_samples, o.radius, to_light, o.no_sun, d.ao_rt, d.sky, d.sun,
                           s);
        } else if (c.trace_kernel == RAW_AO_TRACE_BINARY) {
            r = trace_ao(c, d.position, d.normal, d.mask, w, h, o.rt_samples, o.radius, d.ao_rt, s);
        } else {
            r = trace_ao4(c, d.position, d.normal, d.mask, w, h, o.rt_samples, o.radius, d.ao_rt, s);
        }
        if (r) return r;
    }
    if (!join()) return -2;
    // reconcile: error map and maximum (sc_dev[2])
    float* d_max = reinterpret_cast<float*>(sc_dev + 2);
    if (o.rtao) {
        r = error_map(c, d.ao_ss, d.ao_rt, d.mask, w, h, d.error, d_max, s);
        if (r) return r;
    } else {
        k_fill_u32<<<std::min<uint32_t>(4096, blocks(npix)), FBT, 0, s>>>(reinterpret_cast<uint32_t*>(d.error), npix, 0u);
    }
    // shade + shade_hdr, maxRadiance in sc_dev[3]
    const float* plane_a = model == 2 ? d.sky : (o.rtao ? d.ao_rt : d.ao_ss);
    k_shade<<<reduce_blocks(npix), FBT, 0, s>>>(d.normal, d.albedo, d.mask, npix, model, plane_a, d.sun, ldir[0], ldir[1], ldir[2], li,
                                         d.frame, d.hdr, sc_dev + 3);
    // motion against the previous view-projection (this frame's when none is given)
    Mat cur, prev;
    view_proj(sc.eye, sc.center, sc.up, sc.fovy, sc.aspect, sc.near_z, sc.far_z, cur.m);
    prev = cur;
    if (o.prev_viewproj) std::memcpy(prev.m, o.prev_viewproj, 64);
    k_motion<<<reduce_blocks(npix), FBT, 0, s>>>(d.position, d.mask, npix, cur, prev, d.motion, sc_dev);
    (void)hipEventRecord(c.ev[9], s);
    if (!hip_ok(c, hipGetLastError(), "frame launch")) return -2;

    // read back the error map and the scalars; the reconcile's float64 sum in row-major order
    // (pixels off the mask hold +0 and leave the sum unchanged)
    if (!c.fr_host || c.fr_host_cap < size_t(npix)) {
        if (c.fr_host) (void)hipHostFree(c.fr_host);
        c.fr_host = nullptr;
        c.fr_host_cap = 0;
        if (!hip_ok(c, hipHostMalloc(reinterpret_cast<void**>(&c.fr_host), 4 * size_t(npix) + 64), "pinned error map")) return -2;
        c.fr_host_cap = npix;
    }
    uint32_t* hs = reinterpret_cast<uint32_t*>(c.fr_host + npix);
    if ((o.rtao && !hip_ok(c, hipMemcpyAsync(c.fr_host, d.error, 4 * size_t(npix), hipMemcpyDeviceToHost, s), "error map")) ||
        !hip_ok(c, hipMemcpyAsync(hs, sc_dev, 16, hipMemcpyDeviceToHost, s), "frame scalars") ||
        !hip_ok(c, hipStreamSynchronize(s), "frame"))
        return -2;
    if (stats) {
        std::memset(stats, 0, sizeof *stats);
        stats->motion_valid = int32_t(hs[0]);
        stats->motion_total = int32_t(hs[1]);
        float mr, me;
        std::memcpy(&mr, &hs[3], 4);
        std::memcpy(&me, &hs[2], 4);
        stats->max_radiance = mr;
        if (o.rtao) {
            double sum = 0;
            for (uint32_t i = 0; i < npix; ++i) sum += double(c.fr_host[i]) * c.fr_host[i];
            const int32_t n = int32_t(hs[1]);
            stats->pixels = n;
            stats->rmse = n > 0 ? float(std::sqrt(sum / n)) : 0.0f;
            stats->max_error = me;
            stats->within_tolerance = stats->rmse <= o.tolerance;
        }
        std::memcpy(stats->viewproj, cur.m, 64);
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, c.ev[8], c.ev[9]) == hipSuccess) stats->gpu_ms = ms;
        else (void)hipGetLastError();
    }
    return 0;
}

}  // namespace raw_ao
