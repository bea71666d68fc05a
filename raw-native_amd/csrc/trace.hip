// Ray-traced AO on gfx950: the raw-rt-ao-v1 estimator of computeRTAO
// (src/renderer/ray_ao.cpp:20-58) through the device-built PLOC BVH.
//
// Traversal form (built by k_layout from the binary PLOC tree):
//   node  64 B = 4 x float4: child 0 box, child 1 box (lo xyz hi xyz each), then the two
//         child references (>= 0 an inner node, < 0 ~ a leaf slot). Both boxes arrive in
//         one fetch, so a visit tests the two children together (Aila and Laine 2009).
//   tri   48 B = 3 x float4: a xyz, b xyz, c xyz and the input index, in leaf order
//         (written with the PLOC leaves).
// Boxes are padded outward, and the slab exit and the ray's tMax widened, so the box
// test only ever adds visits: an any-hit answer is the triangle test's alone, i.e. the
// scan of every triangle (accel.cpp:30-35) whatever the tree.
//
// One lane traces one ray; the 64 lanes of a wave are the 64 samples of one pixel when
// spp is 64, so the pixel's open count is one ballot.
#include "ao_math.hpp"
#include "ctx.hpp"
#include <algorithm>
#include <cfloat>

namespace raw_ao {
namespace {

constexpr int BT = 256;
constexpr int STACK = 64;
constexpr float kExit = 1.00001f;       // slab exit widening (the boxes are also padded)
constexpr float kReach = 1.0001f;       // box test reach beyond tMax

struct alignas(16) Node { float4 a, b, c; int4 r; };

__device__ __forceinline__ float pad_of(float l, float h) {
    return (h - l) * 6.103515625e-05f + (fabsf(l) + fabsf(h)) * 9.5367431640625e-07f + 1e-30f;   // 2^-14, 2^-20
}

__global__ __launch_bounds__(BT) void k_layout(const float* __restrict__ boxes, const int32_t* __restrict__ links, uint32_t n,
                                               int32_t root, Node* __restrict__ nodes) {
    const uint32_t i = blockIdx.x * BT + threadIdx.x;
    const uint32_t inner = n > 1 ? n - 1 : 1;
    if (i >= inner) return;
    int32_t ch[2];
    if (n == 1) { ch[0] = 0; ch[1] = -1; }   // a single leaf under a one-record root
    else { ch[0] = links[3 * size_t(n + i)]; ch[1] = links[3 * size_t(n + i) + 1]; }
    float bb[2][6];
    int32_t ref[2];
    for (int k = 0; k < 2; ++k) {
        if (ch[k] < 0) {   // empty slot: a box no ray enters
            for (int a = 0; a < 3; ++a) { bb[k][a] = 1e30f; bb[k][3 + a] = -1e30f; }
            ref[k] = ~int32_t(0x7fffffff);
            continue;
        }
        const float* b = boxes + 6 * size_t(ch[k]);
        for (int a = 0; a < 3; ++a) {
            const float p = pad_of(b[a], b[3 + a]);
            bb[k][a] = b[a] - p;
            bb[k][3 + a] = b[3 + a] + p;
        }
        ref[k] = uint32_t(ch[k]) < n ? ~ch[k] : int32_t(ch[k] - int32_t(n));
    }
    Node r;
    r.a = make_float4(bb[0][0], bb[0][1], bb[0][2], bb[0][3]);
    r.b = make_float4(bb[0][4], bb[0][5], bb[1][0], bb[1][1]);
    r.c = make_float4(bb[1][2], bb[1][3], bb[1][4], bb[1][5]);
    r.r = make_int4(ref[0], ref[1], 0, 0);
    nodes[i] = r;
    (void)root;
}

struct Ray {
    V3 o, d, inv, oinv;
    RayFrame f;
};
__device__ __forceinline__ Ray make_ray(V3 o, V3 d) {
    Ray r;
    r.o = o;
    r.d = d;
    r.inv = v3(d.x != 0.0f ? 1.0f / d.x : copysignf(1e30f, d.x), d.y != 0.0f ? 1.0f / d.y : copysignf(1e30f, d.y),
               d.z != 0.0f ? 1.0f / d.z : copysignf(1e30f, d.z));
    r.oinv = v3(o.x * r.inv.x, o.y * r.inv.y, o.z * r.inv.z);
    r.f = ray_frame(d);
    return r;
}
// Conservative slab test of one box against [0, reach].
__device__ __forceinline__ bool slab(const Ray& r, float lx, float ly, float lz, float hx, float hy, float hz, float reach,
                                     float exit_scale = kExit) {
    const float x0 = fmaf(lx, r.inv.x, -r.oinv.x), x1 = fmaf(hx, r.inv.x, -r.oinv.x);
    const float y0 = fmaf(ly, r.inv.y, -r.oinv.y), y1 = fmaf(hy, r.inv.y, -r.oinv.y);
    const float z0 = fmaf(lz, r.inv.z, -r.oinv.z), z1 = fmaf(hz, r.inv.z, -r.oinv.z);
    const float t0 = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), 0.0f));
    const float t1 = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), reach));
    return t0 <= t1 * exit_scale;
}
__device__ __forceinline__ bool tri_hit(const Ray& r, const float4* __restrict__ ts, uint32_t slot, float& t) {
    const float4 p0 = ts[3 * size_t(slot)], p1 = ts[3 * size_t(slot) + 1], p2 = ts[3 * size_t(slot) + 2];
    return intersect_t(r.o, r.d, r.f, v3(p0.x, p0.y, p0.z), v3(p0.w, p1.x, p1.y), v3(p1.z, p1.w, p2.x), t);
}

// Any hit with t < tmax: true when some triangle's test accepts it.
template <bool COUNT>
__device__ bool occluded(const Ray& r, float tmax, const Node* __restrict__ nodes, const float4* __restrict__ ts,
                         uint32_t root, uint32_t ntri, uint32_t& visits, uint32_t& tests, bool& overflow) {
    const float reach = tmax * kReach;
    uint32_t stk[STACK];
    int sp = 0;
    int32_t node = int32_t(root);
    // A well-formed tree has ntri - 1 inner records; more visits than that means a
    // malformed one: stop and let the caller answer the ray by the scan.
    for (uint32_t guard = 0;; ++guard) {
        if (guard > ntri) { overflow = true; return false; }
        const Node nd = nodes[node];
        if (COUNT) ++visits;
        const bool h0 = slab(r, nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y, reach);
        const bool h1 = slab(r, nd.b.z, nd.b.w, nd.c.x, nd.c.y, nd.c.z, nd.c.w, reach);
        int32_t next = -1;
        int32_t refs[2] = {nd.r.x, nd.r.y};
        bool hs[2] = {h0, h1};
        for (int k = 0; k < 2; ++k) {
            if (!hs[k]) continue;
            const int32_t c = refs[k];
            if (c < 0) {
                float t;
                if (COUNT) ++tests;
                if (uint32_t(~c) < ntri && tri_hit(r, ts, uint32_t(~c), t) && t < tmax) return true;
            } else if (next < 0) {
                next = c;
            } else {
                if (sp == STACK) { overflow = true; return false; }
                stk[sp++] = uint32_t(c);
            }
        }
        if (next >= 0) { node = next; continue; }
        if (sp == 0) return false;
        node = int32_t(stk[--sp]);
    }
}

// The exact answer for a ray whose stack overflowed: every triangle (accel.cpp:30-35).
__device__ bool scan_occluded(const Ray& r, float tmax, const float4* __restrict__ ts, uint32_t ntri) {
    for (uint32_t i = 0; i < ntri; ++i) {
        float t;
        if (tri_hit(r, ts, i, t) && t < tmax) return true;
    }
    return false;
}

struct TraceArgs {
    const float* pos;
    const float* nrm;
    const uint8_t* mask;
    float* ao;
    uint32_t* cnt;          // per-pixel open counts (general spp)
    const Node* nodes;
    const float4* tris;
    const uint8_t* trig;    // host-libm table or null
    TraceCounters* counters;
    uint32_t w, npix, spp, ntri, root;
    float radius;
};

template <bool SPP64, bool COUNT>
__global__ __launch_bounds__(BT) void k_trace_ao(TraceArgs a) {
    const uint64_t gid = uint64_t(blockIdx.x) * BT + threadIdx.x;
    const uint32_t pix = SPP64 ? uint32_t(gid >> 6) : uint32_t(gid / a.spp);
    if (pix >= a.npix) return;
    const uint32_t s = uint32_t(gid - uint64_t(pix) * a.spp);
    if (!a.mask[pix]) {
        if (s == 0) a.ao[pix] = 1.0f;
        return;
    }
    const V3 p = v3(a.pos[3 * size_t(pix)], a.pos[3 * size_t(pix) + 1], a.pos[3 * size_t(pix) + 2]);
    const V3 n = v3(a.nrm[3 * size_t(pix)], a.nrm[3 * size_t(pix) + 1], a.nrm[3 * size_t(pix) + 2]);
    const uint32_t y = pix / a.w, x = pix - y * a.w;
    V3 t, b;
    basis(n, t, b);
    const V3 origin = add(p, scl(n, 1e-3f));
    const uint32_t k1 = hash24(x, y, 2 * s), k2 = hash24(x, y, 2 * s + 1);
    const float u1 = float(k1) / float(0x1000000), u2 = float(k2) / float(0x1000000);
    float sphi, cphi;
    sincos_base(6.2831853f * u2, sphi, cphi);
    if (a.trig) trig_fix(a.trig, k2, sphi, cphi);
    const V3 dir = ao_dir(u1, sphi, cphi, n, t, b);
    bool occ = false;
    uint32_t visits = 0, tests = 0;
    if (a.ntri) {
        const Ray r = make_ray(origin, dir);
        bool overflow = false;
        occ = occluded<COUNT>(r, a.radius, a.nodes, a.tris, a.root, a.ntri, visits, tests, overflow);
        if (overflow) {
            occ = scan_occluded(r, a.radius, a.tris, a.ntri);
            atomicAdd(&a.counters->overflows, 1ull);
        }
    }
    if (COUNT) {
        atomicAdd(&a.counters->visits, (unsigned long long)visits);
        atomicAdd(&a.counters->tests, (unsigned long long)tests);
        atomicAdd(&a.counters->rays, 1ull);
    }
    if (SPP64) {
        const unsigned long long open = __ballot(!occ);
        if (s == 0) a.ao[pix] = float(__popcll(open)) / 64.0f;
    } else if (!occ) {
        atomicAdd(&a.cnt[pix], 1u);
    }
}

__global__ __launch_bounds__(BT) void k_ao_finish(const uint8_t* __restrict__ mask, const uint32_t* __restrict__ cnt,
                                                  uint32_t npix, uint32_t spp, float* __restrict__ ao) {
    const uint32_t i = blockIdx.x * BT + threadIdx.x;
    if (i >= npix) return;
    ao[i] = mask[i] ? float(cnt[i]) / float(spp) : 1.0f;
}

// ---- general rays (traceGpu, src/renderer/gpu/rt_gpu.cpp:97-133) ----------------------
__global__ __launch_bounds__(BT) void k_trace_rays(const float* __restrict__ rays8, uint32_t nrays, int any, float far_scale,
                                                   const Node* __restrict__ nodes, const float4* __restrict__ ts,
                                                   uint32_t root, uint32_t ntri, float* __restrict__ out4) {
    const uint32_t k = blockIdx.x * BT + threadIdx.x;
    if (k >= nrays) return;
    const float* q = rays8 + 8 * size_t(k);
    const Ray r = make_ray(v3(q[0], q[1], q[2]), v3(q[3], q[4], q[5]));
    const float tmax = q[6];
    // the caller's exit factor, never tighter than the AO path's own widening unless it asks
    const float exit_scale = far_scale >= 1.0f ? fmaxf(far_scale, kExit) : far_scale;
    float best = tmax;
    int32_t bi = -1;
    uint32_t visits = 0, tests = 0;
    bool overflow = false;
    if (ntri) {
        uint32_t stk[STACK];
        int sp = 0;
        int32_t node = int32_t(root);
        while (node >= 0) {
            if (visits > ntri) { overflow = true; break; }   // malformed tree (see occluded())
            const Node nd = nodes[node];
            ++visits;
            const float reach = best * kReach;
            int32_t next = -1;
            const int32_t refs[2] = {nd.r.x, nd.r.y};
            const bool hs[2] = {slab(r, nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y, reach, exit_scale),
                                slab(r, nd.b.z, nd.b.w, nd.c.x, nd.c.y, nd.c.z, nd.c.w, reach, exit_scale)};
            bool done = false;
            for (int c = 0; c < 2; ++c) {
                if (!hs[c]) continue;
                const int32_t ch = refs[c];
                if (ch < 0) {
                    const uint32_t slot = uint32_t(~ch);
                    if (slot >= ntri) continue;
                    float t;
                    ++tests;
                    if (!tri_hit(r, ts, slot, t)) continue;
                    const int32_t id = __float_as_int(ts[3 * size_t(slot) + 2].y);
                    if (t < best || (bi >= 0 && t == best && id < bi)) {
                        if (!(t < tmax)) continue;
                        best = t;
                        bi = id;
                        if (any) { done = true; break; }
                    }
                } else if (next < 0) {
                    next = ch;
                } else if (sp == STACK) {
                    overflow = true;
                    done = true;
                    break;
                } else {
                    stk[sp++] = uint32_t(ch);
                }
            }
            if (done) break;
            if (next >= 0) { node = next; continue; }
            node = sp ? int32_t(stk[--sp]) : -1;
        }
    }
    if (overflow) { best = -1.0f; bi = -2; }
    out4[4 * size_t(k)] = bi >= 0 ? best : -1.0f;
    out4[4 * size_t(k) + 1] = float(bi);
    out4[4 * size_t(k) + 2] = float(visits);
    out4[4 * size_t(k) + 3] = float(tests);
}

// sin and cos of every sample angle phi = 6.2831853f * k / 2^24 as the tracer uses them.
__global__ __launch_bounds__(BT) void k_angles(const uint8_t* __restrict__ trig, float* __restrict__ so, float* __restrict__ co) {
    const uint32_t k = blockIdx.x * BT + threadIdx.x;
    if (k >= (1u << 24)) return;
    float s, c;
    sincos_base(6.2831853f * (float(k) / float(0x1000000)), s, c);
    if (trig) trig_fix(trig, k, s, c);
    so[k] = s;
    co[k] = c;
}

uint32_t cdiv(uint64_t a, uint32_t b) { return uint32_t((a + b - 1) / b); }

}  // namespace

int ao_finish(Ctx& c, const uint8_t* d_mask, const uint32_t* d_cnt, uint32_t npix, uint32_t spp, float* d_ao, hipStream_t s) {
    k_ao_finish<<<cdiv(npix, BT), BT, 0, s>>>(d_mask, d_cnt, npix, spp, d_ao);
    return hip_ok(c, hipGetLastError(), "finish launch") ? 0 : -2;
}

int angle_table(Ctx& c, float* d_sin, float* d_cos, hipStream_t s) {
    k_angles<<<(1u << 24) / BT, BT, 0, s>>>(c.host_trig ? c.d_trig : nullptr, d_sin, d_cos);
    return hip_ok(c, hipGetLastError(), "angles launch") ? 0 : -2;
}

int bvh_layout(Ctx& c, hipStream_t s) {
    const uint32_t n = c.ntri;
    if (n == 0) return 0;
    const size_t inner = n > 1 ? n - 1 : 1;
    // the binary records only for the binary traversal; the leaf-order triangles, which the wide
    // one reads too, were written with the PLOC leaves
    const bool binary = c.trace_kernel == RAW_AO_TRACE_BINARY;
    c.trav_root = n > 1 ? uint32_t(c.root) - n : 0u;
    if (!binary) return 0;
    if (!ensure(c, c.nodes, sizeof(Node) * inner)) return -2;
    k_layout<<<cdiv(inner, BT), BT, 0, s>>>(c.boxes.as<float>(), c.links.as<int32_t>(), n, c.root, c.nodes.as<Node>());
    return hip_ok(c, hipGetLastError(), "layout launch") ? 0 : -2;
}

int trace_ao(Ctx& c, const float* d_pos, const float* d_nrm, const uint8_t* d_mask, int w, int h, int spp, float radius,
             float* d_ao, hipStream_t s) {
    const uint32_t npix = uint32_t(w) * uint32_t(h);
    if (!ensure(c, c.counters, sizeof(TraceCounters))) return -2;
    if (!hip_ok(c, hipMemsetAsync(c.counters.p, 0, sizeof(TraceCounters), s), "counter clear")) return -2;
    TraceArgs a{d_pos, d_nrm, d_mask, d_ao, nullptr, c.nodes.as<Node>(), c.tris_sorted.as<float4>(),
                c.host_trig ? c.d_trig : nullptr, c.counters.as<TraceCounters>(), uint32_t(w), npix, uint32_t(spp),
                c.ntri, c.trav_root, radius};
    const bool spp64 = spp == 64;
    if (!spp64) {
        if (!ensure(c, c.st_cnt, 4 * size_t(npix))) return -2;
        a.cnt = c.st_cnt.as<uint32_t>();
        if (!hip_ok(c, hipMemsetAsync(a.cnt, 0, 4 * size_t(npix), s), "count clear")) return -2;
    }
    const uint32_t blocks = cdiv(uint64_t(npix) * uint64_t(spp), BT);
    const bool count = c.count_next;
    (void)hipEventRecord(c.ev[2], s);
    if (spp64) {
        if (count) k_trace_ao<true, true><<<blocks, BT, 0, s>>>(a);
        else k_trace_ao<true, false><<<blocks, BT, 0, s>>>(a);
    } else {
        if (count) k_trace_ao<false, true><<<blocks, BT, 0, s>>>(a);
        else k_trace_ao<false, false><<<blocks, BT, 0, s>>>(a);
    }
    (void)hipEventRecord(c.ev[3], s);
    if (!spp64) k_ao_finish<<<cdiv(npix, BT), BT, 0, s>>>(d_mask, a.cnt, npix, uint32_t(spp), d_ao);
    if (!hip_ok(c, hipGetLastError(), "trace launch")) return -2;
    c.have_counts = count;
    return 0;
}

int trace_rays(Ctx& c, const float* d_rays, uint32_t n, int any, float far_scale, float* d_out, hipStream_t s) {
    if (n == 0) return 0;
    k_trace_rays<<<cdiv(n, BT), BT, 0, s>>>(d_rays, n, any, far_scale, c.nodes.as<Node>(), c.tris_sorted.as<float4>(),
                                             c.trav_root, c.ntri, d_out);
    return hip_ok(c, hipGetLastError(), "trace_rays launch") ? 0 : -2;
}

}  // namespace raw_ao
