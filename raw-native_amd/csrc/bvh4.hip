// The 4-wide traversal form of the PLOC tree and the AO / general-ray kernels over it.
//
// Collapse (top down, one level of the wide tree a step, deterministic): a wide node starts
// from an inner node of the binary tree with its two children as slots, then twice opens
// the inner slot of largest surface area into its two children (ties to the lower slot),
// so it has 2 to 4 slots. Inner slots become the next level's wide nodes; ids come from
// prefix sums over the level in slot order, so the layout never depends on scheduling.
//
// Record (128 B, 8 x 16 B): slot boxes as structure-of-arrays (lo x of the 4 slots, lo y,
// lo z, hi x, hi y, hi z), padded outward as in the binary form; then the 4 slot refs
// (>= 0 a wide node, < 0 ~ a leaf slot = sorted triangle position, EMPTY unused); then
// 16 B spare. One lane reads 7 x 16 B and tests 4 boxes.
//
// Traversal (one ray a lane): the per-lane stack's top STACK_LDS entries live in LDS,
// deeper entries in a private spill array. A node step records the leaf slots its ray
// hits as pending; a lane with pending leaves does no node step until the wave tests
// them together, in a triangle phase the wave enters when at least a quarter of its live
// lanes wait on leaves or none can step, so node and triangle code are not interleaved
// lane by lane.
#include "ao_math.hpp"
#include "batch.hpp"
#include "ctx.hpp"
#include <algorithm>
#include <cfloat>

namespace raw_ao {
namespace {

constexpr int BT = 256;
constexpr int32_t EMPTY = int32_t(0x80000000);
constexpr float kExit = 1.00001f;
constexpr float kReach = 1.0001f;
constexpr float kFar = 3.0e38f;           // box of an empty slot: a point no finite reach gets to
constexpr int STACK_LDS = 16;
constexpr int STACK_SPILL = 48;

struct alignas(16) Node4 { float4 lx, ly, lz, hx, hy, hz; int4 ref; float4 spare; };

struct WideState { uint32_t count, base, next_base, pad; };   // pad: levels run with a frontier

__device__ __forceinline__ float pad_of(float l, float h) {
    return (h - l) * 6.103515625e-05f + (fabsf(l) + fabsf(h)) * 9.5367431640625e-07f + 1e-30f;   // 2^-14, 2^-20
}
__device__ __forceinline__ float box_area(const float* b) {
    const float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
    return 2.0f * (dy * (dx + dz) + dz * dx);
}

// Up to 4 slots (binary node ids) of the wide node rooted at binary inner node f.
__device__ __forceinline__ int expand(const float* __restrict__ boxes, const int32_t* __restrict__ links, uint32_t n,
                                      int32_t f, int32_t slot[4]) {
    slot[0] = links[3 * size_t(f)];
    slot[1] = links[3 * size_t(f) + 1];
    int cnt = 2;
    for (int e = 0; e < 2; ++e) {
        int best = -1;
        float ba = -1.0f;
        for (int k = 0; k < cnt; ++k) {
            if (uint32_t(slot[k]) < n) continue;   // a leaf
            const float a = box_area(boxes + 6 * size_t(slot[k]));
            if (best < 0 || a > ba) { best = k; ba = a; }
        }
        if (best < 0) break;
        const int32_t x = slot[best];
        for (int k = cnt; k > best + 1; --k) slot[k] = slot[k - 1];
        slot[best] = links[3 * size_t(x)];
        slot[best + 1] = links[3 * size_t(x) + 1];
        ++cnt;
    }
    return cnt;
}

// The wide record of a frontier entry at index rec (slots from expand, cnt of them): slot boxes
// padded outward (pad_of), leaf slots as ~triangle position, inner slots as references to the
// next level's records from first_next on, their binary ids appended at next[0..].
__device__ __forceinline__ void write_record(const float* __restrict__ boxes, uint32_t n, const int32_t (&slot)[4], int cnt,
                                             uint32_t rec, uint32_t first_next, uint32_t* __restrict__ next,
                                             Node4* __restrict__ nodes) {
    uint32_t off = 0;
    float bl[3][4], bh[3][4];
    int32_t ref[4];
    uint32_t masks = 0;   // leaf slots in bits 0-3, inner slots in bits 4-7
    for (int k = 0; k < 4; ++k) {
        if (k >= cnt) {
            for (int a = 0; a < 3; ++a) { bl[a][k] = kFar; bh[a][k] = kFar; }
            ref[k] = EMPTY;
            continue;
        }
        masks |= uint32_t(slot[k]) < n ? (1u << k) : (16u << k);
        const float* b = boxes + 6 * size_t(slot[k]);
        for (int a = 0; a < 3; ++a) {
            const float pd = pad_of(b[a], b[3 + a]);
            bl[a][k] = b[a] - pd;
            bh[a][k] = b[3 + a] + pd;
        }
        if (uint32_t(slot[k]) < n) {
            ref[k] = ~slot[k];
        } else {
            ref[k] = int32_t(first_next + off);
            next[off] = uint32_t(slot[k]);
            ++off;
        }
    }
    Node4 r;
    r.lx = make_float4(bl[0][0], bl[0][1], bl[0][2], bl[0][3]);
    r.ly = make_float4(bl[1][0], bl[1][1], bl[1][2], bl[1][3]);
    r.lz = make_float4(bl[2][0], bl[2][1], bl[2][2], bl[2][3]);
    r.hx = make_float4(bh[0][0], bh[0][1], bh[0][2], bh[0][3]);
    r.hy = make_float4(bh[1][0], bh[1][1], bh[1][2], bh[1][3]);
    r.hz = make_float4(bh[2][0], bh[2][1], bh[2][2], bh[2][3]);
    r.ref = make_int4(ref[0], ref[1], ref[2], ref[3]);
    r.spare = make_float4(__uint_as_float(masks), 0, 0, 0);
    nodes[rec] = r;
}

// Pass 1 of a level: each wide node's inner-slot count, summed per block.
// Two launches a level, both with a fixed grid that strides over the tiles of BT frontier
// entries and reads the frontier's size from the device state, so the host enqueues levels
// without reading the size back; a level with an empty frontier does nothing (its state is
// copied).
// Pass 1: each tile's inner-slot count (blk[tile]).
__global__ __launch_bounds__(BT) void k_wide_count(const WideState* __restrict__ st, const uint32_t* __restrict__ frontier,
                                                   const float* __restrict__ boxes, const int32_t* __restrict__ links,
                                                   uint32_t n, uint32_t* __restrict__ blk, uint32_t defer_max) {
    __shared__ uint32_t ws[BT / 64];
    const uint32_t count = st->count;
    if (count <= defer_max) return;   // left to k_wide_small
    const uint32_t ntile = (count + BT - 1) / BT;
    for (uint32_t tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
        __syncthreads();   // the previous tile's ws reads are done
        const uint32_t p = tile * BT + threadIdx.x;
        uint32_t c = 0;
        if (p < count) {
            int32_t slot[4];
            const int cnt = expand(boxes, links, n, int32_t(frontier[p]), slot);
            for (int k = 0; k < cnt; ++k) c += uint32_t(slot[k]) >= n;
        }
        for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
        if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = c;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t s = 0;
            for (int k = 0; k < BT / 64; ++k) s += ws[k];
            blk[tile] = s;
        }
    }
}

// Pass 2: each tile's offset among all tiles (the sum of the earlier tiles' counts, read by the
// tile's own workgroup), then the records of this level and the next level's frontier in entry
// order; the last tile writes the next level's state.
__global__ __launch_bounds__(BT) void k_wide_write(const WideState* __restrict__ st, WideState* __restrict__ st_next,
                                                   const uint32_t* __restrict__ frontier, uint32_t* __restrict__ next,
                                                   const uint32_t* __restrict__ blk, const float* __restrict__ boxes,
                                                   const int32_t* __restrict__ links, uint32_t n, Node4* __restrict__ nodes,
                                                   uint32_t defer_max, uint32_t* __restrict__ hist) {
    __shared__ uint32_t ws[BT / 64], red[BT / 64];
    const WideState s = *st;
    if (s.count <= defer_max) {   // empty, or left to k_wide_small: the state and frontier move on unchanged
        if (blockIdx.x == 0 && threadIdx.x == 0) *st_next = s;
        for (uint32_t i = blockIdx.x * BT + threadIdx.x; i < s.count; i += gridDim.x * BT) next[i] = frontier[i];
        return;
    }
    const uint32_t ntile = (s.count + BT - 1) / BT;
    for (uint32_t tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
        // the tile's offset: the sum of blk[0 .. tile-1] over the workgroup
        uint32_t pre = 0;
        for (uint32_t i = threadIdx.x; i < tile; i += BT) pre += blk[i];
        for (int off = 32; off > 0; off >>= 1) pre += __shfl_xor(pre, off);
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        __syncthreads();   // the previous tile's ws and red reads are done
        if (lane == 0) red[w] = pre;
        const uint32_t p = tile * BT + threadIdx.x;
        int32_t slot[4] = {EMPTY, EMPTY, EMPTY, EMPTY};
        int cnt = 0;
        uint32_t c = 0;
        if (p < s.count) {
            cnt = expand(boxes, links, n, int32_t(frontier[p]), slot);
            for (int k = 0; k < cnt; ++k) c += uint32_t(slot[k]) >= n;
        }
        uint32_t x = c;   // inclusive wave scan of the inner-slot counts
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
        if (lane == 63) ws[w] = x;
        __syncthreads();
        pre = 0;
        for (int k = 0; k < BT / 64; ++k) pre += red[k];
        if (tile == ntile - 1 && threadIdx.x == 0) {   // the last tile: the next level's state
            WideState o;
            o.count = pre + blk[tile];
            o.base = s.base + s.count;
            o.next_base = o.base + o.count;
            o.pad = s.pad + 1u;   // levels that had a frontier
            *st_next = o;
            if (s.pad < uint32_t(kWideHist)) hist[s.pad] = s.count;   // this grid level's frontier, for the next build
        }
        if (p >= s.count) continue;
        uint32_t off = pre + x - c;
        for (int k = 0; k < w; ++k) off += ws[k];
        write_record(boxes, n, slot, cnt, s.base + p, s.next_base + off, next + off, nodes);
    }
}

// The levels of a small frontier (at most SMALL_W entries) in one workgroup: the top of the tree,
// where the frontier grows from the root, and its bottom, where it shrinks again. One entry a
// thread, the next frontier's offsets by a workgroup scan, the frontier kept in LDS between
// levels; the records are the grid kernels' (write_record) and land where theirs would. It stops
// at an empty frontier or one wider than SMALL_W (left to the grid levels) and writes the state
// and frontier back in place. The state's pad (levels the grid kernels ran) is left as it is.
constexpr int SMALL_W = 1024;
__global__ __launch_bounds__(SMALL_W) void k_wide_small(WideState* __restrict__ st, uint32_t* __restrict__ frontier,
                                                        const float* __restrict__ boxes, const int32_t* __restrict__ links,
                                                        uint32_t n, Node4* __restrict__ nodes) {
    __shared__ uint32_t cur[SMALL_W], nxt[4 * SMALL_W];
    __shared__ uint32_t wsum[SMALL_W / 64];
    WideState s = *st;
    if (s.count == 0 || s.count > uint32_t(SMALL_W)) return;
    const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
    if (t < s.count) cur[t] = frontier[t];
    __syncthreads();
    while (s.count > 0 && s.count <= uint32_t(SMALL_W)) {
        int32_t slot[4] = {EMPTY, EMPTY, EMPTY, EMPTY};
        int cnt = 0;
        uint32_t c = 0;
        if (t < s.count) {
            cnt = expand(boxes, links, n, int32_t(cur[t]), slot);
            for (int k = 0; k < cnt; ++k) c += uint32_t(slot[k]) >= n;
        }
        uint32_t x = c;   // inclusive wave scan
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
        if (lane == 63) wsum[w] = x;
        __syncthreads();
        uint32_t off = x - c, total = 0;
        for (uint32_t k = 0; k < SMALL_W / 64; ++k) {
            if (k < w) off += wsum[k];
            total += wsum[k];
        }
        if (t < s.count) write_record(boxes, n, slot, cnt, s.base + t, s.next_base + off, nxt + off, nodes);
        __syncthreads();   // nxt written, wsum and cur read
        WideState o;
        o.count = total;
        o.base = s.base + s.count;
        o.next_base = o.base + total;
        o.pad = s.pad;
        s = o;
        if (s.count <= uint32_t(SMALL_W)) {
            if (t < s.count) cur[t] = nxt[t];
        } else {
            for (uint32_t i = t; i < s.count; i += SMALL_W) frontier[i] = nxt[i];
        }
        __syncthreads();
    }
    if (s.count <= uint32_t(SMALL_W) && t < s.count) frontier[t] = cur[t];
    if (t == 0) *st = s;
}

// A single-triangle scene: one record, the leaf in slot 0.
__global__ void k_wide_single(const float* __restrict__ boxes, Node4* __restrict__ nodes) {
    if (threadIdx.x != 0) return;
    Node4 r;
    float l[3], h[3];
    for (int a = 0; a < 3; ++a) {
        const float pd = pad_of(boxes[a], boxes[3 + a]);
        l[a] = boxes[a] - pd;
        h[a] = boxes[3 + a] + pd;
    }
    r.lx = make_float4(l[0], kFar, kFar, kFar);
    r.ly = make_float4(l[1], kFar, kFar, kFar);
    r.lz = make_float4(l[2], kFar, kFar, kFar);
    r.hx = make_float4(h[0], kFar, kFar, kFar);
    r.hy = make_float4(h[1], kFar, kFar, kFar);
    r.hz = make_float4(h[2], kFar, kFar, kFar);
    r.ref = make_int4(~0, EMPTY, EMPTY, EMPTY);
    r.spare = make_float4(__uint_as_float(1u), 0, 0, 0);
    nodes[0] = r;
}

// ---- quantized records ---------------------------------------------------------------
// Record (64 B, 4 x 16 B), the same node ids as the 128 B records: per axis a base b and a
// power-of-two step s, each slot's lo and hi plane as a byte q (plane = b + q s); the masks
// word; the slot refs. Words: {b x, b y, b z, exps} {q lo x, q lo y, q lo z, q hi x} {q hi y,
// q hi z, masks, 0} {refs}; exps holds the biased exponents of the x, y and z steps in bytes
// 0-2 and, in byte 3, 15 for a node whose planes could not be quantized (every slot entered).
// A q word holds slot k's byte at bits 8k.
//
// The builder takes b a multiple of s and keeps |b / s| + 255 below 2^24, so b + q s is an
// integer times s below 2^24 in magnitude: a float, and fmaf(q, s, b) yields it exactly. It
// rounds each slot's lo plane down and hi plane up to that grid, so every quantized box
// contains the slot's 128 B box. The slab test's plane distances fmaf(plane, 1/d, -o/d) are
// monotone in the plane, so a containing box gives a near distance no larger and a far one no
// smaller: every slot the 128 B record admits, the 64 B one admits too. Visits can only be
// added, and the any-hit answer, which the triangle tests give, stays the same.
struct alignas(16) Node4Q { float bx, by, bz; uint32_t exps; uint32_t qlx, qly, qlz, qhx, qhy, qhz, masks, zero; int4 ref; };
static_assert(sizeof(Node4Q) == 64, "a quantized record is 64 bytes");

// One axis of a node over its slots `valid`: the step's biased exponent e, the base, each slot's
// bytes; false when no step works (a plane infinite or NaN).
__device__ bool quantize_axis(const float (&lo)[4], const float (&hi)[4], uint32_t valid, float& base, uint32_t& e_out,
                              uint32_t& qlo, uint32_t& qhi) {
    float mn = INFINITY, mx = -INFINITY;
    for (int k = 0; k < 4; ++k)
        if ((valid >> k) & 1u) {
            mn = fminf(mn, lo[k]);
            mx = fmaxf(mx, hi[k]);
        }
    if (!(mn <= mx) || !isfinite(mn) || !isfinite(mx)) return false;
    for (int k = 0; k < 4; ++k)
        if (((valid >> k) & 1u) && !(isfinite(lo[k]) && isfinite(hi[k]) && lo[k] <= hi[k])) return false;
    // the smallest step that spans the slots in 255 steps with the grid indices below 2^24 - 256
    const float need = fmaxf((mx - mn) * (1.0f / 253.0f), fmaxf(fabsf(mn), fabsf(mx)) * 0x1p-23f);
    int e = need > 0.0f ? ilogbf(need) + 127 : 1;
    e = e < 1 ? 1 : (e > 254 ? 254 : e);
    for (; e <= 254; ++e) {
        const float st = __uint_as_float(uint32_t(e) << 23);
        // floor and ceil of plane / step on the grid; ldexpf is exact but for a subnormal
        // quotient, which the products below catch
        float fl = floorf(ldexpf(mn, 127 - e));
        if (fl * st > mn) fl -= 1.0f;
        float ch = ceilf(ldexpf(mx, 127 - e));
        if (ch * st < mx) ch += 1.0f;
        if (!(fabsf(fl) <= 16776960.0f && fabsf(ch) <= 16776960.0f && ch - fl <= 255.0f && isfinite(fl * st) && isfinite(ch * st)))
            continue;
        qlo = qhi = 0;
        for (int k = 0; k < 4; ++k) {
            uint32_t a = 255u, b = 0u;   // an unused slot: lo above hi, never entered
            if ((valid >> k) & 1u) {
                float fa = floorf(ldexpf(lo[k], 127 - e));
                if (fa * st > lo[k]) fa -= 1.0f;
                float fb = ceilf(ldexpf(hi[k], 127 - e));
                if (fb * st < hi[k]) fb += 1.0f;
                a = uint32_t(fmaxf(fa, fl) - fl);
                b = uint32_t(fminf(fb, ch) - fl);
            }
            qlo |= a << (8 * k);
            qhi |= b << (8 * k);
        }
        base = fl * st;
        e_out = uint32_t(e);
        return true;
    }
    return false;
}

// The 64 B records of the first *count (or nmax, without count) 128 B records, one a thread.
__global__ __launch_bounds__(BT) void k_quantize4(const Node4* __restrict__ nodes, const WideState* __restrict__ st, uint32_t nmax,
                                                  Node4Q* __restrict__ out) {
    const uint32_t n = st ? min(st->base, nmax) : nmax;
    for (uint32_t i = blockIdx.x * BT + threadIdx.x; i < n; i += gridDim.x * BT) {
        const Node4 nd = nodes[i];
        const uint32_t masks = __float_as_uint(nd.spare.x);
        const uint32_t valid = (masks | (masks >> 4)) & 15u;
        const float lo[3][4] = {{nd.lx.x, nd.lx.y, nd.lx.z, nd.lx.w}, {nd.ly.x, nd.ly.y, nd.ly.z, nd.ly.w}, {nd.lz.x, nd.lz.y, nd.lz.z, nd.lz.w}};
        const float hi[3][4] = {{nd.hx.x, nd.hx.y, nd.hx.z, nd.hx.w}, {nd.hy.x, nd.hy.y, nd.hy.z, nd.hy.w}, {nd.hz.x, nd.hz.y, nd.hz.z, nd.hz.w}};
        Node4Q q;
        float b[3];
        uint32_t e[3], ql[3], qh[3];
        bool ok = true;
        for (int a = 0; a < 3; ++a) ok = ok && quantize_axis(lo[a], hi[a], valid, b[a], e[a], ql[a], qh[a]);
        if (!ok)
            for (int a = 0; a < 3; ++a) { b[a] = 0.0f; e[a] = 127u; ql[a] = 0u; qh[a] = 0u; }
        q.bx = b[0]; q.by = b[1]; q.bz = b[2];
        q.exps = e[0] | (e[1] << 8) | (e[2] << 16) | (ok ? 0u : 15u << 24);
        q.qlx = ql[0]; q.qly = ql[1]; q.qlz = ql[2];
        q.qhx = qh[0]; q.qhy = qh[1]; q.qhz = qh[2];
        q.masks = masks;
        q.zero = 0u;
        q.ref = nd.ref;
        out[i] = q;
    }
}

// ---- traversal -------------------------------------------------------------------
struct Ray {
    V3 o, d, inv, oinv;
    RayFrame f;
};
// The slab tests' 1/d: the hardware reciprocal (within 1 ulp of the correctly rounded quotient),
// +-1e30 for d = 0. The slab test only decides which boxes a ray visits, never a hit: its plane
// distances come out within a relative 2^-22 of those from the exact quotient, far inside the
// 1e-5 the exit is widened by and the boxes' padding, so every box the exact test admits is
// still admitted and the any-hit answer (the triangle test's) is unchanged. An overflow to
// infinity (|d| below 2^-128) makes that axis' distances infinite or NaN, which the min/max
// chain drops, leaving the other axes to decide: again no box is lost.
__device__ __forceinline__ float slab_inv(float d) {
    return d != 0.0f ? __builtin_amdgcn_rcpf(d) : copysignf(1e30f, d);
}
__device__ __forceinline__ Ray make_ray(V3 o, V3 d) {
    Ray r;
    r.o = o;
    r.d = d;
    r.inv = v3(slab_inv(d.x), slab_inv(d.y), slab_inv(d.z));
    r.oinv = v3(o.x * r.inv.x, o.y * r.inv.y, o.z * r.inv.z);
    r.f = ray_frame(d);
    return r;
}
__device__ __forceinline__ bool slab1(const Ray& r, float lx, float ly, float lz, float hx, float hy, float hz, float reach,
                                      float exit_scale) {
    const float x0 = fmaf(lx, r.inv.x, -r.oinv.x), x1 = fmaf(hx, r.inv.x, -r.oinv.x);
    const float y0 = fmaf(ly, r.inv.y, -r.oinv.y), y1 = fmaf(hy, r.inv.y, -r.oinv.y);
    const float z0 = fmaf(lz, r.inv.z, -r.oinv.z), z1 = fmaf(hz, r.inv.z, -r.oinv.z);
    const float t0 = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), 0.0f));
    const float t1 = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), reach));
    return t0 <= t1 * exit_scale;
}
// The 4 slots' hit bits (bit k: slot k's box is entered within reach).
__device__ __forceinline__ uint32_t slab4(const Ray& r, const Node4& nd, float reach, float exit_scale) {
    uint32_t m = 0;
    m |= slab1(r, nd.lx.x, nd.ly.x, nd.lz.x, nd.hx.x, nd.hy.x, nd.hz.x, reach, exit_scale) ? 1u : 0u;
    m |= slab1(r, nd.lx.y, nd.ly.y, nd.lz.y, nd.hx.y, nd.hy.y, nd.hz.y, reach, exit_scale) ? 2u : 0u;
    m |= slab1(r, nd.lx.z, nd.ly.z, nd.lz.z, nd.hx.z, nd.hy.z, nd.hz.z, reach, exit_scale) ? 4u : 0u;
    m |= slab1(r, nd.lx.w, nd.ly.w, nd.lz.w, nd.hx.w, nd.hy.w, nd.hz.w, reach, exit_scale) ? 8u : 0u;
    return m;
}
// The same test with each axis' near and far planes picked by the ray's direction signs
// (near = lo where 1/d >= 0, else hi): for a box with lo <= hi these are exactly the min
// and max of the two plane distances above, so the hit bits are the same, with five
// fewer operations a box. `np` is the node as 8 float4 (lo x, lo y, lo z, hi x, hi y,
// hi z, refs, masks); ox/oy/oz are 0 or 3 by the signs.
__device__ __forceinline__ float4 fma4(float4 p, float s, float c) {
    return make_float4(fmaf(p.x, s, c), fmaf(p.y, s, c), fmaf(p.z, s, c), fmaf(p.w, s, c));
}
__device__ __forceinline__ float4 ld4(const char* __restrict__ base, uint32_t off) {
    return *reinterpret_cast<const float4*>(base + off);
}
// base: the node array; off: the node's byte offset; ox/oy/oz: 0 or 48 (byte offset of
// the far planes' float4 from the near ones' when 1/d < 0).
__device__ __forceinline__ uint32_t slab4s(const Ray& r, const char* __restrict__ base, uint32_t off, uint32_t ox, uint32_t oy,
                                           uint32_t oz, float reach, float exit_scale, float tlo = 0.0f) {
    // 32-bit offsets from the (uniform) base and base + 16, base + 32: one add an address
    const uint32_t o48 = off + 48u;
    const float4 nx = fma4(ld4(base, off + ox), r.inv.x, -r.oinv.x), fx = fma4(ld4(base, o48 - ox), r.inv.x, -r.oinv.x);
    const float4 ny = fma4(ld4(base + 16, off + oy), r.inv.y, -r.oinv.y), fy = fma4(ld4(base + 16, o48 - oy), r.inv.y, -r.oinv.y);
    const float4 nz = fma4(ld4(base + 32, off + oz), r.inv.z, -r.oinv.z), fz = fma4(ld4(base + 32, o48 - oz), r.inv.z, -r.oinv.z);
    const auto one = [&](float a, float b, float c, float d, float e, float f) {
        const float t0 = fmaxf(fmaxf(a, b), fmaxf(c, tlo));
        const float t1 = fminf(fminf(d, e), fminf(f, reach));
        return t0 <= t1 * exit_scale;
    };
    return (one(nx.x, ny.x, nz.x, fx.x, fy.x, fz.x) ? 1u : 0u) | (one(nx.y, ny.y, nz.y, fx.y, fy.y, fz.y) ? 2u : 0u) |
           (one(nx.z, ny.z, nz.z, fx.z, fy.z, fz.z) ? 4u : 0u) | (one(nx.w, ny.w, nz.w, fx.w, fy.w, fz.w) ? 8u : 0u);
}
// slab4s that also returns each slot's entry distance.
__device__ __forceinline__ uint32_t slab4s_t(const Ray& r, const char* __restrict__ base, uint32_t off, uint32_t ox, uint32_t oy,
                                             uint32_t oz, float reach, float exit_scale, float (&t0s)[4], float tlo = 0.0f) {
    // 32-bit offsets from the (uniform) base and base + 16, base + 32: one add an address
    const uint32_t o48 = off + 48u;
    const float4 nx = fma4(ld4(base, off + ox), r.inv.x, -r.oinv.x), fx = fma4(ld4(base, o48 - ox), r.inv.x, -r.oinv.x);
    const float4 ny = fma4(ld4(base + 16, off + oy), r.inv.y, -r.oinv.y), fy = fma4(ld4(base + 16, o48 - oy), r.inv.y, -r.oinv.y);
    const float4 nz = fma4(ld4(base + 32, off + oz), r.inv.z, -r.oinv.z), fz = fma4(ld4(base + 32, o48 - oz), r.inv.z, -r.oinv.z);
    const auto one = [&](float a, float b, float c, float d, float e, float f, float& t0) {
        t0 = fmaxf(fmaxf(a, b), fmaxf(c, tlo));
        const float t1 = fminf(fminf(d, e), fminf(f, reach));
        return t0 <= t1 * exit_scale;
    };
    return (one(nx.x, ny.x, nz.x, fx.x, fy.x, fz.x, t0s[0]) ? 1u : 0u) | (one(nx.y, ny.y, nz.y, fx.y, fy.y, fz.y, t0s[1]) ? 2u : 0u) |
           (one(nx.z, ny.z, nz.z, fx.z, fy.z, fz.z, t0s[2]) ? 4u : 0u) | (one(nx.w, ny.w, nz.w, fx.w, fy.w, fz.w, t0s[3]) ? 8u : 0u);
}
// slab4s_t on a 64 B record at byte offset off of qbase: the same hit bits but for the slots the
// quantized boxes add (see Node4Q), each slot's entry distance; also the refs and masks word.
__device__ __forceinline__ uint32_t slab4q(const Ray& r, const char* __restrict__ qbase, uint32_t off, uint32_t ox, uint32_t oy,
                                           uint32_t oz, float reach, float exit_scale, int4& ref, uint32_t& masks, float (&t0s)[4],
                                           float tlo = 0.0f) {
    const uint4 w0 = *reinterpret_cast<const uint4*>(qbase + off);
    const uint4 w1 = *reinterpret_cast<const uint4*>(qbase + off + 16);
    const uint4 w2 = *reinterpret_cast<const uint4*>(qbase + off + 32);
    ref = *reinterpret_cast<const int4*>(qbase + off + 48);
    masks = w2.z;
    const float sx = __uint_as_float((w0.w & 255u) << 23), sy = __uint_as_float(((w0.w >> 8) & 255u) << 23),
                sz = __uint_as_float(((w0.w >> 16) & 255u) << 23);
    const float bx = __uint_as_float(w0.x), by = __uint_as_float(w0.y), bz = __uint_as_float(w0.z);
    // near planes: lo where 1/d >= 0 (offset 0), else hi
    const uint32_t nxw = ox ? w1.w : w1.x, fxw = ox ? w1.x : w1.w;
    const uint32_t nyw = oy ? w2.x : w1.y, fyw = oy ? w1.y : w2.x;
    const uint32_t nzw = oz ? w2.y : w1.z, fzw = oz ? w1.z : w2.y;
    uint32_t m = w0.w >> 24;   // 15: every slot entered
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const auto pl = [&](uint32_t w, float st, float b) { return fmaf(float((w >> (8 * k)) & 255u), st, b); };
        const float nx = fmaf(pl(nxw, sx, bx), r.inv.x, -r.oinv.x), fx = fmaf(pl(fxw, sx, bx), r.inv.x, -r.oinv.x);
        const float ny = fmaf(pl(nyw, sy, by), r.inv.y, -r.oinv.y), fy = fmaf(pl(fyw, sy, by), r.inv.y, -r.oinv.y);
        const float nz = fmaf(pl(nzw, sz, bz), r.inv.z, -r.oinv.z), fz = fmaf(pl(fzw, sz, bz), r.inv.z, -r.oinv.z);
        const float t0 = fmaxf(fmaxf(nx, ny), fmaxf(nz, tlo));
        const float t1 = fminf(fminf(fx, fy), fminf(fz, reach));
        t0s[k] = t0;
        if (t0 <= t1 * exit_scale) m |= 1u << k;
    }
    return m;
}
__device__ __forceinline__ int32_t ref_of(const int4& r, int k) {
    const int32_t lo = (k & 1) ? r.y : r.x, hi = (k & 1) ? r.w : r.z;
    return (k & 2) ? hi : lo;
}
template <bool SEL = true>
__device__ __forceinline__ bool tri_hit(const Ray& r, const float4* __restrict__ ts, uint32_t slot, float& t) {
    const float4 p0 = ts[3 * size_t(slot)], p1 = ts[3 * size_t(slot) + 1], p2 = ts[3 * size_t(slot) + 2];
    return intersect_t<SEL>(r.o, r.d, r.f, v3(p0.x, p0.y, p0.z), v3(p0.w, p1.x, p1.y), v3(p1.z, p1.w, p2.x), t);
}

constexpr int CUT_N = 32, CUT_LEAF = 16;
constexpr int CUT_WAVE = 7 * (CUT_N + CUT_LEAF);   // LDS words a wave: node ids and boxes, leaf slots and boxes
// A wave's cut of the tree (wave_cut) in LDS.
struct CutView {
    uint32_t* id;      // CUT_N node indices
    float* box;        // 6 rows of CUT_N: lo x, y, z, hi x, y, z
    uint32_t* leaf;    // CUT_LEAF sorted triangle positions
    float* lbox;       // 6 rows of CUT_LEAF
    int n, nleaf;      // entries; n < 0: no cut (start from the root)
    __device__ static bool enters(const Ray& r, const float* bx, int stride, int k, float reach, float exit_scale) {
        const float lx = bx[k], ly = bx[stride + k], lz = bx[2 * stride + k];
        const float hx = bx[3 * stride + k], hy = bx[4 * stride + k], hz = bx[5 * stride + k];
        const bool px = r.inv.x >= 0.0f, py = r.inv.y >= 0.0f, pz = r.inv.z >= 0.0f;
        // the same operations as slab4s on one slot
        const float nx = fmaf(px ? lx : hx, r.inv.x, -r.oinv.x), fx = fmaf(px ? hx : lx, r.inv.x, -r.oinv.x);
        const float ny = fmaf(py ? ly : hy, r.inv.y, -r.oinv.y), fy = fmaf(py ? hy : ly, r.inv.y, -r.oinv.y);
        const float nz = fmaf(pz ? lz : hz, r.inv.z, -r.oinv.z), fz = fmaf(pz ? hz : lz, r.inv.z, -r.oinv.z);
        const float t0 = fmaxf(fmaxf(nx, ny), fmaxf(nz, 0.0f));
        const float t1 = fminf(fminf(fx, fy), fminf(fz, reach));
        return t0 <= t1 * exit_scale;
    }
};

// The per-lane stack: the top STACK_LDS entries live in LDS (a column per lane, entry e
// at lds_col[e * BT]), deeper ones in a private spill array; the stack pointer stays in a
// register.
// Any hit with t < tmax through the wide tree. Returns 1 occluded, 0 open, -1 the stack
// would overflow (the caller answers that ray by the scan).
// NEAR: descend into the hit inner slot with the smallest entry distance first and stack the
// others farthest first; otherwise the first hit slot, the rest in slot order. The answer is
// the same either way.
// cut (optional, the wave's shared cut of the tree, see wave_cut): instead of the root, the
// traversal starts from the first cut node the ray enters and takes the next one whenever its
// stack runs empty.
// tlo: where the slab tests start along the ray (0, or sky_from for a sky ray).
template <bool COUNT, bool NEAR = true>
__device__ int occluded4(const Ray& r, float tmax, const Node4* __restrict__ nodes, const float4* __restrict__ ts,
                         uint32_t ntri, uint32_t* lds_col, uint32_t tri_q, uint32_t& visits, uint32_t& tests,
                         bool active = true, const CutView* cut = nullptr, float tlo = 0.0f) {
    const float reach = tmax * kReach;
    const int ncut = cut ? cut->n : -1;
    int ci = 0;
    const auto next_cut = [&]() -> int32_t {
        while (ci < ncut) {
            const int k = ci++;
            if (CutView::enters(r, cut->box, CUT_N, k, reach, kExit)) return int32_t(cut->id[k]);
        }
        return -1;
    };
    const uint32_t ox = r.inv.x >= 0.0f ? 0u : 48u, oy = r.inv.y >= 0.0f ? 0u : 48u, oz = r.inv.z >= 0.0f ? 0u : 48u;
    const char* const base = reinterpret_cast<const char*>(nodes);
    uint32_t spill[STACK_SPILL];
    int sp = 0;
    int32_t node = ncut < 0 ? 0 : (active ? next_cut() : -1);   // the wide root, or the first cut node entered
    uint32_t pend = 0;                // pending leaf slots of `leafref`
    int4 leafref = make_int4(0, 0, 0, 0);
    int result = 0;
    bool live = active;   // an inactive lane takes part in the wave's votes only
    // A lane's node steps are bounded by the tree's record count; more means a malformed
    // tree, and the lane stops with -1 (answered by the scan). Every iteration some live
    // lane steps or tests, so the loop ends.
    uint32_t work = 0;
    while (true) {
        // node step: lanes with a node and nothing pending
        if (live && pend == 0 && node >= 0) {
            const uint32_t off = uint32_t(node) * uint32_t(sizeof(Node4));
            if (COUNT) ++visits;
            float t0s[4];
            const uint32_t hit = NEAR ? slab4s_t(r, base, off, ox, oy, oz, reach, kExit, t0s, tlo)
                                      : slab4s(r, base, off, ox, oy, oz, reach, kExit, tlo);
            const int4 ref = *reinterpret_cast<const int4*>(base + off + 96);
            const uint32_t masks = *reinterpret_cast<const uint32_t*>(base + off + 112);   // leaf slots bits 0-3, inner 4-7
            pend = hit & masks & 15u;
            leafref = ref;
            uint32_t ih = hit & (masks >> 4);
            if (ih) {
                int first = __builtin_ctz(ih);
                if (NEAR) {
                    float tb = t0s[first];
                    for (int c = first + 1; c < 4; ++c)
                        if (((ih >> c) & 1u) && t0s[c] < tb) { tb = t0s[c]; first = c; }
                }
                node = ref_of(ref, first);
                ih &= ~(1u << first);
                if (NEAR && (ih & (ih - 1u))) {   // two or three more: push the farthest first, so the nearest pops first
                    uint32_t order[3];
                    float dist[3];
                    int m = 0;
                    for (int c = 0; c < 4; ++c)
                        if ((ih >> c) & 1u) { order[m] = uint32_t(ref_of(ref, c)); dist[m] = t0s[c]; ++m; }
                    for (int i = 1; i < m; ++i)   // descending distance
                        for (int j = i; j > 0 && dist[j] > dist[j - 1]; --j) {
                            const float td = dist[j]; dist[j] = dist[j - 1]; dist[j - 1] = td;
                            const uint32_t to = order[j]; order[j] = order[j - 1]; order[j - 1] = to;
                        }
                    for (int i = 0; i < m; ++i) {
                        if (sp < STACK_LDS) lds_col[sp * BT] = order[i];
                        else if (sp < STACK_LDS + STACK_SPILL) spill[sp - STACK_LDS] = order[i];
                        else { result = -1; live = false; }
                        ++sp;
                    }
                    ih = 0;
                }
                while (ih) {
                    const uint32_t v = uint32_t(ref_of(ref, __builtin_ctz(ih)));
                    ih &= ih - 1u;
                    if (sp < STACK_LDS) lds_col[sp * BT] = v;
                    else if (sp < STACK_LDS + STACK_SPILL) spill[sp - STACK_LDS] = v;
                    else { result = -1; live = false; }
                    ++sp;
                }
            } else if (sp > 0) {
                --sp;
                node = int32_t(sp < STACK_LDS ? lds_col[sp * BT] : spill[sp - STACK_LDS]);
            } else {
                node = next_cut();
            }
            if (++work > ntri + 4) { result = -1; live = false; }
        }
        const bool waiting = live && pend != 0;
        const bool can_step = live && pend == 0 && node >= 0;
        const unsigned long long wb = __ballot(waiting), sb = __ballot(can_step), lb = __ballot(live);
        if (lb == 0) break;
        // triangle step: when tri_q/64 of the live lanes wait on leaves, or none can step
        if (sb == 0 || 64u * uint32_t(__popcll(wb)) >= tri_q * uint32_t(__popcll(lb))) {
            if (waiting) {
                const int k = __builtin_ctz(pend);
                pend &= pend - 1u;
                const uint32_t slot = uint32_t(~ref_of(leafref, k));
                if (COUNT) ++tests;
                float t;
                if (slot < ntri && tri_hit(r, ts, slot, t) && t < tmax) { result = 1; live = false; }
            }
        }
        if (live && pend == 0 && node < 0) live = false;   // finished: open
    }
    return result;
}

// Where a sky ray's box tests start. An open AO ray has no triangle with t < radius, so every
// triangle its sky ray (the same origin and direction) hits lies at t >= radius, inside each of
// its ancestors' boxes, whose computed exit is then at least the radius less the slab test's
// float rounding (a few ulp of the coordinates' magnitude). This bound stays 1% of the radius
// plus 1e-4 of that magnitude below it, so boxes the ray leaves before it are skipped and the
// answer is unchanged.
__device__ __forceinline__ float sky_from(V3 o, float radius) {
    const float big = fmaxf(fabsf(o.x), fmaxf(fabsf(o.y), fabsf(o.z)));
    return fmaxf(0.0f, 0.99f * radius - 1e-4f * (big + radius));
}

__device__ bool scan_occluded(const Ray& r, float tmax, const float4* __restrict__ ts, uint32_t ntri) {
    for (uint32_t i = 0; i < ntri; ++i) {
        float t;
        if (tri_hit(r, ts, i, t) && t < tmax) return true;
    }
    return false;
}

// The wave's cut of the tree for rays that all lie in the box [lo, hi] (the 64 rays of one
// pixel: one origin, length <= reach). The wave walks the top of the tree together, one
// frontier node a lane: a slot whose box misses [lo, hi] is dropped, a leaf slot that overlaps
// goes to the cut's leaves (its sorted triangle position and box), an inner slot that overlaps
// goes to the cut's nodes once its box is no wider than `small` on every axis and to the next
// frontier otherwise. When a level would overflow a list, the level is dropped and its frontier
// becomes the rest of the cut's nodes. Every triangle a ray can hit overlaps [lo, hi], and so
// do all its ancestors, so the cut covers every triangle the traversal from the root can reach.
// Each entry keeps its box as the parent slot holds it, so a ray enters an entry exactly when
// the root traversal's test of that slot would pass (CutView::enters). Called by all 64 lanes
// of a wave together; the frontiers live in the wave's rows of the traversal stack (rows 0-6
// and 7-13: node, then lo x, y, z, hi x, y, z), which is free until the rays start.
// Orders one wave's LDS writes before its lanes read each other's entries.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// The tree's slots down to depth TOP_DEPTH (the root's children are depth 1): every slot at that
// depth and every leaf slot above it, with the box its parent holds, in breadth-first order.
// They partition the triangles, so a walk may start from them instead of the root.
constexpr int TOP_DEPTH = 3, TOP_MAX = 64;   // 4^TOP_DEPTH
struct alignas(16) TopEntry {
    float lx, ly, lz, hx, hy, hz;
    int32_t ref;   // >= 0 a node, < 0 ~ a leaf's sorted triangle position
    uint32_t pad;
};
// One wave; four lanes a node, one a slot. Level L holds at most 4^L nodes, so the next
// frontier (at most TOP_MAX / 4 before the last level) and the table (at most TOP_MAX) never
// overflow.
__global__ __launch_bounds__(64) void k_top_table(const Node4* __restrict__ nodes, TopEntry* __restrict__ out,
                                                  uint32_t* __restrict__ count) {
    __shared__ uint32_t ids[2][TOP_MAX / 4];
    const uint32_t lane = threadIdx.x;
    const unsigned long long below = (1ull << lane) - 1ull;
    const char* const base = reinterpret_cast<const char*>(nodes);
    int nf = 1, cur = 0, total = 0;
    if (lane == 0) ids[0][0] = 0u;
    wave_sync();
    for (int level = 0; level < TOP_DEPTH && nf > 0; ++level) {
        int nn = 0;
        for (int i0 = 0; i0 < nf; i0 += 16) {
            const int i = i0 + int(lane >> 2);
            const uint32_t c = lane & 3u;
            bool to_table = false, to_next = false;
            TopEntry e{0, 0, 0, 0, 0, 0, 0, 0};
            if (i < nf) {
                const uint32_t off = ids[cur][i] * uint32_t(sizeof(Node4));
                const uint32_t masks = *reinterpret_cast<const uint32_t*>(base + off + 112);
                if ((masks >> c) & 0x11u) {
                    const float* f = reinterpret_cast<const float*>(base + off);
                    e.lx = f[c];
                    e.ly = f[4 + c];
                    e.lz = f[8 + c];
                    e.hx = f[12 + c];
                    e.hy = f[16 + c];
                    e.hz = f[20 + c];
                    e.ref = reinterpret_cast<const int32_t*>(base + off + 96)[c];
                    to_table = ((masks >> c) & 1u) || level == TOP_DEPTH - 1;
                    to_next = !to_table;
                }
            }
            const unsigned long long bt = __ballot(to_table), bn = __ballot(to_next);
            if (to_table) out[total + __popcll(bt & below)] = e;
            if (to_next) ids[cur ^ 1][nn + __popcll(bn & below)] = uint32_t(e.ref);
            total += __popcll(bt);
            nn += __popcll(bn);
        }
        wave_sync();
        cur ^= 1;
        nf = nn;
    }
    if (lane == 0) *count = uint32_t(total);
}

// The walk starts from the top table when it has entries (else from the root).
template <int SB = BT>
__device__ void wave_cut(const Node4* __restrict__ nodes, const TopEntry* __restrict__ top, uint32_t ntop, V3 lo, V3 hi,
                         float small, uint32_t* rows, CutView& cv, uint32_t& steps) {
    const uint32_t lane = threadIdx.x & 63u;
    const unsigned long long below = (1ull << lane) - 1ull;
    const char* const base = reinterpret_cast<const char*>(nodes);
    // frontier f's row k, entry i: rows[(7 f + k) BT + i]
    const auto fr = [&](int f, int k) { return rows + (7 * f + k) * SB; };
    int cur = 0, nn = 0;
    cv.n = 0;
    cv.nleaf = 0;
    // files one slot (box b, ref) of a lane: dropped, a cut leaf, a cut node or the next frontier
    const auto place = [&](bool valid, const float (&b)[6], int32_t ref, bool is_leaf) {
        const bool ov = valid && b[0] <= hi.x && b[3] >= lo.x && b[1] <= hi.y && b[4] >= lo.y && b[2] <= hi.z && b[5] >= lo.z;
        const bool small_box = b[3] - b[0] <= small && b[4] - b[1] <= small && b[5] - b[2] <= small;
        const bool to_leaf = ov && is_leaf, to_cut = ov && !is_leaf && small_box, to_front = ov && !is_leaf && !small_box;
        const unsigned long long bl = __ballot(to_leaf), bc = __ballot(to_cut), bf = __ballot(to_front);
        if (to_leaf) {
            const int pos = cv.nleaf + __popcll(bl & below);
            if (pos < CUT_LEAF) {
                cv.leaf[pos] = uint32_t(~ref);
                for (int k = 0; k < 6; ++k) cv.lbox[k * CUT_LEAF + pos] = b[k];
            }
        }
        if (to_cut) {
            const int pos = cv.n + __popcll(bc & below);
            if (pos < CUT_N) {
                cv.id[pos] = uint32_t(ref);
                for (int k = 0; k < 6; ++k) cv.box[k * CUT_N + pos] = b[k];
            }
        }
        if (to_front) {
            const int pos = nn + __popcll(bf & below);
            if (pos < 64) {
                fr(cur ^ 1, 0)[pos] = uint32_t(ref);
                for (int k = 0; k < 6; ++k) reinterpret_cast<float*>(fr(cur ^ 1, k + 1))[pos] = b[k];
            }
        }
        cv.nleaf += __popcll(bl);
        cv.n += __popcll(bc);
        nn += __popcll(bf);
    };
    const auto overflow = [&]() { return cv.nleaf > CUT_LEAF || cv.n > CUT_N || nn > 64; };
    int nf;
    bool boxed;   // the frontier's entries carry their boxes (all but the root)
    if (ntop > 0) {
        for (uint32_t e0 = 0; e0 < ntop; e0 += 64) {
            const uint32_t e = e0 + lane;
            float b[6] = {0, 0, 0, 0, 0, 0};
            int32_t ref = 0;
            if (e < ntop) {
                const float4 p = *reinterpret_cast<const float4*>(top + e);
                const float4 q = *(reinterpret_cast<const float4*>(top + e) + 1);
                b[0] = p.x; b[1] = p.y; b[2] = p.z; b[3] = p.w; b[4] = q.x; b[5] = q.y;
                ref = __float_as_int(q.z);
                if ((e & 3u) == 0) ++steps;   // four 32-byte entries: one node record's bytes
            }
            place(e < ntop, b, ref, ref < 0);
        }
        wave_sync();
        if (overflow()) { cv.n = -1; cv.nleaf = 0; return; }
        cur = 1;
        nf = nn;
        boxed = true;
    } else {
        if (lane == 0) fr(0, 0)[0] = 0u;
        wave_sync();
        nf = 1;
        boxed = false;
    }
    for (int level = 0; nf > 0; ++level) {
        if (level == 64) { cv.n = -1; cv.nleaf = 0; return; }   // deeper than any tree this library builds: a malformed tree
        const int n0 = cv.n, l0 = cv.nleaf;
        nn = 0;
        // four lanes a frontier node, one a slot: 16 nodes a pass
        for (int i0 = 0; i0 < nf; i0 += 16) {
            const int i = i0 + int(lane >> 2);
            const uint32_t c = lane & 3u;
            bool valid = false, is_leaf = false;
            int32_t ref = 0;
            float b[6] = {0, 0, 0, 0, 0, 0};
            if (i < nf) {
                // the slot's words load together with the masks (one round trip; an unused slot's
                // words are read but not used)
                const uint32_t off = fr(cur, 0)[i] * uint32_t(sizeof(Node4));
                const uint32_t masks = *reinterpret_cast<const uint32_t*>(base + off + 112);   // leaf slots bits 0-3, inner 4-7
                for (int k = 0; k < 6; ++k) b[k] = *reinterpret_cast<const float*>(base + off + 16 * k + 4 * c);
                ref = *reinterpret_cast<const int32_t*>(base + off + 96 + 4 * c);
                if (c == 0) ++steps;
                is_leaf = (masks >> c) & 1u;
                valid = (masks >> c) & 0x11u;
            }
            place(valid, b, ref, is_leaf);
        }
        wave_sync();
        if (overflow()) {
            // drop this level: its frontier (with the boxes its parents gave it) ends the cut
            cv.n = n0;
            cv.nleaf = l0;
            if (!boxed || n0 + nf > CUT_N) { cv.n = -1; cv.nleaf = 0; return; }
            if (lane < uint32_t(nf)) {
                cv.id[n0 + lane] = fr(cur, 0)[lane];
                for (int k = 0; k < 6; ++k) cv.box[k * CUT_N + n0 + lane] = reinterpret_cast<const float*>(fr(cur, k + 1))[lane];
            }
            cv.n = n0 + nf;
            wave_sync();
            return;
        }
        cur ^= 1;
        nf = nn;
        boxed = true;
    }
}

struct TraceArgs4 {
    const float* pos;
    const float* nrm;
    const uint8_t* mask;
    float* ao;
    uint32_t* cnt;
    const Node4* nodes;
    const float4* tris;
    const uint8_t* trig;
    TraceCounters* counters;
    uint32_t w, npix, spp, ntri, tri_q;
    float radius;
    float cut_small;   // CUT: the wave's cut stops at boxes no wider than this
    float cut_span;    // CUT below 64 spp: no walk when the wave's union box is wider than this
    const TopEntry* top;   // CUT: the tree's top table and its entry count
    const uint32_t* ntop;
    float* sky;            // SKY: sky out (64 spp) or the unblocked counts (below)
    uint32_t* cnt_free;
    uint64_t* open_mask;   // k_trace_ao4rc<.., true> out, k_sky4r in: bit s of a pixel = its AO sample s is open
    const Node4Q* nodesq;  // Q: the 64 B records (the wave cut still walks the 128 B ones)
};

__device__ __forceinline__ float wave_min(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

// CUT: the wave first walks the top of the tree together (wave_cut) and each ray starts from
// the cut instead of the root. With 64 spp the wave is one pixel; otherwise its lanes span
// several pixels, all lanes stay for the walk, and the walk's box is the union of the live
// lanes' reach boxes (no walk when that union is wider than a.cut_span). stk: STACK_LDS x BT
// words; cutmem: CUT_WAVE words a wave.
// SKY (the visibility pass): each open AO ray is traced again to tMax 1e30 from the root (in
// the same lane, right after; its box tests start at sky_from), sky = the unblocked fraction (a.sky; counts in a.cnt_free below
// 64 spp), 1 off the mask.
template <bool SPP64, bool COUNT, bool NEAR, bool CUT, bool SKY = false>
__device__ __forceinline__ void trace_ao4(const TraceArgs4& a, uint32_t* stk, uint32_t* cutmem) {
    constexpr bool STAY = CUT && !SPP64;
    const uint64_t gid = uint64_t(blockIdx.x) * BT + threadIdx.x;
    const uint32_t pix0 = SPP64 ? uint32_t(gid >> 6) : uint32_t(gid / a.spp);
    bool live = pix0 < a.npix;
    if (!STAY && !live) return;
    const uint32_t pix = live ? pix0 : 0u;   // a lane past the end reads pixel 0 and is not live
    const uint32_t s = uint32_t(gid - uint64_t(pix0) * a.spp);
    if (live && !a.mask[pix]) {
        if (s == 0) {
            a.ao[pix] = 1.0f;
            if (SKY) a.sky[pix] = 1.0f;
        }
        live = false;
    }
    if (!STAY && !live) return;
    if (STAY && __ballot(live) == 0ull) return;   // the same in every lane of the wave
    const V3 p = v3(a.pos[3 * size_t(pix)], a.pos[3 * size_t(pix) + 1], a.pos[3 * size_t(pix) + 2]);
    const V3 n = v3(a.nrm[3 * size_t(pix)], a.nrm[3 * size_t(pix) + 1], a.nrm[3 * size_t(pix) + 2]);
    const uint32_t y = pix / a.w, x = pix - y * a.w;
    V3 t, b;
    basis(n, t, b);
    const V3 origin = add(p, scl(n, 1e-3f));
    uint32_t visits = 0, tests = 0;
    CutView cv;
    if (CUT && a.ntri) {   // before the ray's own set-up, which need not stay live across the walk
        // Every ray of the pixel stays inside origin +- radius on each axis; the margin covers
        // the slab test's reach (radius x 1.0001 x 1.00001) and float rounding near the origin.
        const float big = fmaxf(fabsf(origin.x), fmaxf(fabsf(origin.y), fabsf(origin.z)));
        const float half = a.radius * 1.001f + (big + a.radius) * 1e-5f;
        V3 lo = v3(origin.x - half, origin.y - half, origin.z - half);
        V3 hi = v3(origin.x + half, origin.y + half, origin.z + half);
        bool walk = true;
        if (!SPP64) {
            if (!live) {
                lo = v3(kFar, kFar, kFar);
                hi = v3(-kFar, -kFar, -kFar);
            }
            lo = v3(wave_min(lo.x), wave_min(lo.y), wave_min(lo.z));
            hi = v3(wave_max(hi.x), wave_max(hi.y), wave_max(hi.z));
            walk = hi.x - lo.x <= a.cut_span && hi.y - lo.y <= a.cut_span && hi.z - lo.z <= a.cut_span;
        }
        uint32_t* mine = cutmem + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) * CUT_WAVE;
        cv.id = mine;
        cv.box = reinterpret_cast<float*>(mine + CUT_N);
        cv.leaf = mine + 7 * CUT_N;
        cv.lbox = reinterpret_cast<float*>(mine + 7 * CUT_N + CUT_LEAF);
        cv.n = -1;   // no walk: every ray starts from the root
        cv.nleaf = 0;
        if (walk) {
            uint32_t steps = 0;
            wave_cut(a.nodes, a.top, *a.ntop, lo, hi, a.cut_small, stk + (threadIdx.x & ~63u), cv, steps);
            if (COUNT && live) visits += steps;
        }
    }
    const uint32_t k1 = hash24(x, y, 2 * s), k2 = hash24(x, y, 2 * s + 1);
    const float u1 = float(k1) / float(0x1000000), u2 = float(k2) / float(0x1000000);
    float sphi, cphi;
    sincos_base(6.2831853f * u2, sphi, cphi);
    if (a.trig) trig_fix(a.trig, k2, sphi, cphi);
    const V3 dir = ao_dir(u1, sphi, cphi, n, t, b);
    bool occ = false;
    if (a.ntri) {
        const Ray r = make_ray(origin, dir);
        int res;
        if (CUT) {
            bool hit = false;
            const float reach = a.radius * kReach;
            // the cut leaves this ray enters (nleaf is 0 when the cut overflowed), then their
            // triangles two at a time, both loads in flight together
            uint32_t lp = 0;
            if (live)
                for (int i = 0; i < cv.nleaf; ++i)
                    if (cv.leaf[i] < a.ntri && CutView::enters(r, cv.lbox, CUT_LEAF, i, reach, kExit)) lp |= 1u << i;
            while (lp && !hit) {
                const uint32_t s0 = cv.leaf[__builtin_ctz(lp)];
                lp &= lp - 1u;
                const bool two = lp != 0;
                const uint32_t s1 = two ? cv.leaf[__builtin_ctz(lp)] : s0;
                if (two) lp &= lp - 1u;
                const float4* q0 = a.tris + 3 * size_t(s0);
                const float4* q1 = a.tris + 3 * size_t(s1);
                const float4 a0 = q0[0], a1 = q0[1], a2 = q0[2], b0 = q1[0], b1 = q1[1], b2 = q1[2];
                if (COUNT) tests += two ? 2u : 1u;
                float t0, t1;
                hit = (intersect_t(r.o, r.d, r.f, v3(a0.x, a0.y, a0.z), v3(a0.w, a1.x, a1.y), v3(a1.z, a1.w, a2.x), t0) &&
                       t0 < a.radius) ||
                      (two && intersect_t(r.o, r.d, r.f, v3(b0.x, b0.y, b0.z), v3(b0.w, b1.x, b1.y), v3(b1.z, b1.w, b2.x), t1) &&
                       t1 < a.radius);
            }
            res = occluded4<COUNT, NEAR>(r, a.radius, a.nodes, a.tris, a.ntri, stk + threadIdx.x, a.tri_q, visits, tests,
                                         live && !hit, &cv);
            if (hit) res = 1;
        } else {
            res = occluded4<COUNT, NEAR>(r, a.radius, a.nodes, a.tris, a.ntri, stk + threadIdx.x, a.tri_q, visits, tests);
        }
        if (res < 0) {
            occ = scan_occluded(r, a.radius, a.tris, a.ntri);
            atomicAdd(&a.counters->overflows, 1ull);
        } else {
            occ = res == 1;
        }
    }
    bool far_occ = occ;
    if (SKY && a.ntri) {   // blocked within the radius: blocked for the sky too
        constexpr float kFarT = 1e30f;
        const Ray r = make_ray(origin, dir);
        // children in slot order: faster here than nearest first (fewer registers, no sort)
        const int res2 = occluded4<false, false>(r, kFarT, a.nodes, a.tris, a.ntri, stk + threadIdx.x, a.tri_q, visits,
                                                 tests, live && !occ, nullptr, sky_from(origin, a.radius));
        far_occ = occ || (res2 < 0 ? (live && scan_occluded(r, kFarT, a.tris, a.ntri)) : res2 == 1);
    }
    if (STAY && !live) return;   // past the wave's collective steps
    if (COUNT) {
        atomicAdd(&a.counters->visits, (unsigned long long)visits);
        atomicAdd(&a.counters->tests, (unsigned long long)tests);
        atomicAdd(&a.counters->rays, 1ull);
    }
    if (SPP64) {
        const unsigned long long open = __ballot(!occ);
        if (s == 0) a.ao[pix] = float(__popcll(open)) / 64.0f;
        if (SKY) {
            const unsigned long long free = __ballot(!far_occ);
            if (s == 0) a.sky[pix] = float(__popcll(free)) / 64.0f;
        }
    } else {
        if (!occ) atomicAdd(&a.cnt[pix], 1u);
        if (SKY && !far_occ) atomicAdd(&a.cnt_free[pix], 1u);
    }
}

template <bool SPP64, bool COUNT, bool NEAR = true>
__global__ __launch_bounds__(BT) void k_trace_ao4(TraceArgs4 a) {
    __shared__ uint32_t stk[STACK_LDS * BT];
    trace_ao4<SPP64, COUNT, NEAR, false>(a, stk, nullptr);
}
// From the wave's cut, children in slot order (measured faster here than nearest first,
// whose sort costs more than it saves on the cut's short subtrees); held to 7 waves a SIMD
// (72 registers), which its LDS allows. Eight (64 registers, a smaller stack or cut) measured
// slower.
template <bool SPP64, bool COUNT>
__global__ __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(7, 7))) void k_trace_ao4c(TraceArgs4 a) {
    __shared__ uint32_t stk[STACK_LDS * BT];
    __shared__ uint32_t cutmem[CUT_WAVE * (BT / 64)];
    trace_ao4<SPP64, COUNT, false, true>(a, stk, cutmem);
}

// The visibility pass with the AO rays from the wave's cut (the sky rays from the root).
template <bool SPP64>
__global__ __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(7, 7))) void k_vis4c(TraceArgs4 a) {
    __shared__ uint32_t stk[STACK_LDS * BT];
    __shared__ uint32_t cutmem[CUT_WAVE * (BT / 64)];
    trace_ao4<SPP64, false, false, true, true>(a, stk, cutmem);
}

// Nearest (ties to the lower triangle index) or any hit of general rays.
__global__ __launch_bounds__(BT) void k_trace_rays4(const float* __restrict__ rays8, uint32_t nrays, int any, float far_scale,
                                                    const Node4* __restrict__ nodes, const float4* __restrict__ ts,
                                                    uint32_t ntri, float* __restrict__ out4) {
    __shared__ uint32_t stk[STACK_LDS * BT];
    const uint32_t k = blockIdx.x * BT + threadIdx.x;
    if (k >= nrays) return;
    const float* q = rays8 + 8 * size_t(k);
    const Ray r = make_ray(v3(q[0], q[1], q[2]), v3(q[3], q[4], q[5]));
    const float tmax = q[6];
    const float exit_scale = far_scale >= 1.0f ? fmaxf(far_scale, kExit) : far_scale;
    float best = tmax;
    int32_t bi = -1;
    uint32_t visits = 0, tests = 0;
    bool overflow = false;
    if (ntri) {
        uint32_t* const lds_col = stk + threadIdx.x;
        uint32_t spill[STACK_SPILL];
        int sp = 0;
        int32_t node = 0;
        bool done = false;
        while (node >= 0 && !done) {
            if (visits > 4 * ntri + 4) { overflow = true; break; }   // a malformed tree
            const Node4 nd = nodes[node];
            ++visits;
            const uint32_t hit = slab4(r, nd, best * kReach, exit_scale);
            int32_t next = -1;
            for (int c = 0; c < 4 && !done; ++c) {
                if (!((hit >> c) & 1u)) continue;
                const int32_t ch = ref_of(nd.ref, c);
                if (ch == EMPTY) continue;
                if (ch < 0) {
                    const uint32_t slot = uint32_t(~ch);
                    if (slot >= ntri) continue;
                    float t;
                    ++tests;
                    if (!tri_hit(r, ts, slot, t)) continue;
                    const int32_t id = __float_as_int(ts[3 * size_t(slot) + 2].y);
                    if (t < best || (bi >= 0 && t == best && id < bi)) {
                        best = t;
                        bi = id;
                        if (any) done = true;
                    }
                } else if (next < 0) {
                    next = ch;
                } else if (sp < STACK_LDS) {
                    lds_col[sp++ * BT] = uint32_t(ch);
                } else if (sp < STACK_LDS + STACK_SPILL) {
                    spill[sp++ - STACK_LDS] = uint32_t(ch);
                } else {
                    overflow = true;
                    done = true;
                }
            }
            if (done) break;
            if (next >= 0) node = next;
            else if (sp > 0) { --sp; node = int32_t(sp < STACK_LDS ? lds_col[sp * BT] : spill[sp - STACK_LDS]); }
            else node = -1;
        }
    }
    if (overflow) { best = -1.0f; bi = -2; }
    out4[4 * size_t(k)] = bi >= 0 ? best : -1.0f;
    out4[4 * size_t(k) + 1] = float(bi);
    out4[4 * size_t(k) + 2] = float(visits);
    out4[4 * size_t(k) + 3] = float(tests);
}

// rt::closest (src/renderer/rt_bvh.cpp:125-152) for a batch of rays: the nearest hit with
// t < tMax, ties to the lower input index (a function of the triangles alone, as the box test
// only adds visits), then u and v from intersectTri of that triangle in input order. A ray whose
// stack would overflow is answered by a scan of every triangle. hits: {t, u, v, tri bits};
// a miss is {-1, 0, 0, -1}.
__global__ __launch_bounds__(BT) void k_closest4(const float* __restrict__ rays8, uint32_t nrays, const Node4* __restrict__ nodes,
                                                 const float4* __restrict__ ts, uint32_t ntri, const float* __restrict__ tris_in,
                                                 float4* __restrict__ hits) {
    __shared__ uint32_t stk[STACK_LDS * BT];
    const uint32_t k = blockIdx.x * BT + threadIdx.x;
    if (k >= nrays) return;
    const float* q = rays8 + 8 * size_t(k);
    const Ray r = make_ray(v3(q[0], q[1], q[2]), v3(q[3], q[4], q[5]));
    const float tmax = q[6];
    float best = tmax;
    int32_t bi = -1;
    bool overflow = false;
    const auto take = [&](uint32_t slot) {
        float t;
        if (!tri_hit(r, ts, slot, t)) return;
        const int32_t id = __float_as_int(ts[3 * size_t(slot) + 2].y);
        if (t < best || (bi >= 0 && t == best && id < bi)) {
            best = t;
            bi = id;
        }
    };
    if (ntri) {
        uint32_t* const lds_col = stk + threadIdx.x;
        uint32_t spill[STACK_SPILL];
        int sp = 0;
        int32_t node = 0;
        uint32_t visits = 0;
        while (node >= 0) {
            if (++visits > 4 * ntri + 4) { overflow = true; break; }   // a malformed tree
            const Node4 nd = nodes[node];
            const uint32_t hit = slab4(r, nd, best * kReach, kExit);
            int32_t next = -1;
            for (int c = 0; c < 4 && !overflow; ++c) {
                if (!((hit >> c) & 1u)) continue;
                const int32_t ch = ref_of(nd.ref, c);
                if (ch == EMPTY) continue;
                if (ch < 0) {
                    if (uint32_t(~ch) < ntri) take(uint32_t(~ch));
                } else if (next < 0) {
                    next = ch;
                } else if (sp < STACK_LDS) {
                    lds_col[sp++ * BT] = uint32_t(ch);
                } else if (sp < STACK_LDS + STACK_SPILL) {
                    spill[sp++ - STACK_LDS] = uint32_t(ch);
                } else {
                    overflow = true;
                }
            }
            if (overflow) break;
            if (next >= 0) node = next;
            else if (sp > 0) { --sp; node = int32_t(sp < STACK_LDS ? lds_col[sp * BT] : spill[sp - STACK_LDS]); }
            else node = -1;
        }
        if (overflow) {   // the scan of every triangle
            best = tmax;
            bi = -1;
            for (uint32_t i = 0; i < ntri; ++i) take(i);
        }
    }
    float4 h = make_float4(-1.0f, 0.0f, 0.0f, __int_as_float(-1));
    if (bi >= 0) {
        const float* t9 = tris_in + 9 * size_t(bi);
        float t, u, v;
        if (intersect_tuv(r.o, r.d, r.f, v3(t9[0], t9[1], t9[2]), v3(t9[3], t9[4], t9[5]), v3(t9[6], t9[7], t9[8]), t, u, v))
            h = make_float4(t, u, v, __int_as_float(bi));
    }
    hits[k] = h;
}

uint32_t cdiv(uint64_t a, uint32_t b) { return uint32_t((a + b - 1) / b); }


// 64 spp with refill: a wave owns PW pixels (PW x 64 rays) and a lane that finishes its ray
// takes the wave's next one once refill_min lanes are idle (or none is live), so the lanes of
// a wave stay busy while rays of very different lengths finish. Each lane keeps its open
// counts per pixel; a pixel's AO is their wave sum / 64, as k_trace_ao4 computes it.
template <int PW>
__global__ __launch_bounds__(BT) void k_trace_ao4r(TraceArgs4 a, uint32_t refill_min) {
    __shared__ uint32_t stk[STACK_LDS * BT];
    uint32_t* const lds_col = stk + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t pix0 = (blockIdx.x * (BT / 64) + (threadIdx.x >> 6)) * PW;
    if (pix0 >= a.npix) return;
    const unsigned long long below = (1ull << lane) - 1ull;
    uint32_t cnt[PW];
    for (int k = 0; k < PW; ++k) cnt[k] = 0;
    const auto finish = [&]() {
        for (int k = 0; k < PW; ++k) {
            uint32_t v = cnt[k];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
            const uint32_t pix = pix0 + uint32_t(k);
            if (lane == 0 && pix < a.npix) a.ao[pix] = a.mask[pix] ? float(v) / 64.0f : 1.0f;
        }
    };
    if (a.ntri == 0) {   // nothing to hit: every covered pixel's rays are open
        for (int k = 0; k < PW; ++k) cnt[k] = 1;
        finish();
        return;
    }
    const char* const base = reinterpret_cast<const char*>(a.nodes);
    const float reach = a.radius * kReach;
    Ray r;
    uint32_t ox = 0, oy = 0, oz = 0;
    uint32_t spill[STACK_SPILL];
    int sp = 0, pk = 0;
    int32_t node = -1;
    uint32_t pend = 0, work = 0;
    int4 leafref = make_int4(0, 0, 0, 0);
    // ray ri of the wave: pixel pix0 + ri / 64, sample ri % 64; false for an uncovered pixel
    const auto setup = [&](uint32_t ri) -> bool {
        pk = int(ri >> 6);
        const uint32_t pix = pix0 + (ri >> 6), s = ri & 63u;
        if (pix >= a.npix || !a.mask[pix]) return false;
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
        r = make_ray(origin, ao_dir(u1, sphi, cphi, n, t, b));
        ox = r.inv.x >= 0.0f ? 0u : 48u;
        oy = r.inv.y >= 0.0f ? 0u : 48u;
        oz = r.inv.z >= 0.0f ? 0u : 48u;
        sp = 0;
        node = 0;
        pend = 0;
        work = 0;
        return true;
    };
    const auto count_open = [&]() {
        for (int k = 0; k < PW; ++k) cnt[k] += pk == k ? 1u : 0u;
    };
    const uint32_t total = PW * 64u;
    bool live = setup(lane);
    uint32_t next = 64;
    while (true) {
        if (live && pend == 0 && node >= 0) {
            const uint32_t off = uint32_t(node) * uint32_t(sizeof(Node4));
            const uint32_t hit = slab4s(r, base, off, ox, oy, oz, reach, kExit);
            const int4 ref = *reinterpret_cast<const int4*>(base + off + 96);
            const uint32_t masks = *reinterpret_cast<const uint32_t*>(base + off + 112);
            pend = hit & masks & 15u;
            leafref = ref;
            bool overflow = false;
            uint32_t ih = hit & (masks >> 4);
            if (ih) {
                node = ref_of(ref, __builtin_ctz(ih));
                ih &= ih - 1u;
                while (ih) {
                    const uint32_t v = uint32_t(ref_of(ref, __builtin_ctz(ih)));
                    ih &= ih - 1u;
                    if (sp < STACK_LDS) lds_col[sp * BT] = v;
                    else if (sp < STACK_LDS + STACK_SPILL) spill[sp - STACK_LDS] = v;
                    else overflow = true;
                    ++sp;
                }
            } else if (sp > 0) {
                --sp;
                node = int32_t(sp < STACK_LDS ? lds_col[sp * BT] : spill[sp - STACK_LDS]);
            } else {
                node = -1;
            }
            if (++work > a.ntri + 4) overflow = true;
            if (overflow) {   // answered by the scan of every triangle
                if (!scan_occluded(r, a.radius, a.tris, a.ntri)) count_open();
                live = false;
                pend = 0;
            }
        }
        const bool waiting = live && pend != 0;
        const bool can_step = live && pend == 0 && node >= 0;
        const unsigned long long wb = __ballot(waiting), sb = __ballot(can_step), lb = __ballot(live);
        if (lb != 0 && (sb == 0 || 64u * uint32_t(__popcll(wb)) >= a.tri_q * uint32_t(__popcll(lb)))) {
            if (waiting) {
                const int k = __builtin_ctz(pend);
                pend &= pend - 1u;
                const uint32_t slot = uint32_t(~ref_of(leafref, k));
                float t;
                if (slot < a.ntri && tri_hit(r, a.tris, slot, t) && t < a.radius) {
                    live = false;   // occluded
                    pend = 0;
                }
            }
        }
        if (live && pend == 0 && node < 0) {   // finished: open
            count_open();
            live = false;
        }
        const unsigned long long idle = __ballot(!live);
        const uint32_t nidle = uint32_t(__popcll(idle));
        if (next < total && (nidle >= refill_min || nidle == 64u)) {
            if (!live) {
                const uint32_t ri = next + uint32_t(__popcll(idle & below));
                if (ri < total) live = setup(ri);
            }
            next += nidle;
        }
        if (__ballot(live) == 0 && next >= total) break;
    }
    finish();
}

// k_trace_ao4r from the wave's cut: the wave walks the tree once for the union of its PW pixels'
// reach boxes (no walk when that union is wider than a.cut_span), and every ray starts from the
// cut: its entered cut leaves wait with its node's leaf slots for the wave's triangle phases,
// and when its stack runs empty it takes the next cut node it enters. Open counts and refills
// as k_trace_ao4r. Held to 6 waves a SIMD (80 registers). COUNT: per-lane visit and test counts.
// The per-pixel part of a ray's set-up (origin, basis, x and y) is computed once for each of the
// wave's pixels into LDS, and the host-libm angle steps of all PW x 64 rays are read at the start,
// their loads in flight together, and kept as one nibble a ray (lane L holds sample L of each
// pixel); a refilled ray takes its nibble from that lane with a shuffle.
// MASK: also a.open_mask, each pixel's open samples as bits (gathered by LDS atomics as the rays
// end; 0 for a pixel off the mask), for the visibility pass's sky rays (k_sky4r).
constexpr int PIX_WORDS = 16;   // per pixel: origin, n, t, b (12 floats), x, y
// k_trace_ao4rc's waves a block, measured at 1080p on the 1M-triangle scene. Two waves free a
// finished wave's slot sooner (the 64-spp trace alone 5.45 ms against 5.56 with four), but at 64
// spp without the visibility masks, the frames-in-flight case, four leave the partly idle slots
// that the next frame's build kernels on other streams fill (6.40 ms a frame in flight against
// 6.80 with two); eight: 6.63 ms alone. Below 64 spp, 24 spp runs faster with four.
constexpr int rc_waves(int spp, bool mask) { return (spp == 64 && !mask) || spp == 24 ? 4 : 2; }
// k_sky4r: waves a block; on the 64 B records one (the pass at 1080p 17.21 ms against 17.44 with
// two and 17.63 with four; 4 or 16 pixels a wave instead of 8: 18.5 and 18.3 ms)
constexpr int sky_waves(bool q) { return q ? 1 : 2; }
// One group of PW = 256 / SPP pixels (from pix0) for the calling wave: its PW x SPP (at most 256)
// rays, ray ri = sample ri % SPP of pixel pix0 + ri / SPP. stk: the block's traversal stack (STACK_LDS rows of
// SB words, SB the block's threads); cutmem: the wave's CUT_WAVE words; pm: the wave's
// PIX_WORDS x PW floats; om: the wave's 2 PW words (MASK, 64 spp); cl: the wave's PW open counts
// (below 64 spp). COUNT: the wave's lanes add their node visits, triangle tests and rays to
// visits, tests, rays.
template <int SPP, bool COUNT, bool MASK, int SB, bool Q = false>
__device__ __forceinline__ void trace_group_rc(const TraceArgs4& a, uint32_t refill_min, uint32_t pix0, uint32_t* stk,
                                               uint32_t* cutmem, float* pm, uint32_t* om, uint32_t* cl, uint32_t& visits,
                                               uint32_t& tests, uint32_t& rays) {
    static_assert(SPP == 8 || SPP == 16 || SPP == 18 || SPP == 20 || SPP == 24 || SPP == 32 || SPP == 64,
                  "8, 16, 18, 20, 24, 32 or 64 rays a pixel");
    static_assert(!MASK || SPP == 64, "the open masks are 64 bits a pixel");
    constexpr int PW = 256 / SPP;
    // 64 spp: a lane's open counts in one register, a byte a pixel; below, LDS counters
    constexpr bool REGCNT = PW <= 4;
    uint32_t* const lds_col = stk + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u;
    const unsigned long long below = (1ull << lane) - 1ull;
    uint32_t cnt = 0;   // REGCNT: byte k holds this lane's open rays of pixel pix0 + k (at most 64 each)
    if (MASK && lane < 2u * PW) om[lane] = 0u;
    if (!REGCNT && lane < uint32_t(PW)) cl[lane] = 0u;
    const auto finish = [&]() {
        if (MASK || !REGCNT) wave_sync();
        if (REGCNT) {
            uint32_t all = cnt;   // the wave's sums stay below 256 a byte: no carries between pixels
            for (int off = 32; off > 0; off >>= 1) all += __shfl_xor(all, off);
            for (int k = 0; k < PW; ++k) {
                const uint32_t v = (all >> (8 * k)) & 255u;
                const uint32_t pix = pix0 + uint32_t(k);
                if (lane == 0 && pix < a.npix) a.ao[pix] = a.mask[pix] ? float(v) / float(SPP) : 1.0f;
                if (MASK && lane == 0 && pix < a.npix)
                    a.open_mask[pix] = a.mask[pix] ? (uint64_t(om[2 * k + 1]) << 32) | om[2 * k] : 0ull;
            }
        } else {
            const uint32_t pix = pix0 + lane;
            if (lane < uint32_t(PW) && pix < a.npix) a.ao[pix] = a.mask[pix] ? float(cl[lane]) / float(SPP) : 1.0f;
        }
    };
    if (a.ntri == 0) {   // nothing to hit: every covered pixel's rays are open
        cnt = 0x01010101u;   // 64 lanes x 1 = 64 a pixel
        if (MASK && lane < 2u * PW) om[lane] = ~0u;
        if (!REGCNT && lane < uint32_t(PW)) cl[lane] = SPP;
        finish();
        return;
    }
    // lane k < PW: pixel pix0 + k's set-up (as computeRTAO's per pixel: ray_ao.cpp:24-31) and its
    // reach box; the union of the covered pixels' boxes
    V3 lo = v3(kFar, kFar, kFar), hi = v3(-kFar, -kFar, -kFar);
    const bool cov = lane < uint32_t(PW) && pix0 + lane < a.npix && a.mask[pix0 + lane];
    if (cov) {
        const uint32_t pix = pix0 + lane;
        const V3 p = v3(a.pos[3 * size_t(pix)], a.pos[3 * size_t(pix) + 1], a.pos[3 * size_t(pix) + 2]);
        const V3 n = v3(a.nrm[3 * size_t(pix)], a.nrm[3 * size_t(pix) + 1], a.nrm[3 * size_t(pix) + 2]);
        const V3 o = add(p, scl(n, 1e-3f));
        V3 t, b;
        basis(n, t, b);
        const uint32_t y = pix / a.w, x = pix - y * a.w;
        float* q = pm + lane * PIX_WORDS;
        q[0] = o.x; q[1] = o.y; q[2] = o.z;
        q[3] = n.x; q[4] = n.y; q[5] = n.z;
        q[6] = t.x; q[7] = t.y; q[8] = t.z;
        q[9] = b.x; q[10] = b.y; q[11] = b.z;
        q[12] = __uint_as_float(x);
        q[13] = __uint_as_float(y);
        const float big = fmaxf(fabsf(o.x), fmaxf(fabsf(o.y), fabsf(o.z)));
        const float half = a.radius * 1.001f + (big + a.radius) * 1e-5f;
        lo = v3(o.x - half, o.y - half, o.z - half);
        hi = v3(o.x + half, o.y + half, o.z + half);
    }
    const uint32_t covmask = uint32_t(__ballot(cov));   // bit k: pixel pix0 + k is covered
    lo = v3(wave_min(lo.x), wave_min(lo.y), wave_min(lo.z));
    hi = v3(wave_max(hi.x), wave_max(hi.y), wave_max(hi.z));
    if (covmask == 0u) {   // no covered pixel
        finish();
        return;
    }
    wave_sync();
    // the angle steps: nibble k of lane L = ray L + 64 k (0 where no step applies)
    uint32_t fixes = 0;
    if (a.trig) {
        uint32_t words[4], idx[4];
        for (int k = 0; k < 4; ++k) {
            const uint32_t ri = lane + 64u * uint32_t(k), pk = ri / SPP;
            idx[k] = 0;
            words[k] = 0;
            if (ri < uint32_t(PW * SPP) && ((covmask >> pk) & 1u)) {
                const uint32_t x = __float_as_uint(pm[pk * PIX_WORDS + 12]), y = __float_as_uint(pm[pk * PIX_WORDS + 13]);
                idx[k] = hash24(x, y, 2 * (ri % SPP) + 1);
                words[k] = reinterpret_cast<const uint32_t*>(a.trig)[idx[k] >> 5];
            }
        }
        for (int k = 0; k < 4; ++k)
            if ((words[k] >> (idx[k] & 31u)) & 1u)
                fixes |= ((uint32_t(a.trig[kTrigBitmapBytes + (idx[k] >> 1)]) >> ((idx[k] & 1u) * 4)) & 15u) << (4 * k);
    }
    uint32_t* const mine = cutmem;
    CutView cv;
    cv.id = mine;
    cv.box = reinterpret_cast<float*>(mine + CUT_N);
    cv.leaf = mine + 7 * CUT_N;
    cv.lbox = reinterpret_cast<float*>(mine + 7 * CUT_N + CUT_LEAF);
    cv.n = -1;
    cv.nleaf = 0;
    if (hi.x - lo.x <= a.cut_span && hi.y - lo.y <= a.cut_span && hi.z - lo.z <= a.cut_span) {
        uint32_t steps = 0;
        wave_cut<SB>(a.nodes, a.top, *a.ntop, lo, hi, a.cut_small, stk + (threadIdx.x & ~63u), cv, steps);
        if (COUNT) visits += steps;
    }
    const char* const base = reinterpret_cast<const char*>(a.nodes);
    const float reach = a.radius * kReach;
    const int ncut = cv.n;
    Ray r;
    uint32_t ox = 0, oy = 0, oz = 0;
    uint32_t spill[STACK_SPILL];
    int sp = 0, pk = 0, ci = 0;
    int32_t node = -1;
    uint32_t pend = 0, lp = 0, work = 0, cur_s = 0;
    int4 leafref = make_int4(0, 0, 0, 0);
    const auto next_cut = [&]() -> int32_t {
        while (ci < ncut) {
            const int k = ci++;
            if (CutView::enters(r, cv.box, CUT_N, k, reach, kExit)) return int32_t(cv.id[k]);
        }
        return -1;
    };
    // ray ri of the wave: pixel pix0 + ri / SPP, sample ri % SPP, angle step nib; false for an
    // uncovered pixel (or one past the image)
    const auto setup = [&](uint32_t ri, uint32_t nib) -> bool {
        pk = int(ri / SPP);
        const uint32_t s = ri % SPP;
        if (MASK) cur_s = s;
        if (!((covmask >> pk) & 1u)) return false;
        const float* q = pm + pk * PIX_WORDS;
        const V3 origin = v3(q[0], q[1], q[2]), n = v3(q[3], q[4], q[5]), t = v3(q[6], q[7], q[8]), b = v3(q[9], q[10], q[11]);
        const uint32_t x = __float_as_uint(q[12]), y = __float_as_uint(q[13]);
        const uint32_t k1 = hash24(x, y, 2 * s), k2 = hash24(x, y, 2 * s + 1);
        const float u1 = float(k1) / float(0x1000000), u2 = float(k2) / float(0x1000000);
        float sphi, cphi;
        sincos_base(6.2831853f * u2, sphi, cphi);
        trig_adjust(nib, sphi, cphi);
        r = make_ray(origin, ao_dir(u1, sphi, cphi, n, t, b));
        ox = r.inv.x >= 0.0f ? 0u : 48u;
        oy = r.inv.y >= 0.0f ? 0u : 48u;
        oz = r.inv.z >= 0.0f ? 0u : 48u;
        sp = 0;
        pend = 0;
        work = 0;
        ci = 0;
        lp = 0;
        for (int i = 0; i < cv.nleaf; ++i)   // nleaf is 0 when the cut overflowed
            if (cv.leaf[i] < a.ntri && CutView::enters(r, cv.lbox, CUT_LEAF, i, reach, kExit)) lp |= 1u << i;
        node = ncut < 0 ? 0 : next_cut();   // the root, or the first cut node the ray enters
        if (COUNT) ++rays;
        return true;
    };
    const auto count_open = [&]() {
        if (REGCNT) cnt += 1u << (8 * pk);
        else atomicAdd(&cl[pk], 1u);
        if (MASK) atomicOr(&om[2 * pk + (cur_s >> 5)], 1u << (cur_s & 31u));
    };
    constexpr uint32_t total = PW * SPP;   // 256 (252 at 18 spp, 240 at 20 and 24)
    bool live = setup(lane, fixes & 15u);
    uint32_t next = 64;
    while (true) {
        if (live && pend == 0 && node >= 0) {
            if (COUNT) ++visits;
            uint32_t hit, masks;
            int4 ref;
            if constexpr (Q) {
                float t0s[4];   // unused
                hit = slab4q(r, reinterpret_cast<const char*>(a.nodesq), uint32_t(node) * uint32_t(sizeof(Node4Q)), ox, oy, oz, reach,
                             kExit, ref, masks, t0s);
            } else {
                const uint32_t off = uint32_t(node) * uint32_t(sizeof(Node4));
                hit = slab4s(r, base, off, ox, oy, oz, reach, kExit);
                ref = *reinterpret_cast<const int4*>(base + off + 96);
                masks = *reinterpret_cast<const uint32_t*>(base + off + 112);
            }
            pend = hit & masks & 15u;
            leafref = ref;
            bool overflow = false;
            uint32_t ih = hit & (masks >> 4);
            if (ih) {
                node = ref_of(ref, __builtin_ctz(ih));
                ih &= ih - 1u;
                while (ih) {
                    const uint32_t v = uint32_t(ref_of(ref, __builtin_ctz(ih)));
                    ih &= ih - 1u;
                    if (sp < STACK_LDS) lds_col[sp * SB] = v;
                    else if (sp < STACK_LDS + STACK_SPILL) spill[sp - STACK_LDS] = v;
                    else overflow = true;
                    ++sp;
                }
            } else if (sp > 0) {
                --sp;
                node = int32_t(sp < STACK_LDS ? lds_col[sp * SB] : spill[sp - STACK_LDS]);
            } else {
                node = next_cut();
            }
            if (++work > a.ntri + 4) overflow = true;
            if (overflow) {   // answered by the scan of every triangle
                if (!scan_occluded(r, a.radius, a.tris, a.ntri)) count_open();
                live = false;
                pend = 0;
                lp = 0;
            }
        }
        const bool waiting = live && (pend != 0 || lp != 0);
        const bool can_step = live && pend == 0 && node >= 0;
        const unsigned long long wb = __ballot(waiting), sb = __ballot(can_step), lb = __ballot(live);
        if (lb != 0 && (sb == 0 || 64u * uint32_t(__popcll(wb)) >= a.tri_q * uint32_t(__popcll(lb)))) {
            if (waiting) {
                uint32_t slot;
                if (pend) {
                    const int k = __builtin_ctz(pend);
                    pend &= pend - 1u;
                    slot = uint32_t(~ref_of(leafref, k));
                } else {
                    slot = cv.leaf[__builtin_ctz(lp)];
                    lp &= lp - 1u;
                }
                if (COUNT) ++tests;
                float t;
                if (slot < a.ntri && tri_hit(r, a.tris, slot, t) && t < a.radius) {
                    live = false;   // occluded
                    pend = 0;
                    lp = 0;
                }
            }
        }
        if (live && pend == 0 && lp == 0 && node < 0) {   // finished: open
            count_open();
            live = false;
        }
        const unsigned long long idle = __ballot(!live);
        const uint32_t nidle = uint32_t(__popcll(idle));
        if (next < total && (nidle >= refill_min || nidle == 64u)) {
            // every lane takes part in the shuffle (its source lane may be busy tracing)
            const uint32_t ri = next + uint32_t(__popcll(idle & below));
            const uint32_t nib = (uint32_t(__shfl(int(fixes), int(ri & 63u))) >> (4u * ((ri >> 6) & 3u))) & 15u;
            if (!live && ri < total) live = setup(ri, nib);
            next += nidle;
        }
        if (__ballot(live) == 0 && next >= total) break;
    }
    finish();
}

// The kernel: WB waves a block, each tracing one group of PW consecutive pixels. One wave a block
// frees a finished wave's slot at once for the next group (with more, it waits for the block's
// slowest wave).
template <int SPP, bool COUNT, bool MASK = false, bool Q = false, int WB = rc_waves(SPP, MASK)>
__global__ __launch_bounds__(64 * WB) __attribute__((amdgpu_waves_per_eu(6, 6))) void k_trace_ao4rc(TraceArgs4 a,
                                                                                                     uint32_t refill_min) {
    constexpr int PW = 256 / SPP, SB = 64 * WB;
    static_assert(STACK_LDS >= 14, "the wave cut's two frontiers use 14 rows of the traversal stack");
    __shared__ uint32_t stk[STACK_LDS * SB];
    __shared__ uint32_t cutmem[CUT_WAVE * WB];
    __shared__ float pixmem[PIX_WORDS * PW * WB];
    __shared__ uint32_t openmem[MASK ? 2 * PW * WB : 1];
    __shared__ uint32_t cntmem[PW > 4 ? PW * WB : 1];
    const uint32_t wave = WB == 1 ? 0u : __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t pix0 = (blockIdx.x * WB + wave) * PW;
    if (pix0 >= a.npix) return;
    uint32_t visits = 0, tests = 0, rays = 0;
    trace_group_rc<SPP, COUNT, MASK, SB, Q>(a, refill_min, pix0, stk, cutmem + wave * CUT_WAVE, pixmem + wave * (PIX_WORDS * PW),
                                         openmem + (MASK ? wave * 2 * PW : 0), cntmem + (PW > 4 ? wave * PW : 0), visits, tests,
                                         rays);
    if (COUNT) {
        atomicAdd(&a.counters->visits, (unsigned long long)visits);
        atomicAdd(&a.counters->tests, (unsigned long long)tests);
        atomicAdd(&a.counters->rays, (unsigned long long)rays);
    }
}

// The sky rays of the visibility pass at 64 spp, traced apart from its AO rays: a wave takes G
// consecutive pixels, lists the open AO samples of each (a.open_mask, from
// k_trace_ao4rc<.., true>) in LDS, and traces them one a lane to tMax 1e30 from the root, the
// box tests starting at sky_from; a node step descends into the hit inner slot entered nearest
// and stacks the others in slot order (sorting them all by distance: 21.8 ms against 19.1 for the
// visibility pass at 1080p on the 1M-triangle scene; slot order throughout: 19.3). A lane whose ray ends takes the next
// listed ray once refill_min lanes are idle (or all are), so the lanes stay busy until the list
// runs out, where a wave of one pixel's rays idles while its longest ray runs (measured with 2, 4,
// 8, 16 and 32 pixels a wave at 1080p on the 1M-triangle scene: 8 is the fastest). A pixel's sky is
// its unblocked rays / 64 (counted in LDS), 1 off the mask: the same rays and the same any-hit
// answers as k_vis4c, so the same values.
template <int G, int SL, bool Q = false, int WB = sky_waves(Q)>
__global__ __launch_bounds__(64 * WB) __attribute__((amdgpu_waves_per_eu(Q ? 8 : 6, Q ? 8 : 6))) void k_sky4r(TraceArgs4 a,
                                                                                                        uint32_t refill_min) {
    static_assert(G <= 16 && (G & (G - 1)) == 0, "one pixel a lane for the set-up, 4 bits a pixel in klass");
    constexpr int SB = 64 * WB;   // the block's threads: the stack's row stride
    __shared__ uint32_t stk[SL * SB];
    __shared__ float pixmem[PIX_WORDS * G * WB];
    __shared__ uint16_t raylist[64 * G * WB];
    __shared__ uint32_t nfree[G * WB];
    uint32_t* const lds_col = stk + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = WB == 1 ? 0u : __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t pix0 = (blockIdx.x * WB + wave) * G;
    if (pix0 >= a.npix) return;
    const unsigned long long below = (1ull << lane) - 1ull;
    float* const pm = pixmem + wave * (PIX_WORDS * G);
    uint16_t* const lst = raylist + wave * (64 * G);
    uint32_t* const nf = nfree + wave * G;
    // lane k < G: pixel pix0 + k's set-up (as k_trace_ao4rc's) and its open samples
    const uint32_t pixk = pix0 + lane;
    const bool inimg = lane < uint32_t(G) && pixk < a.npix;
    const bool cov = inimg && a.mask[pixk];
    uint64_t m = 0;
    if (cov) {
        const V3 p = v3(a.pos[3 * size_t(pixk)], a.pos[3 * size_t(pixk) + 1], a.pos[3 * size_t(pixk) + 2]);
        const V3 n = v3(a.nrm[3 * size_t(pixk)], a.nrm[3 * size_t(pixk) + 1], a.nrm[3 * size_t(pixk) + 2]);
        const V3 o = add(p, scl(n, 1e-3f));
        V3 t, b;
        basis(n, t, b);
        const uint32_t y = pixk / a.w, x = pixk - y * a.w;
        float* q = pm + lane * PIX_WORDS;
        q[0] = o.x; q[1] = o.y; q[2] = o.z;
        q[3] = n.x; q[4] = n.y; q[5] = n.z;
        q[6] = t.x; q[7] = t.y; q[8] = t.z;
        q[9] = b.x; q[10] = b.y; q[11] = b.z;
        q[12] = __uint_as_float(x);
        q[13] = __uint_as_float(y);
        m = a.open_mask[pixk];
    }
    if (lane < uint32_t(G)) nf[lane] = a.ntri ? 0u : uint32_t(__popcll(m));   // nothing to hit: every open ray is unblocked
    wave_sync();
    // The wave's list of rays, entries (pixel << 6) | sample, the most grazing first: by the top
    // two bits of the sample's u1 (the squared sine of its angle from the normal), descending. A
    // grazing sky ray crosses more of the scene than a steep one, so the long rays start first and
    // the list ends with short ones, which keeps the lanes busy to the end. Within a class the
    // rays of each half of the azimuth (the top bit of u2) come together, so the lanes of a wave
    // set out in like directions (19.14 -> 19.07 ms the visibility pass; quarters: 19.23). One
    // pass counts the 8 buckets, a second writes them; lane s handles sample s of each pixel.
    constexpr int NB = 8;
    uint32_t total = 0;
    uint64_t klass = 0;   // 4 bits a pixel: 0 closed, else 1 + bucket
    uint32_t base_c[NB];
    for (int c = 0; c < NB; ++c) base_c[c] = 0;
    if (a.ntri) {
        for (int k = 0; k < G; ++k) {
            const uint32_t mlo = uint32_t(__shfl(int(uint32_t(m)), k)), mhi = uint32_t(__shfl(int(uint32_t(m >> 32)), k));
            const bool open = ((lane < 32u ? mlo >> lane : mhi >> (lane - 32u)) & 1u) != 0u;
            const uint32_t x = __float_as_uint(pm[k * PIX_WORDS + 12]), y = __float_as_uint(pm[k * PIX_WORDS + 13]);
            const uint32_t bucket = ((hash24(x, y, 2 * lane) >> 22) << 1) | (hash24(x, y, 2 * lane + 1) >> 23);
            const uint32_t cl = open ? 1u + bucket : 0u;
            klass |= uint64_t(cl) << (4 * k);
            for (int c = 0; c < NB; ++c) base_c[c] += uint32_t(__popcll(__ballot(cl == uint32_t(c) + 1u)));
        }
        // buckets 7, 6, ..., 0 in that order (classes 3, 2, 1, 0)
        uint32_t run = 0;
        for (int c = NB - 1; c >= 0; --c) {
            const uint32_t t = base_c[c];
            base_c[c] = run;
            run += t;
        }
        total = run;
        for (int k = 0; k < G; ++k) {
            const uint32_t cl = uint32_t(klass >> (4 * k)) & 15u;
            for (int c = 0; c < NB; ++c) {
                const unsigned long long bal = __ballot(cl == uint32_t(c) + 1u);
                if (cl == uint32_t(c) + 1u) lst[base_c[c] + uint32_t(__popcll(bal & below))] = uint16_t((uint32_t(k) << 6) | lane);
                base_c[c] += uint32_t(__popcll(bal));
            }
        }
    }
    wave_sync();
    constexpr float kFarT = 1e30f;
    const float reach = kFarT * kReach;
    const char* const base = reinterpret_cast<const char*>(a.nodes);
    Ray r;
    uint32_t ox = 0, oy = 0, oz = 0;
    float tlo = 0.0f;
    uint32_t spill[STACK_SPILL];
    int sp = 0, pk = 0;
    int32_t node = -1;
    uint32_t pend = 0, work = 0;
    int4 leafref = make_int4(0, 0, 0, 0);
    const auto setup = [&](uint32_t e) {
        pk = int(e >> 6);
        const uint32_t s = e & 63u;
        const float* q = pm + pk * PIX_WORDS;
        const V3 origin = v3(q[0], q[1], q[2]), n = v3(q[3], q[4], q[5]), t = v3(q[6], q[7], q[8]), b = v3(q[9], q[10], q[11]);
        const uint32_t x = __float_as_uint(q[12]), y = __float_as_uint(q[13]);
        const uint32_t k1 = hash24(x, y, 2 * s), k2 = hash24(x, y, 2 * s + 1);
        const float u1 = float(k1) / float(0x1000000), u2 = float(k2) / float(0x1000000);
        float sphi, cphi;
        sincos_base(6.2831853f * u2, sphi, cphi);
        if (a.trig) trig_fix(a.trig, k2, sphi, cphi);
        r = make_ray(origin, ao_dir(u1, sphi, cphi, n, t, b));
        ox = r.inv.x >= 0.0f ? 0u : 48u;
        oy = r.inv.y >= 0.0f ? 0u : 48u;
        oz = r.inv.z >= 0.0f ? 0u : 48u;
        tlo = sky_from(origin, a.radius);
        sp = 0;
        pend = 0;
        work = 0;
        node = 0;   // the root
    };
    bool live = false;
    uint32_t next = 0;
    while (true) {
        const unsigned long long idle = __ballot(!live);
        const uint32_t nidle = uint32_t(__popcll(idle));
        if (next < total && (nidle >= refill_min || nidle == 64u)) {
            if (!live) {
                const uint32_t ri = next + uint32_t(__popcll(idle & below));
                if (ri < total) {
                    setup(lst[ri]);
                    live = true;
                }
            }
            next += nidle;
        }
        // with every lane idle the refill above takes the rest of the list, so none live = done
        if (__ballot(live) == 0ull) break;
        if (live && pend == 0 && node >= 0) {
            float t0s[4];
            uint32_t hit, neg;
            int4 ref;
            if constexpr (Q) {
                // the 64 B record: slot kinds from its masks word (an unused slot's quantized box is
                // never entered either)
                uint32_t masks;
                hit = slab4q(r, reinterpret_cast<const char*>(a.nodesq), uint32_t(node) * uint32_t(sizeof(Node4Q)), ox, oy, oz, reach,
                             kExit, ref, masks, t0s, tlo);
                hit &= (masks | (masks >> 4)) & 15u;
                neg = masks & 15u;
            } else {
                const uint32_t off = uint32_t(node) * uint32_t(sizeof(Node4));
                hit = slab4s_t(r, base, off, ox, oy, oz, reach, kExit, t0s, tlo);
                ref = *reinterpret_cast<const int4*>(base + off + 96);
                // leaf slots have negative refs, inner ones non-negative, and an empty slot's box
                // is never entered, so the refs give the slot kinds without the masks word's load
                neg = (ref.x < 0 ? 1u : 0u) | (ref.y < 0 ? 2u : 0u) | (ref.z < 0 ? 4u : 0u) | (ref.w < 0 ? 8u : 0u);
            }
            pend = hit & neg;
            leafref = ref;
            bool overflow = false;
            uint32_t ih = hit & ~neg & 15u;
            if (ih) {
                int first = __builtin_ctz(ih);
                float tb = t0s[first];
                for (int c = 1; c < 4; ++c)
                    if (((ih >> c) & 1u) && t0s[c] < tb) { tb = t0s[c]; first = c; }
                node = ref_of(ref, first);
                ih &= ~(1u << first);
                while (ih) {
                    const uint32_t v = uint32_t(ref_of(ref, __builtin_ctz(ih)));
                    ih &= ih - 1u;
                    if (sp < SL) lds_col[sp * SB] = v;
                    else if (sp < SL + STACK_SPILL) spill[sp - SL] = v;
                    else overflow = true;
                    ++sp;
                }
            } else if (sp > 0) {
                --sp;
                node = int32_t(sp < SL ? lds_col[sp * SB] : spill[sp - SL]);
            } else {
                node = -1;
            }
            if (++work > a.ntri + 4) overflow = true;
            if (overflow) {   // answered by the scan of every triangle
                if (!scan_occluded(r, kFarT, a.tris, a.ntri)) atomicAdd(&nf[pk], 1u);
                atomicAdd(&a.counters->overflows, 1ull);
                live = false;
                pend = 0;
            }
        }
        const bool waiting = live && pend != 0;
        const bool can_step = live && pend == 0 && node >= 0;
        const unsigned long long wb = __ballot(waiting), sb = __ballot(can_step), lb = __ballot(live);
        if (lb != 0 && (sb == 0 || 64u * uint32_t(__popcll(wb)) >= a.tri_q * uint32_t(__popcll(lb)))) {
            if (waiting) {
                const int k = __builtin_ctz(pend);
                pend &= pend - 1u;
                const uint32_t slot = uint32_t(~ref_of(leafref, k));
                float t;
                // the change of frame: per-component selects on the 64 B records' kernel (8 waves a
                // SIMD: the pass 17.03 -> 16.90 ms), whole-vector rotations on the other
                if (slot < a.ntri && tri_hit<Q>(r, a.tris, slot, t) && t < kFarT) {
                    live = false;   // blocked
                    pend = 0;
                }
            }
        }
        if (live && pend == 0 && node < 0) {   // finished: unblocked
            atomicAdd(&nf[pk], 1u);
            live = false;
        }
    }
    wave_sync();
    if (inimg) a.sky[pixk] = cov ? float(nf[lane]) / 64.0f : 1.0f;
}

// Lighting model 2's visibility (computeRTVisibility, src/renderer/ray_ao.cpp:20-63): the AO
// rays of k_trace_ao4, and each open one traced again with tMax 1e30 for the sky.
template <bool SPP64>
__global__ __launch_bounds__(BT) void k_trace_vis4(TraceArgs4 a, float* __restrict__ sky, uint32_t* __restrict__ cnt_free) {
    __shared__ uint32_t stk[STACK_LDS * BT];
    constexpr float kFarT = 1e30f;
    const uint64_t gid = uint64_t(blockIdx.x) * BT + threadIdx.x;
    const uint32_t pix = SPP64 ? uint32_t(gid >> 6) : uint32_t(gid / a.spp);
    if (pix >= a.npix) return;
    const uint32_t s = uint32_t(gid - uint64_t(pix) * a.spp);
    if (!a.mask[pix]) {
        if (s == 0) {
            a.ao[pix] = 1.0f;
            sky[pix] = 1.0f;
        }
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
    bool occ = false, far_occ = false;
    if (a.ntri) {
        const Ray r = make_ray(origin, dir);
        uint32_t visits = 0, tests = 0;
        const int res = occluded4<false>(r, a.radius, a.nodes, a.tris, a.ntri, stk + threadIdx.x, a.tri_q, visits, tests);
        occ = res < 0 ? scan_occluded(r, a.radius, a.tris, a.ntri) : res == 1;
        // blocked within radius: blocked for the sky too
        const int res2 = occluded4<false>(r, kFarT, a.nodes, a.tris, a.ntri, stk + threadIdx.x, a.tri_q, visits, tests, !occ,
                                          nullptr, sky_from(origin, a.radius));
        far_occ = occ || (res2 < 0 ? scan_occluded(r, kFarT, a.tris, a.ntri) : res2 == 1);
    }
    if (SPP64) {
        const unsigned long long open = __ballot(!occ), free = __ballot(!far_occ);
        if (s == 0) {
            a.ao[pix] = float(__popcll(open)) / 64.0f;
            sky[pix] = float(__popcll(free)) / 64.0f;
        }
    } else {
        if (!occ) atomicAdd(&a.cnt[pix], 1u);
        if (!far_occ) atomicAdd(&cnt_free[pix], 1u);
    }
}

// The sun's visibility: one ray a covered pixel from the AO origin towards
// normalize(to_light) with tMax 1e30 (src/renderer/ray_ao.cpp:49, :62); 1 off the mask.
__global__ __launch_bounds__(BT) void k_sun4(TraceArgs4 a, float lx, float ly, float lz, int forced, float* __restrict__ sun) {
    __shared__ uint32_t stk[STACK_LDS * BT];
    const uint32_t pix = blockIdx.x * BT + threadIdx.x;
    if (pix >= a.npix) return;
    if (!a.mask[pix] || forced || a.ntri == 0) {
        sun[pix] = 1.0f;
        return;
    }
    const V3 p = v3(a.pos[3 * size_t(pix)], a.pos[3 * size_t(pix) + 1], a.pos[3 * size_t(pix) + 2]);
    const V3 n = v3(a.nrm[3 * size_t(pix)], a.nrm[3 * size_t(pix) + 1], a.nrm[3 * size_t(pix) + 2]);
    const V3 origin = add(p, scl(n, 1e-3f));
    const Ray r = make_ray(origin, normalize(v3(lx, ly, lz)));
    uint32_t visits = 0, tests = 0;
    const int res = occluded4<false>(r, 1e30f, a.nodes, a.tris, a.ntri, stk + threadIdx.x, a.tri_q, visits, tests);
    const bool occ = res < 0 ? scan_occluded(r, 1e30f, a.tris, a.ntri) : res == 1;
    sun[pix] = occ ? 0.0f : 1.0f;
}

// ---- batched AO: one persistent launch over many views ------------------------------------
__global__ __launch_bounds__(BT) void k_batch_clear(BatchDev d) {
    const uint32_t stride = gridDim.x * BT;
    for (uint32_t i = blockIdx.x * BT + threadIdx.x; i < d.total_pix; i += stride) d.cnt[i] = 0;
    if (blockIdx.x == 0) {
        for (uint32_t v = threadIdx.x; v < d.nviews; v += BT) d.ncov[v] = 0;
        if (threadIdx.x == 0) *d.head = 0;
    }
}
// Covered pixels of view blockIdx.y, listed at the view's pix_base: one atomic a wave, so
// a wave's pixels stay contiguous and in order.
__global__ __launch_bounds__(BT) void k_batch_compact(BatchDev d) {
    const uint32_t v = blockIdx.y;
    const BatchView& bv = d.views[v];
    const uint32_t lane = threadIdx.x & 63u;
    for (uint32_t i0 = blockIdx.x * BT; i0 < bv.npix; i0 += gridDim.x * BT) {
        const uint32_t i = i0 + threadIdx.x;
        const bool cov = i < bv.npix && bv.mask[i] != 0;
        const unsigned long long b = __ballot(cov);
        if (b == 0) continue;
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(&d.ncov[v], uint32_t(__popcll(b)));
        base = __shfl(base, 0);
        if (cov) d.cov[bv.pix_base + base + uint32_t(__popcll(b & ((1ull << lane) - 1ull)))] = i;
    }
}
// First 64-ray group of each view (its rays padded to whole groups), and the total.
__global__ void k_batch_plan(BatchDev d) {
    if (threadIdx.x != 0) return;
    uint64_t g = 0;
    for (uint32_t v = 0; v < d.nviews; ++v) {
        d.gbase[v] = uint32_t(g);
        g += (uint64_t(d.ncov[v]) * d.views[v].spp + 63u) / 64u;
    }
    d.gbase[d.nviews] = uint32_t(g);
}
// Persistent: each wave takes `chunk` groups from the queue until it is empty. A group is
// 64 consecutive rays of one view (pixel-major, the samples of a pixel adjacent), one ray
// a lane; a pixel's open rays are summed in the wave and added once.
// CUT: each 64-ray group walks the tree once for the union of its live lanes' reach boxes
// (as trace_ao4 below 64 spp; no walk when that union is wider than 4 x radius) and its rays
// start from the cut.
template <bool CUT>
__global__ __launch_bounds__(BT) void k_batch_trace(BatchDev d) {
    __shared__ uint32_t stk[STACK_LDS * BT];
    __shared__ uint32_t cutmem[CUT ? CUT_WAVE * (BT / 64) : 1];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t total = d.gbase[d.nviews];
    CutView cv;
    if (CUT) {
        uint32_t* mine = cutmem + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) * CUT_WAVE;
        cv.id = mine;
        cv.box = reinterpret_cast<float*>(mine + CUT_N);
        cv.leaf = mine + 7 * CUT_N;
        cv.lbox = reinterpret_cast<float*>(mine + 7 * CUT_N + CUT_LEAF);
    }
    for (;;) {
        uint32_t g0 = 0;
        if (lane == 0) g0 = atomicAdd(d.head, d.chunk);
        // the group index, and with it the view and scene, is the same in every lane: kept
        // in scalar registers
        g0 = __builtin_amdgcn_readfirstlane(__shfl(g0, 0));
        if (g0 >= total) break;
        const uint32_t g1 = min(total, g0 + d.chunk);
        for (uint32_t g = g0; g < g1; ++g) {
            uint32_t lo = 0, hi = d.nviews;   // the view: the last v with gbase[v] <= g
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (d.gbase[mid] <= g) lo = mid;
                else hi = mid;
            }
            lo = __builtin_amdgcn_readfirstlane(lo);
            const BatchView& bv = d.views[lo];
            const BatchScene& sc = d.scenes[bv.scene];
            const uint64_t rr = uint64_t(g - d.gbase[lo]) * 64u + lane;
            const bool active = rr < uint64_t(d.ncov[lo]) * bv.spp;
            const uint32_t ci = active ? uint32_t(rr / bv.spp) : 0u;
            const uint32_t s = active ? uint32_t(rr - uint64_t(ci) * bv.spp) : 0u;
            const uint32_t pix = active ? d.cov[bv.pix_base + ci] : 0u;
            bool occ = false;
            if (sc.ntri) {
                const V3 p = v3(bv.pos[3 * size_t(pix)], bv.pos[3 * size_t(pix) + 1], bv.pos[3 * size_t(pix) + 2]);
                const V3 n = v3(bv.nrm[3 * size_t(pix)], bv.nrm[3 * size_t(pix) + 1], bv.nrm[3 * size_t(pix) + 2]);
                const V3 origin = add(p, scl(n, 1e-3f));
                uint32_t visits = 0, tests = 0;
                if (CUT) {
                    const float big = fmaxf(fabsf(origin.x), fmaxf(fabsf(origin.y), fabsf(origin.z)));
                    const float half = bv.radius * 1.001f + (big + bv.radius) * 1e-5f;
                    const V3 blo = active ? v3(origin.x - half, origin.y - half, origin.z - half) : v3(kFar, kFar, kFar);
                    const V3 bhi = active ? v3(origin.x + half, origin.y + half, origin.z + half) : v3(-kFar, -kFar, -kFar);
                    const V3 ulo = v3(wave_min(blo.x), wave_min(blo.y), wave_min(blo.z));
                    const V3 uhi = v3(wave_max(bhi.x), wave_max(bhi.y), wave_max(bhi.z));
                    const float span = 4.0f * bv.radius;
                    cv.n = -1;
                    cv.nleaf = 0;
                    if (uhi.x - ulo.x <= span && uhi.y - ulo.y <= span && uhi.z - ulo.z <= span) {
                        uint32_t steps = 0;
                        wave_cut(static_cast<const Node4*>(sc.nodes), static_cast<const TopEntry*>(sc.top), *sc.ntop, ulo, uhi,
                                 d.cut_k * bv.radius, stk + (threadIdx.x & ~63u), cv, steps);
                    }
                }
                const uint32_t y = pix / bv.w, x = pix - y * bv.w;
                V3 t, b;
                basis(n, t, b);
                const uint32_t k1 = hash24(x, y, 2 * s), k2 = hash24(x, y, 2 * s + 1);
                const float u1 = float(k1) / float(0x1000000), u2 = float(k2) / float(0x1000000);
                float sphi, cphi;
                sincos_base(6.2831853f * u2, sphi, cphi);
                if (d.trig) trig_fix(d.trig, k2, sphi, cphi);
                const V3 dir = ao_dir(u1, sphi, cphi, n, t, b);
                const Ray r = make_ray(origin, dir);
                bool hit = false;
                if (CUT) {
                    const float reach = bv.radius * kReach;
                    for (int i = 0; i < cv.nleaf && !hit && active; ++i) {
                        if (!CutView::enters(r, cv.lbox, CUT_LEAF, i, reach, kExit)) continue;
                        float th;
                        if (cv.leaf[i] < sc.ntri && tri_hit(r, sc.tris, cv.leaf[i], th) && th < bv.radius) hit = true;
                    }
                }
                const int res = occluded4<false, !CUT>(r, bv.radius, static_cast<const Node4*>(sc.nodes), sc.tris, sc.ntri,
                                                       stk + threadIdx.x, d.tri_q, visits, tests, active && !hit,
                                                       CUT ? &cv : nullptr);
                occ = hit || (res < 0 ? scan_occluded(r, bv.radius, sc.tris, sc.ntri) : res == 1);
            }
            // segments of equal ci: a lane starts one where ci changes
            const uint32_t prev = __shfl_up(ci, 1);
            const bool start = active && (lane == 0 || prev != ci);
            const unsigned long long sb = __ballot(start), ob = __ballot(active && !occ);
            if (start) {
                const unsigned long long after = lane == 63 ? 0ull : (sb >> (lane + 1)) << (lane + 1);
                const uint32_t end = after ? uint32_t(__builtin_ctzll(after)) : 64u;
                const unsigned long long seg = (end == 64 ? ~0ull : ((1ull << end) - 1ull)) & ~((1ull << lane) - 1ull);
                const uint32_t open = uint32_t(__popcll(ob & seg));
                if (open) atomicAdd(&d.cnt[bv.pix_base + pix], open);
            }
        }
    }
}
// The refill form of the batched trace (the queue's unit a range of whole pixels of one view).
// A range: at most BR_PIX consecutive covered pixels and at most 256 rays (spp < 256: 256 / spp
// pixels, at most BR_PIX; from 256 spp one pixel). The wave walks the cut of the range's union
// reach box, traces the range's rays one a lane from the cut and refills idle lanes with the
// range's next rays (k_trace_ao4rc's loop, with the view's spp, radius and scene at run time); a
// pixel's open rays are counted in LDS and stored once (every ray of a pixel is in its range).
constexpr uint32_t BR_PIX = 16;
__host__ __device__ inline uint32_t range_pixels(uint32_t spp) {
    return spp >= 256u ? 1u : (256u / spp < BR_PIX ? 256u / spp : BR_PIX);
}
__global__ void k_batch_plan_rc(BatchDev d) {
    if (threadIdx.x != 0) return;
    uint64_t g = 0;
    for (uint32_t v = 0; v < d.nviews; ++v) {
        d.gbase[v] = uint32_t(g);
        const uint32_t p = range_pixels(d.views[v].spp);
        g += (uint64_t(d.ncov[v]) + p - 1u) / p;
    }
    d.gbase[d.nviews] = uint32_t(g);
}
// 111 VGPRs, 4 waves a SIMD: capped at 5 or 6 waves it spills (3 and 19 registers) and is 2% and
// 9% slower on the batched workload.
__global__ __launch_bounds__(BT) void k_batch_trace_rc(BatchDev d) {
    __shared__ uint32_t stk[STACK_LDS * BT];
    __shared__ uint32_t cutmem[CUT_WAVE * (BT / 64)];
    __shared__ float pixmem[PIX_WORDS * BR_PIX * (BT / 64)];
    __shared__ uint32_t cntmem[BR_PIX * (BT / 64)];
    const uint32_t lane = threadIdx.x & 63u, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const unsigned long long below = (1ull << lane) - 1ull;
    uint32_t* const lds_col = stk + threadIdx.x;
    float* const pm = pixmem + wave * (PIX_WORDS * BR_PIX);
    uint32_t* const cl = cntmem + wave * BR_PIX;
    CutView cv;
    cv.id = cutmem + wave * CUT_WAVE;
    cv.box = reinterpret_cast<float*>(cv.id + CUT_N);
    cv.leaf = cv.id + 7 * CUT_N;
    cv.lbox = reinterpret_cast<float*>(cv.id + 7 * CUT_N + CUT_LEAF);
    const uint32_t nranges = d.gbase[d.nviews];
    for (;;) {
        uint32_t g0 = 0;
        if (lane == 0) g0 = atomicAdd(d.head, d.chunk);
        g0 = __builtin_amdgcn_readfirstlane(__shfl(g0, 0));
        if (g0 >= nranges) break;
        const uint32_t g1 = min(nranges, g0 + d.chunk);
        for (uint32_t g = g0; g < g1; ++g) {
            uint32_t vlo = 0, vhi = d.nviews;   // the view: the last v with gbase[v] <= g
            while (vhi - vlo > 1) {
                const uint32_t mid = (vlo + vhi) >> 1;
                if (d.gbase[mid] <= g) vlo = mid;
                else vhi = mid;
            }
            vlo = __builtin_amdgcn_readfirstlane(vlo);
            const BatchView& bv = d.views[vlo];
            const BatchScene& sc = d.scenes[bv.scene];
            const uint32_t spp = bv.spp, P = range_pixels(spp);
            const uint32_t ci0 = (g - d.gbase[vlo]) * P, np = min(P, d.ncov[vlo] - ci0);
            const uint32_t total = np * spp;
            const float radius = bv.radius;
            // lane k < np: pixel k's set-up (as k_trace_ao4rc's) and its reach box
            wave_sync();   // the previous range's LDS reads are done
            V3 lo = v3(kFar, kFar, kFar), hi = v3(-kFar, -kFar, -kFar);
            uint32_t mypix = 0;
            if (lane < np) {
                mypix = d.cov[bv.pix_base + ci0 + lane];
                const V3 p = v3(bv.pos[3 * size_t(mypix)], bv.pos[3 * size_t(mypix) + 1], bv.pos[3 * size_t(mypix) + 2]);
                const V3 n = v3(bv.nrm[3 * size_t(mypix)], bv.nrm[3 * size_t(mypix) + 1], bv.nrm[3 * size_t(mypix) + 2]);
                const V3 o = add(p, scl(n, 1e-3f));
                V3 t, b;
                basis(n, t, b);
                const uint32_t y = mypix / bv.w, x = mypix - y * bv.w;
                float* q = pm + lane * PIX_WORDS;
                q[0] = o.x; q[1] = o.y; q[2] = o.z;
                q[3] = n.x; q[4] = n.y; q[5] = n.z;
                q[6] = t.x; q[7] = t.y; q[8] = t.z;
                q[9] = b.x; q[10] = b.y; q[11] = b.z;
                q[12] = __uint_as_float(x);
                q[13] = __uint_as_float(y);
                const float big = fmaxf(fabsf(o.x), fmaxf(fabsf(o.y), fabsf(o.z)));
                const float half = radius * 1.001f + (big + radius) * 1e-5f;
                lo = v3(o.x - half, o.y - half, o.z - half);
                hi = v3(o.x + half, o.y + half, o.z + half);
                cl[lane] = sc.ntri ? 0u : spp;   // nothing to hit: every ray is open
            }
            lo = v3(wave_min(lo.x), wave_min(lo.y), wave_min(lo.z));
            hi = v3(wave_max(hi.x), wave_max(hi.y), wave_max(hi.z));
            wave_sync();
            if (sc.ntri) {
                // the angle steps, when the range's rays fit one a lane four times: nibble k of
                // lane L = ray L + 64 k
                const bool pre = d.trig && total <= 256u;
                uint32_t fixes = 0;
                if (pre) {
                    uint32_t words[4], idx[4];
                    for (int k = 0; k < 4; ++k) {
                        const uint32_t ri = lane + 64u * uint32_t(k);
                        idx[k] = 0;
                        words[k] = 0;
                        if (ri < total) {
                            const uint32_t pk = ri / spp;
                            const uint32_t x = __float_as_uint(pm[pk * PIX_WORDS + 12]), y = __float_as_uint(pm[pk * PIX_WORDS + 13]);
                            idx[k] = hash24(x, y, 2 * (ri - pk * spp) + 1);
                            words[k] = reinterpret_cast<const uint32_t*>(d.trig)[idx[k] >> 5];
                        }
                    }
                    for (int k = 0; k < 4; ++k)
                        if ((words[k] >> (idx[k] & 31u)) & 1u)
                            fixes |= ((uint32_t(d.trig[kTrigBitmapBytes + (idx[k] >> 1)]) >> ((idx[k] & 1u) * 4)) & 15u) << (4 * k);
                }
                cv.n = -1;
                cv.nleaf = 0;
                const float span = 4.0f * radius;
                if (d.cut_k > 0.0f && hi.x - lo.x <= span && hi.y - lo.y <= span && hi.z - lo.z <= span) {
                    uint32_t steps = 0;
                    wave_cut(static_cast<const Node4*>(sc.nodes), static_cast<const TopEntry*>(sc.top), *sc.ntop, lo, hi,
                             d.cut_k * radius, stk + (threadIdx.x & ~63u), cv, steps);
                }
                const char* const base = reinterpret_cast<const char*>(sc.nodes);
                const float4* const tris = sc.tris;
                const uint32_t ntri = sc.ntri;
                const float reach = radius * kReach;
                const int ncut = cv.n;
                Ray r;
                uint32_t ox = 0, oy = 0, oz = 0;
                uint32_t spill[STACK_SPILL];
                int sp = 0, pk = 0, ci = 0;
                int32_t node = -1;
                uint32_t pend = 0, lp = 0, work = 0;
                int4 leafref = make_int4(0, 0, 0, 0);
                const auto next_cut = [&]() -> int32_t {
                    while (ci < ncut) {
                        const int k = ci++;
                        if (CutView::enters(r, cv.box, CUT_N, k, reach, kExit)) return int32_t(cv.id[k]);
                    }
                    return -1;
                };
                const auto setup = [&](uint32_t ri, uint32_t nib) {
                    pk = int(ri / spp);
                    const uint32_t s = ri - uint32_t(pk) * spp;
                    const float* q = pm + pk * PIX_WORDS;
                    const V3 origin = v3(q[0], q[1], q[2]), n = v3(q[3], q[4], q[5]), t = v3(q[6], q[7], q[8]),
                             b = v3(q[9], q[10], q[11]);
                    const uint32_t x = __float_as_uint(q[12]), y = __float_as_uint(q[13]);
                    const uint32_t k1 = hash24(x, y, 2 * s), k2 = hash24(x, y, 2 * s + 1);
                    const float u1 = float(k1) / float(0x1000000), u2 = float(k2) / float(0x1000000);
                    float sphi, cphi;
                    sincos_base(6.2831853f * u2, sphi, cphi);
                    if (pre) trig_adjust(nib, sphi, cphi);
                    else if (d.trig) trig_fix(d.trig, k2, sphi, cphi);
                    r = make_ray(origin, ao_dir(u1, sphi, cphi, n, t, b));
                    ox = r.inv.x >= 0.0f ? 0u : 48u;
                    oy = r.inv.y >= 0.0f ? 0u : 48u;
                    oz = r.inv.z >= 0.0f ? 0u : 48u;
                    sp = 0;
                    pend = 0;
                    work = 0;
                    ci = 0;
                    lp = 0;
                    for (int i = 0; i < cv.nleaf; ++i)   // nleaf is 0 when the cut overflowed
                        if (cv.leaf[i] < ntri && CutView::enters(r, cv.lbox, CUT_LEAF, i, reach, kExit)) lp |= 1u << i;
                    node = ncut < 0 ? 0 : next_cut();
                };
                bool live = lane < total;
                if (live) setup(lane, fixes & 15u);
                uint32_t next = 64;
                while (true) {
                    if (live && pend == 0 && node >= 0) {
                        const uint32_t off = uint32_t(node) * uint32_t(sizeof(Node4));
                        const uint32_t hit = slab4s(r, base, off, ox, oy, oz, reach, kExit);
                        const int4 ref = *reinterpret_cast<const int4*>(base + off + 96);
                        const uint32_t masks = *reinterpret_cast<const uint32_t*>(base + off + 112);
                        pend = hit & masks & 15u;
                        leafref = ref;
                        bool overflow = false;
                        uint32_t ih = hit & (masks >> 4);
                        if (ih) {
                            node = ref_of(ref, __builtin_ctz(ih));
                            ih &= ih - 1u;
                            while (ih) {
                                const uint32_t v = uint32_t(ref_of(ref, __builtin_ctz(ih)));
                                ih &= ih - 1u;
                                if (sp < STACK_LDS) lds_col[sp * BT] = v;
                                else if (sp < STACK_LDS + STACK_SPILL) spill[sp - STACK_LDS] = v;
                                else overflow = true;
                                ++sp;
                            }
                        } else if (sp > 0) {
                            --sp;
                            node = int32_t(sp < STACK_LDS ? lds_col[sp * BT] : spill[sp - STACK_LDS]);
                        } else {
                            node = next_cut();
                        }
                        if (++work > ntri + 4) overflow = true;
                        if (overflow) {   // answered by the scan of every triangle
                            if (!scan_occluded(r, radius, tris, ntri)) atomicAdd(&cl[pk], 1u);
                            live = false;
                            pend = 0;
                            lp = 0;
                        }
                    }
                    const bool waiting = live && (pend != 0 || lp != 0);
                    const bool can_step = live && pend == 0 && node >= 0;
                    const unsigned long long wb = __ballot(waiting), sb = __ballot(can_step), lb = __ballot(live);
                    if (lb != 0 && (sb == 0 || 64u * uint32_t(__popcll(wb)) >= d.tri_q * uint32_t(__popcll(lb)))) {
                        if (waiting) {
                            uint32_t slot;
                            if (pend) {
                                const int k = __builtin_ctz(pend);
                                pend &= pend - 1u;
                                slot = uint32_t(~ref_of(leafref, k));
                            } else {
                                slot = cv.leaf[__builtin_ctz(lp)];
                                lp &= lp - 1u;
                            }
                            float t;
                            if (slot < ntri && tri_hit(r, tris, slot, t) && t < radius) {
                                live = false;   // occluded
                                pend = 0;
                                lp = 0;
                            }
                        }
                    }
                    if (live && pend == 0 && lp == 0 && node < 0) {   // finished: open
                        atomicAdd(&cl[pk], 1u);
                        live = false;
                    }
                    const unsigned long long idle = __ballot(!live);
                    const uint32_t nidle = uint32_t(__popcll(idle));
                    if (next < total && (nidle >= d.refill || nidle == 64u)) {
                        const uint32_t ri = next + uint32_t(__popcll(idle & below));
                        // every lane takes part in the shuffle (its source lane may be busy tracing)
                        const uint32_t nib = (uint32_t(__shfl(int(fixes), int(ri & 63u))) >> (4u * ((ri >> 6) & 3u))) & 15u;
                        if (!live && ri < total) {
                            setup(ri, nib);
                            live = true;
                        }
                        next += nidle;
                    }
                    if (__ballot(live) == 0 && next >= total) break;
                }
            }
            wave_sync();
            if (lane < np) d.cnt[bv.pix_base + mypix] = cl[lane];
        }
    }
}

// AO = open / spp where covered, 1 elsewhere (k_ao_finish's formula).
__global__ __launch_bounds__(BT) void k_batch_finish(BatchDev d) {
    const BatchView& bv = d.views[blockIdx.y];
    for (uint32_t i = blockIdx.x * BT + threadIdx.x; i < bv.npix; i += gridDim.x * BT)
        bv.ao[i] = bv.mask[i] ? float(d.cnt[bv.pix_base + i]) / float(bv.spp) : 1.0f;
}
}  // namespace

size_t top_count_offset() { return TOP_MAX * (sizeof(TopEntry) / 4); }

// The top table of the wide tree in c.nodes4 (TOP_MAX entries, then the count).
// The 64 B records of the first `count` 128 B records (count on the device: st->base, at most
// nmax; without st: nmax).
static int quantize_records(Ctx& c, const WideState* st, uint32_t nmax, hipStream_t s) {
    if (!ensure(c, c.nodes4q, sizeof(Node4Q) * std::max<size_t>(1, nmax))) return -2;
    k_quantize4<<<std::max(1u, std::min<uint32_t>(cdiv(nmax, BT), 2048u)), BT, 0, s>>>(c.nodes4.as<Node4>(), st, nmax,
                                                                                          c.nodes4q.as<Node4Q>());
    if (!hip_ok(c, hipGetLastError(), "quantize launch")) return -2;
    c.q_built = true;
    return 0;
}
// The top table of the wide tree in c.nodes4 (TOP_MAX entries, then the count), and with
// RAW_AO_OPT_QUANTIZED_NODES = 2 the 64 B records (of the records the state at c.wide_parity
// counts, or the one record of a single-triangle tree).
static int top_table(Ctx& c, hipStream_t s, bool single = false) {
    if (!ensure(c, c.top, sizeof(TopEntry) * TOP_MAX + 16)) return -2;
    k_top_table<<<1, 64, 0, s>>>(c.nodes4.as<Node4>(), c.top.as<TopEntry>(), c.top.as<uint32_t>() + TOP_MAX * (sizeof(TopEntry) / 4));
    if (!hip_ok(c, hipGetLastError(), "wide build")) return -2;
    c.q_built = false;
    if (c.node_q < 2) return 0;   // made by the next visibility pass that needs them
    return single ? quantize_records(c, nullptr, 1, s)
                  : quantize_records(c, c.wstate.as<WideState>() + c.wide_parity, uint32_t(std::max<size_t>(1, c.ntri - 1)), s);
}

namespace {
// The first level's state: the frontier is the binary root.
__global__ void k_wide_init(WideState* st, uint32_t* frontier, uint32_t root) {
    st[0] = WideState{1, 0, 1, 0};
    frontier[0] = root;
}
void wide_init(Ctx& c, uint32_t root, hipStream_t s) {
    k_wide_init<<<1, 1, 0, s>>>(c.wstate.as<WideState>(), c.wfront_a.as<uint32_t>(), root);
    c.wide_parity = 0;
    c.wide_in_a = true;
    c.wide_level = 0;
}

uint32_t wide_grid(uint32_t n) { return std::min<uint32_t>(cdiv(n, BT), 2048u); }

// Collapse levels to enqueue for n triangles: as many as the last finished build of n triangles
// needed, else an estimate (20 levels for 1.03 M triangles); a collapse whose levels fall short
// continues with read-backs (wide_continue), so the estimate sets only the speed.
uint32_t predicted_levels(const Ctx& c, uint32_t n) {
    if (c.build_rounds > 0) return uint32_t(c.build_rounds);
    if (c.last_rounds_n == n && c.last_wide_levels) return c.last_wide_levels;
    uint32_t l4 = 0;
    while ((uint64_t(1) << (2 * l4)) < n) ++l4;
    return 2 * l4 + 6;
}

// Enqueues `levels` collapse levels from the state at c.wide_parity (frontier in the buffer
// c.wide_in_a names). The state's pad counts the levels these grid kernels ran. With defer, a
// frontier of at most SMALL_W entries is left to k_wide_small (enqueue_small).
void enqueue_levels(Ctx& c, uint32_t levels, hipStream_t s, bool defer = true) {
    const uint32_t n = c.ntri, defer_max = defer ? uint32_t(SMALL_W) : 0u;
    WideState* d_st = c.wstate.as<WideState>();
    for (uint32_t l = 0; l < levels; ++l) {
        // a block a tile of the frontier the last build of this tree had at this grid level (the
        // kernels stride over the tiles: any grid gives the same records), else the full grid
        const uint32_t lv = c.wide_level++;
        uint32_t G = wide_grid(n);
        if (c.last_rounds_n == n && lv < c.level_m.size()) G = std::max(1u, std::min(G, cdiv(c.level_m[lv], BT)));
        uint32_t* fin = c.wide_in_a ? c.wfront_a.as<uint32_t>() : c.wfront_b.as<uint32_t>();
        uint32_t* fout = c.wide_in_a ? c.wfront_b.as<uint32_t>() : c.wfront_a.as<uint32_t>();
        const int parity = c.wide_parity;
        k_wide_count<<<G, BT, 0, s>>>(d_st + parity, fin, c.boxes.as<float>(), c.links.as<int32_t>(), n, c.wblk.as<uint32_t>(),
                                      defer_max);
        k_wide_write<<<G, BT, 0, s>>>(d_st + parity, d_st + (parity ^ 1), fin, fout, c.wblk.as<uint32_t>(), c.boxes.as<float>(),
                                      c.links.as<int32_t>(), n, c.nodes4.as<Node4>(), defer_max,
                                      c.ploc_hist.as<uint32_t>() + kPlocHist);
        c.wide_parity ^= 1;
        c.wide_in_a = !c.wide_in_a;
    }
}
// The small-frontier levels in one workgroup, on the current state and frontier.
void enqueue_small(Ctx& c, hipStream_t s) {
    k_wide_small<<<1, SMALL_W, 0, s>>>(c.wstate.as<WideState>() + c.wide_parity,
                                       c.wide_in_a ? c.wfront_a.as<uint32_t>() : c.wfront_b.as<uint32_t>(), c.boxes.as<float>(),
                                       c.links.as<int32_t>(), c.ntri, c.nodes4.as<Node4>());
}
}  // namespace

int wide_build(Ctx& c, hipStream_t s) {
    const uint32_t n = c.ntri;
    if (n == 0) return 0;
    if (!ensure(c, c.nodes4, sizeof(Node4) * std::max<size_t>(1, n - 1)) || !ensure(c, c.wfront_a, 4 * size_t(n)) ||
        !ensure(c, c.wfront_b, 4 * size_t(n)) || !ensure(c, c.wblk, 4 * size_t(cdiv(n, BT) + 1)) ||
        !ensure(c, c.wstate, 2 * sizeof(WideState)))
        return -2;
    if (n == 1) {
        k_wide_single<<<1, 64, 0, s>>>(c.boxes.as<float>(), c.nodes4.as<Node4>());
        c.wide_nodes = 1;
        c.plan_wide_nb.clear();
        return hip_ok(c, hipGetLastError(), "wide launch") ? top_table(c, s, true) : -2;
    }
    // The binary root is the last merge's node (2n - 2); build_finish confirms it and that the
    // enqueued levels emptied the frontier.
    wide_init(c, 2 * n - 2, s);
    if (!hip_ok(c, hipGetLastError(), "wide init")) return -2;
    // the top levels in one workgroup, the wide middle on the grid, the bottom in one workgroup
    const uint32_t levels = predicted_levels(c, n);
    enqueue_small(c, s);
    enqueue_levels(c, levels, s);
    enqueue_small(c, s);
    c.plan_wide_nb.assign(levels, wide_grid(n));
    return top_table(c, s);
}

int wide_continue(Ctx& c, hipStream_t s) {
    const uint32_t n = c.ntri;
    WideState h{};
    for (int batch = 0;; ++batch) {
        if (!hip_ok(c, hipMemcpyAsync(&h, c.wstate.as<WideState>() + c.wide_parity, sizeof h, hipMemcpyDeviceToHost, s),
                    "wide state") ||
            !hip_ok(c, hipStreamSynchronize(s), "wide level"))
            return -2;
        if (h.count == 0) break;
        if (batch > 64) { c.err = "wide collapse did not finish"; return -2; }
        enqueue_levels(c, 4, s, false);
        c.plan_wide_nb.insert(c.plan_wide_nb.end(), 4, wide_grid(n));
    }
    c.wide_nodes = h.base;
    c.last_wide_levels = std::max(1u, h.pad);   // grid levels (0 when the small-frontier kernels did all)
    return top_table(c, s);
}

// The collapse with the level count of the last eager wide_build of this tree: no host copies
// or waits, so a stream capture can record it.
int wide_build_fixed(Ctx& c, hipStream_t s) {
    const uint32_t n = c.ntri;
    if (n == 0) return 0;
    if (n == 1) {
        k_wide_single<<<1, 64, 0, s>>>(c.boxes.as<float>(), c.nodes4.as<Node4>());
        return hip_ok(c, hipGetLastError(), "wide launch") ? top_table(c, s, true) : -2;
    }
    if (c.plan_wide_nb.empty()) { c.err = "no recorded wide collapse plan"; return -1; }
    wide_init(c, uint32_t(c.root), s);
    enqueue_small(c, s);
    enqueue_levels(c, uint32_t(c.plan_wide_nb.size()), s);
    enqueue_small(c, s);
    return top_table(c, s);
}

int build_finish(Ctx& c, hipStream_t s) {
    const uint32_t n = c.ntri;
    if (n == 0) return 0;
    const bool wide = c.trace_kernel != RAW_AO_TRACE_BINARY && n > 1;
    // the build's one read-back: the PLOC tail's final state and the collapse's last state
    auto* h_ploc = c.h_state;
    auto* h_wide = reinterpret_cast<WideState*>(reinterpret_cast<char*>(c.h_state) + 128);
    const auto read_wide = [&]() {
        return hip_ok(c, hipMemcpyAsync(h_wide, c.wstate.as<WideState>() + c.wide_parity, sizeof(WideState), hipMemcpyDeviceToHost, s),
                      "wide state");
    };
    // and the grid rounds' cluster counts (the next build of these triangles sizes its rounds' grids by them)
    auto* h_hist = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(c.h_state) + 1024);
    const size_t nhist = std::min<size_t>(c.ploc_hist.cap / 4, kPlocHist + kWideHist);
    if (!hip_ok(c, hipMemcpyAsync(h_ploc, c.state.as<PlocState>() + 2, sizeof(PlocState), hipMemcpyDeviceToHost, s), "state read") ||
        !hip_ok(c, hipMemcpyAsync(h_hist, c.ploc_hist.p, 4 * nhist, hipMemcpyDeviceToHost, s), "round counts") ||
        (wide && !read_wide()) || !hip_ok(c, hipStreamSynchronize(s), "build"))
        return -2;
    bool redo = false;
    if (h_ploc->stalled) { c.err = "PLOC round made no merge (degenerate input)"; return -3; }
    if (h_ploc->m != 1) {   // the enqueued rounds fell short: the rest with read-backs, then the later steps again
        const int r = ploc_continue(c, s);
        if (r) return r;
        redo = true;
    }
    const PlocState f = *h_ploc;
    if (f.m != 1 || int32_t(f.root) != (n > 1 ? int32_t(2 * n - 2) : 0)) {
        c.err = "PLOC build ended with an unexpected root";
        return -3;
    }
    c.root = int32_t(f.root);
    c.rounds = f.rounds;
    c.last_rounds_n = n;
    c.last_grid_rounds = f.grid_rounds;
    if (!redo) {   // round r starts with the clusters round r - 1 left (round 0 with n)
        c.round_m.assign(1, n);
        for (uint32_t r = 1; r < f.grid_rounds && r <= uint32_t(kPlocHist); ++r) c.round_m.push_back(h_hist[r - 1]);
    } else {
        c.round_m.clear();
    }
    if (redo) {
        int r = bvh_layout(c, s);
        if (!r && wide) r = wide_build(c, s);
        if (r) return r;
        if (wide && (!read_wide() || !hip_ok(c, hipStreamSynchronize(s), "build"))) return -2;
    }
    if (!wide) return 0;
    if (h_wide->count != 0) {   // the enqueued levels fell short
        c.level_m.clear();
        return wide_continue(c, s);
    }
    c.wide_nodes = h_wide->base;
    c.last_wide_levels = std::max(1u, h_wide->pad);
    c.level_m.clear();   // grid level l starts with the frontier it recorded
    if (!redo && nhist >= size_t(kPlocHist + kWideHist))
        for (uint32_t l = 0; l < h_wide->pad && l < uint32_t(kWideHist); ++l) c.level_m.push_back(h_hist[kPlocHist + l]);
    return 0;
}

namespace {
// Zeroes n words with a kernel (a stream capture records it as a kernel node, not a memset node).
__global__ __launch_bounds__(BT) void k_zero_u32(uint32_t* __restrict__ p, size_t n) {
    for (size_t i = size_t(blockIdx.x) * BT + threadIdx.x; i < n; i += size_t(gridDim.x) * BT) p[i] = 0u;
}
}  // namespace

int trace_ao4(Ctx& c, const float* d_pos, const float* d_nrm, const uint8_t* d_mask, int w, int h, int spp, float radius,
              float* d_ao, hipStream_t s) {
    const uint32_t npix = uint32_t(w) * uint32_t(h);
    if (!ensure(c, c.counters, sizeof(TraceCounters))) return -2;
    if (spp != 64 && !ensure(c, c.st_cnt, 4 * size_t(npix))) return -2;
    // RAW_AO_OPT_QUANTIZED_NODES = 2 set after the tree was built: its 64 B records now
    if (c.node_q >= 2 && !c.q_built && c.ntri > 0 && quantize_records(c, nullptr, c.wide_nodes, s)) return -2;
    return trace_ao4_into(c, d_pos, d_nrm, d_mask, w, h, spp, radius, d_ao, c.counters.as<TraceCounters>(),
                          spp != 64 ? c.st_cnt.as<uint32_t>() : nullptr, true, s);
}

int trace_ao4_into(Ctx& c, const float* d_pos, const float* d_nrm, const uint8_t* d_mask, int w, int h, int spp, float radius,
                   float* d_ao, TraceCounters* d_counters, uint32_t* d_cnt, bool timed, hipStream_t s) {
    const uint32_t npix = uint32_t(w) * uint32_t(h);
    const bool spp64 = spp == 64;
    if (!spp64 && !d_cnt) { c.err = "no per-pixel count buffer"; return -1; }
    k_zero_u32<<<1, BT, 0, s>>>(reinterpret_cast<uint32_t*>(d_counters), sizeof(TraceCounters) / 4);
    TraceArgs4 a{d_pos, d_nrm, d_mask, d_ao, nullptr, c.nodes4.as<Node4>(), c.tris_sorted.as<float4>(),
                 c.host_trig ? c.d_trig : nullptr, d_counters, uint32_t(w), npix, uint32_t(spp),
                 c.ntri, uint32_t(c.tri_q), radius, float(c.wave_cut) * radius, 4.0f * radius, c.top.as<TopEntry>(),
                 c.top.p ? c.top.as<uint32_t>() + TOP_MAX * (sizeof(TopEntry) / 4) : nullptr};
    // the 64 B records (RAW_AO_OPT_QUANTIZED_NODES = 2) where they are made: the 64-spp refill kernel
    const bool q = c.node_q >= 2 && c.q_built && c.ntri > 0;
    if (q) a.nodesq = c.nodes4q.as<Node4Q>();
    // 8, 16, 18, 20, 24 and 32 spp: the refill kernel with 256 / spp pixels a wave (at most 256 rays),
    // which writes the AO itself
    const bool rc_low = (spp == 8 || spp == 16 || spp == 18 || spp == 20 || spp == 24 || spp == 32) && c.refill > 0 && c.wave_cut;
    if (!spp64 && !rc_low) {
        a.cnt = d_cnt;
        k_zero_u32<<<std::min<uint32_t>(2048, cdiv(npix, BT)), BT, 0, s>>>(a.cnt, npix);
    }
    const uint32_t blocks = cdiv(uint64_t(npix) * uint64_t(spp), BT);
    const bool count = c.count_next && timed;
    if (timed) (void)hipEventRecord(c.ev[2], s);
    constexpr int PW = 4;
    // a group of 256 / spp pixels a wave, rc_waves(spp) waves a block
    const auto rc_launch = [&](auto kernel, int spp_k) {
        const uint32_t wb = uint32_t(rc_waves(spp_k, false));
        kernel<<<cdiv(npix, uint32_t(256 / spp_k) * wb), 64 * wb, 0, s>>>(a, uint32_t(c.refill));
    };
    if (spp64 && c.refill > 0 && c.wave_cut && q) {
        if (count) rc_launch(k_trace_ao4rc<64, true, false, true>, 64);
        else rc_launch(k_trace_ao4rc<64, false, false, true>, 64);
    } else if (spp64 && c.refill > 0 && c.wave_cut) {
        if (count) rc_launch(k_trace_ao4rc<64, true>, 64);
        else rc_launch(k_trace_ao4rc<64, false>, 64);
    } else if (rc_low) {
        if (spp == 8) {
            if (count) rc_launch(k_trace_ao4rc<8, true>, 8);
            else rc_launch(k_trace_ao4rc<8, false>, 8);
        } else if (spp == 16) {
            if (count) rc_launch(k_trace_ao4rc<16, true>, 16);
            else rc_launch(k_trace_ao4rc<16, false>, 16);
        } else if (spp == 18) {
            if (count) rc_launch(k_trace_ao4rc<18, true>, 18);
            else rc_launch(k_trace_ao4rc<18, false>, 18);
        } else if (spp == 20) {
            if (count) rc_launch(k_trace_ao4rc<20, true>, 20);
            else rc_launch(k_trace_ao4rc<20, false>, 20);
        } else if (spp == 24) {
            if (count) rc_launch(k_trace_ao4rc<24, true>, 24);
            else rc_launch(k_trace_ao4rc<24, false>, 24);
        } else {
            if (count) rc_launch(k_trace_ao4rc<32, true>, 32);
            else rc_launch(k_trace_ao4rc<32, false>, 32);
        }
    } else if (spp64 && !count && c.refill > 0) {
        k_trace_ao4r<PW><<<cdiv(npix, uint32_t(PW * (BT / 64))), BT, 0, s>>>(a, uint32_t(c.refill));
    } else if (spp64) {
        if (count) {
            if (c.wave_cut) k_trace_ao4c<true, true><<<blocks, BT, 0, s>>>(a);
            else if (c.near_first) k_trace_ao4<true, true><<<blocks, BT, 0, s>>>(a);
            else k_trace_ao4<true, true, false><<<blocks, BT, 0, s>>>(a);
        } else if (c.wave_cut) {
            k_trace_ao4c<true, false><<<blocks, BT, 0, s>>>(a);
        } else if (c.near_first) {
            k_trace_ao4<true, false><<<blocks, BT, 0, s>>>(a);
        } else {
            k_trace_ao4<true, false, false><<<blocks, BT, 0, s>>>(a);
        }
    } else {
        if (c.wave_cut) {
            if (count) k_trace_ao4c<false, true><<<blocks, BT, 0, s>>>(a);
            else k_trace_ao4c<false, false><<<blocks, BT, 0, s>>>(a);
        } else if (count) {
            k_trace_ao4<false, true><<<blocks, BT, 0, s>>>(a);
        } else {
            k_trace_ao4<false, false><<<blocks, BT, 0, s>>>(a);
        }
    }
    if (timed) (void)hipEventRecord(c.ev[3], s);
    if (!spp64 && !rc_low) ao_finish(c, d_mask, a.cnt, npix, uint32_t(spp), d_ao, s);
    if (!hip_ok(c, hipGetLastError(), "trace launch")) return -2;
    if (timed) c.have_counts = count;
    return 0;
}

int trace_rays4(Ctx& c, const float* d_rays, uint32_t n, int any, float far_scale, float* d_out, hipStream_t s) {
    if (n == 0) return 0;
    k_trace_rays4<<<cdiv(n, BT), BT, 0, s>>>(d_rays, n, any, far_scale, c.nodes4.as<Node4>(), c.tris_sorted.as<float4>(),
                                              c.ntri, d_out);
    return hip_ok(c, hipGetLastError(), "trace_rays launch") ? 0 : -2;
}

int closest_hits4(Ctx& c, const float* d_rays, uint32_t n, float* d_hits, hipStream_t s) {
    if (n == 0) return 0;
    k_closest4<<<cdiv(n, BT), BT, 0, s>>>(d_rays, n, c.nodes4.as<Node4>(), c.tris_sorted.as<float4>(), c.ntri,
                                           c.tris_in.as<float>(), reinterpret_cast<float4*>(d_hits));
    return hip_ok(c, hipGetLastError(), "closest_hits launch") ? 0 : -2;
}

int batch_trace_enqueue(const BatchDev& d, uint32_t persistent_blocks, hipStream_t s) {
    k_batch_clear<<<1024, BT, 0, s>>>(d);
    k_batch_compact<<<dim3(256, d.nviews, 1), BT, 0, s>>>(d);
    if (d.refill > 0 && d.cut_k > 0.0f) {
        k_batch_plan_rc<<<1, 64, 0, s>>>(d);
        k_batch_trace_rc<<<persistent_blocks, BT, 0, s>>>(d);
    } else {
        k_batch_plan<<<1, 64, 0, s>>>(d);
        if (d.cut_k > 0.0f) k_batch_trace<true><<<persistent_blocks, BT, 0, s>>>(d);
        else k_batch_trace<false><<<persistent_blocks, BT, 0, s>>>(d);
    }
    k_batch_finish<<<dim3(256, d.nviews, 1), BT, 0, s>>>(d);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int trace_vis4(Ctx& c, const float* d_pos, const float* d_nrm, const uint8_t* d_mask, int w, int h, int spp, float radius,
               const float to_light[3], int sun_forced, float* d_ao, float* d_sky, float* d_sun, hipStream_t s) {
    const uint32_t npix = uint32_t(w) * uint32_t(h);
    if (!ensure(c, c.counters, sizeof(TraceCounters))) return -2;
    k_zero_u32<<<1, BT, 0, s>>>(c.counters.as<uint32_t>(), sizeof(TraceCounters) / 4);
    TraceArgs4 a{d_pos, d_nrm, d_mask, d_ao, nullptr, c.nodes4.as<Node4>(), c.tris_sorted.as<float4>(),
                 c.host_trig ? c.d_trig : nullptr, c.counters.as<TraceCounters>(), uint32_t(w), npix, uint32_t(spp),
                 c.ntri, uint32_t(c.tri_q), radius};
    const bool spp64 = spp == 64;
    uint32_t* cnt_free = nullptr;
    if (!spp64) {
        if (!ensure(c, c.st_cnt, 8 * size_t(npix))) return -2;
        a.cnt = c.st_cnt.as<uint32_t>();
        cnt_free = a.cnt + npix;
        k_zero_u32<<<std::min<uint32_t>(2048, cdiv(2 * npix, BT)), BT, 0, s>>>(a.cnt, 2 * size_t(npix));
    }
    const uint32_t blocks = cdiv(uint64_t(npix) * uint64_t(spp), BT);
    const bool cut = c.wave_cut > 0 && c.ntri > 0;
    if (cut) {
        a.sky = d_sky;
        a.cnt_free = cnt_free;
        a.cut_small = float(c.wave_cut) * radius;
        a.cut_span = 4.0f * radius;
        a.top = c.top.as<TopEntry>();
        a.ntop = c.top.as<uint32_t>() + top_count_offset();
    }
    // 64 spp: the AO rays by the refill kernel (which also records each pixel's open samples),
    // then the open ones' sky rays by k_sky4r
    const bool split = cut && spp64 && c.refill > 0 && c.sky_refill > 0;
    if (split && !ensure(c, c.vis_open, 8 * size_t(npix))) return -2;
    // the 64 B records (RAW_AO_OPT_QUANTIZED_NODES) for the split form's sky rays (1, the default)
    // and its AO rays (2)
    const bool q = split && c.node_q >= 1 && c.ntri > 0, qa = q && c.node_q >= 2;
    if (q && !c.q_built && quantize_records(c, nullptr, c.wide_nodes, s)) return -2;
    if (q) a.nodesq = c.nodes4q.as<Node4Q>();
    (void)hipEventRecord(c.ev[2], s);
    // the sky rays on the 64 B records: 12 LDS stack rows (9.3 KB a block of 2 waves) and at most 64
    // VGPRs, for 8 waves a SIMD (17.40 ms the pass at 1080p, against 17.82 with 16 rows at 7 waves
    // and 18.35 with 16 rows at 6); on the 128 B records 16 rows at 6 waves (18.97 ms; 21.28 at 8)
    constexpr int PW = 4, SKY_G = 8, SKY_STACK_Q = 12, SKY_STACK = STACK_LDS;
    if (split) {
        a.open_mask = c.vis_open.as<uint64_t>();
        constexpr int WBM = rc_waves(64, true);
        constexpr int SWQ = sky_waves(true), SW = sky_waves(false);
        const dim3 g_ao(cdiv(npix, uint32_t(PW * WBM))), g_sky(cdiv(npix, uint32_t(SKY_G * (q ? SWQ : SW))));
        if (qa) k_trace_ao4rc<64, false, true, true><<<g_ao, 64 * WBM, 0, s>>>(a, uint32_t(c.refill));
        else k_trace_ao4rc<64, false, true><<<g_ao, 64 * WBM, 0, s>>>(a, uint32_t(c.refill));
        if (q) k_sky4r<SKY_G, SKY_STACK_Q, true><<<g_sky, 64 * SWQ, 0, s>>>(a, uint32_t(c.sky_refill));
        else k_sky4r<SKY_G, SKY_STACK><<<g_sky, 64 * SW, 0, s>>>(a, uint32_t(c.sky_refill));
    } else if (cut && spp64) k_vis4c<true><<<blocks, BT, 0, s>>>(a);
    else if (cut) k_vis4c<false><<<blocks, BT, 0, s>>>(a);
    else if (spp64) k_trace_vis4<true><<<blocks, BT, 0, s>>>(a, d_sky, nullptr);
    else k_trace_vis4<false><<<blocks, BT, 0, s>>>(a, d_sky, cnt_free);
    k_sun4<<<cdiv(npix, BT), BT, 0, s>>>(a, to_light[0], to_light[1], to_light[2], sun_forced, d_sun);
    (void)hipEventRecord(c.ev[3], s);
    if (!spp64) {
        if (ao_finish(c, d_mask, a.cnt, npix, uint32_t(spp), d_ao, s) || ao_finish(c, d_mask, cnt_free, npix, uint32_t(spp), d_sky, s))
            return -2;
    }
    c.have_counts = false;
    return hip_ok(c, hipGetLastError(), "visibility launch") ? 0 : -2;
}

}  // namespace raw_ao
