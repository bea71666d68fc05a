// PLOC BVH build on gfx950: the tree of raw::rt::buildPloc (src/renderer/rt_bvh.cpp:54-114),
// node for node and bit for bit.
//
//   bounds   centroid box (two-level reduction) and the power-of-two scale
//   morton   30-bit keys, one thread a triangle
//   sort     stable LSD radix sort of (key, index), three passes of 10-bit digits over the
//            30-bit keys, ranks from wave64 ballots; stability with ascending indices gives
//            std::sort's order
//            of (key, index) pairs (rt_bvh.cpp:73)
//   leaves   one leaf a sorted position
//   rounds   per round three launches: nearest neighbours within the radius from an
//            LDS window with a halo (each block recomputes its halo's neighbours, so
//            mutual pairs are decided without a second pass), a single-block scan of the
//            blocks' merge and survivor counts, and the merge/compaction with in-block
//            ranks from ballots. Ids and positions come from prefix sums: the tree does
//            not depend on scheduling.
#include "ao_math.hpp"
#include "ctx.hpp"
#include <algorithm>
#include <cfloat>
#include <cmath>

namespace raw_ao {
namespace {

constexpr int BT = 256;          // threads a block
constexpr int RS_ITEMS = 8;      // keys a thread in the radix passes
constexpr int RS_TILE = BT * RS_ITEMS;
constexpr int RS_BITS = 10, RS_DIG = 1 << RS_BITS;   // digit width and count
constexpr int RS_PER = RS_DIG / BT;                  // digits a thread
constexpr int MAX_R = 32;        // largest PLOC radius supported

__device__ __forceinline__ unsigned long long lanemask_lt() {
    const unsigned lane = __lane_id();
    return lane ? (~0ull >> (64 - lane)) : 0ull;
}

__device__ __forceinline__ uint32_t spread(uint32_t v) {   // rt_bvh.cpp:25-31
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
__device__ __forceinline__ V3 centroid(const float* __restrict__ t) {   // rt_bvh.cpp:34-37
    const float third = 0.333333343f;
    return v3((t[0] + t[3] + t[6]) * third, (t[1] + t[4] + t[7]) * third, (t[2] + t[5] + t[8]) * third);
}
// An unsigned key in the order of float `<` for non-NaN values, with -0 and +0 equal.
__device__ __forceinline__ uint32_t area_order(float d) {
    const uint32_t u = d == 0.0f ? 0u : __float_as_uint(d);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float area6(float lx, float ly, float lz, float hx, float hy, float hz) {   // rt_bvh.cpp:46-52
    const float dx = hx - lx, dy = hy - ly, dz = hz - lz;
    return 2.0f * (dy * (dx + dz) + dz * dx);
}

// ---- centroid bounds --------------------------------------------------------------
// The reference folds min/max in index order; only the sign of a zero bound could
// depend on the order, and neither the keys nor the scale read it, so a parallel
// reduction gives the same keys.
__global__ __launch_bounds__(BT) void k_bounds_part(const float* __restrict__ tris, uint32_t n, float* __restrict__ part) {
    float l[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, h[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint32_t i = blockIdx.x * BT + threadIdx.x; i < n; i += gridDim.x * BT) {
        const V3 c = centroid(tris + 9 * size_t(i));
        const float x[3] = {c.x, c.y, c.z};
        for (int k = 0; k < 3; ++k) { l[k] = stdmin(l[k], x[k]); h[k] = stdmax(h[k], x[k]); }
    }
    for (int off = 32; off > 0; off >>= 1)
        for (int k = 0; k < 3; ++k) {
            l[k] = stdmin(l[k], __shfl_xor(l[k], off));
            h[k] = stdmax(h[k], __shfl_xor(h[k], off));
        }
    __shared__ float s[BT / 64][6];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 3; ++k) { s[w][k] = l[k]; s[w][3 + k] = h[k]; }
    __syncthreads();
    if (threadIdx.x < 6) {
        float v = s[0][threadIdx.x];
        for (int i = 1; i < BT / 64; ++i) v = threadIdx.x < 3 ? stdmin(v, s[i][threadIdx.x]) : stdmax(v, s[i][threadIdx.x]);
        part[6 * blockIdx.x + threadIdx.x] = v;
    }
}
__global__ __launch_bounds__(BT) void k_bounds_final(const float* __restrict__ part, int nparts, float* __restrict__ g4) {
    float l[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, h[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int p = threadIdx.x; p < nparts; p += BT)
        for (int k = 0; k < 3; ++k) { l[k] = stdmin(l[k], part[6 * p + k]); h[k] = stdmax(h[k], part[6 * p + 3 + k]); }
    for (int off = 32; off > 0; off >>= 1)
        for (int k = 0; k < 3; ++k) {
            l[k] = stdmin(l[k], __shfl_xor(l[k], off));
            h[k] = stdmax(h[k], __shfl_xor(h[k], off));
        }
    __shared__ float s6[BT / 64][6];
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 3; ++k) { s6[threadIdx.x >> 6][k] = l[k]; s6[threadIdx.x >> 6][3 + k] = h[k]; }
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int i = 1; i < BT / 64; ++i)
        for (int k = 0; k < 3; ++k) { l[k] = stdmin(l[k], s6[i][k]); h[k] = stdmax(h[k], s6[i][3 + k]); }
    const float extent = stdmax(h[0] - l[0], stdmax(h[1] - l[1], h[2] - l[2]));
    float s = 1.0f;   // pow2Scale (rt_bvh.cpp:38-44)
    if (extent > 0.0f) {
        while (extent * s > 1.0f) s *= 0.5f;
        while (extent * s <= 0.5f) s *= 2.0f;
    }
    g4[0] = l[0]; g4[1] = l[1]; g4[2] = l[2]; g4[3] = s;
}

__global__ __launch_bounds__(BT) void k_morton(const float* __restrict__ tris, uint32_t n, const float* __restrict__ g4,
                                               uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const uint32_t i = blockIdx.x * BT + threadIdx.x;
    if (i >= n) return;
    const V3 c = centroid(tris + 9 * size_t(i));
    const float s = g4[3];
    const uint32_t qx = min(1023u, uint32_t((c.x - g4[0]) * s * 1023.0f));
    const uint32_t qy = min(1023u, uint32_t((c.y - g4[1]) * s * 1023.0f));
    const uint32_t qz = min(1023u, uint32_t((c.z - g4[2]) * s * 1023.0f));
    keys[i] = (spread(qx) << 2) | (spread(qy) << 1) | spread(qz);
    vals[i] = i;
}

// ---- stable LSD radix sort ----------------------------------------------------------
__global__ __launch_bounds__(BT) void k_radix_hist(const uint32_t* __restrict__ keys, uint32_t n, int shift,
                                                   uint32_t* __restrict__ hist, uint32_t nblk) {
    __shared__ uint32_t cnt[RS_DIG];
    for (int k = 0; k < RS_PER; ++k) cnt[threadIdx.x + BT * k] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * RS_TILE;
    for (int it = 0; it < RS_ITEMS; ++it) {
        const uint32_t i = base + it * BT + threadIdx.x;
        if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & (RS_DIG - 1u)], 1u);
    }
    __syncthreads();
    for (int k = 0; k < RS_PER; ++k) hist[(threadIdx.x + BT * k) * nblk + blockIdx.x] = cnt[threadIdx.x + BT * k];
}

// One block a digit: the exclusive scan of that digit's counts over the tiles (a contiguous
// row of the [digit][tile] histogram), in place, and the digit's total.
__global__ __launch_bounds__(BT) void k_radix_colscan(uint32_t* __restrict__ hist, uint32_t nblk, uint32_t* __restrict__ totals) {
    __shared__ uint32_t wsum[BT / 64];
    uint32_t* row = hist + size_t(blockIdx.x) * nblk;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nblk; base += BT) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < nblk ? row[i] : 0u;
        uint32_t x = v;   // inclusive wave scan
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
        if (lane == 63) wsum[w] = x;
        __syncthreads();
        uint32_t before = carry;
        for (int k = 0; k < w; ++k) before += wsum[k];
        if (i < nblk) row[i] = before + x - v;
        uint32_t tile = 0;
        for (int k = 0; k < BT / 64; ++k) tile += wsum[k];
        carry += tile;
        __syncthreads();
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

// A tile's keys in 8 steps of 256; a step's keys of one digit keep their order (wave, then
// lane). A wave's count of a digit in a step sits in LDS stamped with the step, so entries of
// earlier steps read as zero without clearing 4 x 1,024 counters a step.
__global__ __launch_bounds__(BT) void k_radix_scatter(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                      uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, uint32_t n,
                                                      int shift, const uint32_t* __restrict__ offs, uint32_t nblk,
                                                      const uint32_t* __restrict__ totals) {
    __shared__ uint32_t run[RS_DIG];
    __shared__ uint32_t wcnt[BT / 64][RS_DIG];   // (step + 1) << 16 | count
    __shared__ uint32_t part[BT];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    {   // digit bases: exclusive scan of the 1,024 digit totals (4 consecutive digits a thread)
        uint32_t v[RS_PER], sum = 0;
        for (int k = 0; k < RS_PER; ++k) {
            v[k] = totals[threadIdx.x * RS_PER + k];
            sum += v[k];
        }
        uint32_t x = sum;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
        if (lane == 63) part[w] = x;
        for (int k = 0; k < RS_PER; ++k)
            for (int q = 0; q < BT / 64; ++q) wcnt[q][threadIdx.x + BT * k] = 0;
        __syncthreads();
        uint32_t before = x - sum;
        for (int k = 0; k < w; ++k) before += part[k];
        for (int k = 0; k < RS_PER; ++k) {
            const uint32_t d = threadIdx.x * RS_PER + k;
            run[d] = before + offs[d * nblk + blockIdx.x];
            before += v[k];
        }
        __syncthreads();
    }
    const unsigned long long lt = lanemask_lt();
    for (int it = 0; it < RS_ITEMS; ++it) {
        const uint32_t stamp = uint32_t(it + 1) << 16;
        const uint32_t i = blockIdx.x * RS_TILE + it * BT + threadIdx.x;
        const bool valid = i < n;
        const uint32_t key = valid ? kin[i] : 0u;
        const uint32_t d = (key >> shift) & (RS_DIG - 1u);
        unsigned long long same = __ballot(valid);
        for (int b = 0; b < RS_BITS; ++b) {
            const bool bit = (d >> b) & 1u;
            const unsigned long long bm = __ballot(bit);
            same &= bit ? bm : ~bm;
        }
        const uint32_t rank = __popcll(same & lt), cnt = __popcll(same);
        if (valid && rank == 0) wcnt[w][d] = stamp | cnt;
        __syncthreads();
        if (valid) {
            uint32_t pos = run[d] + rank;
            for (int k = 0; k < w; ++k) {
                const uint32_t e = wcnt[k][d];
                pos += (e & 0xffff0000u) == stamp ? (e & 0xffffu) : 0u;
            }
            kout[pos] = key;
            vout[pos] = vin[i];
        }
        __syncthreads();
        if (valid && rank == 0) atomicAdd(&run[d], cnt);   // integer sums: the order does not matter
        __syncthreads();
    }
}

// ---- leaves (rt_bvh.cpp:74-82) ----------------------------------------------------
// Also the triangles in leaf order (the traversal's triangle records: 3 float4 a triangle,
// the 9 coordinates and the triangle's index).
__global__ __launch_bounds__(BT) void k_leaves(const float* __restrict__ tris, const uint32_t* __restrict__ order, uint32_t n,
                                               float* __restrict__ boxes, int32_t* __restrict__ links,
                                               uint32_t* __restrict__ clus, float4* __restrict__ tsorted) {
    const uint32_t p = blockIdx.x * BT + threadIdx.x;
    if (p >= n) return;
    const uint32_t t = order[p];
    const float* v = tris + 9 * size_t(t);
    float4* o = tsorted + 3 * size_t(p);
    o[0] = make_float4(v[0], v[1], v[2], v[3]);
    o[1] = make_float4(v[4], v[5], v[6], v[7]);
    o[2] = make_float4(v[8], __int_as_float(int32_t(t)), 0.0f, 0.0f);
    float* b = boxes + 6 * size_t(p);
    for (int a = 0; a < 3; ++a) {
        float l = v[a], h = v[a];
        l = stdmin(l, v[3 + a]); h = stdmax(h, v[3 + a]);
        l = stdmin(l, v[6 + a]); h = stdmax(h, v[6 + a]);
        b[a] = l;
        b[3 + a] = h;
    }
    links[3 * size_t(p)] = -1;
    links[3 * size_t(p) + 1] = -1;
    links[3 * size_t(p) + 2] = int32_t(t);
    clus[p] = p;
}

// ---- PLOC rounds (rt_bvh.cpp:84-111) ------------------------------------------------
// Two launches a round, both with a fixed grid that strides over the tiles of BT positions and
// reads the cluster count from the device state, so the host enqueues rounds without reading the
// count back. A round that starts at or below tail_at clusters is a no-op (the state and clusters
// carry over to the next buffer); the single-workgroup tail runs the rest.
// Pass 1: the nearest neighbour of every position of a tile and of a halo of `radius` on each
// side (so mutual pairs are decided inside the tile), then the tile's merge and survivor flags
// and counts (counts[k * maxt + tile]).
// R > 0: the radius is R (the reference's kPlocRadius, 16, unrolled); R = 0: `radius` at run time.
template <int R>
__global__ __launch_bounds__(BT) void k_ploc_nn(const PlocState* __restrict__ st, int radius_rt, const float* __restrict__ boxes,
                                                const uint32_t* __restrict__ clus, uint32_t* __restrict__ nn,
                                                uint8_t* __restrict__ flags, uint32_t* __restrict__ counts, uint32_t maxt,
                                                uint32_t tail_at) {
    __shared__ float sb[6][BT + 4 * MAX_R];
    __shared__ int32_t snn[BT + 2 * MAX_R];
    __shared__ unsigned long long skey[BT + 2 * MAX_R];
    __shared__ uint32_t wsum[BT / 64][2];
    const int radius = R > 0 ? R : radius_rt;
    const int32_t m = int32_t(st->m);
    if (uint32_t(m) <= tail_at) return;
    const uint32_t ntile = (uint32_t(m) + BT - 1) / BT;
    for (uint32_t tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
        __syncthreads();   // the previous tile's LDS reads are done
        const int32_t b0 = int32_t(tile) * BT;
        const int32_t base = b0 - 2 * radius;          // position of sb[*][0]
        const int span = BT + 4 * radius;
        for (int k = threadIdx.x; k < span; k += BT) {
            const int32_t pos = base + k;
            if (pos >= 0 && pos < m) {
                const float* b = boxes + 6 * size_t(clus[pos]);
                for (int a = 0; a < 6; ++a) sb[a][k] = b[a];
            }
        }
        const int nspan = BT + 2 * radius;             // neighbours of positions b0-radius ..
        const int32_t q0 = b0 - radius;                // position of snn[0] and skey[0]
        for (int q = threadIdx.x; q < nspan; q += BT) skey[q] = ~0ull;
        __syncthreads();
        // Each pair (p, p + d), d = 1..radius, that some position of the span needs is evaluated
        // once, by the owner of p (positions b0-2radius .. b0+BT+radius-1): a right candidate of p
        // and a left candidate of p + d, both by an LDS atomic minimum of (area order, candidate
        // position). The minimum is the smallest area at the lowest position: the reference's
        // ascending scan with a strict `<` from FLT_MAX (areas not below FLT_MAX never count).
        for (int k = threadIdx.x; k < BT + 3 * radius; k += BT) {
            const int32_t p = base + k;
            if (p < 0 || p >= m) continue;
            const bool needs = p >= q0;                // p's own neighbour is wanted
            const float al0 = sb[0][k], al1 = sb[1][k], al2 = sb[2][k], ah0 = sb[3][k], ah1 = sb[4][k], ah2 = sb[5][k];
            unsigned long long rb = ~0ull;
            const int32_t j1 = min(m - 1, p + radius);
            // candidates p+1 .. p+radius; with R fixed the loop unrolls and its LDS loads issue
            // together (j - base < BT + 4 radius always: a position past j1 reads an unused entry,
            // whose result is dropped)
#pragma unroll
            for (int d = 1; d <= (R > 0 ? R : MAX_R); ++d) {
                if (R == 0 && d > radius) break;
                const int32_t j = p + d;
                const int jp = j - base;
                const float ar = area6(stdmin(al0, sb[0][jp]), stdmin(al1, sb[1][jp]), stdmin(al2, sb[2][jp]),
                                       stdmax(ah0, sb[3][jp]), stdmax(ah1, sb[4][jp]), stdmax(ah2, sb[5][jp]));
                if (j <= j1 && ar < FLT_MAX) {
                    const unsigned long long kd = (unsigned long long)area_order(ar) << 32;
                    rb = min(rb, kd | uint32_t(j));
                    const int qj = j - q0;
                    if (qj >= 0 && qj < nspan) atomicMin(&skey[qj], kd | uint32_t(p));
                }
            }
            if (needs && rb != ~0ull) atomicMin(&skey[p - q0], rb);
        }
        __syncthreads();
        for (int q = threadIdx.x; q < nspan; q += BT) {
            const int32_t pos = q0 + q;
            if (pos < 0 || pos >= m) continue;
            const unsigned long long best = skey[q];
            snn[q] = best == ~0ull ? pos : int32_t(uint32_t(best));
        }
        __syncthreads();
        const int32_t pos = b0 + int32_t(threadIdx.x);
        bool merge = false, survive = false;
        if (pos < m) {
            const int32_t j = snn[threadIdx.x + radius];
            const bool mutual = j != pos && snn[j - (b0 - radius)] == pos;
            merge = mutual && pos < j;
            survive = !(mutual && pos > j);
            nn[pos] = uint32_t(j);
            flags[pos] = uint8_t((merge ? 1 : 0) | (survive ? 2 : 0));
        }
        const unsigned long long bm = __ballot(merge), bs = __ballot(survive);
        const int w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) { wsum[w][0] = __popcll(bm); wsum[w][1] = __popcll(bs); }
        __syncthreads();
        if (threadIdx.x < 2) {
            uint32_t s = 0;
            for (int k = 0; k < BT / 64; ++k) s += wsum[k][threadIdx.x];
            counts[threadIdx.x * maxt + tile] = s;   // merges then survivors, tile-major
        }
    }
}

// The sums of counts[0 .. t-1] and counts[maxt .. maxt + t-1] over the workgroup (each thread
// a strided share, then wave and workgroup reductions): a tile's offsets among all tiles. Called
// by every thread of the workgroup; red: BT / 64 x 2 words of LDS.
__device__ __forceinline__ void tile_prefix(const uint32_t* __restrict__ counts, uint32_t maxt, uint32_t t, uint32_t (*red)[2],
                                            uint32_t& pm, uint32_t& ps) {
    uint32_t a = 0, b = 0;
    for (uint32_t i = threadIdx.x; i < t; i += BT) {
        a += counts[i];
        b += counts[maxt + i];
    }
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
    }
    const int w = threadIdx.x >> 6;
    __syncthreads();   // red's earlier readers are done
    if ((threadIdx.x & 63) == 0) { red[w][0] = a; red[w][1] = b; }
    __syncthreads();
    pm = ps = 0;
    for (int k = 0; k < BT / 64; ++k) { pm += red[k][0]; ps += red[k][1]; }
}

// Pass 2: each tile's offsets among all tiles (the sums of the earlier tiles' counts, read by the
// tile's own workgroup), then the new nodes from mutual pairs (id = next_id + merge rank, in
// position order) and the survivors compacted in position order; the last tile writes the next
// round's state (a copy of this one in a no-op round).
__global__ __launch_bounds__(BT) void k_ploc_merge(const PlocState* __restrict__ st, PlocState* __restrict__ st_next,
                                                   const uint32_t* __restrict__ nn, const uint8_t* __restrict__ flags,
                                                   const uint32_t* __restrict__ counts, uint32_t maxt,
                                                   const uint32_t* __restrict__ cin, uint32_t* __restrict__ cout,
                                                   float* __restrict__ boxes, int32_t* __restrict__ links, uint32_t tail_at,
                                                   uint32_t* __restrict__ hist) {
    __shared__ uint32_t wsum[BT / 64][2];
    __shared__ uint32_t red[BT / 64][2];
    const PlocState in = *st;
    const uint32_t m = in.m;
    if (m <= tail_at) {   // a no-op round: the clusters move to the next buffer unchanged
        for (uint32_t i = blockIdx.x * BT + threadIdx.x; i < m; i += gridDim.x * BT) cout[i] = cin[i];
        if (blockIdx.x == 0 && threadIdx.x == 0) *st_next = in;
        return;
    }
    const uint32_t ntile = (m + BT - 1) / BT;
    for (uint32_t tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
        uint32_t pm, ps;
        tile_prefix(counts, maxt, tile, red, pm, ps);
        const uint32_t pos = tile * BT + threadIdx.x;
        const uint8_t f = pos < m ? flags[pos] : uint8_t(0);
        const bool merge = f & 1u, survive = (f >> 1) & 1u;
        const unsigned long long lt = lanemask_lt();
        const unsigned long long bm = __ballot(merge), bs = __ballot(survive);
        const int w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) { wsum[w][0] = __popcll(bm); wsum[w][1] = __popcll(bs); }
        __syncthreads();
        if (tile == ntile - 1 && threadIdx.x == 0) {   // the last tile: the next round's state
            PlocState o{};
            o.m = ps + counts[maxt + tile];
            o.merges = pm + counts[tile];
            o.next_id = in.next_id + o.merges;
            o.stalled = in.stalled | ((in.m > 1 && o.merges == 0) ? 1u : 0u);
            o.rounds = in.rounds + (in.m > 1 ? 1u : 0u);
            o.grid_rounds = in.grid_rounds + 1;
            *st_next = o;
            if (in.grid_rounds < uint32_t(kPlocHist)) hist[in.grid_rounds] = o.m;
        }
        if (survive) {
            uint32_t S = pm + __popcll(bm & lt), SV = ps + __popcll(bs & lt);
            for (int k = 0; k < w; ++k) { S += wsum[k][0]; SV += wsum[k][1]; }
            const uint32_t a = cin[pos];
            if (!merge) {
                cout[SV] = a;
            } else {
                const uint32_t id = in.next_id + S;
                const uint32_t b = cin[nn[pos]];
                const float* A = boxes + 6 * size_t(a);
                const float* B = boxes + 6 * size_t(b);
                float* N = boxes + 6 * size_t(id);
                for (int k = 0; k < 3; ++k) { N[k] = stdmin(A[k], B[k]); N[3 + k] = stdmax(A[3 + k], B[3 + k]); }
                links[3 * size_t(id)] = int32_t(a);
                links[3 * size_t(id) + 1] = int32_t(b);
                links[3 * size_t(id) + 2] = -1;
                cout[SV] = id;
            }
        }
        __syncthreads();   // wsum's readers are done before the next tile
    }
}

// All remaining rounds once at most TAIL_M clusters are left, in one workgroup with the
// clusters' ids and boxes in LDS: a thread owns TAIL_PER consecutive positions, so the
// block-wide prefix sums run in position order. Same rounds, same order, same ids.
constexpr int TAIL_T = 1024, TAIL_PER = 4, TAIL_M = TAIL_T * TAIL_PER;   // 512 or 256 threads: 5% or 26% slower builds
__global__ __launch_bounds__(TAIL_T) void k_ploc_tail(const PlocState* __restrict__ st_in, PlocState* __restrict__ st_out,
                                                      int radius, const uint32_t* __restrict__ cin, float* __restrict__ boxes,
                                                      int32_t* __restrict__ links) {
    __shared__ float sb[6][TAIL_M];
    __shared__ uint32_t sid[TAIL_M];
    __shared__ int8_t snn[TAIL_M];
    __shared__ unsigned long long skey[TAIL_M];
    __shared__ uint32_t wsum[TAIL_T / 64][2];
    PlocState s = *st_in;
    uint32_t m = s.m;
    if (m > uint32_t(TAIL_M)) {   // too many clusters left for one workgroup: report them, unfinished
        if (threadIdx.x == 0) {
            s.root = 0xffffffffu;
            *st_out = s;
        }
        return;
    }
    for (uint32_t i = threadIdx.x; i < m; i += TAIL_T) {
        const uint32_t c = cin[i];
        sid[i] = c;
        for (int a = 0; a < 6; ++a) sb[a][i] = boxes[6 * size_t(c) + a];
    }
    __syncthreads();
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i0 = threadIdx.x * TAIL_PER;
    while (m > 1) {
        // Each pair (i, i + d), d = 1..radius, is evaluated once, by the owner of i: it is a
        // right candidate of i (kept in a register) and a left candidate of i + d (an LDS
        // atomic minimum). Keys are (area order, candidate position), so the minimum is the
        // smallest area at the lowest position: the reference's ascending scan with a strict
        // `<` from FLT_MAX (areas that are not below FLT_MAX never count). In this step thread
        // t takes positions t, t + TAIL_T, ..., so a wave's LDS reads and atomics are at
        // consecutive positions (no bank conflicts); the steps after it, which scan in position
        // order, read the result from snn.
        for (int k = 0; k < TAIL_PER; ++k) {
            const int i = int(threadIdx.x) + k * TAIL_T;
            if (i < int(m)) skey[i] = ~0ull;
        }
        __syncthreads();
        unsigned long long rbest[TAIL_PER];
        for (int k = 0; k < TAIL_PER; ++k) {
            const int i = int(threadIdx.x) + k * TAIL_T;
            rbest[k] = ~0ull;
            if (i >= int(m)) continue;
            const float al0 = sb[0][i], al1 = sb[1][i], al2 = sb[2][i], ah0 = sb[3][i], ah1 = sb[4][i], ah2 = sb[5][i];
            const int j1 = min(int(m) - 1, i + radius);
            for (int j = i + 1; j <= j1; ++j) {
                const float d = area6(stdmin(al0, sb[0][j]), stdmin(al1, sb[1][j]), stdmin(al2, sb[2][j]),
                                      stdmax(ah0, sb[3][j]), stdmax(ah1, sb[4][j]), stdmax(ah2, sb[5][j]));
                if (d < FLT_MAX) {
                    const unsigned long long kd = (unsigned long long)area_order(d) << 32;
                    rbest[k] = min(rbest[k], kd | uint32_t(j));
                    atomicMin(&skey[j], kd | uint32_t(i));
                }
            }
        }
        __syncthreads();
        for (int k = 0; k < TAIL_PER; ++k) {
            const int i = int(threadIdx.x) + k * TAIL_T;
            if (i >= int(m)) break;
            const unsigned long long best = min(skey[i], rbest[k]);
            snn[i] = best == ~0ull ? int8_t(0) : int8_t(int(uint32_t(best)) - i);
        }
        __syncthreads();
        bool mg[TAIL_PER], sv[TAIL_PER];
        uint32_t cm = 0, cs = 0;
        for (int k = 0; k < TAIL_PER; ++k) {
            const int i = i0 + k;
            mg[k] = sv[k] = false;
            if (i >= int(m)) continue;
            const int j = i + snn[i];
            const bool mutual = j != i && j + snn[j] == i;
            mg[k] = mutual && i < j;
            sv[k] = !(mutual && i > j);
            cm += mg[k];
            cs += sv[k];
        }
        uint32_t xm = cm, xs = cs;   // inclusive wave scans
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t ym = __shfl_up(xm, off), ys = __shfl_up(xs, off);
            if (lane >= off) { xm += ym; xs += ys; }
        }
        if (lane == 63) { wsum[w][0] = xm; wsum[w][1] = xs; }
        __syncthreads();
        uint32_t S = xm - cm, SV = xs - cs, tm = 0, ts = 0;
        for (int k = 0; k < TAIL_T / 64; ++k) {
            if (k < w) { S += wsum[k][0]; SV += wsum[k][1]; }
            tm += wsum[k][0];
            ts += wsum[k][1];
        }
        uint32_t nid[TAIL_PER], npos[TAIL_PER];
        float nb[TAIL_PER][6];
        for (int k = 0; k < TAIL_PER; ++k) {
            const int i = i0 + k;
            npos[k] = 0xffffffffu;
            if (!sv[k]) continue;
            npos[k] = SV++;
            if (!mg[k]) {
                nid[k] = sid[i];
                for (int a = 0; a < 6; ++a) nb[k][a] = sb[a][i];
                continue;
            }
            const int j = i + snn[i];
            const uint32_t id = s.next_id + S++;
            for (int a = 0; a < 3; ++a) {
                nb[k][a] = stdmin(sb[a][i], sb[a][j]);
                nb[k][3 + a] = stdmax(sb[3 + a][i], sb[3 + a][j]);
            }
            for (int a = 0; a < 6; ++a) boxes[6 * size_t(id) + a] = nb[k][a];
            links[3 * size_t(id)] = int32_t(sid[i]);
            links[3 * size_t(id) + 1] = int32_t(sid[j]);
            links[3 * size_t(id) + 2] = -1;
            nid[k] = id;
        }
        __syncthreads();
        for (int k = 0; k < TAIL_PER; ++k) {
            if (npos[k] == 0xffffffffu) continue;
            sid[npos[k]] = nid[k];
            for (int a = 0; a < 6; ++a) sb[a][npos[k]] = nb[k][a];
        }
        __syncthreads();
        s.next_id += tm;
        s.merges = tm;
        s.rounds += 1;
        m = ts;
        if (tm == 0) { s.stalled = 1; break; }
    }
    if (threadIdx.x == 0) {
        s.m = m;
        s.root = sid[0];
        *st_out = s;
    }
}

uint32_t cdiv(uint64_t a, uint32_t b) { return uint32_t((a + b - 1) / b); }

}  // namespace

namespace {
__global__ void k_ploc_init(PlocState* st, uint32_t n) {
    PlocState init{};
    init.m = n;
    init.next_id = n;
    st[0] = init;
}
}  // namespace

namespace {

// The grid of the round kernels (they stride over the tiles of the current clusters).
uint32_t round_grid(uint32_t n) { return std::min<uint32_t>(cdiv(n, BT), 2048u); }
// Round r's grid: a block a tile of the clusters the last finished build of the same n had at
// the start of round r (PLOC is deterministic: the same scene gives the same counts), else the
// full grid. The kernels stride over the tiles, so any grid gives the same tree; one block a tile
// dispatches no idle blocks, which matters when the build overlaps another frame's trace.
uint32_t round_grid(const Ctx& c, uint32_t n, uint32_t r) {
    const uint32_t G = round_grid(n);
    if (c.last_rounds_n != n || r >= c.round_m.size()) return G;
    return std::max(1u, std::min(G, cdiv(c.round_m[r], BT)));
}
uint32_t tail_threshold(const Ctx& c) { return uint32_t(std::max(1, std::min(c.ploc_tail, TAIL_M))); }

// Grid rounds to enqueue for n triangles: as many as the last finished build of n triangles ran,
// else an estimate. The clusters of these scenes shrink by a factor of about 0.83 a round with
// radius 16 (1.03 M to 4,096 in 30 rounds); 0.87 leaves room, and a build whose rounds fall short
// continues with read-backs (ploc_continue), so the estimate sets only the speed.
uint32_t predicted_rounds(const Ctx& c, uint32_t n) {
    const uint32_t tail_at = tail_threshold(c);
    if (n <= tail_at) return 0;
    if (c.build_rounds > 0) return uint32_t(c.build_rounds);
    if (c.last_rounds_n == n && c.last_grid_rounds) return c.last_grid_rounds;
    return uint32_t(std::ceil(std::log(double(n) / double(tail_at)) / std::log(1.0 / 0.87))) + 2;
}

// Enqueues `rounds` grid rounds from the state at c.ploc_parity (clusters in the buffer
// c.ploc_in_a names), then the tail into state 2.
void enqueue_rounds(Ctx& c, uint32_t rounds, int radius, hipStream_t s) {
    const uint32_t n = c.ntri, maxt = cdiv(n, BT), tail_at = tail_threshold(c);
    PlocState* d_st = c.state.as<PlocState>();
    for (uint32_t r = 0; r < rounds; ++r) {
        const uint32_t G = round_grid(c, n, c.ploc_round++);
        uint32_t* cin = c.ploc_in_a ? c.clus_a.as<uint32_t>() : c.clus_b.as<uint32_t>();
        uint32_t* cout = c.ploc_in_a ? c.clus_b.as<uint32_t>() : c.clus_a.as<uint32_t>();
        const int parity = c.ploc_parity;
        if (radius == 16)
            k_ploc_nn<16><<<G, BT, 0, s>>>(d_st + parity, radius, c.boxes.as<float>(), cin, c.nn.as<uint32_t>(),
                                           c.flags.as<uint8_t>(), c.blk.as<uint32_t>(), maxt, tail_at);
        else
            k_ploc_nn<0><<<G, BT, 0, s>>>(d_st + parity, radius, c.boxes.as<float>(), cin, c.nn.as<uint32_t>(),
                                          c.flags.as<uint8_t>(), c.blk.as<uint32_t>(), maxt, tail_at);
        k_ploc_merge<<<G, BT, 0, s>>>(d_st + parity, d_st + (parity ^ 1), c.nn.as<uint32_t>(), c.flags.as<uint8_t>(),
                                      c.blk.as<uint32_t>(), maxt, cin, cout, c.boxes.as<float>(), c.links.as<int32_t>(), tail_at,
                                      c.ploc_hist.as<uint32_t>());
        c.ploc_parity ^= 1;
        c.ploc_in_a = !c.ploc_in_a;
    }
    const uint32_t* cin = c.ploc_in_a ? c.clus_a.as<uint32_t>() : c.clus_b.as<uint32_t>();
    k_ploc_tail<<<1, TAIL_T, 0, s>>>(d_st + c.ploc_parity, d_st + 2, radius, cin, c.boxes.as<float>(), c.links.as<int32_t>());
}

// Bounds, Morton keys, sort and leaves, then the rounds' initial state.
void enqueue_front(Ctx& c, const float* d_tris, uint32_t n, hipStream_t s) {
    const uint32_t nblk_rs = cdiv(n, RS_TILE);
    const uint32_t nblk = cdiv(n, BT);
    const uint32_t nparts = std::min<uint32_t>(1024, nblk);
    k_bounds_part<<<nparts, BT, 0, s>>>(d_tris, n, c.bounds_part.as<float>());
    k_bounds_final<<<1, BT, 0, s>>>(c.bounds_part.as<float>(), int(nparts), c.g4.as<float>());
    // keys into keys2/vals2: after the three passes the sorted pairs are in keys/vals
    k_morton<<<nblk, BT, 0, s>>>(d_tris, n, c.g4.as<float>(), c.keys2.as<uint32_t>(), c.vals2.as<uint32_t>());
    uint32_t *ki = c.keys2.as<uint32_t>(), *vi = c.vals2.as<uint32_t>(), *ko = c.keys.as<uint32_t>(), *vo = c.vals.as<uint32_t>();
    for (int shift = 0; shift < 30; shift += RS_BITS) {
        k_radix_hist<<<nblk_rs, BT, 0, s>>>(ki, n, shift, c.hist.as<uint32_t>(), nblk_rs);
        k_radix_colscan<<<RS_DIG, BT, 0, s>>>(c.hist.as<uint32_t>(), nblk_rs, c.rs_tot.as<uint32_t>());
        k_radix_scatter<<<nblk_rs, BT, 0, s>>>(ki, vi, ko, vo, n, shift, c.hist.as<uint32_t>(), nblk_rs, c.rs_tot.as<uint32_t>());
        std::swap(ki, ko);
        std::swap(vi, vo);
    }
    k_leaves<<<nblk, BT, 0, s>>>(d_tris, vi, n, c.boxes.as<float>(), c.links.as<int32_t>(), c.clus_a.as<uint32_t>(),
                                 c.tris_sorted.as<float4>());
    k_ploc_init<<<1, 1, 0, s>>>(c.state.as<PlocState>(), n);
    c.ploc_parity = 0;
    c.ploc_in_a = true;
    c.ploc_round = 0;
}

}  // namespace

int ploc_build(Ctx& c, const float* d_tris, uint32_t n, int radius, hipStream_t s) {
    c.ntri = n;
    c.radius = radius;
    c.root = -1;
    c.rounds = 0;
    if (radius < 1 || radius > MAX_R) { c.err = "PLOC radius must be in 1..32"; return -1; }
    if (n == 0) return 0;
    const size_t nn = 2 * size_t(n) - 1;
    const uint32_t nblk_rs = cdiv(n, RS_TILE);
    const uint32_t nblk = cdiv(n, BT);
    if (!ensure(c, c.keys, 4 * size_t(n)) || !ensure(c, c.vals, 4 * size_t(n)) || !ensure(c, c.keys2, 4 * size_t(n)) ||
        !ensure(c, c.vals2, 4 * size_t(n)) || !ensure(c, c.hist, 4 * size_t(RS_DIG) * nblk_rs) ||
        !ensure(c, c.tris_sorted, 48 * size_t(n)) ||
        !ensure(c, c.boxes, 24 * nn) || !ensure(c, c.links, 12 * nn) || !ensure(c, c.clus_a, 4 * size_t(n)) ||
        !ensure(c, c.clus_b, 4 * size_t(n)) || !ensure(c, c.nn, 4 * size_t(n)) || !ensure(c, c.flags, size_t(n)) ||
        !ensure(c, c.blk, 8 * size_t(nblk)) || !ensure(c, c.state, 3 * sizeof(PlocState)) || !ensure(c, c.g4, 16) ||
        !ensure(c, c.bounds_part, 6 * 4 * 1024) || !ensure(c, c.rs_tot, 4 * RS_DIG) || !ensure(c, c.ploc_hist, 4 * (kPlocHist + kWideHist)))
        return -2;
    if (!c.h_state && !hip_ok(c, hipHostMalloc(reinterpret_cast<void**>(&c.h_state), 1024 + 4 * (kPlocHist + kWideHist)), "hipHostMalloc"))
        return -2;
    // Everything from the bounds to the root is enqueued at once: the round kernels read the
    // cluster count on the device, the tail finishes below its threshold, and build_finish reads
    // the final state back once, at the end of the whole build.
    enqueue_front(c, d_tris, n, s);
    const uint32_t rounds = predicted_rounds(c, n);
    enqueue_rounds(c, rounds, radius, s);
    c.plan_ploc_nb.assign(rounds, round_grid(n));
    c.root = n > 1 ? int32_t(2 * n - 2) : 0;   // the last merge's node; build_finish confirms it
    c.plan_n = n;
    c.plan_radius = radius;
    return 0;
}

int ploc_continue(Ctx& c, hipStream_t s) {
    const uint32_t n = c.ntri;
    constexpr int kBatch = 6;
    PlocState* d_st = c.state.as<PlocState>();
    PlocState* h = c.h_state;
    for (int batches = 0;; ++batches) {
        if (!hip_ok(c, hipMemcpyAsync(h, d_st + c.ploc_parity, sizeof(PlocState), hipMemcpyDeviceToHost, s), "state read") ||
            !hip_ok(c, hipStreamSynchronize(s), "PLOC rounds"))
            return -2;
        if (h->stalled) { c.err = "PLOC round made no merge (degenerate input)"; return -3; }
        if (h->m <= tail_threshold(c)) break;
        if (batches > int(n)) { c.err = "PLOC rounds did not finish"; return -3; }
        enqueue_rounds(c, kBatch, c.radius, s);   // its tail reports unfinished while m is above its threshold
        c.plan_ploc_nb.insert(c.plan_ploc_nb.end(), kBatch, round_grid(n));
    }
    enqueue_rounds(c, 0, c.radius, s);
    if (!hip_ok(c, hipGetLastError(), "PLOC tail launch") ||
        !hip_ok(c, hipMemcpyAsync(h, d_st + 2, sizeof(PlocState), hipMemcpyDeviceToHost, s), "state read") ||
        !hip_ok(c, hipStreamSynchronize(s), "PLOC build"))
        return -2;
    if (h->stalled || h->m != 1) { c.err = "PLOC round made no merge (degenerate input)"; return -3; }
    return 0;
}

// The build with the round count of the last eager ploc_build of the same triangles: no host
// copies or waits, so a stream capture can record it. PLOC is deterministic, so the same input
// takes the same rounds; ploc_fixed_check() confirms it after the run.
int ploc_build_fixed(Ctx& c, const float* d_tris, hipStream_t s) {
    const uint32_t n = c.ntri;
    const int radius = c.radius;
    if (n == 0) return 0;
    if (c.plan_n != n || c.plan_radius != radius || c.root < 0) { c.err = "no recorded PLOC plan for this scene"; return -1; }
    enqueue_front(c, d_tris, n, s);
    enqueue_rounds(c, uint32_t(c.plan_ploc_nb.size()), radius, s);
    return hip_ok(c, hipGetLastError(), "PLOC launch") ? 0 : -2;
}

// After a ploc_build_fixed has run: the tree is the recorded one (same root and rounds).
int ploc_fixed_check(Ctx& c) {
    if (c.ntri == 0) return 0;
    PlocState st{};
    if (!hip_ok(c, hipMemcpy(&st, c.state.as<PlocState>() + 2, sizeof st, hipMemcpyDeviceToHost), "state read")) return -2;
    if (st.stalled || st.m != 1 || int32_t(st.root) != c.root || st.rounds != c.rounds) {
        c.err = "the replayed PLOC build differs from the recorded one (the scene changed: build it again)";
        return -3;
    }
    return 0;
}

}  // namespace raw_ao
