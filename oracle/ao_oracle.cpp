// CPU oracle for raw-native's ray-traced AO path. TEST INFRASTRUCTURE ONLY: the
// checker that tests, smoke() and bench.py's cpu_baseline leg compare the HIP path
// against. The product never links, loads or calls this file.
//
// It restates the reference's float32 arithmetic operation by operation, so its
// outputs are bit-identical to the reference's on the same inputs (pinned by the
// reference's golden SHA-256 hashes and PLOC tree hashes, see tests/). Built with
// -ffp-contract=off so no multiply-add is fused.
//
// std::min(a,b) is `b < a ? b : a` and std::max(a,b) is `a < b ? b : a`; both are
// written out here as lo()/hi() because signed zeros and NaNs make the operand
// order observable in the output bits.
#include "oracle_api.h"
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <thread>
#include <utility>
#include <vector>

namespace {

struct V3 { float x, y, z; };
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 scl(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float len(V3 v) { return std::sqrt(dot(v, v)); }
// raw/math/vec.hpp:16 -- a multiply by the reciprocal, not a divide.
inline V3 nrm(V3 v) { float l = len(v); return l > 0 ? scl(v, 1.0f / l) : v; }
inline float lo(float a, float b) { return b < a ? b : a; }   // std::min(a, b)
inline float hi(float a, float b) { return a < b ? b : a; }   // std::max(a, b)
inline float ax(V3 v, int k) { return k == 0 ? v.x : k == 1 ? v.y : v.z; }
inline V3 ld3(const float* p) { return {p[0], p[1], p[2]}; }
inline void st3(float* p, V3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }

struct Tri { V3 a, b, c; };
struct Box {
    V3 mn{1e30f, 1e30f, 1e30f}, mx{-1e30f, -1e30f, -1e30f};
    void grow(V3 p) {   // raw/math/primitives.hpp:10-12
        mn = {lo(mn.x, p.x), lo(mn.y, p.y), lo(mn.z, p.z)};
        mx = {hi(mx.x, p.x), hi(mx.y, p.y), hi(mx.z, p.z)};
    }
};

// Row-interleaved threads (raw/core/parallel.hpp:13-31): output independent of the count.
template <class F> void rows(int n, int threads, F&& fn) {
    if (threads > 1 && n > 1) {
        if (threads > n) threads = n;
        std::vector<std::thread> pool;
        for (int t = 0; t < threads; ++t)
            pool.emplace_back([&, t] { for (int y = t; y < n; y += threads) fn(y); });
        for (auto& th : pool) th.join();
        return;
    }
    for (int y = 0; y < n; ++y) fn(y);
}

// hash01 (src/renderer/ray_ao.cpp:8-12).
inline float hash01(uint32_t x, uint32_t y, uint32_t s) {
    uint32_t h = x * 374761393u + y * 668265263u + s * 2246822519u;
    h = (h ^ (h >> 13)) * 1274126177u;
    h ^= h >> 16;
    return (h & 0xFFFFFFu) / float(0x1000000);
}

// The watertight frame (raw/math/primitives.hpp:17-21).
inline V3 frame(V3 v, int kz, bool swap) {
    const V3 p = kz == 0 ? V3{v.y, v.z, v.x} : kz == 1 ? V3{v.z, v.x, v.y} : v;
    return swap ? V3{p.y, p.x, p.z} : p;
}
// intersectTri (raw/math/primitives.hpp:31-64).
bool intersect(V3 ro, V3 rd, const Tri& tri, float& t, float& u, float& v) {
    const float adx = std::fabs(rd.x), ady = std::fabs(rd.y), adz = std::fabs(rd.z);
    const int kz = adx > ady ? (adx > adz ? 0 : 2) : (ady > adz ? 1 : 2);
    V3 d = frame(rd, kz, false);
    const bool sw = d.z < 0.0f;
    if (sw) d = V3{d.y, d.x, d.z};
    const float sx = d.x / d.z, sy = d.y / d.z, sz = 1.0f / d.z;
    const V3 A = frame(sub(tri.a, ro), kz, sw), B = frame(sub(tri.b, ro), kz, sw), C = frame(sub(tri.c, ro), kz, sw);
    const float Ax = A.x - sx * A.z, Ay = A.y - sy * A.z;
    const float Bx = B.x - sx * B.z, By = B.y - sy * B.z;
    const float Cx = C.x - sx * C.z, Cy = C.y - sy * C.z;
    const float U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const V3 e1 = sub(tri.b, tri.a), e2 = sub(tri.c, tri.a), p = cross(rd, e2);
    const float dm = dot(e1, p);
    if (dm > -1e-7f && dm < 1e-7f) {
        const float T = U * (sz * A.z) + V * (sz * B.z) + W * (sz * C.z), inv = 1.0f / det;
        t = T * inv; u = V * inv; v = W * inv;
        return t > 1e-4f;
    }
    const float inv = 1.0f / dm;
    const V3 tv = sub(ro, tri.a), q = cross(tv, e1);
    u = hi(0.0f, dot(tv, p) * inv);
    v = hi(0.0f, dot(rd, q) * inv);
    t = dot(e2, q) * inv;
    const float s = u + v;
    if (s > 1.0f) { u /= s; v /= s; }
    return t > 1e-4f;
}

inline Tri triAt(const float* t9, size_t i) {
    const float* p = t9 + 9 * i;
    return {ld3(p), ld3(p + 3), ld3(p + 6)};
}

// ---- binned-SAH BVH (src/renderer/bvh.cpp) ---------------------------------------
constexpr int kBins = 12, kLeaf = 4;
float boxArea(const Box& b) {   // bvh.cpp:10-13
    const V3 d = sub(b.mx, b.mn);
    return d.x < 0 ? 0.0f : 2.0f * (d.x * d.y + d.y * d.z + d.z * d.x);
}
Box unite(Box a, const Box& b) { a.grow(b.mn); a.grow(b.mx); return a; }
// The slab entry with the exit widened by 1.0000008 (bvh.cpp:21-28).
float entry(const Box& b, V3 o, V3 inv, float tMax) {
    float t0 = 0.0f, t1 = tMax;
    for (int k = 0; k < 3; ++k) {
        const float i = ax(inv, k), a = (ax(b.mn, k) - ax(o, k)) * i, c = (ax(b.mx, k) - ax(o, k)) * i;
        t0 = hi(t0, lo(a, c));
        t1 = lo(t1, hi(a, c));
    }
    return t0 <= t1 * 1.0000008f ? t0 : std::numeric_limits<float>::infinity();
}
V3 inverse(V3 d) {   // bvh.cpp:29-32
    const auto f = [](float x) { return x != 0.0f ? 1.0f / x : std::copysign(std::numeric_limits<float>::infinity(), x); };
    return {f(d.x), f(d.y), f(d.z)};
}

struct Bvh {
    struct Node { Box box; int first{0}, count{0}, right{0}; };
    std::vector<Tri> tris;
    std::vector<Node> nodes;
    void build(std::vector<Tri> in) {   // bvh.cpp:35-47
        tris = std::move(in);
        nodes.clear();
        std::vector<int> order(tris.size());
        std::vector<V3> centres(tris.size());
        for (size_t i = 0; i < tris.size(); ++i) {
            order[i] = int(i);
            centres[i] = scl(add(add(tris[i].a, tris[i].b), tris[i].c), 1.0f / 3.0f);
        }
        nodes.push_back({{}, 0, int(tris.size()), 0});
        if (!tris.empty()) split(0, order, centres);
        std::vector<Tri> sorted(tris.size());
        for (size_t i = 0; i < order.size(); ++i) sorted[i] = tris[size_t(order[i])];
        tris.swap(sorted);
    }
    static int binOf(float c, float l, float ext) { return std::min(kBins - 1, int((c - l) / ext * kBins)); }
    void split(int ni, std::vector<int>& order, std::vector<V3>& centres) {   // bvh.cpp:49-92
        Box box, cb;
        const int first = nodes[size_t(ni)].first, count = nodes[size_t(ni)].count;
        for (int i = first; i < first + count; ++i) {
            const Tri& t = tris[size_t(order[size_t(i)])];
            box.grow(t.a); box.grow(t.b); box.grow(t.c); cb.grow(centres[size_t(order[size_t(i)])]);
        }
        nodes[size_t(ni)].box = box;
        if (count <= kLeaf) return;
        int bestAxis = -1, bestBin = 0;
        float bestCost = boxArea(box) * float(count);
        for (int k = 0; k < 3; ++k) {
            const float l = ax(cb.mn, k), ext = ax(cb.mx, k) - l;
            if (ext <= 0.0f) continue;
            Box bb[kBins];
            int cnt[kBins] = {};
            for (int i = first; i < first + count; ++i) {
                const int o = order[size_t(i)], b = binOf(ax(centres[size_t(o)], k), l, ext);
                const Tri& t = tris[size_t(o)];
                bb[b].grow(t.a); bb[b].grow(t.b); bb[b].grow(t.c); ++cnt[b];
            }
            for (int s = 1; s < kBins; ++s) {
                Box L, R;
                int cl = 0, cr = 0;
                for (int b = 0; b < s; ++b) if (cnt[b]) { L = unite(L, bb[b]); cl += cnt[b]; }
                for (int b = s; b < kBins; ++b) if (cnt[b]) { R = unite(R, bb[b]); cr += cnt[b]; }
                if (!cl || !cr) continue;
                const float cost = boxArea(L) * float(cl) + boxArea(R) * float(cr);
                if (cost < bestCost) { bestCost = cost; bestAxis = k; bestBin = s; }
            }
        }
        if (bestAxis < 0) return;
        const float l = ax(cb.mn, bestAxis), ext = ax(cb.mx, bestAxis) - l;
        const auto mid = std::stable_partition(order.begin() + first, order.begin() + first + count,
                                               [&](int o) { return binOf(ax(centres[size_t(o)], bestAxis), l, ext) < bestBin; });
        const int leftCount = int(mid - (order.begin() + first));
        const int left = int(nodes.size());
        nodes.push_back({{}, first, leftCount, 0});
        nodes.push_back({{}, first + leftCount, count - leftCount, 0});
        nodes[size_t(ni)].count = 0;
        nodes[size_t(ni)].first = left;
        nodes[size_t(ni)].right = left + 1;
        split(left, order, centres);
        split(left + 1, order, centres);
    }
    bool occluded(V3 o, V3 d, float tMax) const {   // bvh.cpp:119-138
        if (nodes.empty() || tris.empty()) return false;
        const V3 inv = inverse(d);
        int stack[128], sp = 0;
        stack[sp++] = 0;
        while (sp) {
            const Node& n = nodes[size_t(stack[--sp])];
            if (entry(n.box, o, inv, tMax) == std::numeric_limits<float>::infinity()) continue;
            if (n.count) {
                for (int i = n.first; i < n.first + n.count; ++i) {
                    float t, u, v;
                    if (intersect(o, d, tris[size_t(i)], t, u, v) && t < tMax) return true;
                }
            } else {
                if (sp > 125) std::abort();
                stack[sp++] = n.right;
                stack[sp++] = n.first;
            }
        }
        return false;
    }
};

bool scanOccluded(const float* t9, size_t n, V3 o, V3 d, float tMax) {   // accel.cpp:30-35
    float t, u, v;
    for (size_t i = 0; i < n; ++i)
        if (intersect(o, d, triAt(t9, i), t, u, v) && t < tMax) return true;
    return false;
}

// basis (src/renderer/ray_ao.cpp:14-17).
void basis(V3 n, V3& t, V3& b) {
    V3 a = std::fabs(n.x) > 0.9f ? V3{0, 1, 0} : V3{1, 0, 0};
    t = nrm(cross(a, n));
    b = cross(n, t);
}
// One AO sample's ray (src/renderer/ray_ao.cpp:29-38).
inline V3 aoDir(uint32_t x, uint32_t y, uint32_t s, V3 n, V3 t, V3 b) {
    float u1 = hash01(x, y, 2 * s);
    float u2 = hash01(x, y, 2 * s + 1);
    float r = std::sqrt(u1);
    float phi = 6.2831853f * u2;
    float lx = r * std::cos(phi), ly = r * std::sin(phi);
    float lz = std::sqrt(hi(0.0f, 1.0f - u1));
    return nrm(add(add(scl(t, lx), scl(b, ly)), scl(n, lz)));
}

// ---- rasterizer (src/renderer/raster.cpp, raw/math/mat.hpp) ----------------------
struct M4 { float m[16]{}; };
struct V4 { float x, y, z, w; };
M4 ident() { M4 r; r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1; return r; }
M4 mul(const M4& a, const M4& b) {   // mat.hpp:7-12: s starts at 0 and accumulates
    M4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            float s = 0;
            for (int k = 0; k < 4; ++k) s += a.m[i * 4 + k] * b.m[k * 4 + j];
            r.m[i * 4 + j] = s;
        }
    return r;
}
V4 mul(const M4& a, V4 v) {
    return {a.m[0] * v.x + a.m[1] * v.y + a.m[2] * v.z + a.m[3] * v.w,
            a.m[4] * v.x + a.m[5] * v.y + a.m[6] * v.z + a.m[7] * v.w,
            a.m[8] * v.x + a.m[9] * v.y + a.m[10] * v.z + a.m[11] * v.w,
            a.m[12] * v.x + a.m[13] * v.y + a.m[14] * v.z + a.m[15] * v.w};
}
M4 perspective(float fovy, float aspect, float n, float f) {   // mat.hpp:21-28
    float t = 1.0f / std::tan(fovy * 0.5f);
    M4 r;
    r.m[0] = t / aspect; r.m[5] = t;
    r.m[10] = (f + n) / (n - f); r.m[11] = (2 * f * n) / (n - f);
    r.m[14] = -1;
    return r;
}
M4 lookAt(V3 eye, V3 center, V3 up) {   // mat.hpp:29-38
    V3 f = nrm(sub(center, eye));
    V3 s = nrm(cross(f, up));
    V3 u = cross(s, f);
    M4 r = ident();
    r.m[0] = s.x; r.m[1] = s.y; r.m[2] = s.z; r.m[3] = -dot(s, eye);
    r.m[4] = u.x; r.m[5] = u.y; r.m[6] = u.z; r.m[7] = -dot(u, eye);
    r.m[8] = -f.x; r.m[9] = -f.y; r.m[10] = -f.z; r.m[11] = dot(f, eye);
    return r;
}
struct CV { V4 cs; V3 wp, wn; };
CV lerp(const CV& a, const CV& b, float t) {   // raster.cpp:10-14
    const auto l3 = [t](V3 x, V3 y) { return V3{x.x + (y.x - x.x) * t, x.y + (y.y - x.y) * t, x.z + (y.z - x.z) * t}; };
    return {{a.cs.x + (b.cs.x - a.cs.x) * t, a.cs.y + (b.cs.y - a.cs.y) * t, a.cs.z + (b.cs.z - a.cs.z) * t, a.cs.w + (b.cs.w - a.cs.w) * t},
            l3(a.wp, b.wp), l3(a.wn, b.wn)};
}
int clipNear(const CV in[3], float wNear, CV out[4]) {   // raster.cpp:17-27
    int n = 0;
    for (int k = 0; k < 3; ++k) {
        const CV& a = in[k];
        const CV& b = in[(k + 1) % 3];
        const bool ain = a.cs.w >= wNear, bin = b.cs.w >= wNear;
        if (ain) out[n++] = a;
        if (ain != bin) out[n++] = lerp(a, b, (wNear - a.cs.w) / (b.cs.w - a.cs.w));
    }
    return n;
}
struct GB { int w, h; float* depth; float* pos; float* nrm; uint8_t* mask; float* alb{nullptr}; V3 albedo{}; };
void drawTri(GB& g, const CV v[3]) {   // raster.cpp:30-90 with default options
    const int w = g.w, h = g.h;
    V3 wp[3] = {v[0].wp, v[1].wp, v[2].wp}, wn[3] = {v[0].wn, v[1].wn, v[2].wn};
    float invw[3];
    V3 sp[3];
    for (int k = 0; k < 3; ++k) {
        invw[k] = 1.0f / v[k].cs.w;
        const V3 ndc{v[k].cs.x * invw[k], v[k].cs.y * invw[k], v[k].cs.z * invw[k]};
        sp[k].x = (ndc.x * 0.5f + 0.5f) * w;
        sp[k].y = (1.0f - (ndc.y * 0.5f + 0.5f)) * h;
        sp[k].z = v[k].cs.w;
        sp[k].x += 0.0f; sp[k].y += 0.0f;   // the zero jitter of the default options
    }
    int minx = std::max(0, (int)std::floor(std::min({sp[0].x, sp[1].x, sp[2].x})));
    int maxx = std::min(w - 1, (int)std::ceil(std::max({sp[0].x, sp[1].x, sp[2].x})));
    int miny = std::max(0, (int)std::floor(std::min({sp[0].y, sp[1].y, sp[2].y})));
    int maxy = std::min(h - 1, (int)std::ceil(std::max({sp[0].y, sp[1].y, sp[2].y})));
    float area = (sp[1].x - sp[0].x) * (sp[2].y - sp[0].y) - (sp[1].y - sp[0].y) * (sp[2].x - sp[0].x);
    if (std::fabs(area) < 1e-9f) return;
    for (int y = miny; y <= maxy; ++y)
        for (int x = minx; x <= maxx; ++x) {
            float px = x + 0.5f, py = y + 0.5f;
            float w0 = ((sp[1].x - px) * (sp[2].y - py) - (sp[1].y - py) * (sp[2].x - px)) / area;
            float w1 = ((sp[2].x - px) * (sp[0].y - py) - (sp[2].y - py) * (sp[0].x - px)) / area;
            float w2 = 1.0f - w0 - w1;
            if (w0 < 0 || w1 < 0 || w2 < 0) continue;
            float iw = w0 * invw[0] + w1 * invw[1] + w2 * invw[2];
            float depth = w0 * sp[0].z + w1 * sp[1].z + w2 * sp[2].z;
            const size_t i = size_t(y) * size_t(w) + size_t(x);
            if (depth >= g.depth[i]) continue;
            auto pc = [&](V3 a, V3 b, V3 c) {
                return V3{(w0 * a.x * invw[0] + w1 * b.x * invw[1] + w2 * c.x * invw[2]) / iw,
                          (w0 * a.y * invw[0] + w1 * b.y * invw[1] + w2 * c.y * invw[2]) / iw,
                          (w0 * a.z * invw[0] + w1 * b.z * invw[1] + w2 * c.z * invw[2]) / iw};
            };
            g.depth[i] = depth;
            st3(g.pos + 3 * i, pc(wp[0], wp[1], wp[2]));
            st3(g.nrm + 3 * i, nrm(pc(wn[0], wn[1], wn[2])));
            if (g.alb) st3(g.alb + 3 * i, g.albedo);   // raster.cpp:82: the mesh material's albedo
            g.mask[i] = 1;
        }
}

// ---- scene builders (src/scene/scene.cpp, src/renderer/swr_builder.hpp) -----------
struct Mesh { std::vector<V3> pos, nrm; std::vector<int32_t> idx; };
void addQuad(Mesh& m, V3 a, V3 b, V3 c, V3 d, V3 n) {   // scene.cpp:4-8
    int base = int(m.pos.size());
    for (V3 p : {a, b, c, d}) { m.pos.push_back(p); m.nrm.push_back(n); }
    for (int i : {0, 1, 2, 0, 2, 3}) m.idx.push_back(base + i);
}
void addBox(Mesh& m, V3 c, float h) {   // scene.cpp:9-17
    V3 p0{c.x - h, c.y - h, c.z - h}, p1{c.x + h, c.y + h, c.z + h};
    addQuad(m, {p0.x, p0.y, p1.z}, {p1.x, p0.y, p1.z}, {p1.x, p1.y, p1.z}, {p0.x, p1.y, p1.z}, {0, 0, 1});
    addQuad(m, {p1.x, p0.y, p0.z}, {p0.x, p0.y, p0.z}, {p0.x, p1.y, p0.z}, {p1.x, p1.y, p0.z}, {0, 0, -1});
    addQuad(m, {p0.x, p1.y, p1.z}, {p1.x, p1.y, p1.z}, {p1.x, p1.y, p0.z}, {p0.x, p1.y, p0.z}, {0, 1, 0});
    addQuad(m, {p0.x, p0.y, p0.z}, {p1.x, p0.y, p0.z}, {p1.x, p0.y, p1.z}, {p0.x, p0.y, p1.z}, {0, -1, 0});
    addQuad(m, {p1.x, p0.y, p1.z}, {p1.x, p0.y, p0.z}, {p1.x, p1.y, p0.z}, {p1.x, p1.y, p1.z}, {1, 0, 0});
    addQuad(m, {p0.x, p0.y, p0.z}, {p0.x, p0.y, p1.z}, {p0.x, p1.y, p1.z}, {p0.x, p1.y, p0.z}, {-1, 0, 0});
}

// The owned-scene geometry builder: positions and indices only (swr_builder.hpp:7-38).
struct Geo {
    std::vector<V3> pos;
    std::vector<uint32_t> idx;
    void quad(V3 a, V3 b, V3 c, V3 d) {
        const uint32_t i = uint32_t(pos.size());
        pos.push_back(a); pos.push_back(b); pos.push_back(c); pos.push_back(d);
        for (uint32_t v : {i, i + 1, i + 2, i, i + 2, i + 3}) idx.push_back(v);
    }
    void box(V3 l, V3 h) {
        quad({l.x, l.y, h.z}, {h.x, l.y, h.z}, {h.x, h.y, h.z}, {l.x, h.y, h.z});
        quad({h.x, l.y, l.z}, {l.x, l.y, l.z}, {l.x, h.y, l.z}, {h.x, h.y, l.z});
        quad({h.x, l.y, h.z}, {h.x, l.y, l.z}, {h.x, h.y, l.z}, {h.x, h.y, h.z});
        quad({l.x, l.y, l.z}, {l.x, l.y, h.z}, {l.x, h.y, h.z}, {l.x, h.y, l.z});
        quad({l.x, h.y, h.z}, {h.x, h.y, h.z}, {h.x, h.y, l.z}, {l.x, h.y, l.z});
        quad({l.x, l.y, l.z}, {h.x, l.y, l.z}, {h.x, l.y, h.z}, {l.x, l.y, h.z});
    }
    void cylinder(V3 base, float r, float h, int seg) {
        for (int i = 0; i < seg; ++i) {
            const float a0 = 6.2831853f * float(i) / float(seg), a1 = 6.2831853f * float(i + 1) / float(seg);
            const V3 p0{base.x + r * std::cos(a0), base.y, base.z - r * std::sin(a0)};
            const V3 p1{base.x + r * std::cos(a1), base.y, base.z - r * std::sin(a1)};
            quad(p0, p1, {p1.x, p1.y + h, p1.z}, {p0.x, p0.y + h, p0.z});
        }
    }
    std::vector<Tri> tris() const {
        std::vector<Tri> t(idx.size() / 3);
        for (size_t i = 0; i < t.size(); ++i) t[i] = {pos[idx[3 * i]], pos[idx[3 * i + 1]], pos[idx[3 * i + 2]]};
        return t;
    }
};
uint32_t mix32(uint32_t x) {   // rt_scenes.cpp:7-10
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}
float unit24(uint32_t h) { return float(h >> 8) / 16777216.0f; }

Geo denseBlock() {   // rt_scenes.cpp:14-38
    Geo b;
    constexpr int N = 36;
    constexpr float lot = 5.0f;
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i) {
            const uint32_t h = mix32(uint32_t(j * N + i) * 2654435761u + 17u);
            const float x0 = float(i - N / 2) * lot, z0 = float(j - N / 2) * lot;
            b.quad({x0, 0, z0 + lot}, {x0 + lot, 0, z0 + lot}, {x0 + lot, 0, z0}, {x0, 0, z0});
            const float height = 2.0f + 12.0f * unit24(h);
            b.box({x0 + 1, 0, z0 + 1}, {x0 + 4, height, z0 + 4});
            for (float y = 3.0f; y + 0.5f < height; y += 3.0f)
                b.box({x0 + 0.85f, y, z0 + 0.85f}, {x0 + 4.15f, y + 0.2f, z0 + 4.15f});
            if (h & 2u) b.cylinder({x0 + 2.5f, height, z0 + 2.5f}, 0.6f, 1.2f, 16);
            b.box({x0 + 4.2f, 0, z0 + 0.2f}, {x0 + 4.7f, 0.5f, z0 + 0.7f});
            b.box({x0 + 0.2f, 0, z0 + 4.2f}, {x0 + 0.6f, 0.4f, z0 + 4.6f});
        }
    return b;
}
std::vector<Tri> soup(size_t count, uint64_t seed) {   // rt_scenes.cpp:52-63
    std::vector<Tri> t(count);
    uint32_t r = uint32_t(seed * 2654435761u + 1u);
    const auto next = [&r] { r = mix32(r + 0x9e3779b9u); return unit24(r); };
    for (Tri& x : t) {
        const V3 c{24.0f * next() - 12.0f, 24.0f * next() - 12.0f, 24.0f * next() - 12.0f};
        x.a = add(c, V3{2 * next() - 1, 2 * next() - 1, 2 * next() - 1});
        x.b = add(c, V3{2 * next() - 1, 2 * next() - 1, 2 * next() - 1});
        x.c = add(c, V3{2 * next() - 1, 2 * next() - 1, 2 * next() - 1});
    }
    return t;
}
Geo retroRoom() {   // swr_scenes.cpp:15-47
    Geo b;
    for (int z = -16; z < 2; ++z)
        for (int x = -4; x < 4; ++x)
            b.quad({float(x), 0, float(z + 1)}, {float(x + 1), 0, float(z + 1)}, {float(x + 1), 0, float(z)}, {float(x), 0, float(z)});
    for (int z = -16; z < 2; z += 2) {
        b.quad({-4, 0, float(z)}, {-4, 0, float(z + 2)}, {-4, 3.5f, float(z + 2)}, {-4, 3.5f, float(z)});
        b.quad({4, 0, float(z + 2)}, {4, 0, float(z)}, {4, 3.5f, float(z)}, {4, 3.5f, float(z + 2)});
        b.quad({-4, 3.5f, float(z + 2)}, {4, 3.5f, float(z + 2)}, {4, 3.5f, float(z)}, {-4, 3.5f, float(z)});
    }
    b.quad({4, 0, -16}, {-4, 0, -16}, {-4, 3.5f, -16}, {4, 3.5f, -16});
    for (float z : {-3.0f, -7.0f, -11.0f})
        for (float x : {-2.5f, 2.5f}) b.box({x - 0.25f, 0, z - 0.25f}, {x + 0.25f, 3.5f, z + 0.25f});
    for (int i = 0; i < 6; ++i)
        b.box({-1.5f, 0, -16.0f + 0.4f * float(5 - i)}, {1.5f, 0.25f * float(i + 1), -16.0f + 0.4f * float(6 - i)});
    b.box({-2.2f, 0.75f, -5.6f}, {-0.8f, 0.85f, -4.6f});
    for (float x : {-2.1f, -0.95f})
        for (float z : {-5.5f, -4.75f}) b.box({x, 0, z}, {x + 0.08f, 0.75f, z + 0.08f});
    b.cylinder({-1.5f, 0.85f, -5.1f}, 0.08f, 0.35f, 8);
    const float fx = 0.8f, fz = -5.0f;
    b.box({fx - 0.22f, 0, fz - 0.1f}, {fx - 0.04f, 0.85f, fz + 0.1f});
    b.box({fx + 0.04f, 0, fz - 0.1f}, {fx + 0.22f, 0.85f, fz + 0.1f});
    b.box({fx - 0.28f, 0.85f, fz - 0.14f}, {fx + 0.28f, 1.5f, fz + 0.14f});
    b.box({fx - 0.42f, 0.9f, fz - 0.08f}, {fx - 0.3f, 1.45f, fz + 0.08f});
    b.box({fx + 0.3f, 0.9f, fz - 0.08f}, {fx + 0.42f, 1.45f, fz + 0.08f});
    b.box({fx - 0.13f, 1.52f, fz - 0.13f}, {fx + 0.13f, 1.8f, fz + 0.13f});
    return b;
}
Geo isoStreet() {   // swr_scenes.cpp:49-74
    Geo b;
    for (int z = -5; z < 5; ++z)
        for (int x = -5; x < 5; ++x) {
            const float h = 0.05f * float((x * 7 + z * 13 + 40) % 5);
            b.box({float(x), -0.5f, float(z)}, {float(x + 1), h, float(z + 1)});
        }
    const float fh[5] = {3.2f, 2.4f, 4.0f, 2.8f, 3.6f};
    for (int i = 0; i < 5; ++i) {
        b.box({-5.0f + 2.0f * float(i), 0, -5.0f}, {-3.1f + 2.0f * float(i), fh[i], -3.8f});
        b.box({-5.1f + 2.0f * float(i), fh[i], -5.1f}, {-3.0f + 2.0f * float(i), fh[i] + 0.15f, -3.7f});
    }
    for (int i = 0; i < 3; ++i) b.box({-5.0f, 0, -3.6f + 2.4f * float(i)}, {-3.8f, 2.2f + 0.6f * float(i), -1.4f + 2.4f * float(i)});
    for (float z : {-2.0f, 1.5f}) {
        b.cylinder({1.5f, 0.2f, z}, 0.06f, 2.6f, 8);
        b.box({1.35f, 2.8f, z - 0.15f}, {1.65f, 3.1f, z + 0.15f});
    }
    b.box({2.6f, 0.2f, 2.4f}, {3.2f, 0.8f, 3.0f});
    b.box({3.0f, 0.2f, 3.1f}, {3.5f, 0.7f, 3.6f});
    return b;
}
Geo clipStress() {   // swr_scenes.cpp:76-89
    Geo b;
    b.quad({-300, 0, 300}, {300, 0, 300}, {300, 0, -300}, {-300, 0, -300});
    b.quad({-0.8f, 0, 1}, {-0.8f, 0, -300}, {-0.8f, 3, -300}, {-0.8f, 3, 1});
    b.box({0.12f, 0, -1.6f}, {0.5f, 0.4f, 0.5f});
    return b;
}

// ---- PLOC (src/renderer/rt_bvh.cpp) ---------------------------------------------
struct PNode { float lo[3]{}, hi[3]{}; int left{-1}, right{-1}, tri{-1}; };
V3 centroid(const Tri& t) {   // rt_bvh.cpp:34-37
    const float third = 0.333333343f;
    return {(t.a.x + t.b.x + t.c.x) * third, (t.a.y + t.b.y + t.c.y) * third, (t.a.z + t.b.z + t.c.z) * third};
}
float pow2Scale(float extent) {   // rt_bvh.cpp:38-44
    float s = 1.0f;
    if (!(extent > 0.0f)) return s;
    while (extent * s > 1.0f) s *= 0.5f;
    while (extent * s <= 0.5f) s *= 2.0f;
    return s;
}
uint32_t spread(uint32_t v) {   // rt_bvh.cpp:25-31
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
float areaLH(const float* l, const float* h) {   // rt_bvh.cpp:46-52, factored form
    const float dx = h[0] - l[0], dy = h[1] - l[1], dz = h[2] - l[2];
    return 2.0f * (dy * (dx + dz) + dz * dx);
}
float unionArea(const PNode& a, const PNode& b) {   // rt_bvh.cpp:20-24
    float l[3], h[3];
    for (int k = 0; k < 3; ++k) { l[k] = lo(a.lo[k], b.lo[k]); h[k] = hi(a.hi[k], b.hi[k]); }
    return areaLH(l, h);
}
void mortonKeys(const std::vector<Tri>& tris, std::vector<std::pair<uint32_t, uint32_t>>& keys, float g4[4]) {   // rt_bvh.cpp:56-72
    const size_t n = tris.size();
    float l[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, h[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    std::vector<V3> c(n);
    for (size_t i = 0; i < n; ++i) {
        c[i] = centroid(tris[i]);
        const float x[3] = {c[i].x, c[i].y, c[i].z};
        for (int k = 0; k < 3; ++k) { l[k] = lo(l[k], x[k]); h[k] = hi(h[k], x[k]); }
    }
    const float s = pow2Scale(hi(h[0] - l[0], hi(h[1] - l[1], h[2] - l[2])));
    keys.resize(n);
    for (size_t i = 0; i < n; ++i) {
        const float x[3] = {c[i].x, c[i].y, c[i].z};
        uint32_t q[3];
        for (int k = 0; k < 3; ++k) q[k] = std::min(1023u, uint32_t((x[k] - l[k]) * s * 1023.0f));
        keys[i] = {(spread(q[0]) << 2) | (spread(q[1]) << 1) | spread(q[2]), uint32_t(i)};
    }
    if (g4) { g4[0] = l[0]; g4[1] = l[1]; g4[2] = l[2]; g4[3] = s; }
}
int plocBuild(const std::vector<Tri>& tris, int radius, std::vector<PNode>& nodes, uint32_t& rounds) {   // rt_bvh.cpp:54-114
    nodes.clear();
    rounds = 0;
    const size_t n = tris.size();
    if (!n) return -1;
    std::vector<std::pair<uint32_t, uint32_t>> keys;
    mortonKeys(tris, keys, nullptr);
    std::sort(keys.begin(), keys.end());
    nodes.assign(2 * n - 1, PNode{});
    std::vector<int> cl(n);
    for (size_t i = 0; i < n; ++i) {
        PNode& L = nodes[i];
        L.tri = int(keys[i].second);
        const Tri& t = tris[keys[i].second];
        const V3 v[3] = {t.a, t.b, t.c};
        for (int k = 0; k < 3; ++k) {
            const float x[3] = {v[k].x, v[k].y, v[k].z};
            for (int a = 0; a < 3; ++a) {
                L.lo[a] = k ? lo(L.lo[a], x[a]) : x[a];
                L.hi[a] = k ? hi(L.hi[a], x[a]) : x[a];
            }
        }
        cl[i] = int(i);
    }
    int next = int(n);
    std::vector<int> nn;
    while (cl.size() > 1) {
        const int m = int(cl.size());
        nn.assign(size_t(m), -1);
        for (int i = 0; i < m; ++i) {
            float best = FLT_MAX;
            nn[size_t(i)] = i;   // no candidate below FLT_MAX: the cluster waits (no merge)
            for (int j = std::max(0, i - radius); j <= std::min(m - 1, i + radius); ++j) {
                if (j == i) continue;
                const float d = unionArea(nodes[size_t(cl[size_t(i)])], nodes[size_t(cl[size_t(j)])]);
                if (d < best) { best = d; nn[size_t(i)] = j; }
            }
        }
        std::vector<int> out;
        out.reserve(size_t(m));
        for (int i = 0; i < m; ++i) {
            const int j = nn[size_t(i)];
            const bool mutual = j != i && nn[size_t(j)] == i;
            if (mutual && i > j) continue;
            if (!mutual) { out.push_back(cl[size_t(i)]); continue; }
            PNode& N = nodes[size_t(next)];
            const PNode& A = nodes[size_t(cl[size_t(i)])];
            const PNode& B = nodes[size_t(cl[size_t(j)])];
            for (int k = 0; k < 3; ++k) { N.lo[k] = lo(A.lo[k], B.lo[k]); N.hi[k] = hi(A.hi[k], B.hi[k]); }
            N.left = cl[size_t(i)];
            N.right = cl[size_t(j)];
            out.push_back(next++);
        }
        if (int(out.size()) == m) return -2;   // a round without a merge never ends
        cl.swap(out);
        ++rounds;
    }
    return cl[0];
}
uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

// ---- binned-SAH cost port (src/renderer/rt_bvh_sah.cpp) ---------------------------
struct SahPort {
    const std::vector<Tri>& tris;
    std::vector<int> order;
    std::vector<V3> centres;
    double cost{0}, rootArea{0};
    size_t nodes{0};
    void split(int first, int count) {
        ++nodes;
        Box box, cb;
        for (int i = first; i < first + count; ++i) {
            const Tri& t = tris[size_t(order[size_t(i)])];
            box.grow(t.a); box.grow(t.b); box.grow(t.c); cb.grow(centres[size_t(order[size_t(i)])]);
        }
        if (rootArea == 0.0) rootArea = boxArea(box);
        int bestAxis = -1, bestBin = 0;
        float bestCost = boxArea(box) * float(count);
        if (count > kLeaf)
            for (int k = 0; k < 3; ++k) {
                const float l = ax(cb.mn, k), ext = ax(cb.mx, k) - l;
                if (ext <= 0.0f) continue;
                Box bb[kBins];
                int cnt[kBins] = {};
                for (int i = first; i < first + count; ++i) {
                    const int o = order[size_t(i)], b = Bvh::binOf(ax(centres[size_t(o)], k), l, ext);
                    const Tri& t = tris[size_t(o)];
                    bb[b].grow(t.a); bb[b].grow(t.b); bb[b].grow(t.c); ++cnt[b];
                }
                for (int s = 1; s < kBins; ++s) {
                    Box L, R;
                    int cl = 0, cr = 0;
                    for (int b = 0; b < s; ++b) if (cnt[b]) { L = unite(L, bb[b]); cl += cnt[b]; }
                    for (int b = s; b < kBins; ++b) if (cnt[b]) { R = unite(R, bb[b]); cr += cnt[b]; }
                    if (!cl || !cr) continue;
                    const float c = boxArea(L) * float(cl) + boxArea(R) * float(cr);
                    if (c < bestCost) { bestCost = c; bestAxis = k; bestBin = s; }
                }
            }
        if (bestAxis < 0) { cost += double(boxArea(box)) * double(count) / rootArea; return; }
        cost += double(boxArea(box)) / rootArea;
        const float l = ax(cb.mn, bestAxis), ext = ax(cb.mx, bestAxis) - l;
        const auto mid = std::stable_partition(order.begin() + first, order.begin() + first + count,
                                               [&](int o) { return Bvh::binOf(ax(centres[size_t(o)], bestAxis), l, ext) < bestBin; });
        const int left = int(mid - (order.begin() + first));
        split(first, left);
        split(first + left, count - left);
    }
};

std::vector<Tri> toTris(const float* t9, size_t n) {
    std::vector<Tri> t(n);
    for (size_t i = 0; i < n; ++i) t[i] = triAt(t9, i);
    return t;
}

}  // namespace

extern "C" {

float orc_hash01(uint32_t x, uint32_t y, uint32_t s) { return hash01(x, y, s); }

void orc_ao_ray(uint32_t x, uint32_t y, uint32_t s, const float* p, const float* n, float* origin, float* dir) {
    V3 N = ld3(n), t, b;
    basis(N, t, b);
    st3(origin, add(ld3(p), scl(N, 1e-3f)));
    st3(dir, aoDir(x, y, s, N, t, b));
}

int orc_intersect(const float* ro, const float* rd, const float* tri9, float* t, float* u, float* v) {
    return intersect(ld3(ro), ld3(rd), triAt(tri9, 0), *t, *u, *v) ? 1 : 0;
}

void orc_rasterize(const float* vpos, const float* vnrm, const int32_t* idx, size_t ntri,
                   const float eye[3], const float center[3], const float up[3],
                   float fovy, float aspect, float nearZ, float farZ, int w, int h,
                   float* depth, float* position, float* normal, uint8_t* mask) {
    const size_t np = size_t(w) * size_t(h);
    for (size_t i = 0; i < np; ++i) {
        depth[i] = std::numeric_limits<float>::infinity();
        mask[i] = 0;
    }
    std::memset(position, 0, np * 3 * sizeof(float));
    std::memset(normal, 0, np * 3 * sizeof(float));
    GB g{w, h, depth, position, normal, mask};
    const M4 vp = mul(perspective(fovy, aspect, nearZ, farZ), lookAt(ld3(eye), ld3(center), ld3(up)));
    for (size_t i = 0; i < ntri; ++i) {
        CV v[3];
        bool behind = false;
        for (int k = 0; k < 3; ++k) {
            const int32_t vi = idx[3 * i + size_t(k)];
            const V3 p = ld3(vpos + 3 * size_t(vi));
            v[k] = {mul(vp, V4{p.x, p.y, p.z, 1}), p, ld3(vnrm + 3 * size_t(vi))};
            behind = behind || v[k].cs.w <= 1e-6f;
        }
        if (!behind) { drawTri(g, v); continue; }
        CV c[4];
        const int n = clipNear(v, nearZ, c);
        for (int k = 1; k + 1 < n; ++k) {
            const CV t[3] = {c[0], c[k], c[k + 1]};
            drawTri(g, t);
        }
    }
}

void orc_rasterize_scene(const float* vpos, const float* vnrm, const int32_t* idx, size_t ntri, const uint32_t* mesh_end,
                         const float* albedo, size_t nmesh, const float eye[3], const float center[3], const float up[3],
                         float fovy, float aspect, float nearZ, float farZ, int w, int h, float* depth, float* position,
                         float* normal, float* albedo_out, uint8_t* mask) {
    const size_t np = size_t(w) * size_t(h);
    for (size_t i = 0; i < np; ++i) {
        depth[i] = std::numeric_limits<float>::infinity();
        mask[i] = 0;
    }
    std::memset(position, 0, np * 3 * sizeof(float));
    std::memset(normal, 0, np * 3 * sizeof(float));
    std::memset(albedo_out, 0, np * 3 * sizeof(float));
    GB g{w, h, depth, position, normal, mask};
    g.alb = albedo_out;
    const M4 vp = mul(perspective(fovy, aspect, nearZ, farZ), lookAt(ld3(eye), ld3(center), ld3(up)));
    size_t i = 0;
    for (size_t m = 0; m < nmesh; ++m) {   // raster.cpp:107-124: meshes in order, triangles in order
        g.albedo = ld3(albedo + 3 * m);
        for (; i < mesh_end[m] && i < ntri; ++i) {
            CV v[3];
            bool behind = false;
            for (int k = 0; k < 3; ++k) {
                const int32_t vi = idx[3 * i + size_t(k)];
                const V3 p = ld3(vpos + 3 * size_t(vi));
                v[k] = {mul(vp, V4{p.x, p.y, p.z, 1}), p, ld3(vnrm + 3 * size_t(vi))};
                behind = behind || v[k].cs.w <= 1e-6f;
            }
            if (!behind) { drawTri(g, v); continue; }
            CV c[4];
            const int n = clipNear(v, nearZ, c);
            for (int k = 1; k + 1 < n; ++k) {
                const CV t[3] = {c[0], c[k], c[k + 1]};
                drawTri(g, t);
            }
        }
    }
}

void orc_view_proj(const float eye[3], const float center[3], const float up[3], float fovy, float aspect, float nearZ,
                   float farZ, float out16[16]) {
    const M4 vp = mul(perspective(fovy, aspect, nearZ, farZ), lookAt(ld3(eye), ld3(center), ld3(up)));
    std::memcpy(out16, vp.m, 64);
}

void orc_shade(const float* normal, const float* albedo, const uint8_t* mask, int w, int h, int model, const float* plane_a,
               const float* plane_sun, const float ldir[3], float li, float* frame, float* hdr) {
    const V3 l = ld3(ldir);
    const size_t np = size_t(w) * size_t(h);
    for (size_t i = 0; i < np; ++i) {
        V3 r{0, 0, 0};
        if (mask[i]) {
            const V3 n = ld3(normal + 3 * i);
            const float ndl = std::max(0.0f, dot(n, V3{l.x * -1.0f, l.y * -1.0f, l.z * -1.0f})) * li;   // composite.cpp:11,21
            float lit;
            if (model == 2) lit = 0.2f * plane_a[i] + ndl * plane_sun[i];   // composite.cpp:22
            else {
                const float ambient = 0.2f;   // composite.cpp:12-14
                lit = (ambient + ndl) * plane_a[i];
            }
            const V3 a = ld3(albedo + 3 * i);
            r = V3{a.x * lit, a.y * lit, a.z * lit};
        }
        st3(hdr + 3 * i, r);
        st3(frame + 3 * i, V3{std::clamp(r.x, 0.0f, 1.0f), std::clamp(r.y, 0.0f, 1.0f), std::clamp(r.z, 0.0f, 1.0f)});
    }
}

float orc_max_radiance(const float* hdr, size_t npix) {   // composite.cpp:82-88
    float m = 0.0f;
    for (size_t i = 0; i < npix; ++i) m = std::max(m, std::max(hdr[3 * i], std::max(hdr[3 * i + 1], hdr[3 * i + 2])));
    return m;
}

int32_t orc_motion(const float* position, const uint8_t* mask, int w, int h, const float cur16[16], const float prev16[16],
                   float* motion) {
    M4 cur, prev;
    std::memcpy(cur.m, cur16, 64);
    std::memcpy(prev.m, prev16, 64);
    const auto project = [](const M4& vp, V3 p, float& u, float& v) {   // motion.cpp:7-17
        const V4 cs = mul(vp, V4{p.x, p.y, p.z, 1.0f});
        if (cs.w <= 1e-6f) return false;
        const float invw = 1.0f / cs.w;
        const float ndcx = cs.x * invw, ndcy = cs.y * invw;
        u = ndcx * 0.5f + 0.5f;
        v = 1.0f - (ndcy * 0.5f + 0.5f);
        return true;
    };
    int32_t valid = 0;
    for (size_t i = 0; i < size_t(w) * size_t(h); ++i) {   // motion.cpp:18-31
        motion[2 * i] = 0.0f;
        motion[2 * i + 1] = 0.0f;
        if (!mask[i]) continue;
        const V3 p = ld3(position + 3 * i);
        float cu, cv, pu, pv;
        if (!project(cur, p, cu, cv) || !project(prev, p, pu, pv)) continue;
        const float mx = cu - pu, my = cv - pv;
        if (!std::isfinite(mx) || !std::isfinite(my)) continue;
        motion[2 * i] = mx;
        motion[2 * i + 1] = my;
        ++valid;
    }
    return valid;
}

size_t orc_test_scene(float* vpos, float* vnrm, int32_t* idx) {
    Mesh m;
    addQuad(m, {-5, 0, -5}, {5, 0, -5}, {5, 0, 5}, {-5, 0, 5}, {0, 1, 0});
    addBox(m, {0, 1, 0}, 1.0f);
    const size_t ntri = m.idx.size() / 3;
    if (vpos) for (size_t i = 0; i < m.pos.size(); ++i) { st3(vpos + 3 * i, m.pos[i]); st3(vnrm + 3 * i, m.nrm[i]); }
    if (idx) std::memcpy(idx, m.idx.data(), m.idx.size() * sizeof(int32_t));
    return ntri;
}

void* orc_bvh_build(const float* tris9, size_t ntri) {
    Bvh* b = new Bvh;
    b->build(toTris(tris9, ntri));
    return b;
}
void orc_bvh_free(void* bvh) { delete static_cast<Bvh*>(bvh); }
size_t orc_bvh_node_count(void* bvh) { return static_cast<Bvh*>(bvh)->nodes.size(); }
int orc_bvh_occluded(void* bvh, const float* o, const float* d, float tmax) {
    return static_cast<Bvh*>(bvh)->occluded(ld3(o), ld3(d), tmax) ? 1 : 0;
}

void orc_rtao(void* bvh, const float* tris9, size_t ntri, int mode,
              const float* position, const float* normal, const uint8_t* mask,
              int w, int h, int y0, int y1, int samples, float radius, int threads, float* ao) {
    orc_rtao_rows(bvh, tris9, ntri, mode, position, normal, mask, w, h, y0, y1, 1, samples, radius, threads, ao);
}

void orc_rtao_rows(void* bvh, const float* tris9, size_t ntri, int mode,
                   const float* position, const float* normal, const uint8_t* mask,
                   int w, int h, int y0, int y1, int ystep, int samples, float radius, int threads, float* ao) {
    const Bvh* B = static_cast<const Bvh*>(bvh);
    if (ystep < 1) ystep = 1;
    const int nrows = y1 > y0 ? (y1 - y0 + ystep - 1) / ystep : 0;
    rows(nrows, threads, [&](int r) {
        const int y = y0 + r * ystep;
        for (int x = 0; x < w; ++x) {
            const size_t i = size_t(y) * size_t(w) + size_t(x);
            if (!mask[i]) { ao[i] = 1.0f; continue; }
            const V3 p = ld3(position + 3 * i), n = ld3(normal + 3 * i);
            V3 t, b;
            basis(n, t, b);
            const V3 origin = add(p, scl(n, 1e-3f));
            int open = 0;
            for (int s = 0; s < samples; ++s) {
                const V3 dir = aoDir(uint32_t(x), uint32_t(y), uint32_t(s), n, t, b);
                bool occ;
                if (mode == 1) occ = scanOccluded(tris9, ntri, origin, dir, radius);
                else occ = B->occluded(origin, dir, mode == 2 ? radius * 0.99f : radius);
                if (!occ) ++open;
            }
            ao[i] = (float)open / (float)samples;
        }
    });
}

void orc_rtvis(void* bvh, const float* tris9, size_t ntri, int mode, const float* position, const float* normal,
               const uint8_t* mask, int w, int h, int samples, float radius, const float to_light[3], int sun_forced,
               int threads, float* ao, float* sky, float* sun) {
    // src/renderer/ray_ao.cpp:20-50 (the shared loop with sky and sun set), :59-63 (toLight normalized)
    const Bvh* B = static_cast<const Bvh*>(bvh);
    constexpr float kFar = 1e30f;
    const V3 L = nrm(ld3(to_light));
    const auto occluded = [&](V3 o, V3 d, float tmax) {
        if (mode == 1) return scanOccluded(tris9, ntri, o, d, tmax);
        return B->occluded(o, d, mode == 2 ? tmax * 0.99f : tmax);
    };
    rows(h, threads, [&](int y) {
        for (int x = 0; x < w; ++x) {
            const size_t i = size_t(y) * size_t(w) + size_t(x);
            if (!mask[i]) { ao[i] = 1.0f; sky[i] = 1.0f; sun[i] = 1.0f; continue; }
            const V3 p = ld3(position + 3 * i), n = ld3(normal + 3 * i);
            V3 t, b;
            basis(n, t, b);
            const V3 origin = add(p, scl(n, 1e-3f));
            int open = 0, free = 0;
            for (int s = 0; s < samples; ++s) {
                const V3 dir = aoDir(uint32_t(x), uint32_t(y), uint32_t(s), n, t, b);
                if (!occluded(origin, dir, radius)) {
                    ++open;
                    if (!occluded(origin, dir, kFar)) ++free;
                }
            }
            ao[i] = (float)open / (float)samples;
            sky[i] = (float)free / (float)samples;
            sun[i] = (sun_forced || !occluded(origin, L, kFar)) ? 1.0f : 0.0f;
        }
    });
}

void orc_ssao(const float* position, const float* normal, const uint8_t* mask,
              int w, int h, int samples, float radius, int threads, float* ao) {
    const int R = 6;
    rows(h, threads, [&](int y) {
        for (int x = 0; x < w; ++x) {
            const size_t i = size_t(y) * size_t(w) + size_t(x);
            if (!mask[i]) { ao[i] = 1.0f; continue; }
            const V3 p = ld3(position + 3 * i), n = ld3(normal + 3 * i);
            int occ = 0, used = 0;
            for (int s = 0; s < samples; ++s) {
                float a = hash01(uint32_t(x), uint32_t(y), uint32_t(s)) * 6.2831853f;
                float rad = (0.3f + 0.7f * hash01(uint32_t(x), uint32_t(y), uint32_t(s + 97))) * R;
                int sx = x + (int)std::lround(std::cos(a) * rad);
                int sy = y + (int)std::lround(std::sin(a) * rad);
                if (sx < 0 || sy < 0 || sx >= w || sy >= h) continue;
                const size_t j = size_t(sy) * size_t(w) + size_t(sx);
                if (!mask[j]) continue;
                ++used;
                V3 d = sub(ld3(position + 3 * j), p);
                float dist = len(d);
                if (dist < 1e-4f || dist > radius) continue;
                float ndl = dot(n, nrm(d));
                if (ndl > 0.15f) ++occ;
            }
            float a = used > 0 ? 1.0f - (float)occ / (float)used : 1.0f;
            ao[i] = a < 0 ? 0 : a;
        }
    });
}

void orc_reconcile(const float* approx, const float* truth, const uint8_t* mask, int w, int h, float tolerance,
                   float* error_map, float* rmse, float* max_error, int32_t* pixels, int32_t* within) {
    double sumSq = 0;
    int n = 0;
    float mx = 0;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const size_t i = size_t(y) * size_t(w) + size_t(x);
            if (!mask[i]) { error_map[i] = 0; continue; }
            float e = std::fabs(approx[i] - truth[i]);
            error_map[i] = e;
            if (e > mx) mx = e;
            sumSq += (double)e * e;
            ++n;
        }
    *pixels = n;
    *max_error = mx;
    *rmse = n > 0 ? (float)std::sqrt(sumSq / n) : 0.0f;
    *within = *rmse <= tolerance ? 1 : 0;
}

int32_t orc_ploc_build(const float* tris9, size_t ntri, int radius, float* boxes6, int32_t* links3, uint32_t* rounds) {
    std::vector<PNode> nodes;
    uint32_t r = 0;
    const int root = plocBuild(toTris(tris9, ntri), radius, nodes, r);
    for (size_t i = 0; i < nodes.size(); ++i) {
        for (int k = 0; k < 3; ++k) { boxes6[6 * i + size_t(k)] = nodes[i].lo[k]; boxes6[6 * i + 3 + size_t(k)] = nodes[i].hi[k]; }
        links3[3 * i] = nodes[i].left; links3[3 * i + 1] = nodes[i].right; links3[3 * i + 2] = nodes[i].tri;
    }
    if (rounds) *rounds = r;
    return root;
}

void orc_ploc_keys(const float* tris9, size_t ntri, uint32_t* keys, uint32_t* order, float* g4) {
    std::vector<std::pair<uint32_t, uint32_t>> k;
    mortonKeys(toTris(tris9, ntri), k, g4);
    for (size_t i = 0; i < k.size(); ++i) keys[i] = k[i].first;
    std::sort(k.begin(), k.end());
    for (size_t i = 0; i < k.size(); ++i) order[i] = k[i].second;
}

uint64_t orc_tree_hash(const float* boxes6, const int32_t* links3, size_t nnodes) {
    uint64_t h = 1469598103934665603ull;
    const auto mixv = [&h](uint32_t v) { for (int i = 0; i < 4; ++i) { h ^= (v >> (8 * i)) & 255u; h *= 1099511628211ull; } };
    for (size_t i = 0; i < nnodes; ++i) {
        mixv(uint32_t(links3[3 * i])); mixv(uint32_t(links3[3 * i + 1])); mixv(uint32_t(links3[3 * i + 2]));
        for (int k = 0; k < 3; ++k) { mixv(fbits(boxes6[6 * i + size_t(k)])); mixv(fbits(boxes6[6 * i + 3 + size_t(k)])); }
    }
    return h;
}

double orc_sah_cost(const float* boxes6, size_t nnodes, int32_t root) {
    if (root < 0) return 0.0;
    const double ra = areaLH(boxes6 + 6 * size_t(root), boxes6 + 6 * size_t(root) + 3);
    double c = 0.0;
    for (size_t i = 0; i < nnodes; ++i) c += double(areaLH(boxes6 + 6 * i, boxes6 + 6 * i + 3)) / ra;
    return c;
}

double orc_binned_sah_cost(const float* tris9, size_t ntri, size_t* node_count) {
    const std::vector<Tri> tris = toTris(tris9, ntri);
    SahPort p{tris, {}, {}};
    p.order.resize(tris.size());
    p.centres.resize(tris.size());
    for (size_t i = 0; i < tris.size(); ++i) {
        p.order[i] = int(i);
        p.centres[i] = scl(add(add(tris[i].a, tris[i].b), tris[i].c), 1.0f / 3.0f);
    }
    if (!tris.empty()) p.split(0, int(tris.size()));
    if (node_count) *node_count = p.nodes;
    return p.cost;
}

void orc_scan_rays(const float* tris9, size_t ntri, const float* rays6, const float* tmax, size_t nrays, int any,
                   int threads, float* t_out, int32_t* tri_out) {
    rows(int(nrays), threads, [&](int k) {
        const V3 o = ld3(rays6 + 6 * size_t(k)), d = ld3(rays6 + 6 * size_t(k) + 3);
        float best = tmax[k];
        int tri = -1;
        for (size_t i = 0; i < ntri; ++i) {
            float t, u, v;
            if (!intersect(o, d, triAt(tris9, i), t, u, v) || !(t < best)) continue;
            best = t;
            tri = int(i);
            if (any) break;
        }
        t_out[k] = tri >= 0 ? best : -1.0f;
        tri_out[k] = tri;
    });
}

// makeRays (src/renderer/gpu/rt_bvh_parity.cpp:36-48): xorshift64 draws; with a camera every odd
// ray leaves the eye through a jittered point of the view, the rest start in a 24^3 cube with a
// random direction. cam_first_u: whether the camera ray's first draw is the s (right) offset,
// the left-to-right evaluation of f + s * (next() - 0.5) + u * (next() - 0.5).
void orc_r2_rays(const float* eye, const float* target, const float* up, size_t n, uint64_t seed, int cam_first_u,
                 float* rays8) {
    uint64_t st = seed;
    const auto next = [&st]() {
        st ^= st << 13; st ^= st >> 7; st ^= st << 17;
        return float(double(st >> 11) * (1.0 / 9007199254740992.0));
    };
    for (size_t k = 0; k < n; ++k) {
        V3 o, d;
        if (eye && (k % 2)) {
            const V3 e = ld3(eye), tg = ld3(target), uu = ld3(up);
            const V3 f = nrm(sub(tg, e)), sv = nrm(cross(f, uu)), u = cross(sv, f);
            float a, b;
            if (cam_first_u) { a = next(); b = next(); }
            else { b = next(); a = next(); }
            const V3 js = V3{sv.x * (a - 0.5f), sv.y * (a - 0.5f), sv.z * (a - 0.5f)};
            const V3 ju = V3{u.x * (b - 0.5f), u.y * (b - 0.5f), u.z * (b - 0.5f)};
            const V3 fs = V3{f.x + js.x, f.y + js.y, f.z + js.z};
            o = e;
            d = nrm(V3{fs.x + ju.x, fs.y + ju.y, fs.z + ju.z});
        } else {
            const float x = 24 * next() - 12, y = 24 * next() - 12, z = 24 * next() - 12;
            o = V3{x, y, z};
            const float dx = next() - 0.5f, dy = next() - 0.5f, dz = next() - 0.5f;
            d = nrm(V3{dx, dy, dz});
        }
        st3(rays8 + 8 * k, o);
        st3(rays8 + 8 * k + 3, d);
        rays8[8 * k + 6] = 1e30f;
        rays8[8 * k + 7] = 0.0f;
    }
}

// bruteForce (rt_bvh_parity.cpp:19-34): the nearest hit of a scan of every triangle (ties to the
// lower index), its u, v, and the second-nearest t (t2, -1 without a hit), over t < 1e30.
void orc_scan_hits(const float* tris9, size_t ntri, const float* rays8, size_t nrays, int threads, float* hits4, float* t2) {
    rows(int(nrays), threads, [&](int k) {
        const V3 o = ld3(rays8 + 8 * size_t(k)), d = ld3(rays8 + 8 * size_t(k) + 3);
        float best = 1e30f, second = 1e30f, bu = 0, bv = 0;
        int tri = -1;
        for (size_t i = 0; i < ntri; ++i) {
            float t, u, v;
            if (!intersect(o, d, triAt(tris9, i), t, u, v)) continue;
            if (t < best) { second = best; best = t; tri = int(i); bu = u; bv = v; }
            else if (t < second) second = t;
        }
        float* h = hits4 + 4 * size_t(k);
        h[0] = tri >= 0 ? best : -1.0f;
        h[1] = tri >= 0 ? bu : 0.0f;
        h[2] = tri >= 0 ? bv : 0.0f;
        std::memcpy(&h[3], &tri, 4);
        if (t2) t2[k] = tri >= 0 ? second : -1.0f;
    });
}

// rt::closest (src/renderer/rt_bvh.cpp:125-152) over a binary tree in the reference's node
// layout (boxes6 {lo, hi}, links3 {left, right, tri}); the ray's own tMax. hits4 as above.
void orc_tree_closest(const float* boxes6, const int32_t* links3, int32_t root, const float* tris9, const float* rays8,
                      size_t nrays, int threads, float* hits4) {
    const float kSlab = 1.0f + 2.0f * (3.0f * 5.96046448e-8f) / (1.0f - 3.0f * 5.96046448e-8f);
    rows(int(nrays), threads, [&](int k) {
        const V3 ro = ld3(rays8 + 8 * size_t(k)), rd = ld3(rays8 + 8 * size_t(k) + 3);
        const float tmax = rays8[8 * size_t(k) + 6];
        float* h = hits4 + 4 * size_t(k);
        h[0] = -1.0f; h[1] = 0.0f; h[2] = 0.0f;
        int32_t none = -1;
        std::memcpy(&h[3], &none, 4);
        if (root < 0) return;
        const auto inv = [](float x) { return x != 0.0f ? 1.0f / x : std::copysign(std::numeric_limits<float>::infinity(), x); };
        const float o[3] = {ro.x, ro.y, ro.z}, iv[3] = {inv(rd.x), inv(rd.y), inv(rd.z)};
        int stack[256], sp = 0;
        stack[sp++] = root;
        bool any = false;
        float best = tmax;
        int32_t htri = -1;
        float hu = 0, hv = 0;
        while (sp) {
            const int32_t ni = stack[--sp];
            const float* b = boxes6 + 6 * size_t(ni);
            float t0 = 0.0f, t1 = best;
            for (int a = 0; a < 3; ++a) {
                const float x = (b[a] - o[a]) * iv[a], y = (b[3 + a] - o[a]) * iv[a];
                t0 = std::max(t0, std::min(x, y));
                t1 = std::min(t1, std::max(x, y));
            }
            if (!(t0 <= t1 * kSlab)) continue;
            const int32_t tri = links3[3 * size_t(ni) + 2];
            if (tri >= 0) {
                float tt, u, v;
                if (intersect(ro, rd, triAt(tris9, size_t(tri)), tt, u, v) && (tt < best || (any && tt == best && tri < htri))) {
                    best = tt; htri = tri; hu = u; hv = v; any = true;
                }
                continue;
            }
            if (sp > 253) std::abort();
            stack[sp++] = links3[3 * size_t(ni) + 1];
            stack[sp++] = links3[3 * size_t(ni)];
        }
        if (any) { h[0] = best; h[1] = hu; h[2] = hv; std::memcpy(&h[3], &htri, 4); }
    });
}

size_t orc_scene_tris(const char* name, size_t count, uint64_t seed, float* tris9) {
    const std::string s(name);
    std::vector<Tri> t;
    if (s == "dense_block") t = denseBlock().tris();
    else if (s == "retro_room") t = retroRoom().tris();
    else if (s == "iso_street") t = isoStreet().tris();
    else if (s == "clip_stress") t = clipStress().tris();
    else if (s == "soup") t = soup(count, seed);
    else return 0;
    if (tris9)
        for (size_t i = 0; i < t.size(); ++i) { st3(tris9 + 9 * i, t[i].a); st3(tris9 + 9 * i + 3, t[i].b); st3(tris9 + 9 * i + 6, t[i].c); }
    return t.size();
}

}  // extern "C"

extern "C" {
// The host libm's std::sin and std::cos of every sample angle phi = 6.2831853f * k / 2^24
// (the argument of src/renderer/ray_ao.cpp:35-36 and src/renderer/ssao.cpp:21-24).
void orc_libm_angles(float* s, float* c) {
    rows(1 << 12, int(std::max(1u, std::thread::hardware_concurrency())), [&](int blk) {
        for (uint32_t k = uint32_t(blk) << 12; k < (uint32_t(blk + 1) << 12); ++k) {
            const float phi = 6.2831853f * (float(k) / float(0x1000000));
            s[k] = std::sin(phi);
            c[k] = std::cos(phi);
        }
    });
}
}
