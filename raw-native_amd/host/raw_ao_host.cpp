// Host side of raw_ao_hip: the reference interface over the C ABI (raw_ao_host.hpp),
// the float64 reconcile sum, the raw-cert/2 writer, and the inputs of the AO path: the
// test scene, a seeded scene generator for the large workloads, and the G-buffer
// rasterizer (src/renderer/raster.cpp:92-126) that feeds every path the same pixels.
// Compiled with -ffp-contract=off: float32 results follow the source order.
#include "raw_ao_host.hpp"
#include "../../include/raw_ao_hip.h"
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <stdexcept>

namespace raw {
namespace {
std::mutex g_mu;
raw_ao_ctx* g_ctx = nullptr;
void check(int rc) {
    if (rc != RAW_AO_OK) throw std::runtime_error(std::string("raw_ao_hip: ") + raw_ao_last_error(g_ctx));
}
const float* fl(const std::vector<Vec3>& v) { return reinterpret_cast<const float*>(v.data()); }
}  // namespace

raw_ao_ctx* device() {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ctx) {
        raw_ao_ctx* c = nullptr;
        const int rc = raw_ao_open(0, &c);
        if (rc != RAW_AO_OK) {
            const std::string m = c ? raw_ao_last_error(c) : "open failed";
            raw_ao_close(c);
            throw std::runtime_error("raw_ao_hip: " + m);
        }
        g_ctx = c;
    }
    return g_ctx;
}

void LinearAccel::build(const std::vector<Tri>& t) {
    tris = t;
    check(raw_ao_build_ploc(device(), reinterpret_cast<const float*>(tris.data()), uint32_t(tris.size()), 16));
    built = true;
}

// The context holds one BVH: computeRTAO traces the accelerator built last.
Buffer<float> computeRTAO(const GBuffer& g, const LinearAccel& accel, int samples, float radius, Arena*, int) {
    if (!accel.built) throw std::runtime_error("raw_ao_hip: LinearAccel not built");
    Buffer<float> ao;
    ao.resize(g.w, g.h);
    check(raw_ao_trace_ao(device(), fl(g.position.px), fl(g.normal.px), g.mask.px.data(), g.w, g.h, samples, radius,
                          ao.px.data()));
    return ao;
}

void computeRTVisibility(const GBuffer& g, const LinearAccel& accel, int samples, float radius, Vec3 toLight, Buffer<float>& ao,
                         Buffer<float>& sky, Buffer<float>& sun, int, bool sunForced) {
    if (!accel.built) throw std::runtime_error("raw_ao_hip: LinearAccel not built");
    ao.resize(g.w, g.h);
    sky.resize(g.w, g.h);
    sun.resize(g.w, g.h);
    const float L[3] = {toLight.x, toLight.y, toLight.z};   // normalized on the device, as the reference does
    check(raw_ao_trace_visibility(device(), fl(g.position.px), fl(g.normal.px), g.mask.px.data(), g.w, g.h, samples, radius, L,
                                  sunForced ? 1 : 0, ao.px.data(), sky.px.data(), sun.px.data()));
}

Buffer<float> computeSSAO(const GBuffer& g, int samples, float radius, Arena*, int) {
    Buffer<float> ao;
    ao.resize(g.w, g.h);
    check(raw_ao_ssao(device(), fl(g.position.px), fl(g.normal.px), g.mask.px.data(), g.w, g.h, samples, radius, ao.px.data()));
    return ao;
}

ReconcileResult reconcile(const Buffer<float>& approx, const Buffer<float>& truth, const Buffer<uint8_t>& mask, float tolerance,
                          Arena*) {
    ReconcileResult r;
    r.errorMap.resize(approx.w, approx.h);
    check(raw_ao_error_map(device(), approx.px.data(), truth.px.data(), mask.px.data(), approx.w, approx.h,
                           r.errorMap.px.data(), &r.maxError));
    double sumSq = 0;   // float64, row-major (src/cert/reconcile.cpp:9-18)
    int n = 0;
    for (size_t i = 0; i < r.errorMap.px.size(); ++i) {
        if (!mask.px[i]) continue;
        const float e = r.errorMap.px[i];
        sumSq += (double)e * e;
        ++n;
    }
    r.pixels = n;
    r.rmse = n > 0 ? (float)std::sqrt(sumSq / n) : 0.0f;
    r.withinTolerance = r.rmse <= tolerance;
    return r;
}

namespace {
std::string f4(double v) { char b[64]; std::snprintf(b, sizeof b, "%.4f", v); return b; }
std::string jnum(const std::optional<double>& v) {
    if (!v.has_value() || !std::isfinite(*v)) return "null";
    char b[64];
    std::snprintf(b, sizeof b, "%.6g", *v);
    return b;
}
std::string jstr(const std::string& s) {
    std::string o = "\"";
    for (unsigned char ch : s) {
        switch (ch) {
            case '"': o += "\\\""; break;
            case '\\': o += "\\\\"; break;
            case '\n': o += "\\n"; break;
            case '\r': o += "\\r"; break;
            case '\t': o += "\\t"; break;
            default:
                if (ch < 0x20) { char b[8]; std::snprintf(b, sizeof b, "\\u%04x", ch); o += b; }
                else o += char(ch);
        }
    }
    return o + "\"";
}
double aoFidelity(double rmse) {   // certificate_from.cpp:9-17
    if (!std::isfinite(rmse) || rmse < 0.0) return 0.0;
    double c = 1.0 / (1.0 + rmse);
    return c < 0.0 ? 0.0 : (c > 1.0 ? 1.0 : c);
}
}  // namespace

std::string exactFloat(float v) {
    if (!std::isfinite(v)) return "null";
    char b[64];
    std::snprintf(b, sizeof b, "%.9g", (double)v);
    return b;
}

std::string certificateJson(const ReconcileResult& r, const CertInputs& in) {
    const char* verdict = r.pixels == 0 ? "unverifiable" : (r.withinTolerance ? "verified" : "refuted");
    std::string o = "{";
    o += "\"claim\":" + jstr("screen-space AO matches ray-traced ground truth within tolerance") + ",";
    o += "\"verdict\":" + jstr(verdict) + ",";
    o += "\"oracle\":" + jstr("raw-rt-ao-v1") + ",";
    o += "\"evidence\":[[\"pixels\"," + jstr(std::to_string(r.pixels)) + "],[\"rmse\"," + jstr(f4(r.rmse)) +
         "],[\"maxError\"," + jstr(f4(r.maxError)) + "],[\"tolerance\"," + jstr(f4(in.tolerance)) + "]]";
    const std::optional<double> aof = r.pixels > 0 ? std::optional<double>(aoFidelity(r.rmse)) : std::nullopt;
    o += ",\"channels\":{\"ao_fidelity\":" + jnum(aof) + ",\"motion_coherence\":" + jnum(in.motionCoherence) +
         ",\"hdr_headroom\":" + jnum(in.hdrHeadroom) + "}";
    o += ",\"schema\":\"raw-cert/2\"";
    o += ",\"renderer\":" + jstr(in.renderer);
    o += ",\"params\":" + (in.paramsJson.empty() ? std::string("{}") : in.paramsJson);
    o += ",\"samples\":{\"rt\":" + std::to_string(in.rtSamples) + ",\"ss\":" + std::to_string(in.ssSamples) + "}";
    o += ",\"exact\":{\"pixels\":" + std::to_string(r.pixels) + ",\"rmse\":" + exactFloat(r.rmse) +
         ",\"maxError\":" + exactFloat(r.maxError) + ",\"tolerance\":" + exactFloat(in.tolerance) + "}";
    o += ",\"outputs\":{";
    for (size_t i = 0; i < in.outputs.size(); ++i) {
        if (i) o += ",";
        o += jstr(in.outputs[i].first) + ":" + jstr(in.outputs[i].second);
    }
    return o + "}}";
}

}  // namespace raw

// ==== inputs of the path: scenes and the G-buffer =====================================
namespace {
using raw::Vec3;
inline Vec3 add(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline Vec3 sub(Vec3 a, Vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline Vec3 mulv(Vec3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vec3 cross(Vec3 a, Vec3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline Vec3 normalize(Vec3 v) { const float l = std::sqrt(dot(v, v)); return l > 0 ? mulv(v, 1.0f / l) : v; }
inline Vec3 ld(const float* p) { return {p[0], p[1], p[2]}; }

struct Mesh {
    std::vector<Vec3> pos, nrm;
    std::vector<int32_t> idx;
    void quad(Vec3 a, Vec3 b, Vec3 c, Vec3 d, Vec3 n) {   // src/scene/scene.cpp:4-8
        const int32_t base = int32_t(pos.size());
        for (Vec3 p : {a, b, c, d}) { pos.push_back(p); nrm.push_back(n); }
        for (int32_t i : {0, 1, 2, 0, 2, 3}) idx.push_back(base + i);
    }
};

// ---- the reference's test scene (src/scene/scene.cpp:18-32) ---------------------------
Mesh testScene() {
    Mesh m;
    m.quad({-5, 0, -5}, {5, 0, -5}, {5, 0, 5}, {-5, 0, 5}, {0, 1, 0});
    const Vec3 c{0, 1, 0};
    const float h = 1.0f;
    const Vec3 p0{c.x - h, c.y - h, c.z - h}, p1{c.x + h, c.y + h, c.z + h};
    m.quad({p0.x, p0.y, p1.z}, {p1.x, p0.y, p1.z}, {p1.x, p1.y, p1.z}, {p0.x, p1.y, p1.z}, {0, 0, 1});
    m.quad({p1.x, p0.y, p0.z}, {p0.x, p0.y, p0.z}, {p0.x, p1.y, p0.z}, {p1.x, p1.y, p0.z}, {0, 0, -1});
    m.quad({p0.x, p1.y, p1.z}, {p1.x, p1.y, p1.z}, {p1.x, p1.y, p0.z}, {p0.x, p1.y, p0.z}, {0, 1, 0});
    m.quad({p0.x, p0.y, p0.z}, {p1.x, p0.y, p0.z}, {p1.x, p0.y, p1.z}, {p0.x, p0.y, p1.z}, {0, -1, 0});
    m.quad({p1.x, p0.y, p1.z}, {p1.x, p0.y, p0.z}, {p1.x, p1.y, p0.z}, {p1.x, p1.y, p1.z}, {1, 0, 0});
    m.quad({p0.x, p0.y, p0.z}, {p0.x, p0.y, p1.z}, {p0.x, p1.y, p1.z}, {p0.x, p1.y, p0.z}, {-1, 0, 0});
    return m;
}

// ---- seeded scene generator for the large workloads ---------------------------------
// A ground quad of side `ground` at y = 0 and `objects` boxes and icospheres placed
// uniformly in [-ground/2, ground/2] x [0, height] x [-ground/2, ground/2], uniformly
// rotated (Shoemake's quaternion), sizes log-uniform in [smin, smax]. A sphere with
// probability p_sphere, subdivided `subdiv` times (20 * 4^subdiv triangles); boxes have
// per-axis aspect in [0.5, 1.5]. splitmix64 draws every number in a fixed order.
struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double u() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
};

struct Ico {
    std::vector<Vec3> v;
    std::vector<int32_t> f;
};
Ico icosphere(int subdiv) {
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    const double base[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                                {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    std::vector<double> vd;
    for (auto& b : base) {
        const double l = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
        vd.insert(vd.end(), {b[0] / l, b[1] / l, b[2] / l});
    }
    std::vector<int32_t> f = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
                              3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
    for (int s = 0; s < subdiv; ++s) {
        std::map<std::pair<int32_t, int32_t>, int32_t> mid;
        auto midpoint = [&](int32_t a, int32_t b) {
            const auto key = std::make_pair(std::min(a, b), std::max(a, b));
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            double m[3];
            for (int k = 0; k < 3; ++k) m[k] = 0.5 * (vd[3 * size_t(a) + size_t(k)] + vd[3 * size_t(b) + size_t(k)]);
            const double l = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
            vd.insert(vd.end(), {m[0] / l, m[1] / l, m[2] / l});
            const int32_t id = int32_t(vd.size() / 3 - 1);
            mid.emplace(key, id);
            return id;
        };
        std::vector<int32_t> nf;
        for (size_t i = 0; i < f.size(); i += 3) {
            const int32_t a = f[i], b = f[i + 1], c = f[i + 2];
            const int32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            nf.insert(nf.end(), {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca});
        }
        f.swap(nf);
    }
    Ico r;
    for (size_t i = 0; i < vd.size(); i += 3) r.v.push_back({float(vd[i]), float(vd[i + 1]), float(vd[i + 2])});
    r.f = f;
    return r;
}

struct GenParams {
    uint64_t seed;
    int objects;
    float p_sphere, smin, smax, ground, height;
    int subdiv;
};
Mesh generate(const GenParams& gp, int32_t* nspheres) {
    Mesh m;
    const float g = gp.ground * 0.5f;
    m.quad({-g, 0, g}, {g, 0, g}, {g, 0, -g}, {-g, 0, -g}, {0, 1, 0});
    const Ico ico = icosphere(gp.subdiv);
    Rng r{gp.seed};
    const double lmin = std::log(double(gp.smin)), lmax = std::log(double(gp.smax));
    int32_t spheres = 0;
    for (int o = 0; o < gp.objects; ++o) {
        const bool sphere = r.u() < gp.p_sphere;
        const float size = float(std::exp(lmin + (lmax - lmin) * r.u()));
        const Vec3 c{float((r.u() * 2.0 - 1.0) * g), float(r.u() * gp.height), float((r.u() * 2.0 - 1.0) * g)};
        // Shoemake, "Uniform random rotations" (Graphics Gems III, 1992)
        const double u1 = r.u(), u2 = r.u() * 6.283185307179586, u3 = r.u() * 6.283185307179586;
        const double a = std::sqrt(1.0 - u1), b = std::sqrt(u1);
        const double qx = a * std::sin(u2), qy = a * std::cos(u2), qz = b * std::sin(u3), qw = b * std::cos(u3);
        const float R[9] = {float(1 - 2 * (qy * qy + qz * qz)), float(2 * (qx * qy - qz * qw)), float(2 * (qx * qz + qy * qw)),
                            float(2 * (qx * qy + qz * qw)), float(1 - 2 * (qx * qx + qz * qz)), float(2 * (qy * qz - qx * qw)),
                            float(2 * (qx * qz - qy * qw)), float(2 * (qy * qz + qx * qw)), float(1 - 2 * (qx * qx + qy * qy))};
        const auto rot = [&R](Vec3 v) {
            return Vec3{R[0] * v.x + R[1] * v.y + R[2] * v.z, R[3] * v.x + R[4] * v.y + R[5] * v.z, R[6] * v.x + R[7] * v.y + R[8] * v.z};
        };
        if (sphere) {
            ++spheres;
            const float rad = 0.5f * size;
            const int32_t base = int32_t(m.pos.size());
            for (const Vec3& v : ico.v) {
                const Vec3 n = rot(v);
                m.pos.push_back(add(c, mulv(n, rad)));
                m.nrm.push_back(n);
            }
            for (int32_t i : ico.f) m.idx.push_back(base + i);
        } else {
            const float hx = 0.5f * size * float(0.5 + r.u()), hy = 0.5f * size * float(0.5 + r.u()), hz = 0.5f * size * float(0.5 + r.u());
            const Vec3 lo{-hx, -hy, -hz}, hi{hx, hy, hz};
            const auto P = [&](float x, float y, float z) { return add(c, rot(Vec3{x, y, z})); };
            const auto face = [&](Vec3 a, Vec3 b, Vec3 cc, Vec3 d, Vec3 n) { m.quad(P(a.x, a.y, a.z), P(b.x, b.y, b.z), P(cc.x, cc.y, cc.z), P(d.x, d.y, d.z), rot(n)); };
            face({lo.x, lo.y, hi.z}, {hi.x, lo.y, hi.z}, {hi.x, hi.y, hi.z}, {lo.x, hi.y, hi.z}, {0, 0, 1});
            face({hi.x, lo.y, lo.z}, {lo.x, lo.y, lo.z}, {lo.x, hi.y, lo.z}, {hi.x, hi.y, lo.z}, {0, 0, -1});
            face({lo.x, hi.y, hi.z}, {hi.x, hi.y, hi.z}, {hi.x, hi.y, lo.z}, {lo.x, hi.y, lo.z}, {0, 1, 0});
            face({lo.x, lo.y, lo.z}, {hi.x, lo.y, lo.z}, {hi.x, lo.y, hi.z}, {lo.x, lo.y, hi.z}, {0, -1, 0});
            face({hi.x, lo.y, hi.z}, {hi.x, lo.y, lo.z}, {hi.x, hi.y, lo.z}, {hi.x, hi.y, hi.z}, {1, 0, 0});
            face({lo.x, lo.y, lo.z}, {lo.x, lo.y, hi.z}, {lo.x, hi.y, hi.z}, {lo.x, hi.y, lo.z}, {-1, 0, 0});
        }
    }
    if (nspheres) *nspheres = spheres;
    return m;
}

// ---- G-buffer rasterizer: raw::rasterize with default RasterOptions ------------------
struct M4 { float m[16]{}; };
struct V4 { float x, y, z, w; };
M4 mul(const M4& a, const M4& b) {   // raw/math/mat.hpp:7-12
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
    return {a.m[0] * v.x + a.m[1] * v.y + a.m[2] * v.z + a.m[3] * v.w, a.m[4] * v.x + a.m[5] * v.y + a.m[6] * v.z + a.m[7] * v.w,
            a.m[8] * v.x + a.m[9] * v.y + a.m[10] * v.z + a.m[11] * v.w, a.m[12] * v.x + a.m[13] * v.y + a.m[14] * v.z + a.m[15] * v.w};
}
M4 perspective(float fovy, float aspect, float n, float f) {   // raw/math/mat.hpp:21-28
    const float t = 1.0f / std::tan(fovy * 0.5f);
    M4 r;
    r.m[0] = t / aspect;
    r.m[5] = t;
    r.m[10] = (f + n) / (n - f);
    r.m[11] = (2 * f * n) / (n - f);
    r.m[14] = -1;
    return r;
}
M4 lookAt(Vec3 eye, Vec3 center, Vec3 up) {   // raw/math/mat.hpp:29-38
    const Vec3 f = normalize(sub(center, eye)), s = normalize(cross(f, up)), u = cross(s, f);
    M4 r;
    r.m[0] = s.x; r.m[1] = s.y; r.m[2] = s.z; r.m[3] = -dot(s, eye);
    r.m[4] = u.x; r.m[5] = u.y; r.m[6] = u.z; r.m[7] = -dot(u, eye);
    r.m[8] = -f.x; r.m[9] = -f.y; r.m[10] = -f.z; r.m[11] = dot(f, eye);
    r.m[15] = 1;
    return r;
}
struct CV { V4 cs; Vec3 wp, wn; };
CV lerpv(const CV& a, const CV& b, float t) {
    const auto l3 = [t](Vec3 x, Vec3 y) { return Vec3{x.x + (y.x - x.x) * t, x.y + (y.y - x.y) * t, x.z + (y.z - x.z) * t}; };
    return {{a.cs.x + (b.cs.x - a.cs.x) * t, a.cs.y + (b.cs.y - a.cs.y) * t, a.cs.z + (b.cs.z - a.cs.z) * t, a.cs.w + (b.cs.w - a.cs.w) * t},
            l3(a.wp, b.wp), l3(a.wn, b.wn)};
}
struct Target { int w, h; float* depth; float* pos; float* nrm; uint8_t* mask; };
void drawTri(Target& g, const CV v[3]) {   // src/renderer/raster.cpp:30-90
    float invw[3];
    Vec3 sp[3];
    for (int k = 0; k < 3; ++k) {
        invw[k] = 1.0f / v[k].cs.w;
        const float nx = v[k].cs.x * invw[k], ny = v[k].cs.y * invw[k];
        sp[k].x = (nx * 0.5f + 0.5f) * g.w + 0.0f;
        sp[k].y = (1.0f - (ny * 0.5f + 0.5f)) * g.h + 0.0f;
        sp[k].z = v[k].cs.w;
    }
    const int minx = std::max(0, (int)std::floor(std::min({sp[0].x, sp[1].x, sp[2].x})));
    const int maxx = std::min(g.w - 1, (int)std::ceil(std::max({sp[0].x, sp[1].x, sp[2].x})));
    const int miny = std::max(0, (int)std::floor(std::min({sp[0].y, sp[1].y, sp[2].y})));
    const int maxy = std::min(g.h - 1, (int)std::ceil(std::max({sp[0].y, sp[1].y, sp[2].y})));
    const float area = (sp[1].x - sp[0].x) * (sp[2].y - sp[0].y) - (sp[1].y - sp[0].y) * (sp[2].x - sp[0].x);
    if (std::fabs(area) < 1e-9f) return;
    for (int y = miny; y <= maxy; ++y)
        for (int x = minx; x <= maxx; ++x) {
            const float px = x + 0.5f, py = y + 0.5f;
            const float w0 = ((sp[1].x - px) * (sp[2].y - py) - (sp[1].y - py) * (sp[2].x - px)) / area;
            const float w1 = ((sp[2].x - px) * (sp[0].y - py) - (sp[2].y - py) * (sp[0].x - px)) / area;
            const float w2 = 1.0f - w0 - w1;
            if (w0 < 0 || w1 < 0 || w2 < 0) continue;
            const float iw = w0 * invw[0] + w1 * invw[1] + w2 * invw[2];
            const float depth = w0 * sp[0].z + w1 * sp[1].z + w2 * sp[2].z;
            const size_t i = size_t(y) * size_t(g.w) + size_t(x);
            if (depth >= g.depth[i]) continue;
            const auto pc = [&](Vec3 a, Vec3 b, Vec3 c) {
                return Vec3{(w0 * a.x * invw[0] + w1 * b.x * invw[1] + w2 * c.x * invw[2]) / iw,
                            (w0 * a.y * invw[0] + w1 * b.y * invw[1] + w2 * c.y * invw[2]) / iw,
                            (w0 * a.z * invw[0] + w1 * b.z * invw[1] + w2 * c.z * invw[2]) / iw};
            };
            g.depth[i] = depth;
            const Vec3 p = pc(v[0].wp, v[1].wp, v[2].wp), n = normalize(pc(v[0].wn, v[1].wn, v[2].wn));
            std::memcpy(g.pos + 3 * i, &p, 12);
            std::memcpy(g.nrm + 3 * i, &n, 12);
            g.mask[i] = 1;
        }
}
int clipNear(const CV in[3], float wNear, CV out[4]) {   // src/renderer/raster.cpp:17-27
    int n = 0;
    for (int k = 0; k < 3; ++k) {
        const CV& a = in[k];
        const CV& b = in[(k + 1) % 3];
        const bool ain = a.cs.w >= wNear, bin = b.cs.w >= wNear;
        if (ain) out[n++] = a;
        if (ain != bin) out[n++] = lerpv(a, b, (wNear - a.cs.w) / (b.cs.w - a.cs.w));
    }
    return n;
}
void copyMesh(const Mesh& m, float* vpos, float* vnrm, int32_t* idx) {
    if (vpos) std::memcpy(vpos, m.pos.data(), 12 * m.pos.size());
    if (vnrm) std::memcpy(vnrm, m.nrm.data(), 12 * m.nrm.size());
    if (idx) std::memcpy(idx, m.idx.data(), 4 * m.idx.size());
}
}  // namespace

extern "C" {

// Root-mean-square of the covered pixels of an error map: float64 sum of e*e in
// row-major order, then float(sqrt(sum / n)) (src/cert/reconcile.cpp:9-18).
float rah_rmse(const float* err, const uint8_t* mask, int w, int h, double* sum_sq, int32_t* pixels) {
    double s = 0;
    int n = 0;
    const size_t np = size_t(w) * size_t(h);
    for (size_t i = 0; i < np; ++i) {
        if (!mask[i]) continue;
        const float e = err[i];
        s += (double)e * e;
        ++n;
    }
    if (sum_sq) *sum_sq = s;
    if (pixels) *pixels = n;
    return n > 0 ? (float)std::sqrt(s / n) : 0.0f;
}

// The raw-cert/2 certificate text into out (cap bytes); returns the length needed.
size_t rah_certificate(float rmse, float max_error, int32_t pixels, float tolerance, int rt, int ss, const char* renderer,
                       const char* params_json, const char* const* output_names, const char* const* output_hashes,
                       int noutputs, int has_motion, double motion, int has_hdr, double hdr, char* out, size_t cap) {
    raw::ReconcileResult r;
    r.rmse = rmse;
    r.maxError = max_error;
    r.pixels = pixels;
    r.withinTolerance = rmse <= tolerance;
    raw::CertInputs in;
    in.renderer = renderer ? renderer : "";
    in.paramsJson = params_json ? params_json : "";
    in.rtSamples = rt;
    in.ssSamples = ss;
    in.tolerance = tolerance;
    if (has_motion) in.motionCoherence = motion;
    if (has_hdr) in.hdrHeadroom = hdr;
    for (int i = 0; i < noutputs; ++i) in.outputs.emplace_back(output_names[i], output_hashes[i]);
    const std::string s = raw::certificateJson(r, in);
    if (out && cap) {
        const size_t k = std::min(cap - 1, s.size());
        std::memcpy(out, s.data(), k);
        out[k] = 0;
    }
    return s.size();
}

// The 14-triangle test scene: returns the triangle count (NULL outputs: sizing only).
size_t rah_test_scene(float* vpos, float* vnrm, int32_t* idx) {
    const Mesh m = testScene();
    copyMesh(m, vpos, vnrm, idx);
    return m.idx.size() / 3;
}

// The seeded scene; call with NULL outputs for the sizes (vertices, triangles).
size_t rah_scene_seeded(uint64_t seed, int objects, float p_sphere, int subdiv, float smin, float smax, float ground,
                        float height, float* vpos, float* vnrm, int32_t* idx, size_t* nverts, int32_t* nspheres) {
    const Mesh m = generate(GenParams{seed, objects, p_sphere, smin, smax, ground, height, subdiv}, nspheres);
    if (nverts) *nverts = m.pos.size();
    copyMesh(m, vpos, vnrm, idx);
    return m.idx.size() / 3;
}

// A seeded orbit camera around the scene: eye at `radius` from the y axis and `height`
// up, looking at `target`.
void rah_orbit_camera(uint64_t seed, float radius, float height, const float* target, float* eye) {
    Rng r{seed ^ 0x6f7262697421ull};
    const double a = r.u() * 6.283185307179586;
    eye[0] = target[0] + float(radius * std::cos(a));
    eye[1] = height;
    eye[2] = target[2] + float(radius * std::sin(a));
}

// raw::rasterize (src/renderer/raster.cpp:92-126) of one indexed mesh, default options.
void rah_rasterize(const float* vpos, const float* vnrm, const int32_t* idx, size_t ntri, const float eye[3],
                   const float center[3], const float up[3], float fovy, float aspect, float nearZ, float farZ, int w, int h,
                   float* depth, float* position, float* normal, uint8_t* mask) {
    const size_t np = size_t(w) * size_t(h);
    std::fill(depth, depth + np, std::numeric_limits<float>::infinity());
    std::memset(position, 0, 12 * np);
    std::memset(normal, 0, 12 * np);
    std::memset(mask, 0, np);
    Target g{w, h, depth, position, normal, mask};
    const M4 vp = mul(perspective(fovy, aspect, nearZ, farZ), lookAt(ld(eye), ld(center), ld(up)));
    for (size_t i = 0; i < ntri; ++i) {
        CV v[3];
        bool behind = false;
        for (int k = 0; k < 3; ++k) {
            const size_t vi = size_t(idx[3 * i + size_t(k)]);
            const Vec3 p = ld(vpos + 3 * vi);
            v[k] = {mul(vp, V4{p.x, p.y, p.z, 1}), p, ld(vnrm + 3 * vi)};
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

}  // extern "C"
