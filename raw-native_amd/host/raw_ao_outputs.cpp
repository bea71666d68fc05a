// The default frame from the GPU and the files a render writes, byte for byte the reference
// CLI's (raw-native 0.8.0):
//   renderGpu            raw/renderer/gpu.hpp:32, over raw_ao_render_frame
//   images               writePPM / writePFM / writePFM1 / writePGM / writeMaskPGM
//                        (src/core/image.cpp), and the readers verify needs
//   certificate.json     raw-cert/2 (src/tools/run.cpp:45-66, src/cert/certificate*.cpp)
//   channels.json        raw-channels/1 (src/tools/channels_json.cpp)
//   receipt.json         superstack.receipt/1 (src/tools/receipt.cpp) in canonical JSON v2,
//                        the rules of superstack v0.2.0 (third_party/superstack, vendored by the
//                        reference): keys sorted by byte order, numbers in the shortest
//                        round-trip digits written positionally, the seal = SHA-256 of the
//                        canonical body
//   verify               raw_native_cli verify DIR (src/tools/verify.cpp)
// Compiled with -ffp-contract=off: float32 results follow the source order.
#include "raw_ao_host.hpp"
#include "../../include/raw_ao_hip.h"
#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <optional>
#include <sstream>
#include <stdexcept>

namespace raw {

// ==== the frame ========================================================================
namespace {
Vec3 vadd(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
Vec3 vscl(Vec3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
float vdot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
Vec3 vnorm(Vec3 v) { const float l = std::sqrt(vdot(v, v)); return l > 0 ? vscl(v, 1.0f / l) : v; }
}  // namespace

Mat4 Camera::viewProj() const {
    Mat4 m;
    // perspective (mat.hpp:21-27)
    Mat4 pr;
    const float t = 1.0f / std::tan(fovy * 0.5f);
    pr.m[0] = t / aspect;
    pr.m[5] = t;
    pr.m[10] = (farZ + nearZ) / (nearZ - farZ);
    pr.m[11] = (2 * farZ * nearZ) / (nearZ - farZ);
    pr.m[14] = -1;
    // lookAt (mat.hpp:28-37)
    const Vec3 E = eye, C = center, U = up;
    const Vec3 f = vnorm(Vec3{C.x - E.x, C.y - E.y, C.z - E.z});
    const Vec3 s = vnorm(Vec3{f.y * U.z - f.z * U.y, f.z * U.x - f.x * U.z, f.x * U.y - f.y * U.x});
    const Vec3 uu{s.y * f.z - s.z * f.y, s.z * f.x - s.x * f.z, s.x * f.y - s.y * f.x};
    Mat4 la;
    la.m[0] = s.x; la.m[1] = s.y; la.m[2] = s.z; la.m[3] = -vdot(s, E);
    la.m[4] = uu.x; la.m[5] = uu.y; la.m[6] = uu.z; la.m[7] = -vdot(uu, E);
    la.m[8] = -f.x; la.m[9] = -f.y; la.m[10] = -f.z; la.m[11] = vdot(f, E);
    la.m[15] = 1;
    for (int i = 0; i < 4; ++i)   // mul (mat.hpp:7-12)
        for (int j = 0; j < 4; ++j) {
            float acc = 0;
            for (int k = 0; k < 4; ++k) acc += pr.m[i * 4 + k] * la.m[k * 4 + j];
            m.m[i * 4 + j] = acc;
        }
    return m;
}

Scene buildTestScene(int w, int h) {
    Scene s;
    const auto quad = [](Mesh& m, Vec3 a, Vec3 b, Vec3 c, Vec3 d, Vec3 n) {   // scene.cpp:4-8
        const int base = int(m.positions.size());
        for (Vec3 p : {a, b, c, d}) { m.positions.push_back(p); m.normals.push_back(n); }
        for (int i : {0, 1, 2, 0, 2, 3}) m.indices.push_back(base + i);
    };
    Mesh plane;
    plane.material.albedo = {0.7f, 0.7f, 0.7f};
    quad(plane, {-5, 0, -5}, {5, 0, -5}, {5, 0, 5}, {-5, 0, 5}, {0, 1, 0});
    Mesh box;
    box.material.albedo = {0.8f, 0.3f, 0.2f};
    const Vec3 c{0, 1, 0};
    const float hh = 1.0f;
    const Vec3 a{c.x - hh, c.y - hh, c.z - hh}, b{c.x + hh, c.y + hh, c.z + hh};   // scene.cpp:9-17
    quad(box, {a.x, a.y, b.z}, {b.x, a.y, b.z}, {b.x, b.y, b.z}, {a.x, b.y, b.z}, {0, 0, 1});
    quad(box, {b.x, a.y, a.z}, {a.x, a.y, a.z}, {a.x, b.y, a.z}, {b.x, b.y, a.z}, {0, 0, -1});
    quad(box, {a.x, b.y, b.z}, {b.x, b.y, b.z}, {b.x, b.y, a.z}, {a.x, b.y, a.z}, {0, 1, 0});
    quad(box, {a.x, a.y, a.z}, {b.x, a.y, a.z}, {b.x, a.y, b.z}, {a.x, a.y, b.z}, {0, -1, 0});
    quad(box, {b.x, a.y, b.z}, {b.x, a.y, a.z}, {b.x, b.y, a.z}, {b.x, b.y, b.z}, {1, 0, 0});
    quad(box, {a.x, a.y, a.z}, {a.x, a.y, b.z}, {a.x, b.y, b.z}, {a.x, b.y, a.z}, {-1, 0, 0});
    s.meshes.push_back(plane);
    s.meshes.push_back(box);
    s.lights.push_back(Light{vnorm(Vec3{-0.4f, -1.0f, -0.3f}), 1.0f});
    s.camera.eye = {4, 4, 6};
    s.camera.center = {0, 1, 0};
    s.camera.up = {0, 1, 0};
    s.camera.fovy = 0.9f;
    s.camera.aspect = (float)w / (float)h;
    return s;
}

int lightingModel(const RenderOptions& o) {
    if (!o.rtao) return 1;
    const char* v = std::getenv("RAW_NATIVE_LIGHTING");
    if (v && std::strcmp(v, "v1") == 0) return 1;
    if (v && (std::strcmp(v, "v2") == 0 || std::strcmp(v, "v2-nosun") == 0)) return 2;
    return o.lighting == 1 ? 1 : 2;
}
bool lightingNoSun() {
    const char* v = std::getenv("RAW_NATIVE_LIGHTING");
    return v && std::strcmp(v, "v2-nosun") == 0;
}

float maxRadiance(const Buffer<Vec3>& hdr) {
    float m = 0.0f;
    for (const Vec3& c : hdr.px) m = std::max(m, std::max(c.x, std::max(c.y, c.z)));
    return m;
}

bool renderGpu(const Scene& scene, int w, int h, const Mat4& prevVP, const RenderOptions& opts, FrameResult& r,
               std::string& err) {
    raw_ao_ctx* ctx = nullptr;
    try {
        ctx = device();
    } catch (const std::exception& e) {
        err = e.what();
        return false;
    }
    // meshes concatenated in scene order; each mesh's indices offset by its first vertex
    std::vector<float> vpos, vnrm, alb;
    std::vector<int32_t> idx;
    std::vector<uint32_t> ends;
    for (const Mesh& m : scene.meshes) {
        const int base = int(vpos.size() / 3);
        for (size_t v = 0; v < m.positions.size(); ++v) {
            vpos.insert(vpos.end(), {m.positions[v].x, m.positions[v].y, m.positions[v].z});
            const Vec3 n = v < m.normals.size() ? m.normals[v] : Vec3{};
            vnrm.insert(vnrm.end(), {n.x, n.y, n.z});
        }
        for (size_t i = 0; i + 2 < m.indices.size(); i += 3)
            for (int k = 0; k < 3; ++k) idx.push_back(base + m.indices[i + size_t(k)]);
        ends.push_back(uint32_t(idx.size() / 3));
        alb.insert(alb.end(), {m.material.albedo.x, m.material.albedo.y, m.material.albedo.z});
    }
    if (ends.empty()) {   // no mesh: one empty one, so the frame still has a table
        ends.push_back(0);
        alb.insert(alb.end(), {0.0f, 0.0f, 0.0f});
    }
    const Camera& c = scene.camera;
    raw_ao_scene sc{};
    sc.vpos = vpos.data();
    sc.vnrm = vnrm.data();
    sc.nverts = uint32_t(vpos.size() / 3);
    sc.idx = idx.data();
    sc.ntri = uint32_t(idx.size() / 3);
    sc.mesh_end = ends.data();
    sc.albedo = alb.data();
    sc.nmesh = uint32_t(ends.size());
    const float e3[3] = {c.eye.x, c.eye.y, c.eye.z}, c3[3] = {c.center.x, c.center.y, c.center.z}, u3[3] = {c.up.x, c.up.y, c.up.z};
    std::memcpy(sc.eye, e3, 12);
    std::memcpy(sc.center, c3, 12);
    std::memcpy(sc.up, u3, 12);
    sc.fovy = c.fovy;
    sc.aspect = c.aspect;
    sc.near_z = c.nearZ;
    sc.far_z = c.farZ;
    sc.has_light = scene.lights.empty() ? 0 : 1;
    if (!scene.lights.empty()) {
        const float d[3] = {scene.lights[0].dir.x, scene.lights[0].dir.y, scene.lights[0].dir.z};
        std::memcpy(sc.light_dir, d, 12);
        sc.light_intensity = scene.lights[0].intensity;
    }
    raw_ao_render_opts o;
    raw_ao_render_opts_default(&o);
    o.tolerance = opts.tolerance;
    o.rtao = opts.rtao ? 1 : 0;
    o.lighting = lightingModel(opts);
    o.no_sun = lightingNoSun() ? 1 : 0;
    o.prev_viewproj = prevVP.m;
    r = FrameResult{};
    r.g.resize(w, h);
    r.aoSS.resize(w, h);
    r.frame.resize(w, h);
    r.hdr.resize(w, h);
    r.rec.errorMap.resize(w, h);
    Buffer<float> sky, sun;
    if (opts.rtao) r.aoRT.resize(w, h);
    if (opts.rtao && o.lighting == 2) { sky.resize(w, h); sun.resize(w, h); }
    raw_ao_frame f{};
    f.depth = r.g.depth.px.data();
    f.position = reinterpret_cast<float*>(r.g.position.px.data());
    f.normal = reinterpret_cast<float*>(r.g.normal.px.data());
    f.albedo = reinterpret_cast<float*>(r.g.albedo.px.data());
    f.mask = r.g.mask.px.data();
    f.motion = reinterpret_cast<float*>(r.g.motion.px.data());
    f.ao_rt = opts.rtao ? r.aoRT.px.data() : nullptr;
    f.ao_ss = r.aoSS.px.data();
    f.sky = sky.px.empty() ? nullptr : sky.px.data();
    f.sun = sun.px.empty() ? nullptr : sun.px.data();
    f.error = r.rec.errorMap.px.data();
    f.frame = reinterpret_cast<float*>(r.frame.px.data());
    f.hdr = reinterpret_cast<float*>(r.hdr.px.data());
    raw_ao_frame_stats st{};
    if (raw_ao_render_frame(ctx, &sc, &o, w, h, &f, &st) != RAW_AO_OK) {
        err = raw_ao_last_error(ctx);
        r = FrameResult{};
        return false;
    }
    if (opts.rtao) {
        r.rec.rmse = st.rmse;
        r.rec.maxError = st.max_error;
        r.rec.pixels = st.pixels;
        r.rec.withinTolerance = st.within_tolerance != 0;
    }
    r.motionValid = st.motion_valid;
    r.motionTotal = st.motion_total;
    std::memcpy(r.viewProj.m, st.viewproj, 64);
    return true;
}

// ==== SHA-256 (FIPS 180-4) ==============================================================
namespace {
// the first 32 bits of the fractional parts of the cube roots of the first 64 primes (FIPS 180-4 4.2.2)
const uint32_t kSha[64] = {
    0x428a2f98u, 0x71374491u, 0xb5c0fbcfu, 0xe9b5dba5u, 0x3956c25bu, 0x59f111f1u, 0x923f82a4u, 0xab1c5ed5u, 0xd807aa98u,
    0x12835b01u, 0x243185beu, 0x550c7dc3u, 0x72be5d74u, 0x80deb1feu, 0x9bdc06a7u, 0xc19bf174u, 0xe49b69c1u, 0xefbe4786u,
    0x0fc19dc6u, 0x240ca1ccu, 0x2de92c6fu, 0x4a7484aau, 0x5cb0a9dcu, 0x76f988dau, 0x983e5152u, 0xa831c66du, 0xb00327c8u,
    0xbf597fc7u, 0xc6e00bf3u, 0xd5a79147u, 0x06ca6351u, 0x14292967u, 0x27b70a85u, 0x2e1b2138u, 0x4d2c6dfcu, 0x53380d13u,
    0x650a7354u, 0x766a0abbu, 0x81c2c92eu, 0x92722c85u, 0xa2bfe8a1u, 0xa81a664bu, 0xc24b8b70u, 0xc76c51a3u, 0xd192e819u,
    0xd6990624u, 0xf40e3585u, 0x106aa070u, 0x19a4c116u, 0x1e376c08u, 0x2748774cu, 0x34b0bcb5u, 0x391c0cb3u, 0x4ed8aa4au,
    0x5b9cca4fu, 0x682e6ff3u, 0x748f82eeu, 0x78a5636fu, 0x84c87814u, 0x8cc70208u, 0x90befffau, 0xa4506cebu, 0xbef9a3f7u,
    0xc67178f2u};
inline uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

struct Sha256 {
    uint32_t st[8] = {0x6a09e667u, 0xbb67ae85u, 0x3c6ef372u, 0xa54ff53au, 0x510e527fu, 0x9b05688cu, 0x1f83d9abu, 0x5be0cd19u};
    uint8_t buf[64];
    size_t fill = 0;
    uint64_t total = 0;
    void block(const uint8_t* p) {
        uint32_t w[64];
        for (int i = 0; i < 16; ++i) w[i] = uint32_t(p[4 * i]) << 24 | uint32_t(p[4 * i + 1]) << 16 | uint32_t(p[4 * i + 2]) << 8 | p[4 * i + 3];
        for (int i = 16; i < 64; ++i) {
            const uint32_t s0 = rotr(w[i - 15], 7) ^ rotr(w[i - 15], 18) ^ (w[i - 15] >> 3);
            const uint32_t s1 = rotr(w[i - 2], 17) ^ rotr(w[i - 2], 19) ^ (w[i - 2] >> 10);
            w[i] = w[i - 16] + s0 + w[i - 7] + s1;
        }
        uint32_t v[8];
        std::memcpy(v, st, 32);
        for (int i = 0; i < 64; ++i) {
            const uint32_t ch = (v[4] & v[5]) ^ (~v[4] & v[6]);
            const uint32_t t1 = v[7] + (rotr(v[4], 6) ^ rotr(v[4], 11) ^ rotr(v[4], 25)) + ch + kSha[i] + w[i];
            const uint32_t maj = (v[0] & v[1]) ^ (v[0] & v[2]) ^ (v[1] & v[2]);
            const uint32_t t2 = (rotr(v[0], 2) ^ rotr(v[0], 13) ^ rotr(v[0], 22)) + maj;
            for (int k = 7; k > 0; --k) v[k] = v[k - 1];
            v[4] += t1;
            v[0] = t1 + t2;
        }
        for (int k = 0; k < 8; ++k) st[k] += v[k];
    }
    void update(const void* data, size_t n) {
        const uint8_t* p = static_cast<const uint8_t*>(data);
        total += n;
        while (n) {
            const size_t k = std::min(n, size_t(64) - fill);
            std::memcpy(buf + fill, p, k);
            fill += k;
            p += k;
            n -= k;
            if (fill == 64) { block(buf); fill = 0; }
        }
    }
    std::string hex() {
        const uint64_t bits = total * 8;
        const uint8_t one = 0x80, zero = 0;
        update(&one, 1);
        while (fill != 56) update(&zero, 1);
        uint8_t len[8];
        for (int i = 0; i < 8; ++i) len[i] = uint8_t(bits >> (56 - 8 * i));
        update(len, 8);
        static const char* hx = "0123456789abcdef";
        std::string o;
        for (uint32_t x : st)
            for (int s = 28; s >= 0; s -= 4) o += hx[(x >> s) & 15u];
        return o;
    }
};
}  // namespace

std::string sha256Hex(const std::string& bytes) {
    Sha256 h;
    h.update(bytes.data(), bytes.size());
    return h.hex();
}
std::string sha256File(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return "";
    Sha256 h;
    char b[1 << 16];
    while (f) {
        f.read(b, sizeof b);
        h.update(b, size_t(f.gcount()));
    }
    return h.hex();
}

// ==== images ============================================================================
namespace {
unsigned char to8(float v) {   // image.cpp:9-10
    v = std::clamp(v, 0.0f, 1.0f);
    return (unsigned char)(v * 255.0f + 0.5f);
}
std::string header(const char* magic, int w, int h, const char* last) {
    return std::string(magic) + "\n" + std::to_string(w) + " " + std::to_string(h) + "\n" + last + "\n";
}
bool put(const std::string& path, const std::string& bytes) {
    std::ofstream f(path, std::ios::binary);
    f.write(bytes.data(), std::streamsize(bytes.size()));
    return bool(f);
}
}  // namespace

std::string rgb8Bytes(const Buffer<Vec3>& img) {
    std::string b;
    b.reserve(img.px.size() * 3);
    for (const Vec3& c : img.px) { b += char(to8(c.x)); b += char(to8(c.y)); b += char(to8(c.z)); }
    return b;
}
std::string f32Bytes(const Buffer<float>& img) {
    std::string b(img.px.size() * 4, '\0');
    for (size_t i = 0; i < img.px.size(); ++i) {
        uint32_t u;
        std::memcpy(&u, &img.px[i], 4);
        for (int k = 0; k < 4; ++k) b[4 * i + size_t(k)] = char((u >> (8 * k)) & 0xFFu);
    }
    return b;
}
void writePPM(const Buffer<Vec3>& img, const std::string& path) { put(path, header("P6", img.w, img.h, "255") + rgb8Bytes(img)); }
void writePFM(const Buffer<Vec3>& img, const std::string& path) {
    std::string b = header("PF", img.w, img.h, "-1.0");
    for (int y = img.h - 1; y >= 0; --y)   // bottom row first, little-endian floats
        for (int x = 0; x < img.w; ++x) {
            const Vec3& c = img.at(x, y);
            const float v[3] = {c.x, c.y, c.z};
            b.append(reinterpret_cast<const char*>(v), 12);
        }
    put(path, b);
}
void writePFM1(const Buffer<float>& img, const std::string& path) {
    std::string b = header("Pf", img.w, img.h, "-1.0");
    for (int y = img.h - 1; y >= 0; --y) b.append(reinterpret_cast<const char*>(&img.px[size_t(y) * size_t(img.w)]), 4 * size_t(img.w));
    put(path, b);
}
void writePGM(const Buffer<float>& img, const std::string& path) {
    std::string b = header("P5", img.w, img.h, "255");
    for (float v : img.px) b += char(to8(v));
    put(path, b);
}
void writeMaskPGM(const Buffer<uint8_t>& mask, const std::string& path) {
    std::string b = header("P5", mask.w, mask.h, "255");
    for (uint8_t v : mask.px) b += char(v ? 255 : 0);
    put(path, b);
}
namespace {
bool readWhole(const std::string& path, std::string& out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    out.assign(std::istreambuf_iterator<char>(f), {});
    return true;
}
// three header lines ("magic", "W H", last), then the body (image.cpp:67-80)
bool splitHeader(const std::string& all, const char* magic, std::string& last, int& w, int& h, std::string& body) {
    const size_t a = all.find('\n');
    if (a == std::string::npos) return false;
    const size_t b = all.find('\n', a + 1);
    if (b == std::string::npos) return false;
    const size_t c = all.find('\n', b + 1);
    if (c == std::string::npos || all.substr(0, a) != magic) return false;
    std::istringstream dims(all.substr(a + 1, b - a - 1));
    if (!(dims >> w >> h) || w <= 0 || h <= 0) return false;
    last = all.substr(b + 1, c - b - 1);
    body = all.substr(c + 1);
    return true;
}
}  // namespace
bool readPFM1(const std::string& path, Buffer<float>& out) {
    std::string all, scale, body;
    int w = 0, h = 0;
    if (!readWhole(path, all) || !splitHeader(all, "Pf", scale, w, h, body) || scale != "-1.0") return false;
    if (body.size() != size_t(w) * size_t(h) * 4) return false;
    out.resize(w, h);
    for (int y = h - 1, k = 0; y >= 0; --y, ++k) std::memcpy(&out.px[size_t(y) * size_t(w)], body.data() + size_t(k) * 4 * size_t(w), 4 * size_t(w));
    return true;
}
bool readMaskPGM(const std::string& path, Buffer<uint8_t>& out) {
    std::string all, maxv, body;
    int w = 0, h = 0;
    if (!readWhole(path, all) || !splitHeader(all, "P5", maxv, w, h, body) || maxv != "255") return false;
    if (body.size() != size_t(w) * size_t(h)) return false;
    out.resize(w, h);
    for (size_t i = 0; i < body.size(); ++i) out.px[i] = body[i] ? 1 : 0;
    return true;
}

// ==== parameters, certificate, channels =================================================
const char* version() { return "raw-native 0.8.0"; }

Camera cameraFromParams(const CliParams& p) {   // cli_params.cpp:12-18
    Camera c;
    c.eye = p.eye;
    c.center = p.center;
    c.up = p.up;
    c.fovy = p.fovy;
    c.aspect = p.height > 0 ? (float)p.width / (float)p.height : 1.0f;
    return c;
}
Camera prevCameraFromParams(const CliParams& p) {   // cli_params.cpp:28-34
    Camera c = cameraFromParams(p);
    if (p.prevEye) c.eye = *p.prevEye;
    if (p.prevCenter) c.center = *p.prevCenter;
    if (p.prevUp) c.up = *p.prevUp;
    return c;
}
namespace {
std::string g9(float v) { char b[32]; std::snprintf(b, sizeof b, "%.9g", (double)v); return b; }
std::string g9v(Vec3 v) { return "[" + g9(v.x) + "," + g9(v.y) + "," + g9(v.z) + "]"; }
std::string g9o(const std::optional<Vec3>& v) { return v ? g9v(*v) : std::string("null"); }
std::string j6(const std::optional<double>& v) {   // %.6g, null when absent or not finite
    if (!v.has_value() || !std::isfinite(*v)) return "null";
    char b[64];
    std::snprintf(b, sizeof b, "%.6g", *v);
    return b;
}
std::string j6(double v) { return j6(std::optional<double>(v)); }
std::string j6v(Vec3 v) { return "[" + j6(v.x) + "," + j6(v.y) + "," + j6(v.z) + "]"; }
std::optional<double> motionCoherence(const FrameResult& o) {
    return o.motionTotal > 0 ? std::optional<double>((double)o.motionValid / o.motionTotal) : std::nullopt;
}
double fidelity(double rmse) {   // certificate_from.cpp:9-17
    if (!std::isfinite(rmse) || rmse < 0.0) return 0.0;
    const double c = 1.0 / (1.0 + rmse);
    return c < 0.0 ? 0.0 : (c > 1.0 ? 1.0 : c);
}
}  // namespace

std::string canonicalParamsJson(const CliParams& p) {   // cli_params.cpp:211-235 (no model, no owned scene)
    std::string o = "{";
    if (p.gpu) o += "\"backend\":\"" + (p.gpuBackend.empty() ? std::string("none") : p.gpuBackend) + "\",";
    o += "\"eye\":" + g9v(p.eye);
    o += ",\"fovy\":" + g9(p.fovy);
    o += ",\"height\":" + std::to_string(p.height);
    o += ",\"prev_eye\":" + g9o(p.prevEye);
    o += ",\"prev_target\":" + g9o(p.prevCenter);
    o += ",\"prev_up\":" + g9o(p.prevUp);
    o += std::string(",\"rt\":") + (p.rtao ? "true" : "false");
    o += ",\"target\":" + g9v(p.center);
    o += ",\"tolerance\":" + g9(p.tolerance);
    o += ",\"up\":" + g9v(p.up);
    o += ",\"width\":" + std::to_string(p.width);
    return o + "}";
}

FileDigests writeFrameFiles(const FrameResult& o, const CliParams& p) {
    const std::string dir = p.out + "/";
    std::vector<std::string> names;
    const auto at = [&](const char* n) { names.emplace_back(n); return dir + n; };
    writePPM(o.frame, at("frame.ppm"));
    writePFM(o.hdr, at("frame_hdr.pfm"));
    writePGM(o.aoSS, at("ao_ss.pgm"));
    writePFM1(o.aoSS, at("ao_ss.pfm"));
    writeMaskPGM(o.g.mask, at("mask.pgm"));
    if (p.rtao) {
        writePGM(o.aoRT, at("ao_rt.pgm"));
        writePFM1(o.aoRT, at("ao_rt.pfm"));
        writePGM(o.rec.errorMap, at("ao_error.pgm"));
    }
    std::sort(names.begin(), names.end());
    FileDigests d;
    for (const auto& n : names) d.emplace_back(n, sha256File(dir + n));
    return d;
}

std::string aoCertificateJson(const FrameResult& o, const CliParams& p, const FileDigests& outputs) {
    CertInputs in;
    in.renderer = version();
    in.paramsJson = canonicalParamsJson(p);
    in.rtSamples = p.rtao ? kRtSamples : 0;
    in.ssSamples = kSsSamples;
    in.tolerance = p.tolerance;
    in.motionCoherence = motionCoherence(o);
    in.hdrHeadroom = double(maxRadiance(o.hdr));
    in.outputs = outputs;
    return certificateJson(o.rec, in);
}

namespace {
// certificate_with_channels without provenance (certificate.cpp:59-84): what channels.json embeds
std::string channelCertificate(const FrameResult& o, float tolerance) {
    CertInputs in;
    in.tolerance = tolerance;
    in.motionCoherence = motionCoherence(o);
    in.hdrHeadroom = double(maxRadiance(o.hdr));
    const std::string full = certificateJson(o.rec, in);
    return full.substr(0, full.find(",\"schema\":")) + "}";   // the raw-cert/2 block is the tail
}
}  // namespace

std::string channelsJson(const FrameResult& o, const CliParams& p, int rw, int rh) {
    // summaries over covered pixels (channels_json.cpp:33-60)
    const GBuffer& g = o.g;
    int covered = 0;
    const int total = g.w * g.h;
    std::optional<double> dmin, dmax, dmean;
    double dsum = 0, msum = 0, mmax = 0, clipFrac = 0;
    Vec3 nsum{0, 0, 0};
    int clip = 0;
    bool anyDepth = false;
    for (int y = 0; y < g.h; ++y)
        for (int x = 0; x < g.w; ++x) {
            if (!g.mask.at(x, y)) continue;
            ++covered;
            const float d = g.depth.at(x, y);
            if (std::isfinite(d)) {
                if (!anyDepth) { dmin = d; dmax = d; anyDepth = true; }
                else { dmin = std::min(*dmin, (double)d); dmax = std::max(*dmax, (double)d); }
                dsum += d;
            }
            nsum = vadd(nsum, g.normal.at(x, y));
            const Vec2 mv = g.motion.at(x, y);
            const float mag = std::sqrt(mv.x * mv.x + mv.y * mv.y);
            mmax = std::max(mmax, (double)mag);
            msum += mag;
            const Vec3 hv = o.hdr.at(x, y);
            if (hv.x > 1.0f || hv.y > 1.0f || hv.z > 1.0f) ++clip;
        }
    std::optional<Vec3> nmean;
    double mmean = 0;
    if (covered > 0) {
        if (anyDepth) dmean = dsum / covered;
        nmean = vnorm(vscl(nsum, 1.0f / (float)covered));
        mmean = msum / covered;
        clipFrac = (double)clip / covered;
    }
    const std::optional<double> coh = motionCoherence(o), head = double(maxRadiance(o.hdr));
    std::string j = "{\"schema\":\"raw-channels/1\",";
    j += "\"frame\":{\"width\":" + std::to_string(g.w) + ",\"height\":" + std::to_string(g.h) +
         ",\"files\":{\"ppm\":\"frame.ppm\",\"pfm\":\"frame_hdr.pfm\",\"ao_rt\":\"ao_rt.pgm\",\"ao_ss\":\"ao_ss.pgm\","
         "\"ao_error\":\"ao_error.pgm\",\"ao_rt_exact\":\"ao_rt.pfm\",\"ao_ss_exact\":\"ao_ss.pfm\",\"mask\":\"mask.pgm\"}},";
    const Camera cam = cameraFromParams(p);
    j += "\"camera\":{\"eye\":" + j6v(cam.eye) + ",\"target\":" + j6v(cam.center) + ",\"up\":" + j6v(cam.up) +
         ",\"fovy\":" + j6(cam.fovy) + ",\"aspect\":" + j6(cam.aspect) + ",\"prev\":";
    if (p.hasPrevCamera()) {
        const Camera pc = prevCameraFromParams(p);
        j += "{\"eye\":" + j6v(pc.eye) + ",\"target\":" + j6v(pc.center) + ",\"up\":" + j6v(pc.up) + "}";
    } else {
        j += "null";
    }
    j += "},";
    j += "\"certificate\":" + channelCertificate(o, p.tolerance) + ",";
    j += "\"channels\":{";
    j += "\"coverage\":" + j6(total > 0 ? (double)covered / total : 0.0) + ",";
    if (dmean.has_value()) j += "\"depth\":{\"min\":" + j6(dmin) + ",\"max\":" + j6(dmax) + ",\"mean\":" + j6(dmean) + "},";
    else j += "\"depth\":null,";
    if (nmean.has_value()) j += "\"normal\":{\"mean\":" + j6v(*nmean) + "},";
    else j += "\"normal\":null,";
    j += "\"motion\":{\"valid\":" + std::to_string(o.motionValid) + ",\"total\":" + std::to_string(o.motionTotal) +
         ",\"coherence\":" + j6(coh) + ",\"max_magnitude\":" + j6(mmax) + ",\"mean_magnitude\":" + j6(mmean) + "},";
    j += "\"hdr\":{\"headroom\":" + j6(head) + ",\"clipping_fraction\":" + j6(clipFrac) + "},";
    const std::optional<double> fid = o.rec.pixels > 0 ? std::optional<double>(fidelity(o.rec.rmse)) : std::nullopt;
    j += "\"ao\":{\"rmse\":" + j6(o.rec.rmse) + ",\"max_error\":" + j6(o.rec.maxError) + ",\"fidelity\":" + j6(fid) +
         ",\"within_tolerance\":" + (o.rec.withinTolerance ? "true" : "false") + "},";
    // the 8 x 8 box-averaged Rec.709 luma of the displayed frame (channels_json.cpp:64-88)
    rw = std::max(1, rw);
    rh = std::max(1, rh);
    const Buffer<Vec3>& fr = o.frame;
    j += "\"readout\":{\"kind\":\"luminance\",\"width\":" + std::to_string(rw) + ",\"height\":" + std::to_string(rh) + ",\"rows\":[";
    for (int cy = 0; cy < rh; ++cy) {
        j += cy ? ",[" : "[";
        for (int cx = 0; cx < rw; ++cx) {
            int x0 = int((long long)cx * fr.w / rw), x1 = int((long long)(cx + 1) * fr.w / rw);
            int y0 = int((long long)cy * fr.h / rh), y1 = int((long long)(cy + 1) * fr.h / rh);
            x1 = std::min(std::max(x1, x0 + 1), fr.w);
            y1 = std::min(std::max(y1, y0 + 1), fr.h);
            double sum = 0;
            int n = 0;
            for (int y = y0; y < y1; ++y)
                for (int x = x0; x < x1; ++x) {
                    const Vec3 c = fr.at(x, y);
                    sum += 0.2126 * c.x + 0.7152 * c.y + 0.0722 * c.z;
                    ++n;
                }
            j += (cx ? "," : "") + j6(n ? sum / n : 0.0);
        }
        j += "]";
    }
    return j + "]}}}";
}

// ==== canonical JSON v2 and the receipt =================================================
namespace {
struct J {   // a JSON value; objects keep insertion order and are written sorted
    enum Kind { Null, Bool, Int, Num, Str, Arr, Obj } k{Null};
    bool b{false};
    long long i{0};
    double d{0};
    std::string s;
    std::vector<J> a;
    std::vector<std::pair<std::string, J>> o;
    J() = default;
    J(std::nullptr_t) {}
    J(bool v) : k(Bool), b(v) {}
    J(int v) : k(Int), i(v) {}
    J(long long v) : k(Int), i(v) {}
    J(double v) : k(Num), d(v) {}
    J(const char* v) : k(Str), s(v) {}
    J(std::string v) : k(Str), s(std::move(v)) {}
    static J arr(std::vector<J> v) { J r; r.k = Arr; r.a = std::move(v); return r; }
    static J obj() { J r; r.k = Obj; return r; }
    J& set(const std::string& key, J v) {
        if (k != Obj) { k = Obj; o.clear(); }
        for (auto& m : o)
            if (m.first == key) { m.second = std::move(v); return *this; }
        o.emplace_back(key, std::move(v));
        return *this;
    }
    const J* find(const std::string& key) const {
        if (k != Obj) return nullptr;
        for (const auto& m : o)
            if (m.first == key) return &m.second;
        return nullptr;
    }
    void erase(const std::string& key) {
        o.erase(std::remove_if(o.begin(), o.end(), [&](const auto& m) { return m.first == key; }), o.end());
    }
};

// a number in the shortest digits that read back as the same double, written positionally
std::string canonNumber(double x) {
    if (!std::isfinite(x)) throw std::runtime_error("non-finite number");
    if (x == 0) return "0";
    char buf[64];
    const auto r = std::to_chars(buf, buf + sizeof buf, x, std::chars_format::scientific);
    std::string t(buf, size_t(r.ptr - buf));
    std::string sign;
    if (t[0] == '-') { sign = "-"; t = t.substr(1); }
    const size_t e = t.find('e');
    const long exp = std::stol(t.substr(e + 1));
    std::string mant = t.substr(0, e), ip = mant, fp;
    const size_t dot = mant.find('.');
    if (dot != std::string::npos) { ip = mant.substr(0, dot); fp = mant.substr(dot + 1); }
    std::string digits = ip + fp;
    long point = long(ip.size()) + exp;
    const size_t lead = digits.find_first_not_of('0');
    if (lead == std::string::npos) return "0";
    point -= long(lead);
    digits = digits.substr(lead);
    digits.erase(digits.find_last_not_of('0') + 1);
    const long n = long(digits.size());
    std::string body;
    if (point <= 0) body = "0." + std::string(size_t(-point), '0') + digits;
    else if (point >= n) body = digits + std::string(size_t(point - n), '0');
    else body = digits.substr(0, size_t(point)) + "." + digits.substr(size_t(point));
    return sign + body;
}
std::string canonString(const std::string& s) {
    static const char* hx = "0123456789abcdef";
    std::string o = "\"";
    for (char ch : s) {
        const unsigned char c = (unsigned char)ch;
        switch (ch) {
            case '"': o += "\\\""; break;
            case '\\': o += "\\\\"; break;
            case '\b': o += "\\b"; break;
            case '\f': o += "\\f"; break;
            case '\n': o += "\\n"; break;
            case '\r': o += "\\r"; break;
            case '\t': o += "\\t"; break;
            default:
                if (c < 0x20) { o += "\\u00"; o += hx[c >> 4]; o += hx[c & 15]; }
                else o += ch;
        }
    }
    return o + "\"";
}
std::string canon(const J& v) {
    switch (v.k) {
        case J::Null: return "null";
        case J::Bool: return v.b ? "true" : "false";
        case J::Int: return std::to_string(v.i);
        case J::Num: return canonNumber(v.d);
        case J::Str: return canonString(v.s);
        case J::Arr: {
            std::string o = "[";
            for (size_t i = 0; i < v.a.size(); ++i) o += (i ? "," : "") + canon(v.a[i]);
            return o + "]";
        }
        case J::Obj: {
            std::vector<const std::pair<std::string, J>*> ms;
            for (const auto& m : v.o) ms.push_back(&m);
            std::sort(ms.begin(), ms.end(), [](const auto* x, const auto* y) { return x->first < y->first; });
            std::string o = "{";
            for (size_t i = 0; i < ms.size(); ++i) o += (i ? "," : "") + canonString(ms[i]->first) + ":" + canon(ms[i]->second);
            return o + "}";
        }
    }
    return "null";
}

// A float as the double its shortest decimal names (receipt.cpp:18-26).
double f2d(float v) {
    char b[48];
    const auto r = std::to_chars(b, b + sizeof b, v);
    double d = 0;
    std::from_chars(b, r.ptr, d);
    return d;
}
J vec3(Vec3 v) { return J::arr({f2d(v.x), f2d(v.y), f2d(v.z)}); }
J vec3d(double x, double y, double z) { return J::arr({x, y, z}); }
J camera(Vec3 eye, Vec3 target, Vec3 up, float fovy) {
    J c = J::obj();
    c.set("eye", vec3(eye)).set("target", vec3(target)).set("up", vec3(up)).set("fovy", f2d(fovy));
    return c;
}
// The test scene in the contract's reference-scene layout (receipt.cpp:39-77).
J sceneValue(const CliParams& p) {
    J frame = J::obj();
    frame.set("width", p.width).set("height", p.height).set("pixel_center", 0.5).set("origin", "top-left");
    J ground = J::obj(), box = J::obj(), light = J::obj(), shade = J::obj(), ao = J::obj();
    ground.set("id", "ground").set("shape", "quad");
    ground.set("corners", J::arr({vec3d(-5, 0, -5), vec3d(5, 0, -5), vec3d(5, 0, 5), vec3d(-5, 0, 5)}));
    ground.set("normal", vec3d(0, 1, 0)).set("albedo", vec3(Vec3{0.7f, 0.7f, 0.7f}));
    box.set("id", "box").set("shape", "box").set("center", vec3d(0, 1, 0)).set("half", 1);
    box.set("albedo", vec3(Vec3{0.8f, 0.3f, 0.2f}));
    light.set("kind", "directional").set("dir", vec3(Vec3{-0.4f, -1.0f, -0.3f})).set("intensity", 1);
    shade.set("model", "raw-lambert-ao/1").set("ambient", 0.2).set("clamp", J::arr({0, 1})).set("encode", "u8-round-half-up");
    ao.set("estimator", p.rtao ? "rt-cosine-hemisphere" : "screen-space");
    ao.set("samples", p.rtao ? kRtSamples : kSsSamples).set("radius", f2d(kAoRadius));
    ao.set("hash", "raw-pixel-hash/1").set("origin_offset", 0.001).set("t_min", 0.0001);
    J s = J::obj();
    s.set("kind", "superstack.scene/1").set("seed", "raw-default").set("t_flicks", 0).set("frame", frame);
    s.set("camera", camera(p.eye, p.center, p.up, p.fovy));
    if (p.hasPrevCamera()) {
        const Camera pc = prevCameraFromParams(p);
        s.set("prev_camera", camera(pc.eye, pc.center, pc.up, pc.fovy));
    }
    s.set("meshes", J::arr({ground, box})).set("lights", J::arr({light})).set("shade", shade).set("ao", ao);
    std::vector<J> outs{"frame.rgb8"};
    if (p.rtao) outs.emplace_back("ao.f32");
    outs.emplace_back("mask.u8");
    s.set("outputs", J::arr(outs));
    return s;
}
// make_receipt + seal of superstack.receipt/1: the body's fields, then receipt_sha256 over
// the canonical body without it.
std::string receipt(const std::string& backend, const J& scene, const J& media, const std::string& content, const J& outputs,
                    const J& reference, const J& reconcile, const std::vector<std::string>& dnp) {
    J rec = J::obj();
    rec.set("schema", "superstack.receipt/1");
    J prod = J::obj();
    std::string ver = version();
    ver = ver.substr(ver.rfind(' ') + 1);
    prod.set("name", "raw-native").set("version", ver);
    rec.set("producer", prod).set("backend", backend).set("scene_sha256", sha256Hex(canon(scene)));
    const J* seed = scene.find("seed");
    const J* t = scene.find("t_flicks");
    rec.set("seed", seed ? *seed : J()).set("seed_rule", "raw-pixel-hash/1").set("seed_u32", J());
    J time = J::obj();
    time.set("base", "flicks").set("per_second", 705600000LL).set("t", t ? *t : J(0));
    rec.set("time", time).set("media", media).set("content_sha256", sha256Hex(content)).set("outputs", outputs);
    J rc = J::obj();
    rc.set("reference", reference);
    for (const auto& m : reconcile.o) rc.set(m.first, m.second);
    rec.set("reconcile", rc);
    std::vector<J> d;
    for (const auto& x : dnp) d.emplace_back(x);
    rec.set("does_not_prove", J::arr(d));
    rec.set("receipt_sha256", sha256Hex(canon(rec)));
    return canon(rec);
}
}  // namespace

std::string aoReceiptJson(const FrameResult& o, const CliParams& p, const std::string& backend, const FileDigests& outputs) {
    const std::string subject = f32Bytes(o.aoSS);
    const std::string refSha = p.rtao ? sha256Hex(f32Bytes(o.aoRT)) : std::string();
    const std::string subjectSha = sha256Hex(subject);
    J metrics = J::obj(), bounds = J::obj();
    bounds.set("ao_rmse_max", f2d(p.tolerance));
    std::string verdict = "unverifiable", reason;
    if (!p.rtao) reason = "the ray-traced reference was skipped (--no-rt)";
    else if (o.rec.pixels == 0) reason = "no covered pixels to compare";
    else {
        metrics.set("ao_rmse", f2d(o.rec.rmse)).set("ao_max_error", f2d(o.rec.maxError)).set("pixels", o.rec.pixels);
        verdict = o.rec.withinTolerance ? "verified" : "refuted";
    }
    J outs = J::obj();
    for (const auto& kv : outputs) outs.set(kv.first, kv.second);
    outs.set("frame.rgb8", sha256Hex(rgb8Bytes(o.frame)));
    J media = J::obj();
    media.set("kind", "image").set("width", p.width).set("height", p.height).set("format", "f32").set("transfer", "linear");
    J ref = J::obj();
    ref.set("backend", backend + "-rt-ao-64spp").set("content_sha256", refSha.empty() ? std::string("none") : refSha);
    J tol = J::obj();
    tol.set("verdict", verdict).set("metrics", metrics).set("bounds", bounds);
    if (!reason.empty()) tol.set("reason", reason);
    J rc = J::obj();
    rc.set("identity", (refSha.empty() ? std::string("none") : refSha) == subjectSha ? "MATCH" : "DRIFT").set("tolerance", tol);
    return receipt(backend + "-ssao", sceneValue(p), media, subject, outs, ref, rc,
                   {"The subject is the screen-space AO and the reference is the ray-traced AO, two different estimators, "
                    "so identity reads DRIFT by design; the tolerance verdict is the claim.",
                    "The ray-traced reference is itself a 64-sample estimate with visible grain, not ground truth.",
                    "An RMSE bound limits the average difference, not the worst pixel.",
                    "One built-in scene. No claim about other geometry, lights or materials."});
}

int emitOutputs(const FrameResult& o, const CliParams& p, const std::string& backend, std::string& err) {
    FileDigests outputs = writeFrameFiles(o, p);
    for (const auto& kv : outputs)
        if (kv.second.empty()) { err = "could not write " + p.out + "/" + kv.first; return 2; }
    const std::string dir = p.out + "/";
    if (!put(dir + "certificate.json", aoCertificateJson(o, p, outputs)) || !put(dir + "channels.json", channelsJson(o, p))) {
        err = "could not write " + p.out + "/certificate.json or channels.json";
        return 2;
    }
    FileDigests all = outputs;
    for (const char* f : {"certificate.json", "channels.json"}) all.emplace_back(f, sha256File(dir + f));
    if (!put(dir + "receipt.json", aoReceiptJson(o, p, backend, all))) {
        err = "could not write " + p.out + "/receipt.json";
        return 2;
    }
    return 0;
}

// ==== raw-gpu-cert/1 =====================================================================
namespace {
// raw/cert/gpu_tolerance.hpp: the bounds the reference committed before its first GPU run
constexpr double kMaskMismatchFraction = 0.002, kDepthRelRmse = 1e-4, kPositionRmse = 1e-3, kNormalRmse = 1e-3,
                 kMotionRmse = 1e-4, kAoSsRmse = 0.01, kAoRtRmse = 0.01, kFrameRmse = 0.01, kReconcileRmseDelta = 0.005;
constexpr bool kVerdictMustMatch = true;

struct Acc {   // gpu_reconcile.cpp:11-20
    double sumSq{0}, max{0};
    long long n{0};
    void add(double d) {
        d = std::fabs(d);
        sumSq += d * d;
        if (d > max) max = d;
        ++n;
    }
    GpuChannelCheck done(const std::string& name, double bound) const {
        GpuChannelCheck c;
        c.name = name;
        c.values = n;
        c.bound = bound;
        c.rmse = n ? std::sqrt(sumSq / (double)n) : 0.0;
        c.maxError = max;
        c.pass = n > 0 && std::isfinite(c.rmse) && c.rmse <= bound;
        return c;
    }
};
std::string gstr(const std::string& s) {   // gpu_reconcile.cpp:21-29
    std::string o = "\"";
    for (char ch : s) {
        if (ch == '"' || ch == '\\') { o += '\\'; o += ch; }
        else if ((unsigned char)ch < 0x20) { char b[8]; std::snprintf(b, sizeof b, "\\u%04x", (unsigned char)ch); o += b; }
        else o += ch;
    }
    return o + "\"";
}
std::string g9d(double v) {
    if (!std::isfinite(v)) return "null";
    char b[40];
    std::snprintf(b, sizeof b, "%.9g", v);
    return b;
}
}  // namespace

AdapterInfo adapterInfo() {
    AdapterInfo a;
    try {
        raw_ao_adapter x{};
        if (raw_ao_adapter_info(device(), &x) == RAW_AO_OK) {
            a.vendor = x.vendor;
            a.architecture = x.architecture;
            a.device = x.device;
            a.description = x.description;
            a.backend = x.backend;
            a.driver = x.driver;
        }
    } catch (const std::exception&) {
    }
    return a;
}

bool GpuReconcile::pass() const {
    if (!sizeMatch || !maskPass) return false;
    for (const auto& c : channels)
        if (!c.pass) return false;
    if (rt && (!reconcilePass || !verdictPass)) return false;
    return true;
}

GpuReconcile reconcileGpuCpu(const FrameResult& gpu, const FrameResult& cpu, bool rt) {
    GpuReconcile r;
    r.rt = rt;
    r.width = cpu.g.w;
    r.height = cpu.g.h;
    if (!gpu.frame.px.empty()) r.gpuFrameSha256 = sha256Hex(rgb8Bytes(gpu.frame));
    if (!cpu.frame.px.empty()) r.cpuFrameSha256 = sha256Hex(rgb8Bytes(cpu.frame));
    r.sizeMatch = gpu.g.w == cpu.g.w && gpu.g.h == cpu.g.h && gpu.frame.w == cpu.frame.w && gpu.frame.h == cpu.frame.h &&
                  (!rt || (gpu.aoRT.w == cpu.aoRT.w && gpu.aoRT.h == cpu.aoRT.h));
    if (!r.sizeMatch) return r;
    Acc depth, pos, nrm, mot, ss, rta, frame;
    for (int y = 0; y < r.height; ++y)
        for (int x = 0; x < r.width; ++x) {
            const bool a = gpu.g.mask.at(x, y) != 0, b = cpu.g.mask.at(x, y) != 0;
            if (a || b) ++r.coveredEither;
            if (a != b) { ++r.maskMismatch; continue; }
            if (!a) continue;
            ++r.coveredBoth;
            const float dc = cpu.g.depth.at(x, y);
            depth.add((gpu.g.depth.at(x, y) - dc) / (std::fabs(dc) > 1e-6f ? dc : 1.0f));
            const Vec3 gp = gpu.g.position.at(x, y), cp = cpu.g.position.at(x, y);
            pos.add(gp.x - cp.x); pos.add(gp.y - cp.y); pos.add(gp.z - cp.z);
            const Vec3 gn = gpu.g.normal.at(x, y), cn = cpu.g.normal.at(x, y);
            nrm.add(gn.x - cn.x); nrm.add(gn.y - cn.y); nrm.add(gn.z - cn.z);
            const Vec2 gm = gpu.g.motion.at(x, y), cm = cpu.g.motion.at(x, y);
            mot.add(gm.x - cm.x); mot.add(gm.y - cm.y);
            ss.add(gpu.aoSS.at(x, y) - cpu.aoSS.at(x, y));
            if (rt) rta.add(gpu.aoRT.at(x, y) - cpu.aoRT.at(x, y));
            const Vec3 gf = gpu.frame.at(x, y), cf = cpu.frame.at(x, y);
            frame.add(gf.x - cf.x); frame.add(gf.y - cf.y); frame.add(gf.z - cf.z);
        }
    r.maskMismatchFraction = r.coveredEither ? (double)r.maskMismatch / (double)r.coveredEither : 0.0;
    r.maskPass = r.coveredEither > 0 && r.maskMismatchFraction <= kMaskMismatchFraction;
    r.channels.push_back(depth.done("depth_rel", kDepthRelRmse));
    r.channels.push_back(pos.done("position", kPositionRmse));
    r.channels.push_back(nrm.done("normal", kNormalRmse));
    r.channels.push_back(mot.done("motion", kMotionRmse));
    r.channels.push_back(ss.done("ao_ss", kAoSsRmse));
    if (rt) r.channels.push_back(rta.done("ao_rt", kAoRtRmse));
    r.channels.push_back(frame.done("frame", kFrameRmse));
    r.cpuAoRmse = cpu.rec.rmse;
    r.gpuAoRmse = gpu.rec.rmse;
    r.cpuWithin = cpu.rec.withinTolerance;
    r.gpuWithin = gpu.rec.withinTolerance;
    r.reconcilePass = std::fabs((double)r.gpuAoRmse - (double)r.cpuAoRmse) <= kReconcileRmseDelta;
    r.verdictPass = !kVerdictMustMatch || r.cpuWithin == r.gpuWithin;
    return r;
}

std::string gpuCertificateJson(const GpuReconcile& r, const AdapterInfo& a, const std::string& renderer,
                               const std::string& paramsJson, double gpuMs, double cpuMs) {   // gpu_reconcile.cpp:85-134
    const char* verdict = !r.sizeMatch ? "unverifiable" : (r.pass() ? "verified" : "refuted");
    std::string o = "{\"schema\":\"raw-gpu-cert/1\"";
    o += ",\"claim\":\"the GPU frame matches the CPU reference within the committed tolerance\"";
    o += std::string(",\"verdict\":\"") + verdict + "\"";
    const bool both = !r.gpuFrameSha256.empty() && !r.cpuFrameSha256.empty();
    o += ",\"identity\":" + (both ? std::string(r.gpuFrameSha256 == r.cpuFrameSha256 ? "\"MATCH\"" : "\"DRIFT\"") : std::string("null"));
    o += ",\"frame_rgb8_sha256\":" +
         (both ? "{\"gpu\":" + gstr(r.gpuFrameSha256) + ",\"cpu\":" + gstr(r.cpuFrameSha256) + "}" : std::string("null"));
    o += ",\"oracle\":\"raw-gpu-cpu-v1\"";
    if (!r.reason.empty()) o += ",\"reason\":" + gstr(r.reason);
    o += ",\"renderer\":" + gstr(renderer);
    o += ",\"params\":" + paramsJson;
    o += ",\"adapter\":{\"vendor\":" + gstr(a.vendor) + ",\"architecture\":" + gstr(a.architecture) + ",\"device\":" +
         gstr(a.device) + ",\"description\":" + gstr(a.description) + ",\"backend\":" + gstr(a.backend) +
         ",\"driver\":" + gstr(a.driver) + "}";
    o += ",\"tolerance_source\":\"raw/gpu_tolerance.hpp, committed before the first GPU run\"";
    o += ",\"coverage\":{\"covered_either\":" + std::to_string(r.coveredEither) + ",\"covered_both\":" +
         std::to_string(r.coveredBoth) + ",\"mismatch\":" + std::to_string(r.maskMismatch) +
         ",\"mismatch_fraction\":" + g9d(r.maskMismatchFraction) + ",\"bound\":" + g9d(kMaskMismatchFraction) +
         ",\"pass\":" + (r.maskPass ? "true" : "false") + "}";
    o += ",\"channels\":[";
    for (size_t i = 0; i < r.channels.size(); ++i) {
        const auto& c = r.channels[i];
        o += std::string(i ? "," : "") + "{\"name\":" + gstr(c.name) + ",\"values\":" + std::to_string(c.values) +
             ",\"rmse\":" + g9d(c.rmse) + ",\"max_error\":" + g9d(c.maxError) + ",\"bound\":" + g9d(c.bound) +
             ",\"pass\":" + (c.pass ? "true" : "false") + "}";
    }
    o += "]";
    if (r.rt && r.sizeMatch) {
        o += ",\"ao_verdict\":{\"cpu_rmse\":" + exactFloat(r.cpuAoRmse) + ",\"gpu_rmse\":" + exactFloat(r.gpuAoRmse) +
             ",\"rmse_delta_bound\":" + g9d(kReconcileRmseDelta) + ",\"cpu\":\"" + (r.cpuWithin ? "verified" : "refuted") +
             "\",\"gpu\":\"" + (r.gpuWithin ? "verified" : "refuted") +
             "\",\"pass\":" + ((r.reconcilePass && r.verdictPass) ? "true" : "false") + "}";
    } else {
        o += ",\"ao_verdict\":null";
    }
    o += ",\"timing_ms\":{\"gpu\":" + g9d(gpuMs) + ",\"cpu\":" + g9d(cpuMs) + ",\"judged\":false}";
    o += ",\"does_not_prove\":["
         "\"Agreement with the CPU path is not agreement with ground truth: the CPU ray-traced AO is itself a 64-sample estimate.\","
         "\"Checked on the adapter and driver named here only. Another GPU, driver or browser may differ.\","
         "\"An RMSE bound limits the average difference, not the worst pixel. Maximum errors are reported, not bounded.\","
         "\"One built-in scene. No claim about other geometry, lights or materials.\","
         "\"The timings are one run each and are not judged.\"]";
    return o + "}";
}

std::string gpuReceiptJson(const FrameResult* gpu, const FrameResult* cpu, const GpuReconcile& r, const CliParams& p,
                           const FileDigests& outputs) {
    const std::string subject = gpu ? rgb8Bytes(gpu->frame) : std::string();
    const std::string subjectSha = sha256Hex(subject);
    const std::string refSha = cpu ? sha256Hex(rgb8Bytes(cpu->frame)) : std::string();
    J metrics = J::obj(), bounds = J::obj();
    bounds.set("coverage_mismatch_frac_max", kMaskMismatchFraction);
    const std::pair<const char*, double> b[] = {{"depth_rel", kDepthRelRmse}, {"position", kPositionRmse},
                                                {"normal", kNormalRmse},      {"motion", kMotionRmse},
                                                {"ao_ss", kAoSsRmse},         {"ao_rt", kAoRtRmse},
                                                {"frame", kFrameRmse}};
    for (const auto& [name, bound] : b) bounds.set(std::string(name) + "_rmse_max", bound);
    bounds.set("ao_rmse_delta_max", kReconcileRmseDelta);
    std::string verdict = "unverifiable", reason = r.reason;
    if (r.sizeMatch && gpu && cpu) {
        metrics.set("coverage_mismatch_frac", r.maskMismatchFraction);
        for (const auto& c : r.channels) {
            metrics.set(c.name + "_rmse", c.rmse);
            metrics.set(c.name + "_max_error", c.maxError);
        }
        if (r.rt) metrics.set("ao_rmse_delta", std::fabs((double)r.gpuAoRmse - (double)r.cpuAoRmse));
        verdict = r.pass() ? "verified" : "refuted";
    } else if (reason.empty()) {
        reason = "the GPU frame and the CPU reference differ in size";
    }
    const std::string backend = "raw-native-" + (p.gpuBackend.empty() ? std::string("none") : p.gpuBackend);
    J outs = J::obj();
    for (const auto& kv : outputs) outs.set(kv.first, kv.second);
    J media = J::obj();
    media.set("kind", "image").set("width", p.width).set("height", p.height).set("format", "rgb8").set("transfer", "linear-u8");
    J ref = J::obj();
    ref.set("backend", "raw-native-cpu").set("content_sha256", refSha.empty() ? std::string("none") : refSha);
    J tol = J::obj();
    tol.set("verdict", verdict).set("metrics", metrics).set("bounds", bounds);
    if (!reason.empty()) tol.set("reason", reason);
    J rc = J::obj();
    rc.set("identity", (refSha.empty() ? std::string("none") : refSha) == subjectSha ? "MATCH" : "DRIFT").set("tolerance", tol);
    return receipt(backend, sceneValue(p), media, subject, outs, ref, rc,
                   {"Identity compares the 8-bit frames only; the tolerance verdict covers every channel at float precision.",
                    "Agreement with the CPU path is not agreement with ground truth: the CPU ray-traced AO is itself a 64-sample "
                    "estimate.",
                    "Checked on the adapter and driver named in gpu_certificate.json only. Another GPU, driver or browser may "
                    "differ.",
                    "An RMSE bound limits the average difference, not the worst pixel. Maximum errors are reported, not bounded.",
                    "One built-in scene. No claim about other geometry, lights or materials."});
}

// ==== verify ============================================================================
namespace {
// A JSON reader for the receipt: numbers without a fraction or exponent within +-2^53 are
// integers, the rest doubles (as superstack reads them).
struct Reader {
    const std::string& t;
    size_t i{0};
    void ws() { while (i < t.size() && (t[i] == ' ' || t[i] == '\n' || t[i] == '\r' || t[i] == '\t')) ++i; }
    [[noreturn]] void fail() { throw std::runtime_error("parse"); }
    J value() {
        ws();
        if (i >= t.size()) fail();
        const char c = t[i];
        if (c == '{') {
            ++i;
            J o = J::obj();
            ws();
            if (i < t.size() && t[i] == '}') { ++i; return o; }
            while (true) {
                ws();
                if (i >= t.size() || t[i] != '"') fail();
                std::string k = str();
                ws();
                if (i >= t.size() || t[i] != ':') fail();
                ++i;
                o.set(k, value());
                ws();
                if (i < t.size() && t[i] == ',') { ++i; continue; }
                if (i < t.size() && t[i] == '}') { ++i; return o; }
                fail();
            }
        }
        if (c == '[') {
            ++i;
            std::vector<J> a;
            ws();
            if (i < t.size() && t[i] == ']') { ++i; return J::arr(a); }
            while (true) {
                a.push_back(value());
                ws();
                if (i < t.size() && t[i] == ',') { ++i; continue; }
                if (i < t.size() && t[i] == ']') { ++i; return J::arr(a); }
                fail();
            }
        }
        if (c == '"') return J(str());
        if (t.compare(i, 4, "true") == 0) { i += 4; return J(true); }
        if (t.compare(i, 5, "false") == 0) { i += 5; return J(false); }
        if (t.compare(i, 4, "null") == 0) { i += 4; return J(); }
        const size_t s = i;
        bool real = false;
        while (i < t.size() && (std::isdigit((unsigned char)t[i]) || t[i] == '-' || t[i] == '+' || t[i] == '.' || t[i] == 'e' || t[i] == 'E')) {
            real = real || t[i] == '.' || t[i] == 'e' || t[i] == 'E';
            ++i;
        }
        if (s == i) fail();
        if (!real) {
            long long n = 0;
            const auto r = std::from_chars(t.data() + s, t.data() + i, n);
            if (r.ec == std::errc() && r.ptr == t.data() + i && n <= 9007199254740991LL && n >= -9007199254740991LL) return J(n);
        }
        double d = 0;
        const auto r = std::from_chars(t.data() + s, t.data() + i, d);
        if (r.ec != std::errc() || r.ptr != t.data() + i) fail();
        return J(d);
    }
    std::string str() {
        ++i;
        std::string o;
        while (i < t.size() && t[i] != '"') {
            if (t[i] == '\\') {
                ++i;
                if (i >= t.size()) fail();
                const char e = t[i++];
                switch (e) {
                    case '"': o += '"'; break;
                    case '\\': o += '\\'; break;
                    case '/': o += '/'; break;
                    case 'b': o += '\b'; break;
                    case 'f': o += '\f'; break;
                    case 'n': o += '\n'; break;
                    case 'r': o += '\r'; break;
                    case 't': o += '\t'; break;
                    case 'u': {
                        if (i + 4 > t.size()) fail();
                        const unsigned cp = unsigned(std::stoul(t.substr(i, 4), nullptr, 16));
                        i += 4;
                        if (cp >= 0x80) fail();   // the receipts written here are ASCII
                        o += char(cp);
                        break;
                    }
                    default: fail();
                }
            } else {
                o += t[i++];
            }
        }
        if (i >= t.size()) fail();
        ++i;
        return o;
    }
};

struct Tally {
    std::string& report;
    int bad{0};
    void check(bool ok, const std::string& what, const std::string& detail) {
        report += (ok ? "MATCH    " : "MISMATCH ") + what + (detail.empty() ? "" : "  " + detail) + "\n";
        if (!ok) ++bad;
    }
};
// the raw text after "key": up to ',' or '}' (verify.cpp:22-30)
bool scalarAfter(const std::string& j, const std::string& key, std::string& out) {
    const size_t k = j.find("\"" + key + "\":");
    if (k == std::string::npos) return false;
    const size_t s = k + key.size() + 3;
    size_t e = s;
    while (e < j.size() && j[e] != ',' && j[e] != '}') ++e;
    out = j.substr(s, e - s);
    if (out.size() >= 2 && out.front() == '"' && out.back() == '"') out = out.substr(1, out.size() - 2);
    return !out.empty();
}
bool objectAfter(const std::string& j, const std::string& key, std::string& out) {
    const size_t k = j.find("\"" + key + "\":{");
    if (k == std::string::npos) return false;
    const size_t s = k + key.size() + 3, e = j.find('}', s);
    if (e == std::string::npos) return false;
    out = j.substr(s, e - s + 1);
    return true;
}
FileDigests parseOutputs(const std::string& obj) {
    FileDigests d;
    size_t i = 0;
    while (true) {
        const size_t a = obj.find('"', i);
        if (a == std::string::npos) break;
        const size_t b = obj.find('"', a + 1);
        if (b == std::string::npos) break;
        const size_t c = obj.find('"', b + 1);
        if (c == std::string::npos) break;
        const size_t e = obj.find('"', c + 1);
        if (e == std::string::npos) break;
        d.emplace_back(obj.substr(a + 1, b - a - 1), obj.substr(c + 1, e - c - 1));
        i = e + 1;
    }
    return d;
}
// the receipt beside a certificate (receipt.cpp:208-250): sealed, its files, its subject and
// reference, its verdict
int checkReceipt(const std::string& dir, const std::string& certVerdict, std::string& report) {
    std::string text;
    if (!readWhole(dir + "/receipt.json", text)) { report += "MISSING  receipt.json\n"; return 1; }
    Tally t{report};
    J rec;
    bool parsed = true;
    try {
        Reader r{text};
        rec = r.value();
    } catch (const std::exception&) {
        parsed = false;
    }
    bool sealed = parsed && rec.k == J::Obj;
    const J* schema = sealed ? rec.find("schema") : nullptr;
    const J* seal = sealed ? rec.find("receipt_sha256") : nullptr;
    sealed = schema && schema->k == J::Str && schema->s == "superstack.receipt/1" && seal && seal->k == J::Str;
    if (sealed) {
        J body = rec;
        body.erase("receipt_sha256");
        sealed = sha256Hex(canon(body)) == seal->s;
    }
    t.check(sealed, "receipt well formed and sealed", parsed ? (sealed ? "" : "seal") : "parse");
    if (!sealed) return t.bad;
    if (const J* outs = rec.find("outputs"))
        for (const auto& m : outs->o) {
            std::string actual;
            if (m.first == "frame.rgb8") {
                std::string ppm;
                if (readWhole(dir + "/frame.ppm", ppm)) {
                    size_t i = 0;
                    for (int line = 0; line < 3 && i != std::string::npos; ++line) i = ppm.find('\n', i) + 1;
                    if (i != 0 && i <= ppm.size()) actual = sha256Hex(ppm.substr(i));
                }
            } else {
                actual = sha256File(dir + "/" + m.first);
            }
            const bool ok = m.second.k == J::Str && actual == m.second.s;
            t.check(ok, "receipt sha256 " + m.first, ok ? "" : actual);
        }
    Buffer<float> ss, rt;
    if (!readPFM1(dir + "/ao_ss.pfm", ss)) { report += "MISSING  ao_ss.pfm\n"; return t.bad + 1; }
    const J* cs = rec.find("content_sha256");
    t.check(cs && cs->k == J::Str && cs->s == sha256Hex(f32Bytes(ss)), "receipt content_sha256 (ao_ss.f32)", "");
    const std::string refSha = readPFM1(dir + "/ao_rt.pfm", rt) ? sha256Hex(f32Bytes(rt)) : std::string("none");
    const J* rc = rec.find("reconcile");
    const J* ref = rc ? rc->find("reference") : nullptr;
    const J* rsha = ref ? ref->find("content_sha256") : nullptr;
    t.check(rsha && rsha->k == J::Str && rsha->s == refSha, "receipt reference (ao_rt.f32)", "");
    const J* tol = rc ? rc->find("tolerance") : nullptr;
    const J* tv = tol ? tol->find("verdict") : nullptr;
    const std::string v = tv && tv->k == J::Str ? tv->s : std::string();
    t.check(v == certVerdict, "receipt tolerance verdict", v + " vs " + certVerdict);
    return t.bad;
}
}  // namespace

namespace {
// verify's own re-derivation of the reconcile from the files (reconcile.cpp:4-21), on the host:
// verify checks a directory anywhere, with or without a GPU.
ReconcileResult recheck(const Buffer<float>& approx, const Buffer<float>& truth, const Buffer<uint8_t>& mask, float tolerance) {
    ReconcileResult r;
    r.errorMap.resize(approx.w, approx.h);
    double sumSq = 0;
    int n = 0;
    for (size_t i = 0; i < approx.px.size(); ++i) {
        if (!mask.px[i]) continue;
        const float e = std::fabs(approx.px[i] - truth.px[i]);
        r.errorMap.px[i] = e;
        if (e > r.maxError) r.maxError = e;
        sumSq += (double)e * e;
        ++n;
    }
    r.pixels = n;
    r.rmse = n > 0 ? (float)std::sqrt(sumSq / n) : 0.0f;
    r.withinTolerance = r.rmse <= tolerance;
    return r;
}
}  // namespace

int verifyDir(const std::string& dir, std::string& report) {
    std::string cert, outputsObj, exact, verdict;
    if (!readWhole(dir + "/certificate.json", cert)) { report += "MISSING  certificate.json\n"; return 2; }
    if (cert.find("\"schema\":\"raw-cert/2\"") == std::string::npos || !objectAfter(cert, "outputs", outputsObj) ||
        !objectAfter(cert, "exact", exact) || !scalarAfter(cert, "verdict", verdict)) {
        report += "MISSING  certificate.json is not raw-cert/2\n";
        return 2;
    }
    Tally t{report};
    const FileDigests outputs = parseOutputs(outputsObj);
    bool haveRt = false;
    for (const auto& [name, digest] : outputs) {
        const std::string actual = sha256File(dir + "/" + name);
        if (actual.empty()) { report += "MISSING  " + name + "\n"; return 2; }
        t.check(actual == digest, "sha256 " + name, actual == digest ? "" : actual + " vs " + digest);
        haveRt = haveRt || name == "ao_rt.pfm";
    }
    if (haveRt) {   // the reconcile again from the full-precision files (verify.cpp:66-86)
        Buffer<float> rt, ss;
        Buffer<uint8_t> mask;
        if (!readPFM1(dir + "/ao_rt.pfm", rt) || !readPFM1(dir + "/ao_ss.pfm", ss) || !readMaskPGM(dir + "/mask.pgm", mask)) {
            report += "MISSING  ao_rt.pfm, ao_ss.pfm or mask.pgm unreadable\n";
            return 2;
        }
        if (rt.w != ss.w || rt.h != ss.h || rt.w != mask.w || rt.h != mask.h) { report += "MISMATCH buffer sizes differ\n"; return 3; }
        std::string pixels, rmse, maxError, tolerance;
        if (!scalarAfter(exact, "pixels", pixels) || !scalarAfter(exact, "rmse", rmse) || !scalarAfter(exact, "maxError", maxError) ||
            !scalarAfter(exact, "tolerance", tolerance)) {
            report += "MISSING  exact block incomplete\n";
            return 2;
        }
        const ReconcileResult r = recheck(ss, rt, mask, std::strtof(tolerance.c_str(), nullptr));
        const char* v = r.pixels == 0 ? "unverifiable" : (r.withinTolerance ? "verified" : "refuted");
        t.check(std::to_string(r.pixels) == pixels, "pixels", std::to_string(r.pixels) + " vs " + pixels);
        t.check(exactFloat(r.rmse) == rmse, "rmse", exactFloat(r.rmse) + " vs " + rmse);
        t.check(exactFloat(r.maxError) == maxError, "maxError", exactFloat(r.maxError) + " vs " + maxError);
        t.check(verdict == v, "verdict", std::string(v) + " vs " + verdict);
    } else {
        t.check(verdict == "unverifiable", "verdict without reference", verdict);
    }
    std::ifstream probe(dir + "/receipt.json");
    if (probe) t.bad += checkReceipt(dir, verdict, report);
    else report += "ABSENT   receipt.json (written from 0.5.0)\n";
    report += t.bad == 0 ? "verify: all checks match\n" : "verify: " + std::to_string(t.bad) + " mismatch(es)\n";
    return t.bad == 0 ? 0 : 3;
}

}  // namespace raw

// ==== C entry points for the bindings ===================================================
// CliParams (the fields the AO frame uses); has_prev_* select the optional previous camera.
struct rah_params {
    int32_t width, height;
    float eye[3], center[3], up[3];
    float fovy, tolerance;
    int32_t rtao, gpu;
    const char* gpu_backend;
    int32_t has_prev_eye, has_prev_center, has_prev_up;
    float prev_eye[3], prev_center[3], prev_up[3];
};
// A frame's planes (host, row-major; position, normal, albedo, frame, hdr 3 floats a pixel,
// motion 2; ao_rt may be NULL without the ray-traced pass) and its scalars.
struct rah_frame {
    const float *depth, *position, *normal, *albedo;
    const uint8_t* mask;
    const float *motion, *ao_rt, *ao_ss, *error, *frame, *hdr;
    float rmse, max_error;
    int32_t pixels, within, motion_valid, motion_total;
};

namespace {
raw::CliParams toParams(const rah_params& q, const char* dir) {
    raw::CliParams p;
    p.out = dir;
    p.width = q.width;
    p.height = q.height;
    p.eye = {q.eye[0], q.eye[1], q.eye[2]};
    p.center = {q.center[0], q.center[1], q.center[2]};
    p.up = {q.up[0], q.up[1], q.up[2]};
    p.fovy = q.fovy;
    p.tolerance = q.tolerance;
    p.rtao = q.rtao != 0;
    p.gpu = q.gpu != 0;
    p.gpuBackend = q.gpu_backend ? q.gpu_backend : "";
    if (q.has_prev_eye) p.prevEye = raw::Vec3{q.prev_eye[0], q.prev_eye[1], q.prev_eye[2]};
    if (q.has_prev_center) p.prevCenter = raw::Vec3{q.prev_center[0], q.prev_center[1], q.prev_center[2]};
    if (q.has_prev_up) p.prevUp = raw::Vec3{q.prev_up[0], q.prev_up[1], q.prev_up[2]};
    return p;
}
template <class T, class S> void fillBuf(raw::Buffer<T>& b, int w, int h, const S* src) {
    b.resize(w, h);
    if (src) std::memcpy(static_cast<void*>(b.px.data()), src, sizeof(T) * b.px.size());
}
raw::FrameResult toFrame(const rah_frame& f, int w, int h) {
    raw::FrameResult o;
    o.g.w = w;
    o.g.h = h;
    fillBuf(o.g.depth, w, h, f.depth);
    fillBuf(o.g.position, w, h, f.position);
    fillBuf(o.g.normal, w, h, f.normal);
    fillBuf(o.g.albedo, w, h, f.albedo);
    fillBuf(o.g.mask, w, h, f.mask);
    fillBuf(o.g.motion, w, h, f.motion);
    if (f.ao_rt) fillBuf(o.aoRT, w, h, f.ao_rt);
    fillBuf(o.aoSS, w, h, f.ao_ss);
    fillBuf(o.rec.errorMap, w, h, f.error);
    fillBuf(o.frame, w, h, f.frame);
    fillBuf(o.hdr, w, h, f.hdr);
    o.rec.rmse = f.rmse;
    o.rec.maxError = f.max_error;
    o.rec.pixels = f.pixels;
    o.rec.withinTolerance = f.within != 0;
    o.motionValid = f.motion_valid;
    o.motionTotal = f.motion_total;
    return o;
}
size_t copyOut(const std::string& s, char* out, size_t cap) {
    if (out && cap) {
        const size_t k = std::min(cap - 1, s.size());
        std::memcpy(out, s.data(), k);
        out[k] = 0;
    }
    return s.size();
}
}  // namespace

extern "C" {

// Writes the render's files under dir (frame.ppm, frame_hdr.pfm, the AO images, mask.pgm,
// certificate.json, channels.json, receipt.json). Returns 0, or 2 with the message in err.
int rah_emit_outputs(const char* dir, const rah_params* q, const rah_frame* f, const char* backend, char* err, size_t cap) {
    if (!dir || !q || !f) return 2;
    const raw::FrameResult o = toFrame(*f, q->width, q->height);
    std::string e;
    int rc = 2;
    try {
        rc = raw::emitOutputs(o, toParams(*q, dir), backend ? backend : "raw-native-cpu", e);
    } catch (const std::exception& x) {
        e = x.what();
    }
    copyOut(e, err, cap);
    return rc;
}

// raw_native_cli verify DIR: the exit code (0 all match, 2 missing, 3 mismatch); the report
// into out (cap bytes), returning its length through *len.
int rah_verify(const char* dir, char* out, size_t cap, size_t* len) {
    std::string report;
    int rc = 2;
    try {
        rc = raw::verifyDir(dir ? dir : ".", report);
    } catch (const std::exception& x) {
        report += std::string("MISSING  ") + x.what() + "\n";
    }
    const size_t n = copyOut(report, out, cap);
    if (len) *len = n;
    return rc;
}

// raw-gpu-cert/1 and gpu_receipt.json for a GPU frame against a CPU frame (either may be NULL:
// no frame, as when no GPU was found; reason then says why). adapter NULL: this device's.
// The certificate is written to dir/gpu_certificate.json, the receipt (listing the
// certificate's digest) to dir/gpu_receipt.json. Returns 0, or 2 when a file cannot be written.
int rah_gpu_certificate(const char* dir, const rah_params* q, const rah_frame* gpu, const rah_frame* cpu, const char* reason,
                        const raw_ao_adapter* adapter, double gpu_ms, double cpu_ms) {
    if (!dir || !q) return 2;
    try {
        const raw::CliParams p = toParams(*q, dir);
        std::optional<raw::FrameResult> g, c;
        if (gpu) g = toFrame(*gpu, q->width, q->height);
        if (cpu) c = toFrame(*cpu, q->width, q->height);
        raw::GpuReconcile r;
        if (g && c) r = raw::reconcileGpuCpu(*g, *c, p.rtao);
        r.rt = p.rtao;
        if (reason) r.reason = reason;
        raw::AdapterInfo a;
        if (adapter) a = {adapter->vendor, adapter->architecture, adapter->device, adapter->description, adapter->backend,
                          adapter->driver};
        else a = raw::adapterInfo();
        const std::string d = std::string(dir) + "/";
        {
            std::ofstream f(d + "gpu_certificate.json", std::ios::binary);
            f << raw::gpuCertificateJson(r, a, raw::version(), raw::canonicalParamsJson(p), gpu_ms, cpu_ms);
            if (!f) return 2;
        }
        const raw::FileDigests outs{{"gpu_certificate.json", raw::sha256File(d + "gpu_certificate.json")}};
        std::ofstream f(d + "gpu_receipt.json", std::ios::binary);
        f << raw::gpuReceiptJson(g ? &*g : nullptr, c ? &*c : nullptr, r, p, outs);
        return f ? 0 : 2;
    } catch (const std::exception&) {
        return 2;
    }
}

// SHA-256 of n bytes as 64 hex digits and a terminating zero.
void rah_sha256(const void* data, size_t n, char* out65) {
    const std::string h = raw::sha256Hex(std::string(static_cast<const char*>(data), n));
    std::memcpy(out65, h.c_str(), 65);
}

}  // extern "C"
