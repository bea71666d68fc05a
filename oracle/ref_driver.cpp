// A C entry layer over the reference's OWN sources, compiled from where they lie under
// the reference tree (see oracle/Makefile). TEST INFRASTRUCTURE ONLY: it lets the tests
// cross-check the oracle restatement against the reference itself on inputs the golden
// hashes do not cover, and gives bench.py a "reference" kind of CPU baseline. The
// product never links or loads it.
//
// Signatures mirror oracle_api.h with a ref_ prefix.
#include "raw/cert/reconcile.hpp"
#include "raw/renderer/accel.hpp"
#include "raw/renderer/composite.hpp"
#include "raw/renderer/raster.hpp"
#include "raw/renderer/render.hpp"
#include "raw/renderer/gpu_reconcile.hpp"
#include "raw/core/version.hpp"
#include "raw/renderer/ray_ao.hpp"
#include "raw/renderer/rt_bvh.hpp"
#include "raw/renderer/rt_scenes.hpp"
#include "raw/renderer/ssao.hpp"
#include "raw/renderer/swr_scenes.hpp"
#include "raw/scene/scene.hpp"
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {
raw::GBuffer toG(const float* position, const float* normal, const uint8_t* mask, int w, int h) {
    raw::GBuffer g;
    g.resize(w, h);
    for (size_t i = 0; i < size_t(w) * size_t(h); ++i) {
        g.position.px[i] = {position[3 * i], position[3 * i + 1], position[3 * i + 2]};
        g.normal.px[i] = {normal[3 * i], normal[3 * i + 1], normal[3 * i + 2]};
        g.mask.px[i] = mask[i];
    }
    return g;
}
raw::Scene triScene(const float* tris9, size_t ntri) {
    raw::Scene s;
    raw::Mesh m;
    for (size_t i = 0; i < ntri; ++i)
        for (int k = 0; k < 3; ++k) {
            m.positions.push_back({tris9[9 * i + 3 * size_t(k)], tris9[9 * i + 3 * size_t(k) + 1], tris9[9 * i + 3 * size_t(k) + 2]});
            m.normals.push_back({0, 1, 0});
            m.indices.push_back(int(3 * i + size_t(k)));
        }
    s.meshes.push_back(std::move(m));
    return s;
}
}  // namespace

extern "C" {

void ref_rasterize(const float* vpos, const float* vnrm, const int32_t* idx, size_t ntri,
                   const float eye[3], const float center[3], const float up[3],
                   float fovy, float aspect, float nearZ, float farZ, int w, int h,
                   float* depth, float* position, float* normal, uint8_t* mask) {
    raw::Scene s;
    raw::Mesh m;
    int32_t maxv = -1;
    for (size_t i = 0; i < 3 * ntri; ++i) { m.indices.push_back(idx[i]); if (idx[i] > maxv) maxv = idx[i]; }
    for (int32_t v = 0; v <= maxv; ++v) {
        m.positions.push_back({vpos[3 * v], vpos[3 * v + 1], vpos[3 * v + 2]});
        m.normals.push_back({vnrm[3 * v], vnrm[3 * v + 1], vnrm[3 * v + 2]});
    }
    s.meshes.push_back(std::move(m));
    s.camera.eye = {eye[0], eye[1], eye[2]};
    s.camera.center = {center[0], center[1], center[2]};
    s.camera.up = {up[0], up[1], up[2]};
    s.camera.fovy = fovy; s.camera.aspect = aspect; s.camera.nearZ = nearZ; s.camera.farZ = farZ;
    raw::GBuffer g = raw::rasterize(s, w, h);
    for (size_t i = 0; i < size_t(w) * size_t(h); ++i) {
        depth[i] = g.depth.px[i];
        position[3 * i] = g.position.px[i].x; position[3 * i + 1] = g.position.px[i].y; position[3 * i + 2] = g.position.px[i].z;
        normal[3 * i] = g.normal.px[i].x; normal[3 * i + 1] = g.normal.px[i].y; normal[3 * i + 2] = g.normal.px[i].z;
        mask[i] = g.mask.px[i];
    }
}

// The test scene (two meshes) rasterized as the reference's CLI does.
void ref_render_test_scene(int w, int h, const float eye[3], const float center[3], float fovy,
                           float* position, float* normal, uint8_t* mask) {
    raw::Scene s = raw::buildTestScene(w, h);
    s.camera.eye = {eye[0], eye[1], eye[2]};
    s.camera.center = {center[0], center[1], center[2]};
    s.camera.fovy = fovy;
    raw::GBuffer g = raw::rasterize(s, w, h);
    for (size_t i = 0; i < size_t(w) * size_t(h); ++i) {
        position[3 * i] = g.position.px[i].x; position[3 * i + 1] = g.position.px[i].y; position[3 * i + 2] = g.position.px[i].z;
        normal[3 * i] = g.normal.px[i].x; normal[3 * i + 1] = g.normal.px[i].y; normal[3 * i + 2] = g.normal.px[i].z;
        mask[i] = g.mask.px[i];
    }
}

// computeRTAO through LinearAccel; mode 0 BVH, 1 scan, 2 skew (set as the reference reads it).
void ref_rtao(const float* tris9, size_t ntri, int mode, const float* position, const float* normal,
              const uint8_t* mask, int w, int h, int samples, float radius, int threads, float* ao) {
    if (mode == 1) setenv("RAW_NATIVE_AO_ORACLE", "scan", 1);
    else if (mode == 2) setenv("RAW_NATIVE_AO_ORACLE", "skew", 1);
    else unsetenv("RAW_NATIVE_AO_ORACLE");
    raw::LinearAccel accel;
    accel.build(triScene(tris9, ntri));
    unsetenv("RAW_NATIVE_AO_ORACLE");
    const raw::GBuffer g = toG(position, normal, mask, w, h);
    const raw::Buffer<float> out = raw::computeRTAO(g, accel, samples, radius, nullptr, threads);
    std::memcpy(ao, out.px.data(), sizeof(float) * size_t(w) * size_t(h));
}

// computeRTVisibility (raw/renderer/ray_ao.hpp:12-14) through LinearAccel, BVH mode.
void ref_rtvis(const float* tris9, size_t ntri, const float* position, const float* normal, const uint8_t* mask, int w,
               int h, int samples, float radius, const float to_light[3], int sun_forced, int threads, float* ao,
               float* sky, float* sun) {
    unsetenv("RAW_NATIVE_AO_ORACLE");
    raw::LinearAccel accel;
    accel.build(triScene(tris9, ntri));
    const raw::GBuffer g = toG(position, normal, mask, w, h);
    raw::Buffer<float> a, s, u;
    raw::computeRTVisibility(g, accel, samples, radius, raw::Vec3{to_light[0], to_light[1], to_light[2]}, a, s, u, threads,
                             sun_forced != 0);
    const size_t np = size_t(w) * size_t(h);
    std::memcpy(ao, a.px.data(), sizeof(float) * np);
    std::memcpy(sky, s.px.data(), sizeof(float) * np);
    std::memcpy(sun, u.px.data(), sizeof(float) * np);
}

// The same, with the accelerator built once (LinearAccel::build, accel.cpp:13-25) and kept,
// so a caller can time the build and the trace apart (BVH mode only).
void* ref_accel_build(const float* tris9, size_t ntri) {
    unsetenv("RAW_NATIVE_AO_ORACLE");
    auto* a = new raw::LinearAccel;
    a->build(triScene(tris9, ntri));
    return a;
}
void ref_accel_free(void* a) { delete static_cast<raw::LinearAccel*>(a); }
void ref_rtao_accel(void* accel, const float* position, const float* normal, const uint8_t* mask, int w, int h,
                    int samples, float radius, int threads, float* ao) {
    const raw::GBuffer g = toG(position, normal, mask, w, h);
    const raw::Buffer<float> out = raw::computeRTAO(g, *static_cast<raw::LinearAccel*>(accel), samples, radius, nullptr, threads);
    std::memcpy(ao, out.px.data(), sizeof(float) * size_t(w) * size_t(h));
}

void ref_ssao(const float* position, const float* normal, const uint8_t* mask, int w, int h, int samples,
              float radius, int threads, float* ao) {
    const raw::GBuffer g = toG(position, normal, mask, w, h);
    const raw::Buffer<float> out = raw::computeSSAO(g, samples, radius, nullptr, threads);
    std::memcpy(ao, out.px.data(), sizeof(float) * size_t(w) * size_t(h));
}

void ref_reconcile(const float* approx, const float* truth, const uint8_t* mask, int w, int h, float tolerance,
                   float* error_map, float* rmse, float* max_error, int32_t* pixels, int32_t* within) {
    raw::Buffer<float> a, t;
    raw::Buffer<uint8_t> m;
    a.resize(w, h); t.resize(w, h); m.resize(w, h);
    std::memcpy(a.px.data(), approx, sizeof(float) * size_t(w) * size_t(h));
    std::memcpy(t.px.data(), truth, sizeof(float) * size_t(w) * size_t(h));
    std::memcpy(m.px.data(), mask, size_t(w) * size_t(h));
    const raw::ReconcileResult r = raw::reconcile(a, t, m, tolerance);
    std::memcpy(error_map, r.errorMap.px.data(), sizeof(float) * size_t(w) * size_t(h));
    *rmse = r.rmse; *max_error = r.maxError; *pixels = r.pixels; *within = r.withinTolerance ? 1 : 0;
}

// renderWithParams (src/renderer/render.cpp:71-105) of a scene of meshes concatenated in order
// (every mesh sees the whole vertex array; its indices select its own vertices): every plane of
// the FrameResult, the reconcile, the motion counts, the view-projection and maxRadiance(hdr).
// lighting 1/2 as RenderOptions::lighting; no_sun sets RAW_NATIVE_LIGHTING=v2-nosun for the call.
// prev16 NULL: this frame's view-projection (zero motion).
void ref_render_frame(const float* vpos, const float* vnrm, size_t nverts, const int32_t* idx, size_t ntri,
                      const uint32_t* mesh_end, const float* albedo, size_t nmesh, const float eye[3], const float center[3],
                      const float up[3], float fovy, float aspect, float nearZ, float farZ, int has_light,
                      const float light_dir[3], float light_intensity, int w, int h, const float* prev16, int lighting,
                      int no_sun, int rtao, float tolerance, int threads, float* depth, float* position, float* normal,
                      float* albedo_out, uint8_t* mask, float* motion, float* ao_rt, float* ao_ss, float* error_map,
                      float* frame, float* hdr, float* stats8, float* viewproj16) {
    raw::Scene s;
    size_t t0 = 0;
    for (size_t m = 0; m < nmesh; ++m) {
        raw::Mesh mesh;
        for (size_t v = 0; v < nverts; ++v) {
            mesh.positions.push_back({vpos[3 * v], vpos[3 * v + 1], vpos[3 * v + 2]});
            mesh.normals.push_back({vnrm[3 * v], vnrm[3 * v + 1], vnrm[3 * v + 2]});
        }
        for (size_t i = 3 * t0; i < 3 * size_t(mesh_end[m]) && i < 3 * ntri; ++i) mesh.indices.push_back(idx[i]);
        t0 = mesh_end[m];
        mesh.material.albedo = {albedo[3 * m], albedo[3 * m + 1], albedo[3 * m + 2]};
        s.meshes.push_back(std::move(mesh));
    }
    if (has_light) s.lights.push_back(raw::Light{{light_dir[0], light_dir[1], light_dir[2]}, light_intensity});
    s.camera.eye = {eye[0], eye[1], eye[2]};
    s.camera.center = {center[0], center[1], center[2]};
    s.camera.up = {up[0], up[1], up[2]};
    s.camera.fovy = fovy; s.camera.aspect = aspect; s.camera.nearZ = nearZ; s.camera.farZ = farZ;
    raw::Mat4 prev = raw::mul(s.camera.proj(), s.camera.view());
    if (prev16) std::memcpy(prev.m, prev16, 64);
    raw::RenderOptions o;
    o.tolerance = tolerance; o.rtao = rtao != 0; o.threads = threads; o.lighting = lighting;
    unsetenv("RAW_NATIVE_AO_ORACLE");
    if (no_sun) setenv("RAW_NATIVE_LIGHTING", "v2-nosun", 1);
    else unsetenv("RAW_NATIVE_LIGHTING");
    const raw::FrameResult r = raw::renderWithParams(s, w, h, prev, nullptr, o);
    unsetenv("RAW_NATIVE_LIGHTING");
    const size_t np = size_t(w) * size_t(h);
    const auto v3 = [np](float* out, const raw::Buffer<raw::Vec3>& b) {
        for (size_t i = 0; i < np; ++i) { out[3 * i] = b.px[i].x; out[3 * i + 1] = b.px[i].y; out[3 * i + 2] = b.px[i].z; }
    };
    std::memcpy(depth, r.g.depth.px.data(), 4 * np);
    v3(position, r.g.position);
    v3(normal, r.g.normal);
    v3(albedo_out, r.g.albedo);
    std::memcpy(mask, r.g.mask.px.data(), np);
    for (size_t i = 0; i < np; ++i) { motion[2 * i] = r.g.motion.px[i].x; motion[2 * i + 1] = r.g.motion.px[i].y; }
    if (rtao) std::memcpy(ao_rt, r.aoRT.px.data(), 4 * np);
    std::memcpy(ao_ss, r.aoSS.px.data(), 4 * np);
    std::memcpy(error_map, r.rec.errorMap.px.data(), 4 * np);
    v3(frame, r.frame);
    v3(hdr, r.hdr);
    stats8[0] = r.rec.rmse; stats8[1] = r.rec.maxError; stats8[2] = float(r.rec.pixels); stats8[3] = r.rec.withinTolerance ? 1.0f : 0.0f;
    stats8[4] = float(r.motionValid); stats8[5] = float(r.motionTotal); stats8[6] = raw::maxRadiance(r.hdr); stats8[7] = 0.0f;
    std::memcpy(viewproj16, r.viewProj.m, 64);
}

// A frame's planes as reconcileGpuCpu reads them (FrameResult, raw/renderer/render.hpp:10-21).
struct ref_frame_planes {
    const float *depth, *position, *normal;
    const uint8_t* mask;
    const float *motion, *ao_rt, *ao_ss, *frame;
    float rmse;
    int32_t within;
};
static raw::FrameResult refFrame(const ref_frame_planes& f, int w, int h) {
    raw::FrameResult r;
    r.g.resize(w, h);
    r.aoSS.resize(w, h);
    r.frame.resize(w, h);
    if (f.ao_rt) r.aoRT.resize(w, h);
    for (size_t i = 0; i < size_t(w) * size_t(h); ++i) {
        r.g.depth.px[i] = f.depth[i];
        r.g.position.px[i] = {f.position[3 * i], f.position[3 * i + 1], f.position[3 * i + 2]};
        r.g.normal.px[i] = {f.normal[3 * i], f.normal[3 * i + 1], f.normal[3 * i + 2]};
        r.g.mask.px[i] = f.mask[i];
        r.g.motion.px[i] = {f.motion[2 * i], f.motion[2 * i + 1]};
        if (f.ao_rt) r.aoRT.px[i] = f.ao_rt[i];
        r.aoSS.px[i] = f.ao_ss[i];
        r.frame.px[i] = {f.frame[3 * i], f.frame[3 * i + 1], f.frame[3 * i + 2]};
    }
    r.rec.rmse = f.rmse;
    r.rec.withinTolerance = f.within != 0;
    return r;
}

// reconcileGpuCpu + gpuCertificateJson (src/renderer/gpu_reconcile.cpp:36-134) of two frames
// (either NULL: no frame, reason says why), with an empty adapter; the text into out.
size_t ref_gpu_certificate(const ref_frame_planes* gpu, const ref_frame_planes* cpu, int w, int h, int rt, const char* reason,
                           const char* params_json, double gpu_ms, double cpu_ms, char* out, size_t cap) {
    raw::GpuReconcile r;
    if (gpu && cpu) r = raw::reconcileGpuCpu(refFrame(*gpu, w, h), refFrame(*cpu, w, h), rt != 0);
    r.rt = rt != 0;
    if (reason) r.reason = reason;
    const std::string s = raw::gpuCertificateJson(r, raw::GpuAdapterInfo{}, raw::version(), params_json, gpu_ms, cpu_ms);
    if (out && cap) {
        const size_t k = std::min(cap - 1, s.size());
        std::memcpy(out, s.data(), k);
        out[k] = 0;
    }
    return s.size();
}

// rt::closest (src/renderer/rt_bvh.cpp:125-152) on the reference's own PLOC tree of tris9:
// hits4 {t, u, v, tri bits} per ray of rays8 {o, d, tMax, pad}.
void ref_tree_closest(const float* tris9, size_t ntri, const float* rays8, size_t nrays, float* hits4) {
    std::vector<raw::Tri> t(ntri);
    for (size_t i = 0; i < ntri; ++i)
        t[i] = {{tris9[9 * i], tris9[9 * i + 1], tris9[9 * i + 2]}, {tris9[9 * i + 3], tris9[9 * i + 4], tris9[9 * i + 5]},
                {tris9[9 * i + 6], tris9[9 * i + 7], tris9[9 * i + 8]}};
    const raw::rt::Tree tree = raw::rt::buildPloc(t);
    for (size_t k = 0; k < nrays; ++k) {
        const float* q = rays8 + 8 * k;
        const raw::Ray r{{q[0], q[1], q[2]}, {q[3], q[4], q[5]}};
        raw::Hit h;
        float* o = hits4 + 4 * k;
        int32_t tri = -1;
        o[0] = -1.0f; o[1] = 0.0f; o[2] = 0.0f;
        if (raw::rt::closest(tree, t, r, q[6], h)) { o[0] = h.t; o[1] = h.u; o[2] = h.v; tri = h.tri; }
        std::memcpy(&o[3], &tri, 4);
    }
}

int32_t ref_ploc_build(const float* tris9, size_t ntri, int radius, float* boxes6, int32_t* links3, uint32_t* rounds) {
    std::vector<raw::Tri> t(ntri);
    for (size_t i = 0; i < ntri; ++i)
        t[i] = {{tris9[9 * i], tris9[9 * i + 1], tris9[9 * i + 2]}, {tris9[9 * i + 3], tris9[9 * i + 4], tris9[9 * i + 5]},
                {tris9[9 * i + 6], tris9[9 * i + 7], tris9[9 * i + 8]}};
    const raw::rt::Tree tree = raw::rt::buildPloc(t, radius);
    for (size_t i = 0; i < tree.nodes.size(); ++i) {
        for (int k = 0; k < 3; ++k) { boxes6[6 * i + size_t(k)] = tree.nodes[i].lo[k]; boxes6[6 * i + 3 + size_t(k)] = tree.nodes[i].hi[k]; }
        links3[3 * i] = tree.nodes[i].left; links3[3 * i + 1] = tree.nodes[i].right; links3[3 * i + 2] = tree.nodes[i].tri;
    }
    if (rounds) *rounds = tree.rounds;
    return tree.root;
}

size_t ref_scene_tris(const char* name, size_t count, uint64_t seed, float* tris9) {
    const std::string s(name);
    std::vector<raw::Tri> t;
    if (s == "dense_block") t = raw::rt::trianglesOf(raw::rt::denseBlock().geo);
    else if (s == "retro_room") t = raw::rt::trianglesOf(raw::swr::retroRoom().geo);
    else if (s == "iso_street") t = raw::rt::trianglesOf(raw::swr::isoStreet().geo);
    else if (s == "clip_stress") t = raw::rt::trianglesOf(raw::swr::clipStress().geo);
    else if (s == "soup") t = raw::rt::soup(count, seed);
    else return 0;
    if (tris9)
        for (size_t i = 0; i < t.size(); ++i) {
            const raw::Vec3 v[3] = {t[i].a, t[i].b, t[i].c};
            for (int k = 0; k < 3; ++k) { tris9[9 * i + 3 * size_t(k)] = v[k].x; tris9[9 * i + 3 * size_t(k) + 1] = v[k].y; tris9[9 * i + 3 * size_t(k) + 2] = v[k].z; }
        }
    return t.size();
}

double ref_binned_sah_cost(const float* tris9, size_t ntri, size_t* node_count) {
    std::vector<raw::Tri> t(ntri);
    for (size_t i = 0; i < ntri; ++i)
        t[i] = {{tris9[9 * i], tris9[9 * i + 1], tris9[9 * i + 2]}, {tris9[9 * i + 3], tris9[9 * i + 4], tris9[9 * i + 5]},
                {tris9[9 * i + 6], tris9[9 * i + 7], tris9[9 * i + 8]}};
    return raw::rt::binnedSahCost(t, node_count);
}

}  // extern "C"
