// Renders the reference's default frame through the C++ host interface (raw::LinearAccel,
// raw::computeRTAO, raw::computeSSAO, raw::reconcile, raw::certificateJson), the way
// src/renderer/render.cpp:26-54 calls them, and writes ao_rt.pfm, ao_ss.pfm and
// certificate.json into argv[1] (PFM as src/core/image.cpp:53-62 writes it).
// Usage: host_interface_frame OUTDIR [W H]
#include "../../raw-native_amd/host/raw_ao_host.hpp"
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

extern "C" {
size_t rah_test_scene(float* vpos, float* vnrm, int32_t* idx);
void rah_rasterize(const float* vpos, const float* vnrm, const int32_t* idx, size_t ntri, const float eye[3],
                   const float center[3], const float up[3], float fovy, float aspect, float nearZ, float farZ, int w, int h,
                   float* depth, float* position, float* normal, uint8_t* mask);
}

static void writePFM1(const raw::Buffer<float>& img, const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "wb");
    std::fprintf(f, "Pf\n%d %d\n-1.0\n", img.w, img.h);
    for (int y = img.h - 1; y >= 0; --y) std::fwrite(&img.px[size_t(y) * size_t(img.w)], 4, size_t(img.w), f);
    std::fclose(f);
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string out = argv[1];
    const int w = argc > 3 ? std::atoi(argv[2]) : 256, h = argc > 3 ? std::atoi(argv[3]) : 256;
    const size_t ntri = rah_test_scene(nullptr, nullptr, nullptr);
    std::vector<float> vpos(6 * ntri), vnrm(6 * ntri);
    std::vector<int32_t> idx(3 * ntri);
    rah_test_scene(vpos.data(), vnrm.data(), idx.data());
    raw::GBuffer g;
    g.resize(w, h);
    const float eye[3] = {4, 4, 6}, center[3] = {0, 1, 0}, up[3] = {0, 1, 0};
    rah_rasterize(vpos.data(), vnrm.data(), idx.data(), ntri, eye, center, up, 0.9f, float(w) / float(h), 0.1f, 100.0f, w, h,
                  g.depth.px.data(), &g.position.px[0].x, &g.normal.px[0].x, g.mask.px.data());
    std::vector<raw::Tri> tris(ntri);
    for (size_t i = 0; i < ntri; ++i)
        for (int k = 0; k < 3; ++k) {
            const float* p = &vpos[3 * size_t(idx[3 * i + size_t(k)])];
            raw::Vec3& v = k == 0 ? tris[i].a : (k == 1 ? tris[i].b : tris[i].c);
            v = {p[0], p[1], p[2]};
        }
    try {
        raw::LinearAccel accel;
        accel.build(tris);
        const raw::Buffer<float> aoRT = raw::computeRTAO(g, accel, 64, 2.0f);
        const raw::Buffer<float> aoSS = raw::computeSSAO(g, 24, 2.0f);
        const raw::ReconcileResult rec = raw::reconcile(aoSS, aoRT, g.mask, 0.12f);
        writePFM1(aoRT, out + "/ao_rt.pfm");
        writePFM1(aoSS, out + "/ao_ss.pfm");
        raw::CertInputs in;
        in.renderer = "raw-ao-mi355x";
        const std::string cert = raw::certificateJson(rec, in);
        FILE* f = std::fopen((out + "/certificate.json").c_str(), "wb");
        std::fwrite(cert.data(), 1, cert.size(), f);
        std::fclose(f);
        std::printf("reconcile: pixels=%d rmse=%s maxError=%s\n", rec.pixels, raw::exactFloat(rec.rmse).c_str(),
                    raw::exactFloat(rec.maxError).c_str());
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 4;
    }
    return 0;
}
