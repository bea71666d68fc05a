// Renders the reference's default frame through the C++ host interface.
//   host_interface_frame OUTDIR [W H]
//     raw::LinearAccel, raw::computeRTAO, raw::computeSSAO, raw::reconcile, raw::certificateJson
//     the way src/renderer/render.cpp:26-54 calls them; writes ao_rt.pfm, ao_ss.pfm and
//     certificate.json into OUTDIR.
//   host_interface_frame OUTDIR W H frame [PREV_EYE_X PREV_EYE_Y PREV_EYE_Z]
//     the CLI's render (src/tools/run.cpp:12-26, app/main.cpp:69-88): raw::renderGpu of the test
//     scene, raw::emitOutputs (images, certificate.json, channels.json, receipt.json), then
//     raw::verifyDir on the directory; prints the verify report, exits with its code.
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

static int cliFrame(const std::string& out, int w, int h, int argc, char** argv) {
    raw::CliParams p;
    p.out = out;
    p.width = w;
    p.height = h;
    if (argc > 7) p.prevEye = raw::Vec3{float(std::atof(argv[5])), float(std::atof(argv[6])), float(std::atof(argv[7]))};
    raw::Scene s = raw::buildTestScene(w, h);   // renderFromParams (src/tools/run.cpp:12-26)
    s.camera = raw::cameraFromParams(p);
    const raw::Mat4 prev = p.hasPrevCamera() ? raw::prevCameraFromParams(p).viewProj() : s.camera.viewProj();
    raw::RenderOptions opts;
    opts.tolerance = p.tolerance;
    opts.rtao = p.rtao;
    raw::FrameResult o;
    std::string err;
    if (!raw::renderGpu(s, w, h, prev, opts, o, err)) {
        std::fprintf(stderr, "%s\n", err.c_str());
        return 4;
    }
    if (raw::emitOutputs(o, p, "raw-native-cpu", err) != 0) {
        std::fprintf(stderr, "%s\n", err.c_str());
        return 2;
    }
    std::string report;
    const int rc = raw::verifyDir(out, report);
    std::printf("%s", report.c_str());
    return rc;
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string out = argv[1];
    const int w = argc > 3 ? std::atoi(argv[2]) : 256, h = argc > 3 ? std::atoi(argv[3]) : 256;
    if (argc > 4 && std::string(argv[4]) == "frame") return cliFrame(out, w, h, argc, argv);
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
        raw::writePFM1(aoRT, out + "/ao_rt.pfm");
        raw::writePFM1(aoSS, out + "/ao_ss.pfm");
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
