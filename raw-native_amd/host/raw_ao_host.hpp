// The C++ host side of raw_ao_hip: the reference's AO/reconcile interface over the C ABI
// of include/raw_ao_hip.h, with the reference's types and semantics, so a caller of
//   raw::computeRTAO   (raw/renderer/ray_ao.hpp:6-8)
//   raw::computeRTVisibility (raw/renderer/ray_ao.hpp:12-14)
//   raw::computeSSAO   (raw/renderer/ssao.hpp:5-6)
//   raw::reconcile     (raw/cert/reconcile.hpp:12-13)
//   raw::LinearAccel   (raw/renderer/accel.hpp:15-23)
// finds the same functions here. The ray tracing, SSAO and error map run on the GPU;
// the float64 row-major sum of the reconcile runs here, in the reference's order.
#pragma once
#include <cstdint>
#include <optional>
#include <string>
#include <utility>
#include <vector>

struct raw_ao_ctx;

namespace raw {

struct Vec3 { float x{}, y{}, z{}; };
struct Vec2 { float x{}, y{}; };
struct Mat4 { float m[16]{}; };   // row-major, element (r, c) = m[4 r + c] (raw/math/mat.hpp:5)
struct Tri { Vec3 a, b, c; };
template <class T> struct Buffer {
    int w{0}, h{0};
    std::vector<T> px;
    void resize(int W, int H) { w = W; h = H; px.assign(size_t(W) * size_t(H), T{}); }
    T& at(int x, int y) { return px[size_t(y) * size_t(w) + size_t(x)]; }
    const T& at(int x, int y) const { return px[size_t(y) * size_t(w) + size_t(x)]; }
};
struct GBuffer {   // raw/renderer/gbuffer.hpp:7-20
    int w{0}, h{0};
    Buffer<float> depth;
    Buffer<Vec3> normal, position, albedo;
    Buffer<uint8_t> mask;
    Buffer<Vec2> motion;
    void resize(int W, int H) {
        w = W; h = H;
        depth.resize(W, H); normal.resize(W, H); position.resize(W, H); albedo.resize(W, H); mask.resize(W, H);
        motion.resize(W, H);
    }
};
struct Arena;   // accepted for signature compatibility; device memory is the context's

// The process-wide GPU context (the reference's rhi::device(), raw/rhi/rhi.hpp:284-286).
// Throws std::runtime_error with the library's message when no GPU is present.
raw_ao_ctx* device();

// The AO accelerator: the triangles and their BVH, built on the GPU (PLOC).
struct LinearAccel {
    std::vector<Tri> tris;
    void build(const std::vector<Tri>& t);
    bool built{false};
};

Buffer<float> computeRTAO(const GBuffer& g, const LinearAccel& accel, int samples, float radius, Arena* arena = nullptr,
                          int threads = 1);
// Lighting model 2's visibility from the same rays (raw/renderer/ray_ao.hpp:12-14): ao as
// computeRTAO, sky the fraction of samples with no hit at any distance, sun 1 when a ray
// towards toLight hits nothing (forced to 1 when sunForced); 1 off the mask.
void computeRTVisibility(const GBuffer& g, const LinearAccel& accel, int samples, float radius, Vec3 toLight,
                         Buffer<float>& ao, Buffer<float>& sky, Buffer<float>& sun, int threads = 1, bool sunForced = false);
Buffer<float> computeSSAO(const GBuffer& g, int samples, float radius, Arena* arena = nullptr, int threads = 1);

struct ReconcileResult {
    Buffer<float> errorMap;
    float rmse{0}, maxError{0};
    int pixels{0};
    bool withinTolerance{false};
};
ReconcileResult reconcile(const Buffer<float>& approx, const Buffer<float>& truth, const Buffer<uint8_t>& mask,
                          float tolerance, Arena* arena = nullptr);

// raw-cert/2 (src/cert/certificate.cpp:38-85, src/cert/certificate_from.cpp:19-56).
struct CertInputs {
    std::string renderer, paramsJson;
    int rtSamples{64}, ssSamples{24};
    float tolerance{0.12f};
    std::optional<double> motionCoherence, hdrHeadroom;   // null when absent, never invented
    std::vector<std::pair<std::string, std::string>> outputs;
};
std::string certificateJson(const ReconcileResult& r, const CertInputs& in);
std::string exactFloat(float v);

// ---- the default frame (raw/renderer/render.hpp, raw/renderer/gpu.hpp) ----------------
struct Material { Vec3 albedo{0.8f, 0.8f, 0.8f}; };
struct Mesh { std::vector<Vec3> positions, normals; std::vector<int> indices; Material material; };
struct Light { Vec3 dir{0, -1, 0}; float intensity{1.0f}; };
struct Camera {   // raw/scene/scene.hpp:18-23
    Vec3 eye{0, 0, 5}, center{0, 0, 0}, up{0, 1, 0};
    float fovy{1.0f}, aspect{1.0f}, nearZ{0.1f}, farZ{100.0f};
    Mat4 viewProj() const;   // mul(proj(), view())
};
struct Scene { std::vector<Mesh> meshes; std::vector<Light> lights; Camera camera; };
Scene buildTestScene(int w, int h);   // src/scene/scene.cpp:18-32

struct FrameResult {   // raw/renderer/render.hpp:10-21
    GBuffer g;
    Buffer<float> aoRT, aoSS;
    ReconcileResult rec;
    Buffer<Vec3> frame, hdr;
    int motionValid{0}, motionTotal{0};
    Mat4 viewProj{};
};
inline constexpr int kRtSamples = 64, kSsSamples = 24;
inline constexpr float kAoRadius = 2.0f;
struct RenderOptions { float tolerance{0.12f}; bool rtao{true}; int threads{1}; int lighting{2}; };
int lightingModel(const RenderOptions& o);   // src/renderer/render.cpp:56-62, RAW_NATIVE_LIGHTING
bool lightingNoSun();
// renderGpu (raw/renderer/gpu.hpp:32): every channel of renderWithParams from the GPU, bit for
// bit the CPU reference's. false with the library's message on a failure.
bool renderGpu(const Scene& scene, int w, int h, const Mat4& prevViewProj, const RenderOptions& opts, FrameResult& out,
               std::string& err);
float maxRadiance(const Buffer<Vec3>& hdr);   // src/renderer/composite.cpp:82-88

// ---- on-disk and wire formats (src/core/image.cpp, src/tools/run.cpp, channels_json.cpp,
// receipt.cpp, verify.cpp; src/cert/certificate*.cpp) -------------------------------------
struct CliParams {   // raw/tools/cli_params.hpp:14-45, the fields the AO frame uses
    std::string out{"."};
    int width{256}, height{256};
    Vec3 eye{4, 4, 6}, center{0, 1, 0}, up{0, 1, 0};
    float fovy{0.9f}, tolerance{0.12f};
    bool rtao{true};
    int threads{1};
    bool gpu{false};
    std::string gpuBackend;
    std::optional<Vec3> prevEye, prevCenter, prevUp;
    bool hasPrevCamera() const { return prevEye.has_value() || prevCenter.has_value() || prevUp.has_value(); }
};
Camera cameraFromParams(const CliParams& p);
Camera prevCameraFromParams(const CliParams& p);
std::string canonicalParamsJson(const CliParams& p);
const char* version();   // "raw-native 0.8.0": the reference release whose formats these are

using FileDigests = std::vector<std::pair<std::string, std::string>>;
std::string sha256Hex(const std::string& bytes);   // FIPS 180-4
std::string sha256File(const std::string& path);   // "" when unreadable
std::string rgb8Bytes(const Buffer<Vec3>& img);
std::string f32Bytes(const Buffer<float>& img);
void writePPM(const Buffer<Vec3>& img, const std::string& path);
void writePFM(const Buffer<Vec3>& img, const std::string& path);
void writePFM1(const Buffer<float>& img, const std::string& path);
void writePGM(const Buffer<float>& img, const std::string& path);
void writeMaskPGM(const Buffer<uint8_t>& mask, const std::string& path);
bool readPFM1(const std::string& path, Buffer<float>& out);
bool readMaskPGM(const std::string& path, Buffer<uint8_t>& out);

FileDigests writeFrameFiles(const FrameResult& o, const CliParams& p);               // run.cpp:27-44
std::string aoCertificateJson(const FrameResult& o, const CliParams& p, const FileDigests& outputs);
std::string channelsJson(const FrameResult& o, const CliParams& p, int readoutW = 8, int readoutH = 8);
std::string aoReceiptJson(const FrameResult& o, const CliParams& p, const std::string& backend, const FileDigests& outputs);
// The files of emitOutputs (app/main.cpp:69-88) but arena_certificate.json: the images,
// certificate.json, channels.json and receipt.json under p.out. 0, or 2 when a file cannot be
// written (err names it).
int emitOutputs(const FrameResult& o, const CliParams& p, const std::string& backend, std::string& err);
// raw_native_cli verify DIR (src/tools/verify.cpp): 0 all match, 2 missing, 3 mismatch.
int verifyDir(const std::string& dir, std::string& report);

// ---- the GPU frame against the CPU reference: raw-gpu-cert/1 (raw/renderer/gpu_reconcile.hpp,
// src/renderer/gpu_reconcile.cpp, raw/cert/gpu_tolerance.hpp) -------------------------------
struct AdapterInfo { std::string vendor, architecture, device, description, backend, driver; };
AdapterInfo adapterInfo();   // this device's (raw_ao_adapter_info); empty without a GPU
struct GpuChannelCheck {
    std::string name;
    long long values{0};
    double rmse{0}, maxError{0}, bound{0};
    bool pass{false};
};
struct GpuReconcile {
    int width{0}, height{0};
    long long coveredEither{0}, coveredBoth{0}, maskMismatch{0};
    double maskMismatchFraction{0};
    bool maskPass{false};
    std::vector<GpuChannelCheck> channels;
    bool rt{true};
    float cpuAoRmse{0}, gpuAoRmse{0};
    bool cpuWithin{false}, gpuWithin{false};
    bool reconcilePass{false}, verdictPass{false};
    bool sizeMatch{false};
    std::string reason;
    std::string gpuFrameSha256, cpuFrameSha256;
    bool pass() const;
};
GpuReconcile reconcileGpuCpu(const FrameResult& gpu, const FrameResult& cpu, bool rt);
std::string gpuCertificateJson(const GpuReconcile& r, const AdapterInfo& a, const std::string& renderer,
                               const std::string& paramsJson, double gpuMs, double cpuMs);
// gpu_receipt.json (src/tools/receipt.cpp:150-196); gpu or cpu may be null (no frame)
std::string gpuReceiptJson(const FrameResult* gpu, const FrameResult* cpu, const GpuReconcile& r, const CliParams& p,
                           const FileDigests& outputs);

}  // namespace raw
