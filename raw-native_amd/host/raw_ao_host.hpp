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
struct Tri { Vec3 a, b, c; };
template <class T> struct Buffer {
    int w{0}, h{0};
    std::vector<T> px;
    void resize(int W, int H) { w = W; h = H; px.assign(size_t(W) * size_t(H), T{}); }
    T& at(int x, int y) { return px[size_t(y) * size_t(w) + size_t(x)]; }
    const T& at(int x, int y) const { return px[size_t(y) * size_t(w) + size_t(x)]; }
};
struct GBuffer {
    int w{0}, h{0};
    Buffer<float> depth;
    Buffer<Vec3> normal, position;
    Buffer<uint8_t> mask;
    void resize(int W, int H) { w = W; h = H; depth.resize(W, H); normal.resize(W, H); position.resize(W, H); mask.resize(W, H); }
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

}  // namespace raw
