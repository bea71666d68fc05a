// The learned AO surrogate's network state (kernels and host driver in unet.hip).
#pragma once
#include "../../include/raw_ao_hip.h"
#include "ctx.hpp"

namespace raw_ao {

// Feature channels a pixel (bf16): covered, eye distance / 20, octahedral normal (2),
// SSAO, noisy RT AO (0 when absent), normal y, noisy-AO present.
constexpr int kNetFeatures = 8;
constexpr int kNetLayers = 14;

struct UNet {
    struct Layer {
        int c0{0}, c1{0}, cout{0}, level{0};
        int K{0}, kpad{0}, kpad_t{0};
        size_t woff{0}, boff{0};        // fp32 master offsets (floats)
        size_t wofs16{0}, wtofs16{0};   // bf16 forward / input-grad weight offsets (elements)
    };
    Ctx* ctx{nullptr};
    Layer layers[kNetLayers];
    size_t head_off{0};
    int step{0};
    DevBuf params, grads, adam_m, adam_v;   // fp32, param_count() each
    DevBuf w16, wt16;                       // packed bf16 weights
    // activations for the current shape
    int an{0}, ah{0}, aw{0};
    bool have_grad_bufs{false};
    DevBuf pred, headpart, lossbuf, cnt, wpart;
    DevBuf act[kNetLayers];
    DevBuf pooled[3];
    DevBuf lvbuf[4][3];                     // backward: three rotating grad buffers a level
    const unsigned short* skip_dx[3]{};     // decoder input grads whose first channels feed a skip
    int skip_cx[3]{};
    // the learned frame: SSAO, features, the probe sub-image and its statistics
    DevBuf f_ssao, f_x, pr_pos, pr_nrm, pr_mask, pr_rt, pr_net, pr_stats;
    float* h_stats{nullptr};                // pinned: sum of squares (as 2 floats of a double), max, count

    size_t param_count() const;
    bool init(Ctx& c);
    bool pack(hipStream_t s);
    bool ensure_acts(int n, int h, int w, bool grad);
    bool features(const float* pos, const float* nrm, const uint8_t* mask, const float* ssao, const float* noisy,
                  const float eye[3], int h, int w, unsigned short* x, hipStream_t s);
    // x: n images of h x w pixels, kNetFeatures bf16 channels a pixel
    bool forward(const unsigned short* x, const uint8_t* mask, int n, int h, int w, float* ao_out, hipStream_t s);
    // after forward on the same input: gradients of the mean squared error over covered pixels
    bool backward(const unsigned short* x, const uint8_t* mask, const float* target, int n, int h, int w, hipStream_t s);
    bool adam(float lr, hipStream_t s);
    void release();
};

struct LearnedArgs {
    const float* pos;
    const float* nrm;
    const uint8_t* mask;
    const float* noisy;
    const float* eye;
    int w, h, ssao_samples;
    float ssao_radius;
    int rt_spp;
    float rt_radius;
    int probe_stride;
    float fallback_rmse;
    float* ao;
};
// SSAO, features and the network into a.ao; with a probe stride, a 64-spp ray-traced
// probe of every stride-th pixel estimates the frame's error, and above fallback_rmse the
// frame is ray traced instead.
int learned_frame(Ctx& c, UNet& u, const LearnedArgs& a, raw_ao_learned_report* rep, hipStream_t s);
// trace_ao / trace_ao4 of the built scene, with the checks of the C ABI
int trace_built(Ctx& c, const float* pos, const float* nrm, const uint8_t* mask, int w, int h, int spp, float radius, float* ao,
                hipStream_t s);

// Single-layer kernels for the tests and the host driver.
// pool (or null): also out's 2x2 average pool, N,H/2,W/2,cout (h and w even)
bool conv3x3(Ctx& c, const unsigned short* in0, const unsigned short* in1, int c0, int c1, const unsigned short* w, int kpad,
             const float* bias, int cout, int relu, int n, int h, int wd, unsigned short* out, hipStream_t s,
             unsigned short* pool = nullptr);
bool conv3x3_wgrad(Ctx& c, DevBuf& part, const unsigned short* dy, const unsigned short* in0, const unsigned short* in1, int c0,
                   int c1, int cout, int n, int h, int wd, float* gw, float* gb, hipStream_t s);

}  // namespace raw_ao
