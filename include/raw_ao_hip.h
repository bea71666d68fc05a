/* raw_ao_hip: the MI355X (gfx950) implementation of raw-native's ray-traced AO path,
 * behind a plain C ABI. Everything is synchronous unless a function takes a stream;
 * every function returns RAW_AO_OK (0) or a negative code and leaves the message in
 * raw_ao_last_error(ctx). No C++ exception crosses this boundary. The caller owns
 * host arrays; the context owns device memory. One context per host thread.
 * Results never depend on launch, tile or queue parameters.
 *
 * Layouts are the reference's: row-major images, y down; a Vec3 is 3 floats (AoS, as
 * raw::Buffer<Vec3> stores it); masks are one byte a pixel (0 off, nonzero covered);
 * triangles are 9 floats (a, b, c) as raw::Tri (raw/math/primitives.hpp:7).
 *
 * Entry points ending in _dev take DEVICE pointers (e.g. from hipMalloc or a torch
 * tensor's data_ptr) and a hipStream_t passed as void* (NULL: the context's stream);
 * they enqueue work and return without waiting.
 */
#ifndef RAW_AO_HIP_H
#define RAW_AO_HIP_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define RAW_AO_ABI_VERSION 1u
/* The largest scene a context builds (2^25 triangles: 32-bit node offsets). */
#define RAW_AO_MAX_TRIANGLES (1u << 25)

enum {
    RAW_AO_OK = 0,
    RAW_AO_ERR_ARG = -1,        /* a bad argument (sizes, null pointers, no scene) */
    RAW_AO_ERR_DEVICE = -2,     /* a HIP call failed */
    RAW_AO_ERR_BUILD = -3,      /* the PLOC build could not finish (a round without a merge) */
    RAW_AO_ERR_NO_DEVICE = -4   /* no GPU: the reference's exit code 4, an "unverifiable" run */
};

typedef struct raw_ao_ctx raw_ao_ctx;

/* The ABI version this library implements (RAW_AO_ABI_VERSION). */
uint32_t raw_ao_abi_version(void);

/* Open a context on device `device_ordinal`. Replaces the reference's process-wide
 * rhi::device() (raw/rhi/rhi.hpp:284-286); RAW_AO_ERR_NO_DEVICE where it would report
 * no adapter (src/tools/gpu_run.cpp:52-59). On failure *out is still a valid context
 * that carries the message (close it). */
int raw_ao_open(int device_ordinal, raw_ao_ctx** out);
void raw_ao_close(raw_ao_ctx* ctx);
/* The first error since the last successful call; "" when none. */
const char* raw_ao_last_error(const raw_ao_ctx* ctx);

/* Options (raw_ao_set_option). */
enum {
    /* 1 (default): sin/cos of the sample angle reproduce the HOST libm bit for bit
     * (the reference calls std::cos/std::sin, src/renderer/ray_ao.cpp:36 and
     * src/renderer/ssao.cpp:21-24); 0: correctly rounded. */
    RAW_AO_OPT_HOST_LIBM_TRIG = 1,
    /* The traversal form, chosen before raw_ao_build_ploc: RAW_AO_TRACE_WIDE (default)
     * or RAW_AO_TRACE_BINARY. Results are the same; only the speed differs. */
    RAW_AO_OPT_TRACE_KERNEL = 2,
    /* Wide kernel tuning, 1..64 (default 16): a wave runs its triangle phase once this
     * many 64ths of its live lanes wait on leaf triangles (or none can take a node step).
     * Results are the same for every value; only the speed differs. */
    RAW_AO_OPT_TRI_PHASE = 3,
    /* The network's convolution kernel: 0 (default) the LDS-tiled kernel wherever a
     * layer's input channels are a multiple of 32 (its float32 sums run chunk-major, so
     * outputs can differ from the others' by one bf16 rounding), in its persistent form
     * (the next tile's data loaded while the current one runs) for layers of 32 inputs and
     * 32 outputs and in its direct-copy persistent form (each chunk's input and weights
     * copied to LDS while the previous chunk runs) for a concatenated input with 32 outputs,
     * 1 the general kernel everywhere, 2 the tap-major kernel where it applies (the general
     * kernel's sum order: identical outputs to 1), 3 as 0 without the persistent forms, 4 the
     * direct-copy form wherever 0 uses the LDS-tiled kernel (3 and 4: identical outputs to 0). */
    RAW_AO_OPT_CONV_KERNEL = 4,
    /* PLOC tuning, 1..4096 (default 1024): at or below this many clusters one workgroup runs every
     * remaining round. The tree is the same for every value; only the speed differs. */
    RAW_AO_OPT_PLOC_TAIL = 5,
    /* 64-spp wide kernel, 0..64 (default 48): with a value n > 0 a wave owns 4 pixels' 256 rays
     * (with the wave cut: from the cut of their union box) and a lane that finishes takes the
     * next ray once n lanes are idle; 0 = one pixel's 64 rays a wave. Results are the same for
     * every value; only the speed differs. */
    RAW_AO_OPT_REFILL = 6,
    /* 64-spp wide kernel, 0 or 1 (default 1): descend into the hit child box with the smallest
     * entry distance first (0: the first in slot order). Results are the same; only the speed
     * differs. */
    RAW_AO_OPT_NEAR_FIRST = 7,
    /* 64-spp wide kernel, 0..64 (default 2): with n > 0 the 64 lanes of a pixel first walk the
     * top of the tree together, keeping the subtrees whose boxes meet the box the pixel's rays
     * can reach and are no wider than n x radius, and every ray starts from those (n = 0: every
     * ray starts from the root; RAW_AO_OPT_NEAR_FIRST applies to that traversal only). Results
     * are the same for every value; only the speed differs. */
    RAW_AO_OPT_WAVE_CUT = 8,
    /* 0 (default) or 1: the reference's skew control (RAW_NATIVE_AO_ORACLE=skew,
     * src/renderer/accel.cpp:6-11,26-28): every AO query distance of raw_ao_trace_ao[_dev] is
     * radius x 0.99f. A negative control for tests: the AO must then differ from the golden. */
    RAW_AO_OPT_SKEW = 9,
    /* Build tuning, 0..4096 (default 0): the build enqueues its PLOC rounds and wide-collapse
     * levels without reading the cluster count back, as many as the last build of as many
     * triangles needed (0: that, or an estimate for a first build), or this many; rounds or
     * levels past the end do nothing, and a build left short continues with read-backs. The tree
     * is the same for every value; only the speed differs. */
    RAW_AO_OPT_BUILD_ROUNDS = 10,
    /* 64-spp visibility (raw_ao_trace_visibility[_dev], the frame's lighting model 2), 0..64
     * (default 20): with n > 0 (and RAW_AO_OPT_REFILL and RAW_AO_OPT_WAVE_CUT above 0) the AO rays
     * are traced first, then the open ones' sky rays apart from them, a wave taking 8 pixels'
     * rays and a lane whose ray ends the next one once n lanes are idle; 0 = each sky ray right
     * after its AO ray in the same lane. Results are the same for every value; only the speed
     * differs. */
    RAW_AO_OPT_SKY_REFILL = 11,
    /* Learned frames with a probe (raw_ao_learned_ao[_dev], probe_stride > 0), 0..1024 (default
     * 0, off): with n > 0, a frame that fell back makes the next calls of the same image size ray
     * trace at once, without the network and the probe (fell_back = 2), for 1, 2, 4, ... up to n
     * calls after consecutive fall-backs, then try the network again; a frame whose probe passes
     * clears it. For serving loops where the network misses the bound for a while: the frames
     * that skip the network cost what plain ray tracing costs. */
    RAW_AO_OPT_LEARNED_BACKOFF = 12,
    /* 0, 1 or 2 (default 1): which kernels read the wide tree as 64 B quantized records (per
     * axis a base and a power-of-two step, one byte a slot plane, every box rounded outward)
     * instead of the 128 B ones: 0 none; 1 the visibility pass's sky rays (64 spp, the split form
     * of RAW_AO_OPT_SKY_REFILL), the records made by the first such pass after each build; 2 also
     * the 64-spp AO rays of the refill kernel (RAW_AO_OPT_REFILL and RAW_AO_OPT_WAVE_CUT above
     * 0), the records made after each wide build, or at the next trace call when the option is
     * set after the build (the no-allocation trace uses them only once made). The boxes only
     * grow, so every result is the same bit for bit; only the speed and the bytes read differ
     * (1080p on the 1M-triangle scene: the latency-bound sky rays run faster, the AO rays, bound
     * by VALU issue, slower). The wave cut and the other kernels read the 128 B records. */
    RAW_AO_OPT_QUANTIZED_NODES = 13
};
enum {
    RAW_AO_TRACE_WIDE = 0,     /* 4-wide nodes, LDS stack, triangle tests batched per wave */
    RAW_AO_TRACE_BINARY = 1    /* the binary PLOC tree, both children's boxes in one 64 B record */
};
int raw_ao_set_option(raw_ao_ctx* ctx, int option, int value);

/* ---- the scene: PLOC BVH build on the device ---------------------------------------
 * Replaces buildPlocGpu (raw/renderer/rt_gpu.hpp:19; src/renderer/gpu/rt_gpu.cpp:25-95)
 * and, as the AO accelerator, LinearAccel::build (raw/renderer/accel.hpp:19,
 * src/renderer/accel.cpp:13-25): triangles in, the context's BVH out. radius: the PLOC
 * search radius (raw::rt::kPlocRadius = 16, raw/renderer/rt_bvh.hpp:30), 1..32. */
int raw_ao_build_ploc(raw_ao_ctx* ctx, const float* tris9, uint32_t ntri, int radius);
int raw_ao_build_ploc_dev(raw_ao_ctx* ctx, const float* d_tris9, uint32_t ntri, int radius, void* stream);

/* The built binary tree as raw::rt::Tree (raw/renderer/rt_bvh.hpp:23-28): node count
 * (2n-1), root id and rounds; then the nodes in the reference's order (leaves 0..n-1 in
 * sorted order, internal nodes in creation order): boxes6 {lo xyz, hi xyz} and links3
 * {left, right, tri} per node. For the node-for-node identity check
 * (src/renderer/gpu/rt_bvh_parity.cpp:80-123). */
int raw_ao_tree_info(raw_ao_ctx* ctx, uint32_t* nnodes, int32_t* root, uint32_t* rounds);
int raw_ao_read_tree(raw_ao_ctx* ctx, float* boxes6, int32_t* links3);
/* The build's Morton keys (keys[i] of input triangle i) and sorted order (order[p], the
 * triangle at sorted position p) (src/renderer/rt_bvh.cpp:56-73). */
int raw_ao_read_keys(raw_ao_ctx* ctx, uint32_t* keys, uint32_t* order);

/* ---- ray-traced AO (the raw-rt-ao-v1 oracle) ---------------------------------------
 * Replaces computeRTAO (raw/renderer/ray_ao.hpp:6-8, src/renderer/ray_ao.cpp:20-58)
 * over the context's BVH: ao[y*w+x] = open/spp for covered pixels, 1.0 elsewhere.
 * position, normal: w*h Vec3 each; mask: w*h bytes; spp 1..4096; radius > 0. */
int raw_ao_trace_ao(raw_ao_ctx* ctx, const float* position, const float* normal, const uint8_t* mask,
                    int w, int h, int spp, float radius, float* ao);
int raw_ao_trace_ao_dev(raw_ao_ctx* ctx, const float* d_position, const float* d_normal, const uint8_t* d_mask,
                        int w, int h, int spp, float radius, float* d_ao, void* stream);

/* No reference counterpart: raw_ao_trace_ao_dev for callers that cannot allocate device memory
 * (code under HIP stream capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * arguments and results, bit for bit; it calls no hipMalloc, hipFree or allocating library
 * routine. Its scratch (trace counters, per-pixel open counts below 64 spp) is one library-wide
 * array per device, so w * h <= RAW_AO_NOALLOC_MAX_PIXELS (larger is RAW_AO_ERR_ARG). Needs the
 * scene built with the wide traversal form (the build sets up everything else it reads). Calls
 * that overlap on one device (several streams or host threads) race on that array. */
#define RAW_AO_NOALLOC_MAX_PIXELS (1u << 22)
int raw_ao_trace_ao_dev_noalloc(raw_ao_ctx* ctx, const float* d_position, const float* d_normal, const uint8_t* d_mask,
                                int w, int h, int spp, float radius, float* d_ao, void* stream);

/* ---- lighting model 2's visibility ---------------------------------------------------
 * Replaces computeRTVisibility (raw/renderer/ray_ao.hpp:12-14, src/renderer/ray_ao.cpp:18-63),
 * the pass the default render runs (src/renderer/render.cpp:31-36): ao exactly as
 * raw_ao_trace_ao; sky = the fraction of the samples open within radius that also hit
 * nothing with tMax 1e30; sun = 1 when the ray from the AO origin towards
 * normalize(to_light) hits nothing with tMax 1e30, or when sun_forced (the no-sun control);
 * all three 1.0 off the mask. Needs the wide traversal form. */
int raw_ao_trace_visibility(raw_ao_ctx* ctx, const float* position, const float* normal, const uint8_t* mask,
                            int w, int h, int spp, float radius, const float to_light[3], int sun_forced,
                            float* ao, float* sky, float* sun);
int raw_ao_trace_visibility_dev(raw_ao_ctx* ctx, const float* d_position, const float* d_normal, const uint8_t* d_mask,
                                int w, int h, int spp, float radius, const float to_light[3], int sun_forced,
                                float* d_ao, float* d_sky, float* d_sun, void* stream);

/* ---- the G-buffer -------------------------------------------------------------------
 * Replaces raw::rasterize (raw/renderer/raster.hpp:9, src/renderer/raster.cpp:92-126) with
 * default RasterOptions, for one indexed mesh (vertex positions and normals, 3 floats each;
 * 3 indices a triangle) and the camera of src/renderer/raster.cpp:94-96 (perspective(fovy,
 * aspect, near, far) * lookAt(eye, center, up), raw/math/mat.hpp:21-38): depth = the
 * screen-linear view distance (+inf where nothing covers the pixel), position and normal
 * (3 floats a pixel, zero where uncovered), mask = 1 where covered. Bit for bit the
 * reference's result: a pixel keeps the nearest fragment, the earliest drawn among equals.
 * The _dev form reads indices as given: each must be below the vertex count. */
int raw_ao_rasterize(raw_ao_ctx* ctx, const float* vpos, const float* vnrm, uint32_t nverts, const int32_t* idx,
                     uint32_t ntri, const float eye[3], const float center[3], const float up[3], float fovy,
                     float aspect, float near_z, float far_z, int w, int h, float* depth, float* position,
                     float* normal, uint8_t* mask);
int raw_ao_rasterize_dev(raw_ao_ctx* ctx, const float* d_vpos, const float* d_vnrm, const int32_t* d_idx, uint32_t ntri,
                         const float eye[3], const float center[3], const float up[3], float fovy, float aspect,
                         float near_z, float far_z, int w, int h, float* d_depth, float* d_position, float* d_normal,
                         uint8_t* d_mask, void* stream);

/* ---- the adapter -------------------------------------------------------------------
 * What the device says about itself, for raw-gpu-cert/1 (rhi::AdapterInfo,
 * raw/rhi/rhi.hpp:39-41): vendor, architecture (the gfx target), device (marketing name),
 * description (compute units and memory), backend ("hip"), driver (HIP runtime and driver
 * versions). Zero-terminated, truncated to the field size. */
typedef struct raw_ao_adapter {
    char vendor[64], architecture[64], device[128], description[128], backend[16], driver[64];
} raw_ao_adapter;
int raw_ao_adapter_info(raw_ao_ctx* ctx, raw_ao_adapter* out);

/* ---- the default frame on the GPU ----------------------------------------------------
 * Replaces renderGpu (raw/renderer/gpu.hpp:32, src/renderer/gpu/gpu_renderer.cpp:129-182) with
 * the passes of the CPU reference frame, renderWithParams (raw/renderer/render.hpp:50-52,
 * src/renderer/render.cpp:71-105), bit for bit: raster (G-buffer with albedo), the AO
 * accelerator over the scene's triangles, lighting model 2's visibility (or model 1's AO),
 * SSAO, the reconcile, shade and shade_hdr (src/renderer/composite.cpp), motion
 * (src/renderer/motion.cpp). A scene is its meshes concatenated in scene order: vertices,
 * indices, and per mesh the end of its triangle range and its albedo. */
typedef struct raw_ao_scene {
    const float* vpos;          /* nverts x 3 */
    const float* vnrm;          /* nverts x 3 */
    uint32_t nverts;
    const int32_t* idx;         /* ntri x 3 */
    uint32_t ntri;
    const uint32_t* mesh_end;   /* host, nmesh: one past each mesh's last triangle (the last = ntri) */
    const float* albedo;        /* host, nmesh x 3 */
    uint32_t nmesh;
    float eye[3], center[3], up[3];
    float fovy, aspect, near_z, far_z;
    int32_t has_light;          /* 0: the reference's default light, direction (0,-1,0), intensity 1 */
    float light_dir[3];
    float light_intensity;
} raw_ao_scene;

/* RenderOptions (raw/renderer/render.hpp:40-45) with lightingModel() resolved by the caller:
 * lighting 2 shades with sky and sun visibility, 1 with (0.2 + sun) x AO; rtao 0 shades model 1
 * with SSAO and leaves the reconcile empty (pixels 0). */
typedef struct raw_ao_render_opts {
    float tolerance;            /* 0.12 */
    int32_t rtao;               /* 1 */
    int32_t lighting;           /* 2 */
    int32_t no_sun;             /* the v2-nosun control: sun visibility forced to 1 */
    int32_t rt_samples;         /* 64 (kRtSamples) */
    int32_t ss_samples;         /* 24 (kSsSamples) */
    float radius;               /* 2.0 (kAoRadius) */
    int32_t bvh_radius;         /* PLOC search radius, 16 */
    const float* prev_viewproj; /* host, 16 floats row-major; NULL: this frame's (zero motion) */
} raw_ao_render_opts;

/* The frame's planes, w x h each (position, normal, albedo, frame, hdr: 3 floats a pixel;
 * motion: 2). raw_ao_render_frame takes host pointers and skips NULL ones;
 * raw_ao_render_frame_dev takes device pointers, all required (sky and sun are written with
 * model 2 only). */
typedef struct raw_ao_frame {
    float* depth;
    float* position;
    float* normal;
    float* albedo;
    uint8_t* mask;
    float* motion;
    float* ao_rt;
    float* ao_ss;
    float* sky;
    float* sun;
    float* error;
    float* frame;
    float* hdr;
} raw_ao_frame;

/* The frame's scalars: the reconcile (rmse from the float64 row-major sum), the motion counts,
 * maxRadiance of the HDR plane (src/renderer/composite.cpp:82-88), the view-projection, and the
 * frame's GPU time. */
typedef struct raw_ao_frame_stats {
    float rmse, max_error;
    int32_t pixels, within_tolerance;
    int32_t motion_valid, motion_total;
    float max_radiance;
    float viewproj[16];
    float gpu_ms;
} raw_ao_frame_stats;

int raw_ao_render_frame(raw_ao_ctx* ctx, const raw_ao_scene* scene, const raw_ao_render_opts* opts, int w, int h,
                        raw_ao_frame* out, raw_ao_frame_stats* stats);
/* Scene vertices and indices in device memory (mesh_end and albedo stay host arrays). */
int raw_ao_render_frame_dev(raw_ao_ctx* ctx, const raw_ao_scene* scene, const raw_ao_render_opts* opts, int w, int h,
                            raw_ao_frame* d_out, raw_ao_frame_stats* stats, void* stream);
/* The reference's defaults: tolerance 0.12, rtao on, lighting 2, 64/24 samples, radius 2, PLOC
 * radius 16, static camera. */
void raw_ao_render_opts_default(raw_ao_render_opts* opts);

/* ---- screen-space AO ---------------------------------------------------------------
 * Replaces computeSSAO (raw/renderer/ssao.hpp:5-6, src/renderer/ssao.cpp:11-39). */
int raw_ao_ssao(raw_ao_ctx* ctx, const float* position, const float* normal, const uint8_t* mask,
                int w, int h, int samples, float radius, float* ao);
int raw_ao_ssao_dev(raw_ao_ctx* ctx, const float* d_position, const float* d_normal, const uint8_t* d_mask,
                    int w, int h, int samples, float radius, float* d_ao, void* stream);

/* ---- the reconcile's per-pixel part ------------------------------------------------
 * The error map and maximum of reconcile (raw/cert/reconcile.hpp:12-13,
 * src/cert/reconcile.cpp:11-16): err[i] = |approx - truth| where covered, else 0;
 * *max_error = the largest. The float64 row-major sum of squares stays with the caller
 * (raw_ao_host reconcile()), so the RMSE bits match the reference's. */
int raw_ao_error_map(raw_ao_ctx* ctx, const float* approx, const float* truth, const uint8_t* mask,
                     int w, int h, float* err, float* max_error);
int raw_ao_error_map_dev(raw_ao_ctx* ctx, const float* d_approx, const float* d_truth, const uint8_t* d_mask,
                         int w, int h, float* d_err, float* d_max_error, void* stream);

/* ---- general rays through the context's BVH ----------------------------------------
 * Replaces traceGpu (raw/renderer/rt_gpu.hpp:27-28, src/renderer/gpu/rt_gpu.cpp:97-133):
 * rays8 {origin xyz, direction xyz, tMax, pad}; any != 0 stops at the first hit with
 * t < tMax; otherwise the nearest (ties to the lower triangle index). out4 per ray:
 * {t or -1, triangle as a float-exact integer (-1 none), nodes visited, triangles tested}.
 * far_scale widens each box's exit (1 + 2 gamma(3) normally; a value below 1 is the
 * reference's short-far control). */
int raw_ao_trace_rays(raw_ao_ctx* ctx, const float* rays8, uint32_t nrays, int any, float far_scale, float* out4);

/* ---- closest hits with barycentrics -------------------------------------------------
 * Replaces rt::closest (raw/renderer/rt_bvh.hpp:55, src/renderer/rt_bvh.cpp:125-152), the
 * query the path tracer runs on the PLOC tree PtScene builds (src/renderer/rt_pt_scene.cpp:36),
 * for a batch of rays through the context's BVH: rays8 {origin xyz, direction xyz, tMax, pad};
 * hits4 per ray {t, u, v, triangle (int32)} as HitOut (raw/rhi/hw.hpp:125): the nearest hit with
 * t < tMax, ties to the lower triangle index, u and v intersectTri's weights of b and c; a miss
 * is {-1, 0, 0, -1}. Bit for bit the reference's Hit. Needs the wide traversal form. */
int raw_ao_closest_hits(raw_ao_ctx* ctx, const float* rays8, uint32_t nrays, void* hits4);
int raw_ao_closest_hits_dev(raw_ao_ctx* ctx, const float* d_rays8, uint32_t nrays, void* d_hits4, void* stream);

/* ---- measurement -------------------------------------------------------------------
 * Node visits and triangle tests of the last trace_ao over its rays (from a separate
 * counting launch when enabled with raw_ao_count_next_trace), and the milliseconds of the
 * last launches of each kernel, timed with HIP events on the launch stream. */
int raw_ao_count_next_trace(raw_ao_ctx* ctx, int enable);
int raw_ao_trace_counts(raw_ao_ctx* ctx, uint64_t* node_visits, uint64_t* tri_tests, uint64_t* rays);
int raw_ao_last_ms(raw_ao_ctx* ctx, float* build_ms, float* trace_ms, float* ssao_ms, float* error_ms);
/* Verification entry point (no reference counterpart): the sin and cos of every sample
 * angle phi = 6.2831853f * k / 2^24, k < 2^24, exactly as the AO and SSAO kernels use
 * them, so a test can compare all of them with the host libm (sin_out, cos_out: 2^24
 * floats each). */
int raw_ao_angle_table(raw_ao_ctx* ctx, float* sin_out, float* cos_out);
/* Bytes of the context's device-resident BVH (nodes + triangles) in its traversal form. */
int raw_ao_bvh_bytes(raw_ao_ctx* ctx, uint64_t* node_bytes, uint64_t* tri_bytes, uint32_t* node_record_bytes);
/* Verification entry point (no reference counterpart): the wide tree's records copied to the
 * host, up to cap of each kind: records128 the 128 B records (6 x 4 slot planes lo x, y, z, hi
 * x, y, z as floats, 4 int32 refs, the masks word, 12 B spare), records64 their 64 B quantized
 * forms (RAW_AO_OPT_QUANTIZED_NODES: base x, y, z as floats, the exponents word, the lo x, y, z
 * and hi x, y, z byte words, the masks word, 0, the 4 refs); either may be NULL. *count = the
 * tree's record count. RAW_AO_ERR_ARG without a wide tree, or for records64 when the quantized
 * records are not made. */
int raw_ao_read_wide_records(raw_ao_ctx* ctx, void* records128, void* records64, uint32_t cap, uint32_t* count);

/* ---- batched AO: many views of several scenes ---------------------------------------
 * No reference function (the reference renders one view of one scene per call,
 * src/renderer/render.cpp:26-54); each view's outputs are exactly computeRTAO's and
 * computeSSAO's for that view, and the reconcile's error map (src/cert/reconcile.cpp:4-21).
 * Scenes are built (eagerly, recording their PLOC and collapse grids) when added; a run
 * traces every view's rays in ONE persistent launch whose waves take 64-ray groups from a
 * global queue, then finishes each view's AO, SSAO and reconcile on the device. The
 * caller's device arrays must outlive the batch. The per-view RMSE sums squares in float64
 * on the device in a fixed order (not the reference's row-major order: the float RMSE can
 * differ in its last bit). */
typedef struct raw_ao_batch raw_ao_batch;
int raw_ao_batch_open(int device_ordinal, raw_ao_batch** out);
void raw_ao_batch_close(raw_ao_batch* batch);
const char* raw_ao_batch_last_error(const raw_ao_batch* batch);
/* d_tris9: ntri triangles (9 floats) on the device, read again by every rebuild. */
int raw_ao_batch_add_scene(raw_ao_batch* batch, const float* d_tris9, uint32_t ntri, int radius, int* scene_id);
/* One view of a scene: G-buffer in, d_ao out (spp rays of length radius a covered pixel);
 * d_ssao (ssao_samples, ssao_radius) and d_err (|ssao - ao| where covered) may be NULL. */
int raw_ao_batch_add_view(raw_ao_batch* batch, int scene, const float* d_pos, const float* d_nrm, const uint8_t* d_mask,
                          int w, int h, int spp, float radius, int ssao_samples, float ssao_radius, float* d_ao,
                          float* d_ssao, float* d_err, int* view_id);
/* Launch parameters: 64-ray groups a wave takes from the queue at a time, the triangle
 * phase threshold (of 64 lanes), persistent blocks per compute unit. Results never depend
 * on them. */
int raw_ao_batch_set_tuning(raw_ao_batch* batch, int chunk, int tri_q, int blocks_per_cu);
int raw_ao_batch_get_tuning(raw_ao_batch* batch, int* chunk, int* tri_q, int* blocks_per_cu);
/* RAW_AO_OPT_HOST_LIBM_TRIG or RAW_AO_OPT_WAVE_CUT for every view of the batch (as
 * raw_ao_set_option; drops a captured graph). Results never depend on the wave cut. */
int raw_ao_batch_set_option(raw_ao_batch* batch, int option, int value);
/* Enqueues one batched frame (rebuild != 0: first rebuild every scene from its triangles
 * with the recorded grids) on stream (NULL: the batch's own). */
int raw_ao_batch_run(raw_ao_batch* batch, int rebuild, void* stream);
/* Records one batched frame (with the scenes' builds when with_builds != 0) into a hipGraph;
 * raw_ao_batch_launch replays it. After replays with builds, raw_ao_batch_check confirms
 * each rebuilt tree is the recorded one (RAW_AO_ERR_BUILD if a scene changed shape). */
int raw_ao_batch_capture(raw_ao_batch* batch, int with_builds);
int raw_ao_batch_launch(raw_ao_batch* batch, void* stream);
int raw_ao_batch_check(raw_ao_batch* batch);
/* Per view (arrays of the view count): RMSE and max error of SSAO against RT AO over the
 * covered pixels, covered pixels; *rays = the rays traced. Waits for the device. */
int raw_ao_batch_results(raw_ao_batch* batch, float* rmse, float* max_error, uint32_t* pixels, uint64_t* rays);
/* Autotuner: times eager runs over queue chunks, triangle phase and persistent blocks
 * (coordinate descent from the defaults) and keeps the fastest; with cache_path, a
 * result for the same device and workload is read from that text file instead, and new
 * results are appended to it. */
int raw_ao_batch_tune(raw_ao_batch* batch, const char* cache_path, float* best_ms);

/* ---- the learned AO surrogate --------------------------------------------------------
 * No reference function: the reference's only learned precedent is its roughness network
 * (scripts/rr_model.py:43-80, src/rhi/vulkan/vk_rrnet.cpp), whose interface this follows:
 * fp32 master parameters owned by the caller's file, a forward pass on the device, a
 * training step in the caller's loop. The network: a 4-level U-Net (32/64/96/128 channels,
 * two 3x3 convolutions with ReLU a level, 2x2 average pooling, nearest 2x upsampling with
 * skip concatenation, a 1x1 head) over RAW_AO_NET_FEATURES bf16 channels a pixel, bf16
 * weights and activations on MFMA with fp32 accumulation. Images are NHWC, w and h
 * multiples of 8. Parameters: per layer (in order) [cout][9*cin] weights with
 * k = (ky*3 + kx)*cin + ci, then cout biases; the head's 32 weights and bias last. */
#define RAW_AO_NET_FEATURES 8
int raw_ao_net_param_count(raw_ao_ctx* ctx, uint64_t* count);
/* Loads fp32 master parameters and resets the optimizer state. */
int raw_ao_net_set_params(raw_ao_ctx* ctx, const float* params, uint64_t count);
int raw_ao_net_get_params(raw_ao_ctx* ctx, float* params, uint64_t count);
/* The gradients of the last training step (fp32, the parameters' layout). */
int raw_ao_net_get_grads(raw_ao_ctx* ctx, float* grads, uint64_t count);
/* Features of one w x h G-buffer (computeRTAO's inputs plus SSAO and, for the denoiser,
 * a low-spp RT AO; d_noisy may be NULL): covered, distance to eye / 20, octahedral
 * normal (2), SSAO, noisy AO, normal y, noisy-AO present; zeros off the mask. */
int raw_ao_net_features_dev(raw_ao_ctx* ctx, const float* d_pos, const float* d_nrm, const uint8_t* d_mask,
                            const float* d_ssao, const float* d_noisy, const float eye[3], int w, int h,
                            uint16_t* d_x, void* stream);
/* AO of n stacked images: clamp(network, 0, 1) where covered, 1 elsewhere. */
int raw_ao_net_forward_dev(raw_ao_ctx* ctx, const uint16_t* d_x, const uint8_t* d_mask, int n, int w, int h,
                           float* d_ao, void* stream);
/* Forward, backward of the mean squared error against d_target over covered pixels, and
 * (lr > 0) one Adam step (beta 0.9, 0.999, eps 1e-8) on the master parameters. */
int raw_ao_net_train_step_dev(raw_ao_ctx* ctx, const uint16_t* d_x, const uint8_t* d_mask, const float* d_target,
                              int n, int w, int h, float lr, void* stream);
/* The loss of the last training step (waits for the device). */
int raw_ao_net_loss(raw_ao_ctx* ctx, float* loss);
/* Loads a weights file written by raw_ao.learned.save_weights: "RAOW", uint32 version 1,
 * uint64 parameter count, uint32 metadata length, the fp32 parameters (little endian), the
 * metadata (JSON text). */
int raw_ao_net_load(raw_ao_ctx* ctx, const char* path);

/* The learned AO frame, for callers of computeRTAO (raw/renderer/ray_ao.hpp:6-8) that
 * accept a learned answer judged against ray tracing: SSAO (ssao_samples, ssao_radius, as
 * computeSSAO), the features (d_noisy: a low-spp computeRTAO result for the denoiser form,
 * or NULL) and the network into d_ao; w, h multiples of 8. With probe_stride > 0, rt_spp-spp
 * ray-traced AO (rt_radius) of every probe_stride-th pixel in x and y of the context's built
 * scene estimates the frame's RMSE against ray tracing (the report; this waits for the
 * device), and when it exceeds fallback_rmse the frame is ray traced instead (fell_back).
 * The probe traces its own sub-image, so its rays differ from a full frame's at the same
 * pixels: the estimate is statistical. */
typedef struct raw_ao_learned_report {
    uint32_t probe_pixels;   /* covered probe pixels */
    float probe_rmse;        /* the network against rt_spp ray tracing on them */
    float probe_max_error;
    int32_t fell_back;       /* 1: d_ao holds the ray-traced frame; 2: the same, the network and the
                              * probe skipped (RAW_AO_OPT_LEARNED_BACKOFF; probe fields 0) */
} raw_ao_learned_report;
int raw_ao_learned_ao_dev(raw_ao_ctx* ctx, const float* d_pos, const float* d_nrm, const uint8_t* d_mask,
                          const float* d_noisy, const float eye[3], int w, int h, int ssao_samples, float ssao_radius,
                          int rt_spp, float rt_radius, int probe_stride, float fallback_rmse, float* d_ao,
                          raw_ao_learned_report* report, void* stream);
int raw_ao_learned_ao(raw_ao_ctx* ctx, const float* pos, const float* nrm, const uint8_t* mask, const float* noisy,
                      const float eye[3], int w, int h, int ssao_samples, float ssao_radius, int rt_spp, float rt_radius,
                      int probe_stride, float fallback_rmse, float* ao, raw_ao_learned_report* report);

/* Verification entry points: a layer's bf16 output from the last forward pass (count =
 * n * (h >> level) * (w >> level) * cout), and the single-layer kernels: a 3x3 convolution
 * (zero padding) of in0 (c0 channels) concatenated with in1 upsampled 2x (c1 channels, may be
 * 0) by w [cout][kpad] bf16, + bias (may be NULL), ReLU when relu != 0; and its weight
 * gradient against the output grad dy: gw [cout][9*(c0+c1)], gb [cout] (fp32). */
int raw_ao_net_read_act(raw_ao_ctx* ctx, int layer, uint16_t* out, uint64_t count);
int raw_ao_conv3x3_dev(raw_ao_ctx* ctx, const uint16_t* d_in0, const uint16_t* d_in1, int c0, int c1,
                       const uint16_t* d_w, int kpad, const float* d_bias, int cout, int relu, int n, int w, int h,
                       uint16_t* d_out, void* stream);
int raw_ao_conv3x3_wgrad_dev(raw_ao_ctx* ctx, const uint16_t* d_dy, const uint16_t* d_in0, const uint16_t* d_in1,
                             int c0, int c1, int cout, int n, int w, int h, float* d_gw, float* d_gb, void* stream);

/* ---- the versioned function table ---------------------------------------------------
 * The ABI as one symbol a loader resolves (the versioned table the reference's GPU AO path
 * would bind; no reference function): raw_ao_hip_api(RAW_AO_ABI_VERSION) returns this
 * library's table, raw_ao_hip_api(v) NULL for a version it does not implement. The entries are
 * the functions declared above, with the same behaviour. */
typedef struct raw_ao_hip_api_v1 {
    uint32_t abi_version;   /* 1 */
    uint32_t size;          /* sizeof(raw_ao_hip_api_v1) */
    int (*open)(int device_ordinal, raw_ao_ctx** out);
    void (*close)(raw_ao_ctx* ctx);
    const char* (*last_error)(const raw_ao_ctx* ctx);
    int (*set_option)(raw_ao_ctx* ctx, int option, int value);
    int (*build_ploc)(raw_ao_ctx* ctx, const float* tris9, uint32_t ntri, int radius);
    int (*build_ploc_dev)(raw_ao_ctx* ctx, const float* d_tris9, uint32_t ntri, int radius, void* stream);
    int (*tree_info)(raw_ao_ctx* ctx, uint32_t* nnodes, int32_t* root, uint32_t* rounds);
    int (*read_tree)(raw_ao_ctx* ctx, float* boxes6, int32_t* links3);
    int (*read_keys)(raw_ao_ctx* ctx, uint32_t* keys, uint32_t* order);
    int (*trace_ao)(raw_ao_ctx* ctx, const float* position, const float* normal, const uint8_t* mask, int w, int h, int spp,
                    float radius, float* ao);
    int (*trace_ao_dev)(raw_ao_ctx* ctx, const float* d_position, const float* d_normal, const uint8_t* d_mask, int w, int h,
                        int spp, float radius, float* d_ao, void* stream);
    int (*trace_visibility)(raw_ao_ctx* ctx, const float* position, const float* normal, const uint8_t* mask, int w, int h,
                            int spp, float radius, const float to_light[3], int sun_forced, float* ao, float* sky, float* sun);
    int (*ssao)(raw_ao_ctx* ctx, const float* position, const float* normal, const uint8_t* mask, int w, int h, int samples,
                float radius, float* ao);
    int (*ssao_dev)(raw_ao_ctx* ctx, const float* d_position, const float* d_normal, const uint8_t* d_mask, int w, int h,
                    int samples, float radius, float* d_ao, void* stream);
    int (*error_map)(raw_ao_ctx* ctx, const float* approx, const float* truth, const uint8_t* mask, int w, int h, float* err,
                     float* max_error);
    int (*error_map_dev)(raw_ao_ctx* ctx, const float* d_approx, const float* d_truth, const uint8_t* d_mask, int w, int h,
                         float* d_err, float* d_max_error, void* stream);
    int (*trace_rays)(raw_ao_ctx* ctx, const float* rays8, uint32_t nrays, int any, float far_scale, float* out4);
    int (*closest_hits)(raw_ao_ctx* ctx, const float* rays8, uint32_t nrays, void* hits4);
    int (*render_frame)(raw_ao_ctx* ctx, const raw_ao_scene* scene, const raw_ao_render_opts* opts, int w, int h,
                        raw_ao_frame* out, raw_ao_frame_stats* stats);
} raw_ao_hip_api_v1;
const raw_ao_hip_api_v1* raw_ao_hip_api(uint32_t abi_version);

#ifdef __cplusplus
}
#endif
#endif
