// The C ABI of include/raw_ao_hip.h over the gfx950 kernels.
#include "../../include/raw_ao_hip.h"
#include "ctx.hpp"
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

namespace raw_ao {

bool hip_ok(Ctx& c, hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    if (c.err.empty()) c.err = std::string(what) + ": " + hipGetErrorString(e);
    return false;
}
bool ensure(Ctx& c, DevBuf& b, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (b.cap >= bytes) return true;
    if (b.p) { (void)hipFree(b.p); b.p = nullptr; b.cap = 0; }
    void* p = nullptr;
    if (!hip_ok(c, hipMalloc(&p, bytes), "hipMalloc")) return false;
    b.p = p;
    b.cap = bytes;
    return true;
}
int trig_table_upload(Ctx& c) {
    if (c.trig_ready) return 0;
    std::vector<uint8_t> t;
    if (!build_trig_table(t, c.err)) return RAW_AO_ERR_DEVICE;
    if (!hip_ok(c, hipMalloc(reinterpret_cast<void**>(&c.d_trig), t.size()), "hipMalloc trig") ||
        !hip_ok(c, hipMemcpy(c.d_trig, t.data(), t.size(), hipMemcpyHostToDevice), "trig upload"))
        return RAW_AO_ERR_DEVICE;
    c.trig_ready = true;
    return 0;
}

}  // namespace raw_ao

using raw_ao::Ctx;
using raw_ao::ensure;
using raw_ao::hip_ok;

namespace {
// The first error stands until a call succeeds.
int fail(Ctx& c, int code, const char* msg) {
    if (c.err.empty()) c.err = msg;
    return code;
}
int code_of(int internal) {
    if (internal == 0) return RAW_AO_OK;
    if (internal == -1) return RAW_AO_ERR_ARG;
    if (internal == -3) return RAW_AO_ERR_BUILD;
    return RAW_AO_ERR_DEVICE;
}
hipStream_t pick(Ctx& c, void* s) { return s ? static_cast<hipStream_t>(s) : c.stream; }
bool begin(Ctx& c) {
    c.err.clear();
    return hip_ok(c, hipSetDevice(c.device), "hipSetDevice");
}
bool image_ok(int w, int h) { return w > 0 && h > 0 && int64_t(w) * int64_t(h) <= (int64_t(1) << 30); }
// Milliseconds between two of the context's events, 0 when either has not been recorded yet.
// The failed query's error is taken off the thread's last-error slot, so the next launch check
// does not report it.
float elapsed(Ctx& c, int a, int b) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, c.ev[a], c.ev[b]) != hipSuccess) {
        (void)hipGetLastError();
        return 0.0f;
    }
    return ms;
}
}  // namespace

extern "C" {

uint32_t raw_ao_abi_version(void) { return RAW_AO_ABI_VERSION; }

int raw_ao_open(int device_ordinal, raw_ao_ctx** out) {
    if (!out) return RAW_AO_ERR_ARG;
    raw_ao_ctx* x = new (std::nothrow) raw_ao_ctx;
    if (!x) return RAW_AO_ERR_ARG;
    *out = x;
    Ctx& c = x->c;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return fail(c, RAW_AO_ERR_NO_DEVICE, "no GPU device");
    if (device_ordinal < 0 || device_ordinal >= count) return fail(c, RAW_AO_ERR_ARG, "device ordinal out of range");
    c.device = device_ordinal;
    if (!hip_ok(c, hipSetDevice(c.device), "hipSetDevice") || !hip_ok(c, hipStreamCreate(&c.stream), "hipStreamCreate"))
        return RAW_AO_ERR_DEVICE;
    for (auto& e : c.ev)
        if (!hip_ok(c, hipEventCreate(&e), "hipEventCreate")) return RAW_AO_ERR_DEVICE;
    return raw_ao::static_scratch_setup(c);
}

void raw_ao_close(raw_ao_ctx* x) {
    if (!x) return;
    Ctx& c = x->c;
    if (c.stream) {
        (void)hipSetDevice(c.device);
        (void)hipStreamSynchronize(c.stream);
        raw_ao::net_release(c);
        raw_ao::DevBuf* bufs[] = {&c.tris_in, &c.keys, &c.vals, &c.keys2, &c.vals2, &c.hist, &c.rs_tot, &c.boxes, &c.links,
                                  &c.clus_a, &c.clus_b, &c.nn, &c.flags, &c.blk, &c.state, &c.g4, &c.bounds_part,
                                  &c.nodes, &c.tris_sorted, &c.nodes4, &c.nodes4q, &c.wfront_a, &c.wfront_b, &c.wblk, &c.wstate, &c.vis_open, &c.st_pos, &c.st_nrm, &c.st_mask, &c.st_out, &c.st_out2,
                                  &c.st_cnt, &c.st_rays, &c.st_max, &c.counters, &c.r_keys, &c.r_tris, &c.r_counts,
                                  &c.r_offs, &c.r_scan_tmp, &c.fr_mesh, &c.fr_scalars, &c.fr_planes, &c.ploc_hist};
        for (auto* b : bufs)
            if (b->p) (void)hipFree(b->p);
        if (c.d_trig) (void)hipFree(c.d_trig);
        if (c.h_state) (void)hipHostFree(c.h_state);
        if (c.fr_host) (void)hipHostFree(c.fr_host);
        for (auto& e : c.ev)
            if (e) (void)hipEventDestroy(e);
        if (c.aux_fork) (void)hipEventDestroy(c.aux_fork);
        if (c.aux_join) (void)hipEventDestroy(c.aux_join);
        if (c.aux) {
            (void)hipStreamSynchronize(c.aux);
            (void)hipStreamDestroy(c.aux);
        }
        (void)hipStreamDestroy(c.stream);
    }
    delete x;
}

const char* raw_ao_last_error(const raw_ao_ctx* x) { return x ? x->c.err.c_str() : "null context"; }

int raw_ao_set_option(raw_ao_ctx* x, int option, int value) {
    if (!x) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    c.err.clear();
    if (option == RAW_AO_OPT_HOST_LIBM_TRIG) { c.host_trig = value != 0; return RAW_AO_OK; }
    if (option == RAW_AO_OPT_TRACE_KERNEL) {
        if (value != RAW_AO_TRACE_WIDE && value != RAW_AO_TRACE_BINARY) return fail(c, RAW_AO_ERR_ARG, "unknown trace kernel");
        c.trace_kernel = value;
        return RAW_AO_OK;
    }
    if (option == RAW_AO_OPT_NEAR_FIRST) { c.near_first = value != 0; return RAW_AO_OK; }
    if (option == RAW_AO_OPT_SKEW) { c.skew = value != 0; return RAW_AO_OK; }
    if (option == RAW_AO_OPT_BUILD_ROUNDS) {
        if (value < 0 || value > 4096) return fail(c, RAW_AO_ERR_ARG, "build rounds must be in 0..4096");
        c.build_rounds = value;
        return RAW_AO_OK;
    }
    if (option == RAW_AO_OPT_WAVE_CUT) {
        if (value < 0 || value > 64) return fail(c, RAW_AO_ERR_ARG, "wave cut must be in 0..64");
        c.wave_cut = value;
        return RAW_AO_OK;
    }
    if (option == RAW_AO_OPT_LEARNED_BACKOFF) {
        if (value < 0 || value > 1024) return fail(c, RAW_AO_ERR_ARG, "learned back-off must be in 0..1024");
        c.learned_backoff = value;
        c.lb_next = c.lb_left = 0;
        return RAW_AO_OK;
    }
    if (option == RAW_AO_OPT_QUANTIZED_NODES) {
        if (value < 0 || value > 2) return fail(c, RAW_AO_ERR_ARG, "quantized nodes must be 0, 1 or 2");
        c.node_q = value;
        return RAW_AO_OK;
    }
    if (option == RAW_AO_OPT_SKY_REFILL) {
        if (value < 0 || value > 64) return fail(c, RAW_AO_ERR_ARG, "sky refill threshold must be in 0..64");
        c.sky_refill = value;
        return RAW_AO_OK;
    }
    if (option == RAW_AO_OPT_REFILL) {
        if (value < 0 || value > 64) return fail(c, RAW_AO_ERR_ARG, "refill threshold must be in 0..64");
        c.refill = value;
        return RAW_AO_OK;
    }
    if (option == RAW_AO_OPT_PLOC_TAIL) {
        if (value < 1 || value > 4096) return fail(c, RAW_AO_ERR_ARG, "PLOC tail threshold must be in 1..4096");
        c.ploc_tail = value;
        return RAW_AO_OK;
    }
    if (option == RAW_AO_OPT_CONV_KERNEL) {
        if (value < 0 || value > 4) return fail(c, RAW_AO_ERR_ARG, "convolution kernel must be in 0..4");
        c.conv_kernel = value;
        return RAW_AO_OK;
    }
    if (option == RAW_AO_OPT_TRI_PHASE) {
        if (value < 1 || value > 64) return fail(c, RAW_AO_ERR_ARG, "triangle phase threshold must be in 1..64");
        c.tri_q = value;
        return RAW_AO_OK;
    }
    return fail(c, RAW_AO_ERR_ARG, "unknown option");
}

int raw_ao_build_ploc_dev(raw_ao_ctx* x, const float* d_tris9, uint32_t ntri, int radius, void* stream) {
    if (!x) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    if (!begin(c)) return RAW_AO_ERR_DEVICE;
    if (ntri && !d_tris9) return fail(c, RAW_AO_ERR_ARG, "null triangles");
    if (ntri > RAW_AO_MAX_TRIANGLES) return fail(c, RAW_AO_ERR_ARG, "more than RAW_AO_MAX_TRIANGLES triangles");
    hipStream_t s = pick(c, stream);
    if (c.host_trig) {   // the traces' angle table, set up here so raw_ao_trace_ao_dev_noalloc allocates nothing
        const int r = raw_ao::trig_table_upload(c);
        if (r) return r;
    }
    if (ntri) {
        if (!ensure(c, c.tris_in, 36 * size_t(ntri))) return RAW_AO_ERR_DEVICE;
        if (d_tris9 != c.tris_in.p &&
            !hip_ok(c, hipMemcpyAsync(c.tris_in.p, d_tris9, 36 * size_t(ntri), hipMemcpyDeviceToDevice, s), "tri copy"))
            return RAW_AO_ERR_DEVICE;
    }
    (void)hipEventRecord(c.ev[0], s);
    int r = raw_ao::ploc_build(c, c.tris_in.as<float>(), ntri, radius, s);
    if (r == 0) r = raw_ao::bvh_layout(c, s);
    if (r == 0 && c.trace_kernel != RAW_AO_TRACE_BINARY) r = raw_ao::wide_build(c, s);
    c.built_kernel = c.trace_kernel;
    (void)hipEventRecord(c.ev[1], s);
    if (r == 0) r = raw_ao::build_finish(c, s);   // the build's one wait
    if (r != 0) {
        c.root = -1;
        return code_of(r);
    }
    c.ms_build = elapsed(c, 0, 1);
    return RAW_AO_OK;
}

int raw_ao_build_ploc(raw_ao_ctx* x, const float* tris9, uint32_t ntri, int radius) {
    if (!x) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    if (!begin(c)) return RAW_AO_ERR_DEVICE;
    if (ntri && !tris9) return fail(c, RAW_AO_ERR_ARG, "null triangles");
    if (ntri > RAW_AO_MAX_TRIANGLES) return fail(c, RAW_AO_ERR_ARG, "more than RAW_AO_MAX_TRIANGLES triangles");
    if (ntri) {
        if (!ensure(c, c.tris_in, 36 * size_t(ntri)) ||
            !hip_ok(c, hipMemcpyAsync(c.tris_in.p, tris9, 36 * size_t(ntri), hipMemcpyHostToDevice, c.stream), "tri upload"))
            return RAW_AO_ERR_DEVICE;
    }
    return raw_ao_build_ploc_dev(x, ntri ? c.tris_in.as<float>() : nullptr, ntri, radius, nullptr);
}

int raw_ao_tree_info(raw_ao_ctx* x, uint32_t* nnodes, int32_t* root, uint32_t* rounds) {
    if (!x) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    c.err.clear();
    if (nnodes) *nnodes = c.ntri ? 2 * c.ntri - 1 : 0;
    if (root) *root = c.root;
    if (rounds) *rounds = c.rounds;
    return RAW_AO_OK;
}

int raw_ao_read_tree(raw_ao_ctx* x, float* boxes6, int32_t* links3) {
    if (!x) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    if (!begin(c)) return RAW_AO_ERR_DEVICE;
    if (!c.ntri) return RAW_AO_OK;
    if (!boxes6 || !links3) return fail(c, RAW_AO_ERR_ARG, "null output");
    const size_t nn = 2 * size_t(c.ntri) - 1;
    if (!hip_ok(c, hipMemcpy(boxes6, c.boxes.p, 24 * nn, hipMemcpyDeviceToHost), "tree read") ||
        !hip_ok(c, hipMemcpy(links3, c.links.p, 12 * nn, hipMemcpyDeviceToHost), "tree read"))
        return RAW_AO_ERR_DEVICE;
    return RAW_AO_OK;
}

int raw_ao_read_keys(raw_ao_ctx* x, uint32_t* keys, uint32_t* order) {
    if (!x) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    if (!begin(c)) return RAW_AO_ERR_DEVICE;
    if (!c.ntri) return RAW_AO_OK;
    if (!keys || !order) return fail(c, RAW_AO_ERR_ARG, "null output");
    // keys are kept sorted; recompute the per-triangle key order on the host side of the copy
    std::vector<uint32_t> sk(c.ntri);
    if (!hip_ok(c, hipMemcpy(sk.data(), c.keys.p, 4 * size_t(c.ntri), hipMemcpyDeviceToHost), "keys read") ||
        !hip_ok(c, hipMemcpy(order, c.vals.p, 4 * size_t(c.ntri), hipMemcpyDeviceToHost), "order read"))
        return RAW_AO_ERR_DEVICE;
    for (uint32_t p = 0; p < c.ntri; ++p) {
        if (order[p] >= c.ntri) return fail(c, RAW_AO_ERR_DEVICE, "sorted order out of range");
        keys[order[p]] = sk[p];
    }
    return RAW_AO_OK;
}

int raw_ao_trace_ao_dev(raw_ao_ctx* x, const float* d_pos, const float* d_nrm, const uint8_t* d_mask, int w, int h, int spp,
                        float radius, float* d_ao, void* stream) {
    if (!x) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    if (!begin(c)) return RAW_AO_ERR_DEVICE;
    if (!image_ok(w, h) || !d_pos || !d_nrm || !d_mask || !d_ao) return fail(c, RAW_AO_ERR_ARG, "bad image arguments");
    if (spp < 1 || spp > 4096) return fail(c, RAW_AO_ERR_ARG, "spp must be in 1..4096");
    if (!(radius > 0.0f)) return fail(c, RAW_AO_ERR_ARG, "radius must be > 0");
    if (c.ntri && c.root < 0) return fail(c, RAW_AO_ERR_ARG, "no BVH: call raw_ao_build_ploc first");
    if (c.host_trig) {
        const int r = raw_ao::trig_table_upload(c);
        if (r) return r;
    }
    if (c.ntri && c.trace_kernel != c.built_kernel)
        return fail(c, RAW_AO_ERR_ARG, "the trace kernel option changed after the build: build again");
    if (c.skew) radius *= 0.99f;   // accel.cpp:28, Mode::Skew
    if (c.trace_kernel == RAW_AO_TRACE_BINARY)
        return code_of(raw_ao::trace_ao(c, d_pos, d_nrm, d_mask, w, h, spp, radius, d_ao, pick(c, stream)));
    return code_of(raw_ao::trace_ao4(c, d_pos, d_nrm, d_mask, w, h, spp, radius, d_ao, pick(c, stream)));
}

int raw_ao_trace_ao(raw_ao_ctx* x, const float* pos, const float* nrm, const uint8_t* mask, int w, int h, int spp,
                    float radius, float* ao) {
    if (!x) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    if (!begin(c)) return RAW_AO_ERR_DEVICE;
    if (!image_ok(w, h) || !pos || !nrm || !mask || !ao) return fail(c, RAW_AO_ERR_ARG, "bad image arguments");
    const size_t np = size_t(w) * size_t(h);
    if (!ensure(c, c.st_pos, 12 * np) || !ensure(c, c.st_nrm, 12 * np) || !ensure(c, c.st_mask, np) ||
        !ensure(c, c.st_out, 4 * np))
        return RAW_AO_ERR_DEVICE;
    hipStream_t s = c.stream;
    if (!hip_ok(c, hipMemcpyAsync(c.st_pos.p, pos, 12 * np, hipMemcpyHostToDevice, s), "upload") ||
        !hip_ok(c, hipMemcpyAsync(c.st_nrm.p, nrm, 12 * np, hipMemcpyHostToDevice, s), "upload") ||
        !hip_ok(c, hipMemcpyAsync(c.st_mask.p, mask, np, hipMemcpyHostToDevice, s), "upload"))
        return RAW_AO_ERR_DEVICE;
    const int r = raw_ao_trace_ao_dev(x, c.st_pos.as<float>(), c.st_nrm.as<float>(), c.st_mask.as<uint8_t>(), w, h, spp,
                                      radius, c.st_out.as<float>(), nullptr);
    if (r) return r;
    if (!hip_ok(c, hipMemcpyAsync(ao, c.st_out.p, 4 * np, hipMemcpyDeviceToHost, s), "download") ||
        !hip_ok(c, hipStreamSynchronize(s), "trace"))
        return RAW_AO_ERR_DEVICE;
    c.ms_trace = elapsed(c, 2, 3);
    if (c.have_counts) {
        if (!hip_ok(c, hipMemcpy(&c.last_counts, c.counters.p, sizeof c.last_counts, hipMemcpyDeviceToHost), "counters"))
            return RAW_AO_ERR_DEVICE;
    }
    raw_ao::TraceCounters tc{};
    if (!hip_ok(c, hipMemcpy(&tc, c.counters.p, sizeof tc, hipMemcpyDeviceToHost), "counters")) return RAW_AO_ERR_DEVICE;
    c.last_counts.overflows = tc.overflows;
    return RAW_AO_OK;
}

int raw_ao_trace_visibility_dev(raw_ao_ctx* x, const float* d_pos, const float* d_nrm, const uint8_t* d_mask, int w, int h,
                                int spp, float radius, const float to_light[3], int sun_forced, float* d_ao, float* d_sky,
                                float* d_sun, void* stream) {
    if (!x) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    if (!begin(c)) return RAW_AO_ERR_DEVICE;
    if (!image_ok(w, h) || !d_pos || !d_nrm || !d_mask || !d_ao || !d_sky || !d_sun || !to_light)
        return fail(c, RAW_AO_ERR_ARG, "bad image arguments");
    if (spp < 1 || spp > 4096) return fail(c, RAW_AO_ERR_ARG, "spp must be in 1..4096");
    if (!(radius > 0.0f)) return fail(c, RAW_AO_ERR_ARG, "radius must be > 0");
    if (c.ntri && c.root < 0) return fail(c, RAW_AO_ERR_ARG, "no BVH: call raw_ao_build_ploc first");
    if (c.ntri && c.built_kernel != RAW_AO_TRACE_WIDE)
        return fail(c, RAW_AO_ERR_ARG, "the visibility pass needs the wide traversal form: build with RAW_AO_TRACE_WIDE");
    if (c.host_trig) {
        const int r = raw_ao::trig_table_upload(c);
        if (r) return r;
    }
    return code_of(raw_ao::trace_vis4(c, d_pos, d_nrm, d_mask, w, h, spp, radius, to_light, sun_forced, d_ao, d_sky, d_sun,
                                      pick(c, stream)));
}

int raw_ao_trace_visibility(raw_ao_ctx* x, const float* pos, const float* nrm, const uint8_t* mask, int w, int h, int spp,
                            float radius, const float to_light[3], int sun_forced, float* ao, float* sky, float* sun) {
    if (!x) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    if (!begin(c)) return RAW_AO_ERR_DEVICE;
    if (!image_ok(w, h) || !pos || !nrm || !mask || !ao || !sky || !sun) return fail(c, RAW_AO_ERR_ARG, "bad image arguments");
    const size_t np = size_t(w) * size_t(h);
    if (!ensure(c, c.st_pos, 12 * np) || !ensure(c, c.st_nrm, 12 * np) || !ensure(c, c.st_mask, np) ||
        !ensure(c, c.st_out, 12 * np))
        return RAW_AO_ERR_DEVICE;
    hipStream_t s = c.stream;
    if (!hip_ok(c, hipMemcpyAsync(c.st_pos.p, pos, 12 * np, hipMemcpyHostToDevice, s), "upload") ||
        !hip_ok(c, hipMemcpyAsync(c.st_nrm.p, nrm, 12 * np, hipMemcpyHostToDevice, s), "upload") ||
        !hip_ok(c, hipMemcpyAsync(c.st_mask.p, mask, np, hipMemcpyHostToDevice, s), "upload"))
        return RAW_AO_ERR_DEVICE;
    float* d = c.st_out.as<float>();
    const int r = raw_ao_trace_visibility_dev(x, c.st_pos.as<float>(), c.st_nrm.as<float>(), c.st_mask.as<uint8_t>(), w, h, spp,
                                              radius, to_light, sun_forced, d, d + np, d + 2 * np, nullptr);
    if (r) return r;
    if (!hip_ok(c, hipMemcpyAsync(ao, d, 4 * np, hipMemcpyDeviceToHost, s), "download") ||
        !hip_ok(c, hipMemcpyAsync(sky, d + np, 4 * np, hipMemcpyDeviceToHost, s), "download") ||
        !hip_ok(c, hipMemcpyAsync(sun, d + 2 * np, 4 * np, hipMemcpyDeviceToHost, s), "download") ||
        !hip_ok(c, hipStreamSynchronize(s), "visibility"))
        return RAW_AO_ERR_DEVICE;
    c.ms_trace = elapsed(c, 2, 3);
    return RAW_AO_OK;
}

namespace {
int raster_args(Ctx& c, const void* vpos, const void* vnrm, const void* idx, uint32_t ntri, const float* eye, const float* center,
                const float* up, float fovy, float aspect, float nearz, float farz, int w, int h) {
    if (!image_ok(w, h)) return fail(c, RAW_AO_ERR_ARG, "bad image size");
    if (ntri && (!vpos || !vnrm || !idx)) return fail(c, RAW_AO_ERR_ARG, "null mesh");
    if (ntri > (1u << 30)) return fail(c, RAW_AO_ERR_ARG, "at most 2^30 triangles");
    if (!eye || !center || !up) return fail(c, RAW_AO_ERR_ARG, "null camera");
    if (!(fovy > 0.0f && fovy < 3.14159f) || !(aspect > 0.0f) || !(nearz > 0.0f) || !(farz > nearz))
        return fail(c, RAW_AO_ERR_ARG, "bad camera: need 0 < fovy < pi, aspect > 0, 0 < near < far");
    return RAW_AO_OK;
}
}  // namespace

int raw_ao_rasterize_dev(raw_ao_ctx* x, const float* d_vpos, const float* d_vnrm, const int32_t* d_idx, uint32_t ntri,
                         const float eye[3], const float center[3], const float up[3], float fovy, float aspect, float near_z,
                         float far_z, int w, int h, float* d_depth, float* d_pos, float* d_nrm, uint8_t* d_mask, void* stream) {
    if (!x) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    if (!begin(c)) return RAW_AO_ERR_DEVICE;
    const int a = raster_args(c, d_vpos, d_vnrm, d_idx, ntri, eye, center, up, fovy, aspect, near_z, far_z, w, h);
    if (a) return a;
    if (!d_depth || !d_pos || !d_nrm || !d_mask) return fail(c, RAW_AO_ERR_ARG, "null output");
    return code_of(raw_ao::rasterize(c, d_vpos, d_vnrm, d_idx, ntri, eye, center, up, fovy, aspect, near_z, far_z, w, h, d_depth,
                                     d_pos, d_nrm, d_mask, pick(c, stream)));
}

int raw_ao_rasterize(raw_ao_ctx* x, const float* vpos, const float* vnrm, uint32_t nverts, const int32_t* idx, uint32_t ntri,
                     const float eye[3], const float center[3], const float up[3], float fovy, float aspect, float near_z,
                     float far_z, int w, int h, float* depth, float* pos, float* nrm, uint8_t* mask) {
    if (!x) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    if (!begin(c)) return RAW_AO_ERR_DEVICE;
    const int a = raster_args(c, vpos, vnrm, idx, ntri, eye, center, up, fovy, aspect, near_z, far_z, w, h);
    if (a) return a;
    if (!depth || !pos || !nrm || !mask) return fail(c, RAW_AO_ERR_ARG, "null output");
    for (size_t i = 0; i < 3 * size_t(ntri); ++i)
        if (idx[i] < 0 || uint32_t(idx[i]) >= nverts) return fail(c, RAW_AO_ERR_ARG, "vertex index out of range");
    const size_t np = size_t(w) * size_t(h);
    const size_t vb = 12 * size_t(nverts), ib = 12 * size_t(ntri);
    // staging: vertices (positions, normals) in st_rays, indices in st_cnt; outputs in st_out, st_mask
    if (!ensure(c, c.st_rays, 2 * vb) || !ensure(c, c.st_cnt, ib) || !ensure(c, c.st_out, 28 * np) || !ensure(c, c.st_mask, np))
        return RAW_AO_ERR_DEVICE;
    hipStream_t s = c.stream;
    char* v = c.st_rays.as<char>();
    if ((vb && (!hip_ok(c, hipMemcpyAsync(v, vpos, vb, hipMemcpyHostToDevice, s), "upload") ||
                !hip_ok(c, hipMemcpyAsync(v + vb, vnrm, vb, hipMemcpyHostToDevice, s), "upload"))) ||
        (ib && !hip_ok(c, hipMemcpyAsync(c.st_cnt.p, idx, ib, hipMemcpyHostToDevice, s), "upload")))
        return RAW_AO_ERR_DEVICE;
    float* o = c.st_out.as<float>();
    const int r = raw_ao_rasterize_dev(x, reinterpret_cast<const float*>(v), reinterpret_cast<const float*>(v + vb),
                                       c.st_cnt.as<int32_t>(), ntri, eye, center, up, fovy, aspect, near_z, far_z, w, h, o, o + np,
                                       o + 4 * np, c.st_mask.as<uint8_t>(), nullptr);
    if (r) return r;
    if (!hip_ok(c, hipMemcpyAsync(depth, o, 4 * np, hipMemcpyDeviceToHost, s), "download") ||
        !hip_ok(c, hipMemcpyAsync(pos, o + np, 12 * np, hipMemcpyDeviceToHost, s), "download") ||
        !hip_ok(c, hipMemcpyAsync(nrm, o + 4 * np, 12 * np, hipMemcpyDeviceToHost, s), "download") ||
        !hip_ok(c, hipMemcpyAsync(mask, c.st_mask.p, np, hipMemcpyDeviceToHost, s), "download") ||
        !hip_ok(c, hipStreamSynchronize(s), "rasterize"))
        return RAW_AO_ERR_DEVICE;
    return RAW_AO_OK;
}

int raw_ao_adapter_info(raw_ao_ctx* x, raw_ao_adapter* out) {
    if (!x || !out) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    if (!begin(c)) return RAW_AO_ERR_DEVICE;
    hipDeviceProp_t p;
    if (!hip_ok(c, hipGetDeviceProperties(&p, c.device), "hipGetDeviceProperties")) return RAW_AO_ERR_DEVICE;
    int rt = 0, drv = 0;
    (void)hipRuntimeGetVersion(&rt);
    (void)hipDriverGetVersion(&drv);
    std::memset(out, 0, sizeof *out);
    std::snprintf(out->vendor, sizeof out->vendor, "AMD");
    std::snprintf(out->architecture, sizeof out->architecture, "%s", p.gcnArchName);
    std::snprintf(out->device, sizeof out->device, "%s", p.name);
    std::snprintf(out->description, sizeof out->description, "%d compute units, %zu MiB", p.multiProcessorCount,
                  size_t(p.totalGlobalMem >> 20));
    std::snprintf(out->backend, sizeof out->backend, "hip");
    std::snprintf(out->driver, sizeof out->driver, "HIP runtime %d, driver %d", rt, drv);
    return RAW_AO_OK;
}

void raw_ao_render_opts_default(raw_ao_render_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->tolerance = 0.12f;
    o->rtao = 1;
    o->lighting = 2;
    o->rt_samples = 64;
    o->ss_samples = 24;
    o->radius = 2.0f;
    o->bvh_radius = 16;
}

namespace {
int frame_args(Ctx& c, const raw_ao_scene* sc, const raw_ao_render_opts& o, int w, int h) {
    if (!sc) return fail(c, RAW_AO_ERR_ARG, "null scene");
    const int a = raster_args(c, sc->vpos, sc->vnrm, sc->idx, sc->ntri, sc->eye, sc->center, sc->up, sc->fovy, sc->aspect,
                              sc->near_z, sc->far_z, w, h);
    if (a) return a;
    if (sc->ntri > RAW_AO_MAX_TRIANGLES) return fail(c, RAW_AO_ERR_ARG, "more than RAW_AO_MAX_TRIANGLES triangles");
    if (sc->nmesh < 1 || !sc->mesh_end || !sc->albedo) return fail(c, RAW_AO_ERR_ARG, "a scene needs at least one mesh");
    for (uint32_t m = 0; m < sc->nmesh; ++m)
        if ((m && sc->mesh_end[m] < sc->mesh_end[m - 1]) || sc->mesh_end[m] > sc->ntri)
            return fail(c, RAW_AO_ERR_ARG, "mesh ends must be non-decreasing and at most ntri");
    if (sc->mesh_end[sc->nmesh - 1] != sc->ntri) return fail(c, RAW_AO_ERR_ARG, "the last mesh must end at ntri");
    if (o.lighting != 1 && o.lighting != 2) return fail(c, RAW_AO_ERR_ARG, "lighting must be 1 or 2");
    if (o.rt_samples < 1 || o.rt_samples > 4096) return fail(c, RAW_AO_ERR_ARG, "rt_samples must be in 1..4096");
    if (o.ss_samples < 0 || o.ss_samples > 4096) return fail(c, RAW_AO_ERR_ARG, "ss_samples must be in 0..4096");
    if (!(o.radius > 0.0f)) return fail(c, RAW_AO_ERR_ARG, "radius must be > 0");
    if (o.bvh_radius < 1 || o.bvh_radius > 32) return fail(c, RAW_AO_ERR_ARG, "PLOC radius must be in 1..32");
    if (o.rtao && o.lighting == 2 && c.trace_kernel == RAW_AO_TRACE_BINARY)
        return fail(c, RAW_AO_ERR_ARG, "lighting model 2 needs the wide traversal form (RAW_AO_TRACE_WIDE)");
    return RAW_AO_OK;
}
}  // namespace

int raw_ao_render_frame_dev(raw_ao_ctx* x, const raw_ao_scene* sc, const raw_ao_render_opts* opts, int w, int h,
                            raw_ao_frame* d, raw_ao_frame_stats* stats, void* stream) {
    if (!x) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    if (!begin(c)) return RAW_AO_ERR_DEVICE;
    raw_ao_render_opts o;
    raw_ao_render_opts_default(&o);
    if (opts) o = *opts;
    const int a = frame_args(c, sc, o, w, h);
    if (a) return a;
    if (!d || !d->depth || !d->position || !d->normal || !d->albedo || !d->mask || !d->motion || !d->ao_rt || !d->ao_ss ||
        !d->error || !d->frame || !d->hdr || (o.rtao && o.lighting == 2 && (!d->sky || !d->sun)))
        return fail(c, RAW_AO_ERR_ARG, "null output plane");
    if (c.host_trig) {
        const int r = raw_ao::trig_table_upload(c);
        if (r) return r;
    }
    return code_of(raw_ao::render_frame(c, *sc, o, w, h, *d, stats, pick(c, stream)));
}

int raw_ao_render_frame(raw_ao_ctx* x, const raw_ao_scene* sc, const raw_ao_render_opts* opts, int w, int h, raw_ao_frame* out,
                        raw_ao_frame_stats* stats) {
    if (!x) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    if (!begin(c)) return RAW_AO_ERR_DEVICE;
    raw_ao_render_opts o;
    raw_ao_render_opts_default(&o);
    if (opts) o = *opts;
    const int a = frame_args(c, sc, o, w, h);
    if (a) return a;
    if (!out) return fail(c, RAW_AO_ERR_ARG, "null output");
    for (size_t i = 0; i < 3 * size_t(sc->ntri); ++i)
        if (sc->idx[i] < 0 || uint32_t(sc->idx[i]) >= sc->nverts) return fail(c, RAW_AO_ERR_ARG, "vertex index out of range");
    const size_t np = size_t(w) * size_t(h);
    const size_t vb = 12 * size_t(sc->nverts), ib = 12 * size_t(sc->ntri);
    // planes in one allocation: 4 B depth, 12 position, 12 normal, 12 albedo, 8 motion, 4 each of
    // ao_rt, ao_ss, sky, sun, error, 12 frame, 12 hdr, then 1 B mask
    const size_t floats_px = 1 + 3 + 3 + 3 + 2 + 1 + 1 + 1 + 1 + 1 + 3 + 3;
    if (!ensure(c, c.st_rays, 2 * vb) || !ensure(c, c.st_cnt, ib) || !ensure(c, c.fr_planes, 4 * floats_px * np + np))
        return RAW_AO_ERR_DEVICE;
    hipStream_t s = c.stream;
    char* v = c.st_rays.as<char>();
    if ((vb && (!hip_ok(c, hipMemcpyAsync(v, sc->vpos, vb, hipMemcpyHostToDevice, s), "upload") ||
                !hip_ok(c, hipMemcpyAsync(v + vb, sc->vnrm, vb, hipMemcpyHostToDevice, s), "upload"))) ||
        (ib && !hip_ok(c, hipMemcpyAsync(c.st_cnt.p, sc->idx, ib, hipMemcpyHostToDevice, s), "upload")))
        return RAW_AO_ERR_DEVICE;
    float* p = c.fr_planes.as<float>();
    raw_ao_frame d;
    d.depth = p; p += np;
    d.position = p; p += 3 * np;
    d.normal = p; p += 3 * np;
    d.albedo = p; p += 3 * np;
    d.motion = p; p += 2 * np;
    d.ao_rt = p; p += np;
    d.ao_ss = p; p += np;
    d.sky = p; p += np;
    d.sun = p; p += np;
    d.error = p; p += np;
    d.frame = p; p += 3 * np;
    d.hdr = p; p += 3 * np;
    d.mask = reinterpret_cast<uint8_t*>(p);
    raw_ao_scene dsc = *sc;
    dsc.vpos = reinterpret_cast<const float*>(v);
    dsc.vnrm = reinterpret_cast<const float*>(v + vb);
    dsc.idx = c.st_cnt.as<int32_t>();
    const int r = raw_ao_render_frame_dev(x, &dsc, &o, w, h, &d, stats, nullptr);
    if (r) return r;
    const bool model2 = o.rtao && o.lighting == 2;
    struct Plane { void* host; const void* dev; size_t bytes; bool ok; } planes[] = {
        {out->depth, d.depth, 4, true}, {out->position, d.position, 12, true}, {out->normal, d.normal, 12, true},
        {out->albedo, d.albedo, 12, true}, {out->mask, d.mask, 1, true}, {out->motion, d.motion, 8, true},
        {out->ao_rt, d.ao_rt, 4, o.rtao != 0}, {out->ao_ss, d.ao_ss, 4, true}, {out->sky, d.sky, 4, model2},
        {out->sun, d.sun, 4, model2}, {out->error, d.error, 4, true}, {out->frame, d.frame, 12, true},
        {out->hdr, d.hdr, 12, true}};
    for (const Plane& q : planes)
        if (q.host && q.ok && !hip_ok(c, hipMemcpyAsync(q.host, q.dev, q.bytes * np, hipMemcpyDeviceToHost, s), "download"))
            return RAW_AO_ERR_DEVICE;
    if (!hip_ok(c, hipStreamSynchronize(s), "frame download")) return RAW_AO_ERR_DEVICE;
    return RAW_AO_OK;
}

int raw_ao_ssao_dev(raw_ao_ctx* x, const float* d_pos, const float* d_nrm, const uint8_t* d_mask, int w, int h, int samples,
                    float radius, float* d_ao, void* stream) {
    if (!x) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    if (!begin(c)) return RAW_AO_ERR_DEVICE;
    if (!image_ok(w, h) || !d_pos || !d_nrm || !d_mask || !d_ao) return fail(c, RAW_AO_ERR_ARG, "bad image arguments");
    if (samples < 0 || samples > 4096) return fail(c, RAW_AO_ERR_ARG, "samples must be in 0..4096");
    if (c.host_trig) {
        const int r = raw_ao::trig_table_upload(c);
        if (r) return r;
    }
    return code_of(raw_ao::ssao(c, d_pos, d_nrm, d_mask, w, h, samples, radius, d_ao, pick(c, stream)));
}

int raw_ao_ssao(raw_ao_ctx* x, const float* pos, const float* nrm, const uint8_t* mask, int w, int h, int samples, float radius,
                float* ao) {
    if (!x) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    if (!begin(c)) return RAW_AO_ERR_DEVICE;
    if (!image_ok(w, h) || !pos || !nrm || !mask || !ao) return fail(c, RAW_AO_ERR_ARG, "bad image arguments");
    const size_t np = size_t(w) * size_t(h);
    if (!ensure(c, c.st_pos, 12 * np) || !ensure(c, c.st_nrm, 12 * np) || !ensure(c, c.st_mask, np) ||
        !ensure(c, c.st_out, 4 * np))
        return RAW_AO_ERR_DEVICE;
    hipStream_t s = c.stream;
    if (!hip_ok(c, hipMemcpyAsync(c.st_pos.p, pos, 12 * np, hipMemcpyHostToDevice, s), "upload") ||
        !hip_ok(c, hipMemcpyAsync(c.st_nrm.p, nrm, 12 * np, hipMemcpyHostToDevice, s), "upload") ||
        !hip_ok(c, hipMemcpyAsync(c.st_mask.p, mask, np, hipMemcpyHostToDevice, s), "upload"))
        return RAW_AO_ERR_DEVICE;
    const int r = raw_ao_ssao_dev(x, c.st_pos.as<float>(), c.st_nrm.as<float>(), c.st_mask.as<uint8_t>(), w, h, samples, radius,
                                  c.st_out.as<float>(), nullptr);
    if (r) return r;
    if (!hip_ok(c, hipMemcpyAsync(ao, c.st_out.p, 4 * np, hipMemcpyDeviceToHost, s), "download") ||
        !hip_ok(c, hipStreamSynchronize(s), "ssao"))
        return RAW_AO_ERR_DEVICE;
    c.ms_ssao = elapsed(c, 4, 5);
    return RAW_AO_OK;
}

int raw_ao_error_map_dev(raw_ao_ctx* x, const float* d_a, const float* d_t, const uint8_t* d_mask, int w, int h, float* d_err,
                         float* d_max, void* stream) {
    if (!x) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    if (!begin(c)) return RAW_AO_ERR_DEVICE;
    if (!image_ok(w, h) || !d_a || !d_t || !d_mask || !d_err || !d_max) return fail(c, RAW_AO_ERR_ARG, "bad image arguments");
    return code_of(raw_ao::error_map(c, d_a, d_t, d_mask, w, h, d_err, d_max, pick(c, stream)));
}

int raw_ao_error_map(raw_ao_ctx* x, const float* a, const float* t, const uint8_t* mask, int w, int h, float* err,
                     float* max_error) {
    if (!x) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    if (!begin(c)) return RAW_AO_ERR_DEVICE;
    if (!image_ok(w, h) || !a || !t || !mask || !err || !max_error) return fail(c, RAW_AO_ERR_ARG, "bad image arguments");
    const size_t np = size_t(w) * size_t(h);
    if (!ensure(c, c.st_pos, 4 * np) || !ensure(c, c.st_nrm, 4 * np) || !ensure(c, c.st_mask, np) ||
        !ensure(c, c.st_out, 4 * np) || !ensure(c, c.st_max, 16))
        return RAW_AO_ERR_DEVICE;
    hipStream_t s = c.stream;
    if (!hip_ok(c, hipMemcpyAsync(c.st_pos.p, a, 4 * np, hipMemcpyHostToDevice, s), "upload") ||
        !hip_ok(c, hipMemcpyAsync(c.st_nrm.p, t, 4 * np, hipMemcpyHostToDevice, s), "upload") ||
        !hip_ok(c, hipMemcpyAsync(c.st_mask.p, mask, np, hipMemcpyHostToDevice, s), "upload"))
        return RAW_AO_ERR_DEVICE;
    const int r = raw_ao_error_map_dev(x, c.st_pos.as<float>(), c.st_nrm.as<float>(), c.st_mask.as<uint8_t>(), w, h,
                                       c.st_out.as<float>(), c.st_max.as<float>(), nullptr);
    if (r) return r;
    if (!hip_ok(c, hipMemcpyAsync(err, c.st_out.p, 4 * np, hipMemcpyDeviceToHost, s), "download") ||
        !hip_ok(c, hipMemcpyAsync(max_error, c.st_max.p, 4, hipMemcpyDeviceToHost, s), "download") ||
        !hip_ok(c, hipStreamSynchronize(s), "error map"))
        return RAW_AO_ERR_DEVICE;
    c.ms_error = elapsed(c, 6, 7);
    return RAW_AO_OK;
}

int raw_ao_trace_rays(raw_ao_ctx* x, const float* rays8, uint32_t nrays, int any, float far_scale, float* out4) {
    if (!x) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    if (!begin(c)) return RAW_AO_ERR_DEVICE;
    if (nrays && (!rays8 || !out4)) return fail(c, RAW_AO_ERR_ARG, "null rays");
    if (c.ntri && c.root < 0) return fail(c, RAW_AO_ERR_ARG, "no BVH: call raw_ao_build_ploc first");
    if (!(far_scale > 0.0f)) return fail(c, RAW_AO_ERR_ARG, "far_scale must be > 0");
    if (!nrays) return RAW_AO_OK;
    if (!ensure(c, c.st_rays, 32 * size_t(nrays)) || !ensure(c, c.st_out2, 16 * size_t(nrays))) return RAW_AO_ERR_DEVICE;
    hipStream_t s = c.stream;
    if (!hip_ok(c, hipMemcpyAsync(c.st_rays.p, rays8, 32 * size_t(nrays), hipMemcpyHostToDevice, s), "upload"))
        return RAW_AO_ERR_DEVICE;
    if (c.ntri && c.trace_kernel != c.built_kernel)
        return fail(c, RAW_AO_ERR_ARG, "the trace kernel option changed after the build: build again");
    const int r = c.trace_kernel == RAW_AO_TRACE_BINARY
                      ? raw_ao::trace_rays(c, c.st_rays.as<float>(), nrays, any, far_scale, c.st_out2.as<float>(), s)
                      : raw_ao::trace_rays4(c, c.st_rays.as<float>(), nrays, any, far_scale, c.st_out2.as<float>(), s);
    if (r) return code_of(r);
    if (!hip_ok(c, hipMemcpyAsync(out4, c.st_out2.p, 16 * size_t(nrays), hipMemcpyDeviceToHost, s), "download") ||
        !hip_ok(c, hipStreamSynchronize(s), "trace_rays"))
        return RAW_AO_ERR_DEVICE;
    return RAW_AO_OK;
}

int raw_ao_closest_hits_dev(raw_ao_ctx* x, const float* d_rays8, uint32_t nrays, void* d_hits4, void* stream) {
    if (!x) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    if (!begin(c)) return RAW_AO_ERR_DEVICE;
    if (nrays && (!d_rays8 || !d_hits4)) return fail(c, RAW_AO_ERR_ARG, "null rays");
    if (c.ntri && c.root < 0) return fail(c, RAW_AO_ERR_ARG, "no BVH: call raw_ao_build_ploc first");
    if (c.ntri && c.built_kernel != RAW_AO_TRACE_WIDE)
        return fail(c, RAW_AO_ERR_ARG, "closest hits need the wide traversal form: build with RAW_AO_TRACE_WIDE");
    return code_of(raw_ao::closest_hits4(c, d_rays8, nrays, static_cast<float*>(d_hits4), pick(c, stream)));
}

int raw_ao_closest_hits(raw_ao_ctx* x, const float* rays8, uint32_t nrays, void* hits4) {
    if (!x) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    if (!begin(c)) return RAW_AO_ERR_DEVICE;
    if (nrays && (!rays8 || !hits4)) return fail(c, RAW_AO_ERR_ARG, "null rays");
    if (!nrays) return RAW_AO_OK;
    if (!ensure(c, c.st_rays, 32 * size_t(nrays)) || !ensure(c, c.st_out2, 16 * size_t(nrays))) return RAW_AO_ERR_DEVICE;
    hipStream_t s = c.stream;
    if (!hip_ok(c, hipMemcpyAsync(c.st_rays.p, rays8, 32 * size_t(nrays), hipMemcpyHostToDevice, s), "upload"))
        return RAW_AO_ERR_DEVICE;
    const int r = raw_ao_closest_hits_dev(x, c.st_rays.as<float>(), nrays, c.st_out2.p, nullptr);
    if (r) return r;
    if (!hip_ok(c, hipMemcpyAsync(hits4, c.st_out2.p, 16 * size_t(nrays), hipMemcpyDeviceToHost, s), "download") ||
        !hip_ok(c, hipStreamSynchronize(s), "closest_hits"))
        return RAW_AO_ERR_DEVICE;
    return RAW_AO_OK;
}

int raw_ao_count_next_trace(raw_ao_ctx* x, int enable) {
    if (!x) return RAW_AO_ERR_ARG;
    x->c.err.clear();
    x->c.count_next = enable != 0;
    return RAW_AO_OK;
}

int raw_ao_trace_counts(raw_ao_ctx* x, uint64_t* visits, uint64_t* tests, uint64_t* rays) {
    if (!x) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    c.err.clear();
    if (!c.have_counts) return fail(c, RAW_AO_ERR_ARG, "no counting trace yet (raw_ao_count_next_trace)");
    if (!hip_ok(c, hipSetDevice(c.device), "hipSetDevice") ||
        !hip_ok(c, hipMemcpy(&c.last_counts, c.counters.p, sizeof c.last_counts, hipMemcpyDeviceToHost), "counters"))
        return RAW_AO_ERR_DEVICE;
    if (visits) *visits = c.last_counts.visits;
    if (tests) *tests = c.last_counts.tests;
    if (rays) *rays = c.last_counts.rays;
    return RAW_AO_OK;
}

int raw_ao_last_ms(raw_ao_ctx* x, float* build_ms, float* trace_ms, float* ssao_ms, float* error_ms) {
    if (!x) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    if (!begin(c)) return RAW_AO_ERR_DEVICE;
    if (!hip_ok(c, hipStreamSynchronize(c.stream), "sync")) return RAW_AO_ERR_DEVICE;
    if (build_ms) *build_ms = c.ms_build;
    if (trace_ms) *trace_ms = elapsed(c, 2, 3);
    if (ssao_ms) *ssao_ms = elapsed(c, 4, 5);
    if (error_ms) *error_ms = elapsed(c, 6, 7);
    return RAW_AO_OK;
}

int raw_ao_angle_table(raw_ao_ctx* x, float* sin_out, float* cos_out) {
    if (!x) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    if (!begin(c)) return RAW_AO_ERR_DEVICE;
    if (!sin_out || !cos_out) return fail(c, RAW_AO_ERR_ARG, "null output");
    if (c.host_trig) {
        const int r = raw_ao::trig_table_upload(c);
        if (r) return r;
    }
    const size_t bytes = 4u << 24;
    if (!ensure(c, c.st_out, bytes) || !ensure(c, c.st_out2, bytes)) return RAW_AO_ERR_DEVICE;
    const int r = raw_ao::angle_table(c, c.st_out.as<float>(), c.st_out2.as<float>(), c.stream);
    if (r) return code_of(r);
    if (!hip_ok(c, hipMemcpyAsync(sin_out, c.st_out.p, bytes, hipMemcpyDeviceToHost, c.stream), "download") ||
        !hip_ok(c, hipMemcpyAsync(cos_out, c.st_out2.p, bytes, hipMemcpyDeviceToHost, c.stream), "download") ||
        !hip_ok(c, hipStreamSynchronize(c.stream), "angles"))
        return RAW_AO_ERR_DEVICE;
    return RAW_AO_OK;
}

int raw_ao_bvh_bytes(raw_ao_ctx* x, uint64_t* node_bytes, uint64_t* tri_bytes, uint32_t* node_record_bytes) {
    if (!x) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    c.err.clear();
    const uint64_t n = c.ntri;
    const bool wide = c.built_kernel == RAW_AO_TRACE_WIDE;
    if (node_bytes) *node_bytes = !n ? 0 : (wide ? 128ull * c.wide_nodes : 64ull * (n > 1 ? n - 1 : 1));
    if (tri_bytes) *tri_bytes = 48ull * n;
    if (node_record_bytes) *node_record_bytes = wide ? 128 : 64;
    return RAW_AO_OK;
}

int raw_ao_read_wide_records(raw_ao_ctx* x, void* records128, void* records64, uint32_t cap, uint32_t* count) {
    if (!x) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    c.err.clear();
    if (c.ntri == 0 || c.root < 0 || c.built_kernel != RAW_AO_TRACE_WIDE) return fail(c, RAW_AO_ERR_ARG, "no wide tree");
    if (records64 && !c.q_built)
        return fail(c, RAW_AO_ERR_ARG, "no quantized records: set RAW_AO_OPT_QUANTIZED_NODES and build or trace first");
    if (!hip_ok(c, hipSetDevice(c.device), "hipSetDevice") || !hip_ok(c, hipStreamSynchronize(c.stream), "sync")) return RAW_AO_ERR_DEVICE;
    const uint32_t n = std::min(cap, c.wide_nodes);
    if (count) *count = c.wide_nodes;
    if (records128 && n && !hip_ok(c, hipMemcpy(records128, c.nodes4.p, 128 * size_t(n), hipMemcpyDeviceToHost), "records"))
        return RAW_AO_ERR_DEVICE;
    if (records64 && n && !hip_ok(c, hipMemcpy(records64, c.nodes4q.p, 64 * size_t(n), hipMemcpyDeviceToHost), "records"))
        return RAW_AO_ERR_DEVICE;
    return RAW_AO_OK;
}

const raw_ao_hip_api_v1* raw_ao_hip_api(uint32_t abi_version) {
    static const raw_ao_hip_api_v1 v1 = {
        1u, uint32_t(sizeof(raw_ao_hip_api_v1)), raw_ao_open, raw_ao_close, raw_ao_last_error, raw_ao_set_option,
        raw_ao_build_ploc, raw_ao_build_ploc_dev, raw_ao_tree_info, raw_ao_read_tree, raw_ao_read_keys, raw_ao_trace_ao,
        raw_ao_trace_ao_dev, raw_ao_trace_visibility, raw_ao_ssao, raw_ao_ssao_dev, raw_ao_error_map, raw_ao_error_map_dev,
        raw_ao_trace_rays, raw_ao_closest_hits, raw_ao_render_frame};
    return abi_version == 1u ? &v1 : nullptr;
}

}  // extern "C"
