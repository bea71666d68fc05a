// Batched AO over many views of several scenes: the host side of one persistent traversal
// launch with a global queue (kernels in bvh4.hip), the per-view reconcile on the device,
// hipGraph capture of build + trace + reconcile, and the tuner of the launch parameters.
#include "../../include/raw_ao_hip.h"
#include "batch.hpp"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <new>
#include <sstream>
#include <string>
#include <vector>

namespace raw_ao {
namespace {

constexpr int RB = 256;   // reconcile block

struct ViewStats {
    double sumsq;
    uint32_t maxbits;
    uint32_t count;
};

__global__ void k_stats_clear(ViewStats* st, uint32_t n) {
    for (uint32_t v = threadIdx.x; v < n; v += blockDim.x) st[v] = ViewStats{0.0, 0u, 0u};
}
// The reconcile's per-pixel part of view blockIdx.y: err = |ssao - ao| where covered, 0
// elsewhere; the view's largest error; per-block partial sums of squares (float64) and
// counts, which k_error_reduce adds in block order.
__global__ __launch_bounds__(RB) void k_batch_error(const BatchView* __restrict__ views, double* __restrict__ part,
                                                    uint32_t* __restrict__ pcount, uint32_t nparts, ViewStats* st) {
    __shared__ double ss[RB];
    __shared__ uint32_t sc[RB];
    const uint32_t v = blockIdx.y;
    const BatchView& bv = views[v];
    double s = 0.0;
    uint32_t cnt = 0;
    float mx = 0.0f;
    if (bv.err)
        for (uint32_t i = blockIdx.x * RB + threadIdx.x; i < bv.npix; i += gridDim.x * RB) {
            const float e = bv.mask[i] ? fabsf(bv.ssao[i] - bv.ao[i]) : 0.0f;
            bv.err[i] = e;
            if (bv.mask[i]) {
                s += double(e) * double(e);
                ++cnt;
                if (e > mx) mx = e;
            }
        }
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
    if ((threadIdx.x & 63) == 0 && mx > 0.0f) atomicMax(&st[v].maxbits, __float_as_uint(mx));
    ss[threadIdx.x] = s;
    sc[threadIdx.x] = cnt;
    __syncthreads();
    for (int k = RB / 2; k > 0; k >>= 1) {
        if (int(threadIdx.x) < k) {
            ss[threadIdx.x] += ss[threadIdx.x + k];
            sc[threadIdx.x] += sc[threadIdx.x + k];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[size_t(v) * nparts + blockIdx.x] = ss[0];
        pcount[size_t(v) * nparts + blockIdx.x] = sc[0];
    }
}
__global__ void k_error_reduce(const double* __restrict__ part, const uint32_t* __restrict__ pcount, uint32_t nparts,
                               uint32_t nviews, ViewStats* st) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nviews) return;
    double s = 0.0;
    uint32_t c = 0;
    for (uint32_t b = 0; b < nparts; ++b) {
        s += part[size_t(v) * nparts + b];
        c += pcount[size_t(v) * nparts + b];
    }
    st[v].sumsq = s;
    st[v].count = c;
}

constexpr uint32_t kErrParts = 64;

}  // namespace

struct Batch {
    raw_ao_ctx* self{nullptr};          // stream, trig table, error text
    std::vector<raw_ao_ctx*> scenes;
    std::vector<const float*> scene_tris;
    std::vector<BatchView> views;
    DevBuf d_views, d_scenes, cov, cnt, ncov, gbase, head, stats, part, pcount;
    BatchTune tune;
    uint32_t total_pix{0};
    uint64_t max_groups{0};   // the views' 64-ray groups if every pixel were covered
    bool dirty{true};
    int cus{256};
    hipGraph_t graph{nullptr};
    hipGraphExec_t exec{nullptr};
    bool graph_has_builds{false};
};

namespace {

Ctx& C(Batch& b) { return b.self->c; }

void drop_graph(Batch& b) {
    if (b.exec) (void)hipGraphExecDestroy(b.exec);
    if (b.graph) (void)hipGraphDestroy(b.graph);
    b.exec = nullptr;
    b.graph = nullptr;
}

int prepare(Batch& b) {
    Ctx& c = C(b);
    if (b.views.empty()) {
        c.err = "the batch has no views";
        return -1;
    }
    if (c.host_trig) {
        const int r = trig_table_upload(c);
        if (r) return r;
    }
    if (!b.dirty) return 0;
    std::vector<BatchScene> sc;
    for (auto* s : b.scenes) {
        const Ctx& x = s->c;
        const uint32_t* ntop = x.top.p ? x.top.as<uint32_t>() + top_count_offset() : nullptr;
        sc.push_back(BatchScene{x.nodes4.p, x.tris_sorted.as<float4>(), x.top.p, ntop, x.ntri, 0});
    }
    const size_t nv = b.views.size();
    if (!ensure(c, b.d_views, sizeof(BatchView) * nv) || !ensure(c, b.d_scenes, sizeof(BatchScene) * std::max<size_t>(1, sc.size())) ||
        !ensure(c, b.cov, 4 * size_t(b.total_pix)) || !ensure(c, b.cnt, 4 * size_t(b.total_pix)) || !ensure(c, b.ncov, 4 * nv) ||
        !ensure(c, b.gbase, 4 * (nv + 1)) || !ensure(c, b.head, 16) || !ensure(c, b.stats, sizeof(ViewStats) * nv) ||
        !ensure(c, b.part, 8 * kErrParts * nv) || !ensure(c, b.pcount, 4 * kErrParts * nv))
        return -2;
    if (!hip_ok(c, hipMemcpy(b.d_views.p, b.views.data(), sizeof(BatchView) * nv, hipMemcpyHostToDevice), "view table") ||
        (!sc.empty() && !hip_ok(c, hipMemcpy(b.d_scenes.p, sc.data(), sizeof(BatchScene) * sc.size(), hipMemcpyHostToDevice),
                                "scene table")))
        return -2;
    b.dirty = false;
    return 0;
}

BatchDev dev_of(Batch& b) {
    Ctx& c = C(b);
    return BatchDev{b.d_views.as<BatchView>(), b.d_scenes.as<BatchScene>(), uint32_t(b.views.size()), b.total_pix,
                    b.cov.as<uint32_t>(), b.cnt.as<uint32_t>(), b.ncov.as<uint32_t>(), b.gbase.as<uint32_t>(),
                    b.head.as<uint32_t>(), c.host_trig ? c.d_trig : nullptr, b.tune.tri_q, b.tune.chunk,
                    float(c.wave_cut), uint32_t(std::max(0, c.refill))};
}

// Everything of one batched frame on stream s, with no host copies, waits or memset nodes.
int enqueue(Batch& b, bool builds, hipStream_t s) {
    Ctx& c = C(b);
    if (builds)
        for (size_t i = 0; i < b.scenes.size(); ++i) {
            Ctx& x = b.scenes[i]->c;
            // from the caller's triangles (a replay sees their current values)
            if (x.ntri && !hip_ok(c, hipMemcpyAsync(x.tris_in.p, b.scene_tris[i], 36 * size_t(x.ntri), hipMemcpyDeviceToDevice, s),
                                  "triangle copy"))
                return -2;
            int r = ploc_build_fixed(x, x.tris_in.as<float>(), s);
            if (!r) r = bvh_layout(x, s);
            if (!r) r = wide_build_fixed(x, s);
            if (r) {
                c.err = x.err;
                return r;
            }
        }
    const uint32_t nv = uint32_t(b.views.size());
    // SSAO of every view (the reconcile's approximation), then the rays
    const uint8_t* trig = c.host_trig ? c.d_trig : nullptr;
    for (const auto& v : b.views)
        if (v.ssao) ssao_enqueue(trig, v.pos, v.nrm, v.mask, int(v.w), int(v.h), int(v.ssao_samples), v.ssao_radius, v.ssao, s);
    const uint32_t blocks = uint32_t(b.cus) * b.tune.blocks_per_cu;
    if (batch_trace_enqueue(dev_of(b), blocks, s)) return -2;
    k_stats_clear<<<1, 256, 0, s>>>(b.stats.as<ViewStats>(), nv);
    k_batch_error<<<dim3(kErrParts, nv, 1), RB, 0, s>>>(b.d_views.as<BatchView>(), b.part.as<double>(), b.pcount.as<uint32_t>(),
                                                        kErrParts, b.stats.as<ViewStats>());
    k_error_reduce<<<(nv + 63) / 64, 64, 0, s>>>(b.part.as<double>(), b.pcount.as<uint32_t>(), kErrParts, nv,
                                                  b.stats.as<ViewStats>());
    return hip_ok(c, hipGetLastError(), "batch launch") ? 0 : -2;
}

float ms_of(const std::chrono::steady_clock::time_point& t0) {
    return std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

int code_of(int internal) {
    if (internal == 0) return RAW_AO_OK;
    if (internal == -1) return RAW_AO_ERR_ARG;
    if (internal == -3) return RAW_AO_ERR_BUILD;
    return RAW_AO_ERR_DEVICE;
}

}  // namespace
}  // namespace raw_ao

struct raw_ao_batch {
    raw_ao::Batch b;
};

using raw_ao::Batch;
using raw_ao::hip_ok;

namespace {
int bfail(Batch& b, int code, const std::string& msg) {
    if (b.self->c.err.empty()) b.self->c.err = msg;
    return code;
}
bool bbegin(Batch& b) {
    b.self->c.err.clear();
    return hip_ok(b.self->c, hipSetDevice(b.self->c.device), "hipSetDevice");
}
hipStream_t bpick(Batch& b, void* s) { return s ? static_cast<hipStream_t>(s) : b.self->c.stream; }
}  // namespace

extern "C" {

int raw_ao_batch_open(int device, raw_ao_batch** out) {
    if (!out) return RAW_AO_ERR_ARG;
    *out = nullptr;
    raw_ao_batch* x = new (std::nothrow) raw_ao_batch;
    if (!x) return RAW_AO_ERR_ARG;
    const int rc = raw_ao_open(device, &x->b.self);
    *out = x;
    if (rc != RAW_AO_OK) return rc;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0) x->b.cus = prop.multiProcessorCount;
    return RAW_AO_OK;
}

void raw_ao_batch_close(raw_ao_batch* x) {
    if (!x) return;
    Batch& b = x->b;
    if (b.self) {
        (void)hipSetDevice(b.self->c.device);
        (void)hipDeviceSynchronize();
        raw_ao::drop_graph(b);
        raw_ao::DevBuf* bufs[] = {&b.d_views, &b.d_scenes, &b.cov, &b.cnt, &b.ncov, &b.gbase, &b.head, &b.stats, &b.part, &b.pcount};
        for (auto* p : bufs)
            if (p->p) (void)hipFree(p->p);
    }
    for (auto* s : b.scenes) raw_ao_close(s);
    raw_ao_close(b.self);
    delete x;
}

const char* raw_ao_batch_last_error(const raw_ao_batch* x) {
    return x && x->b.self ? x->b.self->c.err.c_str() : "null batch";
}

int raw_ao_batch_add_scene(raw_ao_batch* x, const float* d_tris9, uint32_t ntri, int radius, int* scene_id) {
    if (!x || !x->b.self) return RAW_AO_ERR_ARG;
    Batch& b = x->b;
    if (!bbegin(b)) return RAW_AO_ERR_DEVICE;
    if ((ntri && !d_tris9) || ntri > RAW_AO_MAX_TRIANGLES) return bfail(b, RAW_AO_ERR_ARG, "bad scene arguments");
    raw_ao_ctx* s = nullptr;
    int rc = raw_ao_open(b.self->c.device, &s);
    if (rc != RAW_AO_OK) {
        raw_ao_close(s);
        return bfail(b, rc, "scene context");
    }
    s->c.host_trig = b.self->c.host_trig;
    // the eager build records the grids the capturable build replays
    rc = raw_ao_build_ploc_dev(s, d_tris9, ntri, radius, nullptr);
    if (rc != RAW_AO_OK) {
        const std::string e = s->c.err;
        raw_ao_close(s);
        return bfail(b, rc, e);
    }
    b.scenes.push_back(s);
    b.scene_tris.push_back(d_tris9);
    b.dirty = true;
    raw_ao::drop_graph(b);
    if (scene_id) *scene_id = int(b.scenes.size()) - 1;
    return RAW_AO_OK;
}

int raw_ao_batch_add_view(raw_ao_batch* x, int scene, const float* d_pos, const float* d_nrm, const uint8_t* d_mask, int w, int h,
                          int spp, float radius, int ssao_samples, float ssao_radius, float* d_ao, float* d_ssao, float* d_err,
                          int* view_id) {
    if (!x || !x->b.self) return RAW_AO_ERR_ARG;
    Batch& b = x->b;
    b.self->c.err.clear();
    if (scene < 0 || size_t(scene) >= b.scenes.size()) return bfail(b, RAW_AO_ERR_ARG, "no such scene");
    if (!d_pos || !d_nrm || !d_mask || !d_ao || w <= 0 || h <= 0 || int64_t(w) * h > (int64_t(1) << 26) || spp < 1 ||
        spp > 4096 || !(radius > 0.0f) || (d_err && !d_ssao) || ssao_samples < 0 || ssao_samples > 4096)
        return bfail(b, RAW_AO_ERR_ARG, "bad view arguments");
    const uint64_t npix = uint64_t(w) * uint64_t(h);
    if (uint64_t(b.total_pix) + npix >= (uint64_t(1) << 31)) return bfail(b, RAW_AO_ERR_ARG, "the batch's pixels exceed 2^31");
    // the queue counts 64-ray groups in 32 bits, with room for every wave's last grab
    const uint64_t groups = (npix * uint64_t(spp) + 63u) / 64u;
    if (b.max_groups + groups >= (uint64_t(1) << 32) - (uint64_t(1) << 26))
        return bfail(b, RAW_AO_ERR_ARG, "the batch's rays exceed the queue (2^37)");
    b.max_groups += groups;
    raw_ao::BatchView v{};
    v.pos = d_pos;
    v.nrm = d_nrm;
    v.mask = d_mask;
    v.ao = d_ao;
    v.ssao = d_ssao;
    v.err = d_err;
    v.w = uint32_t(w);
    v.h = uint32_t(h);
    v.npix = uint32_t(npix);
    v.spp = uint32_t(spp);
    v.scene = uint32_t(scene);
    v.radius = radius;
    v.ssao_samples = uint32_t(ssao_samples);
    v.ssao_radius = ssao_radius;
    v.pix_base = b.total_pix;
    b.views.push_back(v);
    b.total_pix += uint32_t(npix);
    b.dirty = true;
    raw_ao::drop_graph(b);
    if (view_id) *view_id = int(b.views.size()) - 1;
    return RAW_AO_OK;
}

int raw_ao_batch_set_tuning(raw_ao_batch* x, int chunk, int tri_q, int blocks_per_cu) {
    if (!x || !x->b.self) return RAW_AO_ERR_ARG;
    Batch& b = x->b;
    b.self->c.err.clear();
    if (chunk < 1 || chunk > 1024 || tri_q < 1 || tri_q > 64 || blocks_per_cu < 1 || blocks_per_cu > 64)
        return bfail(b, RAW_AO_ERR_ARG, "tuning out of range (chunk 1..1024, tri_q 1..64, blocks per CU 1..64)");
    b.tune = raw_ao::BatchTune{uint32_t(chunk), uint32_t(tri_q), uint32_t(blocks_per_cu)};
    raw_ao::drop_graph(b);
    return RAW_AO_OK;
}

int raw_ao_batch_set_option(raw_ao_batch* x, int option, int value) {
    if (!x || !x->b.self) return RAW_AO_ERR_ARG;
    if (option != RAW_AO_OPT_HOST_LIBM_TRIG && option != RAW_AO_OPT_WAVE_CUT)
        return bfail(x->b, RAW_AO_ERR_ARG, "a batch takes RAW_AO_OPT_HOST_LIBM_TRIG and RAW_AO_OPT_WAVE_CUT only");
    const int rc = raw_ao_set_option(x->b.self, option, value);
    if (rc == RAW_AO_OK) raw_ao::drop_graph(x->b);
    return rc;
}

int raw_ao_batch_get_tuning(raw_ao_batch* x, int* chunk, int* tri_q, int* blocks_per_cu) {
    if (!x || !x->b.self) return RAW_AO_ERR_ARG;
    if (chunk) *chunk = int(x->b.tune.chunk);
    if (tri_q) *tri_q = int(x->b.tune.tri_q);
    if (blocks_per_cu) *blocks_per_cu = int(x->b.tune.blocks_per_cu);
    return RAW_AO_OK;
}

int raw_ao_batch_run(raw_ao_batch* x, int rebuild, void* stream) {
    if (!x || !x->b.self) return RAW_AO_ERR_ARG;
    Batch& b = x->b;
    if (!bbegin(b)) return RAW_AO_ERR_DEVICE;
    int r = raw_ao::prepare(b);
    if (!r) r = raw_ao::enqueue(b, rebuild != 0, bpick(b, stream));
    return raw_ao::code_of(r);
}

int raw_ao_batch_capture(raw_ao_batch* x, int with_builds) {
    if (!x || !x->b.self) return RAW_AO_ERR_ARG;
    Batch& b = x->b;
    if (!bbegin(b)) return RAW_AO_ERR_DEVICE;
    int r = raw_ao::prepare(b);
    if (r) return raw_ao::code_of(r);
    raw_ao::drop_graph(b);
    raw_ao::Ctx& c = b.self->c;
    hipStream_t s = c.stream;
    if (!hip_ok(c, hipStreamSynchronize(s), "capture") ||
        !hip_ok(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture"))
        return RAW_AO_ERR_DEVICE;
    r = raw_ao::enqueue(b, with_builds != 0, s);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (r) {
        if (g) (void)hipGraphDestroy(g);
        return raw_ao::code_of(r);
    }
    if (!hip_ok(c, e, "end capture")) return RAW_AO_ERR_DEVICE;
    b.graph = g;
    if (!hip_ok(c, hipGraphInstantiate(&b.exec, g, nullptr, nullptr, 0), "graph instantiate")) {
        raw_ao::drop_graph(b);
        return RAW_AO_ERR_DEVICE;
    }
    b.graph_has_builds = with_builds != 0;
    return RAW_AO_OK;
}

int raw_ao_batch_launch(raw_ao_batch* x, void* stream) {
    if (!x || !x->b.self) return RAW_AO_ERR_ARG;
    Batch& b = x->b;
    if (!bbegin(b)) return RAW_AO_ERR_DEVICE;
    if (!b.exec) return bfail(b, RAW_AO_ERR_ARG, "no captured graph: call raw_ao_batch_capture");
    return hip_ok(b.self->c, hipGraphLaunch(b.exec, bpick(b, stream)), "graph launch") ? RAW_AO_OK : RAW_AO_ERR_DEVICE;
}

int raw_ao_batch_check(raw_ao_batch* x) {
    if (!x || !x->b.self) return RAW_AO_ERR_ARG;
    Batch& b = x->b;
    if (!bbegin(b)) return RAW_AO_ERR_DEVICE;
    if (!hip_ok(b.self->c, hipDeviceSynchronize(), "batch")) return RAW_AO_ERR_DEVICE;
    for (auto* s : b.scenes) {
        const int r = raw_ao::ploc_fixed_check(s->c);
        if (r) return bfail(b, raw_ao::code_of(r), s->c.err);
    }
    return RAW_AO_OK;
}

int raw_ao_batch_results(raw_ao_batch* x, float* rmse, float* max_error, uint32_t* pixels, uint64_t* rays) {
    if (!x || !x->b.self) return RAW_AO_ERR_ARG;
    Batch& b = x->b;
    if (!bbegin(b)) return RAW_AO_ERR_DEVICE;
    raw_ao::Ctx& c = b.self->c;
    const size_t nv = b.views.size();
    if (nv == 0 || !b.stats.p) return bfail(b, RAW_AO_ERR_ARG, "the batch has not run");
    std::vector<raw_ao::ViewStats> st(nv);
    std::vector<uint32_t> ncov(nv);
    if (!hip_ok(c, hipDeviceSynchronize(), "batch") ||
        !hip_ok(c, hipMemcpy(st.data(), b.stats.p, sizeof(raw_ao::ViewStats) * nv, hipMemcpyDeviceToHost), "stats") ||
        !hip_ok(c, hipMemcpy(ncov.data(), b.ncov.p, 4 * nv, hipMemcpyDeviceToHost), "counts"))
        return RAW_AO_ERR_DEVICE;
    uint64_t total = 0;
    for (size_t v = 0; v < nv; ++v) {
        total += uint64_t(ncov[v]) * b.views[v].spp;
        if (rmse) rmse[v] = st[v].count ? float(std::sqrt(st[v].sumsq / double(st[v].count))) : 0.0f;
        if (max_error) {
            float m;
            std::memcpy(&m, &st[v].maxbits, 4);
            max_error[v] = m;
        }
        if (pixels) pixels[v] = ncov[v];
    }
    if (rays) *rays = total;
    return RAW_AO_OK;
}

int raw_ao_batch_tune(raw_ao_batch* x, const char* cache_path, float* best_ms) {
    if (!x || !x->b.self) return RAW_AO_ERR_ARG;
    Batch& b = x->b;
    if (!bbegin(b)) return RAW_AO_ERR_DEVICE;
    raw_ao::Ctx& c = b.self->c;
    int r = raw_ao::prepare(b);
    if (r) return raw_ao::code_of(r);
    // the workload's signature: device, CUs, the wave cut, views (size, spp) and scenes (triangles)
    hipDeviceProp_t prop;
    if (!hip_ok(c, hipGetDeviceProperties(&prop, c.device), "device properties")) return RAW_AO_ERR_DEVICE;
    std::ostringstream sig;
    sig << prop.gcnArchName << ":" << b.cus << ":c" << c.wave_cut;
    for (const auto& v : b.views) sig << ":" << v.w << "x" << v.h << "x" << v.spp << "s" << v.scene;
    for (auto* s : b.scenes) sig << ":t" << s->c.ntri;
    const std::string key = sig.str();
    if (cache_path) {
        std::ifstream in(cache_path);
        std::string line;
        while (std::getline(in, line)) {
            std::istringstream ls(line);
            std::string k;
            int ch, tq, bp;
            float ms;
            if (ls >> k >> ch >> tq >> bp >> ms && k == key) {
                b.tune = raw_ao::BatchTune{uint32_t(ch), uint32_t(tq), uint32_t(bp)};
                if (best_ms) *best_ms = ms;
                raw_ao::drop_graph(b);
                return RAW_AO_OK;
            }
        }
    }
    hipStream_t s = c.stream;
    const auto time_one = [&](raw_ao::BatchTune t, float& ms) -> int {
        b.tune = t;
        ms = 1e30f;
        for (int rep = 0; rep < 3; ++rep) {   // the first run also warms the caches
            if (!hip_ok(c, hipStreamSynchronize(s), "tune")) return -2;
            const auto t0 = std::chrono::steady_clock::now();
            const int e = raw_ao::enqueue(b, false, s);
            if (e) return e;
            if (!hip_ok(c, hipStreamSynchronize(s), "tune")) return -2;
            if (rep > 0) ms = std::min(ms, raw_ao::ms_of(t0));
        }
        return 0;
    };
    // coordinate descent from the defaults: queue chunk, then triangle phase, then blocks
    raw_ao::BatchTune best;
    float best_t = 0.0f;
    if ((r = time_one(best, best_t))) return raw_ao::code_of(r);
    const uint32_t chunks[] = {1, 2, 4, 8, 16, 32}, tqs[] = {4, 8, 16, 24, 32, 48}, bps[] = {2, 4, 8, 12, 16};
    for (int axis = 0; axis < 3; ++axis) {
        const uint32_t* vals = axis == 0 ? chunks : (axis == 1 ? tqs : bps);
        const int nvals = axis == 2 ? 5 : 6;
        raw_ao::BatchTune base = best;
        for (int i = 0; i < nvals; ++i) {
            raw_ao::BatchTune t = base;
            (axis == 0 ? t.chunk : (axis == 1 ? t.tri_q : t.blocks_per_cu)) = vals[i];
            if (t.chunk == best.chunk && t.tri_q == best.tri_q && t.blocks_per_cu == best.blocks_per_cu) continue;
            float ms;
            if ((r = time_one(t, ms))) return raw_ao::code_of(r);
            if (ms < best_t) {
                best_t = ms;
                best = t;
            }
        }
    }
    b.tune = best;
    raw_ao::drop_graph(b);
    if (best_ms) *best_ms = best_t;
    if (cache_path) {
        std::ofstream out(cache_path, std::ios::app);
        out << key << " " << best.chunk << " " << best.tri_q << " " << best.blocks_per_cu << " " << best_t << "\n";
    }
    return RAW_AO_OK;
}

}  // extern "C"
