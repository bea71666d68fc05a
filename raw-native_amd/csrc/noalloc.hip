// raw_ao_trace_ao_dev_noalloc: raw_ao_trace_ao_dev for callers that cannot allocate device memory
// (code under stream capture, serving loops, harnesses that forbid hipMalloc).
//
// The trace's scratch (the trace counters, and below 64 spp the per-pixel open counts) lives in
// the file-scope array `scratch` below instead of the context's grow-only buffers. Each device
// that loads this library holds its own copy of the array until the process ends; its address on
// a device is looked up once, with that device current, when the first context opens on it
// (static_scratch_setup), and kept in g_scratch. The array's name in host code is only a host-side
// stand-in, never a device address: every kernel argument below is built from g_scratch.
//
// Every call on one device shares that device's array and leaves its contents behind; the kernels
// clear what must start at zero. Calls that overlap on one device (two streams, two host threads)
// race on the array.
#include "../../include/raw_ao_hip.h"
#include "ctx.hpp"
#include <mutex>

namespace raw_ao {
namespace {

constexpr size_t kCounterBytes = 256;   // TraceCounters, padded to the array's alignment
constexpr size_t kScratchBytes = kCounterBytes + 4 * size_t(RAW_AO_NOALLOC_MAX_PIXELS);
static_assert(sizeof(TraceCounters) <= kCounterBytes, "trace counters exceed their scratch slot");

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

int static_scratch_setup(Ctx& c) {
    if (c.device < 0 || c.device >= kMaxDevices) {
        c.err = "device ordinal beyond the no-allocation scratch table";
        return RAW_AO_ERR_DEVICE;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[c.device]) return RAW_AO_OK;
    void* p = nullptr;
    if (!hip_ok(c, hipGetSymbolAddress(reinterpret_cast<void**>(&p), HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)")) return RAW_AO_ERR_DEVICE;
    if (!p) {
        c.err = "hipGetSymbolAddress(scratch) returned no address";
        return RAW_AO_ERR_DEVICE;
    }
    g_scratch[c.device] = p;
    return RAW_AO_OK;
}

}  // namespace raw_ao

extern "C" int raw_ao_trace_ao_dev_noalloc(raw_ao_ctx* x, const float* d_pos, const float* d_nrm, const uint8_t* d_mask, int w,
                                           int h, int spp, float radius, float* d_ao, void* stream) {
    using namespace raw_ao;
    if (!x) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    c.err.clear();
    if (!hip_ok(c, hipSetDevice(c.device), "hipSetDevice")) return RAW_AO_ERR_DEVICE;
    if (w <= 0 || h <= 0 || !d_pos || !d_nrm || !d_mask || !d_ao) {
        c.err = "bad image arguments";
        return RAW_AO_ERR_ARG;
    }
    if (int64_t(w) * int64_t(h) > int64_t(RAW_AO_NOALLOC_MAX_PIXELS)) {
        c.err = "image larger than RAW_AO_NOALLOC_MAX_PIXELS (4194304 pixels) for the no-allocation trace";
        return RAW_AO_ERR_ARG;
    }
    if (spp < 1 || spp > 4096) {
        c.err = "spp must be in 1..4096";
        return RAW_AO_ERR_ARG;
    }
    if (!(radius > 0.0f)) {
        c.err = "radius must be > 0";
        return RAW_AO_ERR_ARG;
    }
    if (c.ntri && c.root < 0) {
        c.err = "no BVH: call raw_ao_build_ploc first";
        return RAW_AO_ERR_ARG;
    }
    if (c.trace_kernel != RAW_AO_TRACE_WIDE || (c.ntri && c.built_kernel != RAW_AO_TRACE_WIDE)) {
        c.err = "the no-allocation trace needs the wide traversal form (RAW_AO_TRACE_WIDE) built";
        return RAW_AO_ERR_ARG;
    }
    if (c.host_trig && !c.trig_ready) {
        c.err = "the host-libm angle table is not on the device: build the scene (raw_ao_build_ploc) first";
        return RAW_AO_ERR_ARG;
    }
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        base = (c.device >= 0 && c.device < kMaxDevices) ? g_scratch[c.device] : nullptr;
    }
    if (!base) {
        c.err = "no-allocation scratch not set up on this device";
        return RAW_AO_ERR_DEVICE;
    }
    auto* counters = static_cast<TraceCounters*>(base);
    auto* cnt = reinterpret_cast<uint32_t*>(static_cast<unsigned char*>(base) + kCounterBytes);
    if (c.skew) radius *= 0.99f;   // as raw_ao_trace_ao_dev (accel.cpp:28, Mode::Skew)
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : c.stream;
    const int r = trace_ao4_into(c, d_pos, d_nrm, d_mask, w, h, spp, radius, d_ao, counters, spp != 64 ? cnt : nullptr, false, s);
    if (r == 0) return RAW_AO_OK;
    return r == -1 ? RAW_AO_ERR_ARG : RAW_AO_ERR_DEVICE;
}
