// The context behind the C ABI (include/raw_ao_hip.h) and the internal launch API
// shared by the kernel translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "../../include/raw_ao_hip.h"
#include <string>
#include <vector>

namespace raw_ao {

// A grow-only device allocation.
struct DevBuf {
    void* p{nullptr};
    size_t cap{0};
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// Device state of the PLOC rounds, double-buffered by round parity so a round's
// kernels read one copy while the scan writes the next.
struct PlocState {
    uint32_t m;        // clusters this round
    uint32_t next_id;  // id of the next internal node
    uint32_t merges;   // merges made by the round that produced this state
    uint32_t stalled;  // sticky: some round with m > 1 made no merge
    uint32_t rounds;   // rounds with m > 1 so far
    uint32_t root;     // the last cluster (final state only)
    uint32_t grid_rounds;   // rounds run by the grid kernels (the rest by the one-workgroup tail)
    uint32_t pad;
};

struct UNet;

constexpr int kPlocHist = 240;   // grid rounds whose cluster counts a build records (Ctx::ploc_hist)
constexpr int kWideHist = 64;    // and collapse grid levels whose frontier sizes it records (after them)

struct TraceCounters {
    unsigned long long visits, tests, rays, overflows;
};

struct Ctx {
    int device{0};
    hipStream_t stream{nullptr};
    hipStream_t aux{nullptr};            // the whole frame's raster beside its build (made on first use, default priority)
    hipEvent_t aux_fork{nullptr}, aux_join{nullptr};
    std::string err;
    bool host_trig{true};
    int trace_kernel{0};         // RAW_AO_TRACE_WIDE or RAW_AO_TRACE_BINARY
    int built_kernel{0};         // the form the last build made
    int tri_q{16};               // wide kernel: triangle phase when tri_q/64 of live lanes wait
    int ploc_tail{1024};         // the single-workgroup PLOC tail takes over at this many clusters
    int refill{48};              // 64-spp wide kernel: refill lanes once this many are idle (0: off)
    int wave_cut{2};             // 64-spp wide kernel: > 0 starts rays from the wave's cut at wave_cut x radius
    bool near_first{true};       // 64-spp wide kernel: descend into the nearest hit child first
    int learned_backoff{0};      // learned frames: skip the network for up to this many calls after a fall-back (0: off)
    int lb_next{0};              // the skip length after the next fall-back (1, 2, 4, ..., learned_backoff)
    int lb_left{0};              // calls left to skip
    int lb_w{0}, lb_h{0};        // the image size the skipping applies to
    int sky_refill{20};          // 64-spp visibility: sky rays apart from the AO rays, lanes refilled at this many idle (0: off)
    bool skew{false};            // the skew control: AO query distances x 0.99f
    int build_rounds{0};         // PLOC rounds and collapse levels a build enqueues (0: automatic)
    bool count_next{false};

    // host-libm sin/cos correction table: a 2 MB bitmap of the corrected angle indices, then one nibble an index
    uint8_t* d_trig{nullptr};
    bool trig_ready{false};

    // the scene
    uint32_t ntri{0};
    int radius{16};
    int32_t root{-1};
    uint32_t rounds{0};
    DevBuf tris_in;              // n x 9 floats, input order
    DevBuf keys, vals, keys2, vals2, hist, rs_tot;
    DevBuf boxes, links;         // (2n-1) x 6 floats, (2n-1) x 3 ints: the reference's node order
    DevBuf clus_a, clus_b, nn, flags, blk;
    DevBuf state;                // PlocState[2]
    PlocState* h_state{nullptr}; // pinned read-back of the build's final states (PLOC at 0, the collapse's at 128)
                                 // and the grid rounds' cluster counts (at 1024)
    DevBuf g4;                   // centroid box lo xyz and scale
    DevBuf bounds_part;
    // traversal forms
    DevBuf nodes;                // binary records, 64 B each (the ray-per-lane binary kernel)
    DevBuf tris_sorted;          // 48 B each, sorted (leaf) order
    uint32_t trav_root{0};
    DevBuf nodes4;               // 4-wide records, 128 B each, root at 0
    DevBuf nodes4q;              // RAW_AO_OPT_QUANTIZED_NODES: the same records quantized, 64 B each
    int node_q{1};               // RAW_AO_OPT_QUANTIZED_NODES: 0 none, 1 the sky rays, 2 also the 64-spp AO rays
    bool q_built{false};         // nodes4q holds the current tree's records (or is enqueued after its build)
    DevBuf top;                  // the wide tree's top table (slots to depth 3) and its entry count
    DevBuf wfront_a, wfront_b, wblk, wstate;
    DevBuf vis_open;             // the visibility pass's per-pixel open AO samples (64 bits a pixel)
    uint32_t wide_nodes{0};

    // staging for host-pointer entry points
    DevBuf st_pos, st_nrm, st_mask, st_out, st_out2, st_cnt, st_rays, st_max;
    DevBuf counters;             // TraceCounters
    TraceCounters last_counts{};
    bool have_counts{false};

    hipEvent_t ev[10]{};         // 0/1 build, 2/3 trace, 4/5 SSAO, 6/7 error map, 8/9 a whole frame
    float ms_build{0}, ms_trace{0}, ms_ssao{0}, ms_error{0};

    UNet* net{nullptr};          // the learned AO surrogate, made on first use
    int conv_kernel{0};          // 0: LDS-tiled (persistent for 32 in, 32 out; direct-copy persistent for a
                                 // concatenated input with 32 out) where it applies, 1: general, 2: tap-major where
                                 // it applies, 3: LDS-tiled, never persistent, 4: direct-copy persistent where it applies
    int cus{0};                  // compute units of the device (read on first use)

    // grids of the last eager build, replayed by the capturable build (ploc_build_fixed,
    // wide_build_fixed)
    std::vector<uint32_t> plan_ploc_nb, plan_wide_nb;
    DevBuf ploc_hist;                    // clusters left after each grid round of the build in flight, then
                                         // the frontier size of each collapse grid level
    std::vector<uint32_t> round_m;       // clusters at the start of each grid round of the last finished
                                         // build of last_rounds_n triangles (sizes those rounds' grids)
    std::vector<uint32_t> level_m;       // and the frontier at the start of each collapse grid level
    uint32_t ploc_round{0};              // grid rounds enqueued so far in the build in flight
    uint32_t wide_level{0};              // collapse grid levels enqueued so far
    uint32_t plan_n{0};
    int plan_radius{0};
    // the grid rounds and wide-collapse levels of the last finished build, and of how many
    // triangles: the next build of that many enqueues as many (a first build estimates them)
    uint32_t last_rounds_n{0}, last_grid_rounds{0}, last_wide_levels{0};
    // where the enqueued PLOC rounds and collapse levels leave their state (buffer parity) and
    // clusters or frontier (true: the first buffer), for a build that has to continue them
    int ploc_parity{0}, wide_parity{0};
    bool ploc_in_a{true}, wide_in_a{true};

    // the G-buffer rasterizer: per-pixel (depth, draw id) keys, drawn triangles, their box
    // pixel counts and the counts' exclusive scan
    DevBuf r_keys, r_tris, r_counts, r_offs, r_scan_tmp;

    // the frame: mesh table, device scalars, planes for the host-pointer form, and the pinned
    // read-back of the error map (+ the scalars after it)
    DevBuf fr_mesh, fr_scalars, fr_planes;
    float* fr_host{nullptr};
    size_t fr_host_cap{0};
};

bool ensure(Ctx& c, DevBuf& b, size_t bytes);
bool hip_ok(Ctx& c, hipError_t e, const char* what);

// Kernel-side entry points (each in its own translation unit).
// Enqueues the whole PLOC build (bounds to root) with no host wait; build_finish() checks it.
int ploc_build(Ctx& c, const float* d_tris, uint32_t n, int radius, hipStream_t s);
// After ploc_build, bvh_layout and wide_build have been enqueued: one read-back of the final
// PLOC and collapse states (the build's only wait); where the enqueued rounds or levels fell
// short, the rest run with the cluster count read back, and the later steps again.
int build_finish(Ctx& c, hipStream_t s);
// The rest of an unfinished PLOC build (rounds with read-backs, then the tail).
int ploc_continue(Ctx& c, hipStream_t s);
int bvh_layout(Ctx& c, hipStream_t s);
int trig_table_upload(Ctx& c);
int trace_ao(Ctx& c, const float* d_pos, const float* d_nrm, const uint8_t* d_mask, int w, int h, int spp,
             float radius, float* d_ao, hipStream_t s);
int angle_table(Ctx& c, float* d_sin, float* d_cos, hipStream_t s);
int ao_finish(Ctx& c, const uint8_t* d_mask, const uint32_t* d_cnt, uint32_t npix, uint32_t spp, float* d_ao, hipStream_t s);
int wide_build(Ctx& c, hipStream_t s);
// The rest of an unfinished collapse (levels with read-backs), then the top table.
int wide_continue(Ctx& c, hipStream_t s);
// The word offset of the entry count in Ctx::top.
size_t top_count_offset();
int ploc_build_fixed(Ctx& c, const float* d_tris, hipStream_t s);
int ploc_fixed_check(Ctx& c);
int wide_build_fixed(Ctx& c, hipStream_t s);
int trace_ao4(Ctx& c, const float* d_pos, const float* d_nrm, const uint8_t* d_mask, int w, int h, int spp,
              float radius, float* d_ao, hipStream_t s);
// trace_ao4 with its scratch given: d_counters (cleared here), d_cnt (npix words, cleared here; only
// read below 64 spp, may be null at 64). timed: record the context's events and honour
// raw_ao_count_next_trace; otherwise neither, and nothing but kernels is enqueued.
int trace_ao4_into(Ctx& c, const float* d_pos, const float* d_nrm, const uint8_t* d_mask, int w, int h, int spp,
                   float radius, float* d_ao, TraceCounters* d_counters, uint32_t* d_cnt, bool timed, hipStream_t s);
// Looks up this device's copy of the no-allocation entry point's scratch array (once per device,
// with the device current); RAW_AO_ERR_DEVICE when the lookup fails.
int static_scratch_setup(Ctx& c);
int trace_vis4(Ctx& c, const float* d_pos, const float* d_nrm, const uint8_t* d_mask, int w, int h, int spp, float radius,
               const float to_light[3], int sun_forced, float* d_ao, float* d_sky, float* d_sun, hipStream_t s);
int trace_rays4(Ctx& c, const float* d_rays, uint32_t n, int any, float far_scale, float* d_out, hipStream_t s);
int trace_rays(Ctx& c, const float* d_rays, uint32_t n, int any, float far_scale, float* d_out, hipStream_t s);
int closest_hits4(Ctx& c, const float* d_rays, uint32_t n, float* d_hits, hipStream_t s);
int ssao(Ctx& c, const float* d_pos, const float* d_nrm, const uint8_t* d_mask, int w, int h, int samples,
         float radius, float* d_ao, hipStream_t s);
// k_ssao without event records (stream-capture safe); trig may be null (correctly rounded).
void ssao_enqueue(const uint8_t* trig, const float* d_pos, const float* d_nrm, const uint8_t* d_mask, int w, int h, int samples,
                  float radius, float* d_ao, hipStream_t s);
int error_map(Ctx& c, const float* d_a, const float* d_t, const uint8_t* d_mask, int w, int h, float* d_err,
              float* d_max, hipStream_t s);

int rasterize(Ctx& c, const float* d_vpos, const float* d_vnrm, const int32_t* d_idx, uint32_t ntri, const float eye[3],
              const float center[3], const float up[3], float fovy, float aspect, float nearz, float farz, int w, int h,
              float* d_depth, float* d_pos, float* d_nrm, uint8_t* d_mask, hipStream_t s, const uint32_t* d_mesh_end = nullptr,
              const float* d_albedo = nullptr, uint32_t nmesh = 0, float* d_albedo_out = nullptr);
// perspective(fovy, aspect, near, far) * lookAt(eye, center, up) as the reference computes it
// (raw/math/mat.hpp:7-38), row-major.
void view_proj(const float eye[3], const float center[3], const float up[3], float fovy, float aspect, float nearz, float farz,
               float out[16]);

int render_frame(Ctx& c, const raw_ao_scene& sc, const raw_ao_render_opts& o, int w, int h, const raw_ao_frame& d,
                 raw_ao_frame_stats* stats, hipStream_t s);

void net_release(Ctx& c);
bool conv_shape_ok(int c0, int c1, int cout);

// Host-libm correction table (trig_table.cpp): a 2^21-byte bitmap, then 2^23 bytes of nibbles.
bool build_trig_table(std::vector<uint8_t>& out, std::string& err);

}  // namespace raw_ao

struct raw_ao_ctx {
    raw_ao::Ctx c;
};
