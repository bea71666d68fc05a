// Batched AO: many views of several scenes traced by one persistent launch that takes
// work from a global queue of 64-ray groups (kernels in bvh4.hip, host side in batch.hip).
#pragma once
#include "ctx.hpp"

namespace raw_ao {

struct BatchScene {
    const void* nodes;      // the scene's 4-wide records
    const float4* tris;     // its sorted triangles
    const void* top;        // its top table (slots to depth 3)
    const uint32_t* ntop;   // the table's entry count
    uint32_t ntri, pad;
};

struct BatchView {
    const float* pos;
    const float* nrm;
    const uint8_t* mask;
    float* ao;              // computeRTAO's output
    float* ssao;            // computeSSAO's output, the reconcile's approximation (null: neither)
    float* err;             // the reconcile's error map (null: no error map)
    uint32_t w, h, npix, spp, scene;
    float radius;
    uint32_t ssao_samples;
    float ssao_radius;
    uint32_t pix_base;      // the view's first slot in the batch's per-pixel buffers
};

// Device state of one batched launch.
struct BatchDev {
    const BatchView* views;
    const BatchScene* scenes;
    uint32_t nviews;
    uint32_t total_pix;     // sum of the views' pixels
    uint32_t* cov;          // covered pixel indices, the view's at pix_base
    uint32_t* cnt;          // open-ray counts per pixel, the view's at pix_base
    uint32_t* ncov;         // covered pixels per view
    uint32_t* gbase;        // first 64-ray group of each view; [nviews] = the total
    uint32_t* head;         // the queue head (groups handed out)
    const uint8_t* trig;
    uint32_t tri_q;         // triangle phase threshold (of 64)
    uint32_t chunk;         // groups a wave takes from the queue at a time
    float cut_k;            // > 0: a group's rays start from the wave's cut at cut_k x radius
    uint32_t refill;        // > 0 (with cut_k > 0): the refill form, a wave's idle lanes refilled at this many
};

// The tunables of a batched launch.
struct BatchTune {
    uint32_t chunk{4};
    uint32_t tri_q{16};
    uint32_t blocks_per_cu{8};
};

// clear -> compact -> plan -> persistent trace -> finish, all on stream s (graph safe:
// no memset nodes, no host copies).
int batch_trace_enqueue(const BatchDev& d, uint32_t persistent_blocks, hipStream_t s);

}  // namespace raw_ao
