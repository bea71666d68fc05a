// The learned AO surrogate on gfx950: a 4-level convolutional U-Net (32/64/96/128
// channels, two 3x3 convolutions a level, ReLU, 2x2 average pooling down, nearest 2x
// upsampling and skip concatenation up, a 1x1 head) from G-buffer features to AO, trained
// and run with BF16 activations and weights on MFMA with FP32 accumulation.
//
// Layout: activations NHWC bf16 (images stacked, pixel index = (n*H + y)*W + x); a layer's
// weights as [Cout][Kpad] bf16 with k = tap*Cin + ci (tap = ky*3 + kx, offset (ky-1, kx-1)),
// Kpad = K rounded up to 32. FP32 master parameters: per layer [Cout][K] weights, then Cout
// biases; the head's 32 weights and bias last.
//
// Convolution = implicit GEMM on v_mfma_f32_16x16x32_bf16 (lane l of a wave holds A row
// l%16, B column l%16, k 8*(l/16)..+7; D column l%16, rows 4*(l/16)..+3):
//   forward     A = weights (cout x k), B = input patches (k x pixel): a wave computes up to
//               128 output channels for 32 pixels; the decoder's skip + 2x upsampled
//               concatenation is read in the B gather, never materialized; bias and ReLU fused.
//   input grad  the same kernel over the flipped, transposed weights.
//   weight grad A = output grads (cout x pixel), B = input patches (pixel x k), both staged
//               through LDS so lanes read 8 pixels contiguously; per-chunk partial sums
//               reduced over chunks in a fixed order (deterministic). The bias is the weight
//               of a constant-1 input column (k = K).
#include "../../include/raw_ao_hip.h"
#include "unet.hpp"
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

namespace raw_ao {
namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;
typedef unsigned short bf16_t;   // storage type

__device__ __forceinline__ bf16_t f2bf(float f) {   // round to nearest even
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return bf16_t(0x7fc0);
    u += 0x7fffu + ((u >> 16) & 1u);
    return bf16_t(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16_t b) { return __uint_as_float(uint32_t(b) << 16); }
// f2bf's bits from the hardware conversion (v_cvt_pk_bf16_f32 rounds to nearest even): the same for
// every non-NaN float; a NaN gives f2bf's 0x7fc0
__device__ __forceinline__ uint32_t f2bf2(float lo, float hi) {
    typedef __attribute__((ext_vector_type(2))) __bf16 bf16v2;
    typedef __attribute__((ext_vector_type(2))) float f32v2;
    const bf16v2 b = __builtin_convertvector(f32v2{lo, hi}, bf16v2);
    uint32_t u = __builtin_bit_cast(uint32_t, b);
    if (lo != lo) u = (u & 0xffff0000u) | 0x7fc0u;
    if (hi != hi) u = (u & 0x0000ffffu) | 0x7fc00000u;
    return u;
}

union V8 {
    uint4 u;
    bf16x8 v;
};

// ---- forward / input-grad convolution -------------------------------------------
struct ConvArgs {
    const bf16_t* in0;   // N,H,W,C0
    const bf16_t* in1;   // N,H/2,W/2,C1 (upsampled 2x on read) or null
    const bf16_t* w;     // [Cout][Kpad]
    const float* bias;   // [Cout] or null
    bf16_t* out;         // N,H,W,Cout
    int n, h, wd, c0, c1, cout, kpad, relu;
    int co0;             // first output channel of this launch's group
    bf16_t* pool;        // or null: also the 2x2 average pool of out, N,H/2,W/2,Cout (k_conv_lds, k_conv_pf with POOL)
};

// One wave: 32 pixels (2 B tiles) x CT*16 output channels starting at co0.
template <int CT>
__global__ __launch_bounds__(256) void k_conv(ConvArgs a) {
    const int lane = threadIdx.x & 63;
    const int wave = int((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int npix = a.n * a.h * a.wd;
    const int pbase = wave * 32;
    if (pbase >= npix) return;
    const int cin = a.c0 + a.c1, K = 9 * cin;
    const int g = lane >> 4, r = lane & 15;
    int py[2], px[2], pn[2];
    bool pvalid[2];
    for (int t = 0; t < 2; ++t) {
        const int p = pbase + t * 16 + r;
        pvalid[t] = p < npix;
        const int pp = pvalid[t] ? p : 0;
        px[t] = pp % a.wd;
        const int q = pp / a.wd;
        py[t] = q % a.h;
        pn[t] = q / a.h;
    }
    f32x4 acc[CT][2];
    for (int i = 0; i < CT; ++i)
        for (int t = 0; t < 2; ++t) acc[i][t] = f32x4{0, 0, 0, 0};
    const int hh = a.h >> 1, hw = a.wd >> 1;
    for (int kc = 0; kc < a.kpad; kc += 32) {
        const int k = kc + 8 * g;
        V8 bv[2];
        const bool kin = k < K;
        const int tap = kin ? k / cin : 0, ci = kin ? k - tap * cin : 0;
        const int dy = tap / 3 - 1, dx = tap % 3 - 1;
        for (int t = 0; t < 2; ++t) {
            bv[t].u = make_uint4(0, 0, 0, 0);
            const int y = py[t] + dy, x = px[t] + dx;
            if (!kin || !pvalid[t] || y < 0 || y >= a.h || x < 0 || x >= a.wd) continue;
            const bf16_t* src;
            if (ci < a.c0) src = a.in0 + (size_t(pn[t] * a.h + y) * a.wd + x) * a.c0 + ci;
            else src = a.in1 + (size_t(pn[t] * hh + (y >> 1)) * hw + (x >> 1)) * a.c1 + (ci - a.c0);
            bv[t].u = *reinterpret_cast<const uint4*>(src);
        }
        for (int i = 0; i < CT; ++i) {
            V8 av;
            const int co = a.co0 + i * 16 + r;
            av.u = *reinterpret_cast<const uint4*>(a.w + size_t(co) * a.kpad + k);
            for (int t = 0; t < 2; ++t) acc[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av.v, bv[t].v, acc[i][t], 0, 0, 0);
        }
    }
    for (int t = 0; t < 2; ++t) {
        if (!pvalid[t]) continue;
        const size_t p = size_t(pbase + t * 16 + r);
        for (int i = 0; i < CT; ++i) {
            const int co = a.co0 + i * 16 + 4 * g;
            float v[4];
            for (int j = 0; j < 4; ++j) {
                v[j] = acc[i][t][j] + (a.bias ? a.bias[co + j] : 0.0f);
                if (a.relu && v[j] < 0.0f) v[j] = 0.0f;
            }
            const uint32_t lo = uint32_t(f2bf(v[0])) | (uint32_t(f2bf(v[1])) << 16);
            const uint32_t hi = uint32_t(f2bf(v[2])) | (uint32_t(f2bf(v[3])) << 16);
            *reinterpret_cast<uint2*>(a.out + p * a.cout + co) = make_uint2(lo, hi);
        }
    }
}

// The same convolution for cin a multiple of 32 (every layer but the first): a 32-wide k
// step then lies inside one tap, so the tap's pixel offset is wave-uniform and kept
// incrementally (no per-lane divisions), and the 4 waves of a block share the weights
// through LDS, staged KS columns at a time. A wave computes PT tiles of 16 pixels. The
// MFMA sums run over k in the same order as k_conv, so the outputs are identical.
template <int CT, int PT>
__global__ __launch_bounds__(256) void k_conv_t(ConvArgs a) {
    constexpr int KS = 128;
    constexpr int RW = KS / 8 + 1;   // 16-byte units a staged weight row (+1 against bank conflicts)
    __shared__ uint4 wl[CT * 16 * RW];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int g = lane >> 4, r = lane & 15;
    const int npix = a.n * a.h * a.wd;
    const int pbase = (int(blockIdx.x) * 4 + wv) * (16 * PT);
    const int cin = a.c0 + a.c1;
    const int hh = a.h >> 1, hw = a.wd >> 1;
    int py[PT], px[PT], pn[PT];
    bool pvalid[PT];
    for (int t = 0; t < PT; ++t) {
        const int p = pbase + t * 16 + r;
        pvalid[t] = p < npix;
        const int pp = pvalid[t] ? p : 0;
        px[t] = pp % a.wd;
        const int q = pp / a.wd;
        py[t] = q % a.h;
        pn[t] = q / a.h;
    }
    f32x4 acc[CT][PT];
    for (int i = 0; i < CT; ++i)
        for (int t = 0; t < PT; ++t) acc[i][t] = f32x4{0, 0, 0, 0};
    int tap = 0, cb = 0;   // the tap and first channel of the current 32-wide k step
    for (int ks = 0; ks < a.kpad; ks += KS) {
        const int kn = min(KS, a.kpad - ks);
        __syncthreads();
        for (int e = threadIdx.x; e < CT * 16 * (kn / 8); e += 256) {
            const int co = e / (kn / 8), kk = e - co * (kn / 8);
            wl[co * RW + kk] = *reinterpret_cast<const uint4*>(a.w + size_t(a.co0 + co) * a.kpad + ks + kk * 8);
        }
        __syncthreads();
        for (int kc = 0; kc < kn; kc += 32) {
            const int dy = tap / 3 - 1, dx = tap % 3 - 1;
            const int ci = cb + 8 * g;
            const bool lo = cb < a.c0;   // wave-uniform: c0 is a multiple of 32
            V8 bv[PT];
            for (int t = 0; t < PT; ++t) {
                bv[t].u = make_uint4(0, 0, 0, 0);
                const int y = py[t] + dy, x = px[t] + dx;
                if (!pvalid[t] || y < 0 || y >= a.h || x < 0 || x >= a.wd) continue;
                const bf16_t* src = lo ? a.in0 + (size_t(pn[t] * a.h + y) * a.wd + x) * a.c0 + ci
                                       : a.in1 + (size_t(pn[t] * hh + (y >> 1)) * hw + (x >> 1)) * a.c1 + (ci - a.c0);
                bv[t].u = *reinterpret_cast<const uint4*>(src);
            }
            for (int i = 0; i < CT; ++i) {
                V8 av;
                av.u = wl[(i * 16 + r) * RW + (kc >> 3) + g];
                for (int t = 0; t < PT; ++t) acc[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av.v, bv[t].v, acc[i][t], 0, 0, 0);
            }
            cb += 32;
            if (cb == cin) {
                cb = 0;
                ++tap;
            }
        }
    }
    for (int t = 0; t < PT; ++t) {
        if (!pvalid[t]) continue;
        const size_t p = size_t(pbase + t * 16 + r);
        for (int i = 0; i < CT; ++i) {
            const int co = a.co0 + i * 16 + 4 * g;
            float v[4];
            for (int j = 0; j < 4; ++j) {
                v[j] = acc[i][t][j] + (a.bias ? a.bias[co + j] : 0.0f);
                if (a.relu && v[j] < 0.0f) v[j] = 0.0f;
            }
            const uint32_t lo = uint32_t(f2bf(v[0])) | (uint32_t(f2bf(v[1])) << 16);
            const uint32_t hi = uint32_t(f2bf(v[2])) | (uint32_t(f2bf(v[3])) << 16);
            *reinterpret_cast<uint2*>(a.out + p * a.cout + co) = make_uint2(lo, hi);
        }
    }
}

// The convolution from LDS tiles, for cin a multiple of 32. A block computes one work
// item: an 8 x 32 pixel tile of one image (a wave: 2 rows as 4 tiles of 16 pixels) for one
// group of CT*16 output channels (blockIdx.x = group * tiles + tile). Per 32-channel chunk it
// stages the tile's input with a 1-pixel border (10 x 34 pixels, the concatenation's
// upsampled half read in place) and the chunk's weights for all 9 taps in LDS, then runs the
// 9 taps from LDS: every input value is read from global memory once per chunk instead of
// once per tap. Two or more blocks share a CU, so one block's staging overlaps another's
// MFMAs. LDS rows are padded so the 16 lanes of a fragment row hit distinct banks. The sums
// run chunk-major (taps inside a chunk), so the results may differ from k_conv's by the
// rounding of the float32 sums.
constexpr int LT_H = 8, LT_W = 32, LH_H = LT_H + 2, LH_W = LT_W + 2;

// The output of an 8 x 32 tile's wave (rows oy + 2 wv and oy + 2 wv + 1, tile t: row t >> 1,
// columns ox + 16 (t & 1) + r) with bias and ReLU, rounded to bf16; with a.pool also its 2x2
// average pool: a 2x2 block is rows 2 wv, 2 wv + 1 of one tile column and lanes r, r ^ 1, summed
// (x0 + x1) + (x0' + x1') in float from the rounded outputs, times 0.25, rounded (k_pool_fwd's
// operations, so the same bits). With lb (LDS) the group's biases are read from there.
// In two steps, so a persistent kernel can hold a finished tile's outputs in registers and
// store them later: conv_tile_pack rounds the outputs (and forms the pool, lanes r even) into
// registers, conv_tile_store writes what lies inside the image.
template <int CT, bool POOL>
struct TileOut {
    uint2 o[CT][4];
    uint2 p[CT][POOL ? 2 : 1];
};
// the lane's biases (output channels co0 + 16 i + 4 g + j), zero without a bias
template <int CT>
__device__ __forceinline__ void conv_lane_bias(const ConvArgs& a, int co0, int g, float (&b)[CT][4]) {
    for (int i = 0; i < CT; ++i)
        for (int j = 0; j < 4; ++j) b[i][j] = a.bias ? a.bias[co0 + i * 16 + 4 * g + j] : 0.0f;
}
template <int CT, bool POOL>
__device__ __forceinline__ void conv_tile_pack(const ConvArgs& a, const f32x4 (&acc)[CT][4], const float (&bias)[CT][4],
                                               TileOut<CT, POOL>& out) {
    for (int i = 0; i < CT; ++i)
        for (int t = 0; t < 4; ++t) {
            float v[4];
            for (int j = 0; j < 4; ++j) {
                v[j] = acc[i][t][j] + bias[i][j];
                if (a.relu && v[j] < 0.0f) v[j] = 0.0f;
            }
            out.o[i][t] = make_uint2(f2bf2(v[0], v[1]), f2bf2(v[2], v[3]));
        }
    if constexpr (POOL) {
        for (int half = 0; half < 2; ++half)
            for (int i = 0; i < CT; ++i) {
                const uint2 u0 = out.o[i][half], u1 = out.o[i][half + 2];
                const uint32_t n0x = uint32_t(__shfl_xor(int(u0.x), 1)), n0y = uint32_t(__shfl_xor(int(u0.y), 1));
                const uint32_t n1x = uint32_t(__shfl_xor(int(u1.x), 1)), n1y = uint32_t(__shfl_xor(int(u1.y), 1));
                const uint32_t w0[2] = {u0.x, u0.y}, w1[2] = {n0x, n0y}, w2[2] = {u1.x, u1.y}, w3[2] = {n1x, n1y};
                float o[4];
                for (int j = 0; j < 4; ++j) {
                    const int sh = 16 * (j & 1);
                    const float s = (bf2f(bf16_t(w0[j >> 1] >> sh)) + bf2f(bf16_t(w1[j >> 1] >> sh))) +
                                    (bf2f(bf16_t(w2[j >> 1] >> sh)) + bf2f(bf16_t(w3[j >> 1] >> sh)));
                    o[j] = s * 0.25f;
                }
                out.p[i][half] = make_uint2(f2bf2(o[0], o[1]), f2bf2(o[2], o[3]));
            }
    }
}
template <int CT, bool POOL>
__device__ __forceinline__ void conv_tile_store(const ConvArgs& a, const TileOut<CT, POOL>& out, int img, int oy, int ox, int wv,
                                                int g, int r, int co0) {
    for (int t = 0; t < 4; ++t) {
        const int y = oy + 2 * wv + (t >> 1), x = ox + 16 * (t & 1) + r;
        if (y >= a.h || x >= a.wd) continue;
        const size_t p = size_t(img * a.h + y) * a.wd + x;
        for (int i = 0; i < CT; ++i) *reinterpret_cast<uint2*>(a.out + p * a.cout + co0 + i * 16 + 4 * g) = out.o[i][t];
    }
    if constexpr (POOL) {
        const int ph = a.h >> 1, pw = a.wd >> 1;
        const int y = oy + 2 * wv;
        for (int half = 0; half < 2; ++half) {
            const int x = ox + 16 * half + r;
            if ((r & 1) || y + 1 >= a.h || x + 1 >= a.wd) continue;
            const size_t q = size_t(img * ph + (y >> 1)) * pw + (x >> 1);
            for (int i = 0; i < CT; ++i) *reinterpret_cast<uint2*>(a.pool + q * a.cout + co0 + i * 16 + 4 * g) = out.p[i][half];
        }
    }
}
template <int CT, bool POOL>
__device__ __forceinline__ void conv_tile_out(const ConvArgs& a, const f32x4 (&acc)[CT][4], int img, int oy, int ox, int wv,
                                              int g, int r, int co0) {
    float bias[CT][4];
    conv_lane_bias<CT>(a, co0, g, bias);
    TileOut<CT, POOL> out;
    conv_tile_pack<CT, POOL>(a, acc, bias, out);
    conv_tile_store<CT, POOL>(a, out, img, oy, ox, wv, g, r, co0);
}
constexpr int LH_PS = 40;            // bf16 a staged pixel: 32 channels + 8 pad
constexpr int LW_RS = 9 * 32 + 8;    // bf16 a staged weight row: 9 taps x 32 channels + 8 pad
template <int CT, bool POOL = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_conv_lds(ConvArgs a, int ntiles) {
    __shared__ __attribute__((aligned(16))) bf16_t hal[LH_H * LH_W * LH_PS];
    __shared__ __attribute__((aligned(16))) bf16_t wls[CT * 16 * LW_RS];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int g = lane >> 4, r = lane & 15;
    const int tiles_x = (a.wd + LT_W - 1) / LT_W, tiles_y = (a.h + LT_H - 1) / LT_H;
    const int tile = int(blockIdx.x) % ntiles, co0 = a.co0 + (int(blockIdx.x) / ntiles) * CT * 16;
    const int bx = tile % tiles_x, q0 = tile / tiles_x;
    const int by = q0 % tiles_y, img = q0 / tiles_y;
    const int oy = by * LT_H, ox = bx * LT_W;
    const int cin = a.c0 + a.c1, hh = a.h >> 1, hw = a.wd >> 1;
    f32x4 acc[CT][4];
    for (int i = 0; i < CT; ++i)
        for (int t = 0; t < 4; ++t) acc[i][t] = f32x4{0, 0, 0, 0};
    for (int c = 0; c < cin; c += 32) {
        __syncthreads();
#pragma unroll
        for (int k = 0; k < (LH_H * LH_W * 4 + 255) / 256; ++k) {
            const int e = tid + 256 * k;
            if (e < LH_H * LH_W * 4) {
                const int hp = e >> 2, q = e & 3;
                const int gy = oy + hp / LH_W - 1, gx = ox + hp % LH_W - 1;
                uint4 v = make_uint4(0, 0, 0, 0);
                if (gy >= 0 && gy < a.h && gx >= 0 && gx < a.wd) {
                    const int ch = c + 8 * q;
                    const bf16_t* src = ch < a.c0 ? a.in0 + (size_t(img * a.h + gy) * a.wd + gx) * a.c0 + ch
                                                  : a.in1 + (size_t(img * hh + (gy >> 1)) * hw + (gx >> 1)) * a.c1 + (ch - a.c0);
                    v = *reinterpret_cast<const uint4*>(src);
                }
                *reinterpret_cast<uint4*>(&hal[hp * LH_PS + 8 * q]) = v;
            }
        }
#pragma unroll
        for (int k = 0; k < (CT * 16 * 36 + 255) / 256; ++k) {
            const int e = tid + 256 * k;
            if (e < CT * 16 * 36) {
                const int co = e / 36, rem = e - co * 36, tap = rem >> 2, q = rem & 3;
                *reinterpret_cast<uint4*>(&wls[co * LW_RS + tap * 32 + 8 * q]) =
                    *reinterpret_cast<const uint4*>(a.w + size_t(co0 + co) * a.kpad + tap * cin + c + 8 * q);
            }
        }
        __syncthreads();
        for (int tap = 0; tap < 9; ++tap) {
            const int dy = tap / 3, dx = tap % 3;
            V8 bv[4];
            for (int t = 0; t < 4; ++t) {
                const int hp = (2 * wv + (t >> 1) + dy) * LH_W + 16 * (t & 1) + r + dx;
                bv[t].u = *reinterpret_cast<const uint4*>(&hal[hp * LH_PS + 8 * g]);
            }
            for (int i = 0; i < CT; ++i) {
                V8 av;
                av.u = *reinterpret_cast<const uint4*>(&wls[(i * 16 + r) * LW_RS + tap * 32 + 8 * g]);
                for (int t = 0; t < 4; ++t) acc[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av.v, bv[t].v, acc[i][t], 0, 0, 0);
            }
        }
    }
    conv_tile_out<CT, POOL>(a, acc, img, oy, ox, wv, g, r, co0);
}

// The persistent form of k_conv_lds for layers of one 32-channel chunk in and one channel
// group out (cin = 32, cout = CT x 16): the weights are staged once; block b takes tiles b,
// b + gridDim.x, ..., the next tile's halo loaded into registers while the current tile's
// MFMAs run. The staging layout, the fragments and the sum order are k_conv_lds's, so the
// outputs are the same bit for bit. (Measured: the same form for layers of several chunks,
// with each chunk's weights prefetched too, and for 64 outputs ran slower than k_conv_lds.)
constexpr int PF_BLOCKS = 3;   // k_conv_pf<2> blocks a CU: 46 KB of LDS each
template <int CT, bool POOL = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) void k_conv_pf(ConvArgs a, int ntiles) {
    __shared__ __attribute__((aligned(16))) bf16_t hal[LH_H * LH_W * LH_PS];
    __shared__ __attribute__((aligned(16))) bf16_t wls[CT * 16 * LW_RS];
    constexpr int HV = (LH_H * LH_W * 4 + 255) / 256;   // halo pieces (8 channels) a thread
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int g = lane >> 4, r = lane & 15;
    const int tiles_x = (a.wd + LT_W - 1) / LT_W, tiles_y = (a.h + LT_H - 1) / LT_H;
    int tile = int(blockIdx.x);
    if (tile >= ntiles) return;
    for (int e = tid; e < CT * 16 * 36; e += 256) {
        const int co = e / 36, rem = e - co * 36, tap = rem >> 2, q = rem & 3;
        *reinterpret_cast<uint4*>(&wls[co * LW_RS + tap * 32 + 8 * q]) =
            *reinterpret_cast<const uint4*>(a.w + size_t(a.co0 + co) * a.kpad + tap * 32 + 8 * q);
    }
    uint4 hv[HV];
    const auto load = [&](int tl) {
        const int bx = tl % tiles_x, q0 = tl / tiles_x, by = q0 % tiles_y, img = q0 / tiles_y;
#pragma unroll
        for (int k = 0; k < HV; ++k) {
            const int e = tid + 256 * k;
            hv[k] = make_uint4(0, 0, 0, 0);
            if (e >= LH_H * LH_W * 4) continue;
            const int hp = e >> 2, q = e & 3;
            const int gy = by * LT_H + hp / LH_W - 1, gx = bx * LT_W + hp % LH_W - 1;
            if (gy < 0 || gy >= a.h || gx < 0 || gx >= a.wd) continue;
            hv[k] = *reinterpret_cast<const uint4*>(a.in0 + (size_t(img * a.h + gy) * a.wd + gx) * 32 + 8 * q);
        }
    };
    load(tile);
    float bias[CT][4];
    conv_lane_bias<CT>(a, a.co0, g, bias);
    TileOut<CT, POOL> done;   // the previous tile's outputs, stored behind the next tile's loads
    int dtile = -1;
    for (;;) {
        __syncthreads();   // every wave is done with the previous tile's halo
#pragma unroll
        for (int k = 0; k < HV; ++k) {
            const int e = tid + 256 * k;
            if (e < LH_H * LH_W * 4) *reinterpret_cast<uint4*>(&hal[(e >> 2) * LH_PS + 8 * (e & 3)]) = hv[k];
        }
        __syncthreads();
        const int ntile = tile + int(gridDim.x);
        if (ntile < ntiles) load(ntile);   // in flight during this tile's MFMAs
        if (dtile >= 0) {
            const int bx = dtile % tiles_x, q0 = dtile / tiles_x, by = q0 % tiles_y, img = q0 / tiles_y;
            conv_tile_store<CT, POOL>(a, done, img, by * LT_H, bx * LT_W, wv, g, r, a.co0);
        }
        f32x4 acc[CT][4];
        for (int i = 0; i < CT; ++i)
            for (int t = 0; t < 4; ++t) acc[i][t] = f32x4{0, 0, 0, 0};
        for (int tap = 0; tap < 9; ++tap) {
            const int dy = tap / 3, dx = tap % 3;
            V8 bv[4];
            for (int t = 0; t < 4; ++t) {
                const int hp = (2 * wv + (t >> 1) + dy) * LH_W + 16 * (t & 1) + r + dx;
                bv[t].u = *reinterpret_cast<const uint4*>(&hal[hp * LH_PS + 8 * g]);
            }
            for (int i = 0; i < CT; ++i) {
                V8 av;
                av.u = *reinterpret_cast<const uint4*>(&wls[(i * 16 + r) * LW_RS + tap * 32 + 8 * g]);
                for (int t = 0; t < 4; ++t) acc[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av.v, bv[t].v, acc[i][t], 0, 0, 0);
            }
        }
        conv_tile_pack<CT, POOL>(a, acc, bias, done);
        dtile = tile;
        if (ntile >= ntiles) break;
        tile = ntile;
    }
    const int bx = dtile % tiles_x, q0 = dtile / tiles_x, by = q0 % tiles_y, img = q0 / tiles_y;
    conv_tile_store<CT, POOL>(a, done, img, by * LT_H, bx * LT_W, wv, g, r, a.co0);
}

// The persistent, double-buffered form of k_conv_lds for any layer of 32-channel chunks: a
// block of WV waves takes work items (a 2 WV x 32 pixel tile, one group of CT*16 output
// channels) b, b + gridDim.x, ...; the work item's chunks, then the next item's, form one
// sequence of stages, and stage s + 1's input halo and weights are copied from global memory
// straight into the other LDS buffer (global_load_lds_dwordx4: no registers, no LDS writes by
// the waves) while stage s's MFMAs run. A copy instruction fills 64 consecutive 16-byte LDS
// slots, one a lane, so the lanes choose the global pieces: the halo takes 5 slots a pixel
// (its 4 pieces of 8 channels and a pad slot, the 80-byte pixel stride of k_conv_lds) and the
// weights 37 a row (36 pieces and a pad), then the group's biases; pad slots and pixels outside
// the image copy a zeroed global piece. Everything the MFMAs and the epilogue read comes from
// LDS: the vector memory counter retires in order, so waiting on any global load during a stage
// would also wait for the next stage's copies. The fragments, the chunk-major sum and the
// epilogue are k_conv_lds's, so the outputs are the same bit for bit.
__device__ __attribute__((aligned(16))) uint4 g_zero16[1];   // static device memory starts zeroed
typedef __attribute__((address_space(3))) void lds_void;
constexpr int DL_WAVES = 8;   // a block: 16 x 32 pixels (4 waves a block were slower: one block a CU either way)
template <int CT, int WV, bool POOL = false>
__global__ __launch_bounds__(64 * WV) void k_conv_dl(ConvArgs a, int ntiles, int items) {
    constexpr int TH = 2 * WV, HH = TH + 2, HW = LT_W + 2;
    constexpr int NH = HH * HW * 5, NW = CT * 16 * 37 + CT * 4;   // 16-byte slots: halo, weights + biases
    __shared__ uint4 h0[NH], h1[NH], w0[NW], w1[NW];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int g = lane >> 4, r = lane & 15;
    const int tiles_x = (a.wd + LT_W - 1) / LT_W, tiles_y = (a.h + TH - 1) / TH;
    const int cin = a.c0 + a.c1, hh = a.h >> 1, hw = a.wd >> 1, nch = cin / 32;
    const int step = int(gridDim.x);
    int item = int(blockIdx.x);
    if (item >= items) return;
    struct Where { int img, oy, ox, co0; };
    const auto where = [&](int it) __attribute__((always_inline)) {
        const int tile = it % ntiles, q0 = (tile / tiles_x);
        return Where{q0 / tiles_y, (q0 % tiles_y) * TH, (tile % tiles_x) * LT_W, a.co0 + (it / ntiles) * CT * 16};
    };
    const auto issue = [&](uint4* hb, uint4* wb, int it, int ch) __attribute__((always_inline)) {
        const Where p = where(it);
        const int c = 32 * ch;
#pragma unroll
        for (int k = 0; k < (NH + 64 * WV - 1) / (64 * WV); ++k) {
            const int base = (k * WV + wv) * 64, s = base + lane;
            if (base >= NH) break;
            if (s < NH) {
                const int hp = s / 5, q = s - 5 * hp;
                const int gy = p.oy + hp / HW - 1, gx = p.ox + hp % HW - 1;
                const void* src = g_zero16;
                if (q < 4 && gy >= 0 && gy < a.h && gx >= 0 && gx < a.wd) {
                    const int cc = c + 8 * q;
                    src = cc < a.c0 ? a.in0 + (size_t(p.img * a.h + gy) * a.wd + gx) * a.c0 + cc
                                    : a.in1 + (size_t(p.img * hh + (gy >> 1)) * hw + (gx >> 1)) * a.c1 + (cc - a.c0);
                }
                __builtin_amdgcn_global_load_lds(src, (lds_void*)(hb + base), 16, 0, 0);
            }
        }
#pragma unroll
        for (int k = 0; k < (NW + 64 * WV - 1) / (64 * WV); ++k) {
            const int base = (k * WV + wv) * 64, s = base + lane;
            if (base >= NW) break;
            if (s < NW) {
                const int row = s / 37, pc = s - row * 37;
                const void* src = row >= CT * 16 ? (a.bias ? static_cast<const void*>(a.bias + p.co0 + 4 * (s - CT * 16 * 37))
                                                           : static_cast<const void*>(g_zero16))
                                  : pc < 36 ? a.w + size_t(p.co0 + row) * a.kpad + (pc >> 2) * cin + c + 8 * (pc & 3)
                                            : static_cast<const void*>(g_zero16);
                __builtin_amdgcn_global_load_lds(src, (lds_void*)(wb + base), 16, 0, 0);
            }
        }
    };
    f32x4 acc[CT][4];
    TileOut<CT, POOL> done;   // a finished item's outputs, stored during the next stage
    Where dw{};
    bool pending = false;
    int ch = 0;
    // one stage: the current chunk from (hc, wc), the next stage's copies into (hn, wn);
    // false once the block has no work left
    const auto stage = [&](const uint4* hc, const uint4* wc, uint4* hn, uint4* wn) __attribute__((always_inline)) {
        __builtin_amdgcn_s_waitcnt(0);   // this wave's copies into (hc, wc) have landed
        __syncthreads();                 // everyone's have, and nobody still reads (hn, wn)
        const bool last = ch + 1 == nch;
        const int nit = last ? item + step : item, nc = last ? 0 : ch + 1;
        if (nit < items) issue(hn, wn, nit, nc);
        if (pending) {   // stored behind this stage's copies: the next stage's wait finds them done
            conv_tile_store<CT, POOL>(a, done, dw.img, dw.oy, dw.ox, wv, g, r, dw.co0);
            pending = false;
        }
        if (ch == 0)
            for (int i = 0; i < CT; ++i)
                for (int t = 0; t < 4; ++t) acc[i][t] = f32x4{0, 0, 0, 0};
        V8 bv[2][4], av[2][CT];
        const auto frag = [&](int tap, V8 (&b)[4], V8 (&w)[CT]) __attribute__((always_inline)) {
            const int dy = tap / 3, dx = tap % 3;
            for (int t = 0; t < 4; ++t) b[t].u = hc[((2 * wv + (t >> 1) + dy) * HW + 16 * (t & 1) + r + dx) * 5 + g];
            for (int i = 0; i < CT; ++i) w[i].u = wc[(i * 16 + r) * 37 + tap * 4 + g];
        };
        frag(0, bv[0], av[0]);
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
            if (tap < 8) frag(tap + 1, bv[(tap + 1) & 1], av[(tap + 1) & 1]);
            for (int i = 0; i < CT; ++i)
                for (int t = 0; t < 4; ++t)
                    acc[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[tap & 1][i].v, bv[tap & 1][t].v, acc[i][t], 0, 0, 0);
        }
        if (last) {
            const float* lb = reinterpret_cast<const float*>(wc + CT * 16 * 37);   // the group's biases, staged
            float bias[CT][4];
            for (int i = 0; i < CT; ++i)
                for (int j = 0; j < 4; ++j) bias[i][j] = a.bias ? lb[i * 16 + 4 * g + j] : 0.0f;
            conv_tile_pack<CT, POOL>(a, acc, bias, done);
            dw = where(item);
            pending = true;
        }
        item = nit;
        ch = nc;
        return nit < items;
    };
    issue(h0, w0, item, 0);
    for (;;) {
        if (!stage(h0, w0, h1, w1)) break;
        if (!stage(h1, w1, h0, w0)) break;
    }
    if (pending) conv_tile_store<CT, POOL>(a, done, dw.img, dw.oy, dw.ox, wv, g, r, dw.co0);
}

// The first layer's convolution (8 input channels, K = 72 padded to 96): one 32-wide k step
// is 4 taps x 8 channels, so lane group g of a fragment reads tap 4*step + g, shifted, from
// the staged 10 x 34 pixel tile (16 B a pixel). The 3 steps sum k in k_conv's order, so the
// outputs equal k_conv's. Persistent over tiles, the next tile staged through registers
// while the current one runs; the weights (cout x 96) stay in LDS.
constexpr int C8_PS = 8;             // bf16 a staged pixel
template <int CT>
__global__ __launch_bounds__(256) void k_conv_c8(ConvArgs a, int ntiles) {
    __shared__ __attribute__((aligned(16))) bf16_t hal[LH_H * LH_W * C8_PS];
    __shared__ __attribute__((aligned(16))) bf16_t wls[CT * 16 * 104];
    constexpr int HV = (LH_H * LH_W + 255) / 256;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int g = lane >> 4, r = lane & 15;
    const int tiles_x = (a.wd + LT_W - 1) / LT_W, tiles_y = (a.h + LT_H - 1) / LT_H;
    int tile = int(blockIdx.x);
    if (tile >= ntiles) return;
    for (int e = tid; e < CT * 16 * 12; e += 256) {   // 12 pieces of 8 k a row (96 k)
        const int co = e / 12, q = e - co * 12;
        *reinterpret_cast<uint4*>(&wls[co * 104 + 8 * q]) = *reinterpret_cast<const uint4*>(a.w + size_t(a.co0 + co) * a.kpad + 8 * q);
    }
    uint4 hv[HV];
    const auto load_halo = [&](int tl) {
        const int bx = tl % tiles_x, q0 = tl / tiles_x, by = q0 % tiles_y, img = q0 / tiles_y;
        for (int k = 0; k < HV; ++k) {
            const int hp = tid + 256 * k;
            hv[k] = make_uint4(0, 0, 0, 0);
            if (hp >= LH_H * LH_W) continue;
            const int gy = by * LT_H + hp / LH_W - 1, gx = bx * LT_W + hp % LH_W - 1;
            if (gy < 0 || gy >= a.h || gx < 0 || gx >= a.wd) continue;
            hv[k] = *reinterpret_cast<const uint4*>(a.in0 + (size_t(img * a.h + gy) * a.wd + gx) * 8);
        }
    };
    const auto store_halo = [&]() {
        for (int k = 0; k < HV; ++k) {
            const int hp = tid + 256 * k;
            if (hp < LH_H * LH_W) *reinterpret_cast<uint4*>(&hal[hp * C8_PS]) = hv[k];
        }
    };
    load_halo(tile);
    store_halo();
    __syncthreads();
    for (;;) {
        const int ntile = tile + int(gridDim.x);
        if (ntile < ntiles) load_halo(ntile);
        f32x4 acc[CT][4];
        for (int i = 0; i < CT; ++i)
            for (int t = 0; t < 4; ++t) acc[i][t] = f32x4{0, 0, 0, 0};
        for (int step = 0; step < 3; ++step) {
            const int tap = 4 * step + g;
            const int dy = tap < 9 ? tap / 3 : 0, dx = tap < 9 ? tap % 3 : 0;
            V8 bv[4];
            for (int t = 0; t < 4; ++t) {
                const int hp = (2 * wv + (t >> 1) + dy) * LH_W + 16 * (t & 1) + r + dx;
                bv[t].u = tap < 9 ? *reinterpret_cast<const uint4*>(&hal[hp * C8_PS]) : make_uint4(0, 0, 0, 0);
            }
            for (int i = 0; i < CT; ++i) {
                V8 av;
                av.u = *reinterpret_cast<const uint4*>(&wls[(i * 16 + r) * 104 + step * 32 + 8 * g]);
                for (int t = 0; t < 4; ++t) acc[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av.v, bv[t].v, acc[i][t], 0, 0, 0);
            }
        }
        const int bx = tile % tiles_x, q0 = tile / tiles_x, by = q0 % tiles_y, img = q0 / tiles_y;
        for (int t = 0; t < 4; ++t) {
            const int y = by * LT_H + 2 * wv + (t >> 1), x = bx * LT_W + 16 * (t & 1) + r;
            if (y >= a.h || x >= a.wd) continue;
            const size_t p = size_t(img * a.h + y) * a.wd + x;
            for (int i = 0; i < CT; ++i) {
                const int co = a.co0 + i * 16 + 4 * g;
                float v[4];
                for (int j = 0; j < 4; ++j) {
                    v[j] = acc[i][t][j] + (a.bias ? a.bias[co + j] : 0.0f);
                    if (a.relu && v[j] < 0.0f) v[j] = 0.0f;
                }
                const uint32_t lo = uint32_t(f2bf(v[0])) | (uint32_t(f2bf(v[1])) << 16);
                const uint32_t hi = uint32_t(f2bf(v[2])) | (uint32_t(f2bf(v[3])) << 16);
                *reinterpret_cast<uint2*>(a.out + p * a.cout + co) = make_uint2(lo, hi);
            }
        }
        if (ntile >= ntiles) break;
        __syncthreads();
        store_halo();
        __syncthreads();
        tile = ntile;
    }
}

// ---- weight gradient ----------------------------------------------------------------
// dW[co][k] = sum_p dy[p][co] * X[p][k], X the forward's input patches (k < K) and 1 at
// k == K (the bias column). A block: one 64-wide k tile, all cout (<= 128), a chunk of
// pixels; its 4 waves take 32 pixels each per step, then add their tiles in LDS in wave
// order; the block's partial goes to part[chunk][cout][kpg] (kpg = K + 1 rounded up to 64).
struct WgradArgs {
    const bf16_t* dy;    // N,H,W,Cout
    const bf16_t* in0;
    const bf16_t* in1;
    float* part;
    int n, h, wd, c0, c1, cout, kpg, chunk;
};
template <int CT>
__global__ __launch_bounds__(256) void k_wgrad(WgradArgs a) {
    __shared__ bf16_t sa[4][CT * 16][40];   // per wave: [cout][32 pixels] (+8 pad)
    __shared__ bf16_t sb[4][64][40];        // per wave: [k][32 pixels] (+8 pad)
    __shared__ float red[CT * 16 * 64];     // the four waves' sums
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int g = lane >> 4, r = lane & 15;
    const int ktile = blockIdx.x, chunk = blockIdx.y;
    const int npix = a.n * a.h * a.wd;
    const int cin = a.c0 + a.c1, K = 9 * cin;
    const int hh = a.h >> 1, hw = a.wd >> 1;
    f32x4 acc[CT][4];
    for (int i = 0; i < CT; ++i)
        for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0, 0, 0, 0};
    const int p0 = chunk * a.chunk, p1 = min(npix, p0 + a.chunk);
    for (int pb0 = p0; pb0 < p1; pb0 += 4 * 32) {
        const int pb = pb0 + wv * 32;
        // dy: 32 pixels x (CT*16) channels -> sa[wv][co][p]
        for (int e = lane; e < 32 * CT * 2; e += 64) {   // units of 8 channels
            const int pl = e / (CT * 2), cg = e - pl * (CT * 2);
            const int p = pb + pl;
            V8 v;
            v.u = make_uint4(0, 0, 0, 0);
            if (p < p1) v.u = *reinterpret_cast<const uint4*>(a.dy + size_t(p) * a.cout + cg * 8);
            const bf16_t* s = reinterpret_cast<const bf16_t*>(&v.u);
            for (int j = 0; j < 8; ++j) sa[wv][cg * 8 + j][pl] = s[j];
        }
        // X patches: 32 pixels x 64 k -> sb[wv][k][p]
        for (int e = lane; e < 32 * 8; e += 64) {          // units of 8 k
            const int pl = e >> 3, kg = e & 7;
            const int p = pb + pl;
            const int k = ktile * 64 + kg * 8;
            V8 v;
            v.u = make_uint4(0, 0, 0, 0);
            if (p < p1) {
                if (k < K) {
                    const int tap = k / cin, ci = k - tap * cin;
                    const int x0 = p % a.wd, q = p / a.wd, y0 = q % a.h, nn = q / a.h;
                    const int y = y0 + tap / 3 - 1, x = x0 + tap % 3 - 1;
                    if (y >= 0 && y < a.h && x >= 0 && x < a.wd) {
                        const bf16_t* src = ci < a.c0 ? a.in0 + (size_t(nn * a.h + y) * a.wd + x) * a.c0 + ci
                                                      : a.in1 + (size_t(nn * hh + (y >> 1)) * hw + (x >> 1)) * a.c1 + (ci - a.c0);
                        v.u = *reinterpret_cast<const uint4*>(src);
                    }
                } else if (k == K) {
                    v.u.x = 0x3f80u;   // bf16 1.0: the bias column
                }
            }
            const bf16_t* s = reinterpret_cast<const bf16_t*>(&v.u);
            for (int j = 0; j < 8; ++j) sb[wv][kg * 8 + j][pl] = s[j];
        }
        __syncthreads();
        // A = dy^T (cout x 32 pixels), B = X (32 pixels x 64 k)
        V8 bv[4];
        for (int j = 0; j < 4; ++j) bv[j].u = *reinterpret_cast<const uint4*>(&sb[wv][j * 16 + r][8 * g]);
        for (int i = 0; i < CT; ++i) {
            V8 av;
            av.u = *reinterpret_cast<const uint4*>(&sa[wv][i * 16 + r][8 * g]);
            for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av.v, bv[j].v, acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    for (int w = 0; w < 4; ++w) {
        if (wv == w)
            for (int i = 0; i < CT; ++i)
                for (int j = 0; j < 4; ++j)
                    for (int q = 0; q < 4; ++q) {
                        const int co = i * 16 + 4 * g + q, kk = j * 16 + r;
                        if (w == 0) red[co * 64 + kk] = acc[i][j][q];
                        else red[co * 64 + kk] += acc[i][j][q];
                    }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < CT * 16 * 64; e += 256) {
        const int co = e / 64, kk = e - co * 64;
        a.part[(size_t(chunk) * a.cout + co) * a.kpg + ktile * 64 + kk] = red[e];
    }
}

// The chunks' partials summed in a fixed order, two levels so every element's sum runs in
// parallel: group gi of `per` consecutive chunks, in chunk order, to part2[gi][cout][kpg];
// then the groups in group order to grad[co][k] (the bias grad at k == K).
__global__ void k_wgrad_sum_groups(const float* __restrict__ part, int nchunks, int per, int cout, int kpg, int K,
                                   float* __restrict__ part2) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x, gi = blockIdx.y;
    if (e >= cout * (K + 1)) return;
    const int co = e / (K + 1), k = e - co * (K + 1);
    const int c0 = gi * per, c1 = min(nchunks, c0 + per);
    float s = 0.0f;
    for (int c = c0; c < c1; ++c) s += part[(size_t(c) * cout + co) * kpg + k];
    part2[(size_t(gi) * cout + co) * kpg + k] = s;
}
__global__ void k_wgrad_reduce(const float* __restrict__ part2, int ngroups, int cout, int kpg, int K, float* __restrict__ gw,
                               float* __restrict__ gb) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= cout * (K + 1)) return;
    const int co = e / (K + 1), k = e - co * (K + 1);
    float s = 0.0f;
    for (int gi = 0; gi < ngroups; ++gi) s += part2[(size_t(gi) * cout + co) * kpg + k];
    if (k < K) gw[size_t(co) * K + k] = s;
    else gb[co] = s;
}

// ---- elementwise pieces ---------------------------------------------------------------
__global__ void k_relu_bwd(const bf16_t* __restrict__ y, const bf16_t* __restrict__ dy, bf16_t* __restrict__ dpre, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dpre[i] = bf2f(y[i]) > 0.0f ? dy[i] : bf16_t(0);
}
// 2x2 average pool of an h x w level (h, w even)
__global__ void k_pool_fwd(const bf16_t* __restrict__ in, bf16_t* __restrict__ out, int n, int h, int w, int c) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int ho = h / 2, wo = w / 2;
    if (i >= size_t(n) * ho * wo * c) return;
    const int ch = int(i % c);
    const size_t q = i / c;
    const int x = int(q % wo), y = int((q / wo) % ho), nn = int(q / (size_t(wo) * ho));
    const size_t b = (size_t(nn * h + 2 * y) * w + 2 * x) * c + ch;
    const float s = (bf2f(in[b]) + bf2f(in[b + c])) + (bf2f(in[b + size_t(w) * c]) + bf2f(in[b + size_t(w) * c + c]));
    out[i] = f2bf(s * 0.25f);
}
// grad of an encoder output that feeds both a skip and a pool:
// d0 = dx[..., :c0] (the skip half of a decoder's input grad) + 0.25 * dpool (upsampled).
__global__ void k_skip_grad(const bf16_t* __restrict__ dx, int cx, int c0, const bf16_t* __restrict__ dpool,
                            bf16_t* __restrict__ d0, int n, int h, int w) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(n) * h * w * c0) return;
    const int ch = int(i % c0);
    const size_t p = i / c0;
    const int x = int(p % w), y = int((p / w) % h), nn = int(p / (size_t(w) * h));
    const float v = bf2f(dx[p * cx + ch]) + 0.25f * bf2f(dpool[(size_t(nn * (h / 2) + y / 2) * (w / 2) + x / 2) * c0 + ch]);
    d0[i] = f2bf(v);
}
// grad of the lower level read through the 2x upsampling: the sum of its 4 children.
__global__ void k_up_grad(const bf16_t* __restrict__ dx, int cx, int c0, int c1, bf16_t* __restrict__ dlow, int n, int h, int w) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int ho = h / 2, wo = w / 2;
    if (i >= size_t(n) * ho * wo * c1) return;
    const int ch = int(i % c1);
    const size_t q = i / c1;
    const int x = int(q % wo), y = int((q / wo) % ho), nn = int(q / (size_t(wo) * ho));
    const size_t b = (size_t(nn * h + 2 * y) * w + 2 * x) * cx + c0 + ch;
    const float s = (bf2f(dx[b]) + bf2f(dx[b + cx])) + (bf2f(dx[b + size_t(w) * cx]) + bf2f(dx[b + size_t(w) * cx + cx]));
    dlow[i] = f2bf(s);
}

// The feature channels of one image (kNetFeatures bf16 a pixel).
__global__ void k_features(const float* __restrict__ pos, const float* __restrict__ nrm, const uint8_t* __restrict__ mask,
                           const float* __restrict__ ssao, const float* __restrict__ noisy, float ex, float ey, float ez,
                           size_t np, bf16_t* __restrict__ x) {
    const size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= np) return;
    float f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (mask[p]) {
        const float dx = pos[3 * p] - ex, dy = pos[3 * p + 1] - ey, dz = pos[3 * p + 2] - ez;
        const float d = sqrtf(dx * dx + dy * dy + dz * dz);
        const float nx = nrm[3 * p], ny = nrm[3 * p + 1], nz = nrm[3 * p + 2];
        const float l1 = fabsf(nx) + fabsf(ny) + fabsf(nz);
        float ox = l1 > 0.0f ? nx / l1 : 0.0f, oy = l1 > 0.0f ? nz / l1 : 0.0f;   // octahedral about +y
        if (ny < 0.0f) {
            const float tx = (1.0f - fabsf(oy)) * (ox >= 0.0f ? 1.0f : -1.0f);
            const float ty = (1.0f - fabsf(ox)) * (oy >= 0.0f ? 1.0f : -1.0f);
            ox = tx;
            oy = ty;
        }
        f[0] = 1.0f;
        f[1] = d * 0.05f;
        f[2] = ox;
        f[3] = oy;
        f[4] = ssao[p];
        f[5] = noisy ? noisy[p] : 0.0f;
        f[6] = ny;
        f[7] = noisy ? 1.0f : 0.0f;
    }
    uint32_t pk[4];
    for (int j = 0; j < 4; ++j) pk[j] = uint32_t(f2bf(f[2 * j])) | (uint32_t(f2bf(f[2 * j + 1])) << 16);
    *reinterpret_cast<uint4*>(x + p * 8) = make_uint4(pk[0], pk[1], pk[2], pk[3]);
}

// head: pred = b + sum_c w[c] x[p][c] over 32 channels in channel order; ao = clamp(pred,
// 0, 1) on covered pixels and 1 elsewhere (the reference's value for uncovered pixels).
__global__ void k_head(const bf16_t* __restrict__ x, const float* __restrict__ hw, const uint8_t* __restrict__ mask, size_t np,
                       float* __restrict__ pred, float* __restrict__ ao) {
    const size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const uint4* xv = reinterpret_cast<const uint4*>(x + p * 32);
    float s = hw[32];
    for (int q = 0; q < 4; ++q) {
        const uint4 u = xv[q];
        const uint32_t wds[4] = {u.x, u.y, u.z, u.w};
        for (int j = 0; j < 4; ++j) {
            s += hw[8 * q + 2 * j] * __uint_as_float(wds[j] << 16);
            s += hw[8 * q + 2 * j + 1] * __uint_as_float(wds[j] & 0xffff0000u);
        }
    }
    if (pred) pred[p] = s;
    if (ao) ao[p] = mask[p] ? fminf(fmaxf(s, 0.0f), 1.0f) : 1.0f;
}

__global__ void k_zero_u32(uint32_t* p) { p[threadIdx.x] = 0; }
__global__ __launch_bounds__(256) void k_mask_count(const uint8_t* __restrict__ mask, size_t np, uint32_t* cnt) {
    __shared__ uint32_t s;
    if (threadIdx.x == 0) s = 0;
    __syncthreads();
    uint32_t c = 0;
    for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < np; i += size_t(gridDim.x) * 256) c += mask[i] != 0;
    atomicAdd(&s, c);
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(cnt, s);
}

// Mean squared error over covered pixels: dL/dpred = 2 (pred - target) / count; the head's
// input grad dx[p][c] = dL/dpred * w[c]; per-block sums of the head's weight and bias grads
// and of the loss.
__global__ __launch_bounds__(256) void k_head_bwd(const bf16_t* __restrict__ x, const float* __restrict__ hw,
                                                  const float* __restrict__ pred, const float* __restrict__ target,
                                                  const uint8_t* __restrict__ mask, size_t np, const uint32_t* __restrict__ cnt,
                                                  bf16_t* __restrict__ dx, float* __restrict__ part) {
    __shared__ float red[34][256];
    const size_t p = size_t(blockIdx.x) * 256 + threadIdx.x;
    const float inv = 1.0f / float(max(*cnt, 1u));
    float g = 0.0f, l = 0.0f;
    if (p < np && mask[p]) {
        const float e = pred[p] - target[p];
        l = e * e * inv;
        g = 2.0f * e * inv;
    }
    for (int c = 0; c < 32; ++c) {
        const float xv = p < np ? bf2f(x[p * 32 + c]) : 0.0f;
        if (p < np) dx[p * 32 + c] = f2bf(g * hw[c]);
        red[c][threadIdx.x] = g * xv;
    }
    red[32][threadIdx.x] = g;
    red[33][threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s)
            for (int c = 0; c < 34; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 34) part[size_t(blockIdx.x) * 34 + threadIdx.x] = red[threadIdx.x][0];
}
// One block a column (32 head weights, the bias, the loss): strided partial sums, then an
// LDS tree, in a fixed order.
__global__ __launch_bounds__(256) void k_head_reduce(const float* __restrict__ part, int nblk, float* __restrict__ ghead,
                                                     float* __restrict__ loss) {
    __shared__ float red[256];
    const int c = blockIdx.x;
    float s = 0.0f;
    for (int b = threadIdx.x; b < nblk; b += 256) s += part[size_t(b) * 34 + c];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (int(threadIdx.x) < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (c < 33) ghead[c] = red[0];
        else *loss = red[0];
    }
}

// Adam (Kingma and Ba, 2015) on the fp32 master parameters; c1, c2 the bias corrections.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v, size_t n,
                       float lr, float b1, float b2, float eps, float c1, float c2) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// fp32 master [Cout][K] -> bf16 forward weights [Cout][Kpad] and input-grad weights
// [Cin][Kpad_t] with wt[ci][t * Cout + co] = w[co][(8 - t) * Cin + ci].
__global__ void k_pack(const float* __restrict__ m, int cout, int cin, int kpad, int kpad_t, bf16_t* __restrict__ w,
                       bf16_t* __restrict__ wt) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int K = 9 * cin, Kt = 9 * cout;
    if (e < cout * kpad) {
        const int co = e / kpad, k = e - co * kpad;
        w[e] = k < K ? f2bf(m[size_t(co) * K + k]) : bf16_t(0);
    }
    if (e < cin * kpad_t) {
        const int ci = e / kpad_t, kt = e - ci * kpad_t;
        bf16_t v = 0;
        if (kt < Kt) {
            const int tp = kt / cout, co = kt - tp * cout;
            v = f2bf(m[size_t(co) * K + (8 - tp) * cin + ci]);
        }
        wt[e] = v;
    }
}

unsigned cdivu(size_t a, unsigned b) { return unsigned((a + b - 1) / b); }

constexpr int kConvPT = 4;
template <int CT> void launch_conv_t(const ConvArgs& a, hipStream_t s) {
    const size_t npix = size_t(a.n) * a.h * a.wd;
    const size_t waves = (npix + 16 * kConvPT - 1) / (16 * kConvPT);
    k_conv_t<CT, kConvPT><<<cdivu(waves, 4), 256, 0, s>>>(a);
}
template <int CT> void launch_conv(const ConvArgs& a, hipStream_t s) {
    const size_t npix = size_t(a.n) * a.h * a.wd;
    const size_t waves = (npix + 31) / 32;
    k_conv<CT><<<cdivu(waves * 64, 256), 256, 0, s>>>(a);
}

// e1a e1b | e2a e2b | e3a e3b | b4a b4b | d3a d3b | d2a d2b | d1a d1b
struct LayerDef { int c0, c1, cout, level; };
const LayerDef kLayers[kNetLayers] = {{8, 0, 32, 0},   {32, 0, 32, 0},  {32, 0, 64, 1},   {64, 0, 64, 1}, {64, 0, 96, 2},
                                      {96, 0, 96, 2},  {96, 0, 128, 3}, {128, 0, 128, 3}, {96, 128, 96, 2}, {96, 0, 96, 2},
                                      {64, 96, 64, 1}, {64, 0, 64, 1},  {32, 64, 32, 0},  {32, 0, 32, 0}};
int round32(int k) { return (k + 31) / 32 * 32; }

}  // namespace

bool conv_shape_ok(int c0, int c1, int cout) {
    return c0 > 0 && c0 % 8 == 0 && c1 >= 0 && c1 % 8 == 0 && cout > 0 && cout % 16 == 0;
}

bool conv3x3(Ctx& c, const bf16_t* in0, const bf16_t* in1, int c0, int c1, const bf16_t* w, int kpad, const float* bias,
             int cout, int relu, int n, int h, int wd, bf16_t* out, hipStream_t s, bf16_t* pool) {
    ConvArgs a{in0, in1, w, bias, out, n, h, wd, c0, c1, cout, kpad, relu, 0, nullptr};
    // the kernels without the fused pool: a pool pass after them
    const auto pool_after = [&]() {
        if (pool)
            k_pool_fwd<<<cdivu(size_t(n) * (h / 2) * (wd / 2) * cout, 256), 256, 0, s>>>(out, pool, n, h, wd, cout);
    };
    const bool chunked = (c0 + c1) % 32 == 0 && c0 % 32 == 0;
    if (c0 == 8 && c1 == 0 && kpad == 96 && (c.conv_kernel == 0 || c.conv_kernel == 4) && (cout == 16 || cout == 32 || cout == 64)) {
        if (c.cus == 0) {
            int cus = 0;
            c.cus = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c.device) == hipSuccess && cus > 0 ? cus
                                                                                                                     : 256;
        }
        const int ntiles = int(size_t(n) * ((h + LT_H - 1) / LT_H) * ((wd + LT_W - 1) / LT_W));
        const unsigned blocks = unsigned(std::min(ntiles, c.cus * 4));
        a.co0 = 0;
        switch (cout / 16) {
            case 1: k_conv_c8<1><<<blocks, 256, 0, s>>>(a, ntiles); break;
            case 2: k_conv_c8<2><<<blocks, 256, 0, s>>>(a, ntiles); break;
            default: k_conv_c8<4><<<blocks, 256, 0, s>>>(a, ntiles); break;
        }
        pool_after();
        return hip_ok(c, hipGetLastError(), "conv3x3");
    }
    // the direct-copy persistent kernel: by default for a decoder's concatenated input with 32
    // outputs (several chunks, one channel group: its copies overlap the MFMAs); option 4: always
    const bool even = c1 == 0 || (h % 2 == 0 && wd % 2 == 0);
    if (chunked && even && (c.conv_kernel == 4 || (c.conv_kernel == 0 && c1 > 0 && cout == 32))) {
        if (c.cus == 0) {
            int cus = 0;
            c.cus = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c.device) == hipSuccess && cus > 0 ? cus
                                                                                                                     : 256;
        }
        a.pool = pool;
        const int ntiles = int(size_t(n) * ((h + 2 * DL_WAVES - 1) / (2 * DL_WAVES)) * ((wd + LT_W - 1) / LT_W));
        int ct = 3;   // two stages of 8 waves' halo and 64 channels' weights exceed the LDS
        while (ct > 1 && (cout / (16 * ct)) * ntiles < 2 * c.cus) --ct;
        for (int co0 = 0; co0 < cout;) {
            const int w16 = std::min(ct, (cout - co0) / 16);
            const int groups = (cout - co0) / (16 * w16);
            const int items = ntiles * groups;
            const unsigned blocks = unsigned(std::min(items, c.cus));   // one block a CU (its LDS)
            a.co0 = co0;
#define RAW_AO_DL(CT)                                                                                 \
    (pool ? k_conv_dl<CT, DL_WAVES, true><<<blocks, 64 * DL_WAVES, 0, s>>>(a, ntiles, items)         \
          : k_conv_dl<CT, DL_WAVES><<<blocks, 64 * DL_WAVES, 0, s>>>(a, ntiles, items))
            switch (w16) {
                case 1: RAW_AO_DL(1); break;
                case 2: RAW_AO_DL(2); break;
                default: RAW_AO_DL(3); break;
            }
#undef RAW_AO_DL
            co0 += groups * 16 * w16;
        }
        return hip_ok(c, hipGetLastError(), "conv3x3");
    }
    if (chunked && (c.conv_kernel == 0 || c.conv_kernel == 3) && even) {
        if (c.cus == 0) {
            int cus = 0;
            c.cus = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c.device) == hipSuccess && cus > 0 ? cus
                                                                                                                     : 256;
        }
        const int ntiles = int(size_t(n) * ((h + LT_H - 1) / LT_H) * ((wd + LT_W - 1) / LT_W));
        a.pool = pool;
        if (c.conv_kernel == 0 && c0 == 32 && c1 == 0 && cout == 32) {   // one chunk in, one group out: persistent
            a.co0 = 0;
            const unsigned blocks = unsigned(std::min(ntiles, c.cus * PF_BLOCKS));
            if (pool) k_conv_pf<2, true><<<blocks, 256, 0, s>>>(a, ntiles);
            else k_conv_pf<2><<<blocks, 256, 0, s>>>(a, ntiles);
            return hip_ok(c, hipGetLastError(), "conv3x3");
        }
        // the widest channel group (<= 64) that still gives every CU two work items;
        // channel groups of one width run in one launch, a narrower remainder in another
        int ct = 4;
        while (ct > 1 && (cout / (16 * ct)) * ntiles < 2 * c.cus) --ct;
        for (int co0 = 0; co0 < cout;) {
            const int w16 = std::min(ct, (cout - co0) / 16);
            const int groups = (cout - co0) / (16 * w16);
            const unsigned blocks = unsigned(ntiles * groups);
            a.co0 = co0;
            if (pool) {
                switch (w16) {
                    case 1: k_conv_lds<1, true><<<blocks, 256, 0, s>>>(a, ntiles); break;
                    case 2: k_conv_lds<2, true><<<blocks, 256, 0, s>>>(a, ntiles); break;
                    case 3: k_conv_lds<3, true><<<blocks, 256, 0, s>>>(a, ntiles); break;
                    default: k_conv_lds<4, true><<<blocks, 256, 0, s>>>(a, ntiles); break;
                }
            } else {
                switch (w16) {
                    case 1: k_conv_lds<1><<<blocks, 256, 0, s>>>(a, ntiles); break;
                    case 2: k_conv_lds<2><<<blocks, 256, 0, s>>>(a, ntiles); break;
                    case 3: k_conv_lds<3><<<blocks, 256, 0, s>>>(a, ntiles); break;
                    default: k_conv_lds<4><<<blocks, 256, 0, s>>>(a, ntiles); break;
                }
            }
            co0 += groups * 16 * w16;
        }
        return hip_ok(c, hipGetLastError(), "conv3x3");
    }
    const bool tapwise = c.conv_kernel != 1 && chunked && kpad == 9 * (c0 + c1);
    for (int co0 = 0; co0 < cout; co0 += 128) {
        a.co0 = co0;
        if (tapwise) {
            switch (std::min(8, (cout - co0) / 16)) {
                case 1: launch_conv_t<1>(a, s); break;
                case 2: launch_conv_t<2>(a, s); break;
                case 3: launch_conv_t<3>(a, s); break;
                case 4: launch_conv_t<4>(a, s); break;
                case 5: launch_conv_t<5>(a, s); break;
                case 6: launch_conv_t<6>(a, s); break;
                case 7: launch_conv_t<7>(a, s); break;
                default: launch_conv_t<8>(a, s); break;
            }
            continue;
        }
        switch (std::min(8, (cout - co0) / 16)) {
            case 1: launch_conv<1>(a, s); break;
            case 2: launch_conv<2>(a, s); break;
            case 3: launch_conv<3>(a, s); break;
            case 4: launch_conv<4>(a, s); break;
            case 5: launch_conv<5>(a, s); break;
            case 6: launch_conv<6>(a, s); break;
            case 7: launch_conv<7>(a, s); break;
            default: launch_conv<8>(a, s); break;
        }
    }
    pool_after();
    return hip_ok(c, hipGetLastError(), "conv3x3");
}

bool conv3x3_wgrad(Ctx& c, DevBuf& part, const bf16_t* dy, const bf16_t* in0, const bf16_t* in1, int c0, int c1, int cout,
                   int n, int h, int wd, float* gw, float* gb, hipStream_t s) {
    const int K = 9 * (c0 + c1);
    const int kpg = (K + 1 + 63) / 64 * 64;
    const size_t px = size_t(n) * h * wd;
    // chunks: enough blocks to fill the chip, with at least 128 pixels each
    const int ktiles = kpg / 64;
    int chunk = 128;
    while (size_t(ktiles) * ((px + chunk - 1) / chunk) > 2048 && chunk < (1 << 20)) chunk *= 2;
    const int nch = int((px + chunk - 1) / chunk);
    const int per = (nch + 31) / 32, ngroups = (nch + per - 1) / per;   // at most 32 groups, none empty
    const size_t psize = size_t(nch) * cout * kpg;
    if (!ensure(c, part, (psize + size_t(ngroups) * cout * kpg) * 4)) return false;
    WgradArgs wa{dy, in0, in1, part.as<float>(), n, h, wd, c0, c1, cout, kpg, chunk};
    const dim3 grid{unsigned(ktiles), unsigned(nch), 1u};
    switch (cout / 16) {
        case 2: k_wgrad<2><<<grid, 256, 0, s>>>(wa); break;
        case 4: k_wgrad<4><<<grid, 256, 0, s>>>(wa); break;
        case 6: k_wgrad<6><<<grid, 256, 0, s>>>(wa); break;
        case 8: k_wgrad<8><<<grid, 256, 0, s>>>(wa); break;
        default: c.err = "conv3x3 weight grad: output channels must be 32, 64, 96 or 128"; return false;
    }
    float* part2 = part.as<float>() + psize;
    const unsigned eb = cdivu(size_t(cout) * (K + 1), 256);
    k_wgrad_sum_groups<<<dim3{eb, unsigned(ngroups), 1u}, 256, 0, s>>>(part.as<float>(), nch, per, cout, kpg, K, part2);
    k_wgrad_reduce<<<eb, 256, 0, s>>>(part2, ngroups, cout, kpg, K, gw, gb);
    return hip_ok(c, hipGetLastError(), "conv3x3 weight grad");
}

size_t UNet::param_count() const {
    size_t n = 0;
    for (const auto& L : kLayers) n += size_t(L.cout) * 9 * (L.c0 + L.c1) + L.cout;
    return n + 33;
}

bool UNet::init(Ctx& c) {
    ctx = &c;
    size_t off = 0, wb = 0, wtb = 0;
    for (int i = 0; i < kNetLayers; ++i) {
        const LayerDef& L = kLayers[i];
        Layer& l = layers[i];
        l.c0 = L.c0;
        l.c1 = L.c1;
        l.cout = L.cout;
        l.level = L.level;
        l.K = 9 * (L.c0 + L.c1);
        l.kpad = round32(l.K);
        l.kpad_t = round32(9 * L.cout);
        l.woff = off;
        off += size_t(l.cout) * l.K;
        l.boff = off;
        off += size_t(l.cout);
        l.wofs16 = wb;
        wb += size_t(l.cout) * l.kpad;
        l.wtofs16 = wtb;
        wtb += size_t(l.c0 + l.c1) * l.kpad_t;
    }
    head_off = off;
    const size_t np = param_count();
    if (!ensure(c, params, 4 * np) || !ensure(c, grads, 4 * np) || !ensure(c, adam_m, 4 * np) || !ensure(c, adam_v, 4 * np) ||
        !ensure(c, w16, 2 * wb) || !ensure(c, wt16, 2 * wtb))
        return false;
    return hip_ok(c, hipMemset(params.p, 0, 4 * np), "net init") && hip_ok(c, hipMemset(grads.p, 0, 4 * np), "net init") &&
           hip_ok(c, hipMemset(adam_m.p, 0, 4 * np), "net init") && hip_ok(c, hipMemset(adam_v.p, 0, 4 * np), "net init");
}

bool UNet::pack(hipStream_t s) {
    for (auto& l : layers) {
        const int cin = l.c0 + l.c1;
        const int nel = std::max(l.cout * l.kpad, cin * l.kpad_t);
        k_pack<<<cdivu(size_t(nel), 256), 256, 0, s>>>(params.as<float>() + l.woff, l.cout, cin, l.kpad, l.kpad_t,
                                                       w16.as<bf16_t>() + l.wofs16, wt16.as<bf16_t>() + l.wtofs16);
    }
    return hip_ok(*ctx, hipGetLastError(), "net pack");
}

bool UNet::ensure_acts(int n, int h, int w, bool grad) {
    Ctx& c = *ctx;
    const size_t np = size_t(n) * h * w;
    if (!ensure(c, pred, np * 4)) return false;
    for (int i = 0; i < kNetLayers; ++i) {
        const size_t px = size_t(n) * (h >> layers[i].level) * (w >> layers[i].level);
        if (!ensure(c, act[i], px * 2 * layers[i].cout)) return false;
    }
    for (int lv = 0; lv < 3; ++lv) {
        const size_t px = size_t(n) * (h >> (lv + 1)) * (w >> (lv + 1));
        if (!ensure(c, pooled[lv], px * 2 * layers[2 * lv + 1].cout)) return false;
    }
    if (grad) {
        for (int lv = 0; lv < 4; ++lv) {
            // the widest gradient of a level: a decoder input (skip + upsampled) or an output
            int cmax = 0;
            for (const auto& l : layers)
                if (l.level == lv) cmax = std::max(cmax, std::max(l.c0 + l.c1, l.cout));
            const size_t px = size_t(n) * (h >> lv) * (w >> lv);
            for (auto& b : lvbuf[lv])
                if (!ensure(c, b, px * 2 * cmax)) return false;
        }
        if (!ensure(c, headpart, (np / 256 + 2) * 34 * 4) || !ensure(c, lossbuf, 16) || !ensure(c, cnt, 16)) return false;
    }
    an = n;
    ah = h;
    aw = w;
    return true;
}

bool UNet::features(const float* pos, const float* nrm, const uint8_t* mask, const float* ssao, const float* noisy,
                    const float eye[3], int h, int w, bf16_t* x, hipStream_t s) {
    const size_t np = size_t(h) * w;
    k_features<<<cdivu(np, 256), 256, 0, s>>>(pos, nrm, mask, ssao, noisy, eye[0], eye[1], eye[2], np, x);
    return hip_ok(*ctx, hipGetLastError(), "net features");
}

namespace {
// the forward input(s) of layer i
void layer_inputs(const UNet& u, int i, const bf16_t* x, const bf16_t*& in0, const bf16_t*& in1) {
    in1 = nullptr;
    if (i == 0) in0 = x;
    else if (i == 2 || i == 4 || i == 6) in0 = u.pooled[i / 2 - 1].as<bf16_t>();
    else if (i == 8 || i == 10 || i == 12) {
        in0 = u.act[i == 8 ? 5 : (i == 10 ? 3 : 1)].as<bf16_t>();
        in1 = u.act[i - 1].as<bf16_t>();
    } else in0 = u.act[i - 1].as<bf16_t>();
}
}  // namespace

bool UNet::forward(const bf16_t* x, const uint8_t* mask, int n, int h, int w, float* ao_out, hipStream_t s) {
    if (!ensure_acts(n, h, w, false)) return false;
    Ctx& c = *ctx;
    const size_t np = size_t(n) * h * w;
    for (int i = 0; i < kNetLayers; ++i) {
        Layer& l = layers[i];
        const int lh = h >> l.level, lw = w >> l.level;
        const bf16_t *in0, *in1;
        layer_inputs(*this, i, x, in0, in1);
        // an encoder level's last output is pooled for the next level by the same launch
        bf16_t* pool = (i == 1 || i == 3 || i == 5) ? pooled[i / 2].as<bf16_t>() : nullptr;
        if (!conv3x3(c, in0, in1, l.c0, l.c1, w16.as<bf16_t>() + l.wofs16, l.kpad, params.as<float>() + l.boff, l.cout, 1, n, lh,
                     lw, act[i].as<bf16_t>(), s, pool))
            return false;
    }
    k_head<<<cdivu(np, 256), 256, 0, s>>>(act[kNetLayers - 1].as<bf16_t>(), params.as<float>() + head_off, mask, np,
                                          pred.as<float>(), ao_out);
    return hip_ok(c, hipGetLastError(), "net forward");
}

bool UNet::backward(const bf16_t* x, const uint8_t* mask, const float* target, int n, int h, int w, hipStream_t s) {
    Ctx& c = *ctx;
    if (!ensure_acts(n, h, w, true)) return false;
    const size_t np = size_t(n) * h * w;
    const unsigned hb = cdivu(np, 256);
    float* gr = grads.as<float>();
    k_zero_u32<<<1, 4, 0, s>>>(cnt.as<uint32_t>());
    k_mask_count<<<std::min(1024u, hb), 256, 0, s>>>(mask, np, cnt.as<uint32_t>());
    k_head_bwd<<<hb, 256, 0, s>>>(act[kNetLayers - 1].as<bf16_t>(), params.as<float>() + head_off, pred.as<float>(), target,
                                  mask, np, cnt.as<uint32_t>(), lvbuf[0][0].as<bf16_t>(), headpart.as<float>());
    k_head_reduce<<<34, 256, 0, s>>>(headpart.as<float>(), int(hb), gr + head_off, lossbuf.as<float>());
    // Three rotating buffers a level: the grad arriving at a layer's output (cur), its
    // pre-activation grad (cur+1) and the grad it passes to its input (cur+2).
    int cur[4] = {0, 0, 0, 0};
    bf16_t* dy = lvbuf[0][0].as<bf16_t>();
    for (int i = kNetLayers - 1; i >= 0; --i) {
        Layer& l = layers[i];
        const int lv = l.level, lh = h >> lv, lw = w >> lv;
        const size_t px = size_t(n) * lh * lw;
        const int cin = l.c0 + l.c1;
        bf16_t* dpre = lvbuf[lv][(cur[lv] + 1) % 3].as<bf16_t>();
        k_relu_bwd<<<cdivu(px * l.cout, 256), 256, 0, s>>>(act[i].as<bf16_t>(), dy, dpre, px * l.cout);
        const bf16_t *in0, *in1;
        layer_inputs(*this, i, x, in0, in1);
        if (!conv3x3_wgrad(c, wpart, dpre, in0, in1, l.c0, l.c1, l.cout, n, lh, lw, gr + l.woff, gr + l.boff, s)) return false;
        if (i == 0) break;   // the input features take no gradient
        bf16_t* dx = lvbuf[lv][(cur[lv] + 2) % 3].as<bf16_t>();
        if (!conv3x3(c, dpre, nullptr, l.cout, 0, wt16.as<bf16_t>() + l.wtofs16, l.kpad_t, nullptr, cin, 0, n, lh, lw, dx, s))
            return false;
        cur[lv] = (cur[lv] + 2) % 3;
        if (i == 8 || i == 10 || i == 12) {
            // the upsampled half goes to the lower level's last layer; the skip half waits
            // for the encoder output it came from
            const int llv = lv + 1;
            bf16_t* dlow = lvbuf[llv][(cur[llv] + 1) % 3].as<bf16_t>();
            const size_t nl = size_t(n) * (h >> llv) * (w >> llv) * l.c1;
            k_up_grad<<<cdivu(nl, 256), 256, 0, s>>>(dx, cin, l.c0, l.c1, dlow, n, lh, lw);
            cur[llv] = (cur[llv] + 1) % 3;
            skip_dx[lv] = dx;
            skip_cx[lv] = cin;
            dy = dlow;
        } else if (i == 2 || i == 4 || i == 6) {
            // dx is the pooled tensor's grad; the encoder output it pooled also fed a skip
            const int plv = lv - 1;
            const int c0 = layers[i - 1].cout;
            bf16_t* dsk = lvbuf[plv][(cur[plv] + 1) % 3].as<bf16_t>();
            const size_t ns = size_t(n) * (h >> plv) * (w >> plv) * c0;
            k_skip_grad<<<cdivu(ns, 256), 256, 0, s>>>(skip_dx[plv], skip_cx[plv], c0, dx, dsk, n, h >> plv, w >> plv);
            cur[plv] = (cur[plv] + 1) % 3;
            dy = dsk;
        } else {
            dy = dx;
        }
    }
    return hip_ok(c, hipGetLastError(), "net backward");
}

bool UNet::adam(float lr, hipStream_t s) {
    ++step;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, float(step)), c2 = 1.0f - std::pow(b2, float(step));
    const size_t np = param_count();
    k_adam<<<cdivu(np, 256), 256, 0, s>>>(params.as<float>(), grads.as<float>(), adam_m.as<float>(), adam_v.as<float>(), np, lr,
                                          b1, b2, 1e-8f, c1, c2);
    return hip_ok(*ctx, hipGetLastError(), "adam") && pack(s);
}

void UNet::release() {
    auto drop = [](DevBuf& b) {
        if (b.p) (void)hipFree(b.p);
        b.p = nullptr;
        b.cap = 0;
    };
    DevBuf* bufs[] = {&params, &grads, &adam_m, &adam_v, &w16, &wt16, &pred, &headpart, &lossbuf, &cnt, &wpart,
                      &f_ssao, &f_x, &pr_pos, &pr_nrm, &pr_mask, &pr_rt, &pr_net, &pr_stats};
    for (auto* b : bufs) drop(*b);
    if (h_stats) (void)hipHostFree(h_stats);
    h_stats = nullptr;
    for (auto& lv : lvbuf)
        for (auto& b : lv) drop(b);
    for (auto& b : act) drop(b);
    for (auto& b : pooled) drop(b);
}

namespace {
// The probe sub-image: every stride-th pixel in x and y (from stride/2), its G-buffer and
// the network's AO there.
__global__ void k_probe_gather(const float* __restrict__ pos, const float* __restrict__ nrm, const uint8_t* __restrict__ mask,
                               const float* __restrict__ net, int w, int stride, int pw, int ph, float* __restrict__ ppos,
                               float* __restrict__ pnrm, uint8_t* __restrict__ pmask, float* __restrict__ pnet) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= pw * ph) return;
    const int px = i % pw, py = i / pw;
    const size_t src = size_t(py * stride + stride / 2) * w + size_t(px * stride + stride / 2);
    for (int k = 0; k < 3; ++k) {
        ppos[3 * i + k] = pos[3 * src + k];
        pnrm[3 * i + k] = nrm[3 * src + k];
    }
    pmask[i] = mask[src];
    pnet[i] = net[src];
}
// One block: the sum of squared differences (float64), the largest difference and the
// covered count over the probe pixels, in a fixed order.
__global__ __launch_bounds__(1024) void k_probe_stats(const float* __restrict__ rt, const float* __restrict__ net,
                                                      const uint8_t* __restrict__ mask, int n, double* __restrict__ out) {
    __shared__ double ss[1024];
    __shared__ float mx[1024];
    __shared__ unsigned cn[1024];
    double s = 0.0;
    float m = 0.0f;
    unsigned c = 0;
    for (int i = threadIdx.x; i < n; i += 1024)
        if (mask[i]) {
            const float d = fabsf(net[i] - rt[i]);
            s += double(d) * double(d);
            m = fmaxf(m, d);
            ++c;
        }
    ss[threadIdx.x] = s;
    mx[threadIdx.x] = m;
    cn[threadIdx.x] = c;
    __syncthreads();
    for (int k = 512; k > 0; k >>= 1) {
        if (int(threadIdx.x) < k) {
            ss[threadIdx.x] += ss[threadIdx.x + k];
            mx[threadIdx.x] = fmaxf(mx[threadIdx.x], mx[threadIdx.x + k]);
            cn[threadIdx.x] += cn[threadIdx.x + k];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = ss[0];
        out[1] = double(mx[0]);
        out[2] = double(cn[0]);
    }
}
}  // namespace

int trace_built(Ctx& c, const float* pos, const float* nrm, const uint8_t* mask, int w, int h, int spp, float radius, float* ao,
                hipStream_t s) {
    if (c.root < 0) {
        c.err = "no BVH: call raw_ao_build_ploc first";
        return -1;
    }
    if (c.trace_kernel != c.built_kernel) {
        c.err = "the trace kernel option changed after the build: build again";
        return -1;
    }
    if (c.host_trig) {
        const int r = trig_table_upload(c);
        if (r) return r;
    }
    return c.built_kernel == RAW_AO_TRACE_BINARY ? trace_ao(c, pos, nrm, mask, w, h, spp, radius, ao, s)
                                                 : trace_ao4(c, pos, nrm, mask, w, h, spp, radius, ao, s);
}

int learned_frame(Ctx& c, UNet& u, const LearnedArgs& a, raw_ao_learned_report* rep, hipStream_t s) {
    const size_t np = size_t(a.w) * a.h;
    // back-off after fall-backs (RAW_AO_OPT_LEARNED_BACKOFF): ray trace at once while calls to skip remain
    const bool backoff = a.probe_stride > 0 && c.learned_backoff > 0;
    if (backoff && (c.lb_w != a.w || c.lb_h != a.h)) {
        c.lb_w = a.w;
        c.lb_h = a.h;
        c.lb_next = c.lb_left = 0;
    }
    if (backoff && c.lb_left > 0) {
        --c.lb_left;
        if (c.host_trig) {
            const int r = trig_table_upload(c);
            if (r) return r;
        }
        const int r = trace_built(c, a.pos, a.nrm, a.mask, a.w, a.h, a.rt_spp, a.rt_radius, a.ao, s);
        if (r) return r;
        if (rep) {
            *rep = raw_ao_learned_report{};
            rep->fell_back = 2;
        }
        return hip_ok(c, hipGetLastError(), "learned frame") ? 0 : -2;
    }
    if (!ensure(c, u.f_ssao, 4 * np) || !ensure(c, u.f_x, 2 * kNetFeatures * np)) return -2;
    if (c.host_trig) {
        const int r = trig_table_upload(c);
        if (r) return r;
    }
    int r = ssao(c, a.pos, a.nrm, a.mask, a.w, a.h, a.ssao_samples, a.ssao_radius, u.f_ssao.as<float>(), s);
    if (r) return r;
    if (!u.features(a.pos, a.nrm, a.mask, u.f_ssao.as<float>(), a.noisy, a.eye, a.h, a.w, u.f_x.as<bf16_t>(), s) ||
        !u.forward(u.f_x.as<bf16_t>(), a.mask, 1, a.h, a.w, a.ao, s))
        return -2;
    raw_ao_learned_report out{};
    if (a.probe_stride > 0) {
        const int st = a.probe_stride, pw = a.w / st, ph = a.h / st;
        const size_t pn = size_t(pw) * ph;
        if (pn == 0) {
            c.err = "the probe stride exceeds the image";
            return -1;
        }
        if (!ensure(c, u.pr_pos, 12 * pn) || !ensure(c, u.pr_nrm, 12 * pn) || !ensure(c, u.pr_mask, pn) ||
            !ensure(c, u.pr_rt, 4 * pn) || !ensure(c, u.pr_net, 4 * pn) || !ensure(c, u.pr_stats, 32))
            return -2;
        if (!u.h_stats && !hip_ok(c, hipHostMalloc(reinterpret_cast<void**>(&u.h_stats), 32), "pinned stats")) return -2;
        k_probe_gather<<<cdivu(pn, 256), 256, 0, s>>>(a.pos, a.nrm, a.mask, a.ao, a.w, st, pw, ph, u.pr_pos.as<float>(),
                                                      u.pr_nrm.as<float>(), u.pr_mask.as<uint8_t>(), u.pr_net.as<float>());
        r = trace_built(c, u.pr_pos.as<float>(), u.pr_nrm.as<float>(), u.pr_mask.as<uint8_t>(), pw, ph, a.rt_spp, a.rt_radius,
                        u.pr_rt.as<float>(), s);
        if (r) return r;
        k_probe_stats<<<1, 1024, 0, s>>>(u.pr_rt.as<float>(), u.pr_net.as<float>(), u.pr_mask.as<uint8_t>(), int(pn),
                                         u.pr_stats.as<double>());
        if (!hip_ok(c, hipMemcpyAsync(u.h_stats, u.pr_stats.p, 24, hipMemcpyDeviceToHost, s), "probe stats") ||
            !hip_ok(c, hipStreamSynchronize(s), "probe"))
            return -2;
        const double* hs = reinterpret_cast<const double*>(u.h_stats);
        out.probe_pixels = uint32_t(hs[2]);
        out.probe_rmse = out.probe_pixels ? float(std::sqrt(hs[0] / hs[2])) : 0.0f;
        out.probe_max_error = float(hs[1]);
        if (out.probe_rmse > a.fallback_rmse) {
            r = trace_built(c, a.pos, a.nrm, a.mask, a.w, a.h, a.rt_spp, a.rt_radius, a.ao, s);
            if (r) return r;
            out.fell_back = 1;
            if (backoff) {   // skip 1, 2, 4, ... up to learned_backoff calls
                c.lb_next = c.lb_next <= 0 ? 1 : std::min(2 * c.lb_next, c.learned_backoff);
                c.lb_left = c.lb_next;
            }
        } else if (backoff) {
            c.lb_next = c.lb_left = 0;
        }
    }
    if (!hip_ok(c, hipGetLastError(), "learned frame")) return -2;
    if (rep) *rep = out;
    return 0;
}

void net_release(Ctx& c) {
    if (!c.net) return;
    c.net->release();
    delete c.net;
    c.net = nullptr;
}

}  // namespace raw_ao

// ---- C ABI ------------------------------------------------------------------------------
using raw_ao::Ctx;
using raw_ao::UNet;
namespace {
int nfail(Ctx& c, int code, const char* msg) {
    if (c.err.empty()) c.err = msg;
    return code;
}
bool nbegin(Ctx& c) {
    c.err.clear();
    return raw_ao::hip_ok(c, hipSetDevice(c.device), "hipSetDevice");
}
hipStream_t npick(Ctx& c, void* s) { return s ? static_cast<hipStream_t>(s) : c.stream; }
UNet* net_of(Ctx& c) {
    if (c.net) return c.net;
    UNet* u = new (std::nothrow) UNet;
    if (!u) return nullptr;
    if (!u->init(c)) {
        u->release();
        delete u;
        return nullptr;
    }
    c.net = u;
    return u;
}
bool net_shape_ok(int n, int w, int h) {
    return n > 0 && w >= 8 && h >= 8 && w % 8 == 0 && h % 8 == 0 && int64_t(n) * w * h <= (int64_t(1) << 28);
}
int net_read(raw_ao_ctx* x, float* out, uint64_t count, bool grads) {
    if (!x || !out) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    if (!nbegin(c)) return RAW_AO_ERR_DEVICE;
    UNet* u = net_of(c);
    if (!u) return nfail(c, RAW_AO_ERR_DEVICE, "network allocation failed");
    const size_t np = u->param_count();
    if (count != np) return nfail(c, RAW_AO_ERR_ARG, "parameter count mismatch");
    if (!raw_ao::hip_ok(c, hipDeviceSynchronize(), "net read") ||
        !raw_ao::hip_ok(c, hipMemcpy(out, grads ? u->grads.p : u->params.p, 4 * np, hipMemcpyDeviceToHost), "net read"))
        return RAW_AO_ERR_DEVICE;
    return RAW_AO_OK;
}
}  // namespace

extern "C" {

int raw_ao_net_param_count(raw_ao_ctx* x, uint64_t* count) {
    if (!x || !count) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    if (!nbegin(c)) return RAW_AO_ERR_DEVICE;
    UNet* u = net_of(c);
    if (!u) return nfail(c, RAW_AO_ERR_DEVICE, "network allocation failed");
    *count = u->param_count();
    return RAW_AO_OK;
}

int raw_ao_net_set_params(raw_ao_ctx* x, const float* params, uint64_t count) {
    if (!x || !params) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    if (!nbegin(c)) return RAW_AO_ERR_DEVICE;
    UNet* u = net_of(c);
    if (!u) return nfail(c, RAW_AO_ERR_DEVICE, "network allocation failed");
    const size_t np = u->param_count();
    if (count != np) return nfail(c, RAW_AO_ERR_ARG, "parameter count mismatch");
    if (!raw_ao::hip_ok(c, hipDeviceSynchronize(), "param upload") ||
        !raw_ao::hip_ok(c, hipMemcpy(u->params.p, params, 4 * np, hipMemcpyHostToDevice), "param upload") ||
        !raw_ao::hip_ok(c, hipMemset(u->adam_m.p, 0, 4 * np), "adam reset") ||
        !raw_ao::hip_ok(c, hipMemset(u->adam_v.p, 0, 4 * np), "adam reset") || !u->pack(c.stream) ||
        !raw_ao::hip_ok(c, hipStreamSynchronize(c.stream), "param pack"))
        return RAW_AO_ERR_DEVICE;
    u->step = 0;
    return RAW_AO_OK;
}

int raw_ao_net_get_params(raw_ao_ctx* x, float* params, uint64_t count) { return net_read(x, params, count, false); }
int raw_ao_net_get_grads(raw_ao_ctx* x, float* grads, uint64_t count) { return net_read(x, grads, count, true); }

int raw_ao_net_features_dev(raw_ao_ctx* x, const float* d_pos, const float* d_nrm, const uint8_t* d_mask, const float* d_ssao,
                            const float* d_noisy, const float eye[3], int w, int h, uint16_t* d_x, void* stream) {
    if (!x) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    if (!d_pos || !d_nrm || !d_mask || !d_ssao || !eye || !d_x || w <= 0 || h <= 0 || int64_t(w) * h > (int64_t(1) << 28))
        return nfail(c, RAW_AO_ERR_ARG, "invalid feature arguments");
    if (!nbegin(c)) return RAW_AO_ERR_DEVICE;
    UNet* u = net_of(c);
    if (!u) return nfail(c, RAW_AO_ERR_DEVICE, "network allocation failed");
    return u->features(d_pos, d_nrm, d_mask, d_ssao, d_noisy, eye, h, w, d_x, npick(c, stream)) ? RAW_AO_OK : RAW_AO_ERR_DEVICE;
}

int raw_ao_net_forward_dev(raw_ao_ctx* x, const uint16_t* d_x, const uint8_t* d_mask, int n, int w, int h, float* d_ao,
                           void* stream) {
    if (!x) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    if (!d_x || !d_mask || !net_shape_ok(n, w, h)) return nfail(c, RAW_AO_ERR_ARG, "invalid network input (w, h multiples of 8)");
    if (!nbegin(c)) return RAW_AO_ERR_DEVICE;
    UNet* u = net_of(c);
    if (!u) return nfail(c, RAW_AO_ERR_DEVICE, "network allocation failed");
    return u->forward(d_x, d_mask, n, h, w, d_ao, npick(c, stream)) ? RAW_AO_OK : RAW_AO_ERR_DEVICE;
}

int raw_ao_net_train_step_dev(raw_ao_ctx* x, const uint16_t* d_x, const uint8_t* d_mask, const float* d_target, int n, int w,
                              int h, float lr, void* stream) {
    if (!x) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    if (!d_x || !d_mask || !d_target || !net_shape_ok(n, w, h) || !(lr >= 0.0f))
        return nfail(c, RAW_AO_ERR_ARG, "invalid training arguments");
    if (!nbegin(c)) return RAW_AO_ERR_DEVICE;
    UNet* u = net_of(c);
    if (!u) return nfail(c, RAW_AO_ERR_DEVICE, "network allocation failed");
    const hipStream_t s = npick(c, stream);
    if (!u->forward(d_x, d_mask, n, h, w, nullptr, s) || !u->backward(d_x, d_mask, d_target, n, h, w, s)) return RAW_AO_ERR_DEVICE;
    if (lr > 0.0f && !u->adam(lr, s)) return RAW_AO_ERR_DEVICE;
    return RAW_AO_OK;
}

int raw_ao_net_loss(raw_ao_ctx* x, float* loss) {
    if (!x || !loss) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    if (!nbegin(c)) return RAW_AO_ERR_DEVICE;
    if (!c.net || !c.net->lossbuf.p) return nfail(c, RAW_AO_ERR_ARG, "no training step has run");
    if (!raw_ao::hip_ok(c, hipDeviceSynchronize(), "loss read") ||
        !raw_ao::hip_ok(c, hipMemcpy(loss, c.net->lossbuf.p, 4, hipMemcpyDeviceToHost), "loss read"))
        return RAW_AO_ERR_DEVICE;
    return RAW_AO_OK;
}

int raw_ao_net_read_act(raw_ao_ctx* x, int layer, uint16_t* out, uint64_t count) {
    if (!x || !out) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    if (!nbegin(c)) return RAW_AO_ERR_DEVICE;
    if (!c.net || c.net->an == 0 || layer < 0 || layer >= raw_ao::kNetLayers) return nfail(c, RAW_AO_ERR_ARG, "no such activation");
    const UNet& u = *c.net;
    const auto& l = u.layers[layer];
    const size_t n = size_t(u.an) * (u.ah >> l.level) * (u.aw >> l.level) * l.cout;
    if (count != n) return nfail(c, RAW_AO_ERR_ARG, "activation size mismatch");
    if (!raw_ao::hip_ok(c, hipDeviceSynchronize(), "act read") ||
        !raw_ao::hip_ok(c, hipMemcpy(out, u.act[layer].p, 2 * n, hipMemcpyDeviceToHost), "act read"))
        return RAW_AO_ERR_DEVICE;
    return RAW_AO_OK;
}

int raw_ao_conv3x3_dev(raw_ao_ctx* x, const uint16_t* d_in0, const uint16_t* d_in1, int c0, int c1, const uint16_t* d_w, int kpad,
                       const float* d_bias, int cout, int relu, int n, int w, int h, uint16_t* d_out, void* stream) {
    if (!x) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    if (!d_in0 || (c1 > 0 && !d_in1) || !d_w || !d_out || !raw_ao::conv_shape_ok(c0, c1, cout) || kpad < 9 * (c0 + c1) ||
        kpad % 32 != 0 || n <= 0 || w <= 0 || h <= 0 || (c1 > 0 && (w % 2 || h % 2)) || int64_t(n) * w * h > (int64_t(1) << 28))
        return nfail(c, RAW_AO_ERR_ARG, "invalid convolution arguments");
    if (!nbegin(c)) return RAW_AO_ERR_DEVICE;
    return raw_ao::conv3x3(c, d_in0, d_in1, c0, c1, d_w, kpad, d_bias, cout, relu, n, h, w, d_out, npick(c, stream))
               ? RAW_AO_OK
               : RAW_AO_ERR_DEVICE;
}

int raw_ao_conv3x3_wgrad_dev(raw_ao_ctx* x, const uint16_t* d_dy, const uint16_t* d_in0, const uint16_t* d_in1, int c0, int c1,
                             int cout, int n, int w, int h, float* d_gw, float* d_gb, void* stream) {
    if (!x) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    if (!d_dy || !d_in0 || (c1 > 0 && !d_in1) || !d_gw || !d_gb || !raw_ao::conv_shape_ok(c0, c1, cout) ||
        (cout != 32 && cout != 64 && cout != 96 && cout != 128) || n <= 0 || w <= 0 || h <= 0 || (c1 > 0 && (w % 2 || h % 2)) ||
        int64_t(n) * w * h > (int64_t(1) << 28))
        return nfail(c, RAW_AO_ERR_ARG, "invalid weight-gradient arguments");
    if (!nbegin(c)) return RAW_AO_ERR_DEVICE;
    UNet* u = net_of(c);
    if (!u) return nfail(c, RAW_AO_ERR_DEVICE, "network allocation failed");
    return raw_ao::conv3x3_wgrad(c, u->wpart, d_dy, d_in0, d_in1, c0, c1, cout, n, h, w, d_gw, d_gb, npick(c, stream))
               ? RAW_AO_OK
               : RAW_AO_ERR_DEVICE;
}

int raw_ao_net_load(raw_ao_ctx* x, const char* path) {
    if (!x) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    c.err.clear();
    if (!path) return nfail(c, RAW_AO_ERR_ARG, "null path");
    FILE* f = std::fopen(path, "rb");
    if (!f) return nfail(c, RAW_AO_ERR_ARG, "cannot open the weights file");
    char magic[4] = {0, 0, 0, 0};
    uint32_t version = 0, meta_len = 0;
    uint64_t count = 0;
    bool ok = std::fread(magic, 1, 4, f) == 4 && std::fread(&version, 4, 1, f) == 1 && std::fread(&count, 8, 1, f) == 1 &&
              std::fread(&meta_len, 4, 1, f) == 1 && std::memcmp(magic, "RAOW", 4) == 0 && version == 1 &&
              count > 0 && count < (uint64_t(1) << 28);
    std::vector<float> p;
    if (ok) {
        p.resize(size_t(count));
        ok = std::fread(p.data(), 4, p.size(), f) == p.size();
    }
    std::fclose(f);
    if (!ok) return nfail(c, RAW_AO_ERR_ARG, "not a version-1 RAOW weights file");
    return raw_ao_net_set_params(x, p.data(), count);
}

int raw_ao_learned_ao_dev(raw_ao_ctx* x, const float* d_pos, const float* d_nrm, const uint8_t* d_mask, const float* d_noisy,
                          const float eye[3], int w, int h, int ssao_samples, float ssao_radius, int rt_spp, float rt_radius,
                          int probe_stride, float fallback_rmse, float* d_ao, raw_ao_learned_report* report, void* stream) {
    if (!x) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    if (!d_pos || !d_nrm || !d_mask || !eye || !d_ao || !net_shape_ok(1, w, h) || ssao_samples < 0 || ssao_samples > 4096 ||
        probe_stride < 0 || (probe_stride > 0 && (rt_spp < 1 || rt_spp > 4096 || !(rt_radius > 0.0f))))
        return nfail(c, RAW_AO_ERR_ARG, "invalid learned AO arguments (w, h multiples of 8)");
    if (!nbegin(c)) return RAW_AO_ERR_DEVICE;
    UNet* u = net_of(c);
    if (!u) return nfail(c, RAW_AO_ERR_DEVICE, "network allocation failed");
    const raw_ao::LearnedArgs a{d_pos, d_nrm, d_mask, d_noisy, eye, w, h, ssao_samples, ssao_radius, rt_spp, rt_radius,
                                probe_stride, fallback_rmse, d_ao};
    const int r = raw_ao::learned_frame(c, *u, a, report, npick(c, stream));
    if (r == -1) return RAW_AO_ERR_ARG;
    return r ? RAW_AO_ERR_DEVICE : RAW_AO_OK;
}

int raw_ao_learned_ao(raw_ao_ctx* x, const float* pos, const float* nrm, const uint8_t* mask, const float* noisy,
                      const float eye[3], int w, int h, int ssao_samples, float ssao_radius, int rt_spp, float rt_radius,
                      int probe_stride, float fallback_rmse, float* ao, raw_ao_learned_report* report) {
    if (!x) return RAW_AO_ERR_ARG;
    Ctx& c = x->c;
    if (!pos || !nrm || !mask || !ao || !net_shape_ok(1, w, h)) return nfail(c, RAW_AO_ERR_ARG, "bad image arguments");
    if (!nbegin(c)) return RAW_AO_ERR_DEVICE;
    const size_t np = size_t(w) * size_t(h);
    if (!raw_ao::ensure(c, c.st_pos, 12 * np) || !raw_ao::ensure(c, c.st_nrm, 12 * np) || !raw_ao::ensure(c, c.st_mask, np) ||
        !raw_ao::ensure(c, c.st_out, 4 * np) || (noisy && !raw_ao::ensure(c, c.st_out2, 4 * np)))
        return RAW_AO_ERR_DEVICE;
    hipStream_t s = c.stream;
    if (!raw_ao::hip_ok(c, hipMemcpyAsync(c.st_pos.p, pos, 12 * np, hipMemcpyHostToDevice, s), "upload") ||
        !raw_ao::hip_ok(c, hipMemcpyAsync(c.st_nrm.p, nrm, 12 * np, hipMemcpyHostToDevice, s), "upload") ||
        !raw_ao::hip_ok(c, hipMemcpyAsync(c.st_mask.p, mask, np, hipMemcpyHostToDevice, s), "upload") ||
        (noisy && !raw_ao::hip_ok(c, hipMemcpyAsync(c.st_out2.p, noisy, 4 * np, hipMemcpyHostToDevice, s), "upload")))
        return RAW_AO_ERR_DEVICE;
    const int r = raw_ao_learned_ao_dev(x, c.st_pos.as<float>(), c.st_nrm.as<float>(), c.st_mask.as<uint8_t>(),
                                        noisy ? c.st_out2.as<float>() : nullptr, eye, w, h, ssao_samples, ssao_radius, rt_spp,
                                        rt_radius, probe_stride, fallback_rmse, c.st_out.as<float>(), report, nullptr);
    if (r) return r;
    if (!raw_ao::hip_ok(c, hipMemcpyAsync(ao, c.st_out.p, 4 * np, hipMemcpyDeviceToHost, s), "download") ||
        !raw_ao::hip_ok(c, hipStreamSynchronize(s), "learned frame"))
        return RAW_AO_ERR_DEVICE;
    return RAW_AO_OK;
}

}  // extern "C"
