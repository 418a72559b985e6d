// denoise.hip — the learned denoiser: a U-Net (levels of 32/48/64/96
// channels, 3x3 convolutions, ReLU, 2x2 average pooling, nearest upsampling
// with skip concatenation, 1x1 head adding the input's log radiance) trained
// and run with hand-written BF16 MFMA kernels that accumulate in FP32.
//
// The reference has no denoiser; this network is the build's own. Its
// restatement in numpy (oracle/unet.py) is the checker of these kernels.
//
// Layout: activations are NHWC BF16 with the channel count a multiple of 8,
// so a lane's 8-channel group is one 16-byte load. The network image is the
// input padded with zeros to multiples of 8 (three pooling levels).
//
// Convolution = implicit GEMM on v_mfma_f32_16x16x32_bf16 with output
// channels as the M dimension, pixels as N and K = 9 taps x input channels:
//   A (16 x 32)  weights  [Cout][Kpad], k = tap * Cin + c  (lane: row l%16,
//                k = 8 (l/16) .. +7)
//   B (32 x 16)  inputs   pixel (l%16) of the tile, channels of k's tap
//   D (16 x 16)  lane holds pixel l%16, output channels 4 (l/16) .. +3
// The decoder's "upsample, then concatenate with the skip tensor" is not
// materialized: the B loader reads channels [0, C0) from the low-resolution
// tensor at (y/2, x/2) and [C0, C0+C1) from the skip tensor.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <algorithm>
#include <cmath>

#include "internal.hpp"

namespace hh {
namespace dn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

enum Epi { EPI_BIAS_RELU = 0, EPI_MASK = 1, EPI_F32 = 2 };

__device__ inline bf16x8 zero8() {
    bf16x8 z;
#pragma unroll
    for (int i = 0; i < 8; i++) z[i] = (__bf16)0.0f;
    return z;
}

// ---------------------------------------------------------------------------
// 3x3 convolution (forward, and the input gradient with flipped weights).
// A workgroup of kConvRows waves computes kConvRows rows x TW pixels
// (TW = 16 NT) of all COUT channels. Input channels are taken 32 at a time:
// the (kConvRows + 2) x (TW + 2) halo tile of those channels is staged in LDS once
// (coalesced 16-byte loads, zero outside the image and beyond CT; the
// upsample + concat of the decoder is resolved here), then all 9 taps read
// their B operands from LDS. The K order is (channel group, tap, channel):
// the same products as the K order of the weights, summed in another order.
// (Against one wave loading its B operands straight from global memory per
// tap, staging the tile cut a 1080p forward pass from 1.45 to 1.21 ms.)
constexpr int kConvRows = 4;              // output rows per workgroup, one wave each
constexpr int kConvThreads = 64 * kConvRows;
// Waves per SIMD the LDS of one workgroup (halo tile + staged weights) allows
// on a 160 KB CU, one wave per SIMD per workgroup: the register budget is
// held to it, so registers never cost a workgroup that the LDS would admit.
// (Wider outputs, which read their weights from global memory, keep the
// compiler's own budget.)
constexpr int conv_waves(int cout) {
    if (cout > 96) return 1;
    const int tile = (kConvRows + 2) * (16 * 4 + 2) * 4 * 16;
    const int wts = 9 * cout * 5 * 16;
    const int w = 163840 / (tile + wts);
    return w < 1 ? 1 : (w > 8 ? 8 : w);
}
template <int C0, int UP0, int C1, int COUT, int EPI>
__global__ void __launch_bounds__(kConvThreads) __attribute__((amdgpu_waves_per_eu(conv_waves(COUT)))) k_conv3x3(const __bf16* __restrict__ src0, const __bf16* __restrict__ src1,
                                                     const __bf16* __restrict__ wgt, const float* __restrict__ bias,
                                                     const __bf16* __restrict__ mask, void* __restrict__ out, uint32_t N,
                                                     uint32_t H, uint32_t W) {
    constexpr int CT = C0 + C1;
    constexpr int K = 9 * CT;
    constexpr int KPAD = (K + 31) / 32 * 32;
    constexpr int MT = COUT / 16;
    constexpr int NT = COUT <= 96 ? 4 : 2;
    constexpr int TW = 16 * NT;
    constexpr int TR = kConvRows + 2, TC = TW + 2;   // halo tile rows / columns
    constexpr int NG = (CT + 31) / 32;       // channel groups of 32
    static_assert(C0 % 8 == 0 && C1 % 8 == 0 && COUT % 16 == 0, "channel groups");
    constexpr int TILE = TR * TC * 4;
    constexpr int SPT = (TILE + kConvThreads - 1) / kConvThreads;  // staged 16-B chunks per thread
    // [row][col][4 chunks of 8 channels], the chunks of pixel p stored in the
    // order chunk ^ ((p >> 1) & 3), which spreads the fragment reads of
    // neighbouring pixels over the banks
    __shared__ bf16x8 tile[TILE];
    // Up to 48 output channels, the group's weights are staged in LDS too
    // ([tap][co] rows of 32 channels + 8 of padding, so the 16 rows a wave
    // reads at once fall on distinct banks), loaded with the tile.
    constexpr bool LDSW = COUT <= 96;
    constexpr int WROW = 5;                  // 16-B chunks per (tap, co) row
    constexpr int WCH = 9 * COUT * 4;        // weight chunks per group
    constexpr int WPT = LDSW ? (WCH + kConvThreads - 1) / kConvThreads : 1;
    __shared__ bf16x8 wl[LDSW ? 9 * COUT * WROW : 1];
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    const uint32_t r = lane & 15u, g = lane >> 4;
    const uint32_t segs = (W + TW - 1) / TW, rblocks = (H + kConvRows - 1) / kConvRows;
    const uint32_t blk = blockIdx.x;
    const uint32_t seg = blk % segs, rest = blk / segs;
    const uint32_t rb = rest % rblocks, n = rest / rblocks;
    const int y0 = (int)(rb * kConvRows), x0 = (int)(seg * TW);
    const uint32_t H0 = H / UP0, W0 = W / UP0;
    const int y = y0 + (int)wv;

    f32x4 acc[MT][NT];
#pragma unroll
    for (int m = 0; m < MT; m++)
#pragma unroll
        for (int t = 0; t < NT; t++) acc[m][t] = f32x4{0.f, 0.f, 0.f, 0.f};

    // Halo tile of channels [32 cg, 32 cg + 32): every load of a thread is
    // issued before any is stored, and before the barrier that waits for the
    // previous group's MFMAs (the staging registers are dead during them).
    bf16x8 stage[SPT];
    bf16x8 wstage[WPT];
    auto load_group = [&](int cg) {
        if constexpr (LDSW) {
#pragma unroll
            for (int i = 0; i < WPT; i++) {
                const int e = (int)threadIdx.x + i * kConvThreads;
                const int j = e & 3, row = e >> 2;     // 8-channel chunk of row (tap, co)
                const int co = row % COUT, tap = row / COUT;
                const int c = cg * 32 + j * 8;
                bf16x8 v = zero8();
                if (e < WCH && c < CT) v = *(const bf16x8*)(wgt + (size_t)co * KPAD + tap * CT + c);
                wstage[i] = v;
            }
        }
#pragma unroll
        for (int i = 0; i < SPT; i++) {
            const int e = (int)threadIdx.x + i * kConvThreads;
            const int chunk = e & 3, pix = e >> 2;
            const int ty = pix / TC, tx = pix - ty * TC;
            const int sy = y0 - 1 + ty, sx = x0 - 1 + tx;
            const int c = cg * 32 + chunk * 8;
            bf16x8 v = zero8();
            if (e < TILE && c < CT && sy >= 0 && sy < (int)H && sx >= 0 && sx < (int)W && n < N) {
                const __bf16* p;
                if (C1 == 0 || c < C0)
                    p = src0 + (((size_t)n * H0 + (uint32_t)sy / UP0) * W0 + (uint32_t)sx / UP0) * C0 + c;
                else
                    p = src1 + (((size_t)n * H + (uint32_t)sy) * W + (uint32_t)sx) * C1 + (c - C0);
                v = *(const bf16x8*)p;
            }
            stage[i] = v;
        }
    };
    for (int cg = 0; cg < NG; cg++) {
        load_group(cg);
        __syncthreads();  // every wave is done reading the previous group
#pragma unroll
        for (int i = 0; i < SPT; i++) {
            const int e = (int)threadIdx.x + i * kConvThreads;
            if (e < TILE) tile[(e & ~3) | ((e & 3) ^ ((e >> 3) & 3))] = stage[i];
        }
        if constexpr (LDSW) {
#pragma unroll
            for (int i = 0; i < WPT; i++) {
                const int e = (int)threadIdx.x + i * kConvThreads;
                if (e < WCH) wl[(e >> 2) * WROW + (e & 3)] = wstage[i];
            }
        }
        __syncthreads();
        const int c = cg * 32 + (int)g * 8;      // this lane's 8 channels
        const bool cok = c < CT;
#pragma unroll 3
        for (int tap = 0; tap < 9; tap++) {
            const int dy = tap / 3, dx = tap % 3;   // tile offsets (halo origin at -1)
            bf16x8 b[NT];
#pragma unroll
            for (int t = 0; t < NT; t++) {
                const int pix = ((int)wv + dy) * TC + t * 16 + (int)r + dx;
                b[t] = tile[pix * 4 + ((int)g ^ ((pix >> 1) & 3))];
            }
            const int k = tap * CT + c;
#pragma unroll
            for (int m = 0; m < MT; m++) {
                bf16x8 a;
                if constexpr (LDSW) a = wl[(tap * COUT + m * 16 + (int)r) * WROW + (int)g];
                else a = cok ? *(const bf16x8*)(wgt + (size_t)(m * 16 + r) * KPAD + k) : zero8();
#pragma unroll
                for (int t = 0; t < NT; t++) acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[t], acc[m][t], 0, 0, 0);
            }
        }
    }
    if (y >= (int)H || n >= N) return;
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const uint32_t x = (uint32_t)x0 + t * 16 + r;
        if (x >= W) continue;
        const size_t pix = ((size_t)n * H + (uint32_t)y) * W + x;
#pragma unroll
        for (int m = 0; m < MT; m++) {
            const int co = m * 16 + (int)g * 4;
            f32x4 v = acc[m][t];
            if (EPI == EPI_F32) {
                *(f32x4*)((float*)out + pix * COUT + co) = v;
                continue;
            }
            if (EPI == EPI_BIAS_RELU) {
#pragma unroll
                for (int i = 0; i < 4; i++) v[i] = fmaxf(v[i] + bias[co + i], 0.0f);
            } else {
                const bf16x4 mk = *(const bf16x4*)(mask + pix * COUT + co);
#pragma unroll
                for (int i = 0; i < 4; i++) v[i] = (float)mk[i] > 0.0f ? v[i] : 0.0f;
            }
            bf16x4 o;
#pragma unroll
            for (int i = 0; i < 4; i++) o[i] = (__bf16)v[i];
            *(bf16x4*)((__bf16*)out + pix * COUT + co) = o;
        }
    }
}

// ---------------------------------------------------------------------------
// Weight gradient: dW[co][tap][c] = sum over pixels p of dY[p][co] *
// X[p + offset(tap)][c], on MFMA with the pixels as K. One wave owns a 16 x 16
// tile of (co, tap*CT + c) and a slice of the pixels, and adds its partial sum
// to the FP32 gradient with atomics (the gradient is cleared first).
// ---------------------------------------------------------------------------
template <int C0, int UP0, int C1, int COUT>
__global__ void __launch_bounds__(256) k_conv_wgrad(const __bf16* __restrict__ dy, const __bf16* __restrict__ src0,
                                                    const __bf16* __restrict__ src1, float* __restrict__ dw, uint32_t cin_real,
                                                    uint32_t N, uint32_t H, uint32_t W, uint32_t splits) {
    constexpr int CT = C0 + C1;
    constexpr int K = 9 * CT;
    constexpr int MT = COUT / 16;
    constexpr int NTILES = (K + 15) / 16;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t r = lane & 15u, g = lane >> 4;
    const uint32_t wave = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (wave >= (uint32_t)(MT * NTILES) * splits) return;
    const uint32_t split = wave % splits, tile = wave / splits;
    const uint32_t m = tile % MT, ntile = tile / MT;
    const uint32_t P = N * H * W;
    const uint32_t chunk = ((P + splits - 1) / splits + 31) / 32 * 32;
    const uint32_t p_begin = split * chunk, p_end = min(P, p_begin + chunk);
    const uint32_t H0 = H / UP0, W0 = W / UP0;
    const int col = (int)(ntile * 16 + r);
    const bool col_ok = col < K;
    const int tap = col_ok ? col / CT : 0;
    const int c = col_ok ? col - tap * CT : 0;
    const int oy = tap / 3 - 1, ox = tap % 3 - 1;
    const uint32_t co = m * 16 + r;

    f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
    for (uint32_t p0 = p_begin; p0 < p_end; p0 += 32) {
        bf16x8 a, b;
        uint32_t p = p0 + 8 * g;
        uint32_t x = p % W, rest = p / W;
        uint32_t y = rest % H, n = rest / H;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            float av = 0.0f, bv = 0.0f;
            if (p + e < p_end) {
                const size_t pix = ((size_t)n * H + y) * W + x;
                av = (float)dy[pix * COUT + co];
                const int sy = (int)y + oy, sx = (int)x + ox;
                if (col_ok && sy >= 0 && sy < (int)H && sx >= 0 && sx < (int)W) {
                    if (C1 == 0 || c < C0)
                        bv = (float)src0[(((size_t)n * H0 + (uint32_t)sy / UP0) * W0 + (uint32_t)sx / UP0) * C0 + c];
                    else
                        bv = (float)src1[(((size_t)n * H + (uint32_t)sy) * W + (uint32_t)sx) * C1 + (c - C0)];
                }
            }
            a[e] = (__bf16)av;
            b[e] = (__bf16)bv;
            if (++x == W) {
                x = 0;
                if (++y == H) {
                    y = 0;
                    n++;
                }
            }
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
    // D: lane holds column (tap, c) = ntile*16 + r, rows co = m*16 + 4g + i
    const int dcol = (int)(ntile * 16 + r);
    if (dcol >= K) return;
    const int dtap = dcol / CT, dc = dcol - dtap * CT;
    if ((uint32_t)dc >= cin_real) return;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const uint32_t o = m * 16 + 4 * g + i;
        atomicAdd(&dw[((size_t)o * 9 + dtap) * cin_real + dc], acc[i]);
    }
}

// Bias gradient: column sums of dY (one thread per channel and pixel slice).
__global__ void k_colsum(const __bf16* __restrict__ g, uint32_t P, uint32_t C, float* __restrict__ db) {
    const uint32_t c = threadIdx.x;
    if (c >= C) return;
    const uint32_t slices = gridDim.x, per = (P + slices - 1) / slices;
    const uint32_t p0 = blockIdx.x * per, p1 = min(P, p0 + per);
    float s = 0.0f;
    for (uint32_t p = p0; p < p1; p++) s += (float)g[(size_t)p * C + c];
    atomicAdd(&db[c], s);
}

// ---------------------------------------------------------------------------
// Pooling and the gradients of upsampling / pooling
// ---------------------------------------------------------------------------
// 2x2 average; one thread = one output pixel x 8 channels.
__global__ void k_pool2(const __bf16* __restrict__ in, __bf16* __restrict__ out, uint32_t N, uint32_t H, uint32_t W,
                        uint32_t C) {
    const uint32_t Ho = H / 2, Wo = W / 2, G = C / 8;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)N * Ho * Wo * G) return;
    const uint32_t gi = (uint32_t)(i % G);
    size_t q = i / G;
    const uint32_t xo = (uint32_t)(q % Wo);
    q /= Wo;
    const uint32_t yo = (uint32_t)(q % Ho), n = (uint32_t)(q / Ho);
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int d = 0; d < 4; d++) {
        const size_t pix = ((size_t)n * H + 2 * yo + (d >> 1)) * W + 2 * xo + (d & 1);
        const bf16x8 v = *(const bf16x8*)(in + pix * C + gi * 8);
#pragma unroll
        for (int e = 0; e < 8; e++) s[e] += (float)v[e];
    }
    bf16x8 o;
#pragma unroll
    for (int e = 0; e < 8; e++) o[e] = (__bf16)(s[e] * 0.25f);
    *(bf16x8*)(out + i * 8) = o;
}

// Gradient of "upsample then concatenate" for the upsampled part: the sum of
// the four children (channels [0, C0) of the FP32 input gradient, CT wide),
// times the ReLU mask of the low-resolution activation -> BF16.
__global__ void k_upback_mask(const float* __restrict__ gx, uint32_t CT, const __bf16* __restrict__ act,
                              __bf16* __restrict__ gout, uint32_t N, uint32_t Hl, uint32_t Wl, uint32_t C0) {
    const uint32_t G = C0 / 8;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)N * Hl * Wl * G) return;
    const uint32_t gi = (uint32_t)(i % G);
    size_t q = i / G;
    const uint32_t xl = (uint32_t)(q % Wl);
    q /= Wl;
    const uint32_t yl = (uint32_t)(q % Hl), n = (uint32_t)(q / Hl);
    const uint32_t H = 2 * Hl, W = 2 * Wl;
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int d = 0; d < 4; d++) {
        const size_t pix = ((size_t)n * H + 2 * yl + (d >> 1)) * W + 2 * xl + (d & 1);
        const float* p = gx + pix * CT + gi * 8;
#pragma unroll
        for (int e = 0; e < 8; e++) s[e] += p[e];
    }
    const size_t lp = ((size_t)n * Hl + yl) * Wl + xl;
    const bf16x8 a = *(const bf16x8*)(act + lp * C0 + gi * 8);
    bf16x8 o;
#pragma unroll
    for (int e = 0; e < 8; e++) o[e] = (__bf16)((float)a[e] > 0.0f ? s[e] : 0.0f);
    *(bf16x8*)(gout + lp * C0 + gi * 8) = o;
}

// Gradient reaching an encoder output that is both pooled and skipped:
// (pooled-path gradient / 4 + skip gradient) times its ReLU mask -> BF16.
// gpool: FP32, half resolution, C channels; gskip: FP32, channels
// [coff, coff + C) of a CT-wide tensor at full resolution.
__global__ void k_poolback_skip_mask(const float* __restrict__ gpool, const float* __restrict__ gskip, uint32_t CT,
                                     uint32_t coff, const __bf16* __restrict__ act, __bf16* __restrict__ gout, uint32_t N,
                                     uint32_t H, uint32_t W, uint32_t C) {
    const uint32_t G = C / 8;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)N * H * W * G) return;
    const uint32_t gi = (uint32_t)(i % G);
    const size_t pix = i / G;
    const uint32_t x = (uint32_t)(pix % W);
    const size_t q = pix / W;
    const uint32_t y = (uint32_t)(q % H), n = (uint32_t)(q / H);
    const size_t pp = ((size_t)n * (H / 2) + y / 2) * (W / 2) + x / 2;
    const float* gp = gpool + pp * C + gi * 8;
    const float* gs = gskip + pix * CT + coff + gi * 8;
    const bf16x8 a = *(const bf16x8*)(act + pix * C + gi * 8);
    bf16x8 o;
#pragma unroll
    for (int e = 0; e < 8; e++) o[e] = (__bf16)((float)a[e] > 0.0f ? gp[e] * 0.25f + gs[e] : 0.0f);
    *(bf16x8*)(gout + pix * C + gi * 8) = o;
}

// ---------------------------------------------------------------------------
// Input packing, head, loss
// ---------------------------------------------------------------------------
__device__ inline void octahedral(float x, float y, float z, float& ox, float& oy) {
    const float l1 = fabsf(x) + fabsf(y) + fabsf(z);
    if (!(l1 > 1e-12f)) {
        ox = oy = 0.0f;
        return;
    }
    x /= l1;
    y /= l1;
    z /= l1;
    if (z < 0.0f) {
        ox = (1.0f - fabsf(y)) * (x >= 0.0f ? 1.0f : -1.0f);
        oy = (1.0f - fabsf(x)) * (y >= 0.0f ? 1.0f : -1.0f);
    } else {
        ox = x;
        oy = y;
    }
}

// Raw features -> network input (BF16, 16 channels, 10 used) and the FP32
// log radiance for the head's residual. Source: either raw records
// (12 floats per pixel: radiance, albedo, depth code, normal, M2, samples)
// or a surface (radiance plane + two feature planes, `held` samples).
// Pixels of the padded network image outside (h, w) are zero.
__global__ void k_pack(const float* __restrict__ raw, const float4* __restrict__ rad, const float4* __restrict__ aux,
                       uint32_t held, uint32_t N, uint32_t h, uint32_t w, uint32_t Hp, uint32_t Wp,
                       __bf16* __restrict__ xin, float4* __restrict__ resid) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)N * Hp * Wp) return;
    const uint32_t x = (uint32_t)(i % Wp);
    const size_t q = i / Wp;
    const uint32_t y = (uint32_t)(q % Hp), n = (uint32_t)(q / Hp);
    float f[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const bool valid = x < w && y < h;
    if (valid) {
        const size_t src = ((size_t)n * h + y) * w + x;
        if (raw) {
#pragma unroll
            for (int k = 0; k < 12; k++) f[k] = raw[src * 12 + k];
        } else {
            const float4 c = rad[src], a0 = aux[src], a1 = aux[(size_t)w * h + src];
            f[0] = c.x; f[1] = c.y; f[2] = c.z;
            f[3] = a0.x; f[4] = a0.y; f[5] = a0.z; f[6] = a0.w;
            f[7] = a1.x; f[8] = a1.y; f[9] = a1.z; f[10] = a1.w;
            f[11] = (float)held;
        }
    }
    float v[16];
#pragma unroll
    for (int k = 0; k < 16; k++) v[k] = 0.0f;
    if (valid) {
        v[0] = log1pf(fmaxf(f[0], 0.0f));
        v[1] = log1pf(fmaxf(f[1], 0.0f));
        v[2] = log1pf(fmaxf(f[2], 0.0f));
        v[3] = f[3];
        v[4] = f[4];
        v[5] = f[5];
        float nx = f[7], ny = f[8], nz = f[9];
        const float ln = sqrtf(nx * nx + ny * ny + nz * nz);
        if (ln > 1e-12f) {
            nx /= ln;
            ny /= ln;
            nz /= ln;
        } else {
            nx = ny = nz = 0.0f;
        }
        octahedral(nx, ny, nz, v[6], v[7]);
        v[8] = f[6];
        const float nh = f[11];
        v[9] = nh >= 2.0f ? sqrtf(fmaxf(f[10], 0.0f) / fmaxf(nh * (nh - 1.0f), 1.0f)) : 0.0f;
    }
    bf16x8 lo, hi;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        lo[k] = (__bf16)v[k];
        hi[k] = (__bf16)v[8 + k];
    }
    *(bf16x8*)(xin + i * 16) = lo;
    *(bf16x8*)(xin + i * 16 + 8) = hi;
    resid[i] = make_float4(v[0], v[1], v[2], 0.0f);
}

// 1x1 head: out = W_head (3 x 32) . act + b + residual (FP32). Optionally also
// writes the denoised radiance max(expm1(out), 0) into an RGBA32F surface.
__global__ void k_head(const __bf16* __restrict__ act, const __bf16* __restrict__ wh, const float* __restrict__ bh,
                       const float4* __restrict__ resid, float4* __restrict__ out, uint32_t N, uint32_t Hp, uint32_t Wp,
                       float4* __restrict__ surface, uint32_t h, uint32_t w) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)N * Hp * Wp) return;
    float o[3];
#pragma unroll
    for (int k = 0; k < 3; k++) o[k] = 0.0f;
#pragma unroll
    for (int gi = 0; gi < 4; gi++) {
        const bf16x8 a = *(const bf16x8*)(act + i * 32 + gi * 8);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const bf16x8 wv = *(const bf16x8*)(wh + k * 32 + gi * 8);
#pragma unroll
            for (int e = 0; e < 8; e++) o[k] += (float)wv[e] * (float)a[e];
        }
    }
    const float4 rs = resid[i];
    const float4 res = make_float4(o[0] + bh[0] + rs.x, o[1] + bh[1] + rs.y, o[2] + bh[2] + rs.z, 0.0f);
    out[i] = res;
    if (surface) {
        const uint32_t x = (uint32_t)(i % Wp);
        const size_t q = i / Wp;
        const uint32_t y = (uint32_t)(q % Hp), n = (uint32_t)(q / Hp);
        if (n == 0 && x < w && y < h)
            surface[(size_t)y * w + x] =
                make_float4(fmaxf(expm1f(res.x), 0.0f), fmaxf(expm1f(res.y), 0.0f), fmaxf(expm1f(res.z), 0.0f), 1.0f);
    }
}

// Loss over the valid pixels (0 gradient on padding); the sum goes to *loss.
//   kind 0: MSE in log1p space, dout = 2 (out - log1p(t)) / count
//   kind 1: relMSE of y = expm1(out) against t, (y - t)^2 / (t^2 + 0.01),
//           dout = 2 (y - t) (y + 1) / ((t^2 + 0.01) count)
__device__ inline float loss_term(float o, float t, uint32_t kind, float inv, float& g) {
    t = fmaxf(t, 0.0f);
    if (kind == 0u) {
        const float d = o - log1pf(t);
        g = d * inv;
        return d * d;
    }
    const float y = expm1f(o), e = y - t, wgt = 1.0f / (t * t + 0.01f);
    g = e * wgt * (y + 1.0f) * inv;
    return e * e * wgt;
}

__global__ void k_loss(const float4* __restrict__ out, const float* __restrict__ target, uint32_t N, uint32_t h, uint32_t w,
                       uint32_t Hp, uint32_t Wp, uint32_t kind, float4* __restrict__ dout, float* __restrict__ loss) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (i < (size_t)N * Hp * Wp) {
        const uint32_t x = (uint32_t)(i % Wp);
        const size_t q = i / Wp;
        const uint32_t y = (uint32_t)(q % Hp), n = (uint32_t)(q / Hp);
        float4 d = make_float4(0.f, 0.f, 0.f, 0.f);
        if (x < w && y < h) {
            const size_t src = ((size_t)n * h + y) * w + x;
            const float inv = 2.0f / (3.0f * (float)N * (float)h * (float)w);
            const float4 o = out[i];
            float g0, g1, g2;
            l = loss_term(o.x, target[src * 3 + 0], kind, inv, g0) + loss_term(o.y, target[src * 3 + 1], kind, inv, g1) +
                loss_term(o.z, target[src * 3 + 2], kind, inv, g2);
            d = make_float4(g0, g1, g2, 0.0f);
        }
        dout[i] = d;
    }
    for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
    if ((threadIdx.x & 63u) == 0) atomicAdd(loss, l);
}

// Head backward: g_act = (W_head^T dout) * mask(act) -> BF16; dW_head and
// db_head by per-thread partial sums, wave reduction and atomics.
__global__ void __launch_bounds__(256) k_head_backward(const float4* __restrict__ dout, const __bf16* __restrict__ act,
                                                       const __bf16* __restrict__ wh, __bf16* __restrict__ gact,
                                                       float* __restrict__ dwh, float* __restrict__ dbh, size_t P) {
    float sw[3][32];
    float sb[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
        for (int c = 0; c < 32; c++) sw[k][c] = 0.0f;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < P; i += (size_t)gridDim.x * blockDim.x) {
        const float4 d = dout[i];
        const float dv[3] = {d.x, d.y, d.z};
#pragma unroll
        for (int gi = 0; gi < 4; gi++) {
            const bf16x8 a = *(const bf16x8*)(act + i * 32 + gi * 8);
            bf16x8 o;
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int c = gi * 8 + e;
                float s = 0.0f;
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    s += (float)wh[k * 32 + c] * dv[k];
                    sw[k][c] += dv[k] * (float)a[e];
                }
                o[e] = (__bf16)((float)a[e] > 0.0f ? s : 0.0f);
            }
            *(bf16x8*)(gact + i * 32 + gi * 8) = o;
        }
#pragma unroll
        for (int k = 0; k < 3; k++) sb[k] += dv[k];
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
#pragma unroll
        for (int c = 0; c < 32; c++) {
            float v = sw[k][c];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
            if ((threadIdx.x & 63u) == 0) atomicAdd(&dwh[k * 32 + c], v);
        }
        float v = sb[k];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if ((threadIdx.x & 63u) == 0) atomicAdd(&dbh[k], v);
    }
}

// ---------------------------------------------------------------------------
// Parameters: Adam on the FP32 master vector, then BF16 copies in the
// forward layout [Cout][Kpad] (k = tap * CT + c) and the input-gradient
// layout [CT][Kpad'] (k' = tap' * Cout + co, holding W[co][8 - tap'][c]).
// ---------------------------------------------------------------------------
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1, float* __restrict__ m2,
                       size_t n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.0f - b1) * gi;
    const float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    p[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

__global__ void k_relayout(const float* __restrict__ w, uint32_t cout, uint32_t cin, uint32_t ct, uint32_t kpad_f,
                           __bf16* __restrict__ fwd, uint32_t kpad_b, __bf16* __restrict__ bwd) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nf = (size_t)cout * kpad_f;
    if (i < nf) {
        const uint32_t co = (uint32_t)(i / kpad_f), k = (uint32_t)(i % kpad_f);
        const uint32_t tap = k / ct, c = k % ct;
        const float v = (k < 9 * ct && c < cin) ? w[((size_t)co * 9 + tap) * cin + c] : 0.0f;
        fwd[i] = (__bf16)v;
        return;
    }
    const size_t j = i - nf;
    if (!bwd || j >= (size_t)ct * kpad_b) return;
    const uint32_t c = (uint32_t)(j / kpad_b), k = (uint32_t)(j % kpad_b);
    const uint32_t tapb = k / cout, co = k % cout;
    const float v = (k < 9 * cout && c < cin) ? w[((size_t)co * 9 + (8 - tapb)) * cin + c] : 0.0f;
    bwd[j] = (__bf16)v;
}

__global__ void k_to_bf16(const float* __restrict__ src, __bf16* __restrict__ dst, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (__bf16)src[i];
}

__global__ void k_zero_f32(float* __restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static inline uint32_t cdiv(size_t a, uint32_t b) { return (uint32_t)((a + b - 1) / b); }

template <int C0, int UP0, int C1, int COUT, int EPI>
static void conv(hipStream_t st, const __bf16* s0, const __bf16* s1, const __bf16* w, const float* b, const __bf16* mask,
                 void* out, uint32_t N, uint32_t H, uint32_t W) {
    constexpr int NT = COUT <= 96 ? 4 : 2;
    const size_t blocks = (size_t)N * ((H + kConvRows - 1) / kConvRows) * ((W + 16 * NT - 1) / (16 * NT));
    k_conv3x3<C0, UP0, C1, COUT, EPI><<<(uint32_t)blocks, kConvThreads, 0, st>>>(s0, s1, w, b, mask, out, N, H, W);
}

template <int C0, int UP0, int C1, int COUT>
static void wgrad(hipStream_t st, const __bf16* dy, const __bf16* s0, const __bf16* s1, float* dw, uint32_t cin_real, uint32_t N,
                  uint32_t H, uint32_t W) {
    constexpr int K = 9 * (C0 + C1);
    const uint32_t tiles = (COUT / 16) * ((K + 15) / 16);
    const size_t P = (size_t)N * H * W;
    // enough waves to fill the chip, each over >= 512 pixels
    uint32_t splits = (uint32_t)std::max<size_t>(1, std::min<size_t>((P + 511) / 512, (8192 + tiles - 1) / tiles));
    k_conv_wgrad<C0, UP0, C1, COUT><<<cdiv((size_t)tiles * splits, 4), 256, 0, st>>>(dy, s0, s1, dw, cin_real, N, H, W, splits);
}

// Layer table (matches oracle/unet.py LAYERS): c0, up0, c1, cout, level.
struct LayerDesc {
    int c0, up0, c1, cout, level, cin_real;
};
static const LayerDesc kLayers[kDnLayers] = {
    {16, 1, 0, 32, 0, 10},  {32, 1, 0, 32, 0, 32},  {32, 1, 0, 48, 1, 32},  {48, 1, 0, 48, 1, 48},  {48, 1, 0, 64, 2, 48},
    {64, 1, 0, 64, 2, 64},  {64, 1, 0, 96, 3, 64},  {96, 1, 0, 96, 3, 96},  {96, 2, 64, 64, 2, 160}, {64, 1, 0, 64, 2, 64},
    {64, 2, 48, 48, 1, 112}, {48, 1, 0, 48, 1, 48}, {48, 2, 32, 32, 0, 80}, {32, 1, 0, 32, 0, 32},
};

size_t net_num_params() {
    size_t n = 0;
    for (const LayerDesc& L : kLayers) n += (size_t)L.cout * 9 * L.cin_real + L.cout;
    return n + 3 * 32 + 3;
}

// He-normal weights (standard deviation sqrt(2 / fan_in)), zero biases, and
// a head at a tenth of that scale, so an untrained network returns nearly its
// noisy input (the head adds to the input's log radiance); Box-Muller over a
// 64-bit LCG, so the vector depends only on the seed.
void net_he_init(uint64_t seed, float* p) {
    uint64_t s = seed * 6364136223846793005ull + 1442695040888963407ull;
    auto uni = [&]() {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return ((double)(s >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    };
    auto fill = [&](float* w, size_t nw, size_t fan_in, double scale = 1.0) {
        const double sd = scale * std::sqrt(2.0 / (double)fan_in);
        for (size_t i = 0; i < nw; i++) {
            const double u1 = uni(), u2 = uni();
            w[i] = (float)(sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
        }
    };
    size_t off = 0;
    for (const LayerDesc& L : kLayers) {
        const size_t nw = (size_t)L.cout * 9 * L.cin_real;
        fill(p + off, nw, 9 * (size_t)L.cin_real);
        for (int i = 0; i < L.cout; i++) p[off + nw + i] = 0.0f;
        off += nw + L.cout;
    }
    fill(p + off, 96, 32, 0.1);
    for (int i = 0; i < 3; i++) p[off + 96 + i] = 0.0f;
}

static inline int kpad(int k) { return (k + 31) / 32 * 32; }

hipError_t Net::init() {
    if (master) return hipSuccess;
    nparams = net_num_params();
    HH_TRY(hipMalloc(&master, nparams * 4));
    HH_TRY(hipMalloc(&grad, nparams * 4));
    HH_TRY(hipMalloc(&adam_m, nparams * 4));
    HH_TRY(hipMalloc(&adam_v, nparams * 4));
    size_t off = 0;
    for (int l = 0; l < kDnLayers; l++) {
        const LayerDesc& L = kLayers[l];
        const int ct = L.c0 + L.c1;
        w_off[l] = off;
        off += (size_t)L.cout * 9 * L.cin_real + L.cout;
        HH_TRY(hipMalloc(&wf[l], (size_t)L.cout * kpad(9 * ct) * 2));
        HH_TRY(hipMalloc(&wb[l], (size_t)ct * kpad(9 * L.cout) * 2));
    }
    w_off[kDnLayers] = off;
    HH_TRY(hipMalloc(&whead, 96 * 2));
    return hipSuccess;
}

hipError_t Net::zero_state(hipStream_t st) {
    k_zero_f32<<<cdiv(nparams, 256), 256, 0, st>>>(adam_m, nparams);
    k_zero_f32<<<cdiv(nparams, 256), 256, 0, st>>>(adam_v, nparams);
    k_zero_f32<<<cdiv(nparams, 256), 256, 0, st>>>(grad, nparams);
    steps = 0;
    return hipGetLastError();
}

hipError_t Net::refresh_weights(hipStream_t st) {
    for (int l = 0; l < kDnLayers; l++) {
        const LayerDesc& L = kLayers[l];
        const int ct = L.c0 + L.c1;
        const size_t n = (size_t)L.cout * kpad(9 * ct) + (size_t)ct * kpad(9 * L.cout);
        k_relayout<<<cdiv(n, 256), 256, 0, st>>>(master + w_off[l], L.cout, L.cin_real, ct, kpad(9 * ct), wf[l],
                                                 kpad(9 * L.cout), wb[l]);
    }
    k_to_bf16<<<1, 128, 0, st>>>(master + w_off[kDnLayers], whead, 96);
    return hipGetLastError();
}

static hipError_t alloc_bf16(__bf16** p, size_t n) { return hipMalloc(p, n * 2); }

hipError_t Net::ensure(uint32_t N, uint32_t Hp, uint32_t Wp, bool training) {
    const size_t P = (size_t)N * Hp * Wp;
    if (P <= cap_px && (!training || cap_train)) {
        n = N;
        H = Hp;
        W = Wp;
        return hipSuccess;
    }
    release_acts();
    HH_TRY(alloc_bf16(&xin, P * 16));
    HH_TRY(hipMalloc(&resid, P * 16));
    HH_TRY(hipMalloc(&out, P * 16));
    for (int l = 0; l < kDnLayers; l++) {
        const size_t Pl = P >> (2 * kLayers[l].level);
        HH_TRY(alloc_bf16(&act[l], Pl * kLayers[l].cout));
    }
    for (int k = 0; k < 3; k++) HH_TRY(alloc_bf16(&pool[k], (P >> (2 * (k + 1))) * kLayers[2 * k + 1].cout));
    if (training) {
        HH_TRY(hipMalloc(&dout, P * 16));
        size_t gmax = 0;
        for (int l = 0; l < kDnLayers; l++) gmax = std::max(gmax, (P >> (2 * kLayers[l].level)) * kLayers[l].cout);
        HH_TRY(alloc_bf16(&gy[0], gmax));
        HH_TRY(alloc_bf16(&gy[1], gmax));
        for (int k = 0; k < 3; k++) {
            HH_TRY(hipMalloc(&gskip[k], (P >> (2 * k)) * (kLayers[12 - 2 * k].c0 + kLayers[12 - 2 * k].c1) * 4));
            HH_TRY(hipMalloc(&gpool[k], (P >> (2 * (k + 1))) * kLayers[2 * k + 1].cout * 4));
        }
        HH_TRY(hipMalloc(&loss, 64));
        cap_train = true;
    }
    cap_px = P;
    n = N;
    H = Hp;
    W = Wp;
    return hipSuccess;
}

void Net::release_acts() {
    void* ps[] = {xin, resid, out, dout, gy[0], gy[1], loss};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    xin = nullptr;
    resid = nullptr;
    out = nullptr;
    dout = nullptr;
    gy[0] = gy[1] = nullptr;
    loss = nullptr;
    for (int l = 0; l < kDnLayers; l++) {
        if (act[l]) (void)hipFree(act[l]);
        act[l] = nullptr;
    }
    for (int k = 0; k < 3; k++) {
        if (pool[k]) (void)hipFree(pool[k]);
        if (gskip[k]) (void)hipFree(gskip[k]);
        if (gpool[k]) (void)hipFree(gpool[k]);
        pool[k] = nullptr;
        gskip[k] = nullptr;
        gpool[k] = nullptr;
    }
    cap_px = 0;
    cap_train = false;
}

void Net::release() {
    release_acts();
    void* ps[] = {master, grad, adam_m, adam_v, whead};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    master = grad = adam_m = adam_v = nullptr;
    whead = nullptr;
    for (int l = 0; l < kDnLayers; l++) {
        if (wf[l]) (void)hipFree(wf[l]);
        if (wb[l]) (void)hipFree(wb[l]);
        wf[l] = wb[l] = nullptr;
    }
}

hipError_t Net::pack(hipStream_t st, const float* raw, const float4* rad, const float4* aux, uint32_t held, uint32_t h,
                     uint32_t w) {
    k_pack<<<cdiv((size_t)n * H * W, 256), 256, 0, st>>>(raw, rad, aux, held, n, h, w, H, W, xin, (float4*)resid);
    return hipGetLastError();
}

hipError_t Net::forward(hipStream_t st, float4* surface, uint32_t h, uint32_t w) {
    const float* b[kDnLayers];
    for (int l = 0; l < kDnLayers; l++) b[l] = master + w_off[l] + (size_t)kLayers[l].cout * 9 * kLayers[l].cin_real;
    const uint32_t N = n;
    auto Hl = [&](int lv) { return H >> lv; };
    auto Wl = [&](int lv) { return W >> lv; };
    auto pool2 = [&](int k, int l, int lv) {
        const uint32_t C = kLayers[l].cout;
        k_pool2<<<cdiv((size_t)N * Hl(lv + 1) * Wl(lv + 1) * (C / 8), 256), 256, 0, st>>>(act[l], pool[k], N, Hl(lv), Wl(lv), C);
    };
    conv<16, 1, 0, 32, EPI_BIAS_RELU>(st, xin, nullptr, wf[0], b[0], nullptr, act[0], N, Hl(0), Wl(0));
    conv<32, 1, 0, 32, EPI_BIAS_RELU>(st, act[0], nullptr, wf[1], b[1], nullptr, act[1], N, Hl(0), Wl(0));
    pool2(0, 1, 0);
    conv<32, 1, 0, 48, EPI_BIAS_RELU>(st, pool[0], nullptr, wf[2], b[2], nullptr, act[2], N, Hl(1), Wl(1));
    conv<48, 1, 0, 48, EPI_BIAS_RELU>(st, act[2], nullptr, wf[3], b[3], nullptr, act[3], N, Hl(1), Wl(1));
    pool2(1, 3, 1);
    conv<48, 1, 0, 64, EPI_BIAS_RELU>(st, pool[1], nullptr, wf[4], b[4], nullptr, act[4], N, Hl(2), Wl(2));
    conv<64, 1, 0, 64, EPI_BIAS_RELU>(st, act[4], nullptr, wf[5], b[5], nullptr, act[5], N, Hl(2), Wl(2));
    pool2(2, 5, 2);
    conv<64, 1, 0, 96, EPI_BIAS_RELU>(st, pool[2], nullptr, wf[6], b[6], nullptr, act[6], N, Hl(3), Wl(3));
    conv<96, 1, 0, 96, EPI_BIAS_RELU>(st, act[6], nullptr, wf[7], b[7], nullptr, act[7], N, Hl(3), Wl(3));
    conv<96, 2, 64, 64, EPI_BIAS_RELU>(st, act[7], act[5], wf[8], b[8], nullptr, act[8], N, Hl(2), Wl(2));
    conv<64, 1, 0, 64, EPI_BIAS_RELU>(st, act[8], nullptr, wf[9], b[9], nullptr, act[9], N, Hl(2), Wl(2));
    conv<64, 2, 48, 48, EPI_BIAS_RELU>(st, act[9], act[3], wf[10], b[10], nullptr, act[10], N, Hl(1), Wl(1));
    conv<48, 1, 0, 48, EPI_BIAS_RELU>(st, act[10], nullptr, wf[11], b[11], nullptr, act[11], N, Hl(1), Wl(1));
    conv<48, 2, 32, 32, EPI_BIAS_RELU>(st, act[11], act[1], wf[12], b[12], nullptr, act[12], N, Hl(0), Wl(0));
    conv<32, 1, 0, 32, EPI_BIAS_RELU>(st, act[12], nullptr, wf[13], b[13], nullptr, act[13], N, Hl(0), Wl(0));
    k_head<<<cdiv((size_t)N * H * W, 256), 256, 0, st>>>(act[13], whead, master + w_off[kDnLayers] + 96, (const float4*)resid,
                                                         (float4*)out, N, H, W, surface, h, w);
    return hipGetLastError();
}

hipError_t Net::loss_and_backward(hipStream_t st, const float* target, uint32_t h, uint32_t w, uint32_t kind) {
    const uint32_t N = n;
    const size_t P = (size_t)N * H * W;
    auto Hl = [&](int lv) { return H >> lv; };
    auto Wl = [&](int lv) { return W >> lv; };
    k_zero_f32<<<1, 64, 0, st>>>(loss, 1);
    k_zero_f32<<<cdiv(nparams, 256), 256, 0, st>>>(grad, nparams);
    k_loss<<<cdiv(P, 256), 256, 0, st>>>((const float4*)out, target, N, h, w, H, W, kind, (float4*)dout, loss);
    float* gw[kDnLayers];
    float* gb[kDnLayers];
    for (int l = 0; l < kDnLayers; l++) {
        gw[l] = grad + w_off[l];
        gb[l] = gw[l] + (size_t)kLayers[l].cout * 9 * kLayers[l].cin_real;
    }
    float* gh = grad + w_off[kDnLayers];
    __bf16* g = gy[0];
    __bf16* g2 = gy[1];
    // head
    k_head_backward<<<std::max<uint32_t>(1, std::min<uint32_t>(cdiv(P, 256), 1024)), 256, 0, st>>>(
        (const float4*)dout, act[13], whead, g, gh, gh + 96, P);
    auto bias_grad = [&](int l, const __bf16* gyl) {
        const uint32_t Pl = (uint32_t)(P >> (2 * kLayers[l].level));
        k_colsum<<<std::max<uint32_t>(1, std::min<uint32_t>(Pl / 64, 512)), 128, 0, st>>>(gyl, Pl, kLayers[l].cout, gb[l]);
    };
    // dec0b (13): input act[12]
    wgrad<32, 1, 0, 32>(st, g, act[12], nullptr, gw[13], 32, N, Hl(0), Wl(0));
    bias_grad(13, g);
    conv<32, 1, 0, 32, EPI_MASK>(st, g, nullptr, wb[13], nullptr, act[12], g2, N, Hl(0), Wl(0));
    std::swap(g, g2);
    // dec0a (12): input [up(act[11]), act[1]] -> gx (80 ch, FP32)
    wgrad<48, 2, 32, 32>(st, g, act[11], act[1], gw[12], 80, N, Hl(0), Wl(0));
    bias_grad(12, g);
    conv<32, 1, 0, 80, EPI_F32>(st, g, nullptr, wb[12], nullptr, nullptr, gskip[0], N, Hl(0), Wl(0));
    k_upback_mask<<<cdiv((size_t)N * Hl(1) * Wl(1) * 6, 256), 256, 0, st>>>(gskip[0], 80, act[11], g2, N, Hl(1), Wl(1), 48);
    std::swap(g, g2);
    // dec1b (11): input act[10]
    wgrad<48, 1, 0, 48>(st, g, act[10], nullptr, gw[11], 48, N, Hl(1), Wl(1));
    bias_grad(11, g);
    conv<48, 1, 0, 48, EPI_MASK>(st, g, nullptr, wb[11], nullptr, act[10], g2, N, Hl(1), Wl(1));
    std::swap(g, g2);
    // dec1a (10): input [up(act[9]), act[3]] -> gskip[1] (112 ch)
    wgrad<64, 2, 48, 48>(st, g, act[9], act[3], gw[10], 112, N, Hl(1), Wl(1));
    bias_grad(10, g);
    conv<48, 1, 0, 112, EPI_F32>(st, g, nullptr, wb[10], nullptr, nullptr, gskip[1], N, Hl(1), Wl(1));
    k_upback_mask<<<cdiv((size_t)N * Hl(2) * Wl(2) * 8, 256), 256, 0, st>>>(gskip[1], 112, act[9], g2, N, Hl(2), Wl(2), 64);
    std::swap(g, g2);
    // dec2b (9): input act[8]
    wgrad<64, 1, 0, 64>(st, g, act[8], nullptr, gw[9], 64, N, Hl(2), Wl(2));
    bias_grad(9, g);
    conv<64, 1, 0, 64, EPI_MASK>(st, g, nullptr, wb[9], nullptr, act[8], g2, N, Hl(2), Wl(2));
    std::swap(g, g2);
    // dec2a (8): input [up(act[7]), act[5]] -> gskip[2] (160 ch)
    wgrad<96, 2, 64, 64>(st, g, act[7], act[5], gw[8], 160, N, Hl(2), Wl(2));
    bias_grad(8, g);
    conv<64, 1, 0, 160, EPI_F32>(st, g, nullptr, wb[8], nullptr, nullptr, gskip[2], N, Hl(2), Wl(2));
    k_upback_mask<<<cdiv((size_t)N * Hl(3) * Wl(3) * 12, 256), 256, 0, st>>>(gskip[2], 160, act[7], g2, N, Hl(3), Wl(3), 96);
    std::swap(g, g2);
    // midb (7): input act[6]
    wgrad<96, 1, 0, 96>(st, g, act[6], nullptr, gw[7], 96, N, Hl(3), Wl(3));
    bias_grad(7, g);
    conv<96, 1, 0, 96, EPI_MASK>(st, g, nullptr, wb[7], nullptr, act[6], g2, N, Hl(3), Wl(3));
    std::swap(g, g2);
    // mida (6): input pool[2] -> gpool[2] (FP32)
    wgrad<64, 1, 0, 96>(st, g, pool[2], nullptr, gw[6], 64, N, Hl(3), Wl(3));
    bias_grad(6, g);
    conv<96, 1, 0, 64, EPI_F32>(st, g, nullptr, wb[6], nullptr, nullptr, gpool[2], N, Hl(3), Wl(3));
    // enc2b (5): gradient = pooled path + skip of dec2a (channels 96..159)
    k_poolback_skip_mask<<<cdiv((size_t)N * Hl(2) * Wl(2) * 8, 256), 256, 0, st>>>(gpool[2], gskip[2], 160, 96, act[5], g, N,
                                                                                  Hl(2), Wl(2), 64);
    wgrad<64, 1, 0, 64>(st, g, act[4], nullptr, gw[5], 64, N, Hl(2), Wl(2));
    bias_grad(5, g);
    conv<64, 1, 0, 64, EPI_MASK>(st, g, nullptr, wb[5], nullptr, act[4], g2, N, Hl(2), Wl(2));
    std::swap(g, g2);
    // enc2a (4): input pool[1] -> gpool[1]
    wgrad<48, 1, 0, 64>(st, g, pool[1], nullptr, gw[4], 48, N, Hl(2), Wl(2));
    bias_grad(4, g);
    conv<64, 1, 0, 48, EPI_F32>(st, g, nullptr, wb[4], nullptr, nullptr, gpool[1], N, Hl(2), Wl(2));
    // enc1b (3): pooled path + skip of dec1a (channels 64..111)
    k_poolback_skip_mask<<<cdiv((size_t)N * Hl(1) * Wl(1) * 6, 256), 256, 0, st>>>(gpool[1], gskip[1], 112, 64, act[3], g, N,
                                                                                  Hl(1), Wl(1), 48);
    wgrad<48, 1, 0, 48>(st, g, act[2], nullptr, gw[3], 48, N, Hl(1), Wl(1));
    bias_grad(3, g);
    conv<48, 1, 0, 48, EPI_MASK>(st, g, nullptr, wb[3], nullptr, act[2], g2, N, Hl(1), Wl(1));
    std::swap(g, g2);
    // enc1a (2): input pool[0] -> gpool[0]
    wgrad<32, 1, 0, 48>(st, g, pool[0], nullptr, gw[2], 32, N, Hl(1), Wl(1));
    bias_grad(2, g);
    conv<48, 1, 0, 32, EPI_F32>(st, g, nullptr, wb[2], nullptr, nullptr, gpool[0], N, Hl(1), Wl(1));
    // enc0b (1): pooled path + skip of dec0a (channels 48..79)
    k_poolback_skip_mask<<<cdiv((size_t)N * Hl(0) * Wl(0) * 4, 256), 256, 0, st>>>(gpool[0], gskip[0], 80, 48, act[1], g, N,
                                                                                  Hl(0), Wl(0), 32);
    wgrad<32, 1, 0, 32>(st, g, act[0], nullptr, gw[1], 32, N, Hl(0), Wl(0));
    bias_grad(1, g);
    conv<32, 1, 0, 32, EPI_MASK>(st, g, nullptr, wb[1], nullptr, act[0], g2, N, Hl(0), Wl(0));
    std::swap(g, g2);
    // enc0a (0): weights only
    wgrad<16, 1, 0, 32>(st, g, xin, nullptr, gw[0], 10, N, Hl(0), Wl(0));
    bias_grad(0, g);
    return hipGetLastError();
}

hipError_t Net::adam(hipStream_t st, float lr) {
    steps++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)steps), c2 = 1.0f - powf(b2, (float)steps);
    k_adam<<<cdiv(nparams, 256), 256, 0, st>>>(master, grad, adam_m, adam_v, nparams, lr, b1, b2, 1e-8f, c1, c2);
    HH_TRY(hipGetLastError());
    return refresh_weights(st);
}

}  // namespace dn
}  // namespace hh
