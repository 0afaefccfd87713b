// vr_nrc.hpp -- neural radiance cache: a fully fused 96-64-64-64-64-16 ReLU MLP
// on BF16 MFMA (v_mfma_f32_16x16x32_bf16, FP32 accumulation) for inference and
// online training, plus the path-tracing kernels that feed it.
//
// The cache predicts Q(x, w): the radiance a NEE path collects from a real
// collision at x (reached along direction w, throughput 1 on arrival) onwards,
// i.e. the quantity the NEE integrator (include/vr_path_integrator_NEE.h:39-63)
// adds from that vertex on.  Rendering terminates each path at its
// (terminate_after + 1)-th collision and adds W * Q(x, w) there; training
// targets are the suffix radiances of plain NEE paths, one per collision vertex.
//
// Layout.  Weights are stored per layer as Wt[n][k] (output neuron n, input k,
// row-major), in FP32 (master, optimised by Adam) and BF16 (read by the MFMA
// kernels, staged in LDS as rows padded by 8 elements).  A wave works on tiles
// of 16 rows whose activations stay in registers (see the inference section).
//
// MFMA 16x16x32 BF16 lane layout (checked on the device): lane l holds
// A[l%16][8*(l/16) + j] and B[8*(l/16) + j][l%16] for j < 8, and returns
// C[4*(l/16) + j][l%16] for j < 4.
#pragma once

#include "vr_device.hpp"

namespace vr {
namespace nrc {

constexpr int kIn = 96;      // 72 Fourier features + direction (3) + constant 1 + zero padding
constexpr int kH = 64;       // width of the 4 hidden layers
constexpr int kOut = 16;     // output tile; the first 3 are RGB
constexpr int kFreqs = 12;   // Fourier frequencies per position coordinate
constexpr int kTileRows = 16;  // rows per MFMA tile (one row per lane of a 16-lane group)
constexpr int kPad = 8;      // LDS row padding (elements)

// parameter offsets of Wt0 [64][96], Wt1..Wt3 [64][64], Wt4 [16][64]
constexpr int kOff0 = 0;
constexpr int kOff1 = kOff0 + kH * kIn;
constexpr int kOff2 = kOff1 + kH * kH;
constexpr int kOff3 = kOff2 + kH * kH;
constexpr int kOff4 = kOff3 + kH * kH;
constexpr int kParams = kOff4 + kOut * kH;  // 19456

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

VR_DEV float fract(float t) { return t - floorf(t); }

// Features of a point u in the unit cube (the medium box mapped to [0,1]^3) and
// a direction w: sin and cos of pi * 2^f * u_c for f < 12, then w, then 1.
VR_DEV float feature(f3 u, f3 w, int i) {
    if (i < 3 * 2 * kFreqs) {
        const int c = i / (2 * kFreqs), f = (i / 2) % kFreqs;
        const float uc = c == 0 ? u.x : (c == 1 ? u.y : u.z);
        float s, co;
        sincos_turns(fract(uc * __builtin_ldexpf(1.0f, f - 1)), s, co);  // pi 2^f u = 2 pi (2^(f-1) u)
        return (i & 1) ? co : s;
    }
    if (i == 72) return w.x;
    if (i == 73) return w.y;
    if (i == 74) return w.z;
    if (i == 75) return 1.0f;
    return 0.0f;
}

VR_DEV f3 unit_coords(f3 x, f3 bmin, f3 ibsize) {
    return F3((x.x - bmin.x) * ibsize.x, (x.y - bmin.y) * ibsize.y, (x.z - bmin.z) * ibsize.z);
}

// Weight staging.  The global weights (BF16 [kParams], 16-byte aligned) are
// read as 16-byte chunks (8 consecutive k of one row n of a layer Wt [N][K]):
// thread t of a block of THREADS holds chunks t, t + THREADS, ...; the loads
// are issued first (fetch_weights) and the chunks written to LDS later
// (stage_weights), so a kernel can do other work while they are in flight.
// Each layer goes to LDS with row stride K + kPad and, if `wT` is given
// (layers 1..4), transposed [K][N] with row stride N + kPad.  Within a layer
// the chunks run down the rows first (row n = i % N, columns 8 (i / N) ..), so
// a wave's transposed stores cover consecutive n (one LDS row segment) and its
// natural stores start at distinct banks.  Layers hold multiples of 64 chunks,
// so the layer branches are uniform per wave.
constexpr int kWChunks = kParams / 8;
template <int THREADS>
constexpr int kWPer = (kWChunks + THREADS - 1) / THREADS;

// Layer, row and first column of chunk c.
VR_DEV void chunk_rc(int c, int& layer, int& n, int& k) {
    constexpr int c1 = kOff1 / 8, c2 = kOff2 / 8, c3 = kOff3 / 8, c4 = kOff4 / 8;
    static_assert(c1 % 64 == 0 && (c2 - c1) % 64 == 0 && (c4 - c3) % 64 == 0 && (kWChunks - c4) % 64 == 0,
                  "layers hold whole waves of chunks");
    int i, N;
    if (c < c1) { layer = 0; i = c; N = kH; }
    else if (c < c4) { layer = 1 + (c - c1) / (c2 - c1); i = (c - c1) % (c2 - c1); N = kH; }
    else { layer = 4; i = c - c4; N = kOut; }
    n = i % N;
    k = 8 * (i / N);
}

template <int THREADS>
VR_DEV void fetch_weights(const __bf16* __restrict__ g, bf16x8* v) {
    const int offs[5] = {kOff0, kOff1, kOff2, kOff3, kOff4};
#pragma unroll
    for (int t = 0; t < kWPer<THREADS>; t++) {
        const int c = (int)threadIdx.x + t * THREADS;
        if (c < kWChunks) {
            int layer, n, k;
            chunk_rc(c, layer, n, k);
            v[t] = *(const bf16x8*)(g + offs[layer] + n * (layer == 0 ? kIn : kH) + k);
        }
    }
}

template <int K, int N>
VR_DEV void stage_chunk(int n, int k, bf16x8 v, __bf16* __restrict__ wt, __bf16* __restrict__ wT) {
    *(bf16x8*)(wt + n * (K + kPad) + k) = v;
    if (wT) {
#pragma unroll
        for (int j = 0; j < 8; j++) wT[(k + j) * (N + kPad) + n] = v[j];
    }
}

// LDS footprint of the weights, one layout (elements).
constexpr int kWLds = kH * (kIn + kPad) + 3 * kH * (kH + kPad) + kOut * (kH + kPad);
// LDS offsets of the layers in the [N][K+pad] layout
constexpr int kL0 = 0;
constexpr int kL1 = kL0 + kH * (kIn + kPad);
constexpr int kL2 = kL1 + kH * (kH + kPad);
constexpr int kL3 = kL2 + kH * (kH + kPad);
constexpr int kL4 = kL3 + kH * (kH + kPad);
// LDS offsets in the transposed [K][N+pad] layout (Wt0 is never needed transposed)
constexpr int kT1 = 0;
constexpr int kT2 = kT1 + kH * (kH + kPad);
constexpr int kT3 = kT2 + kH * (kH + kPad);
constexpr int kT4 = kT3 + kH * (kH + kPad);
constexpr int kWTLds = kT4 + kH * (kOut + kPad);

template <int THREADS>
VR_DEV void stage_weights(const bf16x8* v, __bf16* __restrict__ w, __bf16* __restrict__ wT) {
#pragma unroll
    for (int t = 0; t < kWPer<THREADS>; t++) {
        const int c = (int)threadIdx.x + t * THREADS;
        if (c >= kWChunks) continue;
        int layer, n, k;
        chunk_rc(c, layer, n, k);
        if (layer == 0) stage_chunk<kIn, kH>(n, k, v[t], w + kL0, nullptr);
        else if (layer == 1) stage_chunk<kH, kH>(n, k, v[t], w + kL1, wT ? wT + kT1 : nullptr);
        else if (layer == 2) stage_chunk<kH, kH>(n, k, v[t], w + kL2, wT ? wT + kT2 : nullptr);
        else if (layer == 3) stage_chunk<kH, kH>(n, k, v[t], w + kL3, wT ? wT + kT3 : nullptr);
        else stage_chunk<kH, kOut>(n, k, v[t], w + kL4, wT ? wT + kT4 : nullptr);
    }
}

// fetch_weights + stage_weights (the caller synchronises the block afterwards).
template <int THREADS>
VR_DEV void load_weights(const __bf16* __restrict__ g, __bf16* __restrict__ w, __bf16* __restrict__ wT) {
    bf16x8 v[kWPer<THREADS>];
    fetch_weights<THREADS>(g, v);
    stage_weights<THREADS>(v, w, wT);
}

// Copies of the FP32 weight gradient the training blocks add into: block b
// adds into copy b % kGradCopies.  Float atomics are performed memory-side, and
// adds to one address serialise there; with one set of adds per block (a step
// of 24576 records runs 192 blocks) 16 copies keep the queues 12 deep (the
// Adam kernel sums the copies).
constexpr int kGradCopies = 16;

// Slots of a training step's loss buffer: the training blocks add their
// partial sums into slots 1.., the Adam kernel leaves their total in slot 0.
constexpr int kLossSlots = 64;

// Zeroes the gradient copies and the loss slots of a training step.
__global__ void k_nrc_zero(float* __restrict__ g, float* __restrict__ loss) {
    const int i0 = blockIdx.x * blockDim.x + threadIdx.x;
    for (int i = i0; i < kGradCopies * kParams; i += gridDim.x * blockDim.x) g[i] = 0.0f;
    if (i0 < kLossSlots) loss[i0] = 0.0f;
}

// Adam on the FP32 master weights (torch.optim.Adam semantics) and the BF16
// copy, on the sum of the gradient copies g[c][i]; the sum is left in copy 0,
// and the loss slots' total in loss[0].
__global__ void k_nrc_adam(float* __restrict__ p, float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                           __bf16* __restrict__ wg, float lr, float b1, float b2, float eps, float bc1, float bc2_sqrt,
                           float* __restrict__ loss) {
    if (blockIdx.x == 0 && threadIdx.x < 64) {
        float t = threadIdx.x < kLossSlots - 1 ? loss[1 + threadIdx.x] : 0.0f;
        for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off);
        if (threadIdx.x == 0) loss[0] = t;
    }
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < kParams; i += gridDim.x * blockDim.x) {
        float gi = g[i];
#pragma unroll
        for (int c = 1; c < kGradCopies; c++) gi += g[(size_t)c * kParams + i];
        g[i] = gi;
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float pi = p[i] - (lr / bc1) * mi / (sqrtf(vi) / bc2_sqrt + eps);
        p[i] = pi;
        wg[i] = (__bf16)pi;
    }
}

// --------------------------------------------------- inference (registers) --
// Inference keeps a tile's activations in registers from layer to layer.  A
// wave runs 16 rows per tile: lane l owns row r = l % 16, and its group
// g = l / 16 holds a quarter of each layer's columns.  Each layer is computed
// transposed, Z^T = Wt X^T: the MFMA's A operand is a 16 x 32 block of Wt (from
// LDS) and its B operand the rows' inputs, so the output tile leaves lane (g, r)
// holding Z[r][16 nt + 4 g + j] (j < 4) -- the layout of the next layer's B
// operand once the 32 inputs of a K-block are taken in the slot order
// pi(8 g + j) = 4 g + j (j < 4), 16 + 4 g + j - 4 (j >= 4).  The next layer's
// weights are read in that order (two 8-byte LDS reads per fragment), so no
// activation passes through LDS.  Layer 0's inputs are the features, computed
// per lane in natural order.

typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;

// A fragment of layer 0 (K = 96, natural slot order): Wt0[16 nt + r][32 ks + 8 g + j].
VR_DEV bf16x8 wfrag_in(const __bf16* __restrict__ w, int nt, int ks) {
    const int l = __lane_id();
    return *(const bf16x8*)(w + (nt * 16 + l % 16) * (kIn + kPad) + ks * 32 + (l / 16) * 8);
}

// A fragment of a layer fed by a previous layer (K = 64, slot order pi).
VR_DEV bf16x8 wfrag_hidden(const __bf16* __restrict__ w, int nt, int ks) {
    const int l = __lane_id();
    const __bf16* row = w + (nt * 16 + l % 16) * (kH + kPad) + ks * 32 + (l / 16) * 4;
    const bf16x4 lo = *(const bf16x4*)row, hi = *(const bf16x4*)(row + 16);
    bf16x8 v;
    v[0] = lo[0]; v[1] = lo[1]; v[2] = lo[2]; v[3] = lo[3];
    v[4] = hi[0]; v[5] = hi[1]; v[6] = hi[2]; v[7] = hi[3];
    return v;
}

// ReLU of two output tiles (columns 32 kb + 4 g + j and 32 kb + 16 + 4 g + j)
// packed as the next layer's B fragment of K-block kb.
VR_DEV bf16x8 relu_pack(floatx4 a, floatx4 b) {
    bf16x8 v;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        v[j] = (__bf16)fmaxf(a[j], 0.0f);
        v[4 + j] = (__bf16)fmaxf(b[j], 0.0f);
    }
    return v;
}

// The network on one 16-row tile whose features are x[0..2] (B fragments of
// the 3 input K-blocks, natural order); returns the output tile: lanes of
// group 0 hold Y[r][0..3] (RGB = the first 3).
VR_DEV floatx4 mlp_tile(const __bf16* __restrict__ w, const bf16x8* x) {
    floatx4 c[4];
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
        c[nt] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ks = 0; ks < 3; ks++) c[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wfrag_in(w + kL0, nt, ks), x[ks], c[nt], 0, 0, 0);
    }
    bf16x8 h[2];
    h[0] = relu_pack(c[0], c[1]);
    h[1] = relu_pack(c[2], c[3]);
    const int offs[3] = {kL1, kL2, kL3};
#pragma unroll
    for (int L = 0; L < 3; L++) {
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            c[nt] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int ks = 0; ks < 2; ks++)
                c[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wfrag_hidden(w + offs[L], nt, ks), h[ks], c[nt], 0, 0, 0);
        }
        h[0] = relu_pack(c[0], c[1]);
        h[1] = relu_pack(c[2], c[3]);
    }
    floatx4 y = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int ks = 0; ks < 2; ks++) y = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wfrag_hidden(w + kL4, 0, ks), h[ks], y, 0, 0, 0);
    return y;
}

// Feature fragments of this lane's row: slots 8 g + j of the 3 K-blocks.
VR_DEV void encode_frags(f3 u, f3 wd, bool valid, bf16x8* x) {
    // lane group g of K-block ks holds features 8 p .. 8 p + 7 with p = 4 ks + g:
    // for p < 9 the sin/cos pairs of frequencies 4 (p % 3) .. + 3 of coordinate
    // p / 3 (the same expressions as feature()), for p = 9 the direction and
    // the constant 1, zeros above
    const int g = (int)__lane_id() / 16;
#pragma unroll
    for (int ks = 0; ks < 3; ks++) {
        const int p = 4 * ks + g;
        if (p < 9) {
            const int c = p / 3, f0 = 4 * (p - 3 * c);
            const float uc = c == 0 ? u.x : (c == 1 ? u.y : u.z);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                float sn, cs;
                sincos_turns(fract(uc * __builtin_ldexpf(1.0f, f0 + q - 1)), sn, cs);
                x[ks][2 * q] = (__bf16)(valid ? sn : 0.0f);
                x[ks][2 * q + 1] = (__bf16)(valid ? cs : 0.0f);
            }
        } else {
            const bool dir = p == 9 && valid;
            x[ks][0] = (__bf16)(dir ? wd.x : 0.0f);
            x[ks][1] = (__bf16)(dir ? wd.y : 0.0f);
            x[ks][2] = (__bf16)(dir ? wd.z : 0.0f);
            x[ks][3] = (__bf16)(dir ? 1.0f : 0.0f);
#pragma unroll
            for (int j = 4; j < 8; j++) x[ks][j] = (__bf16)0.0f;
        }
    }
}

constexpr int kInferWaves = 8;  // waves per block (the block shares the weights in LDS)

// MLP on explicit feature vectors [n][96] (FP32, rounded to BF16) -> out [n][3].
__global__ void __launch_bounds__(64 * kInferWaves) k_nrc_mlp_fw(const float* __restrict__ feats, int64_t n,
                                                                 const __bf16* __restrict__ wg, float* __restrict__ out) {
    __shared__ __bf16 w[kWLds];
    load_weights<64 * kInferWaves>(wg, w, nullptr);
    __syncthreads();
    const int l = __lane_id(), r = l % 16, g = l / 16;
    const int64_t nw = (int64_t)gridDim.x * kInferWaves;
    for (int64_t base = ((int64_t)blockIdx.x * kInferWaves + threadIdx.x / 64) * 16; base < n; base += nw * 16) {
        const int64_t row = base + r;
        const bool valid = row < n;
        bf16x8 x[3];
#pragma unroll
        for (int ks = 0; ks < 3; ks++) {
            const float* f = feats + (valid ? row : 0) * kIn + ks * 32 + g * 8;
            const float4 a = *(const float4*)f, b = *(const float4*)(f + 4);
            const float v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
            for (int j = 0; j < 8; j++) x[ks][j] = (__bf16)(valid ? v[j] : 0.0f);
        }
        const floatx4 y = mlp_tile(w, x);
        if (g == 0 && valid) {
            out[row * 3] = y[0];
            out[row * 3 + 1] = y[1];
            out[row * 3 + 2] = y[2];
        }
    }
}

// Features of explicit (x, w) pairs [n][3] -> [n][96] (FP32).
__global__ void k_nrc_encode(const float* __restrict__ x, const float* __restrict__ w, int64_t n, f3 bmin, f3 ibsize,
                             float* __restrict__ feats) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * kIn) return;
    const int64_t r = i / kIn;
    const int c = (int)(i - r * kIn);
    feats[i] = feature(unit_coords(F3(x[3 * r], x[3 * r + 1], x[3 * r + 2]), bmin, ibsize),
                       F3(w[3 * r], w[3 * r + 1], w[3 * r + 2]), c);
}

// Cache queries (x[3], w[3], weight[3], pixel bits) -> out[pixel] += weight * max(Q(x, w), 0).
__global__ void __launch_bounds__(64 * kInferWaves) k_nrc_apply(const float* __restrict__ queries, const uint32_t* __restrict__ count_p,
                                                                int64_t max_queries, f3 bmin, f3 ibsize,
                                                                const __bf16* __restrict__ wg, float* __restrict__ out) {
    __shared__ __bf16 w[kWLds];
    load_weights<64 * kInferWaves>(wg, w, nullptr);
    __syncthreads();
    const int64_t count = min((int64_t)*count_p, max_queries);
    const int l = __lane_id(), r = l % 16;
    const int64_t stride = (int64_t)gridDim.x * kInferWaves * 16;
    // each lane's query row of the wave's tile, read one tile ahead so that the
    // load overlaps the current tile's network
    auto fetch = [&](int64_t b, float* d) {
        const int64_t row = b + r;
        const float* q = queries + (row < count ? row : 0) * 10;
#pragma unroll
        for (int i = 0; i < 10; i++) d[i] = q[i];
    };
    int64_t base = ((int64_t)blockIdx.x * kInferWaves + threadIdx.x / 64) * 16;
    float q[10];
    if (base < count) fetch(base, q);
    for (; base < count; base += stride) {
        const bool valid = base + r < count;
        float qn[10];
        if (base + stride < count) fetch(base + stride, qn);
        bf16x8 x[3];
        encode_frags(unit_coords(F3(q[0], q[1], q[2]), bmin, ibsize), F3(q[3], q[4], q[5]), valid, x);
        const floatx4 y = mlp_tile(w, x);
        if (l < 16 && valid) {
            const uint32_t pix = __float_as_uint(q[9]);
            atomic_addf(out + 3 * (size_t)pix, q[6] * fmaxf(y[0], 0.0f));
            atomic_addf(out + 3 * (size_t)pix + 1, q[7] * fmaxf(y[1], 0.0f));
            atomic_addf(out + 3 * (size_t)pix + 2, q[8] * fmaxf(y[2], 0.0f));
        }
#pragma unroll
        for (int i = 0; i < 10; i++) q[i] = qn[i];
    }
}

// ------------------------------------------------------------- training --
// A training step runs the inference tile (above) forward, keeping each
// layer's input fragments, then back-propagates in the same transposed form:
// dX^T = Wt^T dZ^T with the transposed weights as the A operand (read in the
// K-slot order of dZ's tiles) and dZ packed as the B operand, so dX lands in
// the output layout and the ReLU masks are lane-local.  The weight gradient
// dW = dZ^T X reduces over the rows of the block's 8 tiles: the waves write dZ
// and X to LDS transposed and share the 16 x 16 output blocks, each one a chain
// of 16 x 16 x 32 MFMAs whose K dimension is the block's rows, added into the
// block's gradient copy (16 consecutive floats per lane group).

constexpr int kTrainWaves = 8;  // waves per block (they share both weight layouts in LDS)
constexpr int kBlockRows = kTileRows * kTrainWaves;  // rows of a block's tiles
constexpr int kBlockStride = kBlockRows + kPad;      // LDS row stride (elements) of a transposed column

// A fragment of a layer's transposed weights wT [K][NIN + kPad] (K-tile kt,
// dZ block kb): lane (g, i) holds wT[16 kt + i][32 kb + pi(8 g + j)], zeros
// past NIN.
template <int NIN>
VR_DEV bf16x8 wtfrag(const __bf16* __restrict__ wT, int kt, int kb) {
    const int l = __lane_id();
    const __bf16* row = wT + (kt * 16 + l % 16) * (NIN + kPad) + kb * 32 + (l / 16) * 4;
    const bf16x4 lo = *(const bf16x4*)row;
    bf16x8 v;
    v[0] = lo[0]; v[1] = lo[1]; v[2] = lo[2]; v[3] = lo[3];
    if (kb * 32 + 16 < NIN) {
        const bf16x4 hi = *(const bf16x4*)(row + 16);
        v[4] = hi[0]; v[5] = hi[1]; v[6] = hi[2]; v[7] = hi[3];
    } else {
        v[4] = v[5] = v[6] = v[7] = (__bf16)0.0f;
    }
    return v;
}

// Two output-layout tiles as one B fragment (columns 32 kb + pi(8 g + j)).
VR_DEV bf16x8 pack2(floatx4 a, floatx4 b) {
    bf16x8 v;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        v[j] = (__bf16)a[j];
        v[4 + j] = (__bf16)b[j];
    }
    return v;
}

// dZ of the layer below: dX (4 output-layout tiles, 64 columns) masked by that
// layer's ReLU outputs h (its B fragments), packed as B fragments.
VR_DEV void mask_pack(const floatx4* dx, const bf16x8* h, bf16x8* dz) {
#pragma unroll
    for (int kb = 0; kb < 2; kb++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            dz[kb][j] = (__bf16)((float)h[kb][j] > 0.0f ? dx[2 * kb][j] : 0.0f);
            dz[kb][4 + j] = (__bf16)((float)h[kb][4 + j] > 0.0f ? dx[2 * kb + 1][j] : 0.0f);
        }
}

// dX^T = Wt^T dZ^T for the 64 inputs of a layer with NIN outputs (NIN = 16: the
// output layer, one dZ block; 64: a hidden layer, two).
template <int NIN>
VR_DEV void back_layer(const __bf16* __restrict__ wT, const bf16x8* dz, floatx4* dx) {
    constexpr int KB = (NIN + 31) / 32;
#pragma unroll
    for (int kt = 0; kt < 4; kt++) {
        dx[kt] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kb = 0; kb < KB; kb++) dx[kt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wtfrag<NIN>(wT, kt, kb), dz[kb], dx[kt], 0, 0, 0);
    }
}

// dW[n][k] += sum over the block's rows of dZ[r][n] X[r][k] for a layer of N
// outputs (dZ: N/32 or one half-filled B fragment) and K inputs (X: K/32 B
// fragments, natural slot order when NATURAL, else pi), into gw (row stride K),
// in two parts.  dw_post: every wave writes its tile's dZ and X to LDS
// transposed (zt [N][kBlockStride], xt [K][kBlockStride]: one row of kBlockRows
// values per column, the wave's 16 rows at columns 16 w ..).  dw_issue: after a
// barrier the waves share the (N/16) x (K/16) output blocks, each reduced over
// all the block's rows by 16 x 16 x 32 MFMAs and added into the gradient copy --
// one set of atomics per block instead of one per wave.  Only the RGB rows of
// the output layer (n < 3) carry gradient; its other rows are not added.  The
// caller alternates two LDS buffers between layers, so the barrier of the next
// layer's dw_issue also orders this layer's reads before that buffer's reuse.
template <int N, int K, bool NATURAL>
VR_DEV void dw_post(const bf16x8* dz, const bf16x8* x, __bf16* __restrict__ zt, __bf16* __restrict__ xt) {
    const int l = __lane_id(), r = l % 16, g = l / 16;
    constexpr int ZB = (N + 31) / 32, XB = K / 32;
    const int col = 16 * (int)(threadIdx.x / 64) + r;
#pragma unroll
    for (int kb = 0; kb < ZB; kb++)
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int n = kb * 32 + (j < 4 ? 4 * g + j : 16 + 4 * g + j - 4);
            if (n < N) zt[n * kBlockStride + col] = dz[kb][j];
        }
#pragma unroll
    for (int kb = 0; kb < XB; kb++)
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int k = kb * 32 + (NATURAL ? 8 * g + j : (j < 4 ? 4 * g + j : 16 + 4 * g + j - 4));
            xt[k * kBlockStride + col] = x[kb][j];
        }
}

template <int N, int K>
VR_DEV void dw_issue(const __bf16* __restrict__ zt, const __bf16* __restrict__ xt, float* __restrict__ gw) {
    const int l = __lane_id(), r = l % 16, g = l / 16;
    __syncthreads();
    constexpr int KT = K / 16, NBLK = (N / 16) * KT;
#pragma unroll 1
    for (int b = (int)(threadIdx.x / 64); b < NBLK; b += kTrainWaves) {
        const int nt = b / KT, kt = b - nt * KT;
        const __bf16* za = zt + (nt * 16 + r) * kBlockStride + 8 * g;
        const __bf16* xb = xt + (kt * 16 + r) * kBlockStride + 8 * g;
        floatx4 c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ks = 0; ks < kBlockRows / 32; ks++)
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(*(const bf16x8*)(za + 32 * ks), *(const bf16x8*)(xb + 32 * ks), c, 0, 0, 0);
        float* gr = gw + (nt * 16 + 4 * g) * K + kt * 16 + r;
#pragma unroll
        for (int jj = 0; jj < 4; jj++)
            if (N > kOut || 4 * g + jj < 3) atomic_addf(gr + jj * K, c[jj]);
    }
}

// Feature fragments of this lane's row of the wave's tile in the block tile at
// bbase (zeros past count).
VR_DEV void tile_features(const float* __restrict__ records, int64_t count, int64_t bbase, f3 bmin, f3 ibsize, bf16x8* x) {
    const int64_t row = bbase + (int64_t)(threadIdx.x / 64) * kTileRows + __lane_id() % 16;
    const bool valid = row < count;
    const float* rec = records + (valid ? row : 0) * 9;
    encode_frags(unit_coords(F3(rec[0], rec[1], rec[2]), bmin, ibsize), F3(rec[3], rec[4], rec[5]), valid, x);
}

// One optimisation step over the training records (x[3], w[3], target[3]) of
// rows [f, min(count, f + max_records)) with f = first % count (so that
// consecutive windows cycle through however many records were written):
// relative L2 loss sum_c (q_c - t_c)^2 / (q_c^2 + 0.01) averaged over rows and
// channels (q_c without gradient in the denominator), its gradient accumulated
// into the kGradCopies copies at `grad` (FP32 [kGradCopies][kParams], zeroed by
// the caller) and the loss into *loss.
__global__ void __launch_bounds__(64 * kTrainWaves) k_nrc_train(const float* __restrict__ records, const uint32_t* __restrict__ count_p,
                                                                int64_t first, int64_t max_records, f3 bmin, f3 ibsize,
                                                                const __bf16* __restrict__ wg, float* __restrict__ grad,
                                                                float* __restrict__ loss) {
    __shared__ __bf16 w_lds[kWLds];
    __shared__ __bf16 wT_lds[kWTLds];
    // two transposition buffers alternated by layer (A: dZ and up to 96
    // inputs, B: dZ and 64 inputs).  Their pointers carry an offset the
    // compiler cannot see through, so it keeps the LDS accesses of consecutive
    // layers in program order (with provably disjoint buffers it interleaves
    // them and runs out of registers).
    __shared__ __bf16 tbuf[(kH + kIn + 2 * kH) * kBlockStride];
    int toff;
    asm volatile("s_mov_b32 %0, 0" : "=s"(toff));
    __bf16* ztA = tbuf + toff;
    __bf16* xtA = ztA + kH * kBlockStride;
    __bf16* ztB = xtA + kIn * kBlockStride;
    __bf16* xtB = ztB + kH * kBlockStride;
    // the weight loads are issued first, staged into LDS after the first
    // tile's features are computed
    bf16x8 wreg[kWPer<64 * kTrainWaves>];
    fetch_weights<64 * kTrainWaves>(wg, wreg);
    grad += (size_t)(blockIdx.x % kGradCopies) * kParams;
    const int64_t written = (int64_t)*count_p;
    if (written <= 0) return;
    const int64_t f = first % written;
    const int64_t count = min(written - f, max_records);
    records += f * 9;
    const float inv = 1.0f / (3.0f * (float)count);
    const int l = __lane_id(), r = l % 16, g = l / 16;
    float lsum = 0.0f;
    const int64_t first_tile = (int64_t)blockIdx.x * kBlockRows;
    bf16x8 x[3];
    tile_features(records, count, first_tile, bmin, ibsize, x);
    stage_weights<64 * kTrainWaves>(wreg, w_lds, wT_lds);
    __syncthreads();
    // block-uniform loop (the weight gradient synchronises the block's waves)
    for (int64_t bbase = first_tile; bbase < count; bbase += (int64_t)gridDim.x * kBlockRows) {
        if (bbase != first_tile) tile_features(records, count, bbase, bmin, ibsize, x);
        const int64_t base = bbase + (threadIdx.x / 64) * kTileRows;
        // the weights are re-read from LDS per tile: an offset the compiler
        // cannot see through keeps it from hoisting every weight fragment of
        // both layouts into registers (~260 VGPRs, spilled)
        int zero;
        asm volatile("s_mov_b32 %0, 0" : "=s"(zero));
        const __bf16* w = w_lds + zero;
        const __bf16* wT = wT_lds + zero;
        const int64_t row = base + r;
        const bool valid = row < count;
        const float* rec = records + (valid ? row : 0) * 9;
        // forward, keeping each layer's input fragments
        floatx4 c[4];
        bf16x8 h1[2], h2[2], h3[2], h4[2];
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            c[nt] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int ks = 0; ks < 3; ks++) c[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wfrag_in(w + kL0, nt, ks), x[ks], c[nt], 0, 0, 0);
        }
        h1[0] = relu_pack(c[0], c[1]);
        h1[1] = relu_pack(c[2], c[3]);
        // one layer at a time: the scheduler would otherwise issue every
        // layer's weight reads up front and hold them in registers
        __builtin_amdgcn_sched_barrier(0);
        bf16x8* hs[4] = {h1, h2, h3, h4};
        const int offs[3] = {kL1, kL2, kL3};
#pragma unroll
        for (int L = 0; L < 3; L++) {
#pragma unroll
            for (int nt = 0; nt < 4; nt++) {
                c[nt] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int ks = 0; ks < 2; ks++)
                    c[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wfrag_hidden(w + offs[L], nt, ks), hs[L][ks], c[nt], 0, 0, 0);
            }
            hs[L + 1][0] = relu_pack(c[0], c[1]);
            hs[L + 1][1] = relu_pack(c[2], c[3]);
            __builtin_amdgcn_sched_barrier(0);
        }
        floatx4 y = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ks = 0; ks < 2; ks++) y = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wfrag_hidden(w + kL4, 0, ks), h4[ks], y, 0, 0, 0);
        // output gradient (group 0 holds outputs 0..3 of each row; RGB = 0..2)
        floatx4 dy = {0.0f, 0.0f, 0.0f, 0.0f};
        if (g == 0 && valid) {
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const float q = y[j], d = q - rec[6 + j], den = q * q + 0.01f;
                dy[j] = 2.0f * d / den * inv;
                lsum += d * d / den * inv;
            }
        }
        bf16x8 dz[2];
        dz[0] = pack2(dy, floatx4{0.0f, 0.0f, 0.0f, 0.0f});
        floatx4 dx[4];
        // output layer (16 x 64), then the hidden layers; each layer's dX is
        // formed between posting its transposed tiles and the barrier
        __builtin_amdgcn_sched_barrier(0);
        dw_post<kOut, kH, false>(dz, h4, ztA, xtA);
        back_layer<kOut>(wT + kT4, dz, dx);
        dw_issue<kOut, kH>(ztA, xtA, grad + kOff4);
        mask_pack(dx, h4, dz);
        __builtin_amdgcn_sched_barrier(0);
        dw_post<kH, kH, false>(dz, h3, ztB, xtB);
        back_layer<kH>(wT + kT3, dz, dx);
        dw_issue<kH, kH>(ztB, xtB, grad + kOff3);
        mask_pack(dx, h3, dz);
        __builtin_amdgcn_sched_barrier(0);
        dw_post<kH, kH, false>(dz, h2, ztA, xtA);
        back_layer<kH>(wT + kT2, dz, dx);
        dw_issue<kH, kH>(ztA, xtA, grad + kOff2);
        mask_pack(dx, h2, dz);
        __builtin_amdgcn_sched_barrier(0);
        dw_post<kH, kH, false>(dz, h1, ztB, xtB);
        back_layer<kH>(wT + kT1, dz, dx);
        dw_issue<kH, kH>(ztB, xtB, grad + kOff1);
        mask_pack(dx, h1, dz);
        __builtin_amdgcn_sched_barrier(0);
        dw_post<kH, kIn, true>(dz, x, ztA, xtA);
        dw_issue<kH, kIn>(ztA, xtA, grad + kOff0);
        // buffer A is written again by the next tile's output layer
        __syncthreads();
    }
    // the loss: one partial per block into slot 1 + blockIdx % (kLossSlots - 1)
    // (float atomics on one address are serialised memory-side)
    for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
    __shared__ float wsum[kTrainWaves];
    if (l == 0) wsum[threadIdx.x / 64] = lsum;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.0f;
#pragma unroll
        for (int i = 0; i < kTrainWaves; i++) t += wsum[i];
        atomic_addf(loss + 1 + blockIdx.x % (kLossSlots - 1), t);
    }
}

// ---------------------------------------------------------------- paths --
// include/vr_path_integrator_NEE.h:1-66 up to the (K+1)-th real collision,
// where the path stops and hands (x, w, W) to the cache.
template <int MAJ>
VR_DEV f3 nee_cached(Tracer<MAJ>& tr, f3 x, f3 w, int K, f3& qx, f3& qw, f3& qW, bool& emit) {
    const MediumView& M = tr.med();
    emit = false;
    float tmin, tmax;
    boundary(M, x, w, tmin, tmax);
    x = x + w * tmin;
    float d = tmax - tmin;
    f3 R = env_eval(M, w) * tr.transmittance(x, w);
    f3 W = F3(1.0f, 1.0f, 1.0f);
    const float EPS = 0.00001f;
    int bounces = 0;
    Walk<MAJ> mj;
    mj.begin(M, x, w);
    while (d > EPS) {
        float md;
        float m = fmaxf(EPS, mj.query(M, md));
        float t = tr.flight(m);
        x = x + w * fminf(t, md);
        if (t > md) { d -= md; mj.cross(M, w); if (d <= kCrossEps) break; continue; }
        if (t > d) break;
        mj.move(t);
        float s = tr.sigma(x);
        if (tr.rnd() * m >= s) { d -= t; continue; }
        if (bounces == K) {
            qx = x; qw = w; qW = W;
            emit = true;
            return R;
        }
        f3 sa = tr.albedo(x);
        R = R + (W * (F3(1.0f, 1.0f, 1.0f) - sa)) * M.emission_c;
        W = W * sa;
        f3 wnee;
        f3 env = env_sample(M, tr.rng, wnee);
        float Tn = tr.transmittance(x, wnee);
        R = R + (W * env) * (Tn * phase_eval(M, x, w, wnee));
        bounces++;
        f3 wo;
        float rho = phase_sample(M, tr.rng, x, w, wo);
        W = W * rho;
        if (W.x < EPS && W.y < EPS && W.z < EPS) return R;
        w = wo;
        boundary(M, x, w, tmin, d);
        mj.begin(M, x, w);
    }
    return R;
}

// The NEE path of a training sample, recording (x, w, R, W) at each real
// collision before its terms are added (at most maxv vertices; the medium's
// max_bounces caps the path).  Returns the vertex count; R_end is the path's
// radiance.
template <int MAJ>
VR_DEV int nee_record(Tracer<MAJ>& tr, f3 x, f3 w, int maxv, float* __restrict__ tmp, f3& R_end) {
    const MediumView& M = tr.med();
    float tmin, tmax;
    boundary(M, x, w, tmin, tmax);
    x = x + w * tmin;
    float d = tmax - tmin;
    f3 R = env_eval(M, w) * tr.transmittance(x, w);
    f3 W = F3(1.0f, 1.0f, 1.0f);
    const float EPS = 0.00001f;
    int bounces = 0, nv = 0;
    Walk<MAJ> mj;
    mj.begin(M, x, w);
    while (d > EPS) {
        float md;
        float m = fmaxf(EPS, mj.query(M, md));
        float t = tr.flight(m);
        x = x + w * fminf(t, md);
        if (t > md) { d -= md; mj.cross(M, w); if (d <= kCrossEps) break; continue; }
        if (t > d) break;
        mj.move(t);
        float s = tr.sigma(x);
        if (tr.rnd() * m >= s) { d -= t; continue; }
        if (nv < maxv) {
            float* o = tmp + nv * 12;
            o[0] = x.x; o[1] = x.y; o[2] = x.z; o[3] = w.x; o[4] = w.y; o[5] = w.z;
            o[6] = R.x; o[7] = R.y; o[8] = R.z; o[9] = W.x; o[10] = W.y; o[11] = W.z;
            nv++;
        }
        f3 sa = tr.albedo(x);
        R = R + (W * (F3(1.0f, 1.0f, 1.0f) - sa)) * M.emission_c;
        W = W * sa;
        f3 wnee;
        f3 env = env_sample(M, tr.rng, wnee);
        float Tn = tr.transmittance(x, wnee);
        R = R + (W * env) * (Tn * phase_eval(M, x, w, wnee));
        bounces++;
        if (M.max_bounces >= 0 && bounces >= M.max_bounces) break;
        f3 wo;
        float rho = phase_sample(M, tr.rng, x, w, wo);
        W = W * rho;
        if (W.x < EPS && W.y < EPS && W.z < EPS) break;
        w = wo;
        boundary(M, x, w, tmin, d);
        mj.begin(M, x, w);
    }
    R_end = R;
    return nv;
}

VR_DEV unsigned long long lanes_below() {
    const unsigned lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Resident waves per SIMD the render kernel and the training-path kernel
// (nested estimators, one thread per sample or path) are compiled for: 6 for the
// render (11.3 ms for 1024^2 x 16 spp against 12.3 at 8, 11.75 at 5, 12.7 at 4
// and 14.6 at 3), 3 for the few, long training paths (training 17.75 -> 17.3 ms).
#ifndef VR_NRC_WAVES_PER_EU
#define VR_NRC_WAVES_PER_EU 6
#endif
#ifndef VR_NRC_PATHS_WAVES_PER_EU
#define VR_NRC_PATHS_WAVES_PER_EU 3
#endif

// Dispatch seeds of the traces of one path launch (kernel-argument array).
constexpr int kMaxTraces = 16;
struct TraceSeeds {
    Seeds s[kMaxTraces];
};

// Training paths: path j of trace blockIdx.y starts at a uniformly drawn sensor
// (the trace's own seeds); each recorded vertex becomes a record (x, w,
// (R_end - R_v) / W_v) appended to the trace's records at its count.  The
// traces of a launch are independent (the paths do not read the cache), so a
// launch of T traces equals T launches of one.  Few, long paths: one wave per
// block so that every CU gets work.
template <int MAJ>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(VR_NRC_PATHS_WAVES_PER_EU))) k_nrc_paths(MediumView M, CameraView cam, TraceSeeds seeds, int64_t total, int n_paths,
                                                   int maxv, float* __restrict__ tmp, float* __restrict__ records,
                                                   int64_t max_records, uint32_t* __restrict__ count) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_paths) return;
    const int trace = blockIdx.y;
    records += (size_t)trace * (size_t)max_records * 9;
    count += trace;
    Tracer<MAJ> tr(first_kernel_arg<MediumView>());
    tr.rng = seed_map(seeds.s[trace], (uint32_t)j);
    const int64_t i = min(total - 1, (int64_t)(tr.rnd() * (float)total));
    int pose, row, col;
    sensor_index(cam, i, pose, row, col);
    if (pose < 0 || pose >= cam.num_poses) return;
    f3 x, w, R_end;
    camera_ray(cam, pose, row, col, tr.rng, x, w);
    float* t = tmp + ((size_t)trace * (size_t)n_paths + (size_t)j) * (size_t)maxv * 12;
    const int nv = nee_record(tr, x, w, maxv, t, R_end);
    const uint32_t base = atomicAdd(count, (uint32_t)nv);
    for (int v = 0; v < nv && (int64_t)base + v < max_records; v++) {
        const float* o = t + v * 12;
        float* r = records + ((size_t)base + v) * 9;
        r[0] = o[0]; r[1] = o[1]; r[2] = o[2]; r[3] = o[3]; r[4] = o[4]; r[5] = o[5];
        r[6] = o[9] > 0.0f ? (R_end.x - o[6]) / o[9] : 0.0f;
        r[7] = o[10] > 0.0f ? (R_end.y - o[7]) / o[10] : 0.0f;
        r[8] = o[11] > 0.0f ? (R_end.z - o[8]) / o[11] : 0.0f;
    }
}

// Capture with cache termination: sensor start + j % cnt, samples [c*chunk, ...)
// as in k_capture_fw; the partial radiance goes to `out` (rows zeroed first when
// chunks > 1) and each terminated sample appends a query (x, w, W/spp, pixel).
template <int MAJ>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(VR_NRC_WAVES_PER_EU))) k_nrc_render(MediumView M, CameraView cam, Seeds seeds, int spp, int chunk, int chunks,
                                                    int K, int64_t total, int64_t start, int64_t cnt,
                                                    float* __restrict__ out, float* __restrict__ queries,
                                                    int64_t max_queries, uint32_t* __restrict__ count) {
    const int64_t jj = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = jj < cnt * chunks;
    const int64_t c = live ? jj / cnt : 0;
    const int64_t i = start + (live ? jj - c * cnt : 0);
    int pose = 0, row = 0, col = 0;
    if (live) sensor_index(cam, i, pose, row, col);
    const bool ok = live && pose >= 0 && pose < cam.num_poses;
    const int s0 = (int)c * chunk, s1 = ok ? min(spp, s0 + chunk) : s0;
    const float inv = 1.0f / (float)spp;
    f3 acc = F3(0.0f, 0.0f, 0.0f);
    Tracer<MAJ> tr(first_kernel_arg<MediumView>());
    // every lane runs `chunk` iterations so that the query slots are taken by
    // wave-wide ballots at the same point of the loop
    for (int s = s0; s < s0 + chunk; s++) {
        bool emit = false;
        f3 qx, qw, qW;
        if (s < s1) {
            tr.rng = seed_capture_fw(seeds, (uint32_t)((int64_t)s * total + i));
            f3 x, w;
            camera_ray(cam, pose, row, col, tr.rng, x, w);
            acc = acc + nee_cached(tr, x, w, K, qx, qw, qW, emit);
        }
        const unsigned long long mask = __ballot(emit);
        if (mask) {
            const int leader = __ffsll((long long)mask) - 1;
            unsigned long long base = 0;
            if ((int)__lane_id() == leader) base = atomicAdd(count, (uint32_t)__popcll(mask));
            base = __shfl(base, leader);
            if (emit) {
                const uint64_t slot = base + (uint64_t)__popcll(mask & lanes_below());
                if ((int64_t)slot < max_queries) {
                    float* q = queries + slot * 10;
                    q[0] = qx.x; q[1] = qx.y; q[2] = qx.z; q[3] = qw.x; q[4] = qw.y; q[5] = qw.z;
                    q[6] = qW.x * inv; q[7] = qW.y * inv; q[8] = qW.z * inv;
                    q[9] = __uint_as_float((uint32_t)i);
                }
            }
        }
    }
    if (!ok) return;
    if (chunks == 1) {
        out[3 * i] = acc.x * inv; out[3 * i + 1] = acc.y * inv; out[3 * i + 2] = acc.z * inv;
    } else {
        atomic_addf(out + 3 * i, acc.x * inv);
        atomic_addf(out + 3 * i + 1, acc.y * inv);
        atomic_addf(out + 3 * i + 2, acc.z * inv);
    }
}

}  // namespace nrc
}  // namespace vr
