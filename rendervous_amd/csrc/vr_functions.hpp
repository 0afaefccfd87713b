// Grid tools of the inverse-rendering loop: random sensor ids, the total
// variation regulariser (forward and adjoint) and the linear grid resampler.
//
// All three are HBM-bound elementwise/stencil passes.  The stencil kernels
// launch one block per row segment of the innermost axis (256 voxels), so the
// block's row and plane come from one scalar division per block and the lanes
// of a wave read consecutive voxels; the +-y and +-z neighbours are re-read
// from L2 / the Infinity Cache, not from HBM, as consecutive blocks walk the
// volume in memory order.
#pragma once

#include "vr_device.hpp"

namespace fn {

using namespace vr;

constexpr int kRowBlock = 256;

// Up to 3 spatial axes, innermost (x) first, plus channels.
struct Shape3 {
    int n[3];      // extent of x, y, z (1 for absent axes)
    int64_t s[3];  // voxel stride of x, y, z
    int C;         // channels
    int xb;        // blocks per row = ceil(n[0] / kRowBlock)
};

struct Voxel {
    int c[3];
    int64_t index;
    bool valid;
};

VR_DEV Voxel voxel_of_block(const Shape3& S) {
    Voxel v;
    const int64_t row = blockIdx.x / S.xb;
    const int xblk = (int)(blockIdx.x - row * S.xb);
    v.c[0] = xblk * kRowBlock + threadIdx.x;
    v.c[1] = (int)(row % S.n[1]);
    v.c[2] = (int)(row / S.n[1]);
    v.valid = v.c[0] < S.n[0];
    v.index = v.c[0] + v.c[1] * S.s[1] + v.c[2] * S.s[2];
    return v;
}

// functions/tools/random_ids.comp.glsl:11-20 with the function seeding of
// functions/common_functions.h:27-49 (rdv_seeds + (0x23F1, 0x3137, 129, i + 129),
// 4 advances, then the float bits of 4 uniforms; the capture-backward init).
__global__ void k_random_ids(int64_t* __restrict__ out, int64_t n, int4 shape, int dim, Seeds sd) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Rng s = seed_capture_bw(sd, (uint32_t)i);
    const int sh[4] = {shape.x, shape.y, shape.z, shape.w};
    for (int k = 0; k < dim; k++) out[i * dim + k] = (int64_t)(int)(rng_uniform(s) * (float)sh[k]);
}

// functions/regularizers/total_variation.comp.glsl:36-74: per voxel, the sum
// over the existing +-1 neighbours of every axis (x first, + before -) of the
// channel-summed absolute differences.
__global__ void __launch_bounds__(kRowBlock) k_tv_fw(const float* __restrict__ in, Shape3 S, int dim, float* __restrict__ out) {
    const Voxel v = voxel_of_block(S);
    if (!v.valid) return;
    const float* ctr = in + v.index * S.C;
    float tv = 0.0f;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        if (d >= dim) break;
#pragma unroll
        for (int sgn = 1; sgn >= -1; sgn -= 2) {
            const int p = v.c[d] + sgn;
            if (p < 0 || p >= S.n[d]) continue;
            const float* adj = ctr + sgn * S.s[d] * S.C;
            float t = 0.0f;
            for (int c = 0; c < S.C; c++) t += fabsf(ctr[c] - adj[c]);
            tv += t;
        }
    }
    out[v.index] = tv;
}

VR_DEV float sign_of(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

// functions/regularizers/total_variation_backward.comp.glsl:36-76:
// dL/dv[i,c] = sum over neighbours j of (g[i] + g[j]) * sign(v[i,c] - v[j,c]),
// written (not accumulated) into in_grad.
__global__ void __launch_bounds__(kRowBlock) k_tv_bw(const float* __restrict__ in, const float* __restrict__ og, Shape3 S,
                                                      int dim, float* __restrict__ grad) {
    const Voxel v = voxel_of_block(S);
    if (!v.valid) return;
    const float* ctr = in + v.index * S.C;
    const float gi = og[v.index];
    constexpr int kMaxC = 4;
    float acc[kMaxC];
    for (int c0 = 0; c0 < S.C; c0 += kMaxC) {
        const int nc = min(kMaxC, S.C - c0);
#pragma unroll
        for (int c = 0; c < kMaxC; c++) acc[c] = 0.0f;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            if (d >= dim) break;
#pragma unroll
            for (int sgn = 1; sgn >= -1; sgn -= 2) {
                const int p = v.c[d] + sgn;
                if (p < 0 || p >= S.n[d]) continue;
                const int64_t j = v.index + sgn * S.s[d];
                const float g = gi + og[j];
                const float* adj = in + j * S.C;
#pragma unroll
                for (int c = 0; c < kMaxC; c++)
                    if (c < nc) acc[c] += g * sign_of(ctr[c0 + c] - adj[c0 + c]);
            }
        }
#pragma unroll
        for (int c = 0; c < kMaxC; c++)
            if (c < nc) grad[v.index * S.C + c0 + c] = acc[c];
    }
}

// Per-axis source coordinate of functions/tools/resample_grid.comp.glsl:14-27:
// coord = (p * (in - 1)) / float(out - 1), split into floor and fraction.
struct Resample {
    Shape3 out;
    int in_n[3];
    int64_t in_s[3];
    int dim;
};

// resample_grid.comp.glsl:30-86: the 2^dim linear taps (the +1 tap clamped to
// the last vertex), weights formed and accumulated in the shader's order with
// no fused multiply-add, so the result does not depend on contraction.
__global__ void __launch_bounds__(kRowBlock) k_resample(const float* __restrict__ in, Resample R, float* __restrict__ out) {
#pragma clang fp contract(off)
    const Voxel v = voxel_of_block(R.out);
    if (!v.valid) return;
    int ic[3] = {0, 0, 0};
    float a[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int i = 0; i < 3; i++) {
        if (i >= R.dim) break;
        const float coord = (float)(v.c[i] * (R.in_n[i] - 1)) / (float)(R.out.n[i] - 1);
        ic[i] = (int)floorf(coord);
        a[i] = coord - (float)ic[i];
    }
    const int C = R.out.C;
    const int taps = 1 << R.dim;
    float w[8];
    int64_t src[8];
#pragma unroll
    for (int t = 0; t < 8; t++) {
        float wt = 1.0f;
        int64_t sv = 0;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            if (i >= R.dim) break;
            const int dx = (t >> i) & 1;
            const float f = dx ? a[i] : 1.0f - a[i];
            wt = i == 0 ? f : wt * f;
            sv += (int64_t)min(ic[i] + dx, R.in_n[i] - 1) * R.in_s[i];
        }
        w[t] = wt;
        src[t] = sv * C;
    }
    constexpr int kMaxC = 4;
    for (int c0 = 0; c0 < C; c0 += kMaxC) {
        const int nc = min(kMaxC, C - c0);
        float acc[kMaxC] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int t = 0; t < 8; t++) {
            if (t >= taps) break;
#pragma unroll
            for (int c = 0; c < kMaxC; c++)
                if (c < nc) acc[c] += w[t] * in[src[t] + c0 + c];
        }
#pragma unroll
        for (int c = 0; c < kMaxC; c++)
            if (c < nc) out[v.index * C + c0 + c] = acc[c];
    }
}

}  // namespace fn
