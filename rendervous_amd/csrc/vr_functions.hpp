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

// Source coordinate of output p along an axis: resample_grid.comp.glsl:14-27
// (vertex-aligned, p (in - 1) / (out - 1)) or resample_img.comp.glsl:14-27
// (pixel-centred, (p + 1/2) / out * in - 1/2, which is -1/2 .. in - 1/2).
template <bool CELL>
VR_DEV float source_coord(int p, int n_in, int n_out) {
#pragma clang fp contract(off)
    if (CELL) return ((float)p + 0.5f) / (float)n_out * (float)n_in - 0.5f;
    return (float)(p * (n_in - 1)) / (float)(n_out - 1);
}

// resample_grid.comp.glsl / resample_img.comp.glsl:30-86: the 2^dim linear taps
// (clamped to the extent), weights formed and accumulated in the shaders' order
// with no fused multiply-add, so the result does not depend on contraction.
template <bool CELL>
__global__ void __launch_bounds__(kRowBlock) k_resample(const float* __restrict__ in, Resample R, float* __restrict__ out) {
#pragma clang fp contract(off)
    const Voxel v = voxel_of_block(R.out);
    if (!v.valid) return;
    int ic[3] = {0, 0, 0};
    float a[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int i = 0; i < 3; i++) {
        if (i >= R.dim) break;
        const float coord = source_coord<CELL>(v.c[i], R.in_n[i], R.out.n[i]);
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
            sv += (int64_t)max(0, min(ic[i] + dx, R.in_n[i] - 1)) * R.in_s[i];
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

// ---- single-channel 3-D fast paths ------------------------------------------
// One lane per 4 consecutive x voxels (float4 loads and stores), a 64 x 4 block
// over 256 x voxels of 4 rows, blockIdx.z the plane.  Neighbour indices are
// clamped and their terms masked, so every load of a lane issues before the
// first wait (7 independent loads per lane in flight instead of one 4-byte
// load): the generic kernels above are latency-bound at ~15 % of HBM.  Outputs
// are written once and streamed past the caches (non-temporal stores), which
// keep the neighbour rows.  Sums keep the shader's neighbour order, so results
// equal the generic kernels'.
struct Dense3 {
    int nx, ny, nz;
    int64_t sy, sz;
};

constexpr int kQuadX = 64, kRowsY = 4;

typedef float f4v __attribute__((ext_vector_type(4)));

VR_DEV void st4_stream(float* p, float a, float b, float c, float d) {
    f4v v = {a, b, c, d};
    __builtin_nontemporal_store(v, reinterpret_cast<f4v*>(p));
}

VR_DEV float4 ld4(const float* p) { return *reinterpret_cast<const float4*>(p); }
VR_DEV float at4(const float4& v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w)); }

struct Nbhd {
    float4 c, ym, yp, zm, zp;
    float xm, xp;
    bool hxm, hxp, hym, hyp, hzm, hzp;
};

VR_DEV bool nbhd_load(const float* __restrict__ in, const Dense3& S, Nbhd& N, int64_t& at) {
    const int x0 = (blockIdx.x * kQuadX + threadIdx.x) * 4;
    // a wave is one threadIdx.y row of the 64 x 4 block: y is wave-uniform
    const int y = blockIdx.y * kRowsY + __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int z = blockIdx.z;
    if (x0 >= S.nx || y >= S.ny) return false;
    at = x0 + y * S.sy + z * S.sz;
    const float* r = in + at;
    N.hxm = x0 > 0; N.hxp = x0 + 4 < S.nx;
    N.hym = y > 0; N.hyp = y + 1 < S.ny;
    N.hzm = z > 0; N.hzp = z + 1 < S.nz;
    N.c = ld4(r);
    N.xm = r[N.hxm ? -1 : 0];
    N.xp = r[N.hxp ? 4 : 3];
    N.yp = ld4(r + (N.hyp ? S.sy : 0));
    N.ym = ld4(r - (N.hym ? S.sy : 0));
    N.zp = ld4(r + (N.hzp ? S.sz : 0));
    N.zm = ld4(r - (N.hzm ? S.sz : 0));
    return true;
}

__global__ void __launch_bounds__(kQuadX * kRowsY) k_tv_fw_x4(const float* __restrict__ in, Dense3 S, float* __restrict__ out) {
    Nbhd N;
    int64_t at;
    if (!nbhd_load(in, S, N, at)) return;
    float r[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float c = at4(N.c, k);
        const float right = k == 3 ? N.xp : at4(N.c, k + 1);
        const float left = k == 0 ? N.xm : at4(N.c, k - 1);
        float tv = 0.0f;
        tv += (k < 3 || N.hxp) ? fabsf(c - right) : 0.0f;
        tv += (k > 0 || N.hxm) ? fabsf(c - left) : 0.0f;
        tv += N.hyp ? fabsf(c - at4(N.yp, k)) : 0.0f;
        tv += N.hym ? fabsf(c - at4(N.ym, k)) : 0.0f;
        tv += N.hzp ? fabsf(c - at4(N.zp, k)) : 0.0f;
        tv += N.hzm ? fabsf(c - at4(N.zm, k)) : 0.0f;
        r[k] = tv;
    }
    st4_stream(out + at, r[0], r[1], r[2], r[3]);
}

__global__ void __launch_bounds__(kQuadX * kRowsY) k_tv_bw_x4(const float* __restrict__ in, const float* __restrict__ og,
                                                               Dense3 S, float* __restrict__ grad) {
    Nbhd V, G;
    int64_t at;
    if (!nbhd_load(in, S, V, at)) return;
    nbhd_load(og, S, G, at);
    float r[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float c = at4(V.c, k), gi = at4(G.c, k);
        const float vr = k == 3 ? V.xp : at4(V.c, k + 1), gr = k == 3 ? G.xp : at4(G.c, k + 1);
        const float vl = k == 0 ? V.xm : at4(V.c, k - 1), gl = k == 0 ? G.xm : at4(G.c, k - 1);
        float acc = 0.0f;
        acc += (k < 3 || V.hxp) ? (gi + gr) * sign_of(c - vr) : 0.0f;
        acc += (k > 0 || V.hxm) ? (gi + gl) * sign_of(c - vl) : 0.0f;
        acc += V.hyp ? (gi + at4(G.yp, k)) * sign_of(c - at4(V.yp, k)) : 0.0f;
        acc += V.hym ? (gi + at4(G.ym, k)) * sign_of(c - at4(V.ym, k)) : 0.0f;
        acc += V.hzp ? (gi + at4(G.zp, k)) * sign_of(c - at4(V.zp, k)) : 0.0f;
        acc += V.hzm ? (gi + at4(G.zm, k)) * sign_of(c - at4(V.zm, k)) : 0.0f;
        r[k] = acc;
    }
    st4_stream(grad + at, r[0], r[1], r[2], r[3]);
}

struct Resample3 {
    Dense3 out;
    int in_n[3];
    int64_t in_sy, in_sz;
};

template <bool CELL>
VR_DEV void axis_coord(int p, int n_in, int n_out, int& i0, int& i1, float& a) {
    const float coord = source_coord<CELL>(p, n_in, n_out);
    const int f = (int)floorf(coord);
    a = coord - (float)f;
    i0 = max(0, min(f, n_in - 1));
    i1 = max(0, min(f + 1, n_in - 1));
}

// The resample_grid taps for 4 consecutive x outputs of one row: y and z taps
// shared, 32 independent loads per lane, one float4 store.
template <bool CELL>
__global__ void __launch_bounds__(kQuadX * kRowsY) k_resample_x4(const float* __restrict__ in, Resample3 R,
                                                                  float* __restrict__ out) {
#pragma clang fp contract(off)
    const int x0 = (blockIdx.x * kQuadX + threadIdx.x) * 4;
    const int y = blockIdx.y * kRowsY + __builtin_amdgcn_readfirstlane(threadIdx.y);  // wave-uniform
    const int z = blockIdx.z;
    if (x0 >= R.out.nx || y >= R.out.ny) return;
    int y0, y1, z0, z1;
    float ay, az;
    axis_coord<CELL>(y, R.in_n[1], R.out.ny, y0, y1, ay);
    axis_coord<CELL>(z, R.in_n[2], R.out.nz, z0, z1, az);
    const float* rows[4] = {in + y0 * R.in_sy + z0 * R.in_sz, in + y1 * R.in_sy + z0 * R.in_sz,
                            in + y0 * R.in_sy + z1 * R.in_sz, in + y1 * R.in_sy + z1 * R.in_sz};
    int xi[4][2];
    float ax[4];
#pragma unroll
    for (int k = 0; k < 4; k++) axis_coord<CELL>(x0 + k, R.in_n[0], R.out.nx, xi[k][0], xi[k][1], ax[k]);
    float v[4][8];
#pragma unroll
    for (int k = 0; k < 4; k++)
#pragma unroll
        for (int t = 0; t < 8; t++) v[k][t] = rows[t >> 1][xi[k][t & 1]];
    float r[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        float acc = 0.0f;
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const float fx = (t & 1) ? ax[k] : 1.0f - ax[k];
            const float fy = (t & 2) ? ay : 1.0f - ay;
            const float fz = (t & 4) ? az : 1.0f - az;
            const float w = (fx * fy) * fz;
            acc += w * v[k][t];
        }
        r[k] = acc;
    }
    st4_stream(out + x0 + y * R.out.sy + z * R.out.sz, r[0], r[1], r[2], r[3]);
}

// LDS-staged form of k_resample_x4 for a dense 3-D grid: a workgroup covers a
// band of kBandY output rows of one output plane.  It first copies the source
// rows the band reads (planes z0 and z1, rows y0(first row) .. y1(last row);
// both are monotone in the output row) into LDS with coalesced loads and tabulates
// the x taps of every output column (two indices and the fraction), then each
// wave forms rows of the band, 4 outputs per lane, from LDS with the taps and
// arithmetic order of k_resample_x4, so the results are bit-identical; each
// source value is read from memory about once per band instead of once per tap.
// Dynamic LDS: [xi0 | xi1 | ax] (nxo each, nxo a multiple of 4) then the rows.
constexpr int kBandY = 8;
constexpr int kResampleLdsBytes = 64 * 1024;  // most dynamic LDS a workgroup asks for

// One output row: 4 consecutive outputs per lane from the source rows (y0,z0),
// (y1,z0), (y0,z1), (y1,z1) and the tabulated x taps.
template <typename RowPtr>
__attribute__((always_inline)) VR_DEV void resample_row4(RowPtr r0, RowPtr r1, RowPtr r2, RowPtr r3, const int* tx0,
                                                         const int* tx1, const float* tax, int nxo, float ay, float az,
                                                         int lane, float* o) {
#pragma clang fp contract(off)
    for (int x0 = lane * 4; x0 < nxo; x0 += 256) {
        const int4 i0 = *reinterpret_cast<const int4*>(tx0 + x0), i1 = *reinterpret_cast<const int4*>(tx1 + x0);
        const float4 a4 = *reinterpret_cast<const float4*>(tax + x0);
        const int xi[4][2] = {{i0.x, i1.x}, {i0.y, i1.y}, {i0.z, i1.z}, {i0.w, i1.w}};
        const float ax[4] = {a4.x, a4.y, a4.z, a4.w};
        float r[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const float v[8] = {r0[xi[k][0]], r0[xi[k][1]], r1[xi[k][0]], r1[xi[k][1]],
                                r2[xi[k][0]], r2[xi[k][1]], r3[xi[k][0]], r3[xi[k][1]]};
            float acc = 0.0f;
#pragma unroll
            for (int t = 0; t < 8; t++) {
                const float fx = (t & 1) ? ax[k] : 1.0f - ax[k];
                const float fy = (t & 2) ? ay : 1.0f - ay;
                const float fz = (t & 4) ? az : 1.0f - az;
                const float w = (fx * fy) * fz;
                acc += w * v[t];
            }
            r[k] = acc;
        }
        st4_stream(o + x0, r[0], r[1], r[2], r[3]);
    }
}

template <bool CELL>
__global__ void __launch_bounds__(256) k_resample_lds(const float* __restrict__ in, Resample3 R, float* __restrict__ out,
                                                      int lds_floats) {
#pragma clang fp contract(off)
    extern __shared__ float lds[];
    const int nxo = R.out.nx, nxi = R.in_n[0];
    int* tx0 = reinterpret_cast<int*>(lds);
    int* tx1 = tx0 + nxo;
    float* tax = lds + 2 * nxo;
    float* tile = lds + 3 * nxo;
    const int Y0 = blockIdx.x * kBandY;
    const int Y1 = min(Y0 + kBandY, R.out.ny) - 1;
    const int z = blockIdx.y;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    int z0, z1, ya, yb, unused;
    float az, fu;
    axis_coord<CELL>(z, R.in_n[2], R.out.nz, z0, z1, az);
    axis_coord<CELL>(Y0, R.in_n[1], R.out.ny, ya, unused, fu);
    axis_coord<CELL>(Y1, R.in_n[1], R.out.ny, unused, yb, fu);
    const int nrows = yb - ya + 1;
    const bool staged = 3 * nxo + 2 * nrows * nxi <= lds_floats;  // else the taps are read from memory
    for (int x = threadIdx.x; x < nxo; x += 256) {
        int a, b;
        float f;
        axis_coord<CELL>(x, nxi, nxo, a, b, f);
        tx0[x] = a;
        tx1[x] = b;
        tax[x] = f;
    }
    if (staged) {
        // row r of plane p (0: z0, 1: z1) at tile[(p * nrows + r) * nxi]
        for (int row = wave; row < 2 * nrows; row += 4) {
            const int p = row >= nrows, r = row - p * nrows;
            const float* src = in + (ya + r) * R.in_sy + (p ? z1 : z0) * R.in_sz;
            float* dst = tile + row * nxi;
            for (int x = lane; x < nxi; x += 64) dst[x] = src[x];
        }
    }
    __syncthreads();
    for (int y = Y0 + wave; y <= Y1; y += 4) {
        int y0, y1;
        float ay;
        axis_coord<CELL>(y, R.in_n[1], R.out.ny, y0, y1, ay);
        float* o = out + y * R.out.sy + z * R.out.sz;
        if (staged)  // separate calls keep the LDS rows in the LDS address space (ds_read)
            resample_row4(tile + (y0 - ya) * nxi, tile + (y1 - ya) * nxi, tile + (nrows + y0 - ya) * nxi,
                          tile + (nrows + y1 - ya) * nxi, tx0, tx1, tax, nxo, ay, az, lane, o);
        else
            resample_row4(in + y0 * R.in_sy + z0 * R.in_sz, in + y1 * R.in_sy + z0 * R.in_sz,
                          in + y0 * R.in_sy + z1 * R.in_sz, in + y1 * R.in_sy + z1 * R.in_sz, tx0, tx1, tax, nxo, ay,
                          az, lane, o);
    }
}

// maths.h:244-251 unpart1by1: compact the even bits.
VR_DEV int unpart1by1(int n) {
    n &= 0x55555555;
    n = (n ^ (n >> 1)) & 0x33333333;
    n = (n ^ (n >> 2)) & 0x0f0f0f0f;
    n = (n ^ (n >> 4)) & 0x00ff00ff;
    n = (n ^ (n >> 8)) & 0x0000ffff;
    return n;
}

// functions/tools/copy_img_to_morton.comp.glsl (TO_MORTON) and
// copy_morton_to_img.comp.glsl: Morton index i <-> pixel (even bits, odd bits).
template <bool TO_MORTON>
__global__ void k_morton_copy(const float* __restrict__ in, int resolution, int C, float* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)resolution * resolution) return;
    const int px = unpart1by1((int)i), py = unpart1by1((int)(i >> 1));
    const int64_t pix = (int64_t)py * resolution + px;
    const int64_t src = (TO_MORTON ? pix : i) * C, dst = (TO_MORTON ? i : pix) * C;
    for (int c = 0; c < C; c++) out[dst + c] = in[src + c];
}

}  // namespace fn
