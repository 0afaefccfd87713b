// vr_device.hpp -- gfx950 device functions of the participating-media path.
//
// FP32 restatements of the reference's GLSL routines (rendervous include/*.h;
// each function cites the lines it follows).  The estimators are templated on
// the majorant kind so the constant-majorant form (the reference default,
// _objects.py:96-101) and the coarse-grid DDA form compile to separate code
// without run-time branching in the inner loops.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/vr_api.h"

#define VR_DEV __device__ __forceinline__
#define VR_HD __host__ __device__ __forceinline__

namespace vr {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kTwoPi = 6.28318530717958647692f;
constexpr float kInf = __builtin_inff();
// Remaining distance below which a majorant-cell crossing ends the tracking:
// where the last coarse cell's exit coincides with the medium's exit the
// crossing test is a rounding tie, and ending there keeps the random stream
// independent of which side of the tie the arithmetic lands on.
constexpr float kCrossEps = 1e-6f;

// ------------------------------------------------------- fast primitives --
// Hardware reciprocal, reciprocal square root, log2 and sine/cosine
// (v_rcp_f32, v_rsq_f32, v_log_f32, v_sin_f32/v_cos_f32; about 1 ulp) in place
// of the correctly rounded division/libm sequences on the tracking paths.
VR_DEV float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
VR_DEV float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
VR_DEV float rsqrt_fast(float x) { return __builtin_amdgcn_rsqf(x); }
// v_sqrt_f32 for a sine from a cosine (argument in [0, 1]; sqrtf lowers to a
// correctly rounded sequence with denormal rescaling)
VR_DEV float sqrt_fast(float x) { return __builtin_amdgcn_sqrtf(x); }
// -log(1-u) of a free-flight draw as -ln2 * log2(1-u): u in [0, 1) on a 2^-23
// lattice keeps 1-u a normal float, so v_log_f32 needs no denormal rescaling
// (__logf lowers to the correctly rounded libm sequence).
VR_DEV float neg_log1m(float u) { return -0.693147182f * __builtin_amdgcn_logf(1.0f - u); }
// sine and cosine of 2*pi*turns
VR_DEV void sincos_turns(float turns, float& s, float& c) {
    s = __builtin_amdgcn_sinf(turns);
    c = __builtin_amdgcn_cosf(turns);
}

// ---------------------------------------------------------------- vectors --
struct f3 {
    float x, y, z;
};
VR_HD f3 F3(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
VR_HD f3 operator+(f3 a, f3 b) { return F3(a.x + b.x, a.y + b.y, a.z + b.z); }
VR_HD f3 operator-(f3 a, f3 b) { return F3(a.x - b.x, a.y - b.y, a.z - b.z); }
VR_HD f3 operator*(f3 a, f3 b) { return F3(a.x * b.x, a.y * b.y, a.z * b.z); }
VR_HD f3 operator*(f3 a, float s) { return F3(a.x * s, a.y * s, a.z * s); }
VR_HD f3 operator*(float s, f3 a) { return F3(a.x * s, a.y * s, a.z * s); }
VR_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
VR_HD f3 cross(f3 a, f3 b) { return F3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
VR_DEV f3 normalize(f3 a) { return a * rsqrt_fast(dot(a, a)); }
VR_HD f3 load3(const float* p) { return F3(p[0], p[1], p[2]); }

// -------------------------------------------------------------------- RNG --
// HybridTaus: three Tausworthe steps and an LCG (include/randoms.h:6-59).
struct Rng {
    uint32_t x, y, z, w;
};

VR_DEV uint32_t taus_step(uint32_t z, int s1, int s2, int s3, uint32_t m) {
    uint32_t b = ((z << s1) ^ z) >> s2;
    return ((z & m) << s3) ^ b;
}

VR_DEV void rng_advance(Rng& s) {
    s.x = taus_step(s.x, 13, 19, 12, 4294967294u);
    s.y = taus_step(s.y, 2, 25, 4, 4294967288u);
    s.z = taus_step(s.z, 3, 11, 17, 4294967280u);
    s.w = 1664525u * s.w + 1013904223u;
}

// randoms.h:53-59: uniform in [0,1) with 23 random mantissa bits.
VR_DEV float rng_uniform(Rng& s) {
    rng_advance(s);
    uint32_t v = s.x ^ s.y ^ s.z ^ s.w;
    return __uint_as_float((v & 0x007FFFFFu) | 0x3F800000u) - 1.0f;
}

struct Seeds {
    uint32_t s[4];
};

// Capture forward per-thread init (_maps.py:549-560).
VR_DEV Rng seed_capture_fw(const Seeds& sd, uint32_t i) {
    Rng s;
    s.x = sd.s[0] + (0x23F1u + i);
    s.y = sd.s[1] + (0x3137u + (i + 14u) * i * 17923113u);
    s.z = sd.s[2] + 129u * i;
    s.w = sd.s[3] + (i + 129u);
    rng_advance(s);
    rng_advance(s);
    float a = rng_uniform(s), b = rng_uniform(s), c = rng_uniform(s), d = rng_uniform(s);
    s.x = __float_as_uint(a); s.y = __float_as_uint(b); s.z = __float_as_uint(c); s.w = __float_as_uint(d);
    rng_advance(s); rng_advance(s); rng_advance(s); rng_advance(s);
    return s;
}

// Capture backward per-thread init (_maps.py:676-685).
VR_DEV Rng seed_capture_bw(const Seeds& sd, uint32_t i) {
    Rng s;
    s.x = sd.s[0] + 0x23F1u;
    s.y = sd.s[1] + 0x3137u;
    s.z = sd.s[2] + 129u;
    s.w = sd.s[3] + (i + 129u);
    rng_advance(s); rng_advance(s); rng_advance(s); rng_advance(s);
    float a = rng_uniform(s), b = rng_uniform(s), c = rng_uniform(s), d = rng_uniform(s);
    s.x = __float_as_uint(a); s.y = __float_as_uint(b); s.z = __float_as_uint(c); s.w = __float_as_uint(d);
    return s;
}

// Map-eval per-thread init (_maps.py:390-393, :465-469).
VR_DEV Rng seed_map(const Seeds& sd, uint32_t i) {
    Rng s;
    s.x = sd.s[0] ^ (i ^ 17u);
    s.y = sd.s[1] ^ (i * 123111171u);
    s.z = sd.s[2] ^ (i + 11u);
    s.w = sd.s[3] ^ (i ^ (sd.s[0] + 13u * sd.s[1]));
    rng_uniform(s); rng_uniform(s); rng_uniform(s);
    return s;
}

// ------------------------------------------------------------------ views --
// Device view of a Grid3D parameter (include/common.h:88-94 + grid3d.h).
struct GridView {
    const float* data;
    float* grad;
    int sx, sy, sz, sc;  // element strides of W, H, D, C
    int nx, ny, nz, C;   // vertices per axis, channels
    f3 scale;            // grid coordinate = c * scale + offset, scale = (n-1)/(bmax-bmin)
    f3 offset;
    f3 dimf;             // n - 1 as floats
    f3 bmin, bmax;
};

// Image2D over 2-D coordinates (include/maps/image2d.h): pixel-centred
// bilinear lookup with border clamping of [H, W, 3].
struct ImageView {
    const float* data;
    int sy, sx, sc;  // element strides of H, W, C
    int w, h;
    float bmin_x, bmin_y, ibs_x, ibs_y;  // bmin, 1 / (bmax - bmin)
};

struct MediumView {
    GridView sigma;
    GridView albedo;
    int albedo_grid;
    f3 albedo_c;
    f3 box_min, box_max;
    f3 emission_c, env_c, sun_dir, sun_radiance;
    float sun_prob;
    int phase_kind;
    float g;
    float maj_value, maj_distance;
    const float* maj_grid;
    int mrx, mry, mrz;
    f3 maj_bmin, maj_csize, maj_icsize;
    int t_kind;
    float ds_eps;
    int max_bounces;
    int medium_filter;
    f3 gathering_c;
    float medium_majorant;
    int bw_mode;
    int env_kind;          // VR_ENV_CONST / VR_ENV_XR
    int env_sampler;       // VR_ENVS_UNIFORM / VR_ENVS_QUADTREE
    ImageView env_img;
    const float* quadtree; // XRQuadtreeRandomDirection densities
    int qt_levels;
    const float4* cells;   // optional sigma cell table (2 float4 per cell), else null
    GridView phase_g;      // HG asymmetry field, read when phase_g_grid
    int phase_g_grid;
};

struct CameraView {
    const float* poses;
    const int64_t* sensors;
    int num_poses, width, height, jittered;
    float fov, znear;
    float znear_over_h;  // znear / height
    float sz;            // znear / tan(fov / 2)
};

// ------------------------------------------------------------- ray / box --
// include/maths.h:28-38 (axis components with |w| <= 1e-6 replaced by 1e-6).
VR_DEV void ray_box_intersection(f3 x, f3 w, f3 bmin, f3 bmax, float& tmin, float& tmax) {
    float ix = rcp(fabsf(w.x) <= 0.000001f ? 0.000001f : w.x);
    float iy = rcp(fabsf(w.y) <= 0.000001f ? 0.000001f : w.y);
    float iz = rcp(fabsf(w.z) <= 0.000001f ? 0.000001f : w.z);
    float c0x = (bmin.x - x.x) * ix, c1x = (bmax.x - x.x) * ix;
    float c0y = (bmin.y - x.y) * iy, c1y = (bmax.y - x.y) * iy;
    float c0z = (bmin.z - x.z) * iz, c1z = (bmax.z - x.z) * iz;
    tmin = fmaxf(fmaxf(fminf(c0x, c1x), fminf(c0y, c1y)), fminf(c0z, c1z));
    tmax = fminf(fminf(fmaxf(c0x, c1x), fmaxf(c0y, c1y)), fmaxf(c0z, c1z));
}

// include/vr_boundary.h:1-11
VR_DEV bool boundary(const MediumView& M, f3 x, f3 w, float& tmin, float& tmax) {
    ray_box_intersection(x, w, M.box_min, M.box_max, tmin, tmax);
    tmin = fmaxf(0.0f, tmin);
    return !(tmax < tmin);
}

// ----------------------------------------------------------------- Grid3D --
// include/maps/grid3d.h:6-37 for a single channel.
VR_DEV float grid_lookup1(const GridView& g, f3 c) {
    float gx = fmaf(c.x, g.scale.x, g.offset.x);
    float gy = fmaf(c.y, g.scale.y, g.offset.y);
    float gz = fmaf(c.z, g.scale.z, g.offset.z);
    if (!(gx >= 0.0f && gy >= 0.0f && gz >= 0.0f && gx < g.dimf.x && gy < g.dimf.y && gz < g.dimf.z)) return 0.0f;
    float fx = floorf(gx), fy = floorf(gy), fz = floorf(gz);
    float ax = gx - fx, ay = gy - fy, az = gz - fz;
    int px = min((int)fx, g.nx - 2), py = min((int)fy, g.ny - 2), pz = min((int)fz, g.nz - 2);
    const float* p = g.data + (pz * g.sz + py * g.sy + px * g.sx);
    float v000 = p[0], v100 = p[g.sx], v010 = p[g.sy], v110 = p[g.sx + g.sy];
    float v001 = p[g.sz], v101 = p[g.sz + g.sx], v011 = p[g.sz + g.sy], v111 = p[g.sz + g.sy + g.sx];
    float bx = 1.0f - ax, by = 1.0f - ay, bz = 1.0f - az;
    float r = v000 * (bx * by * bz);
    r += v100 * (ax * by * bz);
    r += v010 * (bx * ay * bz);
    r += v110 * (ax * ay * bz);
    r += v001 * (bx * by * az);
    r += v101 * (ax * by * az);
    r += v011 * (bx * ay * az);
    r += v111 * (ax * ay * az);
    return r;
}

// grid_lookup1 reading the cell table: the cell's 8 corners are two float4
// (x fastest, then y, then z) at 32-byte index (pz*(ny-1) + py)*(nx-1) + px.
// Same coordinates, weights and summation order as grid_lookup1, so the same
// value bit for bit.
VR_DEV float cells_lookup1(const GridView& g, const float4* __restrict__ cells, f3 c) {
    float gx = fmaf(c.x, g.scale.x, g.offset.x);
    float gy = fmaf(c.y, g.scale.y, g.offset.y);
    float gz = fmaf(c.z, g.scale.z, g.offset.z);
    if (!(gx >= 0.0f && gy >= 0.0f && gz >= 0.0f && gx < g.dimf.x && gy < g.dimf.y && gz < g.dimf.z)) return 0.0f;
    float fx = floorf(gx), fy = floorf(gy), fz = floorf(gz);
    float ax = gx - fx, ay = gy - fy, az = gz - fz;
    int px = min((int)fx, g.nx - 2), py = min((int)fy, g.ny - 2), pz = min((int)fz, g.nz - 2);
    // cell index by 24-bit multiplies (full rate; the host admits tables whose
    // xy planes hold fewer than 2^24 cells)
    const uint32_t cell = __umul24((uint32_t)pz, (uint32_t)((g.ny - 1) * (g.nx - 1))) +
                          __umul24((uint32_t)py, (uint32_t)(g.nx - 1)) + (uint32_t)px;
    const float4* q = cells + 2 * (size_t)cell;
    const float4 lo = q[0], hi = q[1];
    float bx = 1.0f - ax, by = 1.0f - ay, bz = 1.0f - az;
    float r = lo.x * (bx * by * bz);
    r += lo.y * (ax * by * bz);
    r += lo.z * (bx * ay * bz);
    r += lo.w * (ax * ay * bz);
    r += hi.x * (bx * by * az);
    r += hi.y * (ax * by * az);
    r += hi.z * (bx * ay * az);
    r += hi.w * (ax * ay * az);
    return r;
}

// The medium's sigma at x: from the cell table when the medium has one.
VR_DEV float sigma_lookup(const MediumView& M, f3 x) {
    return M.cells ? cells_lookup1(M.sigma, M.cells, x) : grid_lookup1(M.sigma, x);
}

// Multi-channel lookup (C <= 4, out[4]), grid3d.h:6-37.  The channel loops run
// over 4 constant indices guarded by C, so a caller's out[] stays in registers
// (a loop bounded by C indexes it dynamically, which puts it in scratch).
VR_DEV void grid_lookupC(const GridView& g, f3 c, float* out) {
#pragma unroll
    for (int k = 0; k < 4; k++) out[k] = 0.0f;
    float gx = fmaf(c.x, g.scale.x, g.offset.x);
    float gy = fmaf(c.y, g.scale.y, g.offset.y);
    float gz = fmaf(c.z, g.scale.z, g.offset.z);
    if (!(gx >= 0.0f && gy >= 0.0f && gz >= 0.0f && gx < g.dimf.x && gy < g.dimf.y && gz < g.dimf.z)) return;
    float fx = floorf(gx), fy = floorf(gy), fz = floorf(gz);
    float ax = gx - fx, ay = gy - fy, az = gz - fz;
    int px = min((int)fx, g.nx - 2), py = min((int)fy, g.ny - 2), pz = min((int)fz, g.nz - 2);
    float bx = 1.0f - ax, by = 1.0f - ay, bz = 1.0f - az;
    float wgt[8] = {bx * by * bz, ax * by * bz, bx * ay * bz, ax * ay * bz,
                    bx * by * az, ax * by * az, bx * ay * az, ax * ay * az};
    const float* p = g.data + (pz * g.sz + py * g.sy + px * g.sx);
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const float* q = p + ((k & 1) ? g.sx : 0) + ((k & 2) ? g.sy : 0) + ((k & 4) ? g.sz : 0);
#pragma unroll
        for (int ch = 0; ch < 4; ch++)
            if (ch < g.C) out[ch] += q[ch * g.sc] * wgt[k];
    }
}

VR_DEV void atomic_addf(float* p, float v) {
    if (v != 0.0f) unsafeAtomicAdd(p, v);
}

// include/maps/grid3d.h:39-75 (out_grad per channel, og[C]; channel loops as in
// grid_lookupC).
VR_DEV void grid_scatterC(const GridView& g, f3 c, const float* og) {
    if (g.grad == nullptr) return;
    float gx = fmaf(c.x, g.scale.x, g.offset.x);
    float gy = fmaf(c.y, g.scale.y, g.offset.y);
    float gz = fmaf(c.z, g.scale.z, g.offset.z);
    if (!(gx >= 0.0f && gy >= 0.0f && gz >= 0.0f && gx < g.dimf.x && gy < g.dimf.y && gz < g.dimf.z)) return;
    float fx = floorf(gx), fy = floorf(gy), fz = floorf(gz);
    float ax = gx - fx, ay = gy - fy, az = gz - fz;
    int px = min((int)fx, g.nx - 2), py = min((int)fy, g.ny - 2), pz = min((int)fz, g.nz - 2);
    float bx = 1.0f - ax, by = 1.0f - ay, bz = 1.0f - az;
    float wgt[8] = {bx * by * bz, ax * by * bz, bx * ay * bz, ax * ay * bz,
                    bx * by * az, ax * by * az, bx * ay * az, ax * ay * az};
    float* p = g.grad + (pz * g.sz + py * g.sy + px * g.sx);
#pragma unroll
    for (int k = 0; k < 8; k++) {
        float* q = p + ((k & 1) ? g.sx : 0) + ((k & 2) ? g.sy : 0) + ((k & 4) ? g.sz : 0);
#pragma unroll
        for (int ch = 0; ch < 4; ch++)
            if (ch < g.C) atomic_addf(q + ch * g.sc, og[ch] * wgt[k]);
    }
}

// Single channel (the density): the same corners, weights and adds.
VR_DEV void grid_scatter1(const GridView& g, f3 c, float og) {
    if (g.grad == nullptr) return;
    float gx = fmaf(c.x, g.scale.x, g.offset.x);
    float gy = fmaf(c.y, g.scale.y, g.offset.y);
    float gz = fmaf(c.z, g.scale.z, g.offset.z);
    if (!(gx >= 0.0f && gy >= 0.0f && gz >= 0.0f && gx < g.dimf.x && gy < g.dimf.y && gz < g.dimf.z)) return;
    float fx = floorf(gx), fy = floorf(gy), fz = floorf(gz);
    float ax = gx - fx, ay = gy - fy, az = gz - fz;
    int px = min((int)fx, g.nx - 2), py = min((int)fy, g.ny - 2), pz = min((int)fz, g.nz - 2);
    float bx = 1.0f - ax, by = 1.0f - ay, bz = 1.0f - az;
    float wgt[8] = {bx * by * bz, ax * by * bz, bx * ay * bz, ax * ay * bz,
                    bx * by * az, ax * by * az, bx * ay * az, ax * ay * az};
    float* p = g.grad + (pz * g.sz + py * g.sy + px * g.sx);
#pragma unroll
    for (int k = 0; k < 8; k++)
        atomic_addf(p + ((k & 1) ? g.sx : 0) + ((k & 2) ? g.sy : 0) + ((k & 4) ? g.sz : 0), og * wgt[k]);
}

// Wave-cooperative form of grid_scatterC for a grid of C = 1 or 3 channels.
// scatter_post (called by every lane of the wave; lanes with `active` false
// contribute nothing) appends each contributing lane's cell, fractions and og
// to the wave's record buffer `rec` (8 floats per record) after its `n`
// records and returns the new count.  scatter_issue then adds the buffered
// records' corners with the C x 8 adds of one record on consecutive lanes
// (corner major, channel minor), so that the x-adjacent corners of a
// channel-last grid (2C consecutive floats) come from neighbouring lanes of one
// instruction and travel as one memory request instead of one per lane and
// float, and a wave issues its atomics in bursts.  The adds and their values
// are those of grid_scatterC.
template <int C>
VR_DEV int scatter_post(const GridView& g, float* __restrict__ rec, int n, bool active, f3 c, f3 og) {
    static_assert(C == 1 || C == 3, "one or three channels");
    const float gx = fmaf(c.x, g.scale.x, g.offset.x);
    const float gy = fmaf(c.y, g.scale.y, g.offset.y);
    const float gz = fmaf(c.z, g.scale.z, g.offset.z);
    const bool ok = active && gx >= 0.0f && gy >= 0.0f && gz >= 0.0f && gx < g.dimf.x && gy < g.dimf.y && gz < g.dimf.z;
    const unsigned long long mask = __ballot(ok);
    if (!mask) return n;
    if (ok) {
        const unsigned lane = __lane_id();
        const unsigned rank = (unsigned)__popcll(mask & (lane == 0 ? 0ull : (~0ull >> (64 - lane))));
        const float fx = floorf(gx), fy = floorf(gy), fz = floorf(gz);
        const int px = min((int)fx, g.nx - 2), py = min((int)fy, g.ny - 2), pz = min((int)fz, g.nz - 2);
        float4* r = reinterpret_cast<float4*>(rec + (n + (int)rank) * 8);
        r[0] = make_float4(__int_as_float(pz * g.sz + py * g.sy + px * g.sx), gx - fx, gy - fy, gz - fz);
        r[1] = make_float4(og.x, og.y, og.z, 0.0f);
    }
    __builtin_amdgcn_wave_barrier();
    return n + __popcll(mask);
}

template <int C>
VR_DEV void scatter_issue(const GridView& g, const float* __restrict__ rec, int n) {
    constexpr int kPer = 8 * C;        // lanes per record
    constexpr int kGroup = 64 / kPer;  // records per instruction
    const int lane = (int)__lane_id();
    const int j = lane / kPer, m = lane - j * kPer;
    const int k = m / C, ch = m - k * C;
    const int coff = ((k & 1) ? g.sx : 0) + ((k & 2) ? g.sy : 0) + ((k & 4) ? g.sz : 0) + ch * g.sc;
    for (int q0 = 0; q0 < n; q0 += kGroup) {
        const int q = q0 + j;
        if (j < kGroup && q < n) {
            const float4 r0 = reinterpret_cast<const float4*>(rec + q * 8)[0];
            const float v = rec[q * 8 + 4 + ch];
            const float wx = (k & 1) ? r0.y : 1.0f - r0.y;
            const float wy = (k & 2) ? r0.z : 1.0f - r0.z;
            const float wz = (k & 4) ? r0.w : 1.0f - r0.w;
            atomic_addf(g.grad + (__float_as_int(r0.x) + coff), v * (wx * wy * wz));
        }
    }
    __builtin_amdgcn_wave_barrier();
}

// ------------------------------------------------------------- directions --
// include/randoms.h:144-151
VR_DEV void ortho_basis(f3 n, f3& t, f3& b) {
    float sign = n.z >= 0.0f ? 1.0f : -1.0f;
    float a = -rcp(sign + n.z);  // |sign + n.z| >= 1
    float bb = n.x * n.y * a;
    t = F3(1.0f + sign * n.x * n.x * a, sign * bb, -sign * n.x);
    b = F3(bb, sign + n.y * n.y * a, -n.y);
}

// include/randoms.h:153-165 randomDirection(D)
VR_DEV f3 random_direction(Rng& rng, f3 D) {
    float r1 = rng_uniform(rng);
    float r2 = rng_uniform(rng) * 2.0f - 1.0f;
    float s = sqrt_fast(fmaxf(0.0f, 1.0f - r2 * r2));
    float sn, cs;
    sincos_turns(r1, sn, cs);
    f3 t0, t1;
    ortho_basis(D, t0, t1);
    return t0 * (cs * s) + t1 * (sn * s) + D * r2;
}

// include/scattering_AD.h:4-19
VR_DEV f3 hg_sample(Rng& rng, f3 win, float g) {
    float phi_turns = rng_uniform(rng);  // phi = 2 pi u
    float xi = rng_uniform(rng);
    float cos_t;
    if (fabsf(g) < 0.001f) cos_t = 2.0f * xi - 1.0f;
    else {
        float g2 = g * g;
        float t = fdiv(1.0f - g2, 1.0f - g + 2.0f * g * xi);
        cos_t = fdiv(0.5f, g) * (1.0f + g2 - t * t);
    }
    float sin_t = sqrt_fast(fmaxf(0.0f, 1.0f - cos_t * cos_t));
    float sp, cp;
    sincos_turns(phi_turns, sp, cp);
    f3 t0, t1;
    ortho_basis(win, t0, t1);
    return t0 * (sin_t * sp) + t1 * (sin_t * cp) + win * cos_t;
}

// include/scattering_AD.h:39-47
VR_DEV float hg_eval(float cos_t, float g) {
    if (fabsf(g) < 0.001f) return 0.25f / kPi;
    float g2 = g * g;
    float den = 1.0f + g2 - 2.0f * g * cos_t;
    float r = rsqrt_fast(den);
    return (0.25f / kPi) * (1.0f - g2) * (r * r * r);
}

// The HG asymmetry at a scattering position x: the constant g, or the phase_g
// Grid3D read at x (HGPhase and VHGPhaseSampler evaluate their phase_g map at
// the first three inputs, _maps.py:4449,4473).
VR_DEV float phase_g_at(const MediumView& M, f3 x) {
    return M.phase_g_grid ? grid_lookup1(M.phase_g, x) : M.g;
}

// HGPhase (_maps.py:4446-4454) / single HG / isotropic, for asymmetry g.
VR_DEV float phase_eval_g(const MediumView& M, float g, f3 w, f3 wo) {
    float ct = dot(w, wo);
    if (M.phase_kind == VR_PHASE_HG_MIX) {
        float a = hg_eval(ct, g), b = hg_eval(ct, g / 5.0f);
        return a * (1.0f - 0.2f) + b * 0.2f;
    }
    if (M.phase_kind == VR_PHASE_HG) return hg_eval(ct, g);
    return 0.25f / kPi;
}

// VHGPhaseSampler (_maps.py:4463-4481) / HGPhaseSampler (_scenes.py:1984-2004) / uniform.
VR_DEV float phase_sample_g(const MediumView& M, Rng& rng, float g, f3 w, f3& wo) {
    if (M.phase_kind == VR_PHASE_HG_MIX) {
        if (rng_uniform(rng) < 0.8f) g /= 5.0f;
        wo = hg_sample(rng, w, g);
        return 1.0f;
    }
    if (M.phase_kind == VR_PHASE_HG) { wo = hg_sample(rng, w, g); return 1.0f; }
    wo = random_direction(rng, F3(0.0f, 1.0f, 0.0f));
    return 1.0f;
}

// The same at scattering position x.
VR_DEV float phase_eval(const MediumView& M, f3 x, f3 w, f3 wo) { return phase_eval_g(M, phase_g_at(M, x), w, wo); }
VR_DEV float phase_sample(const MediumView& M, Rng& rng, f3 x, f3 w, f3& wo) {
    return phase_sample_g(M, rng, phase_g_at(M, x), w, wo);
}

// include/maths.h:172-178 dir2xr: (atan2(x, z) / pi, -2 asin(y) / pi) in [-1, 1]^2.
VR_DEV void dir2xr(f3 w, float& u, float& v) {
    if (w.x == 0.0f && w.z == 0.0f) w.x += 0.0000001f;
    u = atan2f(w.x, w.z) * (1.0f / kPi);
    v = -2.0f * asinf(fminf(fmaxf(w.y, -1.0f), 1.0f)) * (1.0f / kPi);
}

// include/maps/image2d.h:1-25 FORWARD at c (3 channels).
VR_DEV f3 image_lookup(const ImageView& I, float cx, float cy) {
    const float nx = fminf(fmaxf((cx - I.bmin_x) * I.ibs_x, 0.0f), 1.0f);
    const float ny = fminf(fmaxf((cy - I.bmin_y) * I.ibs_y, 0.0f), 1.0f);
    const float gx = nx * (float)I.w - 0.5f, gy = ny * (float)I.h - 0.5f;
    const float fx = floorf(gx), fy = floorf(gy);
    const int px = (int)fx, py = (int)fy;
    const float ax = gx - fx, ay = gy - fy;
    const int x0 = min(max(px, 0), I.w - 1), x1 = min(max(px + 1, 0), I.w - 1);
    const int y0 = min(max(py, 0), I.h - 1), y1 = min(max(py + 1, 0), I.h - 1);
    const float* p00 = I.data + y0 * I.sy + x0 * I.sx;
    const float* p01 = I.data + y0 * I.sy + x1 * I.sx;
    const float* p10 = I.data + y1 * I.sy + x0 * I.sx;
    const float* p11 = I.data + y1 * I.sy + x1 * I.sx;
    float r[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float a = p00[c * I.sc] * (1.0f - ax) + p01[c * I.sc] * ax;
        const float b = p10[c * I.sc] * (1.0f - ax) + p11[c * I.sc] * ax;
        r[c] = a * (1.0f - ay) + b * ay;
    }
    return F3(r[0], r[1], r[2]);
}

// environment(x, w): a constant, or Image2D(...).after(xr_projection)
// (include/maps/environment_xr.h:3-9, _objects.py:164-173).
VR_DEV f3 env_eval(const MediumView& M, f3 w) {
    if (M.env_kind == VR_ENV_XR) {
        float u, v;
        dir2xr(w, u, v);
        return image_lookup(M.env_img, u, v);
    }
    return M.env_c;
}

// XRQuadtreeRandomDirection (_maps.py:2494-2549): descend the quadtree of
// absolute pixel probabilities (Morton order per level, children of node n at
// 4n .. 4n+3 below it) with one uniform, then a direction uniform in the
// pixel's solid angle (randomDirection(a0, a1, b0, b1), randoms.h:186-199).
// Returns 1 / pdf = pixel solid angle / probability.
VR_DEV float quadtree_direction(const MediumView& M, Rng& rng, f3& wo) {
    float sel = rng_uniform(rng);
    int node = 0;
    float x0 = 0.0f, y0 = 0.0f, x1 = 1.0f, y1 = 1.0f, prob = 1.0f;
    for (int i = 0; i < M.qt_levels; i++) {
        const float* d = M.quadtree + node * 4;
        const float d0 = d[0], d1 = d[1], d2 = d[2], d3 = d[3];
        int child = 3;
        prob = d3;
        if (sel < d0) { child = 0; prob = d0; }
        else {
            sel -= d0;
            if (sel < d1) { child = 1; prob = d1; }
            else {
                sel -= d1;
                if (sel < d2) { child = 2; prob = d2; }
                else sel -= d2;
            }
        }
        const float xm = (x1 + x0) * 0.5f, ym = (y1 + y0) * 0.5f;
        if ((child & 1) == 0) x1 = xm; else x0 = xm;
        if ((child >> 1) == 0) y1 = ym; else y0 = ym;
        node = node * 4 + 1 + child;
    }
    // angles in turns: phi = (x * 2 - 1) pi -> x - 1/2, theta = y pi -> y / 2
    const float cb0 = __builtin_amdgcn_cosf(y0 * 0.5f), cb1 = __builtin_amdgcn_cosf(y1 * 0.5f);
    const float area = 2.0f * kPi * (cb0 - cb1) / (float)(1 << M.qt_levels);
    const float weight = area / fmaxf(0.000000001f, prob);
    const float a0 = x0 - 0.5f, a1 = x1 - 0.5f;
    const float r1 = rng_uniform(rng) * (a1 - a0) + a0;
    const float r2 = cb0 - rng_uniform(rng) * (cb0 - cb1);
    const float s = sqrt_fast(fmaxf(0.0f, 1.0f - r2 * r2));
    float sn, cs;
    sincos_turns(r1, sn, cs);
    wo = F3(sn * s, r2, cs * s);
    return weight;
}

// The environment part of the sampler: _functionsampler(UniformDirectionSampler
// or XRQuadtreeRandomDirection, environment) (_maps.py:2552-2585): L(wo) / pdf.
VR_DEV f3 env_sample_base(const MediumView& M, Rng& rng, f3& wo) {
    if (M.env_sampler == VR_ENVS_QUADTREE) {
        const float wgt = quadtree_direction(M, rng, wo);
        return env_eval(M, wo) * wgt;
    }
    wo = random_direction(rng, F3(0.0f, 1.0f, 0.0f));
    return env_eval(M, wo) * (4.0f * kPi);
}

// The constant-environment sampler alone (uniform directions, optional sun):
// what env_sample does when env_kind == VR_ENV_CONST and the sampler is uniform.
VR_DEV f3 env_sample_const(const MediumView& M, Rng& rng, f3& wo) {
    if (M.sun_prob > 0.0f) {
        if (rng_uniform(rng) < M.sun_prob) {
            wo = M.sun_dir;
            return M.sun_radiance * rcp(M.sun_prob);
        }
        wo = random_direction(rng, F3(0.0f, 1.0f, 0.0f));
        return M.env_c * (4.0f * kPi * rcp(1.0f - M.sun_prob));
    }
    wo = random_direction(rng, F3(0.0f, 1.0f, 0.0f));
    return M.env_c * (4.0f * kPi);
}

// ... mixed with a directional light with MixtureSampler semantics (_maps.py:1541-1571).
VR_DEV f3 env_sample(const MediumView& M, Rng& rng, f3& wo) {
    if (M.sun_prob > 0.0f) {
        if (rng_uniform(rng) < M.sun_prob) {
            wo = M.sun_dir;
            return M.sun_radiance * rcp(M.sun_prob);
        }
        return env_sample_base(M, rng, wo) * rcp(1.0f - M.sun_prob);
    }
    return env_sample_base(M, rng, wo);
}

// ---------------------------------------------------------------- majorant --
// include/vr_majorant.h:1-7 contract (value, distance).
template <int MAJ>
struct Walk;

template <>
struct Walk<VR_MAJ_CONST> {
    static constexpr bool grid = false;
    VR_DEV void begin(const MediumView&, f3, f3) {}
    VR_DEV float query(const MediumView& M, float& md) const { md = M.maj_distance; return M.maj_value; }
    VR_DEV void move(float) {}
    VR_DEV void cross(const MediumView&, f3) {}
};

// The coarse majorant grid staged in LDS: an internal majorant kind for
// kernels that copy a grid of at most kMajLdsCells cells into maj_lds at their
// start (the walk then reads LDS instead of global memory).
constexpr int kMajGridLds = 2;
constexpr int kMajLdsCells = 4096;
__shared__ float maj_lds[kMajLdsCells];

// Coarse max-grid DDA: value = bound of the current coarse cell, distance = to
// that cell's exit along the ray.  Per-lane state is the packed cell index, the
// cell's bound (read when the walk enters the cell, so a query is register
// arithmetic only) and the ray parameters of the next x/y/z faces and of the
// tracked point; the per-axis increments are recomputed from the ray direction
// at each crossing.  LDS: the grid is read from maj_lds.
template <bool LDS>
struct WalkGrid {
    static constexpr bool grid = true;
    int cell;                // cx | cy << 10 | cz << 20; -1 once the ray has left the grid
    float tnx, tny, tnz, t;  // ray parameter of the next x/y/z face and of the tracked point
    float mval;              // bound of the current cell

    static VR_DEV int axis_cell(float xa, float bmin, float ics, int res) {
        int c = (int)floorf((xa - bmin) * ics);
        return max(0, min(res - 1, c));
    }
    static VR_DEV float axis_face(float xa, float wa, float bmin, float cs, int c) {
        // one reciprocal for both signs (a divergent branch each would issue both)
        const float face = bmin + (float)(wa > 0.0f ? c + 1 : c) * cs - xa;
        return (wa > 0.0f || wa < 0.0f) ? face * rcp(wa) : kInf;
    }
    // bound of the packed cell c (>= 0); the grid's axes hold at most 1024
    // cells (the packing), so the index takes 24-bit multiplies
    static VR_DEV float fetch(const MediumView& M, int c) {
        const uint32_t cx = c & 1023, cy = (c >> 10) & 1023, cz = (uint32_t)c >> 20;
        const uint32_t i = __umul24(__umul24(cz, (uint32_t)M.mry) + cy, (uint32_t)M.mrx) + cx;
        return LDS ? maj_lds[i] : M.maj_grid[i];
    }
    VR_DEV void begin(const MediumView& M, f3 x, f3 w) {
        int cx = axis_cell(x.x, M.maj_bmin.x, M.maj_icsize.x, M.mrx);
        int cy = axis_cell(x.y, M.maj_bmin.y, M.maj_icsize.y, M.mry);
        int cz = axis_cell(x.z, M.maj_bmin.z, M.maj_icsize.z, M.mrz);
        tnx = axis_face(x.x, w.x, M.maj_bmin.x, M.maj_csize.x, cx);
        tny = axis_face(x.y, w.y, M.maj_bmin.y, M.maj_csize.y, cy);
        tnz = axis_face(x.z, w.z, M.maj_bmin.z, M.maj_csize.z, cz);
        t = 0.0f;
        cell = cx | (cy << 10) | (cz << 20);
        mval = fetch(M, cell);
    }
    VR_DEV float query(const MediumView&, float& md) const {
        if (cell < 0) { md = kInf; return 0.0f; }
        md = fmaxf(0.0f, fminf(tnx, fminf(tny, tnz)) - t);
        return mval;
    }
    VR_DEV void move(float dt) { t += dt; }
    VR_DEV void cross(const MediumView& M, f3 w) {
        if (cell < 0) return;
        int cx = cell & 1023, cy = (cell >> 10) & 1023, cz = cell >> 20;
        if (tnx <= tny && tnx <= tnz) {
            t = tnx;
            cx += w.x > 0.0f ? 1 : -1;
            tnx += M.maj_csize.x * rcp(fabsf(w.x));
            if (cx < 0 || cx >= M.mrx) { cell = -1; return; }
        } else if (tny <= tnz) {
            t = tny;
            cy += w.y > 0.0f ? 1 : -1;
            tny += M.maj_csize.y * rcp(fabsf(w.y));
            if (cy < 0 || cy >= M.mry) { cell = -1; return; }
        } else {
            t = tnz;
            cz += w.z > 0.0f ? 1 : -1;
            tnz += M.maj_csize.z * rcp(fabsf(w.z));
            if (cz < 0 || cz >= M.mrz) { cell = -1; return; }
        }
        cell = cx | (cy << 10) | (cz << 20);
        mval = fetch(M, cell);
    }
};

template <>
struct Walk<VR_MAJ_GRID> : WalkGrid<false> {};

template <>
struct Walk<kMajGridLds> : WalkGrid<true> {};

// Copies the medium's majorant grid into maj_lds (block-wide; the caller has
// checked that it fits).
VR_DEV void stage_majorant(const MediumView& M) {
    const int n = M.mrx * M.mry * M.mrz;
    for (int i = threadIdx.x; i < n; i += blockDim.x) maj_lds[i] = M.maj_grid[i];
    __syncthreads();
}

// ------------------------------------------------- kernel-argument access --
// Read-only data in the constant address space: loads through such a pointer
// are scalar loads when the address is wave-uniform.
template <typename T>
using ConstPtr = const __attribute__((address_space(4))) T*;

// The running kernel's first parameter, read in place in the kernel-argument
// segment (explicit arguments start at its offset 0).  Valid only in a kernel
// whose first parameter has type T.
template <typename T>
VR_DEV ConstPtr<T> first_kernel_arg() {
    return (ConstPtr<T>)__builtin_amdgcn_kernarg_segment_ptr();
}

// p, passed through an empty asm statement: the compiler can no longer tell
// the result from another address, so the loads made through it stay after
// this point.  Kernels whose loops read many uniform fields (medium, camera,
// seeds) call it at the loop head; hoisted out of the loop instead, every field
// stays live in scalar registers for the whole kernel, more than the 102 a wave
// has, and the excess spills to vector-register lanes (a v_readlane on the
// vector pipe per reload).  Re-reading costs scalar loads that the scalar cache
// serves.
template <typename T>
VR_DEV ConstPtr<T> reread(ConstPtr<T> p) {
    asm volatile("" : "+s"(p));
    return p;
}

// -------------------------------------------------------------- estimators --
template <int MAJ>
struct Tracer {
    ConstPtr<MediumView> Mp;  // the medium, in the constant address space (scalar loads)
    Rng rng;
    int lookups;   // trilinear sigma lookups
    int scatters;  // trilinear sigma gradient scatters
    // the caller draws nothing from rng after the estimator returns (each
    // sample or ray seeds its own stream): a backward may then skip primal
    // estimates whose only use is a gradient that is not requested
    bool stream_ends;

    VR_DEV explicit Tracer(ConstPtr<MediumView> m) : Mp(m), lookups(0), scatters(0), stream_ends(false) {}

    VR_DEV const MediumView& med() const { return *(const MediumView*)Mp; }
    // Loads of the medium's fields after this call are not hoisted above it
    // (reread): a loop that calls it at its head keeps them out of the scalar
    // registers live across the loop.
    VR_DEV void reread() { Mp = vr::reread(Mp); }

    VR_DEV float rnd() { return rng_uniform(rng); }
    VR_DEV float flight(float maj) { return neg_log1m(rnd()) * rcp(maj); }

    VR_DEV float sigma(f3 x) { lookups++; return sigma_lookup(med(), x); }
    VR_DEV void sigma_bw(f3 x, float g) {
        if (g != 0.0f && med().sigma.grad != nullptr) scatters++;
        grid_scatter1(med().sigma, x, g);
    }

    VR_DEV f3 albedo(f3 x) {
        if (!med().albedo_grid) return med().albedo_c;
        float v[4];
        grid_lookupC(med().albedo, x, v);
        if (med().albedo.C == 1) return F3(v[0], v[0], v[0]);
        return F3(v[0], v[1], v[2]);
    }
    VR_DEV bool albedo_has_grad() const { return med().albedo_grid && med().albedo.grad != nullptr; }
    VR_DEV void albedo_bw(f3 x, f3 g) {
        if (!med().albedo_grid || med().albedo.grad == nullptr) return;
        if (g.x == 0.0f && g.y == 0.0f && g.z == 0.0f) return;
        float og[3];
        if (med().albedo.C == 1) og[0] = g.x + g.y + g.z;
        else { og[0] = g.x; og[1] = g.y; og[2] = g.z; }
        grid_scatterC(med().albedo, x, og);
    }

    // include/vr_transmittance_DT.h:1-28
    VR_DEV float t_dt(f3 x, f3 w) {
        float tmin, tmax;
        if (!boundary(med(), x, w, tmin, tmax)) return 1.0f;
        x = x + w * tmin;
        float d = tmax - tmin;
        Walk<MAJ> mj;
        mj.begin(med(), x, w);
        for (;;) {
            float md;
            float maj = fmaxf(0.0000001f, mj.query(med(), md));
            float dt = flight(maj);
            float t = fminf(md, dt);
            x = x + w * t;
            d -= t;
            if (dt > md) { mj.cross(med(), w); if (d <= kCrossEps) return 1.0f; continue; }
            mj.move(t);
            if (d <= 0.0000001f) return 1.0f;
            if (rnd() * maj < sigma(x)) return 0.0f;
        }
    }

    // include/vr_transmittance_RT.h:1-49
    VR_DEV float t_rt(f3 x, f3 w) {
        float T = 1.0f, tmin, tmax;
        if (!boundary(med(), x, w, tmin, tmax)) return T;
        x = x + w * tmin;
        float d = tmax - tmin;
        Walk<MAJ> mj;
        mj.begin(med(), x, w);
        for (;;) {
            float md;
            float maj = fmaxf(0.0000001f, mj.query(med(), md));
            float s = flight(maj);
            if (Walk<MAJ>::grid && md < d && s > md) { x = x + w * md; d -= md; mj.cross(med(), w); if (d <= kCrossEps) return T; continue; }
            float dt = fminf(s, fminf(d, md));
            x = x + w * dt;
            d -= dt;
            mj.move(dt);
            if (d <= 0.0000001f) return T;
            T *= (1.0f - sigma(x) * rcp(maj));
            if (T < 0.01f) {
                if (rnd() < 1.0f - T) return 0.0f;
                T = 1.0f;
            }
        }
    }

    // ---- DDA (include/vr_transmittance_DDA.h), traversal in FP64 -------------
    struct Dda {
        double x[3], d;
        int cell[3], side[3], dim[3];
        double ainc[3], alpha0[3], vn[3], vm[3];
    };

    VR_DEV bool dda_begin(f3 xf, f3 wf, Dda& s) {
        float tminf, tmaxf;
        if (!boundary(med(), xf, wf, tminf, tmaxf)) return false;
        // intersect_ray_box (maths.h:11-26) with the grid box, in double
        double x[3] = {xf.x, xf.y, xf.z}, w[3] = {wf.x, wf.y, wf.z};
        double bmin[3] = {med().sigma.bmin.x, med().sigma.bmin.y, med().sigma.bmin.z};
        double bmax[3] = {med().sigma.bmax.x, med().sigma.bmax.y, med().sigma.bmax.z};
        double gmin = -1e300, gmax = 1e300;
        for (int a = 0; a < 3; a++) {
            double wa = fabs(w[a]) <= 0.000001 ? 0.000001 : w[a];
            double c0 = (bmin[a] - x[a]) / wa, c1 = (bmax[a] - x[a]) / wa;
            gmin = fmax(gmin, fmin(c0, c1));
            gmax = fmin(gmax, fmax(c0, c1));
        }
        gmin = fmax(0.0, gmin);
        if (gmax <= gmin || gmax <= 0.0) return false;
        // the boundary map's interval, recomputed in double (maths.h:28-38)
        double bmn[3] = {med().box_min.x, med().box_min.y, med().box_min.z}, bmx[3] = {med().box_max.x, med().box_max.y, med().box_max.z};
        double tmin = -1e300, tmax = 1e300;
        for (int a = 0; a < 3; a++) {
            double wa = fabs(w[a]) <= 0.000001 ? 0.000001 : w[a];
            double c0 = (bmn[a] - x[a]) / wa, c1 = (bmx[a] - x[a]) / wa;
            tmin = fmax(tmin, fmin(c0, c1));
            tmax = fmin(tmax, fmax(c0, c1));
        }
        tmin = fmax(0.0, tmin);
        if (tmax < tmin) return false;
        tmin = fmax(tmin, gmin);
        tmax = fmin(tmax, gmax);
        if (tmax <= tmin) return false;
        s.d = tmax - tmin;
        s.dim[0] = med().sigma.nx - 1; s.dim[1] = med().sigma.ny - 1; s.dim[2] = med().sigma.nz - 1;
        for (int a = 0; a < 3; a++) {
            double xa = x[a] + w[a] * tmin;
            s.x[a] = xa;
            double bs = bmax[a] - bmin[a];
            double cs = bs / s.dim[a];
            int cl = (int)((xa - bmin[a]) * s.dim[a] / bs);
            cl = max(0, min(s.dim[a] - 1, cl));
            s.cell[a] = cl;
            double aw = fmax(0.000001, fabs(w[a]));
            s.ainc[a] = cs / aw;
            s.side[a] = w[a] > 0 ? 1 : (w[a] < 0 ? -1 : 0);
            double corner = (cl + s.side[a] * 0.5 + 0.5) * cs + bmin[a];
            s.alpha0[a] = fabs(corner - xa) / aw;
            s.vn[a] = (xa - bmin[a]) * s.dim[a] / bs;
            s.vm[a] = w[a] * s.dim[a] / bs;
        }
        return true;
    }

    VR_DEV float dda_sigma_at(int x, int y, int z) {
        const GridView& g = med().sigma;
        x = max(0, min(g.nx - 1, x)); y = max(0, min(g.ny - 1, y)); z = max(0, min(g.nz - 1, z));
        return g.data[z * g.sz + y * g.sy + x * g.sx];
    }

    static VR_DEV double lerpd(double a, double b, double t) { return a * (1 - t) + b * t; }

    static VR_DEV void dda_step(double* alpha, int* cell, const Dda& s) {
        int ix = alpha[0] <= alpha[1] && alpha[0] <= alpha[2];
        int iy = alpha[0] > alpha[1] && alpha[1] <= alpha[2];
        int iz = alpha[0] > alpha[2] && alpha[1] > alpha[2];
        alpha[0] += ix * s.ainc[0]; alpha[1] += iy * s.ainc[1]; alpha[2] += iz * s.ainc[2];
        cell[0] += ix * s.side[0]; cell[1] += iy * s.side[1]; cell[2] += iz * s.side[2];
    }

    VR_DEV double dda_tau(const Dda& s) {
        double alpha[3] = {s.alpha0[0], s.alpha0[1], s.alpha0[2]};
        int cell[3] = {s.cell[0], s.cell[1], s.cell[2]};
        double cur = 0.0, tau = 0.0;
        while (cur < s.d - 0.0001) {
            double nxt = fmin(s.d, fmin(alpha[0], fmin(alpha[1], alpha[2])));
            double v0 = dda_sigma_at(cell[0], cell[1], cell[2]), v1 = dda_sigma_at(cell[0] + 1, cell[1], cell[2]);
            double v2 = dda_sigma_at(cell[0], cell[1] + 1, cell[2]), v3 = dda_sigma_at(cell[0] + 1, cell[1] + 1, cell[2]);
            double v4 = dda_sigma_at(cell[0], cell[1], cell[2] + 1), v5 = dda_sigma_at(cell[0] + 1, cell[1], cell[2] + 1);
            double v6 = dda_sigma_at(cell[0], cell[1] + 1, cell[2] + 1), v7 = dda_sigma_at(cell[0] + 1, cell[1] + 1, cell[2] + 1);
            lookups++;
            double tm = lerpd(cur, nxt, 0.5);
            double ax = s.vm[0] * tm + s.vn[0] - cell[0];
            double ay = s.vm[1] * tm + s.vn[1] - cell[1];
            double az = s.vm[2] * tm + s.vn[2] - cell[2];
            double v = lerpd(lerpd(lerpd(v0, v1, ax), lerpd(v2, v3, ax), ay), lerpd(lerpd(v4, v5, ax), lerpd(v6, v7, ax), ay), az);
            tau += v * (nxt - cur);
            dda_step(alpha, cell, s);
            cur = nxt;
        }
        return tau;
    }

    VR_DEV float t_dda(f3 x, f3 w) {
        Dda s;
        if (!dda_begin(x, w, s)) return 1.0f;
        return (float)exp(-dda_tau(s));
    }

    // include/vr_transmittance_GDT.h:1-84 / _GRT.h:1-83 over the DDA's cells:
    // the cell's minorant (min of its 8 vertices) in closed form, the residual
    // tracked with majorant max - min (GDT ends at a collision, GRT multiplies
    // by 1 - Pc).  Cell walk in FP64 as the DDA's.
    VR_DEV float t_grid_tracking(f3 x, f3 w, bool ratio) {
        Dda s;
        if (!dda_begin(x, w, s)) return 1.0f;
        double alpha[3] = {s.alpha0[0], s.alpha0[1], s.alpha0[2]};
        int cell[3] = {s.cell[0], s.cell[1], s.cell[2]};
        double cur = 0.0;
        float T = 1.0f;
        while (cur < s.d - 0.0001) {
            const double nxt = fmin(s.d, fmin(alpha[0], fmin(alpha[1], alpha[2])));
            float v[8];
            float mx = -kInf, mn = kInf;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                v[k] = dda_sigma_at(cell[0] + (k & 1), cell[1] + ((k >> 1) & 1), cell[2] + ((k >> 2) & 1));
                mx = fmaxf(mx, v[k]);
                mn = fminf(mn, v[k]);
            }
            T *= expf(-mn * (float)(nxt - cur));
            const float r = fmaxf(0.000001f, mx - mn);
            if (!(r < kInf)) return 0.0f;  // an infinite density in the cell absorbs (and ends the walk)
            double ct = cur;
            for (;;) {
                const float dt = flight(r);
                if (ct + dt > nxt) break;
                ct += dt;
                const float ax = (float)(s.vm[0] * ct + s.vn[0] - cell[0]);
                const float ay = (float)(s.vm[1] * ct + s.vn[1] - cell[1]);
                const float az = (float)(s.vm[2] * ct + s.vn[2] - cell[2]);
                const float bx = 1.0f - ax, by = 1.0f - ay;
                const float sv = (1.0f - az) * (by * (bx * v[0] + ax * v[1]) + ay * (bx * v[2] + ax * v[3])) +
                                 az * (by * (bx * v[4] + ax * v[5]) + ay * (bx * v[6] + ax * v[7]));
                lookups++;
                const float Pc = (sv - mn) / r;
                if (!ratio) {
                    if (rnd() < Pc) return 0.0f;
                } else {
                    T *= 1.0f - Pc;
                }
            }
            dda_step(alpha, cell, s);
            cur = nxt;
        }
        return T;
    }

    // include/vr_transmittance_DDA.h:62-156
    VR_DEV void t_dda_bw(f3 x, f3 w, float dL_dT) {
        if (dL_dT == 0.0f || med().sigma.grad == nullptr) return;
        Dda s;
        if (!dda_begin(x, w, s)) return;
        double T = exp(-dda_tau(s));
        double dL_dtau = -(double)dL_dT * T;
        double alpha[3] = {s.alpha0[0], s.alpha0[1], s.alpha0[2]};
        int cell[3] = {s.cell[0], s.cell[1], s.cell[2]};
        double cur = 0.0;
        const GridView& g = med().sigma;
        while (cur < s.d - 0.0001) {
            double nxt = fmin(s.d, fmin(alpha[0], fmin(alpha[1], alpha[2])));
            double tm = lerpd(cur, nxt, 0.5);
            double ax = s.vm[0] * tm + s.vn[0] - cell[0];
            double ay = s.vm[1] * tm + s.vn[1] - cell[1];
            double az = s.vm[2] * tm + s.vn[2] - cell[2];
            double gs = dL_dtau * (nxt - cur);
            double bx = 1 - ax, by = 1 - ay, bz = 1 - az;
            double wg[8] = {bx * by * bz, ax * by * bz, bx * ay * bz, ax * ay * bz, bx * by * az, ax * by * az, bx * ay * az, ax * ay * az};
#pragma unroll
            for (int k = 0; k < 8; k++) {
                int xx = max(0, min(g.nx - 1, cell[0] + (k & 1)));
                int yy = max(0, min(g.ny - 1, cell[1] + ((k >> 1) & 1)));
                int zz = max(0, min(g.nz - 1, cell[2] + ((k >> 2) & 1)));
                atomic_addf(g.grad + (zz * g.sz + yy * g.sy + xx * g.sx), (float)(gs * wg[k]));
            }
            dda_step(alpha, cell, s);
            cur = nxt;
        }
    }

    // include/vr_transmittance_DT.h:30-90
    VR_DEV void t_dt_bw(f3 x, f3 w, float dL_dT) {
        if (dL_dT == 0.0f) return;
        float tmin, tmax;
        if (!boundary(med(), x, w, tmin, tmax)) return;
        x = x + w * tmin;
        float d = tmax - tmin;
        f3 x0 = x;
        float d0 = d;
        Rng saved = rng;
        Walk<MAJ> mj;
        mj.begin(med(), x, w);
        for (;;) {
            float md;
            float maj = fmaxf(0.0000001f, mj.query(med(), md) + 0.1f);
            float dt = flight(maj);
            float t = fminf(md, dt);
            x = x + w * t;
            d -= t;
            if (dt > md) { mj.cross(med(), w); if (d <= kCrossEps) break; continue; }
            mj.move(t);
            if (d <= 0.0000001f) break;
            if (rnd() < sigma(x) / maj) return;
        }
        x = x0; d = d0;
        rng = saved;
        mj.begin(med(), x, w);
        for (;;) {
            float md;
            float maj = fmaxf(0.0000001f, mj.query(med(), md) + 0.1f);
            float dt = flight(maj);
            float t = fminf(md, dt);
            x = x + w * t;
            d -= t;
            if (dt > md) { mj.cross(med(), w); if (d <= kCrossEps) break; continue; }
            mj.move(t);
            if (d <= 0.0000001f) break;
            rnd();
            float Pc = sigma(x) / maj;
            sigma_bw(x, -dL_dT / maj / (1.0f - Pc));
        }
    }

    // include/vr_transmittance_RT.h:51-141
    VR_DEV void t_rt_bw(f3 x, f3 w, float dL_dT) {
        if (dL_dT == 0.0f) return;
        float tmin, tmax;
        if (!boundary(med(), x, w, tmin, tmax)) return;
        const float EPS = 0.0000001f;
        float T = 1.0f;
        x = x + w * tmin;
        float d = tmax - tmin;
        f3 x0 = x;
        float d0 = d;
        Rng saved = rng;
        Walk<MAJ> mj;
        mj.begin(med(), x, w);
        for (;;) {
            float md;
            float maj = fmaxf(EPS, mj.query(med(), md) + 0.1f);
            float s = flight(maj);
            if (Walk<MAJ>::grid && md < d && s > md) { x = x + w * md; d -= md; mj.cross(med(), w); if (d <= kCrossEps) break; continue; }
            float dt = fminf(s, fminf(d, md));
            x = x + w * dt;
            d -= dt;
            mj.move(dt);
            if (d <= EPS) break;
            T *= (1.0f - sigma(x) / maj);
            if (T < 0.1f) {
                if (rnd() < 1.0f - T) return;
                T = 1.0f;
            }
        }
        float out_T = T;
        T = 1.0f;
        rng = saved;
        x = x0; d = d0;
        mj.begin(med(), x, w);
        for (;;) {
            float md;
            float maj = fmaxf(EPS, mj.query(med(), md) + 0.1f);
            float s = flight(maj);
            if (Walk<MAJ>::grid && md < d && s > md) { x = x + w * md; d -= md; mj.cross(med(), w); if (d <= kCrossEps) break; continue; }
            float dt = fminf(s, fminf(d, md));
            x = x + w * dt;
            d -= dt;
            mj.move(dt);
            if (d <= EPS) break;
            float r = 1.0f - sigma(x) / maj;
            T *= r;
            sigma_bw(x, -out_T * dL_dT / maj / r);
            if (T < 0.1f) {
                rng_advance(rng);
                T = 1.0f;
            }
        }
    }

    VR_DEV float transmittance(f3 x, f3 w) {
        if (med().t_kind == VR_T_RT) return t_rt(x, w);
        if (med().t_kind == VR_T_DDA) return t_dda(x, w);
        if (med().t_kind == VR_T_GDT || med().t_kind == VR_T_GRT) return t_grid_tracking(x, w, med().t_kind == VR_T_GRT);
        return t_dt(x, w);
    }

    // (GDT and GRT have no backward in the reference: no gradient)
    VR_DEV void transmittance_bw(f3 x, f3 w, float dL_dT) {
        if (med().t_kind == VR_T_RT) t_rt_bw(x, w, dL_dT);
        else if (med().t_kind == VR_T_DDA) t_dda_bw(x, w, dL_dT);
        else if (med().t_kind == VR_T_DT) t_dt_bw(x, w, dL_dT);
    }

    // include/vr_collision_sampler_DT.h:5-44
    VR_DEV float collision_dt(f3 x, f3 w, float& out_t, float& out_T) {
        out_T = 1.0f; out_t = 0.0f;
        float C = 0.0f, tmin, tmax;
        if (!boundary(med(), x, w, tmin, tmax)) return C;
        float d = tmax - tmin;
        x = x + w * tmin;
        float t = 0.0f;
        Walk<MAJ> mj;
        mj.begin(med(), x, w);
        for (;;) {
            float md;
            float maj = mj.query(med(), md);
            if (Walk<MAJ>::grid) maj = fmaxf(0.0000001f, maj);
            float dt = fminf(flight(maj), fminf(d, md));
            x = x + w * dt;
            t += dt;
            if (dt >= md) { mj.cross(med(), w); if (t >= d - kCrossEps) break; continue; }
            mj.move(dt);
            if (t >= d - 0.0000001f) break;
            if (rnd() * maj < sigma(x)) { out_T = 0.0f; C = 1.0f; break; }
        }
        out_t = tmin + t;
        return C;
    }

    // Defensive primal pass of vr_collision_sampler_DT.h:52-90 (majorant
    // m + 2 eps, collision with Pc = (s + eps) / m) from the box entry: whether
    // it collides (C), the distance t from the entry, the product Wpre of the
    // null-collision weights (1 - s/m) / (1 - Pc) before the collision and
    // s_t = (s/m) / Pc at the collision (0 without one).
    struct CollisionPrimal {
        float C, t, Wpre, s_t;
    };
    VR_DEV CollisionPrimal collision_ds_walk(f3 x, f3 w, float eps, float tmin, float tmax) {
        float d = tmax - tmin;
        x = x + w * tmin;
        CollisionPrimal P;
        P.C = 0.0f; P.t = 0.0f; P.Wpre = 1.0f; P.s_t = 0.0f;
        Walk<MAJ> mj;
        mj.begin(med(), x, w);
        for (;;) {
            float md;
            float maj = fmaxf(0.0000000001f, mj.query(med(), md) + 2.0f * eps);
            float dt = fminf(flight(maj), fminf(d, md));
            x = x + w * dt;
            P.t += dt;
            if (dt >= md) { mj.cross(med(), w); if (P.t >= d - kCrossEps) break; continue; }
            mj.move(dt);
            if (P.t >= d - 0.0000001f) break;
            float s = sigma(x);
            float Pc = (s + eps) / maj;
            if (rnd() < Pc) { P.C = 1.0f; P.s_t = (s / maj) / Pc; break; }
            P.Wpre *= (1.0f - s / maj) / (1.0f - Pc);
        }
        return P;
    }

    // Replay pass of vr_collision_sampler_DT.h:92-134 given the primal pass P
    // drawn from the current random state (VR_BW_CONSISTENT: full path weight).
    VR_DEV void collision_dt_replay(f3 x, f3 w, float eps, float tmin, float tmax, const CollisionPrimal& P,
                                    float dL_dC, float dL_dT) {
        const float C = P.C;
        const float out_T = C != 0.0f ? 0.0f : P.Wpre;
        const bool consistent = med().bw_mode == VR_BW_CONSISTENT;
        const float s_t = consistent ? P.s_t * P.Wpre : P.s_t;
        float d = tmax - tmin;
        x = x + w * tmin;
        float t = 0.0f;
        Walk<MAJ> mj;
        mj.begin(med(), x, w);
        for (;;) {
            float md;
            float maj = fmaxf(0.0000000001f, mj.query(med(), md) + 2.0f * eps);
            float dt = fminf(flight(maj), fminf(d, md));
            x = x + w * dt;
            t += dt;
            if (dt >= md) { mj.cross(med(), w); if (t >= d - kCrossEps) break; continue; }
            mj.move(dt);
            if (t >= d - 0.0000001f) break;
            float s = sigma(x);
            float Pc = (s + eps) / maj;
            if (rnd() < Pc) {
                float tr = C / Pc;
                if (consistent) tr *= P.Wpre;
                sigma_bw(x, dL_dC * tr * (1.0f / maj));
                break;
            } else {
                float q = (dL_dT * out_T + dL_dC * C * s_t) / ((1.0f - s / maj) / (1.0f - Pc)) / (1.0f - Pc);
                sigma_bw(x, q * (-1.0f / maj));
            }
        }
    }

    // include/vr_collision_sampler_DT.h:46-134: the primal pass, then the
    // replay from the same random state.
    VR_DEV void collision_dt_bw(f3 x, f3 w, float eps, float dL_dC, float dL_dT) {
        float tmin, tmax;
        if (!boundary(med(), x, w, tmin, tmax)) return;
        Rng saved = rng;
        CollisionPrimal P = collision_ds_walk(x, w, eps, tmin, tmax);
        rng = saved;
        collision_dt_replay(x, w, eps, tmin, tmax, P, dL_dC, dL_dT);
    }

    // include/maps/single_scattered_radiance_mc.h:6-16
    VR_DEV f3 ss_radiance(f3 x, f3 w) {
        f3 wo;
        f3 env = env_sample(med(), rng, wo);
        float ph = phase_eval(med(), x, w * -1.0f, wo);
        f3 sa = albedo(x);
        float T = transmittance(x, wo);
        return (env * sa) * (ph * T);
    }

    // include/maps/single_scattered_radiance_mc.h:19-38
    VR_DEV void ss_radiance_bw(f3 x, f3 w, f3 dL_dR) {
        f3 wo;
        f3 env = env_sample(med(), rng, wo);
        float ph = phase_eval(med(), x, w * -1.0f, wo);
        f3 sa = albedo(x);
        float dL_dT = dot(dL_dR, (env * sa) * ph);
        if (med().bw_mode == VR_BW_CONSISTENT && stream_ends && !albedo_has_grad()) {
            // T enters only the albedo gradient, and no draw follows this
            // estimator: the shadow ray's primal estimate is not needed
            transmittance_bw(x, wo, dL_dT);
            return;
        }
        Rng before_T = rng;
        float T = transmittance(x, wo);
        f3 dL_dsa = (dL_dR * env) * (ph * T);
        albedo_bw(x, dL_dsa);
        if (med().bw_mode == VR_BW_CONSISTENT) {
            Rng after = rng;
            rng = before_T;
            transmittance_bw(x, wo, dL_dT);
            rng = after;
        } else {
            transmittance_bw(x, w, dL_dT);
        }
    }

    // include/maps/collision_integrator_mc.h:10-28 (exitance = SS radiance)
    VR_DEV f3 mc_collision(f3 x, f3 w) {
        float T, t;
        float C = collision_dt(x, w, t, T);
        f3 xt = w * t + x;
        f3 R = env_eval(med(), w) * T;
        if (C != 0.0f) R = R + ss_radiance(xt, w * -1.0f) * C;
        return R;
    }

    // include/maps/collision_integrator_mc.h:30-44
    VR_DEV void mc_collision_bw(f3 x, f3 w, f3 dL_dR) {
        if (med().bw_mode == VR_BW_REFERENCE) {
            float T, t;
            float C = collision_dt(x, w, t, T);
            f3 xt = w * t + x;
            f3 eR = ss_radiance(xt, w * -1.0f);
            collision_dt_bw(x, w, med().ds_eps, dot(dL_dR, eR), dot(dL_dR, env_eval(med(), w)));
            ss_radiance_bw(xt, w * -1.0f, dL_dR * C);
            return;
        }
        float tmin, tmax;
        if (!boundary(med(), x, w, tmin, tmax)) return;
        // one primal pass serves the exitance estimate and the sampler's replay
        // (the nested form of the reference walks it twice from the same state)
        Rng saved = rng;
        const float eps = med().ds_eps;
        CollisionPrimal P = collision_ds_walk(x, w, eps, tmin, tmax);
        const float C = P.C;
        const float Wc = C != 0.0f ? P.Wpre * P.s_t : P.Wpre;
        f3 xt = x + w * (tmin + P.t);
        f3 eR = F3(0.0f, 0.0f, 0.0f);
        if (C != 0.0f) eR = ss_radiance(xt, w * -1.0f);
        Rng after = rng;
        rng = saved;
        collision_dt_replay(x, w, eps, tmin, tmax, P, dot(dL_dR, eR), dot(dL_dR, env_eval(med(), w)));
        rng = after;
        if (C != 0.0f) ss_radiance_bw(xt, w * -1.0f, dL_dR * Wc);
    }

    // include/vr_path_integrator_NEE.h:1-66 (max_bounces >= 0 caps scatter events)
    VR_DEV f3 nee(f3 x, f3 w) {
        float tmin, tmax;
        boundary(med(), x, w, tmin, tmax);
        x = x + w * tmin;
        float d = tmax - tmin;
        f3 R = env_eval(med(), w) * transmittance(x, w);
        f3 W = F3(1.0f, 1.0f, 1.0f);
        const float EPS = 0.00001f;
        int bounces = 0;
        Walk<MAJ> mj;
        mj.begin(med(), x, w);
        while (d > EPS) {
            float md;
            float m = fmaxf(EPS, mj.query(med(), md));
            float t = flight(m);
            x = x + w * fminf(t, md);
            if (t > md) { d -= md; mj.cross(med(), w); if (d <= kCrossEps) break; continue; }
            if (t > d) break;
            mj.move(t);
            float s = sigma(x);
            if (rnd() * m >= s) { d -= t; continue; }
            f3 sa = albedo(x);
            R = R + (W * (F3(1.0f, 1.0f, 1.0f) - sa)) * med().emission_c;
            W = W * sa;
            f3 wnee;
            f3 env = env_sample(med(), rng, wnee);
            float Tn = transmittance(x, wnee);
            R = R + (W * env) * (Tn * phase_eval(med(), x, w, wnee));
            bounces++;
            if (med().max_bounces >= 0 && bounces >= med().max_bounces) return R;
            f3 wo;
            float rho = phase_sample(med(), rng, x, w, wo);
            W = W * rho;
            if (W.x < EPS && W.y < EPS && W.z < EPS) return R;
            w = wo;
            boundary(med(), x, w, tmin, d);
            mj.begin(med(), x, w);
        }
        return R;
    }

    // include/vr_path_integrator_NEE_DS.h:6-206
    VR_DEV void nee_ds_bw(f3 x, f3 w, float eps, f3 dL_dR) {
        Rng before_path = rng;
        f3 x0 = x, w0 = w;
        float tmin, tmax;
        boundary(med(), x, w, tmin, tmax);
        x = x + w * tmin;
        float d = tmax - tmin;
        f3 R = F3(0.0f, 0.0f, 0.0f), R_sp = F3(0.0f, 0.0f, 0.0f), W = F3(1.0f, 1.0f, 1.0f);
        float null_W = 1.0f;
        const float EPS = 0.00000001f;
        bool scattered = false;
        int bounces = 0;
        Walk<MAJ> mj;
        mj.begin(med(), x, w);
        while (d > EPS) {
            float md;
            float m = fmaxf(EPS, mj.query(med(), md) + 2.0f * eps + 0.1f);
            float t = flight(m);
            x = x + w * fminf(t, md);
            if (t > md) { d -= md; mj.cross(med(), w); if (d <= kCrossEps) break; continue; }
            if (t > d) break;
            mj.move(t);
            float s = sigma(x);
            float Pc = (s + eps) / m;
            if (rnd() >= Pc) { d -= t; W = W * ((1.0f - s / m) / (1.0f - Pc)); continue; }
            scattered = true;
            if (s < EPS) {
                if (null_W < EPS) { W = F3(0.0f, 0.0f, 0.0f); break; }
                null_W = 0.0f;
            } else W = W * ((s / m) / Pc);
            f3 sa = albedo(x);
            f3 added = (W * (F3(1.0f, 1.0f, 1.0f) - sa)) * med().emission_c;
            W = W * sa;
            f3 wnee;
            f3 env = env_sample(med(), rng, wnee);
            float Tn = transmittance(x, wnee);
            float phn = phase_eval(med(), x, w, wnee);
            added = added + (W * env) * (Tn * phn);
            if (null_W < EPS) R_sp = R_sp + added;
            else R = R + added;
            bounces++;
            if (med().max_bounces >= 0 && bounces >= med().max_bounces) { W = F3(0.0f, 0.0f, 0.0f); break; }
            f3 wo;
            float rho = phase_sample(med(), rng, x, w, wo);
            W = W * rho;
            if (W.x < EPS && W.y < EPS && W.z < EPS) { W = F3(0.0f, 0.0f, 0.0f); break; }
            w = wo;
            boundary(med(), x, w, tmin, d);
            mj.begin(med(), x, w);
        }
        if (!scattered) R = R + env_eval(med(), w) * W;

        rng = before_path;
        x = x0; w = w0;
        W = F3(1.0f, 1.0f, 1.0f);
        boundary(med(), x, w, tmin, tmax);
        x = x + w * tmin;
        d = tmax - tmin;
        bounces = 0;
        mj.begin(med(), x, w);
        const bool ref = med().bw_mode == VR_BW_REFERENCE;
        while (d > EPS) {
            float md;
            float mq = mj.query(med(), md);
            float m = ref ? fmaxf(EPS, mq * 1.2f + 2.0f * eps) : fmaxf(EPS, mq + 2.0f * eps + 0.1f);
            float t = flight(m);
            x = x + w * fminf(t, md);
            if (t > md) { d -= md; mj.cross(med(), w); if (d <= kCrossEps) break; continue; }
            if (t > d) break;
            mj.move(t);
            float s = sigma(x);
            float Pc = (s + eps) / m;
            if (rnd() >= Pc) {
                sigma_bw(x, dot(dL_dR, R * (-1.0f / m / (1.0f - s / m))));
                W = W * ((1.0f - s / m) / (1.0f - Pc));
                d -= t;
                continue;
            }
            float gs = s < EPS ? dot(dL_dR, R_sp * (1.0f / eps)) : dot(dL_dR, R * (1.0f / s));
            sigma_bw(x, gs);
            W = W * (s / m / Pc);
            if (s < EPS) return;
            f3 sa = albedo(x);
            f3 dL_dsa = (dL_dR * W) * (med().emission_c * -1.0f);
            R = R - (W * (F3(1.0f, 1.0f, 1.0f) - sa)) * med().emission_c;
            dL_dsa = dL_dsa + dL_dR * F3(R.x / fmaxf(EPS, sa.x), R.y / fmaxf(EPS, sa.y), R.z / fmaxf(EPS, sa.z));
            f3 wnee;
            f3 env = env_sample(med(), rng, wnee);
            Rng before_T = rng;
            float Tn = transmittance(x, wnee);
            float phn = phase_eval(med(), x, w, wnee);
            f3 wo;
            float rho = phase_sample(med(), rng, x, w, wo);
            Rng before_next = rng;
            W = W * sa;
            float dL_dTn = dot(dL_dR, (W * env) * phn);
            R = R - (W * env) * (Tn * phn);
            albedo_bw(x, dL_dsa);
            rng = before_T;
            transmittance_bw(x, wnee, dL_dTn);
            rng = before_next;
            bounces++;
            if (med().max_bounces >= 0 && bounces >= med().max_bounces) return;
            W = W * rho;
            if (W.x < EPS && W.y < EPS && W.z < EPS) return;
            w = wo;
            boundary(med(), x, w, tmin, d);
            mj.begin(med(), x, w);
        }
    }

    // include/vr_path_integrator_DT.h:7-56 (max_bounces >= 0 ends the path after
    // that many scatter events, with no escape term)
    VR_DEV f3 path_dt(f3 x, f3 w) {
        float tmin, tmax;
        boundary(med(), x, w, tmin, tmax);
        x = x + w * tmin;
        float d = tmax - tmin;
        f3 R = F3(0.0f, 0.0f, 0.0f), W = F3(1.0f, 1.0f, 1.0f);
        int bounces = 0;
        Walk<MAJ> mj;
        mj.begin(med(), x, w);
        while (d > 0.00000001f) {
            float md;
            float m = fmaxf(0.00000001f, mj.query(med(), md));
            float t = flight(m);
            x = x + w * fminf(t, md);
            if (t > md) { d -= md; mj.cross(med(), w); if (d <= kCrossEps) break; continue; }
            if (t > d) break;
            mj.move(t);
            float s = sigma(x);
            if (rnd() * m >= s) { d -= t; continue; }
            f3 sa = albedo(x);
            R = R + (W * (F3(1.0f, 1.0f, 1.0f) - sa)) * med().emission_c;
            W = W * sa;
            f3 wo;
            float rho = phase_sample(med(), rng, x, w, wo);
            W = W * rho;
            bounces++;
            if (med().max_bounces >= 0 && bounces >= med().max_bounces) return R;
            if (W.x < 0.0000001f && W.y < 0.0000001f && W.z < 0.0000001f) return R;
            w = wo;
            boundary(med(), x, w, tmin, d);
            mj.begin(med(), x, w);
        }
        return W * env_eval(med(), w) + R;
    }

    // include/vr_path_integrator_DT.h:59-235: defensive pass (majorant + 2 eps,
    // collision probability (s + eps)/m) for the path throughput, then a replay
    // of the same stream scattering d<W env>/d sigma at every collision.
    VR_DEV void path_dt_bw(f3 x, f3 w, float eps, f3 dL_dR) {
        Rng before_path = rng;
        f3 x0 = x, w0 = w;
        float tmin, tmax;
        boundary(med(), x, w, tmin, tmax);
        x = x + w * tmin;
        float d = tmax - tmin;
        const float EPS = 0.00000001f;
        f3 W = F3(1.0f, 1.0f, 1.0f);
        float null_W = 1.0f;
        int bounces = 0;
        Walk<MAJ> mj;
        mj.begin(med(), x, w);
        while (d > EPS) {
            float md;
            float m = fmaxf(EPS, mj.query(med(), md) + 2.0f * eps);
            float t = flight(m);
            x = x + w * fminf(t, md);
            if (t > md) { d -= md; mj.cross(med(), w); if (d <= kCrossEps) break; continue; }
            if (t > d) break;
            mj.move(t);
            float s = sigma(x);
            float Pc = (s + eps) / m;
            if (rnd() >= Pc) { d -= t; W = W * ((1.0f - s / m) / (1.0f - Pc)); continue; }
            if (s == 0.0f) {
                if (null_W == 0.0f) { W = F3(0.0f, 0.0f, 0.0f); break; }
                null_W = 0.0f;
            } else W = W * ((s / m) / Pc);
            f3 sa = albedo(x);
            f3 wo;
            float rho = phase_sample(med(), rng, x, w, wo);
            W = (W * sa) * rho;
            bounces++;
            if (med().max_bounces >= 0 && bounces >= med().max_bounces) { W = F3(0.0f, 0.0f, 0.0f); break; }
            if (W.x < EPS && W.y < EPS && W.z < EPS) { W = F3(0.0f, 0.0f, 0.0f); break; }
            w = wo;
            boundary(med(), x, w, tmin, d);
            mj.begin(med(), x, w);
        }
        const f3 dL_dW = dL_dR * env_eval(med(), w);
        const f3 W_out = W;

        rng = before_path;
        x = x0; w = w0;
        boundary(med(), x, w, tmin, tmax);
        x = x + w * tmin;
        d = tmax - tmin;
        bounces = 0;
        mj.begin(med(), x, w);
        while (d > EPS) {
            float md;
            float m = fmaxf(EPS, mj.query(med(), md) + 2.0f * eps);
            float t = flight(m);
            x = x + w * fminf(t, md);
            if (t > md) { d -= md; mj.cross(med(), w); if (d <= kCrossEps) break; continue; }
            if (t > d) break;
            mj.move(t);
            float s = sigma(x);
            float Pc = (s + eps) / m;
            if (rnd() >= Pc) {
                sigma_bw(x, dot(dL_dW, W_out * (-null_W / m / (1.0f - s / m))));
                d -= t;
                continue;
            }
            sigma_bw(x, dot(dL_dW, s == 0.0f ? W_out : W_out * (null_W / s)));
            if (s == 0.0f) break;
            f3 wo;
            phase_sample(med(), rng, x, w, wo);
            bounces++;
            if (med().max_bounces >= 0 && bounces >= med().max_bounces) return;
            w = wo;
            boundary(med(), x, w, tmin, d);
            mj.begin(med(), x, w);
        }
    }

    // randoms.h:131-139
    VR_DEV Rng create_branch_seed() {
        Rng s;
        s.x = __float_as_uint(rnd()) + 129u;
        s.y = __float_as_uint(rnd()) + 129u;
        s.z = __float_as_uint(rnd()) + 129u;
        s.w = __float_as_uint(rnd()) + 129u;
        rng_advance(s); rng_advance(s); rng_advance(s); rng_advance(s);
        return s;
    }

    // include/vr_path_integrator_NEE.h:1-66 with only_scatters = true
    VR_DEV f3 nee_scatters(f3 x, f3 w) {
        float tmin, tmax;
        boundary(med(), x, w, tmin, tmax);
        x = x + w * tmin;
        float d = tmax - tmin;
        f3 R = F3(0.0f, 0.0f, 0.0f), W = F3(1.0f, 1.0f, 1.0f);
        const float EPS = 0.00001f;
        int bounces = 0;
        Walk<MAJ> mj;
        mj.begin(med(), x, w);
        while (d > EPS) {
            float md;
            float m = fmaxf(EPS, mj.query(med(), md));
            float t = flight(m);
            x = x + w * fminf(t, md);
            if (t > md) { d -= md; mj.cross(med(), w); if (d <= kCrossEps) break; continue; }
            if (t > d) break;
            mj.move(t);
            float s = sigma(x);
            if (rnd() * m >= s) { d -= t; continue; }
            f3 sa = albedo(x);
            R = R + (W * (F3(1.0f, 1.0f, 1.0f) - sa)) * med().emission_c;
            W = W * sa;
            f3 wnee;
            f3 env = env_sample(med(), rng, wnee);
            float Tn = transmittance(x, wnee);
            R = R + (W * env) * (Tn * phase_eval(med(), x, w, wnee));
            bounces++;
            if (med().max_bounces >= 0 && bounces >= med().max_bounces) return R;
            f3 wo;
            float rho = phase_sample(med(), rng, x, w, wo);
            W = W * rho;
            if (W.x < EPS && W.y < EPS && W.z < EPS) return R;
            w = wo;
            boundary(med(), x, w, tmin, d);
            mj.begin(med(), x, w);
        }
        return R;
    }

    // include/vr_path_integrator_NEE_DRT.h:8-44 (the majorant is queried along
    // the walk -- the reference queries it at the start point, which is the same
    // for a constant majorant -- and the end of the segment is tested first)
    VR_DEV float sample_transmittance_position(f3 x, f3 w, float d, float& sample_t, float& weight_t) {
        float T = 1.0f, t = 0.0f, res_t = 0.0f, res_dt = 0.0f, C = 0.0f;
        const float EPS = 0.0000001f;
        Walk<MAJ> mj;
        mj.begin(med(), x, w);
        for (;;) {
            float md;
            float m = fmaxf(EPS, mj.query(med(), md));
            float dt = fminf(flight(m), fminf(md, d - t));
            float w_step = T * dt;
            C += w_step;
            if (rnd() < w_step / C) { res_t = t; res_dt = dt; }
            t += dt;
            if (t >= d - EPS) break;  // before a crossing: the last cell's exit can coincide with the end
            if (dt >= md - EPS) { mj.cross(med(), w); continue; }
            mj.move(dt);
            T *= 1.0f - sigma(x + w * t) / m;
        }
        weight_t = C;
        sample_t = res_t + rnd() * res_dt;
        return T;
    }

    // include/vr_path_integrator_NEE_DRT.h:46-269: primal pass with a reservoir
    // over the scattered segments, replay with the MIS share s^2/(1+s^2) of each
    // real collision, then the remaining share at a transmittance-sampled point
    // of the reservoir segment.  A branch seed drawn from the main stream drives
    // every transmittance backward.
    VR_DEV void nee_drt_bw(f3 x, f3 w, f3 dL_dR) {
        float tmin, tmax;
        boundary(med(), x, w, tmin, tmax);
        x = x + w * tmin;
        float d = tmax - tmin;
        if (d <= 0.0f) return;
        Rng secondary = create_branch_seed(), main_s;
        transmittance(x, w);  // the primary term's T: draws from the main stream
        main_s = rng; rng = secondary;
        transmittance_bw(x, w, dot(dL_dR, env_eval(med(), w)));
        secondary = rng; rng = main_s;

        f3 R = F3(0.0f, 0.0f, 0.0f), W = F3(1.0f, 1.0f, 1.0f);
        f3 res_x = x, res_w = w, res_W = W;
        float res_P = 3.0f, total_weight = 3.0f;
        const f3 ini_x = x, ini_w = w;
        const float ini_d = d;
        const float EPS = 0.00000001f;
        const Rng before_path = rng;
        int bounces = 0;
        Walk<MAJ> mj;
        mj.begin(med(), x, w);
        while (d > EPS) {
            float md;
            float m = fmaxf(0.1f, mj.query(med(), md) * 1.05f + 0.1f);  // a NaN cell bound counts as empty
            float t = flight(m);
            x = x + w * fminf(t, md);
            if (t > md) { d -= md; mj.cross(med(), w); if (d <= kCrossEps) break; continue; }
            if (t > d) break;
            mj.move(t);
            float s = sigma(x);
            if (rnd() >= s / m) { d -= t; continue; }
            W = W * albedo(x);
            f3 wnee;
            f3 env = env_sample(med(), rng, wnee);
            float Tn = transmittance(x, wnee);
            float phn = phase_eval(med(), x, w, wnee);
            R = R + (W * env) * (Tn * phn);
            bounces++;
            if (med().max_bounces >= 0 && bounces >= med().max_bounces) break;
            f3 wo;
            float rho = phase_sample(med(), rng, x, w, wo);
            W = W * rho;
            w = wo;
            boundary(med(), x, w, tmin, d);
            mj.begin(med(), x, w);
            float cur_P = W.x + W.y + W.z;
            total_weight += cur_P;
            if (rnd() < cur_P / total_weight) { res_P = cur_P; res_W = W; res_x = x; res_w = w; }
            if (W.x < EPS && W.y < EPS && W.z < EPS) break;
        }

        W = F3(1.0f, 1.0f, 1.0f);
        x = ini_x; w = ini_w; d = ini_d;
        rng = before_path;
        bounces = 0;
        mj.begin(med(), x, w);
        while (d > EPS) {
            float md;
            float m = fmaxf(0.1f, mj.query(med(), md) * 1.05f + 0.1f);  // a NaN cell bound counts as empty
            float t = flight(m);
            x = x + w * fminf(t, md);
            if (t > md) { d -= md; mj.cross(med(), w); if (d <= kCrossEps) break; continue; }
            if (t > d) break;
            mj.move(t);
            float s = sigma(x);
            float Pc = s / m;
            if (rnd() >= Pc) {
                float g = dot(dL_dR, R * (-1.0f / m / (1.0f - s / m)));
                main_s = rng; rng = secondary;
                sigma_bw(x, g);
                secondary = rng; rng = main_s;
                d -= t;
                continue;
            }
            float g = dot(dL_dR, R * ((s * s / (1.0f + s * s)) / s));
            W = W * albedo(x);
            f3 wnee;
            f3 env = env_sample(med(), rng, wnee);
            float Tn = transmittance(x, wnee);
            float phn = phase_eval(med(), x, w, wnee);
            bounces++;
            const bool last = med().max_bounces >= 0 && bounces >= med().max_bounces;
            f3 wo = w;
            float rho = last ? 1.0f : phase_sample(med(), rng, x, w, wo);
            float dL_dTn = dot(dL_dR, (W * env) * phn);
            R = R - (W * env) * (Tn * phn);
            main_s = rng; rng = secondary;
            sigma_bw(x, g);
            transmittance_bw(x, wnee, dL_dTn);
            secondary = rng; rng = main_s;
            if (last) break;
            W = W * rho;
            w = wo;
            boundary(med(), x, w, tmin, d);
            mj.begin(med(), x, w);
            rnd();  // the reservoir draw of the primal pass
            if (W.x < EPS && W.y < EPS && W.z < EPS) break;
        }

        boundary(med(), res_x, res_w, tmin, d);
        float s_t, weight_t;
        sample_transmittance_position(res_x, res_w, d, s_t, weight_t);
        f3 xs = res_x + res_w * s_t;
        f3 lw;
        f3 Lenv = env_sample(med(), rng, lw);
        f3 Ls = Lenv * (transmittance(xs, lw) * phase_eval(med(), xs, res_w, lw));
        f3 ws;
        float rho_s = phase_sample(med(), rng, xs, res_w, ws);
        Ls = Ls + nee_scatters(xs, ws) * rho_s;
        Ls = ((Ls * res_W) * weight_t) * albedo(xs);
        float sxs = sigma(xs);
        float g = dot(dL_dR, Ls) * (1.0f / (1.0f + sxs * sxs)) * total_weight / res_P;
        main_s = rng; rng = secondary;
        sigma_bw(xs, g);
        secondary = rng; rng = main_s;
    }

    // include/sc_medium_path_integrator_DT_fw.h:1-32
    VR_DEV void medium_dt(f3 x, f3 w, float d, float* out) {
        f3 W = F3(1.0f, 1.0f, 1.0f), A = F3(0.0f, 0.0f, 0.0f);
        f3 xo = w * d + x, wo = w;
        float maj = fmaxf(0.00001f, med().medium_majorant);
        for (;;) {
            float t = flight(maj);
            if (t > d - 0.00001f) break;
            x = x + w * t;
            d -= t;
            float s = sigma(x);
            if (rnd() < s / maj) {
                if (med().medium_filter & 1) {
                    xo = x;
                    phase_sample(med(), rng, x, w, wo);
                    W = albedo(x);
                } else W = F3(0.0f, 0.0f, 0.0f);
                if (med().medium_filter & 2) A = med().gathering_c;
                break;
            }
        }
        out[0] = xo.x; out[1] = xo.y; out[2] = xo.z;
        out[3] = wo.x; out[4] = wo.y; out[5] = wo.z;
        out[6] = W.x; out[7] = W.y; out[8] = W.z;
        out[9] = A.x; out[10] = A.y; out[11] = A.z;
    }

    // include/sc_medium_path_integrator_DT_bw.h:1-86
    VR_DEV void medium_dt_bw(f3 x, f3 w, float d, f3 dL_dW, f3 dL_dA) {
        f3 W = F3(1.0f, 1.0f, 1.0f), A = F3(0.0f, 0.0f, 0.0f);
        f3 xi = x, wi = w;
        float di = d;
        Rng before = rng;
        float maj = fmaxf(0.00001f, med().medium_majorant);
        int f = med().medium_filter;
        for (;;) {
            float t = flight(maj);
            if (t > d - 0.00001f) break;
            x = x + w * t;
            d -= t;
            float s = sigma(x);
            if (rnd() < s / maj) {
                if (f & 1) { f3 wo; float wt = phase_sample(med(), rng, x, w, wo); W = albedo(x) * wt; }
                else W = F3(0.0f, 0.0f, 0.0f);
                if (f & 2) A = med().gathering_c;
                break;
            }
        }
        rng = before;
        medium_dt_replay(xi, wi, di, W, A, dL_dW, dL_dA);
    }

    // The replay half of medium_dt_bw given the primal pass's (W, A), which
    // equal the forward's outputs: a caller that has just run medium_dt on the
    // same state (the scene backward, sc_path_tracing.h:281-296) skips the
    // primal pass.
    VR_DEV void medium_dt_replay(f3 x, f3 w, float d, f3 W, f3 A, f3 dL_dW, f3 dL_dA) {
        const float maj = fmaxf(0.00001f, med().medium_majorant);
        const int f = med().medium_filter;
        const bool scat_blk = (f & 1) != 0;
        const bool acc_blk = med().bw_mode == VR_BW_REFERENCE ? (f & 1) != 0 : (f & 2) != 0;
        for (;;) {
            float t = flight(maj);
            if (t > d - 0.00001f) break;
            x = x + w * t;
            d -= t;
            float s = sigma(x);
            float Pc = s / maj;
            if (rnd() < Pc) {
                float g = 0.0f;
                if (scat_blk) {
                    f3 wo;
                    phase_sample(med(), rng, x, w, wo);
                    g += dot(dL_dW, W * (1.0f / s));
                    albedo_bw(x, dL_dW);
                }
                if (acc_blk) g += dot(dL_dA, A * (1.0f / s));
                if (g != 0.0f) sigma_bw(x, g);
                break;
            } else {
                float g = -dot(dL_dW, W) / maj / (1.0f - Pc);
                if (acc_blk) g += -dot(dL_dA, A) / maj / (1.0f - Pc);
                if (g != 0.0f) sigma_bw(x, g);
            }
        }
    }
};

// ------------------------------------------------------------------ camera --
// include/maps/sensor_perspective_camera.h:1-27
VR_DEV void camera_ray(const CameraView& cam, int pose, int row, int col, Rng& rng, f3& xo, f3& wo) {
    const float* P = cam.poses + 9 * pose;
    f3 o = load3(P), d = load3(P + 3), n = load3(P + 6);
    float u = 0.5f, v = 0.5f;
    if (cam.jittered) { u = rng_uniform(rng); v = rng_uniform(rng); }
    float sx = (((float)col + u) * 2.0f - (float)cam.width) * cam.znear_over_h;
    float sy = (((float)row + v) * 2.0f - (float)cam.height) * cam.znear_over_h;
    float sz = cam.sz;
    f3 za = normalize(d);
    f3 xa = normalize(cross(n, za));
    f3 ya = cross(za, xa);
    f3 w = xa * sx + ya * sy + za * sz;
    xo = o + w;
    wo = normalize(w);
}

VR_DEV void sensor_index(const CameraView& cam, int64_t i, int& pose, int& row, int& col) {
    if (cam.sensors) {
        pose = (int)cam.sensors[3 * i];
        row = (int)cam.sensors[3 * i + 1];
        col = (int)cam.sensors[3 * i + 2];
    } else {
        col = (int)(i % cam.width);
        row = (int)((i / cam.width) % cam.height);
        pose = (int)(i / ((int64_t)cam.width * cam.height));
    }
}

}  // namespace vr
