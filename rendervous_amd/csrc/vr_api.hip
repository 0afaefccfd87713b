// vr_api.hip -- kernels and C ABI (include/vr_api.h) of the MI355X
// participating-media library.
//
// Launch model: one work-item per sensor (pixel) or per input row, wave64,
// 256-thread workgroups (the reference dispatches 32-wide workgroups,
// _maps.py:523, which leaves half of each CDNA wave empty).  Each capture
// work-item loops over its samples and re-seeds the HybridTaus state of sample
// s from the flat index s * threads + i, so sample 0 draws exactly the
// reference's per-thread stream (_maps.py:549-560) and every sample can be
// replayed on its own by the backward kernels.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <mutex>
#include <vector>

#include "vr_device.hpp"
#include "vr_nee_wavefront.hpp"
#include "vr_nrc.hpp"
#include "vr_functions.hpp"
#include "vr_scene.hpp"
#include "vr_nee_bw_flat.hpp"
#include "vr_ss_bw_flat.hpp"

using namespace vr;

namespace {

thread_local char g_err[512];

int set_err(int code, const char* fmt, const char* a = "") {
    std::snprintf(g_err, sizeof(g_err), fmt, a);
    return code;
}

int check_hip(hipError_t e, const char* where) {
    if (e == hipSuccess) return VR_OK;
    std::snprintf(g_err, sizeof(g_err), "%s: %s", where, hipGetErrorString(e));
    return VR_ERR_DEVICE;
}

constexpr int kBlock = 256;

inline unsigned grid_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// ------------------------------------------------------------ host views --
int make_grid_view(const vr_grid* g, GridView& v, int want_c, const char* what) {
    if (!g) return set_err(VR_ERR_ARG, "%s: grid is NULL", what);
    const vr_tensor& t = g->t;
    if (!t.data) return set_err(VR_ERR_ARG, "%s: grid data is NULL", what);
    if (t.ndim != 4) return set_err(VR_ERR_ARG, "%s: grid must be 4-D [D,H,W,C]", what);
    if (t.shape[0] < 2 || t.shape[1] < 2 || t.shape[2] < 2)
        return set_err(VR_ERR_ARG, "%s: grid needs at least 2 vertices per axis", what);
    if (t.shape[3] < 1 || t.shape[3] > 4) return set_err(VR_ERR_ARG, "%s: grid channels must be 1..4", what);
    if (want_c > 0 && t.shape[3] != want_c) return set_err(VR_ERR_ARG, "%s: wrong channel count", what);
    int64_t span = 1;
    for (int i = 0; i < 4; i++) {
        if (t.stride[i] < 0) return set_err(VR_ERR_ARG, "%s: negative strides are not supported", what);
        span += (t.shape[i] - 1) * t.stride[i];
    }
    if (span >= (int64_t(1) << 31)) return set_err(VR_ERR_ARG, "%s: grid spans more than 2^31 elements", what);
    v.data = t.data;
    v.grad = t.grad;
    v.sz = (int)t.stride[0]; v.sy = (int)t.stride[1]; v.sx = (int)t.stride[2]; v.sc = (int)t.stride[3];
    v.nz = (int)t.shape[0]; v.ny = (int)t.shape[1]; v.nx = (int)t.shape[2]; v.C = (int)t.shape[3];
    double n[3] = {(double)v.nx - 1, (double)v.ny - 1, (double)v.nz - 1};
    float sc[3], of[3];
    for (int a = 0; a < 3; a++) {
        double ext = (double)g->bmax[a] - (double)g->bmin[a];
        if (!(ext > 0)) return set_err(VR_ERR_ARG, "%s: bmax must exceed bmin", what);
        double s = n[a] / ext;
        sc[a] = (float)s;
        of[a] = (float)(-(double)g->bmin[a] * s);
    }
    v.scale = F3(sc[0], sc[1], sc[2]);
    v.offset = F3(of[0], of[1], of[2]);
    v.dimf = F3((float)n[0], (float)n[1], (float)n[2]);
    v.bmin = F3(g->bmin[0], g->bmin[1], g->bmin[2]);
    v.bmax = F3(g->bmax[0], g->bmax[1], g->bmax[2]);
    return VR_OK;
}

int map_dims(int kind, int& in_dim, int& out_dim) {
    switch (kind) {
    case VR_MAP_RAYBOX: in_dim = 6; out_dim = 2; return 1;
    case VR_MAP_T_DDA: case VR_MAP_T_DT: case VR_MAP_T_RT: case VR_MAP_T_GDT: case VR_MAP_T_GRT:
        in_dim = 6; out_dim = 1; return 1;
    case VR_MAP_COLLISION_DT: case VR_MAP_SS_RADIANCE: case VR_MAP_MC_COLLISION: case VR_MAP_NEE_DT:
    case VR_MAP_PATH_DT: case VR_MAP_NEE_DRT:
        in_dim = 6; out_dim = 3; return 1;
    case VR_MAP_MEDIUM_DT: in_dim = 7; out_dim = 12; return 1;
    default: return 0;
    }
}

int make_env_view(const vr_medium* m, MediumView& v) {
    v.env_kind = m->env_kind;
    if (m->env_kind == VR_ENV_XR) {
        const vr_tensor& t = m->env_image;
        if (!t.data || t.ndim != 3 || t.shape[2] != 3) return set_err(VR_ERR_ARG, "environment image must be [H, W, 3]");
        if (t.shape[0] < 1 || t.shape[1] < 1 || t.shape[0] > (1 << 16) || t.shape[1] > (1 << 16))
            return set_err(VR_ERR_ARG, "environment image extent out of range");
        int64_t span = 1;
        for (int i = 0; i < 3; i++) {
            if (t.stride[i] < 0) return set_err(VR_ERR_ARG, "environment image: negative strides are not supported");
            span += (t.shape[i] - 1) * t.stride[i];
        }
        if (span >= (int64_t(1) << 31)) return set_err(VR_ERR_ARG, "environment image spans more than 2^31 elements");
        ImageView& I = v.env_img;
        I.data = t.data;
        I.h = (int)t.shape[0]; I.w = (int)t.shape[1];
        I.sy = (int)t.stride[0]; I.sx = (int)t.stride[1]; I.sc = (int)t.stride[2];
        for (int a = 0; a < 2; a++)
            if (!(m->env_image_bmax[a] > m->env_image_bmin[a])) return set_err(VR_ERR_ARG, "environment image box is empty");
        I.bmin_x = m->env_image_bmin[0]; I.bmin_y = m->env_image_bmin[1];
        I.ibs_x = 1.0f / (m->env_image_bmax[0] - m->env_image_bmin[0]);
        I.ibs_y = 1.0f / (m->env_image_bmax[1] - m->env_image_bmin[1]);
    } else if (m->env_kind != VR_ENV_CONST) return set_err(VR_ERR_ARG, "unknown environment kind");
    v.env_sampler = m->env_sampler_kind;
    if (m->env_sampler_kind == VR_ENVS_QUADTREE) {
        if (!m->env_quadtree) return set_err(VR_ERR_ARG, "environment quadtree is NULL");
        if (m->env_quadtree_levels < 1 || m->env_quadtree_levels > 15)
            return set_err(VR_ERR_ARG, "environment quadtree levels must be in [1, 15]");
        v.quadtree = m->env_quadtree;
        v.qt_levels = m->env_quadtree_levels;
    } else if (m->env_sampler_kind != VR_ENVS_UNIFORM) return set_err(VR_ERR_ARG, "unknown environment sampler kind");
    return VR_OK;
}

int make_medium_view(const vr_medium* m, int kind, int bw_mode, MediumView& v) {
    if (!m) return set_err(VR_ERR_ARG, "medium is NULL");
    std::memset(&v, 0, sizeof(v));
    int st;
    if (kind != VR_MAP_RAYBOX) {
        if ((st = make_grid_view(&m->sigma, v.sigma, 1, "sigma"))) return st;
    }
    v.albedo_grid = m->albedo_kind == VR_FIELD_GRID;
    if (v.albedo_grid) {
        if ((st = make_grid_view(&m->albedo, v.albedo, 0, "albedo"))) return st;
        if (v.albedo.C != 1 && v.albedo.C != 3) return set_err(VR_ERR_ARG, "albedo grid needs 1 or 3 channels");
    } else if (m->albedo_kind != VR_FIELD_CONST) return set_err(VR_ERR_ARG, "unknown albedo kind");
    v.albedo_c = load3(m->albedo_c);
    v.box_min = load3(m->box_min);
    v.box_max = load3(m->box_max);
    v.emission_c = load3(m->emission_c);
    v.env_c = load3(m->env_c);
    v.sun_dir = load3(m->sun_dir);
    v.sun_radiance = load3(m->sun_radiance);
    v.sun_prob = m->sun_prob;
    if (m->sun_prob < 0 || m->sun_prob >= 1) return set_err(VR_ERR_ARG, "sun_prob must be in [0,1)");
    if (m->phase_kind < VR_PHASE_HG_MIX || m->phase_kind > VR_PHASE_ISO) return set_err(VR_ERR_ARG, "unknown phase kind");
    v.phase_kind = m->phase_kind;
    v.g = m->phase_g;
    v.phase_g_grid = m->phase_g_kind == VR_FIELD_GRID && m->phase_kind != VR_PHASE_ISO;
    if (v.phase_g_grid) {
        if ((st = make_grid_view(&m->phase_g_grid, v.phase_g, 1, "phase_g"))) return st;
        v.phase_g.grad = nullptr;  // the phase asymmetry receives no gradient
    } else if (m->phase_g_kind != VR_FIELD_CONST && m->phase_g_kind != VR_FIELD_GRID)
        return set_err(VR_ERR_ARG, "unknown phase_g kind");
    v.maj_value = m->majorant_value;
    v.maj_distance = m->majorant_distance;
    // Trackers advance by free flights -log(1-u)/majorant within the majorant's
    // distance: an infinite or NaN majorant, or a distance <= 0, would give steps
    // of length 0 and a tracking loop that never ends.  (The DDA-based maps do
    // not read the majorant; the 7->12 medium integrator reads medium_majorant.)
    const bool reads_majorant = kind != VR_MAP_RAYBOX && kind != VR_MAP_T_DDA && kind != VR_MAP_T_GDT &&
                                kind != VR_MAP_T_GRT && kind != VR_MAP_MEDIUM_DT;
    if (reads_majorant && m->majorant_kind == VR_MAJ_CONST) {
        if (!(m->majorant_value >= 0.0f && m->majorant_value <= 3.402823466e38f))
            return set_err(VR_ERR_ARG, "majorant value must be finite and >= 0");
        if (!(m->majorant_distance > 0.0f)) return set_err(VR_ERR_ARG, "majorant distance must be > 0");
    }
    if (kind == VR_MAP_MEDIUM_DT && !(m->medium_majorant >= 0.0f && m->medium_majorant <= 3.402823466e38f))
        return set_err(VR_ERR_ARG, "medium majorant must be finite and >= 0");
    if (m->majorant_kind == VR_MAJ_GRID) {
        if (!m->majorant_grid) return set_err(VR_ERR_ARG, "majorant grid is NULL");
        for (int a = 0; a < 3; a++)
            if (m->majorant_res[a] < 1 || m->majorant_res[a] > 1023)
                return set_err(VR_ERR_ARG, "majorant resolution must be in [1, 1023]");
        for (int a = 0; a < 3; a++)
            if (m->box_min[a] < m->sigma.bmin[a] - 1e-6f || m->box_max[a] > m->sigma.bmax[a] + 1e-6f)
                return set_err(VR_ERR_UNSUPPORTED, "majorant grid requires the boundary box inside the sigma grid box");
        v.maj_grid = m->majorant_grid;
        v.mrz = m->majorant_res[0]; v.mry = m->majorant_res[1]; v.mrx = m->majorant_res[2];
        v.maj_bmin = load3(m->sigma.bmin);
        v.maj_csize = F3((m->sigma.bmax[0] - m->sigma.bmin[0]) / v.mrx, (m->sigma.bmax[1] - m->sigma.bmin[1]) / v.mry,
                         (m->sigma.bmax[2] - m->sigma.bmin[2]) / v.mrz);
        v.maj_icsize = F3(1.0f / v.maj_csize.x, 1.0f / v.maj_csize.y, 1.0f / v.maj_csize.z);
    } else if (m->majorant_kind != VR_MAJ_CONST) return set_err(VR_ERR_ARG, "unknown majorant kind");
    if (m->transmittance_kind < VR_T_DT || m->transmittance_kind > VR_T_GRT) return set_err(VR_ERR_ARG, "unknown transmittance kind");
    v.t_kind = m->transmittance_kind;
    v.ds_eps = m->ds_epsilon;
    v.max_bounces = m->max_bounces;
    v.medium_filter = m->medium_filter;
    v.gathering_c = load3(m->gathering_c);
    v.medium_majorant = m->medium_majorant;
    if (m->sigma_cells && kind != VR_MAP_RAYBOX) {
        if (reinterpret_cast<uintptr_t>(m->sigma_cells) % 16 != 0)
            return set_err(VR_ERR_ARG, "sigma cell table must be 16-byte aligned");
        if ((int64_t)(v.sigma.nx - 1) * (v.sigma.ny - 1) >= (int64_t(1) << 24))
            return set_err(VR_ERR_ARG, "sigma cell table: an xy plane of the grid holds 2^24 cells or more");
        v.cells = reinterpret_cast<const float4*>(m->sigma_cells);
    }
    if (bw_mode != VR_BW_REFERENCE && bw_mode != VR_BW_CONSISTENT) return set_err(VR_ERR_ARG, "unknown bw mode");
    v.bw_mode = bw_mode;
    if ((st = make_env_view(m, v))) return st;
    return VR_OK;
}

int make_camera_view(const vr_camera* c, CameraView& v) {
    if (!c || !c->poses) return set_err(VR_ERR_ARG, "camera or poses is NULL");
    if (c->width <= 0 || c->height <= 0 || c->num_poses <= 0) return set_err(VR_ERR_ARG, "camera sizes must be positive");
    if (c->sensors && c->num_sensors < 0) return set_err(VR_ERR_ARG, "negative sensor count");
    v.poses = c->poses;
    v.sensors = c->sensors;
    v.num_poses = c->num_poses;
    v.width = c->width;
    v.height = c->height;
    v.jittered = c->jittered;
    v.fov = c->fov;
    v.znear = c->znear;
    v.znear_over_h = c->znear / (float)c->height;
    v.sz = c->znear / std::tan(c->fov * 0.5f);
    return VR_OK;
}

Seeds make_seeds(const vr_launch* L) {
    Seeds s;
    for (int i = 0; i < 4; i++) s.s[i] = L->seeds[i];
    return s;
}

// --------------------------------------------------------- field dispatch --
template <int KIND, int MAJ>
__device__ __forceinline__ void field_fw(Tracer<MAJ>& tr, const float* in, float* out) {
    f3 x = F3(in[0], in[1], in[2]), w = F3(in[3], in[4], in[5]);
    if constexpr (KIND == VR_MAP_RAYBOX) {
        ray_box_intersection(x, w, tr.med().box_min, tr.med().box_max, out[0], out[1]);
    } else if constexpr (KIND == VR_MAP_T_DDA) {
        out[0] = tr.t_dda(x, w);
    } else if constexpr (KIND == VR_MAP_T_DT) {
        out[0] = tr.t_dt(x, w);
    } else if constexpr (KIND == VR_MAP_T_RT) {
        out[0] = tr.t_rt(x, w);
    } else if constexpr (KIND == VR_MAP_T_GDT || KIND == VR_MAP_T_GRT) {
        out[0] = tr.t_grid_tracking(x, w, KIND == VR_MAP_T_GRT);
    } else if constexpr (KIND == VR_MAP_COLLISION_DT) {
        float t, T;
        out[0] = tr.collision_dt(x, w, t, T);
        out[1] = t;
        out[2] = T;
    } else if constexpr (KIND == VR_MAP_SS_RADIANCE) {
        f3 r = tr.ss_radiance(x, w);
        out[0] = r.x; out[1] = r.y; out[2] = r.z;
    } else if constexpr (KIND == VR_MAP_MC_COLLISION) {
        f3 r = tr.mc_collision(x, w);
        out[0] = r.x; out[1] = r.y; out[2] = r.z;
    } else if constexpr (KIND == VR_MAP_NEE_DT || KIND == VR_MAP_NEE_DRT) {
        f3 r = tr.nee(x, w);
        out[0] = r.x; out[1] = r.y; out[2] = r.z;
    } else if constexpr (KIND == VR_MAP_PATH_DT) {
        f3 r = tr.path_dt(x, w);
        out[0] = r.x; out[1] = r.y; out[2] = r.z;
    } else if constexpr (KIND == VR_MAP_MEDIUM_DT) {
        tr.medium_dt(x, w, in[6], out);
    }
}

template <int KIND, int MAJ>
__device__ __forceinline__ void field_bw(Tracer<MAJ>& tr, const float* in, const float* og) {
    f3 x = F3(in[0], in[1], in[2]), w = F3(in[3], in[4], in[5]);
    if constexpr (KIND == VR_MAP_T_DDA) {
        tr.t_dda_bw(x, w, og[0]);
    } else if constexpr (KIND == VR_MAP_T_DT) {
        tr.t_dt_bw(x, w, og[0]);
    } else if constexpr (KIND == VR_MAP_T_RT) {
        tr.t_rt_bw(x, w, og[0]);
    } else if constexpr (KIND == VR_MAP_COLLISION_DT) {
        tr.collision_dt_bw(x, w, tr.med().ds_eps, og[0], og[2]);
    } else if constexpr (KIND == VR_MAP_SS_RADIANCE) {
        tr.ss_radiance_bw(x, w, F3(og[0], og[1], og[2]));
    } else if constexpr (KIND == VR_MAP_MC_COLLISION) {
        tr.mc_collision_bw(x, w, F3(og[0], og[1], og[2]));
    } else if constexpr (KIND == VR_MAP_NEE_DT) {
        tr.nee_ds_bw(x, w, tr.med().ds_eps, F3(og[0], og[1], og[2]));
    } else if constexpr (KIND == VR_MAP_PATH_DT) {
        tr.path_dt_bw(x, w, tr.med().ds_eps, F3(og[0], og[1], og[2]));
    } else if constexpr (KIND == VR_MAP_NEE_DRT) {
        tr.nee_drt_bw(x, w, F3(og[0], og[1], og[2]));
    } else if constexpr (KIND == VR_MAP_MEDIUM_DT) {
        tr.medium_dt_bw(x, w, in[6], F3(og[6], og[7], og[8]), F3(og[9], og[10], og[11]));
    }
}

template <int KIND>
struct Dims;
template <> struct Dims<VR_MAP_RAYBOX> { static constexpr int in = 6, out = 2; };
template <> struct Dims<VR_MAP_T_DDA> { static constexpr int in = 6, out = 1; };
template <> struct Dims<VR_MAP_T_DT> { static constexpr int in = 6, out = 1; };
template <> struct Dims<VR_MAP_T_RT> { static constexpr int in = 6, out = 1; };
template <> struct Dims<VR_MAP_COLLISION_DT> { static constexpr int in = 6, out = 3; };
template <> struct Dims<VR_MAP_SS_RADIANCE> { static constexpr int in = 6, out = 3; };
template <> struct Dims<VR_MAP_MC_COLLISION> { static constexpr int in = 6, out = 3; };
template <> struct Dims<VR_MAP_NEE_DT> { static constexpr int in = 6, out = 3; };
template <> struct Dims<VR_MAP_MEDIUM_DT> { static constexpr int in = 7, out = 12; };
template <> struct Dims<VR_MAP_PATH_DT> { static constexpr int in = 6, out = 3; };
template <> struct Dims<VR_MAP_NEE_DRT> { static constexpr int in = 6, out = 3; };
template <> struct Dims<VR_MAP_T_GDT> { static constexpr int in = 6, out = 1; };
template <> struct Dims<VR_MAP_T_GRT> { static constexpr int in = 6, out = 1; };

// ---------------------------------------------------------------- kernels --
__global__ void __launch_bounds__(kBlock) k_grid_fw(GridView g, const float* __restrict__ x, int64_t n, float* __restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    f3 c = load3(x + 3 * i);
    if (g.C == 1) {
        out[i] = grid_lookup1(g, c);
    } else {
        float v[4];
        grid_lookupC(g, c, v);
        for (int k = 0; k < g.C; k++) out[i * g.C + k] = v[k];
    }
}

__global__ void __launch_bounds__(kBlock) k_grid_bw(GridView g, const float* __restrict__ x, const float* __restrict__ og, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    grid_scatterC(g, load3(x + 3 * i), og + i * g.C);
}

// One workgroup per coarse cell: max over the vertices whose trilinear
// footprint can reach the cell (widened by one vertex per side).  A cell whose
// footprint holds a NaN or infinite vertex gets a NaN bound: every tracker
// floors the majorant (fmaxf(floor, bound) is the floor for a NaN bound), so
// its free flights cross such a cell in one step instead of stalling on an
// infinite majorant.
__global__ void __launch_bounds__(kBlock) k_majorant(GridView g, int rz, int ry, int rx, float* __restrict__ out) {
    int cell = blockIdx.x;
    int cx = cell % rx, cy = (cell / rx) % ry, cz = cell / (rx * ry);
    int n[3] = {g.nx - 1, g.ny - 1, g.nz - 1}, r[3] = {rx, ry, rz}, c[3] = {cx, cy, cz}, lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
        lo[a] = max(0, (int)floor((double)c[a] * n[a] / r[a]) - 1);
        hi[a] = min(n[a], (int)ceil((double)(c[a] + 1) * n[a] / r[a]) + 1);
    }
    int ex = hi[0] - lo[0] + 1, ey = hi[1] - lo[1] + 1, ez = hi[2] - lo[2] + 1;
    int total = ex * ey * ez;
    float m = 0.0f;
    int nonfinite = 0;
    for (int k = threadIdx.x; k < total; k += blockDim.x) {
        int xx = lo[0] + k % ex, yy = lo[1] + (k / ex) % ey, zz = lo[2] + k / (ex * ey);
        const float v = g.data[zz * g.sz + yy * g.sy + xx * g.sx];
        nonfinite |= !(fabsf(v) <= 3.402823466e38f);
        m = fmaxf(m, v);
    }
    __shared__ float red[kBlock];
    red[threadIdx.x] = m;
    nonfinite = __syncthreads_or(nonfinite);
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[cell] = nonfinite ? __builtin_nanf("") : red[0];
}

// Resident waves per SIMD the generic capture and map kernels are compiled for.
// The nested estimators are bound by the latency of dependent gathers; left
// alone the compiler gives the NEE backward ~240 VGPRs (2 waves per SIMD).
// Capping the registers at 5 waves (96 VGPRs, the rest spilled to scratch)
// halves the NEE backward's time on the 512^3 inverse-render step (4 and 6
// waves are close, 8 is slower).
#ifndef VR_CAPTURE_WAVES_PER_EU
#define VR_CAPTURE_WAVES_PER_EU 5
#endif
// The deterministic DDA (FP64 traversal, short launches) keeps its registers.
template <int KIND>
constexpr int capture_waves() {
    return (KIND == VR_MAP_T_DDA || KIND == VR_MAP_RAYBOX || VR_CAPTURE_WAVES_PER_EU < 1) ? 1 : VR_CAPTURE_WAVES_PER_EU;
}
#define VR_CAPTURE_OCCUPANCY __attribute__((amdgpu_waves_per_eu(capture_waves<KIND>())))
// The capture forward of the collision integrator (single-scatter exitance)
// runs fastest at 6 waves per SIMD (k_capture_fw<MC_COLLISION, GRID> of the
// 256^3 inverse render, kernel trace: 708 us at 5 waves, 661 us at 6); the
// other kinds keep the capture cap.
#ifndef VR_CAPTURE_MC_FW_WAVES_PER_EU
#define VR_CAPTURE_MC_FW_WAVES_PER_EU 6
#endif
template <int KIND>
constexpr int capture_fw_waves() {
    return KIND == VR_MAP_MC_COLLISION && VR_CAPTURE_MC_FW_WAVES_PER_EU >= 1 ? VR_CAPTURE_MC_FW_WAVES_PER_EU
                                                                            : capture_waves<KIND>();
}

template <int KIND, int MAJ>
__global__ void __launch_bounds__(kBlock) VR_CAPTURE_OCCUPANCY k_map_fw(MediumView M, const float* __restrict__ in, int64_t start, int64_t cnt,
                                                   float* __restrict__ out, Seeds seeds, int* __restrict__ dbg) {
    int64_t i = start + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= start + cnt) return;
    Tracer<MAJ> tr(first_kernel_arg<MediumView>());
    tr.rng = seed_map(seeds, (uint32_t)i);
    float a[Dims<KIND>::in], o[Dims<KIND>::out];
#pragma unroll
    for (int k = 0; k < Dims<KIND>::in; k++) a[k] = in[i * Dims<KIND>::in + k];
    field_fw<KIND, MAJ>(tr, a, o);
#pragma unroll
    for (int k = 0; k < Dims<KIND>::out; k++) out[i * Dims<KIND>::out + k] = o[k];
    if (dbg) dbg[i] = tr.lookups;
}

template <int KIND, int MAJ>
__global__ void __launch_bounds__(kBlock) VR_CAPTURE_OCCUPANCY k_map_bw(MediumView M, const float* __restrict__ in, const float* __restrict__ og,
                                                   int64_t start, int64_t cnt, Seeds seeds) {
    int64_t i = start + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= start + cnt) return;
    Tracer<MAJ> tr(first_kernel_arg<MediumView>());
    tr.rng = seed_map(seeds, (uint32_t)i);
    tr.stream_ends = true;
    float a[Dims<KIND>::in], o[Dims<KIND>::out];
#pragma unroll
    for (int k = 0; k < Dims<KIND>::in; k++) a[k] = in[i * Dims<KIND>::in + k];
#pragma unroll
    for (int k = 0; k < Dims<KIND>::out; k++) o[k] = og[i * Dims<KIND>::out + k];
    field_bw<KIND, MAJ>(tr, a, o);
}

// Capture kernels: thread j handles sensor start + j % cnt and the samples
// [c*chunk, min(spp, (c+1)*chunk)) with c = j / cnt, so that a pixel's samples
// are spread over several threads (chunks = ceil(spp / chunk)).  With one chunk
// the thread writes its pixel; with more, partial sums are added atomically
// into rows zeroed by k_capture_prepare.
struct CaptureSplit {
    int spp, chunk, chunks;
    int debug_lookups;  // the backward's debug_out counts lookups instead of scatters
    int tiled;          // whole images: 1 = consecutive threads take 8x8 pixel tiles (tiled_pixel),
                        // 2 = the 64 lanes of a wave take pixels hw/64 apart
};

// Sensor of capture position i: the position itself, or (whole images
// without a sensor list, P.tiled) the pixel at position i of its image visited
// in 8x8 tiles, so that a wave's 64 rays come from one tile.  i becomes the
// sensor's own index (its seeds, output row and counter).
VR_DEV void capture_sensor(const CameraView& cam, const CaptureSplit& P, int64_t& i, int& pose, int& row, int& col) {
    if (!P.tiled) {
        sensor_index(cam, i, pose, row, col);
        return;
    }
    const int64_t hw = (int64_t)cam.width * cam.height;
    const int64_t p = i / hw;
    const uint32_t k = (uint32_t)(i - p * hw);
    if (P.tiled == 1) {
        tiled_pixel(k, (uint32_t)cam.width, (uint32_t)cam.height, row, col);
    } else {
        const uint32_t q = (k & 63u) * (uint32_t)(hw >> 6) + (k >> 6);
        row = (int)(q / (uint32_t)cam.width);
        col = (int)(q - (uint32_t)row * (uint32_t)cam.width);
    }
    pose = (int)p;
    i = p * hw + (int64_t)row * cam.width + col;
}


template <int KIND, int MAJ>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(capture_fw_waves<KIND>()))) k_capture_fw(MediumView M, CameraView cam, Seeds seeds, CaptureSplit P, int64_t total,
                                                       int64_t start, int64_t cnt, float* __restrict__ out, int* __restrict__ dbg) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cnt * P.chunks) return;
    const int64_t c = j / cnt;
    int64_t i = start + (j - c * cnt);
    const int s0 = (int)c * P.chunk, s1 = min(P.spp, s0 + P.chunk);
    constexpr int OD = Dims<KIND>::out;
    int pose, row, col;
    capture_sensor(cam, P, i, pose, row, col);
    // a sensor's row and column are used as given (the reference does not check
    // them); a pose index outside the pose buffer yields 0 instead of a fault
    if (pose < 0 || pose >= cam.num_poses) {
        if (P.chunks == 1) {
#pragma unroll
            for (int k = 0; k < OD; k++) out[i * OD + k] = 0.0f;
        }
        return;
    }
    float acc[OD];
#pragma unroll
    for (int k = 0; k < OD; k++) acc[k] = 0.0f;
    Tracer<MAJ> tr(first_kernel_arg<MediumView>());
    for (int s = s0; s < s1; s++) {
        tr.rng = seed_capture_fw(seeds, (uint32_t)((int64_t)s * total + i));
        f3 x, w;
        camera_ray(cam, pose, row, col, tr.rng, x, w);
        float a[6] = {x.x, x.y, x.z, w.x, w.y, w.z}, o[OD];
        field_fw<KIND, MAJ>(tr, a, o);
#pragma unroll
        for (int k = 0; k < OD; k++) acc[k] += o[k];
    }
    const float inv = 1.0f / (float)P.spp;
    if (P.chunks == 1) {
#pragma unroll
        for (int k = 0; k < OD; k++) out[i * OD + k] = P.spp == 1 ? acc[k] : acc[k] * inv;
        if (dbg) dbg[i] = tr.lookups;
    } else {
#pragma unroll
        for (int k = 0; k < OD; k++) atomic_addf(out + i * OD + k, acc[k] * inv);
        if (dbg) atomicAdd(dbg + i, tr.lookups);
    }
}

template <int KIND, int MAJ>
__global__ void __launch_bounds__(kBlock) VR_CAPTURE_OCCUPANCY k_capture_bw(MediumView M, CameraView cam, Seeds seeds, CaptureSplit P, int seed_mode,
                                                       int64_t total, int64_t start, int64_t cnt, const float* __restrict__ og,
                                                       int* __restrict__ dbg) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cnt * P.chunks) return;
    const int64_t c = j / cnt;
    int64_t i = start + (j - c * cnt);
    const int s0 = (int)c * P.chunk, s1 = min(P.spp, s0 + P.chunk);
    constexpr int OD = Dims<KIND>::out;
    int pose, row, col;
    capture_sensor(cam, P, i, pose, row, col);
    if (pose < 0 || pose >= cam.num_poses) return;
    float o[OD];
    bool any = false;
#pragma unroll
    for (int k = 0; k < OD; k++) { o[k] = og[i * OD + k] / (float)P.spp; any |= o[k] != 0.0f; }
    if (!any) {
        if (dbg && P.chunks == 1) dbg[i] = 0;
        return;
    }
    Tracer<MAJ> tr(first_kernel_arg<MediumView>());
    // the reference's fresh-seed backward continues one stream over the samples (one chunk)
    if (seed_mode == VR_SEED_CAPTURE_BW) tr.rng = seed_capture_bw(seeds, (uint32_t)i);
    tr.stream_ends = seed_mode != VR_SEED_CAPTURE_BW;  // per-sample streams
    for (int s = s0; s < s1; s++) {
        if (seed_mode != VR_SEED_CAPTURE_BW) tr.rng = seed_capture_fw(seeds, (uint32_t)((int64_t)s * total + i));
        f3 x, w;
        camera_ray(cam, pose, row, col, tr.rng, x, w);
        float a[6] = {x.x, x.y, x.z, w.x, w.y, w.z};
        field_bw<KIND, MAJ>(tr, a, o);
    }
    if (dbg) {
        const int n = P.debug_lookups ? tr.lookups : tr.scatters;
        if (P.chunks == 1) dbg[i] = n;
        else atomicAdd(dbg + i, n);
    }
}

// The NEE capture backward (defensive replay, DT shadow rays) with the flat
// lane state machine of vr_nee_bw_flat.hpp: same thread -> (sensor, samples)
// split and seeds as k_capture_bw, so the same gradients sample by sample.
// Resident waves per SIMD the flat NEE backward is compiled for: 4 leave 128
// VGPRs per lane and spill a few, yet beat 3 waves (148 against 158 ms for the
// backward of a 256^3 / 512x512 / 64 spp frame).
#ifndef VR_NEE_BW_WAVES_PER_EU
#define VR_NEE_BW_WAVES_PER_EU 4
#endif
template <int MAJ, bool ENVMAP>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(VR_NEE_BW_WAVES_PER_EU)))
k_capture_bw_nee_flat(MediumView M, CameraView cam, Seeds seeds, CaptureSplit P, int seed_mode, int64_t total,
                      int64_t start, int64_t cnt, const float* __restrict__ og, int* __restrict__ dbg) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cnt * P.chunks) return;
    const int64_t c = j / cnt;
    int64_t i = start + (j - c * cnt);
    const int s0 = (int)c * P.chunk, s1 = min(P.spp, s0 + P.chunk);
    int pose, row, col;
    capture_sensor(cam, P, i, pose, row, col);
    if (pose < 0 || pose >= cam.num_poses) return;
    const f3 g = F3(og[3 * i] / (float)P.spp, og[3 * i + 1] / (float)P.spp, og[3 * i + 2] / (float)P.spp);
    if (g.x == 0.0f && g.y == 0.0f && g.z == 0.0f) {
        if (dbg && P.chunks == 1) dbg[i] = 0;
        return;
    }
    __shared__ uint16_t tn_rec[NeeDsBwLane<MAJ, ENVMAP>::kTnRecords * kBlock];
    NeeDsBwLane<MAJ, ENVMAP> L(first_kernel_arg<MediumView>());
    L.tn_rec = tn_rec + threadIdx.x;
    L.eps = M.ds_eps;
    L.dL_dR = g;
    // the reference's fresh-seed backward continues one stream over the samples (one chunk)
    if (seed_mode == VR_SEED_CAPTURE_BW) L.tr.rng = seed_capture_bw(seeds, (uint32_t)i);
    int s = s0;
    bool live = s < s1;
    if (live) {
        if (seed_mode != VR_SEED_CAPTURE_BW) L.tr.rng = seed_capture_fw(seeds, (uint32_t)((int64_t)s * total + i));
        camera_ray(cam, pose, row, col, L.tr.rng, L.x0, L.w0);
        L.begin_pass(1);
    }
    // every sample ends (its paths end); the cap only bounds a thread's loop
    for (int it = 0; live && it < (1 << 24); it++) {
        L.tr.reread();
        if (L.state == kBwDone) {
            if (++s >= s1) break;
            if (seed_mode != VR_SEED_CAPTURE_BW) L.tr.rng = seed_capture_fw(seeds, (uint32_t)((int64_t)s * total + i));
            camera_ray(cam, pose, row, col, L.tr.rng, L.x0, L.w0);
            L.begin_pass(1);
        }
        L.events();
        L.flush_collision_lane();
        for (int k = 0; k < 3; k++) {
            if (L.tracking()) L.step();
            L.flush_step_lane();
            if (__ballot(L.tracking()) == 0ull) break;
        }
    }
    if (dbg) {
        const int n = P.debug_lookups ? L.tr.lookups : L.tr.scatters;
        if (P.chunks == 1) dbg[i] = n;
        else atomicAdd(dbg + i, n);
    }
}

// The same backward with the lanes of a wave pulling work items (a sensor and
// `chunk` consecutive samples of it) from a launch-wide counter, as the
// persistent NEE forward does: a lane whose samples end takes the next item,
// so the wave's lanes keep tracking until the counter drains instead of
// waiting for the lane with the longest paths.  Per-sample streams only (the
// fresh or replayed dispatch seeds); each sample's draws and gradients are
// those of k_capture_bw_nee_flat.  Every argument sits in one structure at the
// start of the kernel-argument segment (the lane reads the medium there).
struct NeeBwArgs {
    MediumView M;
    CameraView cam;
    Seeds seeds;
    NeeSchedule S;
    unsigned long long* counter;
    const float* og;
    int* dbg;
    int debug_lookups;
};
VR_DEV const NeeBwArgs& nee_bw_args() { return *(const NeeBwArgs*)reread(first_kernel_arg<NeeBwArgs>()); }

#ifndef VR_NEE_BW_POOL_WAVES_PER_EU
#define VR_NEE_BW_POOL_WAVES_PER_EU 4
#endif
// Tracking steps per pass of the pooled NEE backward's event loop: more steps
// run the divergent event handlers less often, while lanes whose ray ended wait.
// Measured on the 512^3 NEE step backward (4 views of 512x512, 32 spp): 39.4 ms
// with 2, 35.1 with 3, 33.4 with 4, 32.6-33.0 with 6, 33.6 with 8, 35.6 with 12.
constexpr int kNeeBwPoolSteps = 6;
template <int MAJ, bool ENVMAP>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(VR_NEE_BW_POOL_WAVES_PER_EU)))
k_capture_bw_nee_pool(NeeBwArgs args) {
    __shared__ uint16_t tn_rec[NeeDsBwLane<MAJ, ENVMAP>::kTnRecords * kBlock];
    // per wave: the sigma scatter records (kScatterRecords x 8 floats) and one
    // collision's albedo records (64 x 8 floats)
    constexpr int kSRec = NeeDsBwLane<MAJ, ENVMAP>::kScatterRecords * 8;
    __shared__ float scat_records[kBlock / 64 * (kSRec + 512)];
    float* srec = scat_records + (threadIdx.x / 64) * (kSRec + 512);
    float* arec = srec + kSRec;
    int nrec = 0;  // wave-uniform count of buffered sigma records
    NeeDsBwLane<MAJ, ENVMAP> L(first_kernel_arg<MediumView>());
    L.tn_rec = tn_rec + threadIdx.x;
    L.eps = args.M.ds_eps;
    // A lane without work sits in kBwDone with s == s1 (a per-lane "has work"
    // flag guarding the handlers cost ~150 spilled VGPRs).
    L.state = kBwDone;
    bool done = false;
    ItemPool pool;
    uint32_t sensor = 0;
    int s = 0, s1 = 0, pose = 0, row = 0, col = 0, count0 = 0;
    // every sample ends (its paths end); the cap only bounds a wave's loop
    for (uint32_t it = 0; it < (1u << 28); it++) {
        L.tr.reread();
        // end of a sample: the next sample of the item, or the item ends
        bool fresh = false, need = false;
        if (L.state == kBwDone) {
            if (s < s1) {
                if (++s < s1) {
                    fresh = true;
                } else {
                    const NeeBwArgs& A = nee_bw_args();
                    if (A.dbg) atomicAdd(A.dbg + sensor, (A.debug_lookups ? L.tr.lookups : L.tr.scatters) - count0);
                }
            }
            need = !fresh && !done;
        }
        // lanes that need an item take one from the wave's pool of consecutive
        // item indices; an empty pool is refilled with kPool items by one atomic
        const unsigned long long mask = __ballot(need);
        if (mask) {
            const NeeBwArgs& A = nee_bw_args();
            uint32_t item = 0;
            if (pool.take(mask, need, A.counter, A.S.cnt * (uint32_t)A.S.chunks, item, done)) {
                int s0;
                s = s1 = 0;  // no samples unless the item has a gradient
                if (nee_item(A.S, A.cam, 0, item, sensor, s0, s1, pose, row, col)) {
                    const float inv = (float)A.S.spp;
                    const float* o = A.og + 3 * (size_t)sensor;
                    const f3 g = F3(o[0] / inv, o[1] / inv, o[2] / inv);
                    if (g.x != 0.0f || g.y != 0.0f || g.z != 0.0f) {
                        L.dL_dR = g;
                        s = s0;
                        count0 = A.debug_lookups ? L.tr.lookups : L.tr.scatters;
                        fresh = true;
                    }
                }
                if (!fresh) s = s1;
            }
        }
        if (fresh) {
            const NeeBwArgs& A = nee_bw_args();
            L.tr.rng = seed_capture_fw(A.seeds, (uint32_t)s * A.S.total + sensor);
            camera_ray(A.cam, pose, row, col, L.tr.rng, L.x0, L.w0);
            L.begin_pass(1);
        }
        if (__ballot(L.state != kBwDone || !done) == 0ull) break;  // the wave drained the counter
        L.events();
        nrec = L.post_collision(srec, arec, nrec);
        for (int k = 0; k < kNeeBwPoolSteps; k++) {
            if (L.tracking()) L.step();
            nrec = L.post_step(srec, nrec);
            if (__ballot(L.tracking()) == 0ull) break;
        }
    }
    L.flush_sigma(srec, nrec);
}

// The collision-integrator backward with single-scatter exitance and RT
// shadow rays (k_capture_bw of VR_MAP_MC_COLLISION, consistent form, per-sample
// streams) as the lane state machine of vr_ss_bw_flat.hpp, the lanes of a
// wave pulling one-sample work items from a launch-wide counter as
// k_capture_bw_nee_pool does.  Each sample's draws and gradients are those of
// k_capture_bw.
#ifndef VR_SS_BW_POOL_WAVES_PER_EU
#define VR_SS_BW_POOL_WAVES_PER_EU 3
#endif
template <int MAJ>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(VR_SS_BW_POOL_WAVES_PER_EU)))
k_capture_bw_ss_pool(NeeBwArgs args) {
    constexpr int kSRec = SsBwLane<MAJ>::kScatterRecords * 8;
    __shared__ float scat_records[kBlock / 64 * kSRec];
    float* srec = scat_records + (threadIdx.x / 64) * kSRec;
    int nrec = 0;  // wave-uniform count of buffered sigma records
    SsBwLane<MAJ> L(first_kernel_arg<MediumView>());
    L.eps = args.M.ds_eps;
    L.state = kSsDone;  // a lane without work sits in kSsDone with s == s1
    bool done = false;
    ItemPool pool;
    uint32_t sensor = 0;
    int s = 0, s1 = 0, pose = 0, row = 0, col = 0, count0 = 0;
    // every sample ends (its walks end); the cap only bounds a wave's loop
    for (uint32_t it = 0; it < (1u << 28); it++) {
        L.tr.reread();
        bool fresh = false, need = false;
        if (L.state == kSsDone) {
            if (s < s1) {
                if (++s < s1) {
                    fresh = true;
                } else {
                    const NeeBwArgs& A = nee_bw_args();
                    if (A.dbg) atomicAdd(A.dbg + sensor, (A.debug_lookups ? L.tr.lookups : L.tr.scatters) - count0);
                }
            }
            need = !fresh && !done;
        }
        const unsigned long long mask = __ballot(need);
        if (mask) {
            const NeeBwArgs& A = nee_bw_args();
            uint32_t item = 0;
            if (pool.take(mask, need, A.counter, A.S.cnt * (uint32_t)A.S.chunks, item, done)) {
                int s0;
                s = s1 = 0;
                if (nee_item(A.S, A.cam, 0, item, sensor, s0, s1, pose, row, col)) {
                    const float inv = (float)A.S.spp;
                    const float* o = A.og + 3 * (size_t)sensor;
                    const f3 g = F3(o[0] / inv, o[1] / inv, o[2] / inv);
                    if (g.x != 0.0f || g.y != 0.0f || g.z != 0.0f) {
                        L.dL_dR = g;
                        s = s0;
                        count0 = A.debug_lookups ? L.tr.lookups : L.tr.scatters;
                        fresh = true;
                    }
                }
                if (!fresh) s = s1;
            }
        }
        if (fresh) {
            const NeeBwArgs& A = nee_bw_args();
            L.tr.rng = seed_capture_fw(A.seeds, (uint32_t)s * A.S.total + sensor);
            camera_ray(A.cam, pose, row, col, L.tr.rng, L.x0, L.w0);
            L.begin_sample();
        }
        if (__ballot(L.state != kSsDone || !done) == 0ull) break;  // the wave drained the counter
        L.events();
        for (int k = 0; k < 3; k++) {
            if (L.tracking()) L.step();
            nrec = L.post_step(srec, nrec);
            if (__ballot(L.tracking()) == 0ull) break;
        }
    }
    L.flush_sigma(srec, nrec);
}

// Zeroes the pooled backward's work counter and the debug rows [start, start+cnt).
__global__ void k_nee_bw_prepare(unsigned long long* __restrict__ counter, int64_t start, int64_t cnt, int* __restrict__ dbg) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    if (i == 0) counter[0] = 0ull;
    if (!dbg) return;
    for (; i < cnt; i += stride) dbg[start + i] = 0;
}

// Cell table of a single-channel grid: thread c writes cell c's 8 corners.
__global__ void __launch_bounds__(kBlock) k_pack_cells(GridView g, float4* __restrict__ out, int64_t cells) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cells) return;
    const int cx = g.nx - 1, cy = g.ny - 1;
    const int px = (int)(c % cx), py = (int)((c / cx) % cy), pz = (int)(c / ((int64_t)cx * cy));
    const float* p = g.data + (pz * g.sz + py * g.sy + px * g.sx);
    out[2 * c] = make_float4(p[0], p[g.sx], p[g.sy], p[g.sx + g.sy]);
    out[2 * c + 1] = make_float4(p[g.sz], p[g.sz + g.sx], p[g.sz + g.sy], p[g.sz + g.sy + g.sx]);
}

// Zeroes p[0, n) (a kernel rather than hipMemsetAsync, so that callers may
// capture the work in a graph).
__global__ void k_zero_u32(uint32_t* __restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// Zeroes rows [start, start+cnt) of a capture output (OD floats each) and of the debug counters.
__global__ void k_capture_prepare(float* __restrict__ out, int od, int64_t start, int64_t cnt, int* __restrict__ dbg) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < cnt; i += stride) {
        if (out)
            for (int k = 0; k < od; k++) out[(start + i) * od + k] = 0.0f;
        if (dbg) dbg[start + i] = 0;
    }
}

// torch.optim.Adam (no weight decay, no amsgrad), optionally clearing the gradient.
__global__ void __launch_bounds__(kBlock) k_adam(float* __restrict__ p, float* __restrict__ g, float* __restrict__ m,
                                                 float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps,
                                                 float bc1, float bc2_sqrt, float lo, float hi, int zero_grad) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        float gi = g[i];
        float mi = b1 * m[i] + (1.0f - b1) * gi;
        float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        float denom = sqrtf(vi) / bc2_sqrt + eps;
        p[i] = fminf(fmaxf(p[i] - (lr / bc1) * mi / denom, lo), hi);
        if (zero_grad) g[i] = 0.0f;
    }
}

// --------------------------------------------------- persistent NEE queue --
// The persistent NEE kernel's queue counters (kMaxQueues 64-bit words, one per
// 128-byte line) live in the launch's workspace: the caller's buffer
// (vr_launch.workspace, used in stream order) or one allocated on the stream
// for the call, so captures on different streams never share counters.
constexpr int64_t kCounterBytes = (int64_t)kMaxQueues * kQueueStride * sizeof(unsigned long long);
static_assert(kCounterBytes == VR_CAPTURE_WORKSPACE_BYTES, "workspace size of the C ABI");

// Work counters of one capture: the caller's workspace, else a stream-ordered
// allocation released (stream-ordered) by release().
struct Counters {
    unsigned long long* ptr = nullptr;
    hipStream_t owned_on = nullptr;
    bool owned = false;

    int acquire(const vr_launch* L, hipStream_t st) {
        if (L->workspace) {
            if (L->workspace_bytes < kCounterBytes)
                return set_err(VR_ERR_ARG, "vr_capture_fw: workspace smaller than VR_CAPTURE_WORKSPACE_BYTES");
            if (reinterpret_cast<uintptr_t>(L->workspace) % 256 != 0)
                return set_err(VR_ERR_ARG, "vr_capture_fw: workspace must be 256-byte aligned");
            ptr = static_cast<unsigned long long*>(L->workspace);
            return VR_OK;
        }
        void* p = nullptr;
        int rc = check_hip(hipMallocAsync(&p, (size_t)kCounterBytes, st), "hipMallocAsync(work counters)");
        if (rc) return rc;
        ptr = static_cast<unsigned long long*>(p);
        owned = true;
        owned_on = st;
        return VR_OK;
    }
    int release() {
        if (!owned) return VR_OK;
        owned = false;
        return check_hip(hipFreeAsync(ptr, owned_on), "hipFreeAsync(work counters)");
    }
};

// Tuning knob read from the environment (unset: the built-in default).
int env_int(const char* name, int fallback) {
    const char* v = std::getenv(name);
    return v && *v ? std::atoi(v) : fallback;
}

template <typename K>
int persistent_blocks(K kernel, int* blocks) {
    int dev = 0, cus = 0, per_cu = 0;
    int st = check_hip(hipGetDevice(&dev), "hipGetDevice");
    if (!st) st = check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "CU count");
    if (!st) st = check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0), "occupancy");
    if (st) return st;
    // resident blocks per CU from the occupancy query (VR_NEE_BLOCKS_PER_CU overrides)
    *blocks = cus * std::max(1, env_int("VR_NEE_BLOCKS_PER_CU", per_cu));
    return VR_OK;
}

// Cached-render extras of a persistent launch (CACHE): the query buffer.
struct NeeCache {
    float* queries = nullptr;
    int64_t max_queries = 0;
    uint32_t* qcount = nullptr;
    int k = 0;
};

template <int MAJ, int TKIND, bool ENVMAP, bool LEAN = false, bool CACHE = false, bool CMP = false>
int launch_nee_persistent(const MediumView& M, const CameraView& C, Seeds sd, int spp, int64_t total, int64_t start,
                          int64_t cnt, float* out, int* dbg, unsigned long long* counter, hipStream_t st,
                          const NeeCache& cache = NeeCache()) {
    if (total > (int64_t)UINT32_MAX) return set_err(VR_ERR_UNSUPPORTED, "capture has more than 2^32 sensors");
    if (!counter) return set_err(VR_ERR_ARG, "persistent NEE capture without work counters");
    int rc = VR_OK;
    int blocks = 0;
    if ((rc = persistent_blocks(k_nee_persistent<MAJ, TKIND, ENVMAP, LEAN, CACHE, CMP>, &blocks))) return rc;
    const int64_t lanes = (int64_t)blocks * kBlock;
    const int64_t hw = (int64_t)C.width * C.height;
    // launches of at most 2^30 sensors (whole images when possible) keep item indices in 32 bits
    int64_t batch = int64_t(1) << 30;
    if (!C.sensors && hw <= batch) batch = batch / hw * hw;
    for (int64_t b0 = start; b0 < start + cnt; b0 += batch) {
        const int64_t n = std::min(batch, start + cnt - b0);
        // samples per work item: ~64 items per lane, 1 to 8 samples.  Short items
        // keep a wave's lanes on neighbouring pixels (a refilled pool hands out
        // consecutive sensors) and even out the drain; measured per frame: 512^2 x
        // 64 spp 4.90 ms with 1 sample per item, 4.98 with 2, 5.98 with 10; the
        // 1024^2 x 256 spp render on 512^3 96.5 ms with 8, 99.8 with 2, 126 with 42
        int64_t chunk = (n * (int64_t)spp) / (64 * lanes);
        chunk = std::max<int64_t>(1, std::min<int64_t>(chunk, 8));
        chunk = env_int("VR_NEE_CHUNK", (int)chunk);
        chunk = chunk < 1 ? 1 : (chunk > spp ? spp : chunk);
        while (n * ((spp + chunk - 1) / chunk) > (int64_t)UINT32_MAX - kPool) chunk *= 2;
        NeeSchedule S;
        S.total = (uint32_t)total;
        S.start = (uint32_t)b0;
        S.cnt = (uint32_t)n;
        // one work queue per XCD: each XCD's L2 serves a contiguous eighth of the
        // images (headline frame 4.90 ms with one queue, 4.79 with 2, 4.72 with 4,
        // 4.54 with 8; the 1024^2 x 256 spp render 96.3 -> 92.0 ms)
        S.nq = std::max(1, std::min(kMaxQueues, env_int("VR_NEE_QUEUES", kMaxQueues)));
        S.qpix = (uint32_t)((n + S.nq - 1) / S.nq);
        S.spp = spp;
        S.chunk = (int)std::min<int64_t>(chunk, spp);
        S.chunks = (spp + S.chunk - 1) / S.chunk;
        S.tiled = !C.sensors && b0 % hw == 0 && n % hw == 0;
        // tracking steps per pass of the event loop: 3 for the lean form (headline
        // frame 4.42 ms, 4.55 with 4, 4.86 with 6; the 512^3 / 1024^2 x 256 spp
        // render 87.5 ms, 109 ms with 6), 6 for the general forms (512^3 +
        // 128^3-albedo frame at 4 x 512^2 x 32 spp: 17.2 ms with 3, 16.4 with 4,
        // 16.1 with 6) and the radiance-cache form (its 1024^2 render 9.8 -> 7.3 ms)
        S.ksteps = std::max(1, env_int("VR_NEE_STEPS", LEAN && !CACHE ? 3 : 6));
        // more steps while at least VR_NEE_TRACK_MIN lanes still track (default: never)
        S.kmax = std::min(256, std::max(S.ksteps, env_int("VR_NEE_STEPS_MAX", S.ksteps)));
        S.track_min = std::max(0, std::min(65, env_int("VR_NEE_TRACK_MIN", 65)));
        // measured on the headline frame (256^3, 512x512, 64 spp): 1 -> 6.86 ms,
        // 4 -> 6.52, 8 -> 6.55, 12 -> 6.51, 16 -> 6.68, 32 -> 7.48
        S.min_begin = std::max(1, std::min(64, env_int("VR_NEE_MIN_BEGIN", 8)));
        unsigned prep = (unsigned)std::min<int64_t>((n + kBlock - 1) / kBlock, 4096);
        hipLaunchKernelGGL(k_nee_prepare, dim3(prep > 0 ? prep : 1), dim3(kBlock), 0, st, out, b0, n, dbg, counter);
        if ((rc = check_hip(hipGetLastError(), "vr_capture_fw prepare"))) return rc;
        const NeeArgs args{M, C, sd, S, counter, out, dbg, cache.queries, cache.max_queries, cache.qcount, cache.k};
        hipLaunchKernelGGL((k_nee_persistent<MAJ, TKIND, ENVMAP, LEAN, CACHE, CMP>), dim3(blocks), dim3(kBlock), 0, st, args);
        if ((rc = check_hip(hipGetLastError(), "vr_capture_fw launch"))) return rc;
    }
    return VR_OK;
}

// ------------------------------------------------------------- dispatch --
template <int MAJ>
int launch_map_fw_maj(int kind, const MediumView& M, const float* in, int64_t start, int64_t cnt, float* out, Seeds sd, int* dbg,
                      hipStream_t st) {
    dim3 g(grid_for(cnt)), b(kBlock);
    switch (kind) {
#define VR_CASE(K) case K: hipLaunchKernelGGL((k_map_fw<K, MAJ>), g, b, 0, st, M, in, start, cnt, out, sd, dbg); break;
        VR_CASE(VR_MAP_RAYBOX) VR_CASE(VR_MAP_T_DDA) VR_CASE(VR_MAP_T_DT) VR_CASE(VR_MAP_T_RT) VR_CASE(VR_MAP_COLLISION_DT)
        VR_CASE(VR_MAP_SS_RADIANCE) VR_CASE(VR_MAP_MC_COLLISION) VR_CASE(VR_MAP_NEE_DT) VR_CASE(VR_MAP_MEDIUM_DT)
        VR_CASE(VR_MAP_PATH_DT) VR_CASE(VR_MAP_NEE_DRT) VR_CASE(VR_MAP_T_GDT) VR_CASE(VR_MAP_T_GRT)
#undef VR_CASE
    default: return set_err(VR_ERR_UNSUPPORTED, "unknown map kind");
    }
    return check_hip(hipGetLastError(), "vr_map_fw launch");
}

template <int MAJ>
int launch_map_bw_maj(int kind, const MediumView& M, const float* in, const float* og, int64_t start, int64_t cnt, Seeds sd,
                      hipStream_t st) {
    dim3 g(grid_for(cnt)), b(kBlock);
    switch (kind) {
#define VR_CASE(K) case K: hipLaunchKernelGGL((k_map_bw<K, MAJ>), g, b, 0, st, M, in, og, start, cnt, sd); break;
        VR_CASE(VR_MAP_T_DDA) VR_CASE(VR_MAP_T_DT) VR_CASE(VR_MAP_T_RT) VR_CASE(VR_MAP_COLLISION_DT)
        VR_CASE(VR_MAP_SS_RADIANCE) VR_CASE(VR_MAP_MC_COLLISION) VR_CASE(VR_MAP_NEE_DT) VR_CASE(VR_MAP_MEDIUM_DT)
        VR_CASE(VR_MAP_PATH_DT) VR_CASE(VR_MAP_NEE_DRT)
#undef VR_CASE
    case VR_MAP_RAYBOX: return VR_OK;  // no parameters
    case VR_MAP_T_GDT: case VR_MAP_T_GRT: return VR_OK;  // no backward (maps/transmittance_gdt.h, _grt.h)
    default: return set_err(VR_ERR_UNSUPPORTED, "unknown map kind");
    }
    return check_hip(hipGetLastError(), "vr_map_bw launch");
}

// Samples per capture thread: a pixel's spp are spread over up to 8 threads
// (VR_CAPTURE_CHUNK overrides); the fresh-seed backward keeps one chunk.
CaptureSplit capture_split(int spp, bool one_chunk) {
    CaptureSplit P;
    P.spp = spp;
    int chunk = spp <= 4 ? spp : (spp + 7) / 8;
    chunk = env_int("VR_CAPTURE_CHUNK", chunk);
    P.chunk = one_chunk ? spp : std::max(1, std::min(spp, chunk));
    P.chunks = (spp + P.chunk - 1) / P.chunk;
    P.debug_lookups = 0;
    P.tiled = 0;
    return P;
}

// Pixel visiting order of whole images without a sensor list: 0 row-major, 1
// 8x8 tiles, 2 waves spread over the image (environment `var` overrides
// `fallback`).  Measured on the bench workloads: row-major is fastest for the
// medium captures (tiles: NEE backward 152 -> 214 ms, spread: -> 180 ms);
// spread is fastest for scene captures (forward 28.4 -> 25 ms, backward
// 124 -> 107 ms).
void capture_tiling(CaptureSplit& P, const CameraView& C, int64_t start, int64_t cnt, const char* var = "VR_CAPTURE_TILED",
                    int fallback = 0) {
    const int64_t hw = (int64_t)C.width * C.height;
    const int mode = env_int(var, fallback);
    const bool whole = !C.sensors && hw > 0 && hw <= (int64_t)UINT32_MAX && start % hw == 0 && cnt % hw == 0;
    P.tiled = whole && (mode == 1 || (mode == 2 && hw % 64 == 0)) ? mode : 0;
}

int capture_prepare(const CaptureSplit& P, float* out, int od, int64_t start, int64_t cnt, int* dbg, hipStream_t st) {
    if (P.chunks == 1 || (!out && !dbg)) return VR_OK;
    unsigned prep = (unsigned)std::min<int64_t>((cnt + kBlock - 1) / kBlock, 4096);
    hipLaunchKernelGGL(k_capture_prepare, dim3(prep > 0 ? prep : 1), dim3(kBlock), 0, st, out, od, start, cnt, dbg);
    return check_hip(hipGetLastError(), "capture prepare");
}

// The NEE forward (also DRTPathIntegrator's) with a DT or RT shadow estimator
// runs as the persistent lane state machine, which needs work counters.
bool capture_is_persistent(int kind, const MediumView& M) {
    return (kind == VR_MAP_NEE_DT || kind == VR_MAP_NEE_DRT) && (M.t_kind == VR_T_DT || M.t_kind == VR_T_RT);
}

template <int MAJ>
int launch_capture_fw_maj(int kind, const MediumView& M, const CameraView& C, Seeds sd, int spp, int64_t total, int64_t start,
                          int64_t cnt, float* out, int* dbg, unsigned long long* counter, hipStream_t st) {
    if (capture_is_persistent(kind, M)) {
        const bool envmap = M.env_kind != VR_ENV_CONST || M.env_sampler != VR_ENVS_UNIFORM;
        // cell table + constant albedo and phase asymmetry (DT shadow rays): the lean specialisation
        const bool lean = M.cells && !M.albedo_grid && !M.phase_g_grid && env_int("VR_NEE_LEAN", 1) != 0;
#define VR_NEE_ARGS M, C, sd, spp, total, start, cnt, out, dbg, counter, st
        // VR_MAJ_LDS=1: a coarse majorant grid that fits LDS is staged there
        // (constant environments).  Off by default: the headline frame takes
        // 5.25 ms with it against 4.90 ms reading the grid (an L1/L2-resident
        // 16 KiB grid; the LDS reads share the wait counter of the handlers'
        // scalar argument loads).
        if constexpr (MAJ == VR_MAJ_GRID) {
            if (!envmap && M.mrx * M.mry * M.mrz <= kMajLdsCells && env_int("VR_MAJ_LDS", 0) != 0) {
                if (M.t_kind == VR_T_DT && lean) return launch_nee_persistent<kMajGridLds, VR_T_DT, false, true>(VR_NEE_ARGS);
                if (M.t_kind == VR_T_DT) return launch_nee_persistent<kMajGridLds, VR_T_DT, false>(VR_NEE_ARGS);
                return launch_nee_persistent<kMajGridLds, VR_T_RT, false>(VR_NEE_ARGS);
            }
        }
        // VR_NEE_COMPACT=1: light and phase samples computed block-wide (see
        // k_nee_persistent's CMP form).  Off by default: the headline frame
        // takes 4.49-4.50 ms with it against 4.44-4.47 ms without (the handlers'
        // saved VALU issues are spent on the exchange and the block's waves
        // waiting for each other at its barriers).
        if (M.t_kind == VR_T_DT && lean && !envmap && env_int("VR_NEE_COMPACT", 0) != 0)
            return launch_nee_persistent<MAJ, VR_T_DT, false, true, false, true>(VR_NEE_ARGS);
        if (M.t_kind == VR_T_DT && lean)
            return envmap ? launch_nee_persistent<MAJ, VR_T_DT, true, true>(VR_NEE_ARGS)
                          : launch_nee_persistent<MAJ, VR_T_DT, false, true>(VR_NEE_ARGS);
        if (M.t_kind == VR_T_DT)
            return envmap ? launch_nee_persistent<MAJ, VR_T_DT, true>(VR_NEE_ARGS)
                          : launch_nee_persistent<MAJ, VR_T_DT, false>(VR_NEE_ARGS);
        return envmap ? launch_nee_persistent<MAJ, VR_T_RT, true>(VR_NEE_ARGS)
                      : launch_nee_persistent<MAJ, VR_T_RT, false>(VR_NEE_ARGS);
#undef VR_NEE_ARGS
    }
    int in_dim = 0, od = 0;
    map_dims(kind, in_dim, od);
    CaptureSplit P = capture_split(spp, false);
    capture_tiling(P, C, start, cnt);
    int rc = capture_prepare(P, out, od, start, cnt, dbg, st);
    if (rc) return rc;
    dim3 g(grid_for(cnt * P.chunks)), b(kBlock);
    switch (kind) {
#define VR_CASE(K) case K: hipLaunchKernelGGL((k_capture_fw<K, MAJ>), g, b, 0, st, M, C, sd, P, total, start, cnt, out, dbg); break;
        VR_CASE(VR_MAP_RAYBOX) VR_CASE(VR_MAP_T_DDA) VR_CASE(VR_MAP_T_DT) VR_CASE(VR_MAP_T_RT) VR_CASE(VR_MAP_COLLISION_DT)
        VR_CASE(VR_MAP_SS_RADIANCE) VR_CASE(VR_MAP_MC_COLLISION) VR_CASE(VR_MAP_NEE_DT) VR_CASE(VR_MAP_PATH_DT)
        VR_CASE(VR_MAP_NEE_DRT) VR_CASE(VR_MAP_T_GDT) VR_CASE(VR_MAP_T_GRT)
#undef VR_CASE
    default: return set_err(VR_ERR_UNSUPPORTED, "map kind cannot be captured (needs 6 inputs)");
    }
    return check_hip(hipGetLastError(), "vr_capture_fw launch");
}

// The NEE backward with DT shadow rays and per-sample streams runs as the
// pooled lane state machine (k_capture_bw_nee_pool), which needs a work counter.
bool capture_bw_is_pooled(int kind, const MediumView& M, int seed_mode) {
    return kind == VR_MAP_NEE_DT && M.t_kind == VR_T_DT && seed_mode != VR_SEED_CAPTURE_BW &&
           env_int("VR_NEE_BW_FLAT", 1) != 0 && env_int("VR_NEE_BW_POOL", 1) != 0;
}

// With VR_SS_BW_POOL=1 the collision-integrator backward (single-scatter
// exitance) with RT shadow rays, the consistent estimators, per-sample streams
// and no albedo gradient runs as the pooled lane state machine
// k_capture_bw_ss_pool.  Off by default: its walks are a few steps long at the
// densities of the 256^3 inverse render, where the per-event work of the state
// machine costs more than the nested kernel loses to divergence (256^3, 8 views
// of 256x256 at 16 spp, nested -> pooled: 1.43 -> 1.95-2.03 ms at a constant
// density of 5, 1.95 -> 2.31 ms for an fBm density scaled to 40 (2.0 ms with 2
// samples per work item), 3.74 -> 2.97 ms scaled to 120).
bool capture_bw_is_ss_pooled(int kind, const MediumView& M, int seed_mode) {
    return kind == VR_MAP_MC_COLLISION && M.t_kind == VR_T_RT && M.bw_mode == VR_BW_CONSISTENT &&
           seed_mode != VR_SEED_CAPTURE_BW && !(M.albedo_grid && M.albedo.grad != nullptr) &&
           env_int("VR_SS_BW_POOL", 0) != 0;
}

// Launches a pooled backward kernel (k_capture_bw_nee_pool or k_capture_bw_ss_pool).
int launch_bw_pool(void (*kern)(NeeBwArgs), const MediumView& M, const CameraView& C, Seeds sd, int spp, int debug_kind,
                   int64_t total, int64_t start, int64_t cnt, const float* og, int* dbg, unsigned long long* counter,
                   hipStream_t st) {
    if (total > (int64_t)UINT32_MAX) return set_err(VR_ERR_UNSUPPORTED, "capture has more than 2^32 sensors");
    int rc = VR_OK;
    int blocks = 0;
    if ((rc = persistent_blocks(kern, &blocks))) return rc;
    const int64_t hw = (int64_t)C.width * C.height;
    CaptureSplit P = capture_split(spp, false);
    capture_tiling(P, C, start, cnt);
    // launches of at most 2^30 sensors (whole images when possible) keep item indices in 32 bits
    int64_t batch = int64_t(1) << 30;
    if (!C.sensors && hw <= batch) batch = batch / hw * hw;
    for (int64_t b0 = start; b0 < start + cnt; b0 += batch) {
        const int64_t n = std::min(batch, start + cnt - b0);
        // one sample per work item (measured on the 512^3 NEE step backward:
        // 50.6 ms with 1, 50.9 with 2, 52.7 with 4, 56.7 with 8 samples per item)
        int64_t chunk = std::max(1, std::min(spp, env_int("VR_NEE_BW_CHUNK", 1)));
        while (n * ((spp + chunk - 1) / chunk) > (int64_t)UINT32_MAX - kPool) chunk *= 2;
        NeeBwArgs A;
        A.M = M;
        A.cam = C;
        A.seeds = sd;
        NeeSchedule& S = A.S;
        S.total = (uint32_t)total;
        S.start = (uint32_t)b0;
        S.cnt = (uint32_t)n;
        S.qpix = (uint32_t)n;
        S.nq = 1;
        S.spp = spp;
        S.chunk = (int)chunk;
        S.chunks = (int)((spp + chunk - 1) / chunk);
        S.tiled = P.tiled == 1 && n % hw == 0 && b0 % hw == 0 ? 1 : 0;
        S.ksteps = 3;
        S.kmax = 3;
        S.track_min = 0;
        S.min_begin = 1;
        A.counter = counter;
        A.og = og;
        A.dbg = dbg;
        A.debug_lookups = debug_kind == 1;
        const unsigned prep = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + kBlock - 1) / kBlock, 4096));
        hipLaunchKernelGGL(k_nee_bw_prepare, dim3(prep), dim3(kBlock), 0, st, counter, b0, n, dbg);
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(kBlock), 0, st, A);
        if ((rc = check_hip(hipGetLastError(), "vr_capture_bw launch"))) return rc;
    }
    return VR_OK;
}

template <int MAJ>
int launch_capture_bw_maj(int kind, const MediumView& M, const CameraView& C, Seeds sd, int spp, int seed_mode, int debug_kind,
                          int64_t total,
                          int64_t start, int64_t cnt, const float* og, int* dbg, unsigned long long* counter, hipStream_t st) {
    if (counter && capture_bw_is_pooled(kind, M, seed_mode)) {
        if (M.env_kind != VR_ENV_CONST || M.env_sampler != VR_ENVS_UNIFORM)
            return launch_bw_pool(k_capture_bw_nee_pool<MAJ, true>, M, C, sd, spp, debug_kind, total, start, cnt, og, dbg,
                                  counter, st);
        return launch_bw_pool(k_capture_bw_nee_pool<MAJ, false>, M, C, sd, spp, debug_kind, total, start, cnt, og, dbg,
                              counter, st);
    }
    if (counter && capture_bw_is_ss_pooled(kind, M, seed_mode))
        return launch_bw_pool(k_capture_bw_ss_pool<MAJ>, M, C, sd, spp, debug_kind, total, start, cnt, og, dbg, counter, st);
    CaptureSplit P = capture_split(spp, seed_mode == VR_SEED_CAPTURE_BW);
    capture_tiling(P, C, start, cnt);
    P.debug_lookups = debug_kind == 1;
    int rc = capture_prepare(P, nullptr, 0, start, cnt, dbg, st);
    if (rc) return rc;
    dim3 g(grid_for(cnt * P.chunks)), b(kBlock);
    if (kind == VR_MAP_NEE_DT && M.t_kind == VR_T_DT && env_int("VR_NEE_BW_FLAT", 1) != 0) {
        // constant environments compile the uniform + sun sampler only (as the forward)
        if (M.env_kind != VR_ENV_CONST || M.env_sampler != VR_ENVS_UNIFORM)
            hipLaunchKernelGGL((k_capture_bw_nee_flat<MAJ, true>), g, b, 0, st, M, C, sd, P, seed_mode, total, start, cnt, og,
                               dbg);
        else
            hipLaunchKernelGGL((k_capture_bw_nee_flat<MAJ, false>), g, b, 0, st, M, C, sd, P, seed_mode, total, start, cnt, og,
                               dbg);
        return check_hip(hipGetLastError(), "vr_capture_bw launch");
    }
    switch (kind) {
#define VR_CASE(K) case K: hipLaunchKernelGGL((k_capture_bw<K, MAJ>), g, b, 0, st, M, C, sd, P, seed_mode, total, start, cnt, og, dbg); break;
        VR_CASE(VR_MAP_T_DDA) VR_CASE(VR_MAP_T_DT) VR_CASE(VR_MAP_T_RT) VR_CASE(VR_MAP_COLLISION_DT)
        VR_CASE(VR_MAP_SS_RADIANCE) VR_CASE(VR_MAP_MC_COLLISION) VR_CASE(VR_MAP_NEE_DT) VR_CASE(VR_MAP_PATH_DT)
        VR_CASE(VR_MAP_NEE_DRT)
#undef VR_CASE
    case VR_MAP_RAYBOX: case VR_MAP_T_GDT: case VR_MAP_T_GRT: return VR_OK;
    default: return set_err(VR_ERR_UNSUPPORTED, "map kind cannot be captured (needs 6 inputs)");
    }
    return check_hip(hipGetLastError(), "vr_capture_bw launch");
}

// ------------------------------------------------------------- scenes --
constexpr int kSceneBlock = 128;  // LDS traversal stacks: 32 x 128 x 4 B = 16 KB per workgroup
static_assert(sizeof(SceneView) <= 3072, "SceneView must fit the kernel argument segment");

int make_scene_view(const vr_scene* sc, int bw_mode, SceneView& S) {
    if (!sc) return set_err(VR_ERR_ARG, "scene is NULL");
    std::memset(&S, 0, sizeof(S));
    if (sc->num_tris < 0 || sc->num_nodes < 0 || sc->num_patches < 0 || sc->num_vertices < 0)
        return set_err(VR_ERR_ARG, "scene: negative sizes");
    if (sc->num_tris > 0) {
        if (!sc->nodes || !sc->tri_verts || !sc->tri_ids || !sc->positions || !sc->indices || !sc->tri_patch || !sc->patches)
            return set_err(VR_ERR_ARG, "scene: NULL geometry buffer");
        if (sc->num_nodes < 1 || sc->num_nodes > 2 * sc->num_tris) return set_err(VR_ERR_ARG, "scene: node count out of range");
        if (sc->num_patches < 1) return set_err(VR_ERR_ARG, "scene: no patches");
    }
    if (sc->num_media < 0 || sc->num_media > VR_SCENE_MAX_MEDIA)
        return set_err(VR_ERR_UNSUPPORTED, "scene: at most VR_SCENE_MAX_MEDIA media");
    if (sc->num_media > 0 && !sc->media) return set_err(VR_ERR_ARG, "scene: media is NULL");
    S.nodes = reinterpret_cast<const float4*>(sc->nodes);
    S.tri_verts = sc->tri_verts;
    S.tri_ids = sc->tri_ids;
    S.num_nodes = sc->num_nodes;
    S.num_tris = sc->num_tris;
    S.pos = sc->positions;
    S.nrm = sc->normals;
    S.uv = sc->uvs;
    S.idx = sc->indices;
    S.tri_patch = sc->tri_patch;
    S.patches = sc->patches;
    S.num_patches = sc->num_patches;
    S.num_media = sc->num_media;
    int st;
    for (int m = 0; m < sc->num_media; m++)
        if ((st = make_medium_view(&sc->media[m], VR_MAP_MEDIUM_DT, bw_mode, S.media[m]))) return st;
    // environment: the same checks as a medium's environment
    vr_medium env;
    std::memset(&env, 0, sizeof(env));
    env.env_kind = sc->env_kind;
    env.env_image = sc->env_image;
    for (int a = 0; a < 2; a++) { env.env_image_bmin[a] = sc->env_image_bmin[a]; env.env_image_bmax[a] = sc->env_image_bmax[a]; }
    env.env_sampler_kind = VR_ENVS_UNIFORM;
    MediumView ev;
    std::memset(&ev, 0, sizeof(ev));
    if ((st = make_env_view(&env, ev))) return st;
    S.env_kind = sc->env_kind;
    S.env_c = load3(sc->env_c);
    S.env_img = ev.env_img;
    return VR_OK;
}

__global__ void __launch_bounds__(kSceneBlock) k_scene_raycast(SceneView S, const float* __restrict__ rays, int64_t n,
                                                               float* __restrict__ hits) {
    __shared__ int stk[kSceneStack * kSceneBlock];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const f3 x = load3(rays + 6 * i), w = load3(rays + 6 * i + 3);
    float t, bu, bv;
    const int k = scene_trace<kSceneBlock>(S, x, w, t, bu, bv, stk + threadIdx.x);
    float* h = hits + 16 * i;
    if (k < 0) {  // noHit2array (include/common.h:325-336)
        for (int c = 0; c < 15; c++) h[c] = 0.0f;
        h[15] = __int_as_float(-1);
        return;
    }
    int patch;
    const Surfel s = scene_surfel(S, k, bu, bv, x, w, t, patch);
    // surfel2array (include/common.h:312-323)
    h[0] = t;
    h[1] = s.N.x; h[2] = s.N.y; h[3] = s.N.z;
    h[4] = s.G.x; h[5] = s.G.y; h[6] = s.G.z;
    h[7] = s.cu; h[8] = s.cv;
    for (int c = 9; c < 15; c++) h[c] = 0.0f;
    h[15] = __int_as_float(patch);
}

__global__ void __launch_bounds__(kSceneBlock) k_scene_map_fw(SceneView S, const float* __restrict__ rays, int64_t start,
                                                              int64_t cnt, float* __restrict__ out, Seeds seeds,
                                                              int* __restrict__ dbg) {
    __shared__ int stk[kSceneStack * kSceneBlock];
    const int64_t i = start + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= start + cnt) return;
    Rng rng = seed_map(seeds, (uint32_t)i);
    int lookups = 0;
    const f3 A = scene_path<kSceneBlock>(S, rng, load3(rays + 6 * i), load3(rays + 6 * i + 3), false, stk + threadIdx.x, lookups);
    out[3 * i] = A.x; out[3 * i + 1] = A.y; out[3 * i + 2] = A.z;
    if (dbg) dbg[i] = lookups;
}

__global__ void __launch_bounds__(kSceneBlock) k_scene_map_bw(SceneView S, const float* __restrict__ rays,
                                                              const float* __restrict__ og, int64_t start, int64_t cnt,
                                                              Seeds seeds) {
    __shared__ int stk[kSceneStack * kSceneBlock];
    const int64_t i = start + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= start + cnt) return;
    const f3 g = load3(og + 3 * i);
    if (g.x == 0.0f && g.y == 0.0f && g.z == 0.0f) return;
    Rng rng = seed_map(seeds, (uint32_t)i);
    int lookups = 0, scatters = 0;
    scene_map_bw<kSceneBlock>(S, rng, load3(rays + 6 * i), load3(rays + 6 * i + 3), g, stk + threadIdx.x, lookups, scatters);
}

// Resident waves per SIMD the scene capture kernels are compiled for: 5 cap
// them at 96 VGPRs (137-141 uncapped, 3 waves) and 10 workgroups' traversal
// stacks fill the CU's 160 KB of LDS; on the 256^3 cloud scene at 512x512,
// 16 spp: forward 25.6 ms uncapped -> 22.0 ms (4 or 5 waves), backward 109 ->
// 98.6 (4) -> 96.5 ms (5), same images.
#ifndef VR_SCENE_WAVES_PER_EU
#define VR_SCENE_WAVES_PER_EU 5
#endif
#define VR_SCENE_OCCUPANCY __attribute__((amdgpu_waves_per_eu(VR_SCENE_WAVES_PER_EU)))

// Capture of a PathtracedScene: thread j takes sensor start + j % cnt and samples
// [c*chunk, ...) with c = j / cnt, as k_capture_fw.
__global__ void __launch_bounds__(kSceneBlock) VR_SCENE_OCCUPANCY k_scene_capture_fw(SceneView S, CameraView cam, Seeds seeds, CaptureSplit P,
                                                                  int64_t total, int64_t start, int64_t cnt,
                                                                  float* __restrict__ out, int* __restrict__ dbg) {
    __shared__ int stk[kSceneStack * kSceneBlock];
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cnt * P.chunks) return;
    const int64_t c = j / cnt;
    int64_t i = start + (j - c * cnt);
    const int s0 = (int)c * P.chunk, s1 = min(P.spp, s0 + P.chunk);
    int pose, row, col;
    capture_sensor(cam, P, i, pose, row, col);
    if (pose < 0 || pose >= cam.num_poses) {
        if (P.chunks == 1) { out[3 * i] = 0.0f; out[3 * i + 1] = 0.0f; out[3 * i + 2] = 0.0f; }
        return;
    }
    f3 acc = F3(0.0f, 0.0f, 0.0f);
    int lookups = 0;
    for (int s = s0; s < s1; s++) {
        Rng rng = seed_capture_fw(seeds, (uint32_t)((int64_t)s * total + i));
        f3 x, w;
        camera_ray(cam, pose, row, col, rng, x, w);
        acc = acc + scene_path<kSceneBlock>(S, rng, x, w, false, stk + threadIdx.x, lookups);
    }
    const float inv = 1.0f / (float)P.spp;
    if (P.chunks == 1) {
        out[3 * i] = acc.x * inv; out[3 * i + 1] = acc.y * inv; out[3 * i + 2] = acc.z * inv;
        if (dbg) dbg[i] = lookups;
    } else {
        atomic_addf(out + 3 * i, acc.x * inv);
        atomic_addf(out + 3 * i + 1, acc.y * inv);
        atomic_addf(out + 3 * i + 2, acc.z * inv);
        if (dbg) atomicAdd(dbg + i, lookups);
    }
}

__global__ void __launch_bounds__(kSceneBlock) VR_SCENE_OCCUPANCY k_scene_capture_bw(SceneView S, CameraView cam, Seeds seeds, CaptureSplit P,
                                                                  int seed_mode, int64_t total, int64_t start, int64_t cnt,
                                                                  const float* __restrict__ og, int* __restrict__ dbg) {
    __shared__ int stk[kSceneStack * kSceneBlock];
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cnt * P.chunks) return;
    const int64_t c = j / cnt;
    int64_t i = start + (j - c * cnt);
    const int s0 = (int)c * P.chunk, s1 = min(P.spp, s0 + P.chunk);
    int pose, row, col;
    capture_sensor(cam, P, i, pose, row, col);
    if (pose < 0 || pose >= cam.num_poses) return;
    const f3 g = load3(og + 3 * i) * (1.0f / (float)P.spp);
    if (g.x == 0.0f && g.y == 0.0f && g.z == 0.0f) {
        if (dbg && P.chunks == 1) dbg[i] = 0;
        return;
    }
    int lookups = 0, scatters = 0;
    Rng rng;
    if (seed_mode == VR_SEED_CAPTURE_BW) rng = seed_capture_bw(seeds, (uint32_t)i);
    for (int s = s0; s < s1; s++) {
        if (seed_mode != VR_SEED_CAPTURE_BW) rng = seed_capture_fw(seeds, (uint32_t)((int64_t)s * total + i));
        f3 x, w;
        camera_ray(cam, pose, row, col, rng, x, w);
        scene_map_bw<kSceneBlock>(S, rng, x, w, g, stk + threadIdx.x, lookups, scatters);
    }
    if (dbg) {
        const int n = P.debug_lookups ? lookups : scatters;
        if (P.chunks == 1) dbg[i] = n;
        else atomicAdd(dbg + i, n);
    }
}

// ---- host BVH builder (binned SAH over triangle centroids) ----
struct BvhBuilder {
    struct Box {
        float lo[3], hi[3];
        void reset() { for (int a = 0; a < 3; a++) { lo[a] = 3.0e38f; hi[a] = -3.0e38f; } }
        void grow(const float* p) { for (int a = 0; a < 3; a++) { lo[a] = std::min(lo[a], p[a]); hi[a] = std::max(hi[a], p[a]); } }
        void grow(const Box& b) { for (int a = 0; a < 3; a++) { lo[a] = std::min(lo[a], b.lo[a]); hi[a] = std::max(hi[a], b.hi[a]); } }
        float area() const {
            const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
            return (dx < 0 || dy < 0 || dz < 0) ? 0.0f : 2.0f * (dx * dy + dy * dz + dz * dx);
        }
    };
    static constexpr int kBins = 16, kLeaf = 4, kMaxDepth = 30;
    std::vector<Box> tri_box;
    std::vector<float> cent;   // [T,3]
    std::vector<int32_t> order;
    vr_bvh_node* nodes;
    int64_t max_nodes, used;

    int build(int64_t node, int64_t b, int64_t e, int depth) {
        Box bb, cb;
        bb.reset(); cb.reset();
        for (int64_t k = b; k < e; k++) { bb.grow(tri_box[order[k]]); cb.grow(&cent[3 * (size_t)order[k]]); }
        vr_bvh_node& N = nodes[node];
        for (int a = 0; a < 3; a++) { N.bmin[a] = bb.lo[a]; N.bmax[a] = bb.hi[a]; }
        const int64_t n = e - b;
        if (n <= kLeaf || depth >= kMaxDepth) { N.first = (int32_t)b; N.count = (int32_t)n; return VR_OK; }
        // best binned SAH split over the three axes
        int best_axis = -1, best_bin = -1;
        float best_cost = 3.0e38f;
        for (int a = 0; a < 3; a++) {
            const float ext = cb.hi[a] - cb.lo[a];
            if (!(ext > 0.0f)) continue;
            Box bin_box[kBins];
            int64_t bin_n[kBins] = {0};
            for (int q = 0; q < kBins; q++) bin_box[q].reset();
            const float sc = kBins / ext;
            for (int64_t k = b; k < e; k++) {
                const int t = order[k];
                int q = (int)((cent[3 * (size_t)t + a] - cb.lo[a]) * sc);
                q = std::min(std::max(q, 0), kBins - 1);
                bin_n[q]++;
                bin_box[q].grow(tri_box[t]);
            }
            float left_area[kBins];
            int64_t left_n[kBins];
            Box acc;
            acc.reset();
            int64_t cntl = 0;
            for (int q = 0; q < kBins - 1; q++) {
                acc.grow(bin_box[q]); cntl += bin_n[q];
                left_area[q] = acc.area(); left_n[q] = cntl;
            }
            acc.reset();
            int64_t cntr = 0;
            for (int q = kBins - 1; q >= 1; q--) {
                acc.grow(bin_box[q]); cntr += bin_n[q];
                if (left_n[q - 1] == 0 || cntr == 0) continue;
                const float cost = left_area[q - 1] * (float)left_n[q - 1] + acc.area() * (float)cntr;
                if (cost < best_cost) { best_cost = cost; best_axis = a; best_bin = q; }
            }
        }
        int64_t mid;
        if (best_axis < 0) {
            mid = b + n / 2;  // all centroids coincide: split by index
        } else {
            const float ext = cb.hi[best_axis] - cb.lo[best_axis], sc = kBins / ext;
            const float lo = cb.lo[best_axis];
            auto it = std::partition(order.begin() + b, order.begin() + e, [&](int32_t t) {
                int q = (int)((cent[3 * (size_t)t + best_axis] - lo) * sc);
                return std::min(std::max(q, 0), kBins - 1) < best_bin;
            });
            mid = it - order.begin();
            if (mid == b || mid == e) mid = b + n / 2;
        }
        if (used + 2 > max_nodes) return set_err(VR_ERR_ARG, "vr_scene_build_bvh: node buffer too small");
        const int64_t l = used;
        used += 2;
        N.first = (int32_t)l;
        N.count = 0;
        int st = build(l, b, mid, depth + 1);
        if (st) return st;
        return build(l + 1, mid, e, depth + 1);
    }
};

int resolve_range(const vr_launch* L, int64_t total, int64_t& start, int64_t& cnt) {
    start = L->start_index;
    if (start < 0 || start > total) return set_err(VR_ERR_ARG, "start_index out of range");
    cnt = L->count < 0 ? total - start : L->count;
    if (start + cnt > total) return set_err(VR_ERR_ARG, "start_index + count exceeds the thread count");
    return VR_OK;
}

// Sample s of sensor i draws the stream seeded from the 32-bit flat index
// s * total + i (the fresh-seed backward: one stream per sensor, seeded from i).
// Indices that would wrap, and so repeat streams, are rejected.
int check_sample_indices(int64_t total, int spp, int seed_mode, const char* what) {
    const int64_t limit = int64_t(1) << 32;
    if (total > limit) return set_err(VR_ERR_UNSUPPORTED, "%s: more than 2^32 sensors", what);
    if (seed_mode != VR_SEED_CAPTURE_BW && (int64_t)spp * total > limit)
        return set_err(VR_ERR_UNSUPPORTED, "%s: samples x sensors exceeds 2^32 (per-sample random streams would repeat)",
                       what);
    return VR_OK;
}

int64_t camera_sensors(const vr_camera* c) {
    return c->sensors ? c->num_sensors : (int64_t)c->num_poses * c->width * c->height;
}

constexpr int kMaxDevices = 64;

// Work counters of vr_capture_fw_noalloc: a file-scope device array (one copy
// per GPU, held until the process ends) instead of a workspace.  In host code
// `scratch` only names the symbol; its device address on each GPU comes from
// hipGetSymbolAddress, taken once per device with that device current.
constexpr int kScratchWords = kMaxQueues * kQueueStride;
static __device__ __align__(256) unsigned long long scratch[kScratchWords];
static_assert(sizeof(scratch) >= (size_t)kCounterBytes, "scratch holds the work counters of every queue");
std::mutex g_scratch_mu;
unsigned long long* g_scratch_ptr[kMaxDevices] = {};

int scratch_counters(unsigned long long** out) {
    int dev = 0;
    int st = check_hip(hipGetDevice(&dev), "hipGetDevice");
    if (st) return st;
    if (dev < 0 || dev >= kMaxDevices) return set_err(VR_ERR_DEVICE, "vr_capture_fw_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if ((st = check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)"))) return st;
        g_scratch_ptr[dev] = static_cast<unsigned long long*>(p);
    }
    *out = g_scratch_ptr[dev];
    return VR_OK;
}

// vr_capture_fw and vr_capture_fw_noalloc: the persistent NEE kernel's work
// counters come from the launch's workspace (or a stream-ordered allocation),
// or, with `static_counters`, from the device array above.
int capture_fw(const vr_medium* medium, int32_t kind, const vr_camera* camera, const vr_launch* L, float* out, void* stream,
               bool static_counters, const char* who) {
    int id, od;
    if (!map_dims(kind, id, od)) return set_err(VR_ERR_UNSUPPORTED, "%s: unknown map kind", who);
    if (id != 6) return set_err(VR_ERR_UNSUPPORTED, "%s: capture needs a 6-input map", who);
    if (!L) return set_err(VR_ERR_ARG, "%s: launch is NULL", who);
    MediumView M;
    int st = make_medium_view(medium, kind, L->bw_mode, M);
    if (st) return st;
    CameraView C;
    if ((st = make_camera_view(camera, C))) return st;
    const int64_t total = camera_sensors(camera);
    int64_t start, cnt;
    if ((st = resolve_range(L, total, start, cnt))) return st;
    if (cnt == 0) return VR_OK;
    if (!out) return set_err(VR_ERR_ARG, "%s: out is NULL", who);
    int spp = L->samples < 1 ? 1 : L->samples;
    if ((st = check_sample_indices(total, spp, VR_SEED_CAPTURE_FW, who))) return st;
    Seeds sd = make_seeds(L);
    hipStream_t s = (hipStream_t)stream;
    Counters ctr;
    if (capture_is_persistent(kind, M)) {
        if (static_counters) st = scratch_counters(&ctr.ptr);
        else st = ctr.acquire(L, s);
        if (st) return st;
    }
    if (medium->majorant_kind == VR_MAJ_GRID)
        st = launch_capture_fw_maj<VR_MAJ_GRID>(kind, M, C, sd, spp, total, start, cnt, out, L->debug_out, ctr.ptr, s);
    else
        st = launch_capture_fw_maj<VR_MAJ_CONST>(kind, M, C, sd, spp, total, start, cnt, out, L->debug_out, ctr.ptr, s);
    const int rel = ctr.release();
    return st ? st : rel;
}

}  // namespace

// ================================================================== C ABI ==
extern "C" {

int vr_abi_version(void) { return VR_ABI_VERSION; }

const char* vr_last_error(void) { return g_err; }

int vr_grid_fw(const vr_grid* grid, const float* x, int64_t n, float* out, void* stream) {
    GridView g;
    int st = make_grid_view(grid, g, 0, "vr_grid_fw");
    if (st) return st;
    if (n < 0) return set_err(VR_ERR_ARG, "vr_grid_fw: negative n");
    if (n == 0) return VR_OK;
    if (!x || !out) return set_err(VR_ERR_ARG, "vr_grid_fw: null buffer");
    hipLaunchKernelGGL(k_grid_fw, dim3(grid_for(n)), dim3(kBlock), 0, (hipStream_t)stream, g, x, n, out);
    return check_hip(hipGetLastError(), "vr_grid_fw launch");
}

int vr_grid_bw(const vr_grid* grid, const float* x, const float* out_grad, int64_t n, void* stream) {
    GridView g;
    int st = make_grid_view(grid, g, 0, "vr_grid_bw");
    if (st) return st;
    if (n < 0) return set_err(VR_ERR_ARG, "vr_grid_bw: negative n");
    if (n == 0 || g.grad == nullptr) return VR_OK;
    if (!x || !out_grad) return set_err(VR_ERR_ARG, "vr_grid_bw: null buffer");
    hipLaunchKernelGGL(k_grid_bw, dim3(grid_for(n)), dim3(kBlock), 0, (hipStream_t)stream, g, x, out_grad, n);
    return check_hip(hipGetLastError(), "vr_grid_bw launch");
}

int vr_majorant_build(const vr_grid* sigma, const int32_t res[3], float* out, void* stream) {
    GridView g;
    int st = make_grid_view(sigma, g, 1, "vr_majorant_build");
    if (st) return st;
    if (!res || !out) return set_err(VR_ERR_ARG, "vr_majorant_build: null argument");
    if (res[0] < 1 || res[1] < 1 || res[2] < 1 || (int64_t)res[0] * res[1] * res[2] > (1 << 24))
        return set_err(VR_ERR_ARG, "vr_majorant_build: resolution out of range");
    hipLaunchKernelGGL(k_majorant, dim3(res[0] * res[1] * res[2]), dim3(kBlock), 0, (hipStream_t)stream, g, res[0], res[1],
                       res[2], out);
    return check_hip(hipGetLastError(), "vr_majorant_build launch");
}

int vr_map_fw(const vr_medium* medium, int32_t kind, const float* in, int64_t n, float* out, const vr_launch* L, void* stream) {
    int id, od;
    if (!map_dims(kind, id, od)) return set_err(VR_ERR_UNSUPPORTED, "vr_map_fw: unknown map kind");
    if (!L) return set_err(VR_ERR_ARG, "vr_map_fw: launch is NULL");
    MediumView M;
    int st = make_medium_view(medium, kind, L->bw_mode, M);
    if (st) return st;
    int64_t start, cnt;
    if ((st = resolve_range(L, n, start, cnt))) return st;
    if (cnt == 0) return VR_OK;
    if (!in || !out) return set_err(VR_ERR_ARG, "vr_map_fw: null buffer");
    Seeds sd = make_seeds(L);
    hipStream_t s = (hipStream_t)stream;
    if (medium->majorant_kind == VR_MAJ_GRID) return launch_map_fw_maj<VR_MAJ_GRID>(kind, M, in, start, cnt, out, sd, L->debug_out, s);
    return launch_map_fw_maj<VR_MAJ_CONST>(kind, M, in, start, cnt, out, sd, L->debug_out, s);
}

int vr_map_bw(const vr_medium* medium, int32_t kind, const float* in, const float* og, int64_t n, const vr_launch* L, void* stream) {
    int id, od;
    if (!map_dims(kind, id, od)) return set_err(VR_ERR_UNSUPPORTED, "vr_map_bw: unknown map kind");
    if (!L) return set_err(VR_ERR_ARG, "vr_map_bw: launch is NULL");
    MediumView M;
    int st = make_medium_view(medium, kind, L->bw_mode, M);
    if (st) return st;
    int64_t start, cnt;
    if ((st = resolve_range(L, n, start, cnt))) return st;
    if (cnt == 0) return VR_OK;
    if (!in || !og) return set_err(VR_ERR_ARG, "vr_map_bw: null buffer");
    Seeds sd = make_seeds(L);
    hipStream_t s = (hipStream_t)stream;
    if (medium->majorant_kind == VR_MAJ_GRID) return launch_map_bw_maj<VR_MAJ_GRID>(kind, M, in, og, start, cnt, sd, s);
    return launch_map_bw_maj<VR_MAJ_CONST>(kind, M, in, og, start, cnt, sd, s);
}

int vr_capture_fw(const vr_medium* medium, int32_t kind, const vr_camera* camera, const vr_launch* L, float* out, void* stream) {
    return capture_fw(medium, kind, camera, L, out, stream, false, "vr_capture_fw");
}

int vr_capture_fw_noalloc(const vr_medium* medium, int32_t kind, const vr_camera* camera, const vr_launch* L, float* out,
                          void* stream) {
    return capture_fw(medium, kind, camera, L, out, stream, true, "vr_capture_fw_noalloc");
}

int vr_capture_bw(const vr_medium* medium, int32_t kind, const vr_camera* camera, const vr_launch* L, const float* og, void* stream) {
    int id, od;
    if (!map_dims(kind, id, od)) return set_err(VR_ERR_UNSUPPORTED, "vr_capture_bw: unknown map kind");
    if (id != 6) return set_err(VR_ERR_UNSUPPORTED, "vr_capture_bw: capture needs a 6-input map");
    if (!L) return set_err(VR_ERR_ARG, "vr_capture_bw: launch is NULL");
    MediumView M;
    int st = make_medium_view(medium, kind, L->bw_mode, M);
    if (st) return st;
    CameraView C;
    if ((st = make_camera_view(camera, C))) return st;
    const int64_t total = camera_sensors(camera);
    int64_t start, cnt;
    if ((st = resolve_range(L, total, start, cnt))) return st;
    if (cnt == 0) return VR_OK;
    if (!og) return set_err(VR_ERR_ARG, "vr_capture_bw: out_grad is NULL");
    int spp = L->samples < 1 ? 1 : L->samples;
    if ((st = check_sample_indices(total, spp, L->seed_mode, "vr_capture_bw"))) return st;
    Seeds sd = make_seeds(L);
    hipStream_t s = (hipStream_t)stream;
    Counters ctr;
    if ((capture_bw_is_pooled(kind, M, L->seed_mode) || capture_bw_is_ss_pooled(kind, M, L->seed_mode)) &&
        (st = ctr.acquire(L, s)))
        return st;
    if (medium->majorant_kind == VR_MAJ_GRID)
        st = launch_capture_bw_maj<VR_MAJ_GRID>(kind, M, C, sd, spp, L->seed_mode, L->debug_kind, total, start, cnt, og,
                                                L->debug_out, ctr.ptr, s);
    else
        st = launch_capture_bw_maj<VR_MAJ_CONST>(kind, M, C, sd, spp, L->seed_mode, L->debug_kind, total, start, cnt, og,
                                                 L->debug_out, ctr.ptr, s);
    const int rel = ctr.release();
    return st ? st : rel;
}

int vr_adam_step_bounded(float* p, float* g, float* m, float* v, int64_t n, float lr, float beta1, float beta2, float eps,
                         int32_t step, float lo, float hi, int32_t zero_grad, void* stream) {
    if (n < 0) return set_err(VR_ERR_ARG, "vr_adam_step: negative n");
    if (step < 1) return set_err(VR_ERR_ARG, "vr_adam_step: step must be >= 1");
    if (!(lo <= hi)) return set_err(VR_ERR_ARG, "vr_adam_step: bounds must satisfy lo <= hi");
    if (n == 0) return VR_OK;
    if (!p || !g || !m || !v) return set_err(VR_ERR_ARG, "vr_adam_step: null buffer");
    float bc1 = (float)(1.0 - std::pow((double)beta1, step));
    float bc2s = (float)std::sqrt(1.0 - std::pow((double)beta2, step));
    int64_t blocks = (n + kBlock - 1) / kBlock;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_adam, dim3((unsigned)blocks), dim3(kBlock), 0, (hipStream_t)stream, p, g, m, v, n, lr, beta1, beta2, eps,
                       bc1, bc2s, lo, hi, zero_grad);
    return check_hip(hipGetLastError(), "vr_adam_step launch");
}

int vr_adam_step(float* p, float* g, float* m, float* v, int64_t n, float lr, float beta1, float beta2, float eps, int32_t step,
                 int32_t zero_grad, void* stream) {
    return vr_adam_step_bounded(p, g, m, v, n, lr, beta1, beta2, eps, step, -INFINITY, INFINITY, zero_grad, stream);
}

}  // extern "C"

// ------------------------------------------------------------- grid tools --

namespace {

// Spatial dims (outermost first, as given) -> x, y, z innermost first.
int make_shape3(const int64_t* shape, int dim, int64_t channels, fn::Shape3& S, const char* what) {
    if (!shape) return set_err(VR_ERR_ARG, "%s: shape is NULL", what);
    if (dim < 1 || dim > 3) return set_err(VR_ERR_ARG, "%s: 1 to 3 spatial dimensions", what);
    if (channels < 1 || channels > (1 << 20)) return set_err(VR_ERR_ARG, "%s: bad channel count", what);
    int64_t total = 1;
    for (int i = 0; i < 3; i++) {
        const int64_t n = i < dim ? shape[dim - 1 - i] : 1;
        if (n < 1 || n > (int64_t(1) << 30)) return set_err(VR_ERR_ARG, "%s: bad extent", what);
        S.n[i] = (int)n;
        S.s[i] = total;
        total *= n;
        if (total > (int64_t(1) << 40)) return set_err(VR_ERR_ARG, "%s: tensor too large", what);
    }
    S.C = (int)channels;
    S.xb = (S.n[0] + fn::kRowBlock - 1) / fn::kRowBlock;
    if ((int64_t)S.xb * S.n[1] * S.n[2] >= (int64_t(1) << 31)) return set_err(VR_ERR_ARG, "%s: too many rows", what);
    return VR_OK;
}

inline dim3 shape_grid(const fn::Shape3& S) { return dim3((unsigned)((int64_t)S.xb * S.n[1] * S.n[2])); }

// Single-channel 3-D tensors with rows of whole float4s take the x4 kernels.
bool dense3_of(const fn::Shape3& S, int dim, const void* a, const void* b, const void* c, fn::Dense3& D) {
    const auto aligned = [](const void* p) { return ((uintptr_t)p & 15u) == 0; };
    if (dim != 3 || S.C != 1 || (S.n[0] & 3) || !aligned(a) || !aligned(b) || !aligned(c)) return false;
    if ((S.n[1] + fn::kRowsY - 1) / fn::kRowsY > 65535 || S.n[2] > 65535) return false;
    D.nx = S.n[0]; D.ny = S.n[1]; D.nz = S.n[2];
    D.sy = S.s[1]; D.sz = S.s[2];
    return true;
}

inline dim3 dense3_grid(const fn::Dense3& D) {
    return dim3((unsigned)((D.nx / 4 + fn::kQuadX - 1) / fn::kQuadX), (unsigned)((D.ny + fn::kRowsY - 1) / fn::kRowsY),
                (unsigned)D.nz);
}

}  // namespace

extern "C" {

int vr_random_ids(int64_t* out, int64_t n, const int64_t* shape, int32_t dim, const uint32_t seeds[4], void* stream) {
    if (n < 0 || n >= (int64_t(1) << 31)) return set_err(VR_ERR_ARG, "vr_random_ids: n out of range");
    if (dim < 1 || dim > 4) return set_err(VR_ERR_ARG, "vr_random_ids: 1 to 4 dimensions");
    if (!shape || !seeds) return set_err(VR_ERR_ARG, "vr_random_ids: shape or seeds is NULL");
    int sh[4] = {0, 0, 0, 0};
    for (int k = 0; k < dim; k++) {
        if (shape[k] < 1 || shape[k] >= (int64_t(1) << 31)) return set_err(VR_ERR_ARG, "vr_random_ids: bad extent");
        sh[k] = (int)shape[k];
    }
    if (n == 0) return VR_OK;
    if (!out) return set_err(VR_ERR_ARG, "vr_random_ids: out is NULL");
    Seeds sd;
    for (int k = 0; k < 4; k++) sd.s[k] = seeds[k];
    hipLaunchKernelGGL(fn::k_random_ids, dim3(grid_for(n)), dim3(kBlock), 0, (hipStream_t)stream, out, n,
                       make_int4(sh[0], sh[1], sh[2], sh[3]), (int)dim, sd);
    return check_hip(hipGetLastError(), "vr_random_ids launch");
}

int vr_total_variation_fw(const float* in, const int64_t* shape, int32_t dim, float* out, void* stream) {
    fn::Shape3 S;
    int st = make_shape3(shape, dim, shape && dim >= 1 && dim <= 3 ? shape[dim] : 0, S, "vr_total_variation_fw");
    if (st) return st;
    if (!in || !out) return set_err(VR_ERR_ARG, "vr_total_variation_fw: null buffer");
    fn::Dense3 D;
    if (dense3_of(S, dim, in, out, out, D)) {
        hipLaunchKernelGGL(fn::k_tv_fw_x4, dense3_grid(D), dim3(fn::kQuadX, fn::kRowsY), 0, (hipStream_t)stream, in, D, out);
        return check_hip(hipGetLastError(), "vr_total_variation_fw launch");
    }
    hipLaunchKernelGGL(fn::k_tv_fw, shape_grid(S), dim3(fn::kRowBlock), 0, (hipStream_t)stream, in, S, (int)dim, out);
    return check_hip(hipGetLastError(), "vr_total_variation_fw launch");
}

int vr_total_variation_bw(const float* in, const float* out_grad, const int64_t* shape, int32_t dim, float* in_grad,
                          void* stream) {
    fn::Shape3 S;
    int st = make_shape3(shape, dim, shape && dim >= 1 && dim <= 3 ? shape[dim] : 0, S, "vr_total_variation_bw");
    if (st) return st;
    if (!in || !out_grad || !in_grad) return set_err(VR_ERR_ARG, "vr_total_variation_bw: null buffer");
    fn::Dense3 D;
    if (dense3_of(S, dim, in, out_grad, in_grad, D)) {
        hipLaunchKernelGGL(fn::k_tv_bw_x4, dense3_grid(D), dim3(fn::kQuadX, fn::kRowsY), 0, (hipStream_t)stream, in, out_grad,
                           D, in_grad);
        return check_hip(hipGetLastError(), "vr_total_variation_bw launch");
    }
    hipLaunchKernelGGL(fn::k_tv_bw, shape_grid(S), dim3(fn::kRowBlock), 0, (hipStream_t)stream, in, out_grad, S, (int)dim,
                       in_grad);
    return check_hip(hipGetLastError(), "vr_total_variation_bw launch");
}

}  // extern "C"

namespace {

template <bool CELL>
int resample(const float* in, const int64_t* in_shape, const int64_t* out_shape, int32_t dim, int32_t channels, float* out,
             void* stream, const char* what) {
    fn::Resample R;
    fn::Shape3 I;
    int st = make_shape3(in_shape, dim, channels, I, what);
    if (st) return st;
    if ((st = make_shape3(out_shape, dim, channels, R.out, what))) return st;
    for (int i = 0; i < 3; i++) {
        if (!CELL && i < dim && R.out.n[i] < 2) return set_err(VR_ERR_ARG, "%s: output extents must be >= 2", what);
        // the source coordinate is formed as the int p * (in - 1), as in the shader
        if ((int64_t)(R.out.n[i] - 1) * (I.n[i] - 1) >= (int64_t(1) << 31))
            return set_err(VR_ERR_ARG, "%s: (out - 1) * (in - 1) must stay below 2^31", what);
        R.in_n[i] = I.n[i];
        R.in_s[i] = I.s[i];
    }
    R.dim = dim;
    if (!in || !out) return set_err(VR_ERR_ARG, "%s: null buffer", what);
    fn::Resample3 R3;
    if (dense3_of(R.out, dim, out, out, out, R3.out)) {
        for (int i = 0; i < 3; i++) R3.in_n[i] = I.n[i];
        R3.in_sy = I.s[1];
        R3.in_sz = I.s[2];
        // source rows a band of kBandY output rows reads: (kBandY - 1) * ratio + 2,
        // one more for rounding; the kernel itself checks the exact count
        const double ratio = CELL ? (double)I.n[1] / R.out.n[1] : (double)(I.n[1] - 1) / (R.out.n[1] - 1);
        const int64_t band_rows = (int64_t)std::ceil((fn::kBandY - 1) * ratio) + 3;
        const int64_t lds_floats = 3 * (int64_t)R3.out.nx + 2 * band_rows * I.n[0];
        if (env_int("VR_RESAMPLE_LDS", 1) != 0 && lds_floats * 4 <= fn::kResampleLdsBytes) {
            const dim3 grid((unsigned)((R3.out.ny + fn::kBandY - 1) / fn::kBandY), (unsigned)R3.out.nz);
            hipLaunchKernelGGL(fn::k_resample_lds<CELL>, grid, dim3(256), (unsigned)(lds_floats * 4), (hipStream_t)stream,
                               in, R3, out, (int)lds_floats);
            return check_hip(hipGetLastError(), what);
        }
        hipLaunchKernelGGL(fn::k_resample_x4<CELL>, dense3_grid(R3.out), dim3(fn::kQuadX, fn::kRowsY), 0, (hipStream_t)stream, in,
                           R3, out);
        return check_hip(hipGetLastError(), what);
    }
    hipLaunchKernelGGL(fn::k_resample<CELL>, shape_grid(R.out), dim3(fn::kRowBlock), 0, (hipStream_t)stream, in, R, out);
    return check_hip(hipGetLastError(), what);
}

template <bool TO_MORTON>
int morton_copy(const float* in, int64_t resolution, int32_t channels, float* out, void* stream, const char* what) {
    if (resolution < 1 || resolution > (1 << 15) || (resolution & (resolution - 1)))
        return set_err(VR_ERR_ARG, "%s: resolution must be a power of two up to 2^15", what);
    if (channels < 1 || channels > 64) return set_err(VR_ERR_ARG, "%s: 1 to 64 channels", what);
    if (!in || !out) return set_err(VR_ERR_ARG, "%s: null buffer", what);
    hipLaunchKernelGGL(fn::k_morton_copy<TO_MORTON>, dim3(grid_for(resolution * resolution)), dim3(kBlock), 0,
                       (hipStream_t)stream, in, (int)resolution, (int)channels, out);
    return check_hip(hipGetLastError(), what);
}

}  // namespace

extern "C" {

int vr_resample_grid(const float* in, const int64_t* in_shape, const int64_t* out_shape, int32_t dim, int32_t channels,
                     float* out, void* stream) {
    return resample<false>(in, in_shape, out_shape, dim, channels, out, stream, "vr_resample_grid");
}

int vr_resample_img(const float* in, const int64_t* in_shape, const int64_t* out_shape, int32_t dim, int32_t channels,
                    float* out, void* stream) {
    return resample<true>(in, in_shape, out_shape, dim, channels, out, stream, "vr_resample_img");
}

int vr_copy_img_to_morton(const float* in, int64_t resolution, int32_t channels, float* out, void* stream) {
    return morton_copy<true>(in, resolution, channels, out, stream, "vr_copy_img_to_morton");
}

int vr_copy_morton_to_img(const float* in, int64_t resolution, int32_t channels, float* out, void* stream) {
    return morton_copy<false>(in, resolution, channels, out, stream, "vr_copy_morton_to_img");
}

}  // extern "C"

// ------------------------------------------------------------------ cache --
namespace {

__global__ void k_zero_f32(float* __restrict__ p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = 0.0f;
}

int box_transform(const float bmin[3], const float bmax[3], f3& o, f3& ib) {
    if (!bmin || !bmax) return set_err(VR_ERR_ARG, "nrc: box is NULL");
    for (int a = 0; a < 3; a++)
        if (!(bmax[a] > bmin[a])) return set_err(VR_ERR_ARG, "nrc: box must have positive extent");
    o = F3(bmin[0], bmin[1], bmin[2]);
    ib = F3(1.0f / (bmax[0] - bmin[0]), 1.0f / (bmax[1] - bmin[1]), 1.0f / (bmax[2] - bmin[2]));
    return VR_OK;
}

// Blocks of `waves_per_block` waves for one 16-row tile per wave (at most cap;
// the kernels loop over further tiles).
unsigned nrc_blocks(int64_t rows, int waves_per_block, unsigned cap) {
    int64_t b = (rows + (int64_t)nrc::kTileRows * waves_per_block - 1) / ((int64_t)nrc::kTileRows * waves_per_block);
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(b, cap));
}

// Blocks of the gradient/loss zeroing launch: one element per thread.
unsigned nrc_zero_blocks() { return (unsigned)((nrc::kGradCopies * nrc::kParams + kBlock - 1) / kBlock); }

}  // namespace

extern "C" {

int vr_nrc_encode(const float* x, const float* w, int64_t n, const float box_min[3], const float box_max[3],
                  float* features, void* stream) {
    if (n < 0) return set_err(VR_ERR_ARG, "vr_nrc_encode: negative n");
    if (n == 0) return VR_OK;
    if (!x || !w || !features) return set_err(VR_ERR_ARG, "vr_nrc_encode: null buffer");
    f3 o, ib;
    int st = box_transform(box_min, box_max, o, ib);
    if (st) return st;
    hipLaunchKernelGGL(nrc::k_nrc_encode, dim3(grid_for(n * nrc::kIn)), dim3(kBlock), 0, (hipStream_t)stream, x, w, n, o, ib,
                       features);
    return check_hip(hipGetLastError(), "vr_nrc_encode launch");
}

int vr_nrc_mlp_fw(const float* features, int64_t n, const uint16_t* weights, float* out, void* stream) {
    if (n < 0) return set_err(VR_ERR_ARG, "vr_nrc_mlp_fw: negative n");
    if (n == 0) return VR_OK;
    if (!features || !weights || !out) return set_err(VR_ERR_ARG, "vr_nrc_mlp_fw: null buffer");
    if (((uintptr_t)weights & 15u) != 0 || ((uintptr_t)features & 15u) != 0)
        return set_err(VR_ERR_ARG, "vr_nrc_mlp_fw: weights and features must be 16-byte aligned");
    // one resident block per CU, each streaming many tiles (2.83 M rows: 0.24 ms
    // against 0.27 ms with 4096 blocks, which stage the weights 16x as often)
    hipLaunchKernelGGL(nrc::k_nrc_mlp_fw, dim3(nrc_blocks(n, nrc::kInferWaves, (unsigned)env_int("VR_NRC_INFER_BLOCKS", 256))),
                       dim3(64 * nrc::kInferWaves), 0,
                       (hipStream_t)stream, features, n, (const __bf16*)weights, out);
    return check_hip(hipGetLastError(), "vr_nrc_mlp_fw launch");
}

int vr_nrc_train_step(const float* records, const uint32_t* count, int64_t first, int64_t max_records, const float box_min[3],
                      const float box_max[3], float* params, uint16_t* weights, float* grad, float* m, float* v,
                      int32_t step, float lr, float beta1, float beta2, float eps, float* loss, void* stream) {
    if (max_records < 0 || first < 0) return set_err(VR_ERR_ARG, "vr_nrc_train_step: negative sizes");
    if (step < 1) return set_err(VR_ERR_ARG, "vr_nrc_train_step: step must be >= 1");
    if (!records || !count || !params || !weights || !grad || !m || !v || !loss)
        return set_err(VR_ERR_ARG, "vr_nrc_train_step: null buffer");
    f3 o, ib;
    int st = box_transform(box_min, box_max, o, ib);
    if (st) return st;
    if (((uintptr_t)weights & 15u) != 0) return set_err(VR_ERR_ARG, "vr_nrc_train_step: weights must be 16-byte aligned");
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(nrc::k_nrc_zero, dim3(nrc_zero_blocks()), dim3(kBlock), 0, s, grad, loss);
    if (max_records > 0)
        hipLaunchKernelGGL(nrc::k_nrc_train, dim3(nrc_blocks(max_records, nrc::kTrainWaves, 1024)), dim3(64 * nrc::kTrainWaves),
                           0, s, records, count, first, max_records, o, ib, (const __bf16*)weights, grad, loss);
    const float bc1 = (float)(1.0 - std::pow((double)beta1, step));
    const float bc2s = (float)std::sqrt(1.0 - std::pow((double)beta2, step));
    hipLaunchKernelGGL(nrc::k_nrc_adam, dim3((nrc::kParams + kBlock - 1) / kBlock), dim3(kBlock), 0, s, params, grad, m, v,
                       (__bf16*)weights, lr, beta1, beta2, eps, bc1, bc2s, loss);
    return check_hip(hipGetLastError(), "vr_nrc_train_step launch");
}

int vr_nrc_paths_traces(const vr_medium* medium, const vr_camera* camera, const vr_launch* launches, int32_t traces,
                        int32_t n_paths, int32_t max_vertices, float* scratch, float* records, int64_t max_records,
                        uint32_t* counts, void* stream) {
    if (!launches) return set_err(VR_ERR_ARG, "vr_nrc_paths: launch is NULL");
    if (traces < 1 || traces > nrc::kMaxTraces) return set_err(VR_ERR_ARG, "vr_nrc_paths_traces: traces must be in [1, 16]");
    if (n_paths < 0 || max_vertices < 1 || max_records < 0) return set_err(VR_ERR_ARG, "vr_nrc_paths: bad sizes");
    if (!scratch || !records || !counts) return set_err(VR_ERR_ARG, "vr_nrc_paths: null buffer");
    MediumView M;
    int st = make_medium_view(medium, VR_MAP_NEE_DT, launches[0].bw_mode, M);
    if (st) return st;
    CameraView C;
    if ((st = make_camera_view(camera, C))) return st;
    const int64_t total = camera->sensors ? camera->num_sensors : (int64_t)camera->num_poses * camera->width * camera->height;
    if (total <= 0) return set_err(VR_ERR_ARG, "vr_nrc_paths: no sensors");
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, counts, traces);
    if (n_paths > 0) {
        nrc::TraceSeeds sd = {};
        for (int t = 0; t < traces; t++) sd.s[t] = make_seeds(&launches[t]);
        const dim3 g((n_paths + 63) / 64, traces);
        if (medium->majorant_kind == VR_MAJ_GRID)
            hipLaunchKernelGGL(nrc::k_nrc_paths<VR_MAJ_GRID>, g, dim3(64), 0, s, M, C, sd, total, n_paths, max_vertices,
                               scratch, records, max_records, counts);
        else
            hipLaunchKernelGGL(nrc::k_nrc_paths<VR_MAJ_CONST>, g, dim3(64), 0, s, M, C, sd, total, n_paths, max_vertices,
                               scratch, records, max_records, counts);
    }
    return check_hip(hipGetLastError(), "vr_nrc_paths launch");
}

int vr_nrc_paths(const vr_medium* medium, const vr_camera* camera, const vr_launch* L, int32_t n_paths,
                 int32_t max_vertices, float* scratch, float* records, int64_t max_records, uint32_t* count,
                 void* stream) {
    return vr_nrc_paths_traces(medium, camera, L, 1, n_paths, max_vertices, scratch, records, max_records, count, stream);
}

int vr_nrc_render(const vr_medium* medium, const vr_camera* camera, const vr_launch* L, int32_t terminate_after,
                  const uint16_t* weights, float* out, float* queries, int64_t max_queries, uint32_t* count,
                  void* stream) {
    if (!L) return set_err(VR_ERR_ARG, "vr_nrc_render: launch is NULL");
    if (terminate_after < 0 || max_queries < 0) return set_err(VR_ERR_ARG, "vr_nrc_render: bad sizes");
    if (!weights || !out || !queries || !count) return set_err(VR_ERR_ARG, "vr_nrc_render: null buffer");
    if (((uintptr_t)weights & 15u) != 0) return set_err(VR_ERR_ARG, "vr_nrc_render: weights must be 16-byte aligned");
    MediumView M;
    int st = make_medium_view(medium, VR_MAP_NEE_DT, L->bw_mode, M);
    if (st) return st;
    if (M.t_kind != VR_T_DT && M.t_kind != VR_T_RT) return set_err(VR_ERR_UNSUPPORTED, "vr_nrc_render: DT or RT shadow rays only");
    CameraView C;
    if ((st = make_camera_view(camera, C))) return st;
    const int64_t total = camera_sensors(camera);
    int64_t start, cnt;
    if ((st = resolve_range(L, total, start, cnt))) return st;
    const int spp = L->samples < 1 ? 1 : L->samples;
    if ((st = check_sample_indices(total, spp, VR_SEED_CAPTURE_FW, "vr_nrc_render"))) return st;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, count, 1);
    if (cnt == 0) return check_hip(hipGetLastError(), "vr_nrc_render launch");
    const Seeds sd = make_seeds(L);
    const bool envmap = M.env_kind != VR_ENV_CONST || M.env_sampler != VR_ENVS_UNIFORM;
    if (!envmap && env_int("VR_NRC_PERSISTENT", 1) != 0) {
        // the persistent lane state machine with paths ending into cache queries;
        // the cached estimator has no bounce cap of its own
        MediumView Mc = M;
        Mc.max_bounces = -1;
        const bool lean = Mc.cells && !Mc.albedo_grid && !Mc.phase_g_grid && env_int("VR_NEE_LEAN", 1) != 0;
        NeeCache cache;
        cache.queries = queries;
        cache.max_queries = max_queries;
        cache.qcount = count;
        cache.k = terminate_after;
        Counters ctr;
        if ((st = ctr.acquire(L, s))) return st;
#define VR_NRC_ARGS Mc, C, sd, spp, total, start, cnt, out, nullptr, ctr.ptr, s, cache
        if (medium->majorant_kind == VR_MAJ_GRID) {
            if (Mc.t_kind == VR_T_DT && lean) st = launch_nee_persistent<VR_MAJ_GRID, VR_T_DT, false, true, true>(VR_NRC_ARGS);
            else if (Mc.t_kind == VR_T_DT) st = launch_nee_persistent<VR_MAJ_GRID, VR_T_DT, false, false, true>(VR_NRC_ARGS);
            else st = launch_nee_persistent<VR_MAJ_GRID, VR_T_RT, false, false, true>(VR_NRC_ARGS);
        } else {
            if (Mc.t_kind == VR_T_DT && lean) st = launch_nee_persistent<VR_MAJ_CONST, VR_T_DT, false, true, true>(VR_NRC_ARGS);
            else if (Mc.t_kind == VR_T_DT) st = launch_nee_persistent<VR_MAJ_CONST, VR_T_DT, false, false, true>(VR_NRC_ARGS);
            else st = launch_nee_persistent<VR_MAJ_CONST, VR_T_RT, false, false, true>(VR_NRC_ARGS);
        }
#undef VR_NRC_ARGS
        const int rel = ctr.release();
        if (st || (st = rel)) return st;
    } else {
        const CaptureSplit P = capture_split(spp, false);
        if ((st = capture_prepare(P, out, 3, start, cnt, nullptr, s))) return st;
        const dim3 g(grid_for(cnt * P.chunks)), b(kBlock);
        if (medium->majorant_kind == VR_MAJ_GRID)
            hipLaunchKernelGGL(nrc::k_nrc_render<VR_MAJ_GRID>, g, b, 0, s, M, C, sd, spp, P.chunk, P.chunks, (int)terminate_after,
                               total, start, cnt, out, queries, max_queries, count);
        else
            hipLaunchKernelGGL(nrc::k_nrc_render<VR_MAJ_CONST>, g, b, 0, s, M, C, sd, spp, P.chunk, P.chunks, (int)terminate_after,
                               total, start, cnt, out, queries, max_queries, count);
        if ((st = check_hip(hipGetLastError(), "vr_nrc_render trace"))) return st;
    }
    f3 o, ib;
    if ((st = box_transform(medium->box_min, medium->box_max, o, ib))) return st;
    if (max_queries > 0)
        hipLaunchKernelGGL(nrc::k_nrc_apply, dim3(nrc_blocks(max_queries, nrc::kInferWaves, (unsigned)env_int("VR_NRC_APPLY_BLOCKS", 512))),
                           dim3(64 * nrc::kInferWaves), 0,
                           s, queries, count, max_queries, o, ib, (const __bf16*)weights, out);
    return check_hip(hipGetLastError(), "vr_nrc_render apply");
}


int vr_grid_pack_cells(const vr_grid* grid, float* out, void* stream) {
    GridView g;
    int st = make_grid_view(grid, g, 1, "vr_grid_pack_cells");
    if (st) return st;
    if (!out) return set_err(VR_ERR_ARG, "vr_grid_pack_cells: out is NULL");
    if (reinterpret_cast<uintptr_t>(out) % 16 != 0) return set_err(VR_ERR_ARG, "vr_grid_pack_cells: out must be 16-byte aligned");
    const int64_t cells = (int64_t)(g.nx - 1) * (g.ny - 1) * (g.nz - 1);
    if (2 * cells >= (int64_t(1) << 31)) return set_err(VR_ERR_ARG, "vr_grid_pack_cells: more than 2^30 cells");
    hipLaunchKernelGGL(k_pack_cells, dim3(grid_for(cells)), dim3(kBlock), 0, (hipStream_t)stream, g,
                       reinterpret_cast<float4*>(out), cells);
    return check_hip(hipGetLastError(), "vr_grid_pack_cells launch");
}

int vr_scene_build_bvh(const float* positions, int64_t num_vertices, const int32_t* indices, int64_t num_tris,
                       vr_bvh_node* nodes, int64_t max_nodes, int32_t* tri_ids, float* tri_verts, int32_t* num_nodes) {
    if (num_vertices < 0 || num_tris < 0) return set_err(VR_ERR_ARG, "vr_scene_build_bvh: negative sizes");
    if (num_tris > (int64_t(1) << 30)) return set_err(VR_ERR_ARG, "vr_scene_build_bvh: more than 2^30 triangles");
    if (!num_nodes) return set_err(VR_ERR_ARG, "vr_scene_build_bvh: num_nodes is NULL");
    *num_nodes = 0;
    if (num_tris == 0) return VR_OK;
    if (!positions || !indices || !nodes || !tri_ids || !tri_verts) return set_err(VR_ERR_ARG, "vr_scene_build_bvh: NULL buffer");
    if (max_nodes < 1) return set_err(VR_ERR_ARG, "vr_scene_build_bvh: node buffer too small");
    BvhBuilder B;
    B.tri_box.resize((size_t)num_tris);
    B.cent.resize(3 * (size_t)num_tris);
    B.order.resize((size_t)num_tris);
    for (int64_t t = 0; t < num_tris; t++) {
        BvhBuilder::Box& bx = B.tri_box[(size_t)t];
        bx.reset();
        for (int c = 0; c < 3; c++) {
            const int32_t v = indices[3 * t + c];
            if (v < 0 || v >= num_vertices) return set_err(VR_ERR_ARG, "vr_scene_build_bvh: vertex index out of range");
            bx.grow(positions + 3 * (int64_t)v);
        }
        for (int a = 0; a < 3; a++) {
            if (!std::isfinite(bx.lo[a]) || !std::isfinite(bx.hi[a])) return set_err(VR_ERR_ARG, "vr_scene_build_bvh: non-finite position");
            B.cent[3 * (size_t)t + a] = 0.5f * (bx.lo[a] + bx.hi[a]);
        }
        B.order[(size_t)t] = (int32_t)t;
    }
    B.nodes = nodes;
    B.max_nodes = max_nodes;
    B.used = 1;
    int st = B.build(0, 0, num_tris, 0);
    if (st) return st;
    for (int64_t k = 0; k < num_tris; k++) {
        const int32_t t = B.order[(size_t)k];
        tri_ids[k] = t;
        for (int c = 0; c < 3; c++) {
            const float* p = positions + 3 * (int64_t)indices[3 * (int64_t)t + c];
            tri_verts[9 * k + 3 * c] = p[0];
            tri_verts[9 * k + 3 * c + 1] = p[1];
            tri_verts[9 * k + 3 * c + 2] = p[2];
        }
    }
    *num_nodes = (int32_t)B.used;
    return VR_OK;
}

int vr_scene_raycast(const vr_scene* scene, const float* rays, int64_t n, float* hits, void* stream) {
    if (n < 0) return set_err(VR_ERR_ARG, "vr_scene_raycast: negative n");
    SceneView S;
    int st = make_scene_view(scene, VR_BW_CONSISTENT, S);
    if (st) return st;
    if (n == 0) return VR_OK;
    if (!rays || !hits) return set_err(VR_ERR_ARG, "vr_scene_raycast: NULL buffer");
    hipLaunchKernelGGL(k_scene_raycast, dim3((unsigned)((n + kSceneBlock - 1) / kSceneBlock)), dim3(kSceneBlock), 0,
                       (hipStream_t)stream, S, rays, n, hits);
    return check_hip(hipGetLastError(), "vr_scene_raycast launch");
}

int vr_scene_map_fw(const vr_scene* scene, const float* rays, int64_t n, float* out, const vr_launch* L, void* stream) {
    if (!L) return set_err(VR_ERR_ARG, "vr_scene_map_fw: launch is NULL");
    if (n < 0) return set_err(VR_ERR_ARG, "vr_scene_map_fw: negative n");
    SceneView S;
    int st = make_scene_view(scene, L->bw_mode, S);
    if (st) return st;
    int64_t start, cnt;
    if ((st = resolve_range(L, n, start, cnt))) return st;
    if (cnt == 0) return VR_OK;
    if (!rays || !out) return set_err(VR_ERR_ARG, "vr_scene_map_fw: NULL buffer");
    hipLaunchKernelGGL(k_scene_map_fw, dim3((unsigned)((cnt + kSceneBlock - 1) / kSceneBlock)), dim3(kSceneBlock), 0,
                       (hipStream_t)stream, S, rays, start, cnt, out, make_seeds(L), L->debug_out);
    return check_hip(hipGetLastError(), "vr_scene_map_fw launch");
}

int vr_scene_map_bw(const vr_scene* scene, const float* rays, const float* out_grad, int64_t n, const vr_launch* L,
                    void* stream) {
    if (!L) return set_err(VR_ERR_ARG, "vr_scene_map_bw: launch is NULL");
    if (n < 0) return set_err(VR_ERR_ARG, "vr_scene_map_bw: negative n");
    SceneView S;
    int st = make_scene_view(scene, L->bw_mode, S);
    if (st) return st;
    int64_t start, cnt;
    if ((st = resolve_range(L, n, start, cnt))) return st;
    if (cnt == 0) return VR_OK;
    if (!rays || !out_grad) return set_err(VR_ERR_ARG, "vr_scene_map_bw: NULL buffer");
    hipLaunchKernelGGL(k_scene_map_bw, dim3((unsigned)((cnt + kSceneBlock - 1) / kSceneBlock)), dim3(kSceneBlock), 0,
                       (hipStream_t)stream, S, rays, out_grad, start, cnt, make_seeds(L));
    return check_hip(hipGetLastError(), "vr_scene_map_bw launch");
}

int vr_scene_capture_fw(const vr_scene* scene, const vr_camera* camera, const vr_launch* L, float* out, void* stream) {
    if (!L) return set_err(VR_ERR_ARG, "vr_scene_capture_fw: launch is NULL");
    SceneView S;
    int st = make_scene_view(scene, L->bw_mode, S);
    if (st) return st;
    CameraView C;
    if ((st = make_camera_view(camera, C))) return st;
    const int64_t total = camera_sensors(camera);
    int64_t start, cnt;
    if ((st = resolve_range(L, total, start, cnt))) return st;
    if (cnt == 0) return VR_OK;
    if (!out) return set_err(VR_ERR_ARG, "vr_scene_capture_fw: out is NULL");
    hipStream_t s = (hipStream_t)stream;
    const int spp = L->samples < 1 ? 1 : L->samples;
    if ((st = check_sample_indices(total, spp, VR_SEED_CAPTURE_FW, "vr_scene_capture_fw"))) return st;
    CaptureSplit P = capture_split(spp, false);
    capture_tiling(P, C, start, cnt, "VR_SCENE_TILED", 2);
    if ((st = capture_prepare(P, out, 3, start, cnt, L->debug_out, s))) return st;
    const int64_t threads = cnt * P.chunks;
    hipLaunchKernelGGL(k_scene_capture_fw, dim3((unsigned)((threads + kSceneBlock - 1) / kSceneBlock)), dim3(kSceneBlock), 0, s,
                       S, C, make_seeds(L), P, total, start, cnt, out, L->debug_out);
    return check_hip(hipGetLastError(), "vr_scene_capture_fw launch");
}

int vr_scene_capture_bw(const vr_scene* scene, const vr_camera* camera, const vr_launch* L, const float* out_grad,
                        void* stream) {
    if (!L) return set_err(VR_ERR_ARG, "vr_scene_capture_bw: launch is NULL");
    SceneView S;
    int st = make_scene_view(scene, L->bw_mode, S);
    if (st) return st;
    CameraView C;
    if ((st = make_camera_view(camera, C))) return st;
    const int64_t total = camera_sensors(camera);
    int64_t start, cnt;
    if ((st = resolve_range(L, total, start, cnt))) return st;
    if (cnt == 0) return VR_OK;
    if (!out_grad) return set_err(VR_ERR_ARG, "vr_scene_capture_bw: out_grad is NULL");
    hipStream_t s = (hipStream_t)stream;
    const int spp = L->samples < 1 ? 1 : L->samples;
    if ((st = check_sample_indices(total, spp, L->seed_mode, "vr_scene_capture_bw"))) return st;
    CaptureSplit P = capture_split(spp, L->seed_mode == VR_SEED_CAPTURE_BW);
    capture_tiling(P, C, start, cnt, "VR_SCENE_TILED", 2);
    P.debug_lookups = L->debug_kind == 1;
    if ((st = capture_prepare(P, nullptr, 0, start, cnt, L->debug_out, s))) return st;
    const int64_t threads = cnt * P.chunks;
    hipLaunchKernelGGL(k_scene_capture_bw, dim3((unsigned)((threads + kSceneBlock - 1) / kSceneBlock)), dim3(kSceneBlock), 0, s,
                       S, C, make_seeds(L), P, L->seed_mode, total, start, cnt, out_grad, L->debug_out);
    return check_hip(hipGetLastError(), "vr_scene_capture_bw launch");
}

}  // extern "C"
