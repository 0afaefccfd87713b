/*
 * vr_api.h -- C ABI of the MI355X (gfx950) participating-media library
 * (librendervous_amd.so) and of its CPU oracle (oracle/liboracle.so, whose
 * functions carry the prefix "oracle_" instead of "vr_" and take host pointers).
 *
 * Every entry point is extern "C", takes plain pointers and sizes, and returns
 * an int status (VR_OK == 0).  The caller owns every buffer; the library never
 * allocates or frees caller memory.  All device work is ordered on the caller's
 * HIP stream (`stream`, a hipStream_t; NULL = the default stream) and no entry
 * point synchronises the device.
 *
 * Limits: a capture draws the random stream of sample s of sensor i from the
 * 32-bit flat index s * sensors + i, so spp * sensors must not exceed 2^32
 * (VR_ERR_UNSUPPORTED otherwise: the streams would repeat).  Constant majorant
 * values must be finite and >= 0, majorant distances > 0.
 *
 * Reference interfaces replaced (rendervous, files relative to its root):
 *   - the vulky-dispatched generated GLSL kernels of DispatcherEngine
 *       eval_capture_forward   _maps.py:749-821   -> vr_capture_fw
 *       eval_capture_backward  _maps.py:824-882   -> vr_capture_bw
 *       eval_map_forward       _maps.py:885-925   -> vr_map_fw, vr_grid_fw
 *       eval_map_backward      _maps.py:928-976   -> vr_map_bw, vr_grid_bw
 *   - the Parameter tensor view + param_buffer   include/common.h:88-94,110-179 -> vr_tensor
 *   - the 6->2 majorant map contract             include/vr_majorant.h:1-7    -> vr_majorant_build + VR_MAJ_GRID
 * Entry points without a reference counterpart are marked as such below.
 */
#ifndef VR_API_H
#define VR_API_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VR_ABI_VERSION 3

/* ---- status codes ------------------------------------------------------ */
enum vr_status {
    VR_OK = 0,
    VR_ERR_ARG = 1,          /* invalid argument (shape, stride, null pointer, range) */
    VR_ERR_UNSUPPORTED = 2,  /* map kind / option combination not built */
    VR_ERR_DEVICE = 3        /* HIP runtime error (message in vr_last_error) */
};

/* ---- enums --------------------------------------------------------------- */
enum vr_field_kind {         /* scattering albedo source */
    VR_FIELD_CONST = 0,      /* const[r,g,b]                       (_maps.py:1957-1985) */
    VR_FIELD_GRID = 1        /* Grid3D with C == 3 (or C == 1)     (_maps.py:2319-2365) */
};

enum vr_phase_kind {
    VR_PHASE_HG_MIX = 0,     /* HGPhase eval 0.8 HG(g)+0.2 HG(g/5) and VHGPhaseSampler
                                (g/5 with p=0.8), _maps.py:4436-4484 */
    VR_PHASE_HG = 1,         /* single Henyey-Greenstein for eval and sampling
                                (_scenes.py:1969-2004) */
    VR_PHASE_ISO = 2         /* isotropic 1/(4 pi) */
};

enum vr_majorant_kind {
    VR_MAJ_CONST = 0,        /* const[value, distance] (_objects.py:92-114) */
    VR_MAJ_GRID = 1          /* coarse max grid built by vr_majorant_build, traversed by DDA */
};

enum vr_transmittance_kind { /* estimator behind transmittance() in NEE / single scatter */
    VR_T_DT = 0,             /* include/vr_transmittance_DT.h */
    VR_T_RT = 1,             /* include/vr_transmittance_RT.h */
    VR_T_DDA = 2,            /* include/vr_transmittance_DDA.h */
    VR_T_GDT = 3,            /* include/vr_transmittance_GDT.h (no backward in the reference) */
    VR_T_GRT = 4             /* include/vr_transmittance_GRT.h (no backward in the reference) */
};

enum vr_env_kind {           /* environment(x, w) */
    VR_ENV_CONST = 0,        /* const[r,g,b] (env_c) */
    VR_ENV_XR = 1            /* Image2D(env_image).after(xr_projection): maps/environment_xr.h,
                                _objects.py:164-173 (no gradient into the image) */
};

enum vr_env_sampler_kind {   /* the environment part of environment_sampler(x, w) */
    VR_ENVS_UNIFORM = 0,     /* _functionsampler(UniformDirectionSampler, environment) */
    VR_ENVS_QUADTREE = 1     /* _functionsampler(XRQuadtreeRandomDirection, environment),
                                _maps.py:2494-2549, _objects.py:180-199 */
};

enum vr_map_kind {           /* input -> output dims */
    VR_MAP_RAYBOX = 1,       /* 6 -> 2   RayBoxIntersection          _maps.py:4128-4154 */
    VR_MAP_T_DDA = 2,        /* 6 -> 1   GridDDATransmittance        _maps.py:2640-2667 */
    VR_MAP_T_DT = 3,         /* 6 -> 1   DeltatrackingTransmittance  _maps.py:2685-2697 */
    VR_MAP_T_RT = 4,         /* 6 -> 1   RatiotrackingTransmittance  _maps.py:2670-2682 */
    VR_MAP_COLLISION_DT = 5, /* 6 -> 3   DeltatrackingCollisionSampler (C, t, T) _maps.py:2714-2727 */
    VR_MAP_SS_RADIANCE = 6,  /* 6 -> 3   SingleScatteredRadiance     _maps.py:2761-2784 */
    VR_MAP_MC_COLLISION = 7, /* 6 -> 3   MCCollisionIntegrator(collision DT, exitance =
                                         SingleScatteredRadiance, environment) _maps.py:2788-2808 */
    VR_MAP_NEE_DT = 8,       /* 6 -> 3   DeltatrackingNEEPathIntegrator _maps.py:2924-2978 */
    VR_MAP_MEDIUM_DT = 9,    /* 7 -> 12  DTMediumPathIntegrator      _scenes.py:1614-1632 */
    VR_MAP_PATH_DT = 10,     /* 6 -> 3   DeltatrackingPathIntegrator (no NEE; defensive replay bw)
                                         _maps.py:2886-2921, maps/radiance_DT.h */
    VR_MAP_NEE_DRT = 11,     /* 6 -> 3   DRTPathIntegrator: NEE forward, reservoir backward
                                         _maps.py:2981-3022, maps/radiance_NEE_DRT.h */
    VR_MAP_T_GDT = 12,       /* 6 -> 1   GridDeltatrackingTransmittance _maps.py:2619-2637 */
    VR_MAP_T_GRT = 13        /* 6 -> 1   GridRatiotrackingTransmittance _maps.py:2588-2616 */
};

enum vr_seed_mode {          /* per-thread RNG state derivation from the 4 dispatch seeds */
    VR_SEED_MAP = 0,         /* map eval kernels            _maps.py:390-393,465-469 */
    VR_SEED_CAPTURE_FW = 1,  /* capture forward kernel      _maps.py:549-560 */
    VR_SEED_CAPTURE_BW = 2   /* capture backward kernel     _maps.py:676-685 */
};

enum vr_bw_mode {
    VR_BW_REFERENCE = 0,     /* the reference's backward estimators, quirks included */
    VR_BW_CONSISTENT = 1     /* same estimators with the primal/replay majorants matched and the
                                single-scatter transmittance gradient taken along the sampled ray */
};

/* ---- data structures ------------------------------------------------------ */

/* Strided FP32 tensor view (mirrors `Parameter`, include/common.h:88-94).
 * Index order is the torch order [D,H,W,C]; strides are in elements.
 * grad == NULL means "no gradient requested" (include/maps/grid3d.h:41). */
typedef struct vr_tensor {
    float*  data;
    float*  grad;
    int64_t shape[4];
    int64_t stride[4];
    int32_t ndim;
    int32_t reserved;
} vr_tensor;

/* Grid3D map: vertex-centred trilinear field over the box [bmin, bmax]
 * (include/maps/grid3d.h:6-75). */
typedef struct vr_grid {
    vr_tensor t;             /* [D,H,W,C] */
    float bmin[3];           /* x, y, z */
    float bmax[3];
} vr_grid;

/* A heterogeneous medium: the lowered form of the reference's map graph for
 * sigma / boundary / majorant / albedo / emission / environment / phase. */
typedef struct vr_medium {
    vr_grid sigma;                 /* C == 1 */
    float   box_min[3];            /* boundary = RayBoxIntersection(box_min, box_max) */
    float   box_max[3];
    int32_t albedo_kind;           /* vr_field_kind */
    float   albedo_c[3];
    vr_grid albedo;                /* used when albedo_kind == VR_FIELD_GRID */
    float   emission_c[3];         /* constant emission (ZERO in the recipes) */
    float   env_c[3];              /* constant environment radiance */
    float   sun_dir[3];            /* directional light: unit vector pointing TO the light */
    float   sun_radiance[3];
    float   sun_prob;              /* 0 = no sun; else mixture selection probability */
    int32_t phase_kind;            /* vr_phase_kind */
    float   phase_g;
    int32_t majorant_kind;         /* vr_majorant_kind */
    float   majorant_value;        /* VR_MAJ_CONST: [value, distance] */
    float   majorant_distance;
    const float* majorant_grid;    /* VR_MAJ_GRID: [rz, ry, rx] row-major (x fastest) */
    int32_t majorant_res[3];       /* rz, ry, rx */
    int32_t transmittance_kind;    /* vr_transmittance_kind */
    float   ds_epsilon;            /* defensive-sampling epsilon of the backward estimators */
    int32_t max_bounces;           /* < 0: unbounded (reference); else scatter events per path */
    int32_t medium_filter;         /* DTMediumPathIntegrator MEDIUM_FILTER: bit0 scatter, bit1 gather */
    float   gathering_c[3];        /* constant gathering map of the 7->12 integrator */
    float   medium_majorant;       /* scalar majorant of the 7->12 integrator */
    int32_t env_kind;              /* vr_env_kind */
    int32_t env_sampler_kind;      /* vr_env_sampler_kind */
    vr_tensor env_image;           /* VR_ENV_XR: [H, W, 3] (maps/image2d.h) */
    float   env_image_bmin[2];     /* Image2D bmin, bmax over xr coordinates (-1, 1 in the recipes) */
    float   env_image_bmax[2];
    const float* env_quadtree;     /* VR_ENVS_QUADTREE: functions.create_density_quadtree output */
    int32_t env_quadtree_levels;   /* 1 .. 15 */
    int32_t env_reserved;
    const float* sigma_cells;      /* optional cell table of sigma (vr_grid_pack_cells); NULL = read
                                      the grid.  Lookups give bit-identical values either way.
                                      Requires (nx - 1) * (ny - 1) < 2^24 (VR_ERR_ARG otherwise). */
    int32_t phase_g_kind;          /* vr_field_kind: VR_FIELD_CONST = phase_g everywhere;
                                      VR_FIELD_GRID = phase_g_grid read at each scattering
                                      position (HGPhase / VHGPhaseSampler evaluate their phase_g
                                      map there, _maps.py:4449,4473; no gradient flows into it,
                                      the reference's HGPhase has no backward) */
    int32_t phase_reserved;
    vr_grid phase_g_grid;          /* C == 1, used when phase_g_kind == VR_FIELD_GRID */
} vr_medium;

/* PerspectiveCameraSensor (_maps.py:1765-1786, include/maps/sensor_perspective_camera.h). */
typedef struct vr_camera {
    const float*   poses;          /* [num_poses, 9] = origin, direction, up */
    int32_t        num_poses;
    int32_t        width;
    int32_t        height;
    int32_t        jittered;       /* generation_mode: 0 pixel centre, 1 two randoms */
    float          fov;
    float          znear;
    const int64_t* sensors;        /* optional [num_sensors, 3] (pose, row, col); NULL = all */
    int64_t        num_sensors;
} vr_camera;

/* Per-dispatch settings (the uniform block of the generated kernels, _maps.py:530-542). */
typedef struct vr_launch {
    uint32_t seeds[4];             /* dispatch seeds, each in [129, 2^30) (_internal.py:26-34) */
    int32_t  samples;              /* spp (capture); ignored by maps */
    int32_t  bw_mode;              /* vr_bw_mode */
    int32_t  seed_mode;            /* vr_seed_mode used by vr_capture_bw (fw uses CAPTURE_FW) */
    int32_t  debug_kind;           /* what debug_out counts in a backward: 0 = gradient scatters,
                                      1 = trilinear lookups (forward entry points: lookups) */
    int64_t  start_index;          /* first thread index of this batch (batch_size chunking) */
    int64_t  count;                /* threads in this batch; < 0 = all remaining */
    int32_t* debug_out;            /* optional per-thread counter (the reference's DEBUG_COUNTER
                                      slot, _maps.py:368,611-615), see debug_kind */
    void*    workspace;            /* optional device scratch for vr_capture_fw (the work-queue
                                      counters of the persistent NEE kernel): at least
                                      VR_CAPTURE_WORKSPACE_BYTES, 256-byte aligned, used in stream
                                      order (so captures on different streams never share it).
                                      NULL = the library allocates it for the call on `stream`
                                      (hipMallocAsync / hipFreeAsync). */
    int64_t  workspace_bytes;
} vr_launch;

/* Size of vr_launch.workspace (8 queue counters, one 128-byte line each). */
#define VR_CAPTURE_WORKSPACE_BYTES 1024

/* ---- scene-level path tracing -------------------------------------------------
 * PathtracedScene (_scenes.py:2251-2312, include/maps/pathtraced_scene.h,
 * include/sc_path_tracing.h:12-236): a path tracer over triangle-mesh patches
 * whose patches carry a surface scattering sampler, a surface emission and the
 * media on their inner side (DTMediumPathIntegrator segments between surface
 * hits).  The reference traces rays with Vulkan ray queries against a
 * driver-built acceleration structure (maps/geometry_mesh.h); here a
 * binned-SAH BVH2 built by vr_scene_build_bvh is traversed by the kernels. */

enum vr_surface_sampler {    /* patch_info.surface_scattering_sampler */
    VR_SURF_PASS = 0,        /* no sampler map: the path continues unchanged (sc_path_tracing.h:78-82) */
    VR_SURF_LAMBERT = 1,     /* LambertSurfaceScatteringSampler   _scenes.py:906-937 */
    VR_SURF_MIRROR = 2,      /* MirrorSurfaceScatteringSampler    _scenes.py:1003-1022 */
    VR_SURF_FRESNEL = 3,     /* FresnelSurfaceScatteringSampler   _scenes.py:964-1000 */
    VR_SURF_NULL = 4,        /* NullSurfaceScatteringSampler      _scenes.py:1308-1322 */
    VR_SURF_BLOCK = 5        /* NoSurfaceScatteringSampler (zero) _scenes.py:1325-1340 */
};

enum vr_surface_emission {   /* patch_info.surface_gathering = the material's emission */
    VR_EMIT_NONE = 0,
    VR_EMIT_ISOTROPIC = 1,   /* IsotropicSurfaceEmission(const): E * 0.5/pi   _scenes.py:1392-1412 */
    VR_EMIT_LAMBERT = 2      /* LambertSurfaceEmission(const): E * dot(-w, fN)/pi  _scenes.py:1415-1435 */
};

/* One BVH2 node (32 bytes).  Interior: `first` = index of the left child, the
 * right child is first + 1, count == 0.  Leaf: triangles [first, first + count)
 * of the BVH triangle order.  Node 0 is the root. */
typedef struct vr_bvh_node {
    float   bmin[3];
    int32_t first;
    float   bmax[3];
    int32_t count;
} vr_bvh_node;

/* Per-patch information (PTPatchInfo, _scenes.py:2156-2162). */
typedef struct vr_patch {
    int32_t sampler;               /* vr_surface_sampler */
    float   refraction_index;      /* VR_SURF_FRESNEL */
    int32_t emission;              /* vr_surface_emission */
    float   emission_c[3];         /* constant emission map value */
    int32_t inside_medium;         /* index into vr_scene.media; -1 = none */
    int32_t outside_medium;        /* kept as in the reference; the path tracer does not read it */
} vr_patch;

/* A scene: geometry (device buffers), patches (device), media (host array of
 * DTMediumPathIntegrator descriptors, at most VR_SCENE_MAX_MEDIA) and the
 * environment (env_kind / env_c / env_image* of vr_medium). */
#define VR_SCENE_MAX_MEDIA 4
typedef struct vr_scene {
    const vr_bvh_node* nodes;      /* [num_nodes] */
    const float*   tri_verts;      /* [num_tris, 9]: vertices of each triangle in BVH order */
    const int32_t* tri_ids;        /* [num_tris]: triangle index of each BVH-ordered triangle */
    int32_t        num_nodes;
    int32_t        num_tris;
    const float*   positions;      /* [num_vertices, 3] */
    const float*   normals;        /* [num_vertices, 3] or NULL (face normals) */
    const float*   uvs;            /* [num_vertices, 2] or NULL (C = 0) */
    const int32_t* indices;        /* [num_tris, 3] vertex indices */
    const int32_t* tri_patch;      /* [num_tris] patch index of each triangle */
    const vr_patch* patches;       /* [num_patches] */
    int32_t        num_patches;
    int32_t        num_vertices;
    const vr_medium* media;        /* host array [num_media]; medium_filter / medium_majorant /
                                      gathering_c / sigma / albedo / phase used */
    int32_t        num_media;
    int32_t        env_kind;       /* vr_env_kind */
    float          env_c[3];
    float          env_image_bmin[2];
    float          env_image_bmax[2];
    vr_tensor      env_image;      /* VR_ENV_XR: [H, W, 3] */
} vr_scene;

/* ---- entry points ---------------------------------------------------------- */

/* ABI version of the loaded library (VR_ABI_VERSION). No reference counterpart. */
int vr_abi_version(void);

/* Thread-local message for the last non-OK status. No reference counterpart
 * (the reference reports errors with Python asserts, _maps.py:760,902-903). */
const char* vr_last_error(void);

/* Grid3D forward on points x[n,3] -> out[n,C]  (include/maps/grid3d.h:6-37;
 * dispatched by eval_map_forward, _maps.py:885-925). */
int vr_grid_fw(const vr_grid* grid, const float* x, int64_t n, float* out, void* stream);

/* Grid3D backward: accumulates out_grad[n,C] trilinearly into grid->t.grad
 * (include/maps/grid3d.h:39-75; eval_map_backward, _maps.py:928-976). */
int vr_grid_bw(const vr_grid* grid, const float* x, const float* out_grad, int64_t n, void* stream);

/* Coarse majorant grid: out[rz,ry,rx] = max of sigma over the vertices whose
 * trilinear footprint touches each coarse cell.  Implements the 6->2 majorant
 * contract (include/vr_majorant.h:1-7) that the reference fills with
 * const[max(sigma), 1e6] (_objects.py:96-101).  A cell whose footprint holds a
 * NaN or infinite density gets a NaN bound, which every tracker treats as an
 * empty cell (an infinite bound would give zero-length free flights that never
 * end a ray); such densities are the caller's error, not reported here. */
int vr_majorant_build(const vr_grid* sigma, const int32_t res[3], float* out, void* stream);

/* Map forward on rays: in[n,IN] -> out[n,OUT] for the map `kind`
 * (eval_map_forward, _maps.py:885-925; per-thread seeds _maps.py:390-393). */
int vr_map_fw(const vr_medium* medium, int32_t kind, const float* in, int64_t n, float* out,
              const vr_launch* launch, void* stream);

/* Map backward: accumulates parameter gradients (sigma.grad, albedo.grad) for
 * out_grad[n,OUT] (eval_map_backward, _maps.py:928-976). */
int vr_map_bw(const vr_medium* medium, int32_t kind, const float* in, const float* out_grad, int64_t n,
              const vr_launch* launch, void* stream);

/* Capture forward: out[cams,H,W,OUT] (or [num_sensors,OUT]) = spp-mean of the
 * field `kind` over camera rays (eval_capture_forward, _maps.py:749-821). */
int vr_capture_fw(const vr_medium* medium, int32_t kind, const vr_camera* camera,
                  const vr_launch* launch, float* out, void* stream);

/* vr_capture_fw that allocates nothing.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  The persistent NEE kernel's
 * work counters (the only scratch a capture uses: VR_CAPTURE_WORKSPACE_BYTES for
 * at most 8 work queues, whatever the capture's size) live in a file-scope device
 * array of the library, one per GPU, held until the process ends;
 * launch->workspace is not used.  Calls on one GPU share that array: calls that
 * overlap on one GPU (concurrent streams or threads) race and must be
 * serialised by the caller.  Each call clears the counters before use.  Same
 * arguments, limits and results as vr_capture_fw. */
int vr_capture_fw_noalloc(const vr_medium* medium, int32_t kind, const vr_camera* camera,
                          const vr_launch* launch, float* out, void* stream);

/* Capture backward: accumulates into sigma.grad / albedo.grad the gradient of
 * sum(out_grad * out) (eval_capture_backward, _maps.py:824-882). */
int vr_capture_bw(const vr_medium* medium, int32_t kind, const vr_camera* camera,
                  const vr_launch* launch, const float* out_grad, void* stream);

/* Fused Adam step over n parameters (torch.optim.Adam semantics, bias-corrected,
 * `step` >= 1); zero_grad != 0 also clears g.  No reference counterpart: the
 * reference leaves optimisation to torch. */
int vr_adam_step(float* p, float* g, float* m, float* v, int64_t n, float lr, float beta1, float beta2,
                 float eps, int32_t step, int32_t zero_grad, void* stream);

/* vr_adam_step followed by the projection p = min(max(p, lo), hi) in the same
 * pass (densities >= 0, albedos in [0, 1]); lo <= hi.  No reference
 * counterpart, as vr_adam_step. */
int vr_adam_step_bounded(float* p, float* g, float* m, float* v, int64_t n, float lr, float beta1, float beta2,
                         float eps, int32_t step, float lo, float hi, int32_t zero_grad, void* stream);

/* ---- grid tools of the inverse-rendering loop --------------------------------
 * Dense row-major float32 tensors [d0, .., d(dim-1), C] (C channels innermost);
 * shapes are host arrays. */

/* Replaces functions.random_ids (_functions.py:114-142, functions/tools/
 * random_ids.comp.glsl), which SensorsBase.random_sensors (_maps.py:1751-1762)
 * calls: out[n][dim] int64 with out[i][k] = int(u * shape[k]), u the k-th
 * uniform of thread i seeded as the reference's compute functions
 * (functions/common_functions.h:27-49) from the dispatch seeds[4].  1 <= dim <= 4. */
int vr_random_ids(int64_t* out, int64_t n, const int64_t* shape, int32_t dim, const uint32_t seeds[4], void* stream);

/* Replaces the total_variation forward (_functions.py:352-380, functions/
 * regularizers/total_variation.comp.glsl): shape[dim + 1] = (spatial dims..., C),
 * 1 <= dim <= 3; out [spatial dims..., 1] = per-voxel sum over the existing
 * +-1 neighbours of sum_c |v - v_adj|. */
int vr_total_variation_fw(const float* in, const int64_t* shape, int32_t dim, float* out, void* stream);

/* Replaces the total_variation backward (_functions.py:383-413, total_variation_
 * backward.comp.glsl): in_grad [shape] = sum over neighbours j of
 * (out_grad[i] + out_grad[j]) * sign(v[i] - v[j]); in_grad is overwritten. */
int vr_total_variation_bw(const float* in, const float* out_grad, const int64_t* shape, int32_t dim, float* in_grad,
                          void* stream);

/* One linear resampling step of functions.resample_grid (_functions.py:252-297,
 * functions/tools/resample_grid.comp.glsl): in [in_shape..., C] -> out
 * [out_shape..., C], source coordinate p * (in - 1) / (out - 1) per axis
 * (vertex-aligned), 2^dim taps with the +1 tap clamped.  1 <= dim <= 3, every
 * out_shape entry >= 2 (the shader divides by out - 1). */
int vr_resample_grid(const float* in, const int64_t* in_shape, const int64_t* out_shape, int32_t dim, int32_t channels,
                     float* out, void* stream);

/* One pixel-centred linear resampling step of functions.resample_img
 * (_functions.py:300-345, functions/tools/resample_img.comp.glsl): source
 * coordinate (p + 1/2) / out * in - 1/2 per axis, both taps clamped to the
 * extent.  1 <= dim <= 3; every out_shape entry >= 1. */
int vr_resample_img(const float* in, const int64_t* in_shape, const int64_t* out_shape, int32_t dim, int32_t channels,
                    float* out, void* stream);

/* functions.copy_img_to_morton / copy_morton_to_img (_functions.py:434-506,
 * functions/tools/copy_*_morton.comp.glsl): a square [R, R, C] image <-> [R*R, C]
 * in Morton (Z) order, pixel (px, py) = (even bits, odd bits) of the index.
 * R a power of two up to 2^15. */
int vr_copy_img_to_morton(const float* in, int64_t resolution, int32_t channels, float* out, void* stream);
int vr_copy_morton_to_img(const float* in, int64_t resolution, int32_t channels, float* out, void* stream);

/* ---- neural radiance cache ---------------------------------------------------
 * No reference counterpart (the north star's cached renderer): a 96-64-64-64-64-3
 * ReLU MLP in BF16 with FP32 accumulation.  Q(x, w) = radiance a NEE path collects
 * from a real collision at x, reached along w, onwards.  Weights: VR_NRC_PARAMS
 * values, per layer Wt[n][k] (output n, input k) for layers 96->64, 64->64 (x3),
 * 64->16 (outputs 3..15 unused); `params` FP32 master copy, `weights` BF16 bits
 * (16-byte aligned). */
#define VR_NRC_PARAMS 19456
#define VR_NRC_FEATURES 96
/* Copies of the weight gradient a training step accumulates into. */
#define VR_NRC_GRAD_COPIES 16
#define VR_NRC_LOSS_SLOTS 64

/* Input features of n points x and directions w ([n][3] each): sin/cos of
 * pi 2^f u_c (f < 12, u = x mapped from the box to [0,1]^3), w, 1, zeros. */
int vr_nrc_encode(const float* x, const float* w, int64_t n, const float box_min[3], const float box_max[3],
                  float* features, void* stream);

/* MLP on explicit features [n][96] (rounded to BF16; 16-byte aligned) -> out [n][3]. */
int vr_nrc_mlp_fw(const float* features, int64_t n, const uint16_t* weights, float* out, void* stream);

/* One training step on the records (x[3], w[3], target[3]) of rows
 * [f, min(*count, f + max_records)) of `records`, f = first % *count (*count on
 * the device: the rows written), so that windows first = k * max_records cycle
 * through the records: relative L2 loss, its gradient accumulated in `grad`
 * (VR_NRC_GRAD_COPIES * VR_NRC_PARAMS floats, overwritten; on return the first
 * VR_NRC_PARAMS hold the step's gradient), Adam step `step` (>= 1) on
 * `params`/`m`/`v`, BF16 copy into `weights`; `loss` holds VR_NRC_LOSS_SLOTS
 * floats (device, overwritten), the step's mean loss in loss[0].  lr = 0
 * leaves the weights. */
int vr_nrc_train_step(const float* records, const uint32_t* count, int64_t first, int64_t max_records, const float box_min[3],
                      const float box_max[3], float* params, uint16_t* weights, float* grad, float* m, float* v,
                      int32_t step, float lr, float beta1, float beta2, float eps, float* loss, void* stream);

/* Training samples: n_paths NEE paths from uniformly drawn sensors (seeded by
 * launch->seeds), each collision vertex (at most max_vertices per path) giving a
 * record (x, w, suffix radiance / throughput); *count (device) receives the
 * number written (capped at max_records).  scratch: n_paths*max_vertices*12 floats. */
int vr_nrc_paths(const vr_medium* medium, const vr_camera* camera, const vr_launch* launch, int32_t n_paths,
                 int32_t max_vertices, float* scratch, float* records, int64_t max_records, uint32_t* count,
                 void* stream);

/* vr_nrc_paths for `traces` (1..16) independent traces in one launch: trace t
 * draws its paths from launches[t].seeds and writes records
 * [t * max_records, (t + 1) * max_records) of `records` (9 floats each) and
 * counts[t]; scratch: traces * n_paths * max_vertices * 12 floats.  The paths do
 * not read the cache, so this equals `traces` calls of vr_nrc_paths. */
int vr_nrc_paths_traces(const vr_medium* medium, const vr_camera* camera, const vr_launch* launches, int32_t traces,
                        int32_t n_paths, int32_t max_vertices, float* scratch, float* records, int64_t max_records,
                        uint32_t* counts, void* stream);

/* Capture of the NEE radiance (map kind VR_MAP_NEE_DT) whose paths stop at their
 * (terminate_after + 1)-th collision and add W * max(Q, 0) there: out
 * [sensors][3]; queries: scratch of max_queries * 10 floats, *count (device) the
 * queries issued. */
int vr_nrc_render(const vr_medium* medium, const vr_camera* camera, const vr_launch* launch, int32_t terminate_after,
                  const uint16_t* weights, float* out, float* queries, int64_t max_queries, uint32_t* count,
                  void* stream);

/* Cell table of a single-channel grid [D,H,W,1]: for each of the (D-1)(H-1)(W-1)
 * cells (x fastest) its 8 corner values in the order x, then y, then z fastest
 * (32 bytes per cell), so that a trilinear lookup reads one cell with two 16-byte
 * loads.  out: (D-1)(H-1)(W-1) * 8 floats, 16-byte aligned.  No reference
 * counterpart (a layout of the Grid3D tensor, include/maps/grid3d.h:6-37). */
int vr_grid_pack_cells(const vr_grid* grid, float* out, void* stream);

/* Builds the BVH of a triangle mesh on the host (binned SAH, at most 4
 * triangles per leaf, depth <= 30).  positions [num_vertices,3], indices
 * [num_tris,3] and every output are HOST buffers: nodes (2 * num_tris - 1
 * entries at most; *num_nodes receives the count), tri_ids [num_tris] and
 * tri_verts [num_tris,9] in BVH order.  Replaces the acceleration-structure
 * build of MeshGeometry (_scenes.py:470-490, vk.ads_model / ads_scene). */
int vr_scene_build_bvh(const float* positions, int64_t num_vertices, const int32_t* indices, int64_t num_tris,
                       vr_bvh_node* nodes, int64_t max_nodes, int32_t* tri_ids, float* tri_verts, int32_t* num_nodes);

/* Geometry map forward (6 -> 16, _scenes.py:93-100, maps/geometry_mesh.h:1-28):
 * rays [n,6] -> hits [n,16] = (t, N, G, C, T, B, patch index as int bits; -1 = no hit)
 * of the closest triangle with 0 <= t <= 10000 (common.h:297-333, 467-501). */
int vr_scene_raycast(const vr_scene* scene, const float* rays, int64_t n, float* hits, void* stream);

/* PathtracedScene on explicit rays (6 -> 3): eval_map_forward / eval_map_backward
 * (_maps.py:885-976) with the map per-thread seeds.  The backward accumulates
 * into the media's sigma / albedo gradients. */
int vr_scene_map_fw(const vr_scene* scene, const float* rays, int64_t n, float* out, const vr_launch* launch,
                    void* stream);
int vr_scene_map_bw(const vr_scene* scene, const float* rays, const float* out_grad, int64_t n,
                    const vr_launch* launch, void* stream);

/* Capture of a PathtracedScene: eval_capture_forward / _backward (_maps.py:749-882)
 * with map = PathtracedScene; out / out_grad [sensors][3]. */
int vr_scene_capture_fw(const vr_scene* scene, const vr_camera* camera, const vr_launch* launch, float* out,
                        void* stream);
int vr_scene_capture_bw(const vr_scene* scene, const vr_camera* camera, const vr_launch* launch,
                        const float* out_grad, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* VR_API_H */
