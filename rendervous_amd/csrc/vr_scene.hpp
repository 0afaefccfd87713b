// vr_scene.hpp -- scene-level path tracing: triangle meshes with surface
// scattering and emission, and heterogeneous media between surface hits.
//
// Follows rendervous include/sc_path_tracing.h:12-236 (path_trace,
// path_trace_fw, path_trace_bw), include/maps/pathtraced_scene.h (the map's
// FORWARD / BACKWARD), include/dyn_medium_traversal.h (7 -> 12 medium segments)
// and the surface maps of _scenes.py (samplers :882-1340, emission
// :1343-1454).  The reference casts rays with Vulkan ray queries against a
// driver-built acceleration structure (maps/geometry_mesh.h,
// maps/geometry_group.h); here every lane walks a BVH2 (vr_scene_build_bvh)
// with its traversal stack in LDS, one 32-entry column per lane, so the walk
// needs no scratch memory.
#pragma once

#include "vr_device.hpp"

namespace vr {

constexpr int kSceneStack = 32;             // >= BVH depth limit (30) + 1
constexpr float kSceneTMax = 10000.0f;      // rayQueryInitializeEXT tmax, maps/geometry_mesh.h:15
constexpr int kScenePathCap = 4096;         // loop iterations per path (the reference has no cap)

struct SceneView {
    const float4* nodes;      // 2 x float4 per node: (bmin, first), (bmax, count)
    const float* tri_verts;   // [T, 9] BVH order
    const int* tri_ids;       // [T]
    int num_nodes, num_tris;
    const float* pos;         // [V, 3]
    const float* nrm;         // [V, 3] or null
    const float* uv;          // [V, 2] or null
    const int* idx;           // [T, 3]
    const int* tri_patch;     // [T]
    const vr_patch* patches;  // [P]
    int num_patches;
    int num_media;
    int env_kind;
    f3 env_c;
    ImageView env_img;
    MediumView media[VR_SCENE_MAX_MEDIA];
};

// Surfel (include/common.h:277-285); T and B stay zero (MeshGeometry keeps no
// tangents, _scenes.py:436-437).
struct Surfel {
    f3 P, N, G;
    float cu, cv;
};

VR_DEV f3 scene_env(const SceneView& S, f3 w) {
    if (S.env_kind == VR_ENV_XR) {
        float u, v;
        dir2xr(w, u, v);
        return image_lookup(S.env_img, u, v);
    }
    return S.env_c;
}

// Slab test of a node box against [0, tmax]; returns the entry distance.
VR_DEV bool node_hit(float4 lo, float4 hi, f3 o, f3 inv, float tmax, float& tnear) {
    const float ax = (lo.x - o.x) * inv.x, bx = (hi.x - o.x) * inv.x;
    const float ay = (lo.y - o.y) * inv.y, by = (hi.y - o.y) * inv.y;
    const float az = (lo.z - o.z) * inv.z, bz = (hi.z - o.z) * inv.z;
    const float tn = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), 0.0f));
    const float tf = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fminf(fmaxf(az, bz), tmax));
    tnear = tn;
    return tn <= tf;
}

// Moller-Trumbore; barycentrics (u, v) weight vertices 1 and 2, as the ray
// query's committed barycentrics (maps/geometry_mesh.h:23).
VR_DEV bool tri_hit(const float* v, f3 o, f3 d, float tmax, float& t, float& bu, float& bv) {
    const f3 v0 = F3(v[0], v[1], v[2]);
    const f3 e1 = F3(v[3], v[4], v[5]) - v0, e2 = F3(v[6], v[7], v[8]) - v0;
    const f3 p = cross(d, e2);
    const float det = dot(e1, p);
    if (fabsf(det) < 1e-20f) return false;
    const float id = 1.0f / det;
    const f3 s = o - v0;
    const float u = dot(s, p) * id;
    if (u < 0.0f || u > 1.0f) return false;
    const f3 q = cross(s, e1);
    const float w = dot(d, q) * id;
    if (w < 0.0f || u + w > 1.0f) return false;
    const float tt = dot(e2, q) * id;
    if (tt < 0.0f || tt >= tmax) return false;
    t = tt; bu = u; bv = w;
    return true;
}

// Closest hit in [0, kSceneTMax]: BVH-order triangle index or -1.  `stk` is
// the lane's LDS column (stride STRIDE, kSceneStack entries).
template <int STRIDE>
VR_DEV int scene_trace(const SceneView& S, f3 o, f3 d, float& t, float& bu, float& bv, int* stk) {
    if (S.num_tris <= 0) return -1;
    const f3 dd = F3(fabsf(d.x) < 1e-20f ? copysignf(1e-20f, d.x) : d.x, fabsf(d.y) < 1e-20f ? copysignf(1e-20f, d.y) : d.y,
                     fabsf(d.z) < 1e-20f ? copysignf(1e-20f, d.z) : d.z);
    const f3 inv = F3(1.0f / dd.x, 1.0f / dd.y, 1.0f / dd.z);
    float best_t = kSceneTMax;
    int best = -1;
    float tn;
    if (!node_hit(S.nodes[0], S.nodes[1], o, inv, best_t, tn)) return -1;
    int node = 0, sp = 0;
    for (;;) {
        const float4 lo = S.nodes[2 * node], hi = S.nodes[2 * node + 1];
        const int first = __float_as_int(lo.w), count = __float_as_int(hi.w);
        bool descend = false;
        if (count > 0) {
            for (int k = first; k < first + count; k++) {
                float tt, u, v;
                if (tri_hit(S.tri_verts + 9 * (size_t)k, o, d, best_t, tt, u, v)) {
                    best_t = tt; best = k; bu = u; bv = v;
                }
            }
        } else {
            const int l = first, r = first + 1;
            float tl, tr;
            const bool hl = node_hit(S.nodes[2 * l], S.nodes[2 * l + 1], o, inv, best_t, tl);
            const bool hr = node_hit(S.nodes[2 * r], S.nodes[2 * r + 1], o, inv, best_t, tr);
            if (hl && hr) {
                // a tree of depth <= 30 (vr_scene_build_bvh) never fills the stack
                const bool lfirst = tl <= tr;
                if (sp < kSceneStack) {
                    stk[sp * STRIDE] = lfirst ? r : l;
                    sp++;
                }
                node = lfirst ? l : r;
                descend = true;
            } else if (hl || hr) {
                node = hl ? l : r;
                descend = true;
            }
        }
        if (descend) continue;
        if (sp == 0) break;
        sp--;
        node = stk[sp * STRIDE];
    }
    t = best_t;
    return best;
}

// sample_surfel (include/common.h:467-501) of BVH-order triangle k.
VR_DEV Surfel scene_surfel(const SceneView& S, int k, float bu, float bv, f3 x, f3 w, float t, int& patch) {
    const int tri = S.tri_ids[k];
    const int i0 = S.idx[3 * tri], i1 = S.idx[3 * tri + 1], i2 = S.idx[3 * tri + 2];
    const float a0 = 1.0f - bu - bv;
    const f3 p0 = load3(S.pos + 3 * i0), p1 = load3(S.pos + 3 * i1), p2 = load3(S.pos + 3 * i2);
    Surfel s;
    s.P = w * t + x;  // hit2surfel, common.h:303
    f3 g = cross(p1 - p0, p2 - p0);
    s.G = g * (1.0f / sqrtf(dot(g, g)));
    if (S.nrm) {
        f3 n = load3(S.nrm + 3 * i0) * a0 + load3(S.nrm + 3 * i1) * bu + load3(S.nrm + 3 * i2) * bv;
        s.N = n * (1.0f / sqrtf(dot(n, n)));
    } else s.N = s.G;
    if (S.uv) {
        s.cu = S.uv[2 * i0] * a0 + S.uv[2 * i1] * bu + S.uv[2 * i2] * bv;
        s.cv = S.uv[2 * i0 + 1] * a0 + S.uv[2 * i1 + 1] * bu + S.uv[2 * i2 + 1] * bv;
    } else s.cu = s.cv = 0.0f;
    // patch and medium indices are validated by the host layer; clamping keeps
    // a bad index from reading outside the patch table
    patch = min(max(S.tri_patch[tri], 0), S.num_patches - 1);
    return s;
}

// INCLUDE_ADDITIONAL_COLLISION_VARS (include/common.h:509-512).
VR_DEV void collision_vars(f3 win, const Surfel& s, bool& from_outside, f3& fN) {
    from_outside = dot(win, s.G) < 0.0f;
    const bool correct = (dot(win, s.N) < 0.0f) == from_outside;
    const f3 N = correct ? s.N : s.G;
    fN = from_outside ? N : N * -1.0f;
}

// GLSL reflect / refract.
VR_DEV f3 glsl_reflect(f3 i, f3 n) { return i - n * (2.0f * dot(n, i)); }
VR_DEV f3 glsl_refract(f3 i, f3 n, float eta) {
    const float c = dot(n, i);
    const float k = 1.0f - eta * eta * (1.0f - c * c);
    if (k < 0.0f) return F3(0.0f, 0.0f, 0.0f);
    return i * eta - n * (eta * c + sqrtf(k));
}

// randomHSDirectionCosineWeighted (include/randoms.h:215-231): rejection
// sampling of the unit disk (acceptance pi/4 - 0.001 pi).  The loop is capped
// at 64 draws (reaching the cap has probability below 1e-42); the cap returns N.
VR_DEV f3 cosine_hemisphere(Rng& rng, f3 N, float& ndotd) {
    f3 t0, t1;
    ortho_basis(N, t0, t1);
    for (int it = 0; it < 64; it++) {
        const float x = rng_uniform(rng) * 2.0f - 1.0f;
        const float y = rng_uniform(rng) * 2.0f - 1.0f;
        const float d2 = x * x + y * y;
        if (d2 > 0.001f && d2 < 1.0f) {
            const float z = sqrtf(1.0f - d2);
            ndotd = z;
            return t0 * x + t1 * y + N * z;
        }
    }
    ndotd = 1.0f;
    return N;
}

// surface_scattering_sampler(win, surfel) -> (wout, W, pdf) of a patch.
VR_DEV void surface_sample(const vr_patch& P, Rng& rng, f3 win, const Surfel& s, f3& wout, f3& W, float& pdf) {
    const int kind = P.sampler;
    if (kind == VR_SURF_PASS || kind == VR_SURF_NULL) {  // sc_path_tracing.h:78-82; _scenes.py:1308-1322
        wout = win; W = F3(1.0f, 1.0f, 1.0f); pdf = -1.0f;
        return;
    }
    if (kind == VR_SURF_BLOCK) {  // _scenes.py:1325-1340
        wout = F3(0.0f, 0.0f, 0.0f); W = wout; pdf = 0.0f;
        return;
    }
    bool from_outside;
    f3 fN;
    collision_vars(win, s, from_outside, fN);
    if (kind == VR_SURF_LAMBERT) {  // _scenes.py:906-937
        if (!from_outside) {
            wout = F3(0.0f, 0.0f, 0.0f); pdf = 0.0f; W = wout;
        } else {
            float c;
            wout = cosine_hemisphere(rng, fN, c);
            pdf = c / kPi;
            W = F3(1.0f, 1.0f, 1.0f);
        }
        return;
    }
    if (kind == VR_SURF_MIRROR) {  // _scenes.py:1003-1022
        wout = glsl_reflect(win, fN); W = F3(1.0f, 1.0f, 1.0f); pdf = -1.0f;
        return;
    }
    // VR_SURF_FRESNEL, _scenes.py:964-1000 (Schlick reflectance; the uniform is
    // drawn only when the refraction exists: GLSL's || short-circuits)
    const f3 rw = glsl_reflect(win, fN);
    const float eta = from_outside ? 1.0f / P.refraction_index : P.refraction_index;
    const f3 tw = glsl_refract(win, fN, eta);
    const float th = fabsf(dot(win, fN));
    const float f = ((1.0f - eta) / (1.0f + eta)) * ((1.0f - eta) / (1.0f + eta));
    const float om = 1.0f - th;
    const float R = f + (1.0f - f) * (om * om * om * om * om);
    const bool no_t = tw.x == 0.0f && tw.y == 0.0f && tw.z == 0.0f;
    if (no_t || rng_uniform(rng) < R) wout = rw;
    else wout = tw;
    const float iw = 1.0f / (eta * eta);
    W = F3(iw, iw, iw);
    pdf = -1.0f;
}

// surface_gathering = the material's emission (Scene.pathtrace, _scenes.py:2407-2423).
VR_DEV f3 surface_emission(const vr_patch& P, f3 win, const Surfel& s) {
    if (P.emission == VR_EMIT_NONE) return F3(0.0f, 0.0f, 0.0f);
    const f3 E = load3(P.emission_c);
    if (P.emission == VR_EMIT_ISOTROPIC) return E * (0.5f / kPi);
    bool from_outside;
    f3 fN;
    collision_vars(win, s, from_outside, fN);
    return E * (dot(win * -1.0f, fN) / kPi);
}

// Medium segment (include/dyn_medium_traversal.h, DTMediumPathIntegrator fw).
// The medium is read in the kernel-argument segment: only called from kernels
// whose first parameter is the SceneView S.
VR_DEV void medium_traverse(const SceneView& S, int m, Rng& rng, f3 x, f3 w, float d, f3& xo, f3& wo, f3& W, f3& A,
                            int& lookups) {
    Tracer<VR_MAJ_CONST> tr(&first_kernel_arg<SceneView>()->media[m]);
    tr.rng = rng;
    float out[12];
    tr.medium_dt(x, w, d, out);
    rng = tr.rng;
    lookups += tr.lookups;
    xo = F3(out[0], out[1], out[2]);
    wo = F3(out[3], out[4], out[5]);
    W = F3(out[6], out[7], out[8]);
    A = F3(out[9], out[10], out[11]);
}

// medium_traversal_bw (dyn_medium_traversal.h:20-32) after medium_traverse from
// the same state: (W, A) are that forward's outputs, which the DT backward's
// primal pass would recompute bit for bit, so only its replay runs.
VR_DEV void medium_traverse_bw(const SceneView& S, int m, Rng& rng, f3 x, f3 w, float d, f3 W, f3 A, f3 dL_dW, f3 dL_dA,
                               int& lookups, int& scatters) {
    Tracer<VR_MAJ_CONST> tr(&first_kernel_arg<SceneView>()->media[m]);
    tr.rng = rng;
    tr.medium_dt_replay(x, w, d, W, A, dL_dW, dL_dA);
    rng = tr.rng;
    lookups += tr.lookups;
    scatters += tr.scatters;
}

// begin_branch_seed / END_BRANCH around the (zero) direct gathering of
// path_trace_fw and path_trace_bw (sc_path_tracing.h:124-128, 240-249):
// the main stream advances by 4 uniforms.
VR_DEV void skip_branch(Rng& rng) {
    rng_uniform(rng); rng_uniform(rng); rng_uniform(rng); rng_uniform(rng);
}

// The medium stack of sc_path_tracing.h:84-101.
struct MediumStack {
    int current;   // medium index or -1
    int level;     // -1 = outside every medium
    int stacked[4];
    VR_DEV void init() { current = -1; level = -1; stacked[0] = stacked[1] = stacked[2] = stacked[3] = -1; }
    VR_DEV void cross(bool to_inside, int inside) {
        if (to_inside) {
            if (level >= 0) stacked[level] = current;
            if (level < 3) level++;
            current = inside;
        } else {
            level--;
            current = level == -1 ? -1 : stacked[level < 0 ? 0 : level];
        }
    }
};

// Medium on the inner side of a patch (-1 = none or out of range).
VR_DEV int inside_medium(const SceneView& S, const vr_patch& P) {
    return (P.inside_medium >= 0 && P.inside_medium < S.num_media) ? P.inside_medium : -1;
}

VR_DEV bool below(f3 W, float e) { return W.x < e && W.y < e && W.z < e; }
VR_DEV bool same_dir(f3 a, f3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

// path_trace (sc_path_tracing.h:12-120) with BACKWARD's path_trace_fw
// (:122-236, 4 uniforms of branch set-up first) when BRANCH is true.  The
// direct gathering is ZERO (Scene.pathtrace).  Returns A.
template <int STRIDE>
VR_DEV f3 scene_path(const SceneView& S, Rng& rng, f3 x, f3 w, bool branch, int* stk, int& lookups) {
    if (branch) skip_branch(rng);
    f3 A = F3(0.0f, 0.0f, 0.0f), W = F3(1.0f, 1.0f, 1.0f);
    MediumStack ms;
    ms.init();
    int bounces = 0;
    for (int it = 0; it < kScenePathCap; it++) {
        if (below(W, 0.0000001f)) return A;
        float d, bu, bv;
        const int k = scene_trace<STRIDE>(S, x, w, d, bu, bv, stk);
        if (k < 0) return A + W * scene_env(S, w);
        int patch;
        const Surfel s = scene_surfel(S, k, bu, bv, x, w, d, patch);
        const vr_patch P = S.patches[patch];
        const bool from_inside = dot(w, s.G) > 0.0f;
        if (ms.level == -1 && from_inside) { ms.level = 0; ms.current = inside_medium(S, P); }
        if (ms.current >= 0) {
            f3 xo, wo, vW, vA;
            medium_traverse(S, ms.current, rng, x, w, d, xo, wo, vW, vA, lookups);
            A = A + W * vA;
            W = W * vW;
            if (!same_dir(wo, w)) { x = xo; w = wo; continue; }
        }
        bounces++;
        A = A + W * surface_emission(P, w, s);
        f3 ws, Ws;
        float pdf;
        surface_sample(P, rng, w, s, ws, Ws, pdf);
        const bool to_inside = dot(ws, s.G) < 0.0f;
        if (to_inside != from_inside) ms.cross(to_inside, inside_medium(S, P));
        W = W * Ws;
        w = ws;
        if (pdf > 0.0f && bounces > 2) {
            if (rng_uniform(rng) < 0.5f) return A;
            W = W * 2.0f;
        }
        x = s.P + w * 0.0001f;
    }
    return A;
}

// path_trace_bw (sc_path_tracing.h:238-380): replays the path from the state
// rng held before path_trace_fw, peeling each term off out_A, and sends the
// medium segments' gradients into the media grids.
template <int STRIDE>
VR_DEV void scene_path_bw(const SceneView& S, Rng& rng, f3 x, f3 w, f3 out_A, f3 dL_dA, int* stk, int& lookups,
                          int& scatters) {
    skip_branch(rng);
    f3 rem = out_A, W = F3(1.0f, 1.0f, 1.0f);
    MediumStack ms;
    ms.init();
    int bounces = 0;
    for (int it = 0; it < kScenePathCap; it++) {
        if (below(W, 0.0000001f)) return;
        float d, bu, bv;
        const int k = scene_trace<STRIDE>(S, x, w, d, bu, bv, stk);
        if (k < 0) return;  // environment_bw: a constant (or image) environment takes no gradient
        int patch;
        const Surfel s = scene_surfel(S, k, bu, bv, x, w, d, patch);
        const vr_patch P = S.patches[patch];
        const bool from_inside = dot(w, s.G) > 0.0f;
        if (ms.level == -1 && from_inside) { ms.level = 0; ms.current = inside_medium(S, P); }
        if (ms.current >= 0) {
            f3 xo, wo, vW, vA;
            const Rng before = rng;
            medium_traverse(S, ms.current, rng, x, w, d, xo, wo, vW, vA, lookups);
            rem = rem - W * vA;
            const f3 dL_dvA = dL_dA * W;
            f3 g = dL_dA * rem;
            // dL/dvW = dL/dA * rem / vW; a non-finite component (vW == 0) is dropped
            float gv[3] = {g.x / vW.x, g.y / vW.y, g.z / vW.z};
            for (int c = 0; c < 3; c++)
                if (!isfinite(gv[c])) gv[c] = 0.0f;
            rng = before;
            medium_traverse_bw(S, ms.current, rng, x, w, d, vW, vA, F3(gv[0], gv[1], gv[2]), dL_dvA, lookups, scatters);
            W = W * vW;
            if (!same_dir(wo, w)) { x = xo; w = wo; continue; }
        }
        bounces++;
        rem = rem - W * surface_emission(P, w, s);
        f3 ws, Ws;
        float pdf;
        surface_sample(P, rng, w, s, ws, Ws, pdf);
        const bool to_inside = dot(ws, s.G) < 0.0f;
        if (to_inside != from_inside) ms.cross(to_inside, inside_medium(S, P));
        W = W * Ws;
        w = ws;
        if (pdf > 0.0f && bounces > 2) {
            if (rng_uniform(rng) < 0.5f) return;
            W = W * 2.0f;
        }
        x = s.P + w * 0.0001f;
    }
}

// PathtracedScene BACKWARD (maps/pathtraced_scene.h:13-25).
template <int STRIDE>
VR_DEV void scene_map_bw(const SceneView& S, Rng& rng, f3 x, f3 w, f3 dL_dA, int* stk, int& lookups, int& scatters) {
    const Rng before = rng;
    int lk = 0;
    const f3 A = scene_path<STRIDE>(S, rng, x, w, true, stk, lk);
    rng = before;
    scene_path_bw<STRIDE>(S, rng, x, w, A, dL_dA, stk, lookups, scatters);
    lookups += lk;
}

}  // namespace vr
