/*
 * vr_oracle.c -- CPU oracle for the participating-media hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a plain-C restatement of the
 * reference's GLSL algorithms (rendervous, the include/ headers), used by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker the
 * HIP library is compared against.  Nothing in the product links, loads or
 * calls it.
 *
 * Precision policy: every real-valued quantity is IEEE double (the "FP64
 * oracle"); the RNG is the reference's integer HybridTaus generator, restated
 * bit-exactly, and its uniforms are the same float values the GPU draws.  With
 * the same seeds the oracle therefore makes the same stochastic decisions as the
 * FP32 kernels except where a comparison falls within rounding distance.
 *
 * Exported symbols mirror include/vr_api.h with the prefix "oracle_" and host
 * pointers (no stream argument).
 *
 * Parity pins (what checks each function; see tests/):
 *   rng_*                golden LCG sequence, Tausworthe step structure, maximal
 *                        period of each Tausworthe component (primitive transition
 *                        over GF(2)), uniformity; the seed formulas stay unpinned
 *   hg_eval / hg_sample  normalisation and mean cosine g by quadrature, KS test
 *                        of sampled cosines against the analytic CDF, uniform
 *                        azimuth, sampler histogram = evaluated phase
 *   camera_ray           numpy restatement of the reference's torch camera
 *                        (_maps.py:1788-1801), edge-pixel angles from the FOV
 *   env_sample (sun)     single scattering in a homogeneous box lit by the sun
 *                        alone = 1-D quadrature, two selection probabilities
 *   ray_box / boundary   slab-test special cases, axis-parallel rays
 *   grid_lookup / _bw    closed form (affine fields are reproduced exactly),
 *                        torch grid_sample(align_corners=True) in the interior,
 *                        adjoint (dot-product) identity
 *   dda_fw / dda_bw      constant grid => exp(-sigma*chord); field linear along the
 *                        ray => exact; central finite differences of dda_fw
 *   t_dt / t_rt          homogeneous closed form exp(-sigma*d); exact line integral
 *                        (adaptive quadrature) of a heterogeneous field
 *   collision_dt         homogeneous P(C=1) = 1-exp(-sigma*d), KS test of t
 *   nee_fw               white furnace (albedo 1, consistent phase) R == env;
 *                        albedo 0 => R == T*env
 *   *_bw                 central finite differences with common random numbers
 *   majorant_build       brute-force bound of the trilinear field
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/vr_api.h"

#define PI_D 3.1415926535897932384626433832795
#define TWO_PI_D 6.283185307179586476925286766559
/* Remaining distance below which a majorant-cell crossing ends the tracking.
 * Where the last coarse cell's exit coincides with the medium's exit the
 * crossing test is a rounding tie; ending there keeps the random stream of the
 * FP32 and FP64 implementations aligned. */
#define CROSS_EPS 1e-6

static __thread char g_err[512];

static int fail(int code, const char* msg) {
    snprintf(g_err, sizeof(g_err), "%s", msg);
    return code;
}

const char* oracle_last_error(void) { return g_err; }
int oracle_abi_version(void) { return VR_ABI_VERSION; }

/* ------------------------------------------------------------------------- */
/* RNG: include/randoms.h:6-59                                                 */
/* ------------------------------------------------------------------------- */
typedef struct { uint32_t x, y, z, w; } rng_t;

/* randoms.h:6 */
static uint32_t taus_step(uint32_t z, int s1, int s2, int s3, uint32_t m) {
    uint32_t b = ((z << s1) ^ z) >> s2;
    return ((z & m) << s3) ^ b;
}

/* randoms.h:41-47 */
static void rng_advance(rng_t* s) {
    s->x = taus_step(s->x, 13, 19, 12, 4294967294u);
    s->y = taus_step(s->y, 2, 25, 4, 4294967288u);
    s->z = taus_step(s->z, 3, 11, 17, 4294967280u);
    s->w = 1664525u * s->w + 1013904223u;
}

static float bits_to_float(uint32_t b) { float f; memcpy(&f, &b, 4); return f; }
static uint32_t float_to_bits(float f) { uint32_t b; memcpy(&b, &f, 4); return b; }

/* randoms.h:53-59: 23 random mantissa bits in [1,2), minus 1. */
static float rng_random(rng_t* s) {
    rng_advance(s);
    uint32_t v = s->x ^ s->y ^ s->z ^ s->w;
    return bits_to_float((v & 0x007FFFFFu) | 0x3F800000u) - 1.0f;
}

/* Capture forward per-thread init, _maps.py:549-560. */
static rng_t seed_capture_fw(const uint32_t sd[4], uint32_t i) {
    rng_t s;
    s.x = sd[0] + (0x23F1u + i);
    s.y = sd[1] + (0x3137u + (i + 14u) * i * 17923113u);
    s.z = sd[2] + 129u * i;
    s.w = sd[3] + (i + 129u);
    rng_advance(&s);
    rng_advance(&s);
    float a = rng_random(&s), b = rng_random(&s), c = rng_random(&s), d = rng_random(&s);
    s.x = float_to_bits(a); s.y = float_to_bits(b); s.z = float_to_bits(c); s.w = float_to_bits(d);
    rng_advance(&s); rng_advance(&s); rng_advance(&s); rng_advance(&s);
    return s;
}

/* Capture backward per-thread init, _maps.py:676-685. */
static rng_t seed_capture_bw(const uint32_t sd[4], uint32_t i) {
    rng_t s;
    s.x = sd[0] + 0x23F1u;
    s.y = sd[1] + 0x3137u;
    s.z = sd[2] + 129u;
    s.w = sd[3] + (i + 129u);
    rng_advance(&s); rng_advance(&s); rng_advance(&s); rng_advance(&s);
    float a = rng_random(&s), b = rng_random(&s), c = rng_random(&s), d = rng_random(&s);
    s.x = float_to_bits(a); s.y = float_to_bits(b); s.z = float_to_bits(c); s.w = float_to_bits(d);
    return s;
}

/* Map eval per-thread init, _maps.py:390-393 (fw) and :465-469 (bw). */
static rng_t seed_map(const uint32_t sd[4], uint32_t i) {
    rng_t s;
    s.x = sd[0] ^ (i ^ 17u);
    s.y = sd[1] ^ (i * 123111171u);
    s.z = sd[2] ^ (i + 11u);
    s.w = sd[3] ^ (i ^ (sd[0] + 13u * sd[1]));
    rng_random(&s); rng_random(&s); rng_random(&s);
    return s;
}

/* ------------------------------------------------------------------------- */
/* small vector helpers                                                        */
/* ------------------------------------------------------------------------- */
typedef struct { double x, y, z; } v3;
static v3 V(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static v3 vsub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static v3 vmul(v3 a, v3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
static v3 vs(v3 a, double s) { return V(a.x * s, a.y * s, a.z * s); }
static double vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static v3 vcross(v3 a, v3 b) { return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static v3 vnorm(v3 a) { double l = sqrt(vdot(a, a)); return vs(a, 1.0 / l); }
static double vget(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
static v3 vf(const float* p) { return V(p[0], p[1], p[2]); }
static double dmin(double a, double b) { return a < b ? a : b; }
static double dmax(double a, double b) { return a > b ? a : b; }
static int imin(int a, int b) { return a < b ? a : b; }
static int imax(int a, int b) { return a > b ? a : b; }
static int iclamp(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* ------------------------------------------------------------------------- */
/* Ray-box: include/maths.h:11-38 and include/vr_boundary.h:1-11             */
/* ------------------------------------------------------------------------- */
static void ray_box_intersection(v3 x, v3 w, v3 bmin, v3 bmax, double* tmin, double* tmax) {
    /* maths.h:30-38: |w_i| <= 1e-6 is replaced by 1e-6 ("un-parallelize"). */
    double wx = fabs(w.x) <= 1e-6 ? 1e-6 : w.x;
    double wy = fabs(w.y) <= 1e-6 ? 1e-6 : w.y;
    double wz = fabs(w.z) <= 1e-6 ? 1e-6 : w.z;
    double c0x = (bmin.x - x.x) / wx, c1x = (bmax.x - x.x) / wx;
    double c0y = (bmin.y - x.y) / wy, c1y = (bmax.y - x.y) / wy;
    double c0z = (bmin.z - x.z) / wz, c1z = (bmax.z - x.z) / wz;
    *tmin = dmax(dmax(dmin(c0x, c1x), dmin(c0y, c1y)), dmin(c0z, c1z));
    *tmax = dmin(dmin(dmax(c0x, c1x), dmax(c0y, c1y)), dmax(c0z, c1z));
}

/* maths.h:11-26 */
static int intersect_ray_box(v3 x, v3 w, v3 bmin, v3 bmax, double* tmin, double* tmax) {
    ray_box_intersection(x, w, bmin, bmax, tmin, tmax);
    *tmin = dmax(0.0, *tmin);
    if (*tmax <= *tmin || *tmax <= 0) return 0;
    return 1;
}

/* vr_boundary.h:1-11 with boundary = RayBoxIntersection(box_min, box_max). */
static int boundary(const vr_medium* M, v3 x, v3 w, double* tmin, double* tmax) {
    ray_box_intersection(x, w, vf(M->box_min), vf(M->box_max), tmin, tmax);
    *tmin = dmax(0.0, *tmin);
    if (*tmax < *tmin) return 0;
    return 1;
}

/* ------------------------------------------------------------------------- */
/* Grid3D: include/maps/grid3d.h:6-75, Parameter addressing common.h:110-179  */
/* ------------------------------------------------------------------------- */
static int64_t goff(const vr_tensor* t, int z, int y, int x) {
    return t->stride[0] * z + t->stride[1] * y + t->stride[2] * x;
}

/* Grid3D FORWARD: zero outside [0,1)^3 of the normalised box; vertex-centred
 * trilinear with dim = shape-1 (grid3d.h:13-36).  Writes C channels. */
static void grid_lookup(const vr_grid* g, v3 c, double* out) {
    const vr_tensor* t = &g->t;
    int C = (int)t->shape[3];
    for (int k = 0; k < C; k++) out[k] = 0.0;
    double nx = (c.x - g->bmin[0]) / ((double)g->bmax[0] - g->bmin[0]);
    double ny = (c.y - g->bmin[1]) / ((double)g->bmax[1] - g->bmin[1]);
    double nz = (c.z - g->bmin[2]) / ((double)g->bmax[2] - g->bmin[2]);
    if (nx < 0 || ny < 0 || nz < 0 || nx >= 1 || ny >= 1 || nz >= 1) return;
    int dx = (int)t->shape[2] - 1, dy = (int)t->shape[1] - 1, dz = (int)t->shape[0] - 1;
    double gx = nx * dx, gy = ny * dy, gz = nz * dz;
    double ax = gx - floor(gx), ay = gy - floor(gy), az = gz - floor(gz);
    int px = iclamp((int)gx, 0, dx - 1), py = iclamp((int)gy, 0, dy - 1), pz = iclamp((int)gz, 0, dz - 1);
    double wgt[8] = {
        (1 - ax) * (1 - ay) * (1 - az), ax * (1 - ay) * (1 - az), (1 - ax) * ay * (1 - az), ax * ay * (1 - az),
        (1 - ax) * (1 - ay) * az, ax * (1 - ay) * az, (1 - ax) * ay * az, ax * ay * az};
    for (int k = 0; k < 8; k++) {
        int ox = k & 1, oy = (k >> 1) & 1, oz = (k >> 2) & 1;
        const float* p = t->data + goff(t, pz + oz, py + oy, px + ox);
        for (int ch = 0; ch < C; ch++) out[ch] += (double)p[t->stride[3] * ch] * wgt[k];
    }
}

/* Gradient accumulation target: either the grid's float grad (serial), or a
 * private double buffer (parallel capture bw, reduced afterwards). */
typedef struct {
    double* sig;     /* private accumulation for sigma grad, or NULL */
    double* alb;     /* private accumulation for albedo grad, or NULL */
} gacc_t;

static void acc_add(const vr_tensor* t, double* priv, int64_t off, double v) {
    if (priv) priv[off] += v;
    else t->grad[off] += (float)v;
}

/* Grid3D BACKWARD (grid3d.h:39-75): 8 trilinear-weighted accumulations. */
static void grid_scatter(const vr_grid* g, double* priv, v3 c, const double* og) {
    const vr_tensor* t = &g->t;
    if (t->grad == NULL) return;
    int C = (int)t->shape[3];
    double nx = (c.x - g->bmin[0]) / ((double)g->bmax[0] - g->bmin[0]);
    double ny = (c.y - g->bmin[1]) / ((double)g->bmax[1] - g->bmin[1]);
    double nz = (c.z - g->bmin[2]) / ((double)g->bmax[2] - g->bmin[2]);
    if (nx < 0 || ny < 0 || nz < 0 || nx >= 1 || ny >= 1 || nz >= 1) return;
    int dx = (int)t->shape[2] - 1, dy = (int)t->shape[1] - 1, dz = (int)t->shape[0] - 1;
    double gx = nx * dx, gy = ny * dy, gz = nz * dz;
    double ax = gx - floor(gx), ay = gy - floor(gy), az = gz - floor(gz);
    int px = iclamp((int)gx, 0, dx - 1), py = iclamp((int)gy, 0, dy - 1), pz = iclamp((int)gz, 0, dz - 1);
    double wgt[8] = {
        (1 - ax) * (1 - ay) * (1 - az), ax * (1 - ay) * (1 - az), (1 - ax) * ay * (1 - az), ax * ay * (1 - az),
        (1 - ax) * (1 - ay) * az, ax * (1 - ay) * az, (1 - ax) * ay * az, ax * ay * az};
    for (int k = 0; k < 8; k++) {
        int ox = k & 1, oy = (k >> 1) & 1, oz = (k >> 2) & 1;
        int64_t base = goff(t, pz + oz, py + oy, px + ox);
        for (int ch = 0; ch < C; ch++) acc_add(t, priv, base + t->stride[3] * ch, og[ch] * wgt[k]);
    }
}

/* ------------------------------------------------------------------------- */
/* Medium fields: sigma (vr_sigma.h), albedo (vr_scattering_albedo.h),        */
/* emission (vr_emission.h), environment (vr_environment.h)                   */
/* ------------------------------------------------------------------------- */
typedef struct {
    const vr_medium* M;
    rng_t rng;
    gacc_t acc;
    int64_t lookups;  /* trilinear sigma lookups (debug counter) */
    int bw_mode;
} ctx_t;

static double sigma_at(ctx_t* c, v3 x) {
    double v;
    c->lookups++;
    grid_lookup(&c->M->sigma, x, &v);
    return v;
}

static void sigma_bw(ctx_t* c, v3 x, double g) {
    grid_scatter(&c->M->sigma, c->acc.sig, x, &g);
}

static v3 albedo_at(ctx_t* c, v3 x) {
    const vr_medium* M = c->M;
    if (M->albedo_kind == VR_FIELD_CONST) return vf(M->albedo_c);
    double v[4] = {0, 0, 0, 0};
    grid_lookup(&M->albedo, x, v);
    if (M->albedo.t.shape[3] == 1) return V(v[0], v[0], v[0]);  /* promote 1 -> 3 */
    return V(v[0], v[1], v[2]);
}

static void albedo_bw(ctx_t* c, v3 x, v3 g) {
    const vr_medium* M = c->M;
    if (M->albedo_kind != VR_FIELD_GRID || M->albedo.t.grad == NULL) return;
    double og[3];
    if (M->albedo.t.shape[3] == 1) { og[0] = g.x + g.y + g.z; }
    else { og[0] = g.x; og[1] = g.y; og[2] = g.z; }
    grid_scatter(&M->albedo, c->acc.alb, x, og);
}

static double rnd(ctx_t* c) { return (double)rng_random(&c->rng); }

/* ------------------------------------------------------------------------- */
/* Directions: randoms.h:144-165, scattering_AD.h:4-52                       */
/* ------------------------------------------------------------------------- */
/* randoms.h:144-151 */
static void ortho_basis(v3 n, v3* t, v3* b) {
    double sign = n.z >= 0 ? 1.0 : -1.0;
    double a = -1.0 / (sign + n.z);
    double bb = n.x * n.y * a;
    *t = V(1.0 + sign * n.x * n.x * a, sign * bb, -sign * n.x);
    *b = V(bb, sign + n.y * n.y * a, -n.y);
}

/* randoms.h:153-165 randomDirection(D) */
static v3 random_direction(ctx_t* c, v3 D) {
    double r1 = rnd(c);
    double r2 = rnd(c) * 2 - 1;
    double s = sqrt(dmax(0.0, 1.0 - r2 * r2));
    double xx = cos(TWO_PI_D * r1) * s, yy = sin(TWO_PI_D * r1) * s, zz = r2;
    v3 t0, t1;
    ortho_basis(D, &t0, &t1);
    return vadd(vadd(vs(t0, xx), vs(t1, yy)), vs(D, zz));
}

/* scattering_AD.h:4-19 */
static v3 hg_sample(ctx_t* c, v3 win, double g) {
    double phi = rnd(c) * 2 * PI_D;
    double xi = rnd(c);
    double cos_t;
    if (fabs(g) < 0.001) cos_t = 2 * xi - 1;
    else {
        double g2 = g * g;
        double t = (1.0 - g2) / (1.0 - g + 2.0 * g * xi);
        cos_t = (0.5 / g) * (1.0 + g2 - t * t);
    }
    double sin_t = sqrt(dmax(0.0, 1.0 - cos_t * cos_t));
    v3 t0, t1;
    ortho_basis(win, &t0, &t1);
    return vadd(vadd(vs(t0, sin_t * sin(phi)), vs(t1, sin_t * cos(phi))), vs(win, cos_t));
}

/* scattering_AD.h:39-47 */
static double hg_eval(double cos_t, double g) {
    if (fabs(g) < 0.001) return 0.25 / PI_D;
    double g2 = g * g;
    return 0.25 / PI_D * (1.0 - g2) / pow(1.0 + g2 - 2 * g * cos_t, 1.5);
}

/* The HG asymmetry at scattering position x: the constant phase_g or the
 * phase_g grid read there (HGPhase / VHGPhaseSampler evaluate their phase_g map
 * at the first three inputs, _maps.py:4449,4473). */
static double phase_g_at(const vr_medium* M, v3 x) {
    if (M->phase_g_kind != VR_FIELD_GRID || M->phase_kind == VR_PHASE_ISO) return M->phase_g;
    double g;
    grid_lookup(&M->phase_g_grid, x, &g);
    return g;
}

/* phase(x, w, wo): HGPhase (_maps.py:4446-4454) / single HG / isotropic. */
static double phase_eval(const vr_medium* M, v3 x, v3 w, v3 wo) {
    double ct = vdot(w, wo);
    double g = phase_g_at(M, x);
    switch (M->phase_kind) {
    case VR_PHASE_HG_MIX: {
        double a = hg_eval(ct, g), b = hg_eval(ct, g / 5);
        return a * (1 - 0.2) + b * 0.2;
    }
    case VR_PHASE_HG: return hg_eval(ct, g);
    default: return 0.25 / PI_D;
    }
}

/* phase_sampler(x, w) -> (rho, wo): VHGPhaseSampler (_maps.py:4463-4481) /
 * HGPhaseSampler (_scenes.py:1984-2004) / uniform. */
static double phase_sample(ctx_t* c, v3 x, v3 w, v3* wo) {
    const vr_medium* M = c->M;
    double g = phase_g_at(M, x);
    switch (M->phase_kind) {
    case VR_PHASE_HG_MIX: {
        if (rnd(c) < 0.8) g /= 5;
        *wo = hg_sample(c, w, g);
        return 1.0;
    }
    case VR_PHASE_HG: *wo = hg_sample(c, w, g); return 1.0;
    default: *wo = random_direction(c, V(0, 1, 0)); return 1.0;
    }
}

/* maths.h:172-178 dir2xr */
static void dir2xr(v3 w, double* u, double* v) {
    if (w.x == 0.0 && w.z == 0.0) w.x += 0.0000001;
    *u = atan2(w.x, w.z) / PI_D;
    *v = -2.0 * asin(dmin(dmax(w.y, -1.0), 1.0)) / PI_D;
}

/* maps/image2d.h:1-25: pixel-centred bilinear lookup, coordinates clamped to the
 * box, corner indices clamped to the image. */
static v3 image_lookup(const vr_medium* M, double cx, double cy) {
    const vr_tensor* t = &M->env_image;
    int H = (int)t->shape[0], W = (int)t->shape[1];
    double nx = dmin(dmax((cx - M->env_image_bmin[0]) / ((double)M->env_image_bmax[0] - M->env_image_bmin[0]), 0.0), 1.0);
    double ny = dmin(dmax((cy - M->env_image_bmin[1]) / ((double)M->env_image_bmax[1] - M->env_image_bmin[1]), 0.0), 1.0);
    double gx = nx * W - 0.5, gy = ny * H - 0.5;
    int px = (int)floor(gx), py = (int)floor(gy);
    double ax = gx - px, ay = gy - py;
    int xs[2] = {iclamp(px, 0, W - 1), iclamp(px + 1, 0, W - 1)};
    int ys[2] = {iclamp(py, 0, H - 1), iclamp(py + 1, 0, H - 1)};
    double r[3];
    for (int ch = 0; ch < 3; ch++) {
        double v00 = t->data[ys[0] * t->stride[0] + xs[0] * t->stride[1] + ch * t->stride[2]];
        double v01 = t->data[ys[0] * t->stride[0] + xs[1] * t->stride[1] + ch * t->stride[2]];
        double v10 = t->data[ys[1] * t->stride[0] + xs[0] * t->stride[1] + ch * t->stride[2]];
        double v11 = t->data[ys[1] * t->stride[0] + xs[1] * t->stride[1] + ch * t->stride[2]];
        r[ch] = (v00 * (1 - ax) + v01 * ax) * (1 - ay) + (v10 * (1 - ax) + v11 * ax) * ay;
    }
    return V(r[0], r[1], r[2]);
}

/* environment(x, w): constant, or Image2D over dir2xr(w) (maps/environment_xr.h:3-9). */
static v3 env_eval(const vr_medium* M, v3 w) {
    if (M->env_kind == VR_ENV_XR) {
        double u, v;
        dir2xr(w, &u, &v);
        return image_lookup(M, u, v);
    }
    return vf(M->env_c);
}

/* XRQuadtreeRandomDirection (_maps.py:2494-2549): quadtree descent with one
 * uniform over the absolute node probabilities, then randomDirection(a0, a1,
 * b0, b1) (randoms.h:186-199) in the selected pixel; returns 1 / pdf. */
static double quadtree_direction(ctx_t* c, v3* wo) {
    const vr_medium* M = c->M;
    double sel = rnd(c);
    int node = 0;
    double p0x = 0, p0y = 0, p1x = 1, p1y = 1, prob = 1;
    for (int i = 0; i < M->env_quadtree_levels; i++) {
        const float* d = M->env_quadtree + node * 4;
        int child = 3;
        prob = d[3];
        for (int k = 0; k < 3; k++) {
            if (sel < d[k]) { child = k; prob = d[k]; break; }
            sel -= d[k];
        }
        double xm = (p1x + p0x) / 2, ym = (p1y + p0y) / 2;
        if (child % 2 == 0) p1x = xm; else p0x = xm;
        if (child / 2 == 0) p1y = ym; else p0y = ym;
        node = node * 4 + 1 + child;
    }
    double area = 2 * PI_D * (cos(p0y * PI_D) - cos(p1y * PI_D)) / (double)(1 << M->env_quadtree_levels);
    double weight = area / dmax(0.000000001, prob);
    double a0 = (p0x * 2 - 1) * PI_D, a1 = (p1x * 2 - 1) * PI_D, b0 = p0y * PI_D, b1 = p1y * PI_D;
    double r1 = rnd(c) * (a1 - a0) + a0;
    double r2 = cos(b0) - rnd(c) * (cos(b0) - cos(b1));
    double s = sqrt(dmax(0.0, 1.0 - r2 * r2));
    *wo = V(sin(r1) * s, r2, cos(r1) * s);
    return weight;
}

/* _functionsampler(UniformDirectionSampler or XRQuadtreeRandomDirection,
 * environment) (_maps.py:2552-2585): L(wo) / pdf. */
static v3 env_sample_base(ctx_t* c, v3* wo) {
    const vr_medium* M = c->M;
    if (M->env_sampler_kind == VR_ENVS_QUADTREE) {
        double wgt = quadtree_direction(c, wo);
        return vs(env_eval(M, *wo), wgt);
    }
    *wo = random_direction(c, V(0, 1, 0));
    return vs(env_eval(M, *wo), 4 * PI_D);
}

/* environment_sampler(x, w) -> (L/pdf, wo), optionally mixed with a
 * directional light through MixtureSampler semantics (_maps.py:1541-1571;
 * sub-samplers pre-scaled by their selection probability). */
static v3 env_sample(ctx_t* c, v3* wo) {
    const vr_medium* M = c->M;
    if (M->sun_prob > 0) {
        if (rnd(c) < M->sun_prob) {
            *wo = vf(M->sun_dir);
            return vs(vf(M->sun_radiance), 1.0 / M->sun_prob);
        }
        return vs(env_sample_base(c, wo), 1.0 / (1.0 - M->sun_prob));
    }
    return env_sample_base(c, wo);
}

/* ------------------------------------------------------------------------- */
/* Majorant: include/vr_majorant.h:1-7 (value, distance).                     */
/* VR_MAJ_CONST returns the constant pair.  VR_MAJ_GRID walks the coarse max  */
/* grid with a 3-D DDA: value = current coarse cell's bound, distance = to the */
/* cell's exit along the ray.                                                 */
/* ------------------------------------------------------------------------- */
typedef struct {
    int grid;
    double value, dist;          /* const */
    int cell[3], step[3], res[3];
    double tnext[3], tdelta[3], t;
    int inside;
    const float* data;
} maj_t;

static void maj_begin(const vr_medium* M, maj_t* m, v3 x, v3 w) {
    m->grid = M->majorant_kind == VR_MAJ_GRID;
    if (!m->grid) { m->value = M->majorant_value; m->dist = M->majorant_distance; return; }
    m->data = M->majorant_grid;
    m->res[0] = M->majorant_res[2]; m->res[1] = M->majorant_res[1]; m->res[2] = M->majorant_res[0];
    m->t = 0;
    m->inside = 1;
    for (int a = 0; a < 3; a++) {
        double bmin = M->sigma.bmin[a], bmax = M->sigma.bmax[a];
        double cs = (bmax - bmin) / m->res[a];
        double xa = vget(x, a), wa = vget(w, a);
        int cl = (int)floor((xa - bmin) / cs);
        cl = iclamp(cl, 0, m->res[a] - 1);
        m->cell[a] = cl;
        if (wa > 0) { m->step[a] = 1; m->tnext[a] = (bmin + (cl + 1) * cs - xa) / wa; m->tdelta[a] = cs / wa; }
        else if (wa < 0) { m->step[a] = -1; m->tnext[a] = (bmin + cl * cs - xa) / wa; m->tdelta[a] = -cs / wa; }
        else { m->step[a] = 0; m->tnext[a] = INFINITY; m->tdelta[a] = INFINITY; }
    }
}

static double maj_query(const maj_t* m, double* md) {
    if (!m->grid) { *md = m->dist; return m->value; }
    if (!m->inside) { *md = INFINITY; return 0.0; }
    double tn = dmin(m->tnext[0], dmin(m->tnext[1], m->tnext[2]));
    *md = dmax(0.0, tn - m->t);
    return m->data[((int64_t)m->cell[2] * m->res[1] + m->cell[1]) * m->res[0] + m->cell[0]];
}

/* The tracked point moved by dt inside the current cell. */
static void maj_move(maj_t* m, double dt) { if (m->grid) m->t += dt; }

/* The tracked point moved to the current cell's exit: step the DDA. */
static void maj_cross(maj_t* m) {
    if (!m->grid || !m->inside) return;
    int a = (m->tnext[0] <= m->tnext[1] && m->tnext[0] <= m->tnext[2]) ? 0 : (m->tnext[1] <= m->tnext[2] ? 1 : 2);
    m->t = m->tnext[a];
    m->cell[a] += m->step[a];
    m->tnext[a] += m->tdelta[a];
    if (m->cell[a] < 0 || m->cell[a] >= m->res[a]) m->inside = 0;
}

/* ------------------------------------------------------------------------- */
/* Transmittance estimators                                                   */
/* ------------------------------------------------------------------------- */

/* vr_transmittance_DT.h:1-28 */
static double t_dt(ctx_t* c, v3 x, v3 w) {
    double tmin, tmax;
    if (!boundary(c->M, x, w, &tmin, &tmax)) return 1.0;
    x = vadd(x, vs(w, tmin));
    double d = tmax - tmin;
    maj_t mj;
    maj_begin(c->M, &mj, x, w);
    for (;;) {
        double md;
        double maj = dmax(1e-7, maj_query(&mj, &md));
        double dt = -log(1.0 - rnd(c)) / maj;
        double t = dmin(md, dt);
        x = vadd(x, vs(w, t));
        d -= t;
        if (dt > md) { maj_cross(&mj); if (d <= CROSS_EPS) return 1.0; continue; }
        maj_move(&mj, t);
        if (d <= 1e-7) return 1.0;
        if (rnd(c) < sigma_at(c, x) / maj) return 0.0;
    }
}

/* vr_transmittance_RT.h:1-49 (with majorant-cell crossings for VR_MAJ_GRID). */
static double t_rt(ctx_t* c, v3 x, v3 w) {
    double T = 1.0, tmin, tmax;
    if (!boundary(c->M, x, w, &tmin, &tmax)) return T;
    x = vadd(x, vs(w, tmin));
    double d = tmax - tmin;
    maj_t mj;
    maj_begin(c->M, &mj, x, w);
    for (;;) {
        double md;
        double maj = dmax(1e-7, maj_query(&mj, &md));
        double s = -log(1.0 - rnd(c)) / maj;
        if (mj.grid && md < d && s > md) {
            x = vadd(x, vs(w, md)); d -= md; maj_cross(&mj);
            if (d <= CROSS_EPS) return T;
            continue;
        }
        double dt = dmin(s, dmin(d, md));
        x = vadd(x, vs(w, dt));
        d -= dt;
        maj_move(&mj, dt);
        if (d <= 1e-7) return T;
        T *= (1 - sigma_at(c, x) / maj);
        if (T < 0.01) {
            if (rnd(c) < 1 - T) return 0.0;
            T = 1.0;
        }
    }
}

/* vr_sigma_grid.h:1-28 (cell corners clamped to [0, dim]) */
static double dda_sigma_at(const vr_grid* g, int x, int y, int z) {
    const vr_tensor* t = &g->t;
    int dx = (int)t->shape[2] - 1, dy = (int)t->shape[1] - 1, dz = (int)t->shape[0] - 1;
    x = iclamp(x, 0, dx); y = iclamp(y, 0, dy); z = iclamp(z, 0, dz);
    return t->data[goff(t, z, y, x)];
}

static double lerp(double a, double b, double t) { return a * (1 - t) + b * t; }

typedef struct {
    v3 x;                 /* entry point */
    double d;             /* length inside both boxes */
    int cell[3], side[3];
    double ainc[3], alpha0[3];
    double vn[3], vm[3];
    int dim[3];
} dda_t;

/* vr_transmittance_DDA.h:3-33: returns 0 when the ray misses. */
static int dda_begin(const vr_medium* M, v3 x, v3 w, dda_t* s) {
    double tmin, tmax, gmin, gmax;
    if (!boundary(M, x, w, &tmin, &tmax)) return 0;
    const vr_grid* g = &M->sigma;
    v3 bmin = vf(g->bmin), bmax = vf(g->bmax);
    if (!intersect_ray_box(x, w, bmin, bmax, &gmin, &gmax)) return 0;
    tmin = dmax(tmin, gmin);
    tmax = dmin(tmax, gmax);
    if (tmax <= tmin) return 0;
    x = vadd(x, vs(w, tmin));
    s->x = x;
    s->d = tmax - tmin;
    v3 box = vsub(bmax, bmin);
    s->dim[0] = (int)g->t.shape[2] - 1; s->dim[1] = (int)g->t.shape[1] - 1; s->dim[2] = (int)g->t.shape[0] - 1;
    for (int a = 0; a < 3; a++) {
        double xa = vget(x, a), wa = vget(w, a), bm = vget(bmin, a), bs = vget(box, a);
        double cs = bs / s->dim[a];
        int cl = (int)((xa - bm) * s->dim[a] / bs);
        cl = iclamp(cl, 0, s->dim[a] - 1);
        s->cell[a] = cl;
        double aw = dmax(1e-6, fabs(wa));
        s->ainc[a] = cs / aw;
        s->side[a] = wa > 0 ? 1 : (wa < 0 ? -1 : 0);
        double corner = (cl + s->side[a] * 0.5 + 0.5) * cs + bm;
        s->alpha0[a] = fabs(corner - xa) / aw;
        s->vn[a] = (xa - bm) * s->dim[a] / bs;
        s->vm[a] = wa * s->dim[a] / bs;
    }
    return 1;
}

/* The DDA cell step with the axis tie-break of vr_transmittance_DDA.h:48-55. */
static void dda_step(double alpha[3], int cell[3], const dda_t* s) {
    int ix = alpha[0] <= alpha[1] && alpha[0] <= alpha[2];
    int iy = alpha[0] > alpha[1] && alpha[1] <= alpha[2];
    int iz = alpha[0] > alpha[2] && alpha[1] > alpha[2];
    alpha[0] += ix * s->ainc[0]; alpha[1] += iy * s->ainc[1]; alpha[2] += iz * s->ainc[2];
    cell[0] += ix * s->side[0]; cell[1] += iy * s->side[1]; cell[2] += iz * s->side[2];
}

/* vr_transmittance_DDA.h:1-59: midpoint rule per cell segment. */
static double dda_tau(ctx_t* c, const dda_t* s) {
    const vr_grid* g = &c->M->sigma;
    double alpha[3] = {s->alpha0[0], s->alpha0[1], s->alpha0[2]};
    int cell[3] = {s->cell[0], s->cell[1], s->cell[2]};
    double cur = 0, tau = 0;
    while (cur < s->d - 0.0001) {
        double next = dmin(s->d, dmin(alpha[0], dmin(alpha[1], alpha[2])));
        double sg[8];
        for (int k = 0; k < 8; k++) sg[k] = dda_sigma_at(g, cell[0] + (k & 1), cell[1] + ((k >> 1) & 1), cell[2] + ((k >> 2) & 1));
        c->lookups++;
        double tm = lerp(cur, next, 0.5);
        double ax = s->vm[0] * tm + s->vn[0] - cell[0];
        double ay = s->vm[1] * tm + s->vn[1] - cell[1];
        double az = s->vm[2] * tm + s->vn[2] - cell[2];
        double v = lerp(lerp(lerp(sg[0], sg[1], ax), lerp(sg[2], sg[3], ax), ay),
                        lerp(lerp(sg[4], sg[5], ax), lerp(sg[6], sg[7], ax), ay), az);
        tau += v * (next - cur);
        dda_step(alpha, cell, s);
        cur = next;
    }
    return tau;
}

/* include/vr_transmittance_GDT.h:1-84 / _GRT.h:1-83: over the DDA's cells, the
 * cell's minorant (min of its 8 vertices) is integrated in closed form and the
 * residual sigma - minorant tracked with the majorant max - min (delta tracking
 * ends at a collision; ratio tracking multiplies by 1 - Pc). */
static double t_grid_tracking(ctx_t* c, v3 x, v3 w, int ratio) {
    dda_t s;
    if (!dda_begin(c->M, x, w, &s)) return 1.0;
    const vr_grid* g = &c->M->sigma;
    double alpha[3] = {s.alpha0[0], s.alpha0[1], s.alpha0[2]};
    int cell[3] = {s.cell[0], s.cell[1], s.cell[2]};
    double cur = 0, T = 1.0;
    while (cur < s.d - 0.0001) {
        double next = dmin(s.d, dmin(alpha[0], dmin(alpha[1], alpha[2])));
        double sg[8], mx = -INFINITY, mn = INFINITY;
        for (int k = 0; k < 8; k++) {
            sg[k] = dda_sigma_at(g, cell[0] + (k & 1), cell[1] + ((k >> 1) & 1), cell[2] + ((k >> 2) & 1));
            mx = dmax(mx, sg[k]);
            mn = dmin(mn, sg[k]);
        }
        double ct = cur;
        T *= exp(-mn * (next - cur));
        double r = dmax(0.000001, mx - mn);
        if (!(r < INFINITY)) return 0.0;  /* an infinite density in the cell absorbs */
        for (;;) {
            double dt = -log(1.0 - rnd(c)) / r;
            if (ct + dt > next) break;
            ct += dt;
            double ax = s.vm[0] * ct + s.vn[0] - cell[0];
            double ay = s.vm[1] * ct + s.vn[1] - cell[1];
            double az = s.vm[2] * ct + s.vn[2] - cell[2];
            double v = lerp(lerp(lerp(sg[0], sg[1], ax), lerp(sg[2], sg[3], ax), ay),
                            lerp(lerp(sg[4], sg[5], ax), lerp(sg[6], sg[7], ax), ay), az);
            c->lookups++;
            double Pc = (v - mn) / r;
            if (!ratio) { if (rnd(c) < Pc) return 0.0; }
            else T *= 1 - Pc;
        }
        dda_step(alpha, cell, &s);
        cur = next;
    }
    return T;
}

static double t_dda(ctx_t* c, v3 x, v3 w) {
    dda_t s;
    if (!dda_begin(c->M, x, w, &s)) return 1.0;
    return exp(-dda_tau(c, &s));
}

/* vr_transmittance_DDA.h:62-156 */
static void t_dda_bw(ctx_t* c, v3 x, v3 w, double dL_dT) {
    if (dL_dT == 0) return;
    const vr_grid* g = &c->M->sigma;
    if (g->t.grad == NULL) return;
    dda_t s;
    if (!dda_begin(c->M, x, w, &s)) return;
    double T = exp(-dda_tau(c, &s));
    double dL_dtau = -dL_dT * T;
    double alpha[3] = {s.alpha0[0], s.alpha0[1], s.alpha0[2]};
    int cell[3] = {s.cell[0], s.cell[1], s.cell[2]};
    double cur = 0;
    const vr_tensor* t = &g->t;
    while (cur < s.d - 0.0001) {
        double next = dmin(s.d, dmin(alpha[0], dmin(alpha[1], alpha[2])));
        double tm = lerp(cur, next, 0.5);
        double ax = s.vm[0] * tm + s.vn[0] - cell[0];
        double ay = s.vm[1] * tm + s.vn[1] - cell[1];
        double az = s.vm[2] * tm + s.vn[2] - cell[2];
        double gs = dL_dtau * (next - cur);
        double wg[8] = {(1 - ax) * (1 - ay) * (1 - az), ax * (1 - ay) * (1 - az), (1 - ax) * ay * (1 - az), ax * ay * (1 - az),
                        (1 - ax) * (1 - ay) * az, ax * (1 - ay) * az, (1 - ax) * ay * az, ax * ay * az};
        for (int k = 0; k < 8; k++) {
            int xx = iclamp(cell[0] + (k & 1), 0, s.dim[0]);
            int yy = iclamp(cell[1] + ((k >> 1) & 1), 0, s.dim[1]);
            int zz = iclamp(cell[2] + ((k >> 2) & 1), 0, s.dim[2]);
            acc_add(t, c->acc.sig, goff(t, zz, yy, xx), gs * wg[k]);
        }
        dda_step(alpha, cell, &s);
        cur = next;
    }
}

/* vr_transmittance_DT.h:30-90 */
static void t_dt_bw(ctx_t* c, v3 x, v3 w, double dL_dT) {
    if (dL_dT == 0) return;
    double tmin, tmax;
    if (!boundary(c->M, x, w, &tmin, &tmax)) return;
    x = vadd(x, vs(w, tmin));
    double d = tmax - tmin;
    v3 x0 = x;
    double d0 = d;
    rng_t saved = c->rng;
    maj_t mj;
    maj_begin(c->M, &mj, x, w);
    for (;;) {
        double md;
        double maj = dmax(1e-7, maj_query(&mj, &md) + 0.1);
        double dt = -log(1.0 - rnd(c)) / maj;
        double t = dmin(md, dt);
        x = vadd(x, vs(w, t));
        d -= t;
        if (dt > md) { maj_cross(&mj); if (d <= CROSS_EPS) break; continue; }
        maj_move(&mj, t);
        if (d <= 1e-7) break;
        if (rnd(c) < sigma_at(c, x) / maj) return;
    }
    x = x0; d = d0;
    c->rng = saved;
    maj_begin(c->M, &mj, x, w);
    for (;;) {
        double md;
        double maj = dmax(1e-7, maj_query(&mj, &md) + 0.1);
        double dt = -log(1.0 - rnd(c)) / maj;
        double t = dmin(md, dt);
        x = vadd(x, vs(w, t));
        d -= t;
        if (dt > md) { maj_cross(&mj); if (d <= CROSS_EPS) break; continue; }
        maj_move(&mj, t);
        if (d <= 1e-7) break;
        rnd(c);
        double Pc = sigma_at(c, x) / maj;
        sigma_bw(c, x, -dL_dT / maj / (1 - Pc));
    }
}

/* vr_transmittance_RT.h:51-141 */
static void t_rt_bw(ctx_t* c, v3 x, v3 w, double dL_dT) {
    if (dL_dT == 0) return;
    double tmin, tmax;
    if (!boundary(c->M, x, w, &tmin, &tmax)) return;
    const double EPS = 1e-7;
    double T = 1.0;
    x = vadd(x, vs(w, tmin));
    double d = tmax - tmin;
    v3 x0 = x;
    double d0 = d;
    rng_t saved = c->rng;
    maj_t mj;
    maj_begin(c->M, &mj, x, w);
    for (;;) {
        double md;
        double maj = dmax(EPS, maj_query(&mj, &md) + 0.1);
        double s = -log(1.0 - rnd(c)) / maj;
        if (mj.grid && md < d && s > md) { x = vadd(x, vs(w, md)); d -= md; maj_cross(&mj); if (d <= CROSS_EPS) break; continue; }
        double dt = dmin(s, dmin(d, md));
        x = vadd(x, vs(w, dt));
        d -= dt;
        maj_move(&mj, dt);
        if (d <= EPS) break;
        T *= (1 - sigma_at(c, x) / maj);
        if (T < 0.1) {
            if (rnd(c) < 1 - T) return;
            T = 1.0;
        }
    }
    double out_T = T;
    T = 1.0;
    c->rng = saved;
    x = x0; d = d0;
    maj_begin(c->M, &mj, x, w);
    for (;;) {
        double md;
        double maj = dmax(EPS, maj_query(&mj, &md) + 0.1);
        double s = -log(1.0 - rnd(c)) / maj;
        if (mj.grid && md < d && s > md) { x = vadd(x, vs(w, md)); d -= md; maj_cross(&mj); if (d <= CROSS_EPS) break; continue; }
        double dt = dmin(s, dmin(d, md));
        x = vadd(x, vs(w, dt));
        d -= dt;
        maj_move(&mj, dt);
        if (d <= EPS) break;
        double r = 1 - sigma_at(c, x) / maj;
        T *= r;
        sigma_bw(c, x, -out_T * dL_dT / maj / r);
        if (T < 0.1) {
            rng_advance(&c->rng);  /* vr_transmittance_RT.h:125 skips the RR draw */
            T = 1.0;
        }
    }
}

static double transmittance(ctx_t* c, v3 x, v3 w) {
    switch (c->M->transmittance_kind) {
    case VR_T_RT: return t_rt(c, x, w);
    case VR_T_DDA: return t_dda(c, x, w);
    case VR_T_GDT: return t_grid_tracking(c, x, w, 0);
    case VR_T_GRT: return t_grid_tracking(c, x, w, 1);
    default: return t_dt(c, x, w);
    }
}

static void transmittance_bw(ctx_t* c, v3 x, v3 w, double dL_dT) {
    switch (c->M->transmittance_kind) {
    case VR_T_RT: t_rt_bw(c, x, w, dL_dT); break;
    case VR_T_DDA: t_dda_bw(c, x, w, dL_dT); break;
    case VR_T_GDT: case VR_T_GRT: break;  /* the reference has no backward for them */
    default: t_dt_bw(c, x, w, dL_dT); break;
    }
}

/* ------------------------------------------------------------------------- */
/* Collision sampler: include/vr_collision_sampler_DT.h:5-134                 */
/* ------------------------------------------------------------------------- */
static double collision_dt(ctx_t* c, v3 x, v3 w, double* out_t, double* out_T) {
    *out_T = 1.0; *out_t = 0.0;
    double C = 0.0, tmin, tmax;
    if (!boundary(c->M, x, w, &tmin, &tmax)) return C;
    double d = tmax - tmin;
    x = vadd(x, vs(w, tmin));
    double t = 0;
    maj_t mj;
    maj_begin(c->M, &mj, x, w);
    for (;;) {
        double md;
        double maj = maj_query(&mj, &md);
        if (mj.grid) maj = dmax(1e-7, maj);
        double dt = dmin(-log(1.0 - rnd(c)) / maj, dmin(d, md));
        x = vadd(x, vs(w, dt));
        t += dt;
        if (dt >= md) { maj_cross(&mj); if (t >= d - CROSS_EPS) break; continue; }
        maj_move(&mj, dt);
        if (t >= d - 1e-7) break;
        if (rnd(c) < sigma_at(c, x) / maj) { *out_T = 0.0; C = 1.0; break; }
    }
    *out_t = tmin + t;
    return C;
}

/* Defensive primal pass of vr_collision_sampler_DT.h:52-90: returns C, the
 * in-volume distance t and the path weight Wc (or the escape weight out_T in
 * *Wc when C == 0). */
static void collision_ds_primal(ctx_t* c, v3 x, v3 w, double eps, double tmin, double tmax,
                                double* C, double* t_out, double* Wc) {
    double d = tmax - tmin;
    x = vadd(x, vs(w, tmin));
    double t = 0, W = 1.0;
    *C = 0.0;
    maj_t mj;
    maj_begin(c->M, &mj, x, w);
    for (;;) {
        double md;
        double maj = dmax(1e-10, maj_query(&mj, &md) + 2 * eps);
        double dt = dmin(-log(1.0 - rnd(c)) / maj, dmin(d, md));
        x = vadd(x, vs(w, dt));
        t += dt;
        if (dt >= md) { maj_cross(&mj); if (t >= d - CROSS_EPS) break; continue; }
        maj_move(&mj, dt);
        if (t >= d - 1e-7) break;
        double s = sigma_at(c, x);
        double Pc = (s + eps) / maj;
        if (rnd(c) < Pc) { *C = 1.0; W *= (s / maj) / Pc; break; }
        W *= (1 - s / maj) / (1 - Pc);
    }
    *t_out = t;
    *Wc = W;
}

/* vr_collision_sampler_DT.h:46-134.  VR_BW_REFERENCE scales the collision term
 * by the collision ratio only (s_t = s/maj/Pc, :86); VR_BW_CONSISTENT uses the
 * whole path weight (null ratios before the collision included), which makes
 * the estimator unbiased for eps > 0. */
static void collision_dt_bw(ctx_t* c, v3 x, v3 w, double eps, double dL_dC, double dL_dT) {
    double tmin, tmax;
    if (!boundary(c->M, x, w, &tmin, &tmax)) return;
    rng_t saved = c->rng;
    double out_T = 1.0, C = 0.0, s_t = 0.0, Wpre = 1.0;
    double d = tmax - tmin;
    x = vadd(x, vs(w, tmin));
    v3 x0 = x;
    double t = 0;
    maj_t mj;
    maj_begin(c->M, &mj, x, w);
    for (;;) {
        double md;
        double maj = dmax(1e-10, maj_query(&mj, &md) + 2 * eps);
        double dt = dmin(-log(1.0 - rnd(c)) / maj, dmin(d, md));
        x = vadd(x, vs(w, dt));
        t += dt;
        if (dt >= md) { maj_cross(&mj); if (t >= d - CROSS_EPS) break; continue; }
        maj_move(&mj, dt);
        if (t >= d - 1e-7) break;
        double s = sigma_at(c, x);
        double Pc = (s + eps) / maj;
        if (rnd(c) < Pc) { out_T = 0.0; C = 1.0; s_t = s / maj / Pc; break; }
        double r = (1 - s / maj) / (1 - Pc);
        out_T *= r;
        Wpre *= r;
    }
    int consistent = c->bw_mode == VR_BW_CONSISTENT;
    if (consistent) s_t *= Wpre;
    c->rng = saved;
    x = x0; t = 0;
    maj_begin(c->M, &mj, x, w);
    for (;;) {
        double md;
        double maj = dmax(1e-10, maj_query(&mj, &md) + 2 * eps);
        double dt = dmin(-log(1.0 - rnd(c)) / maj, dmin(d, md));
        x = vadd(x, vs(w, dt));
        t += dt;
        if (dt >= md) { maj_cross(&mj); if (t >= d - CROSS_EPS) break; continue; }
        maj_move(&mj, dt);
        if (t >= d - 1e-7) break;
        double s = sigma_at(c, x);
        double Pc = (s + eps) / maj;
        if (rnd(c) < Pc) {
            double tr = C / Pc;
            if (consistent) tr *= Wpre;
            sigma_bw(c, x, dL_dC * tr * (1.0 / maj));
            break;
        } else {
            double q = (dL_dT * out_T + dL_dC * C * s_t) / ((1 - s / maj) / (1 - Pc)) / (1 - Pc);
            sigma_bw(c, x, q * (-1.0 / maj));
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Single scattered radiance: include/maps/single_scattered_radiance_mc.h     */
/* ------------------------------------------------------------------------- */
static v3 ss_radiance(ctx_t* c, v3 x, v3 w) {
    v3 wo;
    v3 env = env_sample(c, &wo);
    double ph = phase_eval(c->M, x, vs(w, -1.0), wo);
    v3 sa = albedo_at(c, x);
    double T = transmittance(c, x, wo);
    return vs(vmul(env, sa), ph * T);
}

static void ss_radiance_bw(ctx_t* c, v3 x, v3 w, v3 dL_dR) {
    rng_t saved = c->rng;
    v3 wo;
    v3 env = env_sample(c, &wo);
    double ph = phase_eval(c->M, x, vs(w, -1.0), wo);
    v3 sa = albedo_at(c, x);
    rng_t before_T = c->rng;
    double T = transmittance(c, x, wo);
    v3 dL_dsa = vs(vmul(dL_dR, env), ph * T);
    double dL_dT = vdot(dL_dR, vs(vmul(env, sa), ph));
    albedo_bw(c, x, dL_dsa);
    if (c->bw_mode == VR_BW_CONSISTENT) {
        /* transmittance gradient along the sampled ray wo, replaying its randoms */
        rng_t after = c->rng;
        c->rng = before_T;
        transmittance_bw(c, x, wo, dL_dT);
        c->rng = after;
    } else {
        /* single_scattered_radiance_mc.h:37 differentiates T(x, w) */
        transmittance_bw(c, x, w, dL_dT);
    }
    (void)saved;
}

/* include/maps/collision_integrator_mc.h:10-44 with exitance = SS radiance.
 * R = C * SS(x_t, -w) + T * env.  The exitance term is skipped when C == 0: it is
 * multiplied by 0 and nothing after it in the sample uses the random stream. */
static v3 mc_collision(ctx_t* c, v3 x, v3 w) {
    double T, t;
    double C = collision_dt(c, x, w, &t, &T);
    v3 xt = vadd(vs(w, t), x);
    v3 R = vmul(env_eval(c->M, w), V(T, T, T));
    if (C != 0.0) R = vadd(R, vs(ss_radiance(c, xt, vs(w, -1.0)), C));
    return R;
}

static void mc_collision_bw(ctx_t* c, v3 x, v3 w, v3 dL_dR) {
    v3 env = env_eval(c->M, w);
    if (c->bw_mode == VR_BW_REFERENCE) {
        /* collision_integrator_mc.h:30-44: forward collision + exitance, then the
         * collision sampler's own defensive backward with dL_dC = <dL_dR, eR>. */
        double T, t;
        double C = collision_dt(c, x, w, &t, &T);
        v3 xt = vadd(vs(w, t), x);
        v3 eR = ss_radiance(c, xt, vs(w, -1.0));
        double dL_dC = vdot(dL_dR, eR);
        double dL_dT = vdot(dL_dR, env);
        collision_dt_bw(c, x, w, c->M->ds_epsilon, dL_dC, dL_dT);
        ss_radiance_bw(c, xt, vs(w, -1.0), vs(dL_dR, C));
        return;
    }
    /* Consistent: the exitance is estimated at the collision the backward's own
     * defensive primal pass samples, and its pathwise gradient carries that
     * collision's path weight. */
    double tmin, tmax;
    if (!boundary(c->M, x, w, &tmin, &tmax)) return;
    rng_t saved = c->rng;
    double C, t, Wc;
    collision_ds_primal(c, x, w, c->M->ds_epsilon, tmin, tmax, &C, &t, &Wc);
    v3 xt = vadd(vs(w, tmin + t), x);
    v3 eR = V(0, 0, 0);
    if (C != 0.0) eR = ss_radiance(c, xt, vs(w, -1.0));
    rng_t after = c->rng;
    double dL_dC = vdot(dL_dR, eR);
    double dL_dT = vdot(dL_dR, env);
    c->rng = saved;
    collision_dt_bw(c, x, w, c->M->ds_epsilon, dL_dC, dL_dT);
    c->rng = after;
    if (C != 0.0) ss_radiance_bw(c, xt, vs(w, -1.0), vs(dL_dR, Wc));
}

/* ------------------------------------------------------------------------- */
/* NEE path integrator: include/vr_path_integrator_NEE.h:1-66                 */
/* ------------------------------------------------------------------------- */
static v3 nee_fw(ctx_t* c, v3 x, v3 w) {
    const vr_medium* M = c->M;
    double tmin, tmax;
    boundary(M, x, w, &tmin, &tmax);
    x = vadd(x, vs(w, tmin));
    double d = tmax - tmin;
    v3 R = V(0, 0, 0), W = V(1, 1, 1);
    R = vadd(R, vs(env_eval(M, w), transmittance(c, x, w)));
    const double EPS = 0.00001;
    int bounces = 0;
    maj_t mj;
    maj_begin(M, &mj, x, w);
    while (d > EPS) {
        double md;
        double m = dmax(EPS, maj_query(&mj, &md));
        double t = -log(1.0 - rnd(c)) / m;
        x = vadd(x, vs(w, dmin(t, md)));
        if (t > md) { d -= md; maj_cross(&mj); if (d <= CROSS_EPS) break; continue; }
        if (t > d) break;
        maj_move(&mj, t);
        double s = sigma_at(c, x);
        if (rnd(c) >= s / m) { d -= t; continue; }
        v3 sa = albedo_at(c, x);
        v3 em = vf(M->emission_c);
        R = vadd(R, vmul(vmul(W, V(1 - sa.x, 1 - sa.y, 1 - sa.z)), em));
        W = vmul(W, sa);
        v3 wnee;
        v3 env = env_sample(c, &wnee);
        double Tn = transmittance(c, x, wnee);
        R = vadd(R, vs(vmul(W, env), Tn * phase_eval(M, x, w, wnee)));
        bounces++;
        if (M->max_bounces >= 0 && bounces >= M->max_bounces) return R;
        v3 wo;
        double rho = phase_sample(c, x, w, &wo);
        W = vs(W, rho);
        if (W.x < EPS && W.y < EPS && W.z < EPS) return R;
        w = wo;
        boundary(M, x, w, &tmin, &d);
        maj_begin(M, &mj, x, w);
    }
    return R;
}

/* include/vr_path_integrator_NEE_DS.h:6-206 (VR_BW_REFERENCE: replay majorant
 * 1.2 m + 2 eps as in :120; VR_BW_CONSISTENT: the primal majorant in both). */
static void nee_ds_bw(ctx_t* c, v3 x, v3 w, double eps, v3 dL_dR) {
    const vr_medium* M = c->M;
    rng_t before_path = c->rng;
    v3 x0 = x, w0 = w;
    double tmin, tmax;
    boundary(M, x, w, &tmin, &tmax);
    x = vadd(x, vs(w, tmin));
    double d = tmax - tmin;
    v3 R = V(0, 0, 0), R_after_sp = V(0, 0, 0), W = V(1, 1, 1);
    double null_W = 1.0;
    const double EPS = 0.00000001;
    int scattered = 0, bounces = 0;
    maj_t mj;
    maj_begin(M, &mj, x, w);
    while (d > EPS) {
        double md;
        double m = dmax(EPS, maj_query(&mj, &md) + 2 * eps + 0.1);
        double t = -log(1.0 - rnd(c)) / m;
        x = vadd(x, vs(w, dmin(t, md)));
        if (t > md) { d -= md; maj_cross(&mj); if (d <= CROSS_EPS) break; continue; }
        if (t > d) break;
        maj_move(&mj, t);
        double s = sigma_at(c, x);
        double Pc = (s + eps) / m;
        if (rnd(c) >= Pc) { d -= t; W = vs(W, (1 - s / m) / (1 - Pc)); continue; }
        scattered = 1;
        if (s < EPS) {
            if (null_W < EPS) { W = V(0, 0, 0); break; }
            null_W = 0;
        } else W = vs(W, (s / m) / Pc);
        v3 em = vf(M->emission_c);
        v3 sa = albedo_at(c, x);
        v3 added = vmul(vmul(W, V(1 - sa.x, 1 - sa.y, 1 - sa.z)), em);
        W = vmul(W, sa);
        v3 wnee;
        v3 env = env_sample(c, &wnee);
        double Tn = transmittance(c, x, wnee);
        double phn = phase_eval(M, x, w, wnee);
        added = vadd(added, vs(vmul(W, env), Tn * phn));
        if (null_W < EPS) R_after_sp = vadd(R_after_sp, added);
        else R = vadd(R, added);
        bounces++;
        if (M->max_bounces >= 0 && bounces >= M->max_bounces) { W = V(0, 0, 0); break; }
        v3 wo;
        double rho = phase_sample(c, x, w, &wo);
        W = vs(W, rho);
        if (W.x < EPS && W.y < EPS && W.z < EPS) { W = V(0, 0, 0); break; }
        w = wo;
        boundary(M, x, w, &tmin, &d);
        maj_begin(M, &mj, x, w);
    }
    if (!scattered) R = vadd(R, vmul(env_eval(M, w), W));

    /* replay */
    c->rng = before_path;
    x = x0; w = w0;
    W = V(1, 1, 1);
    boundary(M, x, w, &tmin, &tmax);
    x = vadd(x, vs(w, tmin));
    d = tmax - tmin;
    bounces = 0;
    maj_begin(M, &mj, x, w);
    while (d > EPS) {
        double md;
        double mq = maj_query(&mj, &md);
        double m = c->bw_mode == VR_BW_REFERENCE ? dmax(EPS, mq * 1.2 + 2 * eps) : dmax(EPS, mq + 2 * eps + 0.1);
        double t = -log(1.0 - rnd(c)) / m;
        x = vadd(x, vs(w, dmin(t, md)));
        if (t > md) { d -= md; maj_cross(&mj); if (d <= CROSS_EPS) break; continue; }
        if (t > d) break;
        maj_move(&mj, t);
        double s = sigma_at(c, x);
        double Pc = (s + eps) / m;
        if (rnd(c) >= Pc) {
            sigma_bw(c, x, vdot(dL_dR, vs(R, -1.0 / m / (1 - s / m))));
            W = vs(W, (1 - s / m) / (1 - Pc));
            d -= t;
            continue;
        }
        double gs = s < EPS ? vdot(dL_dR, vs(R_after_sp, 1.0 / eps)) : vdot(dL_dR, vs(R, 1.0 / s));
        sigma_bw(c, x, gs);
        W = vs(W, s / m / Pc);
        if (s < EPS) return;
        v3 em = vf(M->emission_c);
        v3 sa = albedo_at(c, x);
        v3 dL_dsa = vmul(vmul(dL_dR, W), vs(em, -1.0));
        R = vsub(R, vmul(vmul(W, V(1 - sa.x, 1 - sa.y, 1 - sa.z)), em));
        dL_dsa = vadd(dL_dsa, vmul(dL_dR, V(R.x / dmax(EPS, sa.x), R.y / dmax(EPS, sa.y), R.z / dmax(EPS, sa.z))));
        rng_t before_env = c->rng;
        v3 wnee;
        v3 env = env_sample(c, &wnee);
        rng_t before_T = c->rng;
        double Tn = transmittance(c, x, wnee);
        double phn = phase_eval(M, x, w, wnee);
        rng_t before_ps = c->rng;
        v3 wo;
        double rho = phase_sample(c, x, w, &wo);
        rng_t before_next = c->rng;
        W = vmul(W, sa);
        double dL_dTn = vdot(dL_dR, vs(vmul(W, env), phn));
        R = vsub(R, vs(vmul(W, env), Tn * phn));
        albedo_bw(c, x, dL_dsa);
        c->rng = before_T;
        transmittance_bw(c, x, wnee, dL_dTn);
        (void)before_env; (void)before_ps;
        c->rng = before_next;
        bounces++;
        if (M->max_bounces >= 0 && bounces >= M->max_bounces) return;
        W = vs(W, rho);
        if (W.x < EPS && W.y < EPS && W.z < EPS) return;
        w = wo;
        boundary(M, x, w, &tmin, &d);
        maj_begin(M, &mj, x, w);
    }
}

/* ------------------------------------------------------------------------- */
/* Delta-tracking path integrator: include/vr_path_integrator_DT.h            */
/* ------------------------------------------------------------------------- */
/* vr_path_integrator_DT.h:7-56 (max_bounces >= 0 ends a path after that many
 * scatter events, with no escape term). */
static v3 path_dt_fw(ctx_t* c, v3 x, v3 w) {
    const vr_medium* M = c->M;
    double tmin, tmax;
    boundary(M, x, w, &tmin, &tmax);
    x = vadd(x, vs(w, tmin));
    double d = tmax - tmin;
    v3 R = V(0, 0, 0), W = V(1, 1, 1);
    int bounces = 0;
    maj_t mj;
    maj_begin(M, &mj, x, w);
    while (d > 0.00000001) {
        double md;
        double m = dmax(0.00000001, maj_query(&mj, &md));
        double t = -log(1.0 - rnd(c)) / m;
        x = vadd(x, vs(w, dmin(t, md)));
        if (t > md) { d -= md; maj_cross(&mj); if (d <= CROSS_EPS) break; continue; }
        if (t > d) break;
        maj_move(&mj, t);
        double s = sigma_at(c, x);
        if (rnd(c) >= s / m) { d -= t; continue; }
        v3 sa = albedo_at(c, x);
        v3 em = vf(M->emission_c);
        R = vadd(R, vmul(vmul(W, V(1 - sa.x, 1 - sa.y, 1 - sa.z)), em));
        W = vmul(W, sa);
        v3 wo;
        double rho = phase_sample(c, x, w, &wo);
        W = vs(W, rho);
        bounces++;
        if (M->max_bounces >= 0 && bounces >= M->max_bounces) return R;
        if (W.x < 0.0000001 && W.y < 0.0000001 && W.z < 0.0000001) return R;
        w = wo;
        boundary(M, x, w, &tmin, &d);
        maj_begin(M, &mj, x, w);
    }
    return vadd(vmul(W, env_eval(M, w)), R);
}

/* vr_path_integrator_DT.h:59-235: a defensive-sampling pass (majorant + 2 eps,
 * collision probability (s + eps)/m) that computes the path throughput, then a
 * replay from the same stream that scatters d<W env>/d sigma at every null and
 * real collision.  The gradient of the emission term is not propagated (the
 * reference's emission/albedo backward lines are commented out). */
static void path_dt_bw(ctx_t* c, v3 x, v3 w, double eps, v3 dL_dR) {
    const vr_medium* M = c->M;
    rng_t before_path = c->rng;
    v3 x0 = x, w0 = w;
    double tmin, tmax;
    boundary(M, x, w, &tmin, &tmax);
    x = vadd(x, vs(w, tmin));
    double d = tmax - tmin;
    const double EPS = 0.00000001;
    v3 W = V(1, 1, 1);
    double null_W = 1.0;
    int bounces = 0;
    maj_t mj;
    maj_begin(M, &mj, x, w);
    while (d > EPS) {
        double md;
        double m = dmax(EPS, maj_query(&mj, &md) + 2 * eps);
        double t = -log(1.0 - rnd(c)) / m;
        x = vadd(x, vs(w, dmin(t, md)));
        if (t > md) { d -= md; maj_cross(&mj); if (d <= CROSS_EPS) break; continue; }
        if (t > d) break;
        maj_move(&mj, t);
        double s = sigma_at(c, x);
        double Pc = (s + eps) / m;
        if (rnd(c) >= Pc) { d -= t; W = vs(W, (1 - s / m) / (1 - Pc)); continue; }
        if (s == 0.0) {
            if (null_W == 0.0) { W = V(0, 0, 0); break; }
            null_W = 0.0;
        } else W = vs(W, (s / m) / Pc);
        v3 sa = albedo_at(c, x);
        v3 wo;
        double rho = phase_sample(c, x, w, &wo);
        W = vs(vmul(W, sa), rho);
        bounces++;
        if (M->max_bounces >= 0 && bounces >= M->max_bounces) { W = V(0, 0, 0); break; }
        if (W.x < EPS && W.y < EPS && W.z < EPS) { W = V(0, 0, 0); break; }
        w = wo;
        boundary(M, x, w, &tmin, &d);
        maj_begin(M, &mj, x, w);
    }
    v3 dL_dW = vmul(dL_dR, env_eval(M, w));
    v3 W_out = W;

    c->rng = before_path;
    x = x0; w = w0;
    boundary(M, x, w, &tmin, &tmax);
    x = vadd(x, vs(w, tmin));
    d = tmax - tmin;
    bounces = 0;
    maj_begin(M, &mj, x, w);
    while (d > EPS) {
        double md;
        double m = dmax(EPS, maj_query(&mj, &md) + 2 * eps);
        double t = -log(1.0 - rnd(c)) / m;
        x = vadd(x, vs(w, dmin(t, md)));
        if (t > md) { d -= md; maj_cross(&mj); if (d <= CROSS_EPS) break; continue; }
        if (t > d) break;
        maj_move(&mj, t);
        double s = sigma_at(c, x);
        double Pc = (s + eps) / m;
        if (rnd(c) >= Pc) {
            sigma_bw(c, x, vdot(dL_dW, vs(W_out, -null_W / m / (1 - s / m))));
            d -= t;
            continue;
        }
        sigma_bw(c, x, vdot(dL_dW, s == 0.0 ? W_out : vs(W_out, null_W / s)));
        if (s == 0.0) break;
        v3 wo;
        phase_sample(c, x, w, &wo);
        bounces++;
        if (M->max_bounces >= 0 && bounces >= M->max_bounces) return;
        w = wo;
        boundary(M, x, w, &tmin, &d);
        maj_begin(M, &mj, x, w);
    }
}

/* ------------------------------------------------------------------------- */
/* NEE with reservoir backward: include/vr_path_integrator_NEE_DRT.h           */
/* ------------------------------------------------------------------------- */
/* randoms.h:131-139: a seed built from 4 uniforms of the current stream. */
static rng_t create_branch_seed(ctx_t* c) {
    rng_t s;
    s.x = float_to_bits(rng_random(&c->rng)) + 129u;
    s.y = float_to_bits(rng_random(&c->rng)) + 129u;
    s.z = float_to_bits(rng_random(&c->rng)) + 129u;
    s.w = float_to_bits(rng_random(&c->rng)) + 129u;
    for (int k = 0; k < 4; k++) rng_advance(&s);
    return s;
}

/* vr_path_integrator_NEE.h:1-66 with only_scatters = true (no primary
 * transmittance term): the continuation path of the reservoir vertex. */
static v3 nee_scatters(ctx_t* c, v3 x, v3 w) {
    const vr_medium* M = c->M;
    double tmin, tmax;
    boundary(M, x, w, &tmin, &tmax);
    x = vadd(x, vs(w, tmin));
    double d = tmax - tmin;
    v3 R = V(0, 0, 0), W = V(1, 1, 1);
    const double EPS = 0.00001;
    int bounces = 0;
    maj_t mj;
    maj_begin(M, &mj, x, w);
    while (d > EPS) {
        double md;
        double m = dmax(EPS, maj_query(&mj, &md));
        double t = -log(1.0 - rnd(c)) / m;
        x = vadd(x, vs(w, dmin(t, md)));
        if (t > md) { d -= md; maj_cross(&mj); if (d <= CROSS_EPS) break; continue; }
        if (t > d) break;
        maj_move(&mj, t);
        double s = sigma_at(c, x);
        if (rnd(c) >= s / m) { d -= t; continue; }
        v3 sa = albedo_at(c, x);
        v3 em = vf(M->emission_c);
        R = vadd(R, vmul(vmul(W, V(1 - sa.x, 1 - sa.y, 1 - sa.z)), em));
        W = vmul(W, sa);
        v3 wnee;
        v3 env = env_sample(c, &wnee);
        double Tn = transmittance(c, x, wnee);
        R = vadd(R, vs(vmul(W, env), Tn * phase_eval(M, x, w, wnee)));
        bounces++;
        if (M->max_bounces >= 0 && bounces >= M->max_bounces) return R;
        v3 wo;
        double rho = phase_sample(c, x, w, &wo);
        W = vs(W, rho);
        if (W.x < EPS && W.y < EPS && W.z < EPS) return R;
        w = wo;
        boundary(M, x, w, &tmin, &d);
        maj_begin(M, &mj, x, w);
    }
    return R;
}

/* vr_path_integrator_NEE_DRT.h:8-44: ratio-tracking walk over [0, d) that picks
 * one free-flight segment with probability proportional to T * dt and a
 * uniform position inside it; weight = sum of T * dt (<T>/pdf).  The majorant
 * is queried along the walk (the reference queries it at the start point,
 * which is the same for a constant majorant), and the end of the segment is
 * tested before a cell crossing. */
static double sample_transmittance_position(ctx_t* c, v3 x, v3 w, double d, double* sample_t, double* weight_t) {
    const vr_medium* M = c->M;
    double T = 1.0, t = 0.0, res_t = 0.0, res_dt = 0.0, C = 0.0;
    const double EPS = 0.0000001;
    maj_t mj;
    maj_begin(M, &mj, x, w);
    for (;;) {
        double md;
        double m = dmax(EPS, maj_query(&mj, &md));
        double dt = dmin(-log(1.0 - rnd(c)) / m, dmin(md, d - t));
        double w_step = T * dt;
        C += w_step;
        if (rnd(c) < w_step / C) { res_t = t; res_dt = dt; }
        t += dt;
        /* end of the segment first: with a majorant grid the last cell's exit
         * can coincide with it (the reference's constant majorant never crosses) */
        if (t >= d - EPS) break;
        if (dt >= md - EPS) { maj_cross(&mj); continue; }
        maj_move(&mj, dt);
        T *= 1 - sigma_at(c, vadd(x, vs(w, t))) / m;
    }
    *weight_t = C;
    *sample_t = res_t + rnd(c) * res_dt;
    return T;
}

/* vr_path_integrator_NEE_DRT.h:46-269.  The secondary stream (a branch seed
 * drawn from the main stream) drives every transmittance backward; the main
 * stream drives the primal pass, its replay and the reservoir-vertex term. */
static void nee_drt_bw(ctx_t* c, v3 x, v3 w, v3 dL_dR) {
    const vr_medium* M = c->M;
    double tmin, tmax;
    boundary(M, x, w, &tmin, &tmax);
    x = vadd(x, vs(w, tmin));
    double d = tmax - tmin;
    if (d <= 0) return;
    rng_t secondary = create_branch_seed(c), main_s;

    double direct_T = transmittance(c, x, w);
    (void)direct_T;  /* dL/denv = dL_dR * direct_T: a constant environment has no parameters */
    double dL_ddT = vdot(dL_dR, env_eval(M, w));
    main_s = c->rng; c->rng = secondary;
    transmittance_bw(c, x, w, dL_ddT);
    secondary = c->rng; c->rng = main_s;

    v3 R = V(0, 0, 0), W = V(1, 1, 1);
    v3 res_x = x, res_w = w, res_W = W;
    double res_P = 3, total_weight = 3;
    v3 ini_x = x, ini_w = w;
    double ini_d = d;
    const double EPS = 0.00000001;
    rng_t before_path = c->rng;
    int bounces = 0;
    maj_t mj;
    maj_begin(M, &mj, x, w);
    while (d > EPS) {
        double md;
        double m = maj_query(&mj, &md) * 1.05 + 0.1;
        double t = -log(1.0 - rnd(c)) / m;
        x = vadd(x, vs(w, dmin(t, md)));
        if (t > md) { d -= md; maj_cross(&mj); if (d <= CROSS_EPS) break; continue; }
        if (t > d) break;
        maj_move(&mj, t);
        double s = sigma_at(c, x);
        if (rnd(c) >= s / m) { d -= t; continue; }
        v3 sa = albedo_at(c, x);
        W = vmul(W, sa);
        v3 wnee;
        v3 env = env_sample(c, &wnee);
        double Tn = transmittance(c, x, wnee);
        double phn = phase_eval(M, x, w, wnee);
        R = vadd(R, vs(vmul(W, env), Tn * phn));
        bounces++;
        if (M->max_bounces >= 0 && bounces >= M->max_bounces) break;
        v3 wo;
        double rho = phase_sample(c, x, w, &wo);
        W = vs(W, rho);
        w = wo;
        boundary(M, x, w, &tmin, &d);
        maj_begin(M, &mj, x, w);
        double cur_P = W.x + W.y + W.z;
        total_weight += cur_P;
        if (rnd(c) < cur_P / total_weight) { res_P = cur_P; res_W = W; res_x = x; res_w = w; }
        if (W.x < EPS && W.y < EPS && W.z < EPS) break;
    }

    /* replay: null collisions and the MIS share s^2/(1+s^2) of real collisions */
    W = V(1, 1, 1);
    x = ini_x; w = ini_w; d = ini_d;
    c->rng = before_path;
    bounces = 0;
    maj_begin(M, &mj, x, w);
    while (d > EPS) {
        double md;
        double m = maj_query(&mj, &md) * 1.05 + 0.1;
        double t = -log(1.0 - rnd(c)) / m;
        x = vadd(x, vs(w, dmin(t, md)));
        if (t > md) { d -= md; maj_cross(&mj); if (d <= CROSS_EPS) break; continue; }
        if (t > d) break;
        maj_move(&mj, t);
        double s = sigma_at(c, x);
        double Pc = s / m;
        if (rnd(c) >= Pc) {
            double g = vdot(dL_dR, vs(R, -1.0 / m / (1 - s / m)));
            main_s = c->rng; c->rng = secondary;
            sigma_bw(c, x, g);
            secondary = c->rng; c->rng = main_s;
            d -= t;
            continue;
        }
        double g = vdot(dL_dR, vs(R, (s * s / (1 + s * s)) / s));
        v3 sa = albedo_at(c, x);
        W = vmul(W, sa);
        v3 wnee;
        v3 env = env_sample(c, &wnee);
        double Tn = transmittance(c, x, wnee);
        double phn = phase_eval(M, x, w, wnee);
        bounces++;
        const int last = M->max_bounces >= 0 && bounces >= M->max_bounces;
        v3 wo = w;
        double rho = last ? 1.0 : phase_sample(c, x, w, &wo);
        double dL_dTn = vdot(dL_dR, vs(vmul(W, env), phn));
        R = vsub(R, vs(vmul(W, env), Tn * phn));
        main_s = c->rng; c->rng = secondary;
        sigma_bw(c, x, g);
        transmittance_bw(c, x, wnee, dL_dTn);
        secondary = c->rng; c->rng = main_s;
        if (last) break;
        W = vs(W, rho);
        w = wo;
        boundary(M, x, w, &tmin, &d);
        maj_begin(M, &mj, x, w);
        rnd(c);  /* the reservoir draw of the primal pass */
        if (W.x < EPS && W.y < EPS && W.z < EPS) break;
    }

    /* reservoir vertex: the remaining 1/(1+s^2) share of the collision term */
    boundary(M, res_x, res_w, &tmin, &d);
    double s_t, weight_t;
    sample_transmittance_position(c, res_x, res_w, d, &s_t, &weight_t);
    v3 xs = vadd(res_x, vs(res_w, s_t));
    v3 lw;
    v3 Lenv = env_sample(c, &lw);
    v3 Ls = vs(Lenv, transmittance(c, xs, lw) * phase_eval(M, xs, res_w, lw));
    v3 ws;
    double rho_s = phase_sample(c, xs, res_w, &ws);
    Ls = vadd(Ls, vs(nee_scatters(c, xs, ws), rho_s));
    Ls = vmul(vs(vmul(Ls, res_W), weight_t), albedo_at(c, xs));
    double sxs = sigma_at(c, xs);
    double g = vdot(dL_dR, Ls) * (1 / (1 + sxs * sxs)) * total_weight / res_P;
    main_s = c->rng; c->rng = secondary;
    sigma_bw(c, xs, g);
    secondary = c->rng; c->rng = main_s;
}

/* ------------------------------------------------------------------------- */
/* Medium path integrator (7 -> 12): include/sc_medium_path_integrator_DT_*.h  */
/* ------------------------------------------------------------------------- */
static void medium_dt_fw(ctx_t* c, v3 x, v3 w, double d, double* out) {
    const vr_medium* M = c->M;
    v3 W = V(1, 1, 1), A = V(0, 0, 0);
    v3 xo = vadd(vs(w, d), x), wo = w;
    double maj = dmax(0.00001, M->medium_majorant);
    for (;;) {
        double t = -log(1.0 - rnd(c)) / maj;
        if (t > d - 0.00001) break;
        x = vadd(x, vs(w, t));
        d -= t;
        double s = sigma_at(c, x);
        if (rnd(c) < s / maj) {
            if (M->medium_filter & 1) {
                xo = x;
                (void)phase_sample(c, x, w, &wo);
                W = albedo_at(c, x);
            } else W = V(0, 0, 0);
            if (M->medium_filter & 2) A = vf(M->gathering_c);
            break;
        }
    }
    out[0] = xo.x; out[1] = xo.y; out[2] = xo.z;
    out[3] = wo.x; out[4] = wo.y; out[5] = wo.z;
    out[6] = W.x; out[7] = W.y; out[8] = W.z;
    out[9] = A.x; out[10] = A.y; out[11] = A.z;
}

static void medium_dt_replay(ctx_t* c, v3 x, v3 w, double d, v3 W, v3 A, v3 dL_dW, v3 dL_dA);

static void medium_dt_bw(ctx_t* c, v3 x, v3 w, double d, v3 dL_dW, v3 dL_dA) {
    const vr_medium* M = c->M;
    v3 W = V(1, 1, 1), A = V(0, 0, 0);
    v3 xi = x, wi = w;
    double di = d;
    rng_t before = c->rng;
    double maj = dmax(0.00001, M->medium_majorant);
    int f = M->medium_filter;
    for (;;) {
        double t = -log(1.0 - rnd(c)) / maj;
        if (t > d - 0.00001) break;
        x = vadd(x, vs(w, t));
        d -= t;
        double s = sigma_at(c, x);
        if (rnd(c) < s / maj) {
            if (f & 1) { v3 wo; double wt = phase_sample(c, x, w, &wo); W = vs(albedo_at(c, x), wt); }
            else W = V(0, 0, 0);
            if (f & 2) A = vf(M->gathering_c);
            break;
        }
    }
    c->rng = before;
    medium_dt_replay(c, xi, wi, di, W, A, dL_dW, dL_dA);
}

/* The replay half of medium_dt_bw given the primal pass's (W, A). */
static void medium_dt_replay(ctx_t* c, v3 x, v3 w, double d, v3 W, v3 A, v3 dL_dW, v3 dL_dA) {
    const vr_medium* M = c->M;
    double maj = dmax(0.00001, M->medium_majorant);
    int f = M->medium_filter;
    /* `#if MEDIUM_FILTER & 1 == 1` and `& 2 == 2` both parse as `& 1`
     * (sc_medium_path_integrator_DT_bw.h:85,95,109); VR_BW_CONSISTENT uses the
     * intended bit tests. */
    int scat_blk = f & 1;
    int acc_blk = c->bw_mode == VR_BW_REFERENCE ? (f & 1) : (f & 2);
    for (;;) {
        double t = -log(1.0 - rnd(c)) / maj;
        if (t > d - 0.00001) break;
        x = vadd(x, vs(w, t));
        d -= t;
        double s = sigma_at(c, x);
        double Pc = s / maj;
        if (rnd(c) < Pc) {
            double g = 0;
            if (scat_blk) {
                v3 wo;
                (void)phase_sample(c, x, w, &wo);
                g += vdot(dL_dW, vs(W, 1.0 / s));
                albedo_bw(c, x, dL_dW);
            }
            if (acc_blk) g += vdot(dL_dA, vs(A, 1.0 / s));
            if (g != 0) sigma_bw(c, x, g);
            break;
        } else {
            double g = -vdot(dL_dW, W) / maj / (1 - Pc);
            if (acc_blk) g += -vdot(dL_dA, A) / maj / (1 - Pc);
            if (g != 0) sigma_bw(c, x, g);
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Map dispatch                                                               */
/* ------------------------------------------------------------------------- */
static int map_dims(int kind, int* in_dim, int* out_dim) {
    switch (kind) {
    case VR_MAP_RAYBOX: *in_dim = 6; *out_dim = 2; return 1;
    case VR_MAP_T_DDA: case VR_MAP_T_DT: case VR_MAP_T_RT: case VR_MAP_T_GDT: case VR_MAP_T_GRT:
        *in_dim = 6; *out_dim = 1; return 1;
    case VR_MAP_COLLISION_DT: case VR_MAP_SS_RADIANCE: case VR_MAP_MC_COLLISION: case VR_MAP_NEE_DT:
    case VR_MAP_PATH_DT: case VR_MAP_NEE_DRT:
        *in_dim = 6; *out_dim = 3; return 1;
    case VR_MAP_MEDIUM_DT: *in_dim = 7; *out_dim = 12; return 1;
    default: return 0;
    }
}

static void field_fw(ctx_t* c, int kind, const double* in, double* out) {
    v3 x = V(in[0], in[1], in[2]), w = V(in[3], in[4], in[5]);
    switch (kind) {
    case VR_MAP_RAYBOX: ray_box_intersection(x, w, vf(c->M->box_min), vf(c->M->box_max), &out[0], &out[1]); break;
    case VR_MAP_T_DDA: out[0] = t_dda(c, x, w); break;
    case VR_MAP_T_DT: out[0] = t_dt(c, x, w); break;
    case VR_MAP_T_RT: out[0] = t_rt(c, x, w); break;
    case VR_MAP_T_GDT: out[0] = t_grid_tracking(c, x, w, 0); break;
    case VR_MAP_T_GRT: out[0] = t_grid_tracking(c, x, w, 1); break;
    case VR_MAP_COLLISION_DT: { double t, T; out[0] = collision_dt(c, x, w, &t, &T); out[1] = t; out[2] = T; break; }
    case VR_MAP_SS_RADIANCE: { v3 r = ss_radiance(c, x, w); out[0] = r.x; out[1] = r.y; out[2] = r.z; break; }
    case VR_MAP_MC_COLLISION: { v3 r = mc_collision(c, x, w); out[0] = r.x; out[1] = r.y; out[2] = r.z; break; }
    case VR_MAP_NEE_DT: case VR_MAP_NEE_DRT: { v3 r = nee_fw(c, x, w); out[0] = r.x; out[1] = r.y; out[2] = r.z; break; }
    case VR_MAP_PATH_DT: { v3 r = path_dt_fw(c, x, w); out[0] = r.x; out[1] = r.y; out[2] = r.z; break; }
    case VR_MAP_MEDIUM_DT: medium_dt_fw(c, x, w, in[6], out); break;
    default: break;
    }
}

static void field_bw(ctx_t* c, int kind, const double* in, const double* og) {
    v3 x = V(in[0], in[1], in[2]), w = V(in[3], in[4], in[5]);
    switch (kind) {
    case VR_MAP_T_DDA: t_dda_bw(c, x, w, og[0]); break;
    case VR_MAP_T_DT: t_dt_bw(c, x, w, og[0]); break;
    case VR_MAP_T_RT: t_rt_bw(c, x, w, og[0]); break;
    case VR_MAP_COLLISION_DT: collision_dt_bw(c, x, w, c->M->ds_epsilon, og[0], og[2]); break;
    case VR_MAP_SS_RADIANCE: ss_radiance_bw(c, x, w, V(og[0], og[1], og[2])); break;
    case VR_MAP_MC_COLLISION: mc_collision_bw(c, x, w, V(og[0], og[1], og[2])); break;
    case VR_MAP_NEE_DT: nee_ds_bw(c, x, w, c->M->ds_epsilon, V(og[0], og[1], og[2])); break;
    case VR_MAP_PATH_DT: path_dt_bw(c, x, w, c->M->ds_epsilon, V(og[0], og[1], og[2])); break;
    case VR_MAP_NEE_DRT: nee_drt_bw(c, x, w, V(og[0], og[1], og[2])); break;
    case VR_MAP_MEDIUM_DT: medium_dt_bw(c, x, w, in[6], V(og[6], og[7], og[8]), V(og[9], og[10], og[11])); break;
    default: break;
    }
}

static int check_medium(const vr_medium* M, int kind) {
    if (!M) return fail(VR_ERR_ARG, "medium is NULL");
    int id = 0, od = 0;
    if (!map_dims(kind, &id, &od)) return fail(VR_ERR_UNSUPPORTED, "unknown map kind");
    if (kind != VR_MAP_RAYBOX) {
        if (!M->sigma.t.data || M->sigma.t.ndim != 4 || M->sigma.t.shape[3] != 1)
            return fail(VR_ERR_ARG, "sigma must be a [D,H,W,1] grid");
        for (int a = 0; a < 3; a++) if (M->sigma.t.shape[a] < 2) return fail(VR_ERR_ARG, "sigma grid needs >= 2 vertices per axis");
    }
    if (M->majorant_kind == VR_MAJ_GRID && !M->majorant_grid) return fail(VR_ERR_ARG, "majorant grid is NULL");
    if (M->env_kind == VR_ENV_XR && (!M->env_image.data || M->env_image.ndim != 3 || M->env_image.shape[2] != 3 ||
                                     M->env_image.shape[0] < 1 || M->env_image.shape[1] < 1))
        return fail(VR_ERR_ARG, "environment image must be [H, W, 3]");
    if (M->env_kind != VR_ENV_XR && M->env_kind != VR_ENV_CONST) return fail(VR_ERR_ARG, "unknown environment kind");
    if (M->env_sampler_kind == VR_ENVS_QUADTREE &&
        (!M->env_quadtree || M->env_quadtree_levels < 1 || M->env_quadtree_levels > 15))
        return fail(VR_ERR_ARG, "bad environment quadtree");
    if (M->env_sampler_kind != VR_ENVS_QUADTREE && M->env_sampler_kind != VR_ENVS_UNIFORM)
        return fail(VR_ERR_ARG, "unknown environment sampler kind");
    if (M->albedo_kind == VR_FIELD_GRID && (!M->albedo.t.data || M->albedo.t.ndim != 4))
        return fail(VR_ERR_ARG, "albedo grid must be [D,H,W,C]");
    if (M->phase_g_kind != VR_FIELD_CONST && M->phase_g_kind != VR_FIELD_GRID)
        return fail(VR_ERR_ARG, "unknown phase_g kind");
    if (M->phase_g_kind == VR_FIELD_GRID && M->phase_kind != VR_PHASE_ISO &&
        (!M->phase_g_grid.t.data || M->phase_g_grid.t.ndim != 4 || M->phase_g_grid.t.shape[3] != 1))
        return fail(VR_ERR_ARG, "phase_g grid must be [D,H,W,1]");
    return VR_OK;
}

/* Private double gradient buffers per thread, reduced into the float grads. */
typedef struct { int64_t sig_n, alb_n; } gsize_t;

static int64_t grid_span(const vr_grid* g) {
    const vr_tensor* t = &g->t;
    int64_t span = 1;
    for (int i = 0; i < 4; i++) span += (t->shape[i] - 1) * t->stride[i];
    return span;
}

/* ------------------------------------------------------------------------- */
/* Exported entry points                                                      */
/* ------------------------------------------------------------------------- */
int oracle_grid_fw(const vr_grid* g, const float* x, int64_t n, float* out) {
    if (!g || !x || !out || g->t.ndim != 4) return fail(VR_ERR_ARG, "bad grid arguments");
    int C = (int)g->t.shape[3];
    for (int64_t i = 0; i < n; i++) {
        double v[64];
        grid_lookup(g, vf(x + 3 * i), v);
        for (int k = 0; k < C; k++) out[i * C + k] = (float)v[k];
    }
    return VR_OK;
}

int oracle_grid_fw_f64(const vr_grid* g, const float* x, int64_t n, double* out) {
    if (!g || !x || !out || g->t.ndim != 4) return fail(VR_ERR_ARG, "bad grid arguments");
    int C = (int)g->t.shape[3];
    for (int64_t i = 0; i < n; i++) {
        double v[64];
        grid_lookup(g, vf(x + 3 * i), v);
        for (int k = 0; k < C; k++) out[i * C + k] = v[k];
    }
    return VR_OK;
}

int oracle_grid_bw(const vr_grid* g, const float* x, const float* og, int64_t n) {
    if (!g || !x || !og || g->t.ndim != 4) return fail(VR_ERR_ARG, "bad grid arguments");
    if (!g->t.grad) return VR_OK;
    int C = (int)g->t.shape[3];
    for (int64_t i = 0; i < n; i++) {
        double o[64];
        for (int k = 0; k < C; k++) o[k] = og[i * C + k];
        grid_scatter(g, NULL, vf(x + 3 * i), o);
    }
    return VR_OK;
}

int oracle_majorant_build(const vr_grid* g, const int32_t res[3], float* out) {
    if (!g || !res || !out) return fail(VR_ERR_ARG, "bad majorant arguments");
    const vr_tensor* t = &g->t;
    int rz = res[0], ry = res[1], rx = res[2];
    int n[3] = {(int)t->shape[2] - 1, (int)t->shape[1] - 1, (int)t->shape[0] - 1};  /* cells x,y,z */
    int r[3] = {rx, ry, rz};
    for (int cz = 0; cz < rz; cz++)
        for (int cy = 0; cy < ry; cy++)
            for (int cx = 0; cx < rx; cx++) {
                int c[3] = {cx, cy, cz}, lo[3], hi[3];
                for (int a = 0; a < 3; a++) {
                    /* vertices of the fine cells overlapping [c, c+1) * n/r, widened by one
                     * vertex on each side to absorb rounding of the coarse boundaries */
                    lo[a] = imax(0, (int)floor((double)c[a] * n[a] / r[a]) - 1);
                    hi[a] = imin(n[a], (int)ceil((double)(c[a] + 1) * n[a] / r[a]) + 1);
                }
                float m = 0.0f;
                int nonfinite = 0;
                for (int z = lo[2]; z <= hi[2]; z++)
                    for (int y = lo[1]; y <= hi[1]; y++)
                        for (int x = lo[0]; x <= hi[0]; x++) {
                            float v = t->data[goff(t, z, y, x)];
                            if (!isfinite(v)) nonfinite = 1;
                            else if (v > m) m = v;
                        }
                /* a NaN bound for a footprint with a NaN or infinite density: the
                 * trackers floor the majorant, so the cell is crossed as empty */
                out[((int64_t)cz * ry + cy) * rx + cx] = nonfinite ? NAN : m;
            }
    return VR_OK;
}

static void reduce_priv(const vr_grid* g, double* priv) {
    if (!priv || !g->t.grad) return;
    int64_t span = grid_span(g);
    for (int64_t i = 0; i < span; i++) if (priv[i] != 0) g->t.grad[i] += (float)priv[i];
}

int oracle_map_fw(const vr_medium* M, int32_t kind, const float* in, int64_t n, float* out, const vr_launch* L) {
    int st = check_medium(M, kind);
    if (st) return st;
    if (!in || !out || !L) return fail(VR_ERR_ARG, "null buffer");
    int id = 0, od = 0;
    map_dims(kind, &id, &od);
    int64_t start = L->start_index, cnt = L->count < 0 ? n - start : L->count;
    #pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = start; i < start + cnt; i++) {
        ctx_t c;
        memset(&c, 0, sizeof(c));
        c.M = M; c.bw_mode = L->bw_mode;
        c.rng = seed_map(L->seeds, (uint32_t)i);
        double a[16], o[16];
        for (int k = 0; k < id; k++) a[k] = in[i * id + k];
        field_fw(&c, kind, a, o);
        for (int k = 0; k < od; k++) out[i * od + k] = (float)o[k];
        if (L->debug_out) L->debug_out[i] = (int32_t)c.lookups;
    }
    return VR_OK;
}

int oracle_map_bw(const vr_medium* M, int32_t kind, const float* in, const float* og, int64_t n, const vr_launch* L) {
    int st = check_medium(M, kind);
    if (st) return st;
    if (!in || !og || !L) return fail(VR_ERR_ARG, "null buffer");
    int id = 0, od = 0;
    map_dims(kind, &id, &od);
    int64_t start = L->start_index, cnt = L->count < 0 ? n - start : L->count;
    int64_t ssig = grid_span(&M->sigma);
    int64_t salb = M->albedo_kind == VR_FIELD_GRID ? grid_span(&M->albedo) : 0;
    #pragma omp parallel
    {
        gacc_t acc;
        acc.sig = M->sigma.t.grad ? (double*)calloc((size_t)ssig, sizeof(double)) : NULL;
        acc.alb = (salb && M->albedo.t.grad) ? (double*)calloc((size_t)salb, sizeof(double)) : NULL;
        #pragma omp for schedule(dynamic, 16)
        for (int64_t i = start; i < start + cnt; i++) {
            ctx_t c;
            memset(&c, 0, sizeof(c));
            c.M = M; c.bw_mode = L->bw_mode; c.acc = acc;
            c.rng = seed_map(L->seeds, (uint32_t)i);
            double a[16], o[16];
            for (int k = 0; k < id; k++) a[k] = in[i * id + k];
            for (int k = 0; k < od; k++) o[k] = og[i * od + k];
            field_bw(&c, kind, a, o);
        }
        #pragma omp critical
        {
            reduce_priv(&M->sigma, acc.sig);
            if (acc.alb) reduce_priv(&M->albedo, acc.alb);
        }
        free(acc.sig);
        free(acc.alb);
    }
    return VR_OK;
}

/* include/maps/sensor_perspective_camera.h:1-27 */
static void camera_ray(const vr_camera* cam, int pose, int row, int col, ctx_t* c, v3* xo, v3* wo) {
    const float* P = cam->poses + 9 * pose;
    v3 o = vf(P), d = vf(P + 3), n = vf(P + 6);
    double u = 0.5, v = 0.5;
    if (cam->jittered) { u = rnd(c); v = rnd(c); }
    double sx = ((col + u) * 2 - cam->width) * cam->znear / cam->height;
    double sy = ((row + v) * 2 - cam->height) * cam->znear / cam->height;
    double sz = cam->znear / tan(cam->fov * 0.5);
    v3 za = vnorm(d);
    v3 xa = vnorm(vcross(n, za));
    v3 ya = vcross(za, xa);
    v3 w = vadd(vadd(vs(xa, sx), vs(ya, sy)), vs(za, sz));
    *xo = vadd(o, w);
    *wo = vnorm(w);
}

static void sensor_index(const vr_camera* cam, int64_t i, int* pose, int* row, int* col) {
    if (cam->sensors) {
        *pose = (int)cam->sensors[3 * i]; *row = (int)cam->sensors[3 * i + 1]; *col = (int)cam->sensors[3 * i + 2];
    } else {
        *col = (int)(i % cam->width);
        *row = (int)((i / cam->width) % cam->height);
        *pose = (int)(i / ((int64_t)cam->width * cam->height));
    }
}

static int check_camera(const vr_camera* cam) {
    if (!cam || !cam->poses || cam->width <= 0 || cam->height <= 0 || cam->num_poses <= 0)
        return fail(VR_ERR_ARG, "bad camera");
    return VR_OK;
}

static int64_t camera_threads(const vr_camera* cam) {
    return cam->sensors ? cam->num_sensors : (int64_t)cam->num_poses * cam->width * cam->height;
}

int oracle_capture_fw(const vr_medium* M, int32_t kind, const vr_camera* cam, const vr_launch* L, float* out) {
    int st = check_medium(M, kind);
    if (st) return st;
    if ((st = check_camera(cam))) return st;
    if (!L || !out) return fail(VR_ERR_ARG, "null buffer");
    int id = 0, od = 0;
    map_dims(kind, &id, &od);
    if (id != 6) return fail(VR_ERR_UNSUPPORTED, "capture needs a 6-input map");
    int64_t total = camera_threads(cam);
    int64_t start = L->start_index, cnt = L->count < 0 ? total - start : L->count;
    int spp = L->samples < 1 ? 1 : L->samples;
    #pragma omp parallel for schedule(dynamic, 4)
    for (int64_t i = start; i < start + cnt; i++) {
        int pose, row, col;
        sensor_index(cam, i, &pose, &row, &col);
        double acc[16] = {0};
        int64_t lk = 0;
        if (pose < 0 || pose >= cam->num_poses) {  /* outside the pose buffer: 0 (not read) */
            for (int k = 0; k < od; k++) out[i * od + k] = 0.0f;
            if (L->debug_out) L->debug_out[i] = 0;
            continue;
        }
        for (int s = 0; s < spp; s++) {
            ctx_t c;
            memset(&c, 0, sizeof(c));
            c.M = M; c.bw_mode = L->bw_mode;
            c.rng = seed_capture_fw(L->seeds, (uint32_t)((int64_t)s * total + i));
            v3 x, w;
            camera_ray(cam, pose, row, col, &c, &x, &w);
            double a[6] = {x.x, x.y, x.z, w.x, w.y, w.z}, o[16];
            field_fw(&c, kind, a, o);
            for (int k = 0; k < od; k++) acc[k] += o[k];
            lk += c.lookups;
        }
        for (int k = 0; k < od; k++) out[i * od + k] = (float)(acc[k] / spp);
        if (L->debug_out) L->debug_out[i] = (int32_t)lk;
    }
    return VR_OK;
}

int oracle_capture_bw(const vr_medium* M, int32_t kind, const vr_camera* cam, const vr_launch* L, const float* og) {
    int st = check_medium(M, kind);
    if (st) return st;
    if ((st = check_camera(cam))) return st;
    if (!L || !og) return fail(VR_ERR_ARG, "null buffer");
    int id = 0, od = 0;
    map_dims(kind, &id, &od);
    if (id != 6) return fail(VR_ERR_UNSUPPORTED, "capture needs a 6-input map");
    int64_t total = camera_threads(cam);
    int64_t start = L->start_index, cnt = L->count < 0 ? total - start : L->count;
    int spp = L->samples < 1 ? 1 : L->samples;
    int64_t ssig = grid_span(&M->sigma);
    int64_t salb = M->albedo_kind == VR_FIELD_GRID ? grid_span(&M->albedo) : 0;
    #pragma omp parallel
    {
        gacc_t acc;
        acc.sig = M->sigma.t.grad ? (double*)calloc((size_t)ssig, sizeof(double)) : NULL;
        acc.alb = (salb && M->albedo.t.grad) ? (double*)calloc((size_t)salb, sizeof(double)) : NULL;
        #pragma omp for schedule(dynamic, 4)
        for (int64_t i = start; i < start + cnt; i++) {
            int pose, row, col;
            sensor_index(cam, i, &pose, &row, &col);
            if (pose < 0 || pose >= cam->num_poses) continue;  /* outside the pose buffer: no gradient */
            double o[16];
            for (int k = 0; k < od; k++) o[k] = og[i * od + k] / spp;
            ctx_t c;
            memset(&c, 0, sizeof(c));
            c.M = M; c.bw_mode = L->bw_mode; c.acc = acc;
            if (L->seed_mode == VR_SEED_CAPTURE_BW) c.rng = seed_capture_bw(L->seeds, (uint32_t)i);
            for (int s = 0; s < spp; s++) {
                if (L->seed_mode != VR_SEED_CAPTURE_BW)
                    c.rng = seed_capture_fw(L->seeds, (uint32_t)((int64_t)s * total + i));
                v3 x, w;
                camera_ray(cam, pose, row, col, &c, &x, &w);
                double a[6] = {x.x, x.y, x.z, w.x, w.y, w.z};
                field_bw(&c, kind, a, o);
            }
        }
        #pragma omp critical
        {
            reduce_priv(&M->sigma, acc.sig);
            if (acc.alb) reduce_priv(&M->albedo, acc.alb);
        }
        free(acc.sig);
        free(acc.alb);
    }
    return VR_OK;
}

int oracle_adam_step_bounded(float* p, float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
                             float eps, int32_t step, float lo, float hi, int32_t zero_grad) {
    if (step < 1) return fail(VR_ERR_ARG, "step must be >= 1");
    double bc1 = 1.0 - pow((double)b1, step), bc2 = 1.0 - pow((double)b2, step);
    for (int64_t i = 0; i < n; i++) {
        double gi = g[i];
        double mi = b1 * (double)m[i] + (1.0 - b1) * gi;
        double vi = b2 * (double)v[i] + (1.0 - b2) * gi * gi;
        m[i] = (float)mi; v[i] = (float)vi;
        double denom = sqrt(vi / bc2) + eps;
        p[i] = (float)dmin(dmax((double)p[i] - lr / bc1 * mi / denom, lo), hi);
        if (zero_grad) g[i] = 0.0f;
    }
    return VR_OK;
}

int oracle_adam_step(float* p, float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
                     float eps, int32_t step, int32_t zero_grad) {
    return oracle_adam_step_bounded(p, g, m, v, n, lr, b1, b2, eps, step, -INFINITY, INFINITY, zero_grad);
}

/* Raw RNG access for known-answer tests: n uniforms from a state. */
int oracle_rng_uniforms(const uint32_t state[4], int64_t n, float* out, uint32_t* final_state) {
    rng_t s = {state[0], state[1], state[2], state[3]};
    for (int64_t i = 0; i < n; i++) out[i] = rng_random(&s);
    if (final_state) { final_state[0] = s.x; final_state[1] = s.y; final_state[2] = s.z; final_state[3] = s.w; }
    return VR_OK;
}

/* Per-thread initial states for known-answer tests. */
int oracle_rng_seed_states(int32_t mode, const uint32_t seeds[4], int64_t first, int64_t n, uint32_t* out) {
    for (int64_t k = 0; k < n; k++) {
        uint32_t i = (uint32_t)(first + k);
        rng_t s = mode == VR_SEED_MAP ? seed_map(seeds, i) : (mode == VR_SEED_CAPTURE_BW ? seed_capture_bw(seeds, i) : seed_capture_fw(seeds, i));
        out[4 * k] = s.x; out[4 * k + 1] = s.y; out[4 * k + 2] = s.z; out[4 * k + 3] = s.w;
    }
    return VR_OK;
}

/* ------------------------------------------------------------------------- */
/* Grid tools (compute functions of the reference)                            */
/* ------------------------------------------------------------------------- */

/* functions/tools/random_ids.comp.glsl:11-20; seeding of
 * functions/common_functions.h:27-49 (seeds + (0x23F1, 0x3137, 129, i + 129), four
 * discarded steps, then the bits of four uniforms as the new state). */
int oracle_random_ids(int64_t* out, int64_t n, const int64_t* shape, int32_t dim, const uint32_t seeds[4]) {
    if (n < 0 || dim < 1 || dim > 4 || !shape || !seeds || (n > 0 && !out)) return fail(VR_ERR_ARG, "bad random_ids arguments");
    for (int64_t i = 0; i < n; i++) {
        rng_t s;
        s.x = seeds[0] + 0x23F1u;
        s.y = seeds[1] + 0x3137u;
        s.z = seeds[2] + 129u;
        s.w = seeds[3] + ((uint32_t)i + 129u);
        for (int k = 0; k < 4; k++) rng_random(&s);
        float a = rng_random(&s), b = rng_random(&s), c = rng_random(&s), d = rng_random(&s);
        s.x = float_to_bits(a); s.y = float_to_bits(b); s.z = float_to_bits(c); s.w = float_to_bits(d);
        for (int k = 0; k < dim; k++) {
            float u = rng_random(&s);
            float prod = u * (float)(int)shape[k];
            out[i * dim + k] = (int64_t)(int)prod;
        }
    }
    return VR_OK;
}

/* total_variation*.comp.glsl: adjacent_index (lines 13-31) walks the
 * coordinates innermost first; shape[0..dim-1] outermost first as stored. */
static int tv_adjacent(int64_t index, const int64_t* shape, int dim, int axis, int delta, int64_t* adj) {
    int64_t size = 1, a = 0;
    for (int i = 0; i < dim; i++) {
        int64_t id = shape[dim - i - 1];
        int64_t pc = index % id;
        if (i == axis) pc += delta;
        if (pc < 0 || pc >= id) return 0;
        a += pc * size;
        index /= id;
        size *= id;
    }
    *adj = a;
    return 1;
}

static int tv_check(const int64_t* shape, int32_t dim, int64_t* voxels) {
    if (!shape || dim < 1 || dim > 3) return fail(VR_ERR_ARG, "total variation: 1 to 3 spatial dimensions");
    int64_t v = 1;
    for (int i = 0; i <= dim; i++) {
        if (shape[i] < 1) return fail(VR_ERR_ARG, "total variation: bad extent");
        if (i < dim) v *= shape[i];
    }
    *voxels = v;
    return VR_OK;
}

/* total_variation.comp.glsl:33-74 (float arithmetic, the shader's order). */
int oracle_total_variation_fw(const float* in, const int64_t* shape, int32_t dim, float* out) {
    int64_t V;
    int st = tv_check(shape, dim, &V);
    if (st) return st;
    const int64_t C = shape[dim];
    for (int64_t i = 0; i < V; i++) {
        float tv = 0.0f;
        for (int d = 0; d < dim; d++) {
            for (int k = 0; k < 2; k++) {
                int64_t j;
                if (!tv_adjacent(i, shape, dim, d, k == 0 ? 1 : -1, &j)) continue;
                float t = 0.0f;
                for (int64_t c = 0; c < C; c++) t += fabsf(in[i * C + c] - in[j * C + c]);
                tv += t;
            }
        }
        out[i] = tv;
    }
    return VR_OK;
}

static float signf_glsl(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

/* total_variation_backward.comp.glsl:33-76. */
int oracle_total_variation_bw(const float* in, const float* out_grad, const int64_t* shape, int32_t dim, float* in_grad) {
    int64_t V;
    int st = tv_check(shape, dim, &V);
    if (st) return st;
    const int64_t C = shape[dim];
    for (int64_t i = 0; i < V; i++) {
        for (int64_t c = 0; c < C; c++) in_grad[i * C + c] = 0.0f;
        for (int d = 0; d < dim; d++) {
            for (int k = 0; k < 2; k++) {
                int64_t j;
                if (!tv_adjacent(i, shape, dim, d, k == 0 ? 1 : -1, &j)) continue;
                for (int64_t c = 0; c < C; c++)
                    in_grad[i * C + c] += (out_grad[i] + out_grad[j]) * signf_glsl(in[i * C + c] - in[j * C + c]);
            }
        }
    }
    return VR_OK;
}

/* functions/tools/resample_grid.comp.glsl:14-86 (vertex-aligned, cell == 0)
 * and resample_img.comp.glsl:14-86 (pixel-centred, cell == 1): float
 * arithmetic, taps and weights in the shaders' order. */
static int resample_any(const float* in, const int64_t* in_shape, const int64_t* out_shape, int32_t dim, int32_t channels,
                        float* out, int cell) {
    if (!in_shape || !out_shape || dim < 1 || dim > 3 || channels < 1) return fail(VR_ERR_ARG, "bad resample arguments");
    int64_t V = 1;
    for (int i = 0; i < dim; i++) {
        if (in_shape[i] < 1 || out_shape[i] < (cell ? 1 : 2)) return fail(VR_ERR_ARG, "resample: output extents too small");
        if ((out_shape[i] - 1) * (in_shape[i] - 1) >= ((int64_t)1 << 31)) return fail(VR_ERR_ARG, "resample: sizes too large");
        V *= out_shape[i];
    }
    for (int64_t index = 0; index < V; index++) {
        int ic[4] = {0, 0, 0, 0};
        float al[4] = {0, 0, 0, 0};
        int64_t r = index;
        for (int i = 0; i < dim; i++) {
            int id = (int)in_shape[dim - i - 1], od = (int)out_shape[dim - i - 1];
            int pc = (int)(r % od);
            float coord = cell ? ((float)pc + 0.5f) / (float)od * (float)id - 0.5f : (float)(pc * (id - 1)) / (float)(od - 1);
            ic[i] = (int)floorf(coord);
            al[i] = coord - (float)ic[i];
            r /= od;
        }
        for (int c = 0; c < channels; c++) out[index * channels + c] = 0.0f;
        for (int t = 0; t < (1 << dim); t++) {
            float w = 0.0f;
            int64_t src = 0, u = 1;
            for (int i = 0; i < dim; i++) {
                int dx = (t >> i) & 1;
                float f = dx ? al[i] : 1.0f - al[i];
                w = i == 0 ? f : w * f;
                int d = (int)in_shape[dim - i - 1];
                int cc = ic[i] + dx < d - 1 ? ic[i] + dx : d - 1;
                if (cc < 0) cc = 0;
                src += (int64_t)cc * u;
                u *= d;
            }
            for (int c = 0; c < channels; c++) out[index * channels + c] += w * in[src * channels + c];
        }
    }
    return VR_OK;
}

int oracle_resample_grid(const float* in, const int64_t* in_shape, const int64_t* out_shape, int32_t dim, int32_t channels,
                         float* out) {
    return resample_any(in, in_shape, out_shape, dim, channels, out, 0);
}

int oracle_resample_img(const float* in, const int64_t* in_shape, const int64_t* out_shape, int32_t dim, int32_t channels,
                        float* out) {
    return resample_any(in, in_shape, out_shape, dim, channels, out, 1);
}

/* maths.h:244-251 unpart1by1: the even bits of n, compacted. */
static int unpart1by1(int n) {
    n &= 0x55555555;
    n = (n ^ (n >> 1)) & 0x33333333;
    n = (n ^ (n >> 2)) & 0x0f0f0f0f;
    n = (n ^ (n >> 4)) & 0x00ff00ff;
    n = (n ^ (n >> 8)) & 0x0000ffff;
    return n;
}

static int morton_check(int64_t resolution, int32_t channels) {
    if (resolution < 1 || resolution > (1 << 15) || (resolution & (resolution - 1)) || channels < 1)
        return fail(VR_ERR_ARG, "morton copy: resolution must be a power of two up to 2^15");
    return VR_OK;
}

/* functions/tools/copy_img_to_morton.comp.glsl: out[i] = img[py][px], (px, py) = (even, odd) bits of i. */
int oracle_copy_img_to_morton(const float* in, int64_t resolution, int32_t channels, float* out) {
    int st = morton_check(resolution, channels);
    if (st) return st;
    for (int64_t i = 0; i < resolution * resolution; i++) {
        int px = unpart1by1((int)i), py = unpart1by1((int)(i >> 1));
        for (int c = 0; c < channels; c++) out[i * channels + c] = in[((int64_t)py * resolution + px) * channels + c];
    }
    return VR_OK;
}

/* functions/tools/copy_morton_to_img.comp.glsl: the inverse permutation. */
int oracle_copy_morton_to_img(const float* in, int64_t resolution, int32_t channels, float* out) {
    int st = morton_check(resolution, channels);
    if (st) return st;
    for (int64_t i = 0; i < resolution * resolution; i++) {
        int px = unpart1by1((int)i), py = unpart1by1((int)(i >> 1));
        for (int c = 0; c < channels; c++) out[((int64_t)py * resolution + px) * channels + c] = in[i * channels + c];
    }
    return VR_OK;
}

/* Oracle-only probe of the environment sampler (no device twin): sample i
 * draws from the map-eval stream of thread i and returns (wo, L(wo) / pdf). */
int oracle_env_sample(const vr_medium* M, const uint32_t seeds[4], int64_t n, float* out) {
    int st = check_medium(M, VR_MAP_RAYBOX);
    if (st) return st;
    if (!seeds || !out || n < 0) return fail(VR_ERR_ARG, "bad env_sample arguments");
    for (int64_t i = 0; i < n; i++) {
        ctx_t c;
        memset(&c, 0, sizeof(c));
        c.M = M;
        c.rng = seed_map(seeds, (uint32_t)i);
        v3 wo;
        v3 L = env_sample(&c, &wo);
        out[6 * i + 0] = (float)wo.x; out[6 * i + 1] = (float)wo.y; out[6 * i + 2] = (float)wo.z;
        out[6 * i + 3] = (float)L.x; out[6 * i + 4] = (float)L.y; out[6 * i + 5] = (float)L.z;
    }
    return VR_OK;
}

/* Oracle-only probe of the phase sampler and the phase function (no device
 * twin): sample i draws from the map-eval stream of thread i, samples wo for the
 * propagation direction w[i] and returns (wo, rho, phase_eval(w, wo)). */
int oracle_phase_sample(const vr_medium* M, const float* w, const uint32_t seeds[4], int64_t n, float* out) {
    int st = check_medium(M, VR_MAP_RAYBOX);
    if (st) return st;
    if (!w || !seeds || !out || n < 0) return fail(VR_ERR_ARG, "bad phase_sample arguments");
    for (int64_t i = 0; i < n; i++) {
        ctx_t c;
        memset(&c, 0, sizeof(c));
        c.M = M;
        c.rng = seed_map(seeds, (uint32_t)i);
        v3 wi = vf(w + 3 * i), wo;
        double rho = phase_sample(&c, V(0, 0, 0), wi, &wo);
        out[5 * i + 0] = (float)wo.x; out[5 * i + 1] = (float)wo.y; out[5 * i + 2] = (float)wo.z;
        out[5 * i + 3] = (float)rho;
        out[5 * i + 4] = (float)phase_eval(M, V(0, 0, 0), wi, wo);
    }
    return VR_OK;
}

/* Oracle-only probe: phase(x, w, wo) for n direction pairs, in double. */
int oracle_phase_eval(const vr_medium* M, const double* w, const double* wo, int64_t n, double* out) {
    int st = check_medium(M, VR_MAP_RAYBOX);
    if (st) return st;
    if (!w || !wo || !out || n < 0) return fail(VR_ERR_ARG, "bad phase_eval arguments");
    for (int64_t i = 0; i < n; i++)
        out[i] = phase_eval(M, V(0, 0, 0), V(w[3 * i], w[3 * i + 1], w[3 * i + 2]), V(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]));
    return VR_OK;
}

/* Oracle-only probe of the camera: (x, w) of sample 0 of every sensor, drawn
 * (when jittered) from that sample's capture-forward stream, as the capture
 * drivers do. */
int oracle_camera_rays(const vr_camera* cam, const uint32_t seeds[4], double* out) {
    int st = check_camera(cam);
    if (st) return st;
    if (!seeds || !out) return fail(VR_ERR_ARG, "bad camera_rays arguments");
    const int64_t n = camera_threads(cam);
    for (int64_t i = 0; i < n; i++) {
        ctx_t c;
        memset(&c, 0, sizeof(c));
        c.rng = seed_capture_fw(seeds, (uint32_t)i);
        int pose, row, col;
        sensor_index(cam, i, &pose, &row, &col);
        v3 x, w;
        camera_ray(cam, pose, row, col, &c, &x, &w);
        out[6 * i + 0] = x.x; out[6 * i + 1] = x.y; out[6 * i + 2] = x.z;
        out[6 * i + 3] = w.x; out[6 * i + 4] = w.y; out[6 * i + 5] = w.z;
    }
    return VR_OK;
}

/* ------------------------------------------------------------------------- */
/* Scene-level path tracing: PathtracedScene (_scenes.py:2251-2312),          */
/* include/sc_path_tracing.h:12-380, maps/pathtraced_scene.h.  The closest    */
/* hit is found by brute force over every triangle (no BVH): it pins the      */
/* library's BVH build and traversal.                                         */
/* ------------------------------------------------------------------------- */
#define SCENE_TMAX 10000.0     /* maps/geometry_mesh.h:15 */
#define SCENE_PATH_CAP 4096    /* loop iterations per path (the reference has no cap) */

typedef struct { v3 P, N, G; double cu, cv; } surfel_t;

/* Moller-Trumbore in double; (u, v) weight vertices 1 and 2. */
static int o_tri_hit(v3 v0, v3 v1, v3 v2, v3 o, v3 d, double tmax, double* t, double* bu, double* bv) {
    v3 e1 = vsub(v1, v0), e2 = vsub(v2, v0);
    v3 p = vcross(d, e2);
    double det = vdot(e1, p);
    if (fabs(det) < 1e-30) return 0;
    double id = 1.0 / det;
    v3 s = vsub(o, v0);
    double u = vdot(s, p) * id;
    if (u < 0 || u > 1) return 0;
    v3 q = vcross(s, e1);
    double w = vdot(d, q) * id;
    if (w < 0 || u + w > 1) return 0;
    double tt = vdot(e2, q) * id;
    if (tt < 0 || tt >= tmax) return 0;
    *t = tt; *bu = u; *bv = w;
    return 1;
}

static v3 o_vertex(const vr_scene* S, int tri, int c) {
    return vf(S->positions + 3 * (int64_t)S->indices[3 * (int64_t)tri + c]);
}

/* Closest triangle (original index) with 0 <= t < 10000, or -1. */
static int o_scene_trace(const vr_scene* S, v3 o, v3 d, double* t, double* bu, double* bv) {
    int best = -1;
    double bt = SCENE_TMAX;
    for (int k = 0; k < S->num_tris; k++) {
        double tt, u, v;
        if (o_tri_hit(o_vertex(S, k, 0), o_vertex(S, k, 1), o_vertex(S, k, 2), o, d, bt, &tt, &u, &v)) {
            bt = tt; best = k; *bu = u; *bv = v;
        }
    }
    *t = bt;
    return best;
}

/* sample_surfel (include/common.h:467-501) + hit2surfel (:297-310). */
static surfel_t o_scene_surfel(const vr_scene* S, int tri, double bu, double bv, v3 x, v3 w, double t, int* patch) {
    surfel_t s;
    int i0 = S->indices[3 * tri], i1 = S->indices[3 * tri + 1], i2 = S->indices[3 * tri + 2];
    double a0 = 1 - bu - bv;
    v3 p0 = vf(S->positions + 3 * i0), p1 = vf(S->positions + 3 * i1), p2 = vf(S->positions + 3 * i2);
    s.P = vadd(vs(w, t), x);
    s.G = vnorm(vcross(vsub(p1, p0), vsub(p2, p0)));
    if (S->normals)
        s.N = vnorm(vadd(vadd(vs(vf(S->normals + 3 * i0), a0), vs(vf(S->normals + 3 * i1), bu)), vs(vf(S->normals + 3 * i2), bv)));
    else s.N = s.G;
    if (S->uvs) {
        s.cu = S->uvs[2 * i0] * a0 + S->uvs[2 * i1] * bu + S->uvs[2 * i2] * bv;
        s.cv = S->uvs[2 * i0 + 1] * a0 + S->uvs[2 * i1 + 1] * bu + S->uvs[2 * i2 + 1] * bv;
    } else s.cu = s.cv = 0;
    *patch = S->tri_patch[tri];
    return s;
}

/* INCLUDE_ADDITIONAL_COLLISION_VARS (include/common.h:509-512). */
static void o_collision_vars(v3 win, const surfel_t* s, int* from_outside, v3* fN) {
    *from_outside = vdot(win, s->G) < 0;
    int correct = (vdot(win, s->N) < 0) == *from_outside;
    v3 N = correct ? s->N : s->G;
    *fN = *from_outside ? N : vs(N, -1.0);
}

static v3 o_reflect(v3 i, v3 n) { return vsub(i, vs(n, 2 * vdot(n, i))); }
static v3 o_refract(v3 i, v3 n, double eta) {
    double c = vdot(n, i);
    double k = 1 - eta * eta * (1 - c * c);
    if (k < 0) return V(0, 0, 0);
    return vsub(vs(i, eta), vs(n, eta * c + sqrt(k)));
}

/* randomHSDirectionCosineWeighted (randoms.h:215-231), capped at 64 draws. */
static v3 o_cosine_hemisphere(ctx_t* c, v3 N, double* ndotd) {
    v3 t0, t1;
    ortho_basis(N, &t0, &t1);
    for (int it = 0; it < 64; it++) {
        /* the GPU forms x, y and d2 in float; so does the oracle, so that the
         * acceptance test takes the same branch */
        float x = (float)rnd(c) * 2.0f - 1.0f;
        float y = (float)rnd(c) * 2.0f - 1.0f;
        float d2 = x * x + y * y;
        if (d2 > 0.001f && d2 < 1.0f) {
            double z = sqrt(1.0 - (double)d2);
            *ndotd = z;
            return vadd(vadd(vs(t0, x), vs(t1, y)), vs(N, z));
        }
    }
    *ndotd = 1;
    return N;
}

/* surface_scattering_sampler of a patch (_scenes.py:906-1340). */
static void o_surface_sample(const vr_patch* P, ctx_t* c, v3 win, const surfel_t* s, v3* wout, v3* W, double* pdf) {
    int kind = P->sampler;
    if (kind == VR_SURF_PASS || kind == VR_SURF_NULL) { *wout = win; *W = V(1, 1, 1); *pdf = -1; return; }
    if (kind == VR_SURF_BLOCK) { *wout = V(0, 0, 0); *W = V(0, 0, 0); *pdf = 0; return; }
    int from_outside;
    v3 fN;
    o_collision_vars(win, s, &from_outside, &fN);
    if (kind == VR_SURF_LAMBERT) {
        if (!from_outside) { *wout = V(0, 0, 0); *pdf = 0; *W = V(0, 0, 0); }
        else {
            double ct;
            *wout = o_cosine_hemisphere(c, fN, &ct);
            *pdf = ct / PI_D;
            *W = V(1, 1, 1);
        }
        return;
    }
    if (kind == VR_SURF_MIRROR) { *wout = o_reflect(win, fN); *W = V(1, 1, 1); *pdf = -1; return; }
    /* Fresnel: the uniform is drawn only when the refraction exists (|| short-circuits) */
    v3 rw = o_reflect(win, fN);
    double eta = from_outside ? 1.0 / P->refraction_index : P->refraction_index;
    v3 tw = o_refract(win, fN, eta);
    double th = fabs(vdot(win, fN));
    double f = pow((1 - eta) / (1 + eta), 2);
    double R = f + (1 - f) * pow(1 - th, 5);
    int no_t = tw.x == 0 && tw.y == 0 && tw.z == 0;
    if (no_t || rnd(c) < R) *wout = rw;
    else *wout = tw;
    *W = V(1 / (eta * eta), 1 / (eta * eta), 1 / (eta * eta));
    *pdf = -1;
}

/* The material's emission as the patch's surface gathering (_scenes.py:1392-1435, 2407-2423). */
static v3 o_surface_emission(const vr_patch* P, v3 win, const surfel_t* s) {
    if (P->emission == VR_EMIT_NONE) return V(0, 0, 0);
    v3 E = vf(P->emission_c);
    if (P->emission == VR_EMIT_ISOTROPIC) return vs(E, 0.5 / PI_D);
    int from_outside;
    v3 fN;
    o_collision_vars(win, s, &from_outside, &fN);
    return vs(E, vdot(vs(win, -1), fN) / PI_D);
}

static v3 o_scene_env(const vr_scene* S, v3 w) {
    vr_medium m;
    memset(&m, 0, sizeof(m));
    m.env_kind = S->env_kind;
    memcpy(m.env_c, S->env_c, sizeof(m.env_c));
    m.env_image = S->env_image;
    memcpy(m.env_image_bmin, S->env_image_bmin, sizeof(m.env_image_bmin));
    memcpy(m.env_image_bmax, S->env_image_bmax, sizeof(m.env_image_bmax));
    return env_eval(&m, w);
}

typedef struct { gacc_t acc[VR_SCENE_MAX_MEDIA]; } sacc_t;

static int o_inside(const vr_scene* S, const vr_patch* P) {
    return (P->inside_medium >= 0 && P->inside_medium < S->num_media) ? P->inside_medium : -1;
}

static void o_use_medium(ctx_t* c, const vr_scene* S, const sacc_t* a, int m) {
    c->M = &S->media[m];
    if (a) c->acc = a->acc[m];
}

typedef struct { int current, level, stacked[4]; } mstack_t;

/* sc_path_tracing.h:84-101 */
static void o_mstack_cross(mstack_t* m, int to_inside, int inside) {
    if (to_inside) {
        if (m->level >= 0) m->stacked[m->level] = m->current;
        if (m->level < 3) m->level++;
        m->current = inside;
    } else {
        m->level--;
        m->current = m->level == -1 ? -1 : m->stacked[m->level < 0 ? 0 : m->level];
    }
}

static int o_below(v3 W) { return W.x < 1e-7 && W.y < 1e-7 && W.z < 1e-7; }

/* path_trace (sc_path_tracing.h:12-120); branch = path_trace_fw's 4 uniforms of
 * branch set-up around the (zero) direct gathering (:124-128). */
static v3 o_scene_path(const vr_scene* S, ctx_t* c, v3 x, v3 w, int branch) {
    if (branch) { rnd(c); rnd(c); rnd(c); rnd(c); }
    v3 A = V(0, 0, 0), W = V(1, 1, 1);
    mstack_t ms = {-1, -1, {-1, -1, -1, -1}};
    int bounces = 0;
    for (int it = 0; it < SCENE_PATH_CAP; it++) {
        if (o_below(W)) return A;
        double d, bu, bv;
        int tri = o_scene_trace(S, x, w, &d, &bu, &bv);
        if (tri < 0) return vadd(A, vmul(W, o_scene_env(S, w)));
        int patch;
        surfel_t s = o_scene_surfel(S, tri, bu, bv, x, w, d, &patch);
        const vr_patch* P = &S->patches[patch];
        int from_inside = vdot(w, s.G) > 0;
        if (ms.level == -1 && from_inside) { ms.level = 0; ms.current = o_inside(S, P); }
        if (ms.current >= 0) {
            double o[12];
            o_use_medium(c, S, NULL, ms.current);
            medium_dt_fw(c, x, w, d, o);
            v3 xo = V(o[0], o[1], o[2]), wo = V(o[3], o[4], o[5]);
            A = vadd(A, vmul(W, V(o[9], o[10], o[11])));
            W = vmul(W, V(o[6], o[7], o[8]));
            if (!(wo.x == w.x && wo.y == w.y && wo.z == w.z)) { x = xo; w = wo; continue; }
        }
        bounces++;
        A = vadd(A, vmul(W, o_surface_emission(P, w, &s)));
        v3 ws, Ws;
        double pdf;
        o_surface_sample(P, c, w, &s, &ws, &Ws, &pdf);
        int to_inside = vdot(ws, s.G) < 0;
        if (to_inside != from_inside) o_mstack_cross(&ms, to_inside, o_inside(S, P));
        W = vmul(W, Ws);
        w = ws;
        if (pdf > 0 && bounces > 2) {
            if (rnd(c) < 0.5) return A;
            W = vs(W, 2);
        }
        x = vadd(s.P, vs(w, 0.0001));
    }
    return A;
}

/* path_trace_bw (sc_path_tracing.h:238-380). */
static void o_scene_path_bw(const vr_scene* S, ctx_t* c, const sacc_t* acc, v3 x, v3 w, v3 out_A, v3 dL_dA) {
    rnd(c); rnd(c); rnd(c); rnd(c);
    v3 rem = out_A, W = V(1, 1, 1);
    mstack_t ms = {-1, -1, {-1, -1, -1, -1}};
    int bounces = 0;
    for (int it = 0; it < SCENE_PATH_CAP; it++) {
        if (o_below(W)) return;
        double d, bu, bv;
        int tri = o_scene_trace(S, x, w, &d, &bu, &bv);
        if (tri < 0) return;
        int patch;
        surfel_t s = o_scene_surfel(S, tri, bu, bv, x, w, d, &patch);
        const vr_patch* P = &S->patches[patch];
        int from_inside = vdot(w, s.G) > 0;
        if (ms.level == -1 && from_inside) { ms.level = 0; ms.current = o_inside(S, P); }
        if (ms.current >= 0) {
            double o[12];
            o_use_medium(c, S, acc, ms.current);
            rng_t before = c->rng;
            medium_dt_fw(c, x, w, d, o);
            v3 xo = V(o[0], o[1], o[2]), wo = V(o[3], o[4], o[5]), vW = V(o[6], o[7], o[8]), vA = V(o[9], o[10], o[11]);
            rem = vsub(rem, vmul(W, vA));
            v3 dL_dvA = vmul(dL_dA, W);
            v3 g = vmul(dL_dA, rem);
            double gv[3] = {g.x / vW.x, g.y / vW.y, g.z / vW.z};
            for (int k = 0; k < 3; k++) if (!isfinite(gv[k])) gv[k] = 0;  /* vW == 0: dropped */
            c->rng = before;
            medium_dt_replay(c, x, w, d, vW, vA, V(gv[0], gv[1], gv[2]), dL_dvA);
            W = vmul(W, vW);
            if (!(wo.x == w.x && wo.y == w.y && wo.z == w.z)) { x = xo; w = wo; continue; }
        }
        bounces++;
        rem = vsub(rem, vmul(W, o_surface_emission(P, w, &s)));
        v3 ws, Ws;
        double pdf;
        o_surface_sample(P, c, w, &s, &ws, &Ws, &pdf);
        int to_inside = vdot(ws, s.G) < 0;
        if (to_inside != from_inside) o_mstack_cross(&ms, to_inside, o_inside(S, P));
        W = vmul(W, Ws);
        w = ws;
        if (pdf > 0 && bounces > 2) {
            if (rnd(c) < 0.5) return;
            W = vs(W, 2);
        }
        x = vadd(s.P, vs(w, 0.0001));
    }
}

/* PathtracedScene BACKWARD (maps/pathtraced_scene.h:13-25). */
static void o_scene_map_bw(const vr_scene* S, ctx_t* c, const sacc_t* acc, v3 x, v3 w, v3 dL_dA) {
    rng_t before = c->rng;
    v3 A = o_scene_path(S, c, x, w, 1);
    c->rng = before;
    o_scene_path_bw(S, c, acc, x, w, A, dL_dA);
}

static int o_check_scene(const vr_scene* S) {
    if (!S) return fail(VR_ERR_ARG, "scene is NULL");
    if (S->num_tris < 0 || S->num_patches < 0) return fail(VR_ERR_ARG, "scene: negative sizes");
    if (S->num_tris > 0 && (!S->positions || !S->indices || !S->tri_patch || !S->patches || S->num_patches < 1))
        return fail(VR_ERR_ARG, "scene: NULL geometry buffer");
    if (S->num_media < 0 || S->num_media > VR_SCENE_MAX_MEDIA) return fail(VR_ERR_UNSUPPORTED, "scene: too many media");
    for (int k = 0; k < S->num_tris; k++)
        if (S->tri_patch[k] < 0 || S->tri_patch[k] >= S->num_patches) return fail(VR_ERR_ARG, "scene: patch index out of range");
    return VR_OK;
}

static void o_sacc_alloc(const vr_scene* S, sacc_t* a) {
    memset(a, 0, sizeof(*a));
    for (int m = 0; m < S->num_media; m++) {
        const vr_medium* M = &S->media[m];
        if (M->sigma.t.grad) a->acc[m].sig = (double*)calloc((size_t)grid_span(&M->sigma), sizeof(double));
        if (M->albedo_kind == VR_FIELD_GRID && M->albedo.t.grad)
            a->acc[m].alb = (double*)calloc((size_t)grid_span(&M->albedo), sizeof(double));
    }
}

static void o_sacc_reduce_free(const vr_scene* S, sacc_t* a) {
    for (int m = 0; m < S->num_media; m++) {
        reduce_priv(&S->media[m].sigma, a->acc[m].sig);
        if (a->acc[m].alb) reduce_priv(&S->media[m].albedo, a->acc[m].alb);
        free(a->acc[m].sig);
        free(a->acc[m].alb);
    }
}

int oracle_scene_build_bvh(const float* positions, int64_t num_vertices, const int32_t* indices, int64_t num_tris,
                           vr_bvh_node* nodes, int64_t max_nodes, int32_t* tri_ids, float* tri_verts, int32_t* num_nodes) {
    /* The oracle needs no hierarchy: one leaf over every triangle in input order
     * (its traversal is the brute-force loop above). */
    if (!num_nodes || num_vertices < 0 || num_tris < 0) return fail(VR_ERR_ARG, "bad sizes");
    *num_nodes = 0;
    if (num_tris == 0) return VR_OK;
    if (!positions || !indices || !nodes || !tri_ids || !tri_verts || max_nodes < 1) return fail(VR_ERR_ARG, "null buffer");
    float lo[3] = {3e38f, 3e38f, 3e38f}, hi[3] = {-3e38f, -3e38f, -3e38f};
    for (int64_t t = 0; t < num_tris; t++) {
        tri_ids[t] = (int32_t)t;
        for (int c = 0; c < 3; c++) {
            int32_t v = indices[3 * t + c];
            if (v < 0 || v >= num_vertices) return fail(VR_ERR_ARG, "vertex index out of range");
            for (int a = 0; a < 3; a++) {
                float p = positions[3 * (int64_t)v + a];
                tri_verts[9 * t + 3 * c + a] = p;
                if (p < lo[a]) lo[a] = p;
                if (p > hi[a]) hi[a] = p;
            }
        }
    }
    for (int a = 0; a < 3; a++) { nodes[0].bmin[a] = lo[a]; nodes[0].bmax[a] = hi[a]; }
    nodes[0].first = 0;
    nodes[0].count = (int32_t)num_tris;
    *num_nodes = 1;
    return VR_OK;
}

int oracle_scene_raycast(const vr_scene* S, const float* rays, int64_t n, float* hits) {
    int st = o_check_scene(S);
    if (st) return st;
    if (n > 0 && (!rays || !hits)) return fail(VR_ERR_ARG, "null buffer");
    #pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; i++) {
        v3 x = vf(rays + 6 * i), w = vf(rays + 6 * i + 3);
        double t, bu, bv;
        int tri = o_scene_trace(S, x, w, &t, &bu, &bv);
        float* h = hits + 16 * i;
        for (int k = 0; k < 15; k++) h[k] = 0;
        if (tri < 0) { int32_t m1 = -1; memcpy(h + 15, &m1, 4); continue; }
        int patch;
        surfel_t s = o_scene_surfel(S, tri, bu, bv, x, w, t, &patch);
        h[0] = (float)t;
        h[1] = (float)s.N.x; h[2] = (float)s.N.y; h[3] = (float)s.N.z;
        h[4] = (float)s.G.x; h[5] = (float)s.G.y; h[6] = (float)s.G.z;
        h[7] = (float)s.cu; h[8] = (float)s.cv;
        int32_t p32 = patch;
        memcpy(h + 15, &p32, 4);
    }
    return VR_OK;
}

int oracle_scene_map_fw(const vr_scene* S, const float* rays, int64_t n, float* out, const vr_launch* L) {
    int st = o_check_scene(S);
    if (st) return st;
    if (!L || (n > 0 && (!rays || !out))) return fail(VR_ERR_ARG, "null buffer");
    int64_t start = L->start_index, cnt = L->count < 0 ? n - start : L->count;
    #pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = start; i < start + cnt; i++) {
        ctx_t c;
        memset(&c, 0, sizeof(c));
        c.bw_mode = L->bw_mode;
        c.rng = seed_map(L->seeds, (uint32_t)i);
        v3 A = o_scene_path(S, &c, vf(rays + 6 * i), vf(rays + 6 * i + 3), 0);
        out[3 * i] = (float)A.x; out[3 * i + 1] = (float)A.y; out[3 * i + 2] = (float)A.z;
        if (L->debug_out) L->debug_out[i] = (int32_t)c.lookups;
    }
    return VR_OK;
}

int oracle_scene_map_bw(const vr_scene* S, const float* rays, const float* og, int64_t n, const vr_launch* L) {
    int st = o_check_scene(S);
    if (st) return st;
    if (!L || (n > 0 && (!rays || !og))) return fail(VR_ERR_ARG, "null buffer");
    int64_t start = L->start_index, cnt = L->count < 0 ? n - start : L->count;
    #pragma omp parallel
    {
        sacc_t acc;
        o_sacc_alloc(S, &acc);
        #pragma omp for schedule(dynamic, 16)
        for (int64_t i = start; i < start + cnt; i++) {
            v3 g = vf(og + 3 * i);
            if (g.x == 0 && g.y == 0 && g.z == 0) continue;
            ctx_t c;
            memset(&c, 0, sizeof(c));
            c.bw_mode = L->bw_mode;
            c.rng = seed_map(L->seeds, (uint32_t)i);
            o_scene_map_bw(S, &c, &acc, vf(rays + 6 * i), vf(rays + 6 * i + 3), g);
        }
        #pragma omp critical
        o_sacc_reduce_free(S, &acc);
    }
    return VR_OK;
}

int oracle_scene_capture_fw(const vr_scene* S, const vr_camera* cam, const vr_launch* L, float* out) {
    int st = o_check_scene(S);
    if (st) return st;
    if ((st = check_camera(cam))) return st;
    if (!L || !out) return fail(VR_ERR_ARG, "null buffer");
    int64_t total = camera_threads(cam);
    int64_t start = L->start_index, cnt = L->count < 0 ? total - start : L->count;
    int spp = L->samples < 1 ? 1 : L->samples;
    #pragma omp parallel for schedule(dynamic, 4)
    for (int64_t i = start; i < start + cnt; i++) {
        int pose, row, col;
        sensor_index(cam, i, &pose, &row, &col);
        if (pose < 0 || pose >= cam->num_poses) {
            out[3 * i] = out[3 * i + 1] = out[3 * i + 2] = 0;
            if (L->debug_out) L->debug_out[i] = 0;
            continue;
        }
        v3 acc = V(0, 0, 0);
        int64_t lk = 0;
        for (int s = 0; s < spp; s++) {
            ctx_t c;
            memset(&c, 0, sizeof(c));
            c.bw_mode = L->bw_mode;
            c.rng = seed_capture_fw(L->seeds, (uint32_t)((int64_t)s * total + i));
            v3 x, w;
            camera_ray(cam, pose, row, col, &c, &x, &w);
            acc = vadd(acc, o_scene_path(S, &c, x, w, 0));
            lk += c.lookups;
        }
        out[3 * i] = (float)(acc.x / spp); out[3 * i + 1] = (float)(acc.y / spp); out[3 * i + 2] = (float)(acc.z / spp);
        if (L->debug_out) L->debug_out[i] = (int32_t)lk;
    }
    return VR_OK;
}

int oracle_scene_capture_bw(const vr_scene* S, const vr_camera* cam, const vr_launch* L, const float* og) {
    int st = o_check_scene(S);
    if (st) return st;
    if ((st = check_camera(cam))) return st;
    if (!L || !og) return fail(VR_ERR_ARG, "null buffer");
    int64_t total = camera_threads(cam);
    int64_t start = L->start_index, cnt = L->count < 0 ? total - start : L->count;
    int spp = L->samples < 1 ? 1 : L->samples;
    #pragma omp parallel
    {
        sacc_t acc;
        o_sacc_alloc(S, &acc);
        #pragma omp for schedule(dynamic, 4)
        for (int64_t i = start; i < start + cnt; i++) {
            int pose, row, col;
            sensor_index(cam, i, &pose, &row, &col);
            if (pose < 0 || pose >= cam->num_poses) continue;
            v3 g = vs(vf(og + 3 * i), 1.0 / spp);
            if (g.x == 0 && g.y == 0 && g.z == 0) continue;
            ctx_t c;
            memset(&c, 0, sizeof(c));
            c.bw_mode = L->bw_mode;
            if (L->seed_mode == VR_SEED_CAPTURE_BW) c.rng = seed_capture_bw(L->seeds, (uint32_t)i);
            for (int s = 0; s < spp; s++) {
                if (L->seed_mode != VR_SEED_CAPTURE_BW)
                    c.rng = seed_capture_fw(L->seeds, (uint32_t)((int64_t)s * total + i));
                v3 x, w;
                camera_ray(cam, pose, row, col, &c, &x, &w);
                o_scene_map_bw(S, &c, &acc, x, w, g);
            }
        }
        #pragma omp critical
        o_sacc_reduce_free(S, &acc);
    }
    return VR_OK;
}

/* Cell table of a single-channel grid (vr_grid_pack_cells): cell c = (pz*(ny-1)
 * + py)*(nx-1) + px holds its 8 corners, x fastest, then y, then z. */
int oracle_grid_pack_cells(const vr_grid* g, float* out) {
    if (!g || !out || g->t.ndim != 4 || g->t.shape[3] != 1) return fail(VR_ERR_ARG, "bad grid arguments");
    const vr_tensor* t = &g->t;
    int64_t nx = t->shape[2], ny = t->shape[1], nz = t->shape[0];
    if (nx < 2 || ny < 2 || nz < 2) return fail(VR_ERR_ARG, "grid needs 2 vertices per axis");
    int64_t c = 0;
    for (int64_t z = 0; z < nz - 1; z++)
        for (int64_t y = 0; y < ny - 1; y++)
            for (int64_t x = 0; x < nx - 1; x++, c++)
                for (int k = 0; k < 8; k++)
                    out[8 * c + k] = t->data[goff(t, (int)(z + ((k >> 2) & 1)), (int)(y + ((k >> 1) & 1)), (int)(x + (k & 1)))];
    return VR_OK;
}
