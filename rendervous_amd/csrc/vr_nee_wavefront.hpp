// vr_nee_wavefront.hpp -- NEE capture forward as a persistent state machine.
//
// The reference integrator (include/vr_path_integrator_NEE.h:1-66) nests a
// transmittance loop (the primary T(x,w)*env term and one shadow ray per
// collision) inside the free-flight loop, and the reference capture kernel
// gives each thread one pixel and all its samples (_maps.py:596-605).  On a
// 64-wide wave both serialise: lanes in free flight wait for lanes running
// shadow rays, and a wave lives as long as its most expensive pixel.
//
// Here every lane runs a small state machine (NeeLane) whose step performs
// exactly one tracking step of whichever ray it is on -- the primary
// transmittance ray, the path ray or a shadow ray -- with one majorant query,
// one free-flight draw and one trilinear gather shared by all lanes.  The
// events between steps (new sample, real collision, end of a transmittance
// ray) only prepare the next ray; the box intersection and majorant-walk set-up
// of every new ray run in one shared block, so a wave pays for that code once
// per iteration however many handlers its lanes went through.  Lanes pull work
// items (a pixel and a run of consecutive samples) from a global counter with
// one atomic per wave, so every lane stays busy until the queue drains.  Each
// sample still draws the stream seeded from its flat index s * pixels + p, in
// the same order as the nested form (Tracer::nee), so results are identical
// sample by sample.
#pragma once

#include "vr_device.hpp"

namespace vr {

enum : int { kModePath = 0, kModeTrack = 1 };

// Lane events: what a lane needs before it can take its next tracking step.
enum : int { kEvNone = 0, kEvBegin = 1, kEvStart = 2, kEvCollision = 3, kEvTrackEnd = 4, kEvPathEnd = 5 };

// ENVMAP: environment image and/or quadtree sampler; false compiles the
// constant-environment (uniform + directional light) sampler only.
// LEAN: the medium has a sigma cell table, a constant albedo and a constant
// phase asymmetry, so the kernel compiles neither the strided grid read nor the
// albedo and phase_g grids (fewer uniform values held in scalar registers).
template <int MAJ, int TKIND, bool ENVMAP, bool LEAN = false>
struct NeeLane {
    static constexpr float EPS = 0.00001f;  // vr_path_integrator_NEE.h:16

    Rng rng;
    Walk<MAJ> walk;
    f3 R, W, contrib;
    f3 rx, rw;        // current ray: point and direction
    float rd;         // remaining distance on the current ray
    f3 px, pw;        // path vertex and incoming direction while a transmittance ray runs
    float Tt;         // ratio-tracking estimate of the tracked ray
    float pd;         // path distance inside the medium while the primary term runs
    float track_T;    // transmittance of a finished tracked ray (kEvTrackEnd)
    int mode, bnd, primary, bounces, lookups, event;

    // Camera ray of a new sample (kEvBegin): the path starts at the box entry
    // (vr_path_integrator_NEE.h:3-6) and its first ray is the primary
    // transmittance term T(x, w) * env(w) (:13-14).
    VR_DEV void begin_sample(const MediumView& M, const CameraView& cam, const Seeds& seeds, uint32_t seed_index, int pose,
                             int row, int col) {
        rng = seed_capture_fw(seeds, seed_index);
        f3 x, w;
        camera_ray(cam, pose, row, col, rng, x, w);
        begin_path(M, x, w);
    }

    // A path along the ray (x, w), the generator already seeded.
    VR_DEV void begin_path(const MediumView& M, f3 x, f3 w) {
        float tmin, tmax;
        boundary(M, x, w, tmin, tmax);
        px = x + w * tmin;
        pw = w;
        pd = tmax - tmin;
        R = F3(0.0f, 0.0f, 0.0f);
        W = F3(1.0f, 1.0f, 1.0f);
        bounces = 0;
        contrib = ENVMAP ? env_eval(M, w) : M.env_c;
        primary = 1;
        rx = px;
        rw = pw;
        mode = kModeTrack;
        bnd = 1;
        event = kEvStart;
    }

    // Set-up of the ray (rx, rw) prepared by a handler (kEvStart).  A
    // transmittance ray is clipped to the medium box (vr_transmittance_DT.h:3-8,
    // vr_transmittance_RT.h:95-100) and one that misses it ends at once with
    // T = 1; a path ray keeps its origin and runs to the box exit (or over the
    // distance already in rd when bnd == 0).
    VR_DEV void start_ray(const MediumView& M) {
        event = kEvNone;
        // set for every ray (only read after kEvTrackEnd, which each ray end
        // sets with its own value): a store in the miss branch alone would be
        // merged with the walk's stores into one through a pointer select,
        // which keeps both fields in scratch memory
        track_T = 1.0f;
        if (bnd) {
            float tmin, tmax;
            bool hit = boundary(M, rx, rw, tmin, tmax);
            if (mode == kModeTrack) {
                if (!hit) {
                    event = kEvTrackEnd;
                    return;
                }
                rx = rx + rw * tmin;
                rd = tmax - tmin;
            } else {
                rd = tmax;
            }
        }
        Tt = 1.0f;
        walk.begin(M, rx, rw);
    }

    // One tracking step of the current ray; sets `event` when the lane needs a
    // handler (collision, end of a transmittance ray, end of the path).
    VR_DEV void step(const MediumView& M) {
        if (TKIND == VR_T_DT) {
            step_dt(M);
            return;
        }
        float md;
        float mq = walk.query(M, md);
        bool need_sigma = false;
        float m = 1.0f, t = 0.0f;
        if (mode == kModePath) {
            if (!(rd > EPS)) {
                event = kEvPathEnd;
            } else {
                m = fmaxf(EPS, mq);
                t = neg_log1m(rng_uniform(rng)) * rcp(m);
                rx = rx + rw * fminf(t, md);
                if (t > md) {
                    rd -= md;
                    walk.cross(M, rw);
                    if (rd <= kCrossEps) event = kEvPathEnd;
                } else if (t > rd) {
                    event = kEvPathEnd;
                } else {
                    walk.move(t);
                    need_sigma = true;
                }
            }
        } else {
            m = fmaxf(0.0000001f, mq);
            if (TKIND == VR_T_DT) {  // vr_transmittance_DT.h:10-27
                float dt = neg_log1m(rng_uniform(rng)) * rcp(m);
                float tt = fminf(md, dt);
                rx = rx + rw * tt;
                rd -= tt;
                if (dt > md) {
                    walk.cross(M, rw);
                    if (rd <= kCrossEps) { event = kEvTrackEnd; track_T = 1.0f; }
                } else {
                    walk.move(tt);
                    if (rd <= 0.0000001f) { event = kEvTrackEnd; track_T = 1.0f; }
                    else need_sigma = true;
                }
            } else {  // vr_transmittance_RT.h:102-119
                float sd = neg_log1m(rng_uniform(rng)) * rcp(m);
                if (Walk<MAJ>::grid && md < rd && sd > md) {
                    rx = rx + rw * md;
                    rd -= md;
                    walk.cross(M, rw);
                    if (rd <= kCrossEps) { event = kEvTrackEnd; track_T = Tt; }
                } else {
                    float dt = fminf(sd, fminf(rd, md));
                    rx = rx + rw * dt;
                    rd -= dt;
                    walk.move(dt);
                    if (rd <= 0.0000001f) { event = kEvTrackEnd; track_T = Tt; }
                    else need_sigma = true;
                }
            }
        }

        if (need_sigma) {
            // the shared trilinear gather
            float sg = LEAN ? cells_lookup1(M.sigma, M.cells, rx) : sigma_lookup(M, rx);
            lookups++;
            if (mode == kModePath) {
                if (rng_uniform(rng) * m >= sg) rd -= t;  // null collision
                else event = kEvCollision;
            } else if (TKIND == VR_T_DT) {
                if (rng_uniform(rng) * m < sg) { event = kEvTrackEnd; track_T = 0.0f; }
            } else {
                // ratio tracking draws its roulette uniform only when T < 0.01
                Tt *= (1.0f - sg * rcp(m));
                if (Tt < 0.01f) {
                    if (rng_uniform(rng) < 1.0f - Tt) { event = kEvTrackEnd; track_T = 0.0f; }
                    else Tt = 1.0f;
                }
            }
        }
    }

    // step() with DT shadow rays: the path's free flight (vr_path_integrator_NEE.h:18-38)
    // and the DT shadow ray's (vr_transmittance_DT.h:10-27) differ only in the
    // majorant floor, the end-of-ray tests and when the remaining distance is
    // decremented, so both modes run one code path with per-lane selects (a wave
    // holding both would otherwise execute both branches).  Same draws and
    // arithmetic per lane as the two branches.
    VR_DEV void step_dt(const MediumView& M) {
        const bool path = mode == kModePath;
        if (path && !(rd > EPS)) {
            event = kEvPathEnd;
            return;
        }
        float md;
        const float mq = walk.query(M, md);
        const float m = fmaxf(path ? EPS : 0.0000001f, mq);
        const float t = neg_log1m(rng_uniform(rng)) * rcp(m);
        rx = rx + rw * fminf(t, md);
        if (t > md) {  // majorant-cell crossing
            rd -= md;
            walk.cross(M, rw);
            if (rd <= kCrossEps) {
                event = path ? kEvPathEnd : kEvTrackEnd;
                track_T = 1.0f;
            }
            return;
        }
        // path: the tentative collision lies beyond the exit (rd is reduced only
        // at a null collision); shadow ray: rd is reduced first, then tested
        const float rest = rd - t;
        if (path ? t > rd : rest <= 0.0000001f) {
            event = path ? kEvPathEnd : kEvTrackEnd;
            track_T = 1.0f;
            return;
        }
        walk.move(t);
        if (!path) rd = rest;
        const float sg = LEAN ? cells_lookup1(M.sigma, M.cells, rx) : sigma_lookup(M, rx);
        lookups++;
        const bool real = rng_uniform(rng) * m < sg;
        if (path) {
            if (real) event = kEvCollision;
            else rd = rest;  // null collision
        } else if (real) {
            event = kEvTrackEnd;
            track_T = 0.0f;
        }
    }

    // Real collision of the path (vr_path_integrator_NEE.h:39-52): emission,
    // albedo, light sample; the shadow ray starts at the collision.
    VR_DEV void on_collision(const MediumView& M) {
        collision_terms(M);
        f3 wnee, env;
        float ph;
        light_sample(M, rng, LEAN ? M.g : phase_g_at(M, px), pw, wnee, env, ph);
        collision_light(wnee, env, ph);
    }

    // The parts of on_collision: the collision's own terms, the light sample
    // (direction, environment and phase value; the generator's only draws) and
    // the shadow ray it starts.
    VR_DEV void collision_terms(const MediumView& M) {
        f3 sa;
        if (LEAN || !M.albedo_grid) sa = M.albedo_c;
        else {
            float v[4];
            grid_lookupC(M.albedo, rx, v);
            sa = M.albedo.C == 1 ? F3(v[0], v[0], v[0]) : F3(v[0], v[1], v[2]);
        }
        R = R + (W * (F3(1.0f, 1.0f, 1.0f) - sa)) * M.emission_c;
        W = W * sa;
        px = rx;
        pw = rw;
    }
    static VR_DEV void light_sample(const MediumView& M, Rng& g, float pg, f3 wi, f3& wnee, f3& env, float& ph) {
        env = ENVMAP ? env_sample(M, g, wnee) : env_sample_const(M, g, wnee);
        ph = phase_eval_g(M, pg, wi, wnee);
    }
    VR_DEV void collision_light(f3 wnee, f3 env, float ph) {
        contrib = (W * env) * ph;
        rw = wnee;
        mode = kModeTrack;
        bnd = 1;
        event = kEvStart;
    }

    // A transmittance ray ended: add its term, then start the path (after the
    // primary term) or continue it after the NEE estimate (vr_path_integrator_NEE.h:53-62).
    VR_DEV void on_track_end(const MediumView& M) {
        if (!track_end_terms(M)) return;
        f3 wo;
        const float rho = phase_sample_g(M, rng, LEAN ? M.g : phase_g_at(M, px), pw, wo);
        track_end_bounce(wo, rho);
    }

    // The parts of on_track_end: the ray's term and the bounce bookkeeping
    // (true: the path continues and needs a phase-function sample at px along
    // pw), then the scattered direction and weight.
    VR_DEV bool track_end_terms(const MediumView& M) {
        R = R + contrib * track_T;
        rx = px;
        mode = kModePath;
        event = kEvStart;
        if (primary) {
            primary = 0;
            rw = pw;
            rd = pd;
            bnd = 0;
            return false;
        }
        bounces++;
        if (M.max_bounces >= 0 && bounces >= M.max_bounces) { event = kEvPathEnd; return false; }
        return true;
    }
    VR_DEV void track_end_bounce(f3 wo, float rho) {
        W = W * rho;
        if (W.x < EPS && W.y < EPS && W.z < EPS) event = kEvPathEnd;
        // stored on both branches (unused once the path ends): stores on one
        // branch only would be merged through a pointer select into scratch
        rw = wo;
        bnd = 1;
    }

    // Runs one event handler or one tracking step (host emulation and tests).
    VR_DEV void advance(const MediumView& M) {
        switch (event) {
        case kEvStart: start_ray(M); break;
        case kEvNone: step(M); break;
        case kEvCollision: on_collision(M); break;
        case kEvTrackEnd: on_track_end(M); break;
        default: break;
        }
    }
};

// Launch-wide schedule of a persistent NEE capture.  Sensors [start, start+cnt)
// of a capture over `total` sensors are cut into `nq` contiguous ranges of `qpix`
// sensors (the last one shorter), one work queue each.  Queue q hands out items
// of `chunk` consecutive samples: its item l covers sensor q*qpix + l % n_q and
// samples [c*chunk, min(spp, (c+1)*chunk)) with c = l / n_q, n_q the queue's
// sensor count.  All counts fit 32 bits (the host splits larger captures).
struct NeeSchedule {
    uint32_t total, start, cnt;
    uint32_t qpix;  // sensors per queue
    int nq;         // work queues (one per XCD); waves start on queue blockIdx.x % nq
    int spp, chunk;
    int chunks;     // ceil(spp / chunk)
    int tiled;      // whole images without a sensor list: visit each pose in 8x8 tiles
    int ksteps;     // tracking steps per iteration of the event loop
    int kmax;       // at most this many steps when ...
    int track_min;  // ... at least this many lanes of the wave are still tracking after ksteps
    int min_begin;  // new samples start once this many lanes wait for one (or no lane tracks)
};

// Sensors and items of queue q.
VR_DEV uint32_t nee_queue_sensors(const NeeSchedule& S, int q) {
    const uint32_t b = (uint32_t)q * S.qpix;
    return b >= S.cnt ? 0u : min(S.qpix, S.cnt - b);
}

// Pixel of the k-th position of a width x height image visited in bands of 8
// rows and, within a band, in blocks of 8 columns (row-major inside a block),
// so that 64 consecutive positions cover an 8x8 tile.
VR_DEV void tiled_pixel(uint32_t k, uint32_t width, uint32_t height, int& row, int& col) {
    const uint32_t band = k / (8u * width);
    const uint32_t m = k - band * 8u * width;
    const uint32_t bh = min(8u, height - band * 8u);
    const uint32_t blk = m / (8u * bh);
    const uint32_t q = m - blk * 8u * bh;
    const uint32_t bw = min(8u, width - blk * 8u);
    const uint32_t r = q / bw;
    row = (int)(band * 8u + r);
    col = (int)(blk * 8u + (q - r * bw));
}

// Decodes item l of queue q; false if its pose index lies outside the pose
// buffer (rows and columns are used as given, as the reference does).
VR_DEV bool nee_item(const NeeSchedule& S, const CameraView& cam, int q, uint32_t l, uint32_t& sensor, int& s0, int& s1,
                     int& pose, int& row, int& col) {
    const uint32_t nqs = nee_queue_sensors(S, q);
    const uint32_t c = l / nqs;
    const uint32_t idx = S.start + (uint32_t)q * S.qpix + (l - c * nqs);
    s0 = (int)c * S.chunk;
    s1 = min(S.spp, s0 + S.chunk);
    sensor = idx;
    if (cam.sensors) {
        pose = (int)cam.sensors[3 * (size_t)idx];
        row = (int)cam.sensors[3 * (size_t)idx + 1];
        col = (int)cam.sensors[3 * (size_t)idx + 2];
        return pose >= 0 && pose < cam.num_poses;
    }
    const uint32_t w = (uint32_t)cam.width, hw = w * (uint32_t)cam.height;
    const uint32_t p = idx / hw;
    const uint32_t k = idx - p * hw;
    pose = (int)p;
    if (S.tiled) {
        tiled_pixel(k, w, (uint32_t)cam.height, row, col);
        sensor = p * hw + (uint32_t)row * w + (uint32_t)col;
    } else {
        const uint32_t r = k / w;
        row = (int)r;
        col = (int)(k - r * w);
    }
    return true;
}

// Resident waves per SIMD the persistent kernel is compiled for, which caps its
// registers: 4 waves (128 VGPRs) for the constant-environment forms with DT
// shadow rays, 3 (168 VGPRs) for the larger forms.
template <int MAJ, int TKIND, bool ENVMAP, bool LEAN>
constexpr int nee_waves_per_eu() {
    return (!ENVMAP && TKIND == VR_T_DT && MAJ == VR_MAJ_GRID) ? 4 : 3;
}

// Item indices a wave takes from a queue counter per atomic.
constexpr int kPool = 64;
// Work queues of a launch (one per XCD) and the spacing of their counters
// (one 128-byte cache line each).
constexpr int kMaxQueues = 8;
constexpr int kQueueStride = 16;

VR_DEV unsigned long long lane_mask_lt() {
    unsigned lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// A wave's share of one launch-wide work counter (the pooled backward
// kernels): the lanes that need an item take consecutive indices from the
// wave's pool, which one atomic refills with kPool indices when it is empty.
// Every lane of the wave calls take() with the wave's ballot of `need`; a lane
// that needs an item gets `item` (true) or, once the counter is drained and the
// pool empty, is marked `done`.
struct ItemPool {
    uint32_t next = 0, end = 0;  // wave-uniform
    bool drained = false;

    VR_DEV bool take(unsigned long long mask, bool need, unsigned long long* counter, uint32_t items, uint32_t& item,
                     bool& done) {
        while (next >= end && !drained) {
            unsigned long long base = 0;
            if (__lane_id() == 0) base = atomicAdd(counter, (unsigned long long)kPool);
            base = __shfl(base, 0);
            if (base < items) {
                next = (uint32_t)base;
                end = items - next > (uint32_t)kPool ? next + (uint32_t)kPool : items;
            } else {
                drained = true;
            }
        }
        const uint32_t avail = end - next;
        const uint32_t rank = (uint32_t)__popcll(mask & lane_mask_lt());
        bool got = false;
        if (need) {
            if (rank < avail) {
                item = next + rank;
                got = true;
            } else if (drained) {
                done = true;
            }
        }
        const uint32_t taken = (uint32_t)__popcll(mask);
        next += taken < avail ? taken : avail;
        return got;
    }
};

// Persistent kernel: every lane pulls work items until the queues drain.  A
// wave starts on the queue of its XCD slot (blockIdx.x % nq: workgroups are
// dispatched to the XCDs round-robin), so each XCD's L2 serves one contiguous
// part of the image, and moves on to the next queue when that one is empty.
// The output must be zero on entry (k_nee_prepare); item sums are added
// atomically.
// Every argument of the persistent kernel, passed as its only argument, so the
// structure starts at offset 0 of the kernel-argument segment.
struct NeeArgs {
    MediumView M;
    CameraView cam;
    Seeds seeds;
    NeeSchedule S;
    unsigned long long* counters;
    float* out;
    int* dbg;
    // cached render (CACHE): a path stops at its (cache_k + 1)-th real collision
    // and appends the query (x, w, W / spp, sensor) at *qcount
    float* queries;
    int64_t max_queries;
    uint32_t* qcount;
    int cache_k;
};

// The kernel's arguments, re-read where a handler uses them (reread): the
// loads of the fields each handler reads stay in that handler.
VR_DEV const NeeArgs& nee_args() { return *(const NeeArgs*)reread(first_kernel_arg<NeeArgs>()); }

// Collision and end-of-transmittance handlers with their samples computed
// block-wide: each lane that needs a light sample (collision) or a phase
// sample (continuing path) posts its generator state and incoming direction to
// LDS at a slot ranked over the block (light samples first), the block's first
// total / 64 waves compute one slot per lane, and the posting lanes read back
// the generator state and the sample.  Every lane draws exactly what it would
// draw in the handler itself.  Called by all threads of the block.
constexpr int kCmpWaves = 4;  // waves per block of the persistent kernel (256 threads)
template <typename Lane>
VR_DEV void nee_compacted_samples(Lane& lane, bool have) {
    __shared__ float xb[11][64 * kCmpWaves];
    __shared__ int cnt[2][kCmpWaves];
    const MediumView& M = nee_args().M;
    const bool rq0 = have && lane.event == kEvCollision;
    if (rq0) lane.collision_terms(M);
    bool rq1 = false;
    if (have && lane.event == kEvTrackEnd) rq1 = lane.track_end_terms(M);
    const unsigned long long m0 = __ballot(rq0), m1 = __ballot(rq1);
    const int wv = (int)(threadIdx.x / 64);
    if (__lane_id() == 0) {
        cnt[0][wv] = __popcll(m0);
        cnt[1][wv] = __popcll(m1);
    }
    __syncthreads();
    int n0 = 0, n1 = 0, b0 = 0, b1 = 0;
#pragma unroll
    for (int v = 0; v < kCmpWaves; v++) {
        const int c0 = cnt[0][v], c1 = cnt[1][v];
        if (v < wv) {
            b0 += c0;
            b1 += c1;
        }
        n0 += c0;
        n1 += c1;
    }
    const int total = n0 + n1;
    if (total == 0) return;  // the same for the whole block
    const unsigned long long lt = lane_mask_lt();
    const int slot = rq0 ? b0 + __popcll(m0 & lt) : n0 + b1 + __popcll(m1 & lt);
    if (rq0 || rq1) {
        xb[0][slot] = __uint_as_float(lane.rng.x);
        xb[1][slot] = __uint_as_float(lane.rng.y);
        xb[2][slot] = __uint_as_float(lane.rng.z);
        xb[3][slot] = __uint_as_float(lane.rng.w);
        xb[4][slot] = lane.pw.x;
        xb[5][slot] = lane.pw.y;
        xb[6][slot] = lane.pw.z;
    }
    __syncthreads();
    const int s = (int)threadIdx.x;
    if (s < total) {
        Rng g;
        g.x = __float_as_uint(xb[0][s]);
        g.y = __float_as_uint(xb[1][s]);
        g.z = __float_as_uint(xb[2][s]);
        g.w = __float_as_uint(xb[3][s]);
        const f3 wi = F3(xb[4][s], xb[5][s], xb[6][s]);
        f3 d, e = F3(0.0f, 0.0f, 0.0f);
        float v;
        if (s < n0) Lane::light_sample(M, g, M.g, wi, d, e, v);
        else v = phase_sample_g(M, g, M.g, wi, d);
        xb[0][s] = __uint_as_float(g.x);
        xb[1][s] = __uint_as_float(g.y);
        xb[2][s] = __uint_as_float(g.z);
        xb[3][s] = __uint_as_float(g.w);
        xb[4][s] = d.x;
        xb[5][s] = d.y;
        xb[6][s] = d.z;
        xb[7][s] = v;
        xb[8][s] = e.x;
        xb[9][s] = e.y;
        xb[10][s] = e.z;
    }
    __syncthreads();
    if (rq0 || rq1) {
        lane.rng.x = __float_as_uint(xb[0][slot]);
        lane.rng.y = __float_as_uint(xb[1][slot]);
        lane.rng.z = __float_as_uint(xb[2][slot]);
        lane.rng.w = __float_as_uint(xb[3][slot]);
        const f3 d = F3(xb[4][slot], xb[5][slot], xb[6][slot]);
        const float v = xb[7][slot];
        if (rq0) lane.collision_light(d, F3(xb[8][slot], xb[9][slot], xb[10][slot]), v);
        else lane.track_end_bounce(d, v);
    }
}

// CACHE: the radiance-cache render -- paths end at their (cache_k + 1)-th real
// collision, which becomes a query of the cache (the medium's max_bounces must
// be -1: the cached estimator has no other cap), as nrc::nee_cached.
// CMP (lean constant-environment forms): the light samples of the collision
// handler and the phase samples of the end-of-transmittance handler are
// gathered block-wide through LDS and computed by the block's first waves
// (one lane per request), then handed back; the 4 waves of a block run the
// event loop in step (three barriers per iteration, one more for its end).
template <int MAJ, int TKIND, bool ENVMAP, bool LEAN, bool CACHE = false, bool CMP = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(nee_waves_per_eu<MAJ, TKIND, ENVMAP, LEAN>())))
k_nee_persistent(NeeArgs args) {
    static_assert(!CMP || (LEAN && !ENVMAP), "compacted samples: lean constant-environment forms only");
    if constexpr (MAJ == kMajGridLds) stage_majorant(args.M);
    const NeeSchedule& S = args.S;
    unsigned long long* __restrict__ counters = args.counters;
    float* __restrict__ out = args.out;
    int* __restrict__ dbg = args.dbg;
    const float inv = S.spp == 1 ? 1.0f : 1.0f / (float)S.spp;
    NeeLane<MAJ, TKIND, ENVMAP, LEAN> lane;
    lane.lookups = 0;
    lane.event = kEvNone;
    bool have = false, done = false;
    // wave-uniform queue state: current queue, queues tried, pool of item indices
    int q = (int)(blockIdx.x % (unsigned)S.nq), tried = 1;
    bool drained = false;
    uint32_t pool_next = 0, pool_end = 0;
    uint32_t sensor = 0;
    int s = 0, s1 = 0, pose = 0, row = 0, col = 0;
    f3 acc = F3(0.0f, 0.0f, 0.0f);
    for (;;) {
        // end of a sample: accumulate, then the next sample of the item or a new item
        if (have && lane.event == kEvPathEnd) {
            acc = acc + lane.R;
            s++;
            if (s < s1) {
                lane.event = kEvBegin;
            } else {
                float* o = out + 3 * (size_t)sensor;
                atomic_addf(o, acc.x * inv);
                atomic_addf(o + 1, acc.y * inv);
                atomic_addf(o + 2, acc.z * inv);
                if (dbg) atomicAdd(dbg + sensor, lane.lookups);
                have = false;
                lane.event = kEvNone;
            }
        }
        // Lanes that need an item take one from the wave's pool of consecutive
        // item indices; an empty pool is refilled with kPool items by one atomic.
        const bool need = !have && !done;
        const unsigned long long mask = __ballot(need);
        if (mask) {
            while (pool_next >= pool_end && !drained) {
                const uint32_t items = nee_queue_sensors(S, q) * (uint32_t)S.chunks;
                unsigned long long base = 0;
                if (__lane_id() == 0) base = atomicAdd(counters + q * kQueueStride, (unsigned long long)kPool);
                base = __shfl(base, 0);
                if (base < items) {
                    pool_next = (uint32_t)base;
                    pool_end = items - pool_next > (uint32_t)kPool ? pool_next + (uint32_t)kPool : items;
                } else if (tried < S.nq) {
                    q = q + 1 == S.nq ? 0 : q + 1;
                    tried++;
                } else {
                    drained = true;
                }
            }
            const uint32_t avail = pool_end - pool_next;
            const uint32_t rank = (uint32_t)__popcll(mask & lane_mask_lt());
            if (need) {
                if (rank < avail) {
                    int s0;
                    if (nee_item(S, nee_args().cam, q, pool_next + rank, sensor, s0, s1, pose, row, col)) {
                        have = true;
                        s = s0;
                        acc = F3(0.0f, 0.0f, 0.0f);
                        lane.lookups = 0;
                        lane.event = kEvBegin;
                    }
                } else if (drained) {
                    done = true;
                }
            }
            const uint32_t taken = (uint32_t)__popcll(mask);
            pool_next += taken < avail ? taken : avail;
        }
        if constexpr (CMP) {
            // the block iterates while any of its lanes has or awaits work
            if (!__syncthreads_or((int)(have || !done))) break;
        } else {
            if (__ballot(have) == 0ull) {
                if (__ballot(!done) == 0ull) break;  // every lane of the wave has drained the queues
                continue;
            }
        }
        // Starting a sample (seeding, camera ray, box entry) is the costliest
        // handler; lanes waiting for one idle while the others track until
        // S.min_begin of them can start together.
        const bool tracks = __ballot(have && lane.event == kEvNone) != 0ull;
        if (!tracks || __popcll(__ballot(have && lane.event == kEvBegin)) >= S.min_begin)
            if (have && lane.event == kEvBegin) {
                const NeeArgs& A = nee_args();
                lane.begin_sample(A.M, A.cam, A.seeds, (uint32_t)s * S.total + sensor, pose, row, col);
            }
        if (have && lane.event == kEvStart) lane.start_ray(nee_args().M);
        const MediumView& Mt = nee_args().M;
        for (int k = 0;;) {
            if (have && lane.event == kEvNone) lane.step(Mt);
            ++k;
            const unsigned long long tracking = __ballot(have && lane.event == kEvNone);
            if (tracking == 0ull || k >= S.kmax) break;
            if (k >= S.ksteps && __popcll(tracking) < S.track_min) break;
        }
        if constexpr (CACHE) {
            // the (cache_k + 1)-th collision: the lanes there append their queries
            // (one atomic per wave) and end their paths
            const bool query = have && lane.event == kEvCollision && lane.bounces == nee_args().cache_k;
            const unsigned long long qm = __ballot(query);
            if (qm) {
                const NeeArgs& A = nee_args();
                const int leader = __ffsll((long long)qm) - 1;
                unsigned long long base = 0;
                if ((int)__lane_id() == leader) base = atomicAdd(A.qcount, (uint32_t)__popcll(qm));
                base = __shfl(base, leader);
                if (query) {
                    const uint64_t slot = base + (uint64_t)__popcll(qm & lane_mask_lt());
                    if ((int64_t)slot < A.max_queries) {
                        float* q = A.queries + slot * 10;
                        q[0] = lane.rx.x; q[1] = lane.rx.y; q[2] = lane.rx.z;
                        q[3] = lane.rw.x; q[4] = lane.rw.y; q[5] = lane.rw.z;
                        q[6] = lane.W.x * inv; q[7] = lane.W.y * inv; q[8] = lane.W.z * inv;
                        q[9] = __uint_as_float(sensor);
                    }
                    lane.event = kEvPathEnd;
                }
            }
        }
        if constexpr (CMP) {
            nee_compacted_samples(lane, have);
        } else {
            if (have && lane.event == kEvCollision) lane.on_collision(nee_args().M);
            if (have && lane.event == kEvTrackEnd) lane.on_track_end(nee_args().M);
        }
    }
}

// Zeroes the capture output rows [start, start+cnt) (3 floats each) and the queue counters.
__global__ void k_nee_prepare(float* __restrict__ out, int64_t start, int64_t cnt, int* __restrict__ dbg,
                              unsigned long long* __restrict__ counters) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    if (i < kMaxQueues) counters[i * kQueueStride] = 0ull;
    for (; i < cnt; i += stride) {
        out[3 * (start + i)] = 0.0f;
        out[3 * (start + i) + 1] = 0.0f;
        out[3 * (start + i) + 2] = 0.0f;
        if (dbg) dbg[start + i] = 0;
    }
}

}  // namespace vr
