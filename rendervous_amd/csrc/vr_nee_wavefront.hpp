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
// one free-flight draw and one trilinear gather shared by all lanes.  Lanes
// pull work items (a pixel and a run of consecutive samples) from a global
// counter with one atomic per wave, so every lane stays busy until the queue
// drains.  Each sample still draws the stream seeded from its flat index
// s * pixels + p, in the same order as the nested form (Tracer::nee), so
// results are identical sample by sample.
#pragma once

#include "vr_device.hpp"

namespace vr {

enum : int { kModePath = 0, kModeTrack = 1 };

// Lane events: what a lane needs before it can take its next tracking step.
enum : int { kEvNone = 0, kEvBegin = 1, kEvCollision = 2, kEvTrackEnd = 3, kEvPathEnd = 4 };

template <int MAJ, int TKIND>
struct NeeLane {
    static constexpr float EPS = 0.00001f;  // vr_path_integrator_NEE.h:16

    Rng rng;
    Walk<MAJ> walk;
    f3 R, W, contrib;
    f3 rx, rw;        // tracked ray: current point and direction
    float rd;         // remaining distance on the tracked ray
    f3 px, pw;        // path vertex and incoming direction while a shadow ray runs
    float Tt;         // ratio-tracking estimate of the tracked ray
    float pd;         // path distance inside the medium while the primary term runs
    float track_T;    // transmittance of a finished tracked ray (kEvTrackEnd)
    int mode, primary, bounces, lookups, event;

    // Starts the transmittance estimate of ray (o, dir) (vr_transmittance_DT.h:3-8,
    // vr_transmittance_RT.h:95-100); false when the boundary is missed (T = 1).
    VR_DEV bool start_track(const MediumView& M, f3 o, f3 dir) {
        float tmin, tmax;
        if (!boundary(M, o, dir, tmin, tmax)) return false;
        rx = o + dir * tmin;
        rw = dir;
        rd = tmax - tmin;
        walk.begin(M, rx, rw);
        Tt = 1.0f;
        mode = kModeTrack;
        return true;
    }

    VR_DEV void start_path_segment(const MediumView& M, f3 x, f3 w, float d) {
        rx = x;
        rw = w;
        rd = d;
        walk.begin(M, rx, rw);
        mode = kModePath;
    }

    // Camera ray and primary transmittance term of a new sample (kEvBegin).
    VR_DEV void begin_sample(const MediumView& M, const CameraView& cam, const Seeds& seeds, uint32_t seed_index, int pose,
                             int row, int col) {
        rng = seed_capture_fw(seeds, seed_index);
        f3 x, w;
        camera_ray(cam, pose, row, col, rng, x, w);
        float tmin, tmax;
        boundary(M, x, w, tmin, tmax);
        px = x + w * tmin;  // the path starts at the box entry (vr_path_integrator_NEE.h:3-6)
        pw = w;
        pd = tmax - tmin;
        R = F3(0.0f, 0.0f, 0.0f);
        W = F3(1.0f, 1.0f, 1.0f);
        bounces = 0;
        contrib = M.env_c;
        primary = 1;
        event = kEvNone;
        if (!start_track(M, px, pw)) {
            R = R + contrib;  // T = 1
            primary = 0;
            start_path_segment(M, px, pw, pd);
        }
    }

    // One tracking step of the current ray; sets `event` when the lane needs a
    // handler (collision, end of a transmittance ray, end of the path).
    VR_DEV void step(const MediumView& M) {
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
            float sg = grid_lookup1(M.sigma, rx);
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

    // Real collision of the path (vr_path_integrator_NEE.h:39-52): emission, albedo,
    // light sample and the start of its shadow ray.
    VR_DEV void on_collision(const MediumView& M) {
        f3 sa;
        if (!M.albedo_grid) sa = M.albedo_c;
        else {
            float v[4];
            grid_lookupC(M.albedo, rx, v);
            sa = M.albedo.C == 1 ? F3(v[0], v[0], v[0]) : F3(v[0], v[1], v[2]);
        }
        R = R + (W * (F3(1.0f, 1.0f, 1.0f) - sa)) * M.emission_c;
        W = W * sa;
        f3 wnee;
        f3 env = env_sample(M, rng, wnee);
        px = rx;
        pw = rw;
        contrib = (W * env) * phase_eval(M, rw, wnee);
        event = kEvNone;
        if (!start_track(M, px, wnee)) { event = kEvTrackEnd; track_T = 1.0f; }
    }

    // A transmittance ray ended: add its term, then start the path (after the
    // primary term) or continue it after the NEE estimate (vr_path_integrator_NEE.h:53-62).
    VR_DEV void on_track_end(const MediumView& M) {
        R = R + contrib * track_T;
        event = kEvNone;
        if (primary) {
            primary = 0;
            start_path_segment(M, px, pw, pd);
            return;
        }
        bounces++;
        if (M.max_bounces >= 0 && bounces >= M.max_bounces) { event = kEvPathEnd; return; }
        f3 wo;
        float rho = phase_sample(M, rng, pw, wo);
        W = W * rho;
        if (W.x < EPS && W.y < EPS && W.z < EPS) { event = kEvPathEnd; return; }
        float tmin, d;
        boundary(M, px, wo, tmin, d);
        start_path_segment(M, px, wo, d);
    }

    // Runs the handlers until the lane can step again or its sample has ended.
    VR_DEV void settle(const MediumView& M) {
        while (event == kEvCollision || event == kEvTrackEnd) {
            if (event == kEvCollision) on_collision(M);
            else on_track_end(M);
        }
    }
};

// Work item j of a capture over `pixels` sensors in chunks of `chunk` samples:
// sensor start + j % pixels, samples [c*chunk, min(spp, (c+1)*chunk)) with c = j / pixels.
struct NeeItem {
    int64_t sensor;
    int s0, s1;
};

VR_DEV NeeItem nee_item(int64_t j, int64_t start, int64_t pixels, int chunk, int spp) {
    NeeItem it;
    int64_t c = j / pixels;
    it.sensor = start + (j - c * pixels);
    it.s0 = (int)c * chunk;
    it.s1 = min(spp, it.s0 + chunk);
    return it;
}

// Bijection of [0, width*height) that visits an image in bands of 8 rows and,
// within a band, in blocks of 8 columns (row-major inside a block), so that 64
// consecutive indices cover an 8x8 tile.  Returns the row-major pixel index.
VR_DEV int64_t tiled_pixel(int64_t k, int width, int height) {
    const int64_t band = k / (8 * (int64_t)width);
    const int64_t m = k - band * 8 * (int64_t)width;
    const int bh = min(8, height - (int)band * 8);
    const int64_t blk = m / (8 * bh);
    const int q = (int)(m - blk * 8 * bh);
    const int bw = min(8, width - (int)blk * 8);
    const int r = q / bw, c = q - (q / bw) * bw;
    return (band * 8 + r) * (int64_t)width + blk * 8 + c;
}

// Item indices a wave takes from the global counter per atomic.
constexpr int kPool = 64;

VR_DEV unsigned long long lane_mask_lt() {
    unsigned lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Persistent kernel: every lane pulls work items until the queue drains.  The
// output must be zero on entry (k_fill_zero); item sums are added atomically.
template <int MAJ, int TKIND>
__global__ void __launch_bounds__(256) k_nee_persistent(MediumView M, CameraView cam, Seeds seeds, int spp, int chunk,
                                                        int64_t total, int64_t start, int64_t cnt,
                                                        unsigned long long* __restrict__ counter, float* __restrict__ out,
                                                        int* __restrict__ dbg) {
    const int64_t chunks = (spp + chunk - 1) / chunk;
    const bool tiled = cam.sensors == nullptr && start == 0 && cnt == total;
    const unsigned long long items = (unsigned long long)(cnt * chunks);
    const float inv = spp == 1 ? 1.0f : 1.0f / (float)spp;
    NeeLane<MAJ, TKIND> lane;
    lane.lookups = 0;
    lane.event = kEvNone;
    bool have = false, done = false;
    bool drained = false;                              // wave-uniform: the global queue is empty
    unsigned long long pool_next = 0ull, pool_end = 0ull;  // wave-uniform pool of item indices
    NeeItem it;
    it.sensor = 0; it.s0 = 0; it.s1 = 0;
    int s = 0, pose = 0, row = 0, col = 0;
    f3 acc = F3(0.0f, 0.0f, 0.0f);
    for (;;) {
        // end of a sample: accumulate, then the next sample of the item or a new item
        if (have && lane.event == kEvPathEnd) {
            acc = acc + lane.R;
            s++;
            if (s < it.s1) {
                lane.event = kEvBegin;
            } else {
                float* o = out + 3 * it.sensor;
                atomic_addf(o, acc.x * inv);
                atomic_addf(o + 1, acc.y * inv);
                atomic_addf(o + 2, acc.z * inv);
                if (dbg) atomicAdd(dbg + it.sensor, lane.lookups);
                have = false;
                lane.event = kEvNone;
            }
        }
        // Lanes that need an item take one from the wave's pool of consecutive
        // item indices; an empty pool is refilled with kPool items by one atomic.
        const bool need = !have && !done;
        const unsigned long long mask = __ballot(need);
        if (mask) {
            if (pool_next >= pool_end && !drained) {
                unsigned long long base = 0;
                if (__lane_id() == 0) base = atomicAdd(counter, (unsigned long long)kPool);
                base = __shfl(base, 0);
                pool_next = base;
                pool_end = base + kPool < items ? base + kPool : items;
                if (base >= items) drained = true;
            }
            const unsigned long long avail = pool_end > pool_next ? pool_end - pool_next : 0ull;
            const unsigned long long rank = (unsigned long long)__popcll(mask & lane_mask_lt());
            if (need) {
                if (rank < avail) {
                    it = nee_item((int64_t)(pool_next + rank), start, cnt, chunk, spp);
                    if (tiled) {  // whole images: visit each pose's pixels in 8x8 tiles
                        const int64_t hw = (int64_t)cam.width * cam.height;
                        const int64_t pose_base = (it.sensor / hw) * hw;
                        it.sensor = pose_base + tiled_pixel(it.sensor - pose_base, cam.width, cam.height);
                    }
                    sensor_index(cam, it.sensor, pose, row, col);
                    if (pose >= 0 && pose < cam.num_poses && row >= 0 && row < cam.height && col >= 0 && col < cam.width) {
                        have = true;
                        s = it.s0;
                        acc = F3(0.0f, 0.0f, 0.0f);
                        lane.lookups = 0;
                        lane.event = kEvBegin;
                    }
                } else if (drained) {
                    done = true;
                }
            }
            const unsigned long long taken = (unsigned long long)__popcll(mask);
            pool_next += taken < avail ? taken : avail;
        }
        if (__ballot(have) == 0ull) {
            if (__ballot(!done) == 0ull) break;  // every lane of the wave has drained the queue
            continue;
        }
        if (have && lane.event == kEvBegin)
            lane.begin_sample(M, cam, seeds, (uint32_t)((int64_t)s * total + it.sensor), pose, row, col);
        if (have && lane.event == kEvNone) lane.step(M);
        if (have && lane.event == kEvCollision) lane.on_collision(M);
        if (have && lane.event == kEvTrackEnd) lane.on_track_end(M);
    }
}

// Zeroes the capture output rows [start, start+cnt) (3 floats each) and the work counter.
__global__ void k_nee_prepare(float* __restrict__ out, int64_t start, int64_t cnt, int* __restrict__ dbg,
                              unsigned long long* __restrict__ counter) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    if (i == 0) *counter = 0ull;
    for (; i < cnt; i += stride) {
        out[3 * (start + i)] = 0.0f;
        out[3 * (start + i) + 1] = 0.0f;
        out[3 * (start + i) + 2] = 0.0f;
        if (dbg) dbg[start + i] = 0;
    }
}

}  // namespace vr
