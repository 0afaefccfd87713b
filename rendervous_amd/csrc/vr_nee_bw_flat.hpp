// vr_nee_bw_flat.hpp -- the defensive-replay NEE backward as a flat lane state
// machine.
//
// Tracer::nee_ds_bw (include/vr_path_integrator_NEE_DS.h:6-206) nests, inside
// each of its two path passes, a shadow-ray loop per collision, and in the
// second pass two more (the DT transmittance backward's primal pass and its
// replay, vr_transmittance_DT.h:30-90).  On a 64-wide wave the nesting
// serialises: lanes on their path wait while other lanes run shadow rays.
// Here each iteration takes a lane through the events it waits on (a real
// collision, the end of a shadow ray or of its backward, the end of a path) and
// then up to 3 tracking steps of whichever ray it is on (pass-1 path, pass-2
// path, shadow ray, shadow-ray backward primal or replay), so the lanes of a
// wave step their rays together; a thread runs its samples back to back.  The random draws
// of every sample are those of Tracer::nee_ds_bw, in the same order, so the
// gradients are the same sample by sample (DT shadow rays; the other
// transmittance kinds keep the nested form).
#pragma once

#include "vr_device.hpp"

namespace vr {

enum : int {
    kBwPath = 0,          // a tracking step of the path ray (pass 1 or 2)
    kBwShadowFw = 1,      // a tracking step of the shadow ray's DT estimate
    kBwShadowPrimal = 2,  // ... of the shadow-ray backward's primal pass (majorant + 0.1)
    kBwShadowReplay = 3,  // ... of its replay, scattering at every null collision
    kBwPathEnd = 4,       // the path of the current pass ended
    kBwAfterT1 = 5,       // pass 1: the shadow ray's T is known
    kBwAfterT2 = 6,       // pass 2: T known, start the shadow-ray backward
    kBwAfterTbw = 7,      // pass 2: the shadow-ray backward ended
    kBwDone = 8,          // the sample ended
    kBwCollision = 9,     // the path ray has a real collision (s, m, Pc saved)
    kBwStart = 10         // a ray waits for its set-up (start_kind) in the shared start block
};

// Ray set-ups of the shared start block.
enum : int {
    kStPath = 0,      // a pass's first path segment: clipped to the medium box
    kStContinue = 1,  // the path continues from a collision to the box exit
    kStShadowFw = 2,  // the shadow ray's DT estimate (t_dt)
    kStShadowBw = 3,  // the shadow ray's DT backward (t_dt_bw)
    kStReplay = 4     // the shadow-ray backward's replay from its saved start
};

// ENVMAP: environment image and/or quadtree sampler; false compiles the
// constant-environment (uniform + directional light) sampler only, which draws
// the same random numbers for a constant environment.
template <int MAJ, bool ENVMAP = true>
struct NeeDsBwLane {
    static constexpr float EPS = 0.00000001f;  // vr_path_integrator_NEE_DS.h:8

    Tracer<MAJ> tr;
    Walk<MAJ> walk;
    f3 x0, w0;                 // the sample's camera ray
    Rng before_path, before_T, before_next;
    f3 rx, rw;                 // current ray
    float rd;                  // its remaining distance
    f3 px, pw;                 // path vertex and direction while a shadow ray runs
    f3 W, R, R_sp, added, env, wo;
    f3 xs0;                    // start of the shadow-ray backward (for its replay)
    float ds0, null_W, rho, Tn, dL_dTn, eps;
    float cs, cm, cpc;         // sigma, majorant and collision probability of a pending collision
    f3 dL_dR;
    int state, pass, scattered, bounces, start_kind, after;
    // Pass 1's shadow-ray estimates, for pass 2 to reuse (VR_BW_CONSISTENT: both
    // passes draw the same collisions and light samples, so pass 2's estimate
    // at collision b equals pass 1's): per collision, the estimate (0 or 1, top
    // bit) and the uniforms its walk drew (the low 15 bits), kTnRecords
    // collisions per lane in LDS with a stride of the block size; kTnRedo marks
    // one pass 2 walks again.
    static constexpr int kTnRecords = 16;
    static constexpr uint16_t kTnRedo = 0xFFFFu;
    uint16_t* tn_rec;  // this lane's first record (nullptr: pass 2 walks every shadow ray)
    int tn_draws;      // uniforms drawn by the current shadow-ray estimate
    // Gradient scatters are issued by the whole wave (wave_scatter): a tracking
    // step leaves its sigma gradient at rx in pend_s, a pass-2 collision its
    // sigma and albedo gradients at px in pend_cs / pend_ca (zero: nothing).
    float pend_s, pend_cs;
    f3 pend_ca;

    VR_DEV explicit NeeDsBwLane(ConstPtr<MediumView> m)
        : tr(m), tn_rec(nullptr), pend_s(0.0f), pend_cs(0.0f), pend_ca(F3(0.0f, 0.0f, 0.0f)) {}

    // Wave-issued scatters (every lane of the wave calls these).  The sigma
    // gradients of the tracking steps and collisions gather in the wave's
    // record buffer `srec` (kScatterRecords records, count n wave-uniform) and
    // are issued when it fills and when the wave ends (flush_sigma); the albedo
    // gradients of a pass-2 collision are issued at once through `arec`.
    static constexpr int kScatterRecords = 128;
    VR_DEV int post_step(float* srec, int n) {
        const MediumView& M = tr.med();
        if (M.sigma.grad == nullptr) { pend_s = 0.0f; return n; }
        if (n > kScatterRecords - 64) { scatter_issue<1>(M.sigma, srec, n); n = 0; }
        if (pend_s != 0.0f) tr.scatters++;
        n = scatter_post<1>(M.sigma, srec, n, pend_s != 0.0f, rx, F3(pend_s, 0.0f, 0.0f));
        pend_s = 0.0f;
        return n;
    }
    VR_DEV int post_collision(float* srec, float* arec, int n) {
        const MediumView& M = tr.med();
        if (M.sigma.grad != nullptr) {
            if (n > kScatterRecords - 64) { scatter_issue<1>(M.sigma, srec, n); n = 0; }
            if (pend_cs != 0.0f) tr.scatters++;
            n = scatter_post<1>(M.sigma, srec, n, pend_cs != 0.0f, px, F3(pend_cs, 0.0f, 0.0f));
        }
        pend_cs = 0.0f;
        if (M.albedo_grid && M.albedo.grad != nullptr) {
            const bool a = pend_ca.x != 0.0f || pend_ca.y != 0.0f || pend_ca.z != 0.0f;
            if (M.albedo.C == 1) {
                const int na = scatter_post<1>(M.albedo, arec, 0, a, px, F3(pend_ca.x + pend_ca.y + pend_ca.z, 0.0f, 0.0f));
                if (na) scatter_issue<1>(M.albedo, arec, na);
            } else {
                const int na = scatter_post<3>(M.albedo, arec, 0, a, px, pend_ca);
                if (na) scatter_issue<3>(M.albedo, arec, na);
            }
        }
        pend_ca = F3(0.0f, 0.0f, 0.0f);
        return n;
    }
    VR_DEV void flush_sigma(float* srec, int n) {
        if (n) scatter_issue<1>(tr.med().sigma, srec, n);
    }

    // The same scatters issued by this lane alone (a kernel whose lanes leave
    // the loop at different times cannot share the issue).
    VR_DEV void flush_step_lane() {
        if (pend_s != 0.0f) tr.sigma_bw(rx, pend_s);
        pend_s = 0.0f;
    }
    VR_DEV void flush_collision_lane() {
        if (pend_cs != 0.0f) tr.sigma_bw(px, pend_cs);
        pend_cs = 0.0f;
        tr.albedo_bw(px, pend_ca);
        pend_ca = F3(0.0f, 0.0f, 0.0f);
    }

    // Pass 1 (or 2) of the sample from its camera ray: nee_ds_bw's start
    // (vr_path_integrator_NEE_DS.h:6-30) and its replay restart (:110-125).
    VR_DEV void begin_pass(int p) {
        pass = p;
        if (p == 1) {
            before_path = tr.rng;
            R = F3(0.0f, 0.0f, 0.0f);
            R_sp = R;
            null_W = 1.0f;
            scattered = 0;
        } else {
            tr.rng = before_path;
        }
        W = F3(1.0f, 1.0f, 1.0f);
        rx = x0;
        rw = w0;
        bounces = 0;
        pend(kStPath);
    }

    VR_DEV void pend(int kind) {
        start_kind = kind;
        state = kBwStart;
    }

    // The shared ray set-up (one block per iteration for every waiting lane):
    // the box interval of (rx, rw) and the majorant walk; the pass's first path
    // segment and the shadow rays (t_dt, t_dt_bw) are clipped to the box, a
    // continued path runs to the box exit, a shadow ray that misses the box has
    // T = 1 (its backward, nothing to do).
    VR_DEV void start() {
        const MediumView& M = tr.med();
        float tmin, tmax;
        const bool hit = boundary(M, rx, rw, tmin, tmax);
        int next = kBwPath;
        if (start_kind == kStContinue) {
            rd = tmax;
        } else if (start_kind == kStReplay) {
            rx = xs0;
            rd = ds0;
            next = kBwShadowReplay;
        } else {
            if (start_kind == kStShadowFw) next = hit ? kBwShadowFw : after;
            else if (start_kind == kStShadowBw) next = hit ? kBwShadowPrimal : kBwAfterTbw;
            if (start_kind == kStShadowFw && !hit) Tn = 1.0f;
            rx = rx + rw * tmin;
            rd = tmax - tmin;
            if (start_kind == kStShadowBw) {
                xs0 = rx;
                ds0 = rd;
            }
        }
        state = next;
        if (state <= kBwShadowReplay) walk.begin(M, rx, rw);
    }

    // The end of a shadow ray (it left the medium): its DT estimate is 1, its
    // backward's primal pass was not absorbed (replay it from the start), or
    // its replay is done.
    VR_DEV void shadow_end() {
        if (state == kBwShadowFw) {
            Tn = 1.0f;
            state = pass == 1 ? kBwAfterT1 : kBwAfterT2;
        } else if (state == kBwShadowPrimal) {
            tr.rng = before_T;
            pend(kStReplay);
        } else {
            state = kBwAfterTbw;
        }
    }

    // One tracking step of the current ray.  The path's flight (either pass)
    // and the shadow ray's (its estimate, its backward's primal pass and
    // replay) differ in the majorant, the end-of-ray tests and when the
    // remaining distance is decremented, so all kinds run one code path with
    // per-lane selects, one uniform after the lookup and one scatter site (a
    // wave holding several kinds would otherwise run each branch).  Every
    // lane's draws and arithmetic are those of Tracer::nee_ds_bw.
    VR_DEV void step() {
        const MediumView& M = tr.med();
        const bool path = state == kBwPath;
        if (path && !(rd > EPS)) { state = kBwPathEnd; return; }
        float md;
        const float q = walk.query(M, md);
        const bool ref2 = pass == 2 && M.bw_mode == VR_BW_REFERENCE;
        const float m_path = ref2 ? fmaxf(EPS, q * 1.2f + 2.0f * eps) : fmaxf(EPS, q + 2.0f * eps + 0.1f);
        const float m_shadow = fmaxf(0.0000001f, state == kBwShadowFw ? q : q + 0.1f);
        const float m = path ? m_path : m_shadow;
        const float t = tr.flight(m);
        rx = rx + rw * fminf(t, md);
        const bool shadow_fw = state == kBwShadowFw;
        tn_draws += shadow_fw ? 1 : 0;
        bool ended;
        if (t > md) {  // majorant-cell crossing
            rd -= md;
            walk.cross(M, rw);
            ended = rd <= kCrossEps;
        } else {
            // path: the tentative collision lies beyond the exit (rd is reduced
            // only at a null collision); shadow ray: rd is reduced, then tested
            const float rest = rd - t;
            ended = path ? t > rd : rest <= 0.0000001f;
            if (!ended) {
                walk.move(t);
                if (!path) rd = rest;
                // the shared trilinear lookup and the decision of each ray kind
                const float s = tr.sigma(rx);
                const float u = tr.rnd();
                tn_draws += shadow_fw ? 1 : 0;
                float g = 0.0f;
                bool scatter = false;
                if (state == kBwShadowFw) {
                    if (u * m < s) { Tn = 0.0f; state = pass == 1 ? kBwAfterT1 : kBwAfterT2; }
                } else if (state == kBwShadowPrimal) {
                    if (u < s / m) state = kBwAfterTbw;  // absorbed: no gradient
                } else if (state == kBwShadowReplay) {
                    const float Pc = s / m;
                    g = -dL_dTn / m / (1.0f - Pc);
                    scatter = true;
                } else {
                    const float Pc = (s + eps) / m;
                    if (u >= Pc) {  // null collision
                        if (pass == 2) {
                            g = dot(dL_dR, R * (-1.0f / m / (1.0f - s / m)));
                            scatter = true;
                        }
                        W = W * ((1.0f - s / m) / (1.0f - Pc));
                        rd = rest;
                    } else {
                        cs = s;
                        cm = m;
                        cpc = Pc;
                        state = kBwCollision;
                    }
                }
                if (scatter) pend_s = g;
            }
        }
        if (ended) {
            if (path) state = kBwPathEnd;
            else shadow_end();
        }
    }

    // A real collision of the path: pass 1 (vr_path_integrator_NEE_DS.h:40-72)
    // or pass 2 (:140-200; the albedo gradient is sent before the shadow ray,
    // it uses no random number), then the light sample and the shadow ray.
    VR_DEV void collision() {
        const MediumView& M = tr.med();
        const float s = cs, m = cm, Pc = cpc;
        const f3 sa = tr.albedo(rx);
        if (pass == 1) {
            scattered = 1;
            if (s < EPS) {
                if (null_W < EPS) { W = F3(0.0f, 0.0f, 0.0f); state = kBwPathEnd; return; }
                null_W = 0.0f;
            } else W = W * ((s / m) / Pc);
            added = (W * (F3(1.0f, 1.0f, 1.0f) - sa)) * M.emission_c;
        } else {
            const float gs = s < EPS ? dot(dL_dR, R_sp * (1.0f / eps)) : dot(dL_dR, R * (1.0f / s));
            pend_cs = gs;  // scattered at the collision point px by flush_collision
            px = rx;
            W = W * (s / m / Pc);
            if (s < EPS) { state = kBwDone; return; }
            f3 dL_dsa = (dL_dR * W) * (M.emission_c * -1.0f);
            R = R - (W * (F3(1.0f, 1.0f, 1.0f) - sa)) * M.emission_c;
            dL_dsa = dL_dsa + dL_dR * F3(R.x / fmaxf(EPS, sa.x), R.y / fmaxf(EPS, sa.y), R.z / fmaxf(EPS, sa.z));
            pend_ca = dL_dsa;
        }
        W = W * sa;
        f3 wnee;
        env = ENVMAP ? env_sample(M, tr.rng, wnee) : env_sample_const(M, tr.rng, wnee);
        before_T = tr.rng;
        px = rx;
        pw = rw;
        rw = wnee;
        after = pass == 1 ? kBwAfterT1 : kBwAfterT2;
        tn_draws = 0;
        if (pass == 2 && tn_rec && bounces < kTnRecords && M.bw_mode == VR_BW_CONSISTENT) {
            const uint16_t r = tn_rec[bounces * (int)blockDim.x];
            if (r != kTnRedo) {
                // pass 1's estimate at this collision; the stream moves past its walk
                Tn = (r & 0x8000u) ? 1.0f : 0.0f;
                for (int k = (int)(r & 0x7FFFu); k > 0; k--) rng_advance(tr.rng);
                state = kBwAfterT2;
                return;
            }
        }
        pend(kStShadowFw);
    }

    // The shadow ray's estimate Tn is known: pass 1 adds the NEE term and
    // samples the next direction (:53-72); pass 2 samples the next direction,
    // subtracts the term and starts the shadow ray's backward (:170-196).
    VR_DEV void after_t() {
        const MediumView& M = tr.med();
        const float phn = phase_eval(M, px, pw, rw);
        bool more = true;
        if (pass == 1) {
            if (tn_rec && bounces < kTnRecords)
                tn_rec[bounces * (int)blockDim.x] =
                    tn_draws <= 0x7FFE ? (uint16_t)((Tn != 0.0f ? 0x8000u : 0u) | (uint32_t)tn_draws) : kTnRedo;
            added = added + (W * env) * (Tn * phn);
            if (null_W < EPS) R_sp = R_sp + added;
            else R = R + added;
            bounces++;
            if (M.max_bounces >= 0 && bounces >= M.max_bounces) {
                W = F3(0.0f, 0.0f, 0.0f);
                state = kBwPathEnd;
                more = false;
            }
        }
        if (!more) return;
        rho = phase_sample(M, tr.rng, px, pw, wo);
        if (pass == 1) {
            W = W * rho;
            if (W.x < EPS && W.y < EPS && W.z < EPS) { W = F3(0.0f, 0.0f, 0.0f); state = kBwPathEnd; return; }
            continue_path();
        } else {
            before_next = tr.rng;
            dL_dTn = dot(dL_dR, (W * env) * phn);
            R = R - (W * env) * (Tn * phn);
            tr.rng = before_T;
            if (dL_dTn == 0.0f) state = kBwAfterTbw;
            else {
                rx = px;
                pend(kStShadowBw);
            }
        }
    }

    // The next path segment from the collision px along wo.
    VR_DEV void continue_path() {
        rw = wo;
        rx = px;
        pend(kStContinue);
    }

    VR_DEV void after_tbw() {
        const MediumView& M = tr.med();
        tr.rng = before_next;
        bounces++;
        if (M.max_bounces >= 0 && bounces >= M.max_bounces) { state = kBwDone; return; }
        W = W * rho;
        if (W.x < EPS && W.y < EPS && W.z < EPS) { state = kBwDone; return; }
        continue_path();
    }

    VR_DEV void path_end() {
        if (pass == 1) {
            if (!scattered) R = R + (ENVMAP ? env_eval(tr.med(), rw) : tr.med().env_c) * W;
            begin_pass(2);
        } else state = kBwDone;
    }

    VR_DEV bool tracking() const { return state <= kBwShadowReplay; }

    // The events a lane is waiting on, in pipeline order, so that one pass can
    // take a lane through several of them (a collision starts its shadow ray,
    // a shadow ray that misses the medium goes on to its continuation, ...).
    VR_DEV void events() {
        if (state == kBwCollision) collision();
        if (state == kBwAfterT1 || state == kBwAfterT2) after_t();
        if (state == kBwAfterTbw) after_tbw();
        if (state == kBwPathEnd) path_end();
        if (state == kBwStart) start();
    }
};

}  // namespace vr
