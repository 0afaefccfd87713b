// vr_ss_bw_flat.hpp -- the collision-integrator backward with single-scatter
// exitance and RT shadow rays as a flat lane state machine.
//
// Tracer::mc_collision_bw (include/maps/collision_integrator_mc.h:30-44, the
// consistent form) walks five rays one after the other per sample: the
// defensive DT collision sampler's primal pass (vr_collision_sampler_DT.h:52-90),
// the exitance's RT shadow ray (single_scattered_radiance_mc.h:6-16 with
// vr_transmittance_RT.h:1-49), the sampler's replay (:92-134), and the RT
// backward's primal pass and replay (vr_transmittance_RT.h:51-141) of the
// exitance backward's shadow ray.  On a 64-wide wave each of those loops runs
// as long as its longest lane.  Here every iteration takes a lane through the
// walk ends it waits on and then up to 3 tracking steps of whichever of the five
// rays it is on, all kinds through one code path (one majorant query, flight
// and trilinear lookup) with per-lane selects for their majorants, end tests
// and decisions.  The random draws of every sample are those of
// Tracer::mc_collision_bw, in the same order, and so is every gradient.
//
// Scope (the host dispatches everything else to the nested kernel): the
// consistent backward, RT transmittance, per-sample streams (the replayed or
// fresh dispatch seeds), no albedo gradient.
#pragma once

#include "vr_device.hpp"

namespace vr {

enum : int {
    kSsPrimal = 0,     // a step of the collision sampler's primal pass
    kSsReplay = 1,     // ... of its replay (scatters at every event)
    kSsTfw = 2,        // ... of the exitance's RT shadow ray (t_rt)
    kSsTbwPrimal = 3,  // ... of the shadow-ray backward's primal pass
    kSsTbwReplay = 4,  // ... of its replay (scatters at every null collision)
    kSsEndPrimal = 5,  // the primal pass ended (a collision or the box exit)
    kSsEndTfw = 6,     // the exitance's shadow ray ended (T known)
    kSsEndReplay = 7,  // the replay ended
    kSsEndTbw = 8,     // the shadow-ray backward's primal pass ended unabsorbed
    kSsStart = 9,      // a ray waits for its set-up (start_kind)
    kSsDone = 10       // the sample ended
};

// Ray set-ups of the shared start block.
enum : int {
    kSsStCamera = 0,  // the camera ray: box interval, the primal pass
    kSsStTfw = 1,     // the exitance's shadow ray from the collision
    kSsStReplay = 2,  // the sampler's replay over the camera ray's interval
    kSsStTbw = 3,     // the exitance backward's shadow ray
    kSsStTbwR = 4     // its replay from the saved start
};

template <int MAJ>
struct SsBwLane {
    static constexpr float kEpsRt = 0.0000001f;  // vr_transmittance_RT.h

    Tracer<MAJ> tr;
    Walk<MAJ> walk;
    f3 x0, w0;          // the sample's camera ray
    float tmin0, d0;    // its box interval
    Rng saved, after;   // the stream at the primal pass and after the exitance estimate
    f3 rx, rw;          // current ray
    float rd, rt;       // DS walks: the interval length and the distance walked; RT walks: the remaining distance
    float T;            // RT estimate of the current shadow ray
    float C, Wpre, s_t; // the primal pass's collision, null-collision weights and collision ratio
    f3 xt;              // collision point
    f3 env, dL_dR;
    f3 sa;
    float ph;
    float dL_dC, dL_dTc, out_Tc, s_tc;  // the replay's gradients and primal figures
    float out_T, dL_dT;                  // the shadow-ray backward's estimate and gradient
    f3 xs0;
    float ds0;
    Rng saved2;
    float eps;
    int state, start_kind;
    float pend_s;  // a step's sigma gradient at rx (0: none), issued by the wave

    VR_DEV explicit SsBwLane(ConstPtr<MediumView> m) : tr(m), pend_s(0.0f) {}

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
    VR_DEV void flush_sigma(float* srec, int n) {
        if (n) scatter_issue<1>(tr.med().sigma, srec, n);
    }

    // mc_collision_bw from its camera ray (x0, w0 set, the stream after camera_ray)
    VR_DEV void begin_sample() {
        rx = x0;
        rw = w0;
        pend(kSsStCamera);
    }

    VR_DEV void pend(int kind) {
        start_kind = kind;
        state = kSsStart;
    }

    // The shared ray set-up: the box interval of (rx, rw) where the walk
    // starts at the box, and the majorant walk.
    VR_DEV void start() {
        const MediumView& M = tr.med();
        float tmin = 0.0f, tmax = 0.0f;
        const bool hit = boundary(M, rx, rw, tmin, tmax);
        int next;
        if (start_kind == kSsStCamera) {
            // mc_collision_bw: no box hit, nothing to do
            next = hit ? kSsPrimal : kSsDone;
            tmin0 = tmin;
            d0 = tmax - tmin;
            saved = tr.rng;
            C = 0.0f; Wpre = 1.0f; s_t = 0.0f;
            rx = rx + rw * tmin;
            rd = d0;
            rt = 0.0f;
        } else if (start_kind == kSsStReplay) {
            next = kSsReplay;
            rx = x0 + w0 * tmin0;
            rd = d0;
            rt = 0.0f;
        } else if (start_kind == kSsStTfw) {
            // t_rt: a shadow ray that misses the box has T = 1 (no collision:
            // no shadow ray, the replay follows)
            T = 1.0f;
            next = hit && C != 0.0f ? kSsTfw : kSsEndTfw;
            rx = rx + rw * tmin;
            rd = tmax - tmin;
        } else if (start_kind == kSsStTbw) {
            // t_rt_bw: no box hit, no gradient
            next = hit ? kSsTbwPrimal : kSsDone;
            rx = rx + rw * tmin;
            rd = tmax - tmin;
            xs0 = rx;
            ds0 = rd;
            saved2 = tr.rng;
            T = 1.0f;
        } else {
            next = kSsTbwReplay;
            rx = xs0;
            rd = ds0;
            T = 1.0f;
            tr.rng = saved2;
        }
        state = next;
        if (state <= kSsTbwReplay) walk.begin(M, rx, rw);
    }

    // One tracking step of the current ray.  The collision sampler's passes
    // measure the distance walked against the interval (rt against rd), the
    // RT walks decrement the remaining distance; the flight is clamped by rd
    // in both, as in the reference.  Every lane's draws and arithmetic are
    // those of its nested loop.
    VR_DEV void step() {
        const MediumView& M = tr.med();
        const bool ds = state <= kSsReplay;
        float md;
        const float q = walk.query(M, md);
        const float m = ds ? fmaxf(0.0000000001f, q + 2.0f * eps)
                           : fmaxf(kEpsRt, state == kSsTfw ? q : q + 0.1f);
        const float fl = tr.flight(m);
        const float dt = fminf(fl, fminf(rd, md));
        const bool cross = ds ? dt >= md : (Walk<MAJ>::grid && md < rd && fl > md);
        const float adv = cross ? md : dt;
        rx = rx + rw * adv;
        if (ds) rt += adv;
        else rd -= adv;
        bool ended;
        if (cross) {
            walk.cross(M, rw);
            ended = ds ? rt >= rd - kCrossEps : rd <= kCrossEps;
        } else {
            walk.move(adv);
            ended = ds ? rt >= rd - 0.0000001f : rd <= kEpsRt;
        }
        if (ended) {
            state = state == kSsPrimal ? kSsEndPrimal
                  : state == kSsReplay ? kSsEndReplay
                  : state == kSsTfw ? kSsEndTfw
                  : state == kSsTbwPrimal ? kSsEndTbw : kSsDone;
            return;
        }
        if (cross) return;
        const float s = tr.sigma(rx);
        if (ds) {
            const float Pc = (s + eps) / m;
            const float u = tr.rnd();
            if (state == kSsPrimal) {
                if (u < Pc) {
                    C = 1.0f;
                    s_t = (s / m) / Pc;
                    state = kSsEndPrimal;
                } else {
                    Wpre *= (1.0f - s / m) / (1.0f - Pc);
                }
            } else if (u < Pc) {
                float trc = C / Pc;
                trc *= Wpre;
                pend_s = dL_dC * trc * (1.0f / m);
                state = kSsEndReplay;
            } else {
                const float qn = (dL_dTc * out_Tc + dL_dC * C * s_tc) / ((1.0f - s / m) / (1.0f - Pc)) / (1.0f - Pc);
                pend_s = qn * (-1.0f / m);
            }
        } else if (state == kSsTfw) {
            T *= (1.0f - s * rcp(m));
            if (T < 0.01f) {
                if (tr.rnd() < 1.0f - T) { T = 0.0f; state = kSsEndTfw; }
                else T = 1.0f;
            }
        } else if (state == kSsTbwPrimal) {
            T *= (1.0f - s / m);
            if (T < 0.1f) {
                if (tr.rnd() < 1.0f - T) state = kSsDone;  // absorbed: no gradient
                else T = 1.0f;
            }
        } else {
            const float r = 1.0f - s / m;
            T *= r;
            pend_s = -out_T * dL_dT / m / r;
            if (T < 0.1f) {
                rng_advance(tr.rng);
                T = 1.0f;
            }
        }
    }

    // The primal pass ended: with a collision, the exitance's light sample and
    // its shadow ray (ss_radiance at the collision); without, the replay (the
    // set-up block sends such a lane on with T unused).  The fields are written
    // on both paths: stores to different fields on the two sides of a branch
    // are merged into one store through a selected address, which keeps the
    // lane's state in scratch memory.
    VR_DEV void end_primal() {
        const MediumView& M = tr.med();
        xt = x0 + w0 * (tmin0 + rt);
        f3 e = F3(0.0f, 0.0f, 0.0f), a = e, wl = w0;
        float p = 0.0f;
        if (C != 0.0f) {
            e = env_sample(M, tr.rng, wl);
            p = phase_eval(M, xt, (w0 * -1.0f) * -1.0f, wl);
            a = tr.albedo(xt);
        }
        env = e;
        ph = p;
        sa = a;
        rx = xt;
        rw = wl;
        pend(kSsStTfw);
    }

    VR_DEV void end_tfw() { to_replay(C != 0.0f ? (env * sa) * (ph * T) : F3(0.0f, 0.0f, 0.0f)); }

    // The sampler's replay from the primal pass's stream, with the exitance eR.
    VR_DEV void to_replay(f3 eR) {
        const MediumView& M = tr.med();
        after = tr.rng;
        tr.rng = saved;
        dL_dC = dot(dL_dR, eR);
        dL_dTc = dot(dL_dR, env_eval(M, w0));
        out_Tc = C != 0.0f ? 0.0f : Wpre;
        s_tc = s_t * Wpre;
        rw = w0;
        pend(kSsStReplay);
    }

    // The replay ended: the exitance backward at the collision
    // (ss_radiance_bw with the path weight Wc), whose only gradient is its
    // shadow ray's RT backward.
    VR_DEV void end_replay() {
        const MediumView& M = tr.med();
        tr.rng = after;
        if (C == 0.0f) { state = kSsDone; return; }
        const float Wc = Wpre * s_t;
        const f3 g = dL_dR * Wc;
        f3 w2;
        const f3 e2 = env_sample(M, tr.rng, w2);
        const float ph2 = phase_eval(M, xt, (w0 * -1.0f) * -1.0f, w2);
        const f3 sa2 = tr.albedo(xt);
        dL_dT = dot(g, (e2 * sa2) * ph2);
        if (dL_dT == 0.0f) { state = kSsDone; return; }
        rx = xt;
        rw = w2;
        pend(kSsStTbw);
    }

    VR_DEV void end_tbw() {
        out_T = T;
        pend(kSsStTbwR);
    }

    VR_DEV bool tracking() const { return state <= kSsTbwReplay; }

    // The walk ends a lane waits on, in pipeline order.
    VR_DEV void events() {
        if (state == kSsEndPrimal) end_primal();
        if (state == kSsEndTfw) end_tfw();
        if (state == kSsEndReplay) end_replay();
        if (state == kSsEndTbw) end_tbw();
        if (state == kSsStart) start();
    }
};

}  // namespace vr
