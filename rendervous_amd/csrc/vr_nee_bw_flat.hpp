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
    kBwCollision = 9      // the path ray has a real collision (s, m, Pc saved)
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
    int state, pass, scattered, bounces;

    VR_DEV explicit NeeDsBwLane(const MediumView& m) : tr(m) {}

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
        float tmin, tmax;
        boundary(tr.M, rx, rw, tmin, tmax);
        rx = rx + rw * tmin;
        rd = tmax - tmin;
        bounces = 0;
        walk.begin(tr.M, rx, rw);
        state = kBwPath;
    }

    // Start of the shadow ray from the path vertex px along rw (t_dt, :616-635).
    VR_DEV void start_shadow_fw(int after) {
        rx = px;
        float tmin, tmax;
        if (!boundary(tr.M, rx, rw, tmin, tmax)) { Tn = 1.0f; state = after; return; }
        rx = rx + rw * tmin;
        rd = tmax - tmin;
        walk.begin(tr.M, rx, rw);
        state = kBwShadowFw;
    }

    // Start of the shadow ray's DT backward (t_dt_bw), with tr.rng = before_T.
    VR_DEV void start_shadow_bw() {
        if (dL_dTn == 0.0f) { state = kBwAfterTbw; return; }
        rx = px;
        float tmin, tmax;
        if (!boundary(tr.M, rx, rw, tmin, tmax)) { state = kBwAfterTbw; return; }
        rx = rx + rw * tmin;
        rd = tmax - tmin;
        xs0 = rx;
        ds0 = rd;
        walk.begin(tr.M, rx, rw);
        state = kBwShadowPrimal;
    }

    // The next path segment from the collision px along wo.
    VR_DEV void continue_path() {
        rw = wo;
        rx = px;
        float tmin;
        boundary(tr.M, rx, rw, tmin, rd);
        walk.begin(tr.M, rx, rw);
        state = kBwPath;
    }

    // One tracking step of the current ray.
    VR_DEV void step() {
        const MediumView& M = tr.M;
        float md;
        const float q = walk.query(M, md);
        float m;
        float t = 0.0f;
        if (state == kBwPath) {
            if (!(rd > EPS)) { state = kBwPathEnd; return; }
            const bool ref = M.bw_mode == VR_BW_REFERENCE;
            m = (pass == 2 && ref) ? fmaxf(EPS, q * 1.2f + 2.0f * eps) : fmaxf(EPS, q + 2.0f * eps + 0.1f);
            t = tr.flight(m);
            rx = rx + rw * fminf(t, md);
            if (t > md) { rd -= md; walk.cross(M, rw); if (rd <= kCrossEps) state = kBwPathEnd; return; }
            if (t > rd) { state = kBwPathEnd; return; }
            walk.move(t);
        } else {
            m = fmaxf(0.0000001f, state == kBwShadowFw ? q : q + 0.1f);
            const float dt = tr.flight(m);
            const float tt = fminf(md, dt);
            rx = rx + rw * tt;
            rd -= tt;
            bool end = false;
            if (dt > md) {
                walk.cross(M, rw);
                if (rd > kCrossEps) return;
                end = true;
            } else {
                walk.move(tt);
                end = rd <= 0.0000001f;
            }
            if (end) {  // the shadow ray left the medium
                if (state == kBwShadowFw) { Tn = 1.0f; state = pass == 1 ? kBwAfterT1 : kBwAfterT2; }
                else if (state == kBwShadowPrimal) {  // not absorbed: replay from the start
                    rx = xs0;
                    rd = ds0;
                    tr.rng = before_T;
                    walk.begin(M, rx, rw);
                    state = kBwShadowReplay;
                } else state = kBwAfterTbw;
                return;
            }
        }
        // the shared trilinear lookup and the decision of each ray kind
        const float s = tr.sigma(rx);
        if (state == kBwShadowFw) {
            if (tr.rnd() * m < s) { Tn = 0.0f; state = pass == 1 ? kBwAfterT1 : kBwAfterT2; }
            return;
        }
        if (state == kBwShadowPrimal) {
            if (tr.rnd() < s / m) state = kBwAfterTbw;  // absorbed: no gradient
            return;
        }
        if (state == kBwShadowReplay) {
            tr.rnd();
            const float Pc = s / m;
            tr.sigma_bw(rx, -dL_dTn / m / (1.0f - Pc));
            return;
        }
        const float Pc = (s + eps) / m;
        if (tr.rnd() >= Pc) {  // null collision
            if (pass == 2) tr.sigma_bw(rx, dot(dL_dR, R * (-1.0f / m / (1.0f - s / m))));
            W = W * ((1.0f - s / m) / (1.0f - Pc));
            rd -= t;
            return;
        }
        cs = s;
        cm = m;
        cpc = Pc;
        state = kBwCollision;
    }

    VR_DEV void collision() {
        if (pass == 1) collision1(cs, cm, cpc);
        else collision2(cs, cm, cpc);
    }

    // Pass-1 real collision (vr_path_integrator_NEE_DS.h:40-72).
    VR_DEV void collision1(float s, float m, float Pc) {
        const MediumView& M = tr.M;
        scattered = 1;
        if (s < EPS) {
            if (null_W < EPS) { W = F3(0.0f, 0.0f, 0.0f); state = kBwPathEnd; return; }
            null_W = 0.0f;
        } else W = W * ((s / m) / Pc);
        const f3 sa = tr.albedo(rx);
        added = (W * (F3(1.0f, 1.0f, 1.0f) - sa)) * M.emission_c;
        W = W * sa;
        f3 wnee;
        env = ENVMAP ? env_sample(M, tr.rng, wnee) : env_sample_const(M, tr.rng, wnee);
        px = rx;
        pw = rw;
        rw = wnee;
        start_shadow_fw(kBwAfterT1);
    }

    VR_DEV void after_t1() {
        const MediumView& M = tr.M;
        const float phn = phase_eval(M, pw, rw);
        added = added + (W * env) * (Tn * phn);
        if (null_W < EPS) R_sp = R_sp + added;
        else R = R + added;
        bounces++;
        if (M.max_bounces >= 0 && bounces >= M.max_bounces) { W = F3(0.0f, 0.0f, 0.0f); state = kBwPathEnd; return; }
        rho = phase_sample(M, tr.rng, pw, wo);
        W = W * rho;
        if (W.x < EPS && W.y < EPS && W.z < EPS) { W = F3(0.0f, 0.0f, 0.0f); state = kBwPathEnd; return; }
        continue_path();
    }

    // Pass-2 real collision (vr_path_integrator_NEE_DS.h:140-200): the albedo
    // gradient is sent before the shadow ray (it uses no random number).
    VR_DEV void collision2(float s, float m, float Pc) {
        const MediumView& M = tr.M;
        const float gs = s < EPS ? dot(dL_dR, R_sp * (1.0f / eps)) : dot(dL_dR, R * (1.0f / s));
        tr.sigma_bw(rx, gs);
        W = W * (s / m / Pc);
        if (s < EPS) { state = kBwDone; return; }
        const f3 sa = tr.albedo(rx);
        f3 dL_dsa = (dL_dR * W) * (M.emission_c * -1.0f);
        R = R - (W * (F3(1.0f, 1.0f, 1.0f) - sa)) * M.emission_c;
        dL_dsa = dL_dsa + dL_dR * F3(R.x / fmaxf(EPS, sa.x), R.y / fmaxf(EPS, sa.y), R.z / fmaxf(EPS, sa.z));
        tr.albedo_bw(rx, dL_dsa);
        W = W * sa;
        f3 wnee;
        env = ENVMAP ? env_sample(M, tr.rng, wnee) : env_sample_const(M, tr.rng, wnee);
        before_T = tr.rng;
        px = rx;
        pw = rw;
        rw = wnee;
        start_shadow_fw(kBwAfterT2);
    }

    VR_DEV void after_t2() {
        const MediumView& M = tr.M;
        const float phn = phase_eval(M, pw, rw);
        rho = phase_sample(M, tr.rng, pw, wo);
        before_next = tr.rng;
        dL_dTn = dot(dL_dR, (W * env) * phn);
        R = R - (W * env) * (Tn * phn);
        tr.rng = before_T;
        start_shadow_bw();
    }

    VR_DEV void after_tbw() {
        const MediumView& M = tr.M;
        tr.rng = before_next;
        bounces++;
        if (M.max_bounces >= 0 && bounces >= M.max_bounces) { state = kBwDone; return; }
        W = W * rho;
        if (W.x < EPS && W.y < EPS && W.z < EPS) { state = kBwDone; return; }
        continue_path();
    }

    VR_DEV void path_end() {
        if (pass == 1) {
            if (!scattered) R = R + (ENVMAP ? env_eval(tr.M, rw) : tr.M.env_c) * W;
            begin_pass(2);
        } else state = kBwDone;
    }

    VR_DEV bool tracking() const { return state <= kBwShadowReplay; }

    // The events a lane is waiting on, in pipeline order, so that one pass can
    // take a lane through several of them (a collision starts its shadow ray,
    // a shadow ray that misses the medium goes on to its continuation, ...).
    VR_DEV void events() {
        if (state == kBwCollision) collision();
        if (state == kBwAfterT1) after_t1();
        if (state == kBwAfterT2) after_t2();
        if (state == kBwAfterTbw) after_tbw();
        if (state == kBwPathEnd) path_end();
    }
};

}  // namespace vr
