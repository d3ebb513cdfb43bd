// brt_steer_common.hpp -- per-neighbour behaviour sums and the per-agent
// force assembly + integration, shared by the standalone steering kernel
// (k_steer) and the fused neighbour+steering kernel, so both produce the same
// bits.
//
// Reference: inBoidNeighborhood, calcSeparation/Cohesion/Alignment,
// xxxsteerForSeek, applySteeringForceSingle
// (PlugIns/OpenSteerWrapper/OpenSteerWrapper_resources.cu:67-379), the steering
// kernels and integrator (OpenSteerWrapper_kernel.cu:62-333), the force
// accumulation order of BehaveRT_Flocking::update
// (Samples/OpensteerDemo/plugins/BehaveRT_Flocking.cpp:183-196).
//
// Cohesion is summed as offsets p_j - p_i (normalize(mean offset)) rather than
// the reference's normalize(sum p_j / n - p_i): equal in real arithmetic, and
// free of the FP32 cancellation of a centroid minus a nearby large position.
#pragma once

#include "brt_device.hpp"

namespace brt {

// inBoidNeighborhood given the (agent, neighbour) pair's dist and cosAngle.
__device__ __forceinline__ bool inBoid(float dist, float cosAngle, float minR, float maxR, float cosMax) {
    if (dist < minR) return true;
    if (dist > maxR) return false;
    return cosAngle > cosMax;
}

// Behaviour weights (separation, alignment, cohesion): the steer params' own,
// or a flock's in a batch.
struct Weights {
    float s, a, c;
};
__device__ __forceinline__ Weights weights_of(const brt_steer_params& sp) {
    return Weights{sp.separation[0], sp.alignment[0], sp.cohesion[0]};
}

// The radius tests of inBoidNeighborhood as exact cuts on the squared
// distance (steer_cuts, host): sqrtf(d2) < minR <=> d2 < tmin and
// sqrtf(d2) > maxR_b <=> d2 > tmax[b] for every float d2, because sqrtf is
// correctly rounded and monotone; cmax[b] = cosMax of behaviour b (separation,
// alignment, cohesion); same: the three share maxR and cosMax.
struct SteerCuts {
    float tmin;
    float tmax[3];
    float cmax[3];
    uint32_t same;
};

// inBoidNeighborhood of one behaviour, decided without a square root when the
// answer is certain: the radius tests by the cuts, the angle test
// cosAngle > cosMax by ca = dot(fwd, off) * rsq(d2).  The forward's cosAngle =
// dot(fwd, off * (1/sqrtf(d2))) and ca both lie within about 10 roundings
// (2^-24 relative) of dot(fwd, off)/sqrt(d2), i.e. within 2e-6 |fwd| of each
// other; the caller passes margin = 1e-5 |fwd|_1 and fast = d2 normal-range and
// ca finite.  1: in, 0: out, 2: undecided (evaluate the forward's expression).
// (selects, no branches)
__device__ __forceinline__ uint32_t cut_test(float d2, float tmin, float tmax, float ca, float cmax, float margin,
                                             bool fast) {
    const float gap = ca - cmax;
    const bool sure = fast & (fabsf(gap) > margin);
    const uint32_t ang = sure ? (gap > 0.0f ? 1u : 0u) : 2u;
    return d2 < tmin ? 1u : (d2 > tmax ? 0u : ang);
}

struct SteerAccum {
    V3 sSep, sAli, sCoh;
    uint32_t nSep, nAli, nCoh;

    __device__ __forceinline__ void init() {
        sSep = v3(0.f, 0.f, 0.f);
        sAli = v3(0.f, 0.f, 0.f);
        sCoh = v3(0.f, 0.f, 0.f);
        nSep = nAli = nCoh = 0;
    }

    // offset = p_j - p_i, d2 = dot(offset, offset) as the list test computed it
    __device__ __forceinline__ void add(const brt_steer_params& sp, Weights w, V3 myFwd, V3 offset, float d2,
                                        V3 fwdj) {
        const float dist = sqrtf(d2);
        const V3 nrm = offset * (1.0f / dist);
        const float cosAngle = dot(myFwd, nrm);
        const float minR = sp.sensing_min_radius;
        if (w.s != 0.0f && inBoid(dist, cosAngle, minR, sp.separation[1], sp.separation[2])) {
            sSep = sSep + offset / (-d2);
            ++nSep;
        }
        if (w.a != 0.0f && inBoid(dist, cosAngle, minR, sp.alignment[1], sp.alignment[2])) {
            sAli = sAli + fwdj;
            ++nAli;
        }
        if (w.c != 0.0f && inBoid(dist, cosAngle, minR, sp.cohesion[1], sp.cohesion[2])) {
            sCoh = sCoh + offset;
            ++nCoh;
        }
    }

    // add() with the neighbourhood tests decided by the cuts (same decisions,
    // same sums); margin = 1e-5 (|fwd.x| + |fwd.y| + |fwd.z|) of the agent.
    __device__ __forceinline__ void add_cut(const brt_steer_params& sp, const SteerCuts& c, Weights w, V3 myFwd,
                                            float margin, V3 offset, float d2, V3 fwdj) {
        const float ca = dot(myFwd, offset) * __builtin_amdgcn_rsqf(d2);
        const bool fast = (d2 >= 0x1p-100f) & (d2 <= 0x1p+100f) & __builtin_isfinite(ca);
        uint32_t ts = cut_test(d2, c.tmin, c.tmax[0], ca, c.cmax[0], margin, fast), ta = ts, tc = ts;
        if (!c.same) {
            ta = cut_test(d2, c.tmin, c.tmax[1], ca, c.cmax[1], margin, fast);
            tc = cut_test(d2, c.tmin, c.tmax[2], ca, c.cmax[2], margin, fast);
        }
        if ((ts | ta | tc) & 2u) {
            const float dist = sqrtf(d2);
            const V3 nrm = offset * (1.0f / dist);
            const float cosAngle = dot(myFwd, nrm);
            const float minR = sp.sensing_min_radius;
            if (ts == 2u) ts = inBoid(dist, cosAngle, minR, sp.separation[1], sp.separation[2]) ? 1u : 0u;
            if (ta == 2u) ta = inBoid(dist, cosAngle, minR, sp.alignment[1], sp.alignment[2]) ? 1u : 0u;
            if (tc == 2u) tc = inBoid(dist, cosAngle, minR, sp.cohesion[1], sp.cohesion[2]) ? 1u : 0u;
        }
        if (w.s != 0.0f && ts) {
            sSep = sSep + offset / (-d2);
            ++nSep;
        }
        if (w.a != 0.0f && ta) {
            sAli = sAli + fwdj;
            ++nAli;
        }
        if (w.c != 0.0f && tc) {
            sCoh = sCoh + offset;
            ++nCoh;
        }
    }
};

__device__ __forceinline__ float cut_margin(V3 fwd) { return 1e-5f * ((fabsf(fwd.x) + fabsf(fwd.y)) + fabsf(fwd.z)); }

// host: the cuts of the steer params
SteerCuts steer_cuts(const brt_steer_params& sp);

struct AgentOut {
    float4* pos_out;
    float4* fwd_out;
    float4* sa_out;
    float4* force_out;  // sorted, optional
    float2* wander;     // by ID, in/out, optional
};

// The behaviour forces in the reference's accumulation order (each term added
// as term + force, the blendIntoSteeringForce order, OpenSteerWrapper_resources.cu:113-129):
// separation, alignment, cohesion, then the build's target seek and wander.
// `id` indexes the state arrays; `lid` (the ID within the agent's flock, equal
// to id outside batches) picks the seek target and keys the wander stream.
__device__ __forceinline__ V3 behaviour_force(const brt_steer_params& sp, Weights w, float dt, uint32_t step,
                                              uint32_t id, uint32_t lid, V3 myPos, V3 myFwd, float mySpeed,
                                              const SteerAccum& acc, float2* wander) {
    V3 force = v3(0.f, 0.f, 0.f);
    if (w.s != 0.0f) {
        V3 s = acc.nSep > 0 ? normalize(acc.sSep / (float)acc.nSep) : acc.sSep;
        force = s * w.s + force;
    }
    if (w.a != 0.0f) {
        V3 s = acc.nAli > 0 ? normalize(acc.sAli / (float)acc.nAli - myFwd) : acc.sAli;
        force = s * w.a + force;
    }
    if (w.c != 0.0f) {
        V3 s = acc.nCoh > 0 ? normalize(acc.sCoh / (float)acc.nCoh) : acc.sCoh;
        force = s * w.c + force;
    }
    if (sp.num_targets > 0 && sp.target_weight != 0.0f) {
        const float* t = sp.targets[lid % sp.num_targets];
        V3 s = steerForSeek(myPos, mySpeed, myFwd, v3(t[0], t[1], t[2]), sp.max_speed);
        force = s * sp.target_weight + force;
    }
    if (sp.wander_weight != 0.0f && wander) {
        uint32_t c[4] = {step, 0u, 0u, 0u};
        philox4x32_10(c, sp.wander_seed, lid);
        const float u0 = u24(c[0]), u1 = u24(c[1]);
        const float walk = 12.0f * dt;
        const float2 wsu = wander[id];
        float ws = wsu.x + ((u0 * 2.0f - 1.0f) * walk);
        ws = ws < -1.0f ? -1.0f : (ws > 1.0f ? 1.0f : ws);
        float wu = wsu.y + ((u1 * 2.0f - 1.0f) * walk);
        wu = wu < -1.0f ? -1.0f : (wu > 1.0f ? 1.0f : wu);
        wander[id] = make_float2(ws, wu);
        // OpenSteer's right-handed local space (LocalSpace.h:277-305) with the
        // world up (0,1,0) as the previous up: side = normalize(forward x up),
        // up = side x forward.  A heading along the world up has no such side;
        // (1,0,0) stands in for it.
        const V3 cr = cross(myFwd, v3(0.f, 1.f, 0.f));
        const V3 side = dot(cr, cr) == 0.0f ? v3(1.f, 0.f, 0.f) : normalize(cr);
        const V3 up = cross(side, myFwd);
        V3 s = side * ws + up * wu;
        force = s * sp.wander_weight + force;
    }
    return force;
}

// computeSeekingsWorldCenter (added last), then applySteeringForceSingle, then
// the scatter of the new state to agent-ID order.  force_out (sorted, optional)
// receives the summed force before the integrator.
__device__ __forceinline__ void apply_force(const brt_grid_params& g, const brt_steer_params& sp, float dt,
                                            uint32_t i, uint32_t id, V3 force, V3 myPos, V3 myFwd, float mySpeed,
                                            V3 mySa, float saW, const AgentOut& o) {
    if (sp.seek_world_center) {
        const float cx = g.world_center[0], cy = g.world_center[1], cz = g.world_center[2];
        const float rx = g.world_radius[0], ry = g.world_radius[1], rz = g.world_radius[2];
        const bool inside = (myPos.x > cx - rx) && (myPos.x < cx + rx) && (myPos.y > cy - ry) &&
                            (myPos.y < cy + ry) && (myPos.z > cz - rz) && (myPos.z < cz + rz);
        V3 s = inside ? v3(0.f, 0.f, 0.f) : steerForSeek(myPos, mySpeed, myFwd, v3(cx, cy, cz), sp.max_speed);
        force = s + force;
    }
    if (o.force_out) o.force_out[i] = make_float4(force.x, force.y, force.z, 0.0f);

    if (sp.use_2d_projection) force.y = 0.0f;
    const V3 adjusted = adjustRawSteeringForce(force, mySpeed, myFwd, sp.max_speed);
    const V3 clipped = truncateLength(adjusted, sp.max_force);
    const V3 newAccel = clipped / sp.mass;
    if (dt > 0.0f) {
        float smoothRate = clampf(9.0f * dt, 0.15f, 0.4f);
        smoothRate = clampf(smoothRate, 0.0f, 1.0f);
        mySa = mySa + (newAccel - mySa) * smoothRate;
    }
    V3 newVel = myFwd * mySpeed + mySa * dt;
    newVel = truncateLength(newVel, sp.max_speed);
    mySpeed = length(newVel);
    V3 newPos = myPos + newVel * dt;
    if (sp.use_2d_projection) newVel.y = 0.0f;
    const V3 newFwd = normalize(newVel);
    if (sp.use_2d_projection) newPos.y = 0.5f;

    o.pos_out[id] = make_float4(newPos.x, newPos.y, newPos.z, 1.0f);
    o.sa_out[id] = make_float4(mySa.x, mySa.y, mySa.z, saW);
    o.fwd_out[id] = make_float4(newFwd.x, newFwd.y, newFwd.z, mySpeed);
}

// The whole per-agent tail of the steering pass.
__device__ __forceinline__ void finish_agent(const brt_grid_params& g, const brt_steer_params& sp, Weights w,
                                             float dt, uint32_t step, uint32_t i, uint32_t id, uint32_t lid,
                                             V3 myPos, V3 myFwd, float mySpeed, V3 mySa, float saW,
                                             const SteerAccum& acc, const AgentOut& o) {
    const V3 force = behaviour_force(sp, w, dt, step, id, lid, myPos, myFwd, mySpeed, acc, o.wander);
    apply_force(g, sp, dt, i, id, force, myPos, myFwd, mySpeed, mySa, saW, o);
}

}  // namespace brt
