// brt_opensteer.hip -- OpenSteerWrapper + Body on gfx950: one fused pass per step.
//
// The reference runs, per step, resetSteerForce + one kernel per behaviour
// (each re-reading the neighbour list and re-gathering neighbour state and
// read-modify-writing a double-buffered float4 force) + the integrator
// (PlugIns/OpenSteerWrapper/include/OpenSteerWrapperPlugIn.h:287-424,
//  OpenSteerWrapper_kernel.cu:62-333).  Here one lane per sorted agent walks its
// list once and keeps the three behaviour sums in registers.  The sums are still
// formed in list order and combined into the force in the reference's order
//   force = (((0 + sep*w_s) + align*w_a) + coh*w_c) [+ seek*w_t] [+ wander*w_w] [+ worldSeek]
// so the result equals the unfused sequence bit for bit.  The new state is
// written by agent ID (newX[hash[i].y] = ..., OpenSteerWrapper_kernel.cu:324-331).
//
// BehaviorClustering feature update (build-defined, no GPU counterpart in the
// reference): f'[id(i)] = f[id(i)] + rate * (mean_j f[id(j)] - f[id(i)]), the list
// walked in order with the sorted->ID map hash[j].y as in
// BehaviorClustering_computeNeighNaiveSimilarities (BehaviorClustering_kernel.cu:306-353).
#include "brt_device.hpp"
#include "brt_internal.hpp"

namespace brt {

namespace {

struct BehaviourTest {
    float minR;
    float r[3];    // sep, align, coh radius
    float cos_[3]; // sep, align, coh cosMaxAngle
};

// inBoidNeighborhood (OpenSteerWrapper_resources.cu:67-92) given the shared
// dist and cosAngle of one (agent, neighbour) pair.
__device__ __forceinline__ bool inBoid(float dist, float cosAngle, float minR, float maxR, float cosMax) {
    if (dist < minR) return true;
    if (dist > maxR) return false;
    return cosAngle > cosMax;
}

__global__ void __launch_bounds__(256) k_steer(brt_grid_params g, uint32_t W, brt_steer_params sp, float dt,
                                               uint32_t step, uint32_t n, const float4* __restrict__ spos,
                                               const float4* __restrict__ sfwd, const float4* __restrict__ ssa,
                                               const uint2* __restrict__ hash,
                                               const uint32_t* __restrict__ lists, float2* __restrict__ wander,
                                               float4* __restrict__ pos_out, float4* __restrict__ fwd_out,
                                               float4* __restrict__ sa_out, float4* __restrict__ force_out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 P = spos[i];
    const float4 Fw = sfwd[i];
    const float4 SA = ssa[i];
    const uint32_t id = hash[i].y;
    if (id >= n) return;
    const V3 myPos = xyz(P);
    const V3 myFwd = xyz(Fw);
    float mySpeed = Fw.w;
    V3 mySa = xyz(SA);

    const uint4* L = reinterpret_cast<const uint4*>(lists + (size_t)i * W);
    uint4 word = L[0];
    const uint32_t count = word.x < W - 1u ? word.x : W - 1u;

    const bool doSep = sp.separation[0] != 0.0f;
    const bool doAli = sp.alignment[0] != 0.0f;
    const bool doCoh = sp.cohesion[0] != 0.0f;
    const float minR = sp.sensing_min_radius;

    V3 sSep = v3(0.f, 0.f, 0.f), sAli = v3(0.f, 0.f, 0.f), sCoh = v3(0.f, 0.f, 0.f);
    uint32_t nSep = 0, nAli = 0, nCoh = 0;
    if (doSep || doAli || doCoh) {
        for (uint32_t t = 0; t < count; ++t) {
            const uint32_t e = t + 1;
            if ((e & 3u) == 0u) word = L[e >> 2];
            const uint32_t sub = e & 3u;
            const uint32_t j = sub == 0 ? word.x : (sub == 1 ? word.y : (sub == 2 ? word.z : word.w));
            if (j >= n) continue;
            const float4 Q = spos[j];
            const V3 other = xyz(Q);
            const V3 offset = other - myPos;
            const float d2 = dot(offset, offset);
            const float dist = sqrtf(d2);
            const V3 nrm = offset * (1.0f / sqrtf(d2));
            const float cosAngle = dot(myFwd, nrm);
            if (doSep && inBoid(dist, cosAngle, minR, sp.separation[1], sp.separation[2])) {
                sSep = sSep + offset / (-d2);
                ++nSep;
            }
            if (doAli && inBoid(dist, cosAngle, minR, sp.alignment[1], sp.alignment[2])) {
                sAli = sAli + xyz(sfwd[j]);
                ++nAli;
            }
            if (doCoh && inBoid(dist, cosAngle, minR, sp.cohesion[1], sp.cohesion[2])) {
                sCoh = sCoh + other;
                ++nCoh;
            }
        }
    }

    V3 force = v3(0.f, 0.f, 0.f);
    if (doSep) {
        V3 s = nSep > 0 ? normalize(sSep / (float)nSep) : sSep;
        force = s * sp.separation[0] + force;
    }
    if (doAli) {
        V3 s = nAli > 0 ? normalize(sAli / (float)nAli - myFwd) : sAli;
        force = s * sp.alignment[0] + force;
    }
    if (doCoh) {
        V3 s = nCoh > 0 ? normalize(sCoh / (float)nCoh - myPos) : sCoh;
        force = s * sp.cohesion[0] + force;
    }
    if (sp.num_targets > 0 && sp.target_weight != 0.0f) {
        const float* t = sp.targets[id % sp.num_targets];
        V3 s = steerForSeek(myPos, mySpeed, myFwd, v3(t[0], t[1], t[2]), sp.max_speed);
        force = s * sp.target_weight + force;
    }
    if (sp.wander_weight != 0.0f && wander) {
        uint32_t c[4] = {step, 0u, 0u, 0u};
        philox4x32_10(c, sp.wander_seed, id);
        const float u0 = u24(c[0]), u1 = u24(c[1]);
        const float walk = 12.0f * dt;
        float2 wsu = wander[id];
        float ws = wsu.x + ((u0 * 2.0f - 1.0f) * walk);
        ws = ws < -1.0f ? -1.0f : (ws > 1.0f ? 1.0f : ws);
        float wu = wsu.y + ((u1 * 2.0f - 1.0f) * walk);
        wu = wu < -1.0f ? -1.0f : (wu > 1.0f ? 1.0f : wu);
        wander[id] = make_float2(ws, wu);
        const V3 cr = cross(v3(0.f, 1.f, 0.f), myFwd);
        const V3 side = dot(cr, cr) == 0.0f ? v3(1.f, 0.f, 0.f) : normalize(cr);
        const V3 up = cross(myFwd, side);
        V3 s = side * ws + up * wu;
        force = s * sp.wander_weight + force;
    }
    if (sp.seek_world_center) {
        const float cx = g.world_center[0], cy = g.world_center[1], cz = g.world_center[2];
        const float rx = g.world_radius[0], ry = g.world_radius[1], rz = g.world_radius[2];
        const bool inside = (myPos.x > cx - rx) && (myPos.x < cx + rx) && (myPos.y > cy - ry) &&
                            (myPos.y < cy + ry) && (myPos.z > cz - rz) && (myPos.z < cz + rz);
        V3 s = inside ? v3(0.f, 0.f, 0.f) : steerForSeek(myPos, mySpeed, myFwd, v3(cx, cy, cz), sp.max_speed);
        force = s + force;
    }
    if (force_out) force_out[i] = make_float4(force.x, force.y, force.z, 0.0f);

    // applySteeringForceSingle (OpenSteerWrapper_resources.cu:326-379)
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

    pos_out[id] = make_float4(newPos.x, newPos.y, newPos.z, 1.0f);
    sa_out[id] = make_float4(mySa.x, mySa.y, mySa.z, SA.w);
    fwd_out[id] = make_float4(newFwd.x, newFwd.y, newFwd.z, mySpeed);
}

__global__ void __launch_bounds__(256) k_features(uint32_t W, float rate, uint32_t n, const uint2* __restrict__ hash,
                                                  const uint32_t* __restrict__ lists,
                                                  const float4* __restrict__ fin, float4* __restrict__ fout) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t id = hash[i].y;
    if (id >= n) return;
    const uint4* L = reinterpret_cast<const uint4*>(lists + (size_t)i * W);
    uint4 word = L[0];
    const uint32_t count = word.x < W - 1u ? word.x : W - 1u;
    const float4 f = fin[id];
    if (count == 0) {
        fout[id] = f;
        return;
    }
    float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
    for (uint32_t t = 0; t < count; ++t) {
        const uint32_t e = t + 1;
        if ((e & 3u) == 0u) word = L[e >> 2];
        const uint32_t sub = e & 3u;
        const uint32_t j = sub == 0 ? word.x : (sub == 1 ? word.y : (sub == 2 ? word.z : word.w));
        const uint32_t jid = j < n ? hash[j].y : n;
        if (jid >= n) continue;
        const float4 q = fin[jid];
        s.x = s.x + q.x;
        s.y = s.y + q.y;
        s.z = s.z + q.z;
        s.w = s.w + q.w;
    }
    const float c = (float)count;
    float4 o;
    o.x = f.x + (s.x / c - f.x) * rate;
    o.y = f.y + (s.y / c - f.y) * rate;
    o.z = f.z + (s.z / c - f.z) * rate;
    o.w = f.w + (s.w / c - f.w) * rate;
    fout[id] = o;
}

}  // namespace

hipError_t launch_steer(const brt_grid_params& g, uint32_t K, const brt_steer_params& sp, float dt, uint32_t step,
                        uint32_t n, const float4* spos, const float4* sfwd, const float4* ssa, const uint2* hash,
                        const uint32_t* lists, float2* wander, float4* pos_out, float4* fwd_out, float4* sa_out,
                        float4* force_out, hipStream_t s) {
    hipLaunchKernelGGL(k_steer, dim3((n + 255) / 256), dim3(256), 0, s, g, list_stride(K), sp, dt, step, n, spos, sfwd,
                       ssa, hash, lists, wander, pos_out, fwd_out, sa_out, force_out);
    return hipGetLastError();
}

hipError_t launch_features(uint32_t K, float rate, uint32_t n, const uint2* hash, const uint32_t* lists,
                           const float4* fin, float4* fout, hipStream_t s) {
    hipLaunchKernelGGL(k_features, dim3((n + 255) / 256), dim3(256), 0, s, list_stride(K), rate, n, hash, lists, fin,
                       fout);
    return hipGetLastError();
}

}  // namespace brt
