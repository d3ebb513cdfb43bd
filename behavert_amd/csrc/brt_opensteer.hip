// brt_opensteer.hip -- OpenSteerWrapper + Body on gfx950: one fused pass per step.
//
// The reference runs, per step, resetSteerForce + one kernel per behaviour
// (each re-reading the neighbour list and re-gathering neighbour state and
// read-modify-writing a double-buffered float4 force) + the integrator
// (PlugIns/OpenSteerWrapper/include/OpenSteerWrapperPlugIn.h:288-426,
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
#include "brt_lists.hpp"
#include "brt_steer_common.hpp"

namespace brt {

namespace {

// One lane per sorted agent: walk the list in order (8 entries per round, their
// gathers issued together), accumulate the behaviour sums, then assemble the
// force and integrate (APPLY), or only store the behaviour force in sorted
// order (the steer-force accumulator of the decomposed sequence).
// The feature update of the agent at sorted index i (ID id):
// f + (mean of the listed neighbours' features - f) * rate, the mean over the
// list's count; the entries are visited eight at a time with their ID and
// feature gathers in flight together and summed in list order.
__device__ __forceinline__ float4 feature_update(uint32_t W, float rate, uint32_t n, const uint2* __restrict__ hash,
                                                 const uint32_t* __restrict__ lists, const float4* __restrict__ fin,
                                                 uint32_t i, uint32_t id) {
    const float4 f = fin[id];
    float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
    const uint32_t count =
        for_list_octets(lists, W, W - 1u, i, n, [&](uint32_t, const uint32_t(&jj)[8], uint32_t mask) {
            uint32_t jid[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) jid[k] = hash[jj[k]].y;
            float4 q[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) q[k] = fin[jid[k] < n ? jid[k] : id];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                if (((mask >> k) & 1u) == 0u || jid[k] >= n) continue;
                s.x = s.x + q[k].x;
                s.y = s.y + q[k].y;
                s.z = s.z + q[k].z;
                s.w = s.w + q[k].w;
            }
        });
    if (count == 0) return f;
    const float inv = 1.0f / (float)count;  // cutil float4 / float
    float4 r;
    r.x = f.x + (s.x * inv - f.x) * rate;
    r.y = f.y + (s.y * inv - f.y) * rate;
    r.z = f.z + (s.z * inv - f.z) * rate;
    r.w = f.w + (s.w * inv - f.w) * rate;
    return r;
}

__global__ void __launch_bounds__(256) k_features(uint32_t W, float rate, uint32_t n, const uint2* __restrict__ hash,
                                                  const uint32_t* __restrict__ lists,
                                                  const float4* __restrict__ fin, float4* __restrict__ fout) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t id = hash[i].y;
    if (id >= n) return;
    fout[id] = feature_update(W, rate, n, hash, lists, fin, i, id);
}

// FEAT: the same pass also writes the agent's feature update (the result of
// feature_update), summing the neighbours' features in the same list walk as
// the steering sums; with all steering weights zero it calls feature_update.
// SFEAT: the neighbours' features are read from sfeat, the features in sorted
// order (sfeat[j] = fin[hash[j].y], gathered with the state), so the walk reads
// them next to their positions instead of at random agent IDs.
template <bool APPLY, bool FEAT = false, bool SFEAT = false>
__global__ void __launch_bounds__(256) k_steer(brt_grid_params g, uint32_t W, brt_steer_params sp, SteerCuts cuts,
                                               float dt,
                                               uint32_t step, uint32_t n, const float4* __restrict__ spos,
                                               const float4* __restrict__ sfwd, const float4* __restrict__ ssa,
                                               const uint2* __restrict__ hash,
                                               const uint32_t* __restrict__ lists, AgentOut o, uint32_t xcd_remap,
                                               float feature_rate = 0.0f, const float4* __restrict__ fin = nullptr,
                                               float4* __restrict__ fout = nullptr,
                                               const float4* __restrict__ sfeat = nullptr) {
    const uint32_t i = xcd_block(xcd_remap != 0) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t id = hash[i].y;
    if (id >= n) return;
    const float4 P = spos[i];
    const float4 Fw = sfwd[i];
    // no sorted copy of the smoothed acceleration: the agent's own, by ID
    const float4 SA = ssa ? ssa[i] : (APPLY ? o.sa_out[id] : make_float4(0.f, 0.f, 0.f, 0.f));
    const V3 myPos = xyz(P);
    const V3 myFwd = xyz(Fw);

    const Weights wts = weights_of(sp);
    const bool any = wts.s != 0.0f || wts.a != 0.0f || wts.c != 0.0f;
    const float margin = cut_margin(myFwd);

    SteerAccum acc;
    acc.init();
    // FEAT: the feature sum of the same entries, in the same order as
    // feature_update (their IDs load with the positions and headings)
    float4 fsum = make_float4(0.f, 0.f, 0.f, 0.f);
    uint32_t count = 0;
    if (any) {
        // the list in order, eight entries at a time with their gathers in flight;
        // the feature variants hold four more gathers per entry and read the
        // row in pairs (whole K 32 rows: 213 -> 162 us for the two-kernel
        // step's steering on 1M blobs, 185 -> 117 us for the behaviour forces)
        count = for_list_octets<FEAT ? kRowPairs : kRowWhole>(lists, W, W - 1u, i, n, [&](uint32_t, const uint32_t(&jj)[8], uint32_t mask) {
            float4 qp[8], qf[8], fq[8];
            uint32_t jid[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                qp[t] = spos[jj[t]];
                qf[t] = sfwd[jj[t]];
                if (FEAT && SFEAT) fq[t] = sfeat[jj[t]];
                if (FEAT && !SFEAT) jid[t] = hash[jj[t]].y;
            }
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                if ((mask >> t) & 1u) {
                    const V3 offset = xyz(qp[t]) - myPos;
                    acc.add_cut(sp, cuts, wts, myFwd, margin, offset, dot(offset, offset), xyz(qf[t]));
                }
            }
            if (FEAT) {
                if (!SFEAT) {
#pragma unroll
                    for (int t = 0; t < 8; ++t) fq[t] = fin[jid[t] < n ? jid[t] : id];
                }
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    if (((mask >> t) & 1u) == 0u || (!SFEAT && jid[t] >= n)) continue;
                    fsum.x = fsum.x + fq[t].x;
                    fsum.y = fsum.y + fq[t].y;
                    fsum.z = fsum.z + fq[t].z;
                    fsum.w = fsum.w + fq[t].w;
                }
            }
        });
    }
    if (FEAT) {
        if (!any) {
            fout[id] = feature_update(W, feature_rate, n, hash, lists, fin, i, id);
        } else {
            const float4 f = SFEAT ? sfeat[i] : fin[id];
            float4 r = f;
            if (count != 0) {
                const float inv = 1.0f / (float)count;  // cutil float4 / float
                r.x = f.x + (fsum.x * inv - f.x) * feature_rate;
                r.y = f.y + (fsum.y * inv - f.y) * feature_rate;
                r.z = f.z + (fsum.z * inv - f.z) * feature_rate;
                r.w = f.w + (fsum.w * inv - f.w) * feature_rate;
            }
            fout[id] = r;
        }
    }
    if (APPLY) {
        finish_agent(g, sp, wts, dt, step, i, id, id, myPos, myFwd, Fw.w, xyz(SA), SA.w, acc, o);
    } else {
        const V3 f = behaviour_force(sp, wts, dt, step, id, id, myPos, myFwd, Fw.w, acc, o.wander);
        o.force_out[i] = make_float4(f.x, f.y, f.z, 0.0f);
    }
}

// applySteeringForces from an accumulated force (sorted float4): the world-centre
// seek and the integrator of the fused pass, on a force built by other passes.
__global__ void __launch_bounds__(256) k_apply(brt_grid_params g, brt_steer_params sp, float dt, uint32_t n,
                                               const float4* __restrict__ spos, const float4* __restrict__ sfwd,
                                               const float4* __restrict__ ssa, const uint2* __restrict__ hash,
                                               const float4* __restrict__ force_in, AgentOut o) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t id = hash[i].y;
    if (id >= n) return;
    const float4 Fw = sfwd[i], SA = ssa[i];
    apply_force(g, sp, dt, i, id, xyz(force_in[i]), xyz(spos[i]), xyz(Fw), Fw.w, xyz(SA), SA.w, o);
}


}  // namespace

namespace {

// The smallest float t >= 0 with sqrtf(t) >= m, so that sqrtf(d2) < m <=> d2 < t
// for every float d2 (NaN included: both false).
float below_cut(float m) {
    if (m != m) return -INFINITY;
    if (m <= 0.0f) return 0.0f;
    if (m == INFINITY) return INFINITY;
    float t = m * m;
    if (t == INFINITY) t = 3.40282347e38f;
    while (t > 0.0f && sqrtf(nextafterf(t, -INFINITY)) >= m) t = nextafterf(t, -INFINITY);
    while (sqrtf(t) < m) t = nextafterf(t, INFINITY);
    return t;
}

}  // namespace

SteerCuts steer_cuts(const brt_steer_params& sp) {
    SteerCuts c{};
    c.tmin = below_cut(sp.sensing_min_radius);
    const float* b[3] = {sp.separation, sp.alignment, sp.cohesion};
    for (int k = 0; k < 3; ++k) {
        c.tmax[k] = brt_radius_threshold(b[k][1]);
        c.cmax[k] = b[k][2];
    }
    c.same = (c.tmax[0] == c.tmax[1] && c.tmax[0] == c.tmax[2] && c.cmax[0] == c.cmax[1] && c.cmax[0] == c.cmax[2])
                 ? 1u
                 : 0u;
    return c;
}

hipError_t launch_steer(const brt_grid_params& g, uint32_t K, const brt_steer_params& sp, float dt, uint32_t step,
                        uint32_t n, const float4* spos, const float4* sfwd, const float4* ssa, const uint2* hash,
                        const uint32_t* lists, float2* wander, float4* pos_out, float4* fwd_out, float4* sa_out,
                        float4* force_out, uint32_t xcd_remap, hipStream_t s, float feature_rate,
                        const float4* features_in, float4* features_out, const float4* sorted_features) {
    AgentOut o{pos_out, fwd_out, sa_out, force_out, wander};
    const SteerCuts cuts = steer_cuts(sp);
    if (features_in && sorted_features)
        BRT_LAUNCH((k_steer<true, true, true>), dim3((n + 255) / 256), dim3(256), 0, s, g, list_stride(K), sp,
                           cuts, dt, step, n, spos, sfwd, ssa, hash, lists, o, xcd_remap, feature_rate, features_in,
                           features_out, sorted_features);
    else if (features_in)
        BRT_LAUNCH((k_steer<true, true>), dim3((n + 255) / 256), dim3(256), 0, s, g, list_stride(K), sp, cuts, dt,
                           step, n, spos, sfwd, ssa, hash, lists, o, xcd_remap, feature_rate, features_in,
                           features_out, nullptr);
    else
        BRT_LAUNCH((k_steer<true>), dim3((n + 255) / 256), dim3(256), 0, s, g, list_stride(K), sp, cuts, dt,
                           step,
                           n, spos, sfwd, ssa, hash, lists, o, xcd_remap, 0.0f, nullptr, nullptr, nullptr);
    return hipGetLastError();
}

hipError_t launch_steer_forces(uint32_t K, const brt_steer_params& sp, float dt, uint32_t step, uint32_t n,
                               const float4* spos, const float4* sfwd, const uint2* hash, const uint32_t* lists,
                               float2* wander, float4* force_out, uint32_t xcd_remap, hipStream_t s) {
    AgentOut o{nullptr, nullptr, nullptr, force_out, wander};
    brt_grid_params g{};
    hipLaunchKernelGGL((k_steer<false>), dim3((n + 255) / 256), dim3(256), 0, s, g, list_stride(K), sp,
                       steer_cuts(sp), dt, step, n,
                       spos, sfwd, sfwd, hash, lists, o, xcd_remap, 0.0f, nullptr, nullptr);
    return hipGetLastError();
}

hipError_t launch_apply(const brt_grid_params& g, const brt_steer_params& sp, float dt, uint32_t n, const float4* spos,
                        const float4* sfwd, const float4* ssa, const uint2* hash, const float4* force_in,
                        float4* pos_out, float4* fwd_out, float4* sa_out, float4* force_out, hipStream_t s) {
    AgentOut o{pos_out, fwd_out, sa_out, force_out, nullptr};
    hipLaunchKernelGGL(k_apply, dim3((n + 255) / 256), dim3(256), 0, s, g, sp, dt, n, spos, sfwd, ssa, hash, force_in,
                       o);
    return hipGetLastError();
}

hipError_t launch_features(uint32_t K, float rate, uint32_t n, const uint2* hash, const uint32_t* lists,
                           const float4* fin, float4* fout, hipStream_t s) {
    BRT_LAUNCH(k_features, dim3((n + 255) / 256), dim3(256), 0, s, list_stride(K), rate, n, hash, lists, fin,
                       fout);
    return hipGetLastError();
}

}  // namespace brt
