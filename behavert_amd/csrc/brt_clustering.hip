// brt_clustering.hip -- BehaviorClustering's similarity and cluster-steering
// kernels on gfx950.
//
// Reference semantics:
//   computeSimilarity(FromTex) ....... PlugIns/BehaviorClustering/BehaviorClustering_resources.cu:156-198
//   updateSimilarityStats ............ BehaviorClustering_resources.cu:84-150
//   interpolateSimilarity ............ BehaviorClustering_resources.cu:61-80
//   computeNeighNaiveSimilarities .... BehaviorClustering_kernel.cu:306-353
//   computeAdaptiveSimilarities ...... BehaviorClustering_kernel.cu:438-497
//   steerForClusterCohesion/Alignment  BehaviorClustering_resources.cu:282-384
//   xxxsteerForSeek/Flee ............. PlugIns/OpenSteerWrapper/OpenSteerWrapper_resources.cu:294-322
// The reference's pipeline (BehaviorClusteringPlugIn::steerForClustering,
// BehaviorClusteringPlugIn.h:737-817) runs the naive similarities, then the
// adaptive pass over them, then the two steering kernels.  Here the naive and
// adaptive passes are one kernel: the adaptive pass consumes each naive value
// once, in list order, so fusing them changes no bit.
//
// One lane per sorted agent.  The agent's own feature vector is held in
// registers; a neighbour's is gathered by ID.  Sums run in the reference's
// order (k = 0..F-1 for the dot products, list order for the forces), with
// one rounding per operation (-ffp-contract=off).
#include "brt_device.hpp"
#include "brt_internal.hpp"

namespace brt {

namespace {

// cutil_math clamp(f, a, b) = fmaxf(a, fminf(f, b))
__device__ __forceinline__ float cclamp(float f, float a, float b) { return fmaxf(a, fminf(f, b)); }

__device__ __forceinline__ float linear_interp(float x, float xa, float ya, float xb, float yb) {
    return ya * ((x - xb) / (xa - xb)) + yb * ((x - xa) / (xb - xa));
}

__device__ __forceinline__ float interpolate_similarity(float v, float mn, float avg, float mx) {
    if (v > avg) return cclamp(linear_interp(v, avg, 0.5f, mx, 1.0f), 0.0f, 1.0f);
    return cclamp(linear_interp(v, mn, 0.0f, avg, 0.5f), 0.0f, 1.0f);
}

__device__ __forceinline__ float4 update_similarity_stats(float sim, float4 st, const brt_bc_params& bc) {
    float mn = st.x, avg = st.y, mx = st.z;
    const float ctrl = bc.avg_similarity_controller;
    const float blend = sim * ctrl + ((mx + mn) / 2.0f) * (1.0f - ctrl);
    avg = avg + (blend - st.y) * bc.stats_params[0];
    avg = cclamp(avg, 0.0f, 1.0f);
    if (sim < mn) {
        mn = mn + (sim - mn) * bc.stats_params[1];
        if (mn > avg) mn = cclamp(mn, 0.0f, sim);
    } else if (mn < avg) {
        mn = mn + bc.stats_params[2];
    }
    if (sim > mx) {
        mx = mx + (sim - mx) * bc.stats_params[1];
        if (mx < avg) mx = cclamp(mx, sim, 1.0f);
    } else if (mx > avg) {
        mx = mx - bc.stats_params[2];
    }
    st.x = mn;
    st.y = avg;
    st.z = mx;
    return st;
}

template <bool ADAPTIVE>
__global__ void __launch_bounds__(256) k_bc_similarities(uint32_t K, uint32_t W, brt_bc_params bc, uint32_t n,
                                                         const uint2* __restrict__ hash,
                                                         const uint32_t* __restrict__ lists,
                                                         const float* __restrict__ features,
                                                         const float4* __restrict__ stats_in,
                                                         float4* __restrict__ stats_out, float* __restrict__ sims) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t F = bc.feature_count;
    const uint32_t id = hash[i].y;
    const uint32_t* row = lists + (size_t)i * W;
    const uint32_t count = row[0] < K ? row[0] : K;
    float* out = sims + (size_t)i * (K + 1u);
    if (id >= n) {
        for (uint32_t t = 0; t <= K; ++t) out[t] = 0.0f;
        return;
    }
    float mine[BRT_MAX_FEATURES];
    const float* fi = features + (size_t)id * F;
#pragma unroll
    for (uint32_t k = 0; k < BRT_MAX_FEATURES; ++k)
        if (k < F) mine[k] = fi[k];
    float4 st = ADAPTIVE ? stats_in[id] : make_float4(0.f, 0.f, 0.f, 0.f);
    for (uint32_t t = 0; t < count; ++t) {
        const uint32_t j = row[1 + t];
        const uint32_t jid = j < n ? hash[j].y : n;
        float sim = 0.0f;
        if (jid < n) {
            const float* fj = features + (size_t)jid * F;
            float n1 = 0.0f, n2 = 0.0f, d = 0.0f;
#pragma unroll
            for (uint32_t k = 0; k < BRT_MAX_FEATURES; ++k) {
                if (k < F) {
                    const float a = mine[k], b = fj[k];
                    d = d + a * b;
                    n1 = n1 + a * a;
                    n2 = n2 + b * b;
                }
            }
            n1 = sqrtf(n1);
            n2 = sqrtf(n2);
            sim = (d / (n1 * n2) + 1.0f) / 2.0f;
            if (ADAPTIVE) {
                st = update_similarity_stats(sim, st, bc);
                sim = interpolate_similarity(sim, st.x, st.y, st.z);
            }
        }
        out[t] = sim;
    }
    for (uint32_t t = count; t <= K; ++t) out[t] = 0.0f;
    if (ADAPTIVE) stats_out[id] = st;
}

__device__ __forceinline__ V3 seek_or_flee(V3 myPos, float mySpeed, V3 myFwd, V3 target, float maxSpeed,
                                           bool flee) {
    const V3 offset = flee ? myPos - target : target - myPos;
    return truncateLength(offset, maxSpeed) - myFwd * mySpeed;
}

__global__ void __launch_bounds__(256) k_bc_cluster_steer(uint32_t K, uint32_t W, brt_steer_params sp,
                                                          brt_bc_params bc, uint32_t n,
                                                          const float4* __restrict__ spos,
                                                          const float4* __restrict__ sfwd,
                                                          const uint32_t* __restrict__ lists,
                                                          const float* __restrict__ sims,
                                                          const float4* force_in, float4* force_out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 P = spos[i], Fw = sfwd[i];
    const V3 myPos = xyz(P), myFwd = xyz(Fw);
    const float mySpeed = Fw.w;
    const uint32_t* row = lists + (size_t)i * W;
    const uint32_t count = row[0] < K ? row[0] : K;
    const float* s = sims + (size_t)i * (K + 1u);
    V3 coh = v3(0.f, 0.f, 0.f), ali = coh;
    for (uint32_t t = 0; t < count; ++t) {
        const uint32_t j = row[1 + t];
        if (j >= n) continue;
        const float sim = s[t];
        const V3 op = xyz(spos[j]);
        const V3 seek = seek_or_flee(myPos, mySpeed, myFwd, op, sp.max_speed, false);
        const V3 flee = seek_or_flee(myPos, mySpeed, myFwd, op, sp.max_speed, true);
        coh = coh + (seek * sim + flee * (1.0f - sim));
        ali = ali + xyz(sfwd[j]) * sim;
    }
    if (count > 0) {
        coh = normalize(coh / (float)count) * bc.cluster_weight;
        ali = normalize(ali / (float)count - myFwd) * bc.alignment_weight;
    }
    const float4 f0 = force_in[i];
    const V3 f1 = coh + xyz(f0);
    const V3 f2 = ali + f1;
    force_out[i] = make_float4(f2.x, f2.y, f2.z, 0.0f);
}

// ---------------------------------------------------------------- cluster membership
// Local propagation (BehaviorClusteringPlugIn::computeClusterMembership,
// BehaviorClusteringPlugIn.h:823-865): every agent's label starts as its sorted
// index (resetClusterMembership, BehaviorClustering_kernel.cu:632-652), then T
// Jacobi iterations replace it by the least label among itself and its listed
// neighbours (computeClusterMembership_kernel, :583-625), then the connector
// pairs each agent's position with its label's (CMConnection_kernel, :657-689).
// The reference keeps the labels in float3.x by agent ID and gathers a
// neighbour's through the sorted->ID map; here they live in sorted order during
// the iterations (a neighbour's label is then lbl[j], next to the agent's own
// in memory), so each iteration reads the list row and one float per entry.
// The first iteration reads no labels (they are the indices themselves) and
// the last writes the float3 by ID and the connector.  Same values: a min of
// the same operands, the same T, no rounding.
template <bool FIRST, bool LAST, bool SCAN>
__global__ void __launch_bounds__(256) k_lp_iterate(uint32_t W, uint32_t n, const uint2* __restrict__ hash,
                                                    const uint32_t* __restrict__ lists,
                                                    const float* __restrict__ lbl_in, float* __restrict__ lbl_out,
                                                    const float* m_in, float* m_out,
                                                    const float4* __restrict__ spos, float4* __restrict__ conn) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float mn = FIRST ? (float)i : lbl_in[i];
    if (SCAN) {
        const uint32_t* row = lists + (size_t)i * W;
        const uint32_t count = row[0] < W - 1u ? row[0] : W - 1u;
        for (uint32_t t = 0; t < count; ++t) {
            const uint32_t j = row[1 + t];
            if (j >= n) continue;
            const float v = FIRST ? (float)j : lbl_in[j];
            if (mn > v) mn = v;
        }
    }
    if (!LAST) {
        lbl_out[i] = mn;
        return;
    }
    const uint32_t id = hash[i].y;
    if (id >= n) return;
    const size_t o = (size_t)id * 3u;
    const float y = m_in[o + 1], z = m_in[o + 2];
    m_out[o] = mn;
    m_out[o + 1] = y;
    m_out[o + 2] = z;
    if (conn) {
        conn[2u * (size_t)id] = spos[i];
        conn[2u * (size_t)id + 1u] = spos[(uint32_t)mn];
    }
}

// analyzeClusterHitRate's kernel (BehaviorClustering_kernel.cu:502-578): among
// the listed neighbours within sameClusterMaxRadius, the fraction sharing the
// agent's label and the mean of their labels (-1 with none); float3 by ID.
__global__ void __launch_bounds__(256) k_hit_rate(uint32_t W, float radius, uint32_t n, const uint2* __restrict__ hash,
                                                  const uint32_t* __restrict__ lists,
                                                  const float4* __restrict__ spos, const float* m_in,
                                                  float* m_out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t id = hash[i].y;
    if (id >= n) return;
    const float my = m_in[(size_t)id * 3u];
    const V3 myPos = xyz(spos[i]);
    const uint32_t* row = lists + (size_t)i * W;
    const uint32_t count = row[0] < W - 1u ? row[0] : W - 1u;
    int close = 0;
    float real = 0.0f, hit = 0.0f;
    for (uint32_t t = 0; t < count; ++t) {
        const uint32_t j = row[1 + t];
        if (j >= n) continue;
        const uint32_t jid = hash[j].y;
        if (jid >= n) continue;
        const float other = m_in[(size_t)jid * 3u];
        const V3 offset = xyz(spos[j]) - myPos;
        if (length(offset) > radius) continue;
        ++close;
        real = real + other;
        if (my == other) hit = hit + 1.0f;
    }
    if (close > 0) {
        hit = hit / (float)close;
        real = real / (float)close;
    } else {
        real = -1.0f;
    }
    const size_t o = (size_t)id * 3u;
    m_out[o] = my;
    m_out[o + 1] = hit;
    m_out[o + 2] = real;
}

}  // namespace

hipError_t launch_cluster_membership(uint32_t K, uint32_t n, uint32_t iterations, const uint2* hash,
                                     const uint32_t* lists, const float4* spos, const float* m_in, float* m_out,
                                     float4* conn, float* lbl_a, float* lbl_b, hipStream_t s) {
    const uint32_t W = list_stride(K);
    const dim3 grid((n + 255) / 256), block(256);
    if (iterations == 0) {
        hipLaunchKernelGGL((k_lp_iterate<true, true, false>), grid, block, 0, s, W, n, hash, lists, nullptr, nullptr,
                           m_in, m_out, spos, conn);
        return hipGetLastError();
    }
    if (iterations == 1) {
        hipLaunchKernelGGL((k_lp_iterate<true, true, true>), grid, block, 0, s, W, n, hash, lists, nullptr, nullptr,
                           m_in, m_out, spos, conn);
        return hipGetLastError();
    }
    hipLaunchKernelGGL((k_lp_iterate<true, false, true>), grid, block, 0, s, W, n, hash, lists, nullptr, lbl_a, m_in,
                       m_out, spos, conn);
    float* src = lbl_a;
    float* dst = lbl_b;
    for (uint32_t t = 1; t + 1 < iterations; ++t) {
        hipLaunchKernelGGL((k_lp_iterate<false, false, true>), grid, block, 0, s, W, n, hash, lists, src, dst, m_in,
                           m_out, spos, conn);
        float* x = src;
        src = dst;
        dst = x;
    }
    hipLaunchKernelGGL((k_lp_iterate<false, true, true>), grid, block, 0, s, W, n, hash, lists, src, nullptr, m_in,
                       m_out, spos, conn);
    return hipGetLastError();
}

hipError_t launch_hit_rate(uint32_t K, float radius, uint32_t n, const uint2* hash, const uint32_t* lists,
                           const float4* spos, const float* m_in, float* m_out, hipStream_t s) {
    hipLaunchKernelGGL(k_hit_rate, dim3((n + 255) / 256), dim3(256), 0, s, list_stride(K), radius, n, hash, lists,
                       spos, m_in, m_out);
    return hipGetLastError();
}

hipError_t launch_bc_similarities(uint32_t K, const brt_bc_params& bc, uint32_t n, const uint2* hash,
                                  const uint32_t* lists, const float* features, const float4* stats_in,
                                  float4* stats_out, float* sims, hipStream_t s) {
    const uint32_t W = list_stride(K);
    if (bc.adaptive)
        hipLaunchKernelGGL(k_bc_similarities<true>, dim3((n + 255) / 256), dim3(256), 0, s, K, W, bc, n, hash, lists,
                           features, stats_in, stats_out, sims);
    else
        hipLaunchKernelGGL(k_bc_similarities<false>, dim3((n + 255) / 256), dim3(256), 0, s, K, W, bc, n, hash, lists,
                           features, stats_in, stats_out, sims);
    return hipGetLastError();
}

hipError_t launch_bc_cluster_steer(uint32_t K, const brt_steer_params& sp, const brt_bc_params& bc, uint32_t n,
                                   const float4* spos, const float4* sfwd, const uint32_t* lists, const float* sims,
                                   const float4* force_in, float4* force_out, hipStream_t s) {
    hipLaunchKernelGGL(k_bc_cluster_steer, dim3((n + 255) / 256), dim3(256), 0, s, K, list_stride(K), sp, bc, n, spos,
                       sfwd, lists, sims, force_in, force_out);
    return hipGetLastError();
}

}  // namespace brt
