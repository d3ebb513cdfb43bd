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
#include "brt_lists.hpp"

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

// Features gathered from agent-ID order into sorted order, padded with zeros
// to FQ float4 per agent, and each agent's feature norm sqrtf(sum_k f_k^2)
// (k ascending, as the reference's computeSimilarity accumulates it,
// BehaviorClustering_resources.cu:156-198).  Neighbours are near each other in
// sorted order, so the similarity pass then reads them with L2 locality and
// without the sorted->ID indirection.
// vec: F is a multiple of 4 and `features` 16-byte aligned, so each row is
// read as float4s.
__global__ void __launch_bounds__(256) k_bc_sort_features(uint32_t F, uint32_t FQ, uint32_t n, uint32_t vec,
                                                          const uint2* __restrict__ hash,
                                                          const float* __restrict__ features,
                                                          float4* __restrict__ sfeat, float* __restrict__ norms) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t id = hash[i].y;
    float4* out = sfeat + (size_t)i * FQ;
    if (id >= n) {
        for (uint32_t q = 0; q < FQ; ++q) out[q] = make_float4(0.f, 0.f, 0.f, 0.f);
        norms[i] = 0.0f;
        return;
    }
    const float* f = features + (size_t)id * F;
    float acc = 0.0f;
    for (uint32_t q = 0; q < FQ; ++q) {
        float v[4];
        if (vec) {
            const float4 x = reinterpret_cast<const float4*>(f)[q];
            v[0] = x.x;
            v[1] = x.y;
            v[2] = x.z;
            v[3] = x.w;
        } else {
#pragma unroll
            for (int c = 0; c < 4; ++c) v[c] = 4u * q + c < F ? f[4u * q + c] : 0.0f;
        }
#pragma unroll
        for (int c = 0; c < 4; ++c)
            if (4u * q + c < F) acc = acc + v[c] * v[c];
        out[q] = make_float4(v[0], v[1], v[2], v[3]);
    }
    norms[i] = sqrtf(acc);
}

// The similarity of agent i (own features `mine`, norm n1) to a gathered
// neighbour: (d / (n1 * n2) + 1) / 2 with d summed over k < F in order.
template <int FQ>
__device__ __forceinline__ float naive_similarity(const float4 (&mine)[FQ], const float4 (&other)[FQ], uint32_t F,
                                                  float n1, float n2) {
    float d = 0.0f;
#pragma unroll
    for (int q = 0; q < FQ; ++q) {
        const float a[4] = {mine[q].x, mine[q].y, mine[q].z, mine[q].w};
        const float b[4] = {other[q].x, other[q].y, other[q].z, other[q].w};
#pragma unroll
        for (int c = 0; c < 4; ++c)
            if ((uint32_t)(4 * q + c) < F) d = d + a[c] * b[c];
    }
    return (d / (n1 * n2) + 1.0f) / 2.0f;
}

// Cluster cohesion + alignment accumulators (steerForClusterCohesion/Alignment,
// BehaviorClustering_resources.cu:282-384).  seek and flee share the truncated
// offset: myPos - p = -(p - myPos) and truncateLength(-v) = -truncateLength(v)
// exactly, so flee = (-t) - v is the reference's flee to the bit.
struct ClusterAccum {
    V3 coh, ali;
    __device__ __forceinline__ void init() {
        coh = v3(0.f, 0.f, 0.f);
        ali = coh;
    }
    __device__ __forceinline__ void add(V3 myPos, V3 vel, float maxSpeed, V3 op, V3 of, float sim) {
        const V3 t = truncateLength(op - myPos, maxSpeed);
        const V3 seek = t - vel;
        const V3 flee = v3(-t.x, -t.y, -t.z) - vel;
        coh = coh + (seek * sim + flee * (1.0f - sim));
        ali = ali + of * sim;
    }
    // force_out = a + (c + force_in), the blend order of the two kernels
    __device__ __forceinline__ float4 finish(uint32_t count, V3 myFwd, const brt_bc_params& bc, float4 f0) const {
        V3 c = coh, a = ali;
        if (count > 0) {
            c = normalize(c / (float)count) * bc.cluster_weight;
            a = normalize(a / (float)count - myFwd) * bc.alignment_weight;
        }
        const V3 f1 = c + xyz(f0);
        const V3 f2 = a + f1;
        return make_float4(f2.x, f2.y, f2.z, 0.0f);
    }
};

// computeNeighNaiveSimilarities (+ computeAdaptiveSimilarities) over sorted
// features; with STEER also steerForClusterCohesion + Alignment on the fly
// (the whole of BehaviorClusteringPlugIn::steerForClustering in one pass).
// One lane per sorted agent; neighbours in groups of G (8 gathers of FQ float4
// in flight at FQ <= 2).  Each lane writes its similarity row into LDS (dynamic,
// blockDim * (K+1) floats); the rows of a wave's 64 agents are contiguous in
// `sims` too, so the wave then stores them with coalesced (float4 when `vec`)
// writes.
extern __shared__ float bc_rows[];

template <bool ADAPTIVE, bool STEER, int FQ>
__device__ __forceinline__ void similarity_agent(uint32_t i, uint32_t K, uint32_t W, const brt_bc_params& bc,
                                                 uint32_t n, const uint2* __restrict__ hash,
                                                 const uint32_t* __restrict__ lists,
                                                 const float4* __restrict__ sfeat, const float* __restrict__ norms,
                                                 const float4* __restrict__ stats_in, float4* __restrict__ stats_out,
                                                 float* out, const brt_steer_params& sp,
                                                 const float4* __restrict__ spos, const float4* __restrict__ sfwd,
                                                 const float4* force_in, float4* force_out) {
    constexpr int G = FQ <= 2 ? 8 : (FQ <= 4 ? 4 : 2);
    const uint32_t F = bc.feature_count;
    const uint32_t id = hash[i].y;
    float4 mine[FQ];
#pragma unroll
    for (int q = 0; q < FQ; ++q) mine[q] = sfeat[(size_t)i * FQ + q];
    const float n1 = norms[i];
    const bool valid = id < n;
    float4 st = ADAPTIVE && valid ? stats_in[id] : make_float4(0.f, 0.f, 0.f, 0.f);
    V3 myPos = v3(0.f, 0.f, 0.f), myFwd = myPos, vel = myPos;
    if (STEER) {
        const float4 P = spos[i], Fw = sfwd[i];
        myPos = xyz(P);
        myFwd = xyz(Fw);
        vel = myFwd * Fw.w;
    }
    ClusterAccum acc;
    acc.init();
    const uint32_t count = for_list_octets<STEER ? kRowPrefetch : kRowWhole>(lists, W, K, i, n, [&](uint32_t base, const uint32_t(&jj)[8], uint32_t mask) {
#pragma unroll
        for (int g0 = 0; g0 < 8; g0 += G) {
            float4 other[G][FQ];
            float n2[G];
            float4 op[G], of[G];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const uint32_t j = jj[g0 + g];
#pragma unroll
                for (int q = 0; q < FQ; ++q) other[g][q] = sfeat[(size_t)j * FQ + q];
                n2[g] = norms[j];
                if (STEER) {
                    op[g] = spos[j];
                    of[g] = sfwd[j];
                }
            }
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const int k = g0 + g;
                if (!(mask & (1u << k))) continue;
                float sim = 0.0f;
                if (valid) {
                    sim = naive_similarity<FQ>(mine, other[g], F, n1, n2[g]);
                    if (ADAPTIVE) {
                        st = update_similarity_stats(sim, st, bc);
                        sim = interpolate_similarity(sim, st.x, st.y, st.z);
                    }
                }
                out[base + k - 1u] = sim;
                if (STEER) acc.add(myPos, vel, sp.max_speed, xyz(op[g]), xyz(of[g]), sim);
            }
        }
    });
    for (uint32_t t = count; t <= K; ++t) out[t] = 0.0f;
    if (ADAPTIVE && valid) stats_out[id] = st;
    if (STEER) force_out[i] = acc.finish(count, myFwd, bc, force_in[i]);
}

template <bool ADAPTIVE, bool STEER, int FQ>
__global__ void __launch_bounds__(256) k_bc_similarities(uint32_t K, uint32_t W, brt_bc_params bc, uint32_t n,
                                                         const uint2* __restrict__ hash,
                                                         const uint32_t* __restrict__ lists,
                                                         const float4* __restrict__ sfeat,
                                                         const float* __restrict__ norms,
                                                         const float4* __restrict__ stats_in,
                                                         float4* __restrict__ stats_out, float* __restrict__ sims,
                                                         brt_steer_params sp, const float4* __restrict__ spos,
                                                         const float4* __restrict__ sfwd, const float4* force_in,
                                                         float4* force_out, uint32_t vec) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) similarity_agent<ADAPTIVE, STEER, FQ>(i, K, W, bc, n, hash, lists, sfeat, norms, stats_in, stats_out,
                                                      bc_rows + (size_t)threadIdx.x * (K + 1u), sp, spos, sfwd,
                                                      force_in, force_out);
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    const uint32_t i0 = blockIdx.x * blockDim.x + wv * 64u;
    if (i0 >= n) return;
    const uint32_t total = (n - i0 < 64u ? n - i0 : 64u) * (K + 1u);
    const float* src = bc_rows + (size_t)wv * 64u * (K + 1u);
    float* dst = sims + (size_t)i0 * (K + 1u);
    uint32_t x = lane;
    if (vec) {
        const uint32_t t4 = total / 4u;
        for (; x < t4; x += 64u) reinterpret_cast<float4*>(dst)[x] = reinterpret_cast<const float4*>(src)[x];
        x = 4u * t4 + lane;
    }
    for (; x < total; x += 64u) dst[x] = src[x];
}

// steerForClusterCohesion + Alignment from stored similarities.
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
    const V3 vel = myFwd * Fw.w;
    const float* s = sims + (size_t)i * (K + 1u);
    ClusterAccum acc;
    acc.init();
    const uint32_t count = for_list_octets(lists, W, K, i, n, [&](uint32_t base, const uint32_t(&jj)[8], uint32_t mask) {
        float4 op[8], of[8];
        float sv[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            op[k] = spos[jj[k]];
            of[k] = sfwd[jj[k]];
            sv[k] = (mask & (1u << k)) ? s[base + k - 1u] : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (mask & (1u << k)) acc.add(myPos, vel, sp.max_speed, xyz(op[k]), xyz(of[k]), sv[k]);
    });
    force_out[i] = acc.finish(count, myFwd, bc, force_in[i]);
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
__global__ void __launch_bounds__(256) k_lp_iterate(uint32_t K, uint32_t W, uint32_t n, const uint2* __restrict__ hash,
                                                    const uint32_t* __restrict__ lists,
                                                    const float* __restrict__ lbl_in, float* __restrict__ lbl_out,
                                                    const float* m_in, float* m_out,
                                                    const float4* __restrict__ spos, float4* __restrict__ conn) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float mn = FIRST ? (float)i : lbl_in[i];
    if (SCAN) {
        for_list_octets(lists, W, K, i, n, [&](uint32_t, const uint32_t(&jj)[8], uint32_t mask) {
            float v[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) v[k] = FIRST ? (float)jj[k] : lbl_in[jj[k]];
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if ((mask & (1u << k)) && mn > v[k]) mn = v[k];
        });
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

// Labels from float3 by agent ID into sorted order (one gather per agent), so
// the hit-rate pass reads its neighbours' labels next to each other.
__global__ void __launch_bounds__(256) k_gather_labels(uint32_t n, const uint2* __restrict__ hash,
                                                       const float* __restrict__ m_in, float* __restrict__ lbl) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t id = hash[i].y;
    lbl[i] = id < n ? m_in[(size_t)id * 3u] : 0.0f;
}

// analyzeClusterHitRate's kernel (BehaviorClustering_kernel.cu:502-578): among
// the listed neighbours within sameClusterMaxRadius, the fraction sharing the
// agent's label and the mean of their labels (-1 with none); float3 by ID.
__global__ void __launch_bounds__(256) k_hit_rate(uint32_t K, uint32_t W, float radius, uint32_t n,
                                                  const uint2* __restrict__ hash, const uint32_t* __restrict__ lists,
                                                  const float4* __restrict__ spos, const float* __restrict__ lbl,
                                                  float* __restrict__ m_out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t id = hash[i].y;
    if (id >= n) return;
    const float my = lbl[i];
    const V3 myPos = xyz(spos[i]);
    int close = 0;
    float real = 0.0f, hit = 0.0f;
    for_list_octets(lists, W, K, i, n, [&](uint32_t, const uint32_t(&jj)[8], uint32_t mask) {
        float4 op[8];
        float v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            op[k] = spos[jj[k]];
            v[k] = lbl[jj[k]];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (!(mask & (1u << k))) continue;
            const V3 offset = xyz(op[k]) - myPos;
            if (length(offset) > radius) continue;
            ++close;
            real = real + v[k];
            if (my == v[k]) hit = hit + 1.0f;
        }
    });
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
        hipLaunchKernelGGL((k_lp_iterate<true, true, false>), grid, block, 0, s, K, W, n, hash, lists, nullptr,
                           nullptr, m_in, m_out, spos, conn);
        return hipGetLastError();
    }
    if (iterations == 1) {
        hipLaunchKernelGGL((k_lp_iterate<true, true, true>), grid, block, 0, s, K, W, n, hash, lists, nullptr,
                           nullptr, m_in, m_out, spos, conn);
        return hipGetLastError();
    }
    hipLaunchKernelGGL((k_lp_iterate<true, false, true>), grid, block, 0, s, K, W, n, hash, lists, nullptr, lbl_a,
                       m_in, m_out, spos, conn);
    float* src = lbl_a;
    float* dst = lbl_b;
    for (uint32_t t = 1; t + 1 < iterations; ++t) {
        hipLaunchKernelGGL((k_lp_iterate<false, false, true>), grid, block, 0, s, K, W, n, hash, lists, src, dst,
                           m_in, m_out, spos, conn);
        float* x = src;
        src = dst;
        dst = x;
    }
    hipLaunchKernelGGL((k_lp_iterate<false, true, true>), grid, block, 0, s, K, W, n, hash, lists, src, nullptr,
                       m_in, m_out, spos, conn);
    return hipGetLastError();
}

hipError_t launch_hit_rate(uint32_t K, float radius, uint32_t n, const uint2* hash, const uint32_t* lists,
                           const float4* spos, const float* m_in, float* m_out, float* lbl, hipStream_t s) {
    const dim3 grid((n + 255) / 256), block(256);
    hipLaunchKernelGGL(k_gather_labels, grid, block, 0, s, n, hash, m_in, lbl);
    hipLaunchKernelGGL(k_hit_rate, grid, block, 0, s, K, list_stride(K), radius, n, hash, lists, spos, lbl, m_out);
    return hipGetLastError();
}

namespace {

template <bool ADAPTIVE, bool STEER>
hipError_t launch_sims_fq(uint32_t FQ, uint32_t K, const brt_bc_params& bc, uint32_t n, const uint2* hash,
                          const uint32_t* lists, const float4* sfeat, const float* norms, const float4* stats_in,
                          float4* stats_out, float* sims, const brt_steer_params& sp, const float4* spos,
                          const float4* sfwd, const float4* force_in, float4* force_out, hipStream_t s) {
    const uint32_t W = list_stride(K);
    // the LDS row tile: 256 agents while it stays within 40 KB, else one wave
    const uint32_t B = (size_t)256u * (K + 1u) * 4u <= 40u * 1024u ? 256u : 64u;
    const size_t lds = (size_t)B * (K + 1u) * 4u;
    const uint32_t vec = ((uintptr_t)sims & 15u) == 0 ? 1u : 0u;
    const dim3 grid((n + B - 1) / B), block(B);
#define BRT_SIMS_CASE(Q)                                                                                            \
    case Q:                                                                                                         \
        hipLaunchKernelGGL((k_bc_similarities<ADAPTIVE, STEER, Q>), grid, block, lds, s, K, W, bc, n, hash, lists, \
                           sfeat, norms, stats_in, stats_out, sims, sp, spos, sfwd, force_in, force_out, vec);     \
        break;
    switch (FQ) {
        BRT_SIMS_CASE(1)
        BRT_SIMS_CASE(2)
        BRT_SIMS_CASE(3)
        BRT_SIMS_CASE(4)
        BRT_SIMS_CASE(5)
        BRT_SIMS_CASE(6)
        BRT_SIMS_CASE(7)
        BRT_SIMS_CASE(8)
        default:
            return hipErrorInvalidValue;
    }
#undef BRT_SIMS_CASE
    return hipGetLastError();
}

}  // namespace

size_t bc_scratch_bytes(uint32_t feature_count, uint32_t n) {
    const size_t FQ = (feature_count + 3u) / 4u;
    return (size_t)n * FQ * 16u + ((size_t)n * 4u + 15u) / 16u * 16u;
}

hipError_t launch_bc_similarities(uint32_t K, const brt_bc_params& bc, uint32_t n, const uint2* hash,
                                  const uint32_t* lists, const float* features, const float4* stats_in,
                                  float4* stats_out, float* sims, void* scratch, const brt_steer_params* sp,
                                  const float4* spos, const float4* sfwd, const float4* force_in, float4* force_out,
                                  hipStream_t s) {
    const uint32_t F = bc.feature_count, FQ = (F + 3u) / 4u;
    float4* sfeat = reinterpret_cast<float4*>(scratch);
    float* norms = reinterpret_cast<float*>(sfeat + (size_t)n * FQ);
    const uint32_t vec = (F % 4u == 0 && ((uintptr_t)features & 15u) == 0) ? 1u : 0u;
    hipLaunchKernelGGL(k_bc_sort_features, dim3((n + 255) / 256), dim3(256), 0, s, F, FQ, n, vec, hash, features,
                       sfeat, norms);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const brt_steer_params none{};
    const brt_steer_params& p = sp ? *sp : none;
    if (bc.adaptive)
        return sp ? launch_sims_fq<true, true>(FQ, K, bc, n, hash, lists, sfeat, norms, stats_in, stats_out, sims, p,
                                               spos, sfwd, force_in, force_out, s)
                  : launch_sims_fq<true, false>(FQ, K, bc, n, hash, lists, sfeat, norms, stats_in, stats_out, sims,
                                                p, spos, sfwd, force_in, force_out, s);
    return sp ? launch_sims_fq<false, true>(FQ, K, bc, n, hash, lists, sfeat, norms, stats_in, stats_out, sims, p,
                                            spos, sfwd, force_in, force_out, s)
              : launch_sims_fq<false, false>(FQ, K, bc, n, hash, lists, sfeat, norms, stats_in, stats_out, sims, p,
                                             spos, sfwd, force_in, force_out, s);
}

hipError_t launch_bc_cluster_steer(uint32_t K, const brt_steer_params& sp, const brt_bc_params& bc, uint32_t n,
                                   const float4* spos, const float4* sfwd, const uint32_t* lists, const float* sims,
                                   const float4* force_in, float4* force_out, hipStream_t s) {
    hipLaunchKernelGGL(k_bc_cluster_steer, dim3((n + 255) / 256), dim3(256), 0, s, K, list_stride(K), sp, bc, n, spos,
                       sfwd, lists, sims, force_in, force_out);
    return hipGetLastError();
}

}  // namespace brt
