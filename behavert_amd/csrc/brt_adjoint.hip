// brt_adjoint.hip -- reverse-mode adjoint of one steering + integration step.
//
// Forward (one agent i, lists fixed as the forward step built them):
//   u_s = normalize(mean_j -(p_j-p_i)/|p_j-p_i|^2), u_a = normalize(mean_j f_j - f_i),
//   u_c = normalize(mean_j (p_j-p_i)), F = w_s u_s + w_a u_a + w_c u_c [+ seeks],
//   then applySteeringForceSingle (PlugIns/OpenSteerWrapper/OpenSteerWrapper_resources.cu:326-379):
//   F' = adjustRawSteeringForce(F, s, f), c = truncate(F', maxF), acc = c / m,
//   a1 = a + (acc - a) r, v0 = f s + a1 dt, v = truncate(v0, maxS),
//   s1 = |v|, p1 = p + v dt, f1 = normalize(v).
// Given dL/d(p1, f1, s1, a1) by agent ID this kernel recomputes the forward
// values of agent i (same arithmetic as the forward kernels), back-propagates
// through the integrator and the behaviour sums, writes agent i's own terms of
// dL/d(p, f, s, a) and adds its neighbours' terms (p_j, f_j appear in i's sums)
// into a scratch accumulator with hardware float atomics; a second kernel adds
// the accumulator to the outputs (only the neighbour terms need atomics).  The
// radius/angle predicates, the list membership and the world-box test are
// piecewise constant and carry no gradient.  dL/dw for the three behaviour
// weights is reduced per workgroup and then in a fixed order.
// The reference has no adjoint; the forward it differentiates is the one
// restated in brt_steer_common.hpp.
#include "brt_device.hpp"
#include "brt_internal.hpp"
#include "brt_steer_common.hpp"

namespace brt {

namespace {

__device__ __forceinline__ V3 normalize_bwd(V3 x, V3 ybar) {
    const float len = length(x);
    const V3 y = x * (1.0f / len);
    return (ybar - y * dot(y, ybar)) * (1.0f / len);
}

__device__ __forceinline__ V3 truncate_bwd(V3 x, float maxLength, V3 ybar) {
    const float len = length(x);
    if (len <= maxLength) return ybar;
    const V3 u = x * (1.0f / len);
    return (ybar - u * dot(u, ybar)) * (maxLength / len);
}

__device__ __forceinline__ void atomic_add3(float4* a, V3 v) {
    float* p = reinterpret_cast<float*>(a);
    unsafeAtomicAdd(p + 0, v.x);
    unsafeAtomicAdd(p + 1, v.y);
    unsafeAtomicAdd(p + 2, v.z);
}

struct AdjArgs {
    brt_grid_params g;
    brt_steer_params sp;
    float dt;
    uint32_t W, n;
    const float4* spos;
    const float4* sfwd;
    const float4* ssa;
    const uint2* hash;
    const uint32_t* lists;
    const float4* gpos_out;
    const float4* gfwd_out;
    const float4* gsa_out;
    float4* gpos_in;
    float4* gfwd_in;
    float4* gsa_in;
    float4* nbr_pos;         // neighbour terms of dL/dp_j by sorted index j, accumulated with atomics
    float4* nbr_fwd;         // neighbour terms of dL/df_j by sorted index j
    float* partials;         // per workgroup: 3 floats
    uint32_t* done;          // workgroups finished (reset to 0 by the last one)
    float* weight_grad;
    uint32_t xcd_remap;
};

__device__ __forceinline__ uint32_t list_entry(const uint32_t* row, uint32_t t) { return row[1 + t]; }

// Sums the per-workgroup partials in a fixed order (strided per-thread sums,
// then a fixed LDS tree) and adds the result to weight_grad; the same inputs
// always give the same bits.
// (256 threads)
__device__ void reduce_partials(const float* __restrict__ partials, uint32_t blocks, float* __restrict__ weight_grad) {
    __shared__ float s_sum[3][256];
    const uint32_t tid = threadIdx.x;
    float v0 = 0.0f, v1 = 0.0f, v2 = 0.0f;
    for (uint32_t b = tid; b < blocks; b += 256u) {
        v0 += partials[(size_t)b * 3 + 0];
        v1 += partials[(size_t)b * 3 + 1];
        v2 += partials[(size_t)b * 3 + 2];
    }
    s_sum[0][tid] = v0;
    s_sum[1][tid] = v1;
    s_sum[2][tid] = v2;
    __syncthreads();
    for (uint32_t o = 128; o > 0; o >>= 1) {
        if (tid < o) {
            s_sum[0][tid] += s_sum[0][tid + o];
            s_sum[1][tid] += s_sum[1][tid + o];
            s_sum[2][tid] += s_sum[2][tid + o];
        }
        __syncthreads();
    }
    if (tid < 3) weight_grad[tid] = weight_grad[tid] + s_sum[tid][0];
}

__global__ void __launch_bounds__(256) k_adjoint(AdjArgs a) {
    __shared__ float s_red[3][4];
    const uint32_t i = xcd_block(a.xcd_remap != 0) * blockDim.x + threadIdx.x;
    float wg[3] = {0.f, 0.f, 0.f};
    const brt_steer_params& sp = a.sp;
    if (i < a.n && a.hash[i].y < a.n) {
        const uint32_t id = a.hash[i].y;
        const float4 P = a.spos[i], Fw = a.sfwd[i], SA = a.ssa[i];
        const V3 p = xyz(P), f = xyz(Fw), sa = xyz(SA);
        const float s = Fw.w;
        const uint32_t* row = a.lists + (size_t)i * a.W;
        const uint32_t count = row[0] < a.W - 1u ? row[0] : a.W - 1u;
        const bool behaviours = sp.separation[0] != 0.0f || sp.alignment[0] != 0.0f || sp.cohesion[0] != 0.0f;

        // ---- forward recompute: behaviour sums (list order) -------------------
        SteerAccum acc;
        acc.init();
        if (behaviours)
            for (uint32_t t = 0; t < count; t += 4) {
                // four neighbours' gathers in flight
                uint32_t jj[4];
                float4 q[4], fq[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const uint32_t j = (t + u < count) ? list_entry(row, t + u) : a.n;
                    jj[u] = j;
                    const uint32_t js = j < a.n ? j : i;
                    q[u] = a.spos[js];
                    fq[u] = a.sfwd[js];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (jj[u] >= a.n) continue;
                    const V3 off = xyz(q[u]) - p;
                    acc.add(sp, weights_of(sp), f, off, dot(off, off), xyz(fq[u]));
                }
            }
        V3 xs = v3(0.f, 0.f, 0.f), xa = xs, xc = xs, us = xs, ua = xs, uc = xs;
        V3 F = v3(0.f, 0.f, 0.f);
        if (sp.separation[0] != 0.0f) {
            if (acc.nSep > 0) {
                xs = acc.sSep / (float)acc.nSep;
                us = normalize(xs);
            }
            F = us * sp.separation[0] + F;
        }
        if (sp.alignment[0] != 0.0f) {
            if (acc.nAli > 0) {
                xa = acc.sAli / (float)acc.nAli - f;
                ua = normalize(xa);
            }
            F = ua * sp.alignment[0] + F;
        }
        if (sp.cohesion[0] != 0.0f) {
            if (acc.nCoh > 0) {
                xc = acc.sCoh / (float)acc.nCoh;
                uc = normalize(xc);
            }
            F = uc * sp.cohesion[0] + F;
        }
        V3 tgt = v3(0.f, 0.f, 0.f), seekOff = tgt;
        const bool doTarget = sp.num_targets > 0 && sp.target_weight != 0.0f;
        if (doTarget) {
            const float* t = sp.targets[id % sp.num_targets];
            tgt = v3(t[0], t[1], t[2]);
            seekOff = tgt - p;
            F = steerForSeek(p, s, f, tgt, sp.max_speed) * sp.target_weight + F;
        }
        bool doCenter = false;
        V3 centerOff = v3(0.f, 0.f, 0.f);
        if (sp.seek_world_center) {
            const brt_grid_params& g = a.g;
            const bool inside = (p.x > g.world_center[0] - g.world_radius[0]) &&
                                (p.x < g.world_center[0] + g.world_radius[0]) &&
                                (p.y > g.world_center[1] - g.world_radius[1]) &&
                                (p.y < g.world_center[1] + g.world_radius[1]) &&
                                (p.z > g.world_center[2] - g.world_radius[2]) &&
                                (p.z < g.world_center[2] + g.world_radius[2]);
            const V3 c = v3(g.world_center[0], g.world_center[1], g.world_center[2]);
            doCenter = !inside;
            centerOff = c - p;
            F = (inside ? v3(0.f, 0.f, 0.f) : steerForSeek(p, s, f, c, sp.max_speed)) + F;
        }

        // ---- forward recompute: integrator ------------------------------------
        const V3 adjF = adjustRawSteeringForce(F, s, f, sp.max_speed);
        const V3 clipped = truncateLength(adjF, sp.max_force);
        const V3 accel = clipped / sp.mass;
        float rate = 0.0f;
        V3 a1 = sa;
        if (a.dt > 0.0f) {
            rate = clampf(clampf(9.0f * a.dt, 0.15f, 0.4f), 0.0f, 1.0f);
            a1 = sa + (accel - sa) * rate;
        }
        const V3 v0 = f * s + a1 * a.dt;
        const V3 v = truncateLength(v0, sp.max_speed);

        // ---- backward: integrator -------------------------------------------
        const float4 gP = a.gpos_out[id], gF = a.gfwd_out[id], gA = a.gsa_out[id];
        const V3 gp1 = xyz(gP), gf1 = xyz(gF), ga1 = xyz(gA);
        const float gs1 = gF.w;
        const float vlen = length(v);
        const V3 f1 = v * (1.0f / vlen);
        V3 vbar = gp1 * a.dt + normalize_bwd(v, gf1) + f1 * gs1;
        const V3 v0bar = truncate_bwd(v0, sp.max_speed, vbar);
        V3 pbar = gp1;
        V3 fbar = v0bar * s;
        float sbar = dot(v0bar, f);
        V3 a1bar = ga1 + v0bar * a.dt;
        V3 abar = a1bar;
        V3 accbar = v3(0.f, 0.f, 0.f);
        if (a.dt > 0.0f) {
            abar = a1bar * (1.0f - rate);
            accbar = a1bar * rate;
        }
        const V3 clipbar = accbar / sp.mass;
        V3 Fbar = truncate_bwd(adjF, sp.max_force, clipbar);
        // adjustRawSteeringForce: identity unless slow and F != 0 and outside the cone
        const float maxAdj = 0.2f * sp.max_speed;
        const bool zeroF = F.x == 0.0f && F.y == 0.0f && F.z == 0.0f;
        if (!(s > maxAdj || zeroF)) {
            const float range = s / maxAdj;
            const float c = 1.0f + (-2.0f * pow20(range));
            const float L = length(F);
            const V3 d = F / L;
            if (dot(d, f) < c) {
                const V3 o = Fbar;  // dL/d(output of the cone limit)
                const float proj = dot(F, f);
                const V3 perp = F - f * proj;
                const float pl = length(perp);
                const V3 u = perp * (1.0f / pl);
                const float pd = sqrtf(1.0f - c * c);
                const V3 inner = f * c + u * pd;
                const float Lbar = dot(o, inner);
                const V3 gi = o * L;
                fbar = fbar + gi * c;
                const V3 ubar = gi * pd;
                const V3 perpbar = (ubar - u * dot(u, ubar)) * (1.0f / pl);
                Fbar = perpbar - f * dot(f, perpbar) + F * (Lbar / L);
                fbar = fbar - perpbar * proj - F * dot(f, perpbar);
                // c = 1 - 2x, x = (s/maxAdj)^20, and pd = sqrt(1 - c^2) = 2 r^10 sqrt(1 - x):
                // dc/ds = -40 r^19 / maxAdj and dpd/ds = 20 c r^9 / (maxAdj sqrt(1 - x)), finite
                // where the FP32 pd itself rounds to 0 (c rounds to 1 for small r)
                float r9 = 1.0f;
                for (int k = 0; k < 9; ++k) r9 = r9 * range;
                const float r10 = r9 * range;
                const float r19 = r9 * r10;
                const float om = 1.0f - r10 * r10;
                const float dc_ds = -40.0f * r19 / maxAdj;
                const float dpd_ds = om > 0.0f ? 20.0f * c * r9 / (maxAdj * sqrtf(om)) : 0.0f;
                sbar = sbar + dot(gi, f) * dc_ds + dot(gi, u) * dpd_ds;
            }
        }
        // seeks: d/d(p, f, s) of truncate(target - p, maxS) - f s
        if (doTarget) {
            const V3 g = Fbar * sp.target_weight;
            pbar = pbar - truncate_bwd(seekOff, sp.max_speed, g);
            fbar = fbar - g * s;
            sbar = sbar - dot(g, f);
        }
        if (doCenter) {
            pbar = pbar - truncate_bwd(centerOff, sp.max_speed, Fbar);
            fbar = fbar - Fbar * s;
            sbar = sbar - dot(Fbar, f);
        }

        // ---- backward: behaviour sums -------------------------------------------
        V3 Sbar = v3(0.f, 0.f, 0.f), Abar = Sbar, Cbar = Sbar;
        if (sp.separation[0] != 0.0f && acc.nSep > 0) {
            wg[0] = dot(Fbar, us);
            Sbar = normalize_bwd(xs, Fbar * sp.separation[0]) * (1.0f / (float)acc.nSep);
        }
        if (sp.alignment[0] != 0.0f && acc.nAli > 0) {
            wg[1] = dot(Fbar, ua);
            const V3 xbar = normalize_bwd(xa, Fbar * sp.alignment[0]);
            fbar = fbar - xbar;
            Abar = xbar * (1.0f / (float)acc.nAli);
        }
        if (sp.cohesion[0] != 0.0f && acc.nCoh > 0) {
            wg[2] = dot(Fbar, uc);
            Cbar = normalize_bwd(xc, Fbar * sp.cohesion[0]) * (1.0f / (float)acc.nCoh);
        }
        if (behaviours)
            for (uint32_t t = 0; t < count; t += 4) {
                uint32_t jj[4];
                float4 q[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const uint32_t j = (t + u < count) ? list_entry(row, t + u) : a.n;
                    jj[u] = j;
                    const uint32_t js = j < a.n ? j : i;
                    q[u] = a.spos[js];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const uint32_t j = jj[u];
                    if (j >= a.n) continue;
                    const V3 off = xyz(q[u]) - p;
                    const float d2 = dot(off, off);
                    const float dist = sqrtf(d2);
                    const float cosA = dot(f, off * (1.0f / dist));
                    const float minR = sp.sensing_min_radius;
                    V3 offbar = v3(0.f, 0.f, 0.f);
                    bool any = false;
                    if (acc.nSep > 0 && sp.separation[0] != 0.0f &&
                        inBoid(dist, cosA, minR, sp.separation[1], sp.separation[2])) {
                        // d/d(off) of -off/|off|^2
                        const float inv = 1.0f / d2;
                        offbar = offbar - Sbar * inv + off * (2.0f * dot(off, Sbar) * inv * inv);
                        any = true;
                    }
                    if (acc.nCoh > 0 && sp.cohesion[0] != 0.0f &&
                        inBoid(dist, cosA, minR, sp.cohesion[1], sp.cohesion[2])) {
                        offbar = offbar + Cbar;
                        any = true;
                    }
                    if (any) {
                        atomic_add3(&a.nbr_pos[j], offbar);
                        pbar = pbar - offbar;
                    }
                    if (acc.nAli > 0 && sp.alignment[0] != 0.0f &&
                        inBoid(dist, cosA, minR, sp.alignment[1], sp.alignment[2]))
                        atomic_add3(&a.nbr_fwd[j], Abar);
                }
            }
        // the agent's own terms are written once (its ID appears once in the
        // sorted hash); the neighbours' terms are added by k_adjoint_combine
        a.gpos_in[id] = make_float4(pbar.x, pbar.y, pbar.z, 0.0f);
        a.gfwd_in[id] = make_float4(fbar.x, fbar.y, fbar.z, sbar);
        a.gsa_in[id] = make_float4(abar.x, abar.y, abar.z, 0.0f);
    }
    // workgroup reduction of dL/dw in a fixed order
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float v = wg[k];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
        if (lane == 0) s_red[k][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const uint32_t nw = blockDim.x / 64u;
        float v = 0.0f;
        for (uint32_t w = 0; w < nw; ++w) v += s_red[threadIdx.x][w];
        a.partials[(size_t)blockIdx.x * 3 + threadIdx.x] = v;
    }
    // the last workgroup to finish sums every workgroup's partials in a fixed
    // order (no wait: it only runs once all others have published).  The
    // partials are written by wave 0, so only wave 0 fences before the count.
    __shared__ bool s_last;
    if (threadIdx.x < 64u) {
        __threadfence();
        if (threadIdx.x == 0) s_last = atomicAdd(a.done, 1u) == gridDim.x - 1u;
    }
    __syncthreads();
    if (s_last) {
        __threadfence();
        reduce_partials(a.partials, gridDim.x, a.weight_grad);
        if (threadIdx.x == 0) __hip_atomic_store(a.done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__global__ void __launch_bounds__(256) k_zero3(float4* __restrict__ a, float4* __restrict__ b,
                                               float4* __restrict__ c, float4* __restrict__ d,
                                               uint32_t n) {
    const uint32_t stride = gridDim.x * blockDim.x;
    const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        a[i] = z;
        b[i] = z;
        c[i] = z;
        d[i] = z;
        d[n + i] = z;
    }
}

// dL/d(p, f) of each agent: its own terms plus the neighbours' terms, which
// were accumulated by sorted index (neighbours are close in sorted order, so
// the atomics of a wave share cache lines)
__global__ void __launch_bounds__(256) k_adjoint_combine(float4* __restrict__ gpos_in, float4* __restrict__ gfwd_in,
                                                         const float4* __restrict__ nbr, const uint2* __restrict__ hash,
                                                         uint32_t n) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t id = hash[j].y;
    if (id >= n) return;
    const float4 np = nbr[j], nf = nbr[n + j];
    float4 p = gpos_in[id], f = gfwd_in[id];
    p.x += np.x;
    p.y += np.y;
    p.z += np.z;
    f.x += nf.x;
    f.y += nf.y;
    f.z += nf.z;
    gpos_in[id] = p;
    gfwd_in[id] = f;
}

}  // namespace

hipError_t launch_adjoint(const brt_grid_params& g, uint32_t K, const brt_steer_params& sp, float dt, uint32_t n,
                          const float4* spos, const float4* sfwd, const float4* ssa, const uint2* hash,
                          const uint32_t* lists, const float4* gpos_out, const float4* gfwd_out,
                          const float4* gsa_out, float4* gpos_in, float4* gfwd_in, float4* gsa_in,
                          float4* nbr, float* partials, uint32_t* done, float* weight_grad, uint32_t xcd_remap,
                          hipStream_t s) {
    const uint32_t blocks = (n + 255) / 256;
    uint32_t zgrid = blocks < 2048 ? blocks : 2048;
    hipLaunchKernelGGL(k_zero3, dim3(zgrid), dim3(256), 0, s, gpos_in, gfwd_in, gsa_in, nbr, n);
    AdjArgs a{};
    a.g = g;
    a.sp = sp;
    a.dt = dt;
    a.W = list_stride(K);
    a.n = n;
    a.spos = spos;
    a.sfwd = sfwd;
    a.ssa = ssa;
    a.hash = hash;
    a.lists = lists;
    a.gpos_out = gpos_out;
    a.gfwd_out = gfwd_out;
    a.gsa_out = gsa_out;
    a.gpos_in = gpos_in;
    a.gfwd_in = gfwd_in;
    a.gsa_in = gsa_in;
    a.nbr_pos = nbr;
    a.nbr_fwd = nbr + n;
    a.partials = partials;
    a.done = done;
    a.weight_grad = weight_grad;
    a.xcd_remap = xcd_remap;
    hipLaunchKernelGGL(k_adjoint, dim3(blocks), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_adjoint_combine, dim3(blocks), dim3(256), 0, s, gpos_in, gfwd_in, nbr, hash, n);
    return hipGetLastError();
}

}  // namespace brt
