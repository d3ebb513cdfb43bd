// brt_adjoint.hip -- reverse-mode adjoint of one steering + integration step.
//
// Forward (one agent i, lists fixed as the forward step built them):
//   u_s = normalize(mean_j -(p_j-p_i)/|p_j-p_i|^2), u_a = normalize(mean_j f_j - f_i),
//   u_c = normalize(mean_j (p_j-p_i)), F = w_s u_s + w_a u_a + w_c u_c [+ seeks],
//   then applySteeringForceSingle (PlugIns/OpenSteerWrapper/OpenSteerWrapper_resources.cu:326-379):
//   F' = adjustRawSteeringForce(F, s, f), c = truncate(F', maxF), acc = c / m,
//   a1 = a + (acc - a) r, v0 = f s + a1 dt, v = truncate(v0, maxS),
//   s1 = |v|, p1 = p + v dt, f1 = normalize(v).
// Given dL/d(p1, f1, s1, a1) by agent ID, two kernels produce dL/d(p, f, s, a)
// by agent ID, with no atomics, so the same inputs always give the same bits:
//   k_adjoint_own (one lane per sorted agent i) recomputes agent i's forward
//     values with the forward kernels' arithmetic, back-propagates through the
//     integrator and its own behaviour sums, and stores by sorted index its own
//     terms and the three vectors its neighbours' terms are made of (the
//     separation, cohesion and alignment sums' adjoints, with flags);
//   k_adjoint_gather (several lanes per sorted agent j) finds every agent i
//     whose list holds j -- j's own list when that list can stand in for them,
//     else the agents of the 27 cells around j's cell -- by repeating the list
//     test (radius, cell range, and, for a full list, j's place in i's
//     visiting order against i's last entry), adds i's terms for p_j and f_j
//     in a fixed order, and writes the sums by ID.
// The radius/angle predicates, the list membership and the world-box test are
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

// Flags of an agent's neighbour-term vectors: which of its behaviour sums have
// neighbour terms, whether its own cell (calcGridPos) lies in the grid, and
// whether its list is full (count == K), and whether that cell is on a low
// face of the grid (a coordinate 0: agents below the grid are keyed at the
// far face, so lists there are not symmetric).
constexpr uint32_t kAdjSep = 1u, kAdjAli = 2u, kAdjCoh = 4u, kAdjInGrid = 8u, kAdjFull = 16u, kAdjLowFace = 32u;

struct AdjArgs {
    brt_grid_params g;
    brt_steer_params sp;
    float dt;
    uint32_t K, W, n;
    const float4* spos;
    const float4* sfwd;
    const float4* ssa;
    const uint2* hash;
    const uint32_t* lists;
    const float4* gpos_out;
    const float4* gfwd_out;
    const float4* gsa_out;
    // by sorted index: 3 float4 of own terms (dL/dp, (dL/df, dL/ds), dL/da) and
    // 3 float4 of neighbour-term vectors (S.xyz C.x | C.yz A.xy | A.z flags)
    float4* own;
    float4* bars;
    float* partials;  // per workgroup: 3 floats
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

// The terms one pair (i, j) of agent i's list contributes to dL/d(off),
// off = p_j - p_i, from i's separation and cohesion sums (anyP: one of them
// holds the pair), and whether i's alignment sum holds it; the forward's
// predicates, recomputed with the forward's arithmetic.
__device__ __forceinline__ V3 pair_terms(const brt_steer_params& sp, uint32_t flags, V3 f, V3 off, V3 Sbar, V3 Cbar,
                                         bool& anyP, bool& ali) {
    const float d2 = dot(off, off);
    const float dist = sqrtf(d2);
    const float cosA = dot(f, off * (1.0f / dist));
    const float minR = sp.sensing_min_radius;
    V3 offbar = v3(0.f, 0.f, 0.f);
    anyP = false;
    if ((flags & kAdjSep) && inBoid(dist, cosA, minR, sp.separation[1], sp.separation[2])) {
        // d/d(off) of -off/|off|^2
        const float inv = 1.0f / d2;
        offbar = offbar - Sbar * inv + off * (2.0f * dot(off, Sbar) * inv * inv);
        anyP = true;
    }
    if ((flags & kAdjCoh) && inBoid(dist, cosA, minR, sp.cohesion[1], sp.cohesion[2])) {
        offbar = offbar + Cbar;
        anyP = true;
    }
    ali = (flags & kAdjAli) && inBoid(dist, cosA, minR, sp.alignment[1], sp.alignment[2]);
    return offbar;
}

__global__ void __launch_bounds__(256) k_adjoint_own(AdjArgs a) {
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
        uint32_t flags = row[0] >= a.K ? kAdjFull : 0u;
        {
            const I3 gpi = calcGridPos(p.x, p.y, p.z, a.g);
            if (inGrid(gpi.x, gpi.y, gpi.z, a.g)) flags |= kAdjInGrid;
            if (gpi.x == 0 || gpi.y == 0 || gpi.z == 0) flags |= kAdjLowFace;
        }
        if (sp.separation[0] != 0.0f && acc.nSep > 0) {
            flags |= kAdjSep;
            wg[0] = dot(Fbar, us);
            Sbar = normalize_bwd(xs, Fbar * sp.separation[0]) * (1.0f / (float)acc.nSep);
        }
        if (sp.alignment[0] != 0.0f && acc.nAli > 0) {
            flags |= kAdjAli;
            wg[1] = dot(Fbar, ua);
            const V3 xbar = normalize_bwd(xa, Fbar * sp.alignment[0]);
            fbar = fbar - xbar;
            Abar = xbar * (1.0f / (float)acc.nAli);
        }
        if (sp.cohesion[0] != 0.0f && acc.nCoh > 0) {
            flags |= kAdjCoh;
            wg[2] = dot(Fbar, uc);
            Cbar = normalize_bwd(xc, Fbar * sp.cohesion[0]) * (1.0f / (float)acc.nCoh);
        }
        // this agent's share of the neighbour terms: p_i appears in every
        // offset p_j - p_i of its own sums (the same per-pair terms
        // k_adjoint_gather adds to p_j)
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
                    if (jj[u] >= a.n) continue;
                    const V3 off = xyz(q[u]) - p;
                    bool anyP, ali;
                    const V3 offbar = pair_terms(sp, flags, f, off, Sbar, Cbar, anyP, ali);
                    if (anyP) pbar = pbar - offbar;
                }
            }
        a.own[3u * i + 0u] = make_float4(pbar.x, pbar.y, pbar.z, 0.0f);
        a.own[3u * i + 1u] = make_float4(fbar.x, fbar.y, fbar.z, sbar);
        a.own[3u * i + 2u] = make_float4(abar.x, abar.y, abar.z, 0.0f);
        a.bars[3u * i + 0u] = make_float4(Sbar.x, Sbar.y, Sbar.z, Cbar.x);
        a.bars[3u * i + 1u] = make_float4(Cbar.y, Cbar.z, Abar.x, Abar.y);
        a.bars[3u * i + 2u] = make_float4(Abar.z, __uint_as_float(flags), 0.0f, 0.0f);
    } else if (i < a.n) {
        // a hash entry without a valid ID: no terms
        const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
        for (uint32_t k = 0; k < 3u; ++k) {
            a.own[3u * i + k] = z;
            a.bars[3u * i + k] = z;
        }
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
}

// Lanes per agent of k_adjoint_gather: lane l takes entries l, l+L, ... of the
// agent's own list (or neighbour cells l, l+L, ... of the 27 when the list
// cannot stand in for them).
#ifndef BRT_ADJ_GATHER_LANES
#define BRT_ADJ_GATHER_LANES 4
#endif
constexpr uint32_t kGatherLanes = BRT_ADJ_GATHER_LANES;
static_assert(kGatherLanes <= 64 && (kGatherLanes & (kGatherLanes - 1)) == 0, "lanes per agent: a power of 2");

struct GatherArgs {
    brt_grid_params g;
    brt_steer_params sp;
    uint32_t K, W, n;
    float T;        // the list test's squared-radius threshold
    uint32_t knn;   // lists built by replacement (useKnn)
    const float4* spos;
    const float4* sfwd;
    const uint2* hash;
    const uint32_t* lists;
    const uint32_t* cell_start;
    const uint32_t* cell_end;
    const float4* own;
    const float4* bars;
    float4* gpos_in;
    float4* gfwd_in;
    float4* gsa_in;
    const float* partials;
    uint32_t partial_blocks;
    float* weight_grad;
};

__device__ __forceinline__ I3 key_cell(uint32_t key, const brt_grid_params& g) {
    const uint32_t gx = g.grid_size[0], gy = g.grid_size[1];
    const uint32_t xy = key % (gx * gy);
    return I3{(int)(xy % gx), (int)(xy / gx), (int)(key / (gx * gy))};
}

// The ordinal of neighbour cell c of an agent in cell gi, in the visiting
// order (x offset outermost, z innermost).
__device__ __forceinline__ int cell_ordinal(I3 c, I3 gi) {
    return 9 * (c.x - gi.x + 1) + 3 * (c.y - gi.y + 1) + (c.z - gi.z + 1);
}

// Whether sorted agent j (key cell cj) is in agent i's list, given that j is
// visible to list tests, is not i and passes the radius test against i: j must
// lie in one of i's 27 cells (known when i lies inside the grid and was found
// in a cell next to j's: near); then it is in the list when the list is not
// full (every candidate that passed is), or, for a full list, when j comes no
// later than the last entry in i's visiting order (cell ordinal, then sorted
// index: the list is the first K that passed), or, for KNN lists
// (replacement), when the row holds j.
// (cj: j's key cell; for j inside the grid it is j's own cell, which from_pos
// computes from Pj only when needed.)
__device__ __forceinline__ bool in_list(const GatherArgs& a, uint32_t i, uint32_t flags_i, float4 Pi, uint32_t j,
                                        I3 cj, bool from_pos, float4 Pj, bool near) {
    const bool full = (flags_i & kAdjFull) != 0u;
    if (near && (flags_i & kAdjInGrid) && !full) return true;
    if (from_pos) cj = calcGridPos(Pj.x, Pj.y, Pj.z, a.g);
    const I3 gi = calcGridPos(Pi.x, Pi.y, Pi.z, a.g);
    if (abs(cj.x - gi.x) > 1 || abs(cj.y - gi.y) > 1 || abs(cj.z - gi.z) > 1) return false;
    if (!full) return true;
    const uint32_t* row = a.lists + (size_t)i * a.W;
    if (a.knn) {
        for (uint32_t t = 1; t <= a.K; ++t)
            if (row[t] == j) return true;
        return false;
    }
    const uint32_t last = row[a.K];
    if (last >= a.n) return false;
    const int oj = cell_ordinal(cj, gi), ol = cell_ordinal(key_cell(a.hash[last].x, a.g), gi);
    return oj < ol || (oj == ol && j <= last);
}

// Adds the terms agent i's list holds for p_j and f_j (off = p_j - p_i).
__device__ __forceinline__ void add_pair(const GatherArgs& a, uint32_t i, float4 B2, V3 off, V3& accP, V3& accF) {
    const float4 B0 = a.bars[3u * i + 0u], B1 = a.bars[3u * i + 1u];
    bool anyP, ali;
    const V3 offbar = pair_terms(a.sp, __float_as_uint(B2.y), xyz(a.sfwd[i]), off, v3(B0.x, B0.y, B0.z),
                                 v3(B0.w, B1.x, B1.y), anyP, ali);
    if (anyP) accP = accP + offbar;
    if (ali) accF = accF + v3(B1.z, B1.w, B2.x);
}

// The key coordinate, on one axis, of the cells whose agents can hold in their
// lists an agent of key coordinate c: c + d for d = -1, 0, 1 inside the grid,
// except that for c = 0 the offset -1 stands for coordinate G-1, where the
// hash keys agents whose coordinate lies below the grid (calcGridHash's
// unsigned clamp); -1: no such cell.
__device__ __forceinline__ int gather_axis(int c, int d, uint32_t G) {
    if (d < 0 && c == 0) return (int)G - 1 > 1 ? (int)G - 1 : -1;
    const int k = c + d;
    return (k >= 0 && k < (int)G) ? k : -1;
}

// dL/d(p_j, f_j, s_j, a_j) by agent ID: agent j's own terms plus the terms of
// every agent i whose list holds j.  When j lies in the grid off its low faces,
// its list is not full and maxBodiesPerCell cuts no cell, those agents are all
// in j's own list (the radius test is symmetric: d2 has the same bits both
// ways), so the lanes walk that list; otherwise they scan the 27 cells around
// j's cell.  Each lane
// sums its part in order and the lanes' sums are combined by a fixed
// butterfly.  Workgroup 0 first adds the weight-gradient partials of
// k_adjoint_own in a fixed order.
__global__ void __launch_bounds__(256) k_adjoint_gather(GatherArgs a) {
    if (blockIdx.x == 0) reduce_partials(a.partials, a.partial_blocks, a.weight_grad);
    constexpr uint32_t L = kGatherLanes;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t j = t / L, l = t % L;
    const brt_steer_params& sp = a.sp;
    const brt_grid_params& g = a.g;
    V3 accP = v3(0.f, 0.f, 0.f), accF = accP;
    uint32_t id = a.n;
    if (j < a.n) {
        const uint2 hj = a.hash[j];
        id = hj.y;
        const uint32_t key = hj.x;
        const bool behaviours = sp.separation[0] != 0.0f || sp.alignment[0] != 0.0f || sp.cohesion[0] != 0.0f;
        const bool unlimited = g.max_bodies_per_cell == 0xFFFFFFFFu;
        // maxBodiesPerCell hides agents past the first ones of their cell from
        // every list test
        const bool visible = key < g.num_cells && (unlimited || j - a.cell_start[key] < g.max_bodies_per_cell);
        if (behaviours && id < a.n && visible) {
            const float4 Pj = a.spos[j];
            const V3 pj = xyz(Pj);
            const uint32_t fj = __float_as_uint(a.bars[3u * j + 2u].y);
            if (unlimited && (fj & kAdjInGrid) && !(fj & (kAdjFull | kAdjLowFace))) {
                const uint32_t* row = a.lists + (size_t)j * a.W;
                const uint32_t cnt = row[0];
                for (uint32_t e = l; e < cnt; e += L) {
                    const uint32_t i = row[1u + e];
                    if (i >= a.n) continue;
                    const float4 Pi = a.spos[i];
                    const float4 B2 = a.bars[3u * i + 2u];
                    if (!in_list(a, i, __float_as_uint(B2.y), Pi, j, I3{0, 0, 0}, true, Pj, true)) continue;
                    add_pair(a, i, B2, pj - xyz(Pi), accP, accF);
                }
            } else {
                const I3 cj = key_cell(key, g);
                for (uint32_t c = l; c < 27u; c += L) {
                    const int nx = gather_axis(cj.x, (int)(c / 9u) - 1, g.grid_size[0]);
                    const int ny = gather_axis(cj.y, (int)((c / 3u) % 3u) - 1, g.grid_size[1]);
                    const int nz = gather_axis(cj.z, (int)(c % 3u) - 1, g.grid_size[2]);
                    if (nx < 0 || ny < 0 || nz < 0) continue;
                    const bool near = abs(nx - cj.x) <= 1 && abs(ny - cj.y) <= 1 && abs(nz - cj.z) <= 1;
                    const uint32_t nk = cellKey(nx, ny, nz, g);
                    const uint32_t s = a.cell_start[nk];
                    if (s == BRT_EMPTY_CELL) continue;
                    const uint32_t e = a.cell_end[nk];
                    for (uint32_t i = s; i < e; ++i) {
                        if (i == j) continue;
                        const float4 Pi = a.spos[i];
                        // the list test of agent i for candidate j (d2 as k_prox forms it)
                        const V3 off = pj - xyz(Pi);
                        if (dot(off, off) > a.T) continue;
                        const float4 B2 = a.bars[3u * i + 2u];
                        if (!in_list(a, i, __float_as_uint(B2.y), Pi, j, cj, false, Pj, near)) continue;
                        add_pair(a, i, B2, off, accP, accF);
                    }
                }
            }
        }
    }
#pragma unroll
    for (uint32_t o = 1; o < L; o <<= 1) {
        accP.x += __shfl_xor(accP.x, (int)o);
        accP.y += __shfl_xor(accP.y, (int)o);
        accP.z += __shfl_xor(accP.z, (int)o);
        accF.x += __shfl_xor(accF.x, (int)o);
        accF.y += __shfl_xor(accF.y, (int)o);
        accF.z += __shfl_xor(accF.z, (int)o);
    }
    if (j < a.n && l == 0u && id < a.n) {
        const float4 op = a.own[3u * j + 0u], of = a.own[3u * j + 1u], oa = a.own[3u * j + 2u];
        a.gpos_in[id] = make_float4(op.x + accP.x, op.y + accP.y, op.z + accP.z, 0.0f);
        a.gfwd_in[id] = make_float4(of.x + accF.x, of.y + accF.y, of.z + accF.z, of.w);
        a.gsa_in[id] = oa;
    }
}

}  // namespace

hipError_t launch_adjoint(const brt_grid_params& g, const brt_prox_params& p, float threshold,
                          const brt_steer_params& sp, float dt, uint32_t n, const float4* spos, const float4* sfwd,
                          const float4* ssa, const uint2* hash, const uint32_t* lists, const uint32_t* cell_start,
                          const uint32_t* cell_end, const float4* gpos_out, const float4* gfwd_out,
                          const float4* gsa_out, float4* gpos_in, float4* gfwd_in, float4* gsa_in, float4* scratch,
                          float* partials, float* weight_grad, uint32_t xcd_remap, hipStream_t s) {
    const uint32_t blocks = (n + 255) / 256;
    AdjArgs a{};
    a.g = g;
    a.sp = sp;
    a.dt = dt;
    a.K = p.max_neighbors;
    a.W = list_stride(p.max_neighbors);
    a.n = n;
    a.spos = spos;
    a.sfwd = sfwd;
    a.ssa = ssa;
    a.hash = hash;
    a.lists = lists;
    a.gpos_out = gpos_out;
    a.gfwd_out = gfwd_out;
    a.gsa_out = gsa_out;
    a.own = scratch;
    a.bars = scratch + 3u * (size_t)n;
    a.partials = partials;
    a.xcd_remap = xcd_remap;
    hipLaunchKernelGGL(k_adjoint_own, dim3(blocks), dim3(256), 0, s, a);
    GatherArgs b{};
    b.g = g;
    b.sp = sp;
    b.K = p.max_neighbors;
    b.W = a.W;
    b.n = n;
    b.T = threshold;
    b.knn = p.use_knn ? 1u : 0u;
    b.spos = spos;
    b.sfwd = sfwd;
    b.hash = hash;
    b.lists = lists;
    b.cell_start = cell_start;
    b.cell_end = cell_end;
    b.own = a.own;
    b.bars = a.bars;
    b.gpos_in = gpos_in;
    b.gfwd_in = gfwd_in;
    b.gsa_in = gsa_in;
    b.partials = partials;
    b.partial_blocks = blocks;
    b.weight_grad = weight_grad;
    const size_t threads = (size_t)n * kGatherLanes;
    hipLaunchKernelGGL(k_adjoint_gather, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, b);
    return hipGetLastError();
}

}  // namespace brt
