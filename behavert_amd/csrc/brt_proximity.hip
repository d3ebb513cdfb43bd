// brt_proximity.hip -- Proximity3D on gfx950: capped 27-cell neighbour lists,
// optionally fused with the OpenSteerWrapper steering + integration pass.
//
// Reference semantics (genericCalcNeighgborhoodD, PlugIns/Proximity3D/Proximity3D_kernel.cu:89-125;
// checkForNeighborsInCell / checkForNeighbor, Proximity3D_resources.cu:59-172):
//   * the agent's cell is calcGridPos(pos) WITHOUT clamping;
//   * cells are visited x-offset outermost, then y, then z innermost, each -1..1,
//     skipping cells outside the grid and empty cells;
//   * inside a cell, agents are visited in sorted order, at most
//     maxBodiesPerCell of them, the agent itself skipped;
//   * a candidate is rejected when sqrtf(dx*dx+dy*dy+dz*dz) > r, dx = other - self;
//   * accepted sorted indices are appended until K are held (later ones are
//     ignored; with useKnn the farthest entry is replaced instead).
// Output layout: the reader layout W = 4*(K/4+1) uints per agent
// (Proximity3D_resources.cu:175-198), [0] = count, then the indices, then padding.
//
// Exactness: the distance test is d2 = (dx*dx + dy*dy) + dz*dz (no FMA) against
// T, the largest float with sqrtf(T) <= r; for a correctly rounded, monotone
// sqrtf, "sqrtf(d2) > r" and "d2 > T" agree for every d2, NaN included.
//
// MI355X design.  One lane per sorted agent, 256 lanes per workgroup; lists are
// staged in LDS and written out with 16-byte coalesced stores.  Per wave64:
//   * uniform mode (few distinct cells in the wave, the dense case): lanes are
//     grouped by cell; a group walks the shared canonical candidate stream with
//     wave-uniform indices, so candidate positions come through the scalar
//     cache (s_load) instead of 64 per-lane vector loads, four at a time;
//   * vector mode (many cells, the sparse case): each lane walks its own stream,
//     four candidate loads in flight per round.
// Both modes skip a cell whose box lies farther than r from the agent (with
// rounding margins so no candidate that passes the exact test is ever skipped),
// only for interior cells: boundary cells can hold clamped out-of-grid or NaN
// agents whose positions are not inside the cell's box.
#include "brt_device.hpp"
#include "brt_internal.hpp"
#include "brt_steer_common.hpp"

namespace brt {

namespace {

// Loads through the constant address space: with a wave-uniform index these
// become scalar-cache loads (s_load_dwordx4).  The arrays are not written by
// the kernel that reads them this way.
typedef float vfloat4 __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(4))) vfloat4* cfloat4_ptr;
typedef const __attribute__((address_space(4))) uint32_t* cuint_ptr;

__device__ __forceinline__ float4 cload(cfloat4_ptr p, uint32_t i) {
    const vfloat4 v = p[i];
    return make_float4(v.x, v.y, v.z, v.w);
}

struct ProxArgs {
    brt_grid_params g;
    uint32_t K, W, n;
    float T;          // exact squared-radius threshold
    float Tskip;      // T * (1 + 1e-5): cell-skip threshold on the box distance
    float slack[3];   // per-axis widening of cell boxes for rounding
    uint32_t uniform_max_cells;
    uint32_t xcd_remap;
    const float4* spos;
    const float4* sfwd;
    const float4* ssa;
    const uint2* hash;
    const uint32_t* cell_start;
    const uint32_t* cell_end;
    uint32_t* lists;
    // fused steering
    brt_steer_params sp;
    float dt;
    uint32_t step;
    AgentOut o;
};

struct Lane {
    float px, py, pz;
    uint32_t i;
    uint32_t cnt;
    uint32_t* row;  // LDS list row
};

__device__ __forceinline__ bool interior(int cx, int cy, int cz, const brt_grid_params& g) {
    return cx >= 1 && cy >= 1 && cz >= 1 && cx <= (int)g.grid_size[0] - 2 && cy <= (int)g.grid_size[1] - 2 &&
           cz <= (int)g.grid_size[2] - 2;
}

// squared distance from (px,py,pz) to the box of cell (cx,cy,cz), each axis gap
// shrunk by the rounding slack
__device__ __forceinline__ float box_d2(const ProxArgs& a, float px, float py, float pz, int cx, int cy, int cz) {
    const brt_grid_params& g = a.g;
    float lo, gx, gy, gz;
    lo = g.world_origin[0] + (float)cx * g.cell_size[0];
    gx = fmaxf(fmaxf(lo - px, px - (lo + g.cell_size[0])) - a.slack[0], 0.0f);
    lo = g.world_origin[1] + (float)cy * g.cell_size[1];
    gy = fmaxf(fmaxf(lo - py, py - (lo + g.cell_size[1])) - a.slack[1], 0.0f);
    lo = g.world_origin[2] + (float)cz * g.cell_size[2];
    gz = fmaxf(fmaxf(lo - pz, pz - (lo + g.cell_size[2])) - a.slack[2], 0.0f);
    return (gx * gx + gy * gy) + gz * gz;
}

__device__ __forceinline__ void accept(Lane& L, uint32_t j) {
    L.row[1 + L.cnt] = j;
    ++L.cnt;
}

// Uniform mode: every active lane shares cell (ux,uy,uz); control flow and
// candidate indices are wave-uniform.
__device__ void scan_uniform(const ProxArgs& a, Lane& L, int ux, int uy, int uz) {
    const brt_grid_params& g = a.g;
    cfloat4_ptr cpos = (cfloat4_ptr)a.spos;
    cuint_ptr cstart = (cuint_ptr)a.cell_start;
    cuint_ptr cend = (cuint_ptr)a.cell_end;
    const uint32_t maxBodies = g.max_bodies_per_cell;
    for (int dx = -1; dx <= 1; ++dx) {
        for (int dy = -1; dy <= 1; ++dy) {
            for (int dz = -1; dz <= 1; ++dz) {
                const int cx = ux + dx, cy = uy + dy, cz = uz + dz;
                if (!inGrid(cx, cy, cz, g)) continue;
                const uint32_t key = cellKey(cx, cy, cz, g);
                const uint32_t s = cstart[key];
                if (s == BRT_EMPTY_CELL) continue;
                const uint32_t e = cend[key];
                uint32_t lim = (e - s) > maxBodies ? s + maxBodies : e;
                lim = lim < a.n ? lim : a.n;
                if (interior(cx, cy, cz, g)) {
                    const bool need = L.cnt < a.K && !(box_d2(a, L.px, L.py, L.pz, cx, cy, cz) > a.Tskip);
                    if (__ballot(need) == 0) continue;
                }
                for (uint32_t j = s; j < lim; j += 4) {
                    float4 q[4];
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        const uint32_t jj = (j + t < lim) ? j + t : lim - 1;
                        q[t] = cload(cpos, jj);
                    }
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        if (j + t < lim && j + t != L.i && L.cnt < a.K) {
                            const float rx = q[t].x - L.px, ry = q[t].y - L.py, rz = q[t].z - L.pz;
                            const float xx = rx * rx, yy = ry * ry, zz = rz * rz;
                            const float d2 = (xx + yy) + zz;
                            if (!(d2 > a.T)) accept(L, j + t);
                        }
                    }
                    if (__ballot(L.cnt < a.K) == 0) return;
                }
            }
        }
    }
}

// Vector mode: each lane walks its own canonical stream.
__device__ void scan_vector(const ProxArgs& a, Lane& L, int gx, int gy, int gz) {
    const brt_grid_params& g = a.g;
    const uint32_t maxBodies = g.max_bodies_per_cell;
    for (int dx = -1; dx <= 1; ++dx) {
        for (int dy = -1; dy <= 1; ++dy) {
            for (int dz = -1; dz <= 1; ++dz) {
                const int cx = gx + dx, cy = gy + dy, cz = gz + dz;
                if (!inGrid(cx, cy, cz, g)) continue;
                const uint32_t key = cellKey(cx, cy, cz, g);
                const uint32_t s = a.cell_start[key];
                if (s == BRT_EMPTY_CELL) continue;
                if (interior(cx, cy, cz, g) && box_d2(a, L.px, L.py, L.pz, cx, cy, cz) > a.Tskip) continue;
                const uint32_t e = a.cell_end[key];
                uint32_t lim = (e - s) > maxBodies ? s + maxBodies : e;
                lim = lim < a.n ? lim : a.n;
                for (uint32_t j = s; j < lim; j += 4) {
                    float4 q[4];
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        const uint32_t jj = (j + t < lim) ? j + t : lim - 1;
                        q[t] = a.spos[jj];
                    }
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        if (j + t < lim && j + t != L.i) {
                            const float rx = q[t].x - L.px, ry = q[t].y - L.py, rz = q[t].z - L.pz;
                            const float xx = rx * rx, yy = ry * ry, zz = rz * rz;
                            const float d2 = (xx + yy) + zz;
                            if (!(d2 > a.T)) {
                                accept(L, j + t);
                                if (L.cnt == a.K) return;
                            }
                        }
                    }
                }
            }
        }
    }
}

template <bool FUSE>
__global__ void __launch_bounds__(256) k_prox(ProxArgs a) {
    extern __shared__ uint32_t s_list[];  // blockDim.x * W
    const uint32_t tid = threadIdx.x;
    const uint32_t B = blockDim.x;
    const uint32_t blk = xcd_block(a.xcd_remap != 0);
    Lane L;
    L.i = blk * B + tid;
    L.cnt = 0;
    L.row = s_list + tid * a.W;
    const bool alive = L.i < a.n;
    float4 P = make_float4(0.f, 0.f, 0.f, 0.f);
    if (alive) P = a.spos[L.i];
    L.px = P.x;
    L.py = P.y;
    L.pz = P.z;
    I3 gp = calcGridPos(P.x, P.y, P.z, a.g);

    // wave-level mode choice: can the wave's lanes be covered by a few cells?
    const uint64_t alive_mask = __ballot(alive);
    uint64_t pending = alive_mask;
    uint32_t groups = 0;
    while (pending != 0 && groups < a.uniform_max_cells) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const int ux = __builtin_amdgcn_readlane(gp.x, leader);
        const int uy = __builtin_amdgcn_readlane(gp.y, leader);
        const int uz = __builtin_amdgcn_readlane(gp.z, leader);
        pending &= ~__ballot(alive && gp.x == ux && gp.y == uy && gp.z == uz);
        ++groups;
    }
    if (pending == 0) {
        pending = alive_mask;
        while (pending != 0) {
            const int leader = __ffsll((unsigned long long)pending) - 1;
            const int ux = __builtin_amdgcn_readlane(gp.x, leader);
            const int uy = __builtin_amdgcn_readlane(gp.y, leader);
            const int uz = __builtin_amdgcn_readlane(gp.z, leader);
            const bool mine = alive && gp.x == ux && gp.y == uy && gp.z == uz;
            pending &= ~__ballot(mine);
            if (mine) scan_uniform(a, L, ux, uy, uz);
        }
    } else if (alive) {
        scan_vector(a, L, gp.x, gp.y, gp.z);
    }

    L.row[0] = L.cnt;
    for (uint32_t q = L.cnt + 1; q < a.W; ++q) L.row[q] = BRT_LIST_PAD;
    __syncthreads();
    const uint32_t row0 = blk * B;
    const uint32_t rows = (a.n - row0) < B ? (a.n - row0) : B;
    const uint32_t words4 = rows * a.W / 4u;
    uint4* dst = reinterpret_cast<uint4*>(a.lists + (size_t)row0 * a.W);
    const uint4* src = reinterpret_cast<const uint4*>(s_list);
    for (uint32_t w = tid; w < words4; w += B) dst[w] = src[w];

    // fused steering: walk the finished list from LDS (all lanes in step, four
    // neighbour gathers in flight), then integrate.  Same order and operations
    // as k_steer over the written list.
    if (FUSE && alive) {
        const uint32_t id = a.hash[L.i].y;
        if (id < a.n) {
            const float4 Fw = a.sfwd[L.i];
            const float4 SA = a.ssa[L.i];
            const V3 myPos = v3(P.x, P.y, P.z), myFwd = xyz(Fw);
            SteerAccum acc;
            acc.init();
            if (a.sp.separation[0] != 0.0f || a.sp.alignment[0] != 0.0f || a.sp.cohesion[0] != 0.0f) {
                for (uint32_t t = 0; t < L.cnt; t += 4) {
                    float4 q[4], f[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const uint32_t j = (t + u < L.cnt) ? L.row[1 + t + u] : L.i;
                        q[u] = a.spos[j];
                        f[u] = a.sfwd[j];
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        if (t + u < L.cnt) {
                            const V3 offset = xyz(q[u]) - myPos;
                            acc.add(a.sp, myFwd, offset, dot(offset, offset), xyz(f[u]));
                        }
                    }
                }
            }
            finish_agent(a.g, a.sp, a.dt, a.step, L.i, id, myPos, myFwd, Fw.w, xyz(SA), SA.w, acc, a.o);
        }
    }
}

// KNN mode (useKnn): the replacement depends on every candidate in order, so a
// plain per-lane scan with the distance list staged in LDS.
__global__ void k_neighbors_knn(brt_grid_params g, uint32_t K, uint32_t W, float T, const float4* __restrict__ spos,
                                const uint32_t* __restrict__ cell_start, const uint32_t* __restrict__ cell_end,
                                uint32_t n, uint32_t* __restrict__ lists) {
    extern __shared__ uint32_t s_mem[];
    const uint32_t B = blockDim.x;
    uint32_t* s_list = s_mem;
    float* s_dist = reinterpret_cast<float*>(s_mem + B * W);
    const uint32_t tid = threadIdx.x;
    const uint32_t i = blockIdx.x * B + tid;
    uint32_t* my = s_list + tid * W;
    float* md = s_dist + tid * K;
    uint32_t cnt = 0;
    if (i < n) {
        const float4 p = spos[i];
        const I3 gp = calcGridPos(p.x, p.y, p.z, g);
        const uint32_t maxBodies = g.max_bodies_per_cell;
        float maxDist = 0.0f;
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz) {
                    const int cx = gp.x + dx, cy = gp.y + dy, cz = gp.z + dz;
                    if (!inGrid(cx, cy, cz, g)) continue;
                    const uint32_t key = cellKey(cx, cy, cz, g);
                    const uint32_t s = cell_start[key];
                    if (s == BRT_EMPTY_CELL) continue;
                    const uint32_t e = cell_end[key];
                    uint32_t lim = (e - s) > maxBodies ? s + maxBodies : e;
                    lim = lim < n ? lim : n;
                    for (uint32_t j = s; j < lim; ++j) {
                        if (j == i) continue;
                        const float4 q = spos[j];
                        const float rx = q.x - p.x, ry = q.y - p.y, rz = q.z - p.z;
                        const float xx = rx * rx, yy = ry * ry, zz = rz * rz;
                        const float d2 = (xx + yy) + zz;
                        if (d2 > T) continue;
                        const float dist = sqrtf(d2);
                        if (cnt < K) {
                            my[1 + cnt] = j;
                            md[cnt] = dist;
                            maxDist = fmaxf(maxDist, dist);
                            ++cnt;
                        } else if (dist < maxDist) {
                            maxDist = md[0];
                            uint32_t mi = 0;
                            for (uint32_t q2 = 1; q2 < cnt; ++q2)
                                if (maxDist < md[q2]) {
                                    maxDist = md[q2];
                                    mi = q2;
                                }
                            my[1 + mi] = j;
                            md[mi] = dist;
                        }
                    }
                }
    }
    my[0] = cnt;
    for (uint32_t q = cnt + 1; q < W; ++q) my[q] = BRT_LIST_PAD;
    __syncthreads();
    const uint32_t row0 = blockIdx.x * B;
    const uint32_t rows = (n - row0) < B ? (n - row0) : B;
    const uint32_t words4 = rows * W / 4u;
    uint4* dst = reinterpret_cast<uint4*>(lists + (size_t)row0 * W);
    const uint4* src = reinterpret_cast<const uint4*>(s_list);
    for (uint32_t w = tid; w < words4; w += B) dst[w] = src[w];
}

// largest workgroup (256, 128 or 64 lanes) whose list staging fits in 40 KB of
// LDS (four workgroups per CU), 64 lanes at least
uint32_t prox_block(uint32_t W) {
    uint32_t B = 256;
    while (B > 64 && (size_t)B * W * 4u > 40u * 1024u) B /= 2;
    return B;
}

ProxArgs make_args(const brt_grid_params& g, const brt_prox_params& p, float threshold, const float4* spos,
                   const uint2* hash, const uint32_t* cell_start, const uint32_t* cell_end, uint32_t n,
                   uint32_t* lists, uint32_t uniform_max_cells, uint32_t xcd_remap) {
    ProxArgs a{};
    a.g = g;
    a.K = p.max_neighbors;
    a.W = list_stride(p.max_neighbors);
    a.n = n;
    a.T = threshold;
    // the skip threshold errs toward scanning: 1e-5 relative on d2 and a
    // per-axis widening of 1e-5 of the grid's coordinate range, both far above
    // the few-ulp rounding of positions, cell bounds and d2
    a.Tskip = threshold * 1.00001f;
    for (int ax = 0; ax < 3; ++ax)
        a.slack[ax] = 1e-5f * (fabsf(g.world_origin[ax]) + (float)g.grid_size[ax] * g.cell_size[ax]);
    a.uniform_max_cells = uniform_max_cells;
    a.xcd_remap = xcd_remap;
    a.spos = spos;
    a.hash = hash;
    a.cell_start = cell_start;
    a.cell_end = cell_end;
    a.lists = lists;
    return a;
}

}  // namespace

hipError_t launch_neighbors(const brt_grid_params& g, const brt_prox_params& p, float threshold,
                            const float4* spos, const uint2* hash, const uint32_t* cell_start,
                            const uint32_t* cell_end, uint32_t n, uint32_t* lists, uint32_t uniform_max_cells,
                            uint32_t xcd_remap, hipStream_t s) {
    const uint32_t K = p.max_neighbors;
    const uint32_t W = list_stride(K);
    if (p.use_knn) {
        uint32_t B = 64;
        while (B > 8 && (size_t)B * (W + K) * 4u > 60u * 1024u) B /= 2;
        const size_t lds = (size_t)B * (W + K) * 4u;
        hipLaunchKernelGGL(k_neighbors_knn, dim3((n + B - 1) / B), dim3(B), lds, s, g, K, W, threshold, spos,
                           cell_start, cell_end, n, lists);
        return hipGetLastError();
    }
    ProxArgs a = make_args(g, p, threshold, spos, hash, cell_start, cell_end, n, lists, uniform_max_cells, xcd_remap);
    const uint32_t B = prox_block(W);
    hipLaunchKernelGGL(k_prox<false>, dim3((n + B - 1) / B), dim3(B), (size_t)B * W * 4u, s, a);
    return hipGetLastError();
}

hipError_t launch_neighbors_steer(const brt_grid_params& g, const brt_prox_params& p, float threshold,
                                  const brt_steer_params& sp, float dt, uint32_t step, const float4* spos,
                                  const float4* sfwd, const float4* ssa, const uint2* hash,
                                  const uint32_t* cell_start, const uint32_t* cell_end, uint32_t n, uint32_t* lists,
                                  float2* wander, float4* pos_out, float4* fwd_out, float4* sa_out,
                                  uint32_t uniform_max_cells, uint32_t xcd_remap, hipStream_t s) {
    ProxArgs a = make_args(g, p, threshold, spos, hash, cell_start, cell_end, n, lists, uniform_max_cells, xcd_remap);
    a.sfwd = sfwd;
    a.ssa = ssa;
    a.sp = sp;
    a.dt = dt;
    a.step = step;
    a.o = AgentOut{pos_out, fwd_out, sa_out, nullptr, wander};
    const uint32_t B = prox_block(a.W);
    hipLaunchKernelGGL(k_prox<true>, dim3((n + B - 1) / B), dim3(B), (size_t)B * a.W * 4u, s, a);
    return hipGetLastError();
}

}  // namespace brt
