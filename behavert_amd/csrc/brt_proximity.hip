// brt_proximity.hip -- Proximity3D on gfx950: capped 27-cell neighbour lists.
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
// The distance test is done on d2 = (dx*dx + dy*dy) + dz*dz (no FMA) against the
// largest float T with sqrtf(T) <= r: for a correctly rounded, monotone sqrtf,
// "sqrtf(d2) > r" and "d2 > T" agree for every d2, NaN included.
#include "brt_device.hpp"
#include "brt_internal.hpp"

namespace brt {

namespace {

template <bool KNN>
__global__ void k_neighbors(brt_grid_params g, uint32_t K, uint32_t W, float T, float r,
                            const float4* __restrict__ spos, const uint2* __restrict__ hash,
                            const uint32_t* __restrict__ cell_start, const uint32_t* __restrict__ cell_end,
                            uint32_t n, uint32_t* __restrict__ lists) {
    extern __shared__ uint32_t s_mem[];
    const uint32_t B = blockDim.x;
    uint32_t* s_list = s_mem;                                   // B * W
    float* s_dist = reinterpret_cast<float*>(s_mem + B * W);    // B * K (KNN only)
    const uint32_t tid = threadIdx.x;
    const uint32_t i = blockIdx.x * B + tid;
    uint32_t* my = s_list + tid * W;
    uint32_t cnt = 0;
    if (i < n) {
        const float4 p = spos[i];
        const I3 gp = calcGridPos(p.x, p.y, p.z, g);
        const uint32_t maxBodies = g.max_bodies_per_cell;
        float maxDist = 0.0f;
        for (int dx = -1; dx <= 1; ++dx) {
            for (int dy = -1; dy <= 1; ++dy) {
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
                        if (!KNN) {
                            my[1 + cnt] = j;
                            ++cnt;
                            if (cnt == K) goto done;
                        } else {
                            const float dist = sqrtf(d2);
                            if (cnt < K) {
                                my[1 + cnt] = j;
                                s_dist[tid * K + cnt] = dist;
                                maxDist = fmaxf(maxDist, dist);
                                ++cnt;
                            } else if (dist < maxDist) {
                                maxDist = s_dist[tid * K];
                                uint32_t mi = 0;
                                for (uint32_t q2 = 1; q2 < cnt; ++q2) {
                                    const float dq = s_dist[tid * K + q2];
                                    if (maxDist < dq) {
                                        maxDist = dq;
                                        mi = q2;
                                    }
                                }
                                my[1 + mi] = j;
                                s_dist[tid * K + mi] = dist;
                            }
                        }
                    }
                }
            }
        }
    }
done:
    my[0] = cnt;
    for (uint32_t q = cnt + 1; q < W; ++q) my[q] = BRT_LIST_PAD;
    __syncthreads();
    // coalesced write-out of the block's rows (uint4 words; W is a multiple of 4)
    const uint32_t row0 = blockIdx.x * B;
    const uint32_t rows = (n - row0) < B ? (n - row0) : B;
    const uint32_t words4 = rows * W / 4u;
    uint4* dst = reinterpret_cast<uint4*>(lists + (size_t)row0 * W);
    const uint4* src = reinterpret_cast<const uint4*>(s_list);
    for (uint32_t w = tid; w < words4; w += B) dst[w] = src[w];
}

}  // namespace

hipError_t launch_neighbors(const brt_grid_params& g, const brt_prox_params& p, float threshold,
                            const float4* spos, const uint2* hash, const uint32_t* cell_start,
                            const uint32_t* cell_end, uint32_t n, uint32_t* lists, hipStream_t s) {
    const uint32_t K = p.max_neighbors;
    const uint32_t W = list_stride(K);
    if (p.use_knn) {
        uint32_t B = 64;
        while (B > 8 && (size_t)B * (W + K) * 4u > 60u * 1024u) B /= 2;
        const size_t lds = (size_t)B * (W + K) * 4u;
        hipLaunchKernelGGL(k_neighbors<true>, dim3((n + B - 1) / B), dim3(B), lds, s, g, K, W, threshold,
                           p.search_radius, spos, hash, cell_start, cell_end, n, lists);
    } else {
        uint32_t B = 256;
        while (B > 64 && (size_t)B * W * 4u > 40u * 1024u) B /= 2;
        const size_t lds = (size_t)B * W * 4u;
        hipLaunchKernelGGL(k_neighbors<false>, dim3((n + B - 1) / B), dim3(B), lds, s, g, K, W, threshold,
                           p.search_radius, spos, hash, cell_start, cell_end, n, lists);
    }
    return hipGetLastError();
}

}  // namespace brt
