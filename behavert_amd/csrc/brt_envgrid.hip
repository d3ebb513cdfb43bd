// brt_envgrid.hip -- EnvGrid3D on gfx950: cell hash, stable key/ID radix sort,
// reorder of the per-agent float4 features and the cell ranges.
//
// Reference semantics:
//   genCalcHashD ........... PlugIns/EnvGrid3D/envgrid3d_kernel.cu:55-68
//   RadixSort (stable) ..... PlugIns/EnvGrid3D/radixsort.cu:50-80 (4 x 8-bit LSD passes,
//                            values = agent IDs in order, so the result is the
//                            (key, ID) lexicographic order)
//   reorderDataFloat4 ...... PlugIns/EnvGrid3D/envgrid3d_kernel.cu:110-122
//   genericFindCellStartD .. PlugIns/Proximity3D/Proximity3D_kernel.cu:54-84
//
// Design (MI355X): the sort is a onesweep LSD radix sort.  The hash kernel also
// builds the digit histograms of every pass; each pass is then ONE kernel: a tile
// of 4096 pairs ranks its keys per wave64 with 64-bit ballots (a digit "match"
// mask per lane), combines the four waves' counts in LDS, publishes its per-digit
// count and finds its global offset by decoupled look-back over the preceding
// tiles' status words, then scatters through LDS so each digit's run is written
// with consecutive addresses.  Only ceil(log2(cells)/8) passes run: keys are
// < cells, so the higher digits are all zero and cannot change a stable order.
#include "brt_device.hpp"
#include "brt_internal.hpp"

namespace brt {

namespace {

constexpr uint32_t kFlagAgg = 1u << 30;
constexpr uint32_t kFlagInc = 2u << 30;
constexpr uint32_t kValueMask = (1u << 30) - 1u;
// Predecessor statuses loaded together per look-back step (onesweep pass at
// 1M keys: 19.4 us with 1, 18.3 with 4, 18.4 with 8, 19.3 with 16; at 4M keys
// 46.3, 43.5, 44.6 and 48.7 us).
#ifndef BRT_SORT_LOOKBACK
#define BRT_SORT_LOOKBACK 4
#endif
constexpr int kLookback = BRT_SORT_LOOKBACK;

__device__ __forceinline__ uint32_t block_excl_scan_256(uint32_t v, uint32_t* s_tmp) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_tmp[wave] = x;
    __syncthreads();
    uint32_t wo = 0;
    for (int w = 0; w < wave; ++w) wo += s_tmp[w];
    __syncthreads();
    return wo + x - v;
}

// Every cell empty (start = end = BRT_EMPTY_CELL) before the reorder writes
// the occupied ones; four cells per thread with 16-byte stores where aligned.
__device__ __forceinline__ void fill_empty_cells(uint32_t* __restrict__ cell_start, uint32_t* __restrict__ cell_end,
                                                 uint32_t num_cells) {
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t n4 = num_cells / 4u;
    const uint4 e = make_uint4(BRT_EMPTY_CELL, BRT_EMPTY_CELL, BRT_EMPTY_CELL, BRT_EMPTY_CELL);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) {
        reinterpret_cast<uint4*>(cell_start)[i] = e;
        reinterpret_cast<uint4*>(cell_end)[i] = e;
    }
    for (uint32_t i = n4 * 4u + blockIdx.x * blockDim.x + threadIdx.x; i < num_cells; i += stride) {
        cell_start[i] = BRT_EMPTY_CELL;
        cell_end[i] = BRT_EMPTY_CELL;
    }
}

__global__ void __launch_bounds__(256) k_prepare(uint32_t* __restrict__ meta, uint32_t meta_words,
                                                 uint32_t* __restrict__ cell_start,
                                                 uint32_t* __restrict__ cell_end, uint32_t num_cells) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < meta_words; i += stride) meta[i] = 0u;
    if (cell_start) fill_empty_cells(cell_start, cell_end, num_cells);
}

// genCalcHashD fused with the per-pass digit histograms (and, when cell_start
// is given, the empty-cell fill).
__global__ void __launch_bounds__(256) k_hash_hist(brt_grid_params g, const float4* __restrict__ pos,
                                                   uint32_t n, uint2* __restrict__ out,
                                                   uint32_t* __restrict__ hist, SortPlan plan,
                                                   uint32_t* __restrict__ cell_start,
                                                   uint32_t* __restrict__ cell_end) {
    __shared__ uint32_t s_hist[kMaxPasses * kRadix];
    if (cell_start) fill_empty_cells(cell_start, cell_end, g.num_cells);
    for (uint32_t t = threadIdx.x; t < plan.passes * kRadix; t += blockDim.x) s_hist[t] = 0u;
    __syncthreads();
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        float4 p = pos[i];
        uint32_t key = calcGridHash(calcGridPos(p.x, p.y, p.z, g), g);
        out[i] = make_uint2(key, i);
        for (uint32_t q = 0; q < plan.passes; ++q)
            atomicAdd(&s_hist[q * kRadix + ((key >> plan.shift[q]) & ((1u << plan.nbits[q]) - 1u))], 1u);
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < plan.passes * kRadix; t += blockDim.x) {
        uint32_t v = s_hist[t];
        if (v) atomicAdd(&hist[(blockIdx.x % kHistCopies) * kHistWords + t], v);
    }
}

// genCalcHashD over a batch of flocks: each wave hashes up to 64 agents of one
// flock with that flock's grid, the key offset by the flock's first global
// cell, so one stable sort orders every flock and keeps each in its own range.
__global__ void __launch_bounds__(256) k_batch_hash_hist(BatchView b, const float4* __restrict__ pos,
                                                         uint2* __restrict__ out, uint32_t* __restrict__ hist,
                                                         SortPlan plan, uint32_t* __restrict__ cell_start,
                                                         uint32_t* __restrict__ cell_end, uint32_t num_cells) {
    __shared__ uint32_t s_hist[kMaxPasses * kRadix];
    if (cell_start) fill_empty_cells(cell_start, cell_end, num_cells);
    for (uint32_t t = threadIdx.x; t < plan.passes * kRadix; t += blockDim.x) s_hist[t] = 0u;
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t wstride = gridDim.x * (blockDim.x >> 6);
    for (uint32_t gw = blockIdx.x * (blockDim.x >> 6) + wid; gw < b.waves; gw += wstride) {
        const uint32_t f = __builtin_amdgcn_readfirstlane(b.wave_flock[gw]);
        const FlockDev& fd = b.flocks[f];
        const uint32_t i = fd.offset + (gw - fd.wave_base) * 64u + lane;
        if (i < fd.offset + fd.n) {
            const float4 p = pos[i];
            const uint32_t key = fd.cell_base + calcGridHash(calcGridPos(p.x, p.y, p.z, fd.g), fd.g);
            out[i] = make_uint2(key, i);
            for (uint32_t q = 0; q < plan.passes; ++q)
                atomicAdd(&s_hist[q * kRadix + ((key >> plan.shift[q]) & ((1u << plan.nbits[q]) - 1u))], 1u);
        }
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < plan.passes * kRadix; t += blockDim.x) {
        uint32_t v = s_hist[t];
        if (v) atomicAdd(&hist[(blockIdx.x % kHistCopies) * kHistWords + t], v);
    }
}

__global__ void k_set_word(uint32_t* p, uint32_t v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// Histograms of existing pairs (standalone sort); optionally copies the pairs.
__global__ void __launch_bounds__(256) k_hist(const uint2* __restrict__ in, uint2* __restrict__ copy_out,
                                              uint32_t n, uint32_t* __restrict__ hist, SortPlan plan) {
    __shared__ uint32_t s_hist[kMaxPasses * kRadix];
    for (uint32_t t = threadIdx.x; t < plan.passes * kRadix; t += blockDim.x) s_hist[t] = 0u;
    __syncthreads();
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint2 v = in[i];
        if (copy_out) copy_out[i] = v;
        for (uint32_t q = 0; q < plan.passes; ++q)
            atomicAdd(&s_hist[q * kRadix + ((v.x >> plan.shift[q]) & ((1u << plan.nbits[q]) - 1u))], 1u);
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < plan.passes * kRadix; t += blockDim.x) {
        uint32_t v = s_hist[t];
        if (v) atomicAdd(&hist[(blockIdx.x % kHistCopies) * kHistWords + t], v);
    }
}

// One onesweep pass.  Items are held warp-striped: item k of lane l of wave w is
// element tile_base + w*64*ITEMS + k*64 + l, so (w, k, l) order is input order and
// ranking in that order keeps the sort stable.
template <int ITEMS>
__global__ void __launch_bounds__(kSortThreads) k_onesweep(const uint2* __restrict__ in, uint2* __restrict__ out,
                                                            uint32_t n, uint32_t shift, uint32_t nbits,
                                                            const uint32_t* __restrict__ ghist,
                                                            uint32_t* __restrict__ status,
                                                            uint32_t* __restrict__ tile_counter,
                                                            uint32_t* __restrict__ dev_flags,
                                                            uint32_t* __restrict__ prev_hist,
                                                            uint32_t* __restrict__ prev_counter,
                                                            uint32_t* __restrict__ prev_status) {
    constexpr int TILE = kSortThreads * ITEMS;
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_whist[4][kRadix];
    __shared__ uint32_t s_digit_start[kRadix];
    __shared__ uint32_t s_base[kRadix];
    __shared__ uint32_t s_gcount[kRadix];
    __shared__ uint32_t s_tmp[4];
    __shared__ uint2 s_items[TILE];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const uint32_t mask = (1u << nbits) - 1u;

    if (tid == 0) s_tile = atomicAdd(tile_counter, 1u);
#pragma unroll
    for (int w = 0; w < 4; ++w) s_whist[w][tid] = 0u;
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t wbase = tile * TILE + wave * (64 * ITEMS);
    // A tile ID past the grid means the counter did not start at zero (the
    // metadata a step relies on was not returned to zero): no tile is
    // processed out of bounds; the flag reports the wrong sort.
    if (tile >= gridDim.x) {
        if (tid == 0)
            __hip_atomic_fetch_or(dev_flags, BRT_DEVFLAG_SORT_LOOKBACK, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        return;
    }
    // the previous pass's metadata back to zero (its tiles use the same IDs)
    if (prev_status) {
        prev_status[(size_t)tile * kRadix + tid] = 0u;
        if (tile == 0u) {
#pragma unroll
            for (uint32_t c = 0; c < kHistCopies; ++c) prev_hist[c * kHistWords + tid] = 0u;
            if (tid == 0) *prev_counter = 0u;
        }
    }

    uint2 it[ITEMS];
    uint32_t rank[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        uint32_t idx = wbase + k * 64 + lane;
        it[k] = idx < n ? in[idx] : make_uint2(0u, 0u);
    }
    // this pass's global count of digit tid, summed over the histogram copies
    // while the items load, and parked in LDS (not in registers) over the
    // ranking
    {
        uint32_t gc = 0u;
#pragma unroll
        for (uint32_t c = 0; c < kHistCopies; ++c) gc += ghist[c * kHistWords + tid];
        s_gcount[tid] = gc;
    }
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const bool valid = (wbase + k * 64 + lane) < n;
        const uint32_t d = (it[k].x >> shift) & mask;
        const uint64_t vm = __ballot(valid);
        // lanes whose digit equals this lane's: the AND over the eight bits of
        // (ballot of bit b) XNOR (this lane's bit b replicated), on the two
        // 32-bit halves (bits above nbits are zero in every lane and leave the
        // mask unchanged)
        uint32_t m_lo = (uint32_t)vm, m_hi = (uint32_t)(vm >> 32);
#pragma unroll
        for (uint32_t b = 0; b < 8u; ++b) {
            const uint32_t rep = (uint32_t)((int32_t)(d << (31u - b)) >> 31);  // bit b: 0 or ~0
            const uint64_t bal = __ballot(rep != 0u);
            m_lo &= ~((uint32_t)bal ^ rep);
            m_hi &= ~((uint32_t)(bal >> 32) ^ rep);
        }
        const uint64_t m = ((uint64_t)m_hi << 32) | m_lo;
        const uint32_t cnt = __popcll(m);
        const uint32_t below = __popcll(m & lt);
        const int leader = __ffsll((unsigned long long)m) - 1;
        uint32_t prev = 0;
        if (valid) prev = s_whist[wave][d];
        __builtin_amdgcn_wave_barrier();
        if (valid && lane == leader) s_whist[wave][d] = prev + cnt;
        __builtin_amdgcn_wave_barrier();
        rank[k] = prev + below;
    }
    __syncthreads();

    // digit = tid: per-wave exclusive offsets and the tile's count
    const uint32_t c0 = s_whist[0][tid], c1 = s_whist[1][tid], c2 = s_whist[2][tid], c3 = s_whist[3][tid];
    const uint32_t tcount = c0 + c1 + c2 + c3;
    s_whist[0][tid] = 0u;
    s_whist[1][tid] = c0;
    s_whist[2][tid] = c0 + c1;
    s_whist[3][tid] = c0 + c1 + c2;
    uint32_t* my_status = status + (size_t)tile * kRadix + tid;
    if (tile == 0)
        __hip_atomic_store(my_status, kFlagInc | tcount, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
        __hip_atomic_store(my_status, kFlagAgg | tcount, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);

    const uint32_t digit_start = block_excl_scan_256(tcount, s_tmp);
    const uint32_t gex = block_excl_scan_256(s_gcount[tid], s_tmp);
    s_digit_start[tid] = digit_start;

    // decoupled look-back over the preceding tiles for digit tid, kLookback
    // tiles per step: the statuses of tiles p, p-1, ..., p-kLookback+1 are
    // loaded together and taken nearest first (aggregates summed up to the
    // first inclusive prefix), so a walk over m tiles waits for about
    // m / kLookback memory latencies instead of m
    uint32_t excl = 0u;
    if (tile > 0) {
        int32_t p = (int32_t)tile - 1;
        uint32_t spins = 0;
        while (p >= 0) {
            uint32_t v[kLookback];
#pragma unroll
            for (int k = 0; k < kLookback; ++k)
                v[k] = (p - k >= 0) ? __hip_atomic_load(status + (size_t)(p - k) * kRadix + tid, __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT)
                                    : kFlagInc;  // before tile 0: never reached (tile 0 is inclusive)
            uint32_t add = 0u, used = 0u;
            bool inc = false, stall = false;
#pragma unroll
            for (int k = 0; k < kLookback; ++k) {
                const uint32_t flag = v[k] & ~kValueMask;
                if (!inc && !stall) {
                    if (flag == 0u) {
                        stall = true;
                    } else {
                        add += v[k] & kValueMask;
                        used = (uint32_t)k + 1u;
                        inc = flag == kFlagInc;
                    }
                }
            }
            excl += add;
            if (inc) break;
            p -= (int32_t)used;
            if (stall) {
                // A predecessor tile has started (tile IDs are taken in start
                // order) and publishes without waiting, so this wait ends; the
                // bound only turns an impossible stall into a flagged wrong
                // result instead of a hang: the sticky device flag is raised
                // for the host to report.
                if (++spins > (1u << 24)) {
                    __hip_atomic_fetch_or(dev_flags, BRT_DEVFLAG_SORT_LOOKBACK, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_SYSTEM);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __hip_atomic_store(my_status, kFlagInc | (excl + tcount), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    s_base[tid] = gex + excl;
    __syncthreads();

    // local scatter into digit order, then coalesced write-out
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const bool valid = (wbase + k * 64 + lane) < n;
        if (valid) {
            const uint32_t d = (it[k].x >> shift) & mask;
            s_items[s_digit_start[d] + s_whist[wave][d] + rank[k]] = it[k];
        }
    }
    __syncthreads();
    const uint32_t tile_begin = tile * TILE;
    const uint32_t tile_n = (n - tile_begin) < (uint32_t)TILE ? (n - tile_begin) : (uint32_t)TILE;
    if (tile_n == (uint32_t)TILE) {
        // full tile: the ITEMS slots of a thread are independent, so their LDS
        // reads overlap
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) {
            const uint32_t j = (uint32_t)(k * kSortThreads + tid);
            const uint2 v = s_items[j];
            const uint32_t d = (v.x >> shift) & mask;
            const uint32_t g = s_base[d] + (j - s_digit_start[d]);
            if (g < n) out[g] = v;
        }
    } else {
        for (uint32_t j = tid; j < tile_n; j += kSortThreads) {
            const uint2 v = s_items[j];
            const uint32_t d = (v.x >> shift) & mask;
            const uint32_t g = s_base[d] + (j - s_digit_start[d]);
            if (g < n) out[g] = v;
        }
    }
}

// reorderDataFloat4 for every feature + cell start/end of the sorted keys.
// With cleanup, the sort's last pass's metadata goes back to zero as well.
__global__ void __launch_bounds__(256) k_reorder_ranges(const uint2* __restrict__ hash, uint32_t n, ReorderArgs f,
                                                        uint32_t num_cells, uint32_t* __restrict__ cell_start,
                                                        uint32_t* __restrict__ cell_end, SortCleanup cleanup) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (cleanup.hist) {
        const uint32_t stride = gridDim.x * blockDim.x;
        for (uint32_t w = i; w < cleanup.status_words; w += stride) cleanup.status[w] = 0u;
        for (uint32_t w = i; w < kHistCopies * kRadix; w += stride)
            cleanup.hist[(w / kRadix) * kHistWords + (w % kRadix)] = 0u;
        if (i == 0) *cleanup.counter = 0u;
    }
    if (i >= n) return;
    const uint2 h = hash[i];
    if (h.y < n) {
        for (uint32_t q = 0; q < f.count; ++q) {
            const float4 v = f.in[q][h.y];
            f.out[q][i] = v;
            if (q == 0 && f.xyz0) {
                float* d = f.xyz0 + 3u * (size_t)i;
                d[0] = v.x;
                d[1] = v.y;
                d[2] = v.z;
            }
        }
    }
    if (cell_start && h.x < num_cells) {
        if (i == 0 || hash[i - 1].x != h.x) cell_start[h.x] = i;
        if (i == n - 1 || hash[i + 1].x != h.x) cell_end[h.x] = i + 1;
    }
}

// Workgroups of the histogram kernels: each adds its per-workgroup histograms
// of every pass into one of the kHistCopies global copies, one atomic per
// nonzero bin.  With a single copy the contended atomics set the kernel's time
// (1M agents: 12.5 us with 256 workgroups against 24.6 us with 1024); spread
// over 8 copies they no longer do, and 1024 workgroups (capped at one per
// 256 keys) measured fastest (1M agents: 8.6 us; 4M: 24.8 us, from 38.7).
#ifndef BRT_HIST_BLOCKS
#define BRT_HIST_BLOCKS 1024u
#endif
inline uint32_t hist_blocks(uint32_t) { return BRT_HIST_BLOCKS; }

inline uint32_t grid_for(uint32_t n, uint32_t block, uint32_t cap) {
    uint32_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    return g < cap ? g : cap;
}

}  // namespace

SortPlan make_sort_plan(uint32_t n, uint32_t num_cells, uint32_t items) {
    SortPlan p{};
    p.items = items == 8u ? 8u : 16u;
    uint32_t bits = 0;
    while (bits < 32 && ((num_cells - 1u) >> bits)) ++bits;
    p.passes = (bits + 7u) / 8u;
    for (uint32_t q = 0; q < p.passes; ++q) {
        p.shift[q] = 8u * q;
        p.nbits[q] = (bits - 8u * q) < 8u ? (bits - 8u * q) : 8u;
    }
    const uint32_t tile = kSortThreads * p.items;
    p.tiles = (n + tile - 1) / tile;
    return p;
}

size_t sort_meta_words(const SortPlan& p) {
    return (size_t)kMetaStatus + (size_t)p.passes * p.tiles * kRadix;
}

SortCleanup sort_cleanup(uint32_t* meta, const SortPlan& plan) {
    SortCleanup c{};
    if (!meta || plan.passes == 0u) return c;
    const uint32_t last = plan.passes - 1u;
    c.hist = meta + kMetaHist + last * kRadix;
    c.counter = meta + kMetaCounters + last;
    c.status = meta + kMetaStatus + (size_t)last * plan.tiles * kRadix;
    c.status_words = plan.tiles * kRadix;
    return c;
}

LaunchMark& launch_mark() {
    static thread_local LaunchMark m;
    return m;
}

hipError_t launch_prepare(uint32_t* meta, size_t meta_words, uint32_t* cell_start, uint32_t* cell_end,
                          uint32_t num_cells, hipStream_t s) {
    size_t work = meta_words > (size_t)num_cells / 4 ? meta_words : (size_t)num_cells / 4;
    uint32_t grid = grid_for((uint32_t)(work > 0xFFFFFFFFull ? 0xFFFFFFFFu : work), 256, 2048);
    BRT_LAUNCH(k_prepare, dim3(grid), dim3(256), 0, s, meta, (uint32_t)meta_words, cell_start, cell_end, num_cells);
    return hipGetLastError();
}

hipError_t launch_hash_hist(const brt_grid_params& g, const float4* pos, uint32_t n, uint2* out, uint32_t* meta,
                            const SortPlan& plan, hipStream_t s, uint32_t* cell_start, uint32_t* cell_end) {
    uint32_t grid = grid_for(n, 256, hist_blocks(n));
    BRT_LAUNCH(k_hash_hist, dim3(grid), dim3(256), 0, s, g, pos, n, out, meta + kMetaHist, plan, cell_start,
               cell_start ? cell_end : nullptr);
    return hipGetLastError();
}

hipError_t launch_batch_hash_hist(const BatchView& b, const float4* pos, uint2* out, uint32_t* meta,
                                  const SortPlan& plan, hipStream_t s, uint32_t* cell_start, uint32_t* cell_end,
                                  uint32_t num_cells) {
    uint32_t grid = grid_for(b.waves, 4, hist_blocks(b.waves * 64u));
    hipLaunchKernelGGL(k_batch_hash_hist, dim3(grid), dim3(256), 0, s, b, pos, out, meta + kMetaHist, plan,
                       cell_start, cell_start ? cell_end : nullptr, num_cells);
    return hipGetLastError();
}

hipError_t launch_set_word(uint32_t* p, uint32_t v, hipStream_t s) {
    hipLaunchKernelGGL(k_set_word, dim3(1), dim3(64), 0, s, p, v);
    return hipGetLastError();
}

hipError_t launch_hist(const uint2* in, uint2* copy_out, uint32_t n, uint32_t* meta, const SortPlan& plan,
                       hipStream_t s) {
    uint32_t grid = grid_for(n, 256, hist_blocks(n));
    hipLaunchKernelGGL(k_hist, dim3(grid), dim3(256), 0, s, in, copy_out, n, meta + kMetaHist, plan);
    return hipGetLastError();
}

hipError_t launch_onesweep(const uint2* in, uint2* out, uint32_t n, uint32_t pass, const SortPlan& plan,
                           uint32_t* meta, uint32_t* dev_flags, hipStream_t s) {
    uint32_t* status = meta + kMetaStatus + (size_t)pass * plan.tiles * kRadix;
    uint32_t* prev_hist = pass ? meta + kMetaHist + (pass - 1u) * kRadix : nullptr;
    uint32_t* prev_counter = pass ? meta + kMetaCounters + (pass - 1u) : nullptr;
    uint32_t* prev_status = pass ? status - (size_t)plan.tiles * kRadix : nullptr;
    if (plan.items == 8u)
        BRT_LAUNCH(k_onesweep<8>, dim3(plan.tiles), dim3(kSortThreads), 0, s, in, out, n, plan.shift[pass],
                   plan.nbits[pass], meta + kMetaHist + pass * kRadix, status, meta + kMetaCounters + pass, dev_flags,
                   prev_hist, prev_counter, prev_status);
    else
        BRT_LAUNCH(k_onesweep<16>, dim3(plan.tiles), dim3(kSortThreads), 0, s, in, out, n, plan.shift[pass],
                   plan.nbits[pass], meta + kMetaHist + pass * kRadix, status, meta + kMetaCounters + pass, dev_flags,
                   prev_hist, prev_counter, prev_status);
    return hipGetLastError();
}

hipError_t launch_reorder_ranges(const uint2* hash, uint32_t n, const ReorderArgs& f, uint32_t num_cells,
                                 uint32_t* cell_start, uint32_t* cell_end, hipStream_t s, const SortCleanup* cleanup) {
    const SortCleanup c = cleanup ? *cleanup : SortCleanup{};
    uint32_t work = n;
    if (c.hist && work < c.status_words) work = c.status_words;
    if (c.hist && work < kHistCopies * kRadix) work = kHistCopies * kRadix;
    uint32_t grid = (work + 255) / 256;
    if (grid == 0) return hipSuccess;
    BRT_LAUNCH(k_reorder_ranges, dim3(grid), dim3(256), 0, s, hash, n, f, num_cells, cell_start, cell_end, c);
    return hipGetLastError();
}

}  // namespace brt
