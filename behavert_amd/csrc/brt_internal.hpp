// brt_internal.hpp -- host-side launchers shared between the kernel TUs and the C ABI.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/behavert_amd.h"

namespace brt {

// Phase timing without marker packets: the caller arms a start and/or stop
// event for the next kernel a launcher enqueues on this thread, and that
// launch binds them to the kernel's own dispatch (hipExtLaunchKernelGGL), so
// the events carry the kernel's start and end timestamps.  A launch consumes
// the mark; launches with no mark armed are plain launches.
struct LaunchMark {
    hipEvent_t start = nullptr;
    hipEvent_t stop = nullptr;
};
LaunchMark& launch_mark();

#define BRT_LAUNCH(kernel, grid, block, shm, stream, ...)                                                \
    do {                                                                                                \
        ::brt::LaunchMark& mark_ = ::brt::launch_mark();                                                \
        if (mark_.start || mark_.stop) {                                                                \
            const ::brt::LaunchMark armed_ = mark_;                                                     \
            mark_ = ::brt::LaunchMark{};                                                                \
            hipExtLaunchKernelGGL(kernel, grid, block, shm, stream, armed_.start, armed_.stop, 0u, __VA_ARGS__); \
        } else {                                                                                        \
            hipLaunchKernelGGL(kernel, grid, block, shm, stream, __VA_ARGS__);                          \
        }                                                                                               \
    } while (0)

constexpr int kSortThreads = 256;            // 4 wave64s per tile
constexpr int kSortItems = 16;               // keys per thread
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kRadix = 256;                  // 8-bit digits
constexpr int kMaxPasses = 4;
// The digit histograms are accumulated in kHistCopies copies (a histogram
// workgroup adds into copy blockIdx % kHistCopies), so that each bin's global
// atomics are spread over that many addresses; a sort pass sums the copies.
#ifndef BRT_HIST_COPIES
#define BRT_HIST_COPIES 8u
#endif
constexpr uint32_t kHistCopies = BRT_HIST_COPIES;
constexpr uint32_t kHistWords = kMaxPasses * kRadix;        // one copy: hist[4][256]
constexpr uint32_t kMetaHist = 0;                           // uint32 hist[copies][4][256]
constexpr uint32_t kMetaCounters = kHistCopies * kHistWords;  // uint32 tile counters[4] (padded to 64)
constexpr uint32_t kMetaStatus = kMetaCounters + 64;          // uint32 status[passes][tiles][256]

struct SortPlan {
    uint32_t passes;
    uint32_t tiles;
    uint32_t items;  // keys per thread: 16 (tile 4096) or 8 (tile 2048)
    uint32_t shift[kMaxPasses];
    uint32_t nbits[kMaxPasses];
};

SortPlan make_sort_plan(uint32_t n, uint32_t num_cells, uint32_t items = kSortItems);
size_t sort_meta_words(const SortPlan& p);

// envgrid
hipError_t launch_prepare(uint32_t* meta, size_t meta_words, uint32_t* cell_start, uint32_t* cell_end,
                          uint32_t num_cells, hipStream_t s);
// cell_start / cell_end (optional): filled with BRT_EMPTY_CELL in the same
// launch (when no k_prepare ran before it)
hipError_t launch_hash_hist(const brt_grid_params& g, const float4* pos, uint32_t n, uint2* out,
                            uint32_t* meta, const SortPlan& plan, hipStream_t s, uint32_t* cell_start = nullptr,
                            uint32_t* cell_end = nullptr);
hipError_t launch_hist(const uint2* in, uint2* copy_out, uint32_t n, uint32_t* meta, const SortPlan& plan,
                       hipStream_t s);
// dev_flags: the context's sticky device flag word (BRT_DEVFLAG_*).  Pass p >= 1
// also returns the previous pass's histogram slice, tile counter and status
// words to zero (nothing reads them any more).
hipError_t launch_onesweep(const uint2* in, uint2* out, uint32_t n, uint32_t pass, const SortPlan& plan,
                           uint32_t* meta, uint32_t* dev_flags, hipStream_t s);

constexpr int kMaxReorderFeatures = 8;
struct ReorderArgs {
    const float4* in[kMaxReorderFeatures];
    float4* out[kMaxReorderFeatures];
    uint32_t count;
    float* xyz0;  // optional: feature 0's x,y,z packed, 12 bytes per agent
};
// The sort metadata a finished sort leaves nonzero: the last pass's histogram
// slice (every copy), tile counter and status words.  A reorder launched right
// after the sort returns them to zero, so that the whole metadata buffer is
// zero between steps and the next step needs no clearing launch.
struct SortCleanup {
    uint32_t* hist;     // meta + kMetaHist + last * kRadix (copies kHistWords apart); NULL: nothing to clear
    uint32_t* counter;  // meta + kMetaCounters + last
    uint32_t* status;   // the last pass's status words
    uint32_t status_words;
};
SortCleanup sort_cleanup(uint32_t* meta, const SortPlan& plan);
hipError_t launch_reorder_ranges(const uint2* hash, uint32_t n, const ReorderArgs& f, uint32_t num_cells,
                                 uint32_t* cell_start, uint32_t* cell_end, hipStream_t s,
                                 const SortCleanup* cleanup = nullptr);

// proximity
hipError_t launch_neighbors(const brt_grid_params& g, const brt_prox_params& p, float threshold,
                            const float4* spos, const uint2* hash, const uint32_t* cell_start,
                            const uint32_t* cell_end, uint32_t n, uint32_t* lists, uint32_t uniform_max_cells,
                            uint32_t xcd_remap, uint32_t lds_lists, hipStream_t s, const float* pxyz = nullptr);
// neighbour lists + steering + integration in one pass (non-KNN lists only)
hipError_t launch_neighbors_steer(const brt_grid_params& g, const brt_prox_params& p, float threshold,
                                  const brt_steer_params& sp, float dt, uint32_t step, const float4* spos,
                                  const float4* sfwd, const float4* ssa, const uint2* hash,
                                  const uint32_t* cell_start, const uint32_t* cell_end, uint32_t n, uint32_t* lists,
                                  float2* wander, float4* pos_out, float4* fwd_out, float4* sa_out,
                                  uint32_t uniform_max_cells, uint32_t xcd_remap, uint32_t lds_lists, hipStream_t s,
                                  const float* pxyz = nullptr);
// pxyz (optional, single-grid launches): the sorted positions packed as x,y,z,
// 12 bytes per agent (ReorderArgs::xyz0)

// opensteer
hipError_t launch_steer(const brt_grid_params& g, uint32_t K, const brt_steer_params& sp, float dt, uint32_t step,
                        uint32_t n, const float4* spos, const float4* sfwd, const float4* ssa, const uint2* hash,
                        const uint32_t* lists, float2* wander, float4* pos_out, float4* fwd_out, float4* sa_out,
                        float4* force_out, uint32_t xcd_remap, hipStream_t s, float feature_rate = 0.0f,
                        const float4* features_in = nullptr, float4* features_out = nullptr,
                        const float4* sorted_features = nullptr);
hipError_t launch_features(uint32_t K, float rate, uint32_t n, const uint2* hash, const uint32_t* lists,
                           const float4* fin, float4* fout, hipStream_t s);
// the decomposed steering sequence: behaviour forces only, then the integrator
hipError_t launch_steer_forces(uint32_t K, const brt_steer_params& sp, float dt, uint32_t step, uint32_t n,
                               const float4* spos, const float4* sfwd, const uint2* hash, const uint32_t* lists,
                               float2* wander, float4* force_out, uint32_t xcd_remap, hipStream_t s);
hipError_t launch_apply(const brt_grid_params& g, const brt_steer_params& sp, float dt, uint32_t n, const float4* spos,
                        const float4* sfwd, const float4* ssa, const uint2* hash, const float4* force_in,
                        float4* pos_out, float4* fwd_out, float4* sa_out, float4* force_out, hipStream_t s);

// BehaviorClustering similarities and cluster steering
// scratch: bc_scratch_bytes(F, n), 16-byte aligned (features in sorted order + norms);
// sp != nullptr also runs the cluster steering on the fly (force_in -> force_out)
size_t bc_scratch_bytes(uint32_t feature_count, uint32_t n);
hipError_t launch_bc_similarities(uint32_t K, const brt_bc_params& bc, uint32_t n, const uint2* hash,
                                  const uint32_t* lists, const float* features, const float4* stats_in,
                                  float4* stats_out, float* sims, void* scratch, const brt_steer_params* sp,
                                  const float4* spos, const float4* sfwd, const float4* force_in, float4* force_out,
                                  hipStream_t s);
hipError_t launch_bc_cluster_steer(uint32_t K, const brt_steer_params& sp, const brt_bc_params& bc, uint32_t n,
                                   const float4* spos, const float4* sfwd, const uint32_t* lists, const float* sims,
                                   const float4* force_in, float4* force_out, hipStream_t s);
hipError_t launch_cluster_membership(uint32_t K, uint32_t n, uint32_t iterations, const uint2* hash,
                                     const uint32_t* lists, const float4* spos, const float* m_in, float* m_out,
                                     float4* conn, float* lbl_a, float* lbl_b, hipStream_t s);
// lbl: float[n] scratch (the labels in sorted order)
hipError_t launch_hit_rate(uint32_t K, float radius, uint32_t n, const uint2* hash, const uint32_t* lists,
                           const float4* spos, const float* m_in, float* m_out, float* lbl, hipStream_t s);

// list consumers outside the step (brt_consumers.hip)
hipError_t launch_slow_avoid(uint32_t K, const brt_steer_params& sp, uint32_t n, const float4* spos,
                             const float4* sfwd, const uint32_t* lists, float4* fwd_out, const float4* force_in,
                             float4* force_out, hipStream_t s);
hipError_t launch_neighborhood_color(uint32_t K, const float base[3], uint32_t vertexes, uint32_t n,
                                     const uint2* hash, const uint32_t* lists, float4* color, hipStream_t s);
hipError_t launch_move_towards_target(int shift, const float base[3], uint32_t n, const uint2* hash,
                                      const float4* spos, const float4* targets, const float4* force_in,
                                      float4* force_out, hipStream_t s);

// adjoint
hipError_t launch_adjoint(const brt_grid_params& g, const brt_prox_params& p, float threshold,
                          const brt_steer_params& sp, float dt, uint32_t n, const float4* spos, const float4* sfwd,
                          const float4* ssa, const uint2* hash, const uint32_t* lists, const uint32_t* cell_start,
                          const uint32_t* cell_end, const float4* gpos_out, const float4* gfwd_out,
                          const float4* gsa_out, float4* gpos_in, float4* gfwd_in, float4* gsa_in, float4* scratch,
                          float* partials, float* weight_grad, uint32_t xcd_remap, hipStream_t s);

inline uint32_t list_stride(uint32_t K) { return 4u * (K / 4u + 1u); }
// Row stride of the LDS list tile of the neighbour kernels: the list stride,
// widened so a round of 8 candidates can append past K without a per-candidate
// count check (rows of W >= K + 8 words, a multiple of 4), unless a 64-lane
// tile of such rows would exceed 64 KB (then the rounds check the count).
inline uint32_t lds_list_stride(uint32_t K) {
    const uint32_t W = list_stride(K), need = (K + 8u + 3u) / 4u * 4u;
    return (W >= need || need > 256u) ? W : need;
}

// ---------------------------------------------------------------- batches
// One flock of a batch as the kernels see it.  Flocks are concatenated: flock f
// owns agent IDs and sorted indices [offset, offset + n) and global cell keys
// [cell_base, cell_base + g.num_cells).  Waves are assigned per flock (a wave
// never spans two flocks): flock f owns waves [wave_base, wave_base + ceil(n/64)).
struct alignas(16) FlockDev {
    brt_grid_params g;
    float slack[3];      // cell-box rounding margins of g
    uint32_t offset, n, cell_base, wave_base;
    float w[3];          // separation / alignment / cohesion weights
    uint32_t pad[4];
};
static_assert(sizeof(FlockDev) == 128, "FlockDev is two 64-byte scalar loads");

struct BatchView {
    const FlockDev* flocks;
    const uint32_t* wave_flock;  // flock of each wave
    uint32_t waves;              // total waves
    const uint32_t* step;        // device step counter (wander), may be NULL
};

// per-axis cell-box rounding margins (see the neighbour kernels)
inline void cell_slack(const brt_grid_params& g, float out[3]) {
    for (int ax = 0; ax < 3; ++ax)
        out[ax] = 1e-5f * (fabsf(g.world_origin[ax]) + (float)g.grid_size[ax] * g.cell_size[ax]);
}

hipError_t launch_batch_hash_hist(const BatchView& b, const float4* pos, uint2* out, uint32_t* meta,
                                  const SortPlan& plan, hipStream_t s, uint32_t* cell_start = nullptr,
                                  uint32_t* cell_end = nullptr, uint32_t num_cells = 0);
hipError_t launch_set_word(uint32_t* p, uint32_t v, hipStream_t s);
hipError_t launch_batch_neighbors_steer(const BatchView& b, uint32_t n, const brt_prox_params& p, float threshold,
                                        const brt_steer_params& sp, float dt, const float4* spos, const float4* sfwd,
                                        const float4* ssa, const uint2* hash, const uint32_t* cell_start,
                                        const uint32_t* cell_end, uint32_t* lists, float2* wander, float4* pos_out,
                                        float4* fwd_out, float4* sa_out, uint32_t block, uint32_t uniform_max_cells,
                                        uint32_t xcd_remap, uint32_t lds_lists, hipStream_t s,
                                        const float* pxyz = nullptr);

}  // namespace brt
