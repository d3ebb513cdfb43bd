/*
 * behavert_amd.h - C ABI of the MI355X (gfx950) per-step agent simulation path.
 *
 * This is the drop-in boundary for BehaveRT's per-step path:
 *   EnvGrid3D hash + stable sort + reorder  ->  Proximity3D cell ranges + capped
 *   27-cell neighbour lists  ->  OpenSteerWrapper steering + Body integration
 *   ->  BehaviorClustering feature update.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns an int
 * status (BRT_OK == 0, negative on error; brt_context_last_error() has the text)
 * and, where it launches device work, takes the hipStream_t to launch on as a
 * void*.  Device pointers are owned by the caller; the context owns only its
 * scratch (sort ping-pong buffer, digit histograms, look-back status words).
 * Scratch grows on demand with hipMalloc, so a call that must run under stream
 * capture needs one earlier call of at least the same size on the same context.
 *
 * The reference's equivalents are parameterless CUDA kernels launched through
 * DeviceInterface::kernelCall (Core/DeviceInterface.cu:380-458) that find their
 * arrays in a __constant__ DataInfo repository (Core/include/devicedata.cuh:51-70)
 * and abort the process on error (CUDA_SAFE_CALL, Core/DeviceInterface.cu:106).
 * Here the arrays and parameters are explicit arguments and errors are returned.
 *
 * Array layouts (N = agents, C = cells, K = max neighbours, W = brt_list_stride(K)):
 *   pos, forward, smooth_accel : float4[N]  (x, y, z, w)
 *       pos.w = 1; forward.xyz = unit heading, forward.w = speed; smooth_accel.w kept.
 *   hash                       : uint2[N]   (.x = cell key, .y = agent ID)
 *   cell_start, cell_end       : uint[C]    (BRT_EMPTY_CELL for empty cells)
 *   lists                      : uint[N*W]  ([0] = count, [1..count] = sorted
 *                                            indices, rest BRT_LIST_PAD)
 *   features                   : float4[N]  (BehaviorClustering 4-D feature, by ID)
 *   wander                     : float2[N]  (wander side/up random-walk state, by ID)
 * "by ID" arrays are indexed by the stable agent ID; "sorted" arrays by the
 * position of the agent in this step's cell-sorted order.
 */
#ifndef BEHAVERT_AMD_H
#define BEHAVERT_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BRT_OK 0
#define BRT_ERR_INVALID_ARGUMENT (-1)
#define BRT_ERR_HIP (-2)
#define BRT_ERR_OUT_OF_MEMORY (-3)
#define BRT_ERR_UNSUPPORTED (-4)
#define BRT_ERR_DEPENDENCY (-5)

#define BRT_EMPTY_CELL 0xFFFFFFFFu
#define BRT_LIST_PAD 0xFFFFFFFFu
#define BRT_UNLIMITED 0xFFFFFFFFu
#define BRT_MAX_TARGETS 16
#define BRT_MAX_NEIGHBORS 252 /* W = 4*(K/4+1) <= 256 */

typedef struct brt_context brt_context;

/* EnvGrid3DParams (PlugIns/EnvGrid3D/include/envgrid3d_kernel.cuh:38-65).
 * max_bodies_per_cell truncates each cell's scan in the neighbour search
 * (PlugIns/Proximity3D/Proximity3D_resources.cu:152); BRT_UNLIMITED = no cap. */
typedef struct brt_grid_params {
    float world_origin[3];
    float cell_size[3];
    uint32_t grid_size[3];
    uint32_t num_cells;
    uint32_t max_bodies_per_cell;
    float world_center[3];
    float world_radius[3];
} brt_grid_params;

/* Proximity3DParams (PlugIns/Proximity3D/include/Proximity3D_kernel.cuh:40-73). */
typedef struct brt_prox_params {
    uint32_t max_neighbors; /* maxNeighbors (K), 1..BRT_MAX_NEIGHBORS */
    float search_radius;    /* commonSearchRadius */
    uint32_t use_knn;       /* useKnn: replace the farthest entry when full */
} brt_prox_params;

/* OpenSteerWrapperParams + Body3DParams subset used on the step path
 * (PlugIns/OpenSteerWrapper/include/OpenSteerWrapperPlugIn.h:134-142,
 *  PlugIns/Body/include/body3d_kernel.cuh:39-58).  separation/alignment/cohesion
 * are (weight, radius, cosMaxAngle) as in the reference's float3 params.
 * The multi-target seek and the wander terms have no GPU counterpart in the
 * reference: seek uses xxxsteerForSeek
 * (PlugIns/OpenSteerWrapper/OpenSteerWrapper_resources.cu:294-306) toward
 * targets[id % num_targets]; wander is OpenSteer's steerForWander
 * (Samples/OpensteerDemo/include/OpenSteer/SteerLibrary.h:319-331) driven by a
 * Philox4x32-10 stream keyed by (wander_seed, agent ID) with counter (step). */
typedef struct brt_steer_params {
    float separation[3];
    float alignment[3];
    float cohesion[3];
    float mass;
    float max_speed;
    float max_force;
    float sensing_min_radius;
    uint32_t seek_world_center; /* computeSeekingsWorldCenter (box form) */
    uint32_t use_2d_projection;
    uint32_t num_targets;       /* 0 disables the multi-target seek */
    float target_weight;
    float targets[BRT_MAX_TARGETS][4];
    float wander_weight;        /* 0 disables wander */
    uint32_t wander_seed;
} brt_steer_params;

/* ---------------------------------------------------------------- context */

/* Replaces CommonResources/DeviceInterface construction
 * (Core/src/CommonResources.cpp:49-64, Core/DeviceInterface.cu:77-102). */
int brt_context_create(int device, brt_context** out);
int brt_context_destroy(brt_context* ctx);
const char* brt_context_last_error(const brt_context* ctx);
/* Library version string; safe to call without a GPU. */
const char* brt_version(void);

/* Context options (launch-shape and fusion choices; none changes any result bit).
 * BRT_OPT_FUSED_STEP: brt_step builds the neighbour lists and runs steering +
 *   integration in one kernel (default 1; KNN lists always use two passes).
 * BRT_OPT_UNIFORM_MAX_CELLS: a wave64 whose agents lie in at most this many
 *   cells scans candidates with wave-uniform (scalar-cache) loads (default 1;
 *   0 = every wave scans per lane).
 * BRT_OPT_XCD_REMAP: gather-heavy kernels map contiguous slices of the sorted
 *   agents to each XCD (one L2 per XCD on MI355X) (default 1).
 * No counterpart in the reference (its launch shapes are fixed per kernel:
 * Proximity3DPlugIn.h:133-134, CommonResources.cpp:78-82). */
#define BRT_OPT_FUSED_STEP 1
#define BRT_OPT_UNIFORM_MAX_CELLS 2
#define BRT_OPT_XCD_REMAP 3
int brt_context_set_option(brt_context* ctx, int option, int value);
int brt_context_get_option(const brt_context* ctx, int option, int* value);

/* Host-side grid set-up: EnvGrid3DPlugIn::setupGrid
 * (PlugIns/EnvGrid3D/include/EnvGrid3DPlugIn.h:478-508).  No device work. */
int brt_grid_setup(const float world_center[3], const float world_radius[3],
                   const float cell_size[3], uint32_t max_bodies_per_cell,
                   brt_grid_params* out);
/* W = 4*(K/4+1): uint slots per neighbour list, the reader/allocation layout
 * (PlugIns/Proximity3D/include/Proximity3DPlugIn.h:148-149,
 *  PlugIns/Proximity3D/Proximity3D_resources.cu:175-198). */
uint32_t brt_list_stride(uint32_t max_neighbors);
/* Largest float d2 with sqrtf(d2) <= r: the exact squared-distance threshold
 * equivalent to the reference's "sqrtf(d2) > r -> reject" test. */
float brt_radius_threshold(float r);

/* ---------------------------------------------------------------- EnvGrid3D */

/* genCalcHashD (PlugIns/EnvGrid3D/envgrid3d_kernel.cu:55-68):
 * hash[i] = (calcGridHash(calcGridPos(pos[i])), i). */
int brt_envgrid_calc_hash(brt_context* ctx, const brt_grid_params* grid,
                          const float* pos, uint32_t n, uint32_t* hash,
                          void* stream);
/* BehaveRT::RadixSort (PlugIns/EnvGrid3D/radixsort.cu:50-80): stable sort of
 * hash pairs by key in place.  Keys must be < num_cells. */
int brt_envgrid_sort(brt_context* ctx, uint32_t* hash, uint32_t n,
                     uint32_t num_cells, void* stream);
/* Fused hash + sort (the first half of EnvGrid3DPlugIn::reorderSimData,
 * PlugIns/EnvGrid3D/include/EnvGrid3DPlugIn.h:292-320). */
int brt_envgrid_hash_sort(brt_context* ctx, const brt_grid_params* grid,
                          const float* pos, uint32_t n, uint32_t* hash,
                          void* stream);
/* reorderDataFloat4 for each feature (PlugIns/EnvGrid3D/envgrid3d_kernel.cu:110-122,
 * loop at EnvGrid3DPlugIn.h:333-411): out[f][i] = in[f][hash[i].y].
 * in/out are HOST arrays of num_features DEVICE pointers (float4[n] each). */
int brt_envgrid_reorder(brt_context* ctx, const uint32_t* hash, uint32_t n,
                        uint32_t num_features, const float* const* in,
                        float* const* out, void* stream);

/* ---------------------------------------------------------------- Proximity3D */

/* genericFindCellStartD (PlugIns/Proximity3D/Proximity3D_kernel.cu:54-84) plus
 * an explicit cell_end; every cell is written (BRT_EMPTY_CELL when empty). */
int brt_prox_cell_ranges(brt_context* ctx, const uint32_t* hash, uint32_t n,
                         uint32_t num_cells, uint32_t* cell_start,
                         uint32_t* cell_end, void* stream);
/* Fused reorder of num_features float4 features + cell ranges (one pass over
 * the sorted hash). */
int brt_envgrid_reorder_ranges(brt_context* ctx, const uint32_t* hash, uint32_t n,
                               uint32_t num_features, const float* const* in,
                               float* const* out, uint32_t num_cells,
                               uint32_t* cell_start, uint32_t* cell_end,
                               void* stream);
/* genericCalcNeighgborhoodD (PlugIns/Proximity3D/Proximity3D_kernel.cu:89-125,
 * Proximity3D_resources.cu:59-172): capped 27-cell neighbour lists in the
 * reference's visitation order (x outermost, z innermost), W-stride layout. */
int brt_prox_neighbors(brt_context* ctx, const brt_grid_params* grid,
                       const brt_prox_params* prox, const float* sorted_pos,
                       const uint32_t* hash, const uint32_t* cell_start,
                       const uint32_t* cell_end, uint32_t n, uint32_t* lists,
                       void* stream);

/* ---------------------------------------------------------------- OpenSteerWrapper */

/* resetSteerForce + computeSeparations/Alignments/Cohesions [+ target seek,
 * wander] [+ computeSeekingsWorldCenter] + applySteeringForces fused in one
 * pass (PlugIns/OpenSteerWrapper/include/OpenSteerWrapperPlugIn.h:287-424,
 * OpenSteerWrapper_kernel.cu:62-333, OpenSteerWrapper_resources.cu:67-379).
 * Reads sorted state and lists, writes the new state by agent ID.
 * wander (float2 by ID) is read and updated in place when wander_weight != 0.
 * force_out (float4 sorted, optional, may be NULL) receives the summed force. */
int brt_osw_steer_integrate(brt_context* ctx, const brt_grid_params* grid,
                            const brt_prox_params* prox,
                            const brt_steer_params* steer, float dt,
                            uint32_t step, uint32_t n, const float* sorted_pos,
                            const float* sorted_forward,
                            const float* sorted_smooth_accel,
                            const uint32_t* hash, const uint32_t* lists,
                            float* wander, float* pos_out, float* forward_out,
                            float* smooth_accel_out, float* force_out,
                            void* stream);

/* ---------------------------------------------------------------- BehaviorClustering */

/* Neighbour-averaged feature update (no GPU counterpart in the reference; the
 * list walk and the sorted->ID map follow
 * BehaviorClustering_computeNeighNaiveSimilarities,
 * PlugIns/BehaviorClustering/BehaviorClustering_kernel.cu:306-353):
 *   f'[id(i)] = f[id(i)] + rate * (mean_{j in list(i)} f[id(j)] - f[id(i)]). */
int brt_bc_feature_update(brt_context* ctx, const brt_prox_params* prox,
                          float rate, uint32_t n, const uint32_t* hash,
                          const uint32_t* lists, const float* features_in,
                          float* features_out, void* stream);

/* ---------------------------------------------------------------- whole step */

/* Device buffers of one flock.  pos/forward/smooth_accel (by ID) are read and
 * overwritten with the new state.  The sorted copies, hash, cell ranges and lists
 * are outputs of the step (the Proximity3D/EnvGrid3D features of the reference).
 * features/features_out and wander are optional (NULL disables). */
typedef struct brt_step_buffers {
    uint32_t n;
    float* pos;
    float* forward;
    float* smooth_accel;
    float* sorted_pos;
    float* sorted_forward;
    float* sorted_smooth_accel;
    uint32_t* hash;
    uint32_t* cell_start;
    uint32_t* cell_end;
    uint32_t* lists;
    float* features;
    float* features_out;
    float* wander;
} brt_step_buffers;

/* One plain step in the reference's order (Samples/OpensteerDemo/plugins/
 * BehaveRT_Flocking.cpp:177-196): hash+sort, reorder+ranges, neighbours,
 * fused steering+integration, then the feature update when features != NULL
 * and feature_rate != 0. */
int brt_step(brt_context* ctx, const brt_grid_params* grid,
             const brt_prox_params* prox, const brt_steer_params* steer,
             float feature_rate, float dt, uint32_t step,
             const brt_step_buffers* buf, void* stream);

/* Per-phase timing of brt_step with hipEvents recorded on the step's stream
 * (the reference harness's per-phase timers, BehaveRT_Flocking.cpp:171-204).
 * Phases: 0 hash+sort, 1 reorder+ranges, 2 neighbours, 3 steering+integration,
 * 4 feature update.  With BRT_OPT_FUSED_STEP, phase 2 holds the fused
 * neighbours+steering kernel and phase 3 is empty.  brt_phase_times() synchronises the recorded events and
 * returns accumulated milliseconds per phase and the number of timed steps. */
#define BRT_NUM_PHASES 5
int brt_phase_timing_enable(brt_context* ctx, int enable);
int brt_phase_times(brt_context* ctx, double* ms_out, uint32_t* steps_out);

#ifdef __cplusplus
}
#endif

#endif /* BEHAVERT_AMD_H */
