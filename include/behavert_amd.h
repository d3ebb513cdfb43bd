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
 *   integration in one kernel (1) or writes the lists and steers from them (0)
 *   (default -1 = automatic: fused for K <= 32, two passes above, where the
 *   fused kernel's occupancy falls; KNN lists always use two passes).
 * BRT_OPT_UNIFORM_MAX_CELLS: a wave64 whose agents lie in at most this many
 *   cells scans candidates with wave-uniform (scalar-cache) loads (default 1;
 *   0 = every wave scans per lane).
 * BRT_OPT_XCD_REMAP: gather-heavy kernels map contiguous slices of the sorted
 *   agents to each XCD (one L2 per XCD on MI355X) (default 1).
 * BRT_OPT_LDS_LISTS: the neighbour kernels stage each workgroup's lists in LDS
 *   and write them out with coalesced 16-byte stores (1), or gather each list's
 *   16-byte chunks in registers and store them as they fill (0), which leaves
 *   LDS free and raises occupancy (default -1 = automatic: LDS for K <= 32).
 * No counterpart in the reference (its launch shapes are fixed per kernel:
 * Proximity3DPlugIn.h:133-134, CommonResources.cpp:78-82). */
#define BRT_OPT_FUSED_STEP 1
#define BRT_OPT_UNIFORM_MAX_CELLS 2
#define BRT_OPT_XCD_REMAP 3
#define BRT_OPT_LDS_LISTS 4
/* BRT_OPT_DEVICE_FLAGS: the context's sticky device flag word (BRT_DEVFLAG_*,
 * see brt_context_check); get synchronises the device and reads it, set writes
 * it (for tests of the error path). */
#define BRT_OPT_DEVICE_FLAGS 5
/* BRT_OPT_STEP_GRAPH: brt_step captures its kernels into a hipGraph on the
 * second call with the same buffers, parameters and options (the step number
 * counts only with wander on) and replays the graph from then on, one launch
 * per step (two graphs are kept, for a feature update's swapped buffers;
 * steps with phase timing on launch their kernels directly).  Default 0: on
 * gfx950 with ROCm 7 a replay measured 5-7 us slower per step than the direct
 * launches (1,048,576 and 65,536 boids). */
#define BRT_OPT_STEP_GRAPH 6
int brt_context_set_option(brt_context* ctx, int option, int value);
int brt_context_get_option(const brt_context* ctx, int option, int* value);
/* The launch shape brt_step and brt_prox_neighbors use for these proximity
 * parameters once automatic options are resolved: fused (0/1), lds_lists (0/1). */
int brt_context_step_shape(const brt_context* ctx, const brt_prox_params* prox, int* fused, int* lds_lists);

/* Failures the kernels record on the device instead of hanging or faulting:
 * BRT_DEVFLAG_SORT_LOOKBACK - a radix-sort pass gave up a look-back wait after
 * its spin bound (2^24 polls); that step's sort permutation, and everything
 * built from it, is wrong.  The flags are sticky until read.
 * brt_context_check synchronises the device, reads and clears them, and
 * returns BRT_ERR_HIP with a message naming the failure (BRT_OK when none was
 * raised).  brt_phase_times, brt_snapshot_save and brt_snapshot_load, which
 * synchronise anyway, check them too.  The reference has no counterpart (its
 * kernels cannot report a failure; CUDA_SAFE_CALL exits on launch errors,
 * Core/DeviceInterface.cu:106). */
#define BRT_DEVFLAG_SORT_LOOKBACK 1u
int brt_context_check(brt_context* ctx);

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
 * reference's visitation order (x outermost, z innermost), W-stride layout.
 * n < 2^28 (as for brt_step and batches). */
int brt_prox_neighbors(brt_context* ctx, const brt_grid_params* grid,
                       const brt_prox_params* prox, const float* sorted_pos,
                       const uint32_t* hash, const uint32_t* cell_start,
                       const uint32_t* cell_end, uint32_t n, uint32_t* lists,
                       void* stream);

/* ---------------------------------------------------------------- OpenSteerWrapper */

/* resetSteerForce + computeSeparations/Alignments/Cohesions [+ target seek,
 * wander] [+ computeSeekingsWorldCenter] + applySteeringForces fused in one
 * pass (PlugIns/OpenSteerWrapper/include/OpenSteerWrapperPlugIn.h:288-426,
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

/* Adjoint (reverse mode) of brt_osw_steer_integrate with the sort and the lists
 * held fixed as the forward step built them.  No counterpart in the reference.
 * Inputs are the forward step's sorted state, hash, lists and cell ranges
 * (exactly what the forward step read and wrote: the lists must be those the
 * neighbour search builds from these sorted positions, hash and ranges, as
 * brt_step records them) and dL/d(new state) by agent ID:
 *   adj_pos_out     float4: xyz = dL/dpos
 *   adj_forward_out float4: xyz = dL/d(unit heading), w = dL/dspeed
 *   adj_smooth_accel_out float4: xyz = dL/dsmooth_accel
 * Outputs (overwritten) are dL/d(previous state) by agent ID in the same
 * layout, and dL/d(separation, alignment, cohesion weight) ADDED to the three
 * floats at weight_grad (device memory).  No atomics: the same inputs give
 * the same bits (the neighbour terms of agent j are gathered from the agents
 * whose lists hold j, found in the 27 cells around j's cell).
 * The hash must be a permutation of the agent IDs (as brt_step writes it).
 * Requires use_2d_projection == 0 and wander_weight == 0. */
int brt_osw_adjoint_step(brt_context* ctx, const brt_grid_params* grid,
                         const brt_prox_params* prox, const brt_steer_params* steer,
                         float dt, uint32_t n, const float* sorted_pos,
                         const float* sorted_forward, const float* sorted_smooth_accel,
                         const uint32_t* hash, const uint32_t* lists,
                         const uint32_t* cell_start, const uint32_t* cell_end,
                         const float* adj_pos_out, const float* adj_forward_out,
                         const float* adj_smooth_accel_out, float* adj_pos_in,
                         float* adj_forward_in, float* adj_smooth_accel_in,
                         float* weight_grad, void* stream);

/* The reference's decomposed steering sequence (OpenSteerWrapperPlugIn.h:288-426:
 * resetSteerForce, one accumulating pass per behaviour, applySteeringForces),
 * for callers that add forces of their own between the behaviours and the
 * integrator (BehaviorClustering's steerForClustering does).
 * brt_osw_steer_forces: force_out (float4 sorted) = the behaviour forces
 * (separation, alignment, cohesion [, target seek, wander]) accumulated from 0.
 * brt_osw_apply_steering_force: adds the world-centre seek when enabled, then
 * integrates (applySteeringForceSingle) and writes the new state by ID;
 * force_out (optional) receives the final force.  steer_forces followed by
 * apply_steering_force equals brt_osw_steer_integrate bit for bit. */
int brt_osw_steer_forces(brt_context* ctx, const brt_prox_params* prox,
                         const brt_steer_params* steer, float dt, uint32_t step,
                         uint32_t n, const float* sorted_pos, const float* sorted_forward,
                         const uint32_t* hash, const uint32_t* lists, float* wander,
                         float* force_out, void* stream);
int brt_osw_apply_steering_force(brt_context* ctx, const brt_grid_params* grid,
                                 const brt_steer_params* steer, float dt, uint32_t n,
                                 const float* sorted_pos, const float* sorted_forward,
                                 const float* sorted_smooth_accel, const uint32_t* hash,
                                 const float* force_in, float* pos_out, float* forward_out,
                                 float* smooth_accel_out, float* force_out, void* stream);

/* OpenSteerWrapperPlugIn::steerAndSlowToAvoidNeighbors (OpenSteerWrapperPlugIn.h:309-325,
 * kernel steerToAvoidNeighbors OpenSteerWrapper_kernel.cu:110-154), per sorted agent:
 *   s = calcSeparation * separation weight; when the list holds c > 16 agents,
 *   s *= (c - 16) * 0.5; speed *= 0.9 when dot(s, forward) < -0.8, *= 1.02 when
 *   > 0.5 (double-precision constants, as compiled for sm_20+);
 *   forward_out = (forward, speed) sorted; force_out = s + force_in (may alias). */
int brt_osw_steer_and_slow_to_avoid_neighbors(brt_context* ctx, const brt_prox_params* prox,
                                              const brt_steer_params* steer, uint32_t n,
                                              const float* sorted_pos, const float* sorted_forward,
                                              const uint32_t* lists, float* forward_out,
                                              const float* force_in, float* force_out, void* stream);

/* Drawable3DPlugIn's neighbourhood colouring (drawBodiesAsPointSprites with
 * neighborhoodColoring, Drawable3DPlugIn.h:712-724; extractColorFromNeighborhoodD
 * drawable3d_kernel.cu:60-93): f = list count / K,
 *   color = (base.x + f, base.y + 0.3 + 1 - f, base.z + f (1 - f), 1)
 * by agent ID; with vertexes_per_agent = 3 (useCUDAGeometry) three float4 per
 * agent, the first minus (0.5, 0.5, 0.5, 0). */
int brt_draw_neighborhood_color(brt_context* ctx, const brt_prox_params* prox,
                                const float color_base[3], uint32_t vertexes_per_agent, uint32_t n,
                                const uint32_t* hash, const uint32_t* lists, float* color,
                                void* stream);

/* Shapes3DPlugIn::moveTowardsTarget (Shapes3DPlugIn.h:552-580, moveTowardsTargetD
 * Shapes3D_kernel.cu:63-87): t = targets[id + index_shift], the index reduced by
 * n when it exceeds n (index n reads as (0,0,0,0), as the reference's texture
 * fetch past the end does); force_out = normalize((t.xyz + target_base) - pos)
 * * 12 * t.w + force_in (sorted, may alias).  targets: float4[n] by shifted ID;
 * 0 <= index_shift < n (the host keeps it there). */
int brt_shapes_move_towards_target(brt_context* ctx, int index_shift, const float target_base[3],
                                   uint32_t n, const uint32_t* hash, const float* sorted_pos,
                                   const float* targets, const float* force_in, float* force_out,
                                   void* stream);

/* ---------------------------------------------------------------- BehaviorClustering */

/* BehaviorClusteringParams (PlugIns/BehaviorClustering/include/BehaviorClustering_kernel.cuh:40-59),
 * the fields the similarity and cluster-steering kernels read. */
#define BRT_MAX_FEATURES 32
typedef struct brt_bc_params {
    uint32_t feature_count;          /* featuresCount, 1..BRT_MAX_FEATURES */
    uint32_t adaptive;               /* 1: adaptive similarities (stats update + interpolation) */
    float avg_similarity_controller; /* avgSimilarityController (the plug-in sets 1.0) */
    float stats_params[3];           /* similarityStatsParams: avg blend, min/max blend, memory */
    float cluster_weight;            /* clusteringForceParams.x */
    float alignment_weight;          /* alignmentClusteringForceParams.x */
} brt_bc_params;

/* computeNeighNaiveSimilarities, then (adaptive) computeAdaptiveSimilarities
 * (BehaviorClustering_kernel.cu:306-353, 438-497; computeSimilarity,
 * updateSimilarityStats, interpolateSimilarity BehaviorClustering_resources.cu:61-180),
 * as one pass: for sorted agent i and list entry t, the cosine similarity of
 * the two feature vectors mapped to [0, 1]; with adaptive, the agent's
 * (min, avg, max) statistics are updated entry by entry and the similarity
 * is interpolated through them.
 *   features      float[n * feature_count] by agent ID
 *   stats_in/out  float4[n] by agent ID (x min, y avg, z max, w kept); read
 *                 and written only when adaptive (the plug-in starts them at
 *                 (1, 0.5, 0, 0))
 *   similarities  float[n * (K + 1)] by sorted index: entry t of agent i at
 *                 i*(K+1) + t (the reference's neighSimilarities stride);
 *                 entries count..K are set to 0. */
int brt_bc_neigh_similarities(brt_context* ctx, const brt_prox_params* prox,
                              const brt_bc_params* bc, uint32_t n, const uint32_t* hash,
                              const uint32_t* lists, const float* features,
                              const float* stats_in, float* stats_out,
                              float* similarities, void* stream);

/* steerForClusterCohesion then steerForClusterAlignment
 * (BehaviorClustering_resources.cu:282-384), blended into the steer-force
 * accumulator as the reference does: per agent
 *   c = sum_t seek(p_t) s_t + flee(p_t) (1 - s_t);  c = normalize(c / n) * cluster_weight
 *   a = sum_t f_t s_t;                               a = normalize(a / n - f) * alignment_weight
 *   force_out = a + (c + force_in)     (terms are zero when the list is empty)
 * with seek/flee = xxxsteerForSeek/Flee (OpenSteerWrapper_resources.cu:294-322).
 * force_in/force_out float4 sorted (may alias). */
int brt_bc_steer_for_clustering(brt_context* ctx, const brt_prox_params* prox,
                                const brt_steer_params* steer, const brt_bc_params* bc,
                                uint32_t n, const float* sorted_pos, const float* sorted_forward,
                                const uint32_t* lists, const float* similarities,
                                const float* force_in, float* force_out, void* stream);

/* BehaviorClusteringPlugIn::steerForClustering (BehaviorClusteringPlugIn.h:740-817)
 * in one pass: brt_bc_neigh_similarities followed by brt_bc_steer_for_clustering,
 * with the same outputs bit for bit (similarities, stats_out, force_out), the
 * cluster sums accumulated as each similarity is produced instead of re-reading
 * the lists and similarities.  force_in/force_out float4 sorted (may alias). */
int brt_bc_similarities_steer(brt_context* ctx, const brt_prox_params* prox,
                              const brt_steer_params* steer, const brt_bc_params* bc, uint32_t n,
                              const uint32_t* hash, const uint32_t* lists, const float* features,
                              const float* stats_in, float* stats_out, float* similarities,
                              const float* sorted_pos, const float* sorted_forward,
                              const float* force_in, float* force_out, void* stream);

/* BehaviorClusteringPlugIn::computeClusterMembership (BehaviorClusteringPlugIn.h:823-865):
 * local propagation of cluster labels over the neighbour lists.
 *   reset (BehaviorClustering_kernel.cu:632-652): label of the agent at sorted
 *     index i = (float)i;
 *   `iterations` Jacobi steps (:583-625; the reference runs iterationLP = 10,
 *     BehaviorClusteringPlugIn.h:683): label = min(own, listed neighbours');
 *   connector (:657-689, optional): connector[2*id] = sorted_pos[i],
 *     connector[2*id + 1] = sorted_pos[label].
 *   membership_in/out  float3[n] by agent ID (x label, y hit rate, z mean
 *                      label: y and z are carried from membership_in); may alias
 *   connector          float4[2n] by agent ID, or NULL
 * n <= 2^24 (labels are floats holding sorted indices exactly). */
int brt_bc_cluster_membership(brt_context* ctx, const brt_prox_params* prox, uint32_t n,
                              uint32_t iterations, const uint32_t* hash, const uint32_t* lists,
                              const float* sorted_pos, const float* membership_in,
                              float* membership_out, float* connector, void* stream);

/* BehaviorClusteringPlugIn::analyzeClusterHitRate's kernel
 * (BehaviorClustering_kernel.cu:502-578): over the listed neighbours within
 * same_cluster_max_radius (the reference fixes 1.5, BehaviorClusteringPlugIn.h:166)
 * of agent i, y = fraction whose label equals the agent's, z = mean of their
 * labels (-1 when there is none); x = the agent's label.  float3 by ID;
 * membership_in and membership_out must differ. */
int brt_bc_cluster_hit_rate(brt_context* ctx, const brt_prox_params* prox,
                            float same_cluster_max_radius, uint32_t n, const uint32_t* hash,
                            const uint32_t* lists, const float* sorted_pos,
                            const float* membership_in, float* membership_out, void* stream);

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
 * features/features_out and wander are optional (NULL disables).
 * sorted_smooth_accel is optional too: with NULL the step does not gather the
 * smoothed acceleration into sorted order (the reference reorders it only for
 * the integrator, EnvGrid3DPlugIn.h:333-411) and the integrator reads each
 * agent's own by ID instead; the new state, hash and lists are the same bits. */
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

/* brt_step with no allocation: the same step, bit for bit, with its scratch
 * (sort metadata, the sort ping-pong / sorted features, the packed positions)
 * in a static device array of the library instead of the context's buffers,
 * so the call makes no hipMalloc/hipFree (and records no phase-timing
 * events, which brt_phase_timing_enable would create).  No counterpart in the
 * reference:
 * for callers that cannot allocate (code under HIP graph capture, serving
 * loops, harnesses that forbid allocation).  Limit: n <= BRT_NOALLOC_MAX_AGENTS
 * (the array holds about 30 MB on each GPU the library runs on, for the life
 * of the process); larger n returns BRT_ERR_INVALID_ARGUMENT.  Every call on
 * one GPU shares that GPU's array: calls that overlap on one GPU (other
 * streams, other host threads, other contexts) race. */
#define BRT_NOALLOC_MAX_AGENTS (1u << 20)
int brt_step_noalloc(brt_context* ctx, const brt_grid_params* grid,
                     const brt_prox_params* prox, const brt_steer_params* steer,
                     float feature_rate, float dt, uint32_t step,
                     const brt_step_buffers* b, void* stream);

/* ---------------------------------------------------------------- state snapshots
 * Binary snapshot of a flock (no counterpart in the reference, which keeps no
 * state on disk; its host round trips are DeviceArrayWrapper::copyArrayFromDevice
 * / copyArrayToDevice, Core/include/DeviceArrayWrapper.h:345-386): a 256-byte
 * header (magic "BRTSNAP\0", version 1, n, K, W, flags, step, grid and proximity
 * parameters, payload size, FNV-1a 64 checksum) then pos, forward and
 * smooth_accel (float4 by ID), with BRT_SNAP_LISTS the sorted hash and the
 * lists, with BRT_SNAP_FEATURES the float4 features, with BRT_SNAP_WANDER the
 * wander state.  Save copies device -> host on `stream` and synchronizes it;
 * load checks magic, version, n, flags and checksum, then copies host -> device
 * the sections asked for (flags must be a subset of the file's). */
#define BRT_SNAP_LISTS 1u
#define BRT_SNAP_FEATURES 2u
#define BRT_SNAP_WANDER 4u
typedef struct brt_snapshot_info {
    uint32_t version, n, max_neighbors, list_stride, flags, step;
    brt_grid_params grid;
    brt_prox_params prox;
    uint64_t payload_bytes;
    uint64_t checksum;
} brt_snapshot_info;
int brt_snapshot_save(brt_context* ctx, const char* path, const brt_grid_params* grid,
                      const brt_prox_params* prox, uint32_t step, uint32_t flags,
                      const brt_step_buffers* buf, void* stream);
/* Header only; needs no device. */
int brt_snapshot_read_info(const char* path, brt_snapshot_info* info);
/* prox: the proximity parameters the destination lists were sized for
 * (brt_list_stride(prox->max_neighbors) slots per agent); with BRT_SNAP_LISTS
 * the file's K must equal prox->max_neighbors.  May be NULL without
 * BRT_SNAP_LISTS.  The header's payload size must follow from its n, K and
 * flags; a file that disagrees is rejected before anything is allocated. */
int brt_snapshot_load(brt_context* ctx, const char* path, uint32_t flags,
                      const brt_prox_params* prox, const brt_step_buffers* buf,
                      brt_snapshot_info* info, void* stream);

/* Per-phase timing of brt_step (the reference harness's per-phase timers,
 * BehaveRT_Flocking.cpp:171-204): each phase's start and stop hipEvents are
 * bound to the phase's first and last kernel launch (hipExtLaunchKernelGGL),
 * so a phase's time is the span of its own kernels, with no marker packets
 * between the phases (the gaps between phases are in no phase).
 * Phases: 0 hash+sort, 1 reorder+ranges, 2 neighbours, 3 steering+integration,
 * 4 feature update.  With BRT_OPT_FUSED_STEP, phase 2 holds the fused
 * neighbours+steering kernel and phase 3 is empty; a phase that launched no
 * kernel reads 0.  Steps enqueued while their stream is being captured into a
 * graph are not timed (and not counted).  brt_phase_times() synchronises the
 * recorded events and returns accumulated milliseconds per phase and the
 * number of timed steps. */
#define BRT_NUM_PHASES 5
int brt_phase_timing_enable(brt_context* ctx, int enable);
int brt_phase_times(brt_context* ctx, double* ms_out, uint32_t* steps_out);

/* ---------------------------------------------------------------- batches */

/* Many independent flocks advanced together.  The reference runs one flock per
 * process (BehaveRT_Flocking.cpp:177-196); a batch has no counterpart there.
 * Each flock keeps its own grid (brt_grid_setup) and behaviour weights; the
 * proximity and the other steering parameters are shared.  Flocks are
 * concatenated in the order given: flock f owns agent IDs and sorted indices
 * [offset_f, offset_f + n_f) and cell keys [cell_base_f, cell_base_f +
 * num_cells_f) of the batch arrays, so that every flock's hash, sort, cell
 * ranges, lists (minus offset_f) and new state equal those of brt_step run on
 * that flock alone. */
typedef struct brt_flock_desc {
    brt_grid_params grid;
    uint32_t n;
    float weights[3]; /* separation, alignment, cohesion (override steer's [0]) */
} brt_flock_desc;

typedef struct brt_batch brt_batch;

/* Tunables of the batched step.  prox_block: lanes per workgroup of the fused
 * neighbour+steering kernel (64, 128 or 256; its LDS list tile is
 * prox_block * brt_list_stride(K) * 4 bytes); uniform_max_cells: a wave whose
 * agents share at most this many cells scans them cell by cell with
 * wave-uniform loads (0..8); xcd_remap: 0/1; sort_items: keys per thread of
 * the radix sort (8 or 16); lds_lists: the list tile lives in LDS (1) or the
 * lists leave through per-lane register chunks (0, no LDS tile; see
 * BRT_OPT_LDS_LISTS); use_graph: capture the step once into a hipGraph and
 * replay it (1) or launch its kernels each call (0). */
typedef struct brt_batch_config {
    uint32_t prox_block;
    uint32_t uniform_max_cells;
    uint32_t xcd_remap;
    uint32_t sort_items;
    uint32_t lds_lists;
    uint32_t use_graph;
} brt_batch_config;

int brt_batch_create(brt_context* ctx, const brt_flock_desc* flocks, uint32_t num_flocks,
                     const brt_prox_params* prox, brt_batch** out);
int brt_batch_destroy(brt_batch* batch);
/* total agents and cells; offsets/cell_bases (num_flocks entries each) may be NULL */
int brt_batch_layout(const brt_batch* batch, uint32_t* total_agents, uint32_t* total_cells,
                     uint32_t* offsets, uint32_t* cell_bases);
int brt_batch_set_config(brt_batch* batch, const brt_batch_config* cfg);
int brt_batch_get_config(const brt_batch* batch, brt_batch_config* cfg);

/* One step of every flock: hash + sort + reorder + ranges + fused neighbour
 * lists and steering + integration (+ the feature update), over the batch
 * arrays described by buf (buf->n = total agents; cell_start/cell_end hold
 * total_cells entries; KNN lists are not supported).  With use_graph the
 * sequence is captured on the first call and replayed while the buffers,
 * parameters and configuration stay the same; the step number reaches the
 * kernels through device memory, so one graph serves every step. */
int brt_batch_step(brt_batch* batch, const brt_steer_params* steer, float feature_rate,
                   float dt, uint32_t step, const brt_step_buffers* buf, void* stream);

/* Times every configuration of the tunables above on the caller's buffers
 * (their state is saved first and restored afterwards), keeps the fastest in
 * the batch and returns it.  With cache_path != NULL the result is looked up
 * in / appended to that text file under a key of (GPU architecture, flock
 * sizes and grids, K), and a cached entry is used without timing. */
int brt_batch_autotune(brt_batch* batch, const brt_steer_params* steer, float dt,
                       const brt_step_buffers* buf, uint32_t steps_per_trial,
                       const char* cache_path, brt_batch_config* chosen, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* BEHAVERT_AMD_H */
