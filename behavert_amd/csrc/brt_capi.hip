// brt_capi.hip -- the extern "C" boundary (include/behavert_amd.h).
//
// Replaces BehaveRT's DeviceInterface/kernelCall launch path
// (Core/DeviceInterface.cu:380-458): explicit pointers and parameters instead of a
// __constant__ repository upload per call, a caller-chosen stream instead of the
// default stream, and status codes instead of exit() on error.
#include <hip/hip_runtime.h>

#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/behavert_amd.h"
#include "brt_host.hpp"
#include "brt_internal.hpp"

namespace brt {
namespace host {

int set_err(brt_context* ctx, int code, const char* fmt, ...) {
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return code;
}

int hip_err(brt_context* ctx, hipError_t e, const char* what) {
    if (e == hipErrorOutOfMemory) return set_err(ctx, BRT_ERR_OUT_OF_MEMORY, "%s: %s", what, hipGetErrorString(e));
    return set_err(ctx, BRT_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

int use_device(brt_context* ctx) {
    int cur = -1;
    BRT_HIP(hipGetDevice(&cur), "hipGetDevice");
    if (cur != ctx->device) BRT_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    return BRT_OK;
}

int check_grid(brt_context* ctx, const brt_grid_params* g) {
    BRT_REQUIRE(g != nullptr, "grid params are NULL");
    uint64_t cells = 1;
    for (int a = 0; a < 3; ++a) {
        BRT_REQUIRE(g->grid_size[a] >= 1, "grid_size[%d] must be >= 1", a);
        BRT_REQUIRE(g->grid_size[a] < (1u << 23), "grid_size[%d] must be < 2^23 (24-bit cell arithmetic)", a);
        BRT_REQUIRE(g->cell_size[a] > 0.0f, "cell_size[%d] must be > 0", a);
        cells *= g->grid_size[a];
    }
    BRT_REQUIRE(cells == g->num_cells, "num_cells (%u) != product of grid_size (%llu)", g->num_cells,
                (unsigned long long)cells);
    BRT_REQUIRE(cells < (1ull << 31), "too many cells (%llu)", (unsigned long long)cells);
    BRT_REQUIRE(g->max_bodies_per_cell >= 1, "max_bodies_per_cell must be >= 1");
    return BRT_OK;
}

int check_n(brt_context* ctx, uint32_t n) {
    BRT_REQUIRE(n < (1u << 30), "n = %u exceeds the 2^30 agent limit", n);
    return BRT_OK;
}

int check_n_lists(brt_context* ctx, uint32_t n) {
    BRT_REQUIRE(n < (1u << 28), "n = %u: the neighbour search handles fewer than 2^28 agents per launch", n);
    return BRT_OK;
}

int check_prox(brt_context* ctx, const brt_prox_params* p) {
    BRT_REQUIRE(p != nullptr, "proximity params are NULL");
    BRT_REQUIRE(p->max_neighbors >= 1 && p->max_neighbors <= BRT_MAX_NEIGHBORS,
                "max_neighbors must be in 1..%d (got %u)", BRT_MAX_NEIGHBORS, p->max_neighbors);
    return BRT_OK;
}

int check_steer(brt_context* ctx, const brt_steer_params* s) {
    BRT_REQUIRE(s != nullptr, "steer params are NULL");
    BRT_REQUIRE(s->num_targets <= BRT_MAX_TARGETS, "num_targets must be <= %d", BRT_MAX_TARGETS);
    BRT_REQUIRE(s->mass != 0.0f, "mass must be non-zero");
    return BRT_OK;
}

int ensure_tmp(brt_context* ctx, size_t need_tmp) {
    if (need_tmp > ctx->tmp_bytes) {
        if (ctx->tmp) BRT_HIP(hipFree(ctx->tmp), "hipFree(tmp)");
        ctx->tmp = nullptr;
        ctx->tmp_bytes = 0;
        size_t bytes = need_tmp + need_tmp / 4 + 4096;
        BRT_HIP(hipMalloc(&ctx->tmp, bytes), "hipMalloc(tmp)");
        ctx->tmp_bytes = bytes;
    }
    return BRT_OK;
}

int ensure_xyz(brt_context* ctx, uint32_t n) {
    const size_t need = (size_t)n * 12u;
    if (need > ctx->xyz_bytes) {
        if (ctx->xyz) BRT_HIP(hipFree(ctx->xyz), "hipFree(xyz)");
        ctx->xyz = nullptr;
        ctx->xyz_bytes = 0;
        const size_t bytes = need + need / 4 + 4096;
        BRT_HIP(hipMalloc((void**)&ctx->xyz, bytes), "hipMalloc(xyz)");
        ctx->xyz_bytes = bytes;
    }
    return BRT_OK;
}

int ensure_scratch(brt_context* ctx, uint32_t n, const brt::SortPlan& plan) {
    int rc = ensure_tmp(ctx, (size_t)n * 8u);
    if (rc) return rc;
    size_t need_meta = brt::sort_meta_words(plan) * 4u;
    if (need_meta > ctx->meta_bytes) {
        if (ctx->meta) BRT_HIP(hipFree(ctx->meta), "hipFree(sort meta)");
        ctx->meta = nullptr;
        ctx->meta_bytes = 0;
        size_t bytes = need_meta + need_meta / 4 + 4096;
        ctx->meta_clean = false;
        BRT_HIP(hipMalloc((void**)&ctx->meta, bytes), "hipMalloc(sort meta)");
        ctx->meta_bytes = bytes;
    }
    return BRT_OK;
}

// hash (+ histograms) then the onesweep passes; the sorted pairs end in `hash`.
// meta: sort metadata (sort_meta_words), tmp: ping-pong uint2[n]; both given
// (meta == nullptr: the context's, allocated as needed).  meta_clean: every
// word of meta is already zero (then no k_prepare runs and the hash kernel
// fills the empty cells).  The last pass leaves its metadata nonzero: a
// reorder with brt::sort_cleanup returns it to zero.
// ev_start / ev_stop (phase timing): bound to the first and the last kernel.
int run_hash_sort(brt_context* ctx, const brt_grid_params* g, const float* pos, uint32_t n, uint32_t* hash,
                  uint32_t* cell_start, uint32_t* cell_end, hipStream_t s, uint32_t* meta = nullptr,
                  void* tmp = nullptr, hipEvent_t ev_start = nullptr, hipEvent_t ev_stop = nullptr,
                  bool meta_clean = false) {
    brt::SortPlan plan = brt::make_sort_plan(n, g->num_cells);
    size_t clear_words = brt::sort_meta_words(plan);
    if (!meta) {
        int rc = ensure_scratch(ctx, n, plan);
        if (rc) return rc;
        meta = ctx->meta;
        tmp = ctx->tmp;
        meta_clean = meta_clean && ctx->meta_clean;
        ctx->meta_clean = false;
        // the whole buffer: an earlier sort of another size may have left words
        // beyond this plan's
        clear_words = ctx->meta_bytes / 4u;
    }
    brt::launch_mark().start = ev_start;
    if (!meta_clean)
        BRT_HIP(brt::launch_prepare(meta, clear_words, cell_start, cell_end, g->num_cells, s), "k_prepare");
    uint2* A = reinterpret_cast<uint2*>(hash);
    uint2* B = reinterpret_cast<uint2*>(tmp);
    uint2* first = (plan.passes & 1u) ? B : A;
    if (plan.passes == 0u) brt::launch_mark().stop = ev_stop;
    BRT_HIP(brt::launch_hash_hist(*g, reinterpret_cast<const float4*>(pos), n, first, meta, plan, s,
                                  meta_clean ? cell_start : nullptr, cell_end),
            "k_hash_hist");
    uint2* src = first;
    uint2* dst = first == A ? B : A;
    for (uint32_t q = 0; q < plan.passes; ++q) {
        if (q + 1u == plan.passes) brt::launch_mark().stop = ev_stop;
        BRT_HIP(brt::launch_onesweep(src, dst, n, q, plan, meta, ctx->dev_flags, s), "k_onesweep");
        std::swap(src, dst);
    }
    return BRT_OK;
}

int check_dev_flags(brt_context* ctx) {
    uint32_t flags = 0;
    BRT_HIP(hipDeviceSynchronize(), "hipDeviceSynchronize");
    BRT_HIP(hipMemcpy(&flags, ctx->dev_flags, sizeof(flags), hipMemcpyDeviceToHost), "hipMemcpy(device flags)");
    if (flags == 0u) return BRT_OK;
    BRT_HIP(hipMemset(ctx->dev_flags, 0, sizeof(uint32_t)), "hipMemset(device flags)");
    if (flags & BRT_DEVFLAG_SORT_LOOKBACK)
        return set_err(ctx, BRT_ERR_HIP,
                       "a radix-sort pass gave up a look-back wait (device flag 0x%x): the sort permutation of "
                       "a step since the last check is wrong",
                       flags);
    return set_err(ctx, BRT_ERR_HIP, "device flags 0x%x raised", flags);
}

void fold_events(brt_context* ctx) {
    for (size_t k = 0; k < ctx->events_used; ++k) {
        auto& ev = ctx->events[k];
        for (int p = 0; p < BRT_NUM_PHASES; ++p) {
            if (!(ctx->events_mask[k] >> p & 1u)) continue;
            float ms = 0.0f;
            (void)hipEventSynchronize(ev[2 * p + 1]);
            if (hipEventElapsedTime(&ms, ev[2 * p], ev[2 * p + 1]) == hipSuccess) ctx->phase_ms[p] += ms;
        }
        ctx->phase_steps += 1;
    }
    ctx->events_used = 0;
}

}  // namespace host
}  // namespace brt

using namespace brt::host;

extern "C" {

const char* brt_version(void) { return "behavert_amd 0.1 (gfx950)"; }

// ------------------------------------------------------------ no-allocation step
// The scratch of brt_step_noalloc: one file-scope device array (each GPU holds
// its own copy, at its own address, for the life of the process), laid out as
// sort metadata, a 16-byte-per-agent buffer and the packed positions, sized for
// BRT_NOALLOC_MAX_AGENTS agents (about 30 MB).  Host code reaches it through
// the address hipGetSymbolAddress returns on each device, kept in the context.
constexpr size_t kNoAllocMetaWords = brt::kMetaStatus + 4 * ((BRT_NOALLOC_MAX_AGENTS + 2047) / 2048) * 256;
constexpr size_t kNoAllocMetaOffset = 0;
constexpr size_t kNoAllocTmpOffset = (kNoAllocMetaOffset + kNoAllocMetaWords * 4 + 255) / 256 * 256;
constexpr size_t kNoAllocXyzOffset = (kNoAllocTmpOffset + (size_t)BRT_NOALLOC_MAX_AGENTS * 16 + 255) / 256 * 256;
constexpr size_t kNoAllocBytes = (kNoAllocXyzOffset + (size_t)BRT_NOALLOC_MAX_AGENTS * 12 + 255) / 256 * 256;

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

int brt_context_create(int device, brt_context** out) {
    if (!out) return BRT_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return BRT_ERR_HIP;
    if (device < 0 || device >= count) return BRT_ERR_INVALID_ARGUMENT;
    if (hipSetDevice(device) != hipSuccess) return BRT_ERR_HIP;
    brt_context* ctx = new brt_context();
    ctx->device = device;
    if (hipMalloc((void**)&ctx->dev_flags, 256) != hipSuccess ||
        hipMemset(ctx->dev_flags, 0, 256) != hipSuccess) {
        if (ctx->dev_flags) (void)hipFree(ctx->dev_flags);
        delete ctx;
        return BRT_ERR_HIP;
    }
    // this device's copy of the no-allocation scratch (device current)
    if (hipGetSymbolAddress(&ctx->scratch, HIP_SYMBOL(scratch)) != hipSuccess) {
        (void)hipFree(ctx->dev_flags);
        delete ctx;
        return BRT_ERR_HIP;
    }
    *out = ctx;
    return BRT_OK;
}

int brt_context_destroy(brt_context* ctx) {
    if (!ctx) return BRT_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    for (auto& ev : ctx->events)
        for (auto e : ev) (void)hipEventDestroy(e);
    if (ctx->tmp) (void)hipFree(ctx->tmp);
    if (ctx->meta) (void)hipFree(ctx->meta);
    if (ctx->partials) (void)hipFree(ctx->partials);
    if (ctx->xyz) (void)hipFree(ctx->xyz);
    if (ctx->dev_flags) (void)hipFree(ctx->dev_flags);
    for (auto& g : ctx->graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.graph) (void)hipGraphDestroy(g.graph);
    }
    if (ctx->cap_stream) (void)hipStreamDestroy(ctx->cap_stream);
    delete ctx;
    return BRT_OK;
}

const char* brt_context_last_error(const brt_context* ctx) { return ctx ? ctx->err.c_str() : "NULL context"; }

int brt_context_set_option(brt_context* ctx, int option, int value) {
    if (!ctx) return BRT_ERR_INVALID_ARGUMENT;
    switch (option) {
        case BRT_OPT_FUSED_STEP:
            BRT_REQUIRE(value >= -1, "fused_step must be -1 (automatic), 0 or 1");
            ctx->fused_step = value < 0 ? -1 : (value != 0);
            return BRT_OK;
        case BRT_OPT_UNIFORM_MAX_CELLS:
            BRT_REQUIRE(value >= 0 && value <= 64, "uniform_max_cells must be in 0..64");
            ctx->uniform_max_cells = (uint32_t)value;
            return BRT_OK;
        case BRT_OPT_XCD_REMAP:
            ctx->xcd_remap = value != 0;
            return BRT_OK;
        case BRT_OPT_LDS_LISTS:
            BRT_REQUIRE(value >= -1, "lds_lists must be -1 (automatic), 0 or 1");
            ctx->lds_lists = value < 0 ? -1 : (value != 0);
            return BRT_OK;
        case BRT_OPT_STEP_GRAPH:
            ctx->step_graph = value != 0;
            return BRT_OK;
        case BRT_OPT_DEVICE_FLAGS: {
            int rc = use_device(ctx);
            if (rc) return rc;
            const uint32_t v = (uint32_t)value;
            BRT_HIP(hipDeviceSynchronize(), "hipDeviceSynchronize");
            BRT_HIP(hipMemcpy(ctx->dev_flags, &v, sizeof(v), hipMemcpyHostToDevice), "hipMemcpy(device flags)");
            return BRT_OK;
        }
        default:
            return set_err(ctx, BRT_ERR_INVALID_ARGUMENT, "unknown option %d", option);
    }
}

int brt_context_get_option(const brt_context* ctx, int option, int* value) {
    if (!ctx || !value) return BRT_ERR_INVALID_ARGUMENT;
    switch (option) {
        case BRT_OPT_FUSED_STEP:
            *value = ctx->fused_step;
            return BRT_OK;
        case BRT_OPT_UNIFORM_MAX_CELLS:
            *value = (int)ctx->uniform_max_cells;
            return BRT_OK;
        case BRT_OPT_XCD_REMAP:
            *value = (int)ctx->xcd_remap;
            return BRT_OK;
        case BRT_OPT_LDS_LISTS:
            *value = ctx->lds_lists;
            return BRT_OK;
        case BRT_OPT_STEP_GRAPH:
            *value = ctx->step_graph;
            return BRT_OK;
        case BRT_OPT_DEVICE_FLAGS: {
            brt_context* c = const_cast<brt_context*>(ctx);
            uint32_t v = 0;
            if (use_device(c) || hipDeviceSynchronize() != hipSuccess ||
                hipMemcpy(&v, ctx->dev_flags, sizeof(v), hipMemcpyDeviceToHost) != hipSuccess)
                return BRT_ERR_HIP;
            *value = (int)v;
            return BRT_OK;
        }
        default:
            return BRT_ERR_INVALID_ARGUMENT;
    }
}

int brt_context_check(brt_context* ctx) {
    if (!ctx) return BRT_ERR_INVALID_ARGUMENT;
    int rc = use_device(ctx);
    if (rc) return rc;
    return check_dev_flags(ctx);
}

int brt_context_step_shape(const brt_context* ctx, const brt_prox_params* prox, int* fused, int* lds_lists) {
    if (!ctx || !prox || !fused || !lds_lists) return BRT_ERR_INVALID_ARGUMENT;
    *fused = resolve_fused(ctx, prox) && !prox->use_knn ? 1 : 0;
    *lds_lists = (int)resolve_lds_lists(ctx, prox);
    return BRT_OK;
}

int brt_grid_setup(const float world_center[3], const float world_radius[3], const float cell_size[3],
                   uint32_t max_bodies_per_cell, brt_grid_params* out) {
    if (!world_center || !world_radius || !cell_size || !out) return BRT_ERR_INVALID_ARGUMENT;
    std::memset(out, 0, sizeof(*out));
    uint64_t cells = 1;
    for (int a = 0; a < 3; ++a) {
        if (!(cell_size[a] > 0.0f)) return BRT_ERR_INVALID_ARGUMENT;
        // getAugmentedRadius + setupGrid: augR = R + 2*cell; size = (uint)(2*augR/cell)
        const float aug = world_radius[a] + cell_size[a] * 2.0f;
        const float cellsf = aug * 2.0f / cell_size[a];
        if (!(cellsf >= 1.0f) || cellsf >= 8388608.0f) return BRT_ERR_INVALID_ARGUMENT;
        out->grid_size[a] = (uint32_t)cellsf;
        out->world_origin[a] = world_center[a] - aug;
        out->cell_size[a] = cell_size[a];
        out->world_center[a] = world_center[a];
        out->world_radius[a] = world_radius[a];
        cells *= out->grid_size[a];
    }
    if (cells >= (1ull << 31)) return BRT_ERR_INVALID_ARGUMENT;
    out->num_cells = (uint32_t)cells;
    out->max_bodies_per_cell = max_bodies_per_cell ? max_bodies_per_cell : BRT_UNLIMITED;
    return BRT_OK;
}

uint32_t brt_list_stride(uint32_t max_neighbors) { return brt::list_stride(max_neighbors); }

float brt_radius_threshold(float r) {
    if (r != r) return INFINITY;
    if (r < 0.0f) return -INFINITY;
    if (r == INFINITY) return INFINITY;
    float t = r * r;
    if (t == INFINITY) t = 3.40282347e38f;
    while (t > 0.0f && sqrtf(t) > r) t = nextafterf(t, -INFINITY);
    while (t < INFINITY && sqrtf(nextafterf(t, INFINITY)) <= r) t = nextafterf(t, INFINITY);
    return t;
}

int brt_envgrid_calc_hash(brt_context* ctx, const brt_grid_params* grid, const float* pos, uint32_t n,
                          uint32_t* hash, void* stream) {
    if (!ctx) return BRT_ERR_INVALID_ARGUMENT;
    int rc = check_grid(ctx, grid);
    if (rc) return rc;
    if ((rc = check_n(ctx, n))) return rc;
    if (n == 0) return BRT_OK;
    BRT_REQUIRE(pos && hash && aligned16(pos) && ((uintptr_t)hash & 7u) == 0, "bad pos/hash pointer");
    if ((rc = use_device(ctx))) return rc;
    hipStream_t s = (hipStream_t)stream;
    brt::SortPlan plan = brt::make_sort_plan(n, grid->num_cells);
    if ((rc = ensure_scratch(ctx, n, plan))) return rc;
    // the histogram side product lands in scratch and is ignored
    ctx->meta_clean = false;
    BRT_HIP(brt::launch_prepare(ctx->meta, brt::kMetaStatus, nullptr, nullptr, 0, s), "k_prepare");
    BRT_HIP(brt::launch_hash_hist(*grid, reinterpret_cast<const float4*>(pos), n, reinterpret_cast<uint2*>(hash),
                                  ctx->meta, plan, s),
            "k_hash_hist");
    return BRT_OK;
}

int brt_envgrid_sort(brt_context* ctx, uint32_t* hash, uint32_t n, uint32_t num_cells, void* stream) {
    if (!ctx) return BRT_ERR_INVALID_ARGUMENT;
    int rc = check_n(ctx, n);
    if (rc) return rc;
    BRT_REQUIRE(num_cells >= 1, "num_cells must be >= 1");
    if (n < 2) return BRT_OK;
    BRT_REQUIRE(hash && ((uintptr_t)hash & 7u) == 0, "bad hash pointer");
    if ((rc = use_device(ctx))) return rc;
    hipStream_t s = (hipStream_t)stream;
    brt::SortPlan plan = brt::make_sort_plan(n, num_cells);
    if (plan.passes == 0) return BRT_OK;
    if ((rc = ensure_scratch(ctx, n, plan))) return rc;
    ctx->meta_clean = false;
    BRT_HIP(brt::launch_prepare(ctx->meta, brt::sort_meta_words(plan), nullptr, nullptr, 0, s), "k_prepare");
    uint2* A = reinterpret_cast<uint2*>(hash);
    uint2* B = reinterpret_cast<uint2*>(ctx->tmp);
    const bool odd = plan.passes & 1u;
    BRT_HIP(brt::launch_hist(A, odd ? B : nullptr, n, ctx->meta, plan, s), "k_hist");
    uint2* src = odd ? B : A;
    uint2* dst = odd ? A : B;
    for (uint32_t q = 0; q < plan.passes; ++q) {
        BRT_HIP(brt::launch_onesweep(src, dst, n, q, plan, ctx->meta, ctx->dev_flags, s), "k_onesweep");
        std::swap(src, dst);
    }
    return BRT_OK;
}

int brt_envgrid_hash_sort(brt_context* ctx, const brt_grid_params* grid, const float* pos, uint32_t n,
                          uint32_t* hash, void* stream) {
    if (!ctx) return BRT_ERR_INVALID_ARGUMENT;
    int rc = check_grid(ctx, grid);
    if (rc) return rc;
    if ((rc = check_n(ctx, n))) return rc;
    if (n == 0) return BRT_OK;
    BRT_REQUIRE(pos && hash && aligned16(pos) && ((uintptr_t)hash & 7u) == 0, "bad pos/hash pointer");
    if ((rc = use_device(ctx))) return rc;
    return run_hash_sort(ctx, grid, pos, n, hash, nullptr, nullptr, (hipStream_t)stream);
}

static int reorder_common(brt_context* ctx, const uint32_t* hash, uint32_t n, uint32_t nf, const float* const* in,
                          float* const* out, uint32_t num_cells, uint32_t* cell_start, uint32_t* cell_end,
                          void* stream) {
    int rc = check_n(ctx, n);
    if (rc) return rc;
    BRT_REQUIRE(nf <= (uint32_t)brt::kMaxReorderFeatures, "at most %d features per call", brt::kMaxReorderFeatures);
    BRT_REQUIRE(nf == 0 || (in && out), "feature pointer arrays are NULL");
    brt::ReorderArgs f{};
    f.count = nf;
    for (uint32_t q = 0; q < nf; ++q) {
        BRT_REQUIRE(in[q] && out[q] && aligned16(in[q]) && aligned16(out[q]), "feature %u pointer bad", q);
        BRT_REQUIRE(in[q] != out[q], "feature %u: in and out must differ (gather)", q);
        f.in[q] = reinterpret_cast<const float4*>(in[q]);
        f.out[q] = reinterpret_cast<float4*>(out[q]);
    }
    if ((rc = use_device(ctx))) return rc;
    hipStream_t s = (hipStream_t)stream;
    if (cell_start) {
        BRT_REQUIRE(cell_end && num_cells >= 1, "cell_end missing");
        BRT_REQUIRE(aligned16(cell_start) && aligned16(cell_end), "cell arrays must be 16-byte aligned");
        BRT_HIP(brt::launch_prepare(nullptr, 0, cell_start, cell_end, num_cells, s), "k_prepare(cells)");
    }
    if (n == 0) return BRT_OK;
    BRT_REQUIRE(hash != nullptr, "hash is NULL");
    BRT_HIP(brt::launch_reorder_ranges(reinterpret_cast<const uint2*>(hash), n, f, num_cells, cell_start, cell_end, s),
            "k_reorder_ranges");
    return BRT_OK;
}

int brt_envgrid_reorder(brt_context* ctx, const uint32_t* hash, uint32_t n, uint32_t num_features,
                        const float* const* in, float* const* out, void* stream) {
    if (!ctx) return BRT_ERR_INVALID_ARGUMENT;
    return reorder_common(ctx, hash, n, num_features, in, out, 0, nullptr, nullptr, stream);
}

int brt_prox_cell_ranges(brt_context* ctx, const uint32_t* hash, uint32_t n, uint32_t num_cells,
                         uint32_t* cell_start, uint32_t* cell_end, void* stream) {
    if (!ctx) return BRT_ERR_INVALID_ARGUMENT;
    if (!cell_start || !cell_end) return set_err(ctx, BRT_ERR_INVALID_ARGUMENT, "cell arrays are NULL");
    return reorder_common(ctx, hash, n, 0, nullptr, nullptr, num_cells, cell_start, cell_end, stream);
}

int brt_envgrid_reorder_ranges(brt_context* ctx, const uint32_t* hash, uint32_t n, uint32_t num_features,
                               const float* const* in, float* const* out, uint32_t num_cells,
                               uint32_t* cell_start, uint32_t* cell_end, void* stream) {
    if (!ctx) return BRT_ERR_INVALID_ARGUMENT;
    if (!cell_start || !cell_end) return set_err(ctx, BRT_ERR_INVALID_ARGUMENT, "cell arrays are NULL");
    return reorder_common(ctx, hash, n, num_features, in, out, num_cells, cell_start, cell_end, stream);
}

int brt_prox_neighbors(brt_context* ctx, const brt_grid_params* grid, const brt_prox_params* prox,
                       const float* sorted_pos, const uint32_t* hash, const uint32_t* cell_start,
                       const uint32_t* cell_end, uint32_t n, uint32_t* lists, void* stream) {
    if (!ctx) return BRT_ERR_INVALID_ARGUMENT;
    int rc = check_grid(ctx, grid);
    if (rc) return rc;
    if ((rc = check_prox(ctx, prox))) return rc;
    if ((rc = check_n(ctx, n))) return rc;
    if ((rc = check_n_lists(ctx, n))) return rc;
    if (n == 0) return BRT_OK;
    BRT_REQUIRE(sorted_pos && hash && cell_start && cell_end && lists, "NULL array");
    BRT_REQUIRE(aligned16(sorted_pos) && aligned16(lists), "sorted_pos and lists must be 16-byte aligned");
    if ((rc = use_device(ctx))) return rc;
    BRT_HIP(brt::launch_neighbors(*grid, *prox, brt_radius_threshold(prox->search_radius),
                                  reinterpret_cast<const float4*>(sorted_pos), reinterpret_cast<const uint2*>(hash),
                                  cell_start, cell_end, n, lists, ctx->uniform_max_cells, ctx->xcd_remap,
                                  resolve_lds_lists(ctx, prox), (hipStream_t)stream),
            "k_neighbors");
    return BRT_OK;
}

int brt_osw_steer_integrate(brt_context* ctx, const brt_grid_params* grid, const brt_prox_params* prox,
                            const brt_steer_params* steer, float dt, uint32_t step, uint32_t n,
                            const float* sorted_pos, const float* sorted_forward, const float* sorted_smooth_accel,
                            const uint32_t* hash, const uint32_t* lists, float* wander, float* pos_out,
                            float* forward_out, float* smooth_accel_out, float* force_out, void* stream) {
    if (!ctx) return BRT_ERR_INVALID_ARGUMENT;
    int rc = check_grid(ctx, grid);
    if (rc) return rc;
    if ((rc = check_prox(ctx, prox))) return rc;
    if ((rc = check_steer(ctx, steer))) return rc;
    if ((rc = check_n(ctx, n))) return rc;
    if (n == 0) return BRT_OK;
    BRT_REQUIRE(sorted_pos && sorted_forward && sorted_smooth_accel && hash && lists && pos_out && forward_out &&
                    smooth_accel_out,
                "NULL array");
    BRT_REQUIRE(aligned16(sorted_pos) && aligned16(sorted_forward) && aligned16(sorted_smooth_accel) &&
                    aligned16(lists) && aligned16(pos_out) && aligned16(forward_out) &&
                    aligned16(smooth_accel_out) && aligned16(force_out),
                "float4 arrays must be 16-byte aligned");
    BRT_REQUIRE(steer->wander_weight == 0.0f || wander != nullptr, "wander_weight != 0 needs the wander state");
    if ((rc = use_device(ctx))) return rc;
    BRT_HIP(brt::launch_steer(*grid, prox->max_neighbors, *steer, dt, step, n,
                              reinterpret_cast<const float4*>(sorted_pos),
                              reinterpret_cast<const float4*>(sorted_forward),
                              reinterpret_cast<const float4*>(sorted_smooth_accel),
                              reinterpret_cast<const uint2*>(hash), lists, reinterpret_cast<float2*>(wander),
                              reinterpret_cast<float4*>(pos_out), reinterpret_cast<float4*>(forward_out),
                              reinterpret_cast<float4*>(smooth_accel_out), reinterpret_cast<float4*>(force_out),
                              ctx->xcd_remap, (hipStream_t)stream),
            "k_steer");
    return BRT_OK;
}

int brt_osw_adjoint_step(brt_context* ctx, const brt_grid_params* grid, const brt_prox_params* prox,
                         const brt_steer_params* steer, float dt, uint32_t n, const float* sorted_pos,
                         const float* sorted_forward, const float* sorted_smooth_accel, const uint32_t* hash,
                         const uint32_t* lists, const uint32_t* cell_start, const uint32_t* cell_end,
                         const float* adj_pos_out, const float* adj_forward_out, const float* adj_smooth_accel_out,
                         float* adj_pos_in, float* adj_forward_in, float* adj_smooth_accel_in, float* weight_grad,
                         void* stream) {
    if (!ctx) return BRT_ERR_INVALID_ARGUMENT;
    int rc = check_grid(ctx, grid);
    if (rc) return rc;
    if ((rc = check_prox(ctx, prox))) return rc;
    if ((rc = check_steer(ctx, steer))) return rc;
    if ((rc = check_n(ctx, n))) return rc;
    if (steer->use_2d_projection || steer->wander_weight != 0.0f)
        return set_err(ctx, BRT_ERR_UNSUPPORTED, "the adjoint supports neither 2-D projection nor wander");
    if (n == 0) return BRT_OK;
    BRT_REQUIRE(sorted_pos && sorted_forward && sorted_smooth_accel && hash && lists && cell_start && cell_end &&
                    adj_pos_out && adj_forward_out && adj_smooth_accel_out && adj_pos_in && adj_forward_in &&
                    adj_smooth_accel_in && weight_grad,
                "NULL array");
    BRT_REQUIRE(aligned16(sorted_pos) && aligned16(sorted_forward) && aligned16(sorted_smooth_accel) &&
                    aligned16(adj_pos_out) && aligned16(adj_forward_out) && aligned16(adj_smooth_accel_out) &&
                    aligned16(adj_pos_in) && aligned16(adj_forward_in) && aligned16(adj_smooth_accel_in),
                "float4 arrays must be 16-byte aligned");
    BRT_REQUIRE(adj_pos_in != adj_pos_out && adj_forward_in != adj_forward_out &&
                    adj_smooth_accel_in != adj_smooth_accel_out,
                "adjoint input and output arrays must differ");
    if ((rc = use_device(ctx))) return rc;
    // per-workgroup weight-gradient partials
    const size_t need = ((size_t)(n + 255) / 256) * 3 * sizeof(float) + 256;
    if (need > ctx->partials_bytes) {
        if (ctx->partials) BRT_HIP(hipFree(ctx->partials), "hipFree(partials)");
        ctx->partials = nullptr;
        ctx->partials_bytes = 0;
        BRT_HIP(hipMalloc((void**)&ctx->partials, need), "hipMalloc(partials)");
        ctx->partials_bytes = need;
    }
    // each agent's own terms and neighbour-term vectors (6 x n float4) in the
    // context's scratch
    if ((rc = ensure_tmp(ctx, (size_t)n * 96u))) return rc;
    BRT_HIP(brt::launch_adjoint(*grid, *prox, brt_radius_threshold(prox->search_radius), *steer, dt, n,
                                reinterpret_cast<const float4*>(sorted_pos),
                                reinterpret_cast<const float4*>(sorted_forward),
                                reinterpret_cast<const float4*>(sorted_smooth_accel),
                                reinterpret_cast<const uint2*>(hash), lists, cell_start, cell_end,
                                reinterpret_cast<const float4*>(adj_pos_out),
                                reinterpret_cast<const float4*>(adj_forward_out),
                                reinterpret_cast<const float4*>(adj_smooth_accel_out),
                                reinterpret_cast<float4*>(adj_pos_in), reinterpret_cast<float4*>(adj_forward_in),
                                reinterpret_cast<float4*>(adj_smooth_accel_in),
                                reinterpret_cast<float4*>(ctx->tmp), ctx->partials, weight_grad, ctx->xcd_remap,
                                (hipStream_t)stream),
            "k_adjoint");
    return BRT_OK;
}

int brt_osw_steer_forces(brt_context* ctx, const brt_prox_params* prox, const brt_steer_params* steer, float dt,
                         uint32_t step, uint32_t n, const float* sorted_pos, const float* sorted_forward,
                         const uint32_t* hash, const uint32_t* lists, float* wander, float* force_out,
                         void* stream) {
    if (!ctx) return BRT_ERR_INVALID_ARGUMENT;
    int rc = check_prox(ctx, prox);
    if (rc) return rc;
    if ((rc = check_steer(ctx, steer))) return rc;
    if ((rc = check_n(ctx, n))) return rc;
    if (n == 0) return BRT_OK;
    BRT_REQUIRE(sorted_pos && sorted_forward && hash && lists && force_out, "NULL array");
    BRT_REQUIRE(aligned16(sorted_pos) && aligned16(sorted_forward) && aligned16(lists) && aligned16(force_out),
                "float4 arrays must be 16-byte aligned");
    BRT_REQUIRE(steer->wander_weight == 0.0f || wander != nullptr, "wander_weight != 0 needs the wander state");
    if ((rc = use_device(ctx))) return rc;
    BRT_HIP(brt::launch_steer_forces(prox->max_neighbors, *steer, dt, step, n,
                                     reinterpret_cast<const float4*>(sorted_pos),
                                     reinterpret_cast<const float4*>(sorted_forward),
                                     reinterpret_cast<const uint2*>(hash), lists, reinterpret_cast<float2*>(wander),
                                     reinterpret_cast<float4*>(force_out), ctx->xcd_remap, (hipStream_t)stream),
            "k_steer<forces>");
    return BRT_OK;
}

int brt_osw_apply_steering_force(brt_context* ctx, const brt_grid_params* grid, const brt_steer_params* steer,
                                 float dt, uint32_t n, const float* sorted_pos, const float* sorted_forward,
                                 const float* sorted_smooth_accel, const uint32_t* hash, const float* force_in,
                                 float* pos_out, float* forward_out, float* smooth_accel_out, float* force_out,
                                 void* stream) {
    if (!ctx) return BRT_ERR_INVALID_ARGUMENT;
    int rc = check_grid(ctx, grid);
    if (rc) return rc;
    if ((rc = check_steer(ctx, steer))) return rc;
    if ((rc = check_n(ctx, n))) return rc;
    if (n == 0) return BRT_OK;
    BRT_REQUIRE(sorted_pos && sorted_forward && sorted_smooth_accel && hash && force_in && pos_out && forward_out &&
                    smooth_accel_out,
                "NULL array");
    BRT_REQUIRE(aligned16(sorted_pos) && aligned16(sorted_forward) && aligned16(sorted_smooth_accel) &&
                    aligned16(force_in) && aligned16(pos_out) && aligned16(forward_out) &&
                    aligned16(smooth_accel_out) && aligned16(force_out),
                "float4 arrays must be 16-byte aligned");
    if ((rc = use_device(ctx))) return rc;
    BRT_HIP(brt::launch_apply(*grid, *steer, dt, n, reinterpret_cast<const float4*>(sorted_pos),
                              reinterpret_cast<const float4*>(sorted_forward),
                              reinterpret_cast<const float4*>(sorted_smooth_accel),
                              reinterpret_cast<const uint2*>(hash), reinterpret_cast<const float4*>(force_in),
                              reinterpret_cast<float4*>(pos_out), reinterpret_cast<float4*>(forward_out),
                              reinterpret_cast<float4*>(smooth_accel_out), reinterpret_cast<float4*>(force_out),
                              (hipStream_t)stream),
            "k_apply");
    return BRT_OK;
}

static int check_bc(brt_context* ctx, const brt_bc_params* bc) {
    BRT_REQUIRE(bc != nullptr, "BehaviorClustering params are NULL");
    BRT_REQUIRE(bc->feature_count >= 1 && bc->feature_count <= BRT_MAX_FEATURES, "feature_count must be in 1..%d",
                BRT_MAX_FEATURES);
    return BRT_OK;
}

int brt_bc_neigh_similarities(brt_context* ctx, const brt_prox_params* prox, const brt_bc_params* bc, uint32_t n,
                              const uint32_t* hash, const uint32_t* lists, const float* features,
                              const float* stats_in, float* stats_out, float* similarities, void* stream) {
    if (!ctx) return BRT_ERR_INVALID_ARGUMENT;
    int rc = check_prox(ctx, prox);
    if (rc) return rc;
    if ((rc = check_bc(ctx, bc))) return rc;
    if ((rc = check_n(ctx, n))) return rc;
    if (n == 0) return BRT_OK;
    BRT_REQUIRE(hash && lists && features && similarities, "NULL array");
    BRT_REQUIRE(!bc->adaptive || (stats_in && stats_out && aligned16(stats_in) && aligned16(stats_out)),
                "adaptive similarities need 16-byte aligned stats_in and stats_out");
    BRT_REQUIRE(!bc->adaptive || stats_in != stats_out, "stats_in and stats_out must differ");
    BRT_REQUIRE(((uintptr_t)features & 3u) == 0 && ((uintptr_t)similarities & 3u) == 0, "float arrays misaligned");
    if ((rc = use_device(ctx))) return rc;
    if ((rc = ensure_tmp(ctx, brt::bc_scratch_bytes(bc->feature_count, n)))) return rc;
    BRT_HIP(brt::launch_bc_similarities(prox->max_neighbors, *bc, n, reinterpret_cast<const uint2*>(hash), lists,
                                        features, reinterpret_cast<const float4*>(stats_in),
                                        reinterpret_cast<float4*>(stats_out), similarities, ctx->tmp, nullptr,
                                        nullptr, nullptr, nullptr, nullptr, (hipStream_t)stream),
            "k_bc_similarities");
    return BRT_OK;
}

int brt_bc_similarities_steer(brt_context* ctx, const brt_prox_params* prox, const brt_steer_params* steer,
                              const brt_bc_params* bc, uint32_t n, const uint32_t* hash, const uint32_t* lists,
                              const float* features, const float* stats_in, float* stats_out, float* similarities,
                              const float* sorted_pos, const float* sorted_forward, const float* force_in,
                              float* force_out, void* stream) {
    if (!ctx) return BRT_ERR_INVALID_ARGUMENT;
    int rc = check_prox(ctx, prox);
    if (rc) return rc;
    if ((rc = check_steer(ctx, steer))) return rc;
    if ((rc = check_bc(ctx, bc))) return rc;
    if ((rc = check_n(ctx, n))) return rc;
    if (n == 0) return BRT_OK;
    BRT_REQUIRE(hash && lists && features && similarities && sorted_pos && sorted_forward && force_in && force_out,
                "NULL array");
    BRT_REQUIRE(!bc->adaptive || (stats_in && stats_out && aligned16(stats_in) && aligned16(stats_out)),
                "adaptive similarities need 16-byte aligned stats_in and stats_out");
    BRT_REQUIRE(!bc->adaptive || stats_in != stats_out, "stats_in and stats_out must differ");
    BRT_REQUIRE(((uintptr_t)features & 3u) == 0 && ((uintptr_t)similarities & 3u) == 0, "float arrays misaligned");
    BRT_REQUIRE(aligned16(sorted_pos) && aligned16(sorted_forward) && aligned16(force_in) && aligned16(force_out),
                "float4 arrays must be 16-byte aligned");
    if ((rc = use_device(ctx))) return rc;
    if ((rc = ensure_tmp(ctx, brt::bc_scratch_bytes(bc->feature_count, n)))) return rc;
    BRT_HIP(brt::launch_bc_similarities(prox->max_neighbors, *bc, n, reinterpret_cast<const uint2*>(hash), lists,
                                        features, reinterpret_cast<const float4*>(stats_in),
                                        reinterpret_cast<float4*>(stats_out), similarities, ctx->tmp, steer,
                                        reinterpret_cast<const float4*>(sorted_pos),
                                        reinterpret_cast<const float4*>(sorted_forward),
                                        reinterpret_cast<const float4*>(force_in),
                                        reinterpret_cast<float4*>(force_out), (hipStream_t)stream),
            "k_bc_similarities<steer>");
    return BRT_OK;
}

int brt_bc_steer_for_clustering(brt_context* ctx, const brt_prox_params* prox, const brt_steer_params* steer,
                                const brt_bc_params* bc, uint32_t n, const float* sorted_pos,
                                const float* sorted_forward, const uint32_t* lists, const float* similarities,
                                const float* force_in, float* force_out, void* stream) {
    if (!ctx) return BRT_ERR_INVALID_ARGUMENT;
    int rc = check_prox(ctx, prox);
    if (rc) return rc;
    if ((rc = check_steer(ctx, steer))) return rc;
    if ((rc = check_bc(ctx, bc))) return rc;
    if ((rc = check_n(ctx, n))) return rc;
    if (n == 0) return BRT_OK;
    BRT_REQUIRE(sorted_pos && sorted_forward && lists && similarities && force_in && force_out, "NULL array");
    BRT_REQUIRE(aligned16(sorted_pos) && aligned16(sorted_forward) && aligned16(force_in) && aligned16(force_out),
                "float4 arrays must be 16-byte aligned");
    if ((rc = use_device(ctx))) return rc;
    BRT_HIP(brt::launch_bc_cluster_steer(prox->max_neighbors, *steer, *bc, n,
                                         reinterpret_cast<const float4*>(sorted_pos),
                                         reinterpret_cast<const float4*>(sorted_forward), lists, similarities,
                                         reinterpret_cast<const float4*>(force_in),
                                         reinterpret_cast<float4*>(force_out), (hipStream_t)stream),
            "k_bc_cluster_steer");
    return BRT_OK;
}

int brt_bc_cluster_membership(brt_context* ctx, const brt_prox_params* prox, uint32_t n, uint32_t iterations,
                              const uint32_t* hash, const uint32_t* lists, const float* sorted_pos,
                              const float* membership_in, float* membership_out, float* connector, void* stream) {
    if (!ctx) return BRT_ERR_INVALID_ARGUMENT;
    int rc = check_prox(ctx, prox);
    if (rc) return rc;
    BRT_REQUIRE(n <= (1u << 24), "n = %u: cluster labels are sorted indices held in floats, n must be <= 2^24", n);
    if (n == 0) return BRT_OK;
    BRT_REQUIRE(hash && lists && sorted_pos && membership_in && membership_out, "NULL array");
    BRT_REQUIRE(aligned16(sorted_pos) && aligned16(connector), "float4 arrays must be 16-byte aligned");
    BRT_REQUIRE(((uintptr_t)membership_in & 3u) == 0 && ((uintptr_t)membership_out & 3u) == 0,
                "membership arrays misaligned");
    if ((rc = use_device(ctx))) return rc;
    float* lbl = nullptr;
    if (iterations >= 2) {
        if ((rc = ensure_tmp(ctx, (size_t)n * 8u))) return rc;
        lbl = reinterpret_cast<float*>(ctx->tmp);
    }
    BRT_HIP(brt::launch_cluster_membership(prox->max_neighbors, n, iterations, reinterpret_cast<const uint2*>(hash),
                                           lists, reinterpret_cast<const float4*>(sorted_pos), membership_in,
                                           membership_out, reinterpret_cast<float4*>(connector), lbl,
                                           lbl ? lbl + n : nullptr, (hipStream_t)stream),
            "k_lp_iterate");
    return BRT_OK;
}

int brt_bc_cluster_hit_rate(brt_context* ctx, const brt_prox_params* prox, float same_cluster_max_radius, uint32_t n,
                            const uint32_t* hash, const uint32_t* lists, const float* sorted_pos,
                            const float* membership_in, float* membership_out, void* stream) {
    if (!ctx) return BRT_ERR_INVALID_ARGUMENT;
    int rc = check_prox(ctx, prox);
    if (rc) return rc;
    if ((rc = check_n(ctx, n))) return rc;
    if (n == 0) return BRT_OK;
    BRT_REQUIRE(hash && lists && sorted_pos && membership_in && membership_out, "NULL array");
    BRT_REQUIRE(membership_in != membership_out, "membership_in and membership_out must differ");
    BRT_REQUIRE(aligned16(sorted_pos), "sorted_pos must be 16-byte aligned");
    BRT_REQUIRE(((uintptr_t)membership_in & 3u) == 0 && ((uintptr_t)membership_out & 3u) == 0,
                "membership arrays misaligned");
    if ((rc = use_device(ctx))) return rc;
    if ((rc = ensure_tmp(ctx, (size_t)n * 4u))) return rc;
    BRT_HIP(brt::launch_hit_rate(prox->max_neighbors, same_cluster_max_radius, n, reinterpret_cast<const uint2*>(hash),
                                 lists, reinterpret_cast<const float4*>(sorted_pos), membership_in, membership_out,
                                 reinterpret_cast<float*>(ctx->tmp), (hipStream_t)stream),
            "k_hit_rate");
    return BRT_OK;
}

int brt_osw_steer_and_slow_to_avoid_neighbors(brt_context* ctx, const brt_prox_params* prox,
                                              const brt_steer_params* steer, uint32_t n, const float* sorted_pos,
                                              const float* sorted_forward, const uint32_t* lists,
                                              float* forward_out, const float* force_in, float* force_out,
                                              void* stream) {
    if (!ctx) return BRT_ERR_INVALID_ARGUMENT;
    int rc = check_prox(ctx, prox);
    if (rc) return rc;
    if ((rc = check_steer(ctx, steer))) return rc;
    if ((rc = check_n(ctx, n))) return rc;
    if (n == 0) return BRT_OK;
    BRT_REQUIRE(sorted_pos && sorted_forward && lists && forward_out && force_in && force_out, "NULL array");
    BRT_REQUIRE(aligned16(sorted_pos) && aligned16(sorted_forward) && aligned16(lists) && aligned16(forward_out) &&
                    aligned16(force_in) && aligned16(force_out),
                "float4 arrays and lists must be 16-byte aligned");
    if ((rc = use_device(ctx))) return rc;
    BRT_HIP(brt::launch_slow_avoid(prox->max_neighbors, *steer, n, reinterpret_cast<const float4*>(sorted_pos),
                                   reinterpret_cast<const float4*>(sorted_forward), lists,
                                   reinterpret_cast<float4*>(forward_out), reinterpret_cast<const float4*>(force_in),
                                   reinterpret_cast<float4*>(force_out), (hipStream_t)stream),
            "k_slow_avoid");
    return BRT_OK;
}

int brt_draw_neighborhood_color(brt_context* ctx, const brt_prox_params* prox, const float color_base[3],
                                uint32_t vertexes_per_agent, uint32_t n, const uint32_t* hash, const uint32_t* lists,
                                float* color, void* stream) {
    if (!ctx) return BRT_ERR_INVALID_ARGUMENT;
    int rc = check_prox(ctx, prox);
    if (rc) return rc;
    if ((rc = check_n(ctx, n))) return rc;
    BRT_REQUIRE(color_base != nullptr, "color_base is NULL");
    BRT_REQUIRE(vertexes_per_agent == 1u || vertexes_per_agent == 3u, "vertexes_per_agent must be 1 or 3");
    if (n == 0) return BRT_OK;
    BRT_REQUIRE(hash && lists && color, "NULL array");
    BRT_REQUIRE(aligned16(color), "color must be 16-byte aligned");
    if ((rc = use_device(ctx))) return rc;
    BRT_HIP(brt::launch_neighborhood_color(prox->max_neighbors, color_base, vertexes_per_agent, n,
                                           reinterpret_cast<const uint2*>(hash), lists,
                                           reinterpret_cast<float4*>(color), (hipStream_t)stream),
            "k_neighborhood_color");
    return BRT_OK;
}

int brt_shapes_move_towards_target(brt_context* ctx, int index_shift, const float target_base[3], uint32_t n,
                                   const uint32_t* hash, const float* sorted_pos, const float* targets,
                                   const float* force_in, float* force_out, void* stream) {
    if (!ctx) return BRT_ERR_INVALID_ARGUMENT;
    int rc = check_n(ctx, n);
    if (rc) return rc;
    BRT_REQUIRE(target_base != nullptr, "target_base is NULL");
    BRT_REQUIRE(index_shift >= 0 && (n == 0 || (uint32_t)index_shift < n), "index_shift must be in [0, n)");
    if (n == 0) return BRT_OK;
    BRT_REQUIRE(hash && sorted_pos && targets && force_in && force_out, "NULL array");
    BRT_REQUIRE(aligned16(sorted_pos) && aligned16(targets) && aligned16(force_in) && aligned16(force_out),
                "float4 arrays must be 16-byte aligned");
    if ((rc = use_device(ctx))) return rc;
    BRT_HIP(brt::launch_move_towards_target(index_shift, target_base, n, reinterpret_cast<const uint2*>(hash),
                                            reinterpret_cast<const float4*>(sorted_pos),
                                            reinterpret_cast<const float4*>(targets),
                                            reinterpret_cast<const float4*>(force_in),
                                            reinterpret_cast<float4*>(force_out), (hipStream_t)stream),
            "k_move_towards_target");
    return BRT_OK;
}

int brt_bc_feature_update(brt_context* ctx, const brt_prox_params* prox, float rate, uint32_t n,
                          const uint32_t* hash, const uint32_t* lists, const float* features_in,
                          float* features_out, void* stream) {
    if (!ctx) return BRT_ERR_INVALID_ARGUMENT;
    int rc = check_prox(ctx, prox);
    if (rc) return rc;
    if ((rc = check_n(ctx, n))) return rc;
    if (n == 0) return BRT_OK;
    BRT_REQUIRE(hash && lists && features_in && features_out, "NULL array");
    BRT_REQUIRE(features_in != features_out, "features_in and features_out must differ");
    BRT_REQUIRE(aligned16(features_in) && aligned16(features_out) && aligned16(lists), "alignment");
    if ((rc = use_device(ctx))) return rc;
    BRT_HIP(brt::launch_features(prox->max_neighbors, rate, n, reinterpret_cast<const uint2*>(hash), lists,
                                 reinterpret_cast<const float4*>(features_in),
                                 reinterpret_cast<float4*>(features_out), (hipStream_t)stream),
            "k_features");
    return BRT_OK;
}

}  // extern "C"

// step_core and the step graphs: C++ helpers of the step entry points, outside
// the extern "C" block (they return and take C++ types)
namespace {

// Scratch of one step given by the caller of step_core instead of the
// context's (the no-allocation entry point): sort metadata, a buffer of 16
// bytes per agent (sort ping-pong, then the sorted features), the packed x,y,z
// positions (12 bytes per agent).
struct StepScratch {
    uint32_t* meta;
    void* tmp;
    float* xyz;
};

int step_core(brt_context* ctx, const brt_grid_params* grid, const brt_prox_params* prox,
              const brt_steer_params* steer, float feature_rate, float dt, uint32_t step,
              const brt_step_buffers* b, void* stream, const StepScratch* given) {
    if (!ctx) return BRT_ERR_INVALID_ARGUMENT;
    int rc = check_grid(ctx, grid);
    if (rc) return rc;
    if ((rc = check_prox(ctx, prox))) return rc;
    if ((rc = check_steer(ctx, steer))) return rc;
    BRT_REQUIRE(b != nullptr, "buffers are NULL");
    if ((rc = check_n(ctx, b->n))) return rc;
    if ((rc = check_n_lists(ctx, b->n))) return rc;
    const uint32_t n = b->n;
    if (n == 0) return BRT_OK;
    BRT_REQUIRE(b->pos && b->forward && b->smooth_accel && b->sorted_pos && b->sorted_forward && b->hash &&
                    b->cell_start && b->cell_end && b->lists,
                "a required step buffer is NULL");
    BRT_REQUIRE(aligned16(b->pos) && aligned16(b->forward) && aligned16(b->smooth_accel) &&
                    aligned16(b->sorted_pos) && aligned16(b->sorted_forward) && aligned16(b->sorted_smooth_accel) &&
                    aligned16(b->lists) && aligned16(b->features) && aligned16(b->features_out) &&
                    aligned16(b->cell_start) && aligned16(b->cell_end),
                "float4 buffers must be 16-byte aligned");
    bool do_features = b->features && b->features_out && feature_rate != 0.0f;
    BRT_REQUIRE(!do_features || b->features != b->features_out, "features and features_out must differ");
    BRT_REQUIRE(steer->wander_weight == 0.0f || b->wander != nullptr, "wander_weight != 0 needs the wander state");
    if ((rc = use_device(ctx))) return rc;
    hipStream_t s = (hipStream_t)stream;

    // phase timing: each phase's start and stop events are bound to its first
    // and last kernel (no marker packets between the phases)
    std::array<hipEvent_t, 2 * BRT_NUM_PHASES>* ev = nullptr;
    uint32_t* ev_mask = nullptr;
    struct MarkGuard {
        ~MarkGuard() { brt::launch_mark() = brt::LaunchMark{}; }
    } mark_guard;
    brt::launch_mark() = brt::LaunchMark{};
    // no phase timing of a step being captured into a caller's graph (the
    // kernel-bound events are not graph nodes)
    hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
    if (ctx->timing && !given && hipStreamIsCapturing(s, &capture) != hipSuccess)
        capture = hipStreamCaptureStatusActive;
    if (ctx->timing && !given && capture == hipStreamCaptureStatusNone) {
        if (ctx->events_used == 4096) fold_events(ctx);
        if (ctx->events_used == ctx->events.size()) {
            std::array<hipEvent_t, 2 * BRT_NUM_PHASES> a;
            for (auto& e : a) BRT_HIP(hipEventCreate(&e), "hipEventCreate");
            ctx->events.push_back(a);
            ctx->events_mask.push_back(0u);
        }
        ev_mask = &ctx->events_mask[ctx->events_used];
        ev = &ctx->events[ctx->events_used++];
        *ev_mask = 0u;
    }
    // arms the phase's events for the next launch (the phase's only kernel)
    auto mark_phase = [&](int phase) {
        if (!ev) return;
        brt::launch_mark() = brt::LaunchMark{(*ev)[2 * phase], (*ev)[2 * phase + 1]};
        *ev_mask |= 1u << phase;
    };
    // the two-pass path gathers the features into sorted order with the state
    // (into the sort's ping-pong buffer, free once the sort is done), so the
    // steering pass reads the neighbours' features by sorted index
    const bool two_pass = !(resolve_fused(ctx, prox) && !prox->use_knn);
    const bool sorted_features = do_features && two_pass;
    if (!given && sorted_features && (rc = ensure_tmp(ctx, (size_t)n * 16u))) return rc;
    // with LDS lists the neighbour scan reads candidate positions from a
    // packed x,y,z copy written by the reorder (three 16-byte loads per four
    // candidates; the register-list scan measured no gain from it)
    const uint32_t lds = resolve_lds_lists(ctx, prox);
    const bool pack = !prox->use_knn && lds != 0u;
    if (!given && pack && (rc = ensure_xyz(ctx, n))) return rc;
    float* xyz = pack ? (given ? given->xyz : ctx->xyz) : nullptr;
    void* tmp = nullptr;
    if (given) {
        tmp = given->tmp;
        if ((rc = run_hash_sort(ctx, grid, b->pos, n, b->hash, b->cell_start, b->cell_end, s, given->meta, tmp)))
            return rc;
    } else {
        // phase 0: hash (+ empty-cell fill) + sort; the clearing launch only
        // when the metadata was not left zero by the previous step
        if ((rc = run_hash_sort(ctx, grid, b->pos, n, b->hash, b->cell_start, b->cell_end, s, nullptr, nullptr,
                                ev ? (*ev)[0] : nullptr, ev ? (*ev)[1] : nullptr, true)))
            return rc;
        if (ev) *ev_mask |= 1u;
        tmp = ctx->tmp;
    }
    // phase 1: reorder pos/forward/smooth_accel + cell ranges
    // (the sorted smoothed acceleration only when the caller asked for it:
    // without it the steering reads each agent's own by ID)
    brt::ReorderArgs f{};
    uint32_t nf = 0;
    f.in[nf] = reinterpret_cast<const float4*>(b->pos);
    f.out[nf++] = reinterpret_cast<float4*>(b->sorted_pos);
    f.in[nf] = reinterpret_cast<const float4*>(b->forward);
    f.out[nf++] = reinterpret_cast<float4*>(b->sorted_forward);
    if (b->sorted_smooth_accel) {
        f.in[nf] = reinterpret_cast<const float4*>(b->smooth_accel);
        f.out[nf++] = reinterpret_cast<float4*>(b->sorted_smooth_accel);
    }
    if (sorted_features) {
        f.in[nf] = reinterpret_cast<const float4*>(b->features);
        f.out[nf++] = reinterpret_cast<float4*>(tmp);
    }
    f.count = nf;
    f.xyz0 = xyz;
    mark_phase(1);
    uint32_t* sort_meta = given ? given->meta : ctx->meta;
    const brt::SortCleanup cleanup = brt::sort_cleanup(sort_meta, brt::make_sort_plan(n, grid->num_cells));
    BRT_HIP(brt::launch_reorder_ranges(reinterpret_cast<const uint2*>(b->hash), n, f, grid->num_cells, b->cell_start,
                                       b->cell_end, s, &cleanup),
            "k_reorder_ranges");
    if (!given) ctx->meta_clean = true;
    const float T = brt_radius_threshold(prox->search_radius);
    if (!two_pass) {
        // phase 2: neighbour lists + steering + integration in one pass
        mark_phase(2);
        BRT_HIP(brt::launch_neighbors_steer(*grid, *prox, T, *steer, dt, step,
                                            reinterpret_cast<const float4*>(b->sorted_pos),
                                            reinterpret_cast<const float4*>(b->sorted_forward),
                                            reinterpret_cast<const float4*>(b->sorted_smooth_accel),
                                            reinterpret_cast<const uint2*>(b->hash), b->cell_start, b->cell_end, n,
                                            b->lists, reinterpret_cast<float2*>(b->wander),
                                            reinterpret_cast<float4*>(b->pos), reinterpret_cast<float4*>(b->forward),
                                            reinterpret_cast<float4*>(b->smooth_accel), ctx->uniform_max_cells,
                                            ctx->xcd_remap, lds, s, xyz),
                "k_prox<fused>");
    } else {
        // phase 2: neighbour lists
        mark_phase(2);
        BRT_HIP(brt::launch_neighbors(*grid, *prox, T, reinterpret_cast<const float4*>(b->sorted_pos),
                                      reinterpret_cast<const uint2*>(b->hash), b->cell_start, b->cell_end, n,
                                      b->lists, ctx->uniform_max_cells, ctx->xcd_remap, lds, s, xyz),
                "k_neighbors");
        // phase 3: fused steering + integration, scattered back to agent-ID order,
        // with the feature update in the same pass over the lists
        mark_phase(3);
        BRT_HIP(brt::launch_steer(*grid, prox->max_neighbors, *steer, dt, step, n,
                                  reinterpret_cast<const float4*>(b->sorted_pos),
                                  reinterpret_cast<const float4*>(b->sorted_forward),
                                  reinterpret_cast<const float4*>(b->sorted_smooth_accel),
                                  reinterpret_cast<const uint2*>(b->hash), b->lists,
                                  reinterpret_cast<float2*>(b->wander), reinterpret_cast<float4*>(b->pos),
                                  reinterpret_cast<float4*>(b->forward), reinterpret_cast<float4*>(b->smooth_accel),
                                  nullptr, ctx->xcd_remap, s, feature_rate,
                                  do_features ? reinterpret_cast<const float4*>(b->features) : nullptr,
                                  do_features ? reinterpret_cast<float4*>(b->features_out) : nullptr,
                                  sorted_features ? reinterpret_cast<const float4*>(tmp) : nullptr),
                "k_steer");
        do_features = false;
    }
    // phase 4: BehaviorClustering feature update (fused path)
    if (do_features) {
        mark_phase(4);
        BRT_HIP(brt::launch_features(prox->max_neighbors, feature_rate, n, reinterpret_cast<const uint2*>(b->hash),
                                     b->lists, reinterpret_cast<const float4*>(b->features),
                                     reinterpret_cast<float4*>(b->features_out), s),
                "k_features");
    }
    return BRT_OK;
}

}  // namespace

namespace {

void append_raw(std::vector<unsigned char>& k, const void* v, size_t bytes) {
    const unsigned char* p = reinterpret_cast<const unsigned char*>(v);
    k.insert(k.end(), p, p + bytes);
}
#define append_key(k, v) append_raw((k), &(v), sizeof(v))

void destroy_step_graph(brt_context::StepGraph& g) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
    g.exec = nullptr;
    g.graph = nullptr;
}

// The key of a step: everything the captured kernels take by value or by
// address, apart from the context's scratch (checked separately).
std::vector<unsigned char> step_key(const brt_context* ctx, const brt_grid_params* grid,
                                    const brt_prox_params* prox, const brt_steer_params* steer, float feature_rate,
                                    float dt, uint32_t step, const brt_step_buffers* b) {
    std::vector<unsigned char> key;
    append_key(key, *b);
    append_key(key, *grid);
    append_key(key, *prox);
    append_key(key, *steer);
    append_key(key, feature_rate);
    append_key(key, dt);
    const uint32_t wstep = steer->wander_weight != 0.0f ? step : 0u;
    append_key(key, wstep);
    append_key(key, ctx->fused_step);
    append_key(key, ctx->lds_lists);
    append_key(key, ctx->uniform_max_cells);
    append_key(key, ctx->xcd_remap);
    return key;
}

// Captures the step into a graph on the context's private stream with the
// context's scratch (grown first, outside the capture) and keeps it.
int capture_step(brt_context* ctx, const brt_grid_params* grid, const brt_prox_params* prox,
                 const brt_steer_params* steer, float feature_rate, float dt, uint32_t step,
                 const brt_step_buffers* b, std::vector<unsigned char>&& key) {
    const uint32_t n = b->n;
    const bool do_features = b->features && b->features_out && feature_rate != 0.0f;
    const bool two_pass = !(resolve_fused(ctx, prox) && !prox->use_knn);
    const bool pack = !prox->use_knn && resolve_lds_lists(ctx, prox) != 0u;
    int rc = ensure_scratch(ctx, n, brt::make_sort_plan(n, grid->num_cells));
    if (rc) return rc;
    if (do_features && two_pass && (rc = ensure_tmp(ctx, (size_t)n * 16u))) return rc;
    if (pack && (rc = ensure_xyz(ctx, n))) return rc;
    if (!ctx->cap_stream)
        BRT_HIP(hipStreamCreateWithFlags(&ctx->cap_stream, hipStreamNonBlocking), "hipStreamCreate");
    if (ctx->graphs.size() == 2) {
        destroy_step_graph(ctx->graphs.front());
        ctx->graphs.erase(ctx->graphs.begin());
    }
    StepScratch sc{ctx->meta, ctx->tmp, ctx->xyz};
    BRT_HIP(hipStreamBeginCapture(ctx->cap_stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    rc = step_core(ctx, grid, prox, steer, feature_rate, dt, step, b, ctx->cap_stream, &sc);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->cap_stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_err(ctx, e, "hipStreamEndCapture");
    brt_context::StepGraph ng;
    ng.graph = g;
    ng.key = std::move(key);
    ng.tmp = ctx->tmp;
    ng.meta = ctx->meta;
    ng.xyz = ctx->xyz;
    const hipError_t ie = hipGraphInstantiate(&ng.exec, ng.graph, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
        destroy_step_graph(ng);
        return hip_err(ctx, ie, "hipGraphInstantiate");
    }
    ctx->graphs.push_back(std::move(ng));
    return BRT_OK;
}

}  // namespace

extern "C" {

int brt_step(brt_context* ctx, const brt_grid_params* grid, const brt_prox_params* prox,
             const brt_steer_params* steer, float feature_rate, float dt, uint32_t step,
             const brt_step_buffers* b, void* stream) {
    if (!ctx) return BRT_ERR_INVALID_ARGUMENT;
    hipStream_t s = (hipStream_t)stream;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    // graphs for steps without phase timing, outside any capture of the caller's
    if (!ctx->step_graph || ctx->timing || !grid || !prox || !steer || !b || b->n == 0 ||
        hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
        return step_core(ctx, grid, prox, steer, feature_rate, dt, step, b, stream, nullptr);
    std::vector<unsigned char> key = step_key(ctx, grid, prox, steer, feature_rate, dt, step, b);
    size_t hit = ctx->graphs.size();
    for (size_t k = 0; k < ctx->graphs.size(); ++k) {
        const brt_context::StepGraph& g = ctx->graphs[k];
        if (g.key == key && g.tmp == ctx->tmp && g.meta == ctx->meta && g.xyz == ctx->xyz) hit = k;
    }
    if (hit == ctx->graphs.size()) {
        // the first call with a key runs (and checks its arguments) directly;
        // the second captures it (a key seen once, e.g. moving seek targets
        // that change the parameters every step, is never captured)
        if (key != ctx->last_key) {
            ctx->last_key = key;
            return step_core(ctx, grid, prox, steer, feature_rate, dt, step, b, stream, nullptr);
        }
        int rc = use_device(ctx);
        if (rc) return rc;
        if ((rc = capture_step(ctx, grid, prox, steer, feature_rate, dt, step, b, std::move(key)))) return rc;
    } else if (hit + 1 != ctx->graphs.size()) {
        std::swap(ctx->graphs[hit], ctx->graphs.back());
    }
    BRT_HIP(hipGraphLaunch(ctx->graphs.back().exec, s), "hipGraphLaunch");
    return BRT_OK;
}

int brt_step_noalloc(brt_context* ctx, const brt_grid_params* grid, const brt_prox_params* prox,
                     const brt_steer_params* steer, float feature_rate, float dt, uint32_t step,
                     const brt_step_buffers* b, void* stream) {
    if (!ctx) return BRT_ERR_INVALID_ARGUMENT;
    BRT_REQUIRE(b != nullptr, "buffers are NULL");
    BRT_REQUIRE(grid != nullptr, "grid params are NULL");
    BRT_REQUIRE(b->n <= BRT_NOALLOC_MAX_AGENTS, "n = %u exceeds the no-allocation step's maximum of %u agents",
                b->n, (unsigned)BRT_NOALLOC_MAX_AGENTS);
    const brt::SortPlan plan = brt::make_sort_plan(b->n, grid->num_cells);
    BRT_REQUIRE(brt::sort_meta_words(plan) <= kNoAllocMetaWords,
                "the sort metadata of this step (%zu words) exceeds the no-allocation scratch (%zu words)",
                brt::sort_meta_words(plan), kNoAllocMetaWords);
    BRT_REQUIRE(ctx->scratch != nullptr, "the device scratch of the no-allocation step is not set up");
    char* base = reinterpret_cast<char*>(ctx->scratch);
    StepScratch sc;
    sc.meta = reinterpret_cast<uint32_t*>(base + kNoAllocMetaOffset);
    sc.tmp = base + kNoAllocTmpOffset;
    sc.xyz = reinterpret_cast<float*>(base + kNoAllocXyzOffset);
    return step_core(ctx, grid, prox, steer, feature_rate, dt, step, b, stream, &sc);
}

int brt_phase_timing_enable(brt_context* ctx, int enable) {
    if (!ctx) return BRT_ERR_INVALID_ARGUMENT;
    ctx->timing = enable != 0;
    for (double& v : ctx->phase_ms) v = 0.0;
    ctx->phase_steps = 0;
    ctx->events_used = 0;
    return BRT_OK;
}

int brt_phase_times(brt_context* ctx, double* ms_out, uint32_t* steps_out) {
    if (!ctx || !ms_out) return BRT_ERR_INVALID_ARGUMENT;
    int rc = use_device(ctx);
    if (rc) return rc;
    fold_events(ctx);
    for (int p = 0; p < BRT_NUM_PHASES; ++p) ms_out[p] = ctx->phase_ms[p];
    if (steps_out) *steps_out = ctx->phase_steps;
    return check_dev_flags(ctx);
}

}  // extern "C"
