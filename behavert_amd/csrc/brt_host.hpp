// brt_host.hpp -- the library context and the host-side helpers shared by the
// C ABI translation units (status codes, validation, scratch management).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <string>
#include <vector>

#include "../../include/behavert_amd.h"
#include "brt_internal.hpp"

struct brt_context {
    int device = 0;
    std::string err;
    void* tmp = nullptr;        // sort ping-pong buffer (uint2[n]); label ping-pong (float[2n])
    size_t tmp_bytes = 0;
    uint32_t* meta = nullptr;   // histograms, tile counters, look-back status
    size_t meta_bytes = 0;
    // every word of meta is zero once the work enqueued so far has run: the
    // last sort on it was a step's, whose passes and reorder return what they
    // used to zero; then the next step's sort needs no clearing launch
    bool meta_clean = false;
    float* partials = nullptr;  // adjoint: per-workgroup weight-gradient partials
    size_t partials_bytes = 0;
    float* xyz = nullptr;       // brt_step: the sorted positions packed as x,y,z
    size_t xyz_bytes = 0;
    uint32_t* dev_flags = nullptr;  // sticky device-side failure flags (BRT_DEVFLAG_*)
    void* scratch = nullptr;        // this device's no-allocation step scratch (a static device array)
    bool timing = false;
    // per timed step: a start and a stop event per phase (bound to the
    // phase's first and last kernel) and the mask of the phases that ran
    std::vector<std::array<hipEvent_t, 2 * BRT_NUM_PHASES>> events;
    std::vector<uint32_t> events_mask;
    size_t events_used = 0;
    double phase_ms[BRT_NUM_PHASES] = {0, 0, 0, 0, 0};
    uint32_t phase_steps = 0;
    int fused_step = -1;        // -1: automatic (resolve_fused)
    uint32_t uniform_max_cells = 1;
    uint32_t xcd_remap = 1;
    int lds_lists = -1;         // -1: automatic (resolve_lds_lists)
    // brt_step's captured graphs (BRT_OPT_STEP_GRAPH), most recently used last,
    // and the key of the last call (a graph is captured on the second call
    // with a key)
    int step_graph = 0;
    hipStream_t cap_stream = nullptr;
    struct StepGraph {
        std::vector<unsigned char> key;
        void* tmp = nullptr;  // the scratch the graph was captured with
        uint32_t* meta = nullptr;
        float* xyz = nullptr;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
    };
    std::vector<StepGraph> graphs;
    std::vector<unsigned char> last_key;
};

// Automatic launch shapes, from the K = 32 and K = 64 workloads measured on
// gfx950: at K = 32 the fused neighbour + steering kernel and LDS-staged lists
// are fastest; at K = 64 the two-pass path with register-chunk lists is (the
// fused kernel's registers and the LDS list tile cut occupancy).
inline bool resolve_fused(const brt_context* ctx, const brt_prox_params* p) {
    return ctx->fused_step >= 0 ? ctx->fused_step != 0 : p->max_neighbors <= 32u;
}
inline uint32_t resolve_lds_lists(const brt_context* ctx, const brt_prox_params* p) {
    return ctx->lds_lists >= 0 ? (uint32_t)(ctx->lds_lists != 0) : (p->max_neighbors <= 32u ? 1u : 0u);
}

namespace brt {
namespace host {

int set_err(brt_context* ctx, int code, const char* fmt, ...);
int hip_err(brt_context* ctx, hipError_t e, const char* what);
int use_device(brt_context* ctx);
int check_grid(brt_context* ctx, const brt_grid_params* g);
int check_n(brt_context* ctx, uint32_t n);
// the neighbour kernels address positions with 32-bit byte offsets
int check_n_lists(brt_context* ctx, uint32_t n);
int check_prox(brt_context* ctx, const brt_prox_params* p);
int check_steer(brt_context* ctx, const brt_steer_params* s);
int ensure_tmp(brt_context* ctx, size_t bytes);
int ensure_scratch(brt_context* ctx, uint32_t n, const brt::SortPlan& plan);
// synchronises the device, reads and clears the sticky flags; BRT_ERR_HIP with
// a message when one was raised
int check_dev_flags(brt_context* ctx);

inline bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

}  // namespace host
}  // namespace brt

#define BRT_HIP(call, what)                              \
    do {                                                 \
        hipError_t e_ = (call);                          \
        if (e_ != hipSuccess) return hip_err(ctx, e_, what); \
    } while (0)

#define BRT_REQUIRE(cond, ...)                                              \
    do {                                                                    \
        if (!(cond)) return set_err(ctx, BRT_ERR_INVALID_ARGUMENT, __VA_ARGS__); \
    } while (0)

