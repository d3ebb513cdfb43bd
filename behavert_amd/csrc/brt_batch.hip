// brt_batch.hip -- many independent flocks in one step: the flock tables, the
// batched step (hash + sort + reorder + ranges + fused neighbours/steering
// [+ features]) captured once into a hipGraph and replayed, and the autotuner
// that picks the step's launch shapes.
//
// Layout: flocks are concatenated in the order given.  Flock f owns agent IDs
// and sorted indices [offset_f, offset_f + n_f) and cell keys [cell_base_f,
// cell_base_f + num_cells_f).  Every key of flock f is below every key of flock
// f+1, so ONE stable sort of (key, ID) over the batch leaves each flock's
// sorted run in its own index range, in the order a sort of that flock alone
// gives; the cell ranges and lists are then the flock's own, shifted by
// offset_f.  Kernels get a flock per wave (waves never straddle flocks) and read
// that flock's grid and weights with wave-uniform (scalar) loads.
//
// The reference has no batched form: it advances one flock per process
// (Samples/OpensteerDemo/plugins/BehaveRT_Flocking.cpp:177-196).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "brt_host.hpp"
#include "brt_internal.hpp"

using namespace brt::host;

struct brt_batch {
    brt_context* ctx = nullptr;
    std::vector<brt_flock_desc> desc;
    std::vector<uint32_t> offsets, cell_bases;
    uint32_t total_n = 0, total_cells = 0, waves = 0;
    brt_prox_params prox{};
    brt::FlockDev* d_flocks = nullptr;
    uint32_t* d_wave_flock = nullptr;
    uint32_t* d_step = nullptr;
    void* tmp = nullptr;       // sort ping-pong (uint2[total_n])
    float* xyz = nullptr;      // sorted positions packed as x,y,z (float[3 total_n])
    uint32_t* meta = nullptr;  // sort histograms / counters / look-back status
    size_t meta_words = 0;
    // every word of meta is zero once the work enqueued so far has run (zeroed
    // at creation; each step's sort passes and reorder return what they used)
    bool meta_clean = false;
    brt_batch_config cfg{};
    hipStream_t cap_stream = nullptr;
    // captured steps, most recently used last; a step with the feature update
    // alternates between two buffer sets (features in/out swapped), so two
    // entries keep both graphs
    struct Graph {
        std::vector<unsigned char> key;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
    };
    std::vector<Graph> graphs;
};

namespace {

uint32_t default_block(uint32_t K) {
    const uint32_t W = brt::lds_list_stride(K);
    uint32_t B = 256;
    while (B > 64 && (size_t)B * W * 4u > 40u * 1024u) B /= 2;
    return B;
}

int check_config(brt_context* ctx, const brt_batch* b, const brt_batch_config* c) {
    BRT_REQUIRE(c != nullptr, "config is NULL");
    BRT_REQUIRE(c->prox_block == 64 || c->prox_block == 128 || c->prox_block == 256,
                "prox_block must be 64, 128 or 256");
    BRT_REQUIRE(!c->lds_lists ||
                    (size_t)c->prox_block * brt::lds_list_stride(b->prox.max_neighbors) * 4u <= 64u * 1024u,
                "prox_block * LDS list stride exceeds 64 KB of LDS");
    BRT_REQUIRE(c->uniform_max_cells <= 8, "uniform_max_cells must be in 0..8");
    BRT_REQUIRE(c->sort_items == 8 || c->sort_items == 16, "sort_items must be 8 or 16");
    return BRT_OK;
}

constexpr size_t kMaxGraphs = 2;

void destroy_graph(brt_batch::Graph& g) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
    g.exec = nullptr;
    g.graph = nullptr;
}

void drop_graph(brt_batch* b) {
    for (brt_batch::Graph& g : b->graphs) destroy_graph(g);
    b->graphs.clear();
}

// The kernels of one batched step, in order, on stream s.
int enqueue_step(brt_batch* b, const brt_steer_params* steer, float feature_rate, float dt,
                 const brt_step_buffers* buf, hipStream_t s) {
    brt_context* ctx = b->ctx;
    const uint32_t n = b->total_n;
    brt::SortPlan plan = brt::make_sort_plan(n, b->total_cells, b->cfg.sort_items);
    // the clearing launch only when the metadata is not known to be zero; the
    // hash kernel fills the empty cells otherwise
    const bool clean = b->meta_clean;
    b->meta_clean = false;
    if (!clean)
        BRT_HIP(brt::launch_prepare(b->meta, b->meta_words, buf->cell_start, buf->cell_end, b->total_cells, s),
                "k_prepare");
    brt::BatchView view{b->d_flocks, b->d_wave_flock, b->waves, b->d_step};
    uint2* A = reinterpret_cast<uint2*>(buf->hash);
    uint2* B = reinterpret_cast<uint2*>(b->tmp);
    uint2* first = (plan.passes & 1u) ? B : A;
    BRT_HIP(brt::launch_batch_hash_hist(view, reinterpret_cast<const float4*>(buf->pos), first, b->meta, plan, s,
                                        clean ? buf->cell_start : nullptr, buf->cell_end, b->total_cells),
            "k_batch_hash_hist");
    uint2* src = first;
    uint2* dst = first == A ? B : A;
    for (uint32_t q = 0; q < plan.passes; ++q) {
        BRT_HIP(brt::launch_onesweep(src, dst, n, q, plan, b->meta, ctx->dev_flags, s), "k_onesweep");
        std::swap(src, dst);
    }
    brt::ReorderArgs f{};
    f.count = 2;
    f.in[0] = reinterpret_cast<const float4*>(buf->pos);
    f.out[0] = reinterpret_cast<float4*>(buf->sorted_pos);
    f.in[1] = reinterpret_cast<const float4*>(buf->forward);
    f.out[1] = reinterpret_cast<float4*>(buf->sorted_forward);
    if (buf->sorted_smooth_accel) {
        // optional: without it the steering reads each agent's own by ID
        f.in[2] = reinterpret_cast<const float4*>(buf->smooth_accel);
        f.out[2] = reinterpret_cast<float4*>(buf->sorted_smooth_accel);
        f.count = 3;
    }
    // the LDS-list scan reads candidates from the packed x,y,z copy
    float* xyz = b->cfg.lds_lists ? b->xyz : nullptr;
    f.xyz0 = xyz;
    const brt::SortCleanup cleanup = brt::sort_cleanup(b->meta, plan);
    BRT_HIP(brt::launch_reorder_ranges(reinterpret_cast<const uint2*>(buf->hash), n, f, b->total_cells,
                                       buf->cell_start, buf->cell_end, s, &cleanup),
            "k_reorder_ranges");
    b->meta_clean = true;
    BRT_HIP(brt::launch_batch_neighbors_steer(
                view, n, b->prox, brt_radius_threshold(b->prox.search_radius), *steer, dt,
                reinterpret_cast<const float4*>(buf->sorted_pos), reinterpret_cast<const float4*>(buf->sorted_forward),
                reinterpret_cast<const float4*>(buf->sorted_smooth_accel), reinterpret_cast<const uint2*>(buf->hash),
                buf->cell_start, buf->cell_end, buf->lists, reinterpret_cast<float2*>(buf->wander),
                reinterpret_cast<float4*>(buf->pos), reinterpret_cast<float4*>(buf->forward),
                reinterpret_cast<float4*>(buf->smooth_accel), b->cfg.prox_block, b->cfg.uniform_max_cells,
                b->cfg.xcd_remap, b->cfg.lds_lists, s, xyz),
            "k_prox<batch>");
    if (buf->features && buf->features_out && feature_rate != 0.0f)
        BRT_HIP(brt::launch_features(b->prox.max_neighbors, feature_rate, n, reinterpret_cast<const uint2*>(buf->hash),
                                     buf->lists, reinterpret_cast<const float4*>(buf->features),
                                     reinterpret_cast<float4*>(buf->features_out), s),
                "k_features");
    return BRT_OK;
}

template <typename T>
void append_bytes(std::vector<unsigned char>& k, const T& v) {
    const unsigned char* p = reinterpret_cast<const unsigned char*>(&v);
    k.insert(k.end(), p, p + sizeof(T));
}

int check_step_args(brt_batch* b, const brt_steer_params* steer, float feature_rate, const brt_step_buffers* buf) {
    brt_context* ctx = b->ctx;
    int rc = check_steer(ctx, steer);
    if (rc) return rc;
    BRT_REQUIRE(buf != nullptr, "buffers are NULL");
    BRT_REQUIRE(buf->n == b->total_n, "buf->n (%u) must equal the batch's total agents (%u)", buf->n, b->total_n);
    BRT_REQUIRE(buf->pos && buf->forward && buf->smooth_accel && buf->sorted_pos && buf->sorted_forward &&
                    buf->hash && buf->cell_start && buf->cell_end && buf->lists,
                "a required step buffer is NULL");
    BRT_REQUIRE(aligned16(buf->pos) && aligned16(buf->forward) && aligned16(buf->smooth_accel) &&
                    aligned16(buf->sorted_pos) && aligned16(buf->sorted_forward) &&
                    aligned16(buf->sorted_smooth_accel) && aligned16(buf->lists) && aligned16(buf->features) &&
                    aligned16(buf->features_out) && aligned16(buf->cell_start) && aligned16(buf->cell_end),
                "float4 buffers must be 16-byte aligned");
    const bool do_features = buf->features && buf->features_out && feature_rate != 0.0f;
    BRT_REQUIRE(!do_features || buf->features != buf->features_out, "features and features_out must differ");
    BRT_REQUIRE(steer->wander_weight == 0.0f || buf->wander != nullptr, "wander_weight != 0 needs the wander state");
    return BRT_OK;
}

int batch_step(brt_batch* b, const brt_steer_params* steer, float feature_rate, float dt, uint32_t step,
               const brt_step_buffers* buf, hipStream_t s) {
    brt_context* ctx = b->ctx;
    if (b->total_n == 0) return BRT_OK;
    if (steer->wander_weight != 0.0f) BRT_HIP(brt::launch_set_word(b->d_step, step, s), "k_set_word");
    if (!b->cfg.use_graph) return enqueue_step(b, steer, feature_rate, dt, buf, s);
    std::vector<unsigned char> key;
    append_bytes(key, *buf);
    append_bytes(key, *steer);
    append_bytes(key, feature_rate);
    append_bytes(key, dt);
    append_bytes(key, b->cfg);
    size_t hit = b->graphs.size();
    for (size_t k = 0; k < b->graphs.size(); ++k)
        if (b->graphs[k].key == key) hit = k;
    if (hit == b->graphs.size()) {
        if (b->graphs.size() == kMaxGraphs) {  // evict the least recently used
            destroy_graph(b->graphs.front());
            b->graphs.erase(b->graphs.begin());
        }
        BRT_HIP(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        const int rc = enqueue_step(b, steer, feature_rate, dt, buf, b->cap_stream);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return hip_err(ctx, e, "hipStreamEndCapture");
        brt_batch::Graph ng;
        ng.graph = g;
        ng.key = key;
        const hipError_t ie = hipGraphInstantiate(&ng.exec, ng.graph, nullptr, nullptr, 0);
        if (ie != hipSuccess) {
            destroy_graph(ng);
            return hip_err(ctx, ie, "hipGraphInstantiate");
        }
        b->graphs.push_back(ng);
    } else if (hit + 1 != b->graphs.size()) {
        std::swap(b->graphs[hit], b->graphs.back());
    }
    // a graph captured from clean metadata holds no clearing launch: clear
    // first if the metadata is not known to be zero (each replay leaves it so)
    if (!b->meta_clean) BRT_HIP(brt::launch_prepare(b->meta, b->meta_words, nullptr, nullptr, 0, s), "k_prepare");
    BRT_HIP(hipGraphLaunch(b->graphs.back().exec, s), "hipGraphLaunch");
    b->meta_clean = true;
    return BRT_OK;
}

// FNV-1a over the tuning-relevant description of the batch
std::string tune_key(const brt_batch* b, int device) {
    hipDeviceProp_t prop{};
    std::string arch = "unknown";
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) arch = prop.gcnArchName;
    uint64_t h = 1469598103934665603ull;
    auto mix = [&h](const void* p, size_t len) {
        const unsigned char* c = static_cast<const unsigned char*>(p);
        for (size_t i = 0; i < len; ++i) {
            h ^= c[i];
            h *= 1099511628211ull;
        }
    };
    mix(&b->prox.max_neighbors, sizeof(uint32_t));
    for (const brt_flock_desc& d : b->desc) {
        mix(&d.n, sizeof(d.n));
        mix(d.grid.grid_size, sizeof(d.grid.grid_size));
        mix(&d.grid.max_bodies_per_cell, sizeof(uint32_t));
    }
    const std::string a = arch.substr(0, arch.find(':'));
    char buf[96];
    snprintf(buf, sizeof(buf), "%s-%u-%016llx", a.c_str(), (unsigned)b->desc.size(), (unsigned long long)h);
    return buf;
}

bool cache_lookup(const char* path, const std::string& key, brt_batch_config* out) {
    FILE* f = fopen(path, "r");
    if (!f) return false;
    char line[256];
    bool found = false;
    while (fgets(line, sizeof(line), f)) {
        char k[128];
        unsigned pb, um, xr, si, ll;
        if (sscanf(line, "%127s %u %u %u %u %u", k, &pb, &um, &xr, &si, &ll) == 6 && key == k) {
            out->prox_block = pb;
            out->uniform_max_cells = um;
            out->xcd_remap = xr;
            out->sort_items = si;
            out->lds_lists = ll;
            found = true;  // the last entry for a key wins
        }
    }
    fclose(f);
    return found;
}

void cache_store(const char* path, const std::string& key, const brt_batch_config& c, double ms) {
    FILE* f = fopen(path, "a");
    if (!f) return;
    fprintf(f, "%s %u %u %u %u %u %.6f\n", key.c_str(), c.prox_block, c.uniform_max_cells, c.xcd_remap,
            c.sort_items, c.lds_lists, ms);
    fclose(f);
}

}  // namespace

extern "C" {

int brt_batch_create(brt_context* ctx, const brt_flock_desc* flocks, uint32_t num_flocks,
                     const brt_prox_params* prox, brt_batch** out) {
    if (!ctx || !out) return BRT_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    BRT_REQUIRE(flocks != nullptr && num_flocks >= 1, "at least one flock is needed");
    int rc = check_prox(ctx, prox);
    if (rc) return rc;
    BRT_REQUIRE(!prox->use_knn, "batches support capped (non-KNN) lists only");
    uint64_t total_n = 0, total_cells = 0, waves = 0;
    for (uint32_t f = 0; f < num_flocks; ++f) {
        if ((rc = check_grid(ctx, &flocks[f].grid))) return rc;
        total_n += flocks[f].n;
        total_cells += flocks[f].grid.num_cells;
        waves += (flocks[f].n + 63u) / 64u;
    }
    BRT_REQUIRE(total_n < (1ull << 28), "a batch holds fewer than 2^28 agents (32-bit byte offsets)");
    BRT_REQUIRE(total_cells < (1ull << 31), "a batch holds fewer than 2^31 cells");
    if ((rc = use_device(ctx))) return rc;

    brt_batch* b = new brt_batch();
    b->ctx = ctx;
    b->desc.assign(flocks, flocks + num_flocks);
    b->prox = *prox;
    b->total_n = (uint32_t)total_n;
    b->total_cells = (uint32_t)total_cells;
    b->waves = (uint32_t)waves;
    std::vector<brt::FlockDev> fd(num_flocks);
    std::vector<uint32_t> wave_flock(waves > 0 ? waves : 1, 0u);
    uint32_t off = 0, cb = 0, wb = 0;
    for (uint32_t f = 0; f < num_flocks; ++f) {
        brt::FlockDev& d = fd[f];
        std::memset(&d, 0, sizeof(d));
        d.g = flocks[f].grid;
        brt::cell_slack(d.g, d.slack);
        d.offset = off;
        d.n = flocks[f].n;
        d.cell_base = cb;
        d.wave_base = wb;
        for (int k = 0; k < 3; ++k) d.w[k] = flocks[f].weights[k];
        b->offsets.push_back(off);
        b->cell_bases.push_back(cb);
        const uint32_t fw = (flocks[f].n + 63u) / 64u;
        for (uint32_t w = 0; w < fw; ++w) wave_flock[wb + w] = f;
        off += flocks[f].n;
        cb += flocks[f].grid.num_cells;
        wb += fw;
    }
    auto fail = [&](int code) {
        brt_batch_destroy(b);
        return code;
    };
    hipError_t e;
    if ((e = hipMalloc((void**)&b->d_flocks, fd.size() * sizeof(brt::FlockDev))) != hipSuccess ||
        (e = hipMalloc((void**)&b->d_wave_flock, wave_flock.size() * 4u)) != hipSuccess ||
        (e = hipMalloc((void**)&b->d_step, 256)) != hipSuccess)
        return fail(hip_err(ctx, e, "hipMalloc(batch tables)"));
    if ((e = hipMemcpy(b->d_flocks, fd.data(), fd.size() * sizeof(brt::FlockDev), hipMemcpyHostToDevice)) !=
            hipSuccess ||
        (e = hipMemcpy(b->d_wave_flock, wave_flock.data(), wave_flock.size() * 4u, hipMemcpyHostToDevice)) !=
            hipSuccess ||
        (e = hipMemset(b->d_step, 0, 256)) != hipSuccess)
        return fail(hip_err(ctx, e, "upload of batch tables"));
    // scratch sized for either sort tile, owned by the batch so that captured
    // graphs never see a reallocation
    for (uint32_t items : {8u, 16u})
        b->meta_words = std::max(b->meta_words, brt::sort_meta_words(brt::make_sort_plan(b->total_n,
                                                                                         b->total_cells, items)));
    if ((e = hipMalloc(&b->tmp, (size_t)(b->total_n > 0 ? b->total_n : 1) * 8u)) != hipSuccess ||
        (e = hipMalloc((void**)&b->xyz, (size_t)(b->total_n > 0 ? b->total_n : 1) * 12u)) != hipSuccess ||
        (e = hipMalloc((void**)&b->meta, b->meta_words * 4u)) != hipSuccess ||
        (e = hipMemset(b->meta, 0, b->meta_words * 4u)) != hipSuccess)
        return fail(hip_err(ctx, e, "hipMalloc(batch sort scratch)"));
    b->meta_clean = true;
    if ((e = hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking)) != hipSuccess)
        return fail(hip_err(ctx, e, "hipStreamCreateWithFlags"));
    b->cfg.prox_block = default_block(prox->max_neighbors);
    b->cfg.uniform_max_cells = 1;
    b->cfg.xcd_remap = 1;
    b->cfg.sort_items = 16;
    b->cfg.lds_lists = 1;
    b->cfg.use_graph = 1;
    *out = b;
    return BRT_OK;
}

int brt_batch_destroy(brt_batch* b) {
    if (!b) return BRT_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(b->ctx->device);
    drop_graph(b);
    if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
    if (b->d_flocks) (void)hipFree(b->d_flocks);
    if (b->d_wave_flock) (void)hipFree(b->d_wave_flock);
    if (b->d_step) (void)hipFree(b->d_step);
    if (b->tmp) (void)hipFree(b->tmp);
    if (b->xyz) (void)hipFree(b->xyz);
    if (b->meta) (void)hipFree(b->meta);
    delete b;
    return BRT_OK;
}

int brt_batch_layout(const brt_batch* b, uint32_t* total_agents, uint32_t* total_cells, uint32_t* offsets,
                     uint32_t* cell_bases) {
    if (!b) return BRT_ERR_INVALID_ARGUMENT;
    if (total_agents) *total_agents = b->total_n;
    if (total_cells) *total_cells = b->total_cells;
    if (offsets) std::copy(b->offsets.begin(), b->offsets.end(), offsets);
    if (cell_bases) std::copy(b->cell_bases.begin(), b->cell_bases.end(), cell_bases);
    return BRT_OK;
}

int brt_batch_set_config(brt_batch* b, const brt_batch_config* cfg) {
    if (!b) return BRT_ERR_INVALID_ARGUMENT;
    brt_context* ctx = b->ctx;
    int rc = check_config(ctx, b, cfg);
    if (rc) return rc;
    b->cfg = *cfg;
    b->cfg.xcd_remap = cfg->xcd_remap != 0;
    b->cfg.lds_lists = cfg->lds_lists != 0;
    b->cfg.use_graph = cfg->use_graph != 0;
    return BRT_OK;
}

int brt_batch_get_config(const brt_batch* b, brt_batch_config* cfg) {
    if (!b || !cfg) return BRT_ERR_INVALID_ARGUMENT;
    *cfg = b->cfg;
    return BRT_OK;
}

int brt_batch_step(brt_batch* b, const brt_steer_params* steer, float feature_rate, float dt, uint32_t step,
                   const brt_step_buffers* buf, void* stream) {
    if (!b) return BRT_ERR_INVALID_ARGUMENT;
    brt_context* ctx = b->ctx;
    int rc = check_step_args(b, steer, feature_rate, buf);
    if (rc) return rc;
    if ((rc = use_device(ctx))) return rc;
    return batch_step(b, steer, feature_rate, dt, step, buf, (hipStream_t)stream);
}

int brt_batch_autotune(brt_batch* b, const brt_steer_params* steer, float dt, const brt_step_buffers* buf,
                       uint32_t steps_per_trial, const char* cache_path, brt_batch_config* chosen, void* stream) {
    if (!b) return BRT_ERR_INVALID_ARGUMENT;
    brt_context* ctx = b->ctx;
    int rc = check_step_args(b, steer, 0.0f, buf);
    if (rc) return rc;
    BRT_REQUIRE(steps_per_trial >= 1 && steps_per_trial <= 1000, "steps_per_trial must be in 1..1000");
    if ((rc = use_device(ctx))) return rc;
    hipStream_t s = (hipStream_t)stream;
    const std::string key = tune_key(b, ctx->device);
    brt_batch_config best = b->cfg;
    if (cache_path && cache_lookup(cache_path, key, &best) && check_config(ctx, b, &best) == BRT_OK) {
        best.use_graph = b->cfg.use_graph;  // the caller's choice, as after a search
        b->cfg = best;
        drop_graph(b);
        if (chosen) *chosen = best;
        return BRT_OK;
    }
    if (b->total_n == 0) {
        if (chosen) *chosen = b->cfg;
        return BRT_OK;
    }
    // save the state the trials overwrite
    const size_t bytes = (size_t)b->total_n * 16u;
    const bool wander = steer->wander_weight != 0.0f && buf->wander;
    float* saved = nullptr;
    BRT_HIP(hipMalloc((void**)&saved, bytes * 3 + (wander ? (size_t)b->total_n * 8u : 0)), "hipMalloc(tune save)");
    float* dst[4] = {saved, saved + 4 * (size_t)b->total_n, saved + 8 * (size_t)b->total_n,
                     saved + 12 * (size_t)b->total_n};
    float* src[4] = {buf->pos, buf->forward, buf->smooth_accel, buf->wander};
    const size_t sizes[4] = {bytes, bytes, bytes, (size_t)b->total_n * 8u};
    const int nsave = wander ? 4 : 3;
    auto copy_all = [&](bool restore) -> hipError_t {
        for (int k = 0; k < nsave; ++k) {
            hipError_t e = hipMemcpyAsync(restore ? src[k] : dst[k], restore ? dst[k] : src[k], sizes[k],
                                          hipMemcpyDeviceToDevice, s);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    };
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t he = copy_all(false);
    if (he == hipSuccess) he = hipEventCreate(&e0);
    if (he == hipSuccess) he = hipEventCreate(&e1);
    double best_ms = 1e30;
    const brt_batch_config keep = b->cfg;
    std::vector<brt_batch_config> space;
    for (uint32_t ll : {1u, 0u})
        for (uint32_t pb : {64u, 128u, 256u})
            for (uint32_t um : {0u, 1u, 2u})
                for (uint32_t xr : {0u, 1u})
                    for (uint32_t si : {8u, 16u}) space.push_back(brt_batch_config{pb, um, xr, si, ll, 1u});
    for (const brt_batch_config& c : space) {
        if (he != hipSuccess) break;
        if (check_config(ctx, b, &c) != BRT_OK) continue;
        b->cfg = c;
        if ((he = copy_all(true)) != hipSuccess) break;
        if ((rc = batch_step(b, steer, 0.0f, dt, 0, buf, s))) break;  // capture + warm-up
        if ((he = hipEventRecord(e0, s)) != hipSuccess) break;
        for (uint32_t t = 0; t < steps_per_trial && !rc; ++t)
            rc = batch_step(b, steer, 0.0f, dt, 1 + t, buf, s);
        if (rc) break;
        if ((he = hipEventRecord(e1, s)) != hipSuccess) break;
        if ((he = hipEventSynchronize(e1)) != hipSuccess) break;
        float ms = 0.0f;
        if ((he = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) break;
        const double per = (double)ms / steps_per_trial;
        if (per < best_ms) {
            best_ms = per;
            best = c;
        }
    }
    if (he == hipSuccess && !rc) he = copy_all(true);
    if (he == hipSuccess) he = hipStreamSynchronize(s);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(saved);
    if (rc || he != hipSuccess) {
        b->cfg = keep;
        drop_graph(b);
        return rc ? rc : hip_err(ctx, he, "autotune");
    }
    best.use_graph = keep.use_graph;
    b->cfg = best;
    drop_graph(b);
    if (cache_path) cache_store(cache_path, key, best, best_ms);
    if (chosen) *chosen = best;
    return BRT_OK;
}

}  // extern "C"
