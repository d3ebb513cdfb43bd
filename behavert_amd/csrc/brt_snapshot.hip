// brt_snapshot.hip -- binary state snapshots of a flock: the per-agent state
// by agent ID, optionally the sorted hash and neighbour lists of the last
// step, the 4-D features and the wander state, behind a fixed 256-byte header.
//
// The reference keeps no state on disk (its only host round trips are
// DeviceArrayWrapper::copyArrayFromDevice/ToDevice,
// Core/include/DeviceArrayWrapper.h:345-386); the format is this library's,
// for fixtures, checkpoints and resuming a run bit for bit.
//
// Layout (little-endian):
//   0   char[8]  magic "BRTSNAP\0"
//   8   u32      version (1)            12  u32 header bytes (256)
//   16  u32      n                      20  u32 max_neighbors (K)
//   24  u32      list stride W          28  u32 flags (BRT_SNAP_*)
//   32  u32      step                   36  u32 sizeof(brt_grid_params) (68)
//   40  brt_grid_params                 108 brt_prox_params (12 bytes)
//   120 u64      payload bytes          128 u64 FNV-1a 64 of the payload
//   136 zeros up to 256
// payload: pos, forward, smooth_accel (float4[n] each, by ID); with
// BRT_SNAP_LISTS hash (uint2[n], sorted) and lists (uint32[n*W]); with
// BRT_SNAP_FEATURES features (float4[n] by ID); with BRT_SNAP_WANDER wander
// (float2[n] by ID).
#include <cstdio>
#include <cstring>
#include <exception>
#include <vector>

#include "brt_host.hpp"

using namespace brt::host;

namespace {

constexpr char kMagic[8] = {'B', 'R', 'T', 'S', 'N', 'A', 'P', '\0'};
constexpr uint32_t kVersion = 1;
constexpr size_t kHeader = 256;

struct Section {
    void* dev;
    size_t bytes;
};

std::vector<Section> sections(const brt_step_buffers* b, uint32_t W, uint32_t flags) {
    const size_t n = b->n;
    std::vector<Section> s = {{b->pos, 16 * n}, {b->forward, 16 * n}, {b->smooth_accel, 16 * n}};
    if (flags & BRT_SNAP_LISTS) {
        s.push_back({b->hash, 8 * n});
        s.push_back({b->lists, 4 * n * W});
    }
    if (flags & BRT_SNAP_FEATURES) s.push_back({b->features, 16 * n});
    if (flags & BRT_SNAP_WANDER) s.push_back({b->wander, 8 * n});
    return s;
}

// bytes of the sections a file with these fields holds (sections() order)
uint64_t payload_size(uint64_t n, uint64_t W, uint32_t flags) {
    uint64_t b = 48 * n;
    if (flags & BRT_SNAP_LISTS) b += 8 * n + 4 * n * W;
    if (flags & BRT_SNAP_FEATURES) b += 16 * n;
    if (flags & BRT_SNAP_WANDER) b += 8 * n;
    return b;
}

uint64_t fnv1a(const unsigned char* p, size_t len, uint64_t h = 0xcbf29ce484222325ull) {
    for (size_t i = 0; i < len; ++i) {
        h ^= p[i];
        h *= 0x100000001b3ull;
    }
    return h;
}

void put32(unsigned char* h, size_t off, uint32_t v) { std::memcpy(h + off, &v, 4); }
void put64(unsigned char* h, size_t off, uint64_t v) { std::memcpy(h + off, &v, 8); }
uint32_t get32(const unsigned char* h, size_t off) {
    uint32_t v;
    std::memcpy(&v, h + off, 4);
    return v;
}
uint64_t get64(const unsigned char* h, size_t off) {
    uint64_t v;
    std::memcpy(&v, h + off, 8);
    return v;
}

// header fields into `info`; BRT_OK or an error message in `err`
int parse_header(const unsigned char* h, brt_snapshot_info* info, std::string& err) {
    if (std::memcmp(h, kMagic, 8) != 0) {
        err = "not a behavert_amd snapshot (bad magic)";
        return BRT_ERR_INVALID_ARGUMENT;
    }
    if (get32(h, 8) != kVersion || get32(h, 12) != kHeader || get32(h, 36) != sizeof(brt_grid_params)) {
        err = "unsupported snapshot version or header layout";
        return BRT_ERR_UNSUPPORTED;
    }
    std::memset(info, 0, sizeof(*info));
    info->version = kVersion;
    info->n = get32(h, 16);
    info->max_neighbors = get32(h, 20);
    info->list_stride = get32(h, 24);
    info->flags = get32(h, 28);
    info->step = get32(h, 32);
    std::memcpy(&info->grid, h + 40, sizeof(brt_grid_params));
    std::memcpy(&info->prox, h + 108, sizeof(brt_prox_params));
    info->payload_bytes = get64(h, 120);
    info->checksum = get64(h, 128);
    if ((info->flags & ~(BRT_SNAP_LISTS | BRT_SNAP_FEATURES | BRT_SNAP_WANDER)) != 0) {
        err = "unknown section flags";
        return BRT_ERR_INVALID_ARGUMENT;
    }
    if (info->max_neighbors < 1 || info->max_neighbors > BRT_MAX_NEIGHBORS ||
        info->list_stride != brt::list_stride(info->max_neighbors)) {
        err = "invalid max_neighbors / list stride";
        return BRT_ERR_INVALID_ARGUMENT;
    }
    // the payload size follows from n, W and the flags; a header that says
    // otherwise is corrupt (and must not size an allocation)
    if (info->payload_bytes != payload_size(info->n, info->list_stride, info->flags)) {
        err = "payload size does not match n, list stride and flags";
        return BRT_ERR_INVALID_ARGUMENT;
    }
    return BRT_OK;
}

}  // namespace

extern "C" {

int brt_snapshot_save(brt_context* ctx, const char* path, const brt_grid_params* grid, const brt_prox_params* prox,
                      uint32_t step, uint32_t flags, const brt_step_buffers* buf, void* stream) {
    if (!ctx) return BRT_ERR_INVALID_ARGUMENT;
    BRT_REQUIRE(path && grid && prox && buf, "NULL argument");
    BRT_REQUIRE((flags & ~(BRT_SNAP_LISTS | BRT_SNAP_FEATURES | BRT_SNAP_WANDER)) == 0, "unknown snapshot flags");
    int rc = check_prox(ctx, prox);
    if (rc) return rc;
    BRT_REQUIRE(buf->pos && buf->forward && buf->smooth_accel, "NULL state array");
    BRT_REQUIRE(!(flags & BRT_SNAP_LISTS) || (buf->hash && buf->lists), "BRT_SNAP_LISTS needs hash and lists");
    BRT_REQUIRE(!(flags & BRT_SNAP_FEATURES) || buf->features, "BRT_SNAP_FEATURES needs features");
    BRT_REQUIRE(!(flags & BRT_SNAP_WANDER) || buf->wander, "BRT_SNAP_WANDER needs wander");
    if ((rc = use_device(ctx))) return rc;
    const uint32_t W = brt::list_stride(prox->max_neighbors);
    const std::vector<Section> secs = sections(buf, W, flags);
    size_t total = 0;
    for (const Section& s : secs) total += s.bytes;
    std::vector<unsigned char> payload(total);
    size_t off = 0;
    for (const Section& s : secs) {
        BRT_HIP(hipMemcpyAsync(payload.data() + off, s.dev, s.bytes, hipMemcpyDeviceToHost, (hipStream_t)stream),
                "hipMemcpyAsync(snapshot)");
        off += s.bytes;
    }
    BRT_HIP(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize(snapshot)");
    if ((rc = check_dev_flags(ctx))) return rc;
    unsigned char h[kHeader];
    std::memset(h, 0, sizeof(h));
    std::memcpy(h, kMagic, 8);
    put32(h, 8, kVersion);
    put32(h, 12, (uint32_t)kHeader);
    put32(h, 16, buf->n);
    put32(h, 20, prox->max_neighbors);
    put32(h, 24, W);
    put32(h, 28, flags);
    put32(h, 32, step);
    put32(h, 36, (uint32_t)sizeof(brt_grid_params));
    std::memcpy(h + 40, grid, sizeof(brt_grid_params));
    std::memcpy(h + 108, prox, sizeof(brt_prox_params));
    put64(h, 120, (uint64_t)total);
    put64(h, 128, fnv1a(payload.data(), total));
    FILE* f = std::fopen(path, "wb");
    BRT_REQUIRE(f != nullptr, "cannot open %s for writing", path);
    const bool ok = std::fwrite(h, 1, kHeader, f) == kHeader && std::fwrite(payload.data(), 1, total, f) == total;
    const bool closed = std::fclose(f) == 0;
    BRT_REQUIRE(ok && closed, "writing %s failed", path);
    return BRT_OK;
}

int brt_snapshot_read_info(const char* path, brt_snapshot_info* info) {
    if (!path || !info) return BRT_ERR_INVALID_ARGUMENT;
    FILE* f = std::fopen(path, "rb");
    if (!f) return BRT_ERR_INVALID_ARGUMENT;
    unsigned char h[kHeader];
    const bool ok = std::fread(h, 1, kHeader, f) == kHeader;
    std::fclose(f);
    if (!ok) return BRT_ERR_INVALID_ARGUMENT;
    std::string err;
    return parse_header(h, info, err);
}

int brt_snapshot_load(brt_context* ctx, const char* path, uint32_t flags, const brt_prox_params* prox,
                      const brt_step_buffers* buf, brt_snapshot_info* info_out, void* stream) {
    if (!ctx) return BRT_ERR_INVALID_ARGUMENT;
    BRT_REQUIRE(path && buf, "NULL argument");
    FILE* f = std::fopen(path, "rb");
    BRT_REQUIRE(f != nullptr, "cannot open %s", path);
    unsigned char h[kHeader];
    if (std::fread(h, 1, kHeader, f) != kHeader) {
        std::fclose(f);
        return set_err(ctx, BRT_ERR_INVALID_ARGUMENT, "%s: truncated header", path);
    }
    brt_snapshot_info info;
    std::string err;
    int rc = parse_header(h, &info, err);
    if (rc) {
        std::fclose(f);
        return set_err(ctx, rc, "%s: %s", path, err.c_str());
    }
    if (info.n != buf->n || (flags & ~info.flags) != 0 ||
        info.list_stride != brt::list_stride(info.max_neighbors)) {
        std::fclose(f);
        return set_err(ctx, BRT_ERR_INVALID_ARGUMENT,
                       "%s: holds n = %u, flags 0x%x; asked for n = %u, flags 0x%x", path, info.n, info.flags,
                       buf->n, flags);
    }
    // the destination lists hold brt_list_stride(prox->max_neighbors) slots per
    // agent: only a file with the same K can be copied into them
    if ((flags & BRT_SNAP_LISTS) && (!prox || prox->max_neighbors != info.max_neighbors)) {
        std::fclose(f);
        return set_err(ctx, BRT_ERR_INVALID_ARGUMENT,
                       "%s: lists were saved with K = %u; the destination lists are sized for K = %u", path,
                       info.max_neighbors, prox ? prox->max_neighbors : 0u);
    }
    std::vector<unsigned char> payload;
    try {
        payload.resize(info.payload_bytes);
    } catch (const std::exception&) {
        std::fclose(f);
        return set_err(ctx, BRT_ERR_OUT_OF_MEMORY, "%s: cannot hold a payload of %llu bytes", path,
                       (unsigned long long)info.payload_bytes);
    }
    const bool ok = std::fread(payload.data(), 1, payload.size(), f) == payload.size();
    std::fclose(f);
    BRT_REQUIRE(ok, "%s: truncated payload", path);
    BRT_REQUIRE(fnv1a(payload.data(), payload.size()) == info.checksum, "%s: checksum mismatch", path);
    // walk every section present in the file; copy the requested ones
    const std::vector<Section> all = sections(buf, info.list_stride, info.flags);
    size_t expect = 0;
    for (const Section& s : all) expect += s.bytes;
    BRT_REQUIRE(expect == payload.size(), "%s: payload size does not match its header", path);
    if ((rc = use_device(ctx))) return rc;
    size_t off = 0;
    int idx = 0;
    auto wanted = [&](int k) {
        // sections: 0..2 state, then lists (2 sections), features, wander in flag order
        if (k < 3) return true;
        int next = 3;
        if (info.flags & BRT_SNAP_LISTS) {
            if (k < next + 2) return (flags & BRT_SNAP_LISTS) != 0;
            next += 2;
        }
        if (info.flags & BRT_SNAP_FEATURES) {
            if (k == next) return (flags & BRT_SNAP_FEATURES) != 0;
            ++next;
        }
        return (flags & BRT_SNAP_WANDER) != 0;
    };
    for (const Section& s : all) {
        if (wanted(idx)) {
            BRT_REQUIRE(s.dev != nullptr, "a requested snapshot section has no destination array");
            BRT_HIP(hipMemcpyAsync(s.dev, payload.data() + off, s.bytes, hipMemcpyHostToDevice, (hipStream_t)stream),
                    "hipMemcpyAsync(snapshot)");
        }
        off += s.bytes;
        ++idx;
    }
    BRT_HIP(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize(snapshot)");
    if ((rc = check_dev_flags(ctx))) return rc;
    if (info_out) *info_out = info;
    return BRT_OK;
}

}  // extern "C"
