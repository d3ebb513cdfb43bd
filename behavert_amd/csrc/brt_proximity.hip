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
// MI355X design.  One lane per sorted agent, 256 lanes per workgroup (64 for
// LDS-list launches of 2^18 agents or more, see prox_block); lists are staged
// in LDS and written out with 16-byte coalesced stores, or gathered in
// registers as 16-byte chunks.  Per wave64:
//   * uniform mode (all lanes in few cells): lanes are grouped by cell; a group
//     walks the shared canonical candidate stream with wave-uniform indices, so
//     candidate positions come through the scalar cache (s_load);
//   * flat mode (LDS lists): each lane walks its own stream of cells with its
//     own cell index, eight (dense waves) or four (sparse waves) candidates
//     per round, their accept bits logged per round;
//   * vector mode (register-chunk lists): each lane walks its own stream, the
//     wave in step cell by cell.
// Every mode skips a cell whose box lies farther than r from the agent (with
// rounding margins so no candidate that passes the exact test is ever skipped),
// only for interior cells: boundary cells can hold clamped out-of-grid or NaN
// agents whose positions are not inside the cell's box.
#include "brt_device.hpp"
#include "brt_internal.hpp"
#include "brt_steer_common.hpp"

// Candidates tested per round of the scan loops (their loads are in flight
// together): per-lane (vector) rounds of dense waves with LDS lists, of sparse
// waves and register-chunk lists, and wave-uniform (scalar) rounds.  Measured
// on gfx950: 8 per round for dense waves with LDS lists (1M-boid blobs, K 32:
// -11 % against 4), 4 for sparse waves (65k boids at ~1 per cell: 8 is 26 %
// slower) and with register-chunk lists (8 spills there).
#ifndef BRT_PROX_GROUP
#define BRT_PROX_GROUP 8
#endif
#ifndef BRT_PROX_GROUP_R
#define BRT_PROX_GROUP_R 4
#endif
#ifndef BRT_PROX_GROUP_U
#define BRT_PROX_GROUP_U 4
#endif
// 1: in dense waves with LDS lists, cells that cannot hold the agent itself run
// rounds without the self test (a second copy of the round code, one ballot
// per cell; elsewhere that measured slower); 0: every round tests for the agent.
#ifndef BRT_PROX_SELF_SPLIT
#define BRT_PROX_SELF_SPLIT 1
#endif
// Neighbours gathered together (positions and headings in flight) by the fused
// steering walk over a finished list.
#ifndef BRT_PROX_STEER_GROUP
#define BRT_PROX_STEER_GROUP 4
#endif
// 1: the loops over the 27 neighbour cells stay loops (one copy of the round
// code per variant; a slab's cell ranges are read from registers by a
// wave-uniform index); 0: they are unrolled, one copy of the round code per cell.
#ifndef BRT_PROX_ROLLED
#define BRT_PROX_ROLLED 1
#endif
// 1: LDS lists are built through a per-round log of accept bits (see
// log_round) when the launch has fewer than 2^24 agents; 0: every accepted
// candidate is appended as it is tested.
#ifndef BRT_PROX_LOG
#define BRT_PROX_LOG 1
#endif
// 1: dense waves with the round log scan in flat mode (see scan_flat).
#ifndef BRT_PROX_FLAT
#define BRT_PROX_FLAT 1
#endif
// 1: sparse waves with the round log scan in flat mode too.
#ifndef BRT_PROX_FLAT_SPARSE
#define BRT_PROX_FLAT_SPARSE 1
#endif
// Workgroup of the single-grid LDS-list neighbour kernels for launches of 2^18
// agents or more (see prox_block).
#ifndef BRT_PROX_BLOCK_LARGE
#define BRT_PROX_BLOCK_LARGE 64
#endif
// A wave is dense (rounds of BRT_PROX_GROUP candidates) when at least
// BRT_PROX_DENSE_LANES of its lanes sit in cells of BRT_PROX_DENSE_OCC agents
// or more.
#ifndef BRT_PROX_DENSE_OCC
#define BRT_PROX_DENSE_OCC 8
#endif
#ifndef BRT_PROX_DENSE_LANES
#define BRT_PROX_DENSE_LANES 32
#endif
#if BRT_PROX_ROLLED
#define BRT_CELL_LOOP _Pragma("nounroll")
#else
#define BRT_CELL_LOOP _Pragma("unroll")
#endif

namespace brt {

namespace {

constexpr int kGroupL = BRT_PROX_GROUP;
constexpr int kGroupR = BRT_PROX_GROUP_R;
constexpr int kGroupU = BRT_PROX_GROUP_U;

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
    brt_grid_params g;  // single-grid launches
    uint32_t K, W, n;   // n = all agents of the launch (a batch: of every flock)
    uint32_t Ws;        // row stride of the LDS list tile (lds_list_stride(K) >= W)
    float T;            // exact squared-radius threshold
    float Tskip;        // T * (1 + 1e-5): cell-skip threshold on the box distance
    float slack[3];     // per-axis widening of cell boxes for rounding (single grid)
    uint32_t uniform_max_cells;
    uint32_t xcd_remap;
    uint32_t log_rounds;  // LDS lists through the round log (n < 2^24)
    const float4* spos;
    const float* pxyz;  // optional: the sorted positions packed as x,y,z (12 bytes each)
    const float4* sfwd;
    const float4* ssa;  // sorted smooth_accel, or NULL: read by agent ID from o.sa_out
    const uint2* hash;
    const uint32_t* cell_start;
    const uint32_t* cell_end;
    uint32_t* lists;
    // fused steering
    brt_steer_params sp;
    SteerCuts cuts;
    float dt;
    uint32_t step;
    AgentOut o;
    // batches: per-flock grids, weights and wave ranges
    BatchView b;
};

// The agents one wave works on and the grid they live in (wave-uniform).
struct Seg {
    brt_grid_params g;
    float slack[3];
    uint32_t cell_base;  // global key of the grid's cell 0
    uint32_t row0;       // sorted index of the wave's lane 0
    uint32_t end;        // one past the segment's last sorted index
    uint32_t id_base;    // agent ID of the segment's first agent
    Weights w;
};

template <bool BATCH>
__device__ __forceinline__ bool wave_segment(const ProxArgs& a, uint32_t gw, Seg& sg) {
    if (!BATCH) {
        sg.g = a.g;
        sg.slack[0] = a.slack[0];
        sg.slack[1] = a.slack[1];
        sg.slack[2] = a.slack[2];
        sg.cell_base = 0;
        sg.row0 = gw * 64u;
        sg.end = a.n;
        sg.id_base = 0;
        sg.w = weights_of(a.sp);
        return sg.row0 < a.n;
    }
    if (gw >= a.b.waves) return false;
    const uint32_t f = __builtin_amdgcn_readfirstlane(a.b.wave_flock[gw]);
    const FlockDev& fd = a.b.flocks[f];
    sg.g = fd.g;
    sg.slack[0] = fd.slack[0];
    sg.slack[1] = fd.slack[1];
    sg.slack[2] = fd.slack[2];
    sg.cell_base = fd.cell_base;
    sg.row0 = fd.offset + (gw - fd.wave_base) * 64u;
    sg.end = fd.offset + fd.n;
    sg.id_base = fd.offset;
    sg.w = Weights{fd.w[0], fd.w[1], fd.w[2]};
    return true;
}

struct Lane {
    float px, py, pz;
    uint32_t i;
    uint32_t cnt;
    uint32_t* row;       // the agent's list row: staged in LDS, or in global memory
    uint32_t* wp;        // LDS lists: the next free entry (row + 1 + cnt) during a round
    uint32_t h0, h1, h2;  // register-chunk lists: entries 0..2 (chunk 0, with the count)
    uint32_t b0, b1, b2;  // register-chunk lists: entries of the pending 16-byte chunk
    // squared distances from the agent to the boxes of the cells one below
    // (lo) and one above (hi) its own on each axis, each gap shrunk by the
    // rounding slack: the box distance of neighbour (dx,dy,dz) is the sum of
    // the three selected terms (offset 0 contributes 0)
    float lo2[3], hi2[3];
};

__device__ __forceinline__ bool interior(int cx, int cy, int cz, const brt_grid_params& g) {
    return cx >= 1 && cy >= 1 && cz >= 1 && cx <= (int)g.grid_size[0] - 2 && cy <= (int)g.grid_size[1] - 2 &&
           cz <= (int)g.grid_size[2] - 2;
}

__device__ __forceinline__ void cell_gaps(const Seg& sg, Lane& L, I3 gp) {
    const float p[3] = {L.px, L.py, L.pz};
    const int c[3] = {gp.x, gp.y, gp.z};
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const float lo = sg.g.world_origin[ax] + (float)c[ax] * sg.g.cell_size[ax];
        const float f = p[ax] - lo;
        const float gl = fmaxf(f - sg.slack[ax], 0.0f);
        const float gh = fmaxf((sg.g.cell_size[ax] - f) - sg.slack[ax], 0.0f);
        L.lo2[ax] = gl * gl;
        L.hi2[ax] = gh * gh;
    }
}

// The lo or hi term of axis ax for offset d (0 for d = 0), chosen by bit masks
// rather than a select of the two fields, which the compiler may turn into a
// select of their addresses and so keep the lane's state in scratch memory
// when d is not a compile-time constant.
__device__ __forceinline__ float gap2(const Lane& L, int ax, int d) {
    const uint32_t mlo = d < 0 ? 0xFFFFFFFFu : 0u, mhi = d > 0 ? 0xFFFFFFFFu : 0u;
    return __uint_as_float((__float_as_uint(L.lo2[ax]) & mlo) | (__float_as_uint(L.hi2[ax]) & mhi));
}

// Appends sorted index c to the lane's list when take.  LDS lists: one LDS
// word per entry.  Register-chunk lists: the row's 16-byte chunk q holds words
// 4q..4q+3, i.e. entries 4q-1..4q+2; entries are gathered in registers and a
// chunk is stored to global memory when its last entry arrives (chunk 0, which
// holds the count, and the final partial chunk are stored by finish_list).
template <bool LDSL>
__device__ __forceinline__ void append(Lane& L, uint32_t c, bool take) {
    if (LDSL) {
        if (take) L.row[1 + L.cnt] = c;
    } else if (take) {
        const uint32_t k = L.cnt;
        if (k < 3u) {
            L.h0 = k == 0u ? c : L.h0;
            L.h1 = k == 1u ? c : L.h1;
            L.h2 = k == 2u ? c : L.h2;
        } else {
            const uint32_t s = (k - 3u) & 3u;
            if (s == 3u) {
                reinterpret_cast<uint4*>(L.row)[(k - 3u) / 4u + 1u] = make_uint4(L.b0, L.b1, L.b2, c);
            } else {
                L.b0 = s == 0u ? c : L.b0;
                L.b1 = s == 1u ? c : L.b1;
                L.b2 = s == 2u ? c : L.b2;
            }
        }
    }
    L.cnt += take ? 1u : 0u;
}

// Count and padding (LDS lists), or the chunks not yet stored (register-chunk
// lists: the pending partial chunk, then chunk 0).  Register-chunk lists
// return the row's first all-padding chunk: those chunks are stored by the
// whole wave together (pad_rows).
template <bool LDSL>
__device__ __forceinline__ uint32_t finish_list(Lane& L, uint32_t W) {
    if (LDSL) {
        L.row[0] = L.cnt;
        for (uint32_t q = L.cnt + 1; q < W; ++q) L.row[q] = BRT_LIST_PAD;
        return W / 4u;
    }
    const uint32_t P = BRT_LIST_PAD, k = L.cnt;
    uint4* row = reinterpret_cast<uint4*>(L.row);
    uint32_t q = (k > 3u ? (k - 3u) / 4u : 0u) + 1u;
    const uint32_t rem = k > 3u ? (k - 3u) & 3u : 0u;
    if (4u * q < W) {
        row[q] = make_uint4(rem > 0u ? L.b0 : P, rem > 1u ? L.b1 : P, rem > 2u ? L.b2 : P, P);
        ++q;
    }
    row[0] = make_uint4(k, k > 0u ? L.h0 : P, k > 1u ? L.h1 : P, k > 2u ? L.h2 : P);
    return q;
}

// Register-chunk lists: the all-padding chunks of the wave's rows (consecutive
// rows of W words from sorted index row0; lane l's row starts padding at its
// chunk pad_from), stored by the whole wave with consecutive 16-byte chunks per
// store instead of each lane's own strided row.  The trip count is
// wave-uniform, so every lane takes part in the shuffle.
__device__ __forceinline__ void pad_rows(uint32_t* lists, uint32_t row0, uint32_t rows, uint32_t W, uint32_t lane,
                                         uint32_t pad_from) {
    const uint32_t W4 = W / 4u, total = rows * W4;
    const uint32_t P = BRT_LIST_PAD;
    uint4* dst = reinterpret_cast<uint4*>(lists + (size_t)row0 * W);
    for (uint32_t base = 0; base < total; base += 64u) {
        const uint32_t w = base + lane;
        const uint32_t r = w / W4;
        const uint32_t first = (uint32_t)__shfl((int)pad_from, (int)(r < 64u ? r : 63u));
        if (w < total && w - r * W4 >= first) dst[w] = make_uint4(P, P, P, P);
    }
}

// Tests the candidates j .. j+G-1 (positions q) for lane L and appends
// those that pass, in order, without branching.  rem = lim - j (candidates
// left in the cell), di = i - j (where the agent itself sits in the round).
// SELF: the agent's own index lies in this cell's range, so it is excluded.
// CAP (some lane may reach K inside this round): 1 = every append re-checks
// the count; 2 = the LDS row has room for G - 1 entries past K (stride
// Ws >= K + G, every lane entering the round below K: the last write lands at
// most at row[K + G - 1]), so the round appends freely and the count is
// clamped to K afterwards: entries past K are overwritten by padding.
// LDS lists append through a write pointer: a predicated store and a pointer
// bump per candidate.
template <bool LDSL, int CAP, bool SELF, int G>
__device__ __forceinline__ void test_round(const ProxArgs& a, Lane& L, uint32_t j, uint32_t rem, uint32_t di,
                                           const float4* q) {
    if (LDSL) L.wp = L.row + 1u + L.cnt;
#pragma unroll
    for (int t = 0; t < G; ++t) {
        const float rx = q[t].x - L.px, ry = q[t].y - L.py, rz = q[t].z - L.pz;
        const float xx = rx * rx, yy = ry * ry, zz = rz * rz;
        const float d2 = (xx + yy) + zz;
        bool take = ((uint32_t)t < rem) & !(d2 > a.T);
        if (SELF) take = take & (di != (uint32_t)t);
        if (LDSL) {
            if (CAP == 1) take = take & (L.wp < L.row + 1u + a.K);
            if (take) *L.wp = j + (uint32_t)t;
            L.wp += take ? 1 : 0;
        } else {
            if (CAP) take = take & (L.cnt < a.K);
            append<false>(L, j + (uint32_t)t, take);
        }
    }
    if (LDSL) {
        const uint32_t c = (uint32_t)(L.wp - (L.row + 1u));
        L.cnt = CAP == 2 ? (c < a.K ? c : a.K) : c;
    }
}

// m = 2m + bit in one VALU add-with-carry whose carry-in is the wave's mask of
// the compare that produced bit (the ballot folds into that compare).
__device__ __forceinline__ uint32_t shift_in(uint32_t m, bool bit) {
    const uint64_t c = __builtin_amdgcn_ballot_w64(bit);
    uint32_t r;
    uint64_t carry_out;
    asm("v_addc_co_u32_e64 %0, %1, %2, %2, %3" : "=v"(r), "=s"(carry_out) : "v"(m), "s"(c));
    return r;
}

// Round log (LDS lists, sorted indices below 2^24): the round's accept bits are
// collected with one add-with-carry per candidate (candidate j+t ends in bit
// 7-t), the limit and self tests are applied to the round's bits at once, and a
// round with any bit set appends the word j << 8 | bits to the agent's row.
// The running count includes every accepted bit; a lane stops before a round
// once it holds K, so at most K words are appended (each carries at least one
// bit) and they fit in the row.  expand_log turns the words into the list.
template <int G, bool SELF>
__device__ __forceinline__ void log_round(const ProxArgs& a, Lane& L, uint32_t j, uint32_t rem, uint32_t di,
                                          const float4* q) {
    static_assert(G >= 1 && G <= 8, "a round's bits must fit in 8 bits");
    uint32_t m = 0;
#pragma unroll
    for (int t = 0; t < G; ++t) {
        const float rx = q[t].x - L.px, ry = q[t].y - L.py, rz = q[t].z - L.pz;
        const float xx = rx * rx, yy = ry * ry, zz = rz * rz;
        const float d2 = (xx + yy) + zz;
        m = shift_in(m, !(d2 > a.T));
    }
    // slots t >= rem (rem >= 1) lie past the cell's limit: bits G-1-t below G-rem
    const uint32_t r = rem < (uint32_t)G ? rem : (uint32_t)G;
    m &= ~((1u << ((uint32_t)G - r)) - 1u);
    if (SELF && di < (uint32_t)G) m &= ~(1u << ((uint32_t)G - 1u - di));
    m <<= 8 - G;
    if (m != 0u) {
        *L.wp = (j << 8) | m;
        ++L.wp;
    }
    L.cnt += __popc(m);
}

// Round log of the flat scan: as log_round, with the self test folded into a
// mask the caller selects per round (all ones unless the round holds the
// agent), and the word stored unconditionally, the write pointer advancing only
// when the word carries a bit.  A word is stored only while the lane holds
// fewer than K, so at most K words are kept and the store lands at most at
// row[K] < W; a stored word without bits is overwritten by the next word or
// by the list's padding.
template <int G>
__device__ __forceinline__ void log_round_flat(const ProxArgs& a, Lane& L, uint32_t j, uint32_t rem, uint32_t sm,
                                               const float4* q) {
    static_assert(G >= 1 && G <= 8, "a round's bits must fit in 8 bits");
    uint32_t m = 0;
#pragma unroll
    for (int t = 0; t < G; ++t) {
        const float rx = q[t].x - L.px, ry = q[t].y - L.py, rz = q[t].z - L.pz;
        const float xx = rx * rx, yy = ry * ry, zz = rz * rz;
        const float d2 = (xx + yy) + zz;
        m = shift_in(m, !(d2 > a.T));
    }
    const uint32_t r = rem < (uint32_t)G ? rem : (uint32_t)G;
    m &= ~((1u << ((uint32_t)G - r)) - 1u) & sm;
    m <<= 8 - G;
    *L.wp = (j << 8) | m;
    L.wp += m != 0u ? 1 : 0;
    L.cnt += __popc(m);
}

// Expands the round log into the list, last word first and in place: word e
// sits at row[1+e] and its entries go to row[1+p] for p >= the bits of words
// 0..e-1 >= e, so no unread word is overwritten.  Entries past K are dropped
// and the count is clamped to K.
__device__ __forceinline__ void expand_log(Lane& L, uint32_t K) {
    const uint32_t nlog = (uint32_t)(L.wp - (L.row + 1u));
    uint32_t pos = L.cnt;
    for (uint32_t e = nlog; e-- > 0u;) {
        const uint32_t w = L.row[1u + e];
        uint32_t m = w & 0xFFu;
        const uint32_t j7 = (w >> 8) + 7u;
        pos -= __popc(m);
        uint32_t p = pos;
        while (m != 0u) {
            const uint32_t b = 31u - __clz(m);  // highest bit: the earliest candidate
            if (p < K) L.row[1u + p] = j7 - b;
            m ^= 1u << b;
            ++p;
        }
    }
    L.cnt = L.cnt < K ? L.cnt : K;
}

template <bool LDSL, bool SELF, int G>
__device__ __forceinline__ void test_round_any(const ProxArgs& a, Lane& L, uint32_t j, uint32_t rem, uint32_t di,
                                               const float4* q) {
    if (LDSL && a.log_rounds) {
        log_round<G, SELF>(a, L, j, rem, di, q);
        return;
    }
    if (__ballot(L.cnt + (uint32_t)G > a.K) == 0)
        test_round<LDSL, 0, SELF, G>(a, L, j, rem, di, q);
    else if (LDSL && a.Ws >= a.K + (uint32_t)G)
        test_round<LDSL, 2, SELF, G>(a, L, j, rem, di, q);
    else
        test_round<LDSL, 1, SELF, G>(a, L, j, rem, di, q);
}

// Buffer resource over a float4 array of n elements: 32-bit offsets with an
// immediate part, and reads past the end return zeros instead of faulting.
// Requires n < 2^28.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t float4_rsrc(const float4* p, uint32_t n) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float4*>(p), (short)0, (int)(n * 16u), 0x00020000);
}

__device__ __forceinline__ float4 bload(__amdgpu_buffer_rsrc_t r, uint32_t byte_offset) {
    const vfloat4 v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)byte_offset, 0, 0);
    return make_float4(v.x, v.y, v.z, v.w);
}

__device__ __forceinline__ __amdgpu_buffer_rsrc_t xyz_rsrc(const float* p, uint32_t n) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(p), (short)0, (int)(n * 12u), 0x00020000);
}

// The positions of candidates j .. j+G-1.  PACKED: from the packed x,y,z copy,
// 3G/4 16-byte loads at byte 12j (4-byte aligned; the wave's load count, not
// its bytes, is what the address unit spends its cycles on); else one 16-byte
// load per candidate.  Reads past the end return zeros.
template <int G, bool PACKED>
__device__ __forceinline__ void load_round(__amdgpu_buffer_rsrc_t rpos, __amdgpu_buffer_rsrc_t rxyz, uint32_t j,
                                           float4 (&q)[G]) {
    if (PACKED) {
        static_assert(G % 4 == 0, "packed rounds take multiples of 4 candidates");
        constexpr int NV = 3 * G / 4;
        float f[3 * G];
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            const vfloat4 v = __builtin_amdgcn_raw_buffer_load_b128(rxyz, (int)(12u * j + 16u * (uint32_t)k), 0, 0);
            f[4 * k + 0] = v.x;
            f[4 * k + 1] = v.y;
            f[4 * k + 2] = v.z;
            f[4 * k + 3] = v.w;
        }
#pragma unroll
        for (int t = 0; t < G; ++t) q[t] = make_float4(f[3 * t], f[3 * t + 1], f[3 * t + 2], 0.0f);
    } else {
#pragma unroll
        for (int t = 0; t < G; ++t) q[t] = bload(rpos, j * 16u + 16u * (uint32_t)t);
    }
}

// Uniform mode: every active lane shares cell (ux,uy,uz); control flow and
// candidate indices are wave-uniform.
template <bool LDSL>
__device__ __forceinline__ void scan_uniform(const ProxArgs& a, const Seg& sg, Lane& L, int ux, int uy, int uz) {
    const brt_grid_params& g = sg.g;
    cfloat4_ptr cpos = (cfloat4_ptr)a.spos;
    cuint_ptr cstart = (cuint_ptr)a.cell_start;
    cuint_ptr cend = (cuint_ptr)a.cell_end;
    const uint32_t maxBodies = g.max_bodies_per_cell;
    BRT_CELL_LOOP
    for (int dx = -1; dx <= 1; ++dx) {
        BRT_CELL_LOOP
        for (int dy = -1; dy <= 1; ++dy) {
            BRT_CELL_LOOP
            for (int dz = -1; dz <= 1; ++dz) {
                const int cx = ux + dx, cy = uy + dy, cz = uz + dz;
                if (!inGrid(cx, cy, cz, g)) continue;
                const uint32_t key = sg.cell_base + cellKey(cx, cy, cz, g);
                const uint32_t s = cstart[key];
                if (s == BRT_EMPTY_CELL) continue;
                const uint32_t e = cend[key];
                uint32_t lim = (e - s) > maxBodies ? s + maxBodies : e;
                lim = lim < a.n ? lim : a.n;
                if (interior(cx, cy, cz, g)) {
                    const float bd = (gap2(L, 0, dx) + gap2(L, 1, dy)) + gap2(L, 2, dz);
                    const bool need = (L.cnt < a.K) & !(bd > a.Tskip);
                    if (__ballot(need) == 0) continue;
                }
                const bool self_here = s <= L.i && L.i < lim;
                const bool any_self = !(LDSL && BRT_PROX_SELF_SPLIT) || __ballot(self_here) != 0;
                for (uint32_t j = s; j < lim; j += kGroupU) {
                    float4 q[kGroupU];
#pragma unroll
                    for (int t = 0; t < kGroupU; ++t) q[t] = cload(cpos, (j + t < lim) ? j + t : lim - 1);
                    if (L.cnt < a.K) {
                        if (any_self)
                            test_round_any<LDSL, true, kGroupU>(a, L, j, lim - j, L.i - j, q);
                        else
                            test_round_any<LDSL, false, kGroupU>(a, L, j, lim - j, L.i - j, q);
                    }
                    if (__ballot(L.cnt < a.K) == 0) return;
                }
            }
        }
    }
}

// Vector mode: each lane walks its own canonical stream.  The nine cell ranges
// of one x-offset slab are loaded together (one memory latency per slab rather
// than two per cell).
template <bool LDSL, int G, bool SPLIT, bool PACKED>
__device__ __forceinline__ void scan_vector(const ProxArgs& a, const Seg& sg, Lane& L, I3 gp) {
    const brt_grid_params& g = sg.g;
    const uint32_t maxBodies = g.max_bodies_per_cell;
    const __amdgpu_buffer_rsrc_t rpos = float4_rsrc(a.spos, a.n);
    const __amdgpu_buffer_rsrc_t rxyz = xyz_rsrc(PACKED ? a.pxyz : nullptr, PACKED ? a.n : 0u);
    BRT_CELL_LOOP
    for (int dx = -1; dx <= 1; ++dx) {
        const int cx = gp.x + dx;
        uint32_t st[9], en[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const int cy = gp.y + k / 3 - 1, cz = gp.z + k % 3 - 1;
            const bool ok = inGrid(cx, cy, cz, g);
            const uint32_t key = sg.cell_base + (ok ? cellKey(cx, cy, cz, g) : 0u);
            st[k] = a.cell_start[key];
            en[k] = a.cell_end[key];
            st[k] = ok ? st[k] : BRT_EMPTY_CELL;
        }
        BRT_CELL_LOOP
        for (int k = 0; k < 9; ++k) {
            const int dy = k / 3 - 1, dz = k % 3 - 1;
            const uint32_t s = st[k];
            if (s == BRT_EMPTY_CELL) continue;
            const float bd = (gap2(L, 0, dx) + gap2(L, 1, dy)) + gap2(L, 2, dz);
            if (bd > a.Tskip && interior(cx, gp.y + dy, gp.z + dz, g)) continue;
            const uint32_t e = en[k];
            uint32_t lim = (e - s) > maxBodies ? s + maxBodies : e;
            lim = lim < a.n ? lim : a.n;
            // the agent's own index is in this range only in its own (stored) cell
            const bool any_self = !SPLIT || __ballot(s <= L.i && L.i < lim) != 0;
            for (uint32_t j = s; j < lim; j += G) {
                // candidates past lim are read (zeros past n) and not taken
                float4 q[G];
                load_round<G, PACKED>(rpos, rxyz, j, q);
                if (any_self)
                    test_round_any<LDSL, true, G>(a, L, j, lim - j, L.i - j, q);
                else
                    test_round_any<LDSL, false, G>(a, L, j, lim - j, L.i - j, q);
                if (L.cnt >= a.K) return;
            }
        }
    }
}

// Flat mode (LDS round log): each lane walks its own canonical stream with its
// own cell index, so a lane that finishes a cell goes on to its next cell while
// the others are still in theirs.  The wave then runs about as many rounds as
// its longest stream needs, instead of the sum over the 27 cells of the longest
// stream in each.  The cells to visit (in the grid and not skipped) are a
// 27-bit mask computed up front; the range of a lane's next cell is loaded
// while it scans the current one.
// Bit d+1 set when coordinate c+d (d = -1, 0, 1) lies in [lo, hi].
__device__ __forceinline__ uint32_t offset_bits(int c, int lo, int hi) {
    return ((c - 1 >= lo && c - 1 <= hi) ? 1u : 0u) | ((c >= lo && c <= hi) ? 2u : 0u) |
           ((c + 1 >= lo && c + 1 <= hi) ? 4u : 0u);
}

// 27-bit masks over the cell ordinals c = 9(dx+1) + 3(dy+1) + (dz+1) from the
// per-axis offset bits: the x bits fill blocks of nine, the y bits groups of
// three in every block, the z bits one position in every group.
__device__ __forceinline__ uint32_t cells_of(uint32_t bx, uint32_t by, uint32_t bz) {
    const uint32_t x = ((bx & 1u) | ((bx & 2u) << 8) | ((bx & 4u) << 16)) * 0x1FFu;
    const uint32_t y = (((by & 1u) | ((by & 2u) << 2) | ((by & 4u) << 4)) * 7u) * 0x40201u;
    const uint32_t z = (bz * 0x49u) * 0x40201u;
    return x & y & z;
}

// The cells of the 27-neighbourhood a lane visits: in the grid, and not an
// interior cell whose box lies farther than r (same tests as the cell-by-cell
// scan, without branches).
__device__ __forceinline__ uint32_t visit_mask(const Seg& sg, const Lane& L, I3 gp, float Tskip) {
    const brt_grid_params& g = sg.g;
    const int gx = (int)g.grid_size[0], gy = (int)g.grid_size[1], gz = (int)g.grid_size[2];
    const uint32_t in_grid = cells_of(offset_bits(gp.x, 0, gx - 1), offset_bits(gp.y, 0, gy - 1),
                                      offset_bits(gp.z, 0, gz - 1));
    const uint32_t inner = cells_of(offset_bits(gp.x, 1, gx - 2), offset_bits(gp.y, 1, gy - 2),
                                    offset_bits(gp.z, 1, gz - 2));
    uint32_t far = 0;
#pragma unroll
    for (int c = 0; c < 27; ++c) {
        const int dx = c / 9 - 1, dy = (c / 3) % 3 - 1, dz = c % 3 - 1;
        const float bd = (gap2(L, 0, dx) + gap2(L, 1, dy)) + gap2(L, 2, dz);
        far |= bd > Tskip ? (1u << c) : 0u;
    }
    return in_grid & ~(far & inner);
}

// Takes the next cell of the mask todo: its ordinal nc (27 when none is left)
// and its cell range, whose loads are left in flight.
__device__ __forceinline__ void flat_fetch(const ProxArgs& a, const Seg& sg, I3 gp, uint32_t& todo, uint32_t& nc,
                                           uint32_t& ns, uint32_t& ne) {
    const bool any = todo != 0u;
    nc = any ? (uint32_t)__builtin_ctz(todo) : 27u;
    todo &= todo - 1u;
    const uint32_t q3 = (nc * 11u) >> 5;  // nc / 3 for nc < 27
    const uint32_t q9 = (nc * 57u) >> 9;  // nc / 9 for nc < 27
    const int dx = (int)q9 - 1, dy = (int)(q3 - 3u * q9) - 1, dz = (int)(nc - 3u * q3) - 1;
    // with no cell left the loads read cell 0 of the grid and are not used
    const uint32_t key = sg.cell_base + (any ? cellKey(gp.x + dx, gp.y + dy, gp.z + dz, sg.g) : 0u);
    ns = a.cell_start[key];
    ne = a.cell_end[key];
}

// own_start: the first sorted index of the agent's own cell.  Rounds of a cell
// start at its first index and advance by G, so the agent's own index i lies
// in the round j_self = own_start + G * floor((i - own_start) / G), in slot
// (i - own_start) mod G, and in no round of another cell.
template <int G, bool PACKED>
__device__ __forceinline__ void scan_flat(const ProxArgs& a, const Seg& sg, Lane& L, I3 gp, uint32_t own_start) {
    static_assert((G & (G - 1)) == 0, "the self round is found with a mask");
    const brt_grid_params& g = sg.g;
    const uint32_t maxBodies = g.max_bodies_per_cell;
    const __amdgpu_buffer_rsrc_t rpos = float4_rsrc(a.spos, a.n);
    const __amdgpu_buffer_rsrc_t rxyz = xyz_rsrc(PACKED ? a.pxyz : nullptr, PACKED ? a.n : 0u);
    const uint32_t ds = (L.i - own_start) & (uint32_t)(G - 1);
    const uint32_t jself = L.i - ds;
    const uint32_t smask = ~(1u << ((uint32_t)G - 1u - ds));
    uint32_t todo = visit_mask(sg, L, gp, a.Tskip);
    // the fetched (next) cell: ordinal nc (27: none left) and its range
    uint32_t nc = 27u, ns = BRT_EMPTY_CELL, ne = 0u;
    flat_fetch(a, sg, gp, todo, nc, ns, ne);
    uint32_t j = 0u, lim = 0u;
    bool live = true;
    while (true) {
        if (live && j >= lim) {
            if (nc == 27u) {
                live = false;
            } else {
                if (ns == BRT_EMPTY_CELL) {
                    j = 0u;
                    lim = 0u;
                } else {
                    lim = (ne - ns) > maxBodies ? ns + maxBodies : ne;
                    lim = lim < a.n ? lim : a.n;
                    j = ns;
                }
                // the old range is consumed before the next one is loaded into
                // the same registers (no copy of an in-flight load)
                __builtin_amdgcn_sched_barrier(0);
                flat_fetch(a, sg, gp, todo, nc, ns, ne);
            }
        }
        if (__ballot(live) == 0) break;
        if (live && j < lim) {
            // candidates past lim are read (zeros past n) and not taken
            float4 q[G];
            load_round<G, PACKED>(rpos, rxyz, j, q);
            log_round_flat<G>(a, L, j, lim - j, j == jself ? smask : ~0u, q);
            j += (uint32_t)G;
            if (L.cnt >= a.K) live = false;
        }
    }
}

template <bool FUSE, bool BATCH, bool LDSL>
__global__ void __launch_bounds__(256) k_prox(ProxArgs a) {
    extern __shared__ uint32_t s_list[];  // blockDim.x * W with LDS lists, else unused
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t wid = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t blk = xcd_block(a.xcd_remap != 0);
    const uint32_t gw = blk * (blockDim.x >> 6) + wid;
    Seg sg;
    const bool wave_live = wave_segment<BATCH>(a, gw, sg);
    Lane L;
    L.i = sg.row0 + lane;
    L.cnt = 0;
    const bool alive = wave_live && L.i < sg.end;
    L.row = LDSL ? s_list + tid * a.Ws : a.lists + (size_t)(alive ? L.i : 0u) * a.W;
    L.wp = L.row + 1u;
    L.h0 = L.h1 = L.h2 = L.b0 = L.b1 = L.b2 = BRT_LIST_PAD;
    float4 P = make_float4(0.f, 0.f, 0.f, 0.f);
    if (alive) P = a.spos[L.i];
    L.px = P.x;
    L.py = P.y;
    L.pz = P.z;
    I3 gp = calcGridPos(P.x, P.y, P.z, sg.g);
    cell_gaps(sg, L, gp);

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
            if (mine) scan_uniform<LDSL>(a, sg, L, ux, uy, uz);
        }
    } else {
        // dense waves (most lanes in cells of >= 8 agents) take rounds of
        // kGroupL candidates and skip the self test outside the agent's own
        // cell; sparse waves, dominated by per-cell overhead, take the short
        // rounds of kGroupR with one code path
        bool dense = false;
        uint32_t own_start = 0;
        if (LDSL) {
            uint32_t occ = 0;
            if (alive) {
                const uint32_t key = a.hash[L.i].x;
                own_start = a.cell_start[key];
                occ = a.cell_end[key] - own_start;
            }
            dense = __popcll(__ballot(alive && occ >= (uint32_t)BRT_PROX_DENSE_OCC)) >= BRT_PROX_DENSE_LANES;
        }
        // with the packed positions a round's loads are 3G/4 instead of G
        const bool packed = a.pxyz != nullptr;
        if (LDSL && BRT_PROX_FLAT && dense && a.log_rounds) {
            if (alive) {
                if (packed)
                    scan_flat<kGroupL, true>(a, sg, L, gp, own_start);
                else
                    scan_flat<kGroupL, false>(a, sg, L, gp, own_start);
            }
        } else if (LDSL && BRT_PROX_FLAT_SPARSE && a.log_rounds) {
            if (alive) {
                if (packed)
                    scan_flat<kGroupR, true>(a, sg, L, gp, own_start);
                else
                    scan_flat<kGroupR, false>(a, sg, L, gp, own_start);
            }
        } else if (alive) {
            if (dense) {
                if (packed)
                    scan_vector<LDSL, kGroupL, BRT_PROX_SELF_SPLIT != 0, true>(a, sg, L, gp);
                else
                    scan_vector<LDSL, kGroupL, BRT_PROX_SELF_SPLIT != 0, false>(a, sg, L, gp);
            } else {
                if (packed)
                    scan_vector<LDSL, kGroupR, false, true>(a, sg, L, gp);
                else
                    scan_vector<LDSL, kGroupR, false, false>(a, sg, L, gp);
            }
        }
    }

    // fused: the agent's own state, loaded before the list is finished so the
    // loads (the smoothed acceleration by agent ID when there is no sorted
    // copy) are in flight meanwhile
    uint32_t own_id = 0u;
    float4 own_fwd = make_float4(0.f, 0.f, 0.f, 0.f), own_sa = make_float4(0.f, 0.f, 0.f, 0.f);
    if (FUSE && alive) {
        own_id = a.hash[L.i].y;
        own_fwd = a.sfwd[L.i];
        if (a.ssa)
            own_sa = a.ssa[L.i];
        else if (own_id < a.n)
            own_sa = a.o.sa_out[own_id];
    }

    if (LDSL && a.log_rounds) expand_log(L, a.K);
    uint32_t pad_from = a.W / 4u;
    if (LDSL || alive) pad_from = finish_list<LDSL>(L, a.W);
    if (!LDSL && wave_live)
        pad_rows(a.lists, sg.row0, (sg.end - sg.row0) < 64u ? (sg.end - sg.row0) : 64u, a.W, lane, pad_from);
    if (LDSL) __syncthreads();
    // LDS lists: each wave writes its own rows, 16 bytes per lane per store
    if (LDSL && wave_live) {
        const uint32_t rows = (sg.end - sg.row0) < 64u ? (sg.end - sg.row0) : 64u;
        const uint32_t W4 = a.W / 4u, Ws4 = a.Ws / 4u;
        const uint32_t words4 = rows * W4;
        uint4* dst = reinterpret_cast<uint4*>(a.lists + (size_t)sg.row0 * a.W);
        const uint4* src = reinterpret_cast<const uint4*>(s_list + (size_t)(wid * 64u) * a.Ws);
        for (uint32_t w = lane; w < words4; w += 64u) {
            const uint32_t r = w / W4;
            dst[w] = src[r * Ws4 + (w - r * W4)];
        }
    }

    // fused steering: walk the finished list (from LDS, or the row this lane
    // just stored), all lanes in step with four neighbour gathers in flight,
    // then integrate.  Same order and operations as k_steer over the written
    // list.
    if (FUSE && alive) {
        const uint32_t id = own_id;
        if (id < a.n) {
            const float4 Fw = own_fwd;
            const float4 SA = own_sa;
            const V3 myPos = v3(P.x, P.y, P.z), myFwd = xyz(Fw);
            SteerAccum acc;
            acc.init();
            const Weights wts = sg.w;
            const float margin = cut_margin(myFwd);
            if (wts.s != 0.0f || wts.a != 0.0f || wts.c != 0.0f) {
                const __amdgpu_buffer_rsrc_t rpos = float4_rsrc(a.spos, a.n);
                const __amdgpu_buffer_rsrc_t rfwd = float4_rsrc(a.sfwd, a.n);
                constexpr int SG = BRT_PROX_STEER_GROUP;
                for (uint32_t t = 0; t < L.cnt; t += SG) {
                    float4 q[SG], f[SG];
#pragma unroll
                    for (int u = 0; u < SG; ++u) {
                        const uint32_t j = (t + u < L.cnt) ? L.row[1 + t + u] : L.i;
                        q[u] = bload(rpos, j * 16u);
                        f[u] = bload(rfwd, j * 16u);
                    }
#pragma unroll
                    for (int u = 0; u < SG; ++u) {
                        if (t + u < L.cnt) {
                            const V3 offset = xyz(q[u]) - myPos;
                            acc.add_cut(a.sp, a.cuts, wts, myFwd, margin, offset, dot(offset, offset), xyz(f[u]));
                        }
                    }
                }
            }
            const uint32_t step = (BATCH && a.b.step) ? *a.b.step : a.step;
            finish_agent(sg.g, a.sp, wts, a.dt, step, L.i, id, id - sg.id_base, myPos, myFwd, Fw.w, xyz(SA), SA.w,
                         acc, a.o);
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

// Workgroup of the single-grid LDS-list neighbour kernels: one wave for large
// launches, so that a wave that finishes early releases its LDS tile at once
// instead of waiting for the slowest wave of a larger workgroup (1M-boid blob
// flock: 64 lanes 4-5 % faster than 256); 256 lanes for small launches (65k
// boids: 3 % faster than 64), halved while the tile exceeds 40 KB.
uint32_t prox_block(uint32_t W, uint32_t n) {
    uint32_t B = n >= (1u << 18) ? BRT_PROX_BLOCK_LARGE : 256u;
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
    a.Ws = lds_list_stride(p.max_neighbors);
    a.n = n;
    a.T = threshold;
    // the skip threshold errs toward scanning: 1e-5 relative on d2 and a
    // per-axis widening of 1e-5 of the grid's coordinate range, both far above
    // the few-ulp rounding of positions, cell bounds and d2
    a.Tskip = threshold * 1.00001f;
    cell_slack(g, a.slack);
    a.uniform_max_cells = uniform_max_cells;
    a.xcd_remap = xcd_remap;
    a.log_rounds = n < (1u << 24) ? BRT_PROX_LOG : 0u;
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
                            uint32_t xcd_remap, uint32_t lds_lists, hipStream_t s, const float* pxyz) {
    const uint32_t K = p.max_neighbors;
    const uint32_t W = list_stride(K);
    if (p.use_knn) {
        uint32_t B = 64;
        while (B > 8 && (size_t)B * (W + K) * 4u > 60u * 1024u) B /= 2;
        const size_t lds = (size_t)B * (W + K) * 4u;
        BRT_LAUNCH(k_neighbors_knn, dim3((n + B - 1) / B), dim3(B), lds, s, g, K, W, threshold, spos,
                           cell_start, cell_end, n, lists);
        return hipGetLastError();
    }
    ProxArgs a = make_args(g, p, threshold, spos, hash, cell_start, cell_end, n, lists, uniform_max_cells, xcd_remap);
    a.pxyz = pxyz;
    const uint32_t B = prox_block(a.Ws, n);
    if (lds_lists)
        BRT_LAUNCH((k_prox<false, false, true>), dim3((n + B - 1) / B), dim3(B), (size_t)B * a.Ws * 4u, s, a);
    else
        BRT_LAUNCH((k_prox<false, false, false>), dim3((n + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_neighbors_steer(const brt_grid_params& g, const brt_prox_params& p, float threshold,
                                  const brt_steer_params& sp, float dt, uint32_t step, const float4* spos,
                                  const float4* sfwd, const float4* ssa, const uint2* hash,
                                  const uint32_t* cell_start, const uint32_t* cell_end, uint32_t n, uint32_t* lists,
                                  float2* wander, float4* pos_out, float4* fwd_out, float4* sa_out,
                                  uint32_t uniform_max_cells, uint32_t xcd_remap, uint32_t lds_lists, hipStream_t s,
                                  const float* pxyz) {
    ProxArgs a = make_args(g, p, threshold, spos, hash, cell_start, cell_end, n, lists, uniform_max_cells, xcd_remap);
    a.pxyz = pxyz;
    a.sfwd = sfwd;
    a.ssa = ssa;
    a.sp = sp;
    a.cuts = steer_cuts(sp);
    a.dt = dt;
    a.step = step;
    a.o = AgentOut{pos_out, fwd_out, sa_out, nullptr, wander};
    const uint32_t B = prox_block(a.Ws, n);
    if (lds_lists)
        BRT_LAUNCH((k_prox<true, false, true>), dim3((n + B - 1) / B), dim3(B), (size_t)B * a.Ws * 4u, s, a);
    else
        BRT_LAUNCH((k_prox<true, false, false>), dim3((n + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_batch_neighbors_steer(const BatchView& b, uint32_t n, const brt_prox_params& p, float threshold,
                                        const brt_steer_params& sp, float dt, const float4* spos, const float4* sfwd,
                                        const float4* ssa, const uint2* hash, const uint32_t* cell_start,
                                        const uint32_t* cell_end, uint32_t* lists, float2* wander, float4* pos_out,
                                        float4* fwd_out, float4* sa_out, uint32_t block, uint32_t uniform_max_cells,
                                        uint32_t xcd_remap, uint32_t lds_lists, hipStream_t s, const float* pxyz) {
    brt_grid_params g0{};
    ProxArgs a = make_args(g0, p, threshold, spos, hash, cell_start, cell_end, n, lists, uniform_max_cells, xcd_remap);
    a.pxyz = pxyz;
    a.sfwd = sfwd;
    a.ssa = ssa;
    a.sp = sp;
    a.cuts = steer_cuts(sp);
    a.dt = dt;
    a.step = 0;
    a.o = AgentOut{pos_out, fwd_out, sa_out, nullptr, wander};
    a.b = b;
    const uint32_t B = block ? block : prox_block(a.Ws, 0u);
    const uint32_t wpb = B / 64u;
    if (lds_lists)
        hipLaunchKernelGGL((k_prox<true, true, true>), dim3((b.waves + wpb - 1) / wpb), dim3(B), (size_t)B * a.Ws * 4u,
                           s, a);
    else
        hipLaunchKernelGGL((k_prox<true, true, false>), dim3((b.waves + wpb - 1) / wpb), dim3(B), 0, s, a);
    return hipGetLastError();
}

}  // namespace brt
