// brt_lists.hpp -- walking a neighbour list row (count, then up to K sorted
// indices, padded; stride W = 4(K/4+1) words, 16-byte aligned rows).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace brt {

// Bit k set for the valid entries e = 4q + k of an octet (1 <= e <= count and
// a sorted index below n); invalid lanes of jj are set to i.  Word entry 0 of
// word 0 is the count; the slot of word entry k is e - 1 = 4q + k - 1.
__device__ __forceinline__ uint32_t octet_mask(uint32_t (&jj)[8], uint32_t q, uint32_t count, uint32_t i,
                                               uint32_t n) {
    uint32_t mask = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t e = 4u * q + k;
        const bool ok = e >= 1u && e <= count && jj[k] < n;
        mask |= ok ? (1u << k) : 0u;
        jj[k] = ok ? jj[k] : i;
    }
    return mask;
}

// How a list walk reads its row.
enum ListRows : int {
    kRowPairs = 0,     // two 16-byte words per octet, after the previous octet's work
    kRowPrefetch = 1,  // two words per octet, the next octet's pair loading with this one
    kRowWhole = 2,     // rows of W = 36 words (K 32..35) whole, else kRowPrefetch
};

// Visits the entries e = 1..count (count capped at K) of agent i's list in
// order, eight at a time: f(base, jj, mask) gets word entries base..base+7
// (list slot base + k - 1), with bit k of mask set for the valid ones.
// Invalid lanes of jj hold i, so gathers through them stay in bounds.
// kRowWhole loads a K 32 row as nine 16-byte words in flight together (1M-boid
// blobs, K 32: label propagation 97 -> 51 us per iteration,
// steer-and-slow-to-avoid 166 -> 84 us, cluster steering 427 -> 319 us).  A
// kernel holding much else in registers reads pairs: the one-pass
// similarities + cluster steering measured 4 % slower with whole rows and 4 %
// faster with prefetched pairs (384 -> 367 us); the steering pass with the
// feature update (K 64) 3 % slower with prefetched pairs than without.
template <int ROWS = kRowWhole, class Fn>
__device__ __forceinline__ uint32_t for_list_octets(const uint32_t* __restrict__ lists, uint32_t W, uint32_t K,
                                                    uint32_t i, uint32_t n, Fn&& f) {
    const uint4* L = reinterpret_cast<const uint4*>(lists + (size_t)i * W);
    if (ROWS == kRowWhole && W == 36u) {
        uint4 c[10];
#pragma unroll
        for (int k = 0; k < 9; ++k) c[k] = L[k];
        c[9] = make_uint4(0u, 0u, 0u, 0u);
        const uint32_t count = c[0].x < K ? c[0].x : K;
#pragma unroll
        for (uint32_t q = 0; q < 10u; q += 2) {
            if (4u * q <= count) {
                const uint4 a = c[q];
                const uint4 b = (4u * (q + 1) <= count) ? c[q + 1] : make_uint4(0u, 0u, 0u, 0u);
                uint32_t jj[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
                const uint32_t mask = octet_mask(jj, q, count, i, n);
                f(4u * q, jj, mask);
            }
        }
        return count;
    }
    const uint4 w0 = L[0];
    const uint32_t count = w0.x < K ? w0.x : K;
    const uint4 z = make_uint4(0u, 0u, 0u, 0u);
    if (ROWS == kRowPairs) {
        for (uint32_t q = 0; 4u * q <= count; q += 2) {
            const uint4 a = q == 0 ? w0 : L[q];
            const uint4 b = (4u * (q + 1) <= count) ? L[q + 1] : z;
            uint32_t jj[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
            const uint32_t mask = octet_mask(jj, q, count, i, n);
            f(4u * q, jj, mask);
        }
        return count;
    }
    uint4 a = w0, b = 4u <= count ? L[1] : z;
    for (uint32_t q = 0; 4u * q <= count; q += 2) {
        // the next pair of words loads while this octet's gathers are in flight
        const uint4 na = 4u * (q + 2) <= count ? L[q + 2] : z;
        const uint4 nb = 4u * (q + 3) <= count ? L[q + 3] : z;
        uint32_t jj[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
        const uint32_t mask = octet_mask(jj, q, count, i, n);
        f(4u * q, jj, mask);
        a = na;
        b = nb;
    }
    return count;
}

}  // namespace brt
