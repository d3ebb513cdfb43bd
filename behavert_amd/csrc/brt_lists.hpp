// brt_lists.hpp -- walking a neighbour list row (count, then up to K sorted
// indices, padded; stride W = 4(K/4+1) words, 16-byte aligned rows).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace brt {

// Visits the entries e = 1..count (count capped at K) of agent i's list in
// order, eight at a time: the row is read as 16-byte words in pairs (as the
// steering pass reads it), and f(base, jj, mask) gets word entries
// base..base+7 (list slot base + k - 1), with bit k of mask set for the valid
// ones.  Invalid lanes of jj hold i, so gathers through them stay in bounds.
// WHOLE_ROW: rows of W = 36 words (K 32..35) are loaded whole, nine 16-byte
// words in flight together, instead of two per octet after the previous
// octet's work (1M-boid blobs, K 32: label propagation 97 -> 51 us per
// iteration, steer-and-slow-to-avoid 166 -> 84 us, cluster steering 427 ->
// 319 us); a kernel holding much else in registers keeps the pairwise reads
// (the one-pass similarities + cluster steering measured 4 % slower whole).
template <bool WHOLE_ROW = true, class Fn>
__device__ __forceinline__ uint32_t for_list_octets(const uint32_t* __restrict__ lists, uint32_t W, uint32_t K,
                                                    uint32_t i, uint32_t n, Fn&& f) {
    const uint4* L = reinterpret_cast<const uint4*>(lists + (size_t)i * W);
    if (WHOLE_ROW && W == 36u) {
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
                uint32_t mask = 0;
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const uint32_t e = 4u * q + k;
                    const bool ok = e >= 1u && e <= count && jj[k] < n;
                    mask |= ok ? (1u << k) : 0u;
                    jj[k] = ok ? jj[k] : i;
                }
                f(4u * q, jj, mask);
            }
        }
        return count;
    }
    const uint4 w0 = L[0];
    const uint32_t count = w0.x < K ? w0.x : K;
    for (uint32_t q = 0; 4u * q <= count; q += 2) {
        const uint4 a = q == 0 ? w0 : L[q];
        const uint4 b = (4u * (q + 1) <= count) ? L[q + 1] : make_uint4(0u, 0u, 0u, 0u);
        uint32_t jj[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
        uint32_t mask = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint32_t e = 4u * q + k;
            const bool ok = e >= 1u && e <= count && jj[k] < n;
            mask |= ok ? (1u << k) : 0u;
            jj[k] = ok ? jj[k] : i;
        }
        // slot of word entry k: e - 1 = 4q + k - 1 (k = 0 of word 0 is the count)
        f(4u * q, jj, mask);
    }
    return count;
}

}  // namespace brt
