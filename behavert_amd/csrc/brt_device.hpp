// brt_device.hpp -- device-side math shared by the gfx950 kernels.
//
// Restates, for the GPU, the reference's helper math in its operation order:
//   calcGridPos / calcGridHash ....... PlugIns/EnvGrid3D/envgrid3d_resources.cu:60-84
//   calcDist4 ........................ Core/common_resources.cu:64-72
//   truncateLength, adjustRawSteeringForce, vecLimitDeviationAngleUtility,
//   blendIntoAccumulator3, interpolate Core/common_resources.cu:140-271
//   inBoidNeighborhood, xxxsteerForSeek PlugIns/OpenSteerWrapper/OpenSteerWrapper_resources.cu:67-92,294-306
// Fixed choices where cutil_math (not in the reference tree) leaves the last bit
// open: normalize(v) = v * (1/sqrtf(dot)), v / s = v * (1/s), powf(x,20) by
// repeated squaring.
// Every TU that includes this is compiled with -ffp-contract=off, so each
// a*b+c below is a rounded multiply followed by a rounded add, and sqrtf and
// '/' are the correctly rounded IEEE operations (HIP's default for FP32).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/behavert_amd.h"

namespace brt {

struct V3 {
    float x, y, z;
};

__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
// cutil_math float3 / float: one correctly rounded reciprocal, three multiplies
__device__ __forceinline__ V3 operator/(V3 a, float s) {
    const float inv = 1.0f / s;
    return a * inv;
}

// cutil dot order: (a.x*b.x + a.y*b.y) + a.z*b.z
__device__ __forceinline__ float dot(V3 a, V3 b) {
    float xx = a.x * b.x;
    float yy = a.y * b.y;
    float zz = a.z * b.z;
    return (xx + yy) + zz;
}
__device__ __forceinline__ float length(V3 v) { return sqrtf(dot(v, v)); }
// normalize: v * (1 / sqrt(dot)), correctly rounded reciprocal of the correctly
// rounded square root (the reference's cutil normalize uses rsqrtf).
__device__ __forceinline__ V3 normalize(V3 v) {
    float inv = 1.0f / sqrtf(dot(v, v));
    return v * inv;
}
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float clampf(float x, float lo, float hi) {
    float t = x < lo ? lo : x;
    return t > hi ? hi : t;
}
__device__ __forceinline__ V3 xyz(float4 a) { return v3(a.x, a.y, a.z); }

__device__ __forceinline__ V3 truncateLength(V3 v, float maxLength) {
    float len = length(v);
    if (len <= maxLength) return v;
    return v * (maxLength / len);
}

__device__ __forceinline__ float pow20(float x) {
    float x2 = x * x;
    float x4 = x2 * x2;
    float x5 = x4 * x;
    float x10 = x5 * x5;
    return x10 * x10;
}

__device__ __forceinline__ V3 vecLimitDeviationAngle(V3 source, float cosCone, V3 basis) {
    float sourceLength = length(source);
    if (sourceLength == 0.0f) return source;
    V3 direction = source / sourceLength;
    float cosSource = dot(direction, basis);
    if (cosSource >= cosCone) return source;
    float projection = dot(source, basis);
    V3 perp = source - basis * projection;
    V3 unitPerp = normalize(perp);
    float perpDist = sqrtf(1.0f - cosCone * cosCone);
    V3 c0 = basis * cosCone;
    V3 c1 = unitPerp * perpDist;
    return (c0 + c1) * sourceLength;
}

__device__ __forceinline__ V3 adjustRawSteeringForce(V3 force, float speed, V3 forward, float maxSpeed) {
    float maxAdjustedSpeed = 0.2f * maxSpeed;
    if (speed > maxAdjustedSpeed || (force.x == 0.0f && force.y == 0.0f && force.z == 0.0f)) return force;
    float range = speed / maxAdjustedSpeed;
    float cosine = 1.0f + (-2.0f * pow20(range));
    return vecLimitDeviationAngle(force, cosine, forward);
}

__device__ __forceinline__ V3 steerForSeek(V3 myPos, float mySpeed, V3 myForward, V3 target, float maxSpeed) {
    V3 offset = target - myPos;
    V3 desired = truncateLength(offset, maxSpeed);
    V3 myVel = myForward * mySpeed;
    return desired - myVel;
}

// ------------------------------------------------------------------ grid math

// float -> int as CUDA's cvt.rzi.s32.f32 (NaN -> 0, saturating): the
// conversion the reference's "int = floor(float)" assignment performs.
__device__ __forceinline__ int f2i(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.0f) return 0x7fffffff;
    if (x <= -2147483648.0f) return (int)0x80000000;
    return (int)x;
}

__device__ __forceinline__ int mul24(int a, int b) {
    // low 32 bits of the product of the sign-extended low 24 bits
    return __mul24(a, b);
}

struct I3 {
    int x, y, z;
};

__device__ __forceinline__ I3 calcGridPos(float px, float py, float pz, const brt_grid_params& g) {
    I3 r;
    r.x = f2i(floorf((px - g.world_origin[0]) / g.cell_size[0]));
    r.y = f2i(floorf((py - g.world_origin[1]) / g.cell_size[1]));
    r.z = f2i(floorf((pz - g.world_origin[2]) / g.cell_size[2]));
    return r;
}

// max(0, min(int, uint)) in the reference compares as unsigned (CUDA's
// mixed-sign min/max overloads): negative -> size-1, above -> size-1.
__device__ __forceinline__ uint32_t clampAxis(int v, uint32_t size) {
    uint32_t u = (uint32_t)v;
    uint32_t hi = size - 1u;
    return u < hi ? u : hi;
}

__device__ __forceinline__ uint32_t cellKey(int x, int y, int z, const brt_grid_params& g) {
    int gx = (int)g.grid_size[0], gy = (int)g.grid_size[1];
    return (uint32_t)(mul24(mul24(z, gy), gx) + mul24(y, gx) + x);
}

__device__ __forceinline__ uint32_t calcGridHash(I3 gp, const brt_grid_params& g) {
    return cellKey((int)clampAxis(gp.x, g.grid_size[0]), (int)clampAxis(gp.y, g.grid_size[1]),
                   (int)clampAxis(gp.z, g.grid_size[2]), g);
}

__device__ __forceinline__ bool inGrid(int x, int y, int z, const brt_grid_params& g) {
    return x >= 0 && x <= (int)g.grid_size[0] - 1 && y >= 0 && y <= (int)g.grid_size[1] - 1 && z >= 0 &&
           z <= (int)g.grid_size[2] - 1;
}

// ------------------------------------------------------------------ launch mapping

// Workgroups are dealt to the 8 XCDs of an MI355X round-robin by workgroup ID,
// and each XCD has its own L2.  Remapping the hardware ID so that XCD k runs a
// contiguous k-th slice of the logical workgroups keeps spatially adjacent
// agents (and the neighbour data they share) in one L2.  A bijection on
// [0, gridDim.x) for any grid size.
__device__ __forceinline__ uint32_t xcd_block(bool remap) {
    const uint32_t b = blockIdx.x, G = gridDim.x;
    if (!remap || G < 16) return b;
    const uint32_t q = G / 8u, r = G % 8u;
    const uint32_t xcd = b % 8u, slot = b / 8u;
    return xcd < r ? xcd * (q + 1u) + slot : r * (q + 1u) + (xcd - r) * q + slot;
}

// ------------------------------------------------------------------ Philox4x32-10
// Salmon, Moraes, Dror, Shaw, SC'11 ("Parallel random numbers: as easy as 1, 2, 3").
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint32_t lo0 = 0xD2511F53u * c[0];
        uint32_t hi0 = __umulhi(0xD2511F53u, c[0]);
        uint32_t lo1 = 0xCD9E8D57u * c[2];
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]);
        uint32_t n0 = hi1 ^ c[1] ^ k0;
        uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
    }
}

__device__ __forceinline__ float u24(uint32_t r) { return (float)(r >> 8) * (1.0f / 16777216.0f); }

}  // namespace brt
