// brt_consumers.hip -- small per-agent consumers of the neighbour lists and of
// the steer force outside the flocking step, on gfx950.
//
//   steerToAvoidNeighbors ....... PlugIns/OpenSteerWrapper/OpenSteerWrapper_kernel.cu:110-154
//                                 (host: OpenSteerWrapperPlugIn.h:309-325)
//   extractColorFromNeighborhoodD PlugIns/Drawable3D/drawable3d_kernel.cu:60-93
//                                 (host: Drawable3DPlugIn.h:712-724)
//   moveTowardsTargetD .......... PlugIns/Shapes3D/Shapes3D_kernel.cu:63-87
//                                 (host: Shapes3DPlugIn.h:552-580)
// The reference writes a few constants as double literals (0.9, 1.02, -0.8,
// 0.5, 0.3): those expressions are evaluated in double and rounded to float,
// as nvcc compiles them for sm_20 and later.
#include "brt_device.hpp"
#include "brt_internal.hpp"
#include "brt_lists.hpp"
#include "brt_steer_common.hpp"

namespace brt {

namespace {

// Separation (calcSeparation, OpenSteerWrapper_resources.cu:166-211) scaled by
// its weight, amplified when the list holds more than 16 agents; the agent's
// speed drops by 10 % when that steering opposes its heading (cos < -0.8) and
// rises by 2 % when it agrees (> 0.5).  forward_out (sorted) keeps the heading;
// force_out = steering + force_in (blendIntoSteeringForce).
__global__ void __launch_bounds__(256) k_slow_avoid(uint32_t K, uint32_t W, brt_steer_params sp, uint32_t n,
                                                    const float4* __restrict__ spos, const float4* sfwd,
                                                    const uint32_t* __restrict__ lists, float4* fwd_out,
                                                    const float4* force_in, float4* force_out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 P = spos[i], Fw = sfwd[i];
    const V3 myPos = xyz(P), myFwd = xyz(Fw);
    const Weights sepOnly{1.0f, 0.0f, 0.0f};
    SteerAccum acc;
    acc.init();
    const uint32_t count = for_list_octets(lists, W, K, i, n, [&](uint32_t, const uint32_t(&jj)[8], uint32_t mask) {
        float4 qp[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) qp[k] = spos[jj[k]];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (!(mask & (1u << k))) continue;
            const V3 offset = xyz(qp[k]) - myPos;
            acc.add(sp, sepOnly, myFwd, offset, dot(offset, offset), myFwd);
        }
    });
    V3 steering = (acc.nSep > 0 ? normalize(acc.sSep / (float)acc.nSep) : acc.sSep) * sp.separation[0];
    if (count > 16u) steering = steering * ((float)(count - 16u) * 0.5f);
    float mySpeed = Fw.w;
    const float c = dot(steering, myFwd);
    if ((double)c < -0.8)
        mySpeed = (float)((double)mySpeed * 0.9);
    else if ((double)c > 0.5)
        mySpeed = (float)((double)mySpeed * 1.02);
    fwd_out[i] = make_float4(Fw.x, Fw.y, Fw.z, mySpeed);
    const float4 f0 = force_in[i];
    force_out[i] = make_float4(steering.x + f0.x, steering.y + f0.y, steering.z + f0.z, 0.0f);
}

// colour by ID from the agent's neighbour count f = count / K:
// (base.x + f, base.y + 0.3 + 1 - f, base.z + f (1 - f), 1); with three
// vertices per agent the first is darkened by 0.5.
__global__ void __launch_bounds__(256) k_neighborhood_color(uint32_t K, uint32_t W, float bx, float by, float bz,
                                                            uint32_t vertexes, uint32_t n,
                                                            const uint2* __restrict__ hash,
                                                            const uint32_t* __restrict__ lists,
                                                            float4* __restrict__ color) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t id = hash[i].y;
    if (id >= n) return;
    const float f = (float)lists[(size_t)i * W] / (float)K;
    const float4 c = make_float4(bx + f, (float)((double)by + 0.3 + 1.0 - (double)f), bz + f * (1.0f - f), 1.0f);
    if (vertexes == 3u) {
        color[3u * (size_t)id] = make_float4(c.x - 0.5f, c.y - 0.5f, c.z - 0.5f, c.w - 0.0f);
        color[3u * (size_t)id + 1u] = c;
        color[3u * (size_t)id + 2u] = c;
    } else {
        color[id] = c;
    }
}

// Attraction toward the agent's shape target: target t = targets[id + shift]
// (wrapped by n when the index exceeds n; index n itself, and anything outside
// [0, n), reads as (0, 0, 0, 0), as the reference's texture fetch past the end
// does); force_out = normalize((t.xyz + base) - pos) * 12 * t.w + force_in.
__global__ void __launch_bounds__(256) k_move_towards_target(int shift, float bx, float by, float bz, uint32_t n,
                                                             const uint2* __restrict__ hash,
                                                             const float4* __restrict__ spos,
                                                             const float4* __restrict__ targets,
                                                             const float4* force_in, float4* force_out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t id = hash[i].y;
    int64_t sh = (int64_t)id + shift;
    if (sh > (int64_t)n) sh -= (int64_t)n;
    const float4 t = (sh >= 0 && sh < (int64_t)n) ? targets[sh] : make_float4(0.f, 0.f, 0.f, 0.f);
    const V3 steering = (xyz(t) + v3(bx, by, bz)) - xyz(spos[i]);
    const V3 f = (normalize(steering) * 12.0f) * t.w;
    const float4 f0 = force_in[i];
    force_out[i] = make_float4(f.x + f0.x, f.y + f0.y, f.z + f0.z, 0.0f);
}

}  // namespace

hipError_t launch_slow_avoid(uint32_t K, const brt_steer_params& sp, uint32_t n, const float4* spos,
                             const float4* sfwd, const uint32_t* lists, float4* fwd_out, const float4* force_in,
                             float4* force_out, hipStream_t s) {
    hipLaunchKernelGGL(k_slow_avoid, dim3((n + 255) / 256), dim3(256), 0, s, K, list_stride(K), sp, n, spos, sfwd,
                       lists, fwd_out, force_in, force_out);
    return hipGetLastError();
}

hipError_t launch_neighborhood_color(uint32_t K, const float base[3], uint32_t vertexes, uint32_t n,
                                     const uint2* hash, const uint32_t* lists, float4* color, hipStream_t s) {
    hipLaunchKernelGGL(k_neighborhood_color, dim3((n + 255) / 256), dim3(256), 0, s, K, list_stride(K), base[0],
                       base[1], base[2], vertexes, n, hash, lists, color);
    return hipGetLastError();
}

hipError_t launch_move_towards_target(int shift, const float base[3], uint32_t n, const uint2* hash,
                                      const float4* spos, const float4* targets, const float4* force_in,
                                      float4* force_out, hipStream_t s) {
    hipLaunchKernelGGL(k_move_towards_target, dim3((n + 255) / 256), dim3(256), 0, s, shift, base[0], base[1],
                       base[2], n, hash, spos, targets, force_in, force_out);
    return hipGetLastError();
}

}  // namespace brt
