// brt_oracle.cpp -- CPU oracle for the BehaveRT per-step simulation path.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load this library, and only as the checker
// (or as the timed CPU baseline).  The product path never calls it.
//
// It restates, on the CPU, the algorithm of the reference's per-step path:
//   calcGridPos / calcGridHash ....... PlugIns/EnvGrid3D/envgrid3d_resources.cu:60-84
//   genCalcHashD ..................... PlugIns/EnvGrid3D/envgrid3d_kernel.cu:55-68
//   RadixSort (stable, key/value) .... PlugIns/EnvGrid3D/radixsort.cu:50-80
//   reorderDataFloat4 ................ PlugIns/EnvGrid3D/envgrid3d_kernel.cu:110-122
//   genericFindCellStartD ............ PlugIns/Proximity3D/Proximity3D_kernel.cu:54-84
//   genericCalcNeighgborhoodD ........ PlugIns/Proximity3D/Proximity3D_kernel.cu:89-125
//   checkForNeighborsInCell/Neighbor . PlugIns/Proximity3D/Proximity3D_resources.cu:59-172
//   getNeighborsList (W layout) ...... PlugIns/Proximity3D/Proximity3D_resources.cu:175-198
//   inBoidNeighborhood, calcSeparation/Cohesion/Alignment, xxxsteerForSeek,
//   applySteeringForceSingle ......... PlugIns/OpenSteerWrapper/OpenSteerWrapper_resources.cu:67-379
//   steering kernels + integrator .... PlugIns/OpenSteerWrapper/OpenSteerWrapper_kernel.cu:62-333
//   truncateLength, adjustRawSteeringForce, vecLimitDeviationAngleUtility,
//   blendIntoAccumulator3, interpolate Core/common_resources.cu:140-271
//   step order ....................... Samples/OpensteerDemo/plugins/BehaveRT_Flocking.cpp:177-196
//   steerForWander / scalarRandomWalk  Samples/OpensteerDemo/include/OpenSteer/SteerLibrary.h:319-331,
//                                      Samples/OpensteerDemo/include/OpenSteer/Utilities.h:158-167
//
// Two precisions of the floating-point part share one template:
//   * FP32 mirror (R = float): the reference's operation order, IEEE single
//     precision, compiled with -ffp-contract=off (no FMA), correctly rounded
//     sqrt and division.  It also defines the integer outputs (keys, stable
//     permutation, cell ranges, neighbour lists) bit for bit.
//   * FP64 (R = double): the same function evaluated in double precision from
//     the same FP32 inputs and constants; the reference point for the float
//     tolerances.
//
// Build: oracle/build.sh (g++ -O2 -ffp-contract=off -fopenmp).
//
// Choices this oracle fixes where the reference leaves the last bit open
// (cutil_math is not in the reference tree):
//   normalize(v)  = v * (1 / sqrt(dot(v,v)))   (reference: v * rsqrtf(dot))
//   v / s         = v * (1 / s)                 (cutil_math's float3 / float)
//   powf(x, 20)   = ((x^2)^2 * x)^2)^2          (repeated squaring)
// and one reformulation for accuracy: cohesion is normalize(mean(p_j - p_i))
// instead of the reference's normalize(sum(p_j)/n - p_i).  The two are equal in
// real arithmetic; the reference's form cancels catastrophically in FP32 once
// |p| >> r (a centroid of ~32 positions near 200 minus a position near 200),
// while the offsets p_j - p_i are exact (Sterbenz) and small.
// Each is compared against the FP64 evaluation by the tests, not bit-pinned to
// the reference GPU (which cannot run here).

#include "../include/behavert_amd.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ------------------------------------------------------------------ vectors

template <typename R>
struct V3 {
    R x, y, z;
};

template <typename R>
inline V3<R> v3(R x, R y, R z) { return V3<R>{x, y, z}; }
template <typename R>
inline V3<R> operator+(V3<R> a, V3<R> b) { return v3<R>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <typename R>
inline V3<R> operator-(V3<R> a, V3<R> b) { return v3<R>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <typename R>
inline V3<R> operator*(V3<R> a, R s) { return v3<R>(a.x * s, a.y * s, a.z * s); }
// cutil_math float3 / float: one reciprocal, three multiplies.
template <typename R>
inline V3<R> operator/(V3<R> a, R s) {
    R inv = R(1) / s;
    return a * inv;
}

inline float rsqrt_sqrt(float x) { return sqrtf(x); }
inline double rsqrt_sqrt(double x) { return std::sqrt(x); }

// cutil_math dot: a.x*b.x + a.y*b.y + a.z*b.z, evaluated left to right.
template <typename R>
inline R dot(V3<R> a, V3<R> b) {
    R xx = a.x * b.x;
    R yy = a.y * b.y;
    R zz = a.z * b.z;
    return (xx + yy) + zz;
}
template <typename R>
inline R length(V3<R> v) { return rsqrt_sqrt(dot(v, v)); }
template <typename R>
inline V3<R> normalize(V3<R> v) {
    R inv = R(1) / rsqrt_sqrt(dot(v, v));
    return v * inv;
}
template <typename R>
inline V3<R> cross(V3<R> a, V3<R> b) {
    return v3<R>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <typename R>
inline R clampr(R x, R lo, R hi) {  // cutil clamp = fminf(fmaxf(x, lo), hi)
    R t = x < lo ? lo : x;
    return t > hi ? hi : t;
}

// Core/common_resources.cu:260-271
template <typename R>
inline V3<R> truncateLength(V3<R> v, R maxLength) {
    R len = length(v);
    if (len <= maxLength) return v;
    return v * (maxLength / len);
}

template <typename R>
inline R pow20(R x) {
    R x2 = x * x;
    R x4 = x2 * x2;
    R x5 = x4 * x;
    R x10 = x5 * x5;
    return x10 * x10;
}

// Core/common_resources.cu:186-228 (insideOrOutside == true)
template <typename R>
inline V3<R> vecLimitDeviationAngle(V3<R> source, R cosineOfConeAngle, V3<R> basis) {
    R sourceLength = length(source);
    if (sourceLength == R(0)) return source;
    V3<R> direction = source / sourceLength;
    R cosineOfSourceAngle = dot(direction, basis);
    if (cosineOfSourceAngle >= cosineOfConeAngle) return source;
    R projection = dot(source, basis);
    V3<R> perp = source - basis * projection;
    V3<R> unitPerp = normalize(perp);
    R perpDist = rsqrt_sqrt(R(1) - cosineOfConeAngle * cosineOfConeAngle);
    V3<R> c0 = basis * cosineOfConeAngle;
    V3<R> c1 = unitPerp * perpDist;
    return (c0 + c1) * sourceLength;
}

// Core/common_resources.cu:234-254
template <typename R>
inline V3<R> adjustRawSteeringForce(V3<R> force, R speed, V3<R> forward, R maxSpeed) {
    R maxAdjustedSpeed = R(0.2f) * maxSpeed;
    if (speed > maxAdjustedSpeed || (force.x == R(0) && force.y == R(0) && force.z == R(0)))
        return force;
    R range = speed / maxAdjustedSpeed;
    R cosine = R(1) + (R(-2) * pow20(range));  // interpolate(p, 1, -1)
    return vecLimitDeviationAngle(force, cosine, forward);
}

// OpenSteerWrapper_resources.cu:67-92.  A discrete decision: both precisions
// evaluate it in FP32 on the FP32-rounded state, so the FP64 path differs from
// the FP32 one only in its continuous arithmetic (a neighbour at d == r exactly
// is in or out for both).
template <typename R>
inline V3<float> f32v(V3<R> v) { return v3<float>((float)v.x, (float)v.y, (float)v.z); }

template <typename R>
inline bool inBoidNeighborhood(V3<R> myPosR, V3<R> otherPosR, float minDist, float maxDist,
                               float cosMaxAngle, V3<R> forwardR) {
    V3<float> myPos = f32v(myPosR), otherPos = f32v(otherPosR), forward = f32v(forwardR);
    V3<float> offset = otherPos - myPos;
    float dist = length(offset);
    if (dist < minDist) return true;
    if (dist > maxDist) return false;
    float cosAngle = dot(forward, normalize(offset));
    return cosAngle > cosMaxAngle;
}

// OpenSteerWrapper_resources.cu:294-306
template <typename R>
inline V3<R> steerForSeek(V3<R> myPos, R mySpeed, V3<R> myForward, V3<R> target, R maxSpeed) {
    V3<R> offset = target - myPos;
    V3<R> desired = truncateLength(offset, maxSpeed);
    V3<R> myVel = myForward * mySpeed;
    return desired - myVel;
}

// ------------------------------------------------------------------ grid math

// float -> int conversion as CUDA's cvt.rzi.s32.f32 does it (NaN -> 0, saturating).
inline int f2i(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.0f) return INT_MAX;
    if (x <= -2147483648.0f) return INT_MIN;
    return (int)x;
}

// __mul24: low 32 bits of the product of the sign-extended low 24 bits.
inline int mul24(int a, int b) {
    int64_t sa = (int64_t)(int32_t)((uint32_t)a << 8) >> 8;
    int64_t sb = (int64_t)(int32_t)((uint32_t)b << 8) >> 8;
    return (int)(uint32_t)(uint64_t)(sa * sb);
}

struct I3 {
    int x, y, z;
};

// envgrid3d_resources.cu:60-67
inline I3 calcGridPos(const float* p, const brt_grid_params& g) {
    I3 r;
    r.x = f2i(floorf((p[0] - g.world_origin[0]) / g.cell_size[0]));
    r.y = f2i(floorf((p[1] - g.world_origin[1]) / g.cell_size[1]));
    r.z = f2i(floorf((p[2] - g.world_origin[2]) / g.cell_size[2]));
    return r;
}

// envgrid3d_resources.cu:75-84.  max(0, min(int, uint)) resolves to CUDA's
// mixed-sign overloads, which compare as unsigned: a negative coordinate maps
// to gridSize-1, a coordinate above the grid to gridSize-1.
inline uint32_t clampAxis(int v, uint32_t size) {
    uint32_t u = (uint32_t)v;
    uint32_t hi = size - 1u;
    return u < hi ? u : hi;
}

inline uint32_t calcGridHash(I3 gp, const brt_grid_params& g) {
    int x = (int)clampAxis(gp.x, g.grid_size[0]);
    int y = (int)clampAxis(gp.y, g.grid_size[1]);
    int z = (int)clampAxis(gp.z, g.grid_size[2]);
    int gx = (int)g.grid_size[0], gy = (int)g.grid_size[1];
    return (uint32_t)(mul24(mul24(z, gy), gx) + mul24(y, gx) + x);
}

inline bool inGrid(I3 c, const brt_grid_params& g) {
    return c.x >= 0 && c.x <= (int)g.grid_size[0] - 1 && c.y >= 0 &&
           c.y <= (int)g.grid_size[1] - 1 && c.z >= 0 && c.z <= (int)g.grid_size[2] - 1;
}

inline uint32_t listStride(uint32_t K) { return 4u * (K / 4u + 1u); }

// ------------------------------------------------------------------ Philox4x32-10
// Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11.
inline void philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

inline float u24(uint32_t r) { return (float)(r >> 8) * (1.0f / 16777216.0f); }

int g_threads = 0;

inline void set_threads() {
#ifdef _OPENMP
    if (g_threads > 0) omp_set_num_threads(g_threads);
#endif
}

// ------------------------------------------------------------------ steering (one agent)

template <typename R>
struct SteerArrays {
    const R* pos;    // sorted, 4 per agent
    const R* fwd;    // sorted
    const R* sa;     // sorted
    const uint32_t* hash;
    const uint32_t* lists;
    uint32_t W;
    R* wander;       // by ID, 2 per agent (in/out), or null
    R* pos_out;      // by ID
    R* fwd_out;
    R* sa_out;
    R* force_out;    // sorted, optional
    R* cond_out;     // sorted, optional: max over behaviours of sum|terms| / |sum|
    const float* pred_pos;  // sorted, optional: FP32 state the discrete decisions
    const float* pred_fwd;  // (radius/angle predicates, world box) are taken from
    // decomposed steering (the reference's separate kernels, each summing onto
    // the steer force): forces_only stops after the behaviours and writes
    // force_out; force_in (sorted) replaces the behaviours by a given force
    // before the world-centre seek and the integration
    bool forces_only;
    const R* force_in;
};

template <typename R>
inline V3<R> ld3(const R* a, uint32_t i) { return v3<R>(a[4 * i], a[4 * i + 1], a[4 * i + 2]); }

template <typename R>
void steer_agent(uint32_t i, const brt_grid_params& g, const brt_steer_params& sp, R dt,
                 uint32_t step, const SteerArrays<R>& A) {
    // state the discrete decisions are evaluated on: the agent's own (rounded
    // to FP32), or a given FP32 state (finite-difference rollouts hold the
    // forward's decisions fixed)
    auto ppos = [&](uint32_t k) -> V3<R> {
        return A.pred_pos ? v3<R>(R(A.pred_pos[4 * k]), R(A.pred_pos[4 * k + 1]), R(A.pred_pos[4 * k + 2]))
                          : ld3(A.pos, k);
    };
    auto pfwd = [&](uint32_t k) -> V3<R> {
        return A.pred_fwd ? v3<R>(R(A.pred_fwd[4 * k]), R(A.pred_fwd[4 * k + 1]), R(A.pred_fwd[4 * k + 2]))
                          : ld3(A.fwd, k);
    };
    const V3<R> myPos = ld3(A.pos, i);
    const V3<R> myFwd = ld3(A.fwd, i);
    R mySpeed = A.fwd[4 * i + 3];
    V3<R> mySa = A.sa ? ld3(A.sa, i) : v3<R>(0, 0, 0);
    const R saW = A.sa ? A.sa[4 * i + 3] : R(0);
    const uint32_t id = A.hash[2 * i + 1];
    const uint32_t* L = A.lists ? A.lists + (size_t)i * A.W : nullptr;
    const uint32_t count = L ? L[0] : 0u;
    const float minR = sp.sensing_min_radius;

    V3<R> force = A.force_in ? ld3(A.force_in, i) : v3<R>(0, 0, 0);
    R cond = R(0);
    auto upd_cond = [&](R mag_sum, V3<R> sum) {
        R l = length(sum);
        R k = l > R(0) ? mag_sum / l : R(INFINITY);
        if (mag_sum > R(0) && k > cond) cond = k;
    };

    if (!A.force_in) {
        // genericSeparationD (OpenSteerWrapper_kernel.cu:86-105), calcSeparation (:166-211)
        if (sp.separation[0] != 0.0f) {
            V3<R> s = v3<R>(0, 0, 0);
            R mag = R(0);
            uint32_t nb = 0;
            for (uint32_t t = 0; t < count; ++t) {
                uint32_t j = L[1 + t];
                V3<R> other = ld3(A.pos, j);
                if (inBoidNeighborhood(ppos(i), ppos(j), minR, sp.separation[1], sp.separation[2], pfwd(i))) {
                    V3<R> offset = other - myPos;
                    R d2 = dot(offset, offset);
                    s = s + offset / (-d2);
                    mag = mag + R(1) / length(offset);
                    ++nb;
                }
            }
            if (nb > 0) {
                if (A.cond_out) upd_cond(mag, s);
                s = normalize(s / R(nb));
            }
            force = s * R(sp.separation[0]) + force;
        }
        // genericAlignmentD (:158-178), calcAlignment (:253-290)
        if (sp.alignment[0] != 0.0f) {
            V3<R> s = v3<R>(0, 0, 0);
            R mag = R(0);
            uint32_t nb = 0;
            for (uint32_t t = 0; t < count; ++t) {
                uint32_t j = L[1 + t];
                if (inBoidNeighborhood(ppos(i), ppos(j), minR, sp.alignment[1], sp.alignment[2], pfwd(i))) {
                    s = s + ld3(A.fwd, j);
                    mag = mag + length(ld3(A.fwd, j));
                    ++nb;
                }
            }
            if (nb > 0) {
                if (A.cond_out) upd_cond(mag + R(nb) * length(myFwd), s - myFwd * R(nb));
                s = normalize(s / R(nb) - myFwd);
            }
            force = s * R(sp.alignment[0]) + force;
        }
        // genericCohesionD (:62-81), calcCohesion (:216-249)
        if (sp.cohesion[0] != 0.0f) {
            V3<R> s = v3<R>(0, 0, 0);
            R mag = R(0);
            uint32_t nb = 0;
            for (uint32_t t = 0; t < count; ++t) {
                uint32_t j = L[1 + t];
                V3<R> other = ld3(A.pos, j);
                if (inBoidNeighborhood(ppos(i), ppos(j), minR, sp.cohesion[1], sp.cohesion[2], pfwd(i))) {
                    s = s + (other - myPos);
                    mag = mag + length(other - myPos);
                    ++nb;
                }
            }
            if (nb > 0) {
                if (A.cond_out) upd_cond(mag, s);
                s = normalize(s / R(nb));
            }
            force = s * R(sp.cohesion[0]) + force;
        }
        // multi-target seek (build-defined; xxxsteerForSeek toward targets[id % T])
        if (sp.num_targets > 0 && sp.target_weight != 0.0f) {
            const float* t = sp.targets[id % sp.num_targets];
            V3<R> tgt = v3<R>(R(t[0]), R(t[1]), R(t[2]));
            V3<R> s = steerForSeek(myPos, mySpeed, myFwd, tgt, R(sp.max_speed));
            force = s * R(sp.target_weight) + force;
        }
        // wander (build-defined; OpenSteer steerForWander with a counter-based RNG)
        if (sp.wander_weight != 0.0f && A.wander) {
            uint32_t ctr[4] = {step, 0u, 0u, 0u};
            uint32_t key[2] = {sp.wander_seed, id};
            uint32_t r[4];
            philox4x32_10(ctr, key, r);
            R u0 = R(u24(r[0])), u1 = R(u24(r[1]));
            R walk = R(12) * dt;
            R ws = A.wander[2 * id], wu = A.wander[2 * id + 1];
            ws = ws + ((u0 * R(2) - R(1)) * walk);
            ws = ws < R(-1) ? R(-1) : (ws > R(1) ? R(1) : ws);
            wu = wu + ((u1 * R(2) - R(1)) * walk);
            wu = wu < R(-1) ? R(-1) : (wu > R(1) ? R(1) : wu);
            A.wander[2 * id] = ws;
            A.wander[2 * id + 1] = wu;
            // OpenSteer's right-handed basis (LocalSpace.h:277-305,
            // setUnitSideFromForwardAndUp / regenerateOrthonormalBasisUF) with
            // the world up as the previous up: side = normalize(fwd x up),
            // up = side x fwd; (1,0,0) when the heading is along the world up.
            V3<R> c = cross(myFwd, v3<R>(0, 1, 0));
            V3<R> side = dot(c, c) == R(0) ? v3<R>(1, 0, 0) : normalize(c);
            V3<R> up = cross(side, myFwd);
            V3<R> s = side * ws + up * wu;
            force = s * R(sp.wander_weight) + force;
        }
    }
    if (A.forces_only) {
        A.force_out[4 * i] = force.x;
        A.force_out[4 * i + 1] = force.y;
        A.force_out[4 * i + 2] = force.z;
        A.force_out[4 * i + 3] = R(0);
        return;
    }
    // genericSeekingWorldCenterD<false> (OpenSteerWrapper_kernel.cu:219-259)
    if (sp.seek_world_center) {
        // the box test is a discrete decision: FP32 on the FP32-rounded position
        const float cx = g.world_center[0], cy = g.world_center[1], cz = g.world_center[2];
        const float rx = g.world_radius[0], ry = g.world_radius[1], rz = g.world_radius[2];
        const V3<float> q = f32v(ppos(i));
        bool inside = (q.x > cx - rx) && (q.x < cx + rx) && (q.y > cy - ry) && (q.y < cy + ry) &&
                      (q.z > cz - rz) && (q.z < cz + rz);
        V3<R> s = inside ? v3<R>(0, 0, 0)
                         : steerForSeek(myPos, mySpeed, myFwd, v3<R>(R(cx), R(cy), R(cz)), R(sp.max_speed));
        force = s + force;
    }
    if (A.cond_out) A.cond_out[i] = cond;
    if (A.force_out) {
        A.force_out[4 * i] = force.x;
        A.force_out[4 * i + 1] = force.y;
        A.force_out[4 * i + 2] = force.z;
        A.force_out[4 * i + 3] = R(0);
    }

    // genericApplySteeringForceD (OpenSteerWrapper_kernel.cu:265-333) +
    // applySteeringForceSingle (OpenSteerWrapper_resources.cu:326-379)
    if (sp.use_2d_projection) force.y = R(0);
    const R maxSpeed = R(sp.max_speed);
    V3<R> adjusted = adjustRawSteeringForce(force, mySpeed, myFwd, maxSpeed);
    V3<R> clipped = truncateLength(adjusted, R(sp.max_force));
    V3<R> newAccel = clipped / R(sp.mass);
    if (dt > R(0)) {
        R smoothRate = clampr(R(9) * dt, R(0.15f), R(0.4f));
        smoothRate = clampr(smoothRate, R(0), R(1));
        mySa = mySa + (newAccel - mySa) * smoothRate;
    }
    V3<R> newVel = myFwd * mySpeed + mySa * dt;
    newVel = truncateLength(newVel, maxSpeed);
    mySpeed = length(newVel);
    V3<R> newPos = myPos + newVel * dt;
    if (sp.use_2d_projection) newVel.y = R(0);
    V3<R> newFwd = normalize(newVel);
    if (sp.use_2d_projection) newPos.y = R(0.5f);

    R* P = A.pos_out + 4 * (size_t)id;
    P[0] = newPos.x; P[1] = newPos.y; P[2] = newPos.z; P[3] = R(1);
    R* S = A.sa_out + 4 * (size_t)id;
    S[0] = mySa.x; S[1] = mySa.y; S[2] = mySa.z; S[3] = saW;
    R* F = A.fwd_out + 4 * (size_t)id;
    F[0] = newFwd.x; F[1] = newFwd.y; F[2] = newFwd.z; F[3] = mySpeed;
}

template <typename R>
void feature_agent(uint32_t i, R rate, const uint32_t* hash, const uint32_t* lists, uint32_t W,
                   const R* fin, R* fout) {
    const uint32_t id = hash[2 * i + 1];
    const uint32_t* L = lists + (size_t)i * W;
    const uint32_t count = L[0];
    const R* f = fin + 4 * (size_t)id;
    R* o = fout + 4 * (size_t)id;
    if (count == 0) {
        for (int c = 0; c < 4; ++c) o[c] = f[c];
        return;
    }
    R s[4] = {0, 0, 0, 0};
    for (uint32_t t = 0; t < count; ++t) {
        uint32_t jid = hash[2 * L[1 + t] + 1];
        const R* q = fin + 4 * (size_t)jid;
        for (int c = 0; c < 4; ++c) s[c] = s[c] + q[c];
    }
    const R inv = R(1) / R(count);  // cutil float4 / float
    for (int c = 0; c < 4; ++c) {
        R mean = s[c] * inv;
        o[c] = f[c] + (mean - f[c]) * rate;
    }
}

// ------------------------------------------------------------------ sort

// ------------------------------------------------------------------ BehaviorClustering

// cutil_math clamp(f, a, b) = fmaxf(a, fminf(f, b)) (IEEE minNum/maxNum: a NaN
// operand yields the other one)
template <typename R>
inline R cutil_clamp(R f, R a, R b) { return std::fmax(a, std::fmin(f, b)); }

// BehaviorClustering_resources.cu:61-80
template <typename R>
inline R linear_interp(R x, R xa, R ya, R xb, R yb) {
    return ya * ((x - xb) / (xa - xb)) + yb * ((x - xa) / (xb - xa));
}
template <typename R>
inline R interpolate_similarity(R v, R mn, R avg, R mx) {
    if (v > avg) return cutil_clamp(linear_interp(v, avg, R(0.5f), mx, R(1)), R(0), R(1));
    return cutil_clamp(linear_interp(v, mn, R(0), avg, R(0.5f)), R(0), R(1));
}

// BehaviorClustering_resources.cu:84-150; st = (min, avg, max)
template <typename R>
inline void update_similarity_stats(R sim, R st[3], const brt_bc_params& bc) {
    R mn = st[0], avg = st[1], mx = st[2];
    const R ctrl = (R)bc.avg_similarity_controller;
    const R blend = sim * ctrl + ((mx + mn) / R(2)) * (R(1) - ctrl);
    avg = avg + (blend - st[1]) * (R)bc.stats_params[0];
    avg = cutil_clamp(avg, R(0), R(1));
    if (sim < mn) {
        mn = mn + (sim - mn) * (R)bc.stats_params[1];
        if (mn > avg) mn = cutil_clamp(mn, R(0), sim);
    } else if (mn < avg) {
        mn = mn + (R)bc.stats_params[2];
    }
    if (sim > mx) {
        mx = mx + (sim - mx) * (R)bc.stats_params[1];
        if (mx < avg) mx = cutil_clamp(mx, sim, R(1));
    } else if (mx > avg) {
        mx = mx - (R)bc.stats_params[2];
    }
    st[0] = mn;
    st[1] = avg;
    st[2] = mx;
}

// computeNeighNaiveSimilarities + computeAdaptiveSimilarities for sorted agent i
// (BehaviorClustering_kernel.cu:306-353, 438-497; computeSimilarityFromTex
// BehaviorClustering_resources.cu:180-198)
template <typename R>
void bc_similarity_agent(uint32_t i, uint32_t n, uint32_t K, uint32_t W, const brt_bc_params& bc,
                         const uint32_t* hash, const uint32_t* lists, const R* features, const R* stats_in,
                         R* stats_out, R* sims) {
    const uint32_t F = bc.feature_count;
    const uint32_t id = hash[2 * i + 1];
    const uint32_t* row = lists + (size_t)i * W;
    const uint32_t count = std::min(row[0], K);
    R* out = sims + (size_t)i * (K + 1);
    for (uint32_t t = 0; t <= K; ++t) out[t] = R(0);
    if (id >= n) return;
    const R* fi = features + (size_t)id * F;
    R st[3] = {R(0), R(0), R(0)};
    if (bc.adaptive)
        for (int k = 0; k < 3; ++k) st[k] = stats_in[4 * (size_t)id + k];
    for (uint32_t t = 0; t < count; ++t) {
        const uint32_t j = row[1 + t];
        const uint32_t jid = j < n ? hash[2 * j + 1] : n;
        R sim = R(0);
        if (jid < n) {
            const R* fj = features + (size_t)jid * F;
            R n1 = R(0), n2 = R(0), d = R(0);
            for (uint32_t k = 0; k < F; ++k) {
                d = d + fi[k] * fj[k];
                n1 = n1 + fi[k] * fi[k];
                n2 = n2 + fj[k] * fj[k];
            }
            n1 = rsqrt_sqrt(n1);
            n2 = rsqrt_sqrt(n2);
            sim = (d / (n1 * n2) + R(1)) / R(2);
            if (bc.adaptive) {
                update_similarity_stats(sim, st, bc);
                sim = interpolate_similarity(sim, st[0], st[1], st[2]);
            }
        }
        out[t] = sim;
    }
    if (bc.adaptive) {
        for (int k = 0; k < 3; ++k) stats_out[4 * (size_t)id + k] = st[k];
        stats_out[4 * (size_t)id + 3] = stats_in[4 * (size_t)id + 3];
    }
}

// OpenSteerWrapper_resources.cu:311-322
template <typename R>
inline V3<R> steerForFlee(V3<R> myPos, R mySpeed, V3<R> myForward, V3<R> target, R maxSpeed) {
    V3<R> offset = myPos - target;
    V3<R> desired = truncateLength(offset, maxSpeed);
    V3<R> myVel = myForward * mySpeed;
    return desired - myVel;
}

// steerForClusterCohesion then steerForClusterAlignment for sorted agent i
// (BehaviorClustering_resources.cu:282-384), blended as term + force.
template <typename R>
void bc_cluster_agent(uint32_t i, uint32_t n, uint32_t K, uint32_t W, const brt_steer_params& sp,
                      const brt_bc_params& bc, const R* spos, const R* sfwd, const uint32_t* lists, const R* sims,
                      const R* force_in, R* force_out) {
    const V3<R> myPos = v3<R>(spos[4 * i], spos[4 * i + 1], spos[4 * i + 2]);
    const V3<R> myFwd = v3<R>(sfwd[4 * i], sfwd[4 * i + 1], sfwd[4 * i + 2]);
    const R mySpeed = sfwd[4 * i + 3];
    const R maxS = (R)sp.max_speed;
    const uint32_t* row = lists + (size_t)i * W;
    const uint32_t count = std::min(row[0], K);
    const R* s = sims + (size_t)i * (K + 1);
    V3<R> coh = v3<R>(0, 0, 0), ali = coh;
    for (uint32_t t = 0; t < count; ++t) {
        const uint32_t j = row[1 + t];
        if (j >= n) continue;
        const R sim = s[t];
        const V3<R> op = v3<R>(spos[4 * j], spos[4 * j + 1], spos[4 * j + 2]);
        const V3<R> of = v3<R>(sfwd[4 * j], sfwd[4 * j + 1], sfwd[4 * j + 2]);
        const V3<R> seek = steerForSeek(myPos, mySpeed, myFwd, op, maxS);
        const V3<R> flee = steerForFlee(myPos, mySpeed, myFwd, op, maxS);
        coh = coh + (seek * sim + flee * (R(1) - sim));
        ali = ali + of * sim;
    }
    if (count > 0) {
        coh = normalize(coh / (R)count) * (R)bc.cluster_weight;
        ali = normalize(ali / (R)count - myFwd) * (R)bc.alignment_weight;
    }
    const V3<R> f0 = v3<R>(force_in[4 * i], force_in[4 * i + 1], force_in[4 * i + 2]);
    const V3<R> f1 = coh + f0;
    const V3<R> f2 = ali + f1;
    force_out[4 * i] = f2.x;
    force_out[4 * i + 1] = f2.y;
    force_out[4 * i + 2] = f2.z;
    force_out[4 * i + 3] = R(0);
}

void radix_sort_pairs(uint32_t* pairs, uint32_t n, uint32_t num_cells, std::vector<uint32_t>& tmp) {
    uint32_t bits = 0;
    while (bits < 32 && (num_cells - 1u) >> bits) ++bits;
    uint32_t passes = (bits + 7) / 8;
    if (passes == 0 || n < 2) return;
    tmp.resize(2 * (size_t)n);
    uint32_t* src = pairs;
    uint32_t* dst = tmp.data();
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    std::vector<uint32_t> hist((size_t)nt * 256);
    for (uint32_t p = 0; p < passes; ++p) {
        const uint32_t shift = 8 * p;
        std::fill(hist.begin(), hist.end(), 0u);
#pragma omp parallel num_threads(nt)
        {
            int t = 0, T = 1;
#ifdef _OPENMP
            t = omp_get_thread_num();
            T = omp_get_num_threads();
#endif
            size_t lo = (size_t)n * t / T, hi = (size_t)n * (t + 1) / T;
            uint32_t* h = &hist[(size_t)t * 256];
            for (size_t i = lo; i < hi; ++i) h[(src[2 * i] >> shift) & 255u]++;
#pragma omp barrier
#pragma omp single
            {
                uint32_t run = 0;
                for (uint32_t d = 0; d < 256; ++d)
                    for (int u = 0; u < T; ++u) {
                        uint32_t c = hist[(size_t)u * 256 + d];
                        hist[(size_t)u * 256 + d] = run;
                        run += c;
                    }
            }
            for (size_t i = lo; i < hi; ++i) {
                uint32_t d = (src[2 * i] >> shift) & 255u;
                uint32_t o = h[d]++;
                dst[2 * (size_t)o] = src[2 * i];
                dst[2 * (size_t)o + 1] = src[2 * i + 1];
            }
        }
        std::swap(src, dst);
    }
    if (src != pairs) std::memcpy(pairs, src, sizeof(uint32_t) * 2 * (size_t)n);
}

}  // namespace

// ====================================================================== C ABI

// The behaviours alone (separation, alignment, cohesion, targets, wander):
// force_out (sorted) holds their sum, as the reference's steer force holds it
// after computeCohesions (OpenSteerWrapperPlugIn.h:288-375).
template <typename R>
void steer_forces(const brt_prox_params* p, const brt_steer_params* s, float dt, uint32_t step, uint32_t n,
                  const R* spos, const R* sfwd, const uint32_t* hash, const uint32_t* lists, R* wander,
                  R* force_out) {
    set_threads();
    brt_grid_params g{};
    SteerArrays<R> A{spos, sfwd, nullptr, hash, lists, listStride(p->max_neighbors), wander, nullptr, nullptr,
                     nullptr, force_out, nullptr, nullptr, nullptr, true, nullptr};
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) steer_agent<R>((uint32_t)i, g, *s, R(dt), step, A);
}

// computeSeekingsWorldCenter + applySteeringForces on a given steer force
// (OpenSteerWrapperPlugIn.h:380-419).
template <typename R>
void apply_steering_force(const brt_grid_params* g, const brt_steer_params* s, float dt, uint32_t n,
                          const R* spos, const R* sfwd, const R* ssa, const uint32_t* hash, const R* force_in,
                          R* pos_out, R* fwd_out, R* sa_out, R* force_out) {
    set_threads();
    SteerArrays<R> A{spos, sfwd, ssa, hash, nullptr, 0, nullptr, pos_out, fwd_out, sa_out, force_out, nullptr,
                     nullptr, nullptr, false, force_in};
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) steer_agent<R>((uint32_t)i, *g, *s, R(dt), 0, A);
}

extern "C" {

void orc_set_threads(int threads) { g_threads = threads; }

int orc_num_threads(void) {
    set_threads();
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    philox4x32_10(ctr, key, out);
}

// EnvGrid3DPlugIn::setupGrid (PlugIns/EnvGrid3D/include/EnvGrid3DPlugIn.h:478-508)
void orc_grid_setup(const float center[3], const float radius[3], const float cell[3],
                    uint32_t max_bodies_per_cell, brt_grid_params* out) {
    std::memset(out, 0, sizeof(*out));
    for (int a = 0; a < 3; ++a) {
        float aug = radius[a] + cell[a] * 2.0f;
        out->grid_size[a] = (uint32_t)(aug * 2.0f / cell[a]);
        out->world_origin[a] = center[a] - aug;
        out->cell_size[a] = cell[a];
        out->world_center[a] = center[a];
        out->world_radius[a] = radius[a];
    }
    out->num_cells = out->grid_size[0] * out->grid_size[1] * out->grid_size[2];
    out->max_bodies_per_cell = max_bodies_per_cell;
}

void orc_calc_grid_pos(const brt_grid_params* g, const float* pos, uint32_t n, int32_t* out3) {
    for (uint32_t i = 0; i < n; ++i) {
        I3 c = calcGridPos(pos + 4 * (size_t)i, *g);
        out3[3 * i] = c.x;
        out3[3 * i + 1] = c.y;
        out3[3 * i + 2] = c.z;
    }
}

// genCalcHashD
void orc_calc_hash(const brt_grid_params* g, const float* pos, uint32_t n, uint32_t* hash) {
    set_threads();
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        hash[2 * i] = calcGridHash(calcGridPos(pos + 4 * i, *g), *g);
        hash[2 * i + 1] = (uint32_t)i;
    }
}

// Textbook stable sort by key (the semantics of the reference's LSD radix sort).
void orc_sort_stable(uint32_t* hash, uint32_t n) {
    std::vector<uint64_t> v(n);
    for (uint32_t i = 0; i < n; ++i) v[i] = ((uint64_t)hash[2 * i] << 32) | hash[2 * i + 1];
    std::stable_sort(v.begin(), v.end(),
                     [](uint64_t a, uint64_t b) { return (a >> 32) < (b >> 32); });
    for (uint32_t i = 0; i < n; ++i) {
        hash[2 * i] = (uint32_t)(v[i] >> 32);
        hash[2 * i + 1] = (uint32_t)v[i];
    }
}

// Parallel LSD radix sort (stable), 8-bit digits over the significant key bits.
void orc_sort_radix(uint32_t* hash, uint32_t n, uint32_t num_cells) {
    set_threads();
    std::vector<uint32_t> tmp;
    radix_sort_pairs(hash, n, num_cells, tmp);
}

// reorderDataFloat4
void orc_reorder(const uint32_t* hash, uint32_t n, const float* in, float* out) {
    set_threads();
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        uint32_t id = hash[2 * i + 1];
        std::memcpy(out + 4 * i, in + 4 * (size_t)id, 16);
    }
}

// genericFindCellStartD with explicit empty-cell value and cell end.
void orc_cell_ranges(const uint32_t* hash, uint32_t n, uint32_t num_cells, uint32_t* start,
                     uint32_t* end) {
    set_threads();
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < (int64_t)num_cells; ++c) {
        start[c] = BRT_EMPTY_CELL;
        end[c] = BRT_EMPTY_CELL;
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        uint32_t k = hash[2 * i];
        if (i == 0 || hash[2 * (i - 1)] != k) start[k] = (uint32_t)i;
        if (i == (int64_t)n - 1 || hash[2 * (i + 1)] != k) end[k] = (uint32_t)(i + 1);
    }
}

// genericCalcNeighgborhoodD: 27 cells, x outermost, z innermost, each cell in
// sorted order, truncated at max_bodies_per_cell, self excluded,
// "sqrtf(d2) > r -> reject", capped at K (or KNN replacement).
void orc_neighbors(const brt_grid_params* g, const brt_prox_params* p, const float* spos,
                   const uint32_t* hash, const uint32_t* start, const uint32_t* end, uint32_t n,
                   uint32_t* lists) {
    set_threads();
    const uint32_t K = p->max_neighbors;
    const uint32_t W = listStride(K);
    const float r = p->search_radius;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const float* pi = spos + 4 * (size_t)i;
        I3 gp = calcGridPos(pi, *g);
        uint32_t cnt = 0;
        uint32_t nl[BRT_MAX_NEIGHBORS + 4];
        float dl[BRT_MAX_NEIGHBORS + 4];
        float maxDist = 0.0f;
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz) {
                    I3 c = {gp.x + dx, gp.y + dy, gp.z + dz};
                    if (!inGrid(c, *g)) continue;
                    uint32_t key = calcGridHash(c, *g);
                    uint32_t s = start[key];
                    if (s == BRT_EMPTY_CELL) continue;
                    for (uint32_t t = 0; t < g->max_bodies_per_cell; ++t) {
                        uint32_t j = s + t;
                        if (j >= n || hash[2 * (size_t)j] != key) break;
                        if (j == i) continue;
                        const float* pj = spos + 4 * (size_t)j;
                        float rx = pj[0] - pi[0], ry = pj[1] - pi[1], rz = pj[2] - pi[2];
                        float xx = rx * rx, yy = ry * ry, zz = rz * rz;
                        float dist = sqrtf((xx + yy) + zz);
                        if (dist > r) continue;
                        if (cnt < K) {
                            nl[cnt] = j;
                            if (p->use_knn) {
                                dl[cnt] = dist;
                                maxDist = std::max(maxDist, dist);
                            }
                            ++cnt;
                        } else if (p->use_knn && dist < maxDist) {
                            maxDist = dl[0];
                            uint32_t mi = 0;
                            for (uint32_t q = 1; q < cnt; ++q)
                                if (maxDist < dl[q]) {
                                    maxDist = dl[q];
                                    mi = q;
                                }
                            nl[mi] = j;
                            dl[mi] = dist;
                        }
                    }
                }
        uint32_t* L = lists + (size_t)i * W;
        L[0] = cnt;
        for (uint32_t q = 0; q < cnt; ++q) L[1 + q] = nl[q];
        for (uint32_t q = cnt + 1; q < W; ++q) L[q] = BRT_LIST_PAD;
    }
}

// Independent O(N^2) formulation for tiny N: enumerate every pair, place each
// accepted candidate at (canonical cell index, sorted index), sort, keep K.
// Non-KNN only.
void orc_neighbors_brute(const brt_grid_params* g, const brt_prox_params* p, const float* spos,
                         const uint32_t* hash, uint32_t n, uint32_t* lists) {
    const uint32_t K = p->max_neighbors;
    const uint32_t W = listStride(K);
    std::vector<uint32_t> first(n);
    for (uint32_t j = 0; j < n; ++j) {
        uint32_t f = j;
        while (f > 0 && hash[2 * (f - 1)] == hash[2 * j]) --f;
        first[j] = f;
    }
    std::vector<std::pair<uint32_t, uint32_t>> cand;
    for (uint32_t i = 0; i < n; ++i) {
        I3 gp = calcGridPos(spos + 4 * (size_t)i, *g);
        cand.clear();
        for (uint32_t j = 0; j < n; ++j) {
            if (j == i) continue;
            int order = -1;
            for (int o = 0; o < 27 && order < 0; ++o) {
                I3 c = {gp.x + o / 9 - 1, gp.y + (o / 3) % 3 - 1, gp.z + o % 3 - 1};
                if (inGrid(c, *g) && calcGridHash(c, *g) == hash[2 * j]) order = o;
            }
            if (order < 0) continue;
            if (j - first[j] >= g->max_bodies_per_cell) continue;
            // single-precision test exactly as the reference evaluates it
            float rx = spos[4 * j] - spos[4 * i], ry = spos[4 * j + 1] - spos[4 * i + 1],
                  rz = spos[4 * j + 2] - spos[4 * i + 2];
            float xx = rx * rx, yy = ry * ry, zz = rz * rz;
            if (sqrtf((xx + yy) + zz) > p->search_radius) continue;
            cand.push_back({(uint32_t)order, j});
        }
        std::sort(cand.begin(), cand.end());
        uint32_t cnt = (uint32_t)std::min<size_t>(cand.size(), K);
        uint32_t* L = lists + (size_t)i * W;
        L[0] = cnt;
        for (uint32_t q = 0; q < cnt; ++q) L[1 + q] = cand[q].second;
        for (uint32_t q = cnt + 1; q < W; ++q) L[q] = BRT_LIST_PAD;
    }
}

void orc_steer_f32(const brt_grid_params* g, const brt_prox_params* p, const brt_steer_params* s,
                   float dt, uint32_t step, uint32_t n, const float* spos, const float* sfwd,
                   const float* ssa, const uint32_t* hash, const uint32_t* lists, float* wander,
                   float* pos_out, float* fwd_out, float* sa_out, float* force_out) {
    set_threads();
    SteerArrays<float> A{spos, sfwd, ssa, hash, lists, listStride(p->max_neighbors),
                         wander, pos_out, fwd_out, sa_out, force_out, nullptr, nullptr, nullptr, false, nullptr};
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) steer_agent<float>((uint32_t)i, *g, *s, dt, step, A);
}

void orc_steer_f64(const brt_grid_params* g, const brt_prox_params* p, const brt_steer_params* s,
                   float dt, uint32_t step, uint32_t n, const double* spos, const double* sfwd,
                   const double* ssa, const uint32_t* hash, const uint32_t* lists, double* wander,
                   double* pos_out, double* fwd_out, double* sa_out, double* force_out, double* cond_out,
                   const float* pred_pos, const float* pred_fwd) {
    set_threads();
    SteerArrays<double> A{spos, sfwd, ssa, hash, lists, listStride(p->max_neighbors),
                          wander, pos_out, fwd_out, sa_out, force_out, cond_out, pred_pos, pred_fwd, false,
                          nullptr};
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) steer_agent<double>((uint32_t)i, *g, *s, (double)dt, step, A);
}

void orc_steer_forces_f32(const brt_prox_params* p, const brt_steer_params* s, float dt, uint32_t step, uint32_t n,
                          const float* spos, const float* sfwd, const uint32_t* hash, const uint32_t* lists,
                          float* wander, float* force_out) {
    steer_forces<float>(p, s, dt, step, n, spos, sfwd, hash, lists, wander, force_out);
}
void orc_steer_forces_f64(const brt_prox_params* p, const brt_steer_params* s, float dt, uint32_t step, uint32_t n,
                          const double* spos, const double* sfwd, const uint32_t* hash, const uint32_t* lists,
                          double* wander, double* force_out) {
    steer_forces<double>(p, s, dt, step, n, spos, sfwd, hash, lists, wander, force_out);
}
void orc_apply_steering_force_f32(const brt_grid_params* g, const brt_steer_params* s, float dt, uint32_t n,
                                  const float* spos, const float* sfwd, const float* ssa, const uint32_t* hash,
                                  const float* force_in, float* pos_out, float* fwd_out, float* sa_out,
                                  float* force_out) {
    apply_steering_force<float>(g, s, dt, n, spos, sfwd, ssa, hash, force_in, pos_out, fwd_out, sa_out, force_out);
}
void orc_apply_steering_force_f64(const brt_grid_params* g, const brt_steer_params* s, float dt, uint32_t n,
                                  const double* spos, const double* sfwd, const double* ssa, const uint32_t* hash,
                                  const double* force_in, double* pos_out, double* fwd_out, double* sa_out,
                                  double* force_out) {
    apply_steering_force<double>(g, s, dt, n, spos, sfwd, ssa, hash, force_in, pos_out, fwd_out, sa_out, force_out);
}

void orc_feature_update_f32(const brt_prox_params* p, float rate, uint32_t n, const uint32_t* hash,
                            const uint32_t* lists, const float* fin, float* fout) {
    set_threads();
    const uint32_t W = listStride(p->max_neighbors);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) feature_agent<float>((uint32_t)i, rate, hash, lists, W, fin, fout);
}

void orc_feature_update_f64(const brt_prox_params* p, float rate, uint32_t n, const uint32_t* hash,
                            const uint32_t* lists, const double* fin, double* fout) {
    set_threads();
    const uint32_t W = listStride(p->max_neighbors);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i)
        feature_agent<double>((uint32_t)i, (double)rate, hash, lists, W, fin, fout);
}

void orc_bc_similarities_f32(const brt_prox_params* p, const brt_bc_params* bc, uint32_t n, const uint32_t* hash,
                             const uint32_t* lists, const float* features, const float* stats_in,
                             float* stats_out, float* sims) {
    set_threads();
    const uint32_t K = p->max_neighbors, W = listStride(K);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i)
        bc_similarity_agent<float>((uint32_t)i, n, K, W, *bc, hash, lists, features, stats_in, stats_out, sims);
}

void orc_bc_similarities_f64(const brt_prox_params* p, const brt_bc_params* bc, uint32_t n, const uint32_t* hash,
                             const uint32_t* lists, const double* features, const double* stats_in,
                             double* stats_out, double* sims) {
    set_threads();
    const uint32_t K = p->max_neighbors, W = listStride(K);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i)
        bc_similarity_agent<double>((uint32_t)i, n, K, W, *bc, hash, lists, features, stats_in, stats_out, sims);
}

void orc_bc_cluster_steer_f32(const brt_prox_params* p, const brt_steer_params* sp, const brt_bc_params* bc,
                              uint32_t n, const float* spos, const float* sfwd, const uint32_t* lists,
                              const float* sims, const float* force_in, float* force_out) {
    set_threads();
    const uint32_t K = p->max_neighbors, W = listStride(K);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i)
        bc_cluster_agent<float>((uint32_t)i, n, K, W, *sp, *bc, spos, sfwd, lists, sims, force_in, force_out);
}

void orc_bc_cluster_steer_f64(const brt_prox_params* p, const brt_steer_params* sp, const brt_bc_params* bc,
                              uint32_t n, const double* spos, const double* sfwd, const uint32_t* lists,
                              const double* sims, const double* force_in, double* force_out) {
    set_threads();
    const uint32_t K = p->max_neighbors, W = listStride(K);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i)
        bc_cluster_agent<double>((uint32_t)i, n, K, W, *sp, *bc, spos, sfwd, lists, sims, force_in, force_out);
}

// BehaviorClusteringPlugIn::computeClusterMembership (BehaviorClusteringPlugIn.h:823-865)
// as the reference runs it: float3 labels by agent ID in a double buffer;
// resetClusterMembership (BehaviorClustering_kernel.cu:632-652) writes the
// sorted index, each computeClusterMembership_kernel pass (:583-625) the least
// label of the agent and its listed neighbours (gathered through the
// sorted->ID map), CMConnection_kernel (:657-689) the position pairs.
void orc_cluster_membership(const brt_prox_params* p, uint32_t n, uint32_t iterations, const uint32_t* hash,
                            const uint32_t* lists, const float* spos, const float* m_in, float* m_out,
                            float* connector) {
    set_threads();
    const uint32_t W = listStride(p->max_neighbors);
    std::vector<float> rd(m_in, m_in + 3 * (size_t)n), wr(3 * (size_t)n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const uint32_t id = hash[2 * i + 1];
        wr[3 * id] = (float)(uint32_t)i;
        wr[3 * id + 1] = rd[3 * id + 1];
        wr[3 * id + 2] = rd[3 * id + 2];
    }
    std::swap(rd, wr);
    for (uint32_t t = 0; t < iterations; ++t) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < (int64_t)n; ++i) {
            const uint32_t id = hash[2 * i + 1];
            const uint32_t* L = lists + (size_t)i * W;
            float mn = rd[3 * id];
            for (uint32_t q = 0; q < L[0]; ++q) {
                const float other = rd[3 * (size_t)hash[2 * (size_t)L[1 + q] + 1]];
                if (mn > other) mn = other;
            }
            wr[3 * id] = mn;
            wr[3 * id + 1] = rd[3 * id + 1];
            wr[3 * id + 2] = rd[3 * id + 2];
        }
        std::swap(rd, wr);
    }
    std::copy(rd.begin(), rd.end(), m_out);
    if (!connector) return;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const uint32_t id = hash[2 * i + 1];
        const uint32_t leader = (uint32_t)rd[3 * id];
        for (int c = 0; c < 4; ++c) {
            connector[8 * (size_t)id + c] = spos[4 * i + c];
            connector[8 * (size_t)id + 4 + c] = spos[4 * (size_t)leader + c];
        }
    }
}

// BehaviorClustering_clusterHitRate (BehaviorClustering_kernel.cu:502-578).
void orc_cluster_hit_rate(const brt_prox_params* p, float radius, uint32_t n, const uint32_t* hash,
                          const uint32_t* lists, const float* spos, const float* m_in, float* m_out) {
    set_threads();
    const uint32_t W = listStride(p->max_neighbors);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const uint32_t id = hash[2 * i + 1];
        const uint32_t* L = lists + (size_t)i * W;
        const float my = m_in[3 * (size_t)id];
        const V3<float> myPos = ld3(spos, (uint32_t)i);
        int close = 0;
        float real = 0.0f, hit = 0.0f;
        for (uint32_t q = 0; q < L[0]; ++q) {
            const uint32_t j = L[1 + q];
            const float other = m_in[3 * (size_t)hash[2 * (size_t)j + 1]];
            const V3<float> offset = ld3(spos, j) - myPos;
            if (length(offset) > radius) continue;
            ++close;
            real += other;
            if (my == other) hit += 1.0f;
        }
        if (close > 0) {
            hit /= (float)close;
            real /= close;
        } else {
            real = -1.0f;
        }
        m_out[3 * (size_t)id] = my;
        m_out[3 * (size_t)id + 1] = hit;
        m_out[3 * (size_t)id + 2] = real;
    }
}

// steerToAvoidNeighbors (OpenSteerWrapper_kernel.cu:110-154): calcSeparation
// (OpenSteerWrapper_resources.cu:166-211) times its weight, amplified above 16
// listed agents, then the speed rule with the reference's double constants.
void orc_slow_avoid(const brt_prox_params* p, const brt_steer_params* sp, uint32_t n, const float* spos,
                    const float* sfwd, const uint32_t* lists, float* fwd_out, const float* force_in,
                    float* force_out) {
    set_threads();
    const uint32_t W = listStride(p->max_neighbors);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const V3<float> myPos = ld3(spos, (uint32_t)i), myFwd = ld3(sfwd, (uint32_t)i);
        const uint32_t* L = lists + (size_t)i * W;
        const uint32_t count = L[0];
        V3<float> steering = v3<float>(0, 0, 0);
        uint32_t neighbors = 0;
        for (uint32_t q = 0; q < count; ++q) {
            const V3<float> other = ld3(spos, L[1 + q]);
            if (inBoidNeighborhood(myPos, other, sp->sensing_min_radius, sp->separation[1], sp->separation[2],
                                   myFwd)) {
                const V3<float> offset = other - myPos;
                const float d2 = dot(offset, offset);
                steering = steering + offset / (-d2);
                ++neighbors;
            }
        }
        if (neighbors > 0) steering = normalize(steering / (float)neighbors);
        steering = steering * sp->separation[0];
        const int neighNumThreshold = 16;
        const float steerForceMultipler = 0.5f;
        if (count > (uint32_t)neighNumThreshold)
            steering = steering * ((float)(count - neighNumThreshold) * steerForceMultipler);
        float mySpeed = sfwd[4 * i + 3];
        if (dot(steering, myFwd) < -0.8)
            mySpeed *= 0.9;
        else if (dot(steering, myFwd) > 0.5)
            mySpeed *= 1.02;
        fwd_out[4 * i] = myFwd.x;
        fwd_out[4 * i + 1] = myFwd.y;
        fwd_out[4 * i + 2] = myFwd.z;
        fwd_out[4 * i + 3] = mySpeed;
        force_out[4 * i] = steering.x + force_in[4 * i];
        force_out[4 * i + 1] = steering.y + force_in[4 * i + 1];
        force_out[4 * i + 2] = steering.z + force_in[4 * i + 2];
        force_out[4 * i + 3] = 0.0f;
    }
}

// extractColorFromNeighborhoodD (drawable3d_kernel.cu:60-93), by agent ID.
void orc_neighborhood_color(const brt_prox_params* p, const float* base, uint32_t vertexes, uint32_t n,
                            const uint32_t* hash, const uint32_t* lists, float* color) {
    const uint32_t W = listStride(p->max_neighbors);
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t id = hash[2 * i + 1];
        const float colorFactor = ((float)lists[(size_t)i * W]) / (float)p->max_neighbors;
        const float c[4] = {base[0] + colorFactor, (float)(base[1] + 0.3 + 1 - colorFactor),
                            base[2] + colorFactor * (1 - colorFactor), 1.0f};
        if (vertexes == 3) {
            for (int k = 0; k < 4; ++k) {
                color[12 * (size_t)id + k] = c[k] - (k < 3 ? 0.5f : 0.0f);
                color[12 * (size_t)id + 4 + k] = c[k];
                color[12 * (size_t)id + 8 + k] = c[k];
            }
        } else {
            for (int k = 0; k < 4; ++k) color[4 * (size_t)id + k] = c[k];
        }
    }
}

// moveTowardsTargetD (Shapes3D_kernel.cu:63-87); index n (and beyond) reads as zeros.
void orc_move_towards_target(int shift, const float* base, uint32_t n, const uint32_t* hash, const float* spos,
                             const float* targets, const float* force_in, float* force_out) {
    for (uint32_t i = 0; i < n; ++i) {
        int shiftedIndex = (int)hash[2 * i + 1] + shift;
        if (shiftedIndex > (int)n) shiftedIndex -= (int)n;
        float t[4] = {0, 0, 0, 0};
        if (shiftedIndex >= 0 && shiftedIndex < (int)n)
            for (int k = 0; k < 4; ++k) t[k] = targets[4 * (size_t)shiftedIndex + k];
        const V3<float> steering = (v3<float>(t[0], t[1], t[2]) + v3<float>(base[0], base[1], base[2])) -
                                   ld3(spos, i);
        const V3<float> f = (normalize(steering) * 12.0f) * t[3];
        force_out[4 * i] = f.x + force_in[4 * i];
        force_out[4 * i + 1] = f.y + force_in[4 * i + 1];
        force_out[4 * i + 2] = f.z + force_in[4 * i + 2];
        force_out[4 * i + 3] = 0.0f;
    }
}

// Whole FP32 step on host buffers laid out as brt_step_buffers (host pointers).
// Phase wall times (seconds) are added into phase_s[0..4] when non-null.
void orc_step_f32(const brt_grid_params* g, const brt_prox_params* p, const brt_steer_params* s,
                  float feature_rate, float dt, uint32_t step, const brt_step_buffers* b,
                  double* phase_s) {
    set_threads();
    const uint32_t n = b->n;
    auto now = []() {
#ifdef _OPENMP
        return omp_get_wtime();
#else
        return 0.0;
#endif
    };
    double t0 = now();
    orc_calc_hash(g, b->pos, n, b->hash);
    orc_sort_radix(b->hash, n, g->num_cells);
    double t1 = now();
    orc_reorder(b->hash, n, b->pos, b->sorted_pos);
    orc_reorder(b->hash, n, b->forward, b->sorted_forward);
    orc_reorder(b->hash, n, b->smooth_accel, b->sorted_smooth_accel);
    orc_cell_ranges(b->hash, n, g->num_cells, b->cell_start, b->cell_end);
    double t2 = now();
    orc_neighbors(g, p, b->sorted_pos, b->hash, b->cell_start, b->cell_end, n, b->lists);
    double t3 = now();
    orc_steer_f32(g, p, s, dt, step, n, b->sorted_pos, b->sorted_forward, b->sorted_smooth_accel,
                  b->hash, b->lists, b->wander, b->pos, b->forward, b->smooth_accel, nullptr);
    double t4 = now();
    if (b->features && b->features_out && feature_rate != 0.0f)
        orc_feature_update_f32(p, feature_rate, n, b->hash, b->lists, b->features, b->features_out);
    double t5 = now();
    if (phase_s) {
        phase_s[0] += t1 - t0;
        phase_s[1] += t2 - t1;
        phase_s[2] += t3 - t2;
        phase_s[3] += t4 - t3;
        phase_s[4] += t5 - t4;
    }
}

}  // extern "C"
