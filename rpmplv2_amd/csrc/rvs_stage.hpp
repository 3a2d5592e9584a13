// Device helpers shared by the kernel files: obstacle staging in LDS,
// oriented-box frame changes and per-link constants.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rvs_geometry.hpp"
#include "rvs_internal.hpp"

namespace rvs {

// LDS image of an obstacle set:
//   xf   [M][12]  centre, rotation rows (oriented boxes only; 16-byte aligned)
//   box  [M]      AABB, or [-h, h] in the box frame for oriented boxes
//   tol  [M]      min_dist_tol
//   gnd  [M]      ground flag
template <typename T>
struct Staged {
    const float* xf;
    Box<T>* box;
    T* tol;
    uint8_t* gnd;
    bool obb;
};

template <typename T>
__host__ __device__ inline size_t staged_bytes(const EnvDev& env) {
    const size_t M = (size_t)env.num_obj;
    return M * ((env.use_obb ? 12 * sizeof(float) : 0) + sizeof(Box<T>) + sizeof(T) + 1);
}

// Copy the obstacle set into LDS (whole workgroup participates).  OBB is a
// compile-time choice so that the axis-aligned kernels keep a constant layout;
// it must equal env.use_obb.
template <typename T, bool OBB>
__device__ inline Staged<T> stage_env(const EnvDev& env, unsigned char* smem) {
    const int M = env.num_obj;
    constexpr bool obb = OBB;
    float* xf = reinterpret_cast<float*>(smem);
    Staged<T> s;
    s.xf = xf;
    s.box = reinterpret_cast<Box<T>*>(smem + (obb ? (size_t)M * 12 * sizeof(float) : 0));
    s.tol = reinterpret_cast<T*>(s.box + M);
    s.gnd = reinterpret_cast<uint8_t*>(s.tol + M);
    s.obb = obb;
    for (int j = threadIdx.x; j < M; j += blockDim.x) {
        Box<T> b;
        if (obb) {
            const float* o = env.obb + 15 * j;
#pragma unroll
            for (int k = 0; k < 12; ++k) xf[12 * j + k] = o[k];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                b.lo[k] = -(T)o[12 + k];
                b.hi[k] = (T)o[12 + k];
            }
        } else {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                b.lo[k] = (T)env.aabb[6 * j + k];
                b.hi[k] = (T)env.aabb[6 * j + 3 + k];
            }
        }
        s.box[j] = b;
        s.tol[j] = (T)env.tol[j];
        s.gnd[j] = env.flags[j] & kObjGround;
    }
    __syncthreads();
    return s;
}

// World point -> box frame: local_k = R[0][k] (x - c_x) + R[1][k] (y - c_y)
// + R[2][k] (z - c_z), summed left to right (x = centre, rotation rows).
template <typename T>
__device__ __attribute__((always_inline)) inline Vec3<T> to_box(const float* x, const Vec3<T>& P) {
    const T vx = P.x - (T)x[0], vy = P.y - (T)x[1], vz = P.z - (T)x[2];
    return Vec3<T>{(T)x[3] * vx + (T)x[6] * vy + (T)x[9] * vz, (T)x[4] * vx + (T)x[7] * vy + (T)x[10] * vz,
                   (T)x[5] * vx + (T)x[8] * vy + (T)x[11] * vz};
}

// Box frame -> world: c_i + (R[i][0] l_0 + R[i][1] l_1 + R[i][2] l_2).
template <typename T>
__device__ __attribute__((always_inline)) inline Vec3<T> from_box(const float* x, const Vec3<T>& l) {
    return Vec3<T>{(T)x[0] + ((T)x[3] * l.x + (T)x[4] * l.y + (T)x[5] * l.z),
                   (T)x[1] + ((T)x[6] * l.x + (T)x[7] * l.y + (T)x[8] * l.z),
                   (T)x[2] + ((T)x[9] * l.x + (T)x[10] * l.y + (T)x[11] * l.z)};
}

// Broad phase.  The distance the case tree returns for a (capsule, box) pair
// is at least the raw segment-box distance minus r, except for its snaps to
// 0 / -r, which need a raw distance below EQUALITY_THRESHOLD (or contact).
// The gap between the segment's bounding box [slo, shi] and a box containing
// the obstacle (cb = lo[3], hi[3]) is a lower bound of the raw distance, so a
// pair whose gap exceeds max(running link minimum + r, EQUALITY_THRESHOLD)
// plus a margin that covers float rounding can neither lower the link's
// minimum nor end the loop: skipping it leaves every result unchanged.
// Compares squared values (no square root); a NaN end point never culls.
constexpr float kCullMargin = 1e-5f;

template <typename T>
__device__ __attribute__((always_inline)) inline bool cull_pair(const float* __restrict__ cb, const Vec3<T>& slo,
                                                                 const Vec3<T>& shi, T bound) {
    T g2 = T(0);
    {
        T g = (T)cb[0] - shi.x;
        const T g1 = slo.x - (T)cb[3];
        g = g1 > g ? g1 : g;
        g2 = g > T(0) ? g * g : T(0);
    }
    {
        T g = (T)cb[1] - shi.y;
        const T g1 = slo.y - (T)cb[4];
        g = g1 > g ? g1 : g;
        g2 = g > T(0) ? g2 + g * g : g2;
    }
    {
        T g = (T)cb[2] - shi.z;
        const T g1 = slo.z - (T)cb[5];
        g = g1 > g ? g1 : g;
        g2 = g > T(0) ? g2 + g * g : g2;
    }
    return g2 > bound * bound;
}

template <typename T>
__device__ __attribute__((always_inline)) inline T cull_bound(T prof, T r, T eq) {
    const T a = prof + r;
    return (a > eq ? a : eq) + (T)kCullMargin;
}

template <typename T>
__device__ __attribute__((always_inline)) inline void seg_bounds(const Vec3<T>& A, const Vec3<T>& B, Vec3<T>& lo,
                                                                 Vec3<T>& hi) {
    lo = Vec3<T>{A.x < B.x ? A.x : B.x, A.y < B.y ? A.y : B.y, A.z < B.z ? A.z : B.z};
    hi = Vec3<T>{A.x < B.x ? B.x : A.x, A.y < B.y ? B.y : A.y, A.z < B.z ? B.z : A.z};
}

template <typename T> __device__ inline T rnew_of(const RobotDev& rb, int i);
template <> __device__ inline float rnew_of<float>(const RobotDev& rb, int i) { return rb.rnew32[i]; }
template <> __device__ inline double rnew_of<double>(const RobotDev& rb, int i) { return rb.rnew64[i]; }

}  // namespace rvs
