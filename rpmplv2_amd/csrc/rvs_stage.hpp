// Device helpers shared by the kernel files: obstacle staging in LDS and
// per-link constants.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rvs_geometry.hpp"
#include "rvs_internal.hpp"

namespace rvs {

template <typename T>
struct Staged {
    Box<T>* box;
    T* tol;
    uint8_t* gnd;
};

template <typename T>
__host__ __device__ inline size_t staged_bytes(int M) {
    return (size_t)M * (sizeof(Box<T>) + sizeof(T) + 1);
}

// Copy the obstacle set into LDS (whole workgroup participates).
template <typename T>
__device__ inline Staged<T> stage_env(const EnvDev& env, unsigned char* smem) {
    const int M = env.num_obj;
    Staged<T> s;
    s.box = reinterpret_cast<Box<T>*>(smem);
    s.tol = reinterpret_cast<T*>(s.box + M);
    s.gnd = reinterpret_cast<uint8_t*>(s.tol + M);
    for (int j = threadIdx.x; j < M; j += blockDim.x) {
        Box<T> b;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            b.lo[k] = (T)env.aabb[6 * j + k];
            b.hi[k] = (T)env.aabb[6 * j + 3 + k];
        }
        s.box[j] = b;
        s.tol[j] = (T)env.tol[j];
        s.gnd[j] = env.flags[j] & kObjGround;
    }
    __syncthreads();
    return s;
}

template <typename T> __device__ inline T rnew_of(const RobotDev& rb, int i);
template <> __device__ inline float rnew_of<float>(const RobotDev& rb, int i) { return rb.rnew32[i]; }
template <> __device__ inline double rnew_of<double>(const RobotDev& rb, int i) { return rb.rnew64[i]; }

}  // namespace rvs
