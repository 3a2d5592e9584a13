// Enclosing radii of the robot skeleton and the Eigen reduction orders the
// bubble arithmetic depends on (device code, compiled without contraction).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#include "rvs_fk.hpp"
#include "rvs_geometry.hpp"

namespace rvs {

// AbstractRobot::computeEnclosingRadii in the reference's float arithmetic:
// R(i, j) = max(R(i, j-1), |P_j - P_i| + r_enc[j-1]) for j > i, zero
// elsewhere (Planar2DOF.cpp:132-148).  Chains measure row 0 in the {x, y}
// plane, |P_j.xy| (Spatial10DOF.cpp:11-30; the same as the planar rule for
// planar chains based at the origin).  xArm6 (xArm6.cpp:237-275): rows
// 0 .. n-2 only; row 0 from j = 2 with |P_j.xy|; row 3 from j = 5 with the
// distance of P_j to its projection on the line P3 P4; capsules_radius_new.
// std::max keeps the running value when the new one is NaN.  The skeleton is
// a MatrixXf, so (col(j) - col(i)).norm() is a dynamic-size reduction,
// summed left to right ((x^2 + y^2) + z^2); the projection in row 3 mixes in
// fixed-size Vector3f operands and sums as x + (y + z).
__device__ inline float norm_dyn3(const Vec3<float>& v) { return sqrtf((v.x * v.x + v.y * v.y) + v.z * v.z); }

__device__ inline float max_std(float a, float b) { return (a < b) ? b : a; }  // std::max(a, b)
__device__ inline float min_std(float a, float b) { return (b < a) ? b : a; }  // std::min(a, b)

// Advance the running radii from column j-1 to column j (1 <= j <= dofs):
// acc(i) holds R(i, j-1) on entry and R(i, j) on exit.  Rows i >= j stay 0.
// pt(k) returns skeleton point k.
template <typename PT, typename ACC>
__device__ inline void radii_column(const RobotDev& rb, PT pt, int j, ACC acc) {
    const int D = rb.num_dofs;
    const bool xarm = rb.kind == 1;
    const Vec3<float> Pj = pt(j);
    const float re = rb.rad_enc[j - 1];
    for (int i = 0; i < j && i < D; ++i) {
        if (xarm && i == D - 1) continue;
        float v;
        if (i == 0) {
            if (xarm && j < 2) continue;
            v = sqrtf(Pj.x * Pj.x + Pj.y * Pj.y);
        } else if (xarm && i == 3) {
            if (j < 5) continue;
            const Vec3<float> P3 = pt(3), P4 = pt(4);
            const Vec3<float> AB = P4 - P3;
            const float t = dot(Pj - P3, AB) / sqnorm(AB);
            const Vec3<float> proj{P3.x + t * AB.x, P3.y + t * AB.y, P3.z + t * AB.z};
            v = sqrtf(sqnorm(Pj - proj));
        } else {
            v = norm_dyn3(Pj - pt(i));
        }
        acc(i) = max_std(acc(i), v + re);
    }
}

// Sum of c(0) .. c(n-1), n >= 1, in the order of Eigen's vectorised redux
// for dynamic-size float expressions with 4-wide SSE packets (no alignment
// peeling for a non-direct-access expression): below 4 terms left to right;
// otherwise two packet accumulators over blocks of 8, folded, plus one more
// packet when a block of 4 remains, reduced as (p0 + p2) + (p1 + p3), then
// the tail left to right.  VectorXf::dot and VectorXf::norm use it.
template <typename F>
__device__ inline float eigen_dyn_sum(F c, int n) {
    if (n < 4) {
        float r = c(0);
        for (int i = 1; i < n; ++i) r = r + c(i);
        return r;
    }
    const int aligned = n & ~3, aligned2 = n & ~7;
    float p0 = c(0), p1 = c(1), p2 = c(2), p3 = c(3);
    if (aligned > 4) {
        float s0 = c(4), s1 = c(5), s2 = c(6), s3 = c(7);
        for (int idx = 8; idx < aligned2; idx += 8) {
            p0 = p0 + c(idx);
            p1 = p1 + c(idx + 1);
            p2 = p2 + c(idx + 2);
            p3 = p3 + c(idx + 3);
            s0 = s0 + c(idx + 4);
            s1 = s1 + c(idx + 5);
            s2 = s2 + c(idx + 6);
            s3 = s3 + c(idx + 7);
        }
        p0 = p0 + s0;
        p1 = p1 + s1;
        p2 = p2 + s2;
        p3 = p3 + s3;
        if (aligned > aligned2) {
            p0 = p0 + c(aligned2);
            p1 = p1 + c(aligned2 + 1);
            p2 = p2 + c(aligned2 + 2);
            p3 = p3 + c(aligned2 + 3);
        }
    }
    float r = (p0 + p2) + (p1 + p3);
    for (int i = aligned; i < n; ++i) r = r + c(i);
    return r;
}

}  // namespace rvs
