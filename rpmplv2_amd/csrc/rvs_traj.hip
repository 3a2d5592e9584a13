// Trajectory safety kernel (gfx950): AbstractTrajectory::isSafe
// (src/planners/trajectory/AbstractTrajectory.cpp:212-256) for a batch of
// candidate trajectories that leave one current state, one lane per
// trajectory.
//
// Along the trajectory points the lane keeps a reference state q_temp with a
// distance profile (first the current state's, from computeDistance).  A
// point is accepted while, for every skeleton point j, the distance it can
// cover from q_temp (enclosing radii R.col(j) . |q - q_temp|) plus the
// distance the obstacles can cover since q_temp's profile was taken
// (max_obs_vel * elapsed time) stays below the link's clearance.  Otherwise
// the point becomes the new q_temp and gets a fresh profile from the
// moving-plane underestimation (planes through the current state's nearest
// points, moved towards the robot for the elapsed time at each obstacle's
// maximal speed); a non-positive result makes the trajectory unsafe.
// Optionally every point is also tested for self-collision (capsule pairs).
//
// No FMA contraction anywhere in this translation unit, headers included:
// the float arithmetic must keep the reference's rounding.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "rvs_fk.hpp"
#include "rvs_geometry.hpp"
#include "rvs_internal.hpp"
#include "rvs_radii.hpp"

namespace rvs {

namespace {

// Floats of one lane's LDS slice: skeleton 3 (D+1), radii column (D),
// reference state (D), its profile (L = D).
__host__ __device__ inline int traj_slice(int D) { return 3 * (D + 1) + 3 * D; }

__global__ __launch_bounds__(64) void k_traj_safe(RobotDev rb, QueryParams qp, TrajArgs a, int64_t T) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int D = rb.num_dofs, L = rb.num_links;
    const int B = blockDim.x, t = threadIdx.x;
    float* sk = reinterpret_cast<float*>(smem);        // [3 (D+1)][B]
    float* ac = sk + (size_t)3 * (D + 1) * B;          // [D][B]
    float* qt = ac + (size_t)D * B;                    // [D][B]
    float* pr = qt + (size_t)D * B;                    // [L][B]
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= T) return;
    const int M = a.num_obj;
    const Params<float> p{qp.eq, qp.maxd};
    auto Q = [&](int k) -> float& { return qt[k * B + t]; };
    auto PR = [&](int k) -> float& { return pr[k * B + t]; };
    auto acc = [&](int i) -> float& { return ac[i * B + t]; };
    auto P = [&](int k) { return Vec3<float>{sk[(3 * k) * B + t], sk[(3 * k + 1) * B + t], sk[(3 * k + 2) * B + t]}; };
    auto skeleton_of = [&](const float* qs, int stride) {
        FKStream fk;
        fk_init(fk);
        for (int k = 0; k <= D; ++k) {
            const Vec3<float> v = skel_point<float>(fk, rb, qs, k, stride);
            sk[(3 * k) * B + t] = v.x;
            sk[(3 * k + 1) * B + t] = v.y;
            sk[(3 * k + 2) * B + t] = v.z;
        }
    };
    for (int k = 0; k < D; ++k) Q(k) = a.q_current[k];
    for (int k = 0; k < L; ++k) PR(k) = a.profile_current[k];
    // the trajectory's max_obs_vel: largest obstacle speed, from 0
    // (AbstractTrajectory.cpp:44-49)
    float max_obs_vel = 0.0f;
    for (int o = 0; o < M; ++o)
        if (a.max_vel[o] > max_obs_vel) max_obs_vel = a.max_vel[o];
    float t_iter = a.t_iter, t_init = 0.0f;
    uint8_t safe = 1;
    const float* pts = a.points + c * (int64_t)a.num_points * D;
    for (int i = 0; i < a.num_points && safe; ++i) {
        const float* qf = pts + (int64_t)i * D;
        if (a.self_collision && rb.num_pairs > 0) {
            // robot->checkSelfCollision(q_final): capsule pairs
            skeleton_of(qf, 1);
            for (int m = 0; m < rb.num_pairs; ++m) {
                const int u = rb.pairs[2 * m], v = rb.pairs[2 * m + 1];
                const float dd = seg_seg<float, false>(P(u), P(u + 1), P(v), P(v + 1), p, nullptr) - rb.radius[u] -
                                 rb.radius[v];
                if (dd < 0.0f) safe = 0;
            }
            if (!safe) break;
        }
        const float rho_obs = max_obs_vel * (t_iter - t_init);
        skeleton_of(qt + t, B);
        for (int k = 0; k < D; ++k) acc(k) = 0.0f;
        for (int j = 1; j <= D; ++j) {
            radii_column(rb, P, j, acc);
            const float rho_robot = eigen_dyn_sum([&](int k) { return acc(k) * fabsf(qf[k] - Q(k)); }, D);
            if (rho_robot + rho_obs >= PR(j - 1)) {
                // possible collision: q_final becomes the reference state,
                // with the moving-plane underestimation as its profile
                for (int k = 0; k < D; ++k) Q(k) = qf[k];
                skeleton_of(qf, 1);
                float d = qp.maxd;
                for (int l = 0; l < L && safe; ++l) {
                    float prof = qp.maxd;
                    for (int o = 0; o < M; ++o) {
                        const float* e = a.nearest_current + ((int64_t)o * L + l) * 6;
                        const Vec3<float> O{e[3], e[4], e[5]};
                        if (vnorm(O) == qp.maxd) continue;
                        const Vec3<float> R{e[0], e[1], e[2]};
                        const float tm = plane_term_moving(P(l), P(l + 1), R, O, rb.radius[l], a.max_vel[o], t_iter);
                        if (tm < 0.0f) {
                            safe = 0;   // returned early: the state keeps no distance
                            break;
                        }
                        prof = min_std(prof, tm);
                    }
                    PR(l) = prof;
                    d = min_std(d, prof);
                }
                t_init = t_iter;
                if (safe && d <= 0.0f) safe = 0;
                break;
            }
        }
        t_iter += a.time_step;
    }
    a.safe[c] = safe;
}

}  // namespace

size_t traj_lds_per_lane(const RobotDev& rb) { return (size_t)traj_slice(rb.num_dofs) * sizeof(float); }

hipError_t launch_traj_safe(const RobotDev& rb, const QueryParams& qp, const TrajArgs& a, int64_t T, hipStream_t st) {
    if (T <= 0) return hipSuccess;
    int B = 64;
    while (B > 16 && traj_lds_per_lane(rb) * B > 64 * 1024) B /= 2;
    const size_t sm = traj_lds_per_lane(rb) * B;
    hipLaunchKernelGGL(k_traj_safe, dim3((unsigned)((T + B - 1) / B)), dim3(B), sm, st, rb, qp, a, T);
    return hipGetLastError();
}

}  // namespace rvs
