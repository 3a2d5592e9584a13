// Trajectory safety kernel (gfx950): AbstractTrajectory::isSafe
// (src/planners/trajectory/AbstractTrajectory.cpp:212-256) for a batch of
// candidate trajectories that leave one current state, one wave per
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

// LDS of one workgroup (one wave, one trajectory), in floats: the reference
// state (D), its profile (L = D), its skeleton 3 (D+1), its radii
// D (D+1), the running radii column (D), and a skeleton slice per lane for
// the self-collision test 64 x 3 (D+1).
__host__ __device__ inline int traj_lds_floats(int D) { return D + D + 3 * (D + 1) + D * (D + 1) + D + 64 * 3 * (D + 1); }

// One wave per trajectory.  The bubble test of a point only reads the
// reference state, so the wave tests 64 points at once and re-anchors at
// the first point that fails; the moving-plane underestimation of the new
// reference state spreads the obstacles over the lanes.  The outcome is the
// reference's sequential walk: a point is tested against the reference
// state the walk would hold there, a minimum of floats is exact in any
// order, and any negative plane term ends the walk as unsafe whichever comes
// first.  Self-collision of any point makes the trajectory unsafe, as the
// walk would stop there or earlier.
__global__ __launch_bounds__(64) void k_traj_safe(RobotDev rb, QueryParams qp, TrajArgs a, int64_t T) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int D = rb.num_dofs, L = rb.num_links;
    const int t = threadIdx.x;
    const int64_t c = blockIdx.x;
    if (c >= T) return;   // uniform across the workgroup
    float* qa = reinterpret_cast<float*>(smem);          // [D]
    float* pr = qa + D;                                   // [L]
    float* ska = pr + L;                                  // [3 (D+1)]
    float* Rm = ska + 3 * (D + 1);                        // [D][D+1]
    float* ac = Rm + D * (D + 1);                         // [D]
    float* skl = ac + D;                                  // [3 (D+1)][64]
    const int M = a.num_obj, K = a.num_points;
    const Params<float> p{qp.eq, qp.maxd};
    const float* pts = a.points + c * (int64_t)K * D;
    auto PA = [&](int k) { return Vec3<float>{ska[3 * k], ska[3 * k + 1], ska[3 * k + 2]}; };
    auto PL = [&](int k) { return Vec3<float>{skl[(3 * k) * 64 + t], skl[(3 * k + 1) * 64 + t], skl[(3 * k + 2) * 64 + t]}; };
    auto anchor_skeleton = [&](const float* qs) {      // lane 0 only
        FKStream fk;
        fk_init(fk);
        for (int k = 0; k <= D; ++k) {
            const Vec3<float> v = skel_point<float>(fk, rb, qs, k);
            ska[3 * k] = v.x;
            ska[3 * k + 1] = v.y;
            ska[3 * k + 2] = v.z;
        }
    };
    // t_iter at point i: t_iter + time_step added i times, in order
    auto time_at = [&](int i) {
        float ti = a.t_iter;
        for (int k = 0; k < i; ++k) ti += a.time_step;
        return ti;
    };
    float max_obs_vel = 0.0f;   // AbstractTrajectory.cpp:44-49
    for (int o = 0; o < M; ++o)
        if (a.max_vel[o] > max_obs_vel) max_obs_vel = a.max_vel[o];

    bool unsafe = false;
    // robot->checkSelfCollision(q_final) at every point: capsule pairs
    if (a.self_collision && rb.num_pairs > 0) {
        bool hit = false;
        for (int i = t; i < K; i += 64) {
            FKStream fk;
            fk_init(fk);
            for (int k = 0; k <= D; ++k) {
                const Vec3<float> v = skel_point<float>(fk, rb, pts + (int64_t)i * D, k);
                skl[(3 * k) * 64 + t] = v.x;
                skl[(3 * k + 1) * 64 + t] = v.y;
                skl[(3 * k + 2) * 64 + t] = v.z;
            }
            for (int m = 0; m < rb.num_pairs; ++m) {
                const int u = rb.pairs[2 * m], w = rb.pairs[2 * m + 1];
                const float dd = seg_seg<float, false>(PL(u), PL(u + 1), PL(w), PL(w + 1), p, nullptr) -
                                 rb.radius[u] - rb.radius[w];
                if (dd < 0.0f) hit = true;
            }
        }
        unsafe = __any(hit);
    }
    if (t == 0) {
        for (int k = 0; k < D; ++k) qa[k] = a.q_current[k];
        for (int k = 0; k < L; ++k) pr[k] = a.profile_current[k];
    }
    float t_init = 0.0f;
    int s = 0;
    while (!unsafe) {
        // radii of the reference state (lane 0), then the points in chunks of 64
        if (t == 0) {
            anchor_skeleton(qa);
            for (int i = 0; i < D; ++i) ac[i] = 0.0f;
            for (int i = 0; i < D; ++i) Rm[i * (D + 1)] = 0.0f;
            auto acc = [&](int i) -> float& { return ac[i]; };
            for (int j = 1; j <= D; ++j) {
                radii_column(rb, PA, j, acc);
                for (int i = 0; i < D; ++i) Rm[i * (D + 1) + j] = ac[i];
            }
        }
        __syncthreads();
        int f = K;
        float tb = time_at(s);   // time at the chunk's first point
        for (int b = s; b < K; b += 64) {
            const int i = b + t;
            bool fail = false;
            float ti = tb;
            for (int k = 0; k < t; ++k) ti += a.time_step;
            for (int k = 0; k < 64; ++k) tb += a.time_step;
            if (i < K) {
                const float* qf = pts + (int64_t)i * D;
                const float rho_obs = max_obs_vel * (ti - t_init);
                for (int j = 0; j < D && !fail; ++j) {
                    const float rho_robot =
                        eigen_dyn_sum([&](int k) { return Rm[k * (D + 1) + j + 1] * fabsf(qf[k] - qa[k]); }, D);
                    if (rho_robot + rho_obs >= pr[j]) fail = true;
                }
            }
            const uint64_t m = __ballot(fail);
            if (m) {
                f = b + __ffsll((unsigned long long)m) - 1;
                break;
            }
        }
        if (f >= K) break;   // every remaining point inside the bubble
        // possible collision at point f: it becomes the reference state
        const float tf = time_at(f);
        const float* qf = pts + (int64_t)f * D;
        __syncthreads();
        if (t == 0) {
            for (int k = 0; k < D; ++k) qa[k] = qf[k];
            anchor_skeleton(qf);
        }
        __syncthreads();
        bool neg = false;
        float d = qp.maxd;
        for (int l = 0; l < L; ++l) {
            float prof = qp.maxd;
            for (int o = t; o < M; o += 64) {
                const float* e = a.nearest_current + ((int64_t)o * L + l) * 6;
                const Vec3<float> O{e[3], e[4], e[5]};
                if (vnorm(O) == qp.maxd) continue;
                const Vec3<float> R{e[0], e[1], e[2]};
                const float tm = plane_term_moving(PA(l), PA(l + 1), R, O, rb.radius[l], a.max_vel[o], tf);
                if (tm < 0.0f) neg = true;
                prof = min_std(prof, tm);
            }
            for (int o = 32; o > 0; o >>= 1) prof = min_std(prof, __shfl_xor(prof, o, 64));
            if (t == 0) pr[l] = prof;
            d = min_std(d, prof);
        }
        // a negative plane term returns 0 without a distance (the state keeps
        // d_c = -1); otherwise the walk stops when the distance is <= 0
        if (__any(neg) || d <= 0.0f) unsafe = true;
        t_init = tf;
        s = f + 1;
        __syncthreads();
    }
    if (t == 0) a.safe[c] = unsafe ? 0 : 1;
}

}  // namespace

size_t traj_lds_bytes(const RobotDev& rb) { return (size_t)traj_lds_floats(rb.num_dofs) * sizeof(float); }

hipError_t launch_traj_safe(const RobotDev& rb, const QueryParams& qp, const TrajArgs& a, int64_t T, hipStream_t st) {
    if (T <= 0) return hipSuccess;
    const size_t sm = traj_lds_bytes(rb);
    hipLaunchKernelGGL(k_traj_safe, dim3((unsigned)T), dim3(64), sm, st, rb, qp, a, T);
    return hipGetLastError();
}

}  // namespace rvs
