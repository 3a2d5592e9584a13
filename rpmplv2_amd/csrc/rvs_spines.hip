// Spine generation kernels (gfx950): the bubble extension of RBT planners.
//
// k_spines: RBTConnect::extendSpine (src/planners/rbt/RBTConnect.cpp:85-150)
// for a batch of spines, one lane per spine.  Each iteration takes the
// enclosing radii of the current state column by column (radii_column), the
// covered-distance bound of every skeleton point, R.col(j).dot(|q_e - q|), in
// Eigen's dynamic-size order, and the largest step whose bound stays inside
// the link's remaining clearance (expanded bubble) or inside the remaining
// distance (plain bubble).  The moved distance of every skeleton point then
// shrinks the clearances for the next iteration.  Skeleton, radii column,
// state and clearances of a lane live in its LDS slice.
//
// k_gen_spines_under: RGBTConnect::extendGenSpine (src/planners/rbt/
// RGBTConnect.cpp:71-90) with underestimation layers, all layers of a spine
// in one lane.  k_gen_layer: the bookkeeping between layers when each
// layer's distance is a separate launch (real distances).
//
// k_edge_self: the capsule part of xArm6::checkSelfCollision(q1, q2)
// (src/robots/xArm6.cpp:284-368) for a batch of edges, one lane per edge:
// radii-bounded steps along the edge while the four link pairs provably stay
// apart, then a walk in EPS_STEP steps that tests the pairs not yet proven
// apart.  A pair collides when its capsules overlap; the FCL mesh
// confirmation of the reference (checkRealSelfCollision, :426-483) is not
// part of this path.
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

constexpr int32_t kAdvanced = 1, kTrapped = 2, kReached = 3;

// Floats of one lane's LDS slice: skeleton of the start state and of the
// current state (3 (D+1) each), radii column, current state, clearance
// losses (D each).
__host__ __device__ inline int spine_slice(int D) { return 6 * (D + 1) + 3 * D; }

// One lane's view of its LDS slice.
struct SpineLane {
    float* sk0;
    float* sk;
    float* ac;
    float* qt;
    float* rp;
    int B, t;
    __device__ SpineLane(float* base, int D, int B_, int t_) : B(B_), t(t_) {
        sk0 = base;
        sk = sk0 + (size_t)3 * (D + 1) * B;
        ac = sk + (size_t)3 * (D + 1) * B;
        qt = ac + (size_t)D * B;
        rp = qt + (size_t)D * B;
    }
    __device__ float& Q(int k) const { return qt[k * B + t]; }
    __device__ float& RP(int k) const { return rp[k * B + t]; }
    __device__ float& acc(int i) const { return ac[i * B + t]; }
    __device__ Vec3<float> P0(int k) const {
        return Vec3<float>{sk0[(3 * k) * B + t], sk0[(3 * k + 1) * B + t], sk0[(3 * k + 2) * B + t]};
    }
    __device__ Vec3<float> P(int k) const {
        return Vec3<float>{sk[(3 * k) * B + t], sk[(3 * k + 1) * B + t], sk[(3 * k + 2) * B + t]};
    }
    __device__ void set_P(float* s, int k, const Vec3<float>& v) const {
        s[(3 * k) * B + t] = v.x;
        s[(3 * k + 1) * B + t] = v.y;
        s[(3 * k + 2) * B + t] = v.z;
    }
};

// extendSpine from the state in the lane's Q (distance dq, link profile
// pc(i)) towards qe; Q holds q_new afterwards.  Returns the status.
template <typename PC>
__device__ int32_t spine_run(const RobotDev& rb, const SpineLane& s, float dq, PC pc, const float* qe,
                             int num_iter, int expanded) {
    const int D = rb.num_dofs;
    if (dq <= 0.0f) return kTrapped;
    {
        FKStream fk;
        fk_init(fk);
        for (int k = 0; k <= D; ++k) {
            const Vec3<float> v = skel_point<float>(fk, rb, s.qt + s.t, k, s.B);
            s.set_P(s.sk0, k, v);
            s.set_P(s.sk, k, v);
        }
    }
    for (int k = 0; k < D; ++k) s.RP(k) = 0.0f;
    float rho = 0.0f;
    int32_t status = kAdvanced;
    auto P = [&](int k) { return s.P(k); };
    auto acc = [&](int i) -> float& { return s.acc(i); };
    for (int it = 1;; ++it) {
        for (int i = 0; i < D; ++i) s.acc(i) = 0.0f;
        // covered distance of skeleton point j: R.col(j).dot(|q_e - q_temp|)
        auto covered = [&]() {
            return eigen_dyn_sum([&](int i) { return s.acc(i) * fabsf(qe[i] - s.Q(i)); }, D);
        };
        float step = INFINITY;
        for (int j = 1; j <= D; ++j) {
            radii_column(rb, P, j, acc);
            if (expanded) step = min_std(step, (pc(j - 1) - s.RP(j - 1)) / covered());
        }
        if (!expanded) step = (dq - rho) / covered();
        if (step > 1.0f) {
            for (int k = 0; k < D; ++k) s.Q(k) = qe[k];
            status = kReached;
        } else {
            for (int k = 0; k < D; ++k) s.Q(k) = s.Q(k) + step * (qe[k] - s.Q(k));
        }
        if (it >= num_iter || status != kAdvanced) break;
        // moved distance of every skeleton point since the start state
        FKStream fk;
        fk_init(fk);
        s.set_P(s.sk, 0, skel_point<float>(fk, rb, s.qt + s.t, 0, s.B));
        float prev = 0.0f;
        for (int k = 1; k <= D; ++k) {
            const Vec3<float> v = skel_point<float>(fk, rb, s.qt + s.t, k, s.B);
            s.set_P(s.sk, k, v);
            const float rk = norm_dyn3(s.P0(k) - v);
            s.RP(k - 1) = max_std(prev, rk);
            prev = rk;
            rho = max_std(rho, rk);
        }
    }
    return status;
}

__global__ __launch_bounds__(64) void k_spines(RobotDev rb, SpineArgs a, int64_t n) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int D = rb.num_dofs;
    const SpineLane s(reinterpret_cast<float*>(smem), D, blockDim.x, threadIdx.x);
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    if (a.active && !a.active[c]) return;
    const float* qc = a.q + c * a.q_stride;
    const float* pc = a.profile + c * a.profile_stride;
    // q_new may alias q (in-place layers): read the whole state first
    for (int k = 0; k < D; ++k) s.Q(k) = qc[k];
    const int32_t status =
        spine_run(rb, s, a.d[c * a.d_stride], [&](int i) { return pc[i]; }, a.q_e + c * D, a.num_iter, a.expanded);
    float* qn = a.q_new + c * D;
    for (int k = 0; k < D; ++k) qn[k] = s.Q(k);
    a.status[c] = status;
}

// All layers of generalized spines with underestimation distances in one
// lane (RGBTConnect::extendGenSpine with use_real_dist = false): spine,
// computeDistanceUnderestimation of its end state against the root's
// nearest points, stop below D_CRIT or when the spine did not advance.  The
// profile of the current layer's start state sits in the lane's slice after
// the spine part.
__global__ __launch_bounds__(64) void k_gen_spines_under(RobotDev rb, GenSpineArgs g, int64_t n) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int D = rb.num_dofs, L = rb.num_links;
    const int B = blockDim.x, t = threadIdx.x;
    float* base = reinterpret_cast<float*>(smem);
    const SpineLane s(base, D, B, t);
    float* pr = base + (size_t)spine_slice(D) * B;    // [L][B]
    auto PR = [&](int k) -> float& { return pr[k * B + t]; };
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const float* qe = g.q_e + c * D;
    const float d0 = g.d_root[0];
    for (int k = 0; k < D; ++k) s.Q(k) = g.q_root[k];
    for (int k = 0; k < L; ++k) PR(k) = g.profile_root[k];
    int32_t status = kTrapped;
    float d_out = d0;
    float dq = d0;
    if (d0 > 0.0f) {
        for (int layer = 0; layer < g.num_layers; ++layer) {
            status = spine_run(rb, s, dq, [&](int i) { return PR(i); }, qe, g.num_iter, g.expanded);
            // computeDistanceUnderestimation(q_new, root nearest points)
            FKStream fk;
            fk_init(fk);
            Vec3<float> P0 = skel_point<float>(fk, rb, s.qt + t, 0, B);
            float d = g.maxd;
            bool neg = false;
            for (int i = 0; i < L && !neg; ++i) {
                const Vec3<float> P1 = skel_point<float>(fk, rb, s.qt + t, i + 1, B);
                float prof = g.maxd;
                for (int j = 0; j < g.num_obj; ++j) {
                    const float* e = g.np_root + ((int64_t)j * L + i) * 6;
                    const Vec3<float> O{e[3], e[4], e[5]};
                    if (vnorm(O) == g.maxd) continue;
                    const Vec3<float> R{e[0], e[1], e[2]};
                    const float tm = plane_term(P0, P1, R, O, rb.radius[i]);
                    if (tm < 0.0f) {
                        neg = true;
                        break;
                    }
                    prof = min_std(prof, tm);
                }
                PR(i) = prof;
                d = min_std(d, prof);
                P0 = P1;
            }
            if (neg) {   // returned early: no profile stored
                d = 0.0f;
                for (int k = 0; k < L; ++k) PR(k) = 0.0f;
            }
            d_out = d;
            dq = d;
            if (d < g.d_crit || status != kAdvanced) break;
        }
    }
    float* qn = g.q_new + c * D;
    for (int k = 0; k < D; ++k) qn[k] = s.Q(k);
    g.status[c] = status;
    g.d_out[c] = d_out;
}

// After the distance of layer 'layer': the reference leaves the loop when
// d_c < D_CRIT or the spine did not advance; at layer 0 a start state in
// contact (d <= 0) stops with its own distance.
__global__ void k_gen_layer(int64_t n, int layer, const float* d_root, const float* d_layer, const int32_t* status,
                            uint8_t* active, float* d_out, float d_crit) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n || !active[c]) return;
    if (layer == 0 && d_root[0] <= 0.0f) {
        d_out[c] = d_root[0];
        active[c] = 0;
        return;
    }
    const float d = d_layer[c];
    d_out[c] = d;
    if (d < d_crit || status[c] != kAdvanced) active[c] = 0;
}

__global__ void k_fill_u8(uint8_t* p, int64_t n, uint8_t v) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) p[c] = v;
}

// Floats of one lane's LDS slice for k_edge_self: skeleton 3 (D+1), radii
// column, current state and next state (D each).
__host__ __device__ inline int edge_self_slice(int D) { return 3 * (D + 1) + 3 * D; }

__global__ __launch_bounds__(64) void k_edge_self(RobotDev rb, QueryParams qp, const float* __restrict__ q1,
                                                  const float* __restrict__ q2, int64_t n, float eps_step,
                                                  uint8_t* collide, float* q_out) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int D = rb.num_dofs;
    const int B = blockDim.x, t = threadIdx.x;
    float* sk = reinterpret_cast<float*>(smem);             // [3 (D+1)][B]
    float* ac = sk + (size_t)3 * (D + 1) * B;               // [D][B]
    float* qt = ac + (size_t)D * B;                         // [D][B]
    float* qn = qt + (size_t)D * B;                         // [D][B]
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const float* a = q1 + c * D;
    const float* b = q2 + c * D;
    float* out = q_out + c * D;
    const Params<float> p{qp.eq, qp.maxd};
    auto Q = [&](int k) -> float& { return qt[k * B + t]; };
    auto QN = [&](int k) -> float& { return qn[k * B + t]; };
    auto acc = [&](int i) -> float& { return ac[i * B + t]; };
    auto P = [&](int k) { return Vec3<float>{sk[(3 * k) * B + t], sk[(3 * k + 1) * B + t], sk[(3 * k + 2) * B + t]}; };
    auto skeleton_of = [&](float* qs) {
        FKStream fk;
        fk_init(fk);
        for (int k = 0; k <= D; ++k) {
            const Vec3<float> v = skel_point<float>(fk, rb, qs + t, k, B);
            sk[(3 * k) * B + t] = v.x;
            sk[(3 * k + 1) * B + t] = v.y;
            sk[(3 * k + 2) * B + t] = v.z;
        }
    };
    // computeCapsulesDistance (xArm6.cpp:417-424) on the skeleton in LDS
    auto caps = [&](int i, int j) {
        return seg_seg<float, false>(P(i), P(i + 1), P(j), P(j + 1), p, nullptr) - rb.radius[i] - rb.radius[j];
    };
    auto finish = [&](uint8_t hit, bool from_state) {
        for (int k = 0; k < D; ++k) out[k] = from_state ? Q(k) : b[k];
        collide[c] = hit;
    };
    for (int k = 0; k < D; ++k) Q(k) = a[k];
    // radii-bounded steps (:300-324)
    float s04 = 0.0f, s05 = 0.0f, s14 = 0.0f, s15 = 0.0f;
    for (int iter = 1;; ++iter) {
        skeleton_of(qt);
        for (int i = 0; i < D; ++i) acc(i) = 0.0f;
        float phi2 = 0.0f, phi4 = 0.0f, phi5 = 0.0f, phi6 = 0.0f;
        for (int j = 1; j <= D; ++j) {
            radii_column(rb, P, j, acc);
            if (j == 2 || j >= 4) {
                const float v = eigen_dyn_sum([&](int i) { return acc(i) * fabsf(b[i] - Q(i)); }, D);
                if (j == 2) phi2 = v;
                else if (j == 4) phi4 = v;
                else if (j == 5) phi5 = v;
                else phi6 = v;
            }
        }
        s04 = caps(0, 4) / max_std(phi4, phi5);
        s05 = caps(0, 5) / max_std(phi5, phi6);
        s14 = caps(1, 4) / (phi2 + max_std(phi4, phi5));
        s15 = caps(1, 5) / (phi2 + max_std(phi5, phi6));
        const float step = min_std(min_std(s04, s05), min_std(s14, s15));
        if (step > 1.0f) {
            finish(0, false);
            return;
        }
        if (iter == 5 || step <= 0.0f) break;
        for (int k = 0; k < D; ++k) Q(k) = Q(k) + step * (b[k] - Q(k));
    }
    // walk in EPS_STEP steps from the last bounded state (:326-366); a pair
    // whose bounded step exceeds the walked fraction is proven apart.  The
    // step count is capped for non-finite or huge coordinates (the walk
    // would not end); such an edge is reported as 255.
    auto norm_to_b = [&](auto qv) {
        return sqrtf(eigen_dyn_sum([&](int i) { const float e = b[i] - qv(i); return e * e; }, D));
    };
    const float Dn = norm_to_b(Q);
    float dist = Dn;
    const float span = Dn / eps_step;
    const int cap = (span < 65536.0f) ? (int)span + 16 : 65552;
    for (int walked = 0;; ++walked) {
        if (walked > cap) {
            finish(255, false);
            return;
        }
        bool reached = false;
        if (dist > eps_step) {
            for (int k = 0; k < D; ++k) QN(k) = Q(k) + (b[k] - Q(k)) / dist * eps_step;
        } else {
            for (int k = 0; k < D; ++k) QN(k) = b[k];
            reached = true;
        }
        dist = norm_to_b(QN);
        const float crit = 1.0f - dist / Dn;
        skeleton_of(qn);
        // xArm6::checkSelfCollision(q, skip) pair order (:384-411)
        const bool hit = (!(s05 > crit) && caps(0, 5) < 0.0f) || (!(s15 > crit) && caps(1, 5) < 0.0f) ||
                         (!(s04 > crit) && caps(0, 4) < 0.0f) || (!(s14 > crit) && caps(1, 4) < 0.0f);
        if (hit) {
            finish(1, true);
            return;
        }
        for (int k = 0; k < D; ++k) Q(k) = QN(k);
        if (reached) break;
    }
    finish(0, false);
}

}  // namespace

size_t spine_lds_per_lane(const RobotDev& rb) { return (size_t)spine_slice(rb.num_dofs) * sizeof(float); }

int spine_block(const RobotDev& rb) {
    const size_t per = spine_lds_per_lane(rb);
    int B = 64;
    while (B > 16 && per * B > 64 * 1024) B /= 2;
    return B;
}

hipError_t launch_spines(const RobotDev& rb, const SpineArgs& a, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int B = spine_block(rb);
    const size_t sm = spine_lds_per_lane(rb) * B;
    hipLaunchKernelGGL(k_spines, dim3((unsigned)((n + B - 1) / B)), dim3(B), sm, st, rb, a, n);
    return hipGetLastError();
}

hipError_t launch_gen_spines_under(const RobotDev& rb, const GenSpineArgs& g, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const size_t per = spine_lds_per_lane(rb) + (size_t)rb.num_links * sizeof(float);
    int B = 64;
    while (B > 16 && per * B > 64 * 1024) B /= 2;
    hipLaunchKernelGGL(k_gen_spines_under, dim3((unsigned)((n + B - 1) / B)), dim3(B), per * B, st, rb, g, n);
    return hipGetLastError();
}

hipError_t launch_gen_layer(int64_t n, int layer, const float* d_root, const float* d_layer, const int32_t* status,
                            uint8_t* active, float* d_out, float d_crit, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gen_layer, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, layer, d_root, d_layer,
                       status, active, d_out, d_crit);
    return hipGetLastError();
}

size_t edge_self_lds_per_lane(const RobotDev& rb) { return (size_t)edge_self_slice(rb.num_dofs) * sizeof(float); }

hipError_t launch_edge_self(const RobotDev& rb, const QueryParams& qp, const float* q1, const float* q2, int64_t n,
                            float eps_step, uint8_t* collide, float* q_out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int B = 64;
    const size_t sm = edge_self_lds_per_lane(rb) * B;
    hipLaunchKernelGGL(k_edge_self, dim3((unsigned)((n + B - 1) / B)), dim3(B), sm, st, rb, qp, q1, q2, n, eps_step,
                       collide, q_out);
    return hipGetLastError();
}

hipError_t launch_fill_u8(uint8_t* p, int64_t n, uint8_t v, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u8, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

}  // namespace rvs
