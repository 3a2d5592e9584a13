// Edge checking kernels (gfx950).
//
// k_edge_valid_q: RealVectorSpace::isValid(q1, q2) (RealVectorSpace.cpp:
// 261-275) for a batch of edges.  The reference walks k = K..1 and stops at
// the first invalid state; the edge is valid exactly when all K states are,
// so here each lane takes one (edge, k) state and a collision clears the
// edge's flag (plain stores of 0; the flags start at 1).
//
// k_edge_first_q: discretised edges of several robots in ONE persistent launch.
// The states of each robot's edges are laid out round-major (state 0 of every
// edge, then state 1 of every edge that has one, ...) after sorting the edges
// by length, so consecutive slots share a robot and a state index.  Items of
// 64 x spl slots of one robot are handed out from a global counter in the
// order of their first slot's state index across the robots, so that early
// states of every robot come first; each wave takes items until the counter
// passes the last one, skips states that lie beyond an already found
// collision of their edge, and records a collision with atomicMin on the
// edge's first index.  The robot is uniform across a wave.
//
// On axis-aligned scenes (k_edge_first_q<false>) each link meets the obstacles
// through an axis-test table first (SlabEntry): collisionCapsuleToBox
// (CollisionAndDistance.cpp:4-42) rejects a capsule whose two end points lie
// beyond one face of the box grown by the radius r on the same side (its
// second test), and its first test, an end point inside the box grown by
// r_new = r sqrt(3) / 3 <= r, can only hold when that one cannot: a point
// below fl(lo - r) is below fl(lo - r_new) as rounding is monotone.  So a
// pair that the axis test separates is a pair the full test rejects, and
// testing it first changes no verdict.  The axis test of a group of obstacles
// uses the least lo and the greatest hi of its members, which separates every
// member it separates.  The tables hold fl(lo - r) and fl(hi + r) as the
// full test rounds them, filled on the device from the scene at launch.
//
// The self pairs get a bound first too.  distanceLineSegToLineSeg
// (CollisionAndDistance.cpp:217-320) returns 0 (below the equality threshold
// eq), max_distance (degenerate input), or the float norm of p1 - p0 for
// points that are end points or a + o (b - a) with o in [0, 1], each within
// 5 eps C of its segment (eps = 2^-24, C the largest coordinate magnitude), so
// the returned distance is at least (D - 21 eps C)(1 - 2 eps) with D the
// distance of the two segments' bounding boxes.  A pair whose D, computed in
// float, exceeds R = max(r_a + r_b, eq) by 4e-6 (C + R) (about 67 eps (C + R),
// which also covers the rounding of D itself) therefore gets a distance of
// at least r_a + r_b and is no collision, provided max_distance >= R; the
// full test runs for every other pair, and for every pair of a state with a
// NaN coordinate.
// No FMA contraction anywhere in this translation unit, headers included:
// the float kernels must keep the reference's rounding.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>

#include "rvs_fk.hpp"
#include "rvs_geometry.hpp"
#include "rvs_internal.hpp"
#include "rvs_stage.hpp"


namespace rvs {

namespace {

// Norm of a dynamic-size float vector in the order Eigen's SSE reduction uses
// for VectorXf (RealVectorSpace::getNorm, RealVectorSpace.cpp:49-52): packets
// of four accumulated pairwise, reduced as (p0 + p2) + (p1 + p3), then the
// remaining coefficients left to right; sizes below four go left to right.
__device__ inline float vecxf_norm_diff(const float* a, const float* b, int n) {
    auto sq = [&](int i) { const float d = a[i] - b[i]; return d * d; };
    if (n < 4) {
        float r = sq(0);
        for (int i = 1; i < n; ++i) r = r + sq(i);
        return sqrtf(r);
    }
    const int aligned = (n / 4) * 4, aligned2 = (n / 8) * 8;
    float p0 = sq(0), p1 = sq(1), p2 = sq(2), p3 = sq(3);
    if (aligned > 4) {
        float s0 = sq(4), s1 = sq(5), s2 = sq(6), s3 = sq(7);
        for (int idx = 8; idx < aligned2; idx += 8) {
            p0 = p0 + sq(idx); p1 = p1 + sq(idx + 1); p2 = p2 + sq(idx + 2); p3 = p3 + sq(idx + 3);
            s0 = s0 + sq(idx + 4); s1 = s1 + sq(idx + 5); s2 = s2 + sq(idx + 6); s3 = s3 + sq(idx + 7);
        }
        p0 = p0 + s0; p1 = p1 + s1; p2 = p2 + s2; p3 = p3 + s3;
        if (aligned > aligned2) {
            p0 = p0 + sq(aligned2); p1 = p1 + sq(aligned2 + 1); p2 = p2 + sq(aligned2 + 2); p3 = p3 + sq(aligned2 + 3);
        }
    }
    float r = (p0 + p2) + (p1 + p3);
    for (int i = aligned; i < n; ++i) r = r + sq(i);
    return sqrtf(r);
}

__global__ void k_fill_u8(uint8_t* p, int64_t n, uint8_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------- first collision
__global__ void k_first_init(int32_t* first, int64_t n, uint32_t* counter) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) first[i] = INT32_MAX;
    if (i == 0) *counter = 0;
}

__global__ void k_first_final(int32_t* first, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && first[i] == INT32_MAX) first[i] = -1;
}

// Binary search: largest k with off[k] <= s (off is non-decreasing, off[0] = 0).
__device__ inline int round_of(const int32_t* off, int nrounds, int32_t s) {
    int lo = 0, hi = nrounds - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= s) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Axis-test tables from the scene's current boxes: block (x, y) fills
// groups of table y (keys[2y] = radius, keys[2y + 1] = 1 when the links of
// this key skip ground boxes).
__global__ void k_slab_build(EnvDev env, const float* __restrict__ keys, int nkeys, SlabEntry* slab) {
    const int M = env.num_obj, G = (M + kSlabGroup - 1) / kSlabGroup;
    const int t = blockIdx.y, g = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nkeys || g >= G) return;
    const float r = keys[2 * t];
    const bool skip = keys[2 * t + 1] != 0.0f;
    SlabEntry* tab = slab + (size_t)t * (G + M);
    SlabEntry u{{INFINITY, INFINITY, INFINITY}, {-INFINITY, -INFINITY, -INFINITY}, 0, 0};
    for (int m = 0; m < kSlabGroup; ++m) {
        const int b = g * kSlabGroup + m;
        if (b >= M) break;
        const int j = env.gorder ? env.gorder[b] : b;
        SlabEntry x{{INFINITY, INFINITY, INFINITY}, {-INFINITY, -INFINITY, -INFINITY}, -1, 0};
        if (!(skip && (env.flags[j] & kObjGround))) {
            x.idx = j;
            for (int k = 0; k < 3; ++k) {
                x.lo[k] = env.aabb[6 * j + k] - r;
                x.hi[k] = env.aabb[6 * j + 3 + k] + r;
                u.lo[k] = fminf(u.lo[k], x.lo[k]);
                u.hi[k] = fmaxf(u.hi[k], x.hi[k]);
            }
        }
        tab[G + b] = x;
    }
    tab[g] = u;
}

// Read-only tables seen through the constant address space: their loads at
// uniform addresses become scalar loads even after the LDS barriers of the
// kernel (which the compiler would otherwise take for possible writes).
template <typename T>
using ConstPtr = const __attribute__((address_space(4))) T*;
template <typename T>
__device__ inline ConstPtr<T> as_const(const T* p) {
    return (ConstPtr<T>)p;
}
__device__ inline SlabEntry load_entry(ConstPtr<SlabEntry> p) {
    SlabEntry x;
    for (int k = 0; k < 3; ++k) {
        x.lo[k] = p->lo[k];
        x.hi[k] = p->hi[k];
    }
    x.idx = p->idx;
    x.pad = 0;
    return x;
}

// One link (P0, P1) of radius r against the obstacles of table tab, for the
// lanes with act set: true where some obstacle collides
// (collisionCapsuleToBox).  The table pointer and the loop counts are
// uniform, so the entries, and the bounds of an obstacle that goes to the
// full test, arrive through scalar loads.
__device__ inline bool link_hits(const Vec3<float>& P0, const Vec3<float>& P1, float r, float rn,
                                 ConstPtr<SlabEntry> tab, int G, int M, ConstPtr<float> aabb,
                                 const Params<float>& p, bool act) {
    // per axis, the larger and the smaller end point coordinate: "both below
    // L" is "the larger below L" unless a coordinate is NaN, and then every
    // obstacle goes to the full test
    const float mxx = fmaxf(P0.x, P1.x), mxy = fmaxf(P0.y, P1.y), mxz = fmaxf(P0.z, P1.z);
    const float mnx = fminf(P0.x, P1.x), mny = fminf(P0.y, P1.y), mnz = fminf(P0.z, P1.z);
    const float sum = ((P0.x + P0.y) + (P0.z + P1.x)) + (P1.y + P1.z);
    const bool exact = !(sum == sum);
    auto apart = [&](const SlabEntry& x) {
        return (mxx < x.lo[0]) | (mxy < x.lo[1]) | (mxz < x.lo[2]) | (mnx > x.hi[0]) | (mny > x.hi[1]) |
               (mnz > x.hi[2]);
    };
    bool hit = false;
    for (int g = 0; g < G; ++g) {
        const SlabEntry gx = load_entry(tab + g);
        if (act && !hit && (exact || !apart(gx))) {
            const int b1 = min(M, (g + 1) * kSlabGroup);
            for (int b = g * kSlabGroup; b < b1; ++b) {
                const SlabEntry x = load_entry(tab + G + b);
                if (!hit && x.idx >= 0 && (exact || !apart(x))) {
                    ConstPtr<float> o = aabb + 6 * x.idx;
                    const Box<float> bx{{o[0], o[1], o[2]}, {o[3], o[4], o[5]}};
                    if (collision_box(P0, P1, r, rn, bx, p)) hit = true;
                }
            }
        }
    }
    return hit;
}

// One link (P0, P1) of radius r against every oriented box, for the lanes
// with act set: true where some box collides (collisionCapsuleToBox in the
// box's frame, as k_edge_first's OBB form evaluates it).  The boxes are read
// with wave-uniform scalar loads (centre and rotation rows, then the half
// extents); links below ground_included skip ground boxes.
__device__ inline bool link_hits_obb(const Vec3<float>& P0, const Vec3<float>& P1, float r, float rn,
                                     ConstPtr<float> obb, ConstPtr<uint8_t> flags, int M, bool skip_ground,
                                     const Params<float>& p, bool act) {
    bool hit = false;
    for (int j = 0; j < M; ++j) {
        if ((j & 7) == 0 && !__any(act && !hit)) break;
        ConstPtr<float> o = obb + 15 * j;
        float x[12];
        for (int k = 0; k < 12; ++k) x[k] = o[k];
        const Box<float> bx{{-o[12], -o[13], -o[14]}, {o[12], o[13], o[14]}};
        const bool skip = skip_ground && (flags[j] & kObjGround);
        if (act && !hit && !skip && collision_box(to_box(x, P0), to_box(x, P1), r, rn, bx, p)) hit = true;
    }
    return hit;
}

// The first-collision search over the items and slots above.  The states
// that are still needed (no collision found yet at an earlier state of their
// edge) are gathered into a queue in LDS, 64 per evaluation, so that slots of
// finished edges take no lanes; the state test of each lane is written as
// uniform loops (a lane that is done, or has no state, rides along masked
// off) so that the tables, or the oriented boxes (OBB: every box through the
// full test in its frame), arrive through scalar loads.  The joint values
// come from the edge's rows (q = q1 + (k step) dir per joint, rounded as the
// reference's interpolation is); LDS holds the queue and the skeletons for
// the self pairs.
constexpr int kQueue = 128;
template <bool OBB>
__global__ __launch_bounds__(64) void k_edge_first_q(EdgeGroups gr, EnvDev env, QueryParams qp,
                                                        const float* __restrict__ q1, const float* __restrict__ dir,
                                                        int32_t row_stride, float step,
                                                        const int32_t* __restrict__ order,
                                                        const int32_t* __restrict__ rounds,
                                                        int32_t* __restrict__ first, uint32_t* __restrict__ counter,
                                                        int32_t total_items, int spl) {
    extern __shared__ __align__(16) unsigned char smem[];
    int2* queue = reinterpret_cast<int2*>(smem);
    float* skel = reinterpret_cast<float*>(smem + kQueue * sizeof(int2));
    const int lane = threadIdx.x;
    const int M = env.num_obj, G = (M + kSlabGroup - 1) / kSlabGroup;
    const Params<float> p{qp.eq, qp.maxd};
    // uniform state: queued entries and their group, the current item and
    // its next 64-slot slice
    int qn = 0, qg = 0, item = 0, sub = spl;
    bool drained = false;
    for (;;) {
        // gather: slices until 64 states wait, the group changes or the items run out
        while (qn < 64 && !drained) {
            if (sub == spl) {
                int it = 0;
                if (lane == 0) it = (int)atomicAdd(counter, 1u);
                item = __builtin_amdgcn_readfirstlane(__shfl(it, 0, 64));
                sub = 0;
                if (item >= total_items) {
                    drained = true;
                    break;
                }
            }
            const int g = as_const(gr.items)[item].x;
            if (qn > 0 && g != qg) break;   // the queued group is evaluated first
            qg = g;
            const int32_t slot = as_const(gr.items)[item].y + sub * 64 + lane;
            ++sub;
            bool live = slot < gr.num_slots[g];
            int k = 0;
            int32_t e = 0;
            if (live) {
                k = round_of(rounds + gr.round_off[g], gr.num_rounds[g] + 1, slot);
                e = order[gr.order_off[g] + (slot - rounds[gr.round_off[g] + k])];
                live = __atomic_load_n(&first[e], __ATOMIC_RELAXED) > k;
            }
            const uint64_t m = __ballot(live);
            if (live) {
                const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                 __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                queue[qn + rank] = make_int2(e, k);
            }
            qn += __builtin_popcountll(m);
            __syncthreads();
        }
        if (qn == 0) {
            if (drained) break;
            continue;
        }
        // evaluate the last min(qn, 64) queued states
        const int take = min(qn, 64);
        qn -= take;
        bool live = lane < take;
        int32_t e = 0;
        int k = 0;
        if (live) {
            const int2 x = queue[qn + lane];
            e = x.x;
            k = x.y;
            live = __atomic_load_n(&first[e], __ATOMIC_RELAXED) > k;
        }
        __syncthreads();
        const RobotDev& rb = gr.rb[qg];
        ConstPtr<int32_t> keys = as_const(gr.link_key) + gr.key_off[qg];
        const float* qa = q1 + (int64_t)e * row_stride;
        const float* qd = dir + (int64_t)e * row_stride;
        const float sk = step * (float)k;
        auto sc = [&](int j, double* st, double* ct) {
            const float qv = qa[j] + sk * qd[j];
            sincos((double)qv, st, ct);
        };
        const bool pairs = rb.num_pairs > 0;
        FKStream fk;
        fk_init(fk);
        Vec3<float> P0 = skel_point_sc<float>(fk, rb, 0, sc);
        // largest coordinate magnitude and NaN check of the skeleton, for the
        // self-pair bound
        float cmax = fmaxf(fmaxf(fabsf(P0.x), fabsf(P0.y)), fabsf(P0.z));
        float csum = (P0.x + P0.y) + P0.z;
        if (pairs) {
            skel[lane] = P0.x;
            skel[64 + lane] = P0.y;
            skel[128 + lane] = P0.z;
        }
        bool done = !live, bad = false;
        for (int i = 0; i < rb.num_links; ++i) {
            if (!__any(!done)) break;
            const Vec3<float> P1 = skel_point_sc<float>(fk, rb, i + 1, sc);
            if (pairs) {
                skel[(3 * i + 3) * 64 + lane] = P1.x;
                skel[(3 * i + 4) * 64 + lane] = P1.y;
                skel[(3 * i + 5) * 64 + lane] = P1.z;
                cmax = fmaxf(cmax, fmaxf(fmaxf(fabsf(P1.x), fabsf(P1.y)), fabsf(P1.z)));
                csum = csum + ((P1.x + P1.y) + P1.z);
            }
            bool h;
            if (OBB) {
                h = link_hits_obb(P0, P1, rb.radius[i], rb.rnew32[i], as_const(env.obb), as_const(env.flags), M,
                                  i < rb.ground_included, p, !done);
            } else {
                ConstPtr<SlabEntry> tab = as_const(gr.slab) + (size_t)keys[i] * (G + M);
                h = link_hits(P0, P1, rb.radius[i], rb.rnew32[i], tab, G, M, as_const(env.aabb), p, !done);
            }
            if (h) {
                bad = true;
                done = true;
                atomicMin(&first[e], k);   // at once: the edge's later states in flight stop at their next link
            }
            // a collision found meanwhile at an earlier state of this edge (by
            // another wave) ends this state's test: its verdict no longer
            // matters.  The first colliding state is never cut short, since
            // only a colliding state before it could cut it.
            if (!done && __atomic_load_n(&first[e], __ATOMIC_RELAXED) <= k) done = true;
            P0 = P1;
        }
        const bool bound_ok = csum == csum;
        for (int m = 0; m < rb.num_pairs; ++m) {
            if (!__any(!done)) break;
            const int a = rb.pairs[2 * m], b = rb.pairs[2 * m + 1];
            const float ra = rb.radius[a], rbb = rb.radius[b];
            auto pt = [&](int kk) {
                return Vec3<float>{skel[(3 * kk) * 64 + lane], skel[(3 * kk + 1) * 64 + lane],
                                   skel[(3 * kk + 2) * 64 + lane]};
            };
            const Vec3<float> A = pt(a), B = pt(a + 1), C = pt(b), D = pt(b + 1);
            // distance of the two bounding boxes, squared (self_pair_hits's
            // bound, written out here: the whole wave tests it, and only the
            // lanes that it does not settle take the full test)
            const float gx = fmaxf(fmaxf(fminf(C.x, D.x) - fmaxf(A.x, B.x), fminf(A.x, B.x) - fmaxf(C.x, D.x)), 0.0f);
            const float gy = fmaxf(fmaxf(fminf(C.y, D.y) - fmaxf(A.y, B.y), fminf(A.y, B.y) - fmaxf(C.y, D.y)), 0.0f);
            const float gz = fmaxf(fmaxf(fminf(C.z, D.z) - fmaxf(A.z, B.z), fminf(A.z, B.z) - fmaxf(C.z, D.z)), 0.0f);
            const float R = fmaxf(ra + rbb, p.eq);
            const float thr = R + 4e-6f * (cmax + R);
            const bool apart = bound_ok && thr <= p.maxd && (gx * gx + gy * gy) + gz * gz > thr * thr;
            if (!done && !apart) {
                const float sd = seg_seg<float, false>(A, B, C, D, p, nullptr);
                if (sd - ra - rbb < 0) {
                    bad = true;
                    done = true;
                    atomicMin(&first[e], k);
                }
            }
            if ((m & 15) == 15 && !done && __atomic_load_n(&first[e], __ATOMIC_RELAXED) <= k) done = true;
        }
    }
}

// isValid(q1, q2): one lane per (edge, state), the collision tree through
// the axis-test tables (axis-aligned scenes) or every oriented box in its
// frame (OBB), as in k_edge_first_q (uniform loops per wave, joint values
// computed where the kinematics reads them), no LDS.
template <bool OBB>
__global__ __launch_bounds__(256) void k_edge_valid_q(RobotDev rb, EnvDev env, QueryParams qp,
                                                         const float* __restrict__ q1, const float* __restrict__ q2,
                                                         int64_t n, uint8_t* __restrict__ valid,
                                                         const int32_t* __restrict__ link_key,
                                                         const SlabEntry* __restrict__ slab) {
    const int K = qp.num_checks;
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = g < n * K;
    const int64_t e = live ? g / K : 0;
    const int kk = live ? K - (int)(g % K) : K;   // k = K .. 1
    const int D = rb.num_dofs;
    const float* a = q1 + e * D;
    const float* b = q2 + e * D;
    const float dist = vecxf_norm_diff(a, b, D);
    // RealVectorSpace::interpolateEdge (RealVectorSpace.cpp:75-88) with
    // step = k / num_checks * dist
    const float step = (float)kk / (float)K * dist;
    const float f = dist > 0 ? step / dist : 0.0f;
    auto sc = [&](int j, double* st, double* ct) {
        const float qv = dist > 0 ? a[j] + (b[j] - a[j]) * f : b[j];
        sincos((double)qv, st, ct);
    };
    const int M = env.num_obj, G = (M + kSlabGroup - 1) / kSlabGroup;
    const Params<float> p{qp.eq, qp.maxd};
    FKStream fk;
    fk_init(fk);
    Vec3<float> P0 = skel_point_sc<float>(fk, rb, 0, sc);
    bool done = !live;
    for (int i = 0; i < rb.num_links; ++i) {
        if (!__any(!done)) break;
        const Vec3<float> P1 = skel_point_sc<float>(fk, rb, i + 1, sc);
        bool h;
        if (OBB) {
            h = link_hits_obb(P0, P1, rb.radius[i], rb.rnew32[i], as_const(env.obb), as_const(env.flags), M,
                              i < rb.ground_included, p, !done);
        } else {
            ConstPtr<SlabEntry> tab = as_const(slab) + (size_t)as_const(link_key)[i] * (G + M);
            h = link_hits(P0, P1, rb.radius[i], rb.rnew32[i], tab, G, M, as_const(env.aabb), p, !done);
        }
        if (h) {
            valid[e] = 0;
            done = true;
        }
        P0 = P1;
    }
}

inline dim3 blocks_for(int64_t n, int b) { return dim3((unsigned)((n + b - 1) / b)); }

}  // namespace

hipError_t launch_edge_valid(const RobotDev& rb, const EnvDev& env, const QueryParams& qp, const float* q1,
                             const float* q2, int64_t n, uint8_t* valid, const int32_t* link_key,
                             const float* keys, int nkeys, SlabEntry* slab, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u8, blocks_for(n, 256), dim3(256), 0, st, valid, n, (uint8_t)1);
    if (!env.use_obb) {
        const int G = (env.num_obj + kSlabGroup - 1) / kSlabGroup;
        if (nkeys > 0 && G > 0)
            hipLaunchKernelGGL(k_slab_build, dim3((unsigned)((G + 63) / 64), (unsigned)nkeys), dim3(64), 0, st, env,
                               keys, nkeys, slab);
        hipLaunchKernelGGL(k_edge_valid_q<false>, blocks_for(n * qp.num_checks, 256), dim3(256), 0, st, rb, env, qp,
                           q1, q2, n, valid, link_key, slab);
    } else {
        hipLaunchKernelGGL(k_edge_valid_q<true>, blocks_for(n * qp.num_checks, 256), dim3(256), 0, st, rb, env, qp,
                           q1, q2, n, valid, link_key, slab);
    }
    return hipGetLastError();
}

size_t edge_first_lds_bytes(const EdgeGroups& gr, const EnvDev& env) {
    (void)env;
    return kQueue * sizeof(int2) + (size_t)64 * 4 * 3 * (gr.max_links + 1);
}

hipError_t launch_edge_first(const EdgeGroups& gr, const EnvDev& env, const QueryParams& qp, const float* q1,
                             const float* dir, int32_t row_stride, float step, const int32_t* order,
                             const int32_t* rounds, int32_t* first, int64_t num_edges, uint32_t* counter,
                             int32_t total_items, int waves, int spl, const float* keys, int nkeys,
                             hipStream_t st) {
    if (num_edges <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_first_init, blocks_for(num_edges, 256), dim3(256), 0, st, first, num_edges, counter);
    if (total_items > 0) {
        const int grid = std::max(1, std::min(total_items, waves));
        const size_t sm = edge_first_lds_bytes(gr, env);
        if (!env.use_obb) {
            const int G = (env.num_obj + kSlabGroup - 1) / kSlabGroup;
            if (nkeys > 0 && G > 0)
                hipLaunchKernelGGL(k_slab_build, dim3((unsigned)((G + 63) / 64), (unsigned)nkeys), dim3(64), 0, st,
                                   env, keys, nkeys, const_cast<SlabEntry*>(gr.slab));
            hipLaunchKernelGGL(k_edge_first_q<false>, dim3(grid), dim3(64), sm, st, gr, env, qp, q1, dir, row_stride,
                               step, order, rounds, first, counter, total_items, spl);
        } else {
            hipLaunchKernelGGL(k_edge_first_q<true>, dim3(grid), dim3(64), sm, st, gr, env, qp, q1, dir, row_stride,
                               step, order, rounds, first, counter, total_items, spl);
        }
    }
    hipLaunchKernelGGL(k_first_final, blocks_for(num_edges, 256), dim3(256), 0, st, first, num_edges);
    return hipGetLastError();
}

}  // namespace rvs
