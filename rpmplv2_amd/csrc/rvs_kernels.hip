// gfx950 kernels of the capsule collision / distance engine.
//
// Layout in HBM (configuration-major):
//   q        [N][n]       float joint coordinates
//   d        [N]          float (or double) minimum distance
//   profile  [N][L]       per-link distance profile
//   nearest  [N][M][L][6] nearest points (robot xyz, obstacle xyz)
//   verdict  [N]          uint8, 1 = collision-free
// One lane owns one configuration.  The robot (segments, skeleton recipe,
// radii) is read with wave-uniform addresses, so it comes through the scalar
// cache; the obstacle set is staged once per workgroup in LDS and every lane
// of a wave reads the same obstacle at the same time (an LDS broadcast).
//
// This file is compiled without FMA contraction, with IEEE division and
// square root: the float kernels reproduce the reference's float rounding
// bit for bit, the double kernels run the same algorithm in FP64.
// No FMA contraction anywhere in this translation unit, headers included:
// the float kernels must keep the reference's rounding.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <climits>

#include "rvs_fk.hpp"
#include "rvs_geometry.hpp"
#include "rvs_internal.hpp"
#include "rvs_radii.hpp"
#include "rvs_stage.hpp"


namespace rvs {

namespace {

constexpr int kBlock = 256;

// ---------------------------------------------------------------- distance
// RealVectorSpace::computeDistance (RealVectorSpace.cpp:318-392) for a fresh
// state: per-link profile, global minimum, early exit on the first pair whose
// running minimum is <= 0 (the profile keeps that value, later links stay 0,
// the distance is 0 and the nearest points are dropped).
// With an index list, thread g handles configuration idx[g] for g < *count
// (the refinement pass of RVS_PREC_FP32_REFINE).
template <typename T, bool WANT_NP, bool OBB>
__device__ __attribute__((always_inline)) inline void distance_body(const RobotDev& rb, const EnvDev& env,
                                                                    const QueryParams& qp, const float* __restrict__ q,
                                                                    int64_t n, const DistOut& out,
                                                                    const int32_t* __restrict__ perm) {
    extern __shared__ __align__(16) unsigned char smem[];
    const Staged<T> s = stage_env<T, OBB>(env, smem);
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    // lane g takes configuration perm[g] of the walk order (similar poses in
    // one wave): the results do not depend on the order
    if (perm) c = perm[c];
    const int M = env.num_obj;
    const int L = rb.num_links;
    const float* qc = q + c * rb.num_dofs;
    const Params<T> p{(T)qp.eq, (T)qp.maxd};

    FKStream fk;
    fk_init(fk);
    Vec3<T> P0 = skel_point<T>(fk, rb, qc, 0);
    T d = p.maxd;
    bool hit = false;
    int i = 0;
    for (; i < L; ++i) {
        const Vec3<T> P1 = skel_point<T>(fk, rb, qc, i + 1);
        const T r = (T)rb.radius[i];
        T prof = p.maxd;
        Vec3<T> slo, shi;
        seg_bounds(P0, P1, slo, shi);
        for (int j = 0; j < M; ++j) {
            T dt;
            NP<T> np;
            // broad phase (nearest points are stored for every pair, so not with them)
            if (!WANT_NP && cull_pair(env.cull + 6 * j, slo, shi, cull_bound(prof, r, p.eq))) continue;
            if (s.gnd[j] && i < rb.ground_included) {
                dt = p.maxd;
                if (WANT_NP) {
                    np.r = Vec3<T>{T(0), T(0), T(0)};
                    np.o = Vec3<T>{T(0), T(0), -p.maxd};
                    np.valid = true;
                }
            } else if (OBB) {
                // oriented box: the pair is solved in the box frame
                const float* x = s.xf + 12 * j;
                const Box<T> bx = s.box[j];
                dt = capsule_box<T, WANT_NP>(to_box(x, P0), to_box(x, P1), r, bx, p, &np, nullptr);
                if (WANT_NP && np.valid) {
                    np.r = from_box(x, np.r);
                    np.o = from_box(x, np.o);
                }
            } else {
                const Box<T> bx = s.box[j];
                dt = capsule_box<T, WANT_NP>(P0, P1, r, bx, p, &np, nullptr);
            }
            if (dt > s.tol[j]) dt = p.maxd;
            prof = smin(prof, dt);
            if (prof <= 0) { hit = true; break; }
            if (WANT_NP) {
                float* e = out.nearest + ((c * M + j) * L + i) * 6;
                if (np.valid) {
                    e[0] = (float)np.r.x; e[1] = (float)np.r.y; e[2] = (float)np.r.z;
                    e[3] = (float)np.o.x; e[4] = (float)np.o.y; e[5] = (float)np.o.z;
                } else {
                    for (int k = 0; k < 6; ++k) e[k] = NAN;
                }
            }
        }
        if (out.profile) out.profile[c * L + i] = (float)prof;
        if (out.profile64) out.profile64[c * L + i] = (double)prof;
        if (hit) break;
        d = smin(d, prof);
        P0 = P1;
    }
    for (int k = i + 1; k < L; ++k) {
        if (out.profile) out.profile[c * L + k] = 0.0f;
        if (out.profile64) out.profile64[c * L + k] = 0.0;
    }
    if (hit) {
        d = T(0);
        if (WANT_NP) {
            float* e = out.nearest + c * (int64_t)M * L * 6;
            for (int k = 0; k < M * L * 6; ++k) e[k] = NAN;
        }
    }
    if (out.d) out.d[c] = (float)d;
    if (out.d64) out.d64[c] = (double)d;
    if (out.verdict) out.verdict[c] = hit ? 0 : 1;
}

// The float kernel is held to 128 VGPRs (4 waves per SIMD); the double
// kernel keeps its natural allocation (a cap would spill).
template <bool WANT_NP, bool OBB>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_distance32(
    RobotDev rb, EnvDev env, QueryParams qp, const float* __restrict__ q, int64_t n, DistOut out,
    const int32_t* __restrict__ perm) {
    distance_body<float, WANT_NP, OBB>(rb, env, qp, q, n, out, perm);
}

// Exact recomputation of the configurations the FP32 bulk pass flagged: one
// wave per configuration, one link at a time.  For link i, lane t evaluates
// obstacles j = t, t + 64, ... with the exact tree; the wave then reduces the
// smallest filtered distance and the first obstacle whose filtered distance
// is <= 0.  That gives exactly what the reference's sequential loop gives:
// its running minimum starts at MAX_DISTANCE, so a NaN distance never enters
// it (std::min keeps the first argument), and it stops at the first pair
// whose distance is <= 0, which then becomes the profile value (the running
// minimum is positive up to that pair).  A minimum of floats is exact in any
// order, and only the first non-positive value can reach the profile, so the
// reduction order cannot change a bit of the result.  Without nearest points
// the obstacles go in rounds of 64 (lane t takes obstacle j0 + t), the wave's
// minimum after each round bounds the broad phase of the next, and the link
// ends after the first round with a hit.
constexpr int kRefineWaves = 4;   // waves per workgroup of the refinement kernel (they share the staged obstacles)

// Entry k of the refinement's work: configuration c and the first link to
// recompute, k0 (entry k itself and k0 = 0 for a direct call, idx null; k0
// = 0 too without a profile, which holds the values of the links before it).
__device__ __attribute__((always_inline)) inline void flag_decode(const int32_t* idx, int64_t k, const QueryParams& qp,
                                                                  const DistOut& out, int64_t& c, int& k0) {
    if (!idx) {
        c = k;
        k0 = 0;
        return;
    }
    const int64_t e = (int64_t)idx[k];
    if (qp.flag_link) {
        c = e & (((int64_t)1 << kFlagLinkShift) - 1);
        k0 = out.profile ? (int)(e >> kFlagLinkShift) : 0;
    } else {
        c = e;
        k0 = 0;
    }
}

// With idx null the kernel serves a direct call: wave k takes configuration
// k < n (latency-bound small batches: one wave per configuration instead of
// one lane).  With nearest points every pair is evaluated (no broad phase)
// and its points stored; a collision leaves NaN points, as the lane kernel.
template <bool OBB, bool WANT_NP>
__global__ __launch_bounds__(64 * kRefineWaves) void k_distance_refine(RobotDev rb, EnvDev env, QueryParams qp,
                                                                       const float* __restrict__ q, DistOut out,
                                                                       const int32_t* __restrict__ idx,
                                                                       const uint32_t* __restrict__ count,
                                                                       int64_t n_direct, uint32_t* count_out) {
    extern __shared__ __align__(16) unsigned char smem[];
    const Staged<float> s = stage_env<float, OBB>(env, smem);
    const int M = env.num_obj, L = rb.num_links;
    const int t = threadIdx.x & 63;
    const uint32_t wave = threadIdx.x >> 6;
    const Params<float> p{qp.eq, qp.maxd};
    const int64_t total = idx ? (int64_t)*count : n_direct;
    if (count_out && blockIdx.x == 0 && threadIdx.x == 0) *count_out = (uint32_t)total;
    for (int64_t k = (int64_t)blockIdx.x * kRefineWaves + wave; k < total; k += (int64_t)gridDim.x * kRefineWaves) {
        int64_t c = 0;
        int k0 = 0;   // the first link to recompute
        flag_decode(idx, k, qp, out, c, k0);
        const float* qc = q + c * rb.num_dofs;
        // joint sines and cosines: with at most 64 joints lane m computes
        // joint m's (the same function of the same value as the serial
        // stream) and the kinematics take them by shuffles, so the chain of
        // frame products no longer waits on each joint's sincos in turn
        const bool par = rb.num_dofs <= 64;
        double my_s = 0.0, my_c = 1.0;
        if (par && t < rb.num_dofs) sincos((double)qc[t], &my_s, &my_c);
        const auto sc = [&](int m, double* st, double* ct) {
            if (par) {
                *st = __shfl(my_s, m, 64);
                *ct = __shfl(my_c, m, 64);
            } else {
                sincos((double)qc[m], st, ct);
            }
        };
        FKStream fk;
        fk_init(fk);
        Vec3<float> P0 = skel_point_sc<float>(fk, rb, 0, sc);
        // the links before k0 keep the bulk pass's values (no hit)
        float d = p.maxd;
        for (int i2 = 0; i2 < k0; ++i2) d = smin(d, out.profile[c * L + i2]);
        bool hit = false;
        int i = 0;
        for (; i < L; ++i) {
            const Vec3<float> P1 = skel_point_sc<float>(fk, rb, i + 1, sc);
            if (i < k0) {
                P0 = P1;
                continue;
            }
            const float r = rb.radius[i];
            float lo = p.maxd;          // smallest filtered distance of this lane's obstacles
            int first = INT_MAX;        // first of them with filtered distance <= 0
            float first_v = 0.0f;
            Vec3<float> slo, shi;
            seg_bounds(P0, P1, slo, shi);
            // filtered distance of (link i, obstacle j), and its nearest points
            auto pair = [&](int j, NP<float>& np) -> float {
                float dt;
                if (s.gnd[j] && i < rb.ground_included) {
                    dt = p.maxd;
                    if (WANT_NP) {
                        np.r = Vec3<float>{0.0f, 0.0f, 0.0f};
                        np.o = Vec3<float>{0.0f, 0.0f, -p.maxd};
                        np.valid = true;
                    }
                } else if (OBB) {
                    const float* x = s.xf + 12 * j;
                    const Box<float> bx = s.box[j];
                    dt = capsule_box<float, WANT_NP>(to_box(x, P0), to_box(x, P1), r, bx, p, &np, nullptr);
                    if (WANT_NP && np.valid) {
                        np.r = from_box(x, np.r);
                        np.o = from_box(x, np.o);
                    }
                } else {
                    const Box<float> bx = s.box[j];
                    dt = capsule_box<float, WANT_NP>(P0, P1, r, bx, p, &np, nullptr);
                }
                return dt > s.tol[j] ? p.maxd : dt;
            };
            if (WANT_NP) {
                // every pair, its points stored
                for (int j = t; j < M; j += 64) {
                    NP<float> np;
                    const float dt = pair(j, np);
                    lo = smin(lo, dt);
                    if (dt <= 0 && first == INT_MAX) {
                        first = j;
                        first_v = dt;
                    }
                    float* e = out.nearest + ((c * M + j) * L + i) * 6;
                    if (np.valid) {
                        e[0] = np.r.x; e[1] = np.r.y; e[2] = np.r.z;
                        e[3] = np.o.x; e[4] = np.o.y; e[5] = np.o.z;
                    } else {
                        for (int k2 = 0; k2 < 6; ++k2) e[k2] = NAN;
                    }
                }
            } else {
                // rounds of 64 obstacles.  The broad phase bound is the
                // wave's minimum over the earlier rounds (an upper bound of
                // the link's minimum; positive while no pair has hit, so a
                // pair that can be <= 0 is never skipped), and the first round
                // with a hit ends the link: later rounds hold only larger
                // obstacle indices, so the first hit is among this round's.
                for (int j0 = 0; j0 < M; j0 += 64) {
                    const int j = j0 + t;
                    if (j < M && !cull_pair(env.cull + 6 * j, slo, shi, cull_bound(lo, r, p.eq))) {
                        NP<float> np;
                        const float dt = pair(j, np);
                        lo = smin(lo, dt);
                        if (dt <= 0) {
                            first = j;
                            first_v = dt;
                        }
                    }
                    if (__ballot(first != INT_MAX)) break;
                    for (int o = 32; o > 0; o >>= 1) lo = smin(lo, __shfl_xor(lo, o, 64));
                }
            }
            for (int o = 32; o > 0; o >>= 1) {
                lo = smin(lo, __shfl_xor(lo, o, 64));
                const int f2 = __shfl_xor(first, o, 64);
                const float v2 = __shfl_xor(first_v, o, 64);
                if (f2 < first) {
                    first = f2;
                    first_v = v2;
                }
            }
            hit = first != INT_MAX;
            const float prof = hit ? first_v : lo;
            if (t == 0 && out.profile) out.profile[c * L + i] = prof;
            if (hit) break;
            d = smin(d, prof);
            P0 = P1;
        }
        if (WANT_NP && hit) {
            // the points this wave stored for the pairs so far must land
            // before the NaN fill of the same addresses by other lanes
            __builtin_amdgcn_s_waitcnt(0);
            float* e = out.nearest + c * (int64_t)M * L * 6;
            for (int k2 = t; k2 < M * L * 6; k2 += 64) e[k2] = NAN;
        }
        if (t == 0) {
            for (int k2 = i + 1; k2 < L; ++k2)
                if (out.profile) out.profile[c * L + k2] = 0.0f;
            if (out.d) out.d[c] = hit ? 0.0f : d;
            if (out.verdict) out.verdict[c] = hit ? 0 : 1;
        }
    }
}

// Exact recomputation for small obstacle sets (links x obstacles <=
// kPairsMax, fewer than 64 links): one wave per configuration evaluates all
// of its (link, obstacle) pairs at once, lane t taking pairs t, t + 64, ...
// of pair = link * M + obstacle, instead of one link per round.  Lane m
// computes joint m's sine and cosine, every lane runs the same forward
// kinematics on them and lane k keeps skeleton point k; a pair fetches its
// end points by lane shuffles.  The filtered pair distances go to
// the wave's LDS row, lane i then walks link i's obstacles in order as the
// reference's loop does (running std::min from MAX_DISTANCE, stop at the
// first value <= 0), and the links combine in order: the first link with a
// hit ends the configuration (later links 0, distance 0), otherwise the
// distance is the smallest link value (a minimum of non-NaN floats, exact in
// any order).  Every pair is evaluated (no broad phase, no early exit), which
// changes no result: the per-link walk over the stored values is the
// reference's loop.  The workgroup's waves step through the configurations
// together so that the LDS row can be handed over with a workgroup barrier.
constexpr int kPairsMax = 256;

__host__ __device__ inline size_t pairs_row_off(const EnvDev& env) {
    return (staged_bytes<float>(env) + 15) & ~size_t(15);
}

inline bool refine_pairs_fit(const RobotDev& rb, const EnvDev& env) {
    return rb.num_links < 64 && rb.num_dofs <= 64 && (int64_t)rb.num_links * env.num_obj <= kPairsMax;
}

template <bool OBB>
__global__ __launch_bounds__(64 * kRefineWaves) void k_distance_refine_pairs(RobotDev rb, EnvDev env,
                                                                             QueryParams qp,
                                                                             const float* __restrict__ q, DistOut out,
                                                                             const int32_t* __restrict__ idx,
                                                                             const uint32_t* __restrict__ count,
                                                                             int64_t n_direct, uint32_t* count_out) {
    extern __shared__ __align__(16) unsigned char smem[];
    const Staged<float> s = stage_env<float, OBB>(env, smem);
    const int M = env.num_obj, L = rb.num_links;
    const int npairs = L * M;
    const int t = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    float* row = reinterpret_cast<float*>(smem + pairs_row_off(env)) + wave * kPairsMax;
    const Params<float> p{qp.eq, qp.maxd};
    const int64_t total = idx ? (int64_t)*count : n_direct;
    if (count_out && blockIdx.x == 0 && threadIdx.x == 0) *count_out = (uint32_t)total;
    for (int64_t kb = (int64_t)blockIdx.x * kRefineWaves; kb < total; kb += (int64_t)gridDim.x * kRefineWaves) {
        const int64_t k = kb + wave;
        const bool live = k < total;   // uniform over the wave
        int64_t c = 0;
        int k0 = 0;   // the first link to recompute; the pairs of the links before it are not evaluated
        if (live) {
            flag_decode(idx, k, qp, out, c, k0);
            const float* qc = q + c * rb.num_dofs;
            // lane m computes the sine and cosine of joint m (the same
            // function of the same value as a serial stream); the FK stream
            // then takes them by shuffles
            double my_s = 0.0, my_c = 1.0;
            if (t < rb.num_dofs) sincos((double)qc[t], &my_s, &my_c);
            const auto sc = [&](int m, double* st, double* ct) {
                *st = __shfl(my_s, m, 64);
                *ct = __shfl(my_c, m, 64);
            };
            FKStream fk;
            fk_init(fk);
            Vec3<float> mine{0.0f, 0.0f, 0.0f};
            for (int pt = 0; pt <= L; ++pt) {
                const Vec3<float> P = skel_point_sc<float>(fk, rb, pt, sc);
                if (t == pt) mine = P;
            }
            for (int base = k0 * M; base < npairs; base += 64) {   // uniform rounds: the shuffles see every lane
                const int pp = base + t;
                const int i = pp < npairs ? pp / M : 0;
                const int j = pp - i * M;
                const Vec3<float> A{__shfl(mine.x, i, 64), __shfl(mine.y, i, 64), __shfl(mine.z, i, 64)};
                const Vec3<float> B{__shfl(mine.x, i + 1, 64), __shfl(mine.y, i + 1, 64), __shfl(mine.z, i + 1, 64)};
                if (pp < npairs) {
                    float dt;
                    if (s.gnd[j] && i < rb.ground_included) {
                        dt = p.maxd;
                    } else if (OBB) {
                        const float* x = s.xf + 12 * j;
                        dt = capsule_box<float, false>(to_box(x, A), to_box(x, B), rb.radius[i], s.box[j], p, nullptr,
                                                       nullptr);
                    } else {
                        dt = capsule_box<float, false>(A, B, rb.radius[i], s.box[j], p, nullptr, nullptr);
                    }
                    if (dt > s.tol[j]) dt = p.maxd;
                    row[pp] = dt;
                }
            }
        }
        __syncthreads();
        if (live) {
            float prof = p.maxd;
            bool hit = false;
            if (t < k0) {
                prof = out.profile[c * L + t];   // a link before k0: the bulk pass's value, no hit
            } else if (t < L) {
                for (int j = 0; j < M; ++j) {
                    prof = smin(prof, row[t * M + j]);
                    if (prof <= 0) {
                        hit = true;
                        break;
                    }
                }
            }
            const uint64_t hm = __ballot(hit);
            const int first_hit = hm ? __ffsll((unsigned long long)hm) - 1 : L;
            float d = (t < first_hit && t < L) ? prof : p.maxd;
            for (int o = 32; o > 0; o >>= 1) d = smin(d, __shfl_xor(d, o, 64));
            if (t >= k0 && t < L && out.profile) out.profile[c * L + t] = t <= first_hit ? prof : 0.0f;
            if (t == 0) {
                if (out.d) out.d[c] = hm ? 0.0f : d;
                if (out.verdict) out.verdict[c] = hm ? 0 : 1;
            }
        }
        __syncthreads();   // the row is rewritten by the next round
    }
}

template <bool WANT_NP, bool OBB>
__global__ __launch_bounds__(kBlock) void k_distance64(RobotDev rb, EnvDev env, QueryParams qp,
                                                       const float* __restrict__ q, int64_t n, DistOut out,
                                                       const int32_t* __restrict__ perm) {
    distance_body<double, WANT_NP, OBB>(rb, env, qp, q, n, out, perm);
}

// ---------------------------------------------------------------- isValid
// RealVectorSpace::isValid(q) (RealVectorSpace.cpp:277-309): collision tree
// per pair, ground skipped for the first ground_included links, stop on the
// first hit.
template <typename T, bool OBB>
__global__ __launch_bounds__(kBlock) void k_valid(RobotDev rb, EnvDev env, QueryParams qp,
                                                  const float* __restrict__ q, int64_t n, uint8_t* valid,
                                                  const int32_t* __restrict__ perm) {
    extern __shared__ __align__(16) unsigned char smem[];
    const Staged<T> s = stage_env<T, OBB>(env, smem);
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    if (perm) c = perm[c];   // walk order: the verdicts do not depend on it
    const int M = env.num_obj;
    const int L = rb.num_links;
    const float* qc = q + c * rb.num_dofs;
    const Params<T> p{(T)qp.eq, (T)qp.maxd};
    FKStream fk;
    fk_init(fk);
    Vec3<T> P0 = skel_point<T>(fk, rb, qc, 0);
    bool ok = true;
    for (int i = 0; i < L && ok; ++i) {
        const Vec3<T> P1 = skel_point<T>(fk, rb, qc, i + 1);
        const T r = (T)rb.radius[i];
        const T rn = rnew_of<T>(rb, i);
        for (int j = 0; j < M; ++j) {
            if (s.gnd[j] && i < rb.ground_included) continue;
            bool col;
            if (OBB) {
                const float* x = s.xf + 12 * j;
                col = collision_box(to_box(x, P0), to_box(x, P1), r, rn, s.box[j], p);
            } else {
                col = collision_box(P0, P1, r, rn, s.box[j], p);
            }
            if (col) { ok = false; break; }
        }
        P0 = P1;
    }
    valid[c] = ok ? 1 : 0;
}

// ---------------------------------------------------------------- underestimation
// RealVectorSpace::computeDistanceUnderestimation (RealVectorSpace.cpp:
// 402-446) on a fresh state, in float.  np_ref holds [M][L][6] per
// configuration (stride np_stride floats; 0 = one set for all).
// With seg_off (nseg + 1 non-decreasing offsets), configuration c uses the
// nearest-point set of its segment r (seg_off[r] <= c < seg_off[r + 1]):
// requests of a served batch each bring their own set.  With max_vel [M]
// the planes first move towards the robot for delta_t at each obstacle's
// maximal speed (AbstractTrajectory.cpp:272-320).
__global__ __launch_bounds__(kBlock) void k_under(RobotDev rb, EnvDev env, QueryParams qp,
                                                  const float* __restrict__ q, int64_t n,
                                                  const float* __restrict__ np_ref, int64_t np_stride, float* d_out,
                                                  float* prof_out, uint8_t* stored,
                                                  const int64_t* __restrict__ seg_off, int nseg,
                                                  const float* __restrict__ max_vel, float delta_t) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int M = env.num_obj;
    const int L = rb.num_links;
    const float* qc = q + c * rb.num_dofs;
    const float* npc = np_ref + c * np_stride;
    if (seg_off) {
        int lo = 0, hi = nseg - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (seg_off[mid] <= c) lo = mid; else hi = mid - 1;
        }
        npc = np_ref + (int64_t)lo * M * L * 6;
    }
    const float maxd = qp.maxd;
    FKStream fk;
    fk_init(fk);
    Vec3<float> P0 = skel_point<float>(fk, rb, qc, 0);
    float d = maxd;
    bool neg = false;
    int i = 0;
    for (; i < L && !neg; ++i) {
        const Vec3<float> P1 = skel_point<float>(fk, rb, qc, i + 1);
        const float r = rb.radius[i];
        float prof = maxd;
        for (int j = 0; j < M; ++j) {
            const float* e = npc + ((int64_t)j * L + i) * 6;
            const Vec3<float> O{e[3], e[4], e[5]};
            if (vnorm(O) == maxd) continue;
            const Vec3<float> R{e[0], e[1], e[2]};
            const float t = max_vel ? plane_term_moving(P0, P1, R, O, r, max_vel[j], delta_t)
                                    : plane_term(P0, P1, R, O, r);
            if (t < 0) { neg = true; break; }
            prof = smin(prof, t);
        }
        if (neg) break;
        if (prof_out) prof_out[c * L + i] = prof;
        d = smin(d, prof);
        P0 = P1;
    }
    if (neg) {
        if (prof_out) for (int k = 0; k < L; ++k) prof_out[c * L + k] = 0.0f;
        d = 0.0f;
    }
    d_out[c] = d;
    if (stored) stored[c] = neg ? 0 : 1;
}

// ---------------------------------------------------------------- self distance
// Capsule-capsule distance of the robot's listed link pairs:
// distanceLineSegToLineSeg - r_a - r_b (xArm6::computeCapsulesDistance,
// xArm6.cpp:417-424).  The skeleton of each lane is kept in LDS, one column
// per lane.
template <typename T>
__global__ __launch_bounds__(128) void k_self(RobotDev rb, QueryParams qp, const float* __restrict__ q, int64_t n,
                                              float* dist, double* dist64, float* min_out, uint8_t* free_out) {
    extern __shared__ __align__(16) unsigned char smem[];
    T* sk = reinterpret_cast<T*>(smem);
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int L = rb.num_links;
    const int B = blockDim.x, t = threadIdx.x;
    const float* qc = q + c * rb.num_dofs;
    const Params<T> p{(T)qp.eq, (T)qp.maxd};
    FKStream fk;
    fk_init(fk);
    for (int k = 0; k <= L; ++k) {
        const Vec3<T> P = skel_point<T>(fk, rb, qc, k);
        sk[(3 * k + 0) * B + t] = P.x;
        sk[(3 * k + 1) * B + t] = P.y;
        sk[(3 * k + 2) * B + t] = P.z;
    }
    auto pt = [&](int k) { return Vec3<T>{sk[(3 * k) * B + t], sk[(3 * k + 1) * B + t], sk[(3 * k + 2) * B + t]}; };
    // smallest pair distance in std::min order from MAX_DISTANCE (a NaN pair
    // distance never replaces the running minimum); colliding = some pair < 0
    T lo = p.maxd;
    for (int m = 0; m < rb.num_pairs; ++m) {
        const int a = rb.pairs[2 * m], b = rb.pairs[2 * m + 1];
        const T sd = seg_seg<T, false>(pt(a), pt(a + 1), pt(b), pt(b + 1), p, nullptr);
        const T v = sd - (T)rb.radius[a] - (T)rb.radius[b];
        if (dist) dist[c * rb.num_pairs + m] = (float)v;
        if (dist64) dist64[c * rb.num_pairs + m] = (double)v;
        lo = smin(lo, v);
    }
    if (min_out) min_out[c] = (float)lo;
    if (free_out) free_out[c] = lo < 0 ? 0 : 1;
}

// ---------------------------------------------------------------- enclosing radii
// R [n][dofs][dofs+1] column by column (radii_column): the skeleton and the
// running column live in LDS, one slice per lane.
__global__ __launch_bounds__(128) void k_radii(RobotDev rb, const float* __restrict__ q, int64_t n, float* R) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int D = rb.num_dofs;
    const int B = blockDim.x, t = threadIdx.x;
    float* sk = reinterpret_cast<float*>(smem);
    float* ac = sk + (size_t)3 * (D + 1) * B;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const float* qc = q + c * D;
    FKStream fk;
    fk_init(fk);
    for (int k = 0; k <= D; ++k) {
        const Vec3<float> P = skel_point<float>(fk, rb, qc, k);
        sk[(3 * k + 0) * B + t] = P.x;
        sk[(3 * k + 1) * B + t] = P.y;
        sk[(3 * k + 2) * B + t] = P.z;
    }
    auto pt = [&](int k) { return Vec3<float>{sk[(3 * k) * B + t], sk[(3 * k + 1) * B + t], sk[(3 * k + 2) * B + t]}; };
    auto acc = [&](int i) -> float& { return ac[i * B + t]; };
    float* out = R + c * (int64_t)D * (D + 1);
    for (int i = 0; i < D; ++i) {
        acc(i) = 0.0f;
        out[i * (D + 1)] = 0.0f;
    }
    for (int j = 1; j <= D; ++j) {
        radii_column(rb, pt, j, acc);
        for (int i = 0; i < D; ++i) out[i * (D + 1) + j] = acc(i);
    }
}

// ---------------------------------------------------------------- moving obstacles
// Positions at a tick: pos = pos0 + vel * (tick * tick_seconds), in float,
// without contraction (a host can reproduce them bit for bit), then FCL's
// world AABB as the engine computes it at set-up: identity rotation
// (-h) + pos, h + pos; any other rotation pos -+ rad.  Oriented-box rows get
// the new centre.  Static boxes keep pos0.
__global__ void k_env_motion(MotionDev m, int64_t tick, float tick_seconds, float* aabb, float* obb, float* cull) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m.num_obj) return;
    const float t = (float)tick * tick_seconds;
    float p[3];
    for (int k = 0; k < 3; ++k) p[k] = m.pos0[3 * j + k] + m.vel[3 * j + k] * t;
    const float* hr = m.half + 3 * j;
    for (int k = 0; k < 3; ++k) {
        if (m.ident[j]) {
            aabb[6 * j + k] = -hr[k] + p[k];
            aabb[6 * j + 3 + k] = hr[k] + p[k];
        } else {
            aabb[6 * j + k] = p[k] - m.rad[j];
            aabb[6 * j + 3 + k] = p[k] + m.rad[j];
        }
    }
    for (int k = 0; k < 3; ++k) obb[15 * j + k] = p[k];
    for (int k = 3; k < 15; ++k) obb[15 * j + k] = m.obb0[15 * j + k];
    for (int k = 0; k < 3; ++k) {
        cull[6 * j + k] = p[k] - m.ext[3 * j + k];
        cull[6 * j + 3 + k] = p[k] + m.ext[3 * j + k];
    }
}

// Environment::updateEnvironment (Environment.cpp:194-377) for the
// deterministic motion laws, one lane per box, all steps in one launch, in
// float as fcl::Vector3f computes them (fixed-size norms and dots sum as
// x + (y + z)); boxes without the dynamic label keep their place.
//   straight (:225-238): x -= delta_time * max_vel
//   circular (:241-257): velocity tangent to the circle about the z axis
//   two tunnels (:260-287): back and forth 0.2 m legs, along x above y = 0
//     and along y below
//   light directions (:330-376): mirror reflection off the ground, the
//     workspace sphere and the keep-out zones around the robot base
// Group boxes of the broad-phase walk after the obstacles moved: thread g
// takes the union of the cull boxes of entries g * kGroupSize .. of order
// (the rule of group_bounds on the host).
__global__ void k_group_bounds(const float* __restrict__ cull, const int32_t* __restrict__ order, int M,
                               float* __restrict__ gcull) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= num_groups(M)) return;
    float lo[3] = {HUGE_VALF, HUGE_VALF, HUGE_VALF}, hi[3] = {-HUGE_VALF, -HUGE_VALF, -HUGE_VALF};
    bool finite = true;
    const int m1 = min(M, (g + 1) * kGroupSize);
    for (int m = g * kGroupSize; m < m1; ++m) {
        const float* b = cull + 6 * order[m];
        for (int k = 0; k < 3; ++k) {
            finite = finite && isfinite(b[k]) && isfinite(b[3 + k]);
            lo[k] = b[k] < lo[k] ? b[k] : lo[k];
            hi[k] = hi[k] < b[3 + k] ? b[3 + k] : hi[k];
        }
    }
    for (int k = 0; k < 3; ++k) {
        gcull[6 * g + k] = finite ? lo[k] : -HUGE_VALF;
        gcull[6 * g + 3 + k] = finite ? hi[k] : HUGE_VALF;
    }
}

__global__ void k_env_step(EnvStepDev m, float dt, int32_t steps, float* aabb, float* obb, float* cull) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m.num_obj) return;
    float p[3] = {m.pos[3 * j], m.pos[3 * j + 1], m.pos[3 * j + 2]};
    float v[3] = {m.vel[3 * j], m.vel[3 * j + 1], m.vel[3 * j + 2]};
    if (m.dynamic[j]) {
        const float vmax = m.max_vel[j];
        float plen = m.path_len[j];
        int32_t sg = m.sign[j];
        auto norm3 = [](float x, float y, float z) { return sqrtf(x * x + (y * y + z * z)); };
        for (int s = 0; s < steps; ++s) {
            if (m.law == 0) {
                p[0] = p[0] - dt * vmax;
            } else if (m.law == 1) {
                const float phi = atan2f(p[1], p[0]);
                v[0] = -sinf(phi) * vmax;
                v[1] = cosf(phi) * vmax;
                v[2] = 0.0f * vmax;
                for (int k = 0; k < 3; ++k) p[k] = p[k] + v[k] * dt;
            } else if (m.law == 2) {
                if (plen > 0.2f) {
                    sg = -sg;
                    plen = -0.2f;
                }
                const float sv = (float)sg;
                const int axis = p[1] > 0.0f ? 0 : 1;
                for (int k = 0; k < 3; ++k) v[k] = ((k == axis) ? 1.0f : 0.0f) * sv * vmax;
                plen = plen + norm3(v[0], v[1], v[2]) * dt;
                for (int k = 0; k < 3; ++k) p[k] = p[k] + v[k] * dt;
            } else {
                const float sp = norm3(v[0], v[1], v[2]);
                const float tol = max_std(sp / m.robot_max_vel, m.base_radius);
                float pn[3];
                for (int k = 0; k < 3; ++k) pn[k] = p[k] + v[k] * dt;
                float n[3];
                bool change = true;
                const float dws = norm3(pn[0] - m.ws[0], pn[1] - m.ws[1], pn[2] - m.ws[2]);
                if (pn[2] < 0.0f) {
                    n[0] = 0.0f; n[1] = 0.0f; n[2] = 1.0f;
                } else if (dws > m.ws_radius) {
                    n[0] = -pn[0]; n[1] = -pn[1]; n[2] = -(pn[2] - m.ws[2]);
                } else if (sqrtf(pn[0] * pn[0] + pn[1] * pn[1]) < tol && pn[2] < m.ws[2]) {
                    n[0] = pn[0]; n[1] = pn[1]; n[2] = 0.0f;
                } else if (dws < tol) {
                    n[0] = pn[0]; n[1] = pn[1]; n[2] = pn[2] - m.ws[2];
                } else {
                    for (int k = 0; k < 3; ++k) p[k] = pn[k];
                    change = false;
                }
                if (change) {
                    const float dd = (pn[0] - p[0]) * n[0] + ((pn[1] - p[1]) * n[1] + (pn[2] - p[2]) * n[2]);
                    const float t = dd / (n[0] * n[0] + (n[1] * n[1] + n[2] * n[2]));
                    for (int k = 0; k < 3; ++k) p[k] = 2.0f * pn[k] - p[k] - 2.0f * t * n[k];
                    for (int k = 0; k < 3; ++k) v[k] = (p[k] - pn[k]) / dt;
                }
            }
        }
        for (int k = 0; k < 3; ++k) {
            m.pos[3 * j + k] = p[k];
            m.vel[3 * j + k] = v[k];
        }
        m.path_len[j] = plen;
        m.sign[j] = sg;
    }
    const float* hr = m.half + 3 * j;
    for (int k = 0; k < 3; ++k) {
        if (m.ident[j]) {
            aabb[6 * j + k] = -hr[k] + p[k];
            aabb[6 * j + 3 + k] = hr[k] + p[k];
        } else {
            aabb[6 * j + k] = p[k] - m.rad[j];
            aabb[6 * j + 3 + k] = p[k] + m.rad[j];
        }
        obb[15 * j + k] = p[k];
        if (cull) {
            cull[6 * j + k] = p[k] - m.ext[3 * j + k];
            cull[6 * j + 3 + k] = p[k] + m.ext[3 * j + k];
        }
    }
}

inline dim3 grid_for(int64_t n, int block) { return dim3((unsigned)((n + block - 1) / block)); }

}  // namespace

hipError_t launch_distance_ref32(const RobotDev& rb, const EnvDev& env, const QueryParams& qp, const float* q,
                                 int64_t n, const DistOut& out, hipStream_t st, const int32_t* perm) {
    if (n <= 0) return hipSuccess;
    const size_t sm = staged_bytes<float>(env);
    const dim3 g = grid_for(n, kBlock), b(kBlock);
    if (env.use_obb) {
        if (out.nearest) hipLaunchKernelGGL((k_distance32<true, true>), g, b, sm, st, rb, env, qp, q, n, out, perm);
        else hipLaunchKernelGGL((k_distance32<false, true>), g, b, sm, st, rb, env, qp, q, n, out, perm);
    } else {
        if (out.nearest) hipLaunchKernelGGL((k_distance32<true, false>), g, b, sm, st, rb, env, qp, q, n, out, perm);
        else hipLaunchKernelGGL((k_distance32<false, false>), g, b, sm, st, rb, env, qp, q, n, out, perm);
    }
    return hipGetLastError();
}

hipError_t launch_distance_fp64(const RobotDev& rb, const EnvDev& env, const QueryParams& qp, const float* q,
                                int64_t n, const DistOut& out, hipStream_t st, const int32_t* perm) {
    if (n <= 0) return hipSuccess;
    const size_t sm = staged_bytes<double>(env);
    const dim3 g = grid_for(n, kBlock), b(kBlock);
    if (env.use_obb) {
        if (out.nearest) hipLaunchKernelGGL((k_distance64<true, true>), g, b, sm, st, rb, env, qp, q, n, out, perm);
        else hipLaunchKernelGGL((k_distance64<false, true>), g, b, sm, st, rb, env, qp, q, n, out, perm);
    } else {
        if (out.nearest) hipLaunchKernelGGL((k_distance64<true, false>), g, b, sm, st, rb, env, qp, q, n, out, perm);
        else hipLaunchKernelGGL((k_distance64<false, false>), g, b, sm, st, rb, env, qp, q, n, out, perm);
    }
    return hipGetLastError();
}

hipError_t launch_distance_ref32_waves(const RobotDev& rb, const EnvDev& env, const QueryParams& qp, const float* q,
                                       int64_t n, const DistOut& out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const unsigned grid = (unsigned)std::min<int64_t>((n + kRefineWaves - 1) / kRefineWaves, 65536);
    const size_t sm = staged_bytes<float>(env);
    const dim3 b(64 * kRefineWaves);
    const int32_t* no_idx = nullptr;
    const uint32_t* no_count = nullptr;
    uint32_t* no_out = nullptr;
    if (!out.nearest && refine_pairs_fit(rb, env)) {
        const size_t smp = pairs_row_off(env) + (size_t)kRefineWaves * kPairsMax * sizeof(float);
        if (env.use_obb)
            hipLaunchKernelGGL((k_distance_refine_pairs<true>), dim3(grid), b, smp, st, rb, env, qp, q, out, no_idx,
                               no_count, n, no_out);
        else
            hipLaunchKernelGGL((k_distance_refine_pairs<false>), dim3(grid), b, smp, st, rb, env, qp, q, out, no_idx,
                               no_count, n, no_out);
        return hipGetLastError();
    }
    if (env.use_obb) {
        if (out.nearest)
            hipLaunchKernelGGL((k_distance_refine<true, true>), dim3(grid), b, sm, st, rb, env, qp, q, out, no_idx,
                               no_count, n, no_out);
        else
            hipLaunchKernelGGL((k_distance_refine<true, false>), dim3(grid), b, sm, st, rb, env, qp, q, out, no_idx,
                               no_count, n, no_out);
    } else {
        if (out.nearest)
            hipLaunchKernelGGL((k_distance_refine<false, true>), dim3(grid), b, sm, st, rb, env, qp, q, out, no_idx,
                               no_count, n, no_out);
        else
            hipLaunchKernelGGL((k_distance_refine<false, false>), dim3(grid), b, sm, st, rb, env, qp, q, out, no_idx,
                               no_count, n, no_out);
    }
    return hipGetLastError();
}

hipError_t launch_distance_refine(const RobotDev& rb, const EnvDev& env, const QueryParams& qp, const float* q,
                                  int64_t n, const DistOut& out, const int32_t* idx, const uint32_t* count,
                                  hipStream_t st, uint32_t* count_out) {
    if (n <= 0) return hipSuccess;
    QueryParams rq = qp;   // the entries' form, as the bulk pass's launch chose it
    rq.flag_link = flags_carry_link(n) ? 1 : 0;
    // enough waves to cover a typical flagged set in one round; the kernel
    // loops when more configurations are flagged
    const unsigned grid = (unsigned)std::min<int64_t>((n + kRefineWaves - 1) / kRefineWaves, 2048);
    const size_t sm = staged_bytes<float>(env);
    const dim3 b(64 * kRefineWaves);
    if (refine_pairs_fit(rb, env)) {
        const size_t smp = pairs_row_off(env) + (size_t)kRefineWaves * kPairsMax * sizeof(float);
        if (env.use_obb)
            hipLaunchKernelGGL((k_distance_refine_pairs<true>), dim3(grid), b, smp, st, rb, env, rq, q, out, idx, count,
                               (int64_t)0, count_out);
        else
            hipLaunchKernelGGL((k_distance_refine_pairs<false>), dim3(grid), b, smp, st, rb, env, rq, q, out, idx,
                               count, (int64_t)0, count_out);
        return hipGetLastError();
    }
    if (env.use_obb)
        hipLaunchKernelGGL((k_distance_refine<true, false>), dim3(grid), b, sm, st, rb, env, rq, q, out, idx, count,
                           (int64_t)0, count_out);
    else
        hipLaunchKernelGGL((k_distance_refine<false, false>), dim3(grid), b, sm, st, rb, env, rq, q, out, idx, count,
                           (int64_t)0, count_out);
    return hipGetLastError();
}

hipError_t launch_is_valid(const RobotDev& rb, const EnvDev& env, const QueryParams& qp, const float* q, int64_t n,
                           uint8_t* valid, int fp64, hipStream_t st, const int32_t* perm) {
    if (n <= 0) return hipSuccess;
    const dim3 g = grid_for(n, kBlock), b(kBlock);
    if (fp64) {
        const size_t sm = staged_bytes<double>(env);
        if (env.use_obb) hipLaunchKernelGGL((k_valid<double, true>), g, b, sm, st, rb, env, qp, q, n, valid, perm);
        else hipLaunchKernelGGL((k_valid<double, false>), g, b, sm, st, rb, env, qp, q, n, valid, perm);
    } else {
        const size_t sm = staged_bytes<float>(env);
        if (env.use_obb) hipLaunchKernelGGL((k_valid<float, true>), g, b, sm, st, rb, env, qp, q, n, valid, perm);
        else hipLaunchKernelGGL((k_valid<float, false>), g, b, sm, st, rb, env, qp, q, n, valid, perm);
    }
    return hipGetLastError();
}

hipError_t launch_underestimation(const RobotDev& rb, const EnvDev& env, const QueryParams& qp, const float* q,
                                  int64_t n, const float* np_ref, int64_t np_stride, float* d, float* profile,
                                  uint8_t* stored, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_under, grid_for(n, kBlock), dim3(kBlock), 0, st, rb, env, qp, q, n, np_ref, np_stride, d,
                       profile, stored, nullptr, 0, nullptr, 0.0f);
    return hipGetLastError();
}

hipError_t launch_underestimation_moving(const RobotDev& rb, const EnvDev& env, const QueryParams& qp, const float* q,
                                         int64_t n, const float* np_ref, int64_t np_stride, const float* max_vel,
                                         float delta_t, float* d, float* profile, uint8_t* stored, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_under, grid_for(n, kBlock), dim3(kBlock), 0, st, rb, env, qp, q, n, np_ref, np_stride, d,
                       profile, stored, nullptr, 0, max_vel, delta_t);
    return hipGetLastError();
}

hipError_t launch_underestimation_seg(const RobotDev& rb, const EnvDev& env, const QueryParams& qp, const float* q,
                                      int64_t n, const float* np_sets, const int64_t* seg_off, int nseg, float* d,
                                      float* profile, uint8_t* stored, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_under, grid_for(n, kBlock), dim3(kBlock), 0, st, rb, env, qp, q, n, np_sets, (int64_t)0, d,
                       profile, stored, seg_off, nseg, nullptr, 0.0f);
    return hipGetLastError();
}

hipError_t launch_enclosing_radii(const RobotDev& rb, const float* q, int64_t n, float* R, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    // per-lane slices of the skeleton and the running column; the block
    // shrinks for long chains so that they stay within 64 KB of LDS
    const size_t per = (size_t)(3 * (rb.num_dofs + 1) + rb.num_dofs) * sizeof(float);
    int B = 128;
    while (B > 32 && per * B > 64 * 1024) B /= 2;
    const size_t sm = per * B;
    hipLaunchKernelGGL(k_radii, grid_for(n, B), dim3(B), sm, st, rb, q, n, R);
    return hipGetLastError();
}

hipError_t launch_group_bounds(const float* cull, const int32_t* order, int M, float* gcull, hipStream_t st) {
    const int ng = num_groups(M);
    if (ng <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_group_bounds, dim3((ng + 63) / 64), dim3(64), 0, st, cull, order, M, gcull);
    return hipGetLastError();
}

hipError_t launch_env_step(const EnvStepDev& m, float delta_time, int32_t steps, float* aabb, float* obb, float* cull,
                           hipStream_t st) {
    if (m.num_obj <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_env_step, dim3((m.num_obj + 255) / 256), dim3(256), 0, st, m, delta_time, steps, aabb, obb,
                       cull);
    return hipGetLastError();
}

hipError_t launch_env_motion(const MotionDev& m, int64_t tick, float tick_seconds, float* aabb, float* obb,
                             float* cull, hipStream_t st) {
    if (m.num_obj <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_env_motion, dim3((m.num_obj + 255) / 256), dim3(256), 0, st, m, tick, tick_seconds, aabb,
                       obb, cull);
    return hipGetLastError();
}

hipError_t launch_self_distance(const RobotDev& rb, const QueryParams& qp, const float* q, int64_t n, float* dist,
                                double* dist64, float* min_out, uint8_t* free_out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int B = 128;
    if (dist64) {
        const size_t sm = (size_t)3 * (rb.num_links + 1) * B * sizeof(double);
        hipLaunchKernelGGL((k_self<double>), grid_for(n, B), dim3(B), sm, st, rb, qp, q, n, dist, dist64, min_out,
                           free_out);
    } else {
        const size_t sm = (size_t)3 * (rb.num_links + 1) * B * sizeof(float);
        hipLaunchKernelGGL((k_self<float>), grid_for(n, B), dim3(B), sm, st, rb, qp, q, n, dist, dist64, min_out,
                           free_out);
    }
    return hipGetLastError();
}

}  // namespace rvs
