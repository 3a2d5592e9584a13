// FP32 bulk pass of RVS_PREC_FP32_REFINE (gfx950).
//
// The reference-exact kernel's early-exit loop over (link, obstacle) pairs in
// relaxed single precision: kinematics in float with the hardware sine and
// cosine, FMA contraction (this translation unit is compiled with
// -ffp-contract=fast), hardware reciprocal and square root, and the
// capsule-box distance by flat_capsule_box, a branch-free form of the
// reference's case tree (rvs_geometry.hpp).  Its results differ from the
// reference arithmetic by up to about 1e-6 m, so every pair it evaluates is
// screened and the configuration is sent to the exact refinement pass when
// any pair is uncertain:
//   end point within band of the box surface  the inside test could differ
//   raw distance < EQ + band      the reference snaps distances below
//                                 EQUALITY_THRESHOLD to 0 (then -r), touches
//                                 a face, or takes its NaN path
//   |d - r| < band                the verdict sign could differ
//   |d - tol| < band              the min_dist_tol filter could differ
//   near-axis-parallel, raw < 0.01  an ill-conditioned segment solve close by
//   zero-length segment           the reference's 0/0 projections
//   non-finite skeleton           NaN / Inf configurations
// A crossing is reported only when the segment meets the box shrunk by band
// on every side, so it is a crossing whatever the rounding.  Configurations
// outside all of these get verdicts identical to the reference (a sign
// decided with margin >= band cannot flip under a 1e-6 m error), and the same
// early-exit pair, so the same partial profile up to rounding.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>

#include "rvs_fk.hpp"
#include "rvs_geometry.hpp"
#include "rvs_internal.hpp"
#include "rvs_stage.hpp"

namespace rvs {

namespace {

constexpr int kBlock = 256;

struct Mat3f {
    float m0, m1, m2, m3, m4, m5, m6, m7, m8;
};

__device__ inline Mat3f mulf(const Mat3f& a, const Mat3f& b) {
    return Mat3f{a.m0 * b.m0 + a.m1 * b.m3 + a.m2 * b.m6, a.m0 * b.m1 + a.m1 * b.m4 + a.m2 * b.m7,
                 a.m0 * b.m2 + a.m1 * b.m5 + a.m2 * b.m8, a.m3 * b.m0 + a.m4 * b.m3 + a.m5 * b.m6,
                 a.m3 * b.m1 + a.m4 * b.m4 + a.m5 * b.m7, a.m3 * b.m2 + a.m4 * b.m5 + a.m5 * b.m8,
                 a.m6 * b.m0 + a.m7 * b.m3 + a.m8 * b.m6, a.m6 * b.m1 + a.m7 * b.m4 + a.m8 * b.m7,
                 a.m6 * b.m2 + a.m7 * b.m5 + a.m8 * b.m8};
}

// Joint values (radians) the FP32 kinematics evaluates with the hardware
// sine and cosine.  Their input, a / 2 pi, is rounded to float, so the
// angle error grows as |a| 6e-8 rad: within 4 pi (twice the +-2 pi joint
// limits of the reference's robots) it stays below 7.5e-7 rad.
constexpr float kTrigRange = 12.566371f;

struct FKf {
    Mat3f M;
    float px, py, pz;
    int seg, qi;
};

// Joint k of the configuration is q[k * qstride].
__device__ __attribute__((always_inline)) inline void fkf_advance(FKf& s, const SegF* seg, const float* q, int upto,
                                                                 int qstride) {
    while (s.seg <= upto) {
        const SegF& g = seg[s.seg];
        Mat3f SM;
        if (g.rev) {
            // the hardware sine and cosine (v_sin_f32 / v_cos_f32 of a / 2 pi,
            // errors of order 1e-6, far inside the screening band); a joint
            // value beyond kTrigRange gives NaN, which flags the configuration
            // for the exact pass
            const float a = q[(s.qi++) * qstride];
            const bool in_range = fabsf(a) <= kTrigRange;
            const float st = in_range ? __sinf(a) : NAN, ct = in_range ? __cosf(a) : NAN;
            const float vt = 1.0f - ct;
            // Rot(axis, q) R = R + sin q K R + (1 - cos q) K^2 R
            SM = Mat3f{g.R[0] + st * g.KR[0] + vt * g.K2R[0], g.R[1] + st * g.KR[1] + vt * g.K2R[1],
                       g.R[2] + st * g.KR[2] + vt * g.K2R[2], g.R[3] + st * g.KR[3] + vt * g.K2R[3],
                       g.R[4] + st * g.KR[4] + vt * g.K2R[4], g.R[5] + st * g.KR[5] + vt * g.K2R[5],
                       g.R[6] + st * g.KR[6] + vt * g.K2R[6], g.R[7] + st * g.KR[7] + vt * g.K2R[7],
                       g.R[8] + st * g.KR[8] + vt * g.K2R[8]};
        } else {
            SM = Mat3f{g.R[0], g.R[1], g.R[2], g.R[3], g.R[4], g.R[5], g.R[6], g.R[7], g.R[8]};
        }
        const float o0 = g.o[0], o1 = g.o[1], o2 = g.o[2];
        const Mat3f& M = s.M;
        const float nx = M.m0 * o0 + M.m1 * o1 + M.m2 * o2 + s.px;
        const float ny = M.m3 * o0 + M.m4 * o1 + M.m5 * o2 + s.py;
        const float nz = M.m6 * o0 + M.m7 * o1 + M.m8 * o2 + s.pz;
        s.M = mulf(M, SM);
        s.px = nx;
        s.py = ny;
        s.pz = nz;
        ++s.seg;
    }
}

__device__ __attribute__((always_inline)) inline Vec3<float> skelf_point(FKf& s, const RobotDev& rb, const float* q,
                                                                       int k, int qstride = 1) {
    const PointDev pt = rb.pts[k];
    if (pt.frame < 0) return Vec3<float>{0.0f, 0.0f, 0.0f};
    fkf_advance(s, rb.segf, q, pt.frame, qstride);
    const float w0 = pt.axis == 0 ? 1.0f : 0.0f, w1 = pt.axis == 1 ? 1.0f : 0.0f, w2 = pt.axis == 2 ? 1.0f : 0.0f;
    const Mat3f& M = s.M;
    const float ax = M.m0 * w0 + M.m1 * w1 + M.m2 * w2;
    const float ay = M.m3 * w0 + M.m4 * w1 + M.m5 * w2;
    const float az = M.m6 * w0 + M.m7 * w1 + M.m8 * w2;
    Vec3<float> P{s.px, s.py, s.pz};
    if (pt.flags & 1) {
        const float pre = (float)pt.pre;
        P.x += pre * ax; P.y += pre * ay; P.z += pre * az;
    }
    if (pt.flags & 2) {
        const float post = (float)pt.post;
        P.x -= post * ax; P.y -= post * ay; P.z -= post * az;
    }
    return P;
}

// Read-only tables indexed by wave-uniform values: read through the
// constant address space, so that they are fetched with scalar loads into
// SGPRs (the tables are not written while a kernel reads them).
template <typename T>
using const_ptr = const __attribute__((address_space(4))) T*;
template <typename T>
__device__ __attribute__((always_inline)) inline const_ptr<T> as_const(const T* p) {
    return (const_ptr<T>)(p);
}

// Broad-phase test of the bulk pass: the gap of cull_pair against a squared
// bound given (kept by the caller while its running minimum is unchanged),
// each axis's gap as max(lo - shi, slo - hi, 0), summed x, y, z with fused
// multiply-adds.  For finite end points it decides as cull_pair does; with a
// non-finite end point it may skip pairs cull_pair would not, but such a
// configuration is sent to the exact pass, which recomputes it whole.
template <class P>
__device__ __attribute__((always_inline)) inline bool cull_fast(P cb, const Vec3<float>& slo, const Vec3<float>& shi,
                                                                float bound2) {
    const float gx = fmaxf(fmaxf(cb[0] - shi.x, slo.x - cb[3]), 0.0f);
    const float gy = fmaxf(fmaxf(cb[1] - shi.y, slo.y - cb[4]), 0.0f);
    const float gz = fmaxf(fmaxf(cb[2] - shi.z, slo.z - cb[5]), 0.0f);
    return fmaf(gz, gz, fmaf(gy, gy, gx * gx)) > bound2;
}

// Square of cull_bound(prof, r, eq).
__device__ __attribute__((always_inline)) inline float cull_bound2(float prof, float r, float eq) {
    const float b = fmaxf(prof + r, eq) + kCullMargin;
    return b * b;
}

// One (link, obstacle) pair of the bulk pass: the filtered distance by the
// case-tree-free segment-box evaluation (flat_capsule_box, in the box frame
// for oriented boxes; the frame change adds a few ulp of the lever arm, far
// inside the band) and the screening rules that send the configuration to
// the exact pass: those of flat_capsule_box, which flag every pair the
// reference's tree returns NaN for (a segment in a face plane) and every
// marginal contact decision, and a value within band of the obstacle's
// tolerance.  A fixed sequence of operations keeps a wave's lanes together
// whatever pairs they evaluate (workload 2, one lane per configuration:
// 0.214 ms per 1,048,576 configurations against 0.266 ms for the tree).
template <bool OBB>
__device__ __attribute__((always_inline)) inline float flat_pair(const Staged<float>& s, int j, const Vec3<float>& P0,
                                                                 const Vec3<float>& P1, float r, const Params<float>& p,
                                                                 float band, bool& uncertain) {
    const Box<float> bx = s.box[j];
    float dt;
    if (OBB) {
        const float* x = s.xf + 12 * j;
        dt = flat_capsule_box<FastMath>(to_box(x, P0), to_box(x, P1), r, bx, p, band, uncertain);
    } else {
        dt = flat_capsule_box<FastMath>(P0, P1, r, bx, p, band, uncertain);
    }
    const float tol = s.tol[j];
    if (tol < p.maxd && fabsf(dt - tol) < band) uncertain = true;
    if (dt > tol) dt = p.maxd;
    return dt;
}

// Smallest filtered distance of one link over the obstacles, with the
// reference's early exit: the first pair in obstacle order whose distance is
// <= 0 ends the link (hit) and its distance is the link's value.
//
// GROUPED (the walk of every scene with group tables): the pair that held
// the previous link's minimum (warm, -1 = none) is evaluated first, since
// neighbouring links are usually nearest to the same obstacle; then the
// obstacles are walked in the scene's Morton order, kGroupSize at a time, a
// group's union box tested before its members.  With a running minimum near
// its final value most groups fail the gap test at once.  The minimum of the
// evaluated pairs is the link's minimum (a skipped pair's distance exceeds
// the running minimum).  A hit found this way need not be the first in
// obstacle order, so the pairs before it are searched for an earlier one
// with the bound of a running minimum of 0, which never skips a pair whose
// exact distance can be <= 0: the value is the in-order one.  On return warm
// holds this link's nearest obstacle.  Otherwise (no group tables) the
// obstacles are walked in order.
//
// Called by the whole wave; lanes with active false evaluate nothing.  Both
// bulk kernels walk a configuration's pairs through this function, so they
// evaluate the same pairs, flag the same configurations for the exact pass
// and return the same bits.
template <bool OBB, bool GROUPED>
__device__ __attribute__((always_inline)) inline float link_min(const Staged<float>& s, const EnvDev& env, int M,
                                                                bool gnd_skip, const Vec3<float>& P0,
                                                                const Vec3<float>& P1, float r, const Params<float>& p,
                                                                float band, bool active, int& warm, bool& hit,
                                                                bool& uncertain) {
    // the tables are read with wave-uniform indices (scalar loads)
    const const_ptr<float> cull = as_const(env.cull);
    const const_ptr<float> gcull = as_const(env.gcull);
    const const_ptr<int32_t> gorder = as_const(env.gorder);
    Vec3<float> slo, shi;
    seg_bounds(P0, P1, slo, shi);
    float prof = p.maxd;
    float bb = cull_bound2(prof, r, p.eq);   // squared bound of the running minimum
    hit = false;
    if (!GROUPED) {   // the plain in-order walk
        bool go = active;
        for (int j = 0; j < M; ++j) {
            if (go && !cull_fast(cull + 6 * j, slo, shi, bb)) {
                const float dt = (s.gnd[j] && gnd_skip) ? p.maxd : flat_pair<OBB>(s, j, P0, P1, r, p, band, uncertain);
                prof = smin(prof, dt);
                bb = cull_bound2(prof, r, p.eq);
                if (prof <= 0) {
                    hit = true;
                    go = false;
                }
            }
            if (__ballot(go) == 0) break;
        }
        return prof;
    }
    int best = -1;
    int jh = -1;         // a hit found out of obstacle order, and its value
    float vh = 0.0f;
    const int w = active ? warm : -1;
    if (w >= 0) {
        const float dt = (s.gnd[w] && gnd_skip) ? p.maxd : flat_pair<OBB>(s, w, P0, P1, r, p, band, uncertain);
        if (dt <= 0) {
            jh = w;
            vh = dt;
        } else if (dt < prof) {
            prof = dt;
            best = w;
            bb = cull_bound2(prof, r, p.eq);
        }
    }
    bool go = active && jh < 0;
    const int ng = num_groups(M);
    for (int g = 0; g < ng; ++g) {
        if (go && !cull_fast(gcull + 6 * g, slo, shi, bb)) {
            const int m1 = min(M, (g + 1) * kGroupSize);
            for (int m = g * kGroupSize; m < m1; ++m) {
                const int j = gorder[m];
                if (go && j != w && !cull_fast(cull + 6 * j, slo, shi, bb)) {
                    const float dt =
                        (s.gnd[j] && gnd_skip) ? p.maxd : flat_pair<OBB>(s, j, P0, P1, r, p, band, uncertain);
                    if (dt < prof) best = j;
                    prof = smin(prof, dt);
                    bb = cull_bound2(prof, r, p.eq);
                    if (prof <= 0) {
                        jh = j;
                        vh = prof;
                        go = false;
                    }
                }
            }
        }
        if (__ballot(go) == 0) break;
    }
    if (jh >= 0) {   // the first hit in obstacle order
        float first = vh;
        const float bb0 = cull_bound2(0.0f, r, p.eq);
        for (int j = 0; j < jh; ++j) {
            if (cull_fast(cull + 6 * j, slo, shi, bb0)) continue;
            const float dj = (s.gnd[j] && gnd_skip) ? p.maxd : flat_pair<OBB>(s, j, P0, P1, r, p, band, uncertain);
            if (dj <= 0) {
                first = dj;
                break;
            }
        }
        prof = first;
        hit = true;
    }
    warm = best;
    return prof;
}

// Wave-aggregated append of configuration c for every lane with flag set
// (called by the whole wave).
__device__ __attribute__((always_inline)) inline void append_flagged(bool flag, int64_t c, int32_t* idx,
                                                                     uint32_t* count) {
    const uint64_t mask = __ballot(flag);
    if (mask) {
        const int lane = __lane_id();
        const int leader = __ffsll((unsigned long long)mask) - 1;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(mask));
        base = __shfl(base, leader, 64);
        if (flag) {
            const uint64_t below = (lane == 0) ? 0ull : (mask & (~0ull >> (64 - lane)));
            idx[base + __popcll(below)] = (int32_t)c;
        }
    }
}

// One configuration of the bulk pass for this lane (configuration c; active
// false: the lane has none and only takes part in the wave's collective
// steps).  With ROWS (a walk order is given) the lane copies the
// configuration's joint values into its LDS slice in one burst before the
// walk and collects its profile row there, writing it out in one burst at the
// end: the rows of a wave's lanes are scattered in memory, and reading or
// writing them a joint or a link at a time, far apart in time, costs a cache
// line each.
template <bool OBB, bool ROWS, bool GROUPED>
__device__ __attribute__((always_inline)) inline void fast_one(const Staged<float>& s, const RobotDev& rb,
                                                               const EnvDev& env, const QueryParams& qp,
                                                               const float* __restrict__ q, int64_t c, bool active,
                                                               const DistOut& out, int32_t* idx, uint32_t* count,
                                                               float* qrow, float* prow, int B) {
    const int M = env.num_obj;
    const int L = rb.num_links;
    const int D = rb.num_dofs;
    const float* qc = q + (active ? c : 0) * D;
    int qs = 1;
    if (ROWS) {
        if (active)
            for (int k = 0; k < D; ++k) qrow[k * B] = qc[k];
        qc = qrow;
        qs = B;
    }
    const Params<float> p{qp.eq, qp.maxd};
    const float band = qp.band;

    FKf fk;
    fk.M = Mat3f{1.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f, 0.0f, 1.0f};
    fk.px = fk.py = fk.pz = 0.0f;
    fk.seg = 0;
    fk.qi = 0;
    Vec3<float> P0 = skelf_point(fk, rb, qc, 0, qs);
    float d = p.maxd;
    bool hit = false, uncertain = false;
    int i = 0, warm = -1, uk = -1;   // uk: the first link with an uncertain decision
    for (; i < L; ++i) {
        const Vec3<float> P1 = skelf_point(fk, rb, qc, i + 1, qs);
        if (!isfinite(P1.x + P1.y + P1.z)) uncertain = true;
        const float r = rb.radius[i];
        const float prof = link_min<OBB, GROUPED>(s, env, M, i < rb.ground_included, P0, P1, r, p, band, active, warm,
                                                  hit, uncertain);
        if (uncertain && uk < 0) uk = i;
        if (ROWS) prow[i * B] = prof;
        else if (active && out.profile) out.profile[c * L + i] = prof;
        if (hit) break;
        d = smin(d, prof);
        P0 = P1;
    }
    if (active) {
        if (out.profile) {
            if (ROWS) {
                for (int k = 0; k < L; ++k) out.profile[c * L + k] = k <= i ? prow[k * B] : 0.0f;
            } else {
                for (int k = i + 1; k < L; ++k) out.profile[c * L + k] = 0.0f;
            }
        }
        if (hit) d = 0.0f;
        if (out.d) out.d[c] = d;
        if (out.verdict) out.verdict[c] = hit ? 0 : 1;
    }
    append_flagged(active && uncertain, flag_entry(c, uk, qp.flag_link != 0), idx, count);
}

// One lane per configuration: lane g takes configuration perm[g] of the walk
// order (g itself without one).  At least 5 waves per SIMD (96 VGPRs, no
// scratch); the walk-order form for axis-aligned boxes takes 71 VGPRs and runs
// 7 waves per SIMD, faster than with its workgroups' LDS padded down to 6 or 5
// (workload 2: 0.182, 0.185 and 0.193 ms).
template <bool OBB, bool ROWS, bool GROUPED>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5))) void k_distance_fast(
    RobotDev rb, EnvDev env, QueryParams qp, const float* __restrict__ q, int64_t n, DistOut out, int32_t* idx,
    uint32_t* count, const int32_t* __restrict__ perm, size_t rows_off) {
    extern __shared__ __align__(16) unsigned char smem[];
    const Staged<float> s = stage_env<float, OBB>(env, smem);
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const int64_t c = perm ? (int64_t)perm[g] : g;
    const int B = blockDim.x;
    float* qrow = reinterpret_cast<float*>(smem + rows_off) + threadIdx.x;   // [D] at stride B
    float* prow = qrow + (size_t)rb.num_dofs * B;                          // [L] at stride B
    fast_one<OBB, ROWS, GROUPED>(s, rb, env, qp, q, c, true, out, idx, count, qrow, prow, B);
}

// The same with persistent waves: each wave takes 64-configuration chunks of
// the walk order from a global counter (next, expected zero) until none is
// left, so a workgroup stages the obstacles once and the launch ends with
// one chunk's imbalance instead of a partial last round of workgroups.
// Every wave leaves once the counter has passed n.  Held to 128 VGPRs (4
// waves per SIMD): at 96 the loop's extra state spills (76 bytes per lane).
template <bool OBB, bool ROWS, bool GROUPED>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_distance_fast_chunks(
    RobotDev rb, EnvDev env, QueryParams qp, const float* __restrict__ q, int64_t n, DistOut out, int32_t* idx,
    uint32_t* count, const int32_t* __restrict__ perm, size_t rows_off, uint32_t* next) {
    extern __shared__ __align__(16) unsigned char smem[];
    const Staged<float> s = stage_env<float, OBB>(env, smem);
    const int B = blockDim.x;
    const int lane = __lane_id();
    float* qrow = reinterpret_cast<float*>(smem + rows_off) + threadIdx.x;   // [D] at stride B
    float* prow = qrow + (size_t)rb.num_dofs * B;                          // [L] at stride B
    for (;;) {
        uint32_t chunk = 0;
        if (lane == 0) chunk = atomicAdd(next, 1u);
        chunk = __shfl(chunk, 0, 64);
        const int64_t g0 = (int64_t)chunk * 64;
        if (g0 >= n) break;   // uniform over the wave
        const int64_t g = g0 + lane;
        const bool active = g < n;
        const int64_t c = active ? (perm ? (int64_t)perm[g] : g) : 0;
        fast_one<OBB, ROWS, GROUPED>(s, rb, env, qp, q, c, active, out, idx, count, qrow, prow, B);
    }
}

// The same pass with lane refill.  The classic kernel keeps a lane on one
// configuration, so a wave runs as long as its slowest configuration while
// lanes whose configuration exited early (first collision) sit idle.  Here a
// persistent wave keeps every lane busy: the obstacle loop runs one link at a
// time for all lanes together (obstacle j is the same for every lane, so the
// LDS reads stay broadcasts), and after each link a lane whose configuration
// is finished takes the next one from a global counter (one atomic per wave
// and link).  Lanes therefore sit on different links of different
// configurations; per configuration the pair order, the early exit and every
// result are those of the classic kernel.
template <bool OBB, bool GROUPED>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_distance_fast_pool(
    RobotDev rb, EnvDev env, QueryParams qp, const float* __restrict__ q, int64_t n, DistOut out, int32_t* idx,
    uint32_t* count, uint32_t* next, const int32_t* __restrict__ perm) {
    extern __shared__ __align__(16) unsigned char smem[];
    const Staged<float> s = stage_env<float, OBB>(env, smem);
    const int M = env.num_obj;
    const int L = rb.num_links;
    const int D = rb.num_dofs;
    const Params<float> p{qp.eq, qp.maxd};
    const float band = qp.band;
    const int lane = __lane_id();

    int64_t c = -1;           // configuration of this lane, -1 = none
    bool drained = false;     // the counter has passed n
    int i = 0;                // current link
    int warm = -1;            // nearest obstacle of the previous link
    int uk = -1;              // the first link with an uncertain decision
    const float* qc = q;
    FKf fk;
    Vec3<float> P0{0.0f, 0.0f, 0.0f};
    float d = p.maxd;
    bool uncertain = false;
    for (;;) {
        // refill the lanes without a configuration
        const bool need = c < 0 && !drained;
        const uint64_t nm = __ballot(need);
        if (nm) {
            const int leader = __ffsll((unsigned long long)nm) - 1;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(next, (uint32_t)__popcll(nm));
            base = __shfl(base, leader, 64);
            if (need) {
                const uint64_t below = (lane == 0) ? 0ull : (nm & (~0ull >> (64 - lane)));
                const int64_t cand = (int64_t)base + __popcll(below);
                if (cand < n) {
                    c = perm ? (int64_t)perm[cand] : cand;
                    qc = q + c * D;
                    fk.M = Mat3f{1.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f, 0.0f, 1.0f};
                    fk.px = fk.py = fk.pz = 0.0f;
                    fk.seg = 0;
                    fk.qi = 0;
                    P0 = skelf_point(fk, rb, qc, 0);
                    i = 0;
                    warm = -1;
                    uk = -1;
                    d = p.maxd;
                    uncertain = false;
                } else {
                    drained = true;
                }
            }
        }
        const bool busy = c >= 0;
        if (__ballot(busy) == 0) break;
        Vec3<float> P1 = P0;
        float r = 0.0f;
        if (busy) {
            P1 = skelf_point(fk, rb, qc, i + 1);
            if (!isfinite(P1.x + P1.y + P1.z)) uncertain = true;
            r = rb.radius[i];
        }
        bool hit = false;
        const float prof = link_min<OBB, GROUPED>(s, env, M, i < rb.ground_included, P0, P1, r, p, band, busy, warm,
                                                  hit, uncertain);
        bool finished = false;
        if (busy) {
            if (uncertain && uk < 0) uk = i;
            if (out.profile) out.profile[c * L + i] = prof;
            if (!hit) d = smin(d, prof);
            if (hit || i == L - 1) {
                for (int k = i + 1; k < L; ++k)
                    if (out.profile) out.profile[c * L + k] = 0.0f;
                if (out.d) out.d[c] = hit ? 0.0f : d;
                if (out.verdict) out.verdict[c] = hit ? 0 : 1;
                finished = true;
            } else {
                ++i;
                P0 = P1;
            }
        }
        append_flagged(finished && uncertain, flag_entry(c, uk, qp.flag_link != 0), idx, count);
        if (finished) c = -1;
    }
}

// The grouped walk with lane refill and packed pair evaluations (tuning
// variant 3).  In the kernels above a lane evaluates a pair where the wave's
// walk meets it, so a wave whose lanes need different obstacles (lanes on
// different poses, as lane refill makes them) pays one evaluation per
// obstacle any lane needs.  Here, per link step (every lane on the current
// link of its own configuration, taken from a global counter as lanes free
// up):
//   1. every lane evaluates its warm pair (its previous link's nearest
//      obstacle), all lanes in one step;
//   2. the wave walks the groups in the scene's walk order together (a
//      group's box and members are read once for all lanes); each lane tests
//      them against its bound and queues the members that pass (an LDS ring
//      of kLaneQueue per lane);
//   3. whenever enough lanes hold queued pairs (and at the end), every lane
//      takes its next queued pair, tests it again against its bound of now
//      and evaluates it if it passes: one pair per lane per step, each its
//      own obstacle, with the flat evaluation (no case tree, so the lanes do
//      not split across branches);
//   4. a lane whose pair ends its link (distance <= 0) stops walking; the
//      first hit in obstacle order is then found among the obstacles before
//      it that the bound of a running minimum of 0 does not rule out, queued
//      and evaluated the same way, the smallest index with distance <= 0
//      winning.
// The walk evaluates exactly the pairs of link_min's walk in its order (the
// queueing bound is never tighter than the walk's, a member's box lies
// inside its group's, and every pair is tested again with exactly the bound
// the walk has when it reaches it); the search evaluates every candidate
// before the hit instead of stopping at the first, which can only add
// screening flags.  The link's value, the warm pair of the next link and the
// verdict are those of link_min.
constexpr int kLaneQueue = 16;         // pending pairs per lane (power of 2, > kGroupSize)
constexpr int kDrainLanes = 40;        // an evaluation step once this many lanes hold a pair

template <bool OBB>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(4))) void k_distance_fast_lanes(
    RobotDev rb, EnvDev env, QueryParams qp, const float* __restrict__ q, int64_t n, DistOut out, int32_t* idx,
    uint32_t* count, uint32_t* next, const int32_t* __restrict__ perm, size_t cull_off, size_t queue_off) {
    extern __shared__ __align__(16) unsigned char smem[];
    float* const lcull = reinterpret_cast<float*>(smem + cull_off);   // [M][6] cull boxes, for the queued pairs
    for (int k = threadIdx.x; k < 6 * env.num_obj; k += blockDim.x) lcull[k] = env.cull[k];
    const Staged<float> s = stage_env<float, OBB>(env, smem);   // ends with a barrier
    const int B = blockDim.x;
    int16_t* const ring = reinterpret_cast<int16_t*>(smem + queue_off) + threadIdx.x;   // slot k: ring[k * B]
    const int M = env.num_obj;
    const int L = rb.num_links;
    const int D = rb.num_dofs;
    const int ng = num_groups(M);
    const const_ptr<float> cull = as_const(env.cull);
    const const_ptr<float> gcull = as_const(env.gcull);
    const const_ptr<int32_t> gorder = as_const(env.gorder);
    const Params<float> p{qp.eq, qp.maxd};
    const float band = qp.band;
    const int lane = __lane_id();

    int64_t c = -1;          // configuration of this lane, -1 = none
    bool drained = false;    // the counter has passed n
    const float* qc = q;
    FKf fk;
    Vec3<float> P0{0.0f, 0.0f, 0.0f};
    float d = 0.0f;
    int i = 0, warm = -1, uk = -1;   // uk: the first link with an uncertain decision
    bool uncertain = false;
    for (;;) {
        // refill the lanes without a configuration
        const bool need = c < 0 && !drained;
        const uint64_t nm = __ballot(need);
        if (nm) {
            const int leader = __ffsll((unsigned long long)nm) - 1;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(next, (uint32_t)__popcll(nm));
            base = __shfl(base, leader, 64);
            if (need) {
                const uint64_t below = (lane == 0) ? 0ull : (nm & (~0ull >> (64 - lane)));
                const int64_t cand = (int64_t)base + __popcll(below);
                if (cand < n) {
                    c = perm ? (int64_t)perm[cand] : cand;
                    qc = q + c * D;
                    fk.M = Mat3f{1.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f, 0.0f, 1.0f};
                    fk.px = fk.py = fk.pz = 0.0f;
                    fk.seg = 0;
                    fk.qi = 0;
                    P0 = skelf_point(fk, rb, qc, 0);
                    i = 0;
                    warm = -1;
                    uk = -1;
                    d = p.maxd;
                    uncertain = false;
                } else {
                    drained = true;
                }
            }
        }
        const bool busy = c >= 0;
        if (__ballot(busy) == 0) break;
        // the link's end point
        Vec3<float> P1 = P0;
        float r = 0.0f;
        if (busy) {
            P1 = skelf_point(fk, rb, qc, i + 1);
            if (!isfinite(P1.x + P1.y + P1.z)) uncertain = true;
            r = rb.radius[i];
        }
        Vec3<float> slo, shi;
        seg_bounds(P0, P1, slo, shi);
        const bool gnd_skip = i < rb.ground_included;
        auto eval = [&](int j) {
            return (s.gnd[j] && gnd_skip) ? p.maxd : flat_pair<OBB>(s, j, P0, P1, r, p, band, uncertain);
        };
        // 1. the warm pair
        float prof = p.maxd, vh = 0.0f;
        int best = -1, jh = -1;
        const int w = busy ? warm : -1;
        if (w >= 0) {
            const float dw = eval(w);
            if (dw <= 0) {
                jh = w;
                vh = dw;
            } else if (dw < prof) {
                prof = dw;
                best = w;
            }
        }
        // 2-3. the walk, its pairs evaluated from the queues
        bool go = busy && jh < 0;
        int qh = 0, qn = 0;
        auto walk_step = [&]() {   // one queued pair per lane
            int task = -1;
            while (qn > 0 && task < 0) {
                const int j = ring[(qh & (kLaneQueue - 1)) * B];
                ++qh;
                --qn;
                if (!cull_fast(lcull + 6 * j, slo, shi, cull_bound2(prof, r, p.eq))) task = j;
            }
            if (task >= 0) {
                const float dt = eval(task);
                if (dt < prof) best = task;
                prof = smin(prof, dt);
                if (prof <= 0) {
                    jh = task;
                    vh = prof;
                    go = false;
                    qn = 0;
                }
            }
        };
        for (int g = 0; g < ng; ++g) {
            if (__ballot(go) == 0) break;
            if (go && !cull_fast(gcull + 6 * g, slo, shi, cull_bound2(prof, r, p.eq))) {
                const int m1 = min(M, (g + 1) * kGroupSize);
                for (int m = g * kGroupSize; m < m1; ++m) {
                    const int j = gorder[m];
                    if (go && j != w && !cull_fast(cull + 6 * j, slo, shi, cull_bound2(prof, r, p.eq))) {
                        ring[((qh + qn) & (kLaneQueue - 1)) * B] = (int16_t)j;
                        ++qn;
                    }
                }
            }
            // evaluate once enough lanes hold pairs; before the next group,
            // every queue must have room for a whole group
            if (__popcll(__ballot(qn > 0)) >= kDrainLanes) walk_step();
            while (__ballot(qn > kLaneQueue - kGroupSize)) walk_step();
        }
        while (__ballot(qn > 0)) walk_step();
        // 4. the first hit in obstacle order before the hit found
        bool srch = busy && jh >= 0;
        if (__ballot(srch)) {
            const float bb0 = cull_bound2(0.0f, r, p.eq);
            int fj = jh;
            float fv = vh;
            auto search_step = [&]() {
                if (qn > 0) {
                    const int j = ring[(qh & (kLaneQueue - 1)) * B];
                    ++qh;
                    --qn;
                    if (j < fj) {
                        const float dt = eval(j);
                        if (dt <= 0) {
                            fj = j;
                            fv = dt;
                        }
                    }
                }
            };
            qh = qn = 0;
            for (int g = 0; g < ng; ++g) {
                if (srch && !cull_fast(gcull + 6 * g, slo, shi, bb0)) {
                    const int m1 = min(M, (g + 1) * kGroupSize);
                    for (int m = g * kGroupSize; m < m1; ++m) {
                        const int j = gorder[m];
                        if (srch && j < fj && !cull_fast(cull + 6 * j, slo, shi, bb0)) {
                            ring[((qh + qn) & (kLaneQueue - 1)) * B] = (int16_t)j;
                            ++qn;
                        }
                    }
                }
                if (__popcll(__ballot(qn > 0)) >= kDrainLanes) search_step();
                while (__ballot(qn > kLaneQueue - kGroupSize)) search_step();
            }
            while (__ballot(qn > 0)) search_step();
            if (srch) vh = fv;
        }
        // the link's value
        bool finished = false;
        if (busy) {
            if (uncertain && uk < 0) uk = i;
            const bool hit = jh >= 0;
            const float lp = hit ? vh : prof;
            if (out.profile) out.profile[c * L + i] = lp;
            if (hit) {
                for (int k = i + 1; k < L; ++k)
                    if (out.profile) out.profile[c * L + k] = 0.0f;
                finished = true;
            } else {
                d = smin(d, lp);
                warm = best;
                if (i == L - 1) {
                    finished = true;
                } else {
                    ++i;
                    P0 = P1;
                }
            }
            if (finished) {
                if (out.d) out.d[c] = hit ? 0.0f : d;
                if (out.verdict) out.verdict[c] = hit ? 0 : 1;
            }
        }
        append_flagged(finished && uncertain, flag_entry(c, uk, qp.flag_link != 0), idx, count);
        if (finished) c = -1;
    }
}

// Wave-wide minimum / maximum of a float (non-finite lanes contribute
// nothing when given as +-infinity by the caller).
__device__ __attribute__((always_inline)) inline float wave_min(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __attribute__((always_inline)) inline float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// Morton code (10 bits per axis, x lowest) of a point quantized over the
// box lo + [0, 1024 / scale) per axis (box = lo[3], scale[3]); points
// outside are clamped, non-finite coordinates give 0.
__device__ __attribute__((always_inline)) inline uint32_t point_key(const Vec3<float>& P, const float* box) {
    const float c[3] = {P.x, P.y, P.z};
    uint32_t key = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float t = (c[a] - box[a]) * box[3 + a];
        const uint32_t u = (t >= 0.0f) ? (uint32_t)fminf(t, 1023.0f) : 0u;   // NaN compares false: 0
#pragma unroll
        for (int b = 0; b < 10; ++b) key |= ((u >> b) & 1u) << (3 * b + a);
    }
    return key;
}

// Set-up of the link-major walk (one workgroup of kBoxThreads):
//   box[0..5]: the box quantizing the survivors' sort keys, the union of the
//     scene's group boxes (finite coordinates only) as lo[3], 1024 /
//     extent[3] (0 for an empty extent);
//   warm0: the first link's warm pair, the obstacle whose cull box is nearest
//     to configuration 0's first skeleton point (smallest index on ties; -1
//     without obstacles).  Any obstacle gives the walk the same results; a
//     near one gives the first link a tight bound from its first pair.
constexpr int kBoxThreads = 256;
__global__ __launch_bounds__(kBoxThreads) void k_links_setup(RobotDev rb, EnvDev env, const float* __restrict__ q,
                                                             float* __restrict__ box, int32_t* __restrict__ warm0) {
    __shared__ float red[6][kBoxThreads];
    __shared__ float bd[kBoxThreads];
    __shared__ int bj[kBoxThreads];
    const int ng = num_groups(env.num_obj);
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int g = threadIdx.x; g < ng; g += kBoxThreads)
        for (int a = 0; a < 3; ++a) {
            const float l = env.gcull[6 * g + a], h = env.gcull[6 * g + 3 + a];
            if (isfinite(l)) lo[a] = fminf(lo[a], l);
            if (isfinite(h)) hi[a] = fmaxf(hi[a], h);
        }
    FKf fk;
    fk.M = Mat3f{1.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f, 0.0f, 1.0f};
    fk.px = fk.py = fk.pz = 0.0f;
    fk.seg = 0;
    fk.qi = 0;
    const Vec3<float> P = skelf_point(fk, rb, q, 0);
    float best = INFINITY;
    int bestj = -1;
    for (int j = threadIdx.x; j < env.num_obj; j += kBoxThreads) {
        const float* cb = env.cull + 6 * j;
        const float gx = fmaxf(fmaxf(cb[0] - P.x, P.x - cb[3]), 0.0f);
        const float gy = fmaxf(fmaxf(cb[1] - P.y, P.y - cb[4]), 0.0f);
        const float gz = fmaxf(fmaxf(cb[2] - P.z, P.z - cb[5]), 0.0f);
        const float g2 = fmaf(gz, gz, fmaf(gy, gy, gx * gx));
        if (g2 < best) {   // ascending j per thread: the first of equals stays
            best = g2;
            bestj = j;
        }
    }
    for (int a = 0; a < 3; ++a) {
        red[a][threadIdx.x] = lo[a];
        red[3 + a][threadIdx.x] = hi[a];
    }
    bd[threadIdx.x] = best;
    bj[threadIdx.x] = bestj;
    __syncthreads();
    for (int w = kBoxThreads / 2; w > 0; w >>= 1) {
        const int t = threadIdx.x;
        if (t < w) {
            for (int a = 0; a < 3; ++a) {
                red[a][t] = fminf(red[a][t], red[a][t + w]);
                red[3 + a][t] = fmaxf(red[3 + a][t], red[3 + a][t + w]);
            }
            const float d2 = bd[t + w];
            const int j2 = bj[t + w];
            if (j2 >= 0 && (bj[t] < 0 || d2 < bd[t] || (d2 == bd[t] && j2 < bj[t]))) {
                bd[t] = d2;
                bj[t] = j2;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 3) {
        const int a = threadIdx.x;
        const float l = red[a][0], h = red[3 + a][0];
        const bool ok = l < h;   // false for an empty scene (infinities)
        box[a] = ok ? l : 0.0f;
        box[3 + a] = ok ? 1024.0f / (h - l) : 0.0f;
    }
    if (threadIdx.x == 0) *warm0 = bj[0];
}

// A round's cell counts and per-wave words (256 + 48 words) reuse the queue
// rings, at least 256 threads' worth.
static_assert(kLaneQueue * sizeof(int16_t) * 256 >= (256 + 48) * sizeof(uint32_t), "queue rings hold the counts");

// The link-major walk (tuning variant 4): one launch per link index i over
// the configurations still alive at link i.  A workgroup takes rounds of
// blockDim.x consecutive entries of the link's list from a counter, 64 per
// wave, so that the lanes of a wave hold the same link of nearby
// configurations.  Per configuration:
//   1. the link's end points: the forward-kinematics walk goes on from the
//      frame the previous link's launch left in the list entry (the same
//      compositions in the same order as the incremental walk, so the same
//      bits);
//   2. the warm pair (the previous link's nearest obstacle; for the first
//      link the obstacle nearest to the first skeleton point), all lanes in
//      one step;
//   3. the wave's candidate groups: 64 groups at a time, one per lane, tested
//      against the union of the lanes' segment boxes with the largest of
//      their bounds (a bound only shrinks during the walk, so a group failing
//      it fails every lane's own test: skipping it changes nothing);
//   4. the wave walks the candidate groups in the walk order together (a
//      group's box and members read once, with scalar loads, for all lanes);
//      each lane queues the members that pass its bound (an LDS ring of
//      kLaneQueue per lane); whenever enough lanes hold queued pairs every
//      lane takes its next one, tests it again against its bound of now and
//      evaluates it if it passes (the flat evaluation, one pair per lane per
//      step, each its own obstacle);
//   5. once a lane's pair ends the link (distance <= 0), the rest of its walk
//      is the search for the first hit in obstacle order: the positions after
//      the hit, with the bound of a running minimum of 0, obstacles of a
//      smaller index than the smallest hit found.  Every position before the
//      hit held a pair evaluated above 0 or culled with a bound at least that
//      of a minimum of 0, which never culls a pair whose distance can be
//      <= 0: none of them can be an earlier hit.
// The link's value (the value at the first hit in obstacle order, or the
// minimum over the pairs, which no culled pair can lower) and the verdict
// are those of link_min whatever the warm pair; the walk evaluates the pairs
// of link_min's walk up to its hit in its order when their warm pairs agree
// (the queueing bound is never tighter than the walk's, a member's box lies
// inside its group's, and every pair is tested again with exactly the bound
// the walk has when it reaches it).  A configuration that goes on keeps its
// running minimum, warm pair, screening flag and walk state in its entry
// of the next link's list; a round's survivors are ordered by the Morton code
// (over box) of their next link's start point, at the resolution of 256 cells
// of the round's own region of the curve, and appended to that list
// together, so that the next launch's waves again hold nearby segments.  The
// lists' order between rounds depends on timing; no configuration's results
// do.  (Rounds of one 64-entry chunk per wave measured fastest: a window
// twice as long per round, 5% slower on workload 5; four times, a smaller
// share of LDS left for resident workgroups, 60% slower.  A full bitonic sort
// of the round's keys was 4% slower than the counting sort.)
template <bool OBB>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(4))) void k_distance_fast_links(
    RobotDev rb, EnvDev env, QueryParams qp, const float* __restrict__ q, int64_t n, DistOut out, int32_t* idx,
    uint32_t* count, int link, const int32_t* __restrict__ perm, const LinkEntry* __restrict__ in,
    const uint32_t* __restrict__ in_count, uint32_t* cursor, LinkEntry* __restrict__ next_list, uint32_t* next_count,
    const float* __restrict__ box, const int32_t* __restrict__ warm0, size_t cull_off, size_t queue_off) {
    // configurations at this link: the first takes all n (in the walk order
    // perm, or the order given), the others the previous link's list
    const int64_t cnt = link == 0 ? n : (int64_t)*in_count;
    const int B = blockDim.x;
    if ((int64_t)blockIdx.x * B >= cnt) return;   // no round for this workgroup (uniform)
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ uint32_t round_id, out_base;
    float* const lcull = reinterpret_cast<float*>(smem + cull_off);   // [M][6] cull boxes, for the queued pairs
    for (int k = threadIdx.x; k < 6 * env.num_obj; k += B) lcull[k] = env.cull[k];
    const Staged<float> s = stage_env<float, OBB>(env, smem);   // ends with a barrier
    int16_t* const ring = reinterpret_cast<int16_t*>(smem + queue_off) + threadIdx.x;   // slot k: ring[k * B]
    const const_ptr<float> gcull = as_const(env.gcull);
    const const_ptr<float> ccull = as_const(env.cull);
    const const_ptr<int32_t> gorder = as_const(env.gorder);
    const float kbox[6] = {box[0], box[1], box[2], box[3], box[4], box[5]};
    const int first_warm = link == 0 ? *warm0 : -1;
    const int M = env.num_obj;
    const int L = rb.num_links;
    const int D = rb.num_dofs;
    const int ng = num_groups(M);
    const Params<float> p{qp.eq, qp.maxd};
    const float band = qp.band;
    const int i = link;
    const float r = rb.radius[i];
    const bool gnd_skip = i < rb.ground_included;
    const float bb0 = cull_bound2(0.0f, r, p.eq);
    const int lane = __lane_id();

    for (;;) {
        // the workgroup's next round
        if (threadIdx.x == 0) round_id = atomicAdd(cursor, 1u);
        __syncthreads();
        const int64_t e = (int64_t)round_id * B + threadIdx.x;
        if ((int64_t)round_id * B >= cnt) break;   // uniform: every thread read the same round
        const bool busy = e < cnt;
        int64_t c = 0;
        float d = p.maxd;
        int warm = first_warm;
        int uk = -1;   // the first link with an uncertain decision
        bool uncertain = false;
        Vec3<float> P0{0.0f, 0.0f, 0.0f}, P1{0.0f, 0.0f, 0.0f};
        FKf fk;
        fk.M = Mat3f{1.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f, 0.0f, 1.0f};
        fk.px = fk.py = fk.pz = 0.0f;
        fk.seg = 0;
        fk.qi = 0;
        if (busy) {
            if (i > 0) {
                // the state the previous link left, its walk included (the
                // walk goes on from the same floats, so the same bits)
                const LinkEntry en = in[e];
                c = en.c;
                d = en.d;
                warm = (int)(en.w >> 1) - 1;
                uncertain = (en.w & 1u) != 0;
                fk.M = Mat3f{en.m[0], en.m[1], en.m[2], en.m[3], en.m[4], en.m[5], en.m[6], en.m[7], en.m[8]};
                fk.px = en.p[0];
                fk.py = en.p[1];
                fk.pz = en.p[2];
                fk.seg = (int)(en.walk & 0xffu);
                fk.qi = (int)((en.walk >> 8) & 0xffu);
                uk = (int)((en.walk >> 16) & 0xffu) - 1;
            } else {
                c = perm ? (int64_t)perm[e] : e;
            }
            const float* qc = q + c * D;
            P0 = skelf_point(fk, rb, qc, i);
            P1 = skelf_point(fk, rb, qc, i + 1);
            if (!isfinite(P1.x + P1.y + P1.z)) uncertain = true;
        }
        Vec3<float> slo, shi;
        seg_bounds(P0, P1, slo, shi);
        auto eval = [&](int j) {
            return (s.gnd[j] && gnd_skip) ? p.maxd : flat_pair<OBB>(s, j, P0, P1, r, p, band, uncertain);
        };
        // 2. the warm pair
        float prof = p.maxd, vh = 0.0f;
        int best = -1, jh = -1;   // jh: the smallest hit index found (-1: walking)
        const int w = busy ? warm : -1;
        if (w >= 0) {
            const float dw = eval(w);
            if (dw <= 0) {
                jh = w;
                vh = dw;
            } else if (dw < prof) {
                prof = dw;
                best = w;
            }
        }
        // the lane's squared bound (running minimum, or 0 once searching; an
        // idle lane's culls everything) and the obstacles it still takes
        // (all while walking, those below the smallest hit once searching)
        float bb = !busy ? -1.0f : (jh < 0 ? cull_bound2(prof, r, p.eq) : bb0);
        int lim = jh < 0 ? M : jh;
        // 3. the union of the lanes' segment boxes and the largest bound
        const Vec3<float> ulo{wave_min(busy ? slo.x : INFINITY), wave_min(busy ? slo.y : INFINITY),
                              wave_min(busy ? slo.z : INFINITY)};
        const Vec3<float> uhi{wave_max(busy ? shi.x : -INFINITY), wave_max(busy ? shi.y : -INFINITY),
                              wave_max(busy ? shi.z : -INFINITY)};
        const float ubb = wave_max(bb);
        // 4-5. the walk, then the search, its pairs evaluated from the queues
        int qh = 0, qt = 0, qn = 0;   // ring head, tail and length
        auto step = [&]() {   // one queued pair per lane
            int task = -1;
            while (qn > 0 && task < 0) {
                const int j = ring[qh * B];
                qh = (qh + 1) & (kLaneQueue - 1);
                --qn;
                if (j < lim && !cull_fast(lcull + 6 * j, slo, shi, bb)) task = j;
            }
            if (task >= 0) {
                const float dt = eval(task);
                if (jh < 0) {
                    if (dt < prof) best = task;
                    prof = smin(prof, dt);
                    if (prof <= 0) {
                        jh = task;
                        vh = prof;
                        lim = task;
                        bb = bb0;
                    } else {
                        bb = cull_bound2(prof, r, p.eq);
                    }
                } else if (dt <= 0) {
                    jh = task;
                    vh = dt;
                    lim = task;
                }
            }
        };
        for (int g0 = 0; g0 < ng; g0 += 64) {
            const int gl = g0 + lane;
            const bool cand = gl < ng && !cull_fast(env.gcull + 6 * gl, ulo, uhi, ubb);
            uint64_t mask = __ballot(cand);
            while (mask) {
                const int g = g0 + __ffsll((unsigned long long)mask) - 1;
                mask &= mask - 1;
                const float gb[6] = {gcull[6 * g], gcull[6 * g + 1], gcull[6 * g + 2],
                                     gcull[6 * g + 3], gcull[6 * g + 4], gcull[6 * g + 5]};
                const bool gpass = !cull_fast(gb, slo, shi, bb);
                if (__ballot(gpass)) {
                    const int m1 = min(M, (g + 1) * kGroupSize);
                    for (int m = g * kGroupSize; m < m1; ++m) {
                        const int j = gorder[m];
                        const float cb[6] = {ccull[6 * j], ccull[6 * j + 1], ccull[6 * j + 2],
                                             ccull[6 * j + 3], ccull[6 * j + 4], ccull[6 * j + 5]};
                        if (gpass && j != w && j < lim && !cull_fast(cb, slo, shi, bb)) {
                            ring[qt * B] = (int16_t)j;
                            qt = (qt + 1) & (kLaneQueue - 1);
                            ++qn;
                        }
                    }
                }
                // evaluate once enough lanes hold pairs; before the next
                // group, every queue must have room for a whole group
                if (__popcll(__ballot(qn > 0)) >= kDrainLanes) step();
                while (__ballot(qn > kLaneQueue - kGroupSize)) step();
            }
        }
        while (__ballot(qn > 0)) step();
        // the link's value
        bool finished = false;
        if (busy) {
            if (uncertain && uk < 0) uk = i;
            const bool hit = jh >= 0;
            const float lp = hit ? vh : prof;
            if (out.profile) out.profile[c * L + i] = lp;
            if (hit) {
                if (out.profile)
                    for (int k = i + 1; k < L; ++k) out.profile[c * L + k] = 0.0f;
                finished = true;
            } else {
                d = smin(d, lp);
                finished = i == L - 1;
            }
            if (finished) {
                if (out.d) out.d[c] = hit ? 0.0f : d;
                if (out.verdict) out.verdict[c] = hit ? 0 : 1;
            }
        }
        append_flagged(busy && finished && uncertain, flag_entry(c, uk, qp.flag_link != 0), idx, count);
        // survivors, with the sort key of their next link's start point
        const bool keep = busy && !finished;
        // the round's survivors, appended together in the order of their next
        // link's start point: a counting sort on the 8 bits of its Morton code
        // just below the round's common prefix (the round's keys lie between
        // its smallest and largest, which share every bit above the highest
        // one where they differ, so these bits cut the round's own region of
        // the curve into 256 cells); ranks within a cell by LDS atomics
        const uint32_t key = keep ? point_key(P1, kbox) : 0u;
        uint32_t kmin = keep ? key : 0xffffffffu, kmax = keep ? key : 0u;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            kmin = min(kmin, (uint32_t)__shfl_xor((int)kmin, o, 64));
            kmax = max(kmax, (uint32_t)__shfl_xor((int)kmax, o, 64));
        }
        uint32_t* const cell = reinterpret_cast<uint32_t*>(smem + queue_off);   // [256] cell counts, then starts
        uint32_t* const wmm = cell + 256;                                       // [16][3] wave min, max, sum
        const int wv = threadIdx.x >> 6, nw = B >> 6;
        __syncthreads();   // every wave has left its walk: the queue rings are free
        if (lane == 0) {
            wmm[3 * wv] = kmin;
            wmm[3 * wv + 1] = kmax;
        }
        if (threadIdx.x < 256) cell[threadIdx.x] = 0u;
        __syncthreads();
        kmin = 0xffffffffu;
        kmax = 0u;
        for (int k = 0; k < nw; ++k) {
            kmin = min(kmin, wmm[3 * k]);
            kmax = max(kmax, wmm[3 * k + 1]);
        }
        const int hb = 31 - __clz((int)((kmin ^ kmax) | 1u));
        const int sh = hb > 7 ? hb - 7 : 0;
        const uint32_t bucket = (key >> sh) & 255u;
        uint32_t rk = 0;
        if (keep) rk = atomicAdd(&cell[bucket], 1u);
        __syncthreads();
        // exclusive scan of the cell counts (one a thread, the first 256)
        const uint32_t v = threadIdx.x < 256 ? cell[threadIdx.x] : 0u;
        uint32_t x = v;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off, 64);
            if (lane >= off) x += y;
        }
        if (lane == 63) wmm[3 * wv + 2] = x;
        __syncthreads();
        uint32_t before = 0, all = 0;
        for (int k = 0; k < nw; ++k) {
            const uint32_t sk = wmm[3 * k + 2];
            before += k < wv ? sk : 0u;
            all += sk;
        }
        if (threadIdx.x < 256) cell[threadIdx.x] = before + x - v;
        if (threadIdx.x == 0) out_base = all ? atomicAdd(next_count, all) : 0u;
        __syncthreads();
        if (keep) {
            LinkEntry en;
            en.c = (int32_t)c;
            en.d = d;
            en.w = ((uint32_t)(best + 1) << 1) | (uncertain ? 1u : 0u);
            en.walk = (uint32_t)fk.seg | ((uint32_t)fk.qi << 8) | ((uint32_t)(uk + 1) << 16);
            en.m[0] = fk.M.m0; en.m[1] = fk.M.m1; en.m[2] = fk.M.m2;
            en.m[3] = fk.M.m3; en.m[4] = fk.M.m4; en.m[5] = fk.M.m5;
            en.m[6] = fk.M.m6; en.m[7] = fk.M.m7; en.m[8] = fk.M.m8;
            en.p[0] = fk.px; en.p[1] = fk.py; en.p[2] = fk.pz;
            next_list[out_base + cell[bucket] + rk] = en;
        }
        __syncthreads();   // round_id and the rings are written again next round
    }
}

__global__ void k_zero_u32(uint32_t* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Coherence order.  The key of a configuration is its first
// min(dofs, 18 / bits) joint angles, each wrapped to [-pi, pi) and cut into
// 2^bits bins (the first joint in the highest bits).  The configurations are sorted by key in two
// passes without global atomics: a counting sort of tiles by the key's top
// (up to) 8 bits into 256 buckets (tile histograms, a scan of each bucket's
// row of tile counts, a scatter ranked by LDS atomics), then every chunk of at
// most kCohChunk configurations of a bucket is sorted by the remaining low
// bits (up to 10) in LDS.  Neighbours in the resulting order hold similar
// poses of the base links, so the lanes of a wave take the same branches of
// the case tree and the same broad-phase decisions more often.  The order
// within a key depends on atomics; a configuration's result does not depend
// on which lane computes it.
__device__ inline uint32_t coh_key(const float* qc, int nj, int bits) {
    const int nb = 1 << bits;
    uint32_t key = 0;
    for (int k = 0; k < nj; ++k) {
        float t = (qc[k] + 3.14159265f) * 0.159154943f;   // revolutions from -pi
        t = t - floorf(t);
        int b = (t >= 0.0f && t < 1.0f) ? (int)(t * (float)nb) : 0;   // non-finite angles: bin 0
        b = b < nb ? b : nb - 1;
        key = key * (uint32_t)nb + (uint32_t)b;
    }
    return key;
}

// Exclusive prefix sum of v over the NT threads of a workgroup; total = the
// sum over all of them.  Every thread of the workgroup must call it.
template <int NT>
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t& total) {
    constexpr int NW = NT / 64;
    __shared__ uint32_t wsum[NW];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
        const uint32_t sk = wsum[k];
        before += k < w ? sk : 0u;
        all += sk;
    }
    __syncthreads();
    total = all;
    return before + x - v;
}

constexpr int kCohThreads = 256;                      // threads of the per-bucket scan
constexpr int kCohTileThreads = 1024;                 // threads of the tile kernels (the first 256: one per bucket)
constexpr int kCohPer = kCohTile / kCohTileThreads;   // configurations per thread of a tile
constexpr int kCohSortThreads = 1024;                 // threads of the low-bit sort (one per low key)
static_assert(kCohBuckets <= kCohTileThreads, "one thread per bucket");
static_assert((1 << kCohLoBits) <= kCohSortThreads, "one thread per low key");

// Pass 1: keys of a tile of kCohTile configurations and the histogram of
// their buckets (key >> lo), stored bucket-major: hist[bucket * tiles + tile].
__global__ __launch_bounds__(kCohTileThreads) void k_coh_hist(const float* __restrict__ q, int64_t n, int D, int nj,
                                                              int bits, int lo, int64_t tiles,
                                                              uint32_t* __restrict__ keys, uint32_t* __restrict__ hist,
                                                              uint32_t* __restrict__ zero, int nzero) {
    __shared__ uint32_t cnt[kCohBuckets];
    const int t = threadIdx.x;
    if (t < kCohBuckets) cnt[t] = 0;
    if (blockIdx.x == 0 && zero)   // the caller's counters, cleared on the way
        for (int k = t; k < nzero; k += kCohTileThreads) zero[k] = 0u;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kCohTile;
    for (int k = 0; k < kCohPer; ++k) {
        const int64_t c = base + k * kCohTileThreads + t;
        if (c < n) {
            const uint32_t key = coh_key(q + c * D, nj, bits);
            keys[c] = key;
            atomicAdd(&cnt[key >> lo], 1u);
        }
    }
    __syncthreads();
    if (t < kCohBuckets) hist[(int64_t)t * tiles + blockIdx.x] = cnt[t];
}

// Pass 2: one workgroup per bucket turns its row of tile counts into
// exclusive offsets within the bucket and stores the bucket's size.
__global__ __launch_bounds__(kCohThreads) void k_coh_scan(uint32_t* __restrict__ hist, int64_t tiles,
                                                          uint32_t* __restrict__ sizes) {
    uint32_t* row = hist + (int64_t)blockIdx.x * tiles;
    uint32_t carry = 0;
    for (int64_t s0 = 0; s0 < tiles; s0 += kCohThreads) {
        const int64_t i = s0 + threadIdx.x;
        const uint32_t v = i < tiles ? row[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan<kCohThreads>(v, tot);
        if (i < tiles) row[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) sizes[blockIdx.x] = carry;
}

// Pass 3: every configuration of a tile goes to its bucket (bucket start +
// the tile's offset in the bucket + its rank among the tile's members); the
// low key bits travel with it.  With lo == 0 this is the final order.
__global__ __launch_bounds__(kCohTileThreads) void k_coh_bucket(const uint32_t* __restrict__ keys, int64_t n,
                                                                int lo, const uint32_t* __restrict__ hist,
                                                                int64_t tiles, const uint32_t* __restrict__ sizes,
                                                                uint16_t* __restrict__ key_lo,
                                                                int32_t* __restrict__ dst) {
    __shared__ uint32_t next[kCohBuckets];
    const int t = threadIdx.x;
    uint32_t tot;
    const uint32_t start = block_exclusive_scan<kCohTileThreads>(t < kCohBuckets ? sizes[t] : 0u, tot);
    if (t < kCohBuckets) next[t] = start + hist[(int64_t)t * tiles + blockIdx.x];
    __syncthreads();
    const uint32_t mask = (1u << lo) - 1u;
    const int64_t base = (int64_t)blockIdx.x * kCohTile;
    for (int k = 0; k < kCohPer; ++k) {
        const int64_t c = base + k * kCohTileThreads + t;
        if (c < n) {
            const uint32_t key = keys[c];
            const uint32_t pos = atomicAdd(&next[key >> lo], 1u);
            if (lo) key_lo[pos] = (uint16_t)(key & mask);
            dst[pos] = (int32_t)c;
        }
    }
}

// Pass 4: workgroup j sorts chunk j of the buckets' chunks (each bucket cut
// into chunks of at most kCohChunk configurations, buckets in order) by the
// low key bits with an LDS counting sort.  Workgroups beyond the last chunk
// return at once; the grid is 256 + ceil(n / kCohChunk), which bounds the
// number of chunks.
__global__ __launch_bounds__(kCohSortThreads) void k_coh_lowsort(const uint32_t* __restrict__ sizes, int lo,
                                                                 const uint16_t* __restrict__ key_lo,
                                                                 const int32_t* __restrict__ idx,
                                                                 int32_t* __restrict__ perm) {
    __shared__ uint32_t cnt[kCohSortThreads];
    __shared__ uint32_t first_chunk[kCohBuckets + 1];
    __shared__ uint32_t bstart[kCohBuckets], bsize[kCohBuckets];
    const int t = threadIdx.x;
    const uint32_t size = t < kCohBuckets ? sizes[t] : 0u;
    const uint32_t chunks = (size + kCohChunk - 1) / kCohChunk;
    uint32_t total, total_chunks;
    const uint32_t start = block_exclusive_scan<kCohSortThreads>(size, total);
    const uint32_t cfirst = block_exclusive_scan<kCohSortThreads>(chunks, total_chunks);
    if (t < kCohBuckets) {
        first_chunk[t] = cfirst;
        bstart[t] = start;
        bsize[t] = size;
    }
    if (t == 0) first_chunk[kCohBuckets] = total_chunks;
    __syncthreads();
    const uint32_t j = blockIdx.x;
    if (j >= total_chunks) return;   // uniform over the workgroup
    // the bucket of chunk j: the largest b with first_chunk[b] <= j (never empty)
    int b = 0;
    for (int step = kCohBuckets / 2; step > 0; step >>= 1)
        if (first_chunk[b + step] <= j) b += step;
    const uint32_t begin = bstart[b] + (j - first_chunk[b]) * (uint32_t)kCohChunk;
    const uint32_t bend = bstart[b] + bsize[b];
    const uint32_t end = bend < begin + (uint32_t)kCohChunk ? bend : begin + (uint32_t)kCohChunk;
    const int nb = 1 << lo;
    if (t < nb) cnt[t] = 0;
    __syncthreads();
    for (uint32_t i = begin + t; i < end; i += kCohSortThreads) atomicAdd(&cnt[key_lo[i]], 1u);
    __syncthreads();
    const uint32_t v = t < nb ? cnt[t] : 0u;
    uint32_t dummy;
    const uint32_t ex = block_exclusive_scan<kCohSortThreads>(v, dummy);
    if (t < nb) cnt[t] = ex;
    __syncthreads();
    for (uint32_t i = begin + t; i < end; i += kCohSortThreads) {
        const uint32_t pos = begin + atomicAdd(&cnt[key_lo[i]], 1u);
        perm[pos] = idx[i];
    }
}

}  // namespace

int fast_pool_blocks_per_cu(const EnvDev& env, int block) {
    int nb = 0;
    const size_t sm = staged_bytes<float>(env);
    const bool warm = env.gorder && env.gcull;
    hipError_t e;
    if (env.use_obb)
        e = warm ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_distance_fast_pool<true, true>, block, sm)
                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_distance_fast_pool<true, false>, block, sm);
    else
        e = warm ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_distance_fast_pool<false, true>, block, sm)
                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_distance_fast_pool<false, false>, block, sm);
    return e == hipSuccess ? nb : 0;
}

size_t device_lds_limit(int dev) {
    static std::mutex mu;
    static std::map<int, size_t> cache;
    if (dev < 0 && hipGetDevice(&dev) != hipSuccess) return 64 * 1024;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(dev);
    if (it != cache.end()) return it->second;
    int v = 0;
    const size_t lim = hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0
                           ? (size_t)v
                           : (size_t)64 * 1024;
    cache[dev] = lim;
    return lim;
}

namespace {

// LDS of the per-lane-walk kernel for workgroups of `block` threads: the
// staged obstacles, then (16-byte aligned) a copy of the cull boxes, then the
// lanes' pair queues; 0 when it cannot run on this obstacle set.
size_t lanes_lds(const EnvDev& env, int block, size_t* cull_off, size_t* queue_off) {
    if (!env.gorder || !env.gcull || env.num_obj > 32767) return 0;
    const size_t coff = (staged_bytes<float>(env) + 15) & ~size_t(15);
    const size_t qoff = coff + (size_t)env.num_obj * 6 * sizeof(float);
    const size_t total = qoff + (size_t)kLaneQueue * block * sizeof(int16_t);
    if (total > device_lds_limit()) return 0;
    if (cull_off) *cull_off = coff;
    if (queue_off) *queue_off = qoff;
    return total;
}

// LDS of the link-major kernel: that of the per-lane-queue kernel (the
// rounds' survivor ranking reuses its queue rings); 0 when it cannot run on
// this obstacle set or with fewer than 256 threads (kStaticLds: room for the
// kernel's own shared variables).
constexpr size_t kStaticLds = 1024;
size_t links_lds(const EnvDev& env, int block, size_t* cull_off, size_t* queue_off) {
    if (block < 256) return 0;
    size_t c = 0, qo = 0;
    const size_t total = lanes_lds(env, block, &c, &qo);
    if (!total || total + kStaticLds > device_lds_limit()) return 0;
    if (cull_off) *cull_off = c;
    if (queue_off) *queue_off = qo;
    return total;
}

}  // namespace

int fast_lanes_blocks_per_cu(const EnvDev& env, int block) {
    const size_t sm = lanes_lds(env, block, nullptr, nullptr);
    if (!sm) return 0;
    int nb = 0;
    const hipError_t e =
        env.use_obb ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_distance_fast_lanes<true>, block, sm)
                    : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_distance_fast_lanes<false>, block, sm);
    return e == hipSuccess ? nb : 0;
}

int fast_links_blocks_per_cu(const EnvDev& env, int block) {
    const size_t sm = links_lds(env, block, nullptr, nullptr);
    if (!sm) return 0;
    int nb = 0;
    const hipError_t e =
        env.use_obb ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_distance_fast_links<true>, block, sm)
                    : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_distance_fast_links<false>, block, sm);
    return e == hipSuccess ? nb : 0;
}

hipError_t launch_zero_u32(uint32_t* p, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_u32, dim3((n + 63) / 64), dim3(64), 0, st, p, n);
    return hipGetLastError();
}

hipError_t launch_coherence_order(const RobotDev& rb, const float* q, int64_t n, const FastTuning& t, uint32_t* coh,
                                  const int32_t** perm, hipStream_t st, uint32_t* zero, int nzero) {
    *perm = nullptr;
    const int bits = std::min<int>(t.coh_bits, kCohKeyBits);
    if (n <= 0 || !coh || bits <= 0 || rb.num_dofs <= 0) return hipSuccess;
    const int nj = std::min(rb.num_dofs, kCohKeyBits / bits);
    if (n > INT32_MAX) return hipErrorInvalidValue;
    const int keybits = nj * bits;
    const int lo = std::max(0, keybits - kCohHiBits);
    const CohLayout L = coh_layout(coh, n);
    const unsigned tiles = (unsigned)L.tiles;
    hipLaunchKernelGGL(k_coh_hist, dim3(tiles), dim3(kCohTileThreads), 0, st, q, n, rb.num_dofs, nj, bits, lo, L.tiles,
                       L.keys, L.hist, zero, nzero);
    hipLaunchKernelGGL(k_coh_scan, dim3(kCohBuckets), dim3(kCohThreads), 0, st, L.hist, L.tiles, L.sizes);
    hipLaunchKernelGGL(k_coh_bucket, dim3(tiles), dim3(kCohTileThreads), 0, st, L.keys, n, lo, L.hist, L.tiles, L.sizes,
                       L.key_lo, lo ? L.idx : L.perm);
    if (lo) {
        const unsigned grid = (unsigned)(kCohBuckets + (n + kCohChunk - 1) / kCohChunk);
        hipLaunchKernelGGL(k_coh_lowsort, dim3(grid), dim3(kCohSortThreads), 0, st, L.sizes, lo, L.key_lo, L.idx,
                           L.perm);
    }
    const hipError_t e = hipGetLastError();
    if (e == hipSuccess) *perm = L.perm;
    return e;
}

namespace {

// Whether a kernel with LDS row slices keeps as many resident workgroups per
// CU as its form without them (answers cached per kernel, block and sizes).
template <typename KRows, typename KPlain>
bool rows_keep_occupancy(KRows krows, KPlain kplain, int block, size_t rows_sm, size_t sm) {
    static std::mutex mu;
    static std::map<std::tuple<const void*, int, size_t, size_t>, bool> cache;
    const auto key = std::make_tuple(reinterpret_cast<const void*>(krows), block, rows_sm, sm);
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int with = 0, without = 0;
    const bool ok = hipOccupancyMaxActiveBlocksPerMultiprocessor(&with, krows, block, rows_sm) == hipSuccess &&
                    hipOccupancyMaxActiveBlocksPerMultiprocessor(&without, kplain, block, sm) == hipSuccess &&
                    with >= without;
    cache[key] = ok;
    return ok;
}

template <bool OBB, bool GROUPED>
hipError_t launch_fast_shape(const RobotDev& rb, const EnvDev& env, const QueryParams& qp, const float* q, int64_t n,
                             const DistOut& out, int32_t* idx, uint32_t* count, const FastTuning& t,
                             const int32_t* perm, uint32_t* links, int num_cus, hipStream_t st) {
    const size_t sm = staged_bytes<float>(env);
    if (t.variant == 4 && GROUPED && links && n <= INT32_MAX) {
        // one launch per link over the configurations alive at it
        const int BL = (t.block == 512 || t.block == 1024) ? t.block : 256;
        size_t cull_off = 0, queue_off = 0;
        const size_t lsm = links_lds(env, BL, &cull_off, &queue_off);
        if (lsm) {
            const LinksLayout ll = links_layout(links, n);
            int per_cu = t.blocks_per_cu > 0 ? t.blocks_per_cu : fast_links_blocks_per_cu(env, BL);
            per_cu = std::max(1, per_cu);
            const int64_t blocks = std::min<int64_t>((int64_t)num_cus * per_cu, (n + BL - 1) / BL);
            const dim3 g((unsigned)std::max<int64_t>(1, blocks)), b(BL);
            hipError_t e = launch_zero_u32(ll.count, kLinksCounters, st);
            if (e == hipSuccess) {
                hipLaunchKernelGGL(k_links_setup, dim3(1), dim3(kBoxThreads), 0, st, rb, env, q, ll.box, ll.warm0);
                e = hipGetLastError();
            }
            for (int i = 0; i < rb.num_links && e == hipSuccess; ++i) {
                const LinkEntry* in = i == 0 ? nullptr : ll.list[(i - 1) & 1];
                hipLaunchKernelGGL((k_distance_fast_links<OBB>), g, b, lsm, st, rb, env, qp, q, n, out, idx, count, i,
                                   perm, in, ll.count + i, ll.cursor + i, ll.list[i & 1], ll.count + i + 1, ll.box,
                                   ll.warm0, cull_off, queue_off);
                e = hipGetLastError();
            }
            return e;
        }
    }
    if ((t.variant == 3 || t.variant == 4) && GROUPED) {
        // persistent per-lane walks; count[32] is the configuration counter
        const int BL = (t.block == 512 || t.block == 1024) ? t.block : 256;
        size_t cull_off = 0, queue_off = 0;
        const size_t lsm = lanes_lds(env, BL, &cull_off, &queue_off);
        if (lsm) {
            int per_cu = t.blocks_per_cu > 0 ? t.blocks_per_cu : fast_lanes_blocks_per_cu(env, BL);
            per_cu = std::max(1, per_cu);
            const int64_t blocks = std::min<int64_t>((int64_t)num_cus * per_cu, (n + BL - 1) / BL);
            const dim3 g((unsigned)std::max<int64_t>(1, blocks)), b(BL);
            hipLaunchKernelGGL((k_distance_fast_lanes<OBB>), g, b, lsm, st, rb, env, qp, q, n, out, idx, count,
                               count + 32, perm, cull_off, queue_off);
            return hipGetLastError();
        }
    }
    const int B = (t.block == 64 || t.block == 128) ? t.block : kBlock;
    if (t.variant == 1 || t.variant == 3 || t.variant == 4) {
        // persistent lane-refill variant; count[32] is its configuration counter
        int per_cu = t.blocks_per_cu > 0 ? t.blocks_per_cu : fast_pool_blocks_per_cu(env, B);
        per_cu = std::max(1, per_cu);
        const int64_t blocks = std::min<int64_t>((int64_t)num_cus * per_cu, (n + B - 1) / B);
        const dim3 g((unsigned)std::max<int64_t>(1, blocks)), b(B);
        hipLaunchKernelGGL((k_distance_fast_pool<OBB, GROUPED>), g, b, sm, st, rb, env, qp, q, n, out, idx, count,
                           count + 32, perm);
        return hipGetLastError();
    }
    // rows staged through LDS when a walk order scatters them, the slices fit,
    // and they cost no resident workgroups (on workload-5-sized obstacle sets
    // the slices would cut the workgroups per CU to one: 3.4x slower)
    const size_t rows_off = (sm + 15) & ~size_t(15);
    const size_t rows_sm = rows_off + (size_t)(rb.num_dofs + rb.num_links) * 4 * B;
    const bool rows = perm && rows_sm <= device_lds_limit() &&
                      (t.variant == 2 ? rows_keep_occupancy(k_distance_fast_chunks<OBB, true, GROUPED>,
                                                            k_distance_fast_chunks<OBB, false, GROUPED>, B, rows_sm, sm)
                                      : rows_keep_occupancy(k_distance_fast<OBB, true, GROUPED>,
                                                            k_distance_fast<OBB, false, GROUPED>, B, rows_sm, sm));
    if (t.variant == 2) {
        // persistent waves over 64-configuration chunks; count[32] is the chunk counter
        const size_t smc = rows ? rows_sm : sm;
        int per_cu = t.blocks_per_cu;
        if (per_cu <= 0) {
            const hipError_t oe =
                rows ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_distance_fast_chunks<OBB, true, GROUPED>,
                                                                    B, smc)
                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_distance_fast_chunks<OBB, false, GROUPED>,
                                                                    B, smc);
            if (oe != hipSuccess) per_cu = 1;
        }
        per_cu = std::max(1, per_cu);
        const int64_t blocks = std::min<int64_t>((int64_t)num_cus * per_cu, (n + B - 1) / B);
        const dim3 g((unsigned)std::max<int64_t>(1, blocks)), b(B);
        if (rows)
            hipLaunchKernelGGL((k_distance_fast_chunks<OBB, true, GROUPED>), g, b, smc, st, rb, env, qp, q, n, out, idx,
                               count, perm, rows_off, count + 32);
        else
            hipLaunchKernelGGL((k_distance_fast_chunks<OBB, false, GROUPED>), g, b, smc, st, rb, env, qp, q, n, out,
                               idx, count, perm, rows_off, count + 32);
        return hipGetLastError();
    }
    const dim3 g((unsigned)((n + B - 1) / B)), b(B);
    if (rows)
        hipLaunchKernelGGL((k_distance_fast<OBB, true, GROUPED>), g, b, rows_sm, st, rb, env, qp, q, n, out, idx, count,
                           perm, rows_off);
    else
        hipLaunchKernelGGL((k_distance_fast<OBB, false, GROUPED>), g, b, sm, st, rb, env, qp, q, n, out, idx, count,
                           perm, rows_off);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_distance_fast(const RobotDev& rb, const EnvDev& env, const QueryParams& qp, const float* q,
                                int64_t n, const DistOut& out, int32_t* idx, uint32_t* count, const FastTuning& t,
                                const int32_t* perm, uint32_t* links, int num_cus, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const bool warm = env.gorder && env.gcull;   // the grouped walk when the scene has group tables
    QueryParams fq = qp;
    fq.flag_link = flags_carry_link(n) ? 1 : 0;
    if (env.use_obb)
        return warm ? launch_fast_shape<true, true>(rb, env, fq, q, n, out, idx, count, t, perm, links, num_cus, st)
                    : launch_fast_shape<true, false>(rb, env, fq, q, n, out, idx, count, t, perm, links, num_cus, st);
    return warm ? launch_fast_shape<false, true>(rb, env, fq, q, n, out, idx, count, t, perm, links, num_cus, st)
                : launch_fast_shape<false, false>(rb, env, fq, q, n, out, idx, count, t, perm, links, num_cus, st);
}

}  // namespace rvs
