// FP32 bulk pass of RVS_PREC_FP32_REFINE (gfx950).
//
// The same case tree and early-exit loop as the reference-exact kernel, in
// relaxed single precision: kinematics in float, FMA contraction (this
// translation unit is compiled with -ffp-contract=fast), hardware reciprocal
// and square root.  Its results differ from the reference arithmetic by a few
// 1e-7 m, so every (link, obstacle) pair it evaluates is screened and the
// configuration is sent to the exact refinement pass when any pair is
// uncertain:
//   |d - r| < band                the verdict sign could differ
//   raw distance < EQ + band      the reference snaps distances below
//                                 EQUALITY_THRESHOLD to 0 (then -r)
//   end point inside, depth < band  the inside test could differ
//   face crossing, margin < band  the crossing test could differ
//   |d - tol| < band              the min_dist_tol filter could differ
//   sin^2 < 1e-6, raw < 0.01      a near-parallel segment solve close by
//   non-finite skeleton           NaN / Inf configurations
// (inside points and face crossings are screened by how far the comparison
// that decided them was from flipping, since their results 0 and -r carry no
// rounding.)  Configurations outside all of these get verdicts identical to the reference
// (a sign decided with margin >= band cannot flip under a 1e-6 m error), and
// the same early-exit pair, so the same partial profile up to rounding.
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

constexpr int kBlock = 256;

struct Mat3f {
    float m0, m1, m2, m3, m4, m5, m6, m7, m8;
};

__device__ inline Mat3f loadf(const double* a) {
    return Mat3f{(float)a[0], (float)a[1], (float)a[2], (float)a[3], (float)a[4],
                 (float)a[5], (float)a[6], (float)a[7], (float)a[8]};
}

__device__ inline Mat3f mulf(const Mat3f& a, const Mat3f& b) {
    return Mat3f{a.m0 * b.m0 + a.m1 * b.m3 + a.m2 * b.m6, a.m0 * b.m1 + a.m1 * b.m4 + a.m2 * b.m7,
                 a.m0 * b.m2 + a.m1 * b.m5 + a.m2 * b.m8, a.m3 * b.m0 + a.m4 * b.m3 + a.m5 * b.m6,
                 a.m3 * b.m1 + a.m4 * b.m4 + a.m5 * b.m7, a.m3 * b.m2 + a.m4 * b.m5 + a.m5 * b.m8,
                 a.m6 * b.m0 + a.m7 * b.m3 + a.m8 * b.m6, a.m6 * b.m1 + a.m7 * b.m4 + a.m8 * b.m7,
                 a.m6 * b.m2 + a.m7 * b.m5 + a.m8 * b.m8};
}

__device__ inline Mat3f rot2f(const double* ad, float st, float ct) {
    const float a0 = (float)ad[0], a1 = (float)ad[1], a2 = (float)ad[2];
    const float vt = 1.0f - ct;
    return Mat3f{ct + vt * a0 * a0,      -a2 * st + vt * a0 * a1, a1 * st + vt * a0 * a2,
                 a2 * st + vt * a0 * a1, ct + vt * a1 * a1,       -a0 * st + vt * a1 * a2,
                 -a1 * st + vt * a0 * a2, a0 * st + vt * a1 * a2, ct + vt * a2 * a2};
}

struct FKf {
    Mat3f M;
    float px, py, pz;
    int seg, qi;
};

__device__ __attribute__((always_inline)) inline void fkf_advance(FKf& s, const SegDev* seg, const float* q, int upto) {
    while (s.seg <= upto) {
        const SegDev& g = seg[s.seg];
        const Mat3f R = loadf(g.R);
        Mat3f SM;
        if (g.rev) {
            float st, ct;
            sincosf(q[s.qi++], &st, &ct);
            SM = mulf(rot2f(g.axis, st, ct), R);
        } else {
            SM = R;
        }
        const float o0 = (float)g.o[0], o1 = (float)g.o[1], o2 = (float)g.o[2];
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
                                                                       int k) {
    const PointDev pt = rb.pts[k];
    if (pt.frame < 0) return Vec3<float>{0.0f, 0.0f, 0.0f};
    fkf_advance(s, rb.seg, q, pt.frame);
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

// One (link, obstacle) pair of the bulk pass: filtered distance, and the
// screening rules that send the configuration to the exact pass.
template <bool OBB>
__device__ __attribute__((always_inline)) inline float fast_pair(const Staged<float>& s, int j, const Vec3<float>& P0,
                                                                 const Vec3<float>& P1, float r, const Params<float>& p,
                                                                 float band, bool& uncertain) {
    const Box<float> bx = s.box[j];
    PairInfo<float> info{p.maxd, p.maxd, -1.0f, 1.0f, false, false};
    float dt;
    if (OBB) {
        // oriented box: solved in the box frame (the frame change adds a few
        // ulp of the lever arm, far inside the band)
        const float* x = s.xf + 12 * j;
        dt = capsule_box<float, false, FastMath>(to_box(x, P0), to_box(x, P1), r, bx, p, nullptr, &info, nullptr);
    } else {
        dt = capsule_box<float, false, FastMath>(P0, P1, r, bx, p, nullptr, &info, nullptr);
    }
    const float tol = s.tol[j];
    // inside points and crossings are decided by comparisons: the result
    // (0 or -r) is exact unless the decision is marginal
    if (info.inside) uncertain |= info.depth < band;
    else if (info.hit) uncertain |= info.hit_margin < band;
    else uncertain |= (info.raw < p.eq + band) || (fabsf(dt) < band) || (info.sin2 < 1e-6f && info.raw < 0.01f);
    if (tol < p.maxd && fabsf(dt - tol) < band) uncertain = true;
    if (dt > tol) dt = p.maxd;
    return dt;
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

template <bool OBB>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_distance_fast(
    RobotDev rb, EnvDev env, QueryParams qp, const float* __restrict__ q, int64_t n, DistOut out, int32_t* idx,
    uint32_t* count) {
    extern __shared__ __align__(16) unsigned char smem[];
    const Staged<float> s = stage_env<float, OBB>(env, smem);
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int M = env.num_obj;
    const int L = rb.num_links;
    const float* qc = q + c * rb.num_dofs;
    const Params<float> p{qp.eq, qp.maxd};
    const float band = qp.band;

    FKf fk;
    fk.M = Mat3f{1.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f, 0.0f, 1.0f};
    fk.px = fk.py = fk.pz = 0.0f;
    fk.seg = 0;
    fk.qi = 0;
    Vec3<float> P0 = skelf_point(fk, rb, qc, 0);
    float d = p.maxd;
    bool hit = false, uncertain = false;
    int i = 0;
    for (; i < L; ++i) {
        const Vec3<float> P1 = skelf_point(fk, rb, qc, i + 1);
        if (!isfinite(P1.x + P1.y + P1.z)) uncertain = true;
        const float r = rb.radius[i];
        float prof = p.maxd;
        Vec3<float> slo, shi;
        seg_bounds(P0, P1, slo, shi);
        for (int j = 0; j < M; ++j) {
            if (cull_pair(env.cull + 6 * j, slo, shi, cull_bound(prof, r, p.eq))) continue;
            const float dt = (s.gnd[j] && i < rb.ground_included) ? p.maxd
                                                                   : fast_pair<OBB>(s, j, P0, P1, r, p, band, uncertain);
            prof = smin(prof, dt);
            if (prof <= 0) { hit = true; break; }
        }
        if (out.profile) out.profile[c * L + i] = prof;
        if (hit) break;
        d = smin(d, prof);
        P0 = P1;
    }
    for (int k = i + 1; k < L; ++k)
        if (out.profile) out.profile[c * L + k] = 0.0f;
    if (hit) d = 0.0f;
    if (out.d) out.d[c] = d;
    if (out.verdict) out.verdict[c] = hit ? 0 : 1;
    append_flagged(uncertain, c, idx, count);
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
template <bool OBB>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_distance_fast_pool(
    RobotDev rb, EnvDev env, QueryParams qp, const float* __restrict__ q, int64_t n, DistOut out, int32_t* idx,
    uint32_t* count, uint32_t* next) {
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
                    c = cand;
                    qc = q + c * D;
                    fk.M = Mat3f{1.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f, 0.0f, 1.0f};
                    fk.px = fk.py = fk.pz = 0.0f;
                    fk.seg = 0;
                    fk.qi = 0;
                    P0 = skelf_point(fk, rb, qc, 0);
                    i = 0;
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
        const bool gnd_skip = i < rb.ground_included;
        float prof = p.maxd;
        bool active = busy, hit = false;
        Vec3<float> slo, shi;
        seg_bounds(P0, P1, slo, shi);
        for (int j = 0; j < M; ++j) {
            if (active && !cull_pair(env.cull + 6 * j, slo, shi, cull_bound(prof, r, p.eq))) {
                const float dt = (s.gnd[j] && gnd_skip) ? p.maxd : fast_pair<OBB>(s, j, P0, P1, r, p, band, uncertain);
                prof = smin(prof, dt);
                if (prof <= 0) { hit = true; active = false; }
            }
            if (__ballot(active) == 0) break;
        }
        bool finished = false;
        if (busy) {
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
        append_flagged(finished && uncertain, c, idx, count);
        if (finished) c = -1;
    }
}

__global__ void k_zero_u32(uint32_t* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

}  // namespace

int fast_pool_blocks_per_cu(const EnvDev& env, int block) {
    int nb = 0;
    const hipError_t e =
        env.use_obb ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_distance_fast_pool<true>, block,
                                                                    staged_bytes<float>(env))
                    : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_distance_fast_pool<false>, block,
                                                                    staged_bytes<float>(env));
    return e == hipSuccess ? nb : 0;
}

hipError_t launch_zero_u32(uint32_t* p, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_u32, dim3((n + 63) / 64), dim3(64), 0, st, p, n);
    return hipGetLastError();
}

hipError_t launch_distance_fast(const RobotDev& rb, const EnvDev& env, const QueryParams& qp, const float* q,
                                int64_t n, const DistOut& out, int32_t* idx, uint32_t* count, const FastTuning& t,
                                int num_cus, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int B = (t.block == 64 || t.block == 128) ? t.block : kBlock;
    const size_t sm = staged_bytes<float>(env);
    if (t.variant == 1) {
        // persistent lane-refill variant; count[32] is its configuration counter
        int per_cu = t.blocks_per_cu > 0 ? t.blocks_per_cu : fast_pool_blocks_per_cu(env, B);
        per_cu = std::max(1, per_cu);
        const int64_t blocks = std::min<int64_t>((int64_t)num_cus * per_cu, (n + B - 1) / B);
        const dim3 g((unsigned)std::max<int64_t>(1, blocks)), b(B);
        if (env.use_obb)
            hipLaunchKernelGGL(k_distance_fast_pool<true>, g, b, sm, st, rb, env, qp, q, n, out, idx, count,
                               count + 32);
        else
            hipLaunchKernelGGL(k_distance_fast_pool<false>, g, b, sm, st, rb, env, qp, q, n, out, idx, count,
                               count + 32);
        return hipGetLastError();
    }
    const dim3 g((unsigned)((n + B - 1) / B)), b(B);
    if (env.use_obb)
        hipLaunchKernelGGL(k_distance_fast<true>, g, b, sm, st, rb, env, qp, q, n, out, idx, count);
    else
        hipLaunchKernelGGL(k_distance_fast<false>, g, b, sm, st, rb, env, qp, q, n, out, idx, count);
    return hipGetLastError();
}

}  // namespace rvs
