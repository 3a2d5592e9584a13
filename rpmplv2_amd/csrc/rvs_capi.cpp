// C ABI of the capsule collision / distance engine (include/rvs.h).
//
// Host side: robots and obstacle sets are converted once into the constant
// tables the kernels read (segment frames in double, skeleton recipe, radii,
// axis-aligned bounds) and kept resident on the GPU; queries stage host
// arrays through a per-context device arena and launch on the context's
// stream.  Every entry point catches its own failures and returns a status.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/rvs.h"
#include "rvs_context.hpp"
#include "rvs_internal.hpp"
#include "rvs_stage.hpp"

using namespace rvs;

namespace {

rvs_status fail(rvs_context* c, rvs_status s, const std::string& msg) {
    if (c) {
        std::lock_guard<std::mutex> lk(c->err_mu);
        c->err = msg;
    }
    return s;
}

rvs_status hip_fail(rvs_context* c, hipError_t e, const char* what) {
    return fail(c, e == hipErrorOutOfMemory ? RVS_ENOMEM : RVS_EDEVICE,
                std::string(what) + ": " + hipGetErrorString(e));
}

// exact-path batches of at least this many configurations run in the walk order
constexpr int64_t kExactSortMin = 65536;

// Scratch of rvs_distance_async_static: the refinement list and counters and
// the walk-order sort of one call of up to kStaticMaxConfigs configurations
// (19.1 MB at that size), in a device array of fixed size that every GPU the
// library runs on holds until the process ends.  Host code never passes the
// array's name to a kernel: it is a host-side stand-in, not a device
// address.  rvs_create looks up the array's address on its device once
// (static_scratch_addr) and the entry point passes that address.
constexpr int64_t kStaticMaxConfigs = int64_t(1) << 20;
constexpr size_t kStaticScratchBytes = size_t(20) << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

std::mutex g_scratch_mu;
std::map<int, void*> g_scratch_addr;   // device -> address of that device's scratch

// The address of the current device's scratch (device dev), looked up once
// per device; called with dev current.
hipError_t static_scratch_resolve(int dev) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_addr.count(dev)) return hipSuccess;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_addr[dev] = ptr;
    return hipSuccess;
}

void* static_scratch_addr(int dev) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_addr.find(dev);
    return it == g_scratch_addr.end() ? nullptr : it->second;
}

// The obstacle set must fit the LDS image of the kernels (oriented boxes
// take 48 more bytes each than axis-aligned ones).
rvs_status env_fits(rvs_context* c, const EnvDev& env, bool fp64, const char* what) {
    const size_t b = fp64 ? staged_bytes<double>(env) : staged_bytes<float>(env);
    if (b > device_lds_limit(c->opt.device))
        return fail(c, RVS_ERANGE, std::string(what) + ": obstacle set too large for LDS staging");
    return RVS_OK;
}

// ---- host restatement of the joint-origin constants (URDF rpy -> quaternion
// -> rotation as urdfdom and kdl_parser do it; KDL::Joint normalises the axis)
void rpy_to_matrix(const double rpy[3], double M[9]) {
    const double phi = rpy[0] / 2.0, the = rpy[1] / 2.0, psi = rpy[2] / 2.0;
    double x = std::sin(phi) * std::cos(the) * std::cos(psi) - std::cos(phi) * std::sin(the) * std::sin(psi);
    double y = std::cos(phi) * std::sin(the) * std::cos(psi) + std::sin(phi) * std::cos(the) * std::sin(psi);
    double z = std::cos(phi) * std::cos(the) * std::sin(psi) - std::sin(phi) * std::sin(the) * std::cos(psi);
    double w = std::cos(phi) * std::cos(the) * std::cos(psi) + std::sin(phi) * std::sin(the) * std::sin(psi);
    const double s = std::sqrt(x * x + y * y + z * z + w * w);
    x /= s; y /= s; z /= s; w /= s;
    const double x2 = x * x, y2 = y * y, z2 = z * z, w2 = w * w;
    M[0] = w2 + x2 - y2 - z2; M[1] = 2 * x * y - 2 * w * z; M[2] = 2 * x * z + 2 * w * y;
    M[3] = 2 * x * y + 2 * w * z; M[4] = w2 - x2 + y2 - z2; M[5] = 2 * y * z - 2 * w * x;
    M[6] = 2 * x * z - 2 * w * y; M[7] = 2 * y * z + 2 * w * x; M[8] = w2 - x2 - y2 + z2;
}

// Overflow-safe vector norm in the form KDL::Vector::Norm uses.
double safe_norm(const double v[3]) {
    const double a = std::fabs(v[0]), b = std::fabs(v[1]), c = std::fabs(v[2]);
    auto sq = [](double t) { return t * t; };
    if (a >= b) {
        if (a >= c) {
            if (a == 0) return 0;
            return a * std::sqrt(1 + sq(v[1] / v[0]) + sq(v[2] / v[0]));
        }
        return c * std::sqrt(1 + sq(v[0] / v[2]) + sq(v[1] / v[2]));
    }
    if (b > c) return b * std::sqrt(1 + sq(v[0] / v[1]) + sq(v[2] / v[1]));
    return c * std::sqrt(1 + sq(v[0] / v[2]) + sq(v[1] / v[2]));
}

template <typename T>
hipError_t upload(std::vector<void*>& allocs, const T* src, size_t count, const T** dst) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) return e;
    allocs.push_back(p);
    if (count) e = hipMemcpy(p, src, count * sizeof(T), hipMemcpyHostToDevice);
    *dst = static_cast<const T*>(p);
    return e;
}

int prec_of(const rvs_context* c, uint32_t flags) {
    const uint32_t v = (flags >> RVS_F_PREC_SHIFT) & 3u;
    return v ? (int)v - 1 : c->opt.precision;
}


// Device arena carving: host arrays are staged into one allocation.
bool robot_ok(rvs_context* c, int32_t id) { return id >= 0 && id < (int32_t)c->robots.size(); }
bool env_ok(rvs_context* c, int32_t id) { return id >= 0 && id < (int32_t)c->envs.size(); }

}  // namespace

namespace rvs {

// FCL CollisionObject::computeAABB for a box (restated in workloads.py and
// include/rvs.h): identity rotation -> pos -+ dim/2; any other rotation ->
// pos -+ |dim/2| on every axis.
// Eigen's isIdentity(1e-5) on the float rotation matrix of the box's
// quaternion (Quaternion::toRotationMatrix order).
bool box_identity(const rvs_box& b) {
    const float w = b.quat_wxyz[0], x = b.quat_wxyz[1], y = b.quat_wxyz[2], z = b.quat_wxyz[3];
    const float tx = 2.0f * x, ty = 2.0f * y, tz = 2.0f * z;
    const float twx = tx * w, twy = ty * w, twz = tz * w;
    const float txx = tx * x, txy = ty * x, txz = tz * x;
    const float tyy = ty * y, tyz = tz * y, tzz = tz * z;
    const float R[9] = {1.0f - (tyy + tzz), txy - twz, txz + twy, txy + twz, 1.0f - (txx + tzz),
                        tyz - twx, txz - twy, tyz + twx, 1.0f - (txx + tyy)};
    const float prec = 1e-5f;
    bool ident = true;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const float v = R[3 * i + j];
            if (i == j) ident = ident && (std::fabs(v - 1.0f) <= prec * std::min(std::fabs(v), 1.0f));
            else ident = ident && (std::fabs(v) <= prec);
        }
    return ident;
}

void box_bounds(const rvs_box& b, float out[6]) {
    const float h[3] = {0.5f * b.dim[0], 0.5f * b.dim[1], 0.5f * b.dim[2]};
    if (box_identity(b)) {
        for (int k = 0; k < 3; ++k) { out[k] = -h[k] + b.pos[k]; out[3 + k] = h[k] + b.pos[k]; }
    } else {
        const float m0 = -h[0], m1 = -h[1], m2 = -h[2];
        const float rad = std::sqrt(m0 * m0 + (m1 * m1 + m2 * m2));
        for (int k = 0; k < 3; ++k) { out[k] = b.pos[k] - rad; out[3 + k] = b.pos[k] + rad; }
    }
}

void box_obb(const rvs_box& b, float out[15]) {
    // centre, rotation rows (unit quaternion -> matrix), half extents
    double w = b.quat_wxyz[0], x = b.quat_wxyz[1], y = b.quat_wxyz[2], z = b.quat_wxyz[3];
    const double n = std::sqrt(w * w + x * x + y * y + z * z);
    w /= n; x /= n; y /= n; z /= n;
    const double R[9] = {1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y),
                         2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
                         2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)};
    for (int k = 0; k < 3; ++k) out[k] = b.pos[k];
    for (int k = 0; k < 9; ++k) out[3 + k] = (float)R[k];
    for (int k = 0; k < 3; ++k) out[12 + k] = 0.5f * b.dim[k];
}

void obb_extent(const float o[15], float ext[3]) {
    for (int i = 0; i < 3; ++i) {
        const float e = std::fabs(o[3 + 3 * i]) * o[12] + std::fabs(o[4 + 3 * i]) * o[13] + std::fabs(o[5 + 3 * i]) * o[14];
        ext[i] = e * (1.0f + 1e-6f) + 1e-7f;
    }
}

void group_order(const float* cull, int M, int32_t* order) {
    std::vector<double> cen(3 * (size_t)M);
    double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
    for (int j = 0; j < M; ++j)
        for (int k = 0; k < 3; ++k) {
            const double c = 0.5 * ((double)cull[6 * j + k] + (double)cull[6 * j + 3 + k]);
            cen[3 * j + k] = c;
            if (std::isfinite(c)) {
                lo[k] = std::min(lo[k], c);
                hi[k] = std::max(hi[k], c);
            }
        }
    std::vector<uint64_t> key(M);
    for (int j = 0; j < M; ++j) {
        uint64_t code = 0;
        for (int k = 0; k < 3; ++k) {
            const double c = cen[3 * j + k];
            uint64_t u = 0;
            if (std::isfinite(c) && hi[k] > lo[k])
                u = (uint64_t)std::min(1023.0, std::max(0.0, std::floor((c - lo[k]) / (hi[k] - lo[k]) * 1024.0)));
            for (int b = 0; b < 10; ++b) code |= ((u >> b) & 1u) << (3 * b + k);
        }
        key[j] = code;
        order[j] = j;
    }
    std::stable_sort(order, order + M, [&](int32_t a, int32_t b) { return key[a] < key[b]; });
}

void group_bounds(const float* cull, const int32_t* order, int M, float* gcull) {
    for (int g = 0; g < num_groups(M); ++g) {
        float lo[3] = {HUGE_VALF, HUGE_VALF, HUGE_VALF}, hi[3] = {-HUGE_VALF, -HUGE_VALF, -HUGE_VALF};
        bool finite = true;
        for (int m = g * kGroupSize; m < std::min(M, (g + 1) * kGroupSize); ++m) {
            const float* b = cull + 6 * (size_t)order[m];
            for (int k = 0; k < 3; ++k) {
                finite = finite && std::isfinite(b[k]) && std::isfinite(b[3 + k]);
                lo[k] = std::min(lo[k], b[k]);
                hi[k] = std::max(hi[k], b[3 + k]);
            }
        }
        for (int k = 0; k < 3; ++k) {
            gcull[6 * g + k] = finite ? lo[k] : -HUGE_VALF;
            gcull[6 * g + 3 + k] = finite ? hi[k] : HUGE_VALF;
        }
    }
}

QueryParams params_of(const rvs_context* c) {
    QueryParams qp{};
    qp.eq = c->opt.equality_threshold;
    qp.maxd = c->opt.max_distance;
    qp.band = c->opt.band;
    qp.num_checks = c->opt.num_checks;
    return qp;
}

}  // namespace rvs

// ================================================================ C API
extern "C" {

const char* rvs_version(void) { return "rvs-mi355x 0.1 (gfx950)"; }

void rvs_default_options(rvs_options* o) {
    if (!o) return;
    o->device = 0;
    o->precision = RVS_PREC_REF32;
    o->equality_threshold = 1e-4f;
    o->max_distance = 1000.0f;
    o->num_checks = 10;
    o->band = 1e-4f;
}

rvs_status rvs_create(const rvs_options* opt, rvs_context** out) {
    if (!out) return RVS_EINVAL;
    *out = nullptr;
    rvs_context* c = new (std::nothrow) rvs_context();
    if (!c) return RVS_ENOMEM;
    if (opt) c->opt = *opt;
    else rvs_default_options(&c->opt);
    if (c->opt.precision < 0 || c->opt.precision > RVS_PREC_FP32_REFINE) { delete c; return RVS_EINVAL; }
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0 || c->opt.device < 0 || c->opt.device >= ndev) {
        delete c;
        return RVS_EDEVICE;
    }
    e = hipSetDevice(c->opt.device);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&c->num_cus, hipDeviceAttributeMultiprocessorCount, c->opt.device);
    if (const char* v = std::getenv("RVS_FAST_VARIANT")) c->fast_variant = std::atoi(v);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    for (int k = 0; k < 4 && e == hipSuccess; ++k) e = hipEventCreate(&c->ev[k]);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->async_ev, hipEventDisableTiming);
    if (e == hipSuccess) e = static_scratch_resolve(c->opt.device);
    if (e == hipSuccess) (void)device_lds_limit(c->opt.device);   // read once here, not inside a query
    if (e != hipSuccess) { delete c; return RVS_EDEVICE; }
    *out = c;
    return RVS_OK;
}

void rvs_destroy(rvs_context* c) {
    if (!c) return;
    (void)hipSetDevice(c->opt.device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->drain_async();
    for (auto& r : c->robots) for (void* p : r.allocs) (void)hipFree(p);
    for (auto& e : c->envs) for (void* p : e.allocs) (void)hipFree(p);
    for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.exec);
    if (c->count_host) (void)hipHostFree(c->count_host);
    c->arena.release();
    c->staging.release();
    for (int k = 0; k < 4; ++k)
        if (c->ev[k]) (void)hipEventDestroy(c->ev[k]);
    if (c->async_ev) (void)hipEventDestroy(c->async_ev);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* rvs_last_error(const rvs_context* c) {
    if (!c) return "null context";
    // a copy per calling thread: valid until this thread's next call
    thread_local std::string copy;
    std::lock_guard<std::mutex> lk(c->err_mu);
    copy = c->err;
    return copy.c_str();
}
void* rvs_stream(rvs_context* c) { return c ? (void*)c->stream : nullptr; }
int64_t rvs_last_refined(const rvs_context* c) { return c ? c->last_refined : 0; }

int32_t rvs_last_timing(const rvs_context* c, float* stage_ms, int32_t max_stages) {
    if (!c || !stage_ms) return 0;
    const int32_t k = std::min<int32_t>(max_stages, c->nstages);
    for (int32_t i = 0; i < k; ++i) stage_ms[i] = c->stage_ms[i];
    return k;
}

rvs_status rvs_robot_load(rvs_context* c, const rvs_robot_desc* d, int32_t* robot_id) {
    if (!c || !d || !robot_id || !d->joints) return fail(c, RVS_EINVAL, "rvs_robot_load: null argument");
    CtxLock lk(c);
    if (d->num_joints <= 0 || d->num_joints > kMaxSeg) return fail(c, RVS_ERANGE, "rvs_robot_load: joint count");
    if (d->num_links <= 0 || d->num_links > kMaxLinks) return fail(c, RVS_ERANGE, "rvs_robot_load: link count");
    if (!d->capsule_radius) return fail(c, RVS_EINVAL, "rvs_robot_load: capsule radii missing");
    std::vector<SegDev> seg(d->num_joints);
    int ndof = 0;
    for (int k = 0; k < d->num_joints; ++k) {
        const rvs_joint& j = d->joints[k];
        SegDev& s = seg[k];
        std::memset(&s, 0, sizeof(s));
        rpy_to_matrix(j.rpy, s.R);
        for (int i = 0; i < 3; ++i) s.o[i] = j.xyz[i];
        s.rev = (j.type == RVS_JOINT_REVOLUTE) ? 1 : 0;
        ndof += s.rev;
        double a[3];
        for (int i = 0; i < 3; ++i) a[i] = s.R[3 * i] * j.axis[0] + s.R[3 * i + 1] * j.axis[1] + s.R[3 * i + 2] * j.axis[2];
        const double nrm = safe_norm(a);
        if (s.rev && !(nrm > 0)) return fail(c, RVS_EINVAL, "rvs_robot_load: zero joint axis");
        for (int i = 0; i < 3; ++i) s.axis[i] = s.rev ? a[i] / nrm : 0.0;
    }
    // the FP32 bulk pass's float segments: R, K R and K^2 R in double, rounded
    std::vector<SegF> segf(d->num_joints);
    for (int k = 0; k < d->num_joints; ++k) {
        const SegDev& s = seg[k];
        SegF& f = segf[k];
        std::memset(&f, 0, sizeof(f));
        const double* a = s.axis;
        const double K[9] = {0.0, -a[2], a[1], a[2], 0.0, -a[0], -a[1], a[0], 0.0};
        double KR[9], K2R[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                KR[3 * i + j] = K[3 * i] * s.R[j] + K[3 * i + 1] * s.R[3 + j] + K[3 * i + 2] * s.R[6 + j];
            }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                K2R[3 * i + j] = K[3 * i] * KR[j] + K[3 * i + 1] * KR[3 + j] + K[3 * i + 2] * KR[6 + j];
        for (int m = 0; m < 9; ++m) {
            f.R[m] = (float)s.R[m];
            f.KR[m] = (float)KR[m];
            f.K2R[m] = (float)K2R[m];
        }
        for (int i = 0; i < 3; ++i) f.o[i] = (float)s.o[i];
        f.rev = s.rev;
    }
    const int L = d->num_links;
    std::vector<PointDev> pts(L + 1);
    for (auto& p : pts) std::memset(&p, 0, sizeof(p));
    if (d->skeleton_kind == RVS_SKEL_CHAIN) {
        if (d->num_joints < L + 1) return fail(c, RVS_EINVAL, "rvs_robot_load: chain needs num_links+1 segments");
        for (int k = 0; k <= L; ++k) { pts[k].frame = k; pts[k].axis = 0; pts[k].flags = 0; }
    } else if (d->skeleton_kind == RVS_SKEL_XARM6) {
        if (L != 6 || ndof != 6) return fail(c, RVS_EINVAL, "rvs_robot_load: xArm6 recipe needs 6 revolute links");
        const double g = (double)d->gripper_length;
        pts[0].frame = -1;
        pts[1].frame = 1;
        pts[2].frame = 2;
        pts[3].frame = 3; pts[3].axis = 2; pts[3].pre = -0.25; pts[3].flags = 1;
        pts[4].frame = 4;
        pts[5].frame = 4; pts[5].axis = 0; pts[5].pre = 0.076; pts[5].flags = 1;
        pts[6].frame = 5; pts[6].axis = 2;
        if (g != 0) { pts[6].pre = g; pts[6].post = 0.3 * g; pts[6].flags = 3; }
    } else {
        return fail(c, RVS_EINVAL, "rvs_robot_load: unknown skeleton kind");
    }
    std::vector<float> rad(d->capsule_radius, d->capsule_radius + L);
    std::vector<float> rn32(L);
    std::vector<double> rn64(L);
    for (int i = 0; i < L; ++i) {
        if (!(rad[i] >= 0)) return fail(c, RVS_EINVAL, "rvs_robot_load: negative capsule radius");
        // collisionCapsuleToBox's inflated radius (CollisionAndDistance.cpp:7)
        rn32[i] = (float)((double)rad[i] * std::sqrt(3.0) / 3.0);
        rn64[i] = (double)rad[i] * std::sqrt(3.0) / 3.0;
    }
    // radii added in computeEnclosingRadii: xArm6 uses capsules_radius_new,
    // max(r_i, r_i+1) and r_last (xArm6::setCapsulesRadius, xArm6.cpp:98-105);
    // the chains use capsules_radius (Planar2DOF.cpp:144, Spatial10DOF.cpp:18-24)
    std::vector<float> renc(rad);
    if (d->skeleton_kind == RVS_SKEL_XARM6)
        for (int i = 0; i + 1 < L; ++i) renc[i] = std::max(rad[i], rad[i + 1]);
    std::vector<int32_t> pairs;
    for (int m = 0; m < d->num_self_pairs; ++m) {
        const int a = d->self_pairs[2 * m], b = d->self_pairs[2 * m + 1];
        if (a < 0 || b < 0 || a >= L || b >= L) return fail(c, RVS_EINVAL, "rvs_robot_load: self pair out of range");
        pairs.push_back(a);
        pairs.push_back(b);
    }
    (void)hipSetDevice(c->opt.device);
    RobotRec r;
    hipError_t e = upload(r.allocs, seg.data(), seg.size(), &r.dev.seg);
    if (e == hipSuccess) e = upload(r.allocs, segf.data(), segf.size(), &r.dev.segf);
    if (e == hipSuccess) e = upload(r.allocs, pts.data(), pts.size(), &r.dev.pts);
    if (e == hipSuccess) e = upload(r.allocs, rad.data(), rad.size(), &r.dev.radius);
    if (e == hipSuccess) e = upload(r.allocs, rn32.data(), rn32.size(), &r.dev.rnew32);
    if (e == hipSuccess) e = upload(r.allocs, rn64.data(), rn64.size(), &r.dev.rnew64);
    if (e == hipSuccess) e = upload(r.allocs, pairs.data(), pairs.size(), &r.dev.pairs);
    if (e == hipSuccess) e = upload(r.allocs, renc.data(), renc.size(), &r.dev.rad_enc);
    if (e != hipSuccess) {
        for (void* p : r.allocs) (void)hipFree(p);
        return hip_fail(c, e, "rvs_robot_load");
    }
    r.dev.kind = d->skeleton_kind;
    r.dev.num_segments = d->num_joints;
    r.dev.num_dofs = ndof;
    r.dev.num_links = L;
    r.dev.ground_included = d->ground_included;
    r.dev.num_pairs = (int32_t)pairs.size() / 2;
    r.num_dofs = ndof;
    r.num_links = L;
    r.num_pairs = r.dev.num_pairs;
    r.radius = rad;
    c->robots.push_back(r);
    *robot_id = (int32_t)c->robots.size() - 1;
    return RVS_OK;
}

rvs_status rvs_env_create(rvs_context* c, const rvs_box* boxes, int32_t num, int32_t semantics, int32_t* env_id) {
    if (!c || !env_id || (num > 0 && !boxes)) return fail(c, RVS_EINVAL, "rvs_env_create: null argument");
    if (num < 0 || num > kMaxLdsObstacles) return fail(c, RVS_ERANGE, "rvs_env_create: at most 1024 boxes");
    if (semantics != RVS_ENV_AABB && semantics != RVS_ENV_OBB) return fail(c, RVS_EINVAL, "rvs_env_create: semantics");
    CtxLock lk(c);
    EnvRec r;
    r.semantics = semantics;
    r.boxes.assign(boxes, boxes + num);
    r.aabb.resize((size_t)num * 6);
    r.obb.resize((size_t)num * 15);
    r.cull.resize((size_t)num * 6);
    std::vector<float> tol(num);
    std::vector<uint8_t> fl(num);
    for (int j = 0; j < num; ++j) {
        box_bounds(boxes[j], &r.aabb[6 * j]);
        box_obb(boxes[j], &r.obb[15 * j]);
        float ext[3];
        obb_extent(&r.obb[15 * j], ext);
        for (int k = 0; k < 3; ++k) {
            r.cull[6 * j + k] = r.obb[15 * j + k] - ext[k];
            r.cull[6 * j + 3 + k] = r.obb[15 * j + k] + ext[k];
        }
        tol[j] = boxes[j].min_dist_tol;
        fl[j] = (uint8_t)((boxes[j].flags & RVS_BOX_GROUND ? kObjGround : 0) |
                          (boxes[j].flags & RVS_BOX_DYNAMIC ? kObjDynamic : 0));
    }
    (void)hipSetDevice(c->opt.device);
    hipError_t e = upload(r.allocs, r.aabb.data(), r.aabb.size(), &r.dev.aabb);
    if (e == hipSuccess) e = upload(r.allocs, tol.data(), tol.size(), &r.dev.tol);
    if (e == hipSuccess) e = upload(r.allocs, fl.data(), fl.size(), &r.dev.flags);
    if (e == hipSuccess) e = upload(r.allocs, r.obb.data(), r.obb.size(), &r.dev.obb);
    if (e == hipSuccess) e = upload(r.allocs, r.cull.data(), r.cull.size(), &r.dev.cull);
    // the grouped walk of the broad phase over the boxes the kernels bound
    // each obstacle by: the AABB itself (reference semantics) or the oriented
    // box's tight world AABB
    const float* cull_src = (semantics == RVS_ENV_AABB) ? r.aabb.data() : r.cull.data();
    r.gorder.resize(num);
    r.gcull.resize((size_t)num_groups(num) * 6);
    if (num > 0) {
        group_order(cull_src, num, r.gorder.data());
        group_bounds(cull_src, r.gorder.data(), num, r.gcull.data());
    }
    if (e == hipSuccess) e = upload(r.allocs, r.gorder.data(), r.gorder.size(), &r.dev.gorder);
    if (e == hipSuccess) e = upload(r.allocs, r.gcull.data(), r.gcull.size(), &r.dev.gcull);
    if (e != hipSuccess) {
        for (void* p : r.allocs) (void)hipFree(p);
        return hip_fail(c, e, "rvs_env_create");
    }
    // reference semantics: the broad phase bounds each box by its AABB itself
    if (semantics == RVS_ENV_AABB) r.dev.cull = r.dev.aabb;
    r.dev.num_obj = num;
    r.dev.use_obb = semantics == RVS_ENV_OBB;
    c->envs.push_back(std::move(r));
    *env_id = (int32_t)c->envs.size() - 1;
    return RVS_OK;
}

rvs_status rvs_env_update(rvs_context* c, int32_t env_id, const int32_t* idx, const float* pos, int32_t count,
                          uint32_t flags) {
    if (!c || (count > 0 && (!idx || !pos))) return fail(c, RVS_EINVAL, "rvs_env_update: null argument");
    if (flags & RVS_F_DEVICE) return fail(c, RVS_ENOTSUP, "rvs_env_update: host arrays only");
    CtxLock lk(c);
    if (!env_ok(c, env_id)) return fail(c, RVS_EINVAL, "rvs_env_update: bad env id");
    EnvRec& r = c->envs[env_id];
    for (int k = 0; k < count; ++k)
        if (idx[k] < 0 || idx[k] >= (int)r.boxes.size()) return fail(c, RVS_EINVAL, "rvs_env_update: bad box index");
    if (count == 0) return RVS_OK;
    (void)hipSetDevice(c->opt.device);
    for (int k = 0; k < count; ++k) {
        const int j = idx[k];
        for (int i = 0; i < 3; ++i) r.boxes[j].pos[i] = pos[3 * k + i];
        box_bounds(r.boxes[j], &r.aabb[6 * j]);
        box_obb(r.boxes[j], &r.obb[15 * j]);
        float ext[3];
        obb_extent(&r.obb[15 * j], ext);
        for (int i = 0; i < 3; ++i) {
            r.cull[6 * j + i] = r.obb[15 * j + i] - ext[i];
            r.cull[6 * j + 3 + i] = r.obb[15 * j + i] + ext[i];
        }
    }
    // the tables are small (<= 1024 boxes): copy them whole, once
    hipError_t e = hipMemcpyAsync((void*)r.dev.aabb, r.aabb.data(), r.aabb.size() * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync((void*)r.dev.obb, r.obb.data(), r.obb.size() * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && r.dev.cull != r.dev.aabb)
        e = hipMemcpyAsync((void*)r.dev.cull, r.cull.data(), r.cull.size() * 4, hipMemcpyHostToDevice, c->stream);
    // group boxes of the broad phase (the walk order stays)
    group_bounds(r.dev.cull == r.dev.aabb ? r.aabb.data() : r.cull.data(), r.gorder.data(), (int)r.boxes.size(),
                 r.gcull.data());
    if (e == hipSuccess && !r.gcull.empty())
        e = hipMemcpyAsync((void*)r.dev.gcull, r.gcull.data(), r.gcull.size() * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "rvs_env_update");
    return RVS_OK;
}

rvs_status rvs_env_set_motion_law(rvs_context* c, int32_t env_id, const rvs_env_motion* law, const float* vel,
                                  const float* max_vel) {
    if (!c || !law || !vel || !max_vel) return fail(c, RVS_EINVAL, "rvs_env_set_motion_law: null argument");
    if (law->motion_type < RVS_MOTION_STRAIGHT || law->motion_type > RVS_MOTION_LIGHT_DIRECTIONS)
        return fail(c, RVS_EINVAL, "rvs_env_set_motion_law: unknown motion type");
    CtxLock lk(c);
    if (!env_ok(c, env_id)) return fail(c, RVS_EINVAL, "rvs_env_set_motion_law: bad env id");
    EnvRec& r = c->envs[env_id];
    const int M = (int)r.boxes.size();
    (void)hipSetDevice(c->opt.device);
    std::vector<float> pos(3 * M), v(vel, vel + 3 * (size_t)M), mv(max_vel, max_vel + M), half(3 * M), rad(M),
        ext(3 * M), plen(M, 0.0f);
    std::vector<int32_t> sign(M, 1);
    std::vector<uint8_t> dyn(M), ident(M);
    for (int j = 0; j < M; ++j) {
        const rvs_box& b = r.boxes[j];
        for (int k = 0; k < 3; ++k) {
            pos[3 * j + k] = b.pos[k];
            half[3 * j + k] = 0.5f * b.dim[k];
        }
        const float m0 = -half[3 * j], m1 = -half[3 * j + 1], m2 = -half[3 * j + 2];
        rad[j] = std::sqrt(m0 * m0 + (m1 * m1 + m2 * m2));
        ident[j] = box_identity(b) ? 1 : 0;
        dyn[j] = (b.flags & RVS_BOX_DYNAMIC) ? 1 : 0;
        obb_extent(&r.obb[15 * (size_t)j], &ext[3 * j]);
    }
    EnvStepDev m{};
    m.law = law->motion_type;
    m.num_obj = M;
    for (int k = 0; k < 3; ++k) m.ws[k] = law->ws_center[k];
    m.ws_radius = law->ws_radius;
    m.base_radius = law->base_radius;
    m.robot_max_vel = law->robot_max_vel;
    hipError_t e = upload(r.allocs, pos.data(), pos.size(), const_cast<const float**>(&m.pos));
    if (e == hipSuccess) e = upload(r.allocs, v.data(), v.size(), const_cast<const float**>(&m.vel));
    if (e == hipSuccess) e = upload(r.allocs, plen.data(), plen.size(), const_cast<const float**>(&m.path_len));
    if (e == hipSuccess) e = upload(r.allocs, sign.data(), sign.size(), const_cast<const int32_t**>(&m.sign));
    if (e == hipSuccess) e = upload(r.allocs, mv.data(), mv.size(), &m.max_vel);
    if (e == hipSuccess) e = upload(r.allocs, dyn.data(), dyn.size(), &m.dynamic);
    if (e == hipSuccess) e = upload(r.allocs, half.data(), half.size(), &m.half);
    if (e == hipSuccess) e = upload(r.allocs, rad.data(), rad.size(), &m.rad);
    if (e == hipSuccess) e = upload(r.allocs, ident.data(), ident.size(), &m.ident);
    if (e == hipSuccess) e = upload(r.allocs, ext.data(), ext.size(), &m.ext);
    if (e != hipSuccess) return hip_fail(c, e, "rvs_env_set_motion_law");
    r.step = m;
    r.has_motion = true;
    return RVS_OK;
}

rvs_status rvs_env_step(rvs_context* c, int32_t env_id, float delta_time, int32_t steps, float* pos, float* vel) {
    if (!c) return RVS_EINVAL;
    if (steps < 0) return fail(c, RVS_EINVAL, "rvs_env_step: negative step count");
    CtxLock lk(c);
    if (!env_ok(c, env_id)) return fail(c, RVS_EINVAL, "rvs_env_step: bad env id");
    EnvRec& r = c->envs[env_id];
    if (!r.has_motion) return fail(c, RVS_EINVAL, "rvs_env_step: no motion law set");
    const int M = (int)r.boxes.size();
    (void)hipSetDevice(c->opt.device);
    float* aabb = const_cast<float*>(r.dev.aabb);
    float* obb = const_cast<float*>(r.dev.obb);
    float* cull = (r.dev.cull == r.dev.aabb) ? nullptr : const_cast<float*>(r.dev.cull);
    hipError_t e = launch_env_step(r.step, delta_time, steps, aabb, obb, cull, c->stream);
    if (e == hipSuccess && M > 0)
        e = launch_group_bounds(r.dev.cull, r.dev.gorder, M, const_cast<float*>(r.dev.gcull), c->stream);
    // keep the host copies (bounds, box positions) in step with the device
    std::vector<float> p(3 * (size_t)M);
    if (e == hipSuccess) e = hipMemcpyAsync(p.data(), r.step.pos, p.size() * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && vel) e = hipMemcpyAsync(vel, r.step.vel, p.size() * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(r.aabb.data(), r.dev.aabb, r.aabb.size() * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(r.obb.data(), r.dev.obb, r.obb.size() * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && cull)
        e = hipMemcpyAsync(r.cull.data(), r.dev.cull, r.cull.size() * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "rvs_env_step");
    for (int j = 0; j < M; ++j) {
        for (int k = 0; k < 3; ++k) r.boxes[j].pos[k] = p[3 * j + k];
        if (!cull) {   // the host copy of the oriented boxes' bounds
            float ext[3];
            obb_extent(&r.obb[15 * (size_t)j], ext);
            for (int k = 0; k < 3; ++k) {
                r.cull[6 * j + k] = r.obb[15 * j + k] - ext[k];
                r.cull[6 * j + 3 + k] = r.obb[15 * j + k] + ext[k];
            }
        }
    }
    if (pos) std::memcpy(pos, p.data(), p.size() * 4);
    return RVS_OK;
}

rvs_status rvs_env_bounds(rvs_context* c, int32_t env_id, float* aabb, int32_t num) {
    if (!c || !aabb) return fail(c, RVS_EINVAL, "rvs_env_bounds: null argument");
    CtxLock lk(c);
    if (!env_ok(c, env_id)) return fail(c, RVS_EINVAL, "rvs_env_bounds: bad env id");
    const EnvRec& r = c->envs[env_id];
    if (num != (int)r.boxes.size()) return fail(c, RVS_EINVAL, "rvs_env_bounds: size mismatch");
    std::memcpy(aabb, r.aabb.data(), r.aabb.size() * sizeof(float));
    return RVS_OK;
}

// ---------------------------------------------------------------- queries
namespace {

rvs_status check_query(rvs_context* c, int32_t robot, int32_t env, const void* q, int64_t n, bool need_env) {
    if (!c) return RVS_EINVAL;
    if (!robot_ok(c, robot)) return fail(c, RVS_EINVAL, "bad robot id");
    if (need_env && !env_ok(c, env)) return fail(c, RVS_EINVAL, "bad env id");
    if (n < 0) return fail(c, RVS_EINVAL, "negative configuration count");
    if (n > 0 && !q) return fail(c, RVS_EINVAL, "null configuration array");
    return RVS_OK;
}

hipError_t h2d(void* dst, const void* src, size_t bytes, hipStream_t s) {
    return bytes ? hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s) : hipSuccess;
}
hipError_t d2h(void* dst, const void* src, size_t bytes, hipStream_t s) {
    return (bytes && dst) ? hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s) : hipSuccess;
}

}  // namespace

namespace {

// The FP32 + refinement pipeline of a device-array rvs_distance call (zero
// the counters, bulk pass, refinement, read back the refinement count into
// pinned memory) is captured once per (robot, environment, arrays, size,
// arena, launch shape) into a HIP graph and replayed; the cache keeps the 8
// most recently used.  Kernel arguments hold pointers only, so in-place
// updates of the inputs, of the obstacle tables and of the robot are seen
// by every replay.  The graph holds no memset nodes (counters are cleared by
// a kernel).
constexpr size_t kMaxGraphs = 8;

hipError_t fp32_graph_run(rvs_context* c, int32_t robot, int32_t env, const RobotDev& rb, const EnvDev& ev,
                          const QueryParams& qp, const float* q, int64_t n, const DistOut& out, int32_t* ridx,
                          uint32_t* rcount, uint32_t* coh, uint32_t* links, const FastTuning& ft, uint32_t* refined) {
    hipError_t e = hipSuccess;
    if (!c->count_host) {
        e = hipHostMalloc(reinterpret_cast<void**>(&c->count_host), 256, hipHostMallocMapped | hipHostMallocCoherent);
        if (e != hipSuccess) return e;
        e = hipHostGetDevicePointer(reinterpret_cast<void**>(&c->count_dev), c->count_host, 0);
        if (e != hipSuccess) {
            (void)hipHostFree(c->count_host);
            c->count_host = nullptr;
            c->count_dev = nullptr;
            return e;
        }
    }
    rvs_context::Graph* hit = nullptr;
    for (auto& g : c->graphs)
        if (g.robot == robot && g.env == env && g.q == q && g.d == out.d && g.profile == out.profile &&
            g.verdict == out.verdict && g.arena == c->arena.p && g.n == n && g.t.variant == ft.variant &&
            g.t.block == ft.block && g.t.blocks_per_cu == ft.blocks_per_cu && g.t.coh_bits == ft.coh_bits) {
            hit = &g;
            break;
        }
    if (!hit) {
        if (c->graphs.size() == kMaxGraphs) {
            auto old = std::min_element(c->graphs.begin(), c->graphs.end(),
                                        [](const rvs_context::Graph& a, const rvs_context::Graph& b) {
                                            return a.used < b.used;
                                        });
            (void)hipGraphExecDestroy(old->exec);
            c->graphs.erase(old);
        }
        hipGraph_t g = nullptr;
        e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return e;
        // the walk order's first kernel clears the counters (a separate
        // clear without it), and the refinement stores the flagged count in
        // the mapped host word: no copy node
        const int32_t* perm = nullptr;
        hipError_t le = launch_coherence_order(rb, q, n, ft, coh, &perm, c->stream, rcount, 64);
        if (le == hipSuccess && !perm) le = launch_zero_u32(rcount, 64, c->stream);
        if (le == hipSuccess)
            le = launch_distance_fast(rb, ev, qp, q, n, out, ridx, rcount, ft, perm, links, c->num_cus, c->stream);
        if (le == hipSuccess)
            le = launch_distance_refine(rb, ev, qp, q, n, out, ridx, rcount, c->stream, c->count_dev);
        e = hipStreamEndCapture(c->stream, &g);
        if (le != hipSuccess) e = le;
        hipGraphExec_t exec = nullptr;
        if (e == hipSuccess) e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        if (e != hipSuccess) return e;
        c->graphs.push_back(rvs_context::Graph{robot, env, q, out.d, out.profile, out.verdict, c->arena.p, n, ft,
                                               exec, 0});
        hit = &c->graphs.back();
    }
    hit->used = ++c->graph_clock;
    e = hipGraphLaunch(hit->exec, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) *refined = *c->count_host;
    return e;
}

}  // namespace

rvs_status rvs_distance(rvs_context* c, int32_t robot, int32_t env, const float* q, int64_t n, float* d,
                        float* profile, float* nearest, uint8_t* verdict, uint32_t flags) {
    rvs_status s = check_query(c, robot, env, q, n, true);
    if (s != RVS_OK) return s;
    CtxLock lk(c);
    const RobotRec& rb = c->robots[robot];
    const EnvRec& ev = c->envs[env];
    int prec = prec_of(c, flags);
    s = env_fits(c, ev.dev, prec == RVS_PREC_FP64, "rvs_distance");
    if (s != RVS_OK) return s;
    // nearest points come from the reference-exact pass
    if (prec == RVS_PREC_FP32_REFINE && nearest) prec = RVS_PREC_REF32;
    if (n == 0) return RVS_OK;
    if (prec == RVS_PREC_FP32_REFINE && n > INT32_MAX) return fail(c, RVS_ERANGE, "rvs_distance: batch too large");
    (void)hipSetDevice(c->opt.device);
    const int L = rb.num_links, D = rb.num_dofs, M = ev.dev.num_obj;
    const QueryParams qp = params_of(c);
    DistOut out{};
    const float* qd = q;
    hipError_t e = hipSuccess;
    // refinement list (FP32_REFINE) lives at the start of the arena
    const size_t refine_bytes = (prec == RVS_PREC_FP32_REFINE) ? (size_t)n * 4 + 256 : 0;
    // the exact paths walk large batches in the walk order too (lanes of a
    // wave then hold similar poses and take the same branches); below this
    // size the sort costs more than it saves
    const bool sort_exact = prec != RVS_PREC_FP32_REFINE && n >= kExactSortMin && n <= INT32_MAX &&
                            (prec == RVS_PREC_FP64 || n > c->wave_max);
    const size_t coh_need = (prec == RVS_PREC_FP32_REFINE || sort_exact) ? coh_bytes(n, D) : 0;
    const FastTuning ft = c->tuning_for(robot, env, L, M);
    const size_t links_need = (prec == RVS_PREC_FP32_REFINE && ft.variant == 4) ? fast_links_bytes(n) : 0;
    int32_t* ridx = nullptr;
    uint32_t* rcount = nullptr;
    uint32_t* coh = nullptr;
    uint32_t* links = nullptr;
    if (flags & RVS_F_DEVICE) {
        e = c->arena.ensure(Carve::need({refine_bytes, coh_need, links_need}));
        if (e != hipSuccess) return hip_fail(c, e, "rvs_distance: staging");
        Carve cv(c->arena.p);
        if (refine_bytes) { rcount = cv.take<uint32_t>(64); ridx = cv.take<int32_t>(n); }
        if (coh_need) coh = cv.take<uint32_t>(coh_need / 4);
        if (links_need) links = cv.take<uint32_t>(links_need / 4);
        out.d = d; out.profile = profile; out.nearest = nearest; out.verdict = verdict;
    } else {
        const size_t need = Carve::need({refine_bytes, coh_need, links_need, (size_t)n * D * 4, (size_t)n * 4,
                                         profile ? (size_t)n * L * 4 : 0, nearest ? (size_t)n * M * L * 24 : 0,
                                         (size_t)n});
        e = c->arena.ensure(need);
        if (e != hipSuccess) return hip_fail(c, e, "rvs_distance: staging");
        Carve cv(c->arena.p);
        if (refine_bytes) { rcount = cv.take<uint32_t>(64); ridx = cv.take<int32_t>(n); }
        if (coh_need) coh = cv.take<uint32_t>(coh_need / 4);
        if (links_need) links = cv.take<uint32_t>(links_need / 4);
        float* qs = cv.take<float>((size_t)n * D);
        out.d = cv.take<float>(n);
        out.profile = profile ? cv.take<float>((size_t)n * L) : nullptr;
        out.nearest = nearest ? cv.take<float>((size_t)n * M * L * 6) : nullptr;
        out.verdict = cv.take<uint8_t>(n);
        e = h2d(qs, q, (size_t)n * D * 4, c->stream);
        qd = qs;
    }
    uint32_t refined = 0;
    const bool timing = (flags & RVS_F_TIMING) != 0;
    if (e == hipSuccess) {
        const bool graph = prec == RVS_PREC_FP32_REFINE && (flags & RVS_F_DEVICE) && !timing &&
                           !(flags & RVS_F_EAGER);
        if (graph) {
            e = fp32_graph_run(c, robot, env, rb.dev, ev.dev, qp, qd, n, out, ridx, rcount, coh, links, ft,
                               &refined);
        } else if (prec == RVS_PREC_FP32_REFINE) {
            e = launch_zero_u32(rcount, 64, c->stream);
            const int32_t* perm = nullptr;
            if (e == hipSuccess && timing) e = hipEventRecord(c->ev[3], c->stream);
            if (e == hipSuccess) e = launch_coherence_order(rb.dev, qd, n, ft, coh, &perm, c->stream);
            if (e == hipSuccess && timing) e = hipEventRecord(c->ev[0], c->stream);
            if (e == hipSuccess)
                e = launch_distance_fast(rb.dev, ev.dev, qp, qd, n, out, ridx, rcount, ft, perm, links, c->num_cus,
                                         c->stream);
            if (e == hipSuccess && timing) e = hipEventRecord(c->ev[1], c->stream);
            if (e == hipSuccess) e = launch_distance_refine(rb.dev, ev.dev, qp, qd, n, out, ridx, rcount, c->stream);
            if (e == hipSuccess && timing) e = hipEventRecord(c->ev[2], c->stream);
            if (e == hipSuccess) e = d2h(&refined, rcount, 4, c->stream);
        } else {
            if (e == hipSuccess && timing) e = hipEventRecord(c->ev[0], c->stream);
            const int32_t* perm = nullptr;
            if (e == hipSuccess && sort_exact)
                e = launch_coherence_order(rb.dev, qd, n, c->tuning_for(robot, env, L, M), coh, &perm, c->stream);
            if (e == hipSuccess)
                e = (prec == RVS_PREC_FP64)
                        ? launch_distance_fp64(rb.dev, ev.dev, qp, qd, n, out, c->stream, perm)
                        : launch_distance_ref32_auto(rb.dev, ev.dev, qp, qd, n, out, c->wave_max, c->stream, perm);
            if (e == hipSuccess && timing) e = hipEventRecord(c->ev[1], c->stream);
        }
    }
    if (e == hipSuccess && !(flags & RVS_F_DEVICE)) {
        e = d2h(d, out.d, (size_t)n * 4, c->stream);
        if (e == hipSuccess) e = d2h(profile, out.profile, (size_t)n * L * 4, c->stream);
        if (e == hipSuccess) e = d2h(nearest, out.nearest, (size_t)n * M * L * 24, c->stream);
        if (e == hipSuccess) e = d2h(verdict, out.verdict, (size_t)n, c->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "rvs_distance");
    c->last_refined = (prec == RVS_PREC_FP32_REFINE) ? (int64_t)refined : 0;
    if (timing) {
        c->nstages = prec == RVS_PREC_FP32_REFINE ? 3 : 2;
        c->stage_ms[0] = c->stage_ms[1] = c->stage_ms[2] = 0;
        (void)hipEventElapsedTime(&c->stage_ms[0], c->ev[0], c->ev[1]);
        if (prec == RVS_PREC_FP32_REFINE) {
            (void)hipEventElapsedTime(&c->stage_ms[1], c->ev[1], c->ev[2]);
            (void)hipEventElapsedTime(&c->stage_ms[2], c->ev[3], c->ev[0]);
        }
    }
    return RVS_OK;
}

namespace {

// Stream-ordered start of an asynchronous call: after every earlier one.
hipError_t async_begin(rvs_context* c, hipStream_t st) {
    return c->async_pending ? hipStreamWaitEvent(st, c->async_ev, 0) : hipSuccess;
}

// End of an asynchronous call: the context's event and the caller's.
hipError_t async_end(rvs_context* c, hipStream_t st, void* done) {
    hipError_t e = hipEventRecord(c->async_ev, st);
    if (e == hipSuccess) c->async_pending = true;
    if (e == hipSuccess && done) e = hipEventRecord(static_cast<hipEvent_t>(done), st);
    return e;
}

}  // namespace

namespace {

// Walk order for a large exact-path batch of an async call (nullptr below
// kExactSortMin): the sort's scratch comes from the arena, which async calls
// use one after another (async_begin orders each after the previous one).
hipError_t async_exact_order(rvs_context* c, int32_t robot, int32_t env, const RobotRec& rb, const EnvRec& ev,
                             const float* q, int64_t n, hipStream_t st, const int32_t** perm) {
    *perm = nullptr;
    if (n < kExactSortMin || n > INT32_MAX) return hipSuccess;
    const size_t need = Carve::need({coh_bytes(n, rb.num_dofs)});
    if (need > c->arena.cap) c->drain_async();   // growing frees the block earlier calls use
    hipError_t e = c->arena.ensure(need);
    if (e != hipSuccess) return e;
    Carve cv(c->arena.p);
    uint32_t* coh = cv.take<uint32_t>(coh_bytes(n, rb.num_dofs) / 4);
    return launch_coherence_order(rb.dev, q, n, c->tuning_for(robot, env, rb.num_links, ev.dev.num_obj), coh, perm,
                                  st);
}

}  // namespace

rvs_status rvs_distance_async(rvs_context* c, int32_t robot, int32_t env, const float* q, int64_t n, float* d,
                              float* profile, uint8_t* verdict, uint32_t flags, void* stream, void* done) {
    rvs_status s = check_query(c, robot, env, q, n, true);
    if (s != RVS_OK) return s;
    if (!(flags & RVS_F_DEVICE)) return fail(c, RVS_EINVAL, "rvs_distance_async: device arrays only (RVS_F_DEVICE)");
    if (flags & RVS_F_TIMING) return fail(c, RVS_EINVAL, "rvs_distance_async: no timing");
    std::lock_guard<std::mutex> lk(c->mu);
    const RobotRec& rb = c->robots[robot];
    const EnvRec& ev = c->envs[env];
    const int prec = prec_of(c, flags);
    s = env_fits(c, ev.dev, prec == RVS_PREC_FP64, "rvs_distance_async");
    if (s != RVS_OK) return s;
    if (prec == RVS_PREC_FP32_REFINE && n > INT32_MAX) return fail(c, RVS_ERANGE, "rvs_distance_async: batch too large");
    (void)hipSetDevice(c->opt.device);
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
    const int L = rb.num_links, D = rb.num_dofs, M = ev.dev.num_obj;
    const QueryParams qp = params_of(c);
    DistOut out{};
    out.d = d; out.profile = profile; out.verdict = verdict;
    hipError_t e = async_begin(c, st);
    if (e == hipSuccess && n > 0) {
        if (prec == RVS_PREC_FP32_REFINE) {
            const FastTuning ft = c->tuning_for(robot, env, L, M);
            const size_t links_need = ft.variant == 4 ? fast_links_bytes(n) : 0;
            const size_t need = Carve::need({(size_t)n * 4 + 256, coh_bytes(n, D), links_need});
            if (need > c->arena.cap) c->drain_async();   // growing frees the block earlier calls use
            e = c->arena.ensure(need);
            if (e != hipSuccess) return hip_fail(c, e, "rvs_distance_async: staging");
            Carve cv(c->arena.p);
            uint32_t* rcount = cv.take<uint32_t>(64);
            int32_t* ridx = cv.take<int32_t>(n);
            uint32_t* coh = cv.take<uint32_t>(coh_bytes(n, D) / 4);
            uint32_t* links = links_need ? cv.take<uint32_t>(links_need / 4) : nullptr;
            const int32_t* perm = nullptr;
            e = launch_zero_u32(rcount, 64, st);
            if (e == hipSuccess) e = launch_coherence_order(rb.dev, q, n, ft, coh, &perm, st);
            if (e == hipSuccess)
                e = launch_distance_fast(rb.dev, ev.dev, qp, q, n, out, ridx, rcount, ft, perm, links, c->num_cus,
                                         st);
            if (e == hipSuccess) e = launch_distance_refine(rb.dev, ev.dev, qp, q, n, out, ridx, rcount, st);
        } else {
            const int32_t* perm = nullptr;
            if (prec == RVS_PREC_FP64 || n > c->wave_max) e = async_exact_order(c, robot, env, rb, ev, q, n, st, &perm);
            if (e == hipSuccess)
                e = prec == RVS_PREC_FP64 ? launch_distance_fp64(rb.dev, ev.dev, qp, q, n, out, st, perm)
                                          : launch_distance_ref32_auto(rb.dev, ev.dev, qp, q, n, out, c->wave_max, st,
                                                                       perm);
        }
    }
    if (e == hipSuccess) e = async_end(c, st, done);
    if (e != hipSuccess) return hip_fail(c, e, "rvs_distance_async");
    return RVS_OK;
}

rvs_status rvs_distance_async_static(rvs_context* c, int32_t robot, int32_t env, const float* q, int64_t n,
                                     float* d, float* profile, uint8_t* verdict, uint32_t flags, void* stream,
                                     void* done) {
    rvs_status s = check_query(c, robot, env, q, n, true);
    if (s != RVS_OK) return s;
    if (!(flags & RVS_F_DEVICE))
        return fail(c, RVS_EINVAL, "rvs_distance_async_static: device arrays only (RVS_F_DEVICE)");
    if (flags & RVS_F_TIMING) return fail(c, RVS_EINVAL, "rvs_distance_async_static: no timing");
    if (n > kStaticMaxConfigs)
        return fail(c, RVS_ERANGE, "rvs_distance_async_static: at most 1048576 configurations per call "
                                   "(the size of its static scratch)");
    std::lock_guard<std::mutex> lk(c->mu);
    const RobotRec& rb = c->robots[robot];
    const EnvRec& ev = c->envs[env];
    const int prec = prec_of(c, flags);
    s = env_fits(c, ev.dev, prec == RVS_PREC_FP64, "rvs_distance_async_static");
    if (s != RVS_OK) return s;
    void* base = static_scratch_addr(c->opt.device);
    if (!base) return fail(c, RVS_EDEVICE, "rvs_distance_async_static: no scratch address for this device");
    const int L = rb.num_links, D = rb.num_dofs, M = ev.dev.num_obj;
    const bool refine = prec == RVS_PREC_FP32_REFINE;
    const bool sort = refine || (n >= kExactSortMin && (prec == RVS_PREC_FP64 || n > c->wave_max));
    const size_t need = Carve::need({refine ? (size_t)n * 4 + 256 : 0, sort ? coh_bytes(n, D) : 0});
    if (need > kStaticScratchBytes)
        return fail(c, RVS_ERANGE, "rvs_distance_async_static: batch needs more scratch than the static array holds");
    (void)hipSetDevice(c->opt.device);
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
    const QueryParams qp = params_of(c);
    DistOut out{};
    out.d = d; out.profile = profile; out.verdict = verdict;
    const FastTuning ft = c->tuning_for(robot, env, L, M);
    hipError_t e = async_begin(c, st);
    if (e == hipSuccess && n > 0) {
        Carve cv(base);
        uint32_t* rcount = refine ? cv.take<uint32_t>(64) : nullptr;
        int32_t* ridx = refine ? cv.take<int32_t>(n) : nullptr;
        uint32_t* coh = sort ? cv.take<uint32_t>(coh_bytes(n, D) / 4) : nullptr;
        const int32_t* perm = nullptr;
        if (refine) {
            // counters cleared by a kernel: the scratch keeps the last call's contents
            e = launch_zero_u32(rcount, 64, st);
            if (e == hipSuccess) e = launch_coherence_order(rb.dev, q, n, ft, coh, &perm, st);
            // no links scratch in the static array: a link-major tuning runs as variant 3
            if (e == hipSuccess)
                e = launch_distance_fast(rb.dev, ev.dev, qp, q, n, out, ridx, rcount, ft, perm, nullptr, c->num_cus,
                                         st);
            if (e == hipSuccess) e = launch_distance_refine(rb.dev, ev.dev, qp, q, n, out, ridx, rcount, st);
        } else {
            if (sort) e = launch_coherence_order(rb.dev, q, n, ft, coh, &perm, st);
            if (e == hipSuccess)
                e = prec == RVS_PREC_FP64 ? launch_distance_fp64(rb.dev, ev.dev, qp, q, n, out, st, perm)
                                          : launch_distance_ref32_auto(rb.dev, ev.dev, qp, q, n, out, c->wave_max, st,
                                                                       perm);
        }
    }
    if (e == hipSuccess) e = async_end(c, st, done);
    if (e != hipSuccess) return hip_fail(c, e, "rvs_distance_async_static");
    return RVS_OK;
}

rvs_status rvs_is_valid_async(rvs_context* c, int32_t robot, int32_t env, const float* q, int64_t n, uint8_t* valid,
                              uint32_t flags, void* stream, void* done) {
    rvs_status s = check_query(c, robot, env, q, n, true);
    if (s != RVS_OK) return s;
    if (n > 0 && !valid) return fail(c, RVS_EINVAL, "rvs_is_valid_async: null output");
    if (!(flags & RVS_F_DEVICE)) return fail(c, RVS_EINVAL, "rvs_is_valid_async: device arrays only (RVS_F_DEVICE)");
    std::lock_guard<std::mutex> lk(c->mu);
    const RobotRec& rb = c->robots[robot];
    const EnvRec& ev = c->envs[env];
    const int fp64 = prec_of(c, flags) == RVS_PREC_FP64;
    s = env_fits(c, ev.dev, fp64, "rvs_is_valid_async");
    if (s != RVS_OK) return s;
    (void)hipSetDevice(c->opt.device);
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
    hipError_t e = async_begin(c, st);
    const int32_t* perm = nullptr;
    if (e == hipSuccess && n > 0) e = async_exact_order(c, robot, env, rb, ev, q, n, st, &perm);
    if (e == hipSuccess && n > 0) e = launch_is_valid(rb.dev, ev.dev, params_of(c), q, n, valid, fp64, st, perm);
    if (e == hipSuccess) e = async_end(c, st, done);
    if (e != hipSuccess) return hip_fail(c, e, "rvs_is_valid_async");
    return RVS_OK;
}

rvs_status rvs_distance_underestimation_async(rvs_context* c, int32_t robot, int32_t env, const float* q, int64_t n,
                                              const float* nearest_ref, int64_t ref_stride, float* d, float* profile,
                                              uint8_t* stored, uint32_t flags, void* stream, void* done) {
    rvs_status s = check_query(c, robot, env, q, n, true);
    if (s != RVS_OK) return s;
    if (n > 0 && (!nearest_ref || !d)) return fail(c, RVS_EINVAL, "rvs_distance_underestimation_async: null argument");
    if (!(flags & RVS_F_DEVICE))
        return fail(c, RVS_EINVAL, "rvs_distance_underestimation_async: device arrays only (RVS_F_DEVICE)");
    std::lock_guard<std::mutex> lk(c->mu);
    const RobotRec& rb = c->robots[robot];
    const EnvRec& ev = c->envs[env];
    const int64_t set = (int64_t)ev.dev.num_obj * rb.num_links * 6;
    if (ref_stride != 0 && ref_stride != set)
        return fail(c, RVS_EINVAL, "rvs_distance_underestimation_async: ref_stride");
    (void)hipSetDevice(c->opt.device);
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
    hipError_t e = async_begin(c, st);
    if (e == hipSuccess && n > 0)
        e = launch_underestimation(rb.dev, ev.dev, params_of(c), q, n, nearest_ref, ref_stride, d, profile, stored, st);
    if (e == hipSuccess) e = async_end(c, st, done);
    if (e != hipSuccess) return hip_fail(c, e, "rvs_distance_underestimation_async");
    return RVS_OK;
}

rvs_status rvs_self_collision_async(rvs_context* c, int32_t robot, const float* q, int64_t n, float* min_dist,
                                    uint8_t* free_flag, uint32_t flags, void* stream, void* done) {
    rvs_status s = check_query(c, robot, -1, q, n, false);
    if (s != RVS_OK) return s;
    if (n > 0 && !min_dist && !free_flag) return fail(c, RVS_EINVAL, "rvs_self_collision_async: no output");
    if (!(flags & RVS_F_DEVICE)) return fail(c, RVS_EINVAL, "rvs_self_collision_async: device arrays only (RVS_F_DEVICE)");
    std::lock_guard<std::mutex> lk(c->mu);
    const RobotRec& rb = c->robots[robot];
    (void)hipSetDevice(c->opt.device);
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
    hipError_t e = async_begin(c, st);
    if (e == hipSuccess && n > 0)
        e = launch_self_distance(rb.dev, params_of(c), q, n, nullptr, nullptr, min_dist, free_flag, st);
    if (e == hipSuccess) e = async_end(c, st, done);
    if (e != hipSuccess) return hip_fail(c, e, "rvs_self_collision_async");
    return RVS_OK;
}

rvs_status rvs_autotune(rvs_context* c, int32_t robot, int32_t env, const float* q, int64_t n, uint32_t flags,
                        rvs_tuning* best) {
    rvs_status s = check_query(c, robot, env, q, n, true);
    if (s != RVS_OK) return s;
    if (n <= 0) return fail(c, RVS_EINVAL, "rvs_autotune: empty sample");
    CtxLock lk(c);
    const RobotRec& rb = c->robots[robot];
    const EnvRec& ev = c->envs[env];
    s = env_fits(c, ev.dev, false, "rvs_autotune");
    if (s != RVS_OK) return s;
    (void)hipSetDevice(c->opt.device);
    // the sample: the first m configurations.  It must hold several
    // configurations per resident lane, or the lane-refill variant has
    // nothing to refill and the choice would not carry over to the batch.
    const int64_t m = std::min<int64_t>(n, 1 << 21);
    const int L = rb.num_links, D = rb.num_dofs;
    hipError_t e = c->arena.ensure(Carve::need({(size_t)m * D * 4, (size_t)m * 4, (size_t)m * L * 4, (size_t)m,
                                                (size_t)m * 4, 256, coh_bytes(m, D), fast_links_bytes(m)}));
    if (e != hipSuccess) return hip_fail(c, e, "rvs_autotune: staging");
    Carve cv(c->arena.p);
    float* qs = cv.take<float>((size_t)m * D);
    DistOut out{};
    out.d = cv.take<float>(m);
    out.profile = cv.take<float>((size_t)m * L);
    out.verdict = cv.take<uint8_t>(m);
    int32_t* idx = cv.take<int32_t>(m);
    uint32_t* count = cv.take<uint32_t>(64);
    uint32_t* coh = cv.take<uint32_t>(coh_bytes(m, D) / 4);
    uint32_t* links = cv.take<uint32_t>(fast_links_bytes(m) / 4);
    e = hipMemcpyAsync(qs, q, (size_t)m * D * 4, (flags & RVS_F_DEVICE) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                       c->stream);
    const QueryParams qp = params_of(c);
    // candidates: one lane per configuration at three workgroup sizes, as one
    // workgroup per 64..256 configurations and in persistent waves; lane
    // refill at three workgroup sizes, with the occupancy limit and half of it
    std::vector<FastTuning> cand;
    for (int cb : {0, 5}) {   // walk order: as given, or sorted by the first joint angles
        for (int b : {64, 128, 256}) cand.push_back(FastTuning{0, b, 0, cb});
        for (int b : {64, 128, 256}) cand.push_back(FastTuning{2, b, 0, cb});
        for (int b : {64, 128, 256}) {
            const int occ = std::max(1, fast_pool_blocks_per_cu(ev.dev, b));
            cand.push_back(FastTuning{1, b, occ, cb});
            if (occ > 1) cand.push_back(FastTuning{1, b, (occ + 1) / 2, cb});
        }
        for (int b : {256, 512, 1024}) {   // per-lane walks, when the scene's tables fit LDS
            const int occ = fast_lanes_blocks_per_cu(ev.dev, b);
            if (occ <= 0) continue;
            cand.push_back(FastTuning{3, b, occ, cb});
            if (occ > 1) cand.push_back(FastTuning{3, b, (occ + 1) / 2, cb});
        }
        for (int b : {256, 512, 1024}) {   // link-major passes, under the same condition
            const int occ = fast_links_blocks_per_cu(ev.dev, b);
            if (occ > 0) cand.push_back(FastTuning{4, b, occ, cb});
        }
    }
    FastTuning win{};
    float win_ms = 1e30f;
    auto try_shape = [&](const FastTuning& t) -> hipError_t {
        float best_ms = 1e30f;
        for (int rep = 0; rep < 3 && e == hipSuccess; ++rep) {   // first run warms up
            e = launch_zero_u32(count, 64, c->stream);
            const int32_t* perm = nullptr;
            if (e == hipSuccess) e = hipEventRecord(c->ev[0], c->stream);
            if (e == hipSuccess) e = launch_coherence_order(rb.dev, qs, m, t, coh, &perm, c->stream);
            if (e == hipSuccess)
                e = launch_distance_fast(rb.dev, ev.dev, qp, qs, m, out, idx, count, t, perm, links, c->num_cus,
                                         c->stream);
            if (e == hipSuccess) e = hipEventRecord(c->ev[1], c->stream);
            if (e == hipSuccess) e = hipEventSynchronize(c->ev[1]);
            float ms = 0;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
            if (rep > 0) best_ms = std::min(best_ms, ms);
        }
        if (e == hipSuccess && best_ms < win_ms) {
            win_ms = best_ms;
            win = t;
        }
        return e;
    };
    for (const FastTuning& t : cand)
        if (try_shape(t) != hipSuccess) return hip_fail(c, e, "rvs_autotune");
    // then the winning shape with other walk-order key widths
    if (win.coh_bits > 0) {
        const FastTuning base = win;
        for (int cb : {3, 4, 6})
            if (try_shape(FastTuning{base.variant, base.block, base.blocks_per_cu, cb}) != hipSuccess)
                return hip_fail(c, e, "rvs_autotune");
    }
    {
        std::lock_guard<std::mutex> tk(c->tuning_mu);
        c->tuning[{robot, env}] = win;
    }
    if (best) *best = rvs_tuning{win.variant, win.block, win.blocks_per_cu, win.coh_bits, win_ms};
    return RVS_OK;
}

rvs_status rvs_set_tuning(rvs_context* c, int32_t robot, int32_t env, const rvs_tuning* t) {
    if (!c || !t) return fail(c, RVS_EINVAL, "rvs_set_tuning: null argument");
    const bool block_ok = t->variant >= 3 ? (t->block == 256 || t->block == 512 || t->block == 1024)
                                          : (t->block == 64 || t->block == 128 || t->block == 256);
    if (t->variant < 0 || t->variant > 4 || !block_ok || t->blocks_per_cu < 0 || t->coherence_bits < 0 ||
        t->coherence_bits > kCohKeyBits)
        return fail(c, RVS_EINVAL,
                    "rvs_set_tuning: variant 0..4, block 64/128/256 (variants 3 and 4: 256/512/1024), "
                    "blocks_per_cu >= 0, coherence_bits 0..18");
    std::lock_guard<std::mutex> lk(c->tuning_mu);
    c->tuning[{robot, env}] = FastTuning{t->variant, t->block, t->blocks_per_cu, t->coherence_bits};
    return RVS_OK;
}

rvs_status rvs_set_small_batch(rvs_context* c, int64_t max_configs) {
    if (!c) return RVS_EINVAL;
    if (max_configs < 0) return fail(c, RVS_EINVAL, "rvs_set_small_batch: negative size");
    CtxLock lk(c);
    c->wave_max = max_configs;
    return RVS_OK;
}

rvs_status rvs_get_tuning(rvs_context* c, int32_t robot, int32_t env, rvs_tuning* t) {
    if (!c || !t) return fail(c, RVS_EINVAL, "rvs_get_tuning: null argument");
    int32_t L = 0, M = 0;
    {
        CtxLock lk(c);
        if (robot_ok(c, robot)) L = c->robots[robot].num_links;
        if (env_ok(c, env)) M = c->envs[env].dev.num_obj;
    }
    const FastTuning f = c->tuning_for(robot, env, L, M);
    *t = rvs_tuning{f.variant, f.block, f.blocks_per_cu, f.coh_bits, 0.0f};
    return RVS_OK;
}

rvs_status rvs_distance_f64(rvs_context* c, int32_t robot, int32_t env, const float* q, int64_t n, double* d,
                            double* profile, uint8_t* verdict, uint32_t flags) {
    rvs_status s = check_query(c, robot, env, q, n, true);
    if (s != RVS_OK) return s;
    CtxLock lk(c);
    const RobotRec& rb = c->robots[robot];
    const EnvRec& ev = c->envs[env];
    s = env_fits(c, ev.dev, true, "rvs_distance_f64");
    if (s != RVS_OK) return s;
    if (n == 0) return RVS_OK;
    (void)hipSetDevice(c->opt.device);
    const int L = rb.num_links, D = rb.num_dofs;
    const QueryParams qp = params_of(c);
    DistOut out{};
    const float* qd = q;
    hipError_t e = hipSuccess;
    // large batches run in the walk order (as rvs_distance's exact paths)
    const size_t coh_need = (n >= kExactSortMin && n <= INT32_MAX) ? coh_bytes(n, D) : 0;
    uint32_t* coh = nullptr;
    if (flags & RVS_F_DEVICE) {
        out.d64 = d; out.profile64 = profile; out.verdict = verdict;
        if (coh_need) {
            e = c->arena.ensure(Carve::need({coh_need}));
            if (e != hipSuccess) return hip_fail(c, e, "rvs_distance_f64: staging");
            Carve cv(c->arena.p);
            coh = cv.take<uint32_t>(coh_need / 4);
        }
    } else {
        const size_t need = Carve::need({coh_need, (size_t)n * D * 4, (size_t)n * 8, profile ? (size_t)n * L * 8 : 0,
                                         (size_t)n});
        e = c->arena.ensure(need);
        if (e != hipSuccess) return hip_fail(c, e, "rvs_distance_f64: staging");
        Carve cv(c->arena.p);
        if (coh_need) coh = cv.take<uint32_t>(coh_need / 4);
        float* qs = cv.take<float>((size_t)n * D);
        out.d64 = cv.take<double>(n);
        out.profile64 = profile ? cv.take<double>((size_t)n * L) : nullptr;
        out.verdict = cv.take<uint8_t>(n);
        e = h2d(qs, q, (size_t)n * D * 4, c->stream);
        qd = qs;
    }
    const int32_t* perm = nullptr;
    if (e == hipSuccess && coh)
        e = launch_coherence_order(rb.dev, qd, n, c->tuning_for(robot, env, L, ev.dev.num_obj), coh, &perm, c->stream);
    if (e == hipSuccess) e = launch_distance_fp64(rb.dev, ev.dev, qp, qd, n, out, c->stream, perm);
    if (e == hipSuccess && !(flags & RVS_F_DEVICE)) {
        e = d2h(d, out.d64, (size_t)n * 8, c->stream);
        if (e == hipSuccess) e = d2h(profile, out.profile64, (size_t)n * L * 8, c->stream);
        if (e == hipSuccess) e = d2h(verdict, out.verdict, (size_t)n, c->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "rvs_distance_f64");
    return RVS_OK;
}

rvs_status rvs_is_valid(rvs_context* c, int32_t robot, int32_t env, const float* q, int64_t n, uint8_t* valid,
                        uint32_t flags) {
    rvs_status s = check_query(c, robot, env, q, n, true);
    if (s != RVS_OK) return s;
    if (n > 0 && !valid) return fail(c, RVS_EINVAL, "rvs_is_valid: null output");
    CtxLock lk(c);
    const RobotRec& rb = c->robots[robot];
    const EnvRec& ev = c->envs[env];
    const int fp64 = prec_of(c, flags) == RVS_PREC_FP64;
    s = env_fits(c, ev.dev, fp64, "rvs_is_valid");
    if (s != RVS_OK) return s;
    if (n == 0) return RVS_OK;
    (void)hipSetDevice(c->opt.device);
    const int D = rb.num_dofs;
    const QueryParams qp = params_of(c);
    hipError_t e = hipSuccess;
    const float* qd = q;
    uint8_t* vd = valid;
    // large batches run in the walk order (as rvs_distance's exact paths)
    const size_t coh_need = (n >= kExactSortMin && n <= INT32_MAX) ? coh_bytes(n, D) : 0;
    uint32_t* coh = nullptr;
    if (!(flags & RVS_F_DEVICE)) {
        e = c->arena.ensure(Carve::need({coh_need, (size_t)n * D * 4, (size_t)n}));
        if (e != hipSuccess) return hip_fail(c, e, "rvs_is_valid: staging");
        Carve cv(c->arena.p);
        if (coh_need) coh = cv.take<uint32_t>(coh_need / 4);
        float* qs = cv.take<float>((size_t)n * D);
        vd = cv.take<uint8_t>(n);
        e = h2d(qs, q, (size_t)n * D * 4, c->stream);
        qd = qs;
    } else if (coh_need) {
        e = c->arena.ensure(Carve::need({coh_need}));
        if (e != hipSuccess) return hip_fail(c, e, "rvs_is_valid: staging");
        Carve cv(c->arena.p);
        coh = cv.take<uint32_t>(coh_need / 4);
    }
    const int32_t* perm = nullptr;
    const bool timing = (flags & RVS_F_TIMING) != 0;
    if (e == hipSuccess && timing) e = hipEventRecord(c->ev[0], c->stream);
    if (e == hipSuccess && coh)
        e = launch_coherence_order(rb.dev, qd, n, c->tuning_for(robot, env, rb.num_links, ev.dev.num_obj), coh, &perm,
                                   c->stream);
    if (e == hipSuccess && timing) e = hipEventRecord(c->ev[1], c->stream);
    if (e == hipSuccess) e = launch_is_valid(rb.dev, ev.dev, qp, qd, n, vd, fp64, c->stream, perm);
    if (e == hipSuccess && timing) e = hipEventRecord(c->ev[2], c->stream);
    if (e == hipSuccess && !(flags & RVS_F_DEVICE)) e = d2h(valid, vd, (size_t)n, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "rvs_is_valid");
    if (timing) {
        // stage 0: the collision tree; stage 1: the walk-order sort
        c->nstages = 2;
        c->stage_ms[0] = c->stage_ms[1] = 0;
        (void)hipEventElapsedTime(&c->stage_ms[0], c->ev[1], c->ev[2]);
        (void)hipEventElapsedTime(&c->stage_ms[1], c->ev[0], c->ev[1]);
    }
    return RVS_OK;
}

rvs_status rvs_distance_underestimation(rvs_context* c, int32_t robot, int32_t env, const float* q, int64_t n,
                                        const float* nearest_ref, int64_t ref_stride, float* d, float* profile,
                                        uint8_t* stored, uint32_t flags) {
    rvs_status s = check_query(c, robot, env, q, n, true);
    if (s != RVS_OK) return s;
    if (n > 0 && (!nearest_ref || !d)) return fail(c, RVS_EINVAL, "rvs_distance_underestimation: null argument");
    CtxLock lk(c);
    const RobotRec& rb = c->robots[robot];
    const EnvRec& ev = c->envs[env];
    const int L = rb.num_links, D = rb.num_dofs, M = ev.dev.num_obj;
    const int64_t set = (int64_t)M * L * 6;
    if (ref_stride != 0 && ref_stride != set) return fail(c, RVS_EINVAL, "rvs_distance_underestimation: ref_stride");
    if (n == 0) return RVS_OK;
    (void)hipSetDevice(c->opt.device);
    const QueryParams qp = params_of(c);
    hipError_t e = hipSuccess;
    const float* qd = q;
    const float* npd = nearest_ref;
    float *dd = d, *pd = profile;
    uint8_t* sd = stored;
    if (!(flags & RVS_F_DEVICE)) {
        const size_t nref = ref_stride ? (size_t)n * set : (size_t)set;
        e = c->arena.ensure(Carve::need({(size_t)n * D * 4, nref * 4, (size_t)n * 4, profile ? (size_t)n * L * 4 : 0,
                                         (size_t)n}));
        if (e != hipSuccess) return hip_fail(c, e, "rvs_distance_underestimation: staging");
        Carve cv(c->arena.p);
        float* qs = cv.take<float>((size_t)n * D);
        float* ns = cv.take<float>(nref);
        dd = cv.take<float>(n);
        pd = profile ? cv.take<float>((size_t)n * L) : nullptr;
        sd = cv.take<uint8_t>(n);
        e = h2d(qs, q, (size_t)n * D * 4, c->stream);
        if (e == hipSuccess) e = h2d(ns, nearest_ref, nref * 4, c->stream);
        qd = qs;
        npd = ns;
    }
    const bool timing = (flags & RVS_F_TIMING) != 0;
    if (e == hipSuccess && timing) e = hipEventRecord(c->ev[0], c->stream);
    if (e == hipSuccess) e = launch_underestimation(rb.dev, ev.dev, qp, qd, n, npd, ref_stride, dd, pd, sd, c->stream);
    if (e == hipSuccess && timing) e = hipEventRecord(c->ev[1], c->stream);
    if (e == hipSuccess && !(flags & RVS_F_DEVICE)) {
        e = d2h(d, dd, (size_t)n * 4, c->stream);
        if (e == hipSuccess) e = d2h(profile, pd, (size_t)n * L * 4, c->stream);
        if (e == hipSuccess) e = d2h(stored, sd, (size_t)n, c->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "rvs_distance_underestimation");
    if (timing) {
        c->nstages = 1;
        c->stage_ms[0] = 0;
        (void)hipEventElapsedTime(&c->stage_ms[0], c->ev[0], c->ev[1]);
    }
    return RVS_OK;
}

rvs_status rvs_distance_underestimation_moving(rvs_context* c, int32_t robot, int32_t env, const float* q, int64_t n,
                                               const float* nearest_ref, int64_t ref_stride, const float* max_vel,
                                               float delta_t, float* d, float* profile, uint8_t* stored,
                                               uint32_t flags) {
    rvs_status s = check_query(c, robot, env, q, n, true);
    if (s != RVS_OK) return s;
    if (n > 0 && (!nearest_ref || !d || !max_vel))
        return fail(c, RVS_EINVAL, "rvs_distance_underestimation_moving: null argument");
    CtxLock lk(c);
    const RobotRec& rb = c->robots[robot];
    const EnvRec& ev = c->envs[env];
    const int L = rb.num_links, D = rb.num_dofs, M = ev.dev.num_obj;
    const int64_t set = (int64_t)M * L * 6;
    if (ref_stride != 0 && ref_stride != set)
        return fail(c, RVS_EINVAL, "rvs_distance_underestimation_moving: ref_stride");
    if (n == 0) return RVS_OK;
    (void)hipSetDevice(c->opt.device);
    const QueryParams qp = params_of(c);
    hipError_t e = hipSuccess;
    const float* qd = q;
    const float* npd = nearest_ref;
    const float* vd = max_vel;
    float *dd = d, *pd = profile;
    uint8_t* sd = stored;
    if (!(flags & RVS_F_DEVICE)) {
        const size_t nref = ref_stride ? (size_t)n * set : (size_t)set;
        e = c->arena.ensure(Carve::need({(size_t)n * D * 4, nref * 4, (size_t)M * 4, (size_t)n * 4,
                                         profile ? (size_t)n * L * 4 : 0, (size_t)n}));
        if (e != hipSuccess) return hip_fail(c, e, "rvs_distance_underestimation_moving: staging");
        Carve cv(c->arena.p);
        float* qs = cv.take<float>((size_t)n * D);
        float* ns = cv.take<float>(nref);
        float* vs = cv.take<float>(M);
        dd = cv.take<float>(n);
        pd = profile ? cv.take<float>((size_t)n * L) : nullptr;
        sd = cv.take<uint8_t>(n);
        e = h2d(qs, q, (size_t)n * D * 4, c->stream);
        if (e == hipSuccess) e = h2d(ns, nearest_ref, nref * 4, c->stream);
        if (e == hipSuccess) e = h2d(vs, max_vel, (size_t)M * 4, c->stream);
        qd = qs;
        npd = ns;
        vd = vs;
    }
    if (e == hipSuccess)
        e = launch_underestimation_moving(rb.dev, ev.dev, qp, qd, n, npd, ref_stride, vd, delta_t, dd, pd, sd,
                                          c->stream);
    if (e == hipSuccess && !(flags & RVS_F_DEVICE)) {
        e = d2h(d, dd, (size_t)n * 4, c->stream);
        if (e == hipSuccess) e = d2h(profile, pd, (size_t)n * L * 4, c->stream);
        if (e == hipSuccess) e = d2h(stored, sd, (size_t)n, c->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "rvs_distance_underestimation_moving");
    return RVS_OK;
}

rvs_status rvs_trajectory_is_safe(rvs_context* c, int32_t robot, int32_t env, const float* points, int64_t num_traj,
                                  int32_t num_points, const float* q_current, const float* profile_current,
                                  const float* nearest_current, const float* max_vel, float t_iter, float time_step,
                                  int32_t self_collision, uint8_t* safe, uint32_t flags) {
    rvs_status s = check_query(c, robot, env, points, num_traj, true);
    if (s != RVS_OK) return s;
    if (num_points < 0) return fail(c, RVS_EINVAL, "rvs_trajectory_is_safe: negative point count");
    if (num_traj > 0 && (!q_current || !profile_current || !nearest_current || !max_vel || !safe))
        return fail(c, RVS_EINVAL, "rvs_trajectory_is_safe: null argument");
    CtxLock lk(c);
    const RobotRec& rb = c->robots[robot];
    const EnvRec& ev = c->envs[env];
    if (rb.num_links != rb.num_dofs)
        return fail(c, RVS_ENOTSUP, "rvs_trajectory_is_safe: needs one link per revolute joint");
    if (traj_lds_bytes(rb.dev) > device_lds_limit(c->opt.device))
        return fail(c, RVS_ERANGE, "rvs_trajectory_is_safe: robot too long");
    if (num_traj > INT32_MAX) return fail(c, RVS_ERANGE, "rvs_trajectory_is_safe: too many trajectories");
    if (num_traj == 0) return RVS_OK;
    (void)hipSetDevice(c->opt.device);
    const QueryParams qp = params_of(c);
    const int L = rb.num_links, D = rb.num_dofs, M = ev.dev.num_obj;
    const size_t pb = (size_t)num_traj * num_points * D * 4, set = (size_t)M * L * 6;
    TrajArgs a{};
    a.points = points;
    a.num_points = num_points;
    a.num_obj = M;
    a.q_current = q_current;
    a.profile_current = profile_current;
    a.nearest_current = nearest_current;
    a.max_vel = max_vel;
    a.t_iter = t_iter;
    a.time_step = time_step;
    a.self_collision = self_collision ? 1 : 0;
    a.safe = safe;
    hipError_t e = hipSuccess;
    if (!(flags & RVS_F_DEVICE)) {
        e = c->arena.ensure(Carve::need({pb, (size_t)D * 4, (size_t)L * 4, set * 4, (size_t)M * 4, (size_t)num_traj}));
        if (e != hipSuccess) return hip_fail(c, e, "rvs_trajectory_is_safe: staging");
        Carve cv(c->arena.p);
        float* ps = cv.take<float>(pb / 4);
        float* qs = cv.take<float>(D);
        float* prs = cv.take<float>(L);
        float* ns = cv.take<float>(set);
        float* vs = cv.take<float>(M);
        a.safe = cv.take<uint8_t>(num_traj);
        e = h2d(ps, points, pb, c->stream);
        if (e == hipSuccess) e = h2d(qs, q_current, (size_t)D * 4, c->stream);
        if (e == hipSuccess) e = h2d(prs, profile_current, (size_t)L * 4, c->stream);
        if (e == hipSuccess) e = h2d(ns, nearest_current, set * 4, c->stream);
        if (e == hipSuccess) e = h2d(vs, max_vel, (size_t)M * 4, c->stream);
        a.points = ps;
        a.q_current = qs;
        a.profile_current = prs;
        a.nearest_current = ns;
        a.max_vel = vs;
    }
    if (e == hipSuccess) e = launch_traj_safe(rb.dev, qp, a, num_traj, c->stream);
    if (e == hipSuccess && !(flags & RVS_F_DEVICE)) e = d2h(safe, a.safe, (size_t)num_traj, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "rvs_trajectory_is_safe");
    return RVS_OK;
}

rvs_status rvs_self_distance(rvs_context* c, int32_t robot, const float* q, int64_t n, float* dist, uint32_t flags) {
    rvs_status s = check_query(c, robot, -1, q, n, false);
    if (s != RVS_OK) return s;
    CtxLock lk(c);
    const RobotRec& rb = c->robots[robot];
    if (n == 0 || rb.num_pairs == 0) return RVS_OK;
    if (!dist) return fail(c, RVS_EINVAL, "rvs_self_distance: null output");
    (void)hipSetDevice(c->opt.device);
    const int D = rb.num_dofs, P = rb.num_pairs;
    const QueryParams qp = params_of(c);
    hipError_t e = hipSuccess;
    const float* qd = q;
    float* od = dist;
    if (!(flags & RVS_F_DEVICE)) {
        e = c->arena.ensure(Carve::need({(size_t)n * D * 4, (size_t)n * P * 4}));
        if (e != hipSuccess) return hip_fail(c, e, "rvs_self_distance: staging");
        Carve cv(c->arena.p);
        float* qs = cv.take<float>((size_t)n * D);
        od = cv.take<float>((size_t)n * P);
        e = h2d(qs, q, (size_t)n * D * 4, c->stream);
        qd = qs;
    }
    if (e == hipSuccess) e = launch_self_distance(rb.dev, qp, qd, n, od, nullptr, nullptr, nullptr, c->stream);
    if (e == hipSuccess && !(flags & RVS_F_DEVICE)) e = d2h(dist, od, (size_t)n * P * 4, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "rvs_self_distance");
    return RVS_OK;
}

rvs_status rvs_self_collision(rvs_context* c, int32_t robot, const float* q, int64_t n, float* min_dist,
                              uint8_t* free_flag, uint32_t flags) {
    rvs_status s = check_query(c, robot, -1, q, n, false);
    if (s != RVS_OK) return s;
    if (n > 0 && !min_dist && !free_flag) return fail(c, RVS_EINVAL, "rvs_self_collision: no output");
    CtxLock lk(c);
    const RobotRec& rb = c->robots[robot];
    if (n == 0) return RVS_OK;
    (void)hipSetDevice(c->opt.device);
    const int D = rb.num_dofs;
    const QueryParams qp = params_of(c);
    hipError_t e = hipSuccess;
    const float* qd = q;
    float* md = min_dist;
    uint8_t* fd = free_flag;
    if (!(flags & RVS_F_DEVICE)) {
        e = c->arena.ensure(Carve::need({(size_t)n * D * 4, min_dist ? (size_t)n * 4 : 0, free_flag ? (size_t)n : 0}));
        if (e != hipSuccess) return hip_fail(c, e, "rvs_self_collision: staging");
        Carve cv(c->arena.p);
        float* qs = cv.take<float>((size_t)n * D);
        md = min_dist ? cv.take<float>(n) : nullptr;
        fd = free_flag ? cv.take<uint8_t>(n) : nullptr;
        e = h2d(qs, q, (size_t)n * D * 4, c->stream);
        qd = qs;
    }
    if (e == hipSuccess) e = launch_self_distance(rb.dev, qp, qd, n, nullptr, nullptr, md, fd, c->stream);
    if (e == hipSuccess && !(flags & RVS_F_DEVICE)) {
        e = d2h(min_dist, md, min_dist ? (size_t)n * 4 : 0, c->stream);
        if (e == hipSuccess) e = d2h(free_flag, fd, free_flag ? (size_t)n : 0, c->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "rvs_self_collision");
    return RVS_OK;
}

rvs_status rvs_enclosing_radii(rvs_context* c, int32_t robot, const float* q, int64_t n, float* R, uint32_t flags) {
    rvs_status s = check_query(c, robot, -1, q, n, false);
    if (s != RVS_OK) return s;
    if (n > 0 && !R) return fail(c, RVS_EINVAL, "rvs_enclosing_radii: null output");
    CtxLock lk(c);
    const RobotRec& rb = c->robots[robot];
    if (rb.num_links != rb.num_dofs)
        return fail(c, RVS_ENOTSUP, "rvs_enclosing_radii: needs one link per revolute joint");
    if (n == 0) return RVS_OK;
    (void)hipSetDevice(c->opt.device);
    const int D = rb.num_dofs;
    const size_t per = (size_t)D * (D + 1);
    hipError_t e = hipSuccess;
    const float* qd = q;
    float* rd = R;
    if (!(flags & RVS_F_DEVICE)) {
        e = c->arena.ensure(Carve::need({(size_t)n * D * 4, (size_t)n * per * 4}));
        if (e != hipSuccess) return hip_fail(c, e, "rvs_enclosing_radii: staging");
        Carve cv(c->arena.p);
        float* qs = cv.take<float>((size_t)n * D);
        rd = cv.take<float>((size_t)n * per);
        e = h2d(qs, q, (size_t)n * D * 4, c->stream);
        qd = qs;
    }
    if (e == hipSuccess) e = launch_enclosing_radii(rb.dev, qd, n, rd, c->stream);
    if (e == hipSuccess && !(flags & RVS_F_DEVICE)) e = d2h(R, rd, (size_t)n * per * 4, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "rvs_enclosing_radii");
    return RVS_OK;
}

rvs_status rvs_extend_spines(rvs_context* c, int32_t robot, const float* q, const float* d, const float* profile,
                             const float* q_e, int64_t n, int32_t num_iter, int32_t use_expanded_bubble,
                             float* q_new, int32_t* status, uint32_t flags) {
    rvs_status s = check_query(c, robot, -1, q, n, false);
    if (s != RVS_OK) return s;
    if (n > 0 && (!d || !q_e || !q_new || !status || (use_expanded_bubble && !profile)))
        return fail(c, RVS_EINVAL, "rvs_extend_spines: null argument");
    if (num_iter < 1) return fail(c, RVS_EINVAL, "rvs_extend_spines: num_iter must be at least 1");
    CtxLock lk(c);
    const RobotRec& rb = c->robots[robot];
    if (rb.num_links != rb.num_dofs)
        return fail(c, RVS_ENOTSUP, "rvs_extend_spines: needs one link per revolute joint");
    if (n == 0) return RVS_OK;
    (void)hipSetDevice(c->opt.device);
    const int D = rb.num_dofs, L = rb.num_links;
    const bool np = use_expanded_bubble != 0;
    hipError_t e = hipSuccess;
    SpineArgs a{};
    a.q = q;
    a.q_stride = D;
    a.d = d;
    a.d_stride = 1;
    a.profile = np ? profile : d;
    a.profile_stride = np ? L : 0;
    a.q_e = q_e;
    a.q_new = q_new;
    a.status = status;
    a.num_iter = num_iter;
    a.expanded = np ? 1 : 0;
    if (!(flags & RVS_F_DEVICE)) {
        const size_t qb = (size_t)n * D * 4;
        e = c->arena.ensure(Carve::need({qb, (size_t)n * 4, np ? (size_t)n * L * 4 : 0, qb, qb, (size_t)n * 4}));
        if (e != hipSuccess) return hip_fail(c, e, "rvs_extend_spines: staging");
        Carve cv(c->arena.p);
        float* qs = cv.take<float>((size_t)n * D);
        float* ds = cv.take<float>(n);
        float* ps = np ? cv.take<float>((size_t)n * L) : nullptr;
        float* es = cv.take<float>((size_t)n * D);
        a.q_new = cv.take<float>((size_t)n * D);
        a.status = cv.take<int32_t>(n);
        e = h2d(qs, q, qb, c->stream);
        if (e == hipSuccess) e = h2d(ds, d, (size_t)n * 4, c->stream);
        if (e == hipSuccess && np) e = h2d(ps, profile, (size_t)n * L * 4, c->stream);
        if (e == hipSuccess) e = h2d(es, q_e, qb, c->stream);
        a.q = qs;
        a.d = ds;
        a.profile = np ? ps : ds;
        a.q_e = es;
    }
    if (e == hipSuccess) e = launch_spines(rb.dev, a, n, c->stream);
    if (e == hipSuccess && !(flags & RVS_F_DEVICE)) {
        e = d2h(q_new, a.q_new, (size_t)n * D * 4, c->stream);
        if (e == hipSuccess) e = d2h(status, a.status, (size_t)n * 4, c->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "rvs_extend_spines");
    return RVS_OK;
}

rvs_status rvs_extend_gen_spines(rvs_context* c, int32_t robot, int32_t env, const float* q_root, const float* q_e,
                                 int64_t n, int32_t num_layers, int32_t num_iter, int32_t use_expanded_bubble,
                                 int32_t use_real_dist, float d_crit, float* q_new, int32_t* status, float* d,
                                 uint32_t flags) {
    rvs_status s = check_query(c, robot, env, q_e, n, true);
    if (s != RVS_OK) return s;
    if (n > 0 && (!q_root || !q_new || !status || !d)) return fail(c, RVS_EINVAL, "rvs_extend_gen_spines: null argument");
    if (num_layers < 1 || num_iter < 1)
        return fail(c, RVS_EINVAL, "rvs_extend_gen_spines: num_layers and num_iter must be at least 1");
    CtxLock lk(c);
    const RobotRec& rb = c->robots[robot];
    const EnvRec& ev = c->envs[env];
    if (rb.num_links != rb.num_dofs)
        return fail(c, RVS_ENOTSUP, "rvs_extend_gen_spines: needs one link per revolute joint");
    s = env_fits(c, ev.dev, false, "rvs_extend_gen_spines");
    if (s != RVS_OK) return s;
    if (n == 0) return RVS_OK;
    (void)hipSetDevice(c->opt.device);
    const QueryParams qp = params_of(c);
    const int D = rb.num_dofs, L = rb.num_links, M = ev.dev.num_obj;
    const bool host = !(flags & RVS_F_DEVICE);
    const size_t qb = (size_t)n * D * 4, set = (size_t)M * L * 6;
    hipError_t e = c->arena.ensure(Carve::need({4, (size_t)L * 4, set * 4, (size_t)n * 4, (size_t)n * L * 4,
                                                (size_t)n, host ? (size_t)D * 4 : 0, host ? qb : 0, host ? qb : 0,
                                                host ? (size_t)n * 4 : 0, host ? (size_t)n * 4 : 0}));
    if (e != hipSuccess) return hip_fail(c, e, "rvs_extend_gen_spines: staging");
    Carve cv(c->arena.p);
    float* d_root = cv.take<float>(1);
    float* p_root = cv.take<float>(L);
    float* np_root = cv.take<float>(set);
    float* d_layer = cv.take<float>(n);
    float* p_layer = cv.take<float>((size_t)n * L);
    uint8_t* active = cv.take<uint8_t>(n);
    const float* qr = q_root;
    const float* qe = q_e;
    float* qn = q_new;
    int32_t* st = status;
    float* dd = d;
    if (host) {
        float* qrs = cv.take<float>(D);
        float* qes = cv.take<float>((size_t)n * D);
        qn = cv.take<float>((size_t)n * D);
        st = cv.take<int32_t>(n);
        dd = cv.take<float>(n);
        e = h2d(qrs, q_root, (size_t)D * 4, c->stream);
        if (e == hipSuccess) e = h2d(qes, q_e, qb, c->stream);
        qr = qrs;
        qe = qes;
    }
    // the root's distance, profile and nearest points (computeDistance)
    DistOut ro{};
    ro.d = d_root;
    ro.profile = p_root;
    ro.nearest = np_root;
    if (e == hipSuccess) e = launch_distance_ref32_auto(rb.dev, ev.dev, qp, qr, 1, ro, c->wave_max, c->stream);
    if (!use_real_dist) {
        // every layer of a spine in one lane
        GenSpineArgs g{};
        g.q_root = qr;
        g.d_root = d_root;
        g.profile_root = p_root;
        g.np_root = np_root;
        g.q_e = qe;
        g.q_new = qn;
        g.status = st;
        g.d_out = dd;
        g.num_obj = M;
        g.num_layers = num_layers;
        g.num_iter = num_iter;
        g.expanded = use_expanded_bubble ? 1 : 0;
        g.d_crit = d_crit;
        g.maxd = qp.maxd;
        if (e == hipSuccess) e = launch_gen_spines_under(rb.dev, g, n, c->stream);
    } else if (e == hipSuccess) {
        e = launch_fill_u8(active, n, 1, c->stream);
    }
    DistOut lo{};
    lo.d = d_layer;
    lo.profile = p_layer;
    for (int layer = 0; use_real_dist && layer < num_layers && e == hipSuccess; ++layer) {
        SpineArgs a{};
        a.q = layer == 0 ? qr : qn;
        a.q_stride = layer == 0 ? 0 : D;
        a.d = layer == 0 ? d_root : d_layer;
        a.d_stride = layer == 0 ? 0 : 1;
        a.profile = layer == 0 ? p_root : p_layer;
        a.profile_stride = layer == 0 ? 0 : L;
        a.q_e = qe;
        a.active = active;
        a.q_new = qn;
        a.status = st;
        a.num_iter = num_iter;
        a.expanded = use_expanded_bubble ? 1 : 0;
        e = launch_spines(rb.dev, a, n, c->stream);
        if (e != hipSuccess) break;
        e = launch_distance_ref32_auto(rb.dev, ev.dev, qp, qn, n, lo, c->wave_max, c->stream);
        if (e == hipSuccess) e = launch_gen_layer(n, layer, d_root, d_layer, st, active, dd, d_crit, c->stream);
    }
    if (e == hipSuccess && host) {
        e = d2h(q_new, qn, qb, c->stream);
        if (e == hipSuccess) e = d2h(status, st, (size_t)n * 4, c->stream);
        if (e == hipSuccess) e = d2h(d, dd, (size_t)n * 4, c->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "rvs_extend_gen_spines");
    return RVS_OK;
}

rvs_status rvs_edge_self_collision(rvs_context* c, int32_t robot, const float* q1, const float* q2, int64_t n,
                                   float eps_step, uint8_t* collide, float* q2_out, uint32_t flags) {
    rvs_status s = check_query(c, robot, -1, q1, n, false);
    if (s != RVS_OK) return s;
    if (n > 0 && (!q2 || !collide || !q2_out)) return fail(c, RVS_EINVAL, "rvs_edge_self_collision: null argument");
    if (!(eps_step > 0.0f)) return fail(c, RVS_EINVAL, "rvs_edge_self_collision: eps_step must be positive");
    CtxLock lk(c);
    const RobotRec& rb = c->robots[robot];
    if (rb.dev.kind != RVS_SKEL_XARM6 || rb.num_links != 6 || rb.num_dofs != 6)
        return fail(c, RVS_ENOTSUP, "rvs_edge_self_collision: the xArm6 pair rule only");
    if (n == 0) return RVS_OK;
    (void)hipSetDevice(c->opt.device);
    const QueryParams qp = params_of(c);
    const int D = rb.num_dofs;
    const size_t qb = (size_t)n * D * 4;
    hipError_t e = hipSuccess;
    const float *a = q1, *b = q2;
    uint8_t* cd = collide;
    float* od = q2_out;
    if (!(flags & RVS_F_DEVICE)) {
        e = c->arena.ensure(Carve::need({qb, qb, (size_t)n, qb}));
        if (e != hipSuccess) return hip_fail(c, e, "rvs_edge_self_collision: staging");
        Carve cv(c->arena.p);
        float* as = cv.take<float>((size_t)n * D);
        float* bs = cv.take<float>((size_t)n * D);
        cd = cv.take<uint8_t>(n);
        od = cv.take<float>((size_t)n * D);
        e = h2d(as, q1, qb, c->stream);
        if (e == hipSuccess) e = h2d(bs, q2, qb, c->stream);
        a = as;
        b = bs;
    }
    if (e == hipSuccess) e = launch_edge_self(rb.dev, qp, a, b, n, eps_step, cd, od, c->stream);
    if (e == hipSuccess && !(flags & RVS_F_DEVICE)) {
        e = d2h(collide, cd, (size_t)n, c->stream);
        if (e == hipSuccess) e = d2h(q2_out, od, qb, c->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "rvs_edge_self_collision");
    return RVS_OK;
}

namespace {
void link_keys(const RobotRec& rb, std::vector<float>& keys, std::vector<int32_t>& link_key);
}

rvs_status rvs_edge_is_valid(rvs_context* c, int32_t robot, int32_t env, const float* q1, const float* q2,
                             int64_t n, uint8_t* valid, uint32_t flags) {
    rvs_status s = check_query(c, robot, env, q1, n, true);
    if (s != RVS_OK) return s;
    if (n > 0 && (!q2 || !valid)) return fail(c, RVS_EINVAL, "rvs_edge_is_valid: null argument");
    CtxLock lk(c);
    const RobotRec& rb = c->robots[robot];
    const EnvRec& ev = c->envs[env];
    if (prec_of(c, flags) == RVS_PREC_FP64) return fail(c, RVS_ENOTSUP, "rvs_edge_is_valid: reference arithmetic only");
    if (n == 0) return RVS_OK;
    if (c->opt.num_checks <= 0) return fail(c, RVS_EINVAL, "rvs_edge_is_valid: num_checks must be positive");
    const int D = rb.num_dofs;
    (void)hipSetDevice(c->opt.device);
    const QueryParams qp = params_of(c);
    // axis-test tables of the robot's links (axis-aligned scenes)
    std::vector<float> keys;
    std::vector<int32_t> lkey;
    if (!ev.dev.use_obb) link_keys(rb, keys, lkey);
    const int nkeys = (int)keys.size() / 2;
    const size_t rows = (flags & RVS_F_DEVICE) ? 0 : (size_t)n * D;
    hipError_t e = c->arena.ensure(Carve::need({lkey.size() * 4, keys.size() * 4,
                                                (size_t)nkeys * slab_entries(ev.dev.num_obj) * sizeof(SlabEntry),
                                                rows * 4, rows * 4, (flags & RVS_F_DEVICE) ? 0 : (size_t)n}));
    if (e != hipSuccess) return hip_fail(c, e, "rvs_edge_is_valid: staging");
    Carve cv(c->arena.p);
    int32_t* lkd = cv.take<int32_t>(lkey.size());
    float* kd = cv.take<float>(keys.size());
    SlabEntry* slab = cv.take<SlabEntry>((size_t)nkeys * slab_entries(ev.dev.num_obj));
    e = h2d(lkd, lkey.data(), lkey.size() * 4, c->stream);
    if (e == hipSuccess) e = h2d(kd, keys.data(), keys.size() * 4, c->stream);
    const float *a = q1, *b = q2;
    uint8_t* vd = valid;
    if (e == hipSuccess && !(flags & RVS_F_DEVICE)) {
        float* as = cv.take<float>(rows);
        float* bs = cv.take<float>(rows);
        vd = cv.take<uint8_t>(n);
        e = h2d(as, q1, rows * 4, c->stream);
        if (e == hipSuccess) e = h2d(bs, q2, rows * 4, c->stream);
        a = as;
        b = bs;
    }
    const bool timing = (flags & RVS_F_TIMING) != 0;
    if (e == hipSuccess && timing) e = hipEventRecord(c->ev[0], c->stream);
    if (e == hipSuccess) e = launch_edge_valid(rb.dev, ev.dev, qp, a, b, n, vd, lkd, kd, nkeys, slab, c->stream);
    if (e == hipSuccess && timing) e = hipEventRecord(c->ev[1], c->stream);
    if (e == hipSuccess && !(flags & RVS_F_DEVICE)) e = d2h(valid, vd, (size_t)n, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "rvs_edge_is_valid");
    if (timing) {
        c->nstages = 1;
        c->stage_ms[0] = 0;
        (void)hipEventElapsedTime(&c->stage_ms[0], c->ev[0], c->ev[1]);
    }
    return RVS_OK;
}

namespace {

// Launch plan of a first-collision call: robot groups, the edge order inside
// each group (longest first) and the round offsets (slots before each state
// index).  Caller holds c->mu.
struct EdgePlan {
    EdgeGroups gr{};
    std::vector<int32_t> order, rounds;
    std::vector<float> keys;          // axis-test tables: (radius, ground skipped) per table
    std::vector<int32_t> link_key;    // table of each (group, link)
    int2* items_host = nullptr;       // item table of the current run, in pinned staging
    int2* items_dev = nullptr;        // its device copy (room for max_items entries each)
};

// Axis-test table of each link of a robot (appended to link_key): links with
// the same radius and the same ground rule share one; new (radius, ground
// skipped) keys are appended to keys.
void link_keys(const RobotRec& rb, std::vector<float>& keys, std::vector<int32_t>& link_key) {
    for (int i = 0; i < rb.num_links; ++i) {
        const float r = rb.radius[i], skip = i < rb.dev.ground_included ? 1.0f : 0.0f;
        int t = 0;
        const int nk = (int)keys.size() / 2;
        while (t < nk && !(keys[2 * t] == r && keys[2 * t + 1] == skip)) ++t;
        if (t == nk) {
            keys.push_back(r);
            keys.push_back(skip);
        }
        link_key.push_back(t);
    }
}

rvs_status plan_edges(rvs_context* c, const int32_t* robots, const int32_t* K, int64_t num_edges, int32_t row_stride,
                      EdgePlan& pl) {
    std::vector<int32_t> gid;     // robot ids in order of first appearance
    // group of each loaded robot (-1: none yet), and the group sizes: one
    // pass to validate and count, one to fill the member lists in place
    std::vector<int32_t> group_of(c->robots.size(), -1);
    std::vector<int64_t> size;
    for (int64_t e = 0; e < num_edges; ++e) {
        const int32_t r = robots[e];
        if (!robot_ok(c, r)) return fail(c, RVS_EINVAL, "rvs_edge_first_collision: bad robot id");
        if (K[e] < 0 || K[e] > (1 << 20)) return fail(c, RVS_ERANGE, "rvs_edge_first_collision: state count");
        if (group_of[r] < 0) {
            if ((int)gid.size() == kMaxEdgeGroups)
                return fail(c, RVS_ERANGE, "rvs_edge_first_collision: at most 8 robots per call");
            group_of[r] = (int32_t)gid.size();
            gid.push_back(r);
            size.push_back(0);
        }
        ++size[group_of[r]];
    }
    std::vector<std::vector<int32_t>> members(gid.size());
    for (size_t g = 0; g < gid.size(); ++g) members[g].reserve((size_t)size[g]);
    for (int64_t e = 0; e < num_edges; ++e) members[group_of[robots[e]]].push_back((int32_t)e);
    EdgeGroups& gr = pl.gr;
    gr.num = (int32_t)gid.size();
    for (int g = 0; g < gr.num; ++g) {
        const RobotRec& rb = c->robots[gid[g]];
        if (row_stride < rb.num_dofs) return fail(c, RVS_EINVAL, "rvs_edge_first_collision: row_stride < dofs");
        gr.rb[g] = rb.dev;
        gr.max_dofs = std::max(gr.max_dofs, rb.num_dofs);
        gr.max_links = std::max(gr.max_links, rb.num_links);
        auto& m = members[g];
        // longest first, stable: a counting sort when the lengths are few
        int32_t kmx = 0;
        for (int32_t e : m) kmx = std::max(kmx, K[e]);
        if ((size_t)kmx <= 4 * m.size() + 1024) {
            std::vector<int32_t> cnt((size_t)kmx + 2, 0), out(m.size());
            for (int32_t e : m) ++cnt[(size_t)(kmx - K[e]) + 1];
            for (size_t k = 1; k < cnt.size(); ++k) cnt[k] += cnt[k - 1];
            for (int32_t e : m) out[(size_t)cnt[(size_t)(kmx - K[e])]++] = e;
            m.swap(out);
        } else {
            std::stable_sort(m.begin(), m.end(), [&](int32_t x, int32_t y) { return K[x] > K[y]; });
        }
        gr.order_off[g] = (int32_t)pl.order.size();
        pl.order.insert(pl.order.end(), m.begin(), m.end());
        const int32_t kmax = m.empty() ? 0 : K[m[0]];
        gr.round_off[g] = (int32_t)pl.rounds.size();
        gr.num_rounds[g] = kmax;
        // rounds[k] = slots before round k; edges alive in round k are a prefix
        int64_t acc = 0;
        size_t alive = m.size();
        for (int32_t k = 0; k <= kmax; ++k) {
            pl.rounds.push_back((int32_t)acc);
            while (alive > 0 && K[m[alive - 1]] <= k) --alive;
            acc += (int64_t)alive;
        }
        if (acc > INT32_MAX / 8) return fail(c, RVS_ERANGE, "rvs_edge_first_collision: too many states");
        gr.num_slots[g] = (int32_t)acc;
        gr.key_off[g] = (int32_t)pl.link_key.size();
        link_keys(rb, pl.keys, pl.link_key);
    }
    return RVS_OK;
}

// Items of 64 * spl slots, ordered by the state index of their first slot and
// then by group, so that the early states of every robot's edges come before
// the late states of any (the search skips the states past a found collision
// only if that collision was found before it takes them).  Table entries are
// (group, first slot).
size_t item_table(const EdgePlan& pl, int spl, int2* t) {
    const int32_t per = 64 * spl;
    int32_t cur[kMaxEdgeGroups] = {}, rk[kMaxEdgeGroups] = {}, kmax = 0;
    for (int g = 0; g < pl.gr.num; ++g) kmax = std::max(kmax, pl.gr.num_rounds[g]);
    size_t n = 0;
    // round by round, groups in order, each group's items in slot order: the
    // order of a stable sort of the group-by-group list by round
    for (int32_t k = 0; k <= kmax; ++k) {
        for (int g = 0; g < pl.gr.num; ++g) {
            const int32_t* r = pl.rounds.data() + pl.gr.round_off[g];
            while (cur[g] < pl.gr.num_slots[g]) {
                while (rk[g] + 1 <= pl.gr.num_rounds[g] && r[rk[g] + 1] <= cur[g]) ++rk[g];
                if (rk[g] != k) break;
                t[n++] = make_int2(g, cur[g]);
                cur[g] += per;
            }
        }
    }
    return n;
}

// Slots of a plan in 64-slot items (the largest item table of any spl).
size_t max_items(const EdgePlan& pl) {
    size_t n = 0;
    for (int g = 0; g < pl.gr.num; ++g) n += ((size_t)pl.gr.num_slots[g] + 63) / 64;
    return n;
}

// Arena bytes of a plan's device tables (before the per-call rows).
std::vector<size_t> plan_bytes(const EdgePlan& pl, const EnvRec& ev) {
    const size_t nk = ev.dev.use_obb ? 0 : pl.keys.size() / 2;
    return {pl.order.size() * 4, pl.rounds.size() * 4, 256, pl.link_key.size() * 4, pl.keys.size() * 4,
            nk * slab_entries(ev.dev.num_obj) * sizeof(SlabEntry), max_items(pl) * sizeof(int2)};
}

// Carves and uploads a plan's tables (in plan_bytes order); sets the table
// pointers of pl.gr.
struct PlanDev {
    int32_t* order;
    int32_t* rounds;
    uint32_t* counter;
    float* keys;
    int nkeys;
};
hipError_t upload_plan(rvs_context* c, const EnvRec& ev, EdgePlan& pl, Carve& cv, PlanDev& d) {
    const size_t base = cv.off;
    d.order = cv.take<int32_t>(pl.order.size());
    d.rounds = cv.take<int32_t>(pl.rounds.size());
    d.counter = cv.take<uint32_t>(64);
    int32_t* lk = cv.take<int32_t>(pl.link_key.size());
    d.keys = cv.take<float>(pl.keys.size());
    const size_t end = cv.off;
    d.nkeys = ev.dev.use_obb ? 0 : (int)pl.keys.size() / 2;
    pl.gr.slab = cv.take<SlabEntry>((size_t)d.nkeys * slab_entries(ev.dev.num_obj));
    pl.items_dev = cv.take<int2>(max_items(pl));
    pl.gr.link_key = lk;
    // the tables through a pinned mirror of [base, end) in one copy; the item
    // table of each run follows them in the staging buffer
    const size_t tables = (end - base + 255) & ~size_t(255);
    hipError_t e = c->staging.ensure(tables + max_items(pl) * sizeof(int2));
    if (e != hipSuccess) return e;
    char* h = static_cast<char*>(c->staging.p);
    char* dbase = cv.base + base;
    auto put = [&](const void* dst, const void* src, size_t bytes) {
        if (bytes) std::memcpy(h + (static_cast<const char*>(dst) - dbase), src, bytes);
    };
    put(d.order, pl.order.data(), pl.order.size() * 4);
    put(d.rounds, pl.rounds.data(), pl.rounds.size() * 4);
    put(lk, pl.link_key.data(), pl.link_key.size() * 4);
    put(d.keys, pl.keys.data(), pl.keys.size() * 4);
    pl.items_host = reinterpret_cast<int2*>(h + tables);
    return hipMemcpyAsync(dbase, h, end - base, hipMemcpyHostToDevice, c->stream);
}

// Runs a planned call on device arrays (q1, dir, first) with a launch shape.
hipError_t run_edges(rvs_context* c, const EnvRec& ev, EdgePlan& pl, const float* q1, const float* dir,
                     int32_t row_stride, float step, int32_t* first, int64_t num_edges, const PlanDev& d,
                     int waves_per_cu, int spl, bool timing) {
    const int32_t items = (int32_t)item_table(pl, spl, pl.items_host);
    pl.gr.items = pl.items_dev;
    const QueryParams qp = params_of(c);
    hipError_t e = h2d(pl.items_dev, pl.items_host, (size_t)items * sizeof(int2), c->stream);
    if (e == hipSuccess && timing) e = hipEventRecord(c->ev[0], c->stream);
    if (e == hipSuccess)
        e = launch_edge_first(pl.gr, ev.dev, qp, q1, dir, row_stride, step, d.order, d.rounds, first, num_edges,
                              d.counter, items, c->num_cus * waves_per_cu, spl, d.keys, d.nkeys, c->stream);
    if (e == hipSuccess && timing) e = hipEventRecord(c->ev[1], c->stream);
    return e;
}

}  // namespace

rvs_status rvs_edge_first_collision(rvs_context* c, int32_t env, const int32_t* robots, const float* q1,
                                    const float* dir, const int32_t* K, int64_t num_edges, int32_t row_stride,
                                    float step, int32_t* first, uint32_t flags) {
    if (!c) return RVS_EINVAL;
    if (num_edges < 0) return fail(c, RVS_EINVAL, "rvs_edge_first_collision: negative edge count");
    if (num_edges > 0 && (!robots || !q1 || !dir || !K || !first))
        return fail(c, RVS_EINVAL, "rvs_edge_first_collision: null argument");
    if (num_edges > INT32_MAX) return fail(c, RVS_ERANGE, "rvs_edge_first_collision: too many edges");
    CtxLock lk(c);
    if (!env_ok(c, env)) return fail(c, RVS_EINVAL, "rvs_edge_first_collision: bad env id");
    const EnvRec& ev = c->envs[env];
    if (prec_of(c, flags) == RVS_PREC_FP64)
        return fail(c, RVS_ENOTSUP, "rvs_edge_first_collision: reference arithmetic only");
    if (num_edges == 0) return RVS_OK;
    EdgePlan pl;
    rvs_status s = plan_edges(c, robots, K, num_edges, row_stride, pl);
    if (s != RVS_OK) return s;
    const size_t lds = edge_first_lds_bytes(pl.gr, ev.dev);
    if (lds > device_lds_limit(c->opt.device)) return fail(c, RVS_ERANGE, "rvs_edge_first_collision: robots too large for LDS staging");
    (void)hipSetDevice(c->opt.device);
    const size_t rows = (size_t)num_edges * row_stride;
    std::vector<size_t> need = plan_bytes(pl, ev);
    if (!(flags & RVS_F_DEVICE)) need.insert(need.end(), {rows * 4, rows * 4, (size_t)num_edges * 4});
    size_t total = 256;
    for (size_t b : need) total += (b + 255) & ~size_t(255);
    hipError_t e = c->arena.ensure(total);
    if (e != hipSuccess) return hip_fail(c, e, "rvs_edge_first_collision: staging");
    Carve cv(c->arena.p);
    PlanDev pd{};
    const float *qa = q1, *qd = dir;
    int32_t* fd = first;
    e = upload_plan(c, ev, pl, cv, pd);
    if (e == hipSuccess && !(flags & RVS_F_DEVICE)) {
        float* as = cv.take<float>(rows);
        float* ds = cv.take<float>(rows);
        fd = cv.take<int32_t>(num_edges);
        e = h2d(as, q1, rows * 4, c->stream);
        if (e == hipSuccess) e = h2d(ds, dir, rows * 4, c->stream);
        qa = as;
        qd = ds;
    }
    const bool timing = (flags & RVS_F_TIMING) != 0;
    if (e == hipSuccess)
        e = run_edges(c, ev, pl, qa, qd, row_stride, step, fd, num_edges, pd, c->edge_waves_per_cu, c->edge_spl,
                      timing);
    if (e == hipSuccess && !(flags & RVS_F_DEVICE)) e = d2h(first, fd, (size_t)num_edges * 4, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "rvs_edge_first_collision");
    if (timing) {
        c->nstages = 1;
        c->stage_ms[0] = 0;
        (void)hipEventElapsedTime(&c->stage_ms[0], c->ev[0], c->ev[1]);
    }
    return RVS_OK;
}

rvs_status rvs_set_edge_tuning(rvs_context* c, const rvs_edge_tuning* t) {
    if (!c || !t) return fail(c, RVS_EINVAL, "rvs_set_edge_tuning: null argument");
    if (t->waves_per_cu < 1 || t->waves_per_cu > 64 || t->slots_per_lane < 1 || t->slots_per_lane > 16)
        return fail(c, RVS_EINVAL, "rvs_set_edge_tuning: waves_per_cu 1..64, slots_per_lane 1..16");
    CtxLock lk(c);
    c->edge_waves_per_cu = t->waves_per_cu;
    c->edge_spl = t->slots_per_lane;
    return RVS_OK;
}

rvs_status rvs_autotune_edges(rvs_context* c, int32_t env, const int32_t* robots, const float* q1, const float* dir,
                              const int32_t* K, int64_t num_edges, int32_t row_stride, float step, uint32_t flags,
                              rvs_edge_tuning* best) {
    if (!c) return RVS_EINVAL;
    if (num_edges <= 0 || num_edges > INT32_MAX || !robots || !q1 || !dir || !K)
        return fail(c, RVS_EINVAL, "rvs_autotune_edges: bad arguments");
    CtxLock lk(c);
    if (!env_ok(c, env)) return fail(c, RVS_EINVAL, "rvs_autotune_edges: bad env id");
    const EnvRec& ev = c->envs[env];
    EdgePlan pl;
    rvs_status s = plan_edges(c, robots, K, num_edges, row_stride, pl);
    if (s != RVS_OK) return s;
    if (edge_first_lds_bytes(pl.gr, ev.dev) > device_lds_limit(c->opt.device))
        return fail(c, RVS_ERANGE, "rvs_autotune_edges: robots too large for LDS staging");
    (void)hipSetDevice(c->opt.device);
    const size_t rows = (size_t)num_edges * row_stride;
    std::vector<size_t> need = plan_bytes(pl, ev);
    need.insert(need.end(), {rows * 4, rows * 4, (size_t)num_edges * 4});
    size_t total = 256;
    for (size_t b : need) total += (b + 255) & ~size_t(255);
    hipError_t e = c->arena.ensure(total);
    if (e != hipSuccess) return hip_fail(c, e, "rvs_autotune_edges: staging");
    Carve cv(c->arena.p);
    PlanDev pd{};
    e = upload_plan(c, ev, pl, cv, pd);
    float* as = cv.take<float>(rows);
    float* ds = cv.take<float>(rows);
    int32_t* fd = cv.take<int32_t>(num_edges);
    const hipMemcpyKind kind = (flags & RVS_F_DEVICE) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (e == hipSuccess) e = hipMemcpyAsync(as, q1, rows * 4, kind, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(ds, dir, rows * 4, kind, c->stream);
    int win_w = 16, win_s = 1;
    float win_ms = 1e30f;
    for (int w : {4, 8, 16, 32}) {
        for (int spl : {1, 2, 4}) {
            float best_ms = 1e30f;
            for (int rep = 0; rep < 3 && e == hipSuccess; ++rep) {   // first run warms up
                e = run_edges(c, ev, pl, as, ds, row_stride, step, fd, num_edges, pd, w, spl, true);
                if (e == hipSuccess) e = hipEventSynchronize(c->ev[1]);
                float ms = 0;
                if (e == hipSuccess) e = hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
                if (rep > 0) best_ms = std::min(best_ms, ms);
            }
            if (e != hipSuccess) return hip_fail(c, e, "rvs_autotune_edges");
            if (best_ms < win_ms) { win_ms = best_ms; win_w = w; win_s = spl; }
        }
    }
    c->edge_waves_per_cu = win_w;
    c->edge_spl = win_s;
    if (best) *best = rvs_edge_tuning{win_w, win_s, win_ms};
    return RVS_OK;
}

}  // extern "C"
