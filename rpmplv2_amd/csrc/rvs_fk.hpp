// Forward kinematics and capsule skeletons for the gfx950 kernels.
//
// A robot is a serial chain of segments; segment k has the pose
//     S_k(q) = Frame(Rot(a_k, q) * R_k, o_k)          (revolute)
//     S_k    = Frame(R_k, o_k)                         (fixed)
// where R_k/o_k is the URDF joint origin and a_k = R_k * axis normalised.
// This is the pose KDL gives a kdl_parser segment (Joint::pose(q) * f_tip with
// f_tip = Joint::pose(0)^-1 * F_parent_jnt), and the frame of segment k is the
// left-to-right product S_0 * S_1 * ... * S_k that TreeFkSolverPos_recursive
// forms (xArm6.cpp:114-129, Planar2DOF.cpp:89-104).  All of it runs in double,
// as the reference's KDL does.
//
// The capsule skeleton is described by a per-point recipe: point k is the
// origin of frame f_k moved by pre_k along axis a_k of that frame (in double),
// rounded to the working type, then moved by -post_k along the same axis in
// the working type.  Chains use P_k = origin of frame k (Planar2DOF.cpp:
// 117-130); the xArm6 recipe (xArm6.cpp:209-235, gripper shift :130, :231)
// becomes P3 = F3 - 0.25 z3, P5 = F4 + 0.076 x4, P6 = F5 + g z5 then
// - (0.3 g) z5 in float.
//
// The chain is walked once per configuration, front to back: a kernel asks
// for skeleton points in increasing order and the stream applies segments
// only as far as the next point needs, keeping one 3x4 frame in registers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rvs_geometry.hpp"

namespace rvs {

constexpr int kMaxSeg = 72;
constexpr int kMaxLinks = 64;

struct SegDev {
    double R[9];      // joint origin rotation (row major)
    double axis[3];   // rotation axis in the parent frame, unit length
    double o[3];      // joint origin translation
    int32_t rev;      // 1 revolute, 0 fixed
    int32_t pad;
};

// The FP32 bulk pass's copy of a segment, in float: the origin rotation R
// and K R and K^2 R, K the cross-product matrix of the unit axis, so that the
// segment's rotation at angle q, Rot(axis, q) R, is R + sin q K R +
// (1 - cos q) K^2 R (Rodrigues' formula): 18 fused multiply-adds, with no
// conversion from double on the device.
struct SegF {
    float R[9];
    float KR[9];
    float K2R[9];
    float o[3];       // joint origin translation
    int32_t rev;      // 1 revolute, 0 fixed
    int32_t pad;
};

struct PointDev {
    int32_t frame;    // frame index, -1 for the base origin
    int32_t axis;     // 0 x, 1 y, 2 z column of that frame
    double pre;       // double-precision offset along the axis
    double post;      // offset subtracted along the axis after rounding,
                      // itself rounded to the working type first
    int32_t flags;    // bit0: apply pre, bit1: apply post
    int32_t pad;
};

struct RobotDev {
    int32_t num_segments;
    int32_t num_dofs;
    int32_t num_links;
    int32_t ground_included;
    int32_t num_pairs;
    int32_t kind;            // skeleton recipe (RVS_SKEL_CHAIN / RVS_SKEL_XARM6)
    const SegDev* seg;
    const SegF* segf;        // the same segments for the FP32 bulk pass
    const PointDev* pts;     // num_links + 1
    const float* radius;     // num_links
    const float* rnew32;     // collision-tree inflated radius, float
    const double* rnew64;    // same in double
    const int32_t* pairs;    // self-collision pairs (a, b)
    const float* rad_enc;    // radii added by computeEnclosingRadii (capsules_radius, or xArm6's capsules_radius_new)
};

// 3x3 rotation held in scalars (row major: m0 m1 m2 / m3 m4 m5 / m6 m7 m8)
// so that frames stay in registers.
struct Mat3 {
    double m0, m1, m2, m3, m4, m5, m6, m7, m8;
};

__host__ __device__ inline Mat3 mat_load(const double* a) {
    return Mat3{a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8]};
}

// Product with each entry summed left to right, as KDL's Rotation operator*
__host__ __device__ inline Mat3 mat_mul(const Mat3& a, const Mat3& b) {
    return Mat3{a.m0 * b.m0 + a.m1 * b.m3 + a.m2 * b.m6, a.m0 * b.m1 + a.m1 * b.m4 + a.m2 * b.m7,
                a.m0 * b.m2 + a.m1 * b.m5 + a.m2 * b.m8, a.m3 * b.m0 + a.m4 * b.m3 + a.m5 * b.m6,
                a.m3 * b.m1 + a.m4 * b.m4 + a.m5 * b.m7, a.m3 * b.m2 + a.m4 * b.m5 + a.m5 * b.m8,
                a.m6 * b.m0 + a.m7 * b.m3 + a.m8 * b.m6, a.m6 * b.m1 + a.m7 * b.m4 + a.m8 * b.m7,
                a.m6 * b.m2 + a.m7 * b.m5 + a.m8 * b.m8};
}

// KDL::Rotation::Rot2: Rodrigues' formula with the unit axis a.
__host__ __device__ inline Mat3 rot2(const double* a, double st, double ct) {
    const double a0 = a[0], a1 = a[1], a2 = a[2];
    const double vt = 1 - ct;
    const double m_vt_0 = vt * a0, m_vt_1 = vt * a1, m_vt_2 = vt * a2;
    const double m_st_0 = a0 * st, m_st_1 = a1 * st, m_st_2 = a2 * st;
    const double m_vt_0_1 = m_vt_0 * a1, m_vt_0_2 = m_vt_0 * a2, m_vt_1_2 = m_vt_1 * a2;
    return Mat3{ct + m_vt_0 * a0,  -m_st_2 + m_vt_0_1, m_st_1 + m_vt_0_2,
                m_st_2 + m_vt_0_1, ct + m_vt_1 * a1,   -m_st_0 + m_vt_1_2,
                -m_st_1 + m_vt_0_2, m_st_0 + m_vt_1_2, ct + m_vt_2 * a2};
}

struct FKStream {
    Mat3 M;
    double px, py, pz;
    int seg;   // next segment to apply
    int qi;    // next joint coordinate
};

__host__ __device__ __attribute__((always_inline)) inline void fk_init(FKStream& s) {
    s.M = Mat3{1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0};
    s.px = s.py = s.pz = 0.0;
    s.seg = 0;
    s.qi = 0;
}

// Apply segments until frame 'upto' is the current frame.  sc(k, &s, &c)
// gives the sine and cosine of joint k (in double, of the joint's float
// value).
template <class SinCos>
__host__ __device__ __attribute__((always_inline)) inline void fk_advance_sc(FKStream& s, const SegDev* seg, int upto,
                                                                           const SinCos& sc) {
    while (s.seg <= upto) {
        const SegDev& g = seg[s.seg];
        const Mat3 R = mat_load(g.R);
        Mat3 SM;
        if (g.rev) {
            double st, ct;
            sc(s.qi++, &st, &ct);
            SM = mat_mul(rot2(g.axis, st, ct), R);
        } else {
            SM = R;
        }
        const double o0 = g.o[0], o1 = g.o[1], o2 = g.o[2];
        const Mat3& M = s.M;
        const double nx = M.m0 * o0 + M.m1 * o1 + M.m2 * o2 + s.px;
        const double ny = M.m3 * o0 + M.m4 * o1 + M.m5 * o2 + s.py;
        const double nz = M.m6 * o0 + M.m7 * o1 + M.m8 * o2 + s.pz;
        s.M = mat_mul(M, SM);
        s.px = nx;
        s.py = ny;
        s.pz = nz;
        ++s.seg;
    }
}

// The same with the joint values read from q: joint k is q[k * qstride].
__host__ __device__ __attribute__((always_inline)) inline void fk_advance(FKStream& s, const SegDev* seg,
                                                                        const float* q, int upto, int qstride = 1) {
    fk_advance_sc(s, seg, upto, [&](int k, double* st, double* ct) { sincos((double)q[k * qstride], st, ct); });
}

// Skeleton point k from the frames of an FK stream advanced by sc (points
// must be requested in increasing k).
template <typename T, class SinCos>
__host__ __device__ __attribute__((always_inline)) inline Vec3<T> skel_point_sc(FKStream& s, const RobotDev& rb, int k,
                                                                              const SinCos& sc) {
    const PointDev pt = rb.pts[k];
    if (pt.frame < 0) return Vec3<T>{T(0), T(0), T(0)};
    fk_advance_sc(s, rb.seg, pt.frame, sc);
    double x = s.px, y = s.py, z = s.pz;
    // column 'axis' of the frame, picked with one-hot weights (exact: the
    // other two terms are signed zeros) rather than an indexed load, which
    // would keep the frame in scratch memory
    const double w0 = pt.axis == 0 ? 1.0 : 0.0, w1 = pt.axis == 1 ? 1.0 : 0.0, w2 = pt.axis == 2 ? 1.0 : 0.0;
    const Mat3& M = s.M;
    const double ax = M.m0 * w0 + M.m1 * w1 + M.m2 * w2;
    const double ay = M.m3 * w0 + M.m4 * w1 + M.m5 * w2;
    const double az = M.m6 * w0 + M.m7 * w1 + M.m8 * w2;
    if (pt.flags & 1) {
        x = x + pt.pre * ax;
        y = y + pt.pre * ay;
        z = z + pt.pre * az;
    }
    Vec3<T> P{(T)x, (T)y, (T)z};
    if (pt.flags & 2) {
        const T f = (T)pt.post;
        P.x = P.x - f * (T)ax;
        P.y = P.y - f * (T)ay;
        P.z = P.z - f * (T)az;
    }
    return P;
}

// Skeleton point k with the joint values read from q (joint m is
// q[m * qstride]; points must be requested in increasing k).
template <typename T>
__host__ __device__ __attribute__((always_inline)) inline Vec3<T> skel_point(FKStream& s, const RobotDev& rb,
                                                                           const float* q, int k, int qstride = 1) {
    return skel_point_sc<T>(s, rb, k,
                            [&](int m, double* st, double* ct) { sincos((double)q[m * qstride], st, ct); });
}

}  // namespace rvs
