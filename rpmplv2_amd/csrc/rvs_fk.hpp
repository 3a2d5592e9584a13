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

#pragma clang fp contract(off)

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
    int32_t pad;
    const SegDev* seg;
    const PointDev* pts;     // num_links + 1
    const float* radius;     // num_links
    const float* rnew32;     // collision-tree inflated radius, float
    const double* rnew64;    // same in double
    const int32_t* pairs;    // self-collision pairs (a, b)
};

// KDL::Rotation::Rot2: Rodrigues' formula with the unit axis a.
__host__ __device__ inline void rot2(const double* a, double st, double ct, double* M) {
    const double vt = 1 - ct;
    const double m_vt_0 = vt * a[0], m_vt_1 = vt * a[1], m_vt_2 = vt * a[2];
    const double m_st_0 = a[0] * st, m_st_1 = a[1] * st, m_st_2 = a[2] * st;
    const double m_vt_0_1 = m_vt_0 * a[1], m_vt_0_2 = m_vt_0 * a[2], m_vt_1_2 = m_vt_1 * a[2];
    M[0] = ct + m_vt_0 * a[0];
    M[1] = -m_st_2 + m_vt_0_1;
    M[2] = m_st_1 + m_vt_0_2;
    M[3] = m_st_2 + m_vt_0_1;
    M[4] = ct + m_vt_1 * a[1];
    M[5] = -m_st_0 + m_vt_1_2;
    M[6] = -m_st_1 + m_vt_0_2;
    M[7] = m_st_0 + m_vt_1_2;
    M[8] = ct + m_vt_2 * a[2];
}

// 3x3 product, each entry summed left to right as KDL does
__host__ __device__ inline void mat_mul(const double* a, const double* b, double* r) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            r[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
}

struct FKStream {
    double M[9];
    double p[3];
    int seg;   // next segment to apply
    int qi;    // next joint coordinate
};

__host__ __device__ inline void fk_init(FKStream& s) {
#pragma unroll
    for (int i = 0; i < 9; ++i) s.M[i] = (i % 4 == 0) ? 1.0 : 0.0;
    s.p[0] = s.p[1] = s.p[2] = 0.0;
    s.seg = 0;
    s.qi = 0;
}

// Apply segments until frame 'upto' is the current frame.
__host__ __device__ inline void fk_advance(FKStream& s, const SegDev* seg, const float* q, int upto) {
    while (s.seg <= upto) {
        const SegDev& g = seg[s.seg];
        double SM[9];
        if (g.rev) {
            double st, ct;
            sincos((double)q[s.qi++], &st, &ct);
            double Rq[9];
            rot2(g.axis, st, ct, Rq);
            mat_mul(Rq, g.R, SM);
        } else {
#pragma unroll
            for (int i = 0; i < 9; ++i) SM[i] = g.R[i];
        }
        double np[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
            np[i] = s.M[3 * i] * g.o[0] + s.M[3 * i + 1] * g.o[1] + s.M[3 * i + 2] * g.o[2] + s.p[i];
        double nM[9];
        mat_mul(s.M, SM, nM);
#pragma unroll
        for (int i = 0; i < 9; ++i) s.M[i] = nM[i];
#pragma unroll
        for (int i = 0; i < 3; ++i) s.p[i] = np[i];
        ++s.seg;
    }
}

// Skeleton point k (points must be requested in increasing k).
template <typename T>
__host__ __device__ inline Vec3<T> skel_point(FKStream& s, const RobotDev& rb, const float* q, int k) {
    const PointDev pt = rb.pts[k];
    if (pt.frame < 0) return Vec3<T>{T(0), T(0), T(0)};
    fk_advance(s, rb.seg, q, pt.frame);
    double x = s.p[0], y = s.p[1], z = s.p[2];
    const int a = pt.axis;
    if (pt.flags & 1) {
        x = x + pt.pre * s.M[a];
        y = y + pt.pre * s.M[3 + a];
        z = z + pt.pre * s.M[6 + a];
    }
    Vec3<T> P{(T)x, (T)y, (T)z};
    if (pt.flags & 2) {
        const T f = (T)pt.post;
        P.x = P.x - f * (T)s.M[a];
        P.y = P.y - f * (T)s.M[3 + a];
        P.z = P.z - f * (T)s.M[6 + a];
    }
    return P;
}

}  // namespace rvs
