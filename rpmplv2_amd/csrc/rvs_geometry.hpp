// Capsule geometry for the gfx950 kernels: capsule-box distance, capsule-box
// collision, segment-segment and segment-point distance.
//
// Semantics follow RPMPLv2's CollisionAndDistance.cpp (cited per function);
// the control flow is re-shaped for 64-wide wavefronts.  The reference walks
// a deep case tree and calls the segment-segment routine from a dozen call
// sites; here every case first selects a small work list (faces to test for a
// crossing, box edges to measure, or one corner point) and then ONE loop runs
// the segment-segment routine over the list.  Lanes that took different cases
// therefore share the same instructions and only differ in trip count.
//
// Every arithmetic expression keeps the reference's operation order: 3-vector
// dot products sum as e0 + (e1 + e2), 2-vectors as e0 + e1, lerps are
// a + t * (b - a) per component.  Division and square root go through a math
// policy: ExactMath (IEEE division and square root) in translation units
// compiled without FMA contraction gives results bit-identical to the
// reference's float arithmetic (T = float) or the same algorithm in FP64
// (T = double); FastMath (hardware reciprocal and square root, contraction
// allowed by its translation unit) is the relaxed FP32 bulk pass, which runs
// flat_capsule_box, a branch-free form of the capsule-box distance, instead of
// the case tree.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rvs {

// Run-time indexed reads go through a register vector (extractelement),
// never through a stack copy: an indexed load would put the operand in
// scratch memory.
template <typename T> using vec4_t = T __attribute__((ext_vector_type(4)));

template <typename T>
__host__ __device__ inline T pick(T a, T b, T c, int k) {
    const vec4_t<T> v = {a, b, c, T(0)};
    return v[k];
}

template <typename T>
struct Vec3 {
    T x, y, z;
    __host__ __device__ T operator[](int k) const { return pick(x, y, z, k); }
};

template <typename T> __host__ __device__ inline Vec3<T> mk(T x, T y, T z) { return Vec3<T>{x, y, z}; }
template <typename T> __host__ __device__ inline Vec3<T> operator-(const Vec3<T>& a, const Vec3<T>& b) {
    return Vec3<T>{a.x - b.x, a.y - b.y, a.z - b.z};
}
template <typename T> __host__ __device__ inline T dot(const Vec3<T>& a, const Vec3<T>& b) {
    return a.x * b.x + (a.y * b.y + a.z * b.z);
}
template <typename T> __host__ __device__ inline T sqnorm(const Vec3<T>& a) { return dot(a, a); }
// IEEE square root of the operand's own precision
__host__ __device__ inline float vsqrt(float x) { return sqrtf(x); }
__host__ __device__ inline double vsqrt(double x) { return sqrt(x); }
struct ExactMath {
    template <typename T> __host__ __device__ static T div(T a, T b) { return a / b; }
    template <typename T> __host__ __device__ static T root(T x) { return vsqrt(x); }
    template <typename T> __host__ __device__ static T clamp(T x, T lo, T hi) { return fmin(fmax(x, lo), hi); }
};

// Relaxed FP32: v_rcp_f32 / v_sqrt_f32 (about 1 ulp), device only.
struct FastMath {
    __device__ static float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    __device__ static float root(float x) { return __builtin_amdgcn_sqrtf(x); }
    // x clamped to [lo, hi] (lo <= hi): the median of the three, one
    // instruction (the value of min(max(x, lo), hi) for finite x)
    __device__ static float clamp(float x, float lo, float hi) { return __builtin_amdgcn_fmed3f(x, lo, hi); }
};

template <typename T, class MP = ExactMath>
__host__ __device__ inline T vnorm(const Vec3<T>& a) { return MP::root(sqnorm(a)); }
// a + s * (b - a)
template <typename T> __host__ __device__ inline Vec3<T> lerp(const Vec3<T>& a, const Vec3<T>& b, T s) {
    return Vec3<T>{a.x + s * (b.x - a.x), a.y + s * (b.y - a.y), a.z + s * (b.z - a.z)};
}
template <typename T> __host__ __device__ inline T smin(T a, T b) { return (b < a) ? b : a; }  // std::min

// Axis-aligned box (xmin, ymin, zmin, xmax, ymax, zmax).
template <typename T>
struct Box {
    T lo[3], hi[3];
    // run-time indexed accessors written as selects so a box held in
    // registers is never spilled to scratch for indexing
    __host__ __device__ T lo_at(int k) const { return pick(lo[0], lo[1], lo[2], k); }
    __host__ __device__ T hi_at(int k) const { return pick(hi[0], hi[1], hi[2], k); }
    __host__ __device__ T f(int face) const { return face < 3 ? lo_at(face) : hi_at(face - 3); }
    // corner c: bit0 -> x max, bit1 -> y max, bit2 -> z max
    __host__ __device__ Vec3<T> corner(int c) const {
        return Vec3<T>{(c & 1) ? hi[0] : lo[0], (c & 2) ? hi[1] : lo[1], (c & 4) ? hi[2] : lo[2]};
    }
};

template <typename T>
struct Params {
    T eq;    // EQUALITY_THRESHOLD
    T maxd;  // MAX_DISTANCE
};

// Nearest points (robot point, obstacle point) of one pair.
template <typename T>
struct NP {
    Vec3<T> r, o;
    bool valid;
};

// ---------------------------------------------------------------- seg-seg
// distanceLineSegToLineSeg, CollisionAndDistance.cpp:217-320.  Returns the
// distance of AB to CD; np receives (point on AB, point on CD), invalid when
// the reference returns no nearest points.
template <typename T, bool WANT_NP, class MP = ExactMath>
__host__ __device__ inline T seg_seg(const Vec3<T>& A, const Vec3<T>& B, const Vec3<T>& C, const Vec3<T>& D,
                                     const Params<T>& p, NP<T>* np) {
    const Vec3<T> BA = B - A, DC = D - C, CD = C - D, AC = A - C, AB = A - B;
    const T alpha1 = sqnorm(BA);
    const T alpha2 = dot(BA, DC);
    const T beta1 = dot(CD, BA);
    const T beta2 = dot(CD, DC);
    const T gamma1 = dot(AC, AB);
    const T gamma2 = dot(AC, CD);
    const T s = MP::div(alpha1 * gamma2 - alpha2 * gamma1, alpha1 * beta2 - alpha2 * beta1);
    const T t = MP::div(gamma1 - beta1 * s, alpha1);
    T d_c = p.maxd;
    if (t > 0 && t < 1 && s > 0 && s < 1) {
        const Vec3<T> P1 = lerp(A, B, t), P2 = lerp(C, D, s);
        d_c = vnorm<T, MP>(P2 - P1);
        if (d_c < p.eq) {
            if (WANT_NP) np->valid = false;
            return T(0);
        }
        if (WANT_NP) { np->r = P1; np->o = P2; np->valid = true; }
        return d_c;
    }
    // the four clamped candidates: C and D projected on AB, A and B on CD
    const T alpha3 = sqnorm(CD);
    const Vec3<T> AD = A - D, BC = B - C;
    const T opt0 = MP::div(dot(AC, AB), alpha1);
    const T opt1 = MP::div(dot(AD, AB), alpha1);
    const T opt2 = MP::div(dot(AC, DC), alpha3);
    const T opt3 = MP::div(dot(BC, DC), alpha3);
    if (WANT_NP) np->valid = false;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const T o = (i == 0) ? opt0 : (i == 1) ? opt1 : (i == 2) ? opt2 : opt3;
        // end points chosen when the parameter is clamped (lines 251-286)
        Vec3<T> p0, p1;
        if (o < 0) {
            p0 = (i == 3) ? B : A;
            p1 = (i == 1) ? D : C;
        } else if (o > 1) {
            p0 = (i == 2) ? A : B;
            p1 = (i == 0) ? C : D;
        } else if (i < 2) {
            p0 = lerp(A, B, o);
            p1 = (i == 0) ? C : D;
        } else {
            p0 = (i == 2) ? A : B;
            p1 = lerp(C, D, o);
        }
        const T d_tmp = vnorm<T, MP>(p1 - p0);
        if (d_tmp < d_c) {
            d_c = d_tmp;
            if (WANT_NP) { np->r = p0; np->o = p1; np->valid = true; }
        }
    }
    return d_c;
}

// ---------------------------------------------------------------- seg-pt
// distanceLineSegToPoint, CollisionAndDistance.cpp:323-342
template <typename T, bool WANT_NP, class MP = ExactMath>
__host__ __device__ inline T seg_pt(const Vec3<T>& A, const Vec3<T>& B, const Vec3<T>& C, const Params<T>& p,
                                    NP<T>* np) {
    const Vec3<T> BA = B - A;
    const T t_opt = MP::div(dot(C - A, BA), sqnorm(BA));
    const Vec3<T> Pn = (t_opt < 0) ? A : ((t_opt > 1) ? B : lerp(A, B, t_opt));
    const T d_c = vnorm<T, MP>(C - Pn);
    if (d_c < p.eq) {
        if (WANT_NP) np->valid = false;
        return T(0);
    }
    if (WANT_NP) { np->r = Pn; np->o = C; np->valid = true; }
    return d_c;
}

// ---------------------------------------------------------------- faces
// Segment AB crosses the open rectangle of face 'face' (0..5 = xmin, ymin,
// zmin, xmax, ymax, zmax).  This is collisionCapsuleToRectangle
// (CollisionAndDistance.cpp:46-111) at radius 0, where only the plane-crossing
// test of lines 57-65 can report a hit.
template <typename T, class MP = ExactMath>
__host__ __device__ inline bool face_crossed(const Vec3<T>& A, const Vec3<T>& B, const Box<T>& b, int face,
                                             const Params<T>& p) {
    const int c = face % 3;
    const int o1 = (c == 0) ? 1 : 0, o2 = (c == 2) ? 1 : 2;
    const T v = b.f(face);
    const T den = B[c] - A[c];
    const T t = (den != 0) ? MP::div(v - A[c], den) : p.maxd;
    if (!(t > 0 && t < 1)) return false;
    const T m0 = A[o1] + t * (B[o1] - A[o1]);
    const T m1 = A[o2] + t * (B[o2] - A[o2]);
    return m0 > b.lo_at(o1) && m0 < b.hi_at(o1) && m1 > b.lo_at(o2) && m1 < b.hi_at(o2);
}

// Work list entries: an edge joins two box corners (3-bit corner indices),
// encoded as c0 | (c1 << 3).  Lists are generated on the fly from a small
// descriptor so they stay in registers:
//   kListExplicit  up to two edges packed in 'param' (6 bits each)
//   kListRing      the four edges of face 'param' (face 0..5)
//   kListAll       all twelve edges
// Ring order follows checkOtherCases (CollisionAndDistance.cpp:631-634,
// 693-696, 721-724): start at the face's (min, min) corner, walk through the
// corner that is max on the first free axis.  The twelve-edge order is
// bottom ring, top ring, then the vertical edges at (xmin,ymin), (xmax,ymin),
// (xmax,ymax), (xmin,ymax) (lines 747-751).
__host__ __device__ inline int edge_of(int c0, int c1) { return c0 | (c1 << 3); }

enum { kListExplicit = 0, kListRing = 1, kListAll = 2 };

__host__ __device__ inline int ring_corner(int face, int k) {
    const int c = face % 3;
    const int fix = (face >= 3) ? (1 << c) : 0;
    const int b1 = 1 << ((c == 0) ? 1 : 0), b2 = 1 << ((c == 2) ? 1 : 2);
    return fix | ((k == 1 || k == 2) ? b1 : 0) | ((k >= 2) ? b2 : 0);
}

__host__ __device__ inline int ring_edge(int face, int e) {
    return edge_of(ring_corner(face, e), ring_corner(face, (e + 1) & 3));
}

__host__ __device__ inline int list_edge(int kind, int param, int e) {
    if (kind == kListExplicit) return (param >> (6 * e)) & 63;
    if (kind == kListRing) return ring_edge(param, e);
    if (e < 8) return ring_edge(e < 4 ? 2 : 5, e & 3);
    const int v = e - 8;
    const int c = ((v == 1 || v == 2) ? 1 : 0) | ((v >= 2) ? 2 : 0);
    return edge_of(c, c | 4);
}

// getLineSegments (CollisionAndDistance.cpp:542-574): the edge(s) of face
// 'face' on the side where the other end point lies, first along the first
// free axis, then along the second.  Returns the count, edges packed.
template <typename T>
__host__ __device__ inline int face_side_edges(const Box<T>& b, int face, const Vec3<T>& other, int* packed) {
    const int c = face % 3;
    const int fix = (face >= 3) ? (1 << c) : 0;
    const int a1 = (c == 0) ? 1 : 0, a2 = (c == 2) ? 1 : 2;
    const int b1 = 1 << a1, b2 = 1 << a2;
    const T u = other[a1], v = other[a2];
    const T lo1 = b.lo_at(a1), hi1 = b.hi_at(a1);
    const T lo2 = b.lo_at(a2), hi2 = b.hi_at(a2);
    int n = 0, pk = 0;
    if (u < lo1) { pk = edge_of(fix, fix | b2); n = 1; }
    else if (u > hi1) { pk = edge_of(fix | b1, fix | b1 | b2); n = 1; }
    int e2 = -1;
    if (v < lo2) e2 = edge_of(fix, fix | b1);
    else if (v > hi2) e2 = edge_of(fix | b2, fix | b1 | b2);
    if (e2 >= 0) { pk |= e2 << (6 * n); ++n; }
    *packed = pk;
    return n;
}

// ---------------------------------------------------------------- capsule-box
// distanceCapsuleToBox -> CapsuleToBox::compute, CollisionAndDistance.cpp:
// 208-214, 395-755.  Returns d - radius (or 0 when an end point lies strictly
// inside the box).  'status' bit 0 is set when the reference would
// dereference a missing nearest-point matrix (end point projections more than
// MAX_DISTANCE away, or an end point exactly on a face plane with the other
// inside the face's span); the distance is then NaN.
template <typename T, bool WANT_NP, class MP = ExactMath>
__host__ __device__ inline T capsule_box(const Vec3<T>& A, const Vec3<T>& B, T radius, const Box<T>& b,
                                         const Params<T>& p, NP<T>* np, uint32_t* status) {
    // --- projectionLineSegOnSide x3 (lines 411-413, 450-474); bit f of the
    // masks is face f (xmin, ymin, zmin, xmax, ymax, zmax)
    int pmA = 0, pmB = 0;
    T distA = p.maxd, distB = p.maxd;
    bool inside = false;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int o1 = (k == 0) ? 1 : 0, o2 = (k == 2) ? 1 : 2;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const Vec3<T>& X = i ? B : A;
            int& pm = i ? pmB : pmA;
            T& dist = i ? distB : distA;
            if (X[o1] >= b.lo[o1] && X[o1] <= b.hi[o1] && X[o2] >= b.lo[o2] && X[o2] <= b.hi[o2]) {
                if (X[k] > b.lo[k] && X[k] < b.hi[k]) inside = true;
                else if (X[k] <= b.lo[k]) { pm |= 1 << k; dist = b.lo[k] - X[k]; }
                else if (X[k] >= b.hi[k]) { pm |= 8 << k; dist = X[k] - b.hi[k]; }
            }
        }
    }
    if (inside) {
        if (WANT_NP) np->valid = false;
        return T(0);
    }

    // --- work list
    T d_c = p.maxd;
    int nfaces = 0;       // faces to test for a crossing: 0, 1 or 6
    int face1 = 0;        // the single face when nfaces == 1
    int nedges = 0;
    int kind = kListExplicit, param = 0;
    bool direct = false;  // single edge / corner assigned directly (tie(...))
    int corner = -1;      // corner point case (segment-point distance)
    bool no_edge = false;

    const int any = pmA | pmB;
    if (any) {
        const bool useB = !(distA < distB);
        d_c = smin(distA, distB);
        const Vec3<T>& X = useB ? B : A;
        const int m = useB ? pmB : pmA;
        if (WANT_NP) {
            // first face (0..5 order) of the chosen point's projections
            const int fc = (m & 1) ? 0 : (m & 2) ? 1 : (m & 4) ? 2 : (m & 8) ? 3 : (m & 16) ? 4 : (m & 32) ? 5 : 0;
            const int ax = fc % 3;
            np->r = X;
            np->o = X;
            const T v = b.f(fc);
            if (ax == 0) np->o.x = v; else if (ax == 1) np->o.y = v; else np->o.z = v;
            np->valid = true;
        }
        if (!(pmA & pmB)) {
            // checkEdges (lines 476-540): face order xmin, xmax, ymin, ymax, zmin, zmax
            const int face = (m & 1) ? 0 : (m & 8) ? 3 : (m & 2) ? 1 : (m & 16) ? 4 : (m & 4) ? 2 : (m & 32) ? 5 : -1;
            if (face < 0) { if (status) *status |= 1u; return T(NAN); }
            nfaces = 1;
            face1 = face;
            nedges = face_side_edges(b, face, useB ? A : B, &param);
            // no edge: the reference would read an uninitialised matrix
            // (only if the face is not crossed, which is tested first)
            no_edge = (nedges == 0);
        }
    } else {
        // checkOtherCases (lines 597-755): classify each axis as both-below
        // (0), both-above (1) or mixed (2)
        const int s0 = (A.x < b.lo[0] && B.x < b.lo[0]) ? 0 : ((A.x > b.hi[0] && B.x > b.hi[0]) ? 1 : 2);
        const int s1 = (A.y < b.lo[1] && B.y < b.lo[1]) ? 0 : ((A.y > b.hi[1] && B.y > b.hi[1]) ? 1 : 2);
        const int s2 = (A.z < b.lo[2] && B.z < b.lo[2]) ? 0 : ((A.z > b.hi[2] && B.z > b.hi[2]) ? 1 : 2);
        const int nfix = (s0 != 2) + (s1 != 2) + (s2 != 2);
        if (nfix == 3) {
            corner = s0 | (s1 << 1) | (s2 << 2);
            direct = true;
        } else if (nfix == 2) {
            // edge along the free axis through the fixed coordinates
            const int free = (s0 == 2) ? 0 : ((s1 == 2) ? 1 : 2);
            const int base = ((s0 == 1) ? 1 : 0) | ((s1 == 1) ? 2 : 0) | ((s2 == 1) ? 4 : 0);
            param = edge_of(base, base | (1 << free));
            nedges = 1;
            direct = true;
        } else if (nfix == 1) {
            const int k = (s0 != 2) ? 0 : ((s1 != 2) ? 1 : 2);
            const int sk = (k == 0) ? s0 : ((k == 1) ? s1 : s2);
            kind = kListRing;
            param = k + 3 * sk;
            nedges = 4;
        } else {
            nfaces = 6;
            kind = kListAll;
            nedges = 12;
        }
    }

    // --- crossing tests: any hit makes the distance 0 (then -radius)
    bool hit = false;
    if (nfaces == 1) hit = face_crossed<T, MP>(A, B, b, face1, p);
    else if (nfaces == 6)
        for (int f = 0; f < 6; ++f) hit = hit || face_crossed<T, MP>(A, B, b, f, p);

    if (hit) {
        d_c = T(0);
    } else if (no_edge) {
        if (status) *status |= 1u;
        return T(NAN);
    } else if (corner >= 0) {
        NP<T> t;
        d_c = seg_pt<T, WANT_NP, MP>(A, B, b.corner(corner), p, &t);
        if (WANT_NP) *np = t;
    } else {
        for (int e = 0; e < nedges; ++e) {
            const int ed = list_edge(kind, param, e);
            NP<T> t;
            const T d_tmp = seg_seg<T, WANT_NP, MP>(A, B, b.corner(ed & 7), b.corner(ed >> 3), p, &t);
            if (direct) {
                d_c = d_tmp;
                if (WANT_NP) *np = t;
            } else if (d_tmp <= 0) {
                d_c = T(0);
                break;
            } else if (d_tmp < d_c) {
                d_c = d_tmp;
                if (WANT_NP) *np = t;
            }
        }
    }
    return d_c - radius;
}

// Capsule-box distance of the relaxed bulk pass without the case tree: a
// fixed sequence of operations, so that lanes evaluating different pairs in
// one step (the per-lane walk) do not split across the tree's branches.
// Away from contact the reference's case tree returns the exact segment-box
// distance minus r; at contact it returns 0 when an end point lies inside
// the box (CollisionAndDistance.cpp:414-418; a point counts when its
// coordinates are all within the box and one strictly) and -r when the
// segment crosses it (:447 after :525 or :584-587).  Here:
//   inside: the larger of the end points' depths (the smallest distance of
//     a coordinate to its slab bound) is > 0: the value is 0, uncertain
//     within band of the surface;
//   crossing: the segment meets the box shrunk by band on every side (then
//     it crosses the box whatever the rounding): the value is -r;
//   otherwise the exact distance raw of the segment to the box, minus r,
//     uncertain when raw < EQ + band (the snap to -r, a touching crossing,
//     the tree's face-plane path), |raw - r| < band, the segment nearly
//     parallel to a box axis within 1 cm (the reference's segment solves
//     are ill-conditioned there) or of zero length (the reference's
//     projections are 0/0 there).
// raw: the squared distance of P(t) = A + t D to the box is convex with
// derivative g(t) = sum_k D_k (P_k(t) - clamp(P_k(t))), continuous,
// piecewise linear and non-decreasing, with breaks where a coordinate meets a
// slab bound.  Its root lies between the largest break point t_a with
// g <= 0 and the smallest t_b with g > 0 (0 and 1 included), and g is linear
// between them: t* = -sum c_k D_k / sum D_k^2 over the axes outside their
// slab there (c_k = A_k - bound), clamped to [t_a, t_b].
template <class MP>
__host__ __device__ inline float flat_capsule_box(const Vec3<float>& A, const Vec3<float>& B, float r,
                                                  const Box<float>& b, const Params<float>& p, float band,
                                                  bool& uncertain) {
    const float dA = fminf(fminf(fminf(A.x - b.lo[0], b.hi[0] - A.x), fminf(A.y - b.lo[1], b.hi[1] - A.y)),
                           fminf(A.z - b.lo[2], b.hi[2] - A.z));
    const float dB = fminf(fminf(fminf(B.x - b.lo[0], b.hi[0] - B.x), fminf(B.y - b.lo[1], b.hi[1] - B.y)),
                           fminf(B.z - b.lo[2], b.hi[2] - B.z));
    const float din = fmaxf(dA, dB);
    const float Dx = B.x - A.x, Dy = B.y - A.y, Dz = B.z - A.z;
    const float dd = Dx * Dx + (Dy * Dy + Dz * Dz);
    const float ix = MP::div(1.0f, Dx), iy = MP::div(1.0f, Dy), iz = MP::div(1.0f, Dz);
    // crossing of the box shrunk by band (slab clipping).  An axis the
    // segment does not move along has an infinite reciprocal: its slab
    // parameters are -inf and +inf when the coordinate lies strictly inside
    // the shrunk slab (the interval unchanged), both +inf or both -inf when
    // outside (the interval empty), and NaN with an infinity when exactly on
    // a shrunk bound (min / max drop the NaN: the interval empty again).
    float t0 = 0.0f, t1 = 1.0f;
    {
        auto slab = [&](float inv, float a0, float lo, float hi) {
            const float a = (lo + band - a0) * inv, c = (hi - band - a0) * inv;
            t0 = fmaxf(t0, fminf(a, c));
            t1 = fminf(t1, fmaxf(a, c));
        };
        slab(ix, A.x, b.lo[0], b.hi[0]);
        slab(iy, A.y, b.lo[1], b.hi[1]);
        slab(iz, A.z, b.lo[2], b.hi[2]);
    }
    const bool deep = t0 <= t1 && b.hi[0] - b.lo[0] > 2.0f * band && b.hi[1] - b.lo[1] > 2.0f * band &&
                      b.hi[2] - b.lo[2] > 2.0f * band;
    float ta = 0.0f, tb = 1.0f;   // the bracket
    // g (the derivative of half the squared distance) at t = 0 and t = 1,
    // where the points are A and B themselves
    const float g0 = Dx * (A.x - MP::clamp(A.x, b.lo[0], b.hi[0])) +
                     (Dy * (A.y - MP::clamp(A.y, b.lo[1], b.hi[1])) + Dz * (A.z - MP::clamp(A.z, b.lo[2], b.hi[2])));
    const float g1 = Dx * (B.x - MP::clamp(B.x, b.lo[0], b.hi[0])) +
                     (Dy * (B.y - MP::clamp(B.y, b.lo[1], b.hi[1])) + Dz * (B.z - MP::clamp(B.z, b.lo[2], b.hi[2])));
    // every break point evaluated, its bracket update selected (a break
    // point outside (0, 1), or infinite or NaN for an axis the segment does
    // not move along, changes nothing)
    // g at a break point of axis K, without axis K's term (zero there up to
    // rounding: the point lies on that slab bound)
    auto gskip = [&](float t, int K) {
        const float px = A.x + t * Dx, py = A.y + t * Dy, pz = A.z + t * Dz;
        const float ex = Dx * (px - MP::clamp(px, b.lo[0], b.hi[0]));
        const float ey = Dy * (py - MP::clamp(py, b.lo[1], b.hi[1]));
        const float ez = Dz * (pz - MP::clamp(pz, b.lo[2], b.hi[2]));
        return K == 0 ? ey + ez : (K == 1 ? ex + ez : ex + ey);
    };
    auto consider = [&](float t, int K) {
        const bool in = t > 0.0f && t < 1.0f;
        const float gt = gskip(t, K);
        ta = (in && gt <= 0.0f) ? fmaxf(ta, t) : ta;
        tb = (in && !(gt <= 0.0f)) ? fminf(tb, t) : tb;
    };
    consider((b.lo[0] - A.x) * ix, 0);
    consider((b.hi[0] - A.x) * ix, 0);
    consider((b.lo[1] - A.y) * iy, 1);
    consider((b.hi[1] - A.y) * iy, 1);
    consider((b.lo[2] - A.z) * iz, 2);
    consider((b.hi[2] - A.z) * iz, 2);
    float ts;
    if (g0 >= 0.0f) {
        ts = 0.0f;
    } else if (g1 <= 0.0f) {
        ts = 1.0f;
    } else {
        // the axes outside their slab on (ta, tb), at its midpoint
        const float tm = 0.5f * (ta + tb);
        // (an axis is outside where its coordinate differs from its clamp,
        // and then the clamp is the bound it is past)
        const float px = A.x + tm * Dx, py = A.y + tm * Dy, pz = A.z + tm * Dz;
        const float qx = MP::clamp(px, b.lo[0], b.hi[0]);
        const float qy = MP::clamp(py, b.lo[1], b.hi[1]);
        const float qz = MP::clamp(pz, b.lo[2], b.hi[2]);
        const float cx = px != qx ? A.x - qx : 0.0f, wx = px != qx ? Dx : 0.0f;
        const float cy = py != qy ? A.y - qy : 0.0f, wy = py != qy ? Dy : 0.0f;
        const float cz = pz != qz ? A.z - qz : 0.0f, wz = pz != qz ? Dz : 0.0f;
        const float num = cx * wx + (cy * wy + cz * wz);
        const float den = wx * wx + (wy * wy + wz * wz);
        ts = den > 0.0f ? MP::clamp(-MP::div(num, den), ta, tb) : ta;
    }
    const float px = A.x + ts * Dx, py = A.y + ts * Dy, pz = A.z + ts * Dz;
    const float ex = px - MP::clamp(px, b.lo[0], b.hi[0]);
    const float ey = py - MP::clamp(py, b.lo[1], b.hi[1]);
    const float ez = pz - MP::clamp(pz, b.lo[2], b.hi[2]);
    const float raw = MP::root(ex * ex + (ey * ey + ez * ez));
    // screening and the value
    if (din > 0.0f) {
        uncertain |= din < band;
        return 0.0f;
    }
    // an end point within band of the surface: the reference's inside test
    // (which also counts points on a face) could go either way
    uncertain |= din > -band;
    if (deep) return -r;
    const float dt = raw - r;
    const float smin2 = fminf(fminf(Dy * Dy + Dz * Dz, Dx * Dx + Dz * Dz), Dx * Dx + Dy * Dy);   // dd sin^2 to the nearest axis
    uncertain |= (raw < p.eq + band) || (fabsf(dt) < band) || !(dd > 0.0f) || (raw < 0.01f && smin2 < 1e-6f * dd);
    return dt;
}

// ---------------------------------------------------------------- collision
// checkCases, CollisionAndDistance.cpp:113-148: distance of AB to the one or
// two rectangle edges on the side of point 'pt' (2-D coordinates u, v).
template <typename T>
__host__ __device__ inline T check_cases(const Vec3<T>& A, const Vec3<T>& B, const Box<T>& b, int face, T u, T v,
                                         const Params<T>& p) {
    const int c = face % 3;
    const int fix = (face >= 3) ? (1 << c) : 0;
    const int a1 = (c == 0) ? 1 : 0, a2 = (c == 2) ? 1 : 2;
    const int b1 = 1 << a1, b2 = 1 << a2;
    T d1 = p.maxd, d2 = p.maxd;
    int e = -1;
    if (u < b.lo_at(a1)) e = edge_of(fix, fix | b2);
    else if (u > b.hi_at(a1)) e = edge_of(fix | b1, fix | b1 | b2);
    if (e >= 0) d1 = seg_seg<T, false>(A, B, b.corner(e & 7), b.corner(e >> 3), p, nullptr);
    e = -1;
    if (d1 > 0 && v < b.lo_at(a2)) e = edge_of(fix, fix | b1);
    else if (d1 > 0 && v > b.hi_at(a2)) e = edge_of(fix | b2, fix | b1 | b2);
    if (e >= 0) d2 = seg_seg<T, false>(A, B, b.corner(e & 7), b.corner(e >> 3), p, nullptr);
    return smin(d1, d2);
}

// collisionCapsuleToRectangle, CollisionAndDistance.cpp:46-111, general radius
template <typename T>
__host__ __device__ inline bool rect_collision(const Vec3<T>& A, const Vec3<T>& B, T radius, const Box<T>& b,
                                               int face, const Params<T>& p) {
    const int c = face % 3;
    const int o1 = (c == 0) ? 1 : 0, o2 = (c == 2) ? 1 : 2;
    const T v = b.f(face);
    // rec = (min1, min2, max1, max2)
    const T r0 = b.lo_at(o1), r1 = b.lo_at(o2), r2 = b.hi_at(o1), r3 = b.hi_at(o2);
    const T Au = A[o1], Av = A[o2], Bu = B[o1], Bv = B[o2];
    const T den = B[c] - A[c];
    const T t = (den != 0) ? (v - A[c]) / den : p.maxd;
    const T M0 = Au + t * (Bu - Au), M1 = Av + t * (Bv - Av);
    // |A - A_proj| and |B - B_proj|, A_proj = A with coordinate c set to v
    Vec3<T> Ap = A, Bp = B;
    if (c == 0) { Ap.x = v; Bp.x = v; } else if (c == 1) { Ap.y = v; Bp.y = v; } else { Ap.z = v; Bp.z = v; }
    const T nA = vnorm(A - Ap), nB = vnorm(B - Bp);
    if (t > 0 && t < 1) {
        if (M0 > r0 - radius && M0 < r2 + radius && M1 > r1 - radius && M1 < r3 + radius) {
            // the corner tests keep the reference's index pattern (lines 69-72)
            const T c00 = M0 - r0, c01 = M1 - r2, c11 = M1 - r3, c20 = M0 - r1;
            if ((M0 > r0 - radius && M0 < r2 + radius && M1 > r1 && M1 < r3) ||
                (M1 > r1 - radius && M1 < r3 + radius && M0 > r0 && M0 < r2) ||
                (M0 < r0 && M1 < r2 && vsqrt(c00 * c00 + c01 * c01) < radius) ||
                (M0 < r0 && M1 > r3 && vsqrt(c00 * c00 + c11 * c11) < radius) ||
                (M0 > r1 && M1 < r2 && vsqrt(c20 * c20 + c01 * c01) < radius) ||
                (M0 > r1 && M1 > r3 && vsqrt(c20 * c20 + c11 * c11) < radius))
                return true;
        }
    } else if (smin(nA, nB) > radius) {
        return false;
    }
    if (radius > 0) {
        const bool a_in = Au > r0 && Au < r2 && Av > r1 && Av < r3;
        const bool b_in = Bu > r0 && Bu < r2 && Bv > r1 && Bv < r3;
        if (a_in) {
            if (b_in) return smin(nA, nB) < radius;
            return smin(check_cases(A, B, b, face, Bu, Bv, p), nA) < radius;
        }
        if (b_in) return smin(check_cases(A, B, b, face, Au, Av, p), nB) < radius;
        if (check_cases(A, B, b, face, Au, Av, p) < radius) return true;
        return check_cases(A, B, b, face, Bu, Bv, p) < radius;
    }
    return false;
}

// collisionCapsuleToBox, CollisionAndDistance.cpp:4-42.  r_new is the
// inflated radius radius*sqrt(3)/3 evaluated in double and stored in T
// (line 7); it depends only on the link, so callers pass it in.  The
// reference tests "an end point inside the box grown by r_new" (true) before
// "both end points beyond one face of the box grown by radius" (false); here
// the second runs first, since it settles most pairs.  The order changes no
// result for 0 <= r_new <= radius: a point below fl(lo - radius) is below
// fl(lo - r_new) (rounding is monotone), so it is not inside, and likewise
// above fl(hi + radius); the coordinates that decide the first test are
// numbers, whatever the others hold.
template <typename T>
__host__ __device__ inline bool collision_box(const Vec3<T>& A, const Vec3<T>& B, T radius, T r_new, const Box<T>& b,
                                              const Params<T>& p) {
    if ((A.x < b.lo[0] - radius && B.x < b.lo[0] - radius) || (A.x > b.hi[0] + radius && B.x > b.hi[0] + radius) ||
        (A.y < b.lo[1] - radius && B.y < b.lo[1] - radius) || (A.y > b.hi[1] + radius && B.y > b.hi[1] + radius) ||
        (A.z < b.lo[2] - radius && B.z < b.lo[2] - radius) || (A.z > b.hi[2] + radius && B.z > b.hi[2] + radius))
        return false;
    if ((A.x > b.lo[0] - r_new && A.y > b.lo[1] - r_new && A.z > b.lo[2] - r_new &&
         A.x < b.hi[0] + r_new && A.y < b.hi[1] + r_new && A.z < b.hi[2] + r_new) ||
        (B.x > b.lo[0] - r_new && B.y > b.lo[1] - r_new && B.z > b.lo[2] - r_new &&
         B.x < b.hi[0] + r_new && B.y < b.hi[1] + r_new && B.z < b.hi[2] + r_new))
        return true;
    // one face per axis, chosen by A's position, tested in x, y, z order
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        int face = -1;
        if (A[k] < b.lo[k]) face = k;
        else if (A[k] > b.hi[k]) face = k + 3;
        if (face >= 0 && rect_collision(A, B, radius, b, face, p)) return true;
    }
    return false;
}

// ---------------------------------------------------------------- underestimation
// One (link, obstacle) term of computeDistanceUnderestimation,
// RealVectorSpace.cpp:420-427: min over the link's end points of the signed
// distance to the plane through O with normal normalize(R - O), minus radius.
// Eigen's normalized() divides by the norm and leaves a zero vector as is.
template <typename T>
__host__ __device__ inline T plane_term(const Vec3<T>& P0, const Vec3<T>& P1, const Vec3<T>& R, const Vec3<T>& O,
                                        T radius) {
    const Vec3<T> RO = R - O;
    const T z = sqnorm(RO);
    Vec3<T> n = RO;
    if (z > 0) {
        const T s = vsqrt(z);
        n = Vec3<T>{RO.x / s, RO.y / s, RO.z / s};
    }
    return smin(dot(P0 - O, n), dot(P1 - O, n)) - radius;
}

// The same term after the plane has moved towards the robot for delta_t at
// the obstacle's maximal speed (AbstractTrajectory::
// computeDistanceUnderestimation, AbstractTrajectory.cpp:295-306): both
// points shift by normalize(R - O) * max_vel * delta_t, then the plane is
// rebuilt from the shifted points.
template <typename T>
__host__ __device__ inline T plane_term_moving(const Vec3<T>& P0, const Vec3<T>& P1, const Vec3<T>& R, const Vec3<T>& O,
                                               T radius, T max_vel, T delta_t) {
    const Vec3<T> RO = R - O;
    const T z = sqnorm(RO);
    Vec3<T> n = RO;
    if (z > 0) {
        const T s = vsqrt(z);
        n = Vec3<T>{RO.x / s, RO.y / s, RO.z / s};
    }
    const Vec3<T> dl{n.x * max_vel * delta_t, n.y * max_vel * delta_t, n.z * max_vel * delta_t};
    const Vec3<T> R2{R.x + dl.x, R.y + dl.y, R.z + dl.z};
    const Vec3<T> O2{O.x + dl.x, O.y + dl.y, O.z + dl.z};
    return plane_term(P0, P1, R2, O2, radius);
}

}  // namespace rvs
