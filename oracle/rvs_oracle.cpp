// CPU oracle for RPMPLv2's RealVectorSpace capsule path.  TEST INFRASTRUCTURE:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
// always as the checker.  The product library never links or calls it.
//
// This is an independent restatement of the reference's algorithms (no source
// copied).  Every function cites the reference lines it follows.  It is
// instantiated twice: T=float ("ref32", the reference's own arithmetic:
// double FK, float skeleton, float geometry in Eigen's operation order) and
// T=double ("ref64", the same algorithm carried out in double).
//
// Operation-order conventions that the restatement keeps (they decide the
// float rounding and therefore exact verdict parity):
//   * Eigen fixed-size 3-vector reductions (dot, squaredNorm, norm) sum as
//     e0 + (e1 + e2); 2-vectors as e0 + e1                     [Eigen 3.4, 3P]
//   * scalar*vector and vector+-vector are per component, no FMA contraction
//     (the reference Release build targets baseline x86-64: no FMA unit)
//   * std::min(a, b) returns (b < a) ? b : a, so a NaN second argument is
//     ignored and a NaN first argument is kept
//   * KDL frame products sum left to right: (a0*b0 + a1*b3) + a2*b6 [KDL, 3P]
//
// Parity status (what pins each part, what is unpinned) is listed in
// rvs_oracle.h and oracle/README.md.

#include "rvs_oracle.h"

#include <cmath>
#include <cstring>
#include <cstdint>
#include <mutex>
#include <thread>
#include <vector>
#include <algorithm>
#include <limits>

namespace {

// ---------------------------------------------------------------- helpers
template <typename T> inline T dot3(const T* a, const T* b) { return a[0] * b[0] + (a[1] * b[1] + a[2] * b[2]); }
template <typename T> inline T sqn3(const T* a) { return dot3(a, a); }
template <typename T> inline T norm3(const T* a) { return std::sqrt(sqn3(a)); }
template <typename T> inline void sub3(const T* a, const T* b, T* o) { o[0] = a[0] - b[0]; o[1] = a[1] - b[1]; o[2] = a[2] - b[2]; }
template <typename T> inline void cpy3(const T* a, T* o) { o[0] = a[0]; o[1] = a[1]; o[2] = a[2]; }
template <typename T> inline void set3(T* o, T x, T y, T z) { o[0] = x; o[1] = y; o[2] = z; }
// a + s * (b - a), per component (Eigen: A + t * (B - A))
template <typename T> inline void lerp3(const T* a, const T* b, T s, T* o) {
    for (int k = 0; k < 3; ++k) o[k] = a[k] + s * (b[k] - a[k]);
}
template <typename T> inline T stdmin(T a, T b) { return (b < a) ? b : a; }
template <typename T> inline T stdmax(T a, T b) { return (a < b) ? b : a; }
template <typename T> inline void np_set(T* np, const T* p0, const T* p1) {
    cpy3(p0, np); cpy3(p1, np + 3);
}
template <typename T> inline void np_nan(T* np) {
    for (int k = 0; k < 6; ++k) np[k] = std::numeric_limits<T>::quiet_NaN();
}

template <typename T> struct P {
    T eq;    // EQUALITY_THRESHOLD
    T maxd;  // MAX_DISTANCE
    explicit P(const orc_params* p) : eq((T)p->eq_threshold), maxd((T)p->max_distance) {}
};

// get3DPoint: insert 'value' at coordinate 'coord' of a 2-D point
// (CollisionAndDistance.cpp:150-155)
template <typename T> inline void get3d(T u, T v, T value, int coord, T* o) {
    if (coord == 0) set3(o, value, u, v);
    else if (coord == 1) set3(o, u, value, v);
    else set3(o, u, v, value);
}

// ------------------------------------------------ distanceLineSegToLineSeg
// CollisionAndDistance.cpp:217-320.  Returns the distance; np receives
// (point on AB, point on CD); np_valid = 0 when the reference returns a null
// nearest-point pointer (interior solution closer than EQUALITY_THRESHOLD) or
// when no candidate ever improved on MAX_DISTANCE (uninitialised in the
// reference).
template <typename T>
T seg_seg(const T* A, const T* B, const T* C, const T* D, const P<T>& pr, T* np, int* np_valid) {
    T d_c = pr.maxd;
    *np_valid = 0;
    T BA[3], DC[3], CD[3], AC[3], AB[3];
    sub3(B, A, BA); sub3(D, C, DC); sub3(C, D, CD); sub3(A, C, AC); sub3(A, B, AB);
    T alpha1 = sqn3(BA);
    T alpha2 = dot3(BA, DC);
    T beta1 = dot3(CD, BA);
    T beta2 = dot3(CD, DC);
    T gamma1 = dot3(AC, AB);
    T gamma2 = dot3(AC, CD);
    T s = (alpha1 * gamma2 - alpha2 * gamma1) / (alpha1 * beta2 - alpha2 * beta1);
    T t = (gamma1 - beta1 * s) / alpha1;

    if (t > 0 && t < 1 && s > 0 && s < 1) {
        T P1[3], P2[3], d[3];
        lerp3(A, B, t, P1);
        lerp3(C, D, s, P2);
        sub3(P2, P1, d);
        d_c = norm3(d);
        if (d_c < pr.eq) { np_nan(np); *np_valid = 0; return T(0); }
        np_set(np, P1, P2); *np_valid = 1;
    } else {
        T alpha3 = sqn3(CD);
        T AD[3], BC[3];
        sub3(A, D, AD); sub3(B, C, BC);
        T opt[4] = { dot3(AC, AB) / alpha1, dot3(AD, AB) / alpha1,
                     dot3(AC, DC) / alpha3, dot3(BC, DC) / alpha3 };
        for (int i = 0; i < 4; ++i) {
            T p0[3], p1[3];
            if (opt[i] < 0) {
                if (i == 0 || i == 2) { cpy3(A, p0); cpy3(C, p1); }
                else if (i == 1) { cpy3(A, p0); cpy3(D, p1); }
                else { cpy3(B, p0); cpy3(C, p1); }
            } else if (opt[i] > 1) {
                if (i == 1 || i == 3) { cpy3(B, p0); cpy3(D, p1); }
                else if (i == 0) { cpy3(B, p0); cpy3(C, p1); }
                else { cpy3(A, p0); cpy3(D, p1); }
            } else {
                if (i == 0) { lerp3(A, B, opt[i], p0); cpy3(C, p1); }
                else if (i == 1) { lerp3(A, B, opt[i], p0); cpy3(D, p1); }
                else if (i == 2) { cpy3(A, p0); lerp3(C, D, opt[i], p1); }
                else { cpy3(B, p0); lerp3(C, D, opt[i], p1); }
            }
            T dd[3];
            sub3(p1, p0, dd);
            T d_tmp = norm3(dd);
            if (d_tmp < d_c) { d_c = d_tmp; np_set(np, p0, p1); *np_valid = 1; }
        }
        if (!*np_valid) np_nan(np);
    }
    return d_c;
}

// ------------------------------------------------ distanceLineSegToPoint
// CollisionAndDistance.cpp:323-342
template <typename T>
T seg_pt(const T* A, const T* B, const T* C, const P<T>& pr, T* np, int* np_valid) {
    T CA[3], BA[3], Pn[3], d[3];
    sub3(C, A, CA); sub3(B, A, BA);
    T t_opt = dot3(CA, BA) / sqn3(BA);
    if (t_opt < 0) cpy3(A, Pn);
    else if (t_opt > 1) cpy3(B, Pn);
    else lerp3(A, B, t_opt, Pn);
    sub3(C, Pn, d);
    T d_c = norm3(d);
    if (d_c < pr.eq) { np_nan(np); *np_valid = 0; return T(0); }
    np_set(np, Pn, C); *np_valid = 1;
    return d_c;
}

// ------------------------------------------------ collision tree
// checkCases: CollisionAndDistance.cpp:113-148
template <typename T>
T check_cases(const T* A, const T* B, const T* rec, const T* pt, T obs_coord, int coord, const P<T>& pr) {
    T d1 = pr.maxd, d2 = pr.maxd;
    T C[3], D[3], np[6];
    int v;
    if (pt[0] < rec[0]) {
        get3d(rec[0], rec[1], obs_coord, coord, C); get3d(rec[0], rec[3], obs_coord, coord, D);
        d1 = seg_seg(A, B, C, D, pr, np, &v);
    } else if (pt[0] > rec[2]) {
        get3d(rec[2], rec[1], obs_coord, coord, C); get3d(rec[2], rec[3], obs_coord, coord, D);
        d1 = seg_seg(A, B, C, D, pr, np, &v);
    }
    if (d1 > 0 && pt[1] < rec[1]) {
        get3d(rec[0], rec[1], obs_coord, coord, C); get3d(rec[2], rec[1], obs_coord, coord, D);
        d2 = seg_seg(A, B, C, D, pr, np, &v);
    } else if (d1 > 0 && pt[1] > rec[3]) {
        get3d(rec[0], rec[3], obs_coord, coord, C); get3d(rec[2], rec[3], obs_coord, coord, D);
        d2 = seg_seg(A, B, C, D, pr, np, &v);
    }
    return stdmin(d1, d2);
}

// collisionCapsuleToRectangle: CollisionAndDistance.cpp:46-111
// 'coord' in 0..5 selects the face {xmin,ymin,zmin,xmax,ymax,zmax}.
template <typename T>
bool collision_rect(const T* A, const T* B, T radius, const T* obs, int coord, const P<T>& pr) {
    T obs_coord = obs[coord];
    if (coord > 2) coord -= 3;
    // rec = (min1, min2, max1, max2) over the two coordinates other than 'coord'
    T rec[4];
    int o1 = (coord == 0) ? 1 : 0;
    int o2 = (coord == 2) ? 1 : 2;
    rec[0] = obs[o1]; rec[1] = obs[o2]; rec[2] = obs[o1 + 3]; rec[3] = obs[o2 + 3];
    T A_rec[2] = { A[o1], A[o2] };
    T B_rec[2] = { B[o1], B[o2] };
    T den = B[coord] - A[coord];
    T t = (den != 0) ? (obs_coord - A[coord]) / den : pr.maxd;
    T M[2] = { A_rec[0] + t * (B_rec[0] - A_rec[0]), A_rec[1] + t * (B_rec[1] - A_rec[1]) };
    T A_proj[3], B_proj[3], dA[3], dB[3];
    get3d(A_rec[0], A_rec[1], obs_coord, coord, A_proj);
    get3d(B_rec[0], B_rec[1], obs_coord, coord, B_proj);
    sub3(A, A_proj, dA); sub3(B, B_proj, dB);

    if (t > 0 && t < 1) {
        if (M[0] > rec[0] - radius && M[0] < rec[2] + radius &&
            M[1] > rec[1] - radius && M[1] < rec[3] + radius) {
            // corner tests keep the reference's index pattern (rec(2) used as
            // the second coordinate, rec(1) as an x bound): lines 69-72
            T c0[2] = { M[0] - rec[0], M[1] - rec[2] };
            T c1[2] = { M[0] - rec[0], M[1] - rec[3] };
            T c2[2] = { M[0] - rec[1], M[1] - rec[2] };
            T c3[2] = { M[0] - rec[1], M[1] - rec[3] };
            if ((M[0] > rec[0] - radius && M[0] < rec[2] + radius && M[1] > rec[1] && M[1] < rec[3]) ||
                (M[1] > rec[1] - radius && M[1] < rec[3] + radius && M[0] > rec[0] && M[0] < rec[2]) ||
                (M[0] < rec[0] && M[1] < rec[2] && std::sqrt(c0[0] * c0[0] + c0[1] * c0[1]) < radius) ||
                (M[0] < rec[0] && M[1] > rec[3] && std::sqrt(c1[0] * c1[0] + c1[1] * c1[1]) < radius) ||
                (M[0] > rec[1] && M[1] < rec[2] && std::sqrt(c2[0] * c2[0] + c2[1] * c2[1]) < radius) ||
                (M[0] > rec[1] && M[1] > rec[3] && std::sqrt(c3[0] * c3[0] + c3[1] * c3[1]) < radius))
                return true;
        }
    } else if (stdmin(norm3(dA), norm3(dB)) > radius) {
        return false;
    }

    if (radius > 0) {
        bool a_in = A_rec[0] > rec[0] && A_rec[0] < rec[2] && A_rec[1] > rec[1] && A_rec[1] < rec[3];
        bool b_in = B_rec[0] > rec[0] && B_rec[0] < rec[2] && B_rec[1] > rec[1] && B_rec[1] < rec[3];
        if (a_in) {
            if (b_in) {
                if (stdmin(norm3(dA), norm3(dB)) < radius) return true;
            } else {
                if (stdmin(check_cases(A, B, rec, B_rec, obs_coord, coord, pr), norm3(dA)) < radius) return true;
            }
        } else if (b_in) {
            if (stdmin(check_cases(A, B, rec, A_rec, obs_coord, coord, pr), norm3(dB)) < radius) return true;
        } else {
            if (check_cases(A, B, rec, A_rec, obs_coord, coord, pr) < radius) return true;
            if (check_cases(A, B, rec, B_rec, obs_coord, coord, pr) < radius) return true;
        }
    }
    return false;
}

// r_new = radius * sqrt(3) / 3 evaluated in double, stored in T
// (CollisionAndDistance.cpp:7)
template <typename T> inline T inflate_radius(T radius) { return (T)((double)radius * std::sqrt(3.0) / 3.0); }

// collisionCapsuleToBox: CollisionAndDistance.cpp:4-42
template <typename T>
bool collision_box(const T* A, const T* B, T radius, const T* obs, const P<T>& pr) {
    bool collision = false;
    T r_new = inflate_radius(radius);
    if ((A[0] > obs[0] - r_new && A[1] > obs[1] - r_new && A[2] > obs[2] - r_new &&
         A[0] < obs[3] + r_new && A[1] < obs[4] + r_new && A[2] < obs[5] + r_new) ||
        (B[0] > obs[0] - r_new && B[1] > obs[1] - r_new && B[2] > obs[2] - r_new &&
         B[0] < obs[3] + r_new && B[1] < obs[4] + r_new && B[2] < obs[5] + r_new))
        return true;
    else if ((A[0] < obs[0] - radius && B[0] < obs[0] - radius) || (A[0] > obs[3] + radius && B[0] > obs[3] + radius) ||
             (A[1] < obs[1] - radius && B[1] < obs[1] - radius) || (A[1] > obs[4] + radius && B[1] > obs[4] + radius) ||
             (A[2] < obs[2] - radius && B[2] < obs[2] - radius) || (A[2] > obs[5] + radius && B[2] > obs[5] + radius))
        return false;

    if (A[0] < obs[0]) collision = collision_rect(A, B, radius, obs, 0, pr);
    else if (A[0] > obs[3]) collision = collision_rect(A, B, radius, obs, 3, pr);
    if (!collision) {
        if (A[1] < obs[1]) collision = collision_rect(A, B, radius, obs, 1, pr);
        else if (A[1] > obs[4]) collision = collision_rect(A, B, radius, obs, 4, pr);
        if (!collision) {
            if (A[2] < obs[2]) collision = collision_rect(A, B, radius, obs, 2, pr);
            else if (A[2] > obs[5]) collision = collision_rect(A, B, radius, obs, 5, pr);
        }
    }
    return collision;
}

// ------------------------------------------------ class CapsuleToBox
// CollisionAndDistance.cpp:395-755.  'path' records which branch produced
// the result (used by the tests to cover every region case):
//   0 inside, 1 two projections same face, 2 one projection + rectangle hit,
//   3 one projection + edges, 4 corner point, 5 one edge, 6 four edges,
//   7 six-face hit, 8 twelve edges, 9 missing-face (reference dereferences null)
template <typename T>
struct CapsuleToBox {
    const T* A; const T* B; T radius; const T* obs;
    const P<T>& pr;
    T AB[2][3];
    T d_c;
    T np[6]; int np_valid;
    int proj[6][2];
    T dist[2];
    int path;

    CapsuleToBox(const T* A_, const T* B_, T r_, const T* obs_, const P<T>& pr_)
        : A(A_), B(B_), radius(r_), obs(obs_), pr(pr_) {
        cpy3(A, AB[0]); cpy3(B, AB[1]);
        d_c = pr.maxd;
        np_nan(np); np_valid = 0;   // the reference allocates a 3x2 matrix uninitialised
        std::memset(proj, 0, sizeof(proj));
        dist[0] = dist[1] = pr.maxd;
        path = -1;
    }

    // projectionLineSegOnSide: lines 450-474
    void projection(int min1, int min2, int min3, int max1, int max2, int max3) {
        for (int i = 0; i < 2; ++i) {
            const T* X = AB[i];
            if (X[min1] >= obs[min1] && X[min1] <= obs[max1] && X[min2] >= obs[min2] && X[min2] <= obs[max2]) {
                if (X[min3] > obs[min3] && X[min3] < obs[max3]) { d_c = 0; return; }
                else if (X[min3] <= obs[min3]) { proj[min3][i] = 1; dist[i] = obs[min3] - X[min3]; }
                else if (X[min3] >= obs[max3]) { proj[max3][i] = 1; dist[i] = X[min3] - obs[max3]; }
            }
        }
    }

    // distanceToMoreLineSegments: lines 576-595.  segs: list of 3-D points, pairs
    void more_segments(const T (*segs)[3], int nseg) {
        for (int k = 0; k < nseg; k += 2) {
            T npt[6]; int v;
            T d_tmp = seg_seg(A, B, segs[k], segs[k + 1], pr, npt, &v);
            if (d_tmp <= 0) { d_c = 0; return; }
            else if (d_tmp < d_c) {
                d_c = d_tmp;
                // the reference copies *nearest_pts_temp; a null pointer cannot
                // reach here because the null case returns d = 0 above
                std::memcpy(np, npt, sizeof(np)); np_valid = v;
            }
        }
    }

    // getLineSegments: lines 542-574 (returns number of points written)
    int line_segments(const T* pt, T min1, T min2, T max1, T max2, T cv, int coord, T (*segs)[3]) {
        int num = 0;
        if (pt[0] < min1) { get3d(min1, min2, cv, coord, segs[0]); get3d(min1, max2, cv, coord, segs[1]); num += 2; }
        else if (pt[0] > max1) { get3d(max1, min2, cv, coord, segs[0]); get3d(max1, max2, cv, coord, segs[1]); num += 2; }
        if (pt[1] < min2) { get3d(min1, min2, cv, coord, segs[num]); get3d(max1, min2, cv, coord, segs[num + 1]); return num + 2; }
        else if (pt[1] > max2) { get3d(min1, max2, cv, coord, segs[num]); get3d(max1, max2, cv, coord, segs[num + 1]); return num + 2; }
        // when no branch applies the reference keeps a 3x2 uninitialised
        // matrix; it is only reachable with a point lying exactly on a face
        // plane.  The restatement reports it through path = 9.
        return (num == 0) ? -1 : num;
    }

    // checkEdges: lines 476-540
    void check_edges(const T* point, int idx) {
        static const int order[6] = { 0, 3, 1, 4, 2, 5 };
        int face = -1;
        for (int f : order) if (proj[f][idx]) { face = f; break; }
        if (face < 0) { path = 9; d_c = std::numeric_limits<T>::quiet_NaN(); return; }
        if (collision_rect(A, B, T(0), obs, face, pr)) { d_c = 0; path = 2; return; }
        int c = face % 3;
        T segs[4][3];
        T pt2[2];
        int n;
        if (c == 0) { pt2[0] = point[1]; pt2[1] = point[2]; n = line_segments(pt2, obs[1], obs[2], obs[4], obs[5], obs[face], 0, segs); }
        else if (c == 1) { pt2[0] = point[0]; pt2[1] = point[2]; n = line_segments(pt2, obs[0], obs[2], obs[3], obs[5], obs[face], 1, segs); }
        else { pt2[0] = point[0]; pt2[1] = point[1]; n = line_segments(pt2, obs[0], obs[1], obs[3], obs[4], obs[face], 2, segs); }
        if (n < 0) { path = 9; d_c = std::numeric_limits<T>::quiet_NaN(); return; }
        path = 3;
        more_segments(segs, n);
    }

    void corner(T x, T y, T z) {
        T C[3] = { x, y, z };
        d_c = seg_pt(A, B, C, pr, np, &np_valid);
        path = 4;
    }
    void edge(T x0, T y0, T z0, T x1, T y1, T z1) {
        T C[3] = { x0, y0, z0 }, D[3] = { x1, y1, z1 };
        d_c = seg_seg(A, B, C, D, pr, np, &np_valid);
        path = 5;
    }
    // four edges of a face ring given as 4 corners in order c0-c1-c2-c3-c0
    void ring(const T* c0, const T* c1, const T* c2, const T* c3) {
        T segs[8][3];
        cpy3(c0, segs[0]); cpy3(c1, segs[1]); cpy3(c1, segs[2]); cpy3(c2, segs[3]);
        cpy3(c2, segs[4]); cpy3(c3, segs[5]); cpy3(c3, segs[6]); cpy3(c0, segs[7]);
        path = 6;
        more_segments(segs, 8);
    }

    // checkOtherCases: lines 597-755
    void other_cases() {
        const T x0 = obs[0], y0 = obs[1], z0 = obs[2], x1 = obs[3], y1 = obs[4], z1 = obs[5];
        bool zlo = A[2] < z0 && B[2] < z0, zhi = A[2] > z1 && B[2] > z1;
        if (A[0] < x0 && B[0] < x0) {
            if (A[1] < y0 && B[1] < y0) {
                if (zlo) corner(x0, y0, z0); else if (zhi) corner(x0, y0, z1); else edge(x0, y0, z0, x0, y0, z1);
            } else if (A[1] > y1 && B[1] > y1) {
                if (zlo) corner(x0, y1, z0); else if (zhi) corner(x0, y1, z1); else edge(x0, y1, z0, x0, y1, z1);
            } else {
                if (zlo) edge(x0, y0, z0, x0, y1, z0);
                else if (zhi) edge(x0, y0, z1, x0, y1, z1);
                else { T a[3] = { x0, y0, z0 }, b[3] = { x0, y1, z0 }, c[3] = { x0, y1, z1 }, d[3] = { x0, y0, z1 }; ring(a, b, c, d); }
            }
        } else if (A[0] > x1 && B[0] > x1) {
            if (A[1] < y0 && B[1] < y0) {
                if (zlo) corner(x1, y0, z0); else if (zhi) corner(x1, y0, z1); else edge(x1, y0, z0, x1, y0, z1);
            } else if (A[1] > y1 && B[1] > y1) {
                if (zlo) corner(x1, y1, z0); else if (zhi) corner(x1, y1, z1); else edge(x1, y1, z0, x1, y1, z1);
            } else {
                if (zlo) edge(x1, y0, z0, x1, y1, z0);
                else if (zhi) edge(x1, y0, z1, x1, y1, z1);
                else { T a[3] = { x1, y0, z0 }, b[3] = { x1, y1, z0 }, c[3] = { x1, y1, z1 }, d[3] = { x1, y0, z1 }; ring(a, b, c, d); }
            }
        } else {
            if (A[1] < y0 && B[1] < y0) {
                if (zlo) edge(x0, y0, z0, x1, y0, z0);
                else if (zhi) edge(x0, y0, z1, x1, y0, z1);
                else { T a[3] = { x0, y0, z0 }, b[3] = { x1, y0, z0 }, c[3] = { x1, y0, z1 }, d[3] = { x0, y0, z1 }; ring(a, b, c, d); }
            } else if (A[1] > y1 && B[1] > y1) {
                if (zlo) edge(x0, y1, z0, x1, y1, z0);
                else if (zhi) edge(x0, y1, z1, x1, y1, z1);
                else { T a[3] = { x0, y1, z0 }, b[3] = { x1, y1, z0 }, c[3] = { x1, y1, z1 }, d[3] = { x0, y1, z1 }; ring(a, b, c, d); }
            } else {
                if (zlo) { T a[3] = { x0, y0, z0 }, b[3] = { x1, y0, z0 }, c[3] = { x1, y1, z0 }, d[3] = { x0, y1, z0 }; ring(a, b, c, d); }
                else if (zhi) { T a[3] = { x0, y0, z1 }, b[3] = { x1, y0, z1 }, c[3] = { x1, y1, z1 }, d[3] = { x0, y1, z1 }; ring(a, b, c, d); }
                else {
                    for (int kk = 0; kk < 6; ++kk)
                        if (collision_rect(A, B, T(0), obs, kk, pr)) { d_c = 0; path = 7; return; }
                    // bottom ring, top ring, then the four vertical edges
                    T segs[24][3];
                    const T zs[2] = { z0, z1 };
                    for (int h = 0; h < 2; ++h) {
                        T z = zs[h];
                        T a[3] = { x0, y0, z }, b[3] = { x1, y0, z }, c[3] = { x1, y1, z }, d[3] = { x0, y1, z };
                        int o = 8 * h;
                        cpy3(a, segs[o + 0]); cpy3(b, segs[o + 1]); cpy3(b, segs[o + 2]); cpy3(c, segs[o + 3]);
                        cpy3(c, segs[o + 4]); cpy3(d, segs[o + 5]); cpy3(d, segs[o + 6]); cpy3(a, segs[o + 7]);
                    }
                    const T xy[4][2] = { { x0, y0 }, { x1, y0 }, { x1, y1 }, { x0, y1 } };
                    for (int v = 0; v < 4; ++v) {
                        set3(segs[16 + 2 * v], xy[v][0], xy[v][1], z0);
                        set3(segs[17 + 2 * v], xy[v][0], xy[v][1], z1);
                    }
                    path = 8;
                    more_segments(segs, 24);
                }
            }
        }
    }

    // compute: lines 409-448
    void compute() {
        projection(1, 2, 0, 4, 5, 3);
        projection(0, 2, 1, 3, 5, 4);
        projection(0, 1, 2, 3, 4, 5);
        if (d_c == 0) { np_nan(np); np_valid = 0; path = 0; return; }
        int num_proj = 0;
        for (int f = 0; f < 6; ++f) num_proj = std::max(num_proj, proj[f][0] + proj[f][1]);
        if (num_proj > 0) {
            int idx = (dist[0] < dist[1]) ? 0 : 1;
            d_c = stdmin(dist[0], dist[1]);                 // Vector2f::minCoeff
            int idx_coord = 0;                              // first maximal entry
            for (int f = 1; f < 6; ++f) if (proj[f][idx] > proj[idx_coord][idx]) idx_coord = f;
            const T* X = AB[idx];
            cpy3(X, np);
            cpy3(X, np + 3);
            np[3 + (idx_coord % 3)] = obs[idx_coord];
            np_valid = 1;
            path = 1;
            if (num_proj == 1) {
                if (idx == 0) check_edges(B, idx);
                else check_edges(A, idx);
            }
        } else {
            other_cases();
        }
        d_c -= radius;
    }
};

template <typename T>
T capsule_box(const T* A, const T* B, T r, const T* obs, const P<T>& pr, T* np, int* np_valid, int* path) {
    CapsuleToBox<T> cb(A, B, r, obs, pr);
    cb.compute();
    std::memcpy(np, cb.np, sizeof(cb.np));
    *np_valid = cb.np_valid;
    if (path) *path = cb.path;
    return cb.d_c;
}

// ------------------------------------------------ KDL-style FK (double)
struct Frame { double M[9]; double p[3]; };

Frame frame_mul(const Frame& a, const Frame& b) {   // KDL operator*(Frame, Frame)
    Frame r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r.M[3 * i + j] = a.M[3 * i + 0] * b.M[0 + j] + a.M[3 * i + 1] * b.M[3 + j] + a.M[3 * i + 2] * b.M[6 + j];
    for (int i = 0; i < 3; ++i)
        r.p[i] = a.M[3 * i + 0] * b.p[0] + a.M[3 * i + 1] * b.p[1] + a.M[3 * i + 2] * b.p[2] + a.p[i];
    return r;
}

// urdf::Rotation::setFromRPY + normalize, then KDL::Rotation::Quaternion [3P]
void rpy_to_rot(const double rpy[3], double M[9]) {
    double phi = rpy[0] / 2.0, the = rpy[1] / 2.0, psi = rpy[2] / 2.0;
    double x = std::sin(phi) * std::cos(the) * std::cos(psi) - std::cos(phi) * std::sin(the) * std::sin(psi);
    double y = std::cos(phi) * std::sin(the) * std::cos(psi) + std::sin(phi) * std::cos(the) * std::sin(psi);
    double z = std::cos(phi) * std::cos(the) * std::sin(psi) - std::sin(phi) * std::sin(the) * std::cos(psi);
    double w = std::cos(phi) * std::cos(the) * std::cos(psi) + std::sin(phi) * std::sin(the) * std::sin(psi);
    double s = std::sqrt(x * x + y * y + z * z + w * w);
    x /= s; y /= s; z /= s; w /= s;
    double x2 = x * x, y2 = y * y, z2 = z * z, w2 = w * w;
    double R[9] = { w2 + x2 - y2 - z2, 2 * x * y - 2 * w * z, 2 * x * z + 2 * w * y,
                    2 * x * y + 2 * w * z, w2 - x2 + y2 - z2, 2 * y * z - 2 * w * x,
                    2 * x * z - 2 * w * y, 2 * y * z + 2 * w * x, w2 - x2 - y2 + z2 };
    std::memcpy(M, R, sizeof(R));
}

// KDL::Vector::Norm (overflow-safe form) [3P]
double kdl_norm(const double v[3]) {
    double a = std::fabs(v[0]), b = std::fabs(v[1]), c;
    auto sq = [](double x) { return x * x; };
    if (a >= b) {
        c = std::fabs(v[2]);
        if (a >= c) {
            if (a == 0) return 0;
            return a * std::sqrt(1 + sq(v[1] / v[0]) + sq(v[2] / v[0]));
        }
        return c * std::sqrt(1 + sq(v[0] / v[2]) + sq(v[1] / v[2]));
    }
    c = std::fabs(v[2]);
    if (b > c) return b * std::sqrt(1 + sq(v[0] / v[1]) + sq(v[2] / v[1]));
    return c * std::sqrt(1 + sq(v[0] / v[2]) + sq(v[1] / v[2]));
}

// KDL::Rotation::Rot2 (Rodrigues with the unit axis 'a') [3P]
void rot2(const double a[3], double ang, double M[9]) {
    double ct = std::cos(ang), st = std::sin(ang), vt = 1 - ct;
    double m_vt_0 = vt * a[0], m_vt_1 = vt * a[1], m_vt_2 = vt * a[2];
    double m_st_0 = a[0] * st, m_st_1 = a[1] * st, m_st_2 = a[2] * st;
    double m_vt_0_1 = m_vt_0 * a[1], m_vt_0_2 = m_vt_0 * a[2], m_vt_1_2 = m_vt_1 * a[2];
    M[0] = ct + m_vt_0 * a[0]; M[1] = -m_st_2 + m_vt_0_1; M[2] = m_st_1 + m_vt_0_2;
    M[3] = m_st_2 + m_vt_0_1;  M[4] = ct + m_vt_1 * a[1]; M[5] = -m_st_0 + m_vt_1_2;
    M[6] = -m_st_1 + m_vt_0_2; M[7] = m_st_0 + m_vt_1_2;  M[8] = ct + m_vt_2 * a[2];
}

struct Seg { int rev; double R[9]; double axis[3]; double o[3]; };

// kdl_parser: segment pose(q) = Joint(origin, R_pj*axis).pose(q) * f_tip with
// f_tip = pose(0)^-1 * F_parent_jnt, which evaluates to Frame(Rot2(a',q)*R_pj, o)
std::vector<Seg> build_segments(const orc_robot* rb) {
    std::vector<Seg> s(rb->num_segments);
    for (int k = 0; k < rb->num_segments; ++k) {
        const orc_joint& j = rb->joints[k];
        rpy_to_rot(j.rpy, s[k].R);
        for (int i = 0; i < 3; ++i) s[k].o[i] = j.xyz[i];
        s[k].rev = (j.type == ORC_JOINT_REVOLUTE);
        double a[3];
        for (int i = 0; i < 3; ++i) a[i] = s[k].R[3 * i] * j.axis[0] + s[k].R[3 * i + 1] * j.axis[1] + s[k].R[3 * i + 2] * j.axis[2];
        double nrm = kdl_norm(a);
        for (int i = 0; i < 3; ++i) s[k].axis[i] = a[i] / nrm;
    }
    return s;
}

// Forward kinematics: frames[k] = (((I * S_0) * S_1) ... * S_k)
// (TreeFkSolverPos_recursive::JntToCart per segment, xArm6.cpp:114-129,
//  Planar2DOF.cpp:89-104)
void fk_frames(const std::vector<Seg>& segs, const float* q, Frame* frames) {
    Frame T;
    std::memset(&T, 0, sizeof(T));
    T.M[0] = T.M[4] = T.M[8] = 1.0;
    int qi = 0;
    for (size_t k = 0; k < segs.size(); ++k) {
        Frame S;
        if (segs[k].rev) {
            double Rq[9];
            rot2(segs[k].axis, (double)q[qi++], Rq);
            Frame J; std::memcpy(J.M, Rq, sizeof(Rq)); J.p[0] = J.p[1] = J.p[2] = 0;
            Frame F; std::memcpy(F.M, segs[k].R, sizeof(F.M)); F.p[0] = F.p[1] = F.p[2] = 0;
            S.p[0] = segs[k].o[0]; S.p[1] = segs[k].o[1]; S.p[2] = segs[k].o[2];
            Frame RS = frame_mul(J, F);
            std::memcpy(S.M, RS.M, sizeof(S.M));
        } else {
            std::memcpy(S.M, segs[k].R, sizeof(S.M));
            S.p[0] = segs[k].o[0]; S.p[1] = segs[k].o[1]; S.p[2] = segs[k].o[2];
        }
        T = frame_mul(T, S);
        frames[k] = T;
    }
}

// Skeleton recipes.  Chain: P_k = frames[k].p (Planar2DOF.cpp:117-130).
// xArm6: xArm6.cpp:209-235 with the gripper shift of xArm6.cpp:130.
// Points are produced in double; ref32 rounds them to float (MatrixXf).
template <typename T>
void skeleton(const orc_robot* rb, const std::vector<Seg>& segs, const float* q, T* P /* (L+1)*3 */) {
    Frame fr[ORC_MAX_SEG];
    fk_frames(segs, q, fr);
    const int L = rb->num_links;
    if (rb->skeleton_kind == ORC_SKEL_CHAIN) {
        for (int k = 0; k <= L; ++k)
            for (int i = 0; i < 3; ++i) P[3 * k + i] = (T)fr[k].p[i];
        return;
    }
    // xArm6
    Frame* f = fr;
    double g = (double)rb->gripper_length;
    // frames_fk.back().p += gripper_length * frames_fk.back().M.UnitZ()
    for (int i = 0; i < 3; ++i) f[5].p[i] = f[5].p[i] + g * f[5].M[3 * i + 2];
    for (int i = 0; i < 3; ++i) P[i] = (T)0;
    for (int i = 0; i < 3; ++i) P[3 + i] = (T)f[1].p[i];
    for (int i = 0; i < 3; ++i) P[6 + i] = (T)f[2].p[i];
    for (int i = 0; i < 3; ++i) P[9 + i] = (T)(f[3].p[i] - f[3].M[3 * i + 2] * 0.25);
    for (int i = 0; i < 3; ++i) P[12 + i] = (T)f[4].p[i];
    for (int i = 0; i < 3; ++i) P[15 + i] = (T)(f[4].p[i] + f[4].M[3 * i + 0] * 0.076);
    for (int i = 0; i < 3; ++i) P[18 + i] = (T)f[5].p[i];
    // skeleton->col(6) -= 0.3 * gripper_length * Vector3f(UnitZ) (line 231):
    // the scalar is formed in double and converted to the vector's type
    T fac = (T)(0.3 * g);
    for (int i = 0; i < 3; ++i) P[18 + i] = P[18 + i] - fac * (T)f[5].M[3 * i + 2];
}

// ------------------------------------------------ RealVectorSpace loops
template <typename T> struct SceneT {
    int M;
    std::vector<T> obs;
    std::vector<uint8_t> ground;
    std::vector<T> tol;
    std::vector<T> obb;   // empty unless oriented boxes
    SceneT(const orc_scene* sc) : M(sc->num_obj), obs(6 * sc->num_obj), ground(sc->num_obj), tol(sc->num_obj) {
        for (int j = 0; j < M; ++j) {
            for (int k = 0; k < 6; ++k) obs[6 * j + k] = (T)sc->obs[6 * j + k];
            ground[j] = sc->is_ground ? sc->is_ground[j] : 0;
            tol[j] = sc->min_dist_tol ? (T)sc->min_dist_tol[j] : (T)1000.0f;
        }
        if (sc->obb) {
            obb.resize(15 * M);
            for (int k = 0; k < 15 * M; ++k) obb[k] = (T)sc->obb[k];
        }
    }
};

// Capsule-OBB distance (build extension, no reference counterpart): both end
// points are expressed in the box frame, local_k = sum_i R[i][k] (X_i - c_i)
// summed left to right, the capsule-AABB tree runs against [-h, h], and the
// nearest points come back to the world frame as c_i + sum_k R[i][k] local_k.
template <typename T>
T capsule_obb(const T* A, const T* B, T r, const T* obb, const P<T>& pr, T* np, int* np_valid) {
    const T* c = obb;
    const T* R = obb + 3;
    const T* h = obb + 12;
    T a[3], b[3];
    const T va[3] = {A[0] - c[0], A[1] - c[1], A[2] - c[2]};
    const T vb[3] = {B[0] - c[0], B[1] - c[1], B[2] - c[2]};
    for (int k = 0; k < 3; ++k) {
        a[k] = R[k] * va[0] + R[3 + k] * va[1] + R[6 + k] * va[2];
        b[k] = R[k] * vb[0] + R[3 + k] * vb[1] + R[6 + k] * vb[2];
    }
    const T box[6] = {-h[0], -h[1], -h[2], h[0], h[1], h[2]};
    T loc[6];
    const T d = capsule_box(a, b, r, box, pr, loc, np_valid, nullptr);
    if (!*np_valid) { np_nan(np); return d; }
    for (int e = 0; e < 2; ++e)
        for (int i = 0; i < 3; ++i)
            np[3 * e + i] = c[i] + (R[3 * i] * loc[3 * e] + R[3 * i + 1] * loc[3 * e + 1] + R[3 * i + 2] * loc[3 * e + 2]);
    return d;
}

// Collision tree against an oriented box, in the box frame.
template <typename T>
bool collision_obb(const T* A, const T* B, T r, const T* obb, const P<T>& pr) {
    const T* c = obb;
    const T* R = obb + 3;
    const T* h = obb + 12;
    T a[3], b[3];
    const T va[3] = {A[0] - c[0], A[1] - c[1], A[2] - c[2]};
    const T vb[3] = {B[0] - c[0], B[1] - c[1], B[2] - c[2]};
    for (int k = 0; k < 3; ++k) {
        a[k] = R[k] * va[0] + R[3 + k] * va[1] + R[6 + k] * va[2];
        b[k] = R[k] * vb[0] + R[3 + k] * vb[1] + R[6 + k] * vb[2];
    }
    const T box[6] = {-h[0], -h[1], -h[2], h[0], h[1], h[2]};
    return collision_box(a, b, r, box, pr);
}

// Broad phase (build optimisation, no reference counterpart): Euclidean gap
// between the segment's bounding box and a box that contains the obstacle
// (the AABB itself, or the tight world AABB of an oriented box).  It is a
// lower bound of every distance the case tree can return for the pair except
// its snaps to 0 / -r, which need the raw distance below EQUALITY_THRESHOLD
// or contact.  A pair with gap > max(prof + r, EQUALITY_THRESHOLD) + margin
// therefore cannot lower the link's running minimum nor end the loop, and
// may be skipped without changing any result (nearest points excepted: they
// are stored for every pair, so the broad phase is off when they are wanted).
template <typename T> inline T seg_box_gap(const T* A, const T* B, const T* lo, const T* hi) {
    T g2 = 0;
    for (int k = 0; k < 3; ++k) {
        const T smin = std::min(A[k], B[k]), smax = std::max(A[k], B[k]);
        const T g = std::max(std::max(lo[k] - smax, smin - hi[k]), T(0));
        g2 += g * g;
    }
    return std::sqrt(g2);
}

template <typename T> void cull_box(const SceneT<T>& sc, int j, T* lo, T* hi) {
    if (sc.obb.empty()) {
        for (int k = 0; k < 3; ++k) { lo[k] = sc.obs[6 * j + k]; hi[k] = sc.obs[6 * j + 3 + k]; }
        return;
    }
    const T* c = &sc.obb[15 * j];
    const T* R = c + 3;
    const T* h = c + 12;
    for (int i = 0; i < 3; ++i) {
        const T e = std::fabs(R[3 * i]) * h[0] + std::fabs(R[3 * i + 1]) * h[1] + std::fabs(R[3 * i + 2]) * h[2];
        lo[i] = c[i] - e;
        hi[i] = c[i] + e;
    }
}

struct CullCount {
    int64_t evaluated = 0, tests = 0;
    bool from_hit = false;   // the first-hit search continues the walk after the hit (else: obstacle order)
};

// The engine's broad-phase walk order (restated): obstacles sorted stably by
// the Morton code, 10 bits per axis over the scene's extent, of their cull
// box centres (in double; non-finite centres get code 0), in groups of 4
// whose union boxes are tested before their members.
constexpr int kGroup = 4;
template <typename T> struct WalkOrder {
    std::vector<int> order;
    std::vector<T> glo, ghi;   // [groups][3]
};
template <typename T> WalkOrder<T> walk_order(const SceneT<T>& sc) {
    const int M = sc.M;
    WalkOrder<T> wo;
    std::vector<T> lo(3 * (size_t)M), hi(3 * (size_t)M);
    std::vector<double> cen(3 * (size_t)M);
    double blo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, bhi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
    for (int j = 0; j < M; ++j) {
        cull_box(sc, j, &lo[3 * j], &hi[3 * j]);
        for (int k = 0; k < 3; ++k) {
            const double c = 0.5 * ((double)lo[3 * j + k] + (double)hi[3 * j + k]);
            cen[3 * j + k] = c;
            if (std::isfinite(c)) { blo[k] = std::min(blo[k], c); bhi[k] = std::max(bhi[k], c); }
        }
    }
    std::vector<uint64_t> key(M);
    wo.order.resize(M);
    for (int j = 0; j < M; ++j) {
        uint64_t code = 0;
        for (int k = 0; k < 3; ++k) {
            const double c = cen[3 * j + k];
            uint64_t u = 0;
            if (std::isfinite(c) && bhi[k] > blo[k])
                u = (uint64_t)std::min(1023.0, std::max(0.0, std::floor((c - blo[k]) / (bhi[k] - blo[k]) * 1024.0)));
            for (int b = 0; b < 10; ++b) code |= ((u >> b) & 1u) << (3 * b + k);
        }
        key[j] = code;
        wo.order[j] = j;
    }
    std::stable_sort(wo.order.begin(), wo.order.end(), [&](int a, int b) { return key[a] < key[b]; });
    const int ng = (M + kGroup - 1) / kGroup;
    wo.glo.assign(3 * (size_t)ng, std::numeric_limits<T>::infinity());
    wo.ghi.assign(3 * (size_t)ng, -std::numeric_limits<T>::infinity());
    for (int g = 0; g < ng; ++g)
        for (int m = g * kGroup; m < std::min(M, (g + 1) * kGroup); ++m)
            for (int k = 0; k < 3; ++k) {
                wo.glo[3 * g + k] = std::min(wo.glo[3 * g + k], lo[3 * wo.order[m] + k]);
                wo.ghi[3 * g + k] = std::max(wo.ghi[3 * g + k], hi[3 * wo.order[m] + k]);
            }
    return wo;
}

// Filtered distance of one (link, obstacle) pair as the reference's loop
// computes it (ground skipped for the first ground_included links, values
// beyond the obstacle's tolerance replaced by MAX_DISTANCE).
template <typename T>
T pair_filtered(const orc_robot* rb, const P<T>& pr, const SceneT<T>& sc, const T* A, const T* B, int i, int j) {
    T np[6];
    int v = 1;
    T d_tmp;
    if (sc.ground[j] && i < rb->ground_included) d_tmp = pr.maxd;
    else if (!sc.obb.empty()) d_tmp = capsule_obb(A, B, (T)rb->radius[i], &sc.obb[15 * j], pr, np, &v);
    else d_tmp = capsule_box(A, B, (T)rb->radius[i], &sc.obs[6 * j], pr, np, &v, nullptr);
    if (d_tmp > sc.tol[j]) d_tmp = pr.maxd;
    return d_tmp;
}

// computeDistance: RealVectorSpace.cpp:318-392 (fresh state, compute_again).
// With walk (and no nearest points) this is the engine's bulk-pass walk
// instead: each link first evaluates the obstacle that held the previous
// link's minimum (warm start), then walks the obstacles in the walk order,
// skipping a group of 4 whose union box, or an obstacle whose cull box, lies
// farther from the segment's bounding box than max(running link minimum + r,
// EQUALITY_THRESHOLD) + 1e-5; a hit found this way is replaced by the first
// hit in obstacle order, searched among the obstacles before it with the
// bound of a running minimum of 0.  Its results must equal the plain loop's.
template <typename T>
void distance_one(const orc_robot* rb, const P<T>& pr, const SceneT<T>& sc, const T* Pk,
                  T* d_out, T* prof, T* np_out /* M*L*6 or null */, uint8_t* verdict,
                  CullCount* cull = nullptr, const WalkOrder<T>* wo = nullptr) {
    const int L = rb->num_links;
    T d_c = pr.maxd;
    for (int i = 0; i < L; ++i) prof[i] = 0;
    if (np_out) for (int k = 0; k < sc.M * L * 6; ++k) np_out[k] = std::numeric_limits<T>::quiet_NaN();
    int warm = -1;
    for (int i = 0; i < L; ++i) {
        prof[i] = pr.maxd;
        const T* A = Pk + 3 * i;
        const T* B = Pk + 3 * (i + 1);
        T r = (T)rb->radius[i];
        if (cull && wo && !np_out) {
            auto far = [&](const T* lo, const T* hi, T bound) {
                if (seg_box_gap(A, B, lo, hi) > std::max(bound + r, pr.eq) + T(1e-5)) return true;
                return false;
            };
            int best = -1, jh = -1, mhit = -1;   // mhit: walk position of the hit (-1: the warm pair)
            T vh = 0;
            const int w = warm;
            if (w >= 0) {
                cull->evaluated++;
                const T dw = pair_filtered(rb, pr, sc, A, B, i, w);
                if (dw <= 0) { jh = w; vh = dw; }
                else if (dw < prof[i]) { prof[i] = dw; best = w; }
            }
            const int ng = (int)wo->glo.size() / 3;
            for (int g = 0; g < ng && jh < 0; ++g) {
                cull->tests++;
                if (far(&wo->glo[3 * g], &wo->ghi[3 * g], prof[i])) continue;
                for (int m = g * kGroup; m < std::min(sc.M, (g + 1) * kGroup) && jh < 0; ++m) {
                    const int j = wo->order[m];
                    if (j == w) continue;
                    T lo[3], hi[3];
                    cull_box(sc, j, lo, hi);
                    cull->tests++;
                    if (far(lo, hi, prof[i])) continue;
                    cull->evaluated++;
                    const T dt = pair_filtered(rb, pr, sc, A, B, i, j);
                    if (dt < prof[i]) best = j;
                    prof[i] = stdmin(prof[i], dt);
                    if (prof[i] <= 0) { jh = j; vh = prof[i]; mhit = m; }
                }
            }
            if (jh >= 0) {
                T first = vh;
                if (cull->from_hit) {
                    // every walk position before the hit held a pair evaluated
                    // above 0 or culled with a bound at least that of a minimum
                    // of 0: the search walks the positions after it (all of
                    // them after a warm hit) for a smaller index with a
                    // distance <= 0
                    int fj = jh;
                    const int m0 = mhit + 1;
                    for (int g = m0 / kGroup; g < ng; ++g) {
                        cull->tests++;
                        if (far(&wo->glo[3 * g], &wo->ghi[3 * g], T(0))) continue;
                        for (int m = std::max(g * kGroup, m0); m < std::min(sc.M, (g + 1) * kGroup); ++m) {
                            const int j = wo->order[m];
                            if (j == w || j >= fj) continue;
                            T lo[3], hi[3];
                            cull_box(sc, j, lo, hi);
                            cull->tests++;
                            if (far(lo, hi, T(0))) continue;
                            cull->evaluated++;
                            const T dj = pair_filtered(rb, pr, sc, A, B, i, j);
                            if (dj <= 0) { fj = j; first = dj; }
                        }
                    }
                } else {
                    for (int j = 0; j < jh; ++j) {
                        T lo[3], hi[3];
                        cull_box(sc, j, lo, hi);
                        cull->tests++;
                        if (far(lo, hi, T(0))) continue;
                        cull->evaluated++;
                        const T dj = pair_filtered(rb, pr, sc, A, B, i, j);
                        if (dj <= 0) { first = dj; break; }
                    }
                }
                prof[i] = first;
                *d_out = 0;
                *verdict = 0;
                return;
            }
            warm = best;
            d_c = stdmin(d_c, prof[i]);
            continue;
        }
        for (int j = 0; j < sc.M; ++j) {
            T d_tmp;
            T np[6];
            int v = 1;
            if (sc.ground[j] && i < rb->ground_included) {
                d_tmp = pr.maxd;
                np[0] = np[1] = np[2] = 0; np[3] = 0; np[4] = 0; np[5] = -pr.maxd;
            } else if (!sc.obb.empty()) {
                d_tmp = capsule_obb(A, B, r, &sc.obb[15 * j], pr, np, &v);
            } else {
                d_tmp = capsule_box(A, B, r, &sc.obs[6 * j], pr, np, &v, nullptr);
            }
            if (d_tmp > sc.tol[j]) d_tmp = pr.maxd;
            prof[i] = stdmin(prof[i], d_tmp);
            if (prof[i] <= 0) {
                *d_out = 0;
                *verdict = 0;
                if (np_out) for (int k = 0; k < sc.M * L * 6; ++k) np_out[k] = std::numeric_limits<T>::quiet_NaN();
                return;
            }
            if (np_out) std::memcpy(np_out + 6 * (j * L + i), np, sizeof(np));
        }
        d_c = stdmin(d_c, prof[i]);
    }
    *d_out = d_c;
    *verdict = 1;
}

// isValid(q): RealVectorSpace.cpp:277-309
template <typename T>
bool valid_one(const orc_robot* rb, const P<T>& pr, const SceneT<T>& sc, const T* Pk) {
    for (int i = 0; i < rb->num_links; ++i)
        for (int j = 0; j < sc.M; ++j) {
            if (sc.ground[j] && i < rb->ground_included) continue;
            const bool col = sc.obb.empty()
                                 ? collision_box(Pk + 3 * i, Pk + 3 * (i + 1), (T)rb->radius[i], &sc.obs[6 * j], pr)
                                 : collision_obb(Pk + 3 * i, Pk + 3 * (i + 1), (T)rb->radius[i], &sc.obb[15 * j], pr);
            if (col) return false;
        }
    return true;
}

template <typename F>
void parallel_for(int64_t n, int nthreads, F&& fn) {
    if (nthreads <= 1 || n < 2 * nthreads) { fn(0, n); return; }
    std::vector<std::thread> th;
    int64_t chunk = (n + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; ++t) {
        int64_t b = t * chunk, e = std::min<int64_t>(n, b + chunk);
        if (b >= e) break;
        th.emplace_back([=, &fn]() { fn(b, e); });
    }
    for (auto& x : th) x.join();
}

template <typename T>
void distance_batch(const orc_robot* rb, const orc_params* p, const orc_scene* scn, const float* q, int64_t n,
                    T* d, T* profile, T* np, uint8_t* verdict, int nthreads, int64_t* counts = nullptr,
                    bool from_hit = false) {
    P<T> pr(p);
    SceneT<T> sc(scn);
    std::vector<Seg> segs = build_segments(rb);
    const int L = rb->num_links, D = rb->num_dofs;
    WalkOrder<T> wo;
    if (counts) wo = walk_order(sc);
    std::mutex mu;
    parallel_for(n, nthreads, [&](int64_t b, int64_t e) {
        std::vector<T> Pk(3 * (L + 1));
        std::vector<T> prof(L);
        CullCount cc;
        cc.from_hit = from_hit;
        for (int64_t c = b; c < e; ++c) {
            skeleton<T>(rb, segs, q + c * D, Pk.data());
            T dd; uint8_t v;
            distance_one(rb, pr, sc, Pk.data(), &dd, prof.data(), np ? np + c * (int64_t)sc.M * L * 6 : nullptr, &v,
                         counts ? &cc : nullptr, counts ? &wo : nullptr);
            if (d) d[c] = dd;
            if (profile) for (int i = 0; i < L; ++i) profile[c * L + i] = prof[i];
            if (verdict) verdict[c] = v;
        }
        if (counts) {
            std::lock_guard<std::mutex> lk(mu);
            counts[0] += cc.evaluated;
            counts[1] += cc.tests;
        }
    });
}

template <typename T>
void valid_batch(const orc_robot* rb, const orc_params* p, const orc_scene* scn, const float* q, int64_t n,
                 uint8_t* out, int nthreads) {
    P<T> pr(p);
    SceneT<T> sc(scn);
    std::vector<Seg> segs = build_segments(rb);
    const int L = rb->num_links, D = rb->num_dofs;
    parallel_for(n, nthreads, [&](int64_t b, int64_t e) {
        std::vector<T> Pk(3 * (L + 1));
        for (int64_t c = b; c < e; ++c) {
            skeleton<T>(rb, segs, q + c * D, Pk.data());
            out[c] = valid_one(rb, pr, sc, Pk.data()) ? 1 : 0;
        }
    });
}

// Eigen normalized(): v / sqrt(|v|^2) when |v|^2 > 0, else v unchanged [3P]
template <typename T> inline void normalized(const T* v, T* o) {
    T z = sqn3(v);
    if (z > 0) { T s = std::sqrt(z); o[0] = v[0] / s; o[1] = v[1] / s; o[2] = v[2] / s; }
    else cpy3(v, o);
}

// computeDistanceUnderestimation: RealVectorSpace.cpp:402-446 on a fresh
// state (d_c = -1, so the early return at :405-406 never fires).
// 'stored' = 0 when the reference returns 0 without touching the state.
// With max_vel (AbstractTrajectory::computeDistanceUnderestimation,
// AbstractTrajectory.cpp:272-320) R and O first shift by
// normalize(R - O) * max_vel[j] * delta_t (:297-300).
template <typename T>
void under_one(const orc_robot* rb, const P<T>& pr, const SceneT<T>& sc, const T* Pk, const T* npr,
               T* d_out, T* prof, uint8_t* stored, const T* max_vel = nullptr, T delta_t = 0) {
    const int L = rb->num_links;
    T d_c = pr.maxd;
    for (int i = 0; i < L; ++i) prof[i] = 0;
    for (int i = 0; i < L; ++i) {
        prof[i] = pr.maxd;
        for (int j = 0; j < sc.M; ++j) {
            const T* e = npr + 6 * (j * L + i);
            T O[3] = { e[3], e[4], e[5] };
            if (norm3(O) == pr.maxd) continue;
            T R[3] = { e[0], e[1], e[2] };
            T RO[3], n[3], a[3], b[3];
            if (max_vel) {
                sub3(R, O, RO);
                normalized(RO, n);
                for (int k = 0; k < 3; ++k) {
                    const T dl = n[k] * max_vel[j] * delta_t;
                    R[k] = R[k] + dl;
                    O[k] = O[k] + dl;
                }
            }
            sub3(R, O, RO);
            normalized(RO, n);
            sub3(Pk + 3 * i, O, a);
            T da = dot3(a, n);
            normalized(RO, n);
            sub3(Pk + 3 * (i + 1), O, b);
            T db = dot3(b, n);
            T d_tmp = stdmin(da, db) - (T)rb->radius[i];
            if (d_tmp < 0) { *d_out = 0; *stored = 0; return; }
            prof[i] = stdmin(prof[i], d_tmp);
        }
        d_c = stdmin(d_c, prof[i]);
    }
    *d_out = d_c;
    *stored = 1;
}

// capsule-capsule distance for a link pair: xArm6.cpp:417-424
template <typename T>
T capsules_distance(const orc_robot* rb, const P<T>& pr, const T* Pk, int a, int b) {
    T np[6]; int v;
    T d = seg_seg(Pk + 3 * a, Pk + 3 * (a + 1), Pk + 3 * b, Pk + 3 * (b + 1), pr, np, &v);
    return d - (T)rb->radius[a] - (T)rb->radius[b];
}

// float interpolation of RealVectorSpace::interpolateEdge (:75-88) with the
// norm of RealVectorSpace::getNorm (:49-52).  VectorXf sums (norm, dot)
// follow Eigen's SSE reduction order for dynamic sizes [3P]: below 4 terms
// left to right; otherwise packets of 4 accumulated in two lanes over blocks
// of 8, folded, plus a last packet of 4, reduced as (p0 + p2) + (p1 + p3),
// then the tail left to right.
float vecxf_sum(const float* c, int n) {
    if (n < 4) { float r = c[0]; for (int i = 1; i < n; ++i) r = r + c[i]; return r; }
    int aligned = (n / 4) * 4, aligned2 = (n / 8) * 8;
    float p0[4], p1[4];
    for (int k = 0; k < 4; ++k) p0[k] = c[k];
    if (aligned > 4) {
        for (int k = 0; k < 4; ++k) p1[k] = c[4 + k];
        for (int idx = 8; idx < aligned2; idx += 8)
            for (int k = 0; k < 4; ++k) { p0[k] = p0[k] + c[idx + k]; p1[k] = p1[k] + c[idx + 4 + k]; }
        for (int k = 0; k < 4; ++k) p0[k] = p0[k] + p1[k];
        if (aligned > aligned2)
            for (int k = 0; k < 4; ++k) p0[k] = p0[k] + c[aligned2 + k];
    }
    float r = (p0[0] + p0[2]) + (p0[1] + p0[3]);
    for (int i = aligned; i < n; ++i) r = r + c[i];
    return r;
}

float vecxf_norm(const float* v, int n) {
    float sq[ORC_MAX_SEG];
    for (int i = 0; i < n; ++i) sq[i] = v[i] * v[i];
    return std::sqrt(vecxf_sum(sq, n));
}

// AbstractRobot::computeEnclosingRadii for one configuration, R [dofs][dofs+1]:
// xArm6.cpp:237-275 (rows 0..n-2; row 0 from j = 2 in the {x, y} plane; row 3
// from j = 5 against the line P3 P4; capsules_radius_new), Spatial10DOF.cpp:
// 11-30 and Planar2DOF.cpp:132-148 for chains.  Pk is scratch for the skeleton.
void radii32(const orc_robot* rb, const std::vector<Seg>& segs, const float* qc, float* R, float* Pk) {
    const int D = rb->num_dofs;
    float rn[ORC_MAX_SEG];
    for (int i = 0; i < D; ++i) rn[i] = rb->radius[i];
    const bool xarm = rb->skeleton_kind == ORC_SKEL_XARM6;
    if (xarm) {
        for (int i = 0; i + 1 < D; ++i) rn[i] = std::max(rb->radius[i], rb->radius[i + 1]);
        rn[D - 1] = rb->radius[D - 1];
    }
    auto norm2d = [](const float* v) { return std::sqrt(v[0] * v[0] + v[1] * v[1]); };
    // (skeleton->col(j) - skeleton->col(i)).norm(): both operands are
    // dynamic-size MatrixXf columns, so Eigen sums the squares left to right
    auto norm_dyn = [](const float* v) { return std::sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]); };
    skeleton<float>(rb, segs, qc, Pk);
    std::fill(R, R + (size_t)D * (D + 1), 0.0f);
    auto col = [&](int k) { return &Pk[3 * k]; };
    if (xarm) {
        for (int i = 0; i < D - 1; ++i)
            for (int j = i + 1; j <= D; ++j) {
                float* Rij = &R[i * (D + 1) + j];
                const float prev = R[i * (D + 1) + j - 1];
                if (i == 0) {
                    if (j >= 2) *Rij = std::max(prev, norm2d(col(j)) + rn[j - 1]);
                } else if (i == 3) {
                    if (j >= 5) {
                        float AB[3], v[3], proj[3], w[3];
                        sub3(col(4), col(3), AB);
                        sub3(col(j), col(3), v);
                        const float t = dot3(v, AB) / sqn3(AB);
                        for (int k = 0; k < 3; ++k) proj[k] = col(3)[k] + t * AB[k];
                        sub3(col(j), proj, w);
                        *Rij = std::max(prev, norm3(w) + rn[j - 1]);
                    }
                } else {
                    float v[3];
                    sub3(col(j), col(i), v);
                    *Rij = std::max(prev, norm_dyn(v) + rn[j - 1]);
                }
            }
    } else {
        for (int j = 1; j <= D; ++j) R[j] = std::max(R[j - 1], norm2d(col(j)) + rn[j - 1]);
        for (int i = 1; i < D; ++i)
            for (int j = i + 1; j <= D; ++j) {
                float v[3];
                sub3(col(j), col(i), v);
                R[i * (D + 1) + j] = std::max(R[i * (D + 1) + j - 1], norm_dyn(v) + rn[j - 1]);
            }
    }
}

}  // namespace

// ================================================================ C API
extern "C" {

void orc_box_aabb(const float dim[3], const float pos[3], const float quat_wxyz[4], float out6[6]) {
    // Eigen Quaternion::toRotationMatrix (float) [3P]
    const float w = quat_wxyz[0], x = quat_wxyz[1], y = quat_wxyz[2], z = quat_wxyz[3];
    const float tx = 2.0f * x, ty = 2.0f * y, tz = 2.0f * z;
    const float twx = tx * w, twy = ty * w, twz = tz * w;
    const float txx = tx * x, txy = ty * x, txz = tz * x;
    const float tyy = ty * y, tyz = tz * y, tzz = tz * z;
    const float R[9] = { 1.0f - (tyy + tzz), txy - twz, txz + twy,
                         txy + twz, 1.0f - (txx + tzz), tyz - twx,
                         txz - twy, tyz + twx, 1.0f - (txx + tyy) };
    // Eigen isIdentity(dummy_precision = 1e-5f)
    const float prec = 1e-5f;
    bool ident = true;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            float v = R[3 * i + j];
            if (i == j) { if (!(std::fabs(v - 1.0f) <= prec * std::min(std::fabs(v), 1.0f))) ident = false; }
            else if (!(std::fabs(v) <= prec)) ident = false;
        }
    // Box local AABB: half extents 0.5*side (FCL computeBV<AABB,Box> at identity)
    float h[3] = { 0.5f * dim[0], 0.5f * dim[1], 0.5f * dim[2] };
    if (ident) {
        // FCL CollisionObject::computeAABB: translate(aabb_local, T)
        for (int k = 0; k < 3; ++k) { out6[k] = -h[k] + pos[k]; out6[3 + k] = h[k] + pos[k]; }
    } else {
        // center +- aabb_radius, aabb_radius = |min - center| of the local AABB
        float m[3] = { -h[0], -h[1], -h[2] };
        float rad = norm3(m);
        for (int k = 0; k < 3; ++k) { out6[k] = pos[k] - rad; out6[3 + k] = pos[k] + rad; }
    }
}

int orc_skeleton32(const orc_robot* rb, const float* q, int64_t n, float* skel) {
    std::vector<Seg> segs = build_segments(rb);
    const int L = rb->num_links, D = rb->num_dofs;
    for (int64_t c = 0; c < n; ++c) skeleton<float>(rb, segs, q + c * D, skel + c * 3 * (L + 1));
    return 0;
}
int orc_skeleton64(const orc_robot* rb, const float* q, int64_t n, double* skel) {
    std::vector<Seg> segs = build_segments(rb);
    const int L = rb->num_links, D = rb->num_dofs;
    for (int64_t c = 0; c < n; ++c) skeleton<double>(rb, segs, q + c * D, skel + c * 3 * (L + 1));
    return 0;
}

float orc_seg_seg32(const float A[3], const float B[3], const float C[3], const float D[3],
                    const orc_params* p, float np[6], int32_t* np_valid) {
    P<float> pr(p); int v; float d = seg_seg(A, B, C, D, pr, np, &v); *np_valid = v; return d;
}
double orc_seg_seg64(const double A[3], const double B[3], const double C[3], const double D[3],
                     const orc_params* p, double np[6], int32_t* np_valid) {
    P<double> pr(p); int v; double d = seg_seg(A, B, C, D, pr, np, &v); *np_valid = v; return d;
}
float orc_seg_pt32(const float A[3], const float B[3], const float C[3], const orc_params* p, float np[6], int32_t* np_valid) {
    P<float> pr(p); int v; float d = seg_pt(A, B, C, pr, np, &v); *np_valid = v; return d;
}
double orc_seg_pt64(const double A[3], const double B[3], const double C[3], const orc_params* p, double np[6], int32_t* np_valid) {
    P<double> pr(p); int v; double d = seg_pt(A, B, C, pr, np, &v); *np_valid = v; return d;
}
float orc_capsule_box32(const float A[3], const float B[3], float r, const float obs[6],
                        const orc_params* p, float np[6], int32_t* np_valid, int32_t* path) {
    P<float> pr(p); int v, pa; float d = capsule_box(A, B, r, obs, pr, np, &v, &pa);
    *np_valid = v; if (path) *path = pa; return d;
}
double orc_capsule_box64(const double A[3], const double B[3], double r, const double obs[6],
                         const orc_params* p, double np[6], int32_t* np_valid, int32_t* path) {
    P<double> pr(p); int v, pa; double d = capsule_box(A, B, r, obs, pr, np, &v, &pa);
    *np_valid = v; if (path) *path = pa; return d;
}
int32_t orc_collision_capsule_box32(const float A[3], const float B[3], float r, const float obs[6], const orc_params* p) {
    P<float> pr(p); return collision_box(A, B, r, obs, pr) ? 1 : 0;
}
int32_t orc_collision_capsule_box64(const double A[3], const double B[3], double r, const double obs[6], const orc_params* p) {
    P<double> pr(p); return collision_box(A, B, r, obs, pr) ? 1 : 0;
}

void orc_distance32(const orc_robot* rb, const orc_params* p, const orc_scene* sc, const float* q, int64_t n,
                    float* d, float* profile, float* np, uint8_t* verdict, int32_t nthreads) {
    distance_batch<float>(rb, p, sc, q, n, d, profile, np, verdict, nthreads);
}
void orc_distance64(const orc_robot* rb, const orc_params* p, const orc_scene* sc, const float* q, int64_t n,
                    double* d, double* profile, double* np, uint8_t* verdict, int32_t nthreads) {
    distance_batch<double>(rb, p, sc, q, n, d, profile, np, verdict, nthreads);
}
void orc_is_valid32(const orc_robot* rb, const orc_params* p, const orc_scene* sc, const float* q, int64_t n,
                    uint8_t* valid, int32_t nthreads) {
    valid_batch<float>(rb, p, sc, q, n, valid, nthreads);
}
void orc_is_valid64(const orc_robot* rb, const orc_params* p, const orc_scene* sc, const float* q, int64_t n,
                    uint8_t* valid, int32_t nthreads) {
    valid_batch<double>(rb, p, sc, q, n, valid, nthreads);
}

void orc_underestimation32(const orc_robot* rb, const orc_params* p, const orc_scene* scn, const float* q, int64_t n,
                           const float* np_ref, int64_t np_stride, float* d, float* profile, uint8_t* stored,
                           int32_t nthreads) {
    P<float> pr(p);
    SceneT<float> sc(scn);
    std::vector<Seg> segs = build_segments(rb);
    const int L = rb->num_links, D = rb->num_dofs;
    parallel_for(n, nthreads, [&](int64_t b, int64_t e) {
        std::vector<float> Pk(3 * (L + 1)), prof(L);
        for (int64_t c = b; c < e; ++c) {
            skeleton<float>(rb, segs, q + c * D, Pk.data());
            float dd; uint8_t st;
            under_one(rb, pr, sc, Pk.data(), np_ref + c * np_stride, &dd, prof.data(), &st);
            d[c] = dd; stored[c] = st;
            if (profile) for (int i = 0; i < L; ++i) profile[c * L + i] = prof[i];
        }
    });
}

void orc_underestimation_moving32(const orc_robot* rb, const orc_params* p, const orc_scene* scn, const float* q,
                                  int64_t n, const float* np_ref, int64_t np_stride, const float* max_vel,
                                  float delta_t, float* d, float* profile, uint8_t* stored, int32_t nthreads) {
    P<float> pr(p);
    SceneT<float> sc(scn);
    std::vector<Seg> segs = build_segments(rb);
    const int L = rb->num_links, D = rb->num_dofs;
    parallel_for(n, nthreads, [&](int64_t b, int64_t e) {
        std::vector<float> Pk(3 * (L + 1)), prof(L);
        for (int64_t c = b; c < e; ++c) {
            skeleton<float>(rb, segs, q + c * D, Pk.data());
            float dd; uint8_t st;
            under_one(rb, pr, sc, Pk.data(), np_ref + c * np_stride, &dd, prof.data(), &st, max_vel, delta_t);
            d[c] = dd; stored[c] = st;
            if (profile) for (int i = 0; i < L; ++i) profile[c * L + i] = prof[i];
        }
    });
}

// AbstractTrajectory::isSafe, AbstractTrajectory.cpp:212-256, for num_traj
// trajectories of num_points points leaving q_current.  max_obs_vel is the
// largest max_vel from 0 (:44-49).  checkSelfCollision(q_final) is the
// capsule test of the listed pairs when self_collision is set.
void orc_traj_safe32(const orc_robot* rb, const orc_params* p, const orc_scene* scn, const float* points,
                     int64_t num_traj, int32_t num_points, const float* q_current, const float* profile_current,
                     const float* nearest_current, const float* max_vel, float t_iter0, float time_step,
                     const int32_t* pairs, int32_t npairs, uint8_t* safe, int32_t nthreads) {
    P<float> pr(p);
    SceneT<float> sc(scn);
    std::vector<Seg> segs = build_segments(rb);
    const int L = rb->num_links, D = rb->num_dofs;
    float max_obs_vel = 0;
    for (int j = 0; j < sc.M; ++j)
        if (max_vel[j] > max_obs_vel) max_obs_vel = max_vel[j];
    parallel_for(num_traj, nthreads, [&](int64_t b, int64_t e) {
        std::vector<float> Pk(3 * (L + 1)), R((size_t)D * (D + 1)), qt(D), prof(L), pn(L), dq(D), prod(D);
        for (int64_t c = b; c < e; ++c) {
            std::copy(q_current, q_current + D, qt.begin());
            std::copy(profile_current, profile_current + L, prof.begin());
            float t_iter = t_iter0, t_init = 0;
            uint8_t ok = 1;
            for (int i = 0; i < num_points && ok; ++i) {
                const float* qf = points + ((int64_t)c * num_points + i) * D;
                if (npairs > 0) {
                    skeleton<float>(rb, segs, qf, Pk.data());
                    for (int m = 0; m < npairs; ++m)
                        if (capsules_distance<float>(rb, pr, Pk.data(), pairs[2 * m], pairs[2 * m + 1]) < 0) ok = 0;
                    if (!ok) break;
                }
                const float rho_obs = max_obs_vel * (t_iter - t_init);
                for (int k = 0; k < D; ++k) dq[k] = std::fabs(qf[k] - qt[k]);
                radii32(rb, segs, qt.data(), R.data(), Pk.data());
                for (int j = 0; j < D; ++j) {
                    for (int k = 0; k < D; ++k) prod[k] = R[(size_t)k * (D + 1) + j + 1] * dq[k];
                    const float rho_robot = vecxf_sum(prod.data(), D);
                    if (rho_robot + rho_obs >= prof[j]) {
                        std::copy(qf, qf + D, qt.begin());
                        skeleton<float>(rb, segs, qt.data(), Pk.data());
                        float dd; uint8_t st;
                        under_one(rb, pr, sc, Pk.data(), nearest_current, &dd, pn.data(), &st, max_vel, t_iter);
                        t_init = t_iter;
                        if (!st) ok = 0;      // the state keeps d_c = -1
                        else {
                            prof = pn;
                            if (dd <= 0) ok = 0;
                        }
                        break;
                    }
                }
                t_iter += time_step;
            }
            safe[c] = ok;
        }
    });
}

// Environment::updateEnvironment(delta_time), Environment.cpp:194-377, for
// the deterministic laws (0 straight :225-238, 1 circular :241-257, 2 two
// tunnels :260-287, 3 light directions :330-376), 'steps' times, in float as
// fcl::Vector3f computes (fixed-size norms and dots x + (y + z)).  pos, vel
// [M][3], path_len [M], sign [M] are updated in place; boxes with
// dynamic[j] = 0 stay.
void orc_env_step32(int32_t law, const float ws[3], float ws_radius, float base_radius, float robot_max_vel,
                    int32_t M, float* pos, float* vel, const float* max_vel, const uint8_t* dynamic,
                    float* path_len, int32_t* sign, float dt, int32_t steps) {
    auto norm3v = [](const float* a) { return std::sqrt(a[0] * a[0] + (a[1] * a[1] + a[2] * a[2])); };
    for (int j = 0; j < M; ++j) {
        if (!dynamic[j]) continue;
        float* p = pos + 3 * j;
        float* v = vel + 3 * j;
        const float vmax = max_vel[j];
        for (int s = 0; s < steps; ++s) {
            if (law == 0) {
                p[0] -= dt * vmax;
            } else if (law == 1) {
                const float phi = std::atan2(p[1], p[0]);
                const float dir[3] = {-std::sin(phi), std::cos(phi), 0.0f};
                for (int k = 0; k < 3; ++k) v[k] = dir[k] * vmax;
                for (int k = 0; k < 3; ++k) p[k] += v[k] * dt;
            } else if (law == 2) {
                if (path_len[j] > 0.2f) {
                    sign[j] *= -1;
                    path_len[j] = -0.2f;
                }
                const float unit[3] = {p[1] > 0 ? 1.0f : 0.0f, p[1] > 0 ? 0.0f : 1.0f, 0.0f};
                for (int k = 0; k < 3; ++k) v[k] = unit[k] * (float)sign[j] * vmax;
                path_len[j] += norm3v(v) * dt;
                for (int k = 0; k < 3; ++k) p[k] += v[k] * dt;
            } else {
                const float tol = std::max(norm3v(v) / robot_max_vel, base_radius);
                float pn[3], n[3], rel[3];
                for (int k = 0; k < 3; ++k) pn[k] = p[k] + v[k] * dt;
                for (int k = 0; k < 3; ++k) rel[k] = pn[k] - ws[k];
                bool change = true;
                if (pn[2] < 0) { n[0] = 0; n[1] = 0; n[2] = 1; }
                else if (norm3v(rel) > ws_radius) { n[0] = -pn[0]; n[1] = -pn[1]; n[2] = -(pn[2] - ws[2]); }
                else if (std::sqrt(pn[0] * pn[0] + pn[1] * pn[1]) < tol && pn[2] < ws[2]) { n[0] = pn[0]; n[1] = pn[1]; n[2] = 0; }
                else if (norm3v(rel) < tol) { n[0] = pn[0]; n[1] = pn[1]; n[2] = pn[2] - ws[2]; }
                else { for (int k = 0; k < 3; ++k) p[k] = pn[k]; change = false; }
                if (change) {
                    float dlt[3];
                    for (int k = 0; k < 3; ++k) dlt[k] = pn[k] - p[k];
                    const float t = dot3(dlt, n) / sqn3(n);
                    for (int k = 0; k < 3; ++k) p[k] = 2 * pn[k] - p[k] - 2 * t * n[k];
                    for (int k = 0; k < 3; ++k) v[k] = (p[k] - pn[k]) / dt;
                }
            }
        }
    }
}

void orc_self_distance32(const orc_robot* rb, const orc_params* p, const float* q, int64_t n,
                         const int32_t* pairs, int32_t npairs, float* dist, int32_t nthreads) {
    P<float> pr(p);
    std::vector<Seg> segs = build_segments(rb);
    const int L = rb->num_links, D = rb->num_dofs;
    parallel_for(n, nthreads, [&](int64_t b, int64_t e) {
        std::vector<float> Pk(3 * (L + 1));
        for (int64_t c = b; c < e; ++c) {
            skeleton<float>(rb, segs, q + c * D, Pk.data());
            for (int k = 0; k < npairs; ++k)
                dist[c * npairs + k] = capsules_distance(rb, pr, Pk.data(), pairs[2 * k], pairs[2 * k + 1]);
        }
    });
}
void orc_self_distance64(const orc_robot* rb, const orc_params* p, const float* q, int64_t n,
                         const int32_t* pairs, int32_t npairs, double* dist, int32_t nthreads) {
    P<double> pr(p);
    std::vector<Seg> segs = build_segments(rb);
    const int L = rb->num_links, D = rb->num_dofs;
    parallel_for(n, nthreads, [&](int64_t b, int64_t e) {
        std::vector<double> Pk(3 * (L + 1));
        for (int64_t c = b; c < e; ++c) {
            skeleton<double>(rb, segs, q + c * D, Pk.data());
            for (int k = 0; k < npairs; ++k)
                dist[c * npairs + k] = capsules_distance(rb, pr, Pk.data(), pairs[2 * k], pairs[2 * k + 1]);
        }
    });
}

// isValid(q1, q2): RealVectorSpace.cpp:261-275 — K checks from the far end,
// q1 itself not checked.  Environment only (self-collision is a separate
// robot call in the reference).
void orc_edge_valid32(const orc_robot* rb, const orc_params* p, const orc_scene* scn,
                      const float* q1, const float* q2, int64_t n, uint8_t* valid, int32_t nthreads) {
    P<float> pr(p);
    SceneT<float> sc(scn);
    std::vector<Seg> segs = build_segments(rb);
    const int L = rb->num_links, D = rb->num_dofs;
    const int K = p->num_checks;
    parallel_for(n, nthreads, [&](int64_t b, int64_t e) {
        std::vector<float> Pk(3 * (L + 1)), qn(D), diff(D);
        for (int64_t c = b; c < e; ++c) {
            const float* a = q1 + c * D;
            const float* z = q2 + c * D;
            for (int i = 0; i < D; ++i) diff[i] = a[i] - z[i];
            float dist = vecxf_norm(diff.data(), D);
            bool ok = true;
            for (float k = (float)K; k >= 1; k--) {
                float step = k / (float)K * dist;
                if (dist > 0) {
                    float f = step / dist;
                    for (int i = 0; i < D; ++i) qn[i] = a[i] + (z[i] - a[i]) * f;
                } else {
                    for (int i = 0; i < D; ++i) qn[i] = z[i];
                }
                skeleton<float>(rb, segs, qn.data(), Pk.data());
                if (!valid_one(rb, pr, sc, Pk.data())) { ok = false; break; }
            }
            valid[c] = ok ? 1 : 0;
        }
    });
}

// Discretised-edge first collision (build-defined workload): state k of edge
// e is q1 + (step * k) * dir for k = 0..K-1, evaluated in float as
// s = step * (float)k; q_i = q1_i + s * dir_i.  A state collides when the
// environment collision tree reports a hit or any listed capsule pair has
// segment distance - r_a - r_b < 0.  first[e] = first colliding k or -1.
void orc_edge_first_collision32(const orc_robot* rb, const orc_params* p, const orc_scene* scn,
                                const int32_t* pairs, int32_t npairs, float step,
                                const float* q1, const float* dir, const int32_t* K, int64_t E,
                                int32_t* first, int32_t nthreads) {
    P<float> pr(p);
    SceneT<float> sc(scn);
    std::vector<Seg> segs = build_segments(rb);
    const int L = rb->num_links, D = rb->num_dofs;
    parallel_for(E, nthreads, [&](int64_t b, int64_t e) {
        std::vector<float> Pk(3 * (L + 1)), qn(D);
        for (int64_t c = b; c < e; ++c) {
            int32_t hit = -1;
            for (int k = 0; k < K[c]; ++k) {
                float s = step * (float)k;
                for (int i = 0; i < D; ++i) qn[i] = q1[c * D + i] + s * dir[c * D + i];
                skeleton<float>(rb, segs, qn.data(), Pk.data());
                bool bad = !valid_one(rb, pr, sc, Pk.data());
                for (int m = 0; !bad && m < npairs; ++m)
                    if (capsules_distance(rb, pr, Pk.data(), pairs[2 * m], pairs[2 * m + 1]) < 0) bad = true;
                if (bad) { hit = k; break; }
            }
            first[c] = hit;
        }
    });
}

void orc_distance_cull32(const orc_robot* rb, const orc_params* p, const orc_scene* sc, const float* q, int64_t n,
                         float* d, float* profile, uint8_t* verdict, int64_t counts[2], int32_t nthreads,
                         int32_t search) {
    counts[0] = counts[1] = 0;
    distance_batch<float>(rb, p, sc, q, n, d, profile, nullptr, verdict, nthreads, counts, search != 0);
}
void orc_distance_cull64(const orc_robot* rb, const orc_params* p, const orc_scene* sc, const float* q, int64_t n,
                         double* d, double* profile, uint8_t* verdict, int64_t counts[2], int32_t nthreads,
                         int32_t search) {
    counts[0] = counts[1] = 0;
    distance_batch<double>(rb, p, sc, q, n, d, profile, nullptr, verdict, nthreads, counts, search != 0);
}

// AbstractRobot::computeEnclosingRadii, float skeleton (ref32):
//   xArm6 (xArm6.cpp:237-275) with capsules_radius_new (xArm6.cpp:98-105),
//   chains as Spatial10DOF (Spatial10DOF.cpp:11-30) with capsules_radius.
// R is [n][dofs][dofs+1], R[i][j] = the reference's R(i, j).
void orc_enclosing_radii32(const orc_robot* rb, const float* q, int64_t n, float* Rout, int32_t nthreads) {
    std::vector<Seg> segs = build_segments(rb);
    const int D = rb->num_dofs;
    parallel_for(n, nthreads, [&](int64_t b, int64_t e) {
        std::vector<float> Pk(3 * (rb->num_links + 1));
        for (int64_t c = b; c < e; ++c) radii32(rb, segs, q + c * D, Rout + c * (int64_t)D * (D + 1), Pk.data());
    });
}

// RBTConnect::extendSpine, src/planners/rbt/RBTConnect.cpp:85-150, with
// robot->checkSelfCollision(q_temp, q_new) (:129-131) false (a robot without
// self-collision checking).  R->col(i+1).dot(delta_q) is a VectorXf dot
// (vecxf_sum order); (skeleton->col(k) - skeleton_new->col(k)).norm() a
// dynamic-size 3-vector norm summed left to right.
void orc_extend_spines32(const orc_robot* rb, const float* q, const float* d, const float* profile,
                         const float* q_e, int64_t n, int32_t num_iter, int32_t expanded, float* q_new,
                         int32_t* status, int32_t nthreads) {
    std::vector<Seg> segs = build_segments(rb);
    const int D = rb->num_dofs, L = rb->num_links;
    auto norm_dyn = [](const float* v) { return std::sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]); };
    parallel_for(n, nthreads, [&](int64_t b, int64_t e) {
        std::vector<float> P0(3 * (L + 1)), Pn(3 * (L + 1)), Pk(3 * (L + 1)), R((size_t)D * (D + 1));
        std::vector<float> qt(D), dq(D), prod(D), rho_profile(L);
        for (int64_t c = b; c < e; ++c) {
            const float* qc = q + c * D;
            const float* qe = q_e + c * D;
            float* qn = q_new + c * D;
            if (d[c] <= 0) {  // :90-91
                std::copy(qc, qc + D, qn);
                status[c] = 2;
                continue;
            }
            skeleton<float>(rb, segs, qc, P0.data());
            std::fill(rho_profile.begin(), rho_profile.end(), 0.0f);
            float rho = 0;
            int32_t st = 1;
            std::copy(qc, qc + D, qt.begin());
            for (int counter = 1;; ++counter) {
                radii32(rb, segs, qt.data(), R.data(), Pk.data());
                for (int k = 0; k < D; ++k) dq[k] = std::fabs(qe[k] - qt[k]);
                auto covered = [&](int j) {
                    for (int k = 0; k < D; ++k) prod[k] = R[(size_t)k * (D + 1) + j] * dq[k];
                    return vecxf_sum(prod.data(), D);
                };
                float step;
                if (expanded) {
                    step = std::numeric_limits<float>::infinity();
                    for (int i = 0; i < D; ++i)
                        step = std::min(step, (profile[c * L + i] - rho_profile[i]) / covered(i + 1));
                } else {
                    step = (d[c] - rho) / covered(D);
                }
                if (step > 1) {
                    std::copy(qe, qe + D, qn);
                    st = 3;
                } else {
                    for (int k = 0; k < D; ++k) qn[k] = qt[k] + step * (qe[k] - qt[k]);
                }
                if (counter == num_iter || st != 1) break;
                skeleton<float>(rb, segs, qn, Pn.data());
                std::fill(rho_profile.begin(), rho_profile.end(), 0.0f);
                float rho_prev = 0;
                for (int k = 1; k <= L; ++k) {
                    float v[3];
                    sub3(&P0[3 * k], &Pn[3 * k], v);
                    const float rho_k = norm_dyn(v);
                    rho_profile[k - 1] = std::max(rho_prev, rho_k);
                    rho_prev = rho_k;
                    rho = std::max(rho, rho_k);
                }
                std::copy(qn, qn + D, qt.begin());
            }
            status[c] = st;
        }
    });
}

// xArm6::checkSelfCollision(q1, q2), src/robots/xArm6.cpp:284-368, with
// checkRealSelfCollision (the FCL mesh test, :426-483) taken as true: a pair
// collides when its capsules overlap.  Phase 1 (:300-324): at most 5
// radii-bounded steps; phase 2 (:326-366): walk in eps_step steps, skipping
// the pairs whose bounded step exceeds the walked fraction.  The walk is
// capped for non-finite or huge coordinates (collide = 255).
void orc_edge_self_collision32(const orc_robot* rb, const orc_params* p, const float* q1, const float* q2,
                               int64_t n, float eps_step, uint8_t* collide, float* q_out, int32_t nthreads) {
    std::vector<Seg> segs = build_segments(rb);
    const int D = rb->num_dofs;
    const P<float> pr(p);
    parallel_for(n, nthreads, [&](int64_t b, int64_t e) {
        std::vector<float> Pk(3 * (rb->num_links + 1)), R((size_t)D * (D + 1)), qt(D), qn(D), dq(D), prod(D);
        for (int64_t c = b; c < e; ++c) {
            const float* a = q1 + c * D;
            const float* g = q2 + c * D;
            float* out = q_out + c * D;
            auto caps = [&](const float* q, int i, int j) {
                skeleton<float>(rb, segs, q, Pk.data());
                return capsules_distance<float>(rb, pr, Pk.data(), i, j);
            };
            auto covered = [&](int j) {
                for (int k = 0; k < D; ++k) prod[k] = R[(size_t)k * (D + 1) + j] * dq[k];
                return vecxf_sum(prod.data(), D);
            };
            std::copy(a, a + D, qt.begin());
            float s04 = 0, s05 = 0, s14 = 0, s15 = 0;
            bool free_edge = false;
            for (int iter = 1;; ++iter) {
                radii32(rb, segs, qt.data(), R.data(), Pk.data());
                for (int k = 0; k < D; ++k) dq[k] = std::fabs(g[k] - qt[k]);
                const float phi2 = covered(2), phi4 = covered(4), phi5 = covered(5), phi6 = covered(6);
                s04 = caps(qt.data(), 0, 4) / std::max(phi4, phi5);
                s05 = caps(qt.data(), 0, 5) / std::max(phi5, phi6);
                s14 = caps(qt.data(), 1, 4) / (phi2 + std::max(phi4, phi5));
                s15 = caps(qt.data(), 1, 5) / (phi2 + std::max(phi5, phi6));
                const float step = std::min(std::min(s04, s05), std::min(s14, s15));
                if (step > 1) { free_edge = true; break; }
                if (iter == 5 || step <= 0) break;
                for (int k = 0; k < D; ++k) qt[k] = qt[k] + step * (g[k] - qt[k]);
            }
            if (free_edge) { std::copy(g, g + D, out); collide[c] = 0; continue; }
            auto norm_to_g = [&](const std::vector<float>& v) {
                for (int k = 0; k < D; ++k) prod[k] = (g[k] - v[k]) * (g[k] - v[k]);
                return std::sqrt(vecxf_sum(prod.data(), D));
            };
            const float Dn = norm_to_g(qt);
            float dist = Dn;
            const float span = Dn / eps_step;
            const int cap = (span < 65536.0f) ? (int)span + 16 : 65552;
            uint8_t hit = 0;
            for (int walked = 0;; ++walked) {
                if (walked > cap) { hit = 255; break; }
                bool reached = false;
                if (dist > eps_step) {
                    for (int k = 0; k < D; ++k) qn[k] = qt[k] + (g[k] - qt[k]) / dist * eps_step;
                } else {
                    std::copy(g, g + D, qn.begin());
                    reached = true;
                }
                dist = norm_to_g(qn);
                const float crit = 1 - dist / Dn;
                const bool skip[4] = {s05 > crit, s15 > crit, s04 > crit, s14 > crit};
                const int pairs[4][2] = {{0, 5}, {1, 5}, {0, 4}, {1, 4}};
                bool any = false;
                for (int m = 0; m < 4 && !any; ++m)
                    if (!skip[m] && caps(qn.data(), pairs[m][0], pairs[m][1]) < 0) any = true;
                if (any) { hit = 1; break; }
                qt = qn;
                if (reached) break;
            }
            if (hit == 1) std::copy(qt.begin(), qt.end(), out);
            else std::copy(g, g + D, out);
            collide[c] = hit;
        }
    });
}

float orc_capsule_obb32(const float A[3], const float B[3], float r, const float obb[15], const orc_params* p,
                        float np[6], int32_t* np_valid) {
    P<float> pr(p); int v; float d = capsule_obb(A, B, r, obb, pr, np, &v); *np_valid = v; return d;
}
double orc_capsule_obb64(const double A[3], const double B[3], double r, const double obb[15], const orc_params* p,
                         double np[6], int32_t* np_valid) {
    P<double> pr(p); int v; double d = capsule_obb(A, B, r, obb, pr, np, &v); *np_valid = v; return d;
}

// Smallest self-pair capsule distance per configuration (std::min order
// over the pair list, starting from MAX_DISTANCE).
void orc_self_min32(const orc_robot* rb, const orc_params* p, const float* q, int64_t n, const int32_t* pairs,
                    int32_t npairs, float* out, int32_t nthreads) {
    P<float> pr(p);
    std::vector<Seg> segs = build_segments(rb);
    const int L = rb->num_links, D = rb->num_dofs;
    parallel_for(n, nthreads, [&](int64_t b, int64_t e) {
        std::vector<float> Pk(3 * (L + 1));
        for (int64_t c = b; c < e; ++c) {
            skeleton<float>(rb, segs, q + c * D, Pk.data());
            float m = pr.maxd;
            for (int k = 0; k < npairs; ++k)
                m = stdmin(m, capsules_distance(rb, pr, Pk.data(), pairs[2 * k], pairs[2 * k + 1]));
            out[c] = m;
        }
    });
}

// Independent checker: exact distance from segment AB to the axis-aligned box
// in double.  g(t) = dist(P(t), box)^2 is convex and piecewise quadratic with
// breakpoints where a coordinate of P(t) crosses a slab plane; the minimum is
// at a breakpoint, an end point, or the stationary point of one piece.
double orc_true_seg_box(const double A[3], const double B[3], const double obs[6]) {
    double u[3] = { B[0] - A[0], B[1] - A[1], B[2] - A[2] };
    std::vector<double> ts = { 0.0, 1.0 };
    for (int k = 0; k < 3; ++k)
        if (u[k] != 0) {
            for (double pl : { obs[k], obs[k + 3] }) {
                double t = (pl - A[k]) / u[k];
                if (t > 0 && t < 1) ts.push_back(t);
            }
        }
    std::sort(ts.begin(), ts.end());
    auto g = [&](double t) {
        double s = 0;
        for (int k = 0; k < 3; ++k) {
            double x = A[k] + t * u[k];
            double e = (x < obs[k]) ? obs[k] - x : (x > obs[k + 3] ? x - obs[k + 3] : 0.0);
            s += e * e;
        }
        return s;
    };
    double best = std::min(g(0.0), g(1.0));
    for (size_t m = 0; m + 1 < ts.size(); ++m) {
        double a = ts[m], b = ts[m + 1];
        best = std::min(best, g(a));
        // on (a, b) each coordinate keeps its region; solve the quadratic
        double mid = 0.5 * (a + b), num = 0, den = 0;
        for (int k = 0; k < 3; ++k) {
            double x = A[k] + mid * u[k];
            if (x < obs[k]) { num += (obs[k] - A[k]) * u[k]; den += u[k] * u[k]; }
            else if (x > obs[k + 3]) { num += (obs[k + 3] - A[k]) * u[k]; den += u[k] * u[k]; }
        }
        if (den > 0) {
            double t = num / den;
            if (t > a && t < b) best = std::min(best, g(t));
        }
    }
    return std::sqrt(best);
}

}  // extern "C"
