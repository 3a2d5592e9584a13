// CPU check of the kernels' geometry source against the oracle.  The device
// functions in rpmplv2_amd/csrc/rvs_geometry.hpp are __host__ __device__;
// this program compiles them for the host (no GPU needed) and compares them
// bit for bit with the oracle's line-faithful restatement on random and
// adversarial pairs.  Usage: geom_check <num_pairs> <seed>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>

#include "../../oracle/rvs_oracle.h"
#include "../../rpmplv2_amd/csrc/rvs_geometry.hpp"

using namespace rvs;

static bool same(float a, float b) { return (std::isnan(a) && std::isnan(b)) || std::memcmp(&a, &b, 4) == 0 || (a == b); }
static bool same(double a, double b) { return (std::isnan(a) && std::isnan(b)) || (a == b); }

// IEEE arithmetic through the relaxed bulk pass's branch-free capsule-box form
// (flat_capsule_box): checks its algebra and screening on the host, to
// rounding (the bulk pass itself uses hardware reciprocal and square root,
// which exist on the GPU only).
struct AxisMath {
    template <typename U> static U div(U a, U b) { return a / b; }
    template <typename U> static U root(U x) { return vsqrt(x); }
    template <typename U> static U clamp(U x, U lo, U hi) { return std::fmin(std::fmax(x, lo), hi); }
};

template <typename T> struct Rng {
    std::mt19937_64 g;
    explicit Rng(uint64_t s) : g(s) {}
    double u() { return std::uniform_real_distribution<double>(0, 1)(g); }
    double u(double a, double b) { return a + (b - a) * u(); }
    int i(int n) { return (int)(u() * n) % n; }
};

template <typename T>
int run(long n, uint64_t seed, const char* tag) {
    Rng<T> R(seed);
    orc_params op{1e-4f, 1000.0f, 10};
    Params<T> p{(T)op.eq_threshold, (T)op.max_distance};
    long bad_cb = 0, bad_np = 0, bad_col = 0, bad_ss = 0, bad_sp = 0, bad_pl = 0;
    long bad_flat = 0, flag_flat = 0;
    double max_flat = 0;
    long paths[10] = {0};
    for (long it = 0; it < n; ++it) {
        // box
        T lo[3], hi[3];
        for (int k = 0; k < 3; ++k) {
            double c = R.u(-1, 1), h = R.u(0.02, 0.4);
            lo[k] = (T)(c - h); hi[k] = (T)(c + h);
        }
        T A[3], B[3];
        int mode = R.i(6);
        for (int k = 0; k < 3; ++k) {
            A[k] = (T)R.u(-1.6, 1.6);
            B[k] = (T)(A[k] + R.u(-0.6, 0.6));
        }
        if (mode == 1) {  // end point exactly on a face plane or an edge coordinate
            int k = R.i(3);
            A[k] = R.i(2) ? lo[k] : hi[k];
        } else if (mode == 2) {  // segment parallel to an axis
            int k = R.i(3);
            for (int m = 0; m < 3; ++m) if (m != k) B[m] = A[m];
        } else if (mode == 3) {  // near-touching: shift so that it almost grazes an edge
            int k = R.i(3);
            A[k] = hi[k] + (T)R.u(-1e-4, 1e-4);
            B[k] = hi[k] + (T)R.u(-1e-4, 1e-4);
        } else if (mode == 4) {  // degenerate segment
            for (int m = 0; m < 3; ++m) B[m] = A[m];
        } else if (mode == 5) {  // segment lying in a face plane (the tree's NaN paths)
            int k = R.i(3);
            A[k] = B[k] = R.i(2) ? lo[k] : hi[k];
        }
        T r = (T)(R.i(4) == 0 ? 0.0 : R.u(0, 0.1));
        T obs[6] = {lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]};
        Box<T> b;
        for (int k = 0; k < 3; ++k) { b.lo[k] = lo[k]; b.hi[k] = hi[k]; }
        Vec3<T> a{A[0], A[1], A[2]}, bb{B[0], B[1], B[2]};
        // capsule-box distance
        T onp[6]; int32_t ov = 0, opath = 0;
        T od;
        if constexpr (sizeof(T) == 4) od = orc_capsule_box32(A, B, r, obs, &op, onp, &ov, &opath);
        else od = orc_capsule_box64(A, B, r, obs, &op, onp, &ov, &opath);
        if (opath >= 0 && opath < 10) paths[opath]++;
        NP<T> np{};
        uint32_t st = 0;
        T gd = capsule_box<T, true>(a, bb, r, b, p, &np, &st);
        if (!same(gd, od)) {
            if (bad_cb < 5) printf("[%s] capsule_box mismatch it=%ld path=%d gpu=%.9g oracle=%.9g\n", tag, it, opath, (double)gd, (double)od);
            bad_cb++;
        } else if (ov && gd - r > 0) {
            bool ok = np.valid == (ov != 0);
            T g6[6] = {np.r.x, np.r.y, np.r.z, np.o.x, np.o.y, np.o.z};
            for (int k = 0; k < 6 && ok; ++k) ok = same(g6[k], onp[k]);
            if (!ok) { if (bad_np < 5) printf("[%s] nearest-point mismatch it=%ld path=%d\n", tag, it, opath); bad_np++; }
        }
        if constexpr (sizeof(T) == 4) {
            // the flat evaluation of the per-lane walk: where it does not flag
            // the pair, the contact outputs (0 inside, -r crossing) are the
            // reference's exactly and other values within 2e-6 m; every NaN of
            // the tree is flagged
            bool unc = false;
            const float gf = flat_capsule_box<AxisMath>(a, bb, r, b, p, 1e-4f, unc);
            if (unc) {
                flag_flat++;
            } else {
                const bool contact = od == 0.0f || od == -r;
                const double e = std::fabs((double)gf - (double)od);
                const bool ok = !std::isnan(od) && (contact ? gf == od : e <= 2e-6);
                if (!ok) {
                    if (bad_flat < 5)
                        printf("[%s] flat mismatch it=%ld mode=%d path=%d flat=%.9g oracle=%.9g r=%.9g\n", tag, it, mode,
                               opath, (double)gf, (double)od, (double)r);
                    bad_flat++;
                }
                if (!contact && e > max_flat) max_flat = e;
            }
        }
        // collision tree
        T rn = (T)((double)r * std::sqrt(3.0) / 3.0);
        int oc;
        if constexpr (sizeof(T) == 4) oc = orc_collision_capsule_box32(A, B, r, obs, &op);
        else oc = orc_collision_capsule_box64(A, B, r, obs, &op);
        int gc = collision_box(a, bb, r, rn, b, p) ? 1 : 0;
        if (gc != oc) { if (bad_col < 5) printf("[%s] collision mismatch it=%ld gpu=%d oracle=%d\n", tag, it, gc, oc); bad_col++; }
        // segment-segment and segment-point
        T C[3], D[3];
        for (int k = 0; k < 3; ++k) { C[k] = (T)R.u(-1, 1); D[k] = (T)(C[k] + R.u(-0.5, 0.5)); }
        if (mode == 2) for (int k = 0; k < 3; ++k) D[k] = C[k] + (B[k] - A[k]);  // parallel
        Vec3<T> c{C[0], C[1], C[2]}, dd{D[0], D[1], D[2]};
        T snp[6]; int32_t sv;
        T os;
        if constexpr (sizeof(T) == 4) os = orc_seg_seg32(A, B, C, D, &op, snp, &sv);
        else os = orc_seg_seg64(A, B, C, D, &op, snp, &sv);
        NP<T> t{};
        T gs = seg_seg<T, true>(a, bb, c, dd, p, &t);
        if (!same(gs, os)) { if (bad_ss < 5) printf("[%s] seg_seg mismatch it=%ld %.9g %.9g\n", tag, it, (double)gs, (double)os); bad_ss++; }
        T opnt;
        if constexpr (sizeof(T) == 4) opnt = orc_seg_pt32(A, B, C, &op, snp, &sv);
        else opnt = orc_seg_pt64(A, B, C, &op, snp, &sv);
        T gp = seg_pt<T, false>(a, bb, c, p, nullptr);
        if (!same(gp, opnt)) { if (bad_sp < 5) printf("[%s] seg_pt mismatch it=%ld\n", tag, it); bad_sp++; }
        (void)bad_pl;
    }
    printf("[%s] pairs=%ld capsule_box=%ld nearest=%ld collision=%ld seg_seg=%ld seg_pt=%ld paths:", tag, n, bad_cb,
           bad_np, bad_col, bad_ss, bad_sp);
    for (int k = 0; k < 10; ++k) printf(" %ld", paths[k]);
    if (sizeof(T) == 4) printf(" flat=%ld (flagged %ld, max error %.3g m)", bad_flat, flag_flat, max_flat);
    printf("\n");
    return (bad_cb || bad_np || bad_col || bad_ss || bad_sp || bad_flat) ? 1 : 0;
}

int main(int argc, char** argv) {
    long n = argc > 1 ? atol(argv[1]) : 200000;
    uint64_t seed = argc > 2 ? strtoull(argv[2], nullptr, 10) : 1;
    int rc = run<float>(n, seed, "f32");
    rc |= run<double>(n, seed + 1, "f64");
    return rc;
}
