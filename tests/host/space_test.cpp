// Drives rvs::RealVectorSpaceGPU (include/rvs_space.hpp) through its
// single-state RealVectorSpace methods on a problem written by
// tests/test_host_space.py, and writes the results for comparison with the
// batched C ABI and the oracle.  Also checks the State caching rules of
// computeDistance / computeDistanceUnderestimation.
//
//   space_test <in.bin> <out.bin>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "rvs_space.hpp"

namespace {

template <typename T> void rd(FILE* f, T* p, size_t n) {
    if (n && fread(p, sizeof(T), n, f) != n) { std::fprintf(stderr, "short read\n"); std::exit(2); }
}
template <typename T> void wr(FILE* f, const T* p, size_t n) { if (n) fwrite(p, sizeof(T), n, f); }

int fails = 0;
void expect(bool ok, const char* what) {
    if (!ok) { std::fprintf(stderr, "FAILED: %s\n", what); ++fails; }
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) { std::fprintf(stderr, "usage: space_test in.bin out.bin\n"); return 2; }
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) { std::perror("open input"); return 2; }
    int32_t h[8];
    rd(f, h, 8);
    const int32_t kind = h[0], nj = h[1], L = h[2], gi = h[3], npairs = h[4], M = h[5], sem = h[6], D = h[7];
    float grip;
    rd(f, &grip, 1);
    int64_t n;
    rd(f, &n, 1);
    std::vector<rvs_joint> joints(nj);
    rd(f, joints.data(), (size_t)nj);
    std::vector<float> radius(L);
    rd(f, radius.data(), (size_t)L);
    std::vector<int32_t> pairs(2 * (size_t)npairs);
    rd(f, pairs.data(), pairs.size());
    std::vector<rvs_box> boxes(M);
    rd(f, boxes.data(), (size_t)M);
    std::vector<float> q((size_t)n * D);
    rd(f, q.data(), q.size());
    auto np_ref = std::make_shared<std::vector<float>>((size_t)M * L * 6);
    rd(f, np_ref->data(), np_ref->size());
    std::fclose(f);

    rvs_robot_desc desc{kind, nj, joints.data(), L, radius.data(), gi, grip, npairs, pairs.data()};
    rvs::RealVectorSpaceGPU ss(desc, boxes, sem);
    expect(ss.getNumDimensions() == D && ss.getNumLinks() == L && ss.getNumObjects() == M, "dimensions");

    std::vector<uint8_t> valid(n), has_np(n), self_col(n), edge(n), under_stored(n);
    std::vector<float> d(n), prof((size_t)n * L), under(n), np0((size_t)M * L * 6, 0.0f);
    for (int64_t k = 0; k < n; ++k) {
        auto s = std::make_shared<rvs::State>(std::vector<float>(q.begin() + k * D, q.begin() + (k + 1) * D));
        valid[k] = ss.isValid(s) ? 1 : 0;
        d[k] = ss.computeDistance(s);
        expect(s->getIsRealDistance() && s->getDistance() == d[k], "computeDistance stores d_c, is_real");
        for (int i = 0; i < L; ++i) prof[k * L + i] = s->getDistanceProfile()[i];
        has_np[k] = s->getNearestPoints() ? 1 : 0;
        expect((d[k] > 0) == (has_np[k] == 1), "nearest points kept exactly for collision-free states");
        if (k == 0 && has_np[k]) np0 = *s->getNearestPoints();
        // cache: a real positive distance is returned without recomputation
        if (d[k] > 0) {
            s->setDistance(123.0f);
            expect(ss.computeDistance(s) == 123.0f, "computeDistance returns the cached real distance");
            expect(ss.computeDistanceUnderestimation(s, np_ref) == 123.0f, "underestimation returns cached real distance");
            expect(ss.computeDistance(s, true) == d[k], "compute_again recomputes");
        }
        // underestimation on a fresh state
        auto u = std::make_shared<rvs::State>(s->getCoord());
        under[k] = ss.computeDistanceUnderestimation(u, np_ref);
        under_stored[k] = u->getDistance() != -1.0f ? 1 : 0;
        if (under_stored[k]) expect(!u->getIsRealDistance() && u->getDistance() == under[k], "underestimation stores is_real=false");
        else expect(under[k] == 0.0f, "underestimation returns 0 when not stored");
        self_col[k] = ss.checkSelfCollision(s) ? 1 : 0;
        auto s2 = std::make_shared<rvs::State>(
            std::vector<float>(q.begin() + ((k + 1) % n) * D, q.begin() + ((k + 1) % n + 1) * D));
        edge[k] = ss.isValid(s, s2) ? 1 : 0;
    }
    // the batched overloads agree with the single-state calls
    std::vector<float> bd(n), bprof((size_t)n * L);
    std::vector<uint8_t> bver(n), bvalid(n);
    ss.computeDistance(q.data(), n, bd.data(), bprof.data(), nullptr, bver.data(), RVS_F_PREC(RVS_PREC_REF32));
    ss.isValid(q.data(), n, bvalid.data());
    expect(std::memcmp(bd.data(), d.data(), n * sizeof(float)) == 0, "batch distance == single-state distance");
    expect(std::memcmp(bprof.data(), prof.data(), prof.size() * sizeof(float)) == 0, "batch profile == single-state");
    expect(std::memcmp(bvalid.data(), valid.data(), (size_t)n) == 0, "batch isValid == single-state isValid");
    // bubble routines: single-state methods agree with the batched calls
    if (L == D) {
        std::vector<float> R1, Rb((size_t)n * D * (D + 1));
        ss.computeEnclosingRadii(q.data(), n, Rb.data());
        std::vector<float> qe((size_t)n * D), qn_b((size_t)n * D);
        std::vector<int32_t> st_b(n);
        for (int64_t k = 0; k < n; ++k)
            for (int i = 0; i < D; ++i) qe[k * D + i] = q[((k + 1) % n) * D + i];
        ss.extendSpines(q.data(), bd.data(), bprof.data(), qe.data(), n, 1, true, qn_b.data(), st_b.data());
        bool radii_ok = true, spine_ok = true;
        for (int64_t k = 0; k < n; ++k) {
            auto s = std::make_shared<rvs::State>(std::vector<float>(q.begin() + k * D, q.begin() + (k + 1) * D));
            R1 = ss.computeEnclosingRadii(s);
            radii_ok = radii_ok && std::memcmp(R1.data(), Rb.data() + k * D * (D + 1), R1.size() * 4) == 0;
            auto e = std::make_shared<rvs::State>(std::vector<float>(qe.begin() + k * D, qe.begin() + (k + 1) * D));
            auto res = ss.extendSpine(s, e, 1, true);   // computes s's distance first
            spine_ok = spine_ok && res.first == st_b[k] &&
                       std::memcmp(res.second->getCoord().data(), qn_b.data() + k * D, D * 4) == 0;
            if (res.first == RVS_SPINE_TRAPPED) spine_ok = spine_ok && res.second == s;
        }
        expect(radii_ok, "computeEnclosingRadii(state) == batched radii");
        expect(spine_ok, "extendSpine(state) == batched spines");
    }
    if (kind == RVS_SKEL_XARM6 && D == 6) {
        std::vector<uint8_t> col(n);
        std::vector<float> q2o((size_t)n * D);
        std::vector<float> q2((size_t)n * D);
        for (int64_t k = 0; k < n; ++k)
            for (int i = 0; i < D; ++i) q2[k * D + i] = q[((k + 1) % n) * D + i];
        ss.checkSelfCollision(q.data(), q2.data(), n, 0.1f, col.data(), q2o.data());
        bool ok = true;
        for (int64_t k = 0; k < n; ++k) {
            auto a = std::make_shared<rvs::State>(std::vector<float>(q.begin() + k * D, q.begin() + (k + 1) * D));
            auto b = std::make_shared<rvs::State>(std::vector<float>(q2.begin() + k * D, q2.begin() + (k + 1) * D));
            const bool hit = ss.checkSelfCollision(a, b, 0.1f);
            ok = ok && hit == (col[k] == 1) && std::memcmp(b->getCoord().data(), q2o.data() + k * D, D * 4) == 0;
        }
        expect(ok, "checkSelfCollision(q1, q2) == batched edge self-collision");
    }
    // errors surface as rvs::Error
    bool threw = false;
    try {
        ss.isValid(std::make_shared<rvs::State>(std::vector<float>(D + 1, 0.0f)));
    } catch (const rvs::Error& e) {
        threw = e.status == RVS_EINVAL;
    }
    expect(threw, "wrong state dimension raises rvs::Error(RVS_EINVAL)");

    FILE* o = std::fopen(argv[2], "wb");
    if (!o) { std::perror("open output"); return 2; }
    wr(o, valid.data(), valid.size());
    wr(o, d.data(), d.size());
    wr(o, prof.data(), prof.size());
    wr(o, has_np.data(), has_np.size());
    wr(o, under.data(), under.size());
    wr(o, under_stored.data(), under_stored.size());
    wr(o, self_col.data(), self_col.size());
    wr(o, edge.data(), edge.size());
    wr(o, np0.data(), np0.size());
    std::fclose(o);
    std::printf("space_test: %lld states, %d failed checks\n", (long long)n, fails);
    return fails ? 1 : 0;
}
