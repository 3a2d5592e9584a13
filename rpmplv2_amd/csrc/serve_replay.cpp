// Replays a recorded request stream against the request server in real time
// and reports latency and throughput (the serving workload of the engine).
//
//   rvs_serve_replay <problem.bin>
//
// The problem file (written by rpmplv2_amd/serving.py) holds the robots, the
// obstacle set with per-box velocities, one configuration pool per robot,
// the request stream (arrival time, robot, kind, size, pool offset) and the
// server options.  P planner threads take requests i = p, p + P, ... in
// arrival order; each sleeps until its request's arrival time, submits it
// with pinned host buffers, and collects finished requests.  Latency is the
// server's completion time minus the scheduled arrival time, both on the
// server's clock, so a planner that submits late is charged for it.
// Prints one JSON line.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <string>
#include <thread>
#include <vector>

#include "rvs.h"

namespace {

FILE* g_in = nullptr;
template <typename T> void rd(T* p, size_t n) {
    if (n && std::fread(p, sizeof(T), n, g_in) != n) {
        std::fprintf(stderr, "rvs_serve_replay: short read\n");
        std::exit(2);
    }
}
template <typename T> T rd1() { T v; rd(&v, 1); return v; }

void die(const char* what, rvs_status s, const char* msg) {
    std::fprintf(stderr, "rvs_serve_replay: %s failed (%d): %s\n", what, (int)s, msg ? msg : "");
    std::exit(3);
}

struct RobotIn {
    int32_t kind, nj, L, gi, npairs, dofs;
    float grip;
    std::vector<rvs_joint> joints;
    std::vector<float> radius;
    std::vector<int32_t> pairs;
    int64_t pool_n;
    float* pool = nullptr;   // pinned
    int32_t id = -1;
};

struct Done {
    double submit_s, admit_s, done_s;
    int32_t batch, status;
};

double pct(std::vector<double> v, double p) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = p * (double)(v.size() - 1);
    const size_t i = (size_t)pos;
    const double f = pos - (double)i;
    return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: rvs_serve_replay problem.bin\n");
        return 2;
    }
    g_in = std::fopen(argv[1], "rb");
    if (!g_in) {
        std::perror("open problem");
        return 2;
    }
    char magic[4];
    rd(magic, 4);
    if (std::memcmp(magic, "RVSW", 4) != 0 || rd1<int32_t>() != 1) {
        std::fprintf(stderr, "rvs_serve_replay: not a version-1 problem file\n");
        return 2;
    }
    const int32_t R = rd1<int32_t>();
    std::vector<RobotIn> robots(R);
    for (auto& r : robots) {
        r.kind = rd1<int32_t>(); r.nj = rd1<int32_t>(); r.L = rd1<int32_t>(); r.gi = rd1<int32_t>();
        r.npairs = rd1<int32_t>(); r.dofs = rd1<int32_t>(); r.grip = rd1<float>();
        r.joints.resize(r.nj); rd(r.joints.data(), r.joints.size());
        r.radius.resize(r.L); rd(r.radius.data(), r.radius.size());
        r.pairs.resize(2 * (size_t)r.npairs); rd(r.pairs.data(), r.pairs.size());
        r.pool_n = rd1<int64_t>();
        r.pool = static_cast<float*>(rvs_host_alloc(r.pool_n * r.dofs * 4));
        if (!r.pool) die("rvs_host_alloc", RVS_ENOMEM, "pool");
        rd(r.pool, (size_t)r.pool_n * r.dofs);
    }
    const int32_t M = rd1<int32_t>();
    std::vector<rvs_box> boxes(M);
    rd(boxes.data(), boxes.size());
    std::vector<float> vel(3 * (size_t)M);
    rd(vel.data(), vel.size());
    const float tick_seconds = rd1<float>();
    const int64_t K = rd1<int64_t>();
    std::vector<double> arrival(K);
    std::vector<int32_t> rob(K), kind(K);
    std::vector<int64_t> size(K), offset(K);
    rd(arrival.data(), K); rd(rob.data(), K); rd(kind.data(), K); rd(size.data(), K); rd(offset.data(), K);
    rvs_server_options so;
    rvs_server_default_options(&so);
    so.memory_budget = rd1<int64_t>();
    so.num_streams = rd1<int32_t>();
    so.max_batch_configs = rd1<int32_t>();
    so.tick_seconds = tick_seconds;
    const int32_t P = std::max(1, rd1<int32_t>());
    so.launches_per_stream = rd1<int32_t>();
    std::fclose(g_in);

    rvs_options o;
    rvs_default_options(&o);
    o.precision = RVS_PREC_FP32_REFINE;
    rvs_context* ctx = nullptr;
    rvs_status st = rvs_create(&o, &ctx);
    if (st != RVS_OK) die("rvs_create", st, "");
    for (auto& r : robots) {
        rvs_robot_desc d{r.kind, r.nj, r.joints.data(), r.L, r.radius.data(), r.gi, r.grip, r.npairs, r.pairs.data()};
        st = rvs_robot_load(ctx, &d, &r.id);
        if (st != RVS_OK) die("rvs_robot_load", st, rvs_last_error(ctx));
    }
    int32_t env = -1;
    st = rvs_env_create(ctx, boxes.data(), M, RVS_ENV_AABB, &env);
    if (st != RVS_OK) die("rvs_env_create", st, rvs_last_error(ctx));
    rvs_server* srv = nullptr;
    st = rvs_server_create(ctx, &so, &srv);
    if (st != RVS_OK) die("rvs_server_create", st, "");
    st = rvs_server_set_motion(srv, env, vel.data());
    if (st != RVS_OK) die("rvs_server_set_motion", st, rvs_server_last_error(srv));

    int32_t maxL = 0;
    for (auto& r : robots) maxL = std::max(maxL, r.L);
    const size_t per_cfg = 4 + 1 + 4 * (size_t)maxL;
    const size_t ring_bytes = std::max<size_t>((size_t)64 << 20, 4 * 65536 * per_cfg);

    // warm-up: one request per (robot, kind), so first-launch costs stay out
    {
        char* buf = static_cast<char*>(rvs_host_alloc((int64_t)(1024 * per_cfg + 1024)));
        for (auto& r : robots)
            for (int k = 0; k < 2; ++k) {
                rvs_request q{r.id, env, k, 0, 1024, r.pool, reinterpret_cast<float*>(buf),
                              k == RVS_REQ_UNDERESTIMATION ? reinterpret_cast<float*>(buf + 1024 * 5 + 64) : nullptr,
                              reinterpret_cast<uint8_t*>(buf + 4096)};
                uint64_t t;
                st = rvs_server_submit(srv, &q, &t);
                if (st != RVS_OK) die("warm-up submit", st, rvs_server_last_error(srv));
                st = rvs_server_wait(srv, t, -1, nullptr);
                if (st != RVS_OK) die("warm-up wait", st, rvs_server_last_error(srv));
            }
        rvs_host_free(buf);
    }

    std::vector<Done> done(K);
    std::vector<double> lag(K, 0.0);
    std::atomic<int> failures{0};
    const double start = rvs_server_now(srv) + 0.05;
    auto planner = [&](int p) {
        char* ring = static_cast<char*>(rvs_host_alloc((int64_t)ring_bytes));
        if (!ring) { failures++; return; }
        struct Out { uint64_t ticket; int64_t i; size_t off, len; };
        std::deque<Out> fifo;
        size_t head = 0;
        auto reap_front = [&](int32_t timeout_ms) {
            rvs_request_info info;
            const rvs_status s = rvs_server_wait(srv, fifo.front().ticket, timeout_ms, &info);
            if (s == RVS_EBUSY) return false;
            if (s != RVS_OK) failures++;
            done[fifo.front().i] = Done{info.submit_s, info.admit_s, info.done_s, info.batch_requests, info.status};
            fifo.pop_front();
            return true;
        };
        for (int64_t i = p; i < K; i += P) {
            const RobotIn& r = robots[rob[i]];
            // region layout: d [n] floats | flag [n] bytes | profile [n][L] floats (underestimation)
            const size_t n = (size_t)size[i];
            const size_t flag_off = (n * 4 + 255) & ~size_t(255);
            const size_t prof_off = (flag_off + n + 255) & ~size_t(255);
            const bool want_prof = kind[i] == RVS_REQ_UNDERESTIMATION;
            const size_t len = ((want_prof ? prof_off + n * 4 * (size_t)r.L : flag_off + n) + 255) & ~size_t(255);
            if (head + len > ring_bytes) head = 0;
            // the region [head, head + len) must be free: wait for older requests until none overlaps it
            for (;;) {
                bool overlap = false;
                for (const Out& f : fifo) overlap = overlap || (f.off < head + len && head < f.off + f.len);
                if (!overlap) break;
                reap_front(-1);
            }
            // sleep until the arrival time (server clock), spin the last 200 us
            const double target = start + arrival[i];
            for (;;) {
                const double now = rvs_server_now(srv);
                if (now >= target) break;
                if (target - now > 3e-4) std::this_thread::sleep_for(std::chrono::duration<double>(target - now - 2e-4));
            }
            char* b = ring + head;
            float* d = reinterpret_cast<float*>(b);
            uint8_t* flag = reinterpret_cast<uint8_t*>(b + flag_off);
            float* prof = want_prof ? reinterpret_cast<float*>(b + prof_off) : nullptr;
            rvs_request q{r.id, env, kind[i], 0, size[i], r.pool + offset[i] * r.dofs, d, prof, flag};
            uint64_t t = 0;
            if (rvs_server_submit(srv, &q, &t) != RVS_OK) { failures++; continue; }
            lag[i] = rvs_server_now(srv) - target;
            fifo.push_back(Out{t, i, head, len});
            head += len;
            while (!fifo.empty() && reap_front(0)) {
            }
        }
        while (!fifo.empty()) reap_front(-1);
        rvs_host_free(ring);
    };
    std::vector<std::thread> th;
    for (int p = 0; p < P; ++p) th.emplace_back(planner, p);
    for (auto& t : th) t.join();

    rvs_server_stats stats;
    rvs_server_get_stats(srv, &stats);
    std::vector<double> lat(K), lat_small, lat_large, qwait(K), batch(K);
    double last = start;
    int64_t configs = 0;
    for (int64_t i = 0; i < K; ++i) {
        lat[i] = (done[i].done_s - (start + arrival[i])) * 1e3;
        qwait[i] = (done[i].admit_s - (start + arrival[i])) * 1e3;
        batch[i] = done[i].batch;
        (size[i] <= 1024 ? lat_small : lat_large).push_back(lat[i]);
        last = std::max(last, done[i].done_s);
        configs += size[i];
    }
    const double span = last - start;
    double mean = 0;
    for (double v : lat) mean += v;
    mean /= std::max<int64_t>(1, K);
    std::printf(
        "{\"requests\": %lld, \"configs\": %lld, \"duration_s\": %.4f, \"offered_requests_per_s\": %.1f, "
        "\"throughput_configs_per_s\": %.1f, \"latency_ms\": {\"p50\": %.4f, \"p90\": %.4f, \"p99\": %.4f, "
        "\"p999\": %.4f, \"max\": %.4f, \"mean\": %.4f}, \"latency_ms_small_le1024\": {\"p50\": %.4f, \"p99\": %.4f}, "
        "\"latency_ms_large\": {\"p50\": %.4f, \"p99\": %.4f}, \"queue_wait_ms\": {\"p50\": %.4f, \"p99\": %.4f}, "
        "\"requests_per_launch\": {\"mean\": %.2f, \"p99\": %.1f}, \"submit_lag_ms_p99\": %.4f, \"launches\": %lld, "
        "\"max_concurrent_launches\": %d, \"peak_device_bytes\": %lld, \"ticks\": %lld, \"planners\": %d, "
        "\"streams\": %d, \"launches_per_stream\": %d, \"max_batch_configs\": %d, \"memory_budget\": %lld, "
        "\"failures\": %d}\n",
        (long long)K, (long long)configs, span, K / std::max(1e-9, arrival.empty() ? 1.0 : arrival.back()),
        configs / std::max(1e-9, span), pct(lat, 0.5), pct(lat, 0.9), pct(lat, 0.99), pct(lat, 0.999),
        pct(lat, 1.0), mean, pct(lat_small, 0.5), pct(lat_small, 0.99), pct(lat_large, 0.5), pct(lat_large, 0.99),
        pct(qwait, 0.5), pct(qwait, 0.99), [&] { double s = 0; for (double v : batch) s += v; return s / std::max<int64_t>(1, K); }(),
        pct(batch, 0.99), pct([&] { std::vector<double> v(lag.begin(), lag.end()); for (auto& x : v) x *= 1e3; return v; }(), 0.99),
        (long long)stats.batches, stats.max_inflight_batches, (long long)stats.peak_bytes, (long long)stats.ticks, P,
        so.num_streams, so.launches_per_stream, so.max_batch_configs, (long long)so.memory_budget, failures.load());
    rvs_server_destroy(srv);
    for (auto& r : robots) rvs_host_free(r.pool);
    rvs_destroy(ctx);
    return failures.load() ? 1 : 0;
}
