// Internal state of a context (shared by the C ABI layer rvs_capi.cpp and
// the request server rvs_server.cpp).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <initializer_list>
#include <map>
#include <utility>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rvs.h"
#include "rvs_internal.hpp"

namespace rvs {

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        size_t want = std::max<size_t>(bytes, 1 << 20);
        hipError_t e = hipMalloc(&p, want);
        if (e == hipSuccess) cap = want;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

// Pinned host staging for the per-call tables of the edge search (copied to
// the device asynchronously; the call synchronizes before it returns).
struct PinnedBuf {
    void* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap) return hipSuccess;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        size_t want = std::max<size_t>(bytes, 1 << 20);
        hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
        if (e == hipSuccess) cap = want;
        return e;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
};

struct RobotRec {
    RobotDev dev{};
    std::vector<void*> allocs;
    int num_dofs = 0, num_links = 0, num_pairs = 0;
    std::vector<float> radius;   // capsule radius per link (host copy)
};

struct EnvRec {
    EnvDev dev{};
    std::vector<void*> allocs;
    std::vector<rvs_box> boxes;
    std::vector<float> aabb, obb, cull;   // host copies of the device tables
    std::vector<int32_t> gorder;          // broad-phase walk order (host copy)
    std::vector<float> gcull;             // group boxes of the walk (host staging)
    int semantics = RVS_ENV_AABB;
    bool has_motion = false;              // rvs_env_set_motion_law was called
    EnvStepDev step{};                    // motion-law state on the device
};

}  // namespace rvs

// The opaque handle of include/rvs.h.
struct rvs_context {
    rvs_options opt{};
    hipStream_t stream = nullptr;
    std::mutex mu;
    // the last failure's message; err_mu guards it alone, since argument
    // checks fail before mu is taken
    mutable std::mutex err_mu;
    std::string err;
    std::vector<rvs::RobotRec> robots;
    std::vector<rvs::EnvRec> envs;
    rvs::DevBuf arena;
    rvs::PinnedBuf staging;
    int64_t last_refined = 0;
    // Asynchronous calls (rvs_*_async) enqueue on a caller's stream and
    // return; async_ev is recorded after each of them (each first waits for
    // the previous one, so the event covers them all).  Every other call
    // waits for it on the host before it touches the arena or the tables.
    hipEvent_t async_ev = nullptr;
    bool async_pending = false;
    void drain_async() {
        if (async_pending) {
            (void)hipEventSynchronize(async_ev);
            async_pending = false;
        }
    }
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    float stage_ms[3] = {0, 0, 0};
    int32_t nstages = 0;
    int32_t num_cus = 0;
    // FP32 bulk pass variant when no tuning is recorded: -1 = by walk length
    // (lane refill when links x obstacles >= 1024), 0 = one lane per
    // configuration, 1 = lane refill
    int32_t fast_variant = -1;
    // reference-arithmetic distance batches up to this many configurations
    // run one wave per configuration (latency) instead of one lane
    int64_t wave_max = 16384;
    int32_t edge_waves_per_cu = 16;   // persistent first-collision kernel: resident waves per CU
    int32_t edge_spl = 1;             // ... and slots per lane in one 64-lane item
    std::map<std::pair<int32_t, int32_t>, rvs::FastTuning> tuning;   // (robot, env) -> autotuned launch shape
    // captured FP32 + refinement pipelines of device-array rvs_distance calls
    struct Graph {
        int32_t robot, env;
        const void *q, *d, *profile, *verdict, *arena;
        int64_t n;
        rvs::FastTuning t;
        hipGraphExec_t exec;
        uint64_t used;
    };
    std::vector<Graph> graphs;
    uint64_t graph_clock = 0;
    uint32_t* count_host = nullptr;   // pinned, coherent: refinement count the graphs' refinement stores
    uint32_t* count_dev = nullptr;    // the same word's device address
    std::mutex tuning_mu;
    rvs::FastTuning tuning_for(int32_t robot, int32_t env, int32_t links, int32_t objects) {
        std::lock_guard<std::mutex> lk(tuning_mu);
        auto it = tuning.find({robot, env});
        if (it != tuning.end()) return it->second;
        const int32_t v = fast_variant >= 0 ? fast_variant : ((int64_t)links * objects >= 1024 ? 1 : 0);
        // one lane per configuration: 64-thread workgroups measured fastest
        return rvs::FastTuning{v, v == 0 ? 64 : 256, 0, 5};
    }
};

namespace rvs {

// The context's lock for a synchronous call: it also waits for the
// asynchronous calls still in flight, which may use the arena and read the
// robot and obstacle tables.
struct CtxLock {
    std::lock_guard<std::mutex> lk;
    explicit CtxLock(rvs_context* c) : lk(c->mu) { c->drain_async(); }
};

// Sub-allocation of one device block, 256-byte aligned pieces.
struct Carve {
    char* base;
    size_t off = 0;
    explicit Carve(void* b) : base(static_cast<char*>(b)) {}
    template <typename T> T* take(size_t count) {
        off = (off + 255) & ~size_t(255);
        T* p = reinterpret_cast<T*>(base + off);
        off += count * sizeof(T);
        return p;
    }
    static size_t need(std::initializer_list<size_t> bytes) {
        size_t s = 0;
        for (size_t b : bytes) s += ((b + 255) & ~size_t(255));
        return s + 256;
    }
};


// FCL's world AABB of a box (and whether its rotation counts as identity)
// and the oriented-box table of the engine.
bool box_identity(const rvs_box& b);
void box_bounds(const rvs_box& b, float out[6]);
void box_obb(const rvs_box& b, float out[15]);
// Half extents of the tight world AABB of an oriented box (|R| h, rounded up
// by one part in 10^6 so that it contains the box despite float rounding).
void obb_extent(const float obb[15], float ext[3]);
QueryParams params_of(const rvs_context* c);

}  // namespace rvs
