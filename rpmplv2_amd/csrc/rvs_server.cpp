// Request server of the C ABI (rvs_server_*): continuous batching of planner
// requests on one GPU.
//
// Submitters append requests to an arrival-ordered queue.  One server thread
//   - moves obstacles: once per tick it writes the obstacle tables of the new
//     tick into a free slot of a small ring of device copies (the launches
//     already running keep reading their slot), and later launches wait on
//     that update with a device-side event;
//   - retires launches whose completion event has fired: their requests are
//     marked done and their device memory returns to the pool;
//   - admits: the oldest waiting request picks the (robot, environment, kind)
//     of the next launch, every waiting request of the same key joins it (in
//     arrival order, up to max_batch_configs), the launch's state is carved
//     from one device memory pool of memory_budget bytes, and it goes to the
//     least loaded of num_streams streams.  A launch that does not fit the pool
//     waits until earlier ones retire; later arrivals never overtake the
//     oldest request's key.
// Each configuration is computed by the same kernels as a direct call, so
// its result does not depend on what it was batched with.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "rvs_context.hpp"
#include "rvs_internal.hpp"
#include "rvs_stage.hpp"

using namespace rvs;

namespace {

using Clock = std::chrono::steady_clock;
constexpr size_t kGranule = 256;
constexpr int kMotionSlots = 4;
// Requests of at most this many configurations travel through the server's
// pinned staging memory: one copy per launch each way instead of one per
// request and array (large requests are copied directly).
constexpr int64_t kStageMax = 16384;
constexpr size_t kStageBytes = (size_t)256 << 20;

inline size_t up(size_t b) { return (b + kGranule - 1) & ~(kGranule - 1); }

// First-fit sub-allocator over one device block, with coalescing on release.
// (device memory, or pinned host memory for the staging pool)
class Pool {
  public:
    hipError_t init(size_t cap, bool host = false) {
        cap_ = up(cap);
        host_ = host;
        hipError_t e = host ? hipHostMalloc(reinterpret_cast<void**>(&base_), cap_, hipHostMallocDefault)
                            : hipMalloc(reinterpret_cast<void**>(&base_), cap_);
        if (e == hipSuccess) free_[0] = cap_;
        return e;
    }
    void destroy() {
        if (base_) (void)(host_ ? hipHostFree(base_) : hipFree(base_));
        base_ = nullptr;
    }
    bool alloc(size_t bytes, size_t* off) {
        bytes = up(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            const size_t o = it->first, rest = it->second - bytes;
            free_.erase(it);
            if (rest) free_[o + bytes] = rest;
            *off = o;
            used_ += bytes;
            return true;
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = up(bytes);
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    char* base() const { return base_; }
    size_t capacity() const { return cap_; }
    size_t used() const { return used_; }

  private:
    char* base_ = nullptr;
    size_t cap_ = 0, used_ = 0;
    bool host_ = false;
    std::map<size_t, size_t> free_;   // offset -> size
};

struct Req {
    uint64_t ticket = 0;
    rvs_request r{};
    RobotDev rb{};
    EnvDev env{};
    int dofs = 0, links = 0;
    double submit_s = 0, admit_s = 0, done_s = 0;
    int64_t tick = -1;
    int32_t batch = 0;
    int32_t status = RVS_OK;
    bool done = false;
};

struct Motion {
    MotionDev dev{};
    std::vector<void*> allocs;
    const int32_t* gorder = nullptr;   // the environment's broad-phase walk order
    bool use_obb = false;
    struct Slot {
        float* aabb = nullptr;
        float* obb = nullptr;
        float* cull = nullptr;   // broad-phase bounds of oriented boxes
        float* gcull = nullptr;  // group boxes of the broad-phase walk
        int64_t tick = -1;
        int refs = 0;
        hipEvent_t ready = nullptr;
    };
    Slot slot[kMotionSlots];
    int current = -1;
};

struct Batch {
    std::vector<std::shared_ptr<Req>> reqs;
    size_t off = 0, bytes = 0;
    int stream = 0;
    hipEvent_t done = nullptr;
    Motion* motion = nullptr;
    int slot = -1;
    std::vector<int64_t> seg_host;   // source of an asynchronous copy: lives until retirement
    // staging: the first n_staged requests (batch order) go through pinned
    // host memory [hoff, hoff + hbytes) of the staging pool
    size_t n_staged = 0;
    int64_t N_staged = 0;
    size_t hoff = 0, hbytes = 0;
    bool staged = false;
    int L = 0;
    bool prof = false;
};

// Device state of one launch, carved from its pool block.
struct Layout {
    int64_t N = 0;
    int nreq = 0;
    bool prof = false, under = false;
    int D = 0, L = 0, M = 0;
    size_t need() const {
        return Carve::need({(size_t)N * D * 4, (size_t)N * 4, (size_t)N, prof ? (size_t)N * L * 4 : 0,
                            under ? 0 : (size_t)N * 4, 64 * 4,
                            under ? (size_t)nreq * D * 4 : 0, under ? (size_t)nreq * M * L * 24 : 0,
                            under ? (size_t)nreq * 4 : 0, under ? (size_t)nreq : 0,
                            under ? (size_t)(nreq + 1) * 8 : 0});
    }
};

}  // namespace

struct rvs_server {
    rvs_context* ctx = nullptr;
    rvs_server_options opt{};
    Pool pool;
    Pool stage;                                   // pinned host staging (server thread only)
    bool have_stage = false;
    std::vector<hipStream_t> streams;
    std::vector<int> load;
    hipStream_t upd = nullptr;
    std::thread th;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    std::string err;
    std::deque<std::shared_ptr<Req>> pending;
    std::unordered_map<uint64_t, std::shared_ptr<Req>> live;
    std::vector<Batch> inflight;                  // server thread only
    std::vector<hipEvent_t> spare;                // server thread only
    std::map<int32_t, std::unique_ptr<Motion>> motion;
    uint64_t next_ticket = 1;
    Clock::time_point t0;
    rvs_server_stats stats{};
    QueryParams qp{};
    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

void finish(rvs_server* s, Batch& b, int32_t status) {
    const double t = s->now();
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto& r : b.reqs) {
        r->done = true;
        r->done_s = t;
        r->status = status;
        s->stats.completed++;
        s->stats.configs_completed += r->r.n;
    }
    s->cv_done.notify_all();
}

// Host staging layout of a launch (pinned): staged inputs, first
// configurations and segment offsets (underestimation), staged outputs.
struct StageLayout {
    size_t q, qf, seg, d, flag, prof, total;
    StageLayout(int64_t Ns, int nreq, int D, int L, bool under, bool prof_) {
        size_t o = 0;
        auto put = [&](size_t bytes) { const size_t at = o; o = up(o + bytes); return at; };
        q = put((size_t)Ns * D * 4);
        qf = put(under ? (size_t)nreq * D * 4 : 0);
        seg = put(under ? (size_t)(nreq + 1) * 8 : 0);
        d = put((size_t)Ns * 4);
        flag = put((size_t)Ns);
        prof = put(prof_ ? (size_t)Ns * L * 4 : 0);
        total = o;
    }
};

// Issue one launch on its stream.  Returns the first HIP error.
hipError_t issue(rvs_server* s, Batch& b, const Layout& lay) {
    hipStream_t st = s->streams[b.stream];
    const Req& h = *b.reqs.front();
    const RobotDev& rb = h.rb;
    EnvDev env = h.env;
    hipError_t e = hipSuccess;
    if (b.motion) {
        Motion::Slot& sl = b.motion->slot[b.slot];
        env.aabb = sl.aabb;
        env.obb = sl.obb;
        env.cull = env.use_obb ? sl.cull : sl.aabb;
        env.gcull = sl.gcull;
        e = hipStreamWaitEvent(st, sl.ready, 0);
    }
    Carve cv(s->pool.base() + b.off);
    const int64_t N = lay.N;
    const int D = lay.D, L = lay.L, M = lay.M;
    float* q = cv.take<float>((size_t)N * D);
    float* d = cv.take<float>(N);
    uint8_t* flag = cv.take<uint8_t>(N);
    float* prof = lay.prof ? cv.take<float>((size_t)N * L) : nullptr;
    int32_t* idx = lay.under ? nullptr : cv.take<int32_t>(N);
    uint32_t* count = cv.take<uint32_t>(64);
    float* qf = lay.under ? cv.take<float>((size_t)lay.nreq * D) : nullptr;
    float* nps = lay.under ? cv.take<float>((size_t)lay.nreq * M * L * 6) : nullptr;
    float* df = lay.under ? cv.take<float>(lay.nreq) : nullptr;
    uint8_t* vf = lay.under ? cv.take<uint8_t>(lay.nreq) : nullptr;
    int64_t* seg = lay.under ? cv.take<int64_t>(lay.nreq + 1) : nullptr;
    // inputs: staged requests (a prefix of the batch) in one copy, the rest directly
    const StageLayout sl(b.N_staged, lay.nreq, D, L, lay.under, lay.prof);
    char* hs = b.staged ? s->stage.base() + b.hoff : nullptr;
    int64_t off = 0;
    for (size_t k = 0; k < b.reqs.size(); ++k) {
        const rvs_request& r = b.reqs[k]->r;
        if (b.staged && k < b.n_staged) {
            std::memcpy(hs + sl.q + (size_t)off * D * 4, r.q, (size_t)r.n * D * 4);
            if (lay.under) std::memcpy(hs + sl.qf + k * D * 4, r.q, (size_t)D * 4);
        } else {
            if (e == hipSuccess && r.n)
                e = hipMemcpyAsync(q + off * D, r.q, (size_t)r.n * D * 4, hipMemcpyHostToDevice, st);
            if (lay.under && b.staged) std::memcpy(hs + sl.qf + k * D * 4, r.q, (size_t)D * 4);
        }
        off += r.n;
    }
    if (b.staged && b.N_staged && e == hipSuccess)
        e = hipMemcpyAsync(q, hs + sl.q, (size_t)b.N_staged * D * 4, hipMemcpyHostToDevice, st);
    if (!lay.under) {
        DistOut out{};
        out.d = d;
        out.profile = prof;
        out.verdict = flag;
        const int prec = s->opt.precision;
        if (prec == RVS_PREC_FP32_REFINE) {
            const FastTuning ft = s->ctx->tuning_for(h.r.robot, h.r.env, L, M);
            if (e == hipSuccess) e = launch_zero_u32(count, 64, st);
            if (e == hipSuccess)
                // order given; no links scratch (a link-major tuning runs as variant 3)
                e = launch_distance_fast(rb, env, s->qp, q, N, out, idx, count, ft, nullptr, nullptr, s->ctx->num_cus,
                                         st);
            if (e == hipSuccess) e = launch_distance_refine(rb, env, s->qp, q, N, out, idx, count, st);
        } else if (e == hipSuccess) {
            e = launch_distance_ref32_auto(rb, env, s->qp, q, N, out, s->ctx->wave_max, st);
        }
    } else {
        // nearest points of every request's first configuration, then the
        // underestimation of all configurations against their request's set
        if (b.staged) {
            std::memcpy(hs + sl.seg, b.seg_host.data(), b.seg_host.size() * 8);
            if (e == hipSuccess)
                e = hipMemcpyAsync(qf, hs + sl.qf, (size_t)lay.nreq * D * 4, hipMemcpyHostToDevice, st);
            if (e == hipSuccess)
                e = hipMemcpyAsync(seg, hs + sl.seg, b.seg_host.size() * 8, hipMemcpyHostToDevice, st);
        } else {
            off = 0;
            for (size_t k = 0; k < b.reqs.size(); ++k) {
                if (e == hipSuccess)
                    e = hipMemcpyAsync(qf + k * D, q + off * D, (size_t)D * 4, hipMemcpyDeviceToDevice, st);
                off += b.reqs[k]->r.n;
            }
            if (e == hipSuccess)
                e = hipMemcpyAsync(seg, b.seg_host.data(), b.seg_host.size() * 8, hipMemcpyHostToDevice, st);
        }
        DistOut o1{};
        o1.d = df;
        o1.nearest = nps;
        o1.verdict = vf;
        if (e == hipSuccess) e = launch_distance_ref32_auto(rb, env, s->qp, qf, lay.nreq, o1, s->ctx->wave_max, st);
        if (e == hipSuccess)
            e = launch_underestimation_seg(rb, env, s->qp, q, N, nps, seg, lay.nreq, d, prof, flag, st);
    }
    // outputs: the staged prefix into staging (scattered at retirement), the rest directly
    if (b.staged && b.N_staged) {
        if (e == hipSuccess) e = hipMemcpyAsync(hs + sl.d, d, (size_t)b.N_staged * 4, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipMemcpyAsync(hs + sl.flag, flag, (size_t)b.N_staged, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess && lay.prof)
            e = hipMemcpyAsync(hs + sl.prof, prof, (size_t)b.N_staged * L * 4, hipMemcpyDeviceToHost, st);
    }
    off = 0;
    for (size_t k = 0; k < b.reqs.size(); ++k) {
        const rvs_request& r = b.reqs[k]->r;
        const int64_t n = r.n;
        if (n && !(b.staged && k < b.n_staged)) {
            if (e == hipSuccess) e = hipMemcpyAsync(r.d, d + off, (size_t)n * 4, hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipMemcpyAsync(r.flag, flag + off, (size_t)n, hipMemcpyDeviceToHost, st);
            if (e == hipSuccess && r.profile)
                e = hipMemcpyAsync(r.profile, prof + off * L, (size_t)n * L * 4, hipMemcpyDeviceToHost, st);
        }
        off += n;
    }
    if (e == hipSuccess) e = hipEventRecord(b.done, st);
    return e;
}

// Copies the staged outputs of a finished launch to the requests' buffers.
void scatter(rvs_server* s, Batch& b) {
    if (!b.staged || !b.N_staged) return;
    const StageLayout sl(b.N_staged, (int)b.reqs.size(), b.reqs.front()->dofs, b.L, b.reqs.front()->r.kind ==
                         RVS_REQ_UNDERESTIMATION, b.prof);
    const char* hs = s->stage.base() + b.hoff;
    int64_t off = 0;
    for (size_t k = 0; k < b.n_staged; ++k) {
        const rvs_request& r = b.reqs[k]->r;
        std::memcpy(r.d, hs + sl.d + (size_t)off * 4, (size_t)r.n * 4);
        std::memcpy(r.flag, hs + sl.flag + (size_t)off, (size_t)r.n);
        if (r.profile) std::memcpy(r.profile, hs + sl.prof + (size_t)off * b.L * 4, (size_t)r.n * b.L * 4);
        off += r.n;
    }
}

// Called with s->mu held.
void tick_motion(rvs_server* s) {
    const double t = s->now();
    const int64_t tick = (int64_t)std::floor(t / (double)s->opt.tick_seconds);
    for (auto& kv : s->motion) {
        Motion& m = *kv.second;
        if (m.current >= 0 && m.slot[m.current].tick >= tick) continue;
        int k = -1;
        for (int i = 0; i < kMotionSlots; ++i)
            if (i != m.current && m.slot[i].refs == 0) { k = i; break; }
        if (k < 0) continue;   // every other slot still read by running launches: retry next loop
        Motion::Slot& sl = m.slot[k];
        hipError_t e = launch_env_motion(m.dev, tick, s->opt.tick_seconds, sl.aabb, sl.obb, sl.cull, s->upd);
        if (e == hipSuccess && m.dev.num_obj > 0)
            e = launch_group_bounds(m.use_obb ? sl.cull : sl.aabb, m.gorder, m.dev.num_obj, sl.gcull, s->upd);
        if (e == hipSuccess) e = hipEventRecord(sl.ready, s->upd);
        if (e != hipSuccess) {
            s->err = std::string("obstacle update: ") + hipGetErrorString(e);   // caller holds s->mu
            continue;
        }
        sl.tick = tick;
        m.current = k;
        s->stats.ticks++;
    }
}

bool retire(rvs_server* s) {
    bool any = false;
    for (size_t i = 0; i < s->inflight.size();) {
        Batch& b = s->inflight[i];
        const hipError_t q = hipEventQuery(b.done);
        if (q == hipErrorNotReady) { ++i; continue; }
        if (q == hipSuccess) scatter(s, b);
        finish(s, b, q == hipSuccess ? RVS_OK : RVS_EDEVICE);
        if (b.staged) s->stage.release(b.hoff, b.hbytes);
        if (q != hipSuccess) {
            std::lock_guard<std::mutex> lk(s->mu);
            s->err = std::string("launch failed: ") + hipGetErrorString(q);
        }
        s->pool.release(b.off, b.bytes);
        if (b.motion) b.motion->slot[b.slot].refs--;
        s->load[b.stream]--;
        s->spare.push_back(b.done);
        s->inflight.erase(s->inflight.begin() + (long)i);
        any = true;
    }
    return any;
}

// Picks the next launch from the queue (caller holds s->mu).  Returns false
// when nothing can be admitted now.
bool form_batch(rvs_server* s, Batch& b, Layout& lay) {
    if (s->pending.empty()) return false;
    const auto head = s->pending.front();
    const rvs_request& hr = head->r;
    lay = Layout{};
    lay.D = head->dofs;
    lay.L = head->links;
    lay.M = head->env.num_obj;
    lay.under = hr.kind == RVS_REQ_UNDERESTIMATION;
    std::vector<size_t> take;
    for (size_t k = 0; k < s->pending.size(); ++k) {
        const auto& r = s->pending[k];
        if (r->r.robot != hr.robot || r->r.env != hr.env || r->r.kind != hr.kind) continue;
        if (!take.empty() && lay.N + r->r.n > s->opt.max_batch_configs) break;
        take.push_back(k);
        lay.N += r->r.n;
        lay.nreq++;
        lay.prof = lay.prof || r->r.profile != nullptr;
    }
    size_t off = 0;
    while (!s->pool.alloc(lay.need(), &off)) {
        if (take.size() == 1) return false;   // wait for memory
        // shrink to the older half and retry
        const size_t keep = take.size() / 2;
        take.resize(keep);
        lay.N = 0;
        lay.nreq = 0;
        lay.prof = false;
        for (size_t k : take) {
            lay.N += s->pending[k]->r.n;
            lay.nreq++;
            lay.prof = lay.prof || s->pending[k]->r.profile != nullptr;
        }
    }
    b = Batch{};
    b.off = off;
    b.bytes = lay.need();
    for (size_t i = take.size(); i-- > 0;) {
        b.reqs.insert(b.reqs.begin(), s->pending[take[i]]);
        s->pending.erase(s->pending.begin() + (long)take[i]);
    }
    // small requests first: they form the staged prefix of the launch
    std::stable_partition(b.reqs.begin(), b.reqs.end(),
                          [](const std::shared_ptr<Req>& r) { return r->r.n <= kStageMax; });
    b.L = lay.L;
    b.prof = lay.prof;
    if (s->have_stage) {
        for (auto& r : b.reqs) {
            if (r->r.n > kStageMax) break;
            b.n_staged++;
            b.N_staged += r->r.n;
        }
        const StageLayout sl(b.N_staged, lay.nreq, lay.D, lay.L, lay.under, lay.prof);
        if (sl.total && s->stage.alloc(sl.total, &b.hoff)) {
            b.staged = true;
            b.hbytes = sl.total;
        } else {
            b.n_staged = 0;
            b.N_staged = 0;
        }
    }
    if (lay.under) {
        b.seg_host.resize(b.reqs.size() + 1);
        int64_t acc = 0;
        for (size_t k = 0; k < b.reqs.size(); ++k) {
            b.seg_host[k] = acc;
            acc += b.reqs[k]->r.n;
        }
        b.seg_host[b.reqs.size()] = acc;
    }
    return true;
}

void serve(rvs_server* s) {
    (void)hipSetDevice(s->ctx->opt.device);
    const size_t max_inflight = (size_t)s->opt.num_streams * s->opt.launches_per_stream;
    for (;;) {
        bool progress = retire(s);
        std::unique_lock<std::mutex> lk(s->mu);
        tick_motion(s);
        if (s->stop && s->pending.empty() && s->inflight.empty()) break;
        while (s->inflight.size() < max_inflight) {
            Batch b;
            Layout lay;
            if (!form_batch(s, b, lay)) break;
            const double t = s->now();
            int sbest = 0;
            for (int k = 1; k < (int)s->streams.size(); ++k)
                if (s->load[k] < s->load[sbest]) sbest = k;
            b.stream = sbest;
            auto mit = s->motion.find(b.reqs.front()->r.env);
            if (mit != s->motion.end() && mit->second->current >= 0) {
                b.motion = mit->second.get();
                b.slot = b.motion->current;
                b.motion->slot[b.slot].refs++;
            }
            const int64_t tick = b.motion ? b.motion->slot[b.slot].tick : -1;
            for (auto& r : b.reqs) {
                r->admit_s = t;
                r->tick = tick;
                r->batch = (int32_t)b.reqs.size();
            }
            s->stats.peak_bytes = std::max<int64_t>(s->stats.peak_bytes, (int64_t)s->pool.used());
            lk.unlock();
            if (s->spare.empty()) {
                hipEvent_t ev = nullptr;
                (void)hipEventCreateWithFlags(&ev, hipEventDisableTiming);
                s->spare.push_back(ev);
            }
            b.done = s->spare.back();
            s->spare.pop_back();
            const hipError_t e = issue(s, b, lay);
            s->load[b.stream]++;
            s->inflight.push_back(std::move(b));
            lk.lock();
            s->stats.batches++;
            s->stats.max_inflight_batches = std::max<int32_t>(s->stats.max_inflight_batches, (int32_t)s->inflight.size());
            if (e != hipSuccess) {
                s->err = std::string("launch: ") + hipGetErrorString(e);
                // the completion event was not recorded: retire as failed now
                Batch& fb = s->inflight.back();
                lk.unlock();
                (void)hipStreamSynchronize(s->streams[fb.stream]);
                finish(s, fb, RVS_EDEVICE);
                s->pool.release(fb.off, fb.bytes);
                if (fb.staged) s->stage.release(fb.hoff, fb.hbytes);
                if (fb.motion) fb.motion->slot[fb.slot].refs--;
                s->load[fb.stream]--;
                s->spare.push_back(fb.done);
                s->inflight.pop_back();
                lk.lock();
            }
            progress = true;
        }
        if (!progress) {
            if (s->inflight.empty() && s->pending.empty() && s->motion.empty())
                s->cv_work.wait(lk, [&] { return s->stop || !s->pending.empty() || !s->motion.empty(); });
            else
                s->cv_work.wait_for(lk, std::chrono::microseconds(20));
        }
    }
}

rvs_status sfail(rvs_server* s, rvs_status st, const std::string& m) {
    if (s) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->err = m;
    }
    return st;
}

}  // namespace

extern "C" {

void rvs_server_default_options(rvs_server_options* o) {
    if (!o) return;
    o->memory_budget = (int64_t)8 << 30;
    o->num_streams = 2;
    o->max_batch_configs = 1 << 20;
    o->tick_seconds = 0.01f;
    o->precision = RVS_PREC_FP32_REFINE;
    o->launches_per_stream = 1;
}

rvs_status rvs_server_create(rvs_context* c, const rvs_server_options* opt, rvs_server** out) {
    if (!c || !out) return RVS_EINVAL;
    *out = nullptr;
    rvs_server_options o;
    if (opt) o = *opt;
    else rvs_server_default_options(&o);
    if (o.memory_budget < (1 << 20) || o.num_streams < 1 || o.num_streams > 16 || o.max_batch_configs < 1 ||
        o.launches_per_stream < 1 || o.launches_per_stream > 16 ||
        !(o.tick_seconds > 0) || (o.precision != RVS_PREC_FP32_REFINE && o.precision != RVS_PREC_REF32))
        return RVS_EINVAL;
    rvs_server* s = new (std::nothrow) rvs_server();
    if (!s) return RVS_ENOMEM;
    s->ctx = c;
    s->opt = o;
    {
        std::lock_guard<std::mutex> lk(c->mu);
        s->qp = params_of(c);
    }
    hipError_t e = hipSetDevice(c->opt.device);
    if (e == hipSuccess) e = s->pool.init((size_t)o.memory_budget);
    if (e == hipSuccess) s->have_stage = s->stage.init(kStageBytes, true) == hipSuccess;
    for (int k = 0; k < o.num_streams && e == hipSuccess; ++k) {
        hipStream_t st = nullptr;
        e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (e == hipSuccess) s->streams.push_back(st);
    }
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->upd, hipStreamNonBlocking);
    if (e != hipSuccess) {
        for (auto st : s->streams) (void)hipStreamDestroy(st);
        if (s->upd) (void)hipStreamDestroy(s->upd);
        s->pool.destroy();
        delete s;
        return e == hipErrorOutOfMemory ? RVS_ENOMEM : RVS_EDEVICE;
    }
    s->load.assign(s->streams.size(), 0);
    s->t0 = Clock::now();
    s->th = std::thread(serve, s);
    *out = s;
    return RVS_OK;
}

void rvs_server_destroy(rvs_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->th.joinable()) s->th.join();
    (void)hipSetDevice(s->ctx->opt.device);
    for (auto st : s->streams) (void)hipStreamSynchronize(st);
    if (s->upd) (void)hipStreamSynchronize(s->upd);
    for (auto ev : s->spare) (void)hipEventDestroy(ev);
    for (auto& kv : s->motion) {
        for (auto& sl : kv.second->slot)
            if (sl.ready) (void)hipEventDestroy(sl.ready);
        for (void* p : kv.second->allocs) (void)hipFree(p);
    }
    for (auto st : s->streams) (void)hipStreamDestroy(st);
    if (s->upd) (void)hipStreamDestroy(s->upd);
    s->pool.destroy();
    if (s->have_stage) s->stage.destroy();
    delete s;
}

rvs_status rvs_server_set_motion(rvs_server* s, int32_t env, const float* vel) {
    if (!s || !vel) return RVS_EINVAL;
    rvs_context* c = s->ctx;
    std::vector<rvs_box> boxes;
    const int32_t* gorder = nullptr;
    bool use_obb = false;
    {
        std::lock_guard<std::mutex> lk(c->mu);
        if (env < 0 || env >= (int32_t)c->envs.size()) return sfail(s, RVS_EINVAL, "rvs_server_set_motion: bad env id");
        boxes = c->envs[env].boxes;
        gorder = c->envs[env].dev.gorder;
        use_obb = c->envs[env].dev.use_obb != 0;
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->motion.count(env)) return sfail(s, RVS_ENOTSUP, "rvs_server_set_motion: motion already set");
    }
    const int M = (int)boxes.size();
    std::vector<float> pos0(3 * M), v(vel, vel + 3 * M), half(3 * M), rad(M), obb0(15 * (size_t)M), ext(3 * M);
    std::vector<uint8_t> ident(M);
    for (int j = 0; j < M; ++j) {
        box_obb(boxes[j], &obb0[15 * (size_t)j]);
        obb_extent(&obb0[15 * (size_t)j], &ext[3 * j]);
        for (int k = 0; k < 3; ++k) {
            pos0[3 * j + k] = boxes[j].pos[k];
            half[3 * j + k] = 0.5f * boxes[j].dim[k];
        }
        ident[j] = box_identity(boxes[j]) ? 1 : 0;   // the branch box_bounds takes
        const float m0 = -half[3 * j], m1 = -half[3 * j + 1], m2 = -half[3 * j + 2];
        rad[j] = std::sqrt(m0 * m0 + (m1 * m1 + m2 * m2));
    }
    auto m = std::make_unique<Motion>();
    m->gorder = gorder;
    m->use_obb = use_obb;
    (void)hipSetDevice(c->opt.device);
    hipError_t e = hipSuccess;
    auto put = [&](const void* src, size_t bytes, const void** dst) {
        void* p = nullptr;
        if (e == hipSuccess) e = hipMalloc(&p, std::max<size_t>(bytes, 4));
        if (e == hipSuccess) {
            m->allocs.push_back(p);
            e = hipMemcpy(p, src, bytes, hipMemcpyHostToDevice);
        }
        *dst = p;
    };
    m->dev.num_obj = M;
    put(pos0.data(), pos0.size() * 4, reinterpret_cast<const void**>(&m->dev.pos0));
    put(v.data(), v.size() * 4, reinterpret_cast<const void**>(&m->dev.vel));
    put(half.data(), half.size() * 4, reinterpret_cast<const void**>(&m->dev.half));
    put(rad.data(), rad.size() * 4, reinterpret_cast<const void**>(&m->dev.rad));
    put(ident.data(), ident.size(), reinterpret_cast<const void**>(&m->dev.ident));
    put(obb0.data(), obb0.size() * 4, reinterpret_cast<const void**>(&m->dev.obb0));
    put(ext.data(), ext.size() * 4, reinterpret_cast<const void**>(&m->dev.ext));
    for (auto& sl : m->slot) {
        void* a = nullptr;
        void* o = nullptr;
        if (e == hipSuccess) e = hipMalloc(&a, std::max<size_t>((size_t)M * 24, 4));
        if (e == hipSuccess) m->allocs.push_back(a);
        if (e == hipSuccess) e = hipMalloc(&o, std::max<size_t>((size_t)M * 60, 4));
        if (e == hipSuccess) m->allocs.push_back(o);
        void* cb = nullptr;
        if (e == hipSuccess) e = hipMalloc(&cb, std::max<size_t>((size_t)M * 24, 4));
        if (e == hipSuccess) m->allocs.push_back(cb);
        sl.cull = static_cast<float*>(cb);
        void* gb = nullptr;
        if (e == hipSuccess) e = hipMalloc(&gb, std::max<size_t>((size_t)num_groups(M) * 24, 4));
        if (e == hipSuccess) m->allocs.push_back(gb);
        sl.gcull = static_cast<float*>(gb);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.ready, hipEventDisableTiming);
        sl.aabb = static_cast<float*>(a);
        sl.obb = static_cast<float*>(o);
    }
    if (e != hipSuccess) {
        for (auto& sl : m->slot)
            if (sl.ready) (void)hipEventDestroy(sl.ready);
        for (void* p : m->allocs) (void)hipFree(p);
        return sfail(s, e == hipErrorOutOfMemory ? RVS_ENOMEM : RVS_EDEVICE,
                     std::string("rvs_server_set_motion: ") + hipGetErrorString(e));
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->motion[env] = std::move(m);
    }
    s->cv_work.notify_all();
    return RVS_OK;
}

rvs_status rvs_server_submit(rvs_server* s, const rvs_request* req, uint64_t* ticket) {
    if (!s || !req || !ticket) return RVS_EINVAL;
    rvs_context* c = s->ctx;
    auto r = std::make_shared<Req>();
    r->r = *req;
    if (req->n < 0 || (req->n > 0 && (!req->q || !req->d || !req->flag)))
        return sfail(s, RVS_EINVAL, "rvs_server_submit: bad request");
    if (req->kind != RVS_REQ_VERDICT && req->kind != RVS_REQ_UNDERESTIMATION)
        return sfail(s, RVS_EINVAL, "rvs_server_submit: kind");
    {
        std::lock_guard<std::mutex> lk(c->mu);
        if (req->robot < 0 || req->robot >= (int32_t)c->robots.size() || req->env < 0 ||
            req->env >= (int32_t)c->envs.size())
            return sfail(s, RVS_EINVAL, "rvs_server_submit: bad robot or env id");
        r->rb = c->robots[req->robot].dev;
        r->env = c->envs[req->env].dev;
        r->dofs = c->robots[req->robot].num_dofs;
        r->links = c->robots[req->robot].num_links;
    }
    if (staged_bytes<float>(r->env) > device_lds_limit(s->ctx->opt.device))
        return sfail(s, RVS_ERANGE, "rvs_server_submit: obstacle set too large");
    if (req->n > s->opt.max_batch_configs || req->n > INT32_MAX)
        return sfail(s, RVS_ERANGE, "rvs_server_submit: request larger than max_batch_configs");
    Layout lay;
    lay.N = req->n;
    lay.nreq = 1;
    lay.D = r->dofs;
    lay.L = r->links;
    lay.M = r->env.num_obj;
    lay.prof = req->profile != nullptr;
    lay.under = req->kind == RVS_REQ_UNDERESTIMATION;
    if (up(lay.need()) > s->pool.capacity())
        return sfail(s, RVS_ERANGE, "rvs_server_submit: request larger than the memory budget");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) return RVS_EINVAL;
        r->ticket = s->next_ticket++;
        r->submit_s = s->now();
        if (req->n == 0) {
            r->done = true;
            r->done_s = r->admit_s = r->submit_s;
        } else {
            s->pending.push_back(r);
        }
        s->live[r->ticket] = r;
        s->stats.submitted++;
        *ticket = r->ticket;
    }
    s->cv_work.notify_all();
    return RVS_OK;
}

rvs_status rvs_server_wait(rvs_server* s, uint64_t ticket, int32_t timeout_ms, rvs_request_info* info) {
    if (!s) return RVS_EINVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->live.find(ticket);
    if (it == s->live.end()) return RVS_EINVAL;
    std::shared_ptr<Req> r = it->second;
    if (timeout_ms < 0) s->cv_done.wait(lk, [&] { return r->done; });
    else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&] { return r->done; })) return RVS_EBUSY;
    if (info) {
        info->submit_s = r->submit_s;
        info->admit_s = r->admit_s;
        info->done_s = r->done_s;
        info->tick = r->tick;
        info->batch_requests = r->batch;
        info->status = r->status;
    }
    s->live.erase(ticket);
    return (rvs_status)r->status;
}

rvs_status rvs_server_get_stats(rvs_server* s, rvs_server_stats* st) {
    if (!s || !st) return RVS_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    *st = s->stats;
    return RVS_OK;
}

double rvs_server_now(rvs_server* s) { return s ? s->now() : 0.0; }

const char* rvs_server_last_error(rvs_server* s) {
    if (!s) return "null server";
    // a copy per calling thread: the server thread may replace the message
    thread_local std::string copy;
    std::lock_guard<std::mutex> lk(s->mu);
    copy = s->err;
    return copy.c_str();
}

void* rvs_host_alloc(int64_t bytes) {
    void* p = nullptr;
    if (bytes <= 0 || hipHostMalloc(&p, (size_t)bytes, hipHostMallocDefault) != hipSuccess) return nullptr;
    return p;
}

void rvs_host_free(void* p) {
    if (p) (void)hipHostFree(p);
}

}  // extern "C"
