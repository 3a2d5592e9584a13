// Internal interface between the C ABI layer (rvs_capi.cpp) and the kernel
// launchers (rvs_kernels.hip).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "rvs_fk.hpp"

namespace rvs {

enum : uint8_t { kObjGround = 1, kObjDynamic = 2, kObjObb = 4 };

// Obstacle set as the kernels see it.
struct EnvDev {
    int32_t num_obj;
    int32_t use_obb;          // 1: true capsule-OBB distance (build extension)
    const float* aabb;        // [M][6] xmin ymin zmin xmax ymax zmax
    const float* tol;         // [M] min_dist_tol
    const uint8_t* flags;     // [M] kObj*
    const float* obb;         // [M][15] center(3), rotation rows(9), half extents(3)
    const float* cull;        // [M][6] box containing the obstacle for the broad phase:
                              // the AABB (reference semantics) or the OBB's tight world AABB
    const int32_t* gorder;    // [M] broad-phase walk order (obstacles by Morton code of their cull box
                              // centres); null = obstacle order
    const float* gcull;       // [ceil(M / kGroupSize)][6] union of the cull boxes of each group of
                              // kGroupSize consecutive entries of gorder
};

// Obstacles per group of the broad phase's grouped walk.
constexpr int kGroupSize = 4;
__host__ __device__ inline int num_groups(int M) { return (M + kGroupSize - 1) / kGroupSize; }

// Walk order of the grouped broad phase: obstacles sorted (stably) by the
// Morton code, 10 bits per axis over the scene's extent, of their cull box
// centres (computed in double).  Non-finite centres go to code 0.
void group_order(const float* cull, int M, int32_t* order);
// Union of the cull boxes of every group of kGroupSize consecutive entries of
// order; a group with a non-finite member bound gets an infinite box.
void group_bounds(const float* cull, const int32_t* order, int M, float* gcull);
// The same on the device (after the obstacles moved).
hipError_t launch_group_bounds(const float* cull, const int32_t* order, int M, float* gcull, hipStream_t st);

struct QueryParams {
    float eq;       // EQUALITY_THRESHOLD
    float maxd;     // MAX_DISTANCE
    float band;     // near-contact band of the FP32 path
    int32_t num_checks;
    int32_t flag_link;   // set by the FP32 path's launches: flag_entry carries the start link
};

// Entries of the FP32 path's list of configurations for the exact pass.  In
// a batch of at most 2^kFlagLinkShift configurations an entry is c | k <<
// kFlagLinkShift, k the first link whose decision the bulk pass flagged: the
// links before k were decided with the same margin as an unflagged
// configuration's (no hit, values within the bulk pass's error), so the
// exact pass keeps their values and goes on from link k.  Larger batches
// list c alone (exact pass from link 0).  The bulk and refinement launches
// both take the rule from n.
constexpr int kFlagLinkShift = 25;   // 2^25 configurations, 6 bits of link (kMaxLinks = 64)
static_assert(kMaxLinks <= 64, "a flag entry holds the link in 6 bits");
inline bool flags_carry_link(int64_t n) { return n <= ((int64_t)1 << kFlagLinkShift); }
__host__ __device__ inline int64_t flag_entry(int64_t c, int link, bool carry) {
    return carry ? (c | ((int64_t)link << kFlagLinkShift)) : c;
}

// Outputs of the distance query.  Pointers may be null when not wanted.
struct DistOut {
    float* d;          // [N]
    float* profile;    // [N][L]
    float* nearest;    // [N][M][L][6]
    uint8_t* verdict;  // [N]  1 = collision-free (d > 0)
    double* d64;       // [N]     FP64 path only
    double* profile64; // [N][L]  FP64 path only
};

// Largest obstacle count staged whole in LDS by the distance kernels.
constexpr int kMaxLdsObstacles = 1024;

// Edges of several robots for the persistent first-collision kernel.  Item t
// is 64 * spl consecutive slots (spl = slots per lane) of group items[t].x
// from slot items[t].y on; slot s of group g is state k of edge
// order[order_off[g] + s - rounds[round_off[g] + k]] where k is the largest
// round with rounds[round_off[g] + k] <= s.  Items are ordered by the state
// index of their first slot, then by group.
constexpr int kMaxEdgeGroups = 8;

// Axis-test table entry of the first-collision kernel on axis-aligned
// scenes: an obstacle's bounds minus (lo) and plus (hi) one link radius, in
// float as collisionCapsuleToBox's second test rounds them, and the
// obstacle's index (-1: the link skips the obstacle, a ground box below
// ground_included).  One table per distinct (radius, ground skipped) key:
// ceil(M / kSlabGroup) group entries (the element-wise min of lo and max of
// hi over kSlabGroup consecutive obstacles of the walk order) followed by M
// obstacle entries in walk order.
constexpr int kSlabGroup = 8;
struct SlabEntry {
    float lo[3], hi[3];
    int32_t idx, pad;
};
static_assert(sizeof(SlabEntry) == 32, "one 32-byte scalar load per entry");

struct EdgeGroups {
    RobotDev rb[kMaxEdgeGroups];
    int32_t num;
    int32_t max_dofs;
    int32_t max_links;
    int32_t pad;
    const int2* items;
    int32_t num_slots[kMaxEdgeGroups];
    int32_t num_rounds[kMaxEdgeGroups];
    int32_t round_off[kMaxEdgeGroups];
    int32_t order_off[kMaxEdgeGroups];
    // axis-test tables (axis-aligned scenes): link i of group g uses table
    // link_key[key_off[g] + i] of slab
    int32_t key_off[kMaxEdgeGroups];
    const int32_t* link_key;
    const SlabEntry* slab;
};

// Axis-aligned scenes use the axis-test tables: link i of the robot uses
// table link_key[i] of slab (nkeys tables, filled by the launch from keys).
hipError_t launch_edge_valid(const RobotDev& rb, const EnvDev& env, const QueryParams& qp, const float* q1,
                             const float* q2, int64_t n, uint8_t* valid, const int32_t* link_key,
                             const float* keys, int nkeys, SlabEntry* slab, hipStream_t st);
size_t edge_first_lds_bytes(const EdgeGroups& gr, const EnvDev& env);
// keys: (radius, ground skipped as 0 / 1) of each axis-test table, nkeys of
// them; the launch fills gr.slab from the scene before the search.
hipError_t launch_edge_first(const EdgeGroups& gr, const EnvDev& env, const QueryParams& qp, const float* q1,
                             const float* dir, int32_t row_stride, float step, const int32_t* order,
                             const int32_t* rounds, int32_t* first, int64_t num_edges, uint32_t* counter,
                             int32_t total_items, int waves, int spl, const float* keys, int nkeys,
                             hipStream_t st);
inline size_t slab_entries(int num_obj) { return (size_t)(num_obj + kSlabGroup - 1) / kSlabGroup + num_obj; }

// FP32 bulk pass: relaxed arithmetic, writes its results and appends every
// configuration with an uncertain pair to idx (count in *count, which the
// pass expects to be zero).
// Launch shape of the FP32 bulk pass (chosen by the autotuner).
struct FastTuning {
    int32_t variant;        // 0: one lane per configuration; 1: persistent lane refill;
                            // 2: one lane per configuration in persistent waves over 64-configuration chunks;
                            // 3: persistent lane refill with every lane queueing its own pairs;
                            // 4: one launch per link over the configurations alive at it
    int32_t block;          // threads per workgroup: 64, 128 or 256 (variants 3 and 4: 256, 512 or 1024)
    int32_t blocks_per_cu;  // lane refill: resident workgroups per CU, 0 = the occupancy limit
    int32_t coh_bits;       // walk order: 2^bits bins per joint angle of the first 18 / bits joints, 0 = off
};
// Walk-order sort: keys of up to kCohKeyBits bits (bits per joint angle x
// min(dofs, kCohKeyBits / bits) joints); the top kCohHiBits select one of kCohBuckets buckets,
// the rest (at most kCohLoBits) order each chunk of a bucket.  (A sorted copy
// of the configurations instead of the permutation's gather was measured
// slower.)
constexpr int kCohKeyBits = 18;
constexpr int kCohHiBits = 8;
constexpr int kCohLoBits = kCohKeyBits - kCohHiBits;
constexpr int kCohBuckets = 1 << kCohHiBits;
constexpr int kCohTile = 4096;     // configurations per workgroup of the bucket passes
constexpr int kCohChunk = 16384;   // at most this many configurations per low-bit sort
inline int64_t coh_tiles(int64_t n) { return (n + kCohTile - 1) / kCohTile; }
inline size_t coh_words(size_t w) { return (w + 63) & ~size_t(63); }   // 256-byte pieces
// Pieces of the walk-order scratch (256-byte aligned base).
struct CohLayout {
    uint32_t* hist;     // kCohBuckets x tiles tile counts, then offsets
    uint32_t* sizes;    // kCohBuckets bucket sizes
    uint32_t* keys;     // n keys
    int32_t* idx;       // n configuration indices in bucket order
    uint16_t* key_lo;   // n low key bits in bucket order
    int32_t* perm;      // n: the walk order
    int64_t tiles;
};
inline size_t coh_bytes(int64_t n, int dofs) {
    (void)dofs;
    const size_t m = (size_t)std::max<int64_t>(n, 0);
    return 4 * (coh_words(kCohBuckets * (size_t)coh_tiles(n)) + coh_words(kCohBuckets) + 3 * coh_words(m) +
                coh_words((m + 1) / 2));
}
inline CohLayout coh_layout(uint32_t* base, int64_t n) {
    CohLayout L;
    const size_t m = (size_t)std::max<int64_t>(n, 0);
    L.tiles = coh_tiles(n);
    L.hist = base;
    L.sizes = L.hist + coh_words(kCohBuckets * (size_t)L.tiles);
    L.keys = L.sizes + coh_words(kCohBuckets);
    L.idx = reinterpret_cast<int32_t*>(L.keys + coh_words(m));
    L.perm = L.idx + coh_words(m);
    L.key_lo = reinterpret_cast<uint16_t*>(L.perm + coh_words(m));
    return L;
}
// The walk order of the bulk pass: with t.coh_bits > 0 and coh scratch the
// configurations are bucketed by their first joint angles (so that the lanes
// of a wave hold similar poses) and *perm points at the order in coh;
// otherwise *perm is null (the order given).  When it runs, its first kernel
// also clears zero[0, nzero) (zero may be null); with *perm null nothing
// was cleared.
hipError_t launch_coherence_order(const RobotDev& rb, const float* q, int64_t n, const FastTuning& t, uint32_t* coh,
                                  const int32_t** perm, hipStream_t st, uint32_t* zero = nullptr, int nzero = 0);
// Scratch of the link-major bulk pass (tuning variant 4), 256-byte aligned:
// two lists of the configurations alive at a link, used in turn, whose
// entries carry each configuration's state between links (read and written
// in list order, so the state moves in whole lines): its running minimum,
// (warm pair + 1) << 1 | screening flag and the forward-kinematics walk up to
// the link's end point, so that the next link goes on from there; then the
// lists' sizes per link and the round counters per link (kLinksCounters
// words, cleared together), the box that quantizes the sort keys and the
// first link's warm pair.
struct alignas(16) LinkEntry {
    int32_t c;       // configuration
    float d;         // running minimum over the links before
    uint32_t w;      // (warm pair + 1) << 1 | screening flag
    uint32_t walk;   // the walk's next segment | next joint << 8 | (first uncertain link + 1) << 16
    float m[9];      // the walk's frame: orientation (row major)
    float p[3];      // and origin
};
static_assert(sizeof(LinkEntry) == 64, "LinkEntry is four 16-byte pieces");
static_assert(kMaxSeg < 255 && kMaxLinks < 255, "the walk word holds segment, joint and link in 8 bits each");
constexpr int kLinksCounters = 2 * kMaxLinks + 2;
struct LinksLayout {
    LinkEntry* list[2];
    uint32_t* count;    // [kMaxLinks + 1] list sizes, then
    uint32_t* cursor;   // [kMaxLinks + 1] round counters
    float* box;         // lo[3], 1024 / extent[3]
    int32_t* warm0;
};
inline size_t fast_links_bytes(int64_t n) {
    const size_t m = (size_t)std::max<int64_t>(n, 0);
    return 4 * (2 * coh_words(16 * m) + coh_words(kLinksCounters) + coh_words(8));
}
inline LinksLayout links_layout(uint32_t* base, int64_t n) {
    const size_t lw = coh_words(16 * (size_t)std::max<int64_t>(n, 0));
    LinksLayout L;
    L.list[0] = reinterpret_cast<LinkEntry*>(base);
    L.list[1] = reinterpret_cast<LinkEntry*>(base + lw);
    L.count = base + 2 * lw;
    L.cursor = L.count + kMaxLinks + 1;
    L.box = reinterpret_cast<float*>(L.count + coh_words(kLinksCounters));
    L.warm0 = reinterpret_cast<int32_t*>(L.box + 6);
    return L;
}
// The persistent variants (1: lane refill, 2: chunks, 3: per-lane queues)
// use count[32] as their work counter (expected zero, like count[0]).
// perm: the walk order (may be null).  Variants 3 and 4 need the scene's
// group tables and their queues in LDS, and variant 4 the links scratch of
// fast_links_bytes(n) bytes (links, may be null); without them variant 4
// runs as 3, and 3 as 1.
hipError_t launch_distance_fast(const RobotDev& rb, const EnvDev& env, const QueryParams& qp, const float* q,
                                int64_t n, const DistOut& out, int32_t* idx, uint32_t* count, const FastTuning& t,
                                const int32_t* perm, uint32_t* links, int num_cus, hipStream_t st);
// Resident workgroups per CU of the lane-refill kernel for this obstacle set.
int fast_pool_blocks_per_cu(const EnvDev& env, int block);
// The same for the per-lane-walk kernel (0 when it does not run on this
// obstacle set: no group tables, or its tables exceed the LDS limit).
int fast_lanes_blocks_per_cu(const EnvDev& env, int block);
// The same for the link-major kernel.
int fast_links_blocks_per_cu(const EnvDev& env, int block);
// Workgroup LDS of device dev, or of the current device when dev < 0
// (hipDeviceAttributeMaxSharedMemoryPerBlock, read once per device).
size_t device_lds_limit(int dev = -1);
// The exact pass over the list the bulk pass wrote (idx, *count entries);
// count_out (a device address, may be null) receives *count, for a caller
// that reads it from mapped host memory instead of copying it.
hipError_t launch_distance_refine(const RobotDev& rb, const EnvDev& env, const QueryParams& qp, const float* q,
                                  int64_t n, const DistOut& out, const int32_t* idx, const uint32_t* count,
                                  hipStream_t st, uint32_t* count_out = nullptr);
hipError_t launch_zero_u32(uint32_t* p, int n, hipStream_t st);


// One lane per configuration; with perm, lane g takes configuration perm[g]
// (a walk order from launch_coherence_order).
hipError_t launch_distance_ref32(const RobotDev& rb, const EnvDev& env, const QueryParams& qp, const float* q,
                                 int64_t n, const DistOut& out, hipStream_t st, const int32_t* perm = nullptr);
// The same results with one wave per configuration (lanes over obstacles):
// the low-latency form for small batches.
hipError_t launch_distance_ref32_waves(const RobotDev& rb, const EnvDev& env, const QueryParams& qp, const float* q,
                                       int64_t n, const DistOut& out, hipStream_t st);
// Either form by batch size: waves up to wave_max configurations.
inline hipError_t launch_distance_ref32_auto(const RobotDev& rb, const EnvDev& env, const QueryParams& qp,
                                             const float* q, int64_t n, const DistOut& out, int64_t wave_max,
                                             hipStream_t st, const int32_t* perm = nullptr) {
    return n <= wave_max ? launch_distance_ref32_waves(rb, env, qp, q, n, out, st)
                         : launch_distance_ref32(rb, env, qp, q, n, out, st, perm);
}
hipError_t launch_distance_fp64(const RobotDev& rb, const EnvDev& env, const QueryParams& qp, const float* q,
                                int64_t n, const DistOut& out, hipStream_t st, const int32_t* perm = nullptr);
hipError_t launch_is_valid(const RobotDev& rb, const EnvDev& env, const QueryParams& qp, const float* q, int64_t n,
                           uint8_t* valid, int fp64, hipStream_t st, const int32_t* perm = nullptr);
hipError_t launch_underestimation(const RobotDev& rb, const EnvDev& env, const QueryParams& qp, const float* q,
                                  int64_t n, const float* np_ref, int64_t np_stride, float* d, float* profile,
                                  uint8_t* stored, hipStream_t st);
// The same with the planes moved towards the robot for delta_t at max_vel[j]
// of each obstacle (AbstractTrajectory::computeDistanceUnderestimation).
hipError_t launch_underestimation_moving(const RobotDev& rb, const EnvDev& env, const QueryParams& qp, const float* q,
                                         int64_t n, const float* np_ref, int64_t np_stride, const float* max_vel,
                                         float delta_t, float* d, float* profile, uint8_t* stored, hipStream_t st);
// Underestimation for a served batch: configuration c takes nearest-point
// set r of np_sets ([nseg][M][L][6]) where seg_off[r] <= c < seg_off[r + 1].
hipError_t launch_underestimation_seg(const RobotDev& rb, const EnvDev& env, const QueryParams& qp, const float* q,
                                      int64_t n, const float* np_sets, const int64_t* seg_off, int nseg, float* d,
                                      float* profile, uint8_t* stored, hipStream_t st);

// A moving obstacle set on the device (all arrays device pointers).
struct MotionDev {
    int32_t num_obj;
    int32_t pad;
    const float* pos0;    // [M][3]
    const float* vel;     // [M][3] m/s (0 for static boxes)
    const float* half;    // [M][3] dim / 2
    const float* rad;     // [M] |dim / 2| (rotated boxes)
    const uint8_t* ident; // [M] 1 = identity rotation
    const float* obb0;    // [M][15] oriented-box rows at pos0
    const float* ext;     // [M][3] half extents of the oriented box's tight world AABB
};
// Writes the AABB table [M][6], the oriented-box table [M][15] and the
// oriented boxes' broad-phase bounds [M][6] at a tick.
hipError_t launch_env_motion(const MotionDev& m, int64_t tick, float tick_seconds, float* aabb, float* obb,
                             float* cull, hipStream_t st);

// Motion-law state of an environment's boxes (device pointers, [M] or [M][3]).
struct EnvStepDev {
    int32_t law;            // RVS_MOTION_*
    int32_t num_obj;
    float ws[3];            // workspace centre
    float ws_radius;
    float base_radius;
    float robot_max_vel;
    float* pos;
    float* vel;
    float* path_len;        // two tunnels
    int32_t* sign;          // two tunnels
    const float* max_vel;
    const uint8_t* dynamic; // 1 = label "dynamic_obstacle"
    const float* half;
    const float* rad;
    const uint8_t* ident;
    const float* ext;
};
// Environment::updateEnvironment(delta_time) 'steps' times, then the AABB
// [M][6], oriented-box centres [M][15] and (when cull is not null) the
// broad-phase bounds [M][6] at the final positions.
hipError_t launch_env_step(const EnvStepDev& m, float delta_time, int32_t steps, float* aabb, float* obb, float* cull,
                           hipStream_t st);

// AbstractRobot::computeEnclosingRadii for N configurations: R [N][n][n+1]
// (row i = starting skeleton point, column j = final point).
hipError_t launch_enclosing_radii(const RobotDev& rb, const float* q, int64_t n, float* R, hipStream_t st);

// One batch of spines (device pointers).  Spine c starts at q + c*q_stride
// with distance d[c*d_stride] and link profile profile + c*profile_stride
// (stride 0 = one start state shared by all spines) and heads for
// q_e [N][n].  q_new [N][n] may be q itself when q_stride = n.  Spines with
// active[c] = 0 are left untouched (active may be null).
struct SpineArgs {
    const float* q;
    int64_t q_stride;
    const float* d;
    int64_t d_stride;
    const float* profile;
    int64_t profile_stride;
    const float* q_e;
    const uint8_t* active;
    float* q_new;
    int32_t* status;   // 1 advanced, 2 trapped, 3 reached
    int32_t num_iter;
    int32_t expanded;
};
hipError_t launch_spines(const RobotDev& rb, const SpineArgs& a, int64_t n, hipStream_t st);
// Generalized spines with underestimation layers, fused (device pointers):
// root state, its distance [1], profile [L] and nearest points [M][L][6];
// targets q_e [N][n]; outputs q_new [N][n], status [N], d_out [N].
struct GenSpineArgs {
    const float* q_root;
    const float* d_root;
    const float* profile_root;
    const float* np_root;
    const float* q_e;
    float* q_new;
    int32_t* status;
    float* d_out;
    int32_t num_obj;
    int32_t num_layers;
    int32_t num_iter;
    int32_t expanded;
    float d_crit;
    float maxd;
};
hipError_t launch_gen_spines_under(const RobotDev& rb, const GenSpineArgs& g, int64_t n, hipStream_t st);
// Between layers of generalized spines: d_out = d_layer and deactivation of
// spines with d_layer < d_crit or status != advanced (layer 0: a root in
// contact stops every spine with the root's distance).
hipError_t launch_gen_layer(int64_t n, int layer, const float* d_root, const float* d_layer, const int32_t* status,
                            uint8_t* active, float* d_out, float d_crit, hipStream_t st);
hipError_t launch_fill_u8(uint8_t* p, int64_t n, uint8_t v, hipStream_t st);

// A batch of candidate trajectories leaving one current state (device
// pointers): points [T][num_points][n]; the current state's coordinates,
// profile [L] and nearest points [M][L][6]; max_vel [M].
struct TrajArgs {
    const float* points;
    int32_t num_points;
    int32_t num_obj;
    const float* q_current;
    const float* profile_current;
    const float* nearest_current;
    const float* max_vel;
    float t_iter;
    float time_step;
    int32_t self_collision;
    uint8_t* safe;
};
hipError_t launch_traj_safe(const RobotDev& rb, const QueryParams& qp, const TrajArgs& a, int64_t T, hipStream_t st);
size_t traj_lds_bytes(const RobotDev& rb);   // LDS of one trajectory's workgroup
// Capsule part of xArm6::checkSelfCollision(q1, q2) per edge: collide[e] =
// 1 (q_out = last state before the hit), 0 (q_out = q2) or 255 (the walk
// would not end: non-finite or huge coordinates).
hipError_t launch_edge_self(const RobotDev& rb, const QueryParams& qp, const float* q1, const float* q2, int64_t n,
                            float eps_step, uint8_t* collide, float* q_out, hipStream_t st);

// Per-pair distances (dist / dist64, [N][pairs]) and/or the per-configuration
// minimum and collision-free flag (min_out / free_out); null outputs are skipped.
hipError_t launch_self_distance(const RobotDev& rb, const QueryParams& qp, const float* q, int64_t n, float* dist,
                                double* dist64, float* min_out, uint8_t* free_out, hipStream_t st);

}  // namespace rvs
