/*
 * rvs.h - C ABI of the MI355X capsule collision / distance engine.
 *
 * Drop-in for the capsule-primitive RealVectorSpace path of RPMPLv2
 * (include/state_spaces/real_vector_space/RealVectorSpace.h:17-53, the
 * base::StateSpace pure virtuals of include/state_spaces/StateSpace.h:52-56,
 * and the robot-side queries of include/robots/AbstractRobot.h:58-61).
 *
 * Conventions
 *   - Plain C types only; every call returns rvs_status; no exceptions cross
 *     the ABI.  rvs_last_error() describes the last failure of a context;
 *     the string is the calling thread's own copy, valid until that thread
 *     calls rvs_last_error (or rvs_server_last_error) again.
 *   - Buffers are caller-owned.  With RVS_F_DEVICE every array argument is a
 *     device pointer on the context's GPU; otherwise arrays are host memory
 *     and the call stages them through the context's device arena.  No
 *     pointer is kept after a synchronous call returns.
 *   - A context is internally synchronised: calls from several host threads
 *     on one context are serialised; separate contexts run independently.
 *   - Layouts are configuration-major: q [N][n], profile [N][L],
 *     nearest [N][M][L][6] (robot xyz, obstacle xyz), verdict [N] (1 = free).
 */
#ifndef RVS_H
#define RVS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum rvs_status {
    RVS_OK = 0,
    RVS_EINVAL = 1,   /* bad argument (null pointer, size, id) */
    RVS_ENOMEM = 2,   /* device or pinned allocation failed */
    RVS_EDEVICE = 3,  /* HIP runtime or kernel launch error */
    RVS_ERANGE = 4,   /* size beyond a documented limit */
    RVS_ENOTSUP = 5,  /* combination not supported */
    RVS_EBUSY = 6     /* a served request has not finished within the wait */
} rvs_status;

/* Arithmetic of a query. */
typedef enum rvs_precision {
    /* Reference arithmetic: FP64 kinematics, FP32 skeleton and geometry in
     * the reference's operation order.  Bit-identical to RPMPLv2. */
    RVS_PREC_REF32 = 0,
    /* The same algorithm carried out entirely in FP64. */
    RVS_PREC_FP64 = 1,
    /* FP32 bulk pass in relaxed single precision; every configuration with
     * an uncertain pair (inside the near-contact band, near the
     * EQUALITY_THRESHOLD snap, a marginal inside/crossing decision, near a
     * min_dist_tol, a near-parallel segment solve, or non-finite input) is
     * recomputed with the RVS_PREC_REF32 arithmetic (FP64 kinematics,
     * reference-order FP32 geometry) from the first link with an uncertain
     * pair on (from link 0 in batches above 2^25 configurations or without
     * a profile output); the links before it keep the bulk pass's values,
     * decided with the same margin as an unflagged configuration's.
     * Verdicts are those of RVS_PREC_REF32; distances stay within 1e-5 m of
     * RVS_PREC_FP64. */
    RVS_PREC_FP32_REFINE = 2
} rvs_precision;

/* Call flags. */
#define RVS_F_DEVICE 0x1u          /* array arguments are device pointers */
#define RVS_F_TIMING 0x2u          /* time the call's kernel stages with HIP events (rvs_last_timing) */
#define RVS_F_EAGER 0x4u           /* launch kernel by kernel even where a captured HIP graph would be replayed */
#define RVS_F_PREC_SHIFT 4         /* bits 4..5: precision override + 1 (0 = context default) */
#define RVS_F_PREC(p) ((uint32_t)((p) + 1) << RVS_F_PREC_SHIFT)

typedef struct rvs_options {
    int32_t device;             /* HIP device ordinal */
    int32_t precision;          /* rvs_precision used when a call does not override it */
    float equality_threshold;   /* RealVectorSpaceConfig::EQUALITY_THRESHOLD (yaml: 1e-4) */
    float max_distance;         /* RealVectorSpaceConfig::MAX_DISTANCE (1000) */
    int32_t num_checks;         /* RealVectorSpaceConfig::NUM_INTERPOLATION_VALIDITY_CHECKS (yaml: 10) */
    float band;                 /* near-contact band of RVS_PREC_FP32_REFINE, metres */
} rvs_options;

/* Fills the defaults the reference's apps load from
 * data/configurations/configuration_realvectorspace.yaml:1-3. */
void rvs_default_options(rvs_options* opt);

typedef struct rvs_context rvs_context;

/* ---------------------------------------------------------------- robots */
enum { RVS_JOINT_FIXED = 0, RVS_JOINT_REVOLUTE = 1 };
enum { RVS_SKEL_CHAIN = 0, RVS_SKEL_XARM6 = 1 };

typedef struct rvs_joint {
    int32_t type;        /* RVS_JOINT_* */
    int32_t reserved;
    double xyz[3];       /* URDF joint origin */
    double rpy[3];
    double axis[3];      /* URDF joint axis */
} rvs_joint;

/* A serial chain from the base link to the tool, in URDF joint order.
 * skeleton_kind RVS_SKEL_CHAIN: capsule end points are the segment origins
 * (Planar2DOF::computeSkeleton, src/robots/Planar2DOF.cpp:117-130);
 * RVS_SKEL_XARM6: the xArm6 recipe (src/robots/xArm6.cpp:209-235). */
typedef struct rvs_robot_desc {
    int32_t skeleton_kind;
    int32_t num_joints;
    const rvs_joint* joints;
    int32_t num_links;             /* capsules; points = num_links + 1 */
    const float* capsule_radius;   /* [num_links] (AbstractRobot::getCapsuleRadius) */
    int32_t ground_included;       /* links [0, ground_included) ignore "ground" boxes */
    float gripper_length;          /* xArm6 only */
    int32_t num_self_pairs;        /* link pairs (a, b) for self-collision */
    const int32_t* self_pairs;     /* [num_self_pairs][2] */
} rvs_robot_desc;

/* Replaces robots::*::computeSkeleton set-up + AbstractRobot::setCapsulesRadius
 * (include/robots/AbstractRobot.h:46,58). */
rvs_status rvs_robot_load(rvs_context* ctx, const rvs_robot_desc* desc, int32_t* robot_id);

/* ---------------------------------------------------------------- obstacles */
#define RVS_BOX_GROUND 0x1u    /* label "ground" (Environment.cpp:32-33) */
#define RVS_BOX_DYNAMIC 0x2u   /* label "dynamic_obstacle" */

typedef struct rvs_box {
    float dim[3];
    float pos[3];
    float quat_wxyz[4];        /* identity for axis-aligned boxes */
    uint32_t flags;            /* RVS_BOX_* */
    float min_dist_tol;        /* Object::getMinDistTol (default MAX_DISTANCE) */
} rvs_box;

/* Obstacle semantics of an environment. */
enum {
    /* Reference semantics: every box is queried through its FCL world AABB
     * (RealVectorSpace.cpp:293-295, 344-346); a rotated box becomes the cube
     * centre +- |dim|/2. */
    RVS_ENV_AABB = 0,
    /* Build extension: rotated boxes are queried as oriented boxes (the
     * capsule is moved into the box frame; nearest points come back in the
     * world frame).  Accepted by every environment query.  The distance and
     * collision-tree queries stage the obstacle set in the workgroup LDS the
     * device reports (hipDeviceAttributeMaxSharedMemoryPerBlock): with 64 KB,
     * up to 851 oriented boxes for float queries and 624 for RVS_PREC_FP64
     * (RVS_ERANGE beyond); the edge queries read the boxes from device memory
     * and take any environment; environments hold at most 1024 boxes. */
    RVS_ENV_OBB = 1
};

/* Replaces env::Environment construction from a scenario
 * (src/environments/Environment.cpp:13-109, Box.cpp:3-15). */
rvs_status rvs_env_create(rvs_context* ctx, const rvs_box* boxes, int32_t num_boxes, int32_t semantics,
                          int32_t* env_id);
/* In-place move of boxes idx[k] to pos[k] (Object::setPosition,
 * src/environments/Object.cpp:21-26); the device bounds tables are
 * rewritten. */
rvs_status rvs_env_update(rvs_context* ctx, int32_t env_id, const int32_t* idx, const float* pos, int32_t count,
                          uint32_t flags);
/* Axis-aligned bounds [num][6] the queries use (host memory). */
rvs_status rvs_env_bounds(rvs_context* ctx, int32_t env_id, float* aabb, int32_t num);

/* Motion laws of Environment::updateEnvironment (src/environments/
 * Environment.cpp:194-377).  random_directions draws from std::rand
 * (Eigen's Random) and has no counterpart here. */
enum { RVS_MOTION_STRAIGHT = 0, RVS_MOTION_CIRCULAR = 1, RVS_MOTION_TWO_TUNNELS = 2, RVS_MOTION_LIGHT_DIRECTIONS = 3 };

typedef struct rvs_env_motion {
    int32_t motion_type;     /* RVS_MOTION_* (scenario "testing: motion_type") */
    float ws_center[3];      /* Environment::getWSCenter (scenario robot WS_center) */
    float ws_radius;         /* getWSRadius */
    float base_radius;       /* getBaseRadius */
    float robot_max_vel;     /* getRobotMaxVel */
} rvs_env_motion;

/* Sets the motion law of an environment with each box's velocity vel [M][3]
 * (Object::getVelocity) and maximal speed max_vel [M] (Object::getMaxVel),
 * host arrays.  Boxes flagged RVS_BOX_DYNAMIC move; the two-tunnels state
 * starts at sign 1, path length 0 (Environment.cpp:82-83). */
rvs_status rvs_env_set_motion_law(rvs_context* ctx, int32_t env_id, const rvs_env_motion* law, const float* vel,
                                  const float* max_vel);
/* Environment::updateEnvironment(delta_time), 'steps' times, on the GPU;
 * positions, velocities and every bound the queries read are updated in
 * place.  pos and vel ([M][3], host, may be NULL) receive the new state. */
rvs_status rvs_env_step(rvs_context* ctx, int32_t env_id, float delta_time, int32_t steps, float* pos, float* vel);

/* ---------------------------------------------------------------- queries */
/* RealVectorSpace::computeDistance(q, compute_again = true)
 * (RealVectorSpace.cpp:318-392) for N configurations: d, per-link profile,
 * nearest points (optional), verdict = (d > 0).  profile and nearest may be
 * NULL. */
rvs_status rvs_distance(rvs_context* ctx, int32_t robot, int32_t env, const float* q, int64_t n, float* d,
                        float* profile, float* nearest, uint8_t* verdict, uint32_t flags);

/* The FP64 path with double outputs (RVS_PREC_FP64 always). */
rvs_status rvs_distance_f64(rvs_context* ctx, int32_t robot, int32_t env, const float* q, int64_t n, double* d,
                            double* profile, uint8_t* verdict, uint32_t flags);

/* RealVectorSpace::isValid(q) (RealVectorSpace.cpp:277-309): collision tree,
 * environment only.  valid[k] = 1 when configuration k is collision-free.
 * RVS_F_TIMING records the collision tree's kernel time as stage 0 and the
 * walk-order sort's as stage 1 of rvs_last_timing. */
rvs_status rvs_is_valid(rvs_context* ctx, int32_t robot, int32_t env, const float* q, int64_t n, uint8_t* valid,
                        uint32_t flags);

/* Stream-ordered forms of rvs_distance (no nearest points) and rvs_is_valid
 * (and, below, of the underestimation and self-collision queries) for device
 * arrays (flags must hold RVS_F_DEVICE; no RVS_F_TIMING): the
 * work is enqueued on 'stream' (a hipStream_t; NULL = the context's stream)
 * and the call returns at once.  'done' (a hipEvent_t, may be NULL) is
 * recorded on the stream after the work.  The arrays must stay valid until
 * then.  Asynchronous calls run in call order on the device (each waits for
 * the previous one, whatever its stream); any other call on the context
 * first waits on the host for the asynchronous calls in flight.
 * rvs_last_refined does not report asynchronous calls.  The reference's
 * interface has no asynchronous form (RealVectorSpace.h:45-49 are blocking
 * calls); these serve callers that overlap queries with their own GPU work. */
rvs_status rvs_distance_async(rvs_context* ctx, int32_t robot, int32_t env, const float* q, int64_t n, float* d,
                              float* profile, uint8_t* verdict, uint32_t flags, void* stream, void* done);
/* rvs_distance_async that allocates nothing: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid it).  Its
 * scratch (refinement list and counters, walk-order sort) is a fixed device
 * array of 20 MiB that each GPU the library runs on holds until the process
 * ends: at most 1048576 configurations per call (RVS_ERANGE beyond).  Same
 * arguments, results bit-identical to rvs_distance_async.  Every call on one
 * GPU shares that GPU's array, whatever its context: calls that overlap on
 * one GPU (two contexts, or a call enqueued on a stream that runs beside
 * another such call) race on it; calls on one context are ordered like
 * every asynchronous call. */
rvs_status rvs_distance_async_static(rvs_context* ctx, int32_t robot, int32_t env, const float* q, int64_t n,
                                     float* d, float* profile, uint8_t* verdict, uint32_t flags, void* stream,
                                     void* done);
rvs_status rvs_is_valid_async(rvs_context* ctx, int32_t robot, int32_t env, const float* q, int64_t n,
                              uint8_t* valid, uint32_t flags, void* stream, void* done);
/* The same stream-ordered form of rvs_distance_underestimation and of
 * rvs_self_collision (device arrays). */
rvs_status rvs_distance_underestimation_async(rvs_context* ctx, int32_t robot, int32_t env, const float* q,
                                              int64_t n, const float* nearest_ref, int64_t ref_stride, float* d,
                                              float* profile, uint8_t* stored, uint32_t flags, void* stream,
                                              void* done);
rvs_status rvs_self_collision_async(rvs_context* ctx, int32_t robot, const float* q, int64_t n, float* min_dist,
                                    uint8_t* free_flag, uint32_t flags, void* stream, void* done);

/* RealVectorSpace::computeDistanceUnderestimation(q, nearest_points)
 * (RealVectorSpace.cpp:402-446) on fresh states.  nearest_ref is [M][L][6]
 * shared by all configurations (ref_stride = 0) or one set per configuration
 * (ref_stride = M*L*6).  stored[k] = 0 where the reference returns 0 without
 * storing a profile (a negative plane distance); profile rows are then 0.
 * RVS_F_TIMING records the kernel time as stage 0 of rvs_last_timing. */
rvs_status rvs_distance_underestimation(rvs_context* ctx, int32_t robot, int32_t env, const float* q, int64_t n,
                                        const float* nearest_ref, int64_t ref_stride, float* d, float* profile,
                                        uint8_t* stored, uint32_t flags);

/* AbstractTrajectory::computeDistanceUnderestimation(q, nearest_points,
 * delta_t) (src/planners/trajectory/AbstractTrajectory.cpp:272-320): as
 * rvs_distance_underestimation, with every plane first moved towards the
 * robot by max_vel[j] * delta_t (max_vel [M], Object::getMaxVel; 0 for
 * static boxes). */
rvs_status rvs_distance_underestimation_moving(rvs_context* ctx, int32_t robot, int32_t env, const float* q, int64_t n,
                                               const float* nearest_ref, int64_t ref_stride, const float* max_vel,
                                               float delta_t, float* d, float* profile, uint8_t* stored,
                                               uint32_t flags);

/* AbstractTrajectory::isSafe (src/planners/trajectory/AbstractTrajectory.cpp:
 * 212-256) for num_traj candidate trajectories leaving one current state.
 * points [num_traj][num_points][dofs] are the positions at t_iter,
 * t_iter + time_step, ... (TrajectoryConfig::TIME_STEP); q_current [dofs]
 * with its distance profile [links] and nearest points [M][links][6] from
 * rvs_distance; max_vel [M] per obstacle.  safe[k] = 1 when trajectory k
 * is safe.  self_collision = 1 also rejects a point whose self-collision
 * capsule pairs overlap (robot->checkSelfCollision(q_final) without the FCL
 * mesh confirmation, so conservative); 0 for robots without self-collision
 * checking.  Needs one link per revolute joint. */
rvs_status rvs_trajectory_is_safe(rvs_context* ctx, int32_t robot, int32_t env, const float* points, int64_t num_traj,
                                  int32_t num_points, const float* q_current, const float* profile_current,
                                  const float* nearest_current, const float* max_vel, float t_iter, float time_step,
                                  int32_t self_collision, uint8_t* safe, uint32_t flags);

/* Capsule-capsule distances of the robot's self-collision pairs
 * (xArm6::computeCapsulesDistance, src/robots/xArm6.cpp:417-424), [N][P].
 * A pair with distance < 0 is the capsule part of checkSelfCollision
 * (xArm6.cpp:384-411; the FCL mesh confirmation is not part of this path). */
rvs_status rvs_self_distance(rvs_context* ctx, int32_t robot, const float* q, int64_t n, float* dist,
                             uint32_t flags);

/* Capsule part of checkSelfCollision(q) per configuration
 * (xArm6::checkSelfCollision, src/robots/xArm6.cpp:374-411, pair loop at
 * :386-411 with computeCapsulesDistance :417-424): min_dist[N] = smallest pair
 * distance (std::min order from max_distance, NaN pairs ignored) and
 * free_flag[N] = 1 when no pair distance is < 0.  Either output may be null.
 * For chains the pair list (j >= i + 2) is a build extension; the reference
 * chains have no self-collision (Spatial10DOF.cpp:32-43). */
rvs_status rvs_self_collision(rvs_context* ctx, int32_t robot, const float* q, int64_t n, float* min_dist,
                              uint8_t* free_flag, uint32_t flags);

/* AbstractRobot::computeEnclosingRadii for n configurations (xArm6.cpp:237-275,
 * Planar2DOF.cpp:132-148, Spatial10DOF.cpp:11-30): R [n][dofs][dofs+1], row i
 * = starting skeleton point, column j = final point (the transpose of the
 * reference's column-major Eigen::MatrixXf(dofs, dofs+1) read as [i][j]).
 * Chains use the spatial rule (row 0 measured in the {x, y} plane), which is
 * the planar rule for planar chains based at the origin.  Needs one link per
 * revolute joint. */
rvs_status rvs_enclosing_radii(rvs_context* ctx, int32_t robot, const float* q, int64_t n, float* R,
                               uint32_t flags);

/* Capsule part of xArm6::checkSelfCollision(q1, q2) (src/robots/xArm6.cpp:
 * 284-368) for n edges: radii-bounded steps while the link pairs {0-4, 0-5,
 * 1-4, 1-5} provably stay apart (at most 5), then a walk in eps_step
 * (RRTConnectConfig::EPS_STEP) steps testing the pairs not yet proven apart.
 * collide[e] = 1 when a walked state has overlapping capsules, with
 * q2_out[e] = the last state before it (the reference's modified q2);
 * 0 with q2_out[e] = q2[e] otherwise; 255 for an edge whose walk would not
 * end (non-finite or huge coordinates).  A capsule overlap counts as a
 * collision: the FCL mesh confirmation (checkRealSelfCollision, :426-483) is
 * not part of this path, so the answer is conservative.  xArm6 only
 * (RVS_ENOTSUP otherwise; the reference chains never self-collide). */
rvs_status rvs_edge_self_collision(rvs_context* ctx, int32_t robot, const float* q1, const float* q2, int64_t n,
                                   float eps_step, uint8_t* collide, float* q2_out, uint32_t flags);

/* ---------------------------------------------------------------- spines */
/* Spine outcomes, numbered as base::State::Status (include/state_spaces/State.h:22). */
enum { RVS_SPINE_ADVANCED = 1, RVS_SPINE_TRAPPED = 2, RVS_SPINE_REACHED = 3 };

/* RBTConnect::extendSpine (src/planners/rbt/RBTConnect.cpp:85-150) for n
 * spines.  Spine k starts at q[k] ([n][dofs]), whose distance d[k] and
 * per-link profile[k] ([n][links]) come from rvs_distance or
 * rvs_distance_underestimation, and heads for q_e[k].  num_iter =
 * NUM_ITER_SPINE (>= 1), use_expanded_bubble = USE_EXPANDED_BUBBLE (profile
 * may be NULL without it).  Outputs q_new [n][dofs] and status[k]
 * (RVS_SPINE_*; d[k] <= 0 gives TRAPPED with q_new = q).  Reference float
 * arithmetic, bit for bit.  The reference also runs
 * robot->checkSelfCollision(q_temp, q_new) in every iteration (:129-131);
 * these are the spines of a robot without self-collision checking (the
 * chains; xArm6 with self_collision_checking false).  With num_iter = 1 that
 * check is the iteration's last action, so a caller that checks
 * self-collision applies it to (q, q_new) afterwards.  Needs one link per
 * revolute joint. */
rvs_status rvs_extend_spines(rvs_context* ctx, int32_t robot, const float* q, const float* d, const float* profile,
                             const float* q_e, int64_t n, int32_t num_iter, int32_t use_expanded_bubble,
                             float* q_new, int32_t* status, uint32_t flags);

/* RGBTConnect::extendGenSpine (src/planners/rbt/RGBTConnect.cpp:71-90) from
 * one root state q_root ([dofs]) towards n targets q_e ([n][dofs]), the way a
 * generalized bur is built (PatternTree::generateGBur, PatternTree.cpp:20-60;
 * DRGBT::computeReachedState, DRGBT.cpp:475-495): the root's distance
 * (computeDistance, with nearest points), then up to num_layers spines per
 * target (rvs_extend_spines rules), each followed by the distance of its end
 * state: computeDistanceUnderestimation against the root's nearest points
 * (use_real_dist = 0) or computeDistance (use_real_dist = 1).  A target stops
 * when that distance is below d_crit (D_CRIT) or its spine did not advance.
 * Outputs q_new [n][dofs], status [n] and d [n]: the distance the last layer
 * compared with d_crit (0 where the underestimation returns early; the
 * root's distance when the root is in contact).  Environment distances use
 * the reference float arithmetic whatever the context precision. */
rvs_status rvs_extend_gen_spines(rvs_context* ctx, int32_t robot, int32_t env, const float* q_root, const float* q_e,
                                 int64_t n, int32_t num_layers, int32_t num_iter, int32_t use_expanded_bubble,
                                 int32_t use_real_dist, float d_crit, float* q_new, int32_t* status, float* d,
                                 uint32_t flags);

/* RealVectorSpace::isValid(q1, q2) (RealVectorSpace.cpp:261-275): num_checks
 * interpolated states from the far end, q1 not checked.  RVS_F_TIMING
 * records the kernel time as stage 0 of rvs_last_timing. */
rvs_status rvs_edge_is_valid(rvs_context* ctx, int32_t robot, int32_t env, const float* q1, const float* q2,
                             int64_t n, uint8_t* valid, uint32_t flags);

/* Discretised edges: state k of edge e is q1[e] + (step*k) * dir[e],
 * k = 0..K[e]-1.  first[e] = first k whose state collides with the
 * environment (collision tree) or has a self-collision pair with distance
 * < 0, or -1.  Edges of several robots go in one persistent launch:
 * robots[e] selects the robot of edge e, q1/dir rows are padded to
 * row_stride floats.  robots and K are host arrays (the launch is planned on
 * the host); with RVS_F_DEVICE, q1, dir and first are device arrays.
 * RVS_F_TIMING records the kernel time as stage 0 of rvs_last_timing. */
rvs_status rvs_edge_first_collision(rvs_context* ctx, int32_t env, const int32_t* robots, const float* q1,
                                    const float* dir, const int32_t* K, int64_t num_edges, int32_t row_stride,
                                    float step, int32_t* first, uint32_t flags);

/* ---------------------------------------------------------------- autotuning
 * Launch shape of the RVS_PREC_FP32_REFINE bulk pass for one (robot,
 * environment): configurations per wavefront (one per lane for the whole
 * walk, or persistent lane refill that hands a lane the next configuration
 * after each link, or persistent lane refill with the wave walking the
 * broad phase together and every lane queueing its own (link, obstacle)
 * pairs, so that a wave evaluates one pair per lane per step, or one launch
 * per link over the configurations still alive at it, in the walk order,
 * with the same queues) and workgroup size.  The obstacle set is always staged whole in LDS (every lane walks
 * every obstacle of its current link, so a smaller obstacle tile would only
 * add reloads); it must fit the device's workgroup LDS.  Variants 0, 1 and 2
 * evaluate the same pairs and give the same results bit for bit.  Variants 3
 * and 4 search for the first hit in obstacle order among other candidates
 * (more of them, or the walk positions after the hit), so that their
 * screening flags, and with them the refined set, can differ: the same
 * verdicts, and distances within FP32 rounding (a few 1e-7 m) of the other
 * variants'.  Each variant's results do not depend on its block size,
 * workgroups per CU or walk order. */
typedef struct rvs_tuning {
    int32_t variant;        /* 0: one lane per configuration; 1: persistent lane refill; 2: one lane per
                               configuration in persistent waves over 64-configuration chunks;
                               3: persistent lane refill, each lane queueing its own pairs (needs the
                               broad-phase group tables, and the cull boxes and pair queues in LDS
                               beside the obstacles; runs as 1 otherwise); 4: one launch per link
                               over the configurations alive at it, each lane queueing its own pairs
                               (the needs of 3 and 16 bytes of scratch per configuration; runs as 3
                               in rvs_distance_async_static and served requests) */
    int32_t block;          /* threads per workgroup: 64, 128 or 256; variants 3 and 4: 256, 512 or 1024 */
    int32_t blocks_per_cu;  /* variants 1 to 4: resident workgroups per CU (0 = as many as fit) */
    int32_t coherence_bits; /* walk order (0..18): configurations sorted by their first
                               min(dofs, 18 / bits) joint angles, 2^bits bins each, so that a wave's
                               lanes hold similar poses; 0 = the order given.  Results do not depend
                               on it. */
    float ms;               /* bulk-pass time on the sample (rvs_autotune) */
} rvs_tuning;

/* Times every candidate shape on the first min(n, 2097152) configurations of
 * q, keeps the fastest for later RVS_PREC_FP32_REFINE calls on (robot, env)
 * and returns it in best (may be NULL).  No counterpart in the reference. */
rvs_status rvs_autotune(rvs_context* ctx, int32_t robot, int32_t env, const float* q, int64_t n, uint32_t flags,
                        rvs_tuning* best);
rvs_status rvs_set_tuning(rvs_context* ctx, int32_t robot, int32_t env, const rvs_tuning* tuning);
rvs_status rvs_get_tuning(rvs_context* ctx, int32_t robot, int32_t env, rvs_tuning* tuning);

/* Batches of reference-arithmetic distances (RVS_PREC_REF32 calls, calls
 * with nearest points, the distances inside generalized-spine and served
 * underestimation calls) up to max_configs configurations run one wave per
 * configuration, lanes over the obstacles: lower latency for small batches,
 * the same results bit for bit.
 * Larger batches run one lane per configuration.  0 disables; default
 * 16384 (measured crossover on MI355X: 1 configuration 37 us vs 42 us per
 * call, 2048: 58 us vs 267 us, 32768: 429 us vs 218 us).  No counterpart in
 * the reference (a launch-shape choice). */
rvs_status rvs_set_small_batch(rvs_context* ctx, int64_t max_configs);

/* Launch shape of the persistent first-collision kernel: resident waves per
 * CU, and slots (edge states) per lane in one 64-lane work item, i.e. 64 *
 * slots_per_lane configurations per wavefront item. */
typedef struct rvs_edge_tuning {
    int32_t waves_per_cu;
    int32_t slots_per_lane;
    float ms;
} rvs_edge_tuning;
/* Times waves_per_cu in {4, 8, 16, 32} x slots_per_lane in {1, 2, 4} on the
 * given edges and keeps the fastest for later rvs_edge_first_collision calls
 * of the context.  No counterpart in the reference. */
rvs_status rvs_autotune_edges(rvs_context* ctx, int32_t env, const int32_t* robots, const float* q1, const float* dir,
                              const int32_t* K, int64_t num_edges, int32_t row_stride, float step, uint32_t flags,
                              rvs_edge_tuning* best);
rvs_status rvs_set_edge_tuning(rvs_context* ctx, const rvs_edge_tuning* tuning);

/* ---------------------------------------------------------------- serving
 * A device-resident request loop over one context (no counterpart in the
 * reference, whose planners call RealVectorSpace one state at a time).
 * Planner threads submit requests of any size; a server thread admits them
 * in arrival order while the device memory budget allows, fuses the admitted
 * requests of one (robot, environment, kind) into one launch, runs up to
 * num_streams launches concurrently, and keeps moving obstacles up to date
 * on the device.  Each configuration's result is the one rvs_distance /
 * rvs_distance_underestimation computes for it alone against the obstacle
 * positions of the request's tick (rvs_request_info.tick).
 */
typedef struct rvs_server rvs_server;

enum {
    /* computeDistance verdicts: d[n] (0 on collision), flag[n] = 1 when
     * collision-free, profile[n][L] optional; the server's precision. */
    RVS_REQ_VERDICT = 0,
    /* computeDistanceUnderestimation of every configuration against the
     * nearest points of the request's first configuration (computed by the
     * reference arithmetic): d[n], flag[n] = 1 when stored, profile optional. */
    RVS_REQ_UNDERESTIMATION = 1
};

typedef struct rvs_server_options {
    int64_t memory_budget;      /* device bytes for request state (default 8 GiB) */
    int32_t num_streams;        /* streams the launches are spread over (default 2) */
    int32_t max_batch_configs;  /* configurations per fused launch (default 1 << 20) */
    float tick_seconds;         /* obstacle update period (default 0.01 s) */
    int32_t precision;          /* of verdict requests (default RVS_PREC_FP32_REFINE) */
    int32_t launches_per_stream;/* launches in flight per stream (default 1: requests arriving
                                   meanwhile wait and join the next, larger launch) */
} rvs_server_options;

typedef struct rvs_request {
    int32_t robot;
    int32_t env;
    int32_t kind;               /* RVS_REQ_* */
    int32_t reserved;
    int64_t n;                  /* configurations */
    const float* q;             /* [n][dofs] host memory, untouched until the request is done */
    float* d;                   /* [n] host */
    float* profile;             /* [n][L] host, or NULL */
    uint8_t* flag;              /* [n] host */
} rvs_request;

typedef struct rvs_request_info {
    double submit_s;            /* seconds since the server started */
    double admit_s;             /* launched on the device */
    double done_s;              /* results in the caller's buffers */
    int64_t tick;               /* obstacle tick used (-1: environment without motion) */
    int32_t batch_requests;     /* requests fused into the same launch */
    int32_t status;             /* rvs_status of the request */
} rvs_request_info;

typedef struct rvs_server_stats {
    int64_t submitted, completed, configs_completed, batches, ticks;
    int32_t max_inflight_batches;  /* largest number of launches on the device at once */
    int32_t reserved;
    int64_t peak_bytes;            /* largest device memory in use by requests */
} rvs_server_stats;

void rvs_server_default_options(rvs_server_options* opt);
/* The context must outlive the server; other calls on the context stay
 * allowed (they are serialised with the server's use of it). */
rvs_status rvs_server_create(rvs_context* ctx, const rvs_server_options* opt, rvs_server** server);
/* Finishes every submitted request, then stops. */
void rvs_server_destroy(rvs_server* server);
/* From the next tick on, box j of env moves at vel[3j..3j+2] m/s:
 * pos = pos_at_set_up + vel * (tick * tick_seconds), tick counted from the
 * server's start.  Boxes with zero velocity stay put. */
rvs_status rvs_server_set_motion(rvs_server* server, int32_t env, const float* vel);
/* Non-blocking.  Host buffers (pinned ones from rvs_host_alloc give
 * asynchronous copies) must stay valid until the request is done. */
rvs_status rvs_server_submit(rvs_server* server, const rvs_request* req, uint64_t* ticket);
/* Waits up to timeout_ms (-1: no limit) for the request.  RVS_OK: done, info
 * filled (may be NULL) and the ticket released; RVS_EBUSY: not yet. */
rvs_status rvs_server_wait(rvs_server* server, uint64_t ticket, int32_t timeout_ms, rvs_request_info* info);
rvs_status rvs_server_get_stats(rvs_server* server, rvs_server_stats* stats);
/* Seconds since the server started (the clock of rvs_request_info). */
double rvs_server_now(rvs_server* server);
const char* rvs_server_last_error(rvs_server* server);
/* Pinned host memory (hipHostMalloc) for request buffers. */
void* rvs_host_alloc(int64_t bytes);
void rvs_host_free(void* p);

/* ---------------------------------------------------------------- context */
rvs_status rvs_create(const rvs_options* opt, rvs_context** ctx);
void rvs_destroy(rvs_context* ctx);
const char* rvs_last_error(const rvs_context* ctx);
/* The HIP stream (hipStream_t) the context launches on. */
void* rvs_stream(rvs_context* ctx);
/* Statistics of the last RVS_PREC_FP32_REFINE query: configurations
 * recomputed by the refinement pass. */
int64_t rvs_last_refined(const rvs_context* ctx);
/* Stage durations of the last call made with RVS_F_TIMING, in ms, measured
 * with HIP events on the context's stream.  rvs_distance reports
 * [0] the bulk kernel (k_distance_fast for RVS_PREC_FP32_REFINE, the exact
 * kernel otherwise), [1] the refinement kernel (0 when there is none) and,
 * for RVS_PREC_FP32_REFINE, [2] the walk-order kernels (coherence_bits of
 * rvs_tuning; 0 when off).  Returns the number of stages written. */
int32_t rvs_last_timing(const rvs_context* ctx, float* stage_ms, int32_t max_stages);
/* Library version string. */
const char* rvs_version(void);

#ifdef __cplusplus
}
#endif
#endif /* RVS_H */
