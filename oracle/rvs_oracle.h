/*
 * CPU oracle for the capsule collision / distance path of RPMPLv2's
 * RealVectorSpace.  TEST INFRASTRUCTURE ONLY: this library is the checker
 * used by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline
 * leg.  The product (rpmplv2_amd/librvs.so) never links, loads or calls it.
 *
 * Two instantiations of one restatement:
 *   ref32 - reference-faithful: double-precision KDL-style FK, skeleton
 *           rounded to float, all geometry in float in the reference's
 *           operation order (no FMA contraction, IEEE div/sqrt).
 *   ref64 - the same algorithm with the skeleton and all geometry in double.
 *
 * Parity: segment distances, the capsule-box tree (where it is exact), FK
 * closed forms, the loop semantics (reference scenario data) and the
 * oriented-box extension are pinned by tests/test_oracle.py.  Parity
 * unpinned: Eigen 3.4's FP32 reduction order, KDL's FP64 frame rounding,
 * FCL's rotated-box AABB rule and the reference tree's non-exact quirks
 * (corner-index test, face/edge lists, 0 vs -r contact outputs); they are
 * restated from the source without an executable reference.
 */
#ifndef RVS_ORACLE_H
#define RVS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_SEG 72
#define ORC_MAX_LINKS 64

enum { ORC_JOINT_FIXED = 0, ORC_JOINT_REVOLUTE = 1 };
enum { ORC_SKEL_CHAIN = 0, ORC_SKEL_XARM6 = 1 };

typedef struct orc_joint {
    int32_t type;        /* ORC_JOINT_FIXED / ORC_JOINT_REVOLUTE */
    int32_t pad;
    double xyz[3];       /* URDF joint origin xyz */
    double rpy[3];       /* URDF joint origin rpy */
    double axis[3];      /* URDF joint axis (joint frame) */
} orc_joint;

typedef struct orc_robot {
    int32_t skeleton_kind;     /* ORC_SKEL_CHAIN or ORC_SKEL_XARM6 */
    int32_t num_segments;      /* chain segments root->tip, incl. fixed ones */
    int32_t num_dofs;          /* revolute joints */
    int32_t num_links;         /* capsules */
    int32_t ground_included;
    float gripper_length;
    float radius[ORC_MAX_LINKS];
    orc_joint joints[ORC_MAX_SEG];
} orc_robot;

typedef struct orc_params {
    float eq_threshold;        /* RealVectorSpaceConfig::EQUALITY_THRESHOLD */
    float max_distance;        /* RealVectorSpaceConfig::MAX_DISTANCE */
    int32_t num_checks;        /* NUM_INTERPOLATION_VALIDITY_CHECKS */
} orc_params;

typedef struct orc_scene {
    int32_t num_obj;
    int32_t pad;
    const float* obs;          /* num_obj x 6 : xmin ymin zmin xmax ymax zmax */
    const uint8_t* is_ground;  /* num_obj */
    const float* min_dist_tol; /* num_obj */
    const float* obb;          /* NULL, or num_obj x 15: centre, rotation rows, half extents
                                  (oriented-box extension: capsules are moved into the box frame) */
} orc_scene;

/* ---- scene ---- */
void orc_box_aabb(const float dim[3], const float pos[3], const float quat_wxyz[4], float out6[6]);

/* ---- kinematics ---- */
int orc_skeleton32(const orc_robot* rb, const float* q, int64_t n, float* skel);
int orc_skeleton64(const orc_robot* rb, const float* q, int64_t n, double* skel);

/* ---- pair geometry (single pair) ---- */
float orc_seg_seg32(const float A[3], const float B[3], const float C[3], const float D[3],
                    const orc_params* p, float np[6], int32_t* np_valid);
double orc_seg_seg64(const double A[3], const double B[3], const double C[3], const double D[3],
                     const orc_params* p, double np[6], int32_t* np_valid);
float orc_seg_pt32(const float A[3], const float B[3], const float C[3],
                   const orc_params* p, float np[6], int32_t* np_valid);
double orc_seg_pt64(const double A[3], const double B[3], const double C[3],
                    const orc_params* p, double np[6], int32_t* np_valid);
float orc_capsule_box32(const float A[3], const float B[3], float r, const float obs[6],
                        const orc_params* p, float np[6], int32_t* np_valid, int32_t* path);
double orc_capsule_box64(const double A[3], const double B[3], double r, const double obs[6],
                         const orc_params* p, double np[6], int32_t* np_valid, int32_t* path);
int32_t orc_collision_capsule_box32(const float A[3], const float B[3], float r, const float obs[6],
                                    const orc_params* p);
int32_t orc_collision_capsule_box64(const double A[3], const double B[3], double r, const double obs[6],
                                    const orc_params* p);

/* ---- batched RealVectorSpace loops (static partition over nthreads) ---- */
void orc_distance32(const orc_robot* rb, const orc_params* p, const orc_scene* sc,
                    const float* q, int64_t n, float* d, float* profile, float* np,
                    uint8_t* verdict, int32_t nthreads);
void orc_distance64(const orc_robot* rb, const orc_params* p, const orc_scene* sc,
                    const float* q, int64_t n, double* d, double* profile, double* np,
                    uint8_t* verdict, int32_t nthreads);
void orc_is_valid32(const orc_robot* rb, const orc_params* p, const orc_scene* sc,
                    const float* q, int64_t n, uint8_t* valid, int32_t nthreads);
void orc_is_valid64(const orc_robot* rb, const orc_params* p, const orc_scene* sc,
                    const float* q, int64_t n, uint8_t* valid, int32_t nthreads);
void orc_underestimation32(const orc_robot* rb, const orc_params* p, const orc_scene* sc,
                           const float* q, int64_t n, const float* np_ref, int64_t np_stride,
                           float* d, float* profile, uint8_t* stored, int32_t nthreads);
void orc_self_distance32(const orc_robot* rb, const orc_params* p, const float* q, int64_t n,
                         const int32_t* pairs, int32_t npairs, float* dist, int32_t nthreads);
void orc_self_distance64(const orc_robot* rb, const orc_params* p, const float* q, int64_t n,
                         const int32_t* pairs, int32_t npairs, double* dist, int32_t nthreads);
void orc_edge_valid32(const orc_robot* rb, const orc_params* p, const orc_scene* sc,
                      const float* q1, const float* q2, int64_t n, uint8_t* valid, int32_t nthreads);
void orc_edge_first_collision32(const orc_robot* rb, const orc_params* p, const orc_scene* sc,
                                const int32_t* pairs, int32_t npairs, float step,
                                const float* q1, const float* dir, const int32_t* K, int64_t E,
                                int32_t* first, int32_t nthreads);

/* oriented-box extension: obb = centre(3), rotation rows(9), half extents(3) */
float orc_capsule_obb32(const float A[3], const float B[3], float r, const float obb[15], const orc_params* p,
                        float np[6], int32_t* np_valid);
double orc_capsule_obb64(const double A[3], const double B[3], double r, const double obb[15], const orc_params* p,
                         double np[6], int32_t* np_valid);
void orc_self_min32(const orc_robot* rb, const orc_params* p, const float* q, int64_t n, const int32_t* pairs,
                    int32_t npairs, float* out, int32_t nthreads);

/* computeDistance walked as the engine's bulk pass walks it: each link starts
 * at the obstacle that held the previous link's minimum, then walks the
 * obstacles in the engine's walk order (Morton order of the cull box centres,
 * groups of 4 tested by their union box first), skipping boxes farther from
 * the segment's bounding box than max(running link minimum + r,
 * EQUALITY_THRESHOLD) + 1e-5; a hit found out of obstacle order is replaced
 * by the first one in order, searched (search 0) among the obstacles before
 * it in obstacle order, as the bulk kernels of tuning variants 0-3 do, or
 * (search 1) among the walk positions after it, as the link-major kernel
 * does.  Results equal orc_distance32/64; counts = {pairs evaluated,
 * bounding-box tests (groups and obstacles)}. */
void orc_distance_cull32(const orc_robot* rb, const orc_params* p, const orc_scene* sc, const float* q, int64_t n,
                         float* d, float* profile, uint8_t* verdict, int64_t counts[2], int32_t nthreads,
                         int32_t search);
void orc_distance_cull64(const orc_robot* rb, const orc_params* p, const orc_scene* sc, const float* q, int64_t n,
                         double* d, double* profile, uint8_t* verdict, int64_t counts[2], int32_t nthreads,
                         int32_t search);

/* AbstractRobot::computeEnclosingRadii (xArm6 / spatial chains), float:
 * R [n][dofs][dofs+1]. */
void orc_enclosing_radii32(const orc_robot* rb, const float* q, int64_t n, float* R, int32_t nthreads);

/* AbstractTrajectory::computeDistanceUnderestimation(q, np, delta_t): the
 * planes move by max_vel[j] * delta_t towards the robot first. */
void orc_underestimation_moving32(const orc_robot* rb, const orc_params* p, const orc_scene* sc, const float* q,
                                  int64_t n, const float* np_ref, int64_t np_stride, const float* max_vel,
                                  float delta_t, float* d, float* profile, uint8_t* stored, int32_t nthreads);

/* AbstractTrajectory::isSafe for num_traj trajectories [num_traj][num_points][dofs]
 * leaving q_current (profile, nearest points [M][L][6]); pairs = the
 * self-collision pairs to test at every point (npairs 0: none). */
void orc_traj_safe32(const orc_robot* rb, const orc_params* p, const orc_scene* sc, const float* points,
                     int64_t num_traj, int32_t num_points, const float* q_current, const float* profile_current,
                     const float* nearest_current, const float* max_vel, float t_iter0, float time_step,
                     const int32_t* pairs, int32_t npairs, uint8_t* safe, int32_t nthreads);

/* Environment::updateEnvironment(dt) 'steps' times for the deterministic
 * motion laws (0 straight, 1 circular, 2 two tunnels, 3 light directions). */
void orc_env_step32(int32_t law, const float ws[3], float ws_radius, float base_radius, float robot_max_vel,
                    int32_t M, float* pos, float* vel, const float* max_vel, const uint8_t* dynamic,
                    float* path_len, int32_t* sign, float dt, int32_t steps);

/* Capsule part of xArm6::checkSelfCollision(q1, q2) for n edges: collide 1
 * (q_out = last state before the hit), 0 (q_out = q2) or 255 (walk capped). */
void orc_edge_self_collision32(const orc_robot* rb, const orc_params* p, const float* q1, const float* q2,
                               int64_t n, float eps_step, uint8_t* collide, float* q_out, int32_t nthreads);

/* RBTConnect::extendSpine for n spines (robot without self-collision
 * checking), float: q_new [n][dofs], status 1 advanced / 2 trapped /
 * 3 reached. */
void orc_extend_spines32(const orc_robot* rb, const float* q, const float* d, const float* profile,
                         const float* q_e, int64_t n, int32_t num_iter, int32_t expanded, float* q_new,
                         int32_t* status, int32_t nthreads);

/* ---- independent checker: exact segment-AABB distance in double ---- */
double orc_true_seg_box(const double A[3], const double B[3], const double obs[6]);

#ifdef __cplusplus
}
#endif
#endif
