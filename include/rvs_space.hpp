// rvs_space.hpp - C++ host class over the C ABI (rvs.h) that keeps the
// RealVectorSpace interface of RPMPLv2, so a planner written against
// base::RealVectorSpace can call the GPU engine with the same method names,
// arguments and State side effects.
//
// Mirrors (reference files):
//   base::StateSpace pure virtuals           include/state_spaces/StateSpace.h:52-56
//   base::RealVectorSpace overrides          include/state_spaces/real_vector_space/RealVectorSpace.h:45-49
//   isValid(q1, q2)                          src/state_spaces/real_vector_space/RealVectorSpace.cpp:261-275
//   isValid(q)                               RealVectorSpace.cpp:277-309
//   computeDistance(q, compute_again)        RealVectorSpace.cpp:318-392 (cache test :320-321)
//   computeDistanceUnderestimation(q, NP)    RealVectorSpace.cpp:402-446 (cache test :405-406)
//   State fields d_c / d_c_profile /
//     is_real_d_c / nearest_points           include/state_spaces/State.h:25-41, src/state_spaces/State.cpp:9-11
//   robot->checkSelfCollision(q), capsule part  src/robots/xArm6.cpp:374-424
//
// Differences that are deliberate:
//   - nearest points are one flat float array [M][L][6] per State (column i
//     of the reference's Eigen::MatrixXf(6, L) for obstacle j is
//     nearest[(j * L + i) * 6 .. + 6], robot xyz then obstacle xyz);
//   - a failing device call throws rvs::Error (the reference's CPU path has
//     no failure mode to report); results are never silently substituted;
//   - the object is re-entrant: one context serialises its callers;
//   - batch overloads take N configurations at once (host or device arrays).
#ifndef RVS_SPACE_HPP
#define RVS_SPACE_HPP

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "rvs.h"

namespace rvs {

class Error : public std::runtime_error {
  public:
    Error(rvs_status s, const std::string& what) : std::runtime_error(what), status(s) {}
    rvs_status status;
};

// The parts of base::State that RealVectorSpace reads and writes.
class State {
  public:
    explicit State(std::vector<float> coord) : coord_(std::move(coord)) {}
    const std::vector<float>& getCoord() const { return coord_; }
    float getDistance() const { return d_c_; }
    const std::vector<float>& getDistanceProfile() const { return d_c_profile_; }
    bool getIsRealDistance() const { return is_real_d_c_; }
    std::shared_ptr<std::vector<float>> getNearestPoints() const { return nearest_points_; }
    void setDistance(float d) { d_c_ = d; }
    void setDistanceProfile(std::vector<float> p) { d_c_profile_ = std::move(p); }
    void setIsRealDistance(bool r) { is_real_d_c_ = r; }
    void setNearestPoints(std::shared_ptr<std::vector<float>> np) { nearest_points_ = std::move(np); }

  private:
    std::vector<float> coord_;
    float d_c_ = -1.0f;            // State.cpp:9
    std::vector<float> d_c_profile_;
    bool is_real_d_c_ = true;      // State.cpp:11
    std::shared_ptr<std::vector<float>> nearest_points_;
};

class RealVectorSpaceGPU {
  public:
    // One robot and one obstacle set (semantics RVS_ENV_AABB = the
    // reference's FCL-AABB reading of every box, RVS_ENV_OBB = oriented
    // boxes).  opt = nullptr takes rvs_default_options().
    RealVectorSpaceGPU(const rvs_robot_desc& robot, const std::vector<rvs_box>& boxes, int32_t semantics = RVS_ENV_AABB,
                       const rvs_options* opt = nullptr)
        : num_links_(robot.num_links), num_objects_((int32_t)boxes.size()) {
        rvs_options o;
        if (opt) o = *opt;
        else rvs_default_options(&o);
        check(rvs_create(&o, &ctx_), "rvs_create");
        int32_t dofs = 0;
        for (int32_t k = 0; k < robot.num_joints; ++k) dofs += robot.joints[k].type == RVS_JOINT_REVOLUTE;
        num_dofs_ = dofs;
        check(rvs_robot_load(ctx_, &robot, &robot_), "rvs_robot_load");
        check(rvs_env_create(ctx_, boxes.data(), num_objects_, semantics, &env_), "rvs_env_create");
    }
    ~RealVectorSpaceGPU() { rvs_destroy(ctx_); }
    RealVectorSpaceGPU(const RealVectorSpaceGPU&) = delete;
    RealVectorSpaceGPU& operator=(const RealVectorSpaceGPU&) = delete;

    int32_t getNumLinks() const { return num_links_; }
    int32_t getNumDimensions() const { return num_dofs_; }
    int32_t getNumObjects() const { return num_objects_; }
    rvs_context* context() const { return ctx_; }

    // ---------------------------------------------------- single states
    bool isValid(const std::shared_ptr<State>& q) {
        uint8_t v = 0;
        check(rvs_is_valid(ctx_, robot_, env_, coords(q), 1, &v, 0), "rvs_is_valid");
        return v != 0;
    }

    bool isValid(const std::shared_ptr<State>& q1, const std::shared_ptr<State>& q2) {
        uint8_t v = 0;
        check(rvs_edge_is_valid(ctx_, robot_, env_, coords(q1), coords(q2), 1, &v, 0), "rvs_edge_is_valid");
        return v != 0;
    }

    float computeDistance(const std::shared_ptr<State>& q, bool compute_again = false) {
        if (!compute_again && q->getDistance() > 0 && q->getIsRealDistance()) return q->getDistance();
        const size_t L = (size_t)num_links_;
        float d = 0.0f;
        std::vector<float> prof(L);
        auto np = std::make_shared<std::vector<float>>((size_t)num_objects_ * L * 6);
        uint8_t verdict = 0;
        check(rvs_distance(ctx_, robot_, env_, coords(q), 1, &d, prof.data(), np->data(), &verdict,
                           RVS_F_PREC(RVS_PREC_REF32)),
              "rvs_distance");
        q->setDistance(d);
        q->setDistanceProfile(std::move(prof));
        q->setIsRealDistance(true);
        q->setNearestPoints(verdict ? np : nullptr);   // collision: nullptr (RealVectorSpace.cpp:376)
        return d;
    }

    float computeDistanceUnderestimation(const std::shared_ptr<State>& q,
                                         const std::shared_ptr<std::vector<float>>& nearest_points) {
        if (q->getDistance() > 0 && q->getIsRealDistance()) return q->getDistance();
        if (!nearest_points || nearest_points->size() != (size_t)num_objects_ * num_links_ * 6)
            throw Error(RVS_EINVAL, "computeDistanceUnderestimation: nearest_points must hold [M][L][6] floats");
        float d = 0.0f;
        std::vector<float> prof((size_t)num_links_);
        uint8_t stored = 0;
        check(rvs_distance_underestimation(ctx_, robot_, env_, coords(q), 1, nearest_points->data(), 0, &d,
                                           prof.data(), &stored, 0),
              "rvs_distance_underestimation");
        if (!stored) return 0.0f;   // a negative plane distance: returns 0, State untouched (:428-429)
        q->setDistance(d);
        q->setDistanceProfile(std::move(prof));
        q->setIsRealDistance(false);
        return d;
    }

    // Capsule part of robot->checkSelfCollision(q): true when some listed
    // link pair has capsule distance < 0.
    bool checkSelfCollision(const std::shared_ptr<State>& q) {
        uint8_t free_flag = 1;
        check(rvs_self_collision(ctx_, robot_, coords(q), 1, nullptr, &free_flag, 0), "rvs_self_collision");
        return free_flag == 0;
    }

    // AbstractRobot::computeEnclosingRadii(q): R [dofs][dofs+1], row-major
    // (row i = starting skeleton point, column j = final point).
    std::vector<float> computeEnclosingRadii(const std::shared_ptr<State>& q) {
        std::vector<float> R((size_t)num_dofs_ * (num_dofs_ + 1));
        check(rvs_enclosing_radii(ctx_, robot_, coords(q), 1, R.data(), 0), "rvs_enclosing_radii");
        return R;
    }

    // RBTConnect::extendSpine(q, q_e) (RBTConnect.cpp:85-150) for a robot
    // without self-collision checking: (status RVS_SPINE_*, q_new); q gets its
    // distance first when it has none (:88-89); a trapped spine returns q.
    std::pair<int32_t, std::shared_ptr<State>> extendSpine(const std::shared_ptr<State>& q,
                                                           const std::shared_ptr<State>& q_e, int32_t num_iter = 5,
                                                           bool use_expanded_bubble = true) {
        if (q->getDistance() < 0) computeDistance(q);
        const float d = q->getDistance();
        std::vector<float> prof = q->getDistanceProfile();
        prof.resize((size_t)num_links_, 0.0f);
        std::vector<float> qn((size_t)num_dofs_);
        int32_t st = 0;
        check(rvs_extend_spines(ctx_, robot_, coords(q), &d, prof.data(), coords(q_e), 1, num_iter,
                                use_expanded_bubble ? 1 : 0, qn.data(), &st, 0),
              "rvs_extend_spines");
        if (st == RVS_SPINE_TRAPPED) return {st, q};
        return {st, std::make_shared<State>(std::move(qn))};
    }

    // Capsule part of xArm6::checkSelfCollision(q1, q2) (xArm6.cpp:284-368):
    // true on a capsule overlap along the edge, with q2 replaced by the last
    // state before it (the reference's modified q2).
    bool checkSelfCollision(const std::shared_ptr<State>& q1, std::shared_ptr<State>& q2, float eps_step = 0.1f) {
        uint8_t hit = 0;
        std::vector<float> out((size_t)num_dofs_);
        check(rvs_edge_self_collision(ctx_, robot_, coords(q1), coords(q2), 1, eps_step, &hit, out.data(), 0),
              "rvs_edge_self_collision");
        if (hit == 255) throw Error(RVS_ERANGE, "checkSelfCollision: the edge walk would not end");
        if (hit) q2 = std::make_shared<State>(std::move(out));
        return hit != 0;
    }

    // AbstractTrajectory::isSafe(pos_points, q_current, t_iter, time_step)
    // (AbstractTrajectory.cpp:212-256) for one trajectory of num_points
    // positions [num_points][dofs]; q_current needs its real distance
    // (computed here when missing) and max_vel holds one speed per obstacle.
    bool isSafe(const float* points, int32_t num_points, const std::shared_ptr<State>& q_current,
                const std::vector<float>& max_vel, float t_iter, float time_step = 0.01f,
                bool self_collision = false) {
        if ((int32_t)max_vel.size() != num_objects_) throw Error(RVS_EINVAL, "isSafe: one max_vel per obstacle");
        if (!q_current->getNearestPoints() || !q_current->getIsRealDistance()) computeDistance(q_current, true);
        if (!q_current->getNearestPoints()) return false;   // the current state is in collision
        uint8_t safe = 0;
        check(rvs_trajectory_is_safe(ctx_, robot_, env_, points, 1, num_points, coords(q_current),
                                     q_current->getDistanceProfile().data(), q_current->getNearestPoints()->data(),
                                     max_vel.data(), t_iter, time_step, self_collision ? 1 : 0, &safe, 0),
              "rvs_trajectory_is_safe");
        return safe != 0;
    }

    // ---------------------------------------------------- batches
    // Arrays are host memory, or device memory with RVS_F_DEVICE in flags.
    void isValid(const float* q, int64_t n, uint8_t* valid, uint32_t flags = 0) {
        check(rvs_is_valid(ctx_, robot_, env_, q, n, valid, flags), "rvs_is_valid");
    }
    void isValid(const float* q1, const float* q2, int64_t n, uint8_t* valid, uint32_t flags = 0) {
        check(rvs_edge_is_valid(ctx_, robot_, env_, q1, q2, n, valid, flags), "rvs_edge_is_valid");
    }
    void computeDistance(const float* q, int64_t n, float* d, float* profile, float* nearest, uint8_t* verdict,
                         uint32_t flags = 0) {
        check(rvs_distance(ctx_, robot_, env_, q, n, d, profile, nearest, verdict, flags), "rvs_distance");
    }
    void computeDistance(const float* q, int64_t n, double* d, double* profile, uint8_t* verdict, uint32_t flags = 0) {
        check(rvs_distance_f64(ctx_, robot_, env_, q, n, d, profile, verdict, flags), "rvs_distance_f64");
    }
    void computeDistanceUnderestimation(const float* q, int64_t n, const float* nearest_ref, int64_t ref_stride,
                                        float* d, float* profile, uint8_t* stored, uint32_t flags = 0) {
        check(rvs_distance_underestimation(ctx_, robot_, env_, q, n, nearest_ref, ref_stride, d, profile, stored, flags),
              "rvs_distance_underestimation");
    }
    void checkSelfCollision(const float* q, int64_t n, float* min_dist, uint8_t* free_flag, uint32_t flags = 0) {
        check(rvs_self_collision(ctx_, robot_, q, n, min_dist, free_flag, flags), "rvs_self_collision");
    }

    void computeEnclosingRadii(const float* q, int64_t n, float* R, uint32_t flags = 0) {
        check(rvs_enclosing_radii(ctx_, robot_, q, n, R, flags), "rvs_enclosing_radii");
    }
    void extendSpines(const float* q, const float* d, const float* profile, const float* q_e, int64_t n,
                      int32_t num_iter, bool use_expanded_bubble, float* q_new, int32_t* status, uint32_t flags = 0) {
        check(rvs_extend_spines(ctx_, robot_, q, d, profile, q_e, n, num_iter, use_expanded_bubble ? 1 : 0, q_new,
                                status, flags),
              "rvs_extend_spines");
    }
    void extendGenSpines(const float* q_root, const float* q_e, int64_t n, int32_t num_layers, int32_t num_iter,
                         bool use_expanded_bubble, bool use_real_dist, float d_crit, float* q_new, int32_t* status,
                         float* d, uint32_t flags = 0) {
        check(rvs_extend_gen_spines(ctx_, robot_, env_, q_root, q_e, n, num_layers, num_iter,
                                    use_expanded_bubble ? 1 : 0, use_real_dist ? 1 : 0, d_crit, q_new, status, d,
                                    flags),
              "rvs_extend_gen_spines");
    }
    void checkSelfCollision(const float* q1, const float* q2, int64_t n, float eps_step, uint8_t* collide,
                            float* q2_out, uint32_t flags = 0) {
        check(rvs_edge_self_collision(ctx_, robot_, q1, q2, n, eps_step, collide, q2_out, flags),
              "rvs_edge_self_collision");
    }

    // Environment::updateEnvironment(delta_time), steps times, with the
    // motion law set by setMotionLaw; pos / vel [M][3] may be null.
    void setMotionLaw(const rvs_env_motion& law, const std::vector<float>& vel, const std::vector<float>& max_vel) {
        if (vel.size() != (size_t)num_objects_ * 3 || max_vel.size() != (size_t)num_objects_)
            throw Error(RVS_EINVAL, "setMotionLaw: vel [M][3] and max_vel [M]");
        check(rvs_env_set_motion_law(ctx_, env_, &law, vel.data(), max_vel.data()), "rvs_env_set_motion_law");
    }
    void updateEnvironment(float delta_time, int32_t steps = 1, float* pos = nullptr, float* vel = nullptr) {
        check(rvs_env_step(ctx_, env_, delta_time, steps, pos, vel), "rvs_env_step");
    }

    // Move obstacles in place (Environment::updateEnvironment's position
    // update for dynamic obstacles).
    void updateObstacles(const std::vector<int32_t>& idx, const std::vector<float>& pos_xyz) {
        if (pos_xyz.size() != idx.size() * 3) throw Error(RVS_EINVAL, "updateObstacles: pos must hold 3 floats per index");
        check(rvs_env_update(ctx_, env_, idx.data(), pos_xyz.data(), (int32_t)idx.size(), 0), "rvs_env_update");
    }

  private:
    const float* coords(const std::shared_ptr<State>& q) const {
        if (!q || (int32_t)q->getCoord().size() != num_dofs_)
            throw Error(RVS_EINVAL, "state dimension does not match the robot");
        return q->getCoord().data();
    }
    void check(rvs_status s, const char* what) const {
        if (s != RVS_OK) throw Error(s, std::string(what) + ": " + (ctx_ ? rvs_last_error(ctx_) : "no context"));
    }

    rvs_context* ctx_ = nullptr;
    int32_t robot_ = -1, env_ = -1;
    int32_t num_links_ = 0, num_dofs_ = 0, num_objects_ = 0;
};

}  // namespace rvs

#endif
