// collision_amd::CollisionSolver3d -- C++ host API over the C ABI
// (include/collision_amd.h), with the reference's method names and meaning
// (collid.h:128-244).  Where the reference walks a FronTier INTERFACE
// (assembleFromInterface, dcollid3d.cpp:12-52) this class takes a flat mesh;
// where the reference mutates POINT/STATE in place, state moves through
// setState / getState.  Errors throw collision_amd::Error carrying the
// cs_status instead of exiting the process (the reference's clean_up(ERROR)).
//
// Header-only; link with -lcollision_amd.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "../collision_amd.h"

namespace collision_amd {

class Error : public std::runtime_error {
 public:
  Error(int status, const std::string& msg) : std::runtime_error(msg), status_(status) {}
  int status() const { return status_; }

 private:
  int status_;
};

class CollisionSolver3d {
 public:
  CollisionSolver3d() {
    cs_default_params(&p_);
    check_create(cs_create(&h_, &p_));
  }
  explicit CollisionSolver3d(const cs_params& p) : p_(p) { check_create(cs_create(&h_, &p_)); }
  ~CollisionSolver3d() { cs_destroy(h_); }
  CollisionSolver3d(const CollisionSolver3d&) = delete;
  CollisionSolver3d& operator=(const CollisionSolver3d&) = delete;

  // parameter setters / getters (dcollid.cpp:56-89); per instance here
  void setRoundingTolerance(double v) { p_.round_eps = v; push(); }
  double getRoundingTolerance() const { return p_.round_eps; }
  void setFabricThickness(double v) { p_.thickness = v; push(); }
  double getFabricThickness() const { return p_.thickness; }
  void setTimeStepSize(double v) { p_.dt = v; push(); }
  double getTimeStepSize() const { return p_.dt; }
  void setSpringConstant(double v) { p_.spring_k = v; push(); }
  double getSpringConstant() const { return p_.spring_k; }
  void setFrictionConstant(double v) { p_.friction = v; push(); }
  double getFrictionConstant() const { return p_.friction; }
  void setPointMass(double v) { p_.mass = v; push(); }
  double getPointMass() const { return p_.mass; }
  void setRestitutionCoef(double v) { p_.restitution = v; push(); }
  double getRestitutionCoef() const { return p_.restitution; }
  // not in the reference: per-pass timing points (2 phases and kernels, 1 kernels, 0 none);
  // each is a marker on the stream and costs GPU time
  void setPassTiming(int level) { p_.pass_timing = level; push(); }
  int getPassTiming() const { return p_.pass_timing; }
  void setDomainBoundary(const double* L, const double* U) {
    check(cs_set_domain_boundary(h_, L, U));
    for (int i = 0; i < 3; ++i) {
      p_.dom_lo[i] = L[i];
      p_.dom_hi[i] = U[i];
    }
  }
  double getDomainBoundary(int dir, int side) const {
    return side == 0 ? p_.dom_lo[dir] : p_.dom_hi[dir];
  }

  // topology (assembleFromInterface) and state
  void assembleFromMesh(const cs_mesh_view& m) {
    check(cs_assemble(h_, &m));
    nv_ = m.nv;
  }
  void setState(const double* x, const double* vel, const uint8_t* has_collsn = nullptr) {
    check(cs_set_state(h_, x, vel, has_collsn));
  }
  void getState(double* x, double* vel, double* avg_vel = nullptr, double* x_old = nullptr,
                uint8_t* has_collsn = nullptr) {
    check(cs_get_state(h_, x, vel, avg_vel, x_old, has_collsn));
  }
  void recordOriginPosition() { check(cs_record_origin_position(h_)); }

  // the driver's predictor (test.cpp:226-258) and resolveCollision
  void predict() { check(cs_predict(h_)); }
  void resolveCollision() {
    const int rc = cs_resolve(h_, &report_);
    if (rc != CS_OK && rc != CS_WARN_ZONE_CAP) check(rc);
  }
  bool hasCollision() const { return cs_has_collision(h_) != 0; }
  const cs_step_report& lastReport() const { return report_; }
  int numVertices() const { return nv_; }
  cs_solver* handle() { return h_; }

 private:
  void push() { check(cs_set_params(h_, &p_)); }
  void check(int rc) const {
    if (rc != CS_OK) throw Error(rc, cs_last_error(h_));
  }
  static void check_create(int rc) {
    if (rc != CS_OK) throw Error(rc, cs_last_error(nullptr));
  }
  cs_params p_{};
  cs_solver* h_ = nullptr;
  cs_step_report report_{};
  int nv_ = 0;
};

}  // namespace collision_amd
