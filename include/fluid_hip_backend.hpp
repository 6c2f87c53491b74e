// fluid_hip_backend.hpp — the `hip` backend for the reference's C++ solver
// interface, header-only, over the C ABI in fluid_hip.h.
//
// Drop-in for the reference's CUDA backend (cuda/include/fluid/cuda.h:7-9):
//
//   int  fluid::hip_version();                                  // cuda_version
//   bool fluid::hip_device_available(int* count, const char**); // cuda_device_available
//   void fluid::hip_step(const Params&, State&);                // cuda_step
//
// The functions are templates over the solver types, so this header builds
// against the reference's own `fluid::Params` / `fluid::State`
// (core/include/fluid/core.h:11-132) without restating them: any types with
// the same field names work.  Include it after "fluid/core.h".
//
// hip_step has cuda_step's call shape and lifetime (everything is set up and
// torn down inside the call, state crosses as host SoA vectors).
// fluid::HipSession is the device-resident form: one upload, many steps,
// downloads only when the caller asks (e.g. at output frames).
//
// Errors: cuda_step returns void and does not check CUDA errors
// (cuda/src/cuda_stub.cu:764-1099); hip_step keeps the void signature but
// throws std::runtime_error carrying fluid_hip_last_error() instead of
// continuing silently.
#pragma once

#include <cstddef>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "fluid_hip.h"

namespace fluid {

inline int hip_version() { return fluid_hip_version(); }

inline bool hip_device_available(int* count, const char** error) {
  return fluid_hip_device_available(count, error) != 0;
}

namespace hip_detail {

inline void check(int rc, const char* what) {
  if (rc != FLUID_HIP_OK)
    throw std::runtime_error(std::string(what) + ": " + fluid_hip_last_error());
}

// Params -> the flat C struct (field for field; core/include/fluid/core.h:11-79).
template <class P>
fluid_hip_params to_c(const P& p) {
  fluid_hip_params q;
  std::memset(&q, 0, sizeof(q));
  q.dt = p.dt;
  q.density = p.density;
  q.particle_mass = p.particle_mass;
  q.h = p.h;
  q.particle_radius = p.particle_radius;
  q.epsilon = p.epsilon;
  q.solver_iterations = p.solver_iterations;
  q.use_uniform_grid = p.use_uniform_grid ? 1 : 0;
  q.enable_scorr = p.enable_scorr ? 1 : 0;
  q.enable_xsph = p.enable_xsph ? 1 : 0;
  q.enable_vorticity = p.enable_vorticity ? 1 : 0;
  q.scorr_k = p.scorr_k;
  q.scorr_n = p.scorr_n;
  q.scorr_dq_coeff = p.scorr_dq_coeff;
  q.visc_c = p.visc_c;
  q.plane_restitution = p.plane_restitution;
  q.plane_friction = p.plane_friction;
  q.vort_epsilon = p.vort_epsilon;
  q.vort_norm_eps = p.vort_norm_eps;
  q.gravity[0] = p.external_forces.x;
  q.gravity[1] = p.external_forces.y;
  q.gravity[2] = p.external_forces.z;
  const std::size_t np = p.planes.size();
  if (np > FLUID_HIP_MAX_PLANES)
    throw std::runtime_error("hip backend: at most " + std::to_string(FLUID_HIP_MAX_PLANES) +
                             " planes");
  q.n_planes = static_cast<int32_t>(np);
  for (std::size_t i = 0; i < np; ++i) {
    q.planes[i][0] = p.planes.nx[i];
    q.planes[i][1] = p.planes.ny[i];
    q.planes[i][2] = p.planes.nz[i];
    q.planes[i][3] = p.planes.d[i];
  }
  return q;
}

template <class S>
void pack(const S& s, std::vector<float>& pos, std::vector<float>& vel) {
  const std::size_t n = s.pos_x.size();
  pos.resize(3 * n);
  vel.resize(3 * n);
  std::memcpy(pos.data(), s.pos_x.data(), n * sizeof(float));
  std::memcpy(pos.data() + n, s.pos_y.data(), n * sizeof(float));
  std::memcpy(pos.data() + 2 * n, s.pos_z.data(), n * sizeof(float));
  std::memcpy(vel.data(), s.vel_x.data(), n * sizeof(float));
  std::memcpy(vel.data() + n, s.vel_y.data(), n * sizeof(float));
  std::memcpy(vel.data() + 2 * n, s.vel_z.data(), n * sizeof(float));
}

template <class S>
void unpack(const std::vector<float>& pos, const std::vector<float>& vel, S& s) {
  const std::size_t n = s.pos_x.size();
  std::memcpy(s.pos_x.data(), pos.data(), n * sizeof(float));
  std::memcpy(s.pos_y.data(), pos.data() + n, n * sizeof(float));
  std::memcpy(s.pos_z.data(), pos.data() + 2 * n, n * sizeof(float));
  std::memcpy(s.vel_x.data(), vel.data(), n * sizeof(float));
  std::memcpy(s.vel_y.data(), vel.data() + n, n * sizeof(float));
  std::memcpy(s.vel_z.data(), vel.data() + 2 * n, n * sizeof(float));
}

}  // namespace hip_detail

// A device-resident context for one scene.
class HipSession {
 public:
  template <class P, class S>
  HipSession(const P& params, const S& state) {
    const fluid_hip_params q = hip_detail::to_c(params);
    const int64_t n = static_cast<int64_t>(state.pos_x.size());
    hip_detail::check(fluid_hip_create(&ctx_, 1, &n, &q, nullptr), "fluid_hip_create");
    upload(state);
  }
  ~HipSession() { fluid_hip_destroy(ctx_); }
  HipSession(const HipSession&) = delete;
  HipSession& operator=(const HipSession&) = delete;

  template <class P>
  void set_params(const P& params) {
    const fluid_hip_params q = hip_detail::to_c(params);
    hip_detail::check(fluid_hip_set_params(ctx_, &q), "fluid_hip_set_params");
  }
  template <class S>
  void upload(const S& state) {
    hip_detail::pack(state, pos_, vel_);
    hip_detail::check(fluid_hip_upload(ctx_, pos_.data(), vel_.data(), state.time),
                      "fluid_hip_upload");
  }
  void step(int nsteps = 1, bool graph = false) {
    hip_detail::check(fluid_hip_step(ctx_, nsteps, graph ? 1 : 0), "fluid_hip_step");
  }
  template <class S>
  void download(S& state) {
    const std::size_t n = state.pos_x.size();
    pos_.resize(3 * n);
    vel_.resize(3 * n);
    float t = 0.0f;
    hip_detail::check(fluid_hip_download(ctx_, pos_.data(), vel_.data(), &t), "fluid_hip_download");
    hip_detail::unpack(pos_, vel_, state);
    state.time = t;
  }

 private:
  fluid_hip_ctx* ctx_ = nullptr;
  std::vector<float> pos_, vel_;
};

// One substep with the call shape of cuda_step(const Params&, State&).
template <class P, class S>
void hip_step(const P& params, S& state) {
  if (state.pos_x.empty()) {  // core/src/core.cpp:122-125
    state.time += params.dt;
    return;
  }
  HipSession session(params, state);
  session.step(1);
  session.download(state);
}

}  // namespace fluid
