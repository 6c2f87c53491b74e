// TEST INFRASTRUCTURE ONLY — the hip backend behind the reference's own C++
// solver types.  Compiled (oracle/Makefile, target `dropin`) together with
// the reference's core.cpp and init.cpp from where they lie; the scene is
// built with the reference's init_block/init_planes
// (core/src/init.cpp:72-117) and stepped once with the reference's
// fluid::step (core/src/core.cpp:119-615) and once with fluid::hip_step /
// fluid::HipSession (include/fluid_hip_backend.hpp).  Exit status 0 when
// every position, velocity and the time agree bitwise.  With arguments
// SCENE STEPS [FLAGS] the scene comes from a scene file through the
// reference's own init_scene_from_json instead.

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "fluid/core.h"
#include "fluid/init.h"
#include "fluid_hip_backend.hpp"

namespace {

bool same(const std::vector<float>& a, const std::vector<float>& b) {
  return a.size() == b.size() && std::memcmp(a.data(), b.data(), a.size() * sizeof(float)) == 0;
}

bool same_state(const fluid::State& a, const fluid::State& b) {
  return same(a.pos_x, b.pos_x) && same(a.pos_y, b.pos_y) && same(a.pos_z, b.pos_z) &&
         same(a.vel_x, b.vel_x) && same(a.vel_y, b.vel_y) && same(a.vel_z, b.vel_z) &&
         std::memcmp(&a.time, &b.time, sizeof(float)) == 0;
}

void scene(fluid::Params& p, fluid::State& s, bool flags) {
  p.dt = static_cast<float>(1.0 / 120);
  p.density = 1000.0f;
  p.particle_mass = 0.125f;
  p.h = 0.1f;
  p.epsilon = 100.0f;
  p.plane_restitution = 0.05f;
  p.plane_friction = 0.1f;
  if (flags) {
    p.enable_scorr = true;
    p.scorr_k = 0.1f;
    p.scorr_dq_coeff = 0.2f;
    p.enable_xsph = true;
    p.visc_c = 0.01f;
    p.enable_vorticity = true;
    p.vort_epsilon = 5e-4f;
  }
  fluid::init_planes(p, 1.0f, 1.0f, 1.0f);
  fluid::init_block(p, s, 14, 12, 10, 0.025f, 0.025f, 0.025f, 0.05f);
  // a deterministic velocity field so the step is not a plain lattice fall
  for (std::size_t i = 0; i < s.size(); ++i) {
    s.vel_x[i] = 0.5f * std::sin(0.37f * static_cast<float>(i));
    s.vel_y[i] = 0.5f * std::cos(0.23f * static_cast<float>(i));
    s.vel_z[i] = 0.25f * std::sin(0.11f * static_cast<float>(i));
  }
}

}  // namespace

// A scene file through the reference's own loader (init_scene_from_json,
// core/src/init.cpp:158-418) with the app's settings for it, then `steps`
// substeps with fluid::step and with fluid::HipSession.
int run_scene_file(const char* path, int steps, bool flags) {
  fluid::Params p;
  fluid::State ref = fluid::make_state(0);
  std::string error;
  if (!fluid::init_scene_from_json(path, p, ref, &error)) {
    std::printf("scene %s failed to load: %s\n", path, error.c_str());
    return 1;
  }
  p.dt = static_cast<float>(1.0 / 120);  // --steps-per-sec 120
  p.plane_restitution = 0.05f;           // --plane-restitution 0.05
  p.plane_friction = 0.1f;               // --plane-friction 0.1
  if (flags) {                           // --enable-scorr --enable-xsph --enable-vorticity
    p.enable_scorr = true;
    p.enable_xsph = true;
    p.enable_vorticity = true;
  }
  fluid::State one = ref;
  fluid::HipSession session(p, one);
  for (int k = 0; k < steps; ++k) fluid::step(p, ref);
  session.step(steps);
  session.download(one);
  const bool ok = same_state(ref, one);
  std::printf("scene=%s flags=%d particles=%zu session(%d steps) bitwise=%d\n", path, flags ? 1 : 0,
              ref.size(), steps, ok ? 1 : 0);
  return ok ? 0 : 1;
}

int main(int argc, char** argv) {
  int ndev = 0;
  const char* err = nullptr;
  if (!fluid::hip_device_available(&ndev, &err)) {
    std::printf("no HIP device: %s\n", err ? err : "none");
    return 2;
  }
  std::printf("hip_version=%d devices=%d\n", fluid::hip_version(), ndev);
  if (argc >= 3)  // dropin_test SCENE.json STEPS [flags]
    return run_scene_file(argv[1], std::atoi(argv[2]), argc >= 4 && std::atoi(argv[3]) != 0);
  int failures = 0;
  for (int flags = 0; flags < 2; ++flags) {
    fluid::Params p;
    fluid::State ref = fluid::make_state(0);
    scene(p, ref, flags != 0);
    fluid::State one = ref;

    // one substep through the cuda_step call shape
    fluid::step(p, ref);
    fluid::hip_step(p, one);
    const bool ok1 = same_state(ref, one);

    // device-resident: 8 more substeps in one session
    fluid::HipSession session(p, one);
    for (int k = 0; k < 8; ++k) fluid::step(p, ref);
    session.step(8);
    session.download(one);
    const bool ok8 = same_state(ref, one);
    std::printf("flags=%d particles=%zu hip_step bitwise=%d session(8 steps) bitwise=%d\n", flags,
                ref.size(), ok1 ? 1 : 0, ok8 ? 1 : 0);
    failures += (ok1 ? 0 : 1) + (ok8 ? 0 : 1);
  }
  // the empty state only advances time (core/src/core.cpp:122-125)
  fluid::Params p;
  p.h = 0.1f;
  fluid::State empty = fluid::make_state(0);
  fluid::hip_step(p, empty);
  const bool ok0 = empty.time == p.dt;
  std::printf("empty state time advanced=%d\n", ok0 ? 1 : 0);
  failures += ok0 ? 0 : 1;
  return failures == 0 ? 0 : 1;
}
