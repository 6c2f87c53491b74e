// fluid_fit — an inverse-problem driver over the hip backend's adjoint:
// fits per-scene gravity and/or per-particle initial velocities so that the
// final positions of a rollout match a target cloud.  No counterpart in the
// reference, whose solver has no derivatives; it is the use the adjoint
// (fluid_hip_forward_record / fluid_hip_backward) is built for.
//
// Every scene is a block of nx*ny*nz jittered lattice sites in the box
// [0,4]x[0,3]x[0,2] with seeded N(0, sigma^2) initial velocities, s_corr and
// XSPH on and vorticity off, with the reference's default relaxation and
// s_corr / XSPH coefficients (eps 600, k 5e-5, dq 0.3h, c 2e-4; the
// gradient of these rollouts stays bounded).  All scenes run as one batched
// context, so one forward and one reverse pass per iteration serve them all.
//
// The target of scene s is a forward run with gravity g*_s (its x and z
// components and the deviation of y from -9.8 are drawn from a seeded
// stream) and, with --fit velocity, a second seeded velocity field.  The
// loss of scene s is L_s = mean_i |x_T,i - target_i|^2; the scenes'
// losses are independent, so the gradient of their sum gives each scene's
// own.  Parameters are updated by Adam (Kingma & Ba, 2015).
//
//   fluid_fit [--scenes S] [--dims nx,ny,nz] [--steps T] [--iters K]
//             [--fit gravity|velocity|both] [--lr-gravity A] [--lr-velocity B]
//             [--checkpoint-every C] [--seed N] [--sigma V]
//
// Prints one JSON line per iteration and a final summary line.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "fluid_hip.h"
#include "scene.hpp"

namespace {

struct Adam {
  std::vector<double> m, v;
  double lr, b1 = 0.9, b2 = 0.999, eps = 1e-8;
  int t = 0;
  Adam(size_t n, double lr_) : m(n, 0.0), v(n, 0.0), lr(lr_) {}
  template <class T>
  void step(std::vector<T>& x, const std::vector<double>& g) {
    ++t;
    const double c1 = 1.0 - std::pow(b1, t), c2 = 1.0 - std::pow(b2, t);
    for (size_t i = 0; i < x.size(); ++i) {
      m[i] = b1 * m[i] + (1.0 - b1) * g[i];
      v[i] = b2 * v[i] + (1.0 - b2) * g[i] * g[i];
      x[i] = static_cast<T>(static_cast<double>(x[i]) - lr * (m[i] / c1) / (std::sqrt(v[i] / c2) + eps));
    }
  }
};

bool check(int rc, const char* what) {
  if (rc != FLUID_HIP_OK) {
    std::fprintf(stderr, "%s failed: %s\n", what, fluid_hip_last_error());
    return false;
  }
  return true;
}

}  // namespace

int main(int argc, char** argv) {
  int scenes = 4, steps = 20, iters = 40, every = 10, seed = 7;
  size_t dims[3] = {16, 12, 10};
  double lr_g = 0.05, lr_v = 0.02, sigma = 0.5;
  std::string fit = "gravity";
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&]() -> const char* {
      if (i + 1 >= argc) {
        std::fprintf(stderr, "missing value for %s\n", a.c_str());
        std::exit(2);
      }
      return argv[++i];
    };
    if (a == "--scenes") scenes = std::atoi(next());
    else if (a == "--steps") steps = std::atoi(next());
    else if (a == "--iters") iters = std::atoi(next());
    else if (a == "--checkpoint-every") every = std::atoi(next());
    else if (a == "--seed") seed = std::atoi(next());
    else if (a == "--lr-gravity") lr_g = std::atof(next());
    else if (a == "--lr-velocity") lr_v = std::atof(next());
    else if (a == "--sigma") sigma = std::atof(next());
    else if (a == "--fit") fit = next();
    else if (a == "--dims") {
      if (std::sscanf(next(), "%zu,%zu,%zu", &dims[0], &dims[1], &dims[2]) != 3) {
        std::fprintf(stderr, "--dims needs nx,ny,nz\n");
        return 2;
      }
    } else if (a == "-h" || a == "--help") {
      std::printf("usage: %s [--scenes S] [--dims nx,ny,nz] [--steps T] [--iters K] "
                  "[--fit gravity|velocity|both] [--lr-gravity A] [--lr-velocity B] "
                  "[--checkpoint-every C] [--seed N] [--sigma V]\n", argv[0]);
      return 0;
    } else {
      std::fprintf(stderr, "unknown option %s\n", a.c_str());
      return 2;
    }
  }
  const bool fit_g = fit == "gravity" || fit == "both";
  const bool fit_v = fit == "velocity" || fit == "both";
  if ((!fit_g && !fit_v) || scenes < 1 || steps < 1 || iters < 0 || every < 1) {
    std::fprintf(stderr, "bad arguments\n");
    return 2;
  }

  // scene parameters: the differentiable-rollout workload's
  fluid_hip_params base = fluidapp::default_params();
  base.dt = static_cast<float>(1.0 / 120);
  base.density = 1000.0f;
  base.particle_mass = 0.125f;
  base.h = 0.1f;
  base.particle_radius = 0.025f;
  base.epsilon = 600.0f;
  base.enable_scorr = 1;
  base.scorr_k = 5e-5f;
  base.scorr_dq_coeff = 0.3f;
  base.enable_xsph = 1;
  base.visc_c = 2e-4f;
  base.plane_restitution = 0.05f;
  base.plane_friction = 0.1f;
  fluidapp::init_planes(base, 4.0f, 3.0f, 2.0f);

  const size_t per = dims[0] * dims[1] * dims[2];
  const size_t n = per * static_cast<size_t>(scenes);
  std::vector<int64_t> counts(static_cast<size_t>(scenes), static_cast<int64_t>(per));
  std::vector<float> pos(3 * n), vel0(3 * n), vel_t(3 * n);
  const std::vector<double> gu = fluidapp::uniform01(seed, 3 * static_cast<size_t>(scenes), 91);
  std::vector<float> g_true(3 * static_cast<size_t>(scenes));
  for (int s = 0; s < scenes; ++s) {
    std::vector<float> px, py, pz;
    const double origin[3] = {0.0, 0.0, 0.0};
    fluidapp::lattice_block(px, py, pz, dims[0], dims[1], dims[2], 0.05, origin, 0.1, seed, s);
    const std::vector<double> va = fluidapp::normal(seed + 1000 + s, 3 * per, 7, sigma);
    const std::vector<double> vb = fluidapp::normal(seed + 2000 + s, 3 * per, 7, sigma);
    for (size_t i = 0; i < per; ++i) {
      const size_t j = static_cast<size_t>(s) * per + i;
      pos[j] = px[i];
      pos[n + j] = py[i];
      pos[2 * n + j] = pz[i];
      for (int k = 0; k < 3; ++k) {
        vel0[k * n + j] = static_cast<float>(va[k * per + i]);
        vel_t[k * n + j] = fit_v ? static_cast<float>(vb[k * per + i]) : vel0[k * n + j];
      }
    }
    const double* u = &gu[3 * static_cast<size_t>(s)];
    g_true[3 * s + 0] = fit_g ? static_cast<float>(2.0 * u[0] - 1.0) : 0.0f;
    g_true[3 * s + 1] = fit_g ? static_cast<float>(-9.8 + 2.0 * (2.0 * u[1] - 1.0)) : -9.8f;
    g_true[3 * s + 2] = fit_g ? static_cast<float>(2.0 * u[2] - 1.0) : 0.0f;
  }
  std::vector<fluid_hip_params> prm(static_cast<size_t>(scenes), base);
  auto set_gravity = [&](const std::vector<float>& g) {
    for (int s = 0; s < scenes; ++s)
      for (int k = 0; k < 3; ++k) prm[s].gravity[k] = g[3 * s + k];
  };

  fluid_hip_ctx* ctx = nullptr;
  if (!check(fluid_hip_create(&ctx, scenes, counts.data(), prm.data(), nullptr), "fluid_hip_create"))
    return 1;
  bool ok = true;
  // the target: a forward run with the true parameters
  std::vector<float> target(3 * n), out(3 * n);
  set_gravity(g_true);
  ok = ok && check(fluid_hip_set_params(ctx, prm.data()), "fluid_hip_set_params");
  ok = ok && check(fluid_hip_upload(ctx, pos.data(), vel_t.data(), 0.0f), "fluid_hip_upload");
  ok = ok && check(fluid_hip_step(ctx, steps, 0), "fluid_hip_step");
  ok = ok && check(fluid_hip_download(ctx, target.data(), nullptr, nullptr), "fluid_hip_download");

  std::vector<float> g(3 * static_cast<size_t>(scenes));
  for (int s = 0; s < scenes; ++s) {
    g[3 * s + 0] = 0.0f;
    g[3 * s + 1] = -9.8f;
    g[3 * s + 2] = 0.0f;
  }
  std::vector<float> v = vel0;
  Adam opt_g(g.size(), lr_g), opt_v(v.size(), lr_v);
  std::vector<float> seed_grad(3 * n), dv0(3 * n), dg(3 * static_cast<size_t>(scenes));
  double first_loss = -1.0, loss = 0.0, g_err = 0.0;
  for (int it = 0; ok && it <= iters; ++it) {
    set_gravity(g);
    ok = ok && check(fluid_hip_set_params(ctx, prm.data()), "fluid_hip_set_params");
    ok = ok && check(fluid_hip_upload(ctx, pos.data(), v.data(), 0.0f), "fluid_hip_upload");
    ok = ok && check(fluid_hip_forward_record(ctx, steps, every), "fluid_hip_forward_record");
    ok = ok && check(fluid_hip_download(ctx, out.data(), nullptr, nullptr), "fluid_hip_download");
    if (!ok) break;
    loss = 0.0;
    for (int s = 0; s < scenes; ++s) {
      double ls = 0.0;
      for (size_t i = 0; i < per; ++i) {
        const size_t j = static_cast<size_t>(s) * per + i;
        for (int k = 0; k < 3; ++k) {
          const double d = static_cast<double>(out[k * n + j]) - static_cast<double>(target[k * n + j]);
          ls += d * d;
          seed_grad[k * n + j] = static_cast<float>(2.0 * d / static_cast<double>(per));
        }
      }
      loss += ls / static_cast<double>(per);
    }
    loss /= scenes;
    if (first_loss < 0.0) first_loss = loss;
    g_err = 0.0;
    for (size_t k = 0; k < g.size(); ++k)
      g_err = std::max(g_err, std::fabs(static_cast<double>(g[k]) - static_cast<double>(g_true[k])));
    std::printf("{\"iter\": %d, \"loss\": %.9g, \"gravity_max_abs_error\": %.6g}\n", it, loss, g_err);
    std::fflush(stdout);
    if (it == iters) break;
    ok = ok && check(fluid_hip_backward(ctx, seed_grad.data(), nullptr, nullptr, fit_v ? dv0.data() : nullptr,
                                        fit_g ? dg.data() : nullptr),
                     "fluid_hip_backward");
    if (!ok) break;
    if (fit_g) opt_g.step(g, std::vector<double>(dg.begin(), dg.end()));
    if (fit_v) opt_v.step(v, std::vector<double>(dv0.begin(), dv0.end()));
  }
  fluid_hip_destroy(ctx);
  if (!ok) return 1;
  std::printf("{\"summary\": true, \"scenes\": %d, \"particles\": %zu, \"steps\": %d, \"iters\": %d, "
              "\"fit\": \"%s\", \"first_loss\": %.9g, \"final_loss\": %.9g, "
              "\"loss_ratio\": %.6g, \"gravity_max_abs_error\": %.6g}\n",
              scenes, n, steps, iters, fit.c_str(), first_loss, loss,
              first_loss > 0.0 ? loss / first_loss : 0.0, g_err);
  return 0;
}
