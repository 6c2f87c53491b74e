// fluid_framediff — the parity harness's diff tool: compares state frames
// (.fsf, frames.hpp) of two runs, e.g. the reference CPU step and the hip
// backend on the same scene, frame by frame.
//
//   fluid_framediff A B [--pos-tol T] [--vel-tol T] [--rho-rel-tol R] [--json]
//
// A and B are two frame files or two directories (frames matched by file
// name).  Per frame it reports: particle count, step and time agreement,
// bitwise-equal positions/velocities, max |a - b| and RMSE of positions and
// velocities, and, when both frames carry densities, the density-error
// statistics mean |rho/rho0 - 1| and max rho of each side with their
// relative difference (the statistic that stays comparable over long
// rollouts, where chaotic divergence makes per-particle comparison
// meaningless).  Exit status 0 when every frame is within the tolerances
// (defaults: positions and velocities bitwise, density statistics 1%), 1
// otherwise, 2 on usage or read errors.
#include <dirent.h>
#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "frames.hpp"

namespace {

struct Stats {
  double max_abs = 0.0, sq = 0.0;
  size_t unequal = 0, count = 0;
  void add(float a, float b) {
    const double d = std::fabs(static_cast<double>(a) - static_cast<double>(b));
    if (std::memcmp(&a, &b, sizeof(float)) != 0) ++unequal;
    if (std::isnan(d)) {
      max_abs = INFINITY;
    } else {
      max_abs = std::max(max_abs, d);
      sq += d * d;
    }
    ++count;
  }
  double rmse() const { return count ? std::sqrt(sq / static_cast<double>(count)) : 0.0; }
};

struct DensityStats {
  double mean_err = 0.0, max_rho = 0.0;
};

DensityStats density_stats(const fluidapp::Frame& f) {
  DensityStats s;
  const double rho0 = f.rho0 > 0.0f ? static_cast<double>(f.rho0) : 1.0;
  double acc = 0.0;
  for (float r : f.rho) {
    acc += std::fabs(static_cast<double>(r) / rho0 - 1.0);
    s.max_rho = std::max(s.max_rho, static_cast<double>(r));
  }
  s.mean_err = f.rho.empty() ? 0.0 : acc / static_cast<double>(f.rho.size());
  return s;
}

bool is_dir(const std::string& p) {
  struct stat st;
  return ::stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

std::vector<std::string> frame_files(const std::string& dir) {
  std::vector<std::string> out;
  if (DIR* d = ::opendir(dir.c_str())) {
    while (dirent* e = ::readdir(d)) {
      const std::string name = e->d_name;
      if (name.size() > 4 && name.compare(name.size() - 4, 4, ".fsf") == 0) out.push_back(name);
    }
    ::closedir(d);
  }
  std::sort(out.begin(), out.end());
  return out;
}

double rel_diff(double a, double b) {
  const double m = std::max(std::fabs(a), std::fabs(b));
  return m > 0.0 ? std::fabs(a - b) / m : 0.0;
}

}  // namespace

int main(int argc, char** argv) {
  std::vector<std::string> paths;
  double pos_tol = 0.0, vel_tol = 0.0, rho_tol = 0.01;
  bool as_json = false, bitwise_pos = true, bitwise_vel = true;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto value = [&](double& out, bool* exact) {
      if (i + 1 >= argc) return false;
      out = std::atof(argv[++i]);
      if (exact) *exact = false;
      return true;
    };
    if (a == "--pos-tol") {
      if (!value(pos_tol, &bitwise_pos)) return 2;
    } else if (a == "--vel-tol") {
      if (!value(vel_tol, &bitwise_vel)) return 2;
    } else if (a == "--rho-rel-tol") {
      if (!value(rho_tol, nullptr)) return 2;
    } else if (a == "--json") {
      as_json = true;
    } else if (a == "-h" || a == "--help") {
      std::printf("usage: %s A B [--pos-tol T] [--vel-tol T] [--rho-rel-tol R] [--json]\n",
                  argv[0]);
      return 0;
    } else {
      paths.push_back(a);
    }
  }
  if (paths.size() != 2) {
    std::fprintf(stderr, "usage: %s A B [--pos-tol T] [--vel-tol T] [--rho-rel-tol R] [--json]\n",
                 argv[0]);
    return 2;
  }
  std::vector<std::pair<std::string, std::string>> pairs;
  if (is_dir(paths[0]) && is_dir(paths[1])) {
    const std::vector<std::string> fa = frame_files(paths[0]), fb = frame_files(paths[1]);
    for (const std::string& f : fa)
      if (std::find(fb.begin(), fb.end(), f) != fb.end())
        pairs.emplace_back(fluidapp::join_path(paths[0], f), fluidapp::join_path(paths[1], f));
    if (pairs.empty() || fa.size() != fb.size()) {
      std::fprintf(stderr, "frame sets differ: %zu frames in %s, %zu in %s, %zu in common\n",
                   fa.size(), paths[0].c_str(), fb.size(), paths[1].c_str(), pairs.size());
      return pairs.empty() ? 2 : 1;
    }
  } else {
    pairs.emplace_back(paths[0], paths[1]);
  }

  bool all_ok = true;
  if (as_json) std::printf("[");
  for (size_t k = 0; k < pairs.size(); ++k) {
    fluidapp::Frame a, b;
    std::string err;
    if (!fluidapp::read_frame(pairs[k].first, a, &err) ||
        !fluidapp::read_frame(pairs[k].second, b, &err)) {
      std::fprintf(stderr, "%s\n", err.c_str());
      return 2;
    }
    const bool same_n = a.n == b.n;
    Stats sp, sv;
    if (same_n) {
      for (size_t i = 0; i < a.pos.size(); ++i) sp.add(a.pos[i], b.pos[i]);
      for (size_t i = 0; i < a.vel.size(); ++i) sv.add(a.vel[i], b.vel[i]);
    }
    const bool have_rho = !a.rho.empty() && !b.rho.empty();
    const DensityStats da = density_stats(a), db = density_stats(b);
    const double derr = have_rho ? rel_diff(da.mean_err, db.mean_err) : 0.0;
    const double dmax = have_rho ? rel_diff(da.max_rho, db.max_rho) : 0.0;
    const bool pos_ok = bitwise_pos ? sp.unequal == 0 : sp.max_abs <= pos_tol;
    const bool vel_ok = bitwise_vel ? sv.unequal == 0 : sv.max_abs <= vel_tol;
    const bool rho_ok = !have_rho || (derr <= rho_tol && dmax <= rho_tol);
    const bool ok = same_n && a.step == b.step && pos_ok && vel_ok && rho_ok;
    all_ok = all_ok && ok;
    const char* name = std::strrchr(pairs[k].first.c_str(), '/');
    name = name ? name + 1 : pairs[k].first.c_str();
    if (as_json) {
      std::printf(
          "%s{\"frame\": \"%s\", \"n\": %llu, \"step\": %lld, \"same_step\": %s, "
          "\"time_a\": %.9g, \"time_b\": %.9g, \"pos_unequal\": %zu, \"pos_max_abs\": %.9g, "
          "\"pos_rmse\": %.9g, \"vel_unequal\": %zu, \"vel_max_abs\": %.9g, \"vel_rmse\": %.9g",
          k ? ", " : "", name, static_cast<unsigned long long>(a.n),
          static_cast<long long>(a.step), a.step == b.step ? "true" : "false",
          static_cast<double>(a.time), static_cast<double>(b.time), sp.unequal, sp.max_abs,
          sp.rmse(), sv.unequal, sv.max_abs, sv.rmse());
      if (have_rho)
        std::printf(", \"density_error_a\": %.9g, \"density_error_b\": %.9g, "
                    "\"density_error_rel_diff\": %.9g, \"max_rho_a\": %.9g, \"max_rho_b\": %.9g",
                    da.mean_err, db.mean_err, derr, da.max_rho, db.max_rho);
      std::printf(", \"ok\": %s}", ok ? "true" : "false");
    } else {
      std::printf("%s n=%llu step=%lld%s pos: %zu unequal, max_abs %.3g, rmse %.3g | vel: %zu "
                  "unequal, max_abs %.3g, rmse %.3g",
                  name, static_cast<unsigned long long>(a.n), static_cast<long long>(a.step),
                  a.step == b.step ? "" : " (steps differ)", sp.unequal, sp.max_abs, sp.rmse(),
                  sv.unequal, sv.max_abs, sv.rmse());
      if (have_rho)
        std::printf(" | density error %.6g vs %.6g (rel %.3g), max rho %.6g vs %.6g",
                    da.mean_err, db.mean_err, derr, da.max_rho, db.max_rho);
      if (!same_n) std::printf(" | particle counts differ (%llu vs %llu)",
                               static_cast<unsigned long long>(a.n),
                               static_cast<unsigned long long>(b.n));
      std::printf(" %s\n", ok ? "OK" : "FAIL");
    }
  }
  if (as_json) std::printf("]\n");
  return all_ok ? 0 : 1;
}
