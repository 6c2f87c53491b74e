// TEST INFRASTRUCTURE ONLY — a C entry point onto the reference's own CPU
// step, compiled together with the reference sources where they lie
// (core/src/core.cpp, core/src/init.cpp; see oracle/Makefile).  It lets the
// tests and the CPU-baseline leg of the benchmark run the real reference
// `fluid::step` (core/src/core.cpp:119-615) on arrays they choose.  The
// built library goes to oracle/_ref/ and is never part of the product.

#include <cstring>
#include <string>

#include "fluid/core.h"
#include "fluid/init.h"

#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

// Same layout as OracleParams in pbf_oracle.cpp.
struct RefParams {
  double dt;
  double density;
  double particle_mass;
  double h;
  double epsilon;
  int solver_iterations;
  int use_uniform_grid;
  int enable_scorr;
  int enable_xsph;
  int enable_vorticity;
  int scorr_n;
  double scorr_k;
  double scorr_dq_coeff;
  double visc_c;
  double plane_restitution;
  double plane_friction;
  double vort_epsilon;
  double vort_norm_eps;
  double gravity[3];
  int n_planes;
  int mutation;  // ignored here
  const double* planes;
};

struct RefDiag {
  int* cell;
  int* sorted_ids;
  int* nbr_count;
  int* nbr_ids;
  long long nbr_cap;
  long long nbr_total;
  void* rho;
  void* lambda;
  void* pred0;  // not available from the reference (pred is overwritten)
};

}  // extern "C"

namespace {

fluid::Params to_params(const RefParams& p) {
  fluid::Params q;
  q.dt = static_cast<float>(p.dt);
  q.density = static_cast<float>(p.density);
  q.particle_mass = static_cast<float>(p.particle_mass);
  q.h = static_cast<float>(p.h);
  q.epsilon = static_cast<float>(p.epsilon);
  q.solver_iterations = p.solver_iterations;
  q.use_uniform_grid = p.use_uniform_grid != 0;
  q.enable_scorr = p.enable_scorr != 0;
  q.enable_xsph = p.enable_xsph != 0;
  q.enable_vorticity = p.enable_vorticity != 0;
  q.scorr_n = p.scorr_n;
  q.scorr_k = static_cast<float>(p.scorr_k);
  q.scorr_dq_coeff = static_cast<float>(p.scorr_dq_coeff);
  q.visc_c = static_cast<float>(p.visc_c);
  q.plane_restitution = static_cast<float>(p.plane_restitution);
  q.plane_friction = static_cast<float>(p.plane_friction);
  q.vort_epsilon = static_cast<float>(p.vort_epsilon);
  q.vort_norm_eps = static_cast<float>(p.vort_norm_eps);
  q.external_forces.x = static_cast<float>(p.gravity[0]);
  q.external_forces.y = static_cast<float>(p.gravity[1]);
  q.external_forces.z = static_cast<float>(p.gravity[2]);
  q.planes.clear();
  for (int i = 0; i < p.n_planes; ++i)
    q.planes.add(static_cast<float>(p.planes[4 * i + 0]),
                 static_cast<float>(p.planes[4 * i + 1]),
                 static_cast<float>(p.planes[4 * i + 2]),
                 static_cast<float>(p.planes[4 * i + 3]));
  return q;
}

}  // namespace

extern "C" {

int ref_step_f32(const RefParams* P, long long n, float* pos, float* vel,
                 int nsteps, float* time, RefDiag* diag) {
  if (!P || n < 0 || nsteps < 0) return -1;
  const fluid::Params params = to_params(*P);
  fluid::State st = fluid::make_state(static_cast<std::size_t>(n));
  std::memcpy(st.pos_x.data(), pos, sizeof(float) * n);
  std::memcpy(st.pos_y.data(), pos + n, sizeof(float) * n);
  std::memcpy(st.pos_z.data(), pos + 2 * n, sizeof(float) * n);
  std::memcpy(st.vel_x.data(), vel, sizeof(float) * n);
  std::memcpy(st.vel_y.data(), vel + n, sizeof(float) * n);
  std::memcpy(st.vel_z.data(), vel + 2 * n, sizeof(float) * n);
  if (time) st.time = *time;
  for (int s = 0; s < nsteps; ++s) fluid::step(params, st);
  std::memcpy(pos, st.pos_x.data(), sizeof(float) * n);
  std::memcpy(pos + n, st.pos_y.data(), sizeof(float) * n);
  std::memcpy(pos + 2 * n, st.pos_z.data(), sizeof(float) * n);
  std::memcpy(vel, st.vel_x.data(), sizeof(float) * n);
  std::memcpy(vel + n, st.vel_y.data(), sizeof(float) * n);
  std::memcpy(vel + 2 * n, st.vel_z.data(), sizeof(float) * n);
  if (time) *time = st.time;
  if (diag && n > 0 && nsteps > 0) {
    const auto& c = st.cpu;
    if (diag->cell && params.use_uniform_grid)
      for (long long k = 0; k < n; ++k) {
        const auto& e = c.grid_entries[k];
        diag->cell[3 * e.particle + 0] = e.key.x;
        diag->cell[3 * e.particle + 1] = e.key.y;
        diag->cell[3 * e.particle + 2] = e.key.z;
      }
    if (diag->sorted_ids && params.use_uniform_grid)
      for (long long k = 0; k < n; ++k) diag->sorted_ids[k] = c.grid_entries[k].particle;
    if (diag->nbr_count)
      for (long long i = 0; i < n; ++i)
        diag->nbr_count[i] =
            c.neighbor_prefix_sum[i] - (i == 0 ? 0 : c.neighbor_prefix_sum[i - 1]);
    diag->nbr_total = static_cast<long long>(c.neighbor_indices.size());
    if (diag->nbr_ids) {
      const long long m = diag->nbr_cap < diag->nbr_total ? diag->nbr_cap : diag->nbr_total;
      std::memcpy(diag->nbr_ids, c.neighbor_indices.data(), sizeof(int) * m);
    }
    if (diag->rho) std::memcpy(diag->rho, c.rho.data(), sizeof(float) * n);
    if (diag->lambda) std::memcpy(diag->lambda, c.lambda.data(), sizeof(float) * n);
  }
  return 0;
}

int ref_set_threads(int t) {
#ifdef _OPENMP
  if (t > 0) omp_set_num_threads(t);
  return omp_get_max_threads();
#else
  (void)t;
  return 1;
#endif
}

// Loads a scene JSON through the reference loader (core/src/init.cpp:158-418).
// Returns the particle count, or -1 with `err` filled.  With pos/vel null it
// only counts; planes receives up to `plane_cap` (nx, ny, nz, d) rows.
long long ref_load_scene(const char* path, float* pos, float* vel, long long cap,
                         float* scalars /* mass, density, h, eps, k, c, radius, gx, gy, gz, n */,
                         float* planes, int plane_cap, int* n_planes, char* err,
                         int err_len) {
  fluid::Params params;
  fluid::State st = fluid::make_state(0);
  std::string e;
  if (!fluid::init_scene_from_json(path, params, st, &e)) {
    if (err && err_len > 0) {
      std::strncpy(err, e.c_str(), err_len - 1);
      err[err_len - 1] = 0;
    }
    return -1;
  }
  const long long n = static_cast<long long>(st.size());
  if (pos && vel && cap >= n) {
    std::memcpy(pos, st.pos_x.data(), sizeof(float) * n);
    std::memcpy(pos + n, st.pos_y.data(), sizeof(float) * n);
    std::memcpy(pos + 2 * n, st.pos_z.data(), sizeof(float) * n);
    std::memcpy(vel, st.vel_x.data(), sizeof(float) * n);
    std::memcpy(vel + n, st.vel_y.data(), sizeof(float) * n);
    std::memcpy(vel + 2 * n, st.vel_z.data(), sizeof(float) * n);
  }
  if (scalars) {
    scalars[0] = params.particle_mass;
    scalars[1] = params.density;
    scalars[2] = params.h;
    scalars[3] = params.epsilon;
    scalars[4] = params.scorr_k;
    scalars[5] = params.visc_c;
    scalars[6] = params.particle_radius;
    scalars[7] = params.external_forces.x;
    scalars[8] = params.external_forces.y;
    scalars[9] = params.external_forces.z;
    scalars[10] = static_cast<float>(params.scorr_n);
  }
  const int np = static_cast<int>(params.planes.size());
  if (n_planes) *n_planes = np;
  if (planes)
    for (int i = 0; i < np && i < plane_cap; ++i) {
      planes[4 * i + 0] = params.planes.nx[i];
      planes[4 * i + 1] = params.planes.ny[i];
      planes[4 * i + 2] = params.planes.nz[i];
      planes[4 * i + 3] = params.planes.d[i];
    }
  return n;
}

}  // extern "C"
