// TEST INFRASTRUCTURE ONLY — the CPU oracle for the PBF substep.
//
// This file is a restatement, written for this project, of the reference
// CPU step `fluid::step` (core/src/core.cpp:119-615) and of the kernel
// helpers it calls (core/src/core.cpp:28-71).  It follows the reference
// operation for operation — same formulas, same evaluation order, same loop
// orders, the same serial grid/neighbor build — so that, compiled with
// -ffp-contract=off and no -march, its float instantiation reproduces the
// reference bit for bit.  The double instantiation is the FP64 path that
// finite-difference gradient checks run on.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load this library, and only as the checker.  The product path
// (fluidsimulator_amd) never links or calls it.
//
// Pinning (what checks this restatement, see tests/test_oracle.py):
//   * golden vectors written by the reference's own core.cpp compiled in the
//     build container (tests/golden/*.npz): bit-exact for every flag combo;
//   * oracle/_ref (the reference core.cpp built from its own sources) on
//     fresh seeded scenes, when that library is present: bit-exact;
//   * closed forms: poly6/spiky coefficients, the rest-lattice density
//     (about 1009.8 for h=0.1, d=0.05, m=0.125), an isolated particle's
//     lambda, free fall, single-particle plane projection and restitution;
//   * brute force: the O(N^2) neighbor path (core/src/core.cpp:248-268)
//     against the uniform-grid path on tiny scenes.
// The `mutation` knob plants known mistakes so tests can show that each of
// those checks catches them.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// Planted mistakes (tests only).  0 = faithful.
enum Mutation : int {
  kMutNone = 0,
  kMutDropSelfDensity = 1,   // omit the W(0) self term of rho
  kMutSpikySign = 2,         // flip the spiky gradient sign
  kMutSwapDelta = 3,         // use pj - pi instead of pi - pj in the delta pass
  kMutJacobiInPlace = 4,     // Gauss-Seidel update instead of Jacobi
  kMutNeighborLe = 5,        // r2 <= h2 instead of r2 < h2 for the list
};

struct CellKey {
  int x, y, z;
};

inline bool cell_less(const CellKey& a, const CellKey& b) {
  // lexicographic (x, y, z) order — core/src/core.cpp:12-21
  if (a.x != b.x) return a.x < b.x;
  if (a.y != b.y) return a.y < b.y;
  return a.z < b.z;
}

inline bool cell_eq(const CellKey& a, const CellKey& b) {
  return a.x == b.x && a.y == b.y && a.z == b.z;
}

template <typename T>
CellKey cell_of(T x, T y, T z, T cell) {
  // floor(x * (1/cell)) — core/src/core.cpp:28-34 (not x / cell)
  const T inv = T(1) / cell;
  return CellKey{static_cast<int>(std::floor(x * inv)),
                 static_cast<int>(std::floor(y * inv)),
                 static_cast<int>(std::floor(z * inv))};
}

template <typename T>
constexpr T pi_v() {
  return static_cast<T>(3.14159265358979323846);
}

template <typename T>
T poly6(T r2, T h) {
  // core/src/core.cpp:35-46; coefficient recomputed per call as there.
  const T h2 = h * h;
  if (r2 > h2) return T(0);
  const T diff = h2 - r2;
  const T diff3 = diff * diff * diff;
  const T h4 = h2 * h2;
  const T h9 = h4 * h4 * h;
  const T coeff = T(315) / (T(64) * pi_v<T>() * h9);
  return coeff * diff3;
}

template <typename T>
T spiky_factor(T r, T h) {
  // core/src/core.cpp:48-57: the factor multiplies (pi - pj) directly,
  // it is not divided by r.
  if (r > h) return T(0);
  const T h2 = h * h;
  const T h6 = h2 * h2 * h2;
  const T diff = h - r;
  const T coeff = T(-45) / (pi_v<T>() * h6);
  return coeff * diff * diff;
}

// 0: std::pow as the reference calls it; 1: the correctly rounded power (the
// double-precision power rounded to T).  Mode 1 is not the reference: it
// runs a sibling trajectory whose only difference is a pow within 1 ulp, the
// yardstick for how far such a difference carries in a given scene.
int g_pow_mode = 0;

template <typename T>
T pow_n(T ratio, int n) {
  // core/src/core.cpp:59-71
  if (n == 2) return ratio * ratio;
  if (n == 3) return ratio * ratio * ratio;
  if (n == 4) {
    const T r2 = ratio * ratio;
    return r2 * r2;
  }
  if (g_pow_mode == 1)
    return static_cast<T>(std::pow(static_cast<double>(ratio), static_cast<double>(n)));
  if (g_pow_mode == 2) {  // a product chain: within about n/2 ulp
    T r = ratio;
    for (int i = 1; i < n; ++i) r *= ratio;
    return r;
  }
  return std::pow(ratio, static_cast<T>(n));
}

}  // namespace

extern "C" {

// Flattened Params (core/include/fluid/core.h:11-79).  All reals are passed
// as double; the float instantiation narrows them, which is exact for values
// that were float to begin with.
struct OracleParams {
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
  int mutation;
  const double* planes;  // n_planes x (nx, ny, nz, d)
  // Extension beyond the reference (whose restitution and friction hold for
  // every plane): n_planes x (restitution, friction), or null for the two
  // globals above on every plane.
  const double* plane_props;
};

// Optional per-step outputs of the last step, in original particle order.
struct OracleDiag {
  int* cell;             // [3*n] cell (x,y,z) of pred0
  int* sorted_ids;       // [n] grid entries after the sort
  int* nbr_count;        // [n]
  int* nbr_ids;          // [nbr_cap] concatenated neighbor lists
  long long nbr_cap;
  long long nbr_total;   // out
  void* rho;             // [n] of the real type
  void* lambda;          // [n]
  void* pred0;           // [3*n] predicted positions before the solve
};

}  // extern "C"

namespace {

template <typename T>
struct Scratch {
  std::vector<T> pred_x, pred_y, pred_z, del_x, del_y, del_z, lambda, rho;
  std::vector<T> dv_x, dv_y, dv_z, om_x, om_y, om_z, om_mag, eta_x, eta_y, eta_z;
  std::vector<int> nbr, prefix;
  struct Entry {
    CellKey key;
    int particle;
  };
  std::vector<Entry> entries;
  std::vector<CellKey> gkeys;
  std::vector<int> gstart, gend;

  void resize(size_t n) {
    for (auto* v : {&pred_x, &pred_y, &pred_z, &del_x, &del_y, &del_z, &lambda,
                    &rho, &dv_x, &dv_y, &dv_z, &om_x, &om_y, &om_z, &om_mag,
                    &eta_x, &eta_y, &eta_z})
      v->assign(n, T(0));
    prefix.assign(n, 0);
    entries.resize(n);
  }
};

template <typename T>
void oracle_step_impl(const OracleParams& P, size_t n, T* px_, T* py_, T* pz_,
                      T* vx_, T* vy_, T* vz_, Scratch<T>& S, T* pred0_out) {
  // core/src/core.cpp:119-615.  State arrays are the caller's SoA.
  const T dt = static_cast<T>(P.dt);
  const T h = static_cast<T>(P.h);
  const T h2 = h * h;
  const T min_r = T(0.01f) * h;  // 0.01f literal as in the reference
  const T min_r2 = min_r * min_r;
  const bool scorr_on = P.enable_scorr && static_cast<T>(P.scorr_k) != T(0);
  const T dq_coeff = (static_cast<T>(P.scorr_dq_coeff) > T(0))
                         ? static_cast<T>(P.scorr_dq_coeff)
                         : T(0.3f);  // core.cpp:144-145
  const T dq = dq_coeff * h;
  const T wdq = scorr_on ? poly6<T>(dq * dq, h) : T(0);
  const T inv_wdq = (wdq > T(1e-12f)) ? (T(1) / wdq) : T(0);
  const T gx = static_cast<T>(P.gravity[0]);
  const T gy = static_cast<T>(P.gravity[1]);
  const T gz = static_cast<T>(P.gravity[2]);
  const int mut = P.mutation;

  std::vector<T> pl_nx(P.n_planes), pl_ny(P.n_planes), pl_nz(P.n_planes), pl_d(P.n_planes);
  for (int p = 0; p < P.n_planes; ++p) {
    pl_nx[p] = static_cast<T>(P.planes[4 * p + 0]);
    pl_ny[p] = static_cast<T>(P.planes[4 * p + 1]);
    pl_nz[p] = static_cast<T>(P.planes[4 * p + 2]);
    pl_d[p] = static_cast<T>(P.planes[4 * p + 3]);
  }

  S.resize(n);
  S.nbr.clear();
  const long N = static_cast<long>(n);

  // predict — core.cpp:150-161
#pragma omp parallel for
  for (long i = 0; i < N; ++i) {
    vx_[i] += gx * dt;
    vy_[i] += gy * dt;
    vz_[i] += gz * dt;
    S.pred_x[i] = px_[i] + vx_[i] * dt;
    S.pred_y[i] = py_[i] + vy_[i] * dt;
    S.pred_z[i] = pz_[i] + vz_[i] * dt;
  }
  if (pred0_out) {
    std::copy(S.pred_x.begin(), S.pred_x.end(), pred0_out);
    std::copy(S.pred_y.begin(), S.pred_y.end(), pred0_out + n);
    std::copy(S.pred_z.begin(), S.pred_z.end(), pred0_out + 2 * n);
  }

  // neighbor list from pred0 — core.cpp:163-268 (serial, as there)
  if (P.use_uniform_grid && h > T(0)) {
    for (size_t i = 0; i < n; ++i)
      S.entries[i] = {cell_of<T>(S.pred_x[i], S.pred_y[i], S.pred_z[i], h),
                      static_cast<int>(i)};
    std::sort(S.entries.begin(), S.entries.end(),
              [](const typename Scratch<T>::Entry& a,
                 const typename Scratch<T>::Entry& b) {
                if (cell_less(a.key, b.key)) return true;
                if (cell_less(b.key, a.key)) return false;
                return a.particle < b.particle;
              });
    S.gkeys.clear();
    S.gstart.clear();
    S.gend.clear();
    CellKey cur = S.entries.front().key;
    int start = 0;
    for (size_t i = 1; i < n; ++i) {
      if (!cell_eq(cur, S.entries[i].key)) {
        S.gkeys.push_back(cur);
        S.gstart.push_back(start);
        S.gend.push_back(static_cast<int>(i));
        cur = S.entries[i].key;
        start = static_cast<int>(i);
      }
    }
    S.gkeys.push_back(cur);
    S.gstart.push_back(start);
    S.gend.push_back(static_cast<int>(n));

    for (size_t i = 0; i < n; ++i) {
      const T px = S.pred_x[i], py = S.pred_y[i], pz = S.pred_z[i];
      const CellKey base = cell_of<T>(px, py, pz, h);
      for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dx = -1; dx <= 1; ++dx) {
            const CellKey key{base.x + dx, base.y + dy, base.z + dz};
            auto it = std::lower_bound(S.gkeys.begin(), S.gkeys.end(), key, cell_less);
            if (it == S.gkeys.end() || !cell_eq(*it, key)) continue;
            const size_t c = static_cast<size_t>(it - S.gkeys.begin());
            for (int idx = S.gstart[c]; idx < S.gend[c]; ++idx) {
              const int j = S.entries[idx].particle;
              if (j == static_cast<int>(i)) continue;
              const T dxp = px - S.pred_x[j];
              const T dyp = py - S.pred_y[j];
              const T dzp = pz - S.pred_z[j];
              const T r2 = dxp * dxp + dyp * dyp + dzp * dzp;
              const bool in = (mut == kMutNeighborLe) ? (r2 <= h2) : (r2 < h2);
              if (in) S.nbr.push_back(j);
            }
          }
      S.prefix[i] = static_cast<int>(S.nbr.size());
    }
  } else {
    for (size_t i = 0; i < n; ++i) {
      const T px = S.pred_x[i], py = S.pred_y[i], pz = S.pred_z[i];
      for (size_t j = 0; j < n; ++j) {
        if (i == j) continue;
        const T dx = px - S.pred_x[j];
        const T dy = py - S.pred_y[j];
        const T dz = pz - S.pred_z[j];
        const T r2 = dx * dx + dy * dy + dz * dz;
        if (r2 < h2) S.nbr.push_back(static_cast<int>(j));
      }
      S.prefix[i] = static_cast<int>(S.nbr.size());
    }
  }

  const T density = static_cast<T>(P.density);
  const T inv_density = T(1) / density;
  const T mass = static_cast<T>(P.particle_mass);
  const T grad_scale = mass * inv_density;
  const T eps = static_cast<T>(P.epsilon);
  const T scorr_k = static_cast<T>(P.scorr_k);
  const T spiky_sign = (mut == kMutSpikySign) ? T(-1) : T(1);

  for (int iter = 0; iter < P.solver_iterations; ++iter) {
    // lambda pass — core.cpp:277-329
#pragma omp parallel for
    for (long i = 0; i < N; ++i) {
      const int b = (i == 0) ? 0 : S.prefix[i - 1];
      const int e = S.prefix[i];
      const T px = S.pred_x[i], py = S.pred_y[i], pz = S.pred_z[i];
      T rho = T(0), gsx = T(0), gsy = T(0), gsz = T(0), sg2 = T(0);
      for (int k = b; k < e; ++k) {
        const int j = S.nbr[k];
        const T dx = px - S.pred_x[j];
        const T dy = py - S.pred_y[j];
        const T dz = pz - S.pred_z[j];
        const T r2 = dx * dx + dy * dy + dz * dz;
        rho += poly6<T>(r2, h);
        if (r2 < h2) {
          const T r = std::sqrt(r2 < min_r2 ? min_r2 : r2);
          const T f = spiky_sign * spiky_factor<T>(r, h);
          const T gxx = f * dx, gyy = f * dy, gzz = f * dz;
          gsx += gxx;
          gsy += gyy;
          gsz += gzz;
          const T jx = -grad_scale * gxx;
          const T jy = -grad_scale * gyy;
          const T jz = -grad_scale * gzz;
          sg2 += jx * jx + jy * jy + jz * jz;
        }
      }
      if (mut != kMutDropSelfDensity) rho += poly6<T>(T(0), h);
      rho *= mass;
      S.rho[i] = rho;
      const T C = rho * inv_density - T(1);
      const T gix = grad_scale * gsx, giy = grad_scale * gsy, giz = grad_scale * gsz;
      sg2 += gix * gix + giy * giy + giz * giz;
      S.lambda[i] = -C / (sg2 + eps);
    }

    // delta pass with plane projection — core.cpp:331-398
    auto delta_one = [&](long i) {
      const int b = (i == 0) ? 0 : S.prefix[i - 1];
      const int e = S.prefix[i];
      const T px = S.pred_x[i], py = S.pred_y[i], pz = S.pred_z[i];
      const T li = S.lambda[i];
      T ddx = T(0), ddy = T(0), ddz = T(0);
      for (int k = b; k < e; ++k) {
        const int j = S.nbr[k];
        T dx = px - S.pred_x[j];
        T dy = py - S.pred_y[j];
        T dz = pz - S.pred_z[j];
        if (mut == kMutSwapDelta) {
          dx = -dx;
          dy = -dy;
          dz = -dz;
        }
        const T r2 = dx * dx + dy * dy + dz * dz;
        if (r2 < h2) {
          const T r = std::sqrt(r2 < min_r2 ? min_r2 : r2);
          const T f = spiky_sign * spiky_factor<T>(r, h);
          T s = li + S.lambda[j];
          if (scorr_on && inv_wdq > T(0)) {
            const T W = poly6<T>(r2, h);
            const T ratio = W * inv_wdq;
            const T sc = -scorr_k * pow_n<T>(ratio, P.scorr_n);
            s += sc;
          }
          ddx += s * f * dx;
          ddy += s * f * dy;
          ddz += s * f * dz;
        }
      }
      ddx *= inv_density;
      ddy *= inv_density;
      ddz *= inv_density;
      if (P.n_planes > 0) {
        T qx = px + ddx, qy = py + ddy, qz = pz + ddz;
        for (int p = 0; p < P.n_planes; ++p) {
          const T sd = pl_nx[p] * qx + pl_ny[p] * qy + pl_nz[p] * qz - pl_d[p];
          const T pen = -sd;
          if (pen > T(0)) {
            qx += pl_nx[p] * pen;
            qy += pl_ny[p] * pen;
            qz += pl_nz[p] * pen;
          }
        }
        ddx = qx - px;
        ddy = qy - py;
        ddz = qz - pz;
      }
      S.del_x[i] = ddx;
      S.del_y[i] = ddy;
      S.del_z[i] = ddz;
    };
    if (mut == kMutJacobiInPlace) {
      for (long i = 0; i < N; ++i) {
        delta_one(i);
        S.pred_x[i] += S.del_x[i];
        S.pred_y[i] += S.del_y[i];
        S.pred_z[i] += S.del_z[i];
      }
    } else {
#pragma omp parallel for
      for (long i = 0; i < N; ++i) delta_one(i);
      // apply — core.cpp:400-407
#pragma omp parallel for
      for (long i = 0; i < N; ++i) {
        S.pred_x[i] += S.del_x[i];
        S.pred_y[i] += S.del_y[i];
        S.pred_z[i] += S.del_z[i];
      }
    }
  }

  // velocity update and commit — core.cpp:410-421
#pragma omp parallel for
  for (long i = 0; i < N; ++i) {
    vx_[i] = (S.pred_x[i] - px_[i]) / dt;
    vy_[i] = (S.pred_y[i] - py_[i]) / dt;
    vz_[i] = (S.pred_z[i] - pz_[i]) / dt;
    px_[i] = S.pred_x[i];
    py_[i] = S.pred_y[i];
    pz_[i] = S.pred_z[i];
  }

  // XSPH — core.cpp:423-466 (rho is the last lambda pass's)
  const T visc_c = static_cast<T>(P.visc_c);
  if (P.enable_xsph && visc_c != T(0)) {
#pragma omp parallel for
    for (long i = 0; i < N; ++i) {
      const int b = (i == 0) ? 0 : S.prefix[i - 1];
      const int e = S.prefix[i];
      const T vx = vx_[i], vy = vy_[i], vz = vz_[i];
      T ax = T(0), ay = T(0), az = T(0);
      for (int k = b; k < e; ++k) {
        const int j = S.nbr[k];
        const T dx = px_[i] - px_[j];
        const T dy = py_[i] - py_[j];
        const T dz = pz_[i] - pz_[j];
        const T r2 = dx * dx + dy * dy + dz * dz;
        if (r2 < h2) {
          const T W = poly6<T>(r2, h);
          const T irho = (S.rho[j] > T(0)) ? (mass / S.rho[j]) : T(0);
          ax += (vx_[j] - vx) * W * irho;
          ay += (vy_[j] - vy) * W * irho;
          az += (vz_[j] - vz) * W * irho;
        }
      }
      S.dv_x[i] = ax;
      S.dv_y[i] = ay;
      S.dv_z[i] = az;
    }
#pragma omp parallel for
    for (long i = 0; i < N; ++i) {
      vx_[i] += visc_c * S.dv_x[i];
      vy_[i] += visc_c * S.dv_y[i];
      vz_[i] += visc_c * S.dv_z[i];
    }
  }

  // vorticity confinement — core.cpp:468-571
  const T vort_eps = static_cast<T>(P.vort_epsilon);
  if (P.enable_vorticity && vort_eps != T(0)) {
#pragma omp parallel for
    for (long i = 0; i < N; ++i) {
      const int b = (i == 0) ? 0 : S.prefix[i - 1];
      const int e = S.prefix[i];
      const T px = px_[i], py = py_[i], pz = pz_[i];
      const T vx = vx_[i], vy = vy_[i], vz = vz_[i];
      T ox = T(0), oy = T(0), oz = T(0);
      for (int k = b; k < e; ++k) {
        const int j = S.nbr[k];
        const T dx = px - px_[j];
        const T dy = py - py_[j];
        const T dz = pz - pz_[j];
        const T r2 = dx * dx + dy * dy + dz * dz;
        if (r2 < h2) {
          const T r = std::sqrt(r2 < min_r2 ? min_r2 : r2);
          const T f = spiky_sign * spiky_factor<T>(r, h);
          const T gxx = f * dx, gyy = f * dy, gzz = f * dz;
          const T dvx = vx_[j] - vx, dvy = vy_[j] - vy, dvz = vz_[j] - vz;
          ox += dvy * gzz - dvz * gyy;
          oy += dvz * gxx - dvx * gzz;
          oz += dvx * gyy - dvy * gxx;
        }
      }
      S.om_x[i] = ox;
      S.om_y[i] = oy;
      S.om_z[i] = oz;
      S.om_mag[i] = std::sqrt(ox * ox + oy * oy + oz * oz);
    }
#pragma omp parallel for
    for (long i = 0; i < N; ++i) {
      const int b = (i == 0) ? 0 : S.prefix[i - 1];
      const int e = S.prefix[i];
      const T px = px_[i], py = py_[i], pz = pz_[i];
      const T omi = S.om_mag[i];
      T ex = T(0), ey = T(0), ez = T(0);
      for (int k = b; k < e; ++k) {
        const int j = S.nbr[k];
        const T dx = px - px_[j];
        const T dy = py - py_[j];
        const T dz = pz - pz_[j];
        const T r2 = dx * dx + dy * dy + dz * dz;
        if (r2 < h2) {
          const T r = std::sqrt(r2 < min_r2 ? min_r2 : r2);
          const T f = spiky_sign * spiky_factor<T>(r, h);
          const T c = S.om_mag[j] - omi;
          ex += c * (f * dx);
          ey += c * (f * dy);
          ez += c * (f * dz);
        }
      }
      S.eta_x[i] = ex;
      S.eta_y[i] = ey;
      S.eta_z[i] = ez;
    }
    const T norm_eps = static_cast<T>(P.vort_norm_eps);
#pragma omp parallel for
    for (long i = 0; i < N; ++i) {
      const T ex = S.eta_x[i], ey = S.eta_y[i], ez = S.eta_z[i];
      const T len = std::sqrt(ex * ex + ey * ey + ez * ez);
      T nx = T(0), ny = T(0), nz = T(0);
      if (len > norm_eps) {
        const T inv = T(1) / len;
        nx = ex * inv;
        ny = ey * inv;
        nz = ez * inv;
      }
      const T ox = S.om_x[i], oy = S.om_y[i], oz = S.om_z[i];
      const T fx = vort_eps * (ny * oz - nz * oy);
      const T fy = vort_eps * (nz * ox - nx * oz);
      const T fz = vort_eps * (nx * oy - ny * ox);
      vx_[i] += dt * fx;
      vy_[i] += dt * fy;
      vz_[i] += dt * fz;
    }
  }

  // plane restitution / friction — core.cpp:573-612 (with plane_props each
  // plane's own pair; the pass runs when any plane has one above zero)
  const T rest = static_cast<T>(P.plane_restitution);
  const T fric = static_cast<T>(P.plane_friction);
  auto rest_of = [&](int p) { return P.plane_props ? static_cast<T>(P.plane_props[2 * p]) : rest; };
  auto fric_of = [&](int p) {
    return P.plane_props ? static_cast<T>(P.plane_props[2 * p + 1]) : fric;
  };
  bool contact = rest > T(0) || fric > T(0);
  if (P.plane_props) {
    contact = false;
    for (int p = 0; p < P.n_planes; ++p) contact = contact || rest_of(p) > T(0) || fric_of(p) > T(0);
  }
  if (contact && P.n_planes > 0) {
#pragma omp parallel for
    for (long i = 0; i < N; ++i) {
      T vx = vx_[i], vy = vy_[i], vz = vz_[i];
      const T px = px_[i], py = py_[i], pz = pz_[i];
      for (int p = 0; p < P.n_planes; ++p) {
        const T nx = pl_nx[p], ny = pl_ny[p], nz = pl_nz[p];
        const T sd = nx * px + ny * py + nz * pz - pl_d[p];
        if (sd <= T(0)) {
          const T vn = nx * vx + ny * vy + nz * vz;
          T vn_new = vn;
          if (vn < T(0)) vn_new = -rest_of(p) * vn;
          const T tx = vx - vn * nx;
          const T ty = vy - vn * ny;
          const T tz = vz - vn * nz;
          const T scale = T(1) - fric_of(p);
          vx = tx * scale + vn_new * nx;
          vy = ty * scale + vn_new * ny;
          vz = tz * scale + vn_new * nz;
        }
      }
      vx_[i] = vx;
      vy_[i] = vy;
      vz_[i] = vz;
    }
  }
}

template <typename T>
int run_steps(const OracleParams* P, long long n, T* pos, T* vel, int nsteps,
              T* time, OracleDiag* diag) {
  if (!P || n < 0 || nsteps < 0) return -1;
  if (P->n_planes > 0 && !P->planes) return -1;
  const T dt = static_cast<T>(P->dt);
  if (n == 0) {
    for (int s = 0; s < nsteps; ++s)
      if (time) *time += dt;  // core.cpp:122-125
    return 0;
  }
  Scratch<T> S;
  T* px = pos;
  T* py = pos + n;
  T* pz = pos + 2 * n;
  T* vx = vel;
  T* vy = vel + n;
  T* vz = vel + 2 * n;
  for (int s = 0; s < nsteps; ++s) {
    const bool last = (s + 1 == nsteps);
    T* pred0 = (last && diag) ? static_cast<T*>(diag->pred0) : nullptr;
    oracle_step_impl<T>(*P, static_cast<size_t>(n), px, py, pz, vx, vy, vz, S, pred0);
    if (time) *time += dt;
    if (last && diag) {
      if (diag->cell && P->use_uniform_grid) {
        for (long long k = 0; k < n; ++k) {
          const auto& e = S.entries[k];
          diag->cell[3 * e.particle + 0] = e.key.x;
          diag->cell[3 * e.particle + 1] = e.key.y;
          diag->cell[3 * e.particle + 2] = e.key.z;
        }
      }
      if (diag->sorted_ids && P->use_uniform_grid)
        for (long long k = 0; k < n; ++k) diag->sorted_ids[k] = S.entries[k].particle;
      if (diag->nbr_count)
        for (long long i = 0; i < n; ++i)
          diag->nbr_count[i] = S.prefix[i] - (i == 0 ? 0 : S.prefix[i - 1]);
      diag->nbr_total = static_cast<long long>(S.nbr.size());
      if (diag->nbr_ids) {
        const long long m = std::min<long long>(diag->nbr_cap, diag->nbr_total);
        std::memcpy(diag->nbr_ids, S.nbr.data(), sizeof(int) * m);
      }
      if (diag->rho) std::memcpy(diag->rho, S.rho.data(), sizeof(T) * n);
      if (diag->lambda) std::memcpy(diag->lambda, S.lambda.data(), sizeof(T) * n);
    }
  }
  return 0;
}

}  // namespace

extern "C" {

// s_corr's power for exponents outside {2, 3, 4}: 0 as the reference, 1 the
// correctly rounded power (see g_pow_mode).  Returns the previous mode.
int oracle_set_pow_mode(int mode) {
  const int old = g_pow_mode;
  g_pow_mode = mode;
  return old;
}

// pos/vel are SoA blocks of 3*n (x block, y block, z block), updated in place.
int oracle_step_f32(const OracleParams* P, long long n, float* pos, float* vel,
                    int nsteps, float* time, OracleDiag* diag) {
  return run_steps<float>(P, n, pos, vel, nsteps, time, diag);
}

int oracle_step_f64(const OracleParams* P, long long n, double* pos, double* vel,
                    int nsteps, double* time, OracleDiag* diag) {
  return run_steps<double>(P, n, pos, vel, nsteps, time, diag);
}

int oracle_set_threads(int t) {
#ifdef _OPENMP
  if (t > 0) omp_set_num_threads(t);
  return omp_get_max_threads();
#else
  (void)t;
  return 1;
#endif
}

// Closed-form helpers exposed for the kernel-coefficient tests.
double oracle_poly6_f32(double r2, double h) {
  return poly6<float>(static_cast<float>(r2), static_cast<float>(h));
}
double oracle_spiky_f32(double r, double h) {
  return spiky_factor<float>(static_cast<float>(r), static_cast<float>(h));
}

}  // extern "C"
