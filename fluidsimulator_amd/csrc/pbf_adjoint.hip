// Reverse-mode adjoint of one PBF substep on gfx950, with the forward's
// neighbor topology frozen.
//
// Forward map of one substep, particle i, vorticity off
// (core/src/core.cpp:150-612):
//   va = v + g dt,  p0 = x + va dt                               predict
//   for k < I:   rho_i = m (W0 + sum_j W(r2_ij))                  lambda pass
//                S_i = sum_j F_ij d_ij,  Q_i = gs^2 (sum_j |F_ij d_ij|^2 + |S_i|^2)
//                lambda_i = -(rho_i/rho0 - 1) / (Q_i + eps)
//                q_i = p_i + (1/rho0) sum_j (lambda_i + lambda_j + sc_ij) F_ij d_ij
//                p^{k+1}_i = Proj(q_i)                           delta + planes
//   x' = p^I,  vc = (p^I - x) / dt                                commit
//   vx = vc_i + c sum_j (vc_j - vc_i) W_ij m/rho_j                XSPH (rho of the last pass)
//   v' = R(vx) at x'                                              restitution / friction
// with d_ij = p_i - p_j, F = spiky factor (r clamped at 0.01 h), W = poly6,
// sc = -k (W / W(dq))^n, sums over the substep's list (built on p0) and,
// except for rho, only over pairs with r2 < h2 at the current iterate.
//
// The lists are symmetric and every pair term depends on d_ij only, so each
// adjoint is a gather over the same lists: particle i collects the
// derivative of its own terms and of its neighbors' terms with respect to
// p_i.  Per substep (reverse order):
//   seed     adjoint of restitution; without XSPH also of the commit
//   xsph     adjoint of XSPH and of the commit (gather)
//   then for k = I-1 .. 0:
//     delta   lambda-bar and the delta part of p^k-bar (gather); the lambda
//             formula's per-particle adjoint (rho-bar, Q-bar)
//     lambda  the density and Q parts of p^k-bar (gather); then the
//             projection of iteration k-1, or at k = 0 the predict, which
//             writes x-bar and v-bar back in original particle order
//   gravity  per-scene sum of the predict's g-bar (fixed order, double)
// The recorded forward supplies p^k (with lambda^k), S and Q, the plane and
// restitution decisions and the lists; nothing is recomputed here.
// Each gather splits a particle's list over kSplit lanes in a fixed
// interleave and adds their partial sums in a fixed order, and the gravity
// sum runs in a fixed tree, so the gradients are deterministic (bitwise
// repeatable, and independent of how a list is stored).

#include "pbf_ctx.h"
#include "pbf_device.h"

namespace fluidhip {

namespace {

__device__ __forceinline__ float3 f3(const float4& v) { return make_float3(v.x, v.y, v.z); }
__device__ __forceinline__ float dot3(const float3& a, const float3& b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}

// The gathers give each particle kSplit consecutive lanes: lane q of a
// group takes list entries q, q + kSplit, q + 2 kSplit, ... in list order
// (every lane decodes the whole run-length list), so a particle's chain of
// dependent loads is a quarter as long and four times as many wavefronts
// hide each other's latency.  group_sum adds the group's partial sums in
// one fixed order (lanes 0+1 and 2+3, then the two pairs) and gives every
// lane of the group the total.
constexpr int kSplit = 4;

template <typename Load, typename F>
__device__ __forceinline__ void for_each_neighbor_split(const NbrView& v, int s, int q,
                                                        Load&& load, F&& f) {
  const int* __restrict__ L = v.list + v.off[s];
  const int nw = v.nw[s];
  int idx = 0;  // list index of the run's first entry
  for (int w = 0; w < nw;) {
    const int e = L[w++];
    const int t0 = e & 0x7fffffff;
    const int len = e < 0 ? L[w++] : 1;
    // the first entry of this run whose list index is q modulo kSplit
    const int skip = (q - idx) & (kSplit - 1);
    for (int t = t0 + skip; t < t0 + len; t += kSplit) f(t, load(t));
    idx += len;
  }
}

__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
  for (int m = 1; m < kSplit; m <<= 1) v += __shfl_xor(v, m);
  return v;
}

__device__ __forceinline__ float3 group_sum(float3 v) {
  return make_float3(group_sum(v.x), group_sum(v.y), group_sum(v.z));
}

// Geometry of a pair with r2 < h2: the spiky factor F (clamped below
// min_r), dF/dr / r (0 for clamped pairs, where F is constant), poly6 W and
// dW/dr2.
struct PairGeo {
  float F, dFr, W, dW;
};
__device__ __forceinline__ PairGeo pair_geo(float r2, const StepConsts& c) {
  PairGeo g;
  if (r2 < c.min_r2) {
    g.F = c.f_min;
    g.dFr = 0.0f;
  } else {
    const float r = sqrtf(r2);
    const float hr = c.h - r;
    g.F = c.s_coeff * hr * hr;
    g.dFr = (r > 0.0f) ? (-2.0f * c.s_coeff * hr) / r : 0.0f;
  }
  const float diff = c.h2 - r2;
  g.W = c.w_coeff * diff * diff * diff;
  g.dW = -3.0f * c.w_coeff * diff * diff;
  return g;
}

// s_corr term and its derivative with respect to r2.
__device__ __forceinline__ float scorr_and_slope(const PairGeo& g, const StepConsts& c,
                                                 float& slope) {
  const float ratio = g.W * c.scorr_inv_wdq;
  const int n = c.scorr_n;
  float pm1;  // ratio^(n-1)
  if (n == 1)
    pm1 = 1.0f;
  else if (n == 2)
    pm1 = ratio;
  else if (n == 3)
    pm1 = ratio * ratio;
  else if (n == 4)
    pm1 = ratio * ratio * ratio;
  else
    pm1 = (n == 0) ? 0.0f : powf(ratio, static_cast<float>(n - 1));
  slope = c.scorr_neg_k * static_cast<float>(n) * pm1 * c.scorr_inv_wdq * g.dW;
  return c.scorr_neg_k * pow_ratio(ratio, n);
}

// Transpose of the projection's Jacobian: the pushing planes in reverse order,
// each (I - n n^T).
__device__ __forceinline__ float3 project_bwd(float3 v, unsigned mask, const ScenePlanes& sp) {
  for (int p = sp.n_planes - 1; p >= 0; --p) {
    if (!((mask >> p) & 1u)) continue;
    const float4 pl = sp.planes[p];
    const float vn = pl.x * v.x + pl.y * v.y + pl.z * v.z;
    v.x -= vn * pl.x;
    v.y -= vn * pl.y;
    v.z -= vn * pl.z;
  }
  return v;
}

// Transpose of restitution/friction's Jacobian (each contact plane's map is
// symmetric), planes in reverse order.
__device__ __forceinline__ float3 restitute_bwd(float3 v, unsigned mask, const ScenePlanes& sp) {
  for (int p = sp.n_planes - 1; p >= 0; --p) {
    if (!((mask >> p) & 1u)) continue;
    const float t = 1.0f - sp.fric_of[p];
    const float a = ((mask >> (8 + p)) & 1u) ? -sp.rest_of[p] : 1.0f;
    const float4 pl = sp.planes[p];
    const float vn = pl.x * v.x + pl.y * v.y + pl.z * v.z;
    const float k = (a - t) * vn;
    v.x = t * v.x + k * pl.x;
    v.y = t * v.y + k * pl.y;
    v.z = t * v.z + k * pl.z;
  }
  return v;
}

// Adjoint of the commit (x' = p^I, vc = (p^I - x)/dt) and of the last
// iteration's projection: leaves p^{I-1}-side q-bar in PB and w = q-bar/rho0.
__device__ __forceinline__ void commit_bwd(const StepConsts& c, int s, float3 xbar, float3 vcbar,
                                           unsigned mask, const ScenePlanes& sp,
                                           float4* __restrict__ PB, float4* __restrict__ XB,
                                           float4* __restrict__ Wv) {
  const float3 vd = make_float3(vcbar.x / c.dt, vcbar.y / c.dt, vcbar.z / c.dt);
  XB[s] = make_float4(-vd.x, -vd.y, -vd.z, 0.0f);
  const float3 q =
      project_bwd(make_float3(xbar.x + vd.x, xbar.y + vd.y, xbar.z + vd.z), mask, sp);
  PB[s] = make_float4(q.x, q.y, q.z, 0.0f);
  Wv[s] = make_float4(q.x * c.inv_density, q.y * c.inv_density, q.z * c.inv_density, 0.0f);
}

__global__ __launch_bounds__(256) void k_adj_seed(
    StepConsts c, const float4* __restrict__ X, const unsigned* __restrict__ RM,
    const unsigned* __restrict__ Mlast, const int* __restrict__ seg,
    const ScenePlanes* __restrict__ sp, const float4* __restrict__ GX,
    const float4* __restrict__ GV, float4* __restrict__ U, float4* __restrict__ PB,
    float4* __restrict__ XB, float4* __restrict__ Wv, float* __restrict__ RB, int xsph) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= c.n) return;
  const int oid = __float_as_int(X[s].w);
  const float3 xbar = f3(GX[oid]);
  float3 vbar = f3(GV[oid]);
  const ScenePlanes& sc = sp[scene_of(s, seg, c.n_scenes)];
  if (RM) vbar = restitute_bwd(vbar, RM[s], sc);
  if (xsph) {
    U[s] = make_float4(vbar.x, vbar.y, vbar.z, 0.0f);
    PB[s] = make_float4(xbar.x, xbar.y, xbar.z, 0.0f);
    return;
  }
  RB[s] = 0.0f;
  commit_bwd(c, s, xbar, vbar, Mlast[s], sc, PB, XB, Wv);
}

// Neighbor record of the XSPH adjoint: position, commit velocity and rho,
// velocity adjoint.
struct XNb {
  float3 x;
  float4 v;
  float3 u;
};

__global__ __launch_bounds__(256) void k_adj_xsph(
    StepConsts c, NbrView nv, const float4* __restrict__ PI, const float4* __restrict__ VC,
    const float4* __restrict__ U, const unsigned* __restrict__ Mlast, const int* __restrict__ seg,
    const ScenePlanes* __restrict__ sp, float4* __restrict__ PB, float4* __restrict__ XB,
    float4* __restrict__ Wv, float* __restrict__ RB) {
  const int gi = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = gi / kSplit, qi = gi & (kSplit - 1);
  if (s >= c.n) return;  // uniform over a lane group
  const float3 xi = ld3(PI, s);
  const float4 vi = VC[s];
  const float3 ui = ld3(U, s);
  const float mui = (vi.w > 0.0f) ? (c.mass / vi.w) : 0.0f;
  float3 av = make_float3(0.f, 0.f, 0.f), ax = make_float3(0.f, 0.f, 0.f);
  float ar = 0.0f;
  for_each_neighbor_split(
      nv, s, qi, [&](int t) { return XNb{ld3(PI, t), VC[t], ld3(U, t)}; },
      [&](int, const XNb& q) {
        const float3 d = make_float3(xi.x - q.x.x, xi.y - q.x.y, xi.z - q.x.z);
        const float r2 = d.x * d.x + d.y * d.y + d.z * d.z;
        if (!(r2 < c.h2)) return;
        const float diff = c.h2 - r2;
        const float W = c.w_coeff * diff * diff * diff;
        const float dW = -3.0f * c.w_coeff * diff * diff;
        const float muj = (q.v.w > 0.0f) ? (c.mass / q.v.w) : 0.0f;
        // vc-bar_i: W (mu_i u_j - mu_j u_i)
        av.x += W * (mui * q.u.x - muj * ui.x);
        av.y += W * (mui * q.u.y - muj * ui.y);
        av.z += W * (mui * q.u.z - muj * ui.z);
        // x'-bar_i: dW ((v_j - v_i) . (mu_j u_i - mu_i u_j)) d
        const float3 dv = make_float3(q.v.x - vi.x, q.v.y - vi.y, q.v.z - vi.z);
        const float3 e = make_float3(muj * ui.x - mui * q.u.x, muj * ui.y - mui * q.u.y,
                                     muj * ui.z - mui * q.u.z);
        const float k = dW * dot3(dv, e);
        ax.x += k * d.x;
        ax.y += k * d.y;
        ax.z += k * d.z;
        // rho-bar_i: W u_j . (v_i - v_j)
        ar -= W * dot3(q.u, dv);
      });
  av = group_sum(av);
  ax = group_sum(ax);
  ar = group_sum(ar);
  if (qi != 0) return;
  const float cv = c.visc_c;
  const float3 vcbar = make_float3(ui.x + cv * av.x, ui.y + cv * av.y, ui.z + cv * av.z);
  const float4 pb = PB[s];
  const float3 xbar =
      make_float3(pb.x + 2.0f * cv * ax.x, pb.y + 2.0f * cv * ax.y, pb.z + 2.0f * cv * ax.z);
  RB[s] = (vi.w > 0.0f) ? -(mui / vi.w) * cv * ar : 0.0f;
  commit_bwd(c, s, xbar, vcbar, Mlast[s], sp[scene_of(s, seg, c.n_scenes)], PB, XB, Wv);
}

// Neighbor record of the delta adjoint: p^k, lambda^k and w = delta-bar/rho0.
struct DNb {
  float4 p;
  float3 w;
};

__global__ __launch_bounds__(256) void k_adj_delta(StepConsts c, NbrView nv,
                                                   const float4* __restrict__ P,
                                                   const float4* __restrict__ Wv,
                                                   const float4* __restrict__ SQ,
                                                   float4* __restrict__ PB, float* __restrict__ RB,
                                                   float4* __restrict__ LB, int add_rb) {
  const int gi = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = gi / kSplit, qi = gi & (kSplit - 1);
  if (s >= c.n) return;  // uniform over a lane group
  const float4 pi = P[s];
  const float3 wi = ld3(Wv, s);
  float lb = 0.0f;
  float3 gp = make_float3(0.f, 0.f, 0.f);
  for_each_neighbor_split(
      nv, s, qi, [&](int t) { return DNb{P[t], ld3(Wv, t)}; },
      [&](int, const DNb& q) {
        const float3 d = make_float3(pi.x - q.p.x, pi.y - q.p.y, pi.z - q.p.z);
        const float r2 = d.x * d.x + d.y * d.y + d.z * d.z;
        if (!(r2 < c.h2)) return;
        const PairGeo g = pair_geo(r2, c);
        float sv = pi.w + q.p.w;
        float sslope = 0.0f;
        if (c.scorr_on) sv += scorr_and_slope(g, c, sslope);
        const float3 u = make_float3(wi.x - q.w.x, wi.y - q.w.y, wi.z - q.w.z);
        const float ud = dot3(u, d);
        lb += g.F * ud;
        const float a = sv * g.F;
        const float b = ud * (sv * g.dFr + 2.0f * g.F * sslope);
        gp.x += a * u.x + b * d.x;
        gp.y += a * u.y + b * d.y;
        gp.z += a * u.z + b * d.z;
      });
  lb = group_sum(lb);
  gp = group_sum(gp);
  if (qi != 0) return;
  // lambda = -C / (Q + eps), C = rho/rho0 - 1
  const float4 sq = SQ[s];
  const float den = sq.w + c.eps;
  const float cbar = -lb / den;
  const float qbar = -lb * pi.w / den;
  RB[s] = cbar * c.inv_density + (add_rb ? RB[s] : 0.0f);
  LB[s] = make_float4(qbar * sq.x, qbar * sq.y, qbar * sq.z, qbar);
  const float4 pb = PB[s];
  PB[s] = make_float4(pb.x + gp.x, pb.y + gp.y, pb.z + gp.z, 0.0f);
}

// Neighbor record of the lambda adjoint: p^k, (Q-bar S, Q-bar), rho-bar.
struct LNb {
  float3 p;
  float4 l;
  float r;
};

__global__ __launch_bounds__(256) void k_adj_lambda(
    StepConsts c, NbrView nv, const float4* __restrict__ P, const float4* __restrict__ LB,
    const float* __restrict__ RB, float4* __restrict__ PB, const unsigned* __restrict__ Mprev,
    float4* __restrict__ Wv, const float4* __restrict__ X, const float4* __restrict__ XB,
    float4* __restrict__ GX, float4* __restrict__ GV, const int* __restrict__ seg,
    const ScenePlanes* __restrict__ sp) {
  const int gi = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = gi / kSplit, qi = gi & (kSplit - 1);
  if (s >= c.n) return;  // uniform over a lane group
  const float3 pi = ld3(P, s);
  const float4 li = LB[s];
  const float ri = RB[s];
  const float g2 = 2.0f * c.grad_scale * c.grad_scale;
  const float m2 = 2.0f * c.mass;
  float3 gp = make_float3(0.f, 0.f, 0.f);
  for_each_neighbor_split(
      nv, s, qi, [&](int t) { return LNb{ld3(P, t), LB[t], RB[t]}; },
      [&](int, const LNb& q) {
        const float3 d = make_float3(pi.x - q.p.x, pi.y - q.p.y, pi.z - q.p.z);
        const float r2 = d.x * d.x + d.y * d.y + d.z * d.z;
        if (!(r2 < c.h2)) return;
        const PairGeo g = pair_geo(r2, c);
        // density: (rho-bar_i + rho-bar_j) m dW/dr2 2 d
        const float kr = (ri + q.r) * m2 * g.dW;
        // Q: u = 2 gs^2 ((Qb_i + Qb_j) F d + Qb_i S_i - Qb_j S_j)
        const float qs = (li.w + q.l.w) * g.F;
        const float3 u = make_float3(g2 * (qs * d.x + li.x - q.l.x), g2 * (qs * d.y + li.y - q.l.y),
                                     g2 * (qs * d.z + li.z - q.l.z));
        const float b = dot3(u, d) * g.dFr + kr;
        gp.x += g.F * u.x + b * d.x;
        gp.y += g.F * u.y + b * d.y;
        gp.z += g.F * u.z + b * d.z;
      });
  gp = group_sum(gp);
  if (qi != 0) return;
  const float4 pb0 = PB[s];
  const float3 pb = make_float3(pb0.x + gp.x, pb0.y + gp.y, pb0.z + gp.z);
  if (Mprev) {
    const float3 q = project_bwd(pb, Mprev[s], sp[scene_of(s, seg, c.n_scenes)]);
    PB[s] = make_float4(q.x, q.y, q.z, 0.0f);
    Wv[s] = make_float4(q.x * c.inv_density, q.y * c.inv_density, q.z * c.inv_density, 0.0f);
    return;
  }
  // predict: p0 = x + (v + g dt) dt
  const int oid = __float_as_int(X[s].w);
  const float4 xb = XB[s];
  GX[oid] = make_float4(xb.x + pb.x, xb.y + pb.y, xb.z + pb.z, 0.0f);
  GV[oid] = make_float4(c.dt * pb.x, c.dt * pb.y, c.dt * pb.z, 0.0f);
  PB[s] = make_float4(pb.x, pb.y, pb.z, 0.0f);
}

// g-bar of each scene += dt^2 * sum of p0-bar over its slots, summed in a
// fixed order in double: one workgroup of 1024 threads per scene, each thread
// with four accumulators over interleaved slots (independent chains, so the
// loads overlap), folded in a fixed order and then across the workgroup.
__global__ __launch_bounds__(1024) void k_adj_gravity(StepConsts c, const int* __restrict__ seg,
                                                      const float4* __restrict__ PB,
                                                      double* __restrict__ gacc) {
  constexpr int kT = 1024;
  __shared__ double lds[3][kT];
  const int sc = blockIdx.x;
  const int a = seg[sc], b = seg[sc + 1];
  double sx[4] = {0.0, 0.0, 0.0, 0.0}, sy[4] = {0.0, 0.0, 0.0, 0.0},
         sz[4] = {0.0, 0.0, 0.0, 0.0};
  int i = a + static_cast<int>(threadIdx.x);
  for (; i + 3 * kT < b; i += 4 * kT) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const float4 v = PB[i + u * kT];
      sx[u] += v.x;
      sy[u] += v.y;
      sz[u] += v.z;
    }
  }
  for (int u = 0; i < b; i += kT, ++u) {
    const float4 v = PB[i];
    sx[u] += v.x;
    sy[u] += v.y;
    sz[u] += v.z;
  }
  lds[0][threadIdx.x] = (sx[0] + sx[1]) + (sx[2] + sx[3]);
  lds[1][threadIdx.x] = (sy[0] + sy[1]) + (sy[2] + sy[3]);
  lds[2][threadIdx.x] = (sz[0] + sz[1]) + (sz[2] + sz[3]);
  __syncthreads();
  for (int w = kT / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) {
      lds[0][threadIdx.x] += lds[0][threadIdx.x + w];
      lds[1][threadIdx.x] += lds[1][threadIdx.x + w];
      lds[2][threadIdx.x] += lds[2][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double dt2 = static_cast<double>(c.dt) * static_cast<double>(c.dt);
    gacc[3 * sc + 0] += dt2 * lds[0][0];
    gacc[3 * sc + 1] += dt2 * lds[1][0];
    gacc[3 * sc + 2] += dt2 * lds[2][0];
  }
}

// The tape copy of a recorded substep, on the device: the substep's list
// length is a device value, so no host round trip sits between substeps.
__global__ __launch_bounds__(256) void k_tape_copy(
    int n, const float4* __restrict__ posB, const float4* __restrict__ posA,
    const float4* __restrict__ velB, const long long* __restrict__ noff,
    const int* __restrict__ nwords, const int* __restrict__ ncount, const int* __restrict__ nbr,
    const unsigned long long* __restrict__ nbr_total, float4* __restrict__ X,
    float4* __restrict__ PI, float4* __restrict__ VC, long long* __restrict__ tnoff,
    int* __restrict__ tnwords, int* __restrict__ tncount, int* __restrict__ tnbr, long long tcap,
    unsigned long long* __restrict__ need, const unsigned* __restrict__ flags) {
  if (*flags & kFlagFatal) return;  // the substep failed and will run again
  const unsigned long long words = *nbr_total;
  if (words > static_cast<unsigned long long>(tcap)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicMax(need, words);
    return;
  }
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    X[i] = posB[i];
    PI[i] = posA[i];
    if (VC) VC[i] = velB[i];
    tnoff[i] = noff[i];
    tnwords[i] = nwords[i];
    tncount[i] = ncount[i];
  }
  const long long w = static_cast<long long>(words);
  for (long long i = blockIdx.x * blockDim.x + threadIdx.x; i < w; i += stride) tnbr[i] = nbr[i];
}

__global__ void k_soa_to_f4(int n, const float* __restrict__ soa, float4* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t N = static_cast<size_t>(n);
  out[i] = soa ? make_float4(soa[i], soa[N + i], soa[2 * N + i], 0.0f)
               : make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ void k_f4_to_soa(int n, const float4* __restrict__ in, float* __restrict__ soa) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t N = static_cast<size_t>(n);
  const float4 v = in[i];
  soa[i] = v.x;
  soa[N + i] = v.y;
  soa[2 * N + i] = v.z;
}

__global__ void k_zero_f64(int n, double* __restrict__ p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

inline int nblocks(long long n, int bs) { return static_cast<int>((n + bs - 1) / bs); }

}  // namespace

hipError_t enqueue_adjoint_seed(Ctx& c, const float* d_dpos, const float* d_dvel,
                                const AdjBuffers& b) {
  if (c.n > 0) {
    k_soa_to_f4<<<nblocks(c.n, 256), 256, 0, c.stream>>>(c.n, d_dpos, b.GX);
    k_soa_to_f4<<<nblocks(c.n, 256), 256, 0, c.stream>>>(c.n, d_dvel, b.GV);
  }
  k_zero_f64<<<nblocks(3 * c.n_scenes, 256), 256, 0, c.stream>>>(3 * c.n_scenes, b.gacc);
  return hipGetLastError();
}

hipError_t enqueue_adjoint_result(Ctx& c, float* d_dpos, float* d_dvel, const AdjBuffers& b) {
  if (c.n > 0) {
    k_f4_to_soa<<<nblocks(c.n, 256), 256, 0, c.stream>>>(c.n, b.GX, d_dpos);
    k_f4_to_soa<<<nblocks(c.n, 256), 256, 0, c.stream>>>(c.n, b.GV, d_dvel);
  }
  return hipGetLastError();
}

hipError_t enqueue_adjoint_substep(Ctx& c, const SubTape& t, const AdjBuffers& b) {
  if (c.n == 0) return hipSuccess;
  const StepConsts k = c.consts;
  const int I = c.iterations;
  const int n = c.n;
  const int nb = nblocks(n, 256);
  const int ng = nblocks(static_cast<long long>(n) * kSplit, 256);  // the gathers' lane groups
  hipStream_t st = c.stream;
  const NbrView nv{t.noff, t.nwords, t.nbr, t.ncount};
  const size_t N = static_cast<size_t>(n);
  const unsigned* Mlast = t.M + static_cast<size_t>(I - 1) * N;
  k_adj_seed<<<nb, 256, 0, st>>>(k, t.X, c.restitution_any ? t.RM : nullptr, Mlast, c.seg,
                                 c.planes, b.GX, b.GV, b.U, b.PB, b.XB, b.W, b.RB,
                                 c.xsph ? 1 : 0);
  if (c.xsph)
    k_adj_xsph<<<ng, 256, 0, st>>>(k, nv, t.PI, t.VC, b.U, Mlast, c.seg, c.planes, b.PB, b.XB,
                                   b.W, b.RB);
  for (int it = I - 1; it >= 0; --it) {
    const float4* P = t.P + static_cast<size_t>(it) * N;
    const float4* SQ = t.SQ + static_cast<size_t>(it) * N;
    k_adj_delta<<<ng, 256, 0, st>>>(k, nv, P, b.W, SQ, b.PB, b.RB, b.LB,
                                    (it == I - 1 && c.xsph) ? 1 : 0);
    const unsigned* Mprev = it > 0 ? t.M + static_cast<size_t>(it - 1) * N : nullptr;
    k_adj_lambda<<<ng, 256, 0, st>>>(k, nv, P, b.LB, b.RB, b.PB, Mprev, b.W, t.X, b.XB, b.GX,
                                     b.GV, c.seg, c.planes);
  }
  k_adj_gravity<<<c.n_scenes, 1024, 0, st>>>(k, c.seg, b.PB, b.gacc);
  return hipGetLastError();
}

hipError_t enqueue_tape_copy(Ctx& c, SubTape& t, unsigned long long* need) {
  if (c.n == 0) return hipSuccess;
  k_tape_copy<<<1024, 256, 0, c.stream>>>(c.n, c.posB, c.posA, c.velB, c.noff, c.nwords, c.ncount,
                                          c.nbr, c.nbr_total, t.X, t.PI, c.xsph ? t.VC : nullptr,
                                          t.noff, t.nwords, t.ncount, t.nbr, t.nbr_cap, need,
                                          c.flags);
  return hipGetLastError();
}

}  // namespace fluidhip
