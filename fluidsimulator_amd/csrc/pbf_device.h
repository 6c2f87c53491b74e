// Device-side data structures and helpers of the gfx950 PBF step.
//
// Arithmetic in this file follows the reference CPU step
// (core/src/core.cpp:35-71, :119-615) operation for operation; the whole
// library is compiled with -ffp-contract=off and IEEE division/sqrt, so each
// expression rounds exactly as the reference's does.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fluidhip {

constexpr int kMaxPlanes = 8;
constexpr int kCoordLimit = 1 << 20;

// Device error flags (sticky until the host clears them).
constexpr unsigned kFlagRange = 1u;     // non-finite / out-of-range position
constexpr unsigned kFlagCells = 2u;     // cell table capacity exceeded
constexpr unsigned kFlagNeighbors = 4u; // neighbor list capacity exceeded
constexpr unsigned kFlagFatal = kFlagRange | kFlagCells | kFlagNeighbors;

// Loop-invariant constants of one substep, computed on the host with the
// same float expressions the reference evaluates inside its loops.
// Items per workgroup of the radix sort that orders slots by (cell, id).
// A scatter workgroup ranks its items in serial rounds of 256, so small
// tiles (more workgroups per CU) hide more of each round's latency: 1024
// took the 65,536-particle sort from 86 to 54-62 us and the 1,048,576-particle
// one from 148 to 135 us against 4096 (2048: 71 and 133 us).
constexpr int kRadixTile = 1024;

struct StepConsts {
  int n;            // particles, all scenes
  int n_scenes;
  float dt;
  float h;
  float inv_h;      // 1.0f / h            (core.cpp:29)
  float h2;         // h * h               (core.cpp:140)
  float min_r2;     // (0.01f*h)^2         (core.cpp:141-142)
  float f_min;      // spiky(sqrt(min_r2)): the factor of every clamped pair
  float lam_safe;   // |lambda| below which (li + lj + scorr) * f_min is finite
  float w_coeff;    // 315 / (64 pi h^9)   (core.cpp:44)
  float w0;         // poly6(0, h)         (core.cpp:319)
  float s_coeff;    // -45 / (pi h^6)      (core.cpp:55)
  float inv_density;
  float mass;
  float grad_scale; // m * (1/rho0)        (core.cpp:273)
  float eps;
  int scorr_on;     // enable_scorr && k != 0 && inv_wdq > 0
  float scorr_inv_wdq;
  float scorr_neg_k;  // -k
  int scorr_n;
  float visc_c;
  float vort_eps;
  float vort_norm_eps;
  int record_diag;
  int xcd_remap;
  int coop;         // 1: wavefronts sharing a cell use shared candidate tiles
  int coop_bulk;    // 1: ... and whole coincident tiles in one step
  int runs;         // 1: one lane per run of identical consecutive slots
  int full_lists;   // 1: cooperative particles get stored lists too (otherwise
                    //    their passes retest the list's distance predicate)
  int count_only;   // 1: the neighbor walk only counts pairs (statistics)
  int tiles;        // 1: sparse 4x4x4-cell bricks run the LDS-tiled passes
};

struct ScenePlanes {
  float gx, gy, gz;
  int n_planes;
  int restitution_on;  // e > 0 || f > 0 for some plane (core.cpp:573)
  int pad[3];
  float4 planes[kMaxPlanes];  // (nx, ny, nz, d)
  float rest_of[kMaxPlanes];  // restitution e of each plane (the scene's
                              // plane_restitution unless given per plane)
  float fric_of[kMaxPlanes];  // friction f of each plane
};

// Per-scene grid descriptor, rebuilt every substep from the predicted
// positions' bounding box.
struct GridDesc {
  int ox, oy, oz;     // cell coordinate of table origin (bbox min - 1)
  int nbx, nby, nbz;  // bricks per axis
  int cell_off;       // first key of this scene
  int ncells;         // nbx*nby*nbz bricks of 2^(3 bshift) cells
  int bshift;         // log2 of the cells per brick edge (0..3)
};

__device__ __forceinline__ float poly6(float r2, const StepConsts& c) {
  // core.cpp:35-46 with the coefficient hoisted (same float value).
  if (r2 > c.h2) return 0.0f;
  const float diff = c.h2 - r2;
  const float diff3 = diff * diff * diff;
  return c.w_coeff * diff3;
}

// The correctly rounded square root (IEEE sqrt, as sqrtf under
// -fhip-fp32-correctly-rounded-divide-sqrt).  For finite x >= 2^-96 the
// hardware estimate s is corrected by the residuals x - s*s' of its float
// neighbours s' -- the same correction the compiler emits for sqrtf, whose
// other parts (scaling tiny inputs, passing 0 and inf through) do nothing
// there, so both give the same bits; other x go through sqrtf itself.  The
// gather passes call it once per neighbor, where those parts cost a third of
// the square root.
__device__ __forceinline__ float sqrt_core(float x) {
  // the correction alone: finite x >= 2^-96 only
  const float s = __builtin_amdgcn_sqrtf(x);
  const float dn = __int_as_float(__float_as_int(s) - 1);
  const float up = __int_as_float(__float_as_int(s) + 1);
  const float rdn = fmaf(-dn, s, x);
  const float rup = fmaf(-up, s, x);
  const float r = rdn <= 0.0f ? dn : s;
  return rup > 0.0f ? up : r;
}

__device__ __forceinline__ float sqrt_rn(float x) {
  if (x >= 0x1p-96f && x <= 0x1.fffffep+127f) return sqrt_core(x);
  return sqrtf(x);
}

__device__ __forceinline__ float spiky(float r, const StepConsts& c) {
  // core.cpp:48-57 with the coefficient hoisted.
  if (r > c.h) return 0.0f;
  const float diff = c.h - r;
  return c.s_coeff * diff * diff;
}

__device__ __forceinline__ float pow_ratio(float ratio, int n) {
  // core.cpp:59-71.  n outside {2,3,4} goes through powf, whose last bit
  // may differ from the host libm's.
  if (n == 2) return ratio * ratio;
  if (n == 3) return ratio * ratio * ratio;
  if (n == 4) {
    const float r2 = ratio * ratio;
    return r2 * r2;
  }
  return powf(ratio, static_cast<float>(n));
}

// Spiky factor of a pair at squared distance r2 < h2: pairs closer than
// min_r share one precomputed factor (core.cpp:303-304 clamps r there), so
// only the others pay for the IEEE square root.
__device__ __forceinline__ float pair_spiky(float r2, const StepConsts& c) {
  if (r2 < c.min_r2) return c.f_min;
  return spiky(sqrt_rn(r2), c);
}

// The neighbor predicate: candidate q is a neighbor of a particle at p when
// |p - q|^2 < h^2, both at their predicted positions p^0 (core.cpp:140-142
// and 222-245).  Lists and neighbor masks are built with it, and cooperative
// walks retest it instead of reading stored lists, so all must use this one
// expression.
__device__ __forceinline__ bool near0(const float3& p, const float3& q, const StepConsts& c) {
  const float ex = p.x - q.x;
  const float ey = p.y - q.y;
  const float ez = p.z - q.z;
  return ex * ex + ey * ey + ez * ez < c.h2;
}

// Lanes whose own cell holds at least this many particles are served by
// wavefront-cooperative walks (one per distinct dense cell in the
// wavefront); the others walk alone.
constexpr int kDenseCell = 32;
// ... and so are lanes with at least this many neighbors (or, while the
// lists are built, candidates).
constexpr int kHeavyList = 256;

// Z-order of a cell inside its brick: the low `bits` bits of x, y, z
// interleaved as x0 y0 z0 x1 y1 z1 ...
__device__ __forceinline__ int morton_in_brick(int x, int y, int z, int bits) {
  int m = 0;
  for (int b = 0; b < bits; ++b)
    m |= (((x >> b) & 1) << (3 * b)) | (((y >> b) & 1) << (3 * b + 1)) |
         (((z >> b) & 1) << (3 * b + 2));
  return m;
}

// Key of the cell at table-relative coordinates (rx, ry, rz) >= 0: bricks
// in x-fastest order, cells in Z-order inside a brick.  The key only sets
// where a cell's particles sit in the sorted arrays; every sum visits cells
// in (dz, dy, dx) order and particles by id, so the brick edge never
// changes a result.
__device__ __forceinline__ int cell_key(int rx, int ry, int rz, const GridDesc& g) {
  const int s = g.bshift;
  const int m = (1 << s) - 1;
  const int b = ((rz >> s) * g.nby + (ry >> s)) * g.nbx + (rx >> s);
  return g.cell_off + (b << (3 * s)) + morton_in_brick(rx & m, ry & m, rz & m, s);
}

__device__ __forceinline__ int floor_cell(float v, float inv_h, bool& bad) {
  // static_cast<int>(std::floor(v * inv)) — core.cpp:28-34
  const float s = v * inv_h;
  if (!(fabsf(s) < static_cast<float>(kCoordLimit))) {
    bad = true;
    return 0;
  }
  return static_cast<int>(floorf(s));
}

__device__ __forceinline__ uint64_t pack_cell(int x, int y, int z) {
  const uint64_t m = (1ull << 21) - 1ull;
  return ((uint64_t(x + kCoordLimit) & m) << 42) | ((uint64_t(y + kCoordLimit) & m) << 21) |
         (uint64_t(z + kCoordLimit) & m);
}

__device__ __forceinline__ void unpack_cell(uint64_t k, int& x, int& y, int& z) {
  const uint64_t m = (1ull << 21) - 1ull;
  x = int((k >> 42) & m) - kCoordLimit;
  y = int((k >> 21) & m) - kCoordLimit;
  z = int(k & m) - kCoordLimit;
}

// Scene of a slot: scenes are contiguous, fixed-size slot segments.
__device__ __forceinline__ int scene_of(int slot, const int* __restrict__ seg, int n_scenes) {
  if (n_scenes == 1) return 0;
  int lo = 0, hi = n_scenes - 1;  // seg has n_scenes+1 entries
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (seg[mid] <= slot)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Workgroup -> particle tile, so that each XCD (dispatched round-robin)
// streams a contiguous range of the sorted arrays through its own L2.
__device__ __forceinline__ int tile_index(int remap) {
  const int bid = blockIdx.x;
  if (!remap) return bid;
  const int nb = gridDim.x;
  const int q = nb >> 3, r = nb & 7;
  const int xcd = bid & 7, local = bid >> 3;
  return xcd * q + (xcd < r ? xcd : r) + local;
}

__device__ __forceinline__ float3 ld3(const float4* p, int i) {
  const float* f = reinterpret_cast<const float*>(p + i);
  return make_float3(f[0], f[1], f[2]);
}

// Sequential push-out along each plane normal (core.cpp:372-389).  Returns
// the planes that pushed (bit p), which fixes the Jacobian of the
// projection: I - n n^T per pushing plane, applied in plane order.
__device__ __forceinline__ unsigned project_planes(float& qx, float& qy, float& qz,
                                                   const ScenePlanes& sp) {
  unsigned mask = 0;
  for (int p = 0; p < sp.n_planes; ++p) {
    const float4 pl = sp.planes[p];
    const float sd = pl.x * qx + pl.y * qy + pl.z * qz - pl.w;
    const float pen = -sd;
    if (pen > 0.0f) {
      qx += pl.x * pen;
      qy += pl.y * pen;
      qz += pl.z * pen;
      mask |= 1u << p;
    }
  }
  return mask;
}

// Plane restitution and friction (core.cpp:573-612), with each plane's own
// e and f (equal to the scene's globals unless set per plane).  Returns the planes in
// contact (bit p) and, among those, the ones the velocity approached
// (bit 8 + p): with them the map is linear, v -> (1-f) v + (a - (1-f)) (n.v) n
// with a = -e when approaching and 1 otherwise.
__device__ __forceinline__ unsigned restitute(float& vx, float& vy, float& vz, float px, float py,
                                              float pz, const ScenePlanes& sp) {
  if (!sp.restitution_on) return 0;
  unsigned mask = 0;
  for (int p = 0; p < sp.n_planes; ++p) {
    const float4 pl = sp.planes[p];
    const float sd = pl.x * px + pl.y * py + pl.z * pz - pl.w;
    if (sd <= 0.0f) {
      const float vn = pl.x * vx + pl.y * vy + pl.z * vz;
      float vn_new = vn;
      mask |= 1u << p;
      if (vn < 0.0f) {
        vn_new = -sp.rest_of[p] * vn;
        mask |= 1u << (8 + p);
      }
      const float tx = vx - vn * pl.x;
      const float ty = vy - vn * pl.y;
      const float tz = vz - vn * pl.z;
      const float scale = 1.0f - sp.fric_of[p];
      vx = tx * scale + vn_new * pl.x;
      vy = ty * scale + vn_new * pl.y;
      vz = tz * scale + vn_new * pl.z;
    }
  }
  return mask;
}

// Run-length-encoded neighbor lists in slot space: a word t >= 0 is the
// single neighbor t; a word with the top bit set starts a run of
// consecutive slots (t & 0x7fffffff) .. + next word - 1.
struct NbrView {
  const long long* __restrict__ off;
  const int* __restrict__ nw;
  const int* __restrict__ list;
  const int* __restrict__ count;
};

// Calls f(t, load(t)) for each neighbor slot t of slot s, in list order.
// Runs are walked four slots at a time with the four loads issued before
// any use, which keeps several memory requests in flight per lane; the
// callbacks still run strictly in list order, so every sum accumulates in
// the reference's order.
template <typename Load, typename F>
__device__ __forceinline__ void for_each_neighbor(const NbrView& v, int s, Load&& load, F&& f) {
  const int* __restrict__ L = v.list + v.off[s];
  const int nw = v.nw[s];
  for (int w = 0; w < nw;) {
    const int e = L[w++];
    int t = e & 0x7fffffff;
    int end = t + 1;
    if (e < 0) end = t + L[w++];
    for (; t + 4 <= end; t += 4) {
      const auto a = load(t);
      const auto b = load(t + 1);
      const auto c = load(t + 2);
      const auto d = load(t + 3);
      f(t, a);
      f(t + 1, b);
      f(t + 2, c);
      f(t + 3, d);
    }
    for (; t < end; ++t) f(t, load(t));
  }
}

}  // namespace fluidhip
