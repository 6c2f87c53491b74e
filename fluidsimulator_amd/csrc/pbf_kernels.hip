// gfx950 kernels of one PBF substep and the plan that queues them.
//
// Functional spec: the reference CPU step `fluid::step`
// (core/src/core.cpp:119-615).  Everything here is written for CDNA4:
// 64-lane wavefronts, float4 AoS state resident in HBM, a dense per-scene
// cell table of e x e x e-cell bricks (Z-order inside a brick), a stable LSD
// radix sort of the cell keys into the reference's (cell, id) order, and
// run-length-encoded neighbor lists allocated per workgroup.
//
// Phase order of one substep (one launch each unless noted; every launch
// covers all scenes of a batch):
//   predict_bbox  v' = v + g dt, pred0 = x + v' dt, cell = floor(pred0/h),
//                 per-tile bounding boxes of cells        (core.cpp:150-161)
//   grid_setup    per-scene table origin/extent, key offsets, counters
//   cell_keys     key of each particle
//   sort          keys in original-id order, then 8-bit LSD radix passes
//                 (digit histogram, its scan, stable ballot-ranked scatter):
//                 sorting stably by key leaves each cell's particles in
//                 ascending id, the reference's order (core.cpp:164-203)
//   cell_count    cell histogram from the sorted keys' runs, its exclusive
//                 scan -> cell starts (two launches: range sums, then each
//                 range scanned)
//   reorder       gather each particle's state into its slot, recompute pred0
//   tiles         with 4-cell bricks: the bricks that run through LDS (the
//                 particles of their good cells) and their neighbor lists
//                 (pbf_tiles.hip; on the side stream in large batches)
//   runs          runs of identical consecutive particles and the lane map of
//                 the gather passes (dense cells from a wave boundary; tile
//                 slots get no lane)
//   neighbors     27 cells in (dz, dy, dx) order, each by ascending id,
//                 j != i and r^2 < h^2 on pred0       (core.cpp:205-247)
//   lambda/delta  x solver_iterations, Jacobi double buffer (core.cpp:277-408);
//                 the last delta also commits x and v  (core.cpp:410-421)
//   xsph          (core.cpp:423-466)
//   vorticity     omega, then eta + confinement        (core.cpp:468-571)
//   plane restitution/friction fused into whichever kernel ends the chain
//                                                      (core.cpp:573-612)
// Each gather pass also runs its tile kernel, and after it k_run_fill copies
// each run's outputs from its first slot to the others.

#include <climits>

#include "exact_sum.h"
#include "pbf_ctx.h"
#include "pbf_device.h"

namespace fluidhip {

namespace {

constexpr int kWave = 64;
// Bound below which a difference of two floats cannot overflow.
constexpr float kHalfMax = 1.0e38f;

// Kernels launched after the grid setup stop as soon as an earlier phase has
// raised a fatal flag; the committed state is then left untouched.
__device__ __forceinline__ bool fatal(const unsigned* flags) {
  return (*flags & kFlagFatal) != 0u;
}

template <typename T>
__device__ __forceinline__ T wave_min(T v) {
  for (int o = kWave / 2; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, kWave));
  return v;
}
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
  for (int o = kWave / 2; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, kWave));
  return v;
}

// Inclusive prefix sum over the wavefront.
__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = threadIdx.x & (kWave - 1);
  for (int o = 1; o < kWave; o <<= 1) {
    const int u = __shfl_up(v, o, kWave);
    if (lane >= o) v += u;
  }
  return v;
}

// Exclusive prefix sum over the workgroup; *total receives the sum.
// `lds` needs blockDim.x / 64 ints.
__device__ __forceinline__ int block_excl_scan(int v, int* lds, int* total) {
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  const int nw = blockDim.x / kWave;
  const int inc = wave_incl_scan(v);
  if (lane == kWave - 1) lds[w] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int i = 0; i < nw; ++i) {
      const int t = lds[i];
      lds[i] = run;
      run += t;
    }
    lds[nw] = run;
  }
  __syncthreads();
  const int r = lds[w] + inc - v;
  *total = lds[nw];
  __syncthreads();
  return r;
}

// ---------------------------------------------------------------- wavefront

// Lane-to-lane broadcasts.  They must run on the whole wavefront: the
// value of lane j is only there if lane j executed the code producing it,
// so no broadcast sits under a lane-dependent branch.
__device__ __forceinline__ int bcast(int v, int j) { return __builtin_amdgcn_readlane(v, j); }
__device__ __forceinline__ float bcast(float v, int j) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}
__device__ __forceinline__ float3 bcast(const float3& v, int j) {
  return make_float3(bcast(v.x, j), bcast(v.y, j), bcast(v.z, j));
}
__device__ __forceinline__ float4 bcast(const float4& v, int j) {
  return make_float4(bcast(v.x, j), bcast(v.y, j), bcast(v.z, j), bcast(v.w, j));
}

// Bit differences of float values: zero exactly when they are equal bit for
// bit (-0 and +0 differ, a NaN equals itself).  Comparing a record by OR-ing
// its fields' differences and testing once keeps the test free of branches
// (a chain of && over per-lane values can compile to one divergent branch
// per term).
__device__ __forceinline__ unsigned bit_diff(float a, float b) {
  return __float_as_uint(a) ^ __float_as_uint(b);
}

// Bitwise equality of float values (distinguishes -0 and +0, NaN equals itself).
__device__ __forceinline__ bool same_bits(const float3& a, const float3& b) {
  return __float_as_int(a.x) == __float_as_int(b.x) && __float_as_int(a.y) == __float_as_int(b.y) &&
         __float_as_int(a.z) == __float_as_int(b.z);
}
__device__ __forceinline__ bool same_bits(const float4& a, const float4& b) {
  return same_bits(make_float3(a.x, a.y, a.z), make_float3(b.x, b.y, b.z)) &&
         __float_as_int(a.w) == __float_as_int(b.w);
}
__device__ __forceinline__ bool same_bits(float a, float b) {
  return __float_as_int(a) == __float_as_int(b);
}

// ---------------------------------------------------------------- transfers

__global__ void k_upload(int n, const float* __restrict__ pos, const float* __restrict__ vel,
                         float4* __restrict__ posA, float4* __restrict__ velA) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t N = static_cast<size_t>(n);
  posA[i] = make_float4(pos[i], pos[N + i], pos[2 * N + i], __int_as_float(i));
  velA[i] = make_float4(vel[i], vel[N + i], vel[2 * N + i], 0.0f);
}

__global__ void k_download(int n, const float4* __restrict__ posA, const float4* __restrict__ velA,
                           float* __restrict__ pos, float* __restrict__ vel) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const float4 p = posA[s];
  const float4 v = velA[s];
  const size_t id = static_cast<size_t>(__float_as_int(p.w));
  const size_t N = static_cast<size_t>(n);
  pos[id] = p.x;
  pos[N + id] = p.y;
  pos[2 * N + id] = p.z;
  vel[id] = v.x;
  vel[N + id] = v.y;
  vel[2 * N + id] = v.z;
}

__global__ void k_reset_flags(unsigned* flags, int* steps_started) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    *flags = 0u;
    *steps_started = 0;
  }
}

__global__ void k_reset_total(unsigned long long* total) {
  if (threadIdx.x == 0) *total = 0ull;
}

__global__ void k_init_bbox(int n_scenes, int* bbox) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_scenes) return;
  bbox[6 * s + 0] = INT_MAX;
  bbox[6 * s + 1] = INT_MAX;
  bbox[6 * s + 2] = INT_MAX;
  bbox[6 * s + 3] = INT_MIN;
  bbox[6 * s + 4] = INT_MIN;
  bbox[6 * s + 5] = INT_MIN;
}

// ---------------------------------------------------------------- grid build

// Calls fn(leader, mine) once per distinct key among the lanes with
// want == true: `leader` is the lowest such lane holding the key and `mine`
// marks the lanes that share it.  Wavefront-uniform control flow; every
// lane must be present.
template <typename Fn>
__device__ __forceinline__ void for_each_key_group(bool want, int key, Fn&& fn) {
  bool pend = want;
  while (true) {
    const unsigned long long m = __ballot(pend);
    if (m == 0ull) break;
    const int leader = __builtin_ctzll(m);
    const int kl = __builtin_amdgcn_readlane(key, leader);
    const bool mine = pend && key == kl;
    fn(leader, mine);
    pend = pend && !mine;
  }
}

__global__ __launch_bounds__(1024) void k_predict_bbox(
    StepConsts c, const float4* __restrict__ pos, const float4* __restrict__ vel,
    const ScenePlanes* __restrict__ sp, const int* __restrict__ seg,
    uint64_t* __restrict__ cellc, int* __restrict__ bbox, int* __restrict__ tile_bbox,
    unsigned* __restrict__ flags) {
  __shared__ int red[6][1024 / kWave];
  const int tile = tile_index(c.xcd_remap);
  const int base = tile * blockDim.x;
  const int i = base + threadIdx.x;
  const bool active = i < c.n;
  int mn0 = INT_MAX, mn1 = INT_MAX, mn2 = INT_MAX;
  int mx0 = INT_MIN, mx1 = INT_MIN, mx2 = INT_MIN;
  bool bad = false;
  int sc = 0;
  if (active) {
    sc = scene_of(i, seg, c.n_scenes);
    const ScenePlanes& s = sp[sc];
    const float3 x = ld3(pos, i);
    const float3 v = ld3(vel, i);
    const float vx = v.x + s.gx * c.dt;  // core.cpp:155-157
    const float vy = v.y + s.gy * c.dt;
    const float vz = v.z + s.gz * c.dt;
    const float px = x.x + vx * c.dt;    // core.cpp:158-160
    const float py = x.y + vy * c.dt;
    const float pz = x.z + vz * c.dt;
    const int cx = floor_cell(px, c.inv_h, bad);
    const int cy = floor_cell(py, c.inv_h, bad);
    const int cz = floor_cell(pz, c.inv_h, bad);
    cellc[i] = pack_cell(cx, cy, cz);
    mn0 = mx0 = cx;
    mn1 = mx1 = cy;
    mn2 = mx2 = cz;
  }
  if (bad) atomicOr(flags, kFlagRange);
  const int last = min(base + static_cast<int>(blockDim.x), c.n) - 1;
  if (base > last) return;  // whole block past the end (uniform)
  const int s_first = scene_of(base, seg, c.n_scenes);
  const int s_last = scene_of(last, seg, c.n_scenes);
  if (s_first == s_last) {
    // one scene: reduce the workgroup, then one set of atomics
    mn0 = wave_min(mn0), mn1 = wave_min(mn1), mn2 = wave_min(mn2);
    mx0 = wave_max(mx0), mx1 = wave_max(mx1), mx2 = wave_max(mx2);
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    if (lane == 0) {
      red[0][w] = mn0, red[1][w] = mn1, red[2][w] = mn2;
      red[3][w] = mx0, red[4][w] = mx1, red[5][w] = mx2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const int nw = blockDim.x / kWave;
      for (int k = 1; k < nw; ++k) {
        mn0 = min(mn0, red[0][k]), mn1 = min(mn1, red[1][k]), mn2 = min(mn2, red[2][k]);
        mx0 = max(mx0, red[3][k]), mx1 = max(mx1, red[4][k]), mx2 = max(mx2, red[5][k]);
      }
      // one plain store per tile; grid_setup reduces the tiles of a scene
      int* b = tile_bbox + 6 * tile;
      b[0] = mn0, b[1] = mn1, b[2] = mn2, b[3] = mx0, b[4] = mx1, b[5] = mx2;
    }
  } else {
    // a tile that straddles scenes: each wavefront reduces the lanes of each
    // of its scenes and takes one set of atomics per scene (in a batch of
    // small scenes many tiles straddle, and per-particle atomics on one
    // scene's six words serialize)
    const int lane = threadIdx.x & (kWave - 1);
    for_each_key_group(active, sc, [&](int leader, bool mine) {
      const int a0 = wave_min(mine ? mn0 : INT_MAX), a1 = wave_min(mine ? mn1 : INT_MAX),
                a2 = wave_min(mine ? mn2 : INT_MAX);
      const int b0 = wave_max(mine ? mx0 : INT_MIN), b1 = wave_max(mine ? mx1 : INT_MIN),
                b2 = wave_max(mine ? mx2 : INT_MIN);
      if (lane == leader) {
        int* b = bbox + 6 * sc;
        atomicMin(b + 0, a0), atomicMin(b + 1, a1), atomicMin(b + 2, a2);
        atomicMax(b + 3, b0), atomicMax(b + 4, b1), atomicMax(b + 5, b2);
      }
    });
    if (threadIdx.x == 0) {
      int* b = tile_bbox + 6 * tile;
      b[0] = b[1] = b[2] = INT_MAX;
      b[3] = b[4] = b[5] = INT_MIN;
    }
  }
}

// One workgroup of kGridSetupThreads: per-scene table extents and key
// offsets; resets the per-step counters.
constexpr int kGridSetupThreads = 1024;
__global__ __launch_bounds__(kGridSetupThreads) void k_grid_setup(int n_scenes, const int* __restrict__ seg,
                                                    const int* __restrict__ tile_bbox,
                                                    int tile_size, int* __restrict__ bbox,
                                                    GridDesc* __restrict__ grid,
                                                    int* __restrict__ total_cells,
                                                    long long cell_cap, int bshift,
                                                    unsigned* __restrict__ flags,
                                                    unsigned long long* __restrict__ nbr_total,
                                                    unsigned long long* __restrict__ need_cells,
                                                    int* __restrict__ steps_started,
                                                    int* __restrict__ tile_count) {
  __shared__ long long sh[kGridSetupThreads];
  __shared__ long long carry;
  // After a failed substep the counters keep what the host needs to size
  // the retry; only the bounding boxes are reset.
  const bool was_fatal = (*flags & kFlagFatal) != 0u;
  if (threadIdx.x == 0) carry = 0;
  // Fold each scene's tile partials into its box (which already holds the
  // atomics of tiles that straddle two scenes): `per` wavefronts per scene,
  // each folding every per-th stretch of 64 tiles.
  {
    const int lane = threadIdx.x & (kWave - 1);
    const int nw = static_cast<int>(blockDim.x) / kWave;
    const int w = static_cast<int>(threadIdx.x) / kWave;
    const int per = n_scenes >= nw ? 1 : nw / n_scenes;
    for (int sc = w / per; sc < n_scenes; sc += nw / per) {
      const int a = seg[sc], b = seg[sc + 1];
      int v[6] = {INT_MAX, INT_MAX, INT_MAX, INT_MIN, INT_MIN, INT_MIN};
      if (b > a) {
        for (int t = a / tile_size + (w % per) * kWave + lane; t <= (b - 1) / tile_size;
             t += per * kWave) {
          const int* p = tile_bbox + 6 * t;
          v[0] = min(v[0], p[0]), v[1] = min(v[1], p[1]), v[2] = min(v[2], p[2]);
          v[3] = max(v[3], p[3]), v[4] = max(v[4], p[4]), v[5] = max(v[5], p[5]);
        }
      }
      for (int k = 0; k < 3; ++k) v[k] = wave_min(v[k]);
      for (int k = 3; k < 6; ++k) v[k] = wave_max(v[k]);
      if (lane == 0) {
        int* bb = bbox + 6 * sc;
        for (int k = 0; k < 3; ++k) atomicMin(bb + k, v[k]);
        for (int k = 3; k < 6; ++k) atomicMax(bb + k, v[k]);
      }
    }
  }
  __syncthreads();
  for (int b0 = 0; b0 < n_scenes; b0 += blockDim.x) {
    const int s = b0 + threadIdx.x;
    long long nc = 0;
    GridDesc g{};
    if (s < n_scenes) {
      int* bb = bbox + 6 * s;
      const long long lx = bb[0], ly = bb[1], lz = bb[2];
      const long long hx = bb[3], hy = bb[4], hz = bb[5];
      if (lx <= hx && ly <= hy && lz <= hz) {
        const long long sx = hx - lx + 3, sy = hy - ly + 3, sz = hz - lz + 3;
        const long long e = 1ll << bshift;
        const long long bx = (sx + e - 1) / e, by = (sy + e - 1) / e, bz = (sz + e - 1) / e;
        nc = (bx * by * bz) << (3 * bshift);
        g.ox = static_cast<int>(lx - 1);
        g.oy = static_cast<int>(ly - 1);
        g.oz = static_cast<int>(lz - 1);
        g.nbx = static_cast<int>(bx);
        g.nby = static_cast<int>(by);
        g.nbz = static_cast<int>(bz);
      }
      g.bshift = bshift;
      bb[0] = INT_MAX, bb[1] = INT_MAX, bb[2] = INT_MAX;
      bb[3] = INT_MIN, bb[4] = INT_MIN, bb[5] = INT_MIN;
    }
    // exclusive scan of nc over the workgroup
    sh[threadIdx.x] = nc;
    __syncthreads();
    for (int o = 1; o < static_cast<int>(blockDim.x); o <<= 1) {
      const long long t = (threadIdx.x >= static_cast<unsigned>(o)) ? sh[threadIdx.x - o] : 0;
      __syncthreads();
      sh[threadIdx.x] += t;
      __syncthreads();
    }
    const long long excl = sh[threadIdx.x] - nc + carry;
    if (s < n_scenes) {
      g.cell_off = static_cast<int>(excl < cell_cap ? excl : 0);
      g.ncells = static_cast<int>(nc < cell_cap ? nc : 0);
      grid[s] = g;
    }
    __syncthreads();
    if (threadIdx.x == blockDim.x - 1) carry += sh[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0 && !was_fatal) {
    *nbr_total = 0ull;
    if (tile_count) {
      tile_count[0] = 0;  // tiles
      tile_count[1] = 0;  // candidate bricks
    }
    if (carry > cell_cap) {
      atomicOr(flags, kFlagCells);
      *total_cells = 0;
      *need_cells = static_cast<unsigned long long>(carry);
    } else {
      *total_cells = static_cast<int>(carry);
      atomicAdd(steps_started, 1);
    }
  }
}


// Distinct predicted positions p^0 among the lanes of a walk up to which a
// tile's neighbor sets are decided with one ballot per position, and
// distinct p^0 among a tile's candidates up to which they are decided with
// one test per position and lane; beyond both each lane tests each
// candidate.  Measured on the piled-up million-particle dam break
// (substeps/s): both at 3: 35.7, 4: 39.0, 5: 40.5, 6: 40.9, 8: 40.8, 12: 39.6;
// lanes 4 / 8 with candidates at 6: 40.8 / 40.3; candidates 4 / 8 / 10 / 12
// with lanes at 6: 39.0 / 41.3 / 39.8 / 41.2.
constexpr int kP0Classes = 6;
constexpr int kCandPoints = 8;

// Distinct positions of a tile's candidates up to which a pass walks the
// tile as runs at one point (the pair geometry hoisted once per run);
// other tiles are walked as runs of one record.  Measured on the piled-up
// million-particle dam break (substeps/s): 1 point 32.9, 2 points 36.3,
// 3 points 36.0, 4 points 35.7.
constexpr int kTilePoints = 2;


// Each particle's cell key, in the previous substep's slot order.
__global__ __launch_bounds__(1024) void k_cell_keys(StepConsts c,
                                                    const uint64_t* __restrict__ cellc,
                                                    const int* __restrict__ seg,
                                                    const GridDesc* __restrict__ grid,
                                                    int* __restrict__ key,
                                                    const unsigned* __restrict__ flags) {
  const int i = tile_index(c.xcd_remap) * blockDim.x + threadIdx.x;
  if (fatal(flags) || i >= c.n) return;
  const GridDesc g = grid[scene_of(i, seg, c.n_scenes)];
  int x, y, z;
  unpack_cell(cellc[i], x, y, z);
  key[i] = cell_key(x - g.ox, y - g.oy, z - g.oz, g);
}

// The cell histogram from the sorted keys, where each cell's particles are
// one run.  A run inside one wavefront stores its length; a run that crosses
// a wavefront boundary adds each wavefront's part atomically.  A substep
// thus takes about one atomic per wavefront instead of one per cell and
// wavefront (device-scope atomics are carried out in memory on this
// multi-XCD part).  The histogram is all zeros on entry: the scan that reads
// it clears it.
__global__ __launch_bounds__(1024) void k_cell_runs(int n, const int* __restrict__ sk,
                                                    int* __restrict__ count,
                                                    const unsigned* __restrict__ flags) {
  if (fatal(flags)) return;  // uniform
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & (kWave - 1);
  const bool active = s < n;
  const int k = active ? sk[s] : 0;
  const bool head = active && (s == 0 || sk[s - 1] != k);
  const bool tail = active && (s == n - 1 || sk[s + 1] != k);
  // segments: a run's slots within one wavefront; active lanes are a prefix
  const unsigned long long starts = __ballot(active && (head || lane == 0));
  const unsigned long long tails = __ballot(tail);
  const int nact = __popcll(__ballot(active));
  if (!active || ((starts >> lane) & 1ull) == 0ull) return;
  const unsigned long long above = starts & ~((2ull << lane) - 1ull);
  const int next = above != 0ull ? __builtin_ctzll(above) : nact;
  const int len = next - lane;
  if (head && ((tails >> (next - 1)) & 1ull) != 0ull)
    count[k] = len;  // the whole run: its only writer
  else
    atomicAdd(count + k, len);
}

// ---- exclusive scan over a device-sized array (three launches, fixed grid)

// Exclusive scans of one array, or of two arrays of one length at once (in2
// and out2 non-null), in two launches: k_scan_reduce writes each workgroup's
// range sums to partial ([0, G) for in, [G, 2G) for in2), k_scan_down scans.
__global__ __launch_bounds__(256) void k_scan_reduce(const int* __restrict__ in,
                                                     const int* __restrict__ in2,
                                                     const int* __restrict__ d_len,
                                                     int* __restrict__ partial,
                                                     const unsigned* __restrict__ flags) {
  __shared__ int lds[8];
  if (fatal(flags)) return;
  const int len = *d_len;
  const int G = gridDim.x;
  const int chunk = (len + G - 1) / G;
  const int b0 = blockIdx.x * chunk;
  const int b1 = min(len, b0 + chunk);
  int acc = 0, acc2 = 0;
  for (int i = b0 + threadIdx.x; i < b1; i += blockDim.x) {
    acc += in[i];
    if (in2) acc2 += in2[i];
  }
  int total;
  block_excl_scan(acc, lds, &total);
  if (threadIdx.x == 0) partial[blockIdx.x] = total;
  if (in2) {
    block_excl_scan(acc2, lds, &total);
    if (threadIdx.x == 0) partial[G + blockIdx.x] = total;
  }
}

// Sum of the partials of the workgroups before this one (*pre) and of all
// of them (*grand).
__device__ __forceinline__ void partial_prefix(const int* __restrict__ partial, int G, int* lds,
                                               int* pre, int* grand) {
  int p = 0, a = 0;
  for (int j = threadIdx.x; j < G; j += blockDim.x) {
    const int v = partial[j];
    a += v;
    p += j < static_cast<int>(blockIdx.x) ? v : 0;
  }
  block_excl_scan(p, lds, pre);
  block_excl_scan(a, lds, grand);
}

// out[i] = exclusive prefix; out[len] = grand total (and so for in2/out2).
// zero_in clears `in` (the cell histogram is reused by the next substep).
// Each workgroup first sums the partials of the workgroups before it, then
// scans its own range, each thread taking four consecutive elements per
// block scan.
__global__ __launch_bounds__(256) void k_scan_down(int* __restrict__ in,
                                                   const int* __restrict__ in2,
                                                   const int* __restrict__ d_len,
                                                   const int* __restrict__ partial,
                                                   int* __restrict__ out, int* __restrict__ out2,
                                                   int zero_in,
                                                   const unsigned* __restrict__ flags) {
  __shared__ int lds[8];
  if (fatal(flags)) return;
  const int len = *d_len;
  const int G = gridDim.x;
  const int chunk = (len + G - 1) / G;
  const int b0 = blockIdx.x * chunk;
  const int b1 = min(len, b0 + chunk);
  int run, grand, run2 = 0, grand2 = 0;
  partial_prefix(partial, G, lds, &run, &grand);
  if (in2) partial_prefix(partial + G, G, lds, &run2, &grand2);
  constexpr int kPer = 4;
  for (int t0 = b0; t0 < b1; t0 += kPer * static_cast<int>(blockDim.x)) {
    const int i0 = t0 + kPer * static_cast<int>(threadIdx.x);
    int v[kPer], w[kPer];
    int sum = 0, sum2 = 0;
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      v[q] = (i0 + q < b1) ? in[i0 + q] : 0;
      w[q] = (in2 && i0 + q < b1) ? in2[i0 + q] : 0;
      sum += v[q];
      sum2 += w[q];
    }
    int total, total2 = 0;
    int e = run + block_excl_scan(sum, lds, &total);
    int e2 = in2 ? run2 + block_excl_scan(sum2, lds, &total2) : 0;
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      if (i0 + q < b1) {
        out[i0 + q] = e;
        if (zero_in) in[i0 + q] = 0;
        if (in2) out2[i0 + q] = e2;
      }
      e += v[q];
      e2 += w[q];
    }
    run += total;
    run2 += total2;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    out[len] = grand;
    if (in2) out2[len] = grand2;
  }
}

// ---- stable radix sort of the slots by (cell key, original id)
//
// The keys are written in original-id order, so a stable sort by key alone
// leaves the particles of each cell in ascending id: the reference's order.
// Each pass sorts 8 bits: a per-workgroup digit histogram, the device scan
// of the histogram in digit-major order, and a stable scatter.

// Keys and values in original-id order: k[id] = cell key, v[id] = slot.
__global__ __launch_bounds__(1024) void k_sort_keys(StepConsts c, const float4* __restrict__ pos,
                                                    const int* __restrict__ key,
                                                    int* __restrict__ kout, int* __restrict__ vout,
                                                    unsigned* __restrict__ flags) {
  if (fatal(flags)) return;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= c.n) return;
  const int id = __float_as_int(reinterpret_cast<const float*>(pos + p)[3]);
  if (id < 0 || id >= c.n) {  // ids are 0..n-1; refuse rather than write outside
    atomicOr(flags, kFlagRange);
    return;
  }
  kout[id] = key[p];
  vout[id] = p;
}

// Every key lies below the substep's table size, so a pass whose digits
// start at or above its bit length has only digit 0: the stable pass keeps
// the order, its histogram is not needed and its scatter is a copy.
__device__ __forceinline__ bool pass_is_identity(const int* total_cells, int shift) {
  return static_cast<long long>(*total_cells) <= (1ll << shift);
}

// hist[d * G + b]: items of workgroup b (kRadixTile of them) with digit d.
__global__ __launch_bounds__(256) void k_radix_hist(int n, const int* __restrict__ kin, int shift,
                                                    int* __restrict__ hist,
                                                    const int* __restrict__ total_cells,
                                                    const unsigned* __restrict__ flags) {
  __shared__ int h[256];
  if (fatal(flags) || pass_is_identity(total_cells, shift)) return;
  h[threadIdx.x] = 0;
  __syncthreads();
  const int b0 = blockIdx.x * kRadixTile;
  const int b1 = min(n, b0 + kRadixTile);
  for (int i = b0 + static_cast<int>(threadIdx.x); i < b1; i += 256)
    atomicAdd(&h[(kin[i] >> shift) & 255], 1);
  __syncthreads();
  hist[threadIdx.x * gridDim.x + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of one pass: the workgroup's items in index order, 256 per
// round; an item goes to its digit's offset for this workgroup, plus the
// items of that digit in earlier rounds, earlier wavefronts and lower lanes.
__global__ __launch_bounds__(256) void k_radix_scatter(int n, const int* __restrict__ kin,
                                                       const int* __restrict__ vin,
                                                       int* __restrict__ kout,
                                                       int* __restrict__ vout, int shift,
                                                       const int* __restrict__ off,
                                                       const int* __restrict__ total_cells,
                                                       const unsigned* __restrict__ flags) {
  __shared__ int run[256];      // per digit: where the next item of this workgroup goes
  __shared__ int wcnt[4][256];  // per wavefront of the round: items per digit
  if (fatal(flags)) return;
  const int tid = threadIdx.x;
  if (pass_is_identity(total_cells, shift)) {
    const int b1 = min(n, static_cast<int>(blockIdx.x) * kRadixTile + kRadixTile);
    for (int i = blockIdx.x * kRadixTile + tid; i < b1; i += 256) {
      kout[i] = kin[i];
      vout[i] = vin[i];
    }
    return;
  }
  const int lane = tid & (kWave - 1);
  const int w = tid >> 6;
  run[tid] = off[tid * gridDim.x + blockIdx.x];
  const int b0 = blockIdx.x * kRadixTile;
  for (int r = 0; r < kRadixTile / 256; ++r) {
    const int i = b0 + r * 256 + tid;
    const bool ok = i < n;
    const int kk = ok ? kin[i] : 0;
    const int d = (kk >> shift) & 255;
    // the lanes of this wavefront with the same digit
    unsigned long long peers = __ballot(ok);
    for (int bit = 0; bit < 8; ++bit) {
      const bool x = (d >> bit) & 1;
      const unsigned long long bb = __ballot(x);
      peers &= x ? bb : ~bb;
    }
    if (!ok) peers = 0ull;
    const int below = __popcll(peers & ((1ull << lane) - 1ull));
    for (int v = 0; v < 4; ++v) wcnt[v][tid] = 0;
    __syncthreads();
    if (ok && below == 0) wcnt[w][d] = __popcll(peers);
    __syncthreads();
    if (ok) {
      int pos = run[d] + below;
      for (int v = 0; v < w; ++v) pos += wcnt[v][d];
      if (pos < n) {  // always, as the histogram counted these same items
        kout[pos] = kk;
        vout[pos] = vin[i];
      }
    }
    __syncthreads();
    run[tid] += wcnt[0][tid] + wcnt[1][tid] + wcnt[2][tid] + wcnt[3][tid];
    // the next round's zeroing of wcnt[.][tid] follows this read in the same thread
  }
}

// Gathers each particle into its slot of the (cell, id) order (sv: the
// sorted previous slots, sk: their cell keys) and recomputes pred there.
__global__ __launch_bounds__(1024) void k_reorder(
    StepConsts c, const int* __restrict__ sk, const int* __restrict__ sv,
    const int* __restrict__ seg, const ScenePlanes* __restrict__ sp,
    const float4* __restrict__ posA, const float4* __restrict__ velA, float4* __restrict__ posB,
    float4* __restrict__ velB, float4* __restrict__ pred, const uint64_t* __restrict__ cellc,
    uint64_t* __restrict__ slot_cell, int* __restrict__ slot_key,
    unsigned* __restrict__ flags) {
  if (fatal(flags)) return;
  const int dest = blockIdx.x * blockDim.x + threadIdx.x;
  if (dest >= c.n) return;
  const int p = sv[dest];
  if (p < 0 || p >= c.n) {  // never for a complete sort; refuse rather than read outside
    atomicOr(flags, kFlagRange);
    return;
  }
  const ScenePlanes& s = sp[scene_of(dest, seg, c.n_scenes)];
  const float4 x = posA[p];
  const float4 v = velA[p];
  const float vx = v.x + s.gx * c.dt;
  const float vy = v.y + s.gy * c.dt;
  const float vz = v.z + s.gz * c.dt;
  posB[dest] = x;
  velB[dest] = make_float4(vx, vy, vz, 0.0f);
  pred[dest] = make_float4(x.x + vx * c.dt, x.y + vy * c.dt, x.z + vz * c.dt, 0.0f);
  slot_cell[dest] = cellc[p];
  slot_key[dest] = sk[dest];
}

// ---------------------------------------------------------------- neighbors

// Neighbor lists are run-length encoded in slot space: a word t >= 0 is the
// single neighbor t; a word with the top bit set starts a run of
// consecutive slots (t & 0x7fffffff) .. + next word - 1.  Iterating the
// words in order visits neighbors in exactly the order they were found
// (27 cells in (dz, dy, dx) order, each by ascending id), so the encoding is
// invisible to the arithmetic.  Dense cells — where particles pile up on
// walls and corners — produce long runs and shrink by orders of magnitude.
struct RunWriter {
  int* out;
  long long words = 0;
  int count = 0;
  int run_start = -1;
  int run_len = 0;
  template <bool kFill>
  __device__ __forceinline__ void flush() {
    if (run_len == 1) {
      if (kFill) out[words] = run_start;
      words += 1;
    } else if (run_len > 1) {
      if (kFill) {
        out[words] = run_start | static_cast<int>(0x80000000u);
        out[words + 1] = run_len;
      }
      words += 2;
    }
  }
  template <bool kFill>
  __device__ __forceinline__ void add(int t) {
    ++count;
    if (run_len > 0 && t == run_start + run_len) {
      ++run_len;
      return;
    }
    flush<kFill>();
    run_start = t;
    run_len = 1;
  }
  // add(a), add(a+1), ..., add(b-1) in O(1)
  template <bool kFill>
  __device__ __forceinline__ void add_range(int a, int b) {
    if (a >= b) return;
    count += b - a;
    if (run_len > 0 && a == run_start + run_len) {
      run_len += b - a;
      return;
    }
    flush<kFill>();
    run_start = a;
    run_len = b - a;
  }
};


// Candidates of a particle: the 27 cells around its own in (dz, dy, dx)
// order, each cell's slots in ascending order, i.e. ascending original id.
template <bool kFill>
__device__ __forceinline__ void walk_cells(int s, float px, float py, float pz, int rx, int ry,
                                          int rz, const GridDesc& g, const StepConsts& c,
                                          const int* __restrict__ start,
                                          const float4* __restrict__ pred, RunWriter& w) {
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        const int k = cell_key(rx + dx, ry + dy, rz + dz, g);
        const int a = start[k];
        const int b = start[k + 1];
        const float3 me = make_float3(px, py, pz);
        int t = a;
        // four loads in flight before any test; tests and appends stay in order
        for (; t + 4 <= b; t += 4) {
          const float3 q0 = ld3(pred, t), q1 = ld3(pred, t + 1);
          const float3 q2 = ld3(pred, t + 2), q3 = ld3(pred, t + 3);
          if (t != s && near0(me, q0, c)) w.add<kFill>(t);
          if (t + 1 != s && near0(me, q1, c)) w.add<kFill>(t + 1);
          if (t + 2 != s && near0(me, q2, c)) w.add<kFill>(t + 2);
          if (t + 3 != s && near0(me, q3, c)) w.add<kFill>(t + 3);
        }
        for (; t < b; ++t)
          if (t != s && near0(me, ld3(pred, t), c)) w.add<kFill>(t);
      }
  w.flush<kFill>();
}

// Per-slot cell data the cooperative walks need.
struct CellView {
  const uint64_t* __restrict__ slot_cell;  // packed absolute cell
  const int* __restrict__ slot_key;        // table key of that cell
  const GridDesc* __restrict__ grid;
  const int* __restrict__ start;
  const int* __restrict__ seg;
  const int* __restrict__ lane_rep;        // with runs: each lane's run head slot (-1: none)
  const int* __restrict__ lane_info;       // with runs: packed lanes, aligned lanes' range
  const int* __restrict__ run_end;         // with runs: per slot, one past its run's end
  const float4* __restrict__ p0;           // predicted positions p^0 (xyz) of the lists
  const int* __restrict__ in_tile;         // without runs: slots the tile kernels compute
};

// One lane's view: its slot (the last particle for lanes past the end, so
// that every wavefront stays whole), cell and whether it walks
// cooperatively ("dense").  With runs, a lane computes the first slot of
// one run (the run's other slots copy its outputs): each dense (or heavy)
// cell's runs start at a wave boundary, from lane 0, so that a wave walks
// one such cell only and those long waves are dispatched first; the runs of
// the other cells are packed after them; padding lanes are inactive.
struct LaneCell {
  int s;
  bool active;
  bool dense;
  int key;
  uint64_t cell;
};

// `heavy` adds lanes of sparse cells whose own walk is long (next to a
// dense cell): they are grouped by their cell too.  A particle without a
// stored list (nwords < 0) always walks cooperatively.
__device__ __forceinline__ LaneCell lane_cell(const StepConsts& c, const CellView& cv, int s0,
                                              const int* __restrict__ ncount = nullptr,
                                              const int* __restrict__ nwords = nullptr) {
  LaneCell L;
  if (c.runs) {
    const int packed = cv.lane_info[0];
    int slot = -1;
    if (s0 < packed || (s0 >= cv.lane_info[1] && s0 < cv.lane_info[2])) slot = cv.lane_rep[s0];
    L.active = slot >= 0;
    L.s = L.active ? slot : c.n - 1;
  } else {
    L.active = s0 < c.n && !(cv.in_tile && cv.in_tile[s0]);
    L.s = s0 < c.n ? s0 : c.n - 1;
  }
  L.key = cv.slot_key[L.s];
  L.cell = cv.slot_cell[L.s];
  const bool heavy = (ncount != nullptr && ncount[L.s] >= kHeavyList) ||
                     (nwords != nullptr && nwords[L.s] < 0);
  L.dense = c.coop && L.active &&
            (heavy || (cv.start[L.key + 1] - cv.start[L.key]) >= kDenseCell);
  return L;
}

// The lane of this thread in the lane map.  The map puts its longest
// wavefronts first so that they are dispatched first; workgroups take it in
// dispatch order (the XCD-aware order, which spreads consecutive workgroups
// over the map, would start some of them last: the 256-scene ensemble ran
// 9.3 instead of 5.5 ms per substep with it).
__device__ __forceinline__ int map_lane() {
  return static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
}

// With runs the passes launch over the lane map's capacity; workgroups past
// the lanes in use ([0, max(lane_info[0], lane_info[2]))) have nothing to do
// and return at once (uniform over the workgroup).
__device__ __forceinline__ bool past_lanes(const StepConsts& c, const CellView& cv) {
  if (!c.runs) return false;
  const int first = static_cast<int>(blockIdx.x * blockDim.x);
  return first >= max(cv.lane_info[0], cv.lane_info[2]);
}

// Wavefront-uniform description of the group led by `leader`.
struct Group {
  int rx, ry, rz;
  GridDesc g;
};

__device__ __forceinline__ Group group_of(const StepConsts& c, const CellView& cv,
                                          const LaneCell& L, int leader) {
  Group gr;
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(L.cell)), leader);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(L.cell >> 32)),
                                           leader);
  int sc = 0;
  if (c.n_scenes > 1) {
    const int sl = __builtin_amdgcn_readlane(L.s, leader);
    sc = __builtin_amdgcn_readfirstlane(scene_of(sl, cv.seg, c.n_scenes));
  }
  gr.g = cv.grid[sc];
  int x, y, z;
  unpack_cell((static_cast<uint64_t>(static_cast<uint32_t>(hi)) << 32) | static_cast<uint32_t>(lo),
              x, y, z);
  gr.rx = x - gr.g.ox;
  gr.ry = y - gr.g.oy;
  gr.rz = z - gr.g.oz;
  return gr;
}


// Runs of identical slots at least this long are walked as one step.
constexpr int kRunStep = 32;

// Walks the group's 27-cell neighborhood in (dz, dy, dx) order, each cell's
// slots in ascending order.  Each step is body(base, m, tile, whole): with
// run_end (runs on) a run of at least kRunStep slots is one step (whole,
// m = its length, lane 0 of `tile` holding slot base: every slot of the run
// holds the same record); otherwise a tile of up to 64 candidates, `tile`
// holding load(candidate) of lane `lane` (lanes past m repeat the last).
// The next step's load is issued before a step is processed; the 27 cells'
// slot ranges (and the run end of each one's first slot) are loaded up
// front, one cell per lane, instead of one dependent load chain per cell.
template <typename Load, typename Body>
__device__ __forceinline__ void group_steps(const Group& gr, const int* __restrict__ start,
                                            const int* __restrict__ run_end, Load&& load,
                                            Body&& body) {
  const int lane = threadIdx.x & (kWave - 1);
  int ca = 0, cb = 0, ce = 0;
  if (lane < 27) {  // lane = 9 (dz + 1) + 3 (dy + 1) + (dx + 1)
    const int k = cell_key(gr.rx + lane % 3 - 1, gr.ry + (lane / 3) % 3 - 1,
                           gr.rz + lane / 9 - 1, gr.g);
    ca = start[k];
    cb = start[k + 1];
    if (run_end != nullptr && ca < cb) ce = run_end[ca];
  }
  for (int ci = 0; ci < 27; ++ci) {
    const int a = __builtin_amdgcn_readlane(ca, ci);
    const int b = __builtin_amdgcn_readlane(cb, ci);
    if (a >= b) continue;
    int t = a;
    int et = run_end ? __builtin_amdgcn_readlane(ce, ci) : t;
    auto next = load(t + min(lane, min(kWave, b - t) - 1));
    while (t < b) {
      const int er = min(et, b);
      const bool whole = run_end != nullptr && er - t >= kRunStep;
      const int m = whole ? er - t : min(kWave, b - t);
      const auto cur = next;
      const int tn = t + m;
      int en = tn;
      if (tn < b) {
        if (run_end) en = __builtin_amdgcn_readfirstlane(run_end[tn]);
        next = load(tn + min(lane, min(kWave, b - tn) - 1));
      }
      body(t, m, cur, whole);
      t = tn;
      et = en;
    }
  }
}

__device__ __forceinline__ float3 pos_of(const float3& q) { return q; }
__device__ __forceinline__ float3 pos_of(const float4& q) { return make_float3(q.x, q.y, q.z); }

// Lanes of a walk grouped by their p^0 (bit for bit).  Lanes that share p^0
// share their neighbors, but for themselves.  Class k's p^0 is kept in lane
// k of `pt` (uniform values live across the whole walk would crowd out the
// scalar registers), so a tile reads it from a fixed lane.
struct LaneClasses {
  int cls;    // this lane's class, -1 when none
  int n;      // classes found
  bool all;   // every walking lane has a class
  float3 pt;  // lane k < n: class k's p^0
};

__device__ __forceinline__ LaneClasses lane_classes(const float3& me0, bool in, bool on) {
  const int lane = threadIdx.x & (kWave - 1);
  LaneClasses lc;
  lc.cls = -1;
  lc.n = 0;
  lc.pt = me0;
  bool pend = in;
  if (on) {
#pragma unroll
    for (int k = 0; k < kP0Classes; ++k) {
      const unsigned long long pm = __ballot(pend);
      if (pm == 0ull) break;
      const float3 pk = bcast(me0, __builtin_ctzll(pm));
      const bool eq = pend && same_bits(me0, pk);
      if (eq) lc.cls = k;
      if (lane == k) lc.pt = pk;
      pend = pend && !eq;
      ++lc.n;
    }
  }
  lc.all = __ballot(pend) == 0ull;
  return lc;
}

// The candidates of `rest` (a set of a tile's candidates) from its first
// one, j, up to the next run start in `starts` after j.
__device__ __forceinline__ unsigned long long run_from(unsigned long long rest,
                                                       unsigned long long starts, int j) {
  const unsigned long long later = starts & ~((2ull << j) - 1ull);
  return later != 0ull ? rest & ((1ull << __builtin_ctzll(later)) - 1ull) : rest;
}

// The value of v in the previous lane (lane 0: lane 63's), by a permute
// over the whole wavefront; every lane must be active.
__device__ __forceinline__ float prev_lane(float v) {
  const int src = ((threadIdx.x & (kWave - 1)) + kWave - 1) & (kWave - 1);
  return __int_as_float(__builtin_amdgcn_ds_bpermute(src << 2, __float_as_int(v)));
}
__device__ __forceinline__ float3 prev_lane(const float3& v) {
  return make_float3(prev_lane(v.x), prev_lane(v.y), prev_lane(v.z));
}

// Groups the m candidates of a tile by position (bit for bit), up to K
// positions: cls is the index of the position this lane's candidate sits at
// (-1 for lanes past m).  True when every candidate is at one of them.  A
// position's point is its candidates' own, and its mask a ballot of cls (no
// per-position arrays, which would crowd the scalar registers).
template <int K>
__device__ __forceinline__ bool tile_point_ids(const float3& q, int m, int& cls) {
  const int lane = threadIdx.x & (kWave - 1);
  bool pend = lane < m;
  cls = -1;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const unsigned long long pm = __ballot(pend);
    if (pm == 0ull) break;
    const bool eq = pend && same_bits(q, bcast(q, __builtin_ctzll(pm)));
    if (eq) cls = k;
    pend = pend && !eq;
  }
  return __ballot(pend) == 0ull;
}

// This lane's neighbors among the m candidates of a tile (bit j: candidate
// base + j, whose p^0 this lane holds in t0), by the list predicate near0;
// `any` receives the candidates that some walking lane takes (a superset).
// With few lane classes one ballot per class decides the tile; otherwise,
// with few distinct candidate positions, one test per position and lane;
// otherwise every lane tests every candidate.
__device__ __forceinline__ unsigned long long tile_members(const LaneClasses& lc,
                                                           const float3& me0, bool in,
                                                           const float3& t0, int m, int self,
                                                           bool on, const StepConsts& c,
                                                           unsigned long long& any) {
  const int lane = threadIdx.x & (kWave - 1);
  const unsigned long long full = m == kWave ? ~0ull : ((1ull << m) - 1ull);
  unsigned long long mine = 0ull;
  any = 0ull;
  bool done = false;
  if (on && lc.all) {
#pragma unroll
    for (int k = 0; k < kP0Classes; ++k) {
      if (k < lc.n) {
        const float3 pk = bcast(lc.pt, k);
        const unsigned long long mk = __ballot(lane < m && near0(pk, t0, c));
        any |= mk;
        if (lc.cls == k) mine = mk;
      }
    }
    done = true;
  }
  if (on && !done) {
    // distinct candidate positions of the tile, each tested once per lane as
    // it is found
    bool pend = lane < m;
    unsigned long long mm = 0ull, an = 0ull;
#pragma unroll
    for (int k = 0; k < kCandPoints; ++k) {
      const unsigned long long pm = __ballot(pend);
      if (pm == 0ull) break;
      const float3 pk = bcast(t0, __builtin_ctzll(pm));
      const bool eq = pend && same_bits(t0, pk);
      pend = pend && !eq;
      const unsigned long long ck = __ballot(eq);
      const bool nk = in && near0(me0, pk, c);
      if (nk) mm |= ck;
      if (__ballot(nk) != 0ull) an |= ck;
    }
    if (__ballot(pend) == 0ull) {
      mine = mm;
      any = an;
      done = true;
    }
  }
  if (!done) {
    any = full;
    for (int j = 0; j < m; ++j)
      if (near0(me0, bcast(t0, j), c)) mine |= 1ull << j;
  }
  if (!in) mine = 0ull;
  if (self >= 0 && self < m) mine &= ~(1ull << self);
  return mine;
}

template <bool kFill>
__device__ __forceinline__ void group_walk_cells(const Group& gr, bool mine, int s,
                                                float px, float py, float pz,
                                                const StepConsts& c,
                                                const int* __restrict__ start,
                                                const int* __restrict__ run_end,
                                                const float4* __restrict__ pred, RunWriter& w) {
  const float3 me = make_float3(px, py, pz);
  const bool on = c.coop_bulk != 0;
  const LaneClasses lc = lane_classes(me, mine, on);
  group_steps(gr, start, run_end, [&](int t) { return ld3(pred, t); },
              [&](int base, int m, const float3& tile, bool whole) {
                if (whole) {
                  // a run: every slot at one p^0, all taken or none
                  if (mine && near0(me, bcast(tile, 0), c)) {
                    if (s >= base && s < base + m) {
                      w.add_range<kFill>(base, s);
                      w.add_range<kFill>(s + 1, base + m);
                    } else {
                      w.add_range<kFill>(base, base + m);
                    }
                  }
                  return;
                }
                unsigned long long any;
                unsigned long long mask = tile_members(lc, me, mine, tile, m, s - base, on, c, any);
                // the mask's runs of set bits, in order
                while (mask != 0ull) {
                  const int a = __builtin_ctzll(mask);
                  const unsigned long long rest = ~(mask >> a);
                  const int len = rest != 0ull ? __builtin_ctzll(rest) : kWave - a;
                  w.add_range<kFill>(base + a, base + a + len);
                  mask = (a + len >= kWave) ? 0ull : (mask & (~0ull << (a + len)));
                }
              });
  if (mine) w.flush<kFill>();
}

__global__ __launch_bounds__(1024) void k_neighbors(
    StepConsts c, const float4* __restrict__ pred, CellView cv, long long* __restrict__ noff,
    int* __restrict__ nwords, int* __restrict__ ncount, int* __restrict__ nbr, long long nbr_cap,
    unsigned long long* __restrict__ nbr_total, unsigned* __restrict__ flags) {
  __shared__ int lds[1024 / kWave + 1];
  __shared__ long long base_sh;
  __shared__ int abort_sh;
  // Other workgroups of this launch may raise the neighbor flag at any time:
  // read it once per workgroup so the barriers below stay uniform.
  if (past_lanes(c, cv)) return;
  if (threadIdx.x == 0) abort_sh = fatal(flags) ? 1 : 0;
  __syncthreads();
  if (abort_sh) return;
  const LaneCell L = lane_cell(c, cv, map_lane());
  const int s = L.s;
  bool active = L.active;
  const float3 p = ld3(pred, s);
  const GridDesc g = cv.grid[scene_of(s, cv.seg, c.n_scenes)];
  bool bad = false;
  const int rx = floor_cell(p.x, c.inv_h, bad) - g.ox;
  const int ry = floor_cell(p.y, c.inv_h, bad) - g.oy;
  const int rz = floor_cell(p.z, c.inv_h, bad) - g.oz;
  // The table has a one-cell margin around the bounding box of these same
  // cells; anything else would index outside it, so refuse instead.
  const int ext = 1 << g.bshift;  // cells per brick edge
  if (active && (bad || rx < 1 || ry < 1 || rz < 1 || rx > ext * g.nbx - 2 ||
                 ry > ext * g.nby - 2 || rz > ext * g.nbz - 2)) {
    if (!c.count_only) atomicOr(flags, kFlagRange);
    active = false;
  }
  // lanes of sparse cells next to dense ones: long candidate walks
  bool heavy = false;
  if (c.coop && active && !L.dense) {
    int cand = 0;
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int k = cell_key(rx + dx, ry + dy, rz + dz, g);
          cand += cv.start[k + 1] - cv.start[k];
        }
    heavy = cand >= kHeavyList;
  }
  // every dense group lies inside the table: its cell is some valid lane's
  const bool dense = (L.dense || heavy) && active;
  // Particles of dense cells store no list unless full lists are asked for:
  // the passes always walk them cooperatively and retest near0 on p^0 tile
  // by tile, in the same order.  Particles of sparse cells keep theirs: a
  // heavy one (next to a dense cell) may still have a short list, which the
  // passes then walk on their own.
  const bool listed = active && (!L.dense || c.full_lists || c.count_only);
  RunWriter w{nullptr};
  if (c.coop)
    for_each_key_group(dense && listed, L.key, [&](int leader, bool mine) {
      group_walk_cells<false>(group_of(c, cv, L, leader), mine, s, p.x, p.y, p.z, c, cv.start,
                              cv.run_end, pred, w);
    });
  if (active && !dense) walk_cells<false>(s, p.x, p.y, p.z, rx, ry, rz, g, c, cv.start, pred, w);
  if (c.count_only) {
    // statistics: add this wavefront's pairs to *nbr_total and keep each
    // particle's count; no list is stored
    unsigned long long v = active ? static_cast<unsigned long long>(w.count) : 0ull;
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & (kWave - 1)) == 0 && v != 0ull) atomicAdd(nbr_total, v);
    if (L.active) ncount[s] = active ? w.count : 0;
    return;
  }
  // The encoded length of one list stays below 2^31 words (at most two words
  // per neighbor and fewer than 2^29 particles).
  const int nw = listed ? static_cast<int>(w.words) : 0;
  int total;
  const int pre = block_excl_scan(nw, lds, &total);
  if (threadIdx.x == 0)
    base_sh = static_cast<long long>(atomicAdd(nbr_total, static_cast<unsigned long long>(total)));
  __syncthreads();
  const long long off = base_sh + pre;
  const bool fits = off + nw <= nbr_cap;
  if (listed && !fits) atomicOr(flags, kFlagNeighbors);
  const bool ok = listed && fits;
  RunWriter f{nbr + (ok ? off : 0)};
  if (c.coop)
    for_each_key_group(dense && ok, L.key, [&](int leader, bool mine) {
      group_walk_cells<true>(group_of(c, cv, L, leader), mine, s, p.x, p.y, p.z, c, cv.start,
                             cv.run_end, pred, f);
    });
  if (ok && !dense) walk_cells<true>(s, p.x, p.y, p.z, rx, ry, rz, g, c, cv.start, pred, f);
  if (L.active) {
    // nwords = -1: no stored list, the passes walk this particle cooperatively
    const bool unlisted = active && !listed;
    noff[s] = ok ? off : 0;
    nwords[s] = ok ? nw : (unlisted ? -1 : 0);
    ncount[s] = ok ? f.count : 0;
  }
}

// The pair arithmetic of one gather pass as five operations:
//   same(leader)  this lane's pair state (what its terms depend on besides
//                 the candidate) equals the leader lane's, bit for bit
//   bind(leader)  take the leader's state as the group's
//   gterm(q)      the term of candidate q against the group's state
//   add(t)        accumulate one term (sums run in list order)
//   pair(q)       add(the term of q against this lane's own state)
// same and bind broadcast, so they run on the whole wavefront.
template <typename Same, typename Bind, typename GTerm, typename Add, typename Pair>
struct PairOps {
  Same same;
  Bind bind;
  GTerm gterm;
  Add add;
  Pair pair;
};
template <typename Same, typename Bind, typename GTerm, typename Add, typename Pair>
__device__ __forceinline__ PairOps<Same, Bind, GTerm, Add, Pair> pair_ops(Same s, Bind b, GTerm g,
                                                                          Add a, Pair p) {
  return {s, b, g, a, p};
}


// Lanes of one group that share their pair state bit for bit form a class;
// a group with at most this many classes walks its tiles once per class
// with terms computed once per candidate, otherwise once with every lane
// evaluating every candidate.  Walks of one wavefront run one after the
// other, and holding several classes' terms at once costs registers (fewer
// resident wavefronts): on the piled-up million-particle dam break one
// class is fastest (4.26 substeps/s; 4 class walks 3.38; 4 classes in one
// walk 3.64; no class walks 3.99; measured again with the current walk:
// 2 class walks 27.9 against 41.3 for one).
constexpr int kMaxClasses = 1;

// A candidate tile with the predicted positions p^0 of the same slots.
template <typename T>
struct TileP0 {
  T q;
  float3 p0;
};

// Gathers over the neighbors of every lane of a wavefront, in each lane's
// list order:
//  * lanes in dense cells: cooperative walks of each distinct dense cell of
//    the wavefront; candidate tiles are loaded once (coalesced) and handed
//    lane to lane with readlane.  A lane's neighbors in a tile are the
//    candidates t != s with near0(p^0_s, p^0_t), the predicate its list is
//    built with, so the walk visits exactly the list, in its order, without
//    reading it; when the group shares one p^0 a single ballot decides
//    every candidate for all of its lanes;
//  * when the candidates of a tile sit at no more than kTilePoints points,
//    the tile is walked as maximal runs of consecutive candidates at one
//    point cp: for each run prep(cp, flag) hoists the pair geometry once
//    per lane and says how the lane's neighbors in the run contribute:
//    0 one by one through pair_u (or, when the run's candidates hold one
//    record, that one term `count` times exactly through pair_rep),
//    1 nothing at all (exact signed zeros), 2 `count` identical terms
//    (on_count); flag is tflag(tile, m), evaluated by the whole wavefront;
//  * consecutive tiles whose candidates all hold one record (bit for bit,
//    p^0 included) form one range: each lane only counts its members, and
//    the range then contributes like one run at one point, with mode 0
//    adding pair_u's term `count` times exactly (pair_rep);
//  * lanes that share their pair state bit for bit (particles piled up at
//    one point) get the same term from each candidate: in their walk, lane
//    j computes candidate j's term once and every lane adds the broadcast
//    terms of its own neighbors in order;
//  * other tiles are walked as maximal runs of consecutive candidates
//    holding one record (bit for bit), each with prep at its point and
//    then, in mode 0, pair_rep once for the lane's members in the run;
//  * other lanes walk their own run-length-encoded lists.
// Every lane belongs to exactly one walk, so each sum still runs over the
// lane's list in order.
template <typename Load, typename Ops, typename TFlag, typename Prep, typename PairU,
          typename PairRep, typename OnCount>
__device__ __forceinline__ void gather(const StepConsts& c, const NbrView& nv, const CellView& cv,
                                       const LaneCell& L, Load&& load, Ops&& ops,
                                       TFlag&& tflag, Prep&& prep, PairU&& pair_u,
                                       PairRep&& pair_rep, OnCount&& on_count) {
  const float3 me0 = ld3(cv.p0, L.s);
  using Q = decltype(load(0));
  // one walk of the group `gr` by the lanes with `in` set
  auto walk = [&](const Group& gr, bool in, bool uni) {
    const bool on = c.coop_bulk != 0;
    const LaneClasses lc = lane_classes(me0, in, on);
    // a pending range of candidates with one record: the record and its
    // p^0 and tile flag (uniform), this lane's members in it
    bool pend = false, pend_tf = false, pend_near = false;
    Q pend_q{};
    float3 pend_p0 = make_float3(0.0f, 0.0f, 0.0f);
    int pend_cnt = 0;
    auto flush = [&]() {
      if (!pend) return;
      pend = false;
      const int mode = pend_cnt > 0 ? prep(pos_of(pend_q.p), pend_tf) : 1;
      if (mode == 2)
        on_count(pend_cnt);
      else if (mode == 0)
        pair_rep(pend_q, pend_cnt);
    };
    group_steps(
        gr, cv.start, cv.run_end, [&](int t) { return TileP0<Q>{load(t), ld3(cv.p0, t)}; },
        [&](int base, int m, const TileP0<Q>& tt, bool whole) {
          const Q& tile = tt.q;
          // one record across the step (then r, r0), continuing the pending range?
          const int lane = threadIdx.x & (kWave - 1);
          const Q r = bcast(tile, 0);
          const float3 r0 = bcast(tt.p0, 0);
          const bool one =
              whole || (on && __all(lane >= m || (same_rec(tile, r) && same_bits(tt.p0, r0))));
          const bool cont = one && pend && same_rec(r, pend_q) && same_bits(r0, pend_p0);
          if (!cont) flush();
          if (one) {
            if (!cont) {
              pend = true;
              pend_q = r;
              pend_p0 = r0;
              pend_tf = tflag(tile, whole ? 1 : m);
              pend_near = in && near0(me0, r0, c);
              pend_cnt = 0;
            }
            const int self = L.s - base;
            if (pend_near) pend_cnt += m - ((self >= 0 && self < m) ? 1 : 0);
            return;
          }
          // this lane's neighbors among the tile's candidates, and the
          // candidates some lane takes
          unsigned long long any;
          const unsigned long long mine =
              tile_members(lc, me0, in, tt.p0, m, L.s - base, on, c, any);
          int cls;
          if (c.coop_bulk && tile_point_ids<kTilePoints>(pos_of(tile.p), m, cls)) {
            // Candidates at a few points: walk maximal runs of consecutive
            // candidates at one point; each run hoists the pair geometry once.
            const bool tf = tflag(tile, m);
            unsigned long long rest = any;
            while (rest != 0ull) {
              const int j = __builtin_ctzll(rest);
              // the run's point: candidate j's position; its candidates: one ballot
              const float3 cpk = bcast(pos_of(tile.p), j);
              const unsigned long long cmk = __ballot(cls == __builtin_amdgcn_readlane(cls, j));
              const unsigned long long other = rest & ~cmk;
              const unsigned long long upto =
                  other != 0ull ? ((1ull << __builtin_ctzll(other)) - 1ull) : ~0ull;
              const unsigned long long run = rest & upto;
              rest &= ~upto;
              const unsigned long long mr = mine & run;
              const int mode = mr != 0ull ? prep(cpk, tf) : 1;
              if (__all(mode != 0)) {
                if (mode == 2) on_count(__popcll(mr));
                continue;
              }
              // a run whose candidates hold one record gives one term, k times
              const auto q0 = bcast(tile, j);
              if (__all(((run >> lane) & 1ull) == 0ull || same_rec(tile, q0))) {
                if (mode == 0)
                  pair_rep(q0, __popcll(mr));
                else if (mode == 2)
                  on_count(__popcll(mr));
                continue;
              }
              for (unsigned long long b = run; b != 0ull; b &= b - 1ull) {
                const int jj = __builtin_ctzll(b);
                const auto q = bcast(tile, jj);
                if ((mr >> jj) & 1ull) {
                  if (mode == 0)
                    pair_u(q);
                  else if (mode == 2)
                    on_count(1);
                }
              }
            }
            return;
          }
          if (uni) {
            const auto t = ops.gterm(tile);
            for (unsigned long long b = any; b != 0ull; b &= b - 1ull) {
              const int j = __builtin_ctzll(b);
              const auto tj = bcast(t, j);
              if ((mine >> j) & 1ull) ops.add(tj);
            }
            return;
          }
          if (!c.coop_bulk) {
            for (unsigned long long b = any; b != 0ull; b &= b - 1ull) {
              const int j = __builtin_ctzll(b);
              const auto q = bcast(tile, j);
              if ((mine >> j) & 1ull) ops.pair(q);
            }
            return;
          }
          // other tiles: maximal runs of consecutive candidates holding one
          // record (one position), each through prep once per lane; a run
          // starts where a candidate's record differs from the previous one's
          const bool tf = tflag(tile, m);
          // (the permute runs in every lane: inactive lanes give no data)
          const auto prev = prev_lane(tile);
          const bool differs = !same_rec(tile, prev);
          const unsigned long long starts = __ballot(lane < m && (lane == 0 || differs));
          for (unsigned long long rr = any; rr != 0ull;) {
            const int j = __builtin_ctzll(rr);
            const auto q0 = bcast(tile, j);
            const unsigned long long sub = run_from(rr, starts, j);
            rr &= ~sub;
            const unsigned long long mr = mine & sub;
            const int mode = mr != 0ull ? prep(pos_of(q0.p), tf) : 1;
            if (mode == 0)
              pair_rep(q0, __popcll(mr));
            else if (mode == 2)
              on_count(__popcll(mr));
          }
        });
    flush();
  };
  if (c.coop)
    for_each_key_group(L.dense, L.key, [&](int leader, bool mine) {
      const Group gr = group_of(c, cv, L, leader);
      // count the group's state classes (up to one past the limit)
      int classes = 0;
      if (c.coop_bulk) {
        bool pend = mine;
        while (classes <= kMaxClasses) {
          const unsigned long long pm = __ballot(pend);
          if (pm == 0ull) break;
          const bool eq = ops.same(__builtin_ctzll(pm));
          pend = pend && !eq;
          ++classes;
        }
      }
      if (classes == 0 || classes > kMaxClasses) {
        walk(gr, mine, false);
        return;
      }
      bool pend = mine;
      while (true) {
        const unsigned long long pm = __ballot(pend);
        if (pm == 0ull) break;
        const int cl = __builtin_ctzll(pm);
        const bool eq = ops.same(cl);  // broadcasts: whole wavefront
        const bool in = pend && eq;
        ops.bind(cl);
        walk(gr, in, true);
        pend = pend && !in;
      }
    });
  if (L.active && !L.dense)
    for_each_neighbor(nv, L.s, load, [&](int, const auto& q) { ops.pair(q); });
}


// Candidate records loaded once per tile and broadcast lane to lane.
struct EPos {
  float3 p;
};
struct EPosL {
  float3 p;
  float l;  // lambda
};
struct EPosV {
  float3 p;
  float3 v;
  float irho;  // m / rho (0 when rho <= 0)
};
struct EPosV3 {
  float3 p;
  float3 v;
};
struct EPosW {
  float3 p;
  float w;  // |omega|
};
__device__ __forceinline__ EPos bcast(const EPos& e, int j) { return EPos{bcast(e.p, j)}; }
__device__ __forceinline__ EPosL bcast(const EPosL& e, int j) {
  return EPosL{bcast(e.p, j), bcast(e.l, j)};
}
__device__ __forceinline__ EPosV bcast(const EPosV& e, int j) {
  return EPosV{bcast(e.p, j), bcast(e.v, j), bcast(e.irho, j)};
}
__device__ __forceinline__ EPosV3 bcast(const EPosV3& e, int j) {
  return EPosV3{bcast(e.p, j), bcast(e.v, j)};
}
__device__ __forceinline__ EPosW bcast(const EPosW& e, int j) {
  return EPosW{bcast(e.p, j), bcast(e.w, j)};
}

// The previous lane's candidate record.
__device__ __forceinline__ EPos prev_lane(const EPos& e) { return EPos{prev_lane(e.p)}; }
__device__ __forceinline__ EPosL prev_lane(const EPosL& e) {
  return EPosL{prev_lane(e.p), prev_lane(e.l)};
}
__device__ __forceinline__ EPosV prev_lane(const EPosV& e) {
  return EPosV{prev_lane(e.p), prev_lane(e.v), prev_lane(e.irho)};
}
__device__ __forceinline__ EPosV3 prev_lane(const EPosV3& e) {
  return EPosV3{prev_lane(e.p), prev_lane(e.v)};
}
__device__ __forceinline__ EPosW prev_lane(const EPosW& e) {
  return EPosW{prev_lane(e.p), prev_lane(e.w)};
}

// Bitwise equality of candidate records.
__device__ __forceinline__ bool same_rec(const EPos& a, const EPos& b) { return same_bits(a.p, b.p); }
__device__ __forceinline__ bool same_rec(const EPosL& a, const EPosL& b) {
  return same_bits(a.p, b.p) && same_bits(a.l, b.l);
}
__device__ __forceinline__ bool same_rec(const EPosV& a, const EPosV& b) {
  return same_bits(a.p, b.p) && same_bits(a.v, b.v) && same_bits(a.irho, b.irho);
}
__device__ __forceinline__ bool same_rec(const EPosV3& a, const EPosV3& b) {
  return same_bits(a.p, b.p) && same_bits(a.v, b.v);
}
__device__ __forceinline__ bool same_rec(const EPosW& a, const EPosW& b) {
  return same_bits(a.p, b.p) && same_bits(a.w, b.w);
}

__device__ __forceinline__ bool tile_all(bool ok) { return __ballot(ok) == ~0ull; }

// Per-candidate terms of the passes (what one neighbor adds to the sums).
struct LTerm {  // lambda: rho, gradient sum, squared gradient
  float w, gx, gy, gz, j2;
  int grad;
};
struct V3Term {  // delta, XSPH, omega, eta: a 3-vector, or nothing
  float x, y, z;
  int on;
};
__device__ __forceinline__ LTerm bcast(const LTerm& t, int j) {
  return LTerm{bcast(t.w, j), bcast(t.gx, j), bcast(t.gy, j), bcast(t.gz, j), bcast(t.j2, j),
               bcast(t.grad, j)};
}
__device__ __forceinline__ V3Term bcast(const V3Term& t, int j) {
  return V3Term{bcast(t.x, j), bcast(t.y, j), bcast(t.z, j), bcast(t.on, j)};
}
// k successive additions of one 3-vector term, exactly (exact_sum.h): short
// runs in one loop over the components (unrolled: the trip count varies by
// lane), long ones by the repeated addition.
__device__ __forceinline__ void add_rep3(float& x, float& y, float& z, const V3Term& t, int k) {
  if (!t.on) return;
  if (k <= kAddInPlace) {
#pragma unroll 4
    for (int i = 0; i < k; ++i) {
      x += t.x;
      y += t.y;
      z += t.z;
    }
    return;
  }
  x = add_repeat(x, t.x, k);
  y = add_repeat(y, t.y, k);
  z = add_repeat(z, t.z, k);
}

__device__ __forceinline__ bool same_term(const LTerm& a, const LTerm& b) {
  return (static_cast<unsigned>(a.grad ^ b.grad) | bit_diff(a.w, b.w) | bit_diff(a.gx, b.gx) |
          bit_diff(a.gy, b.gy) | bit_diff(a.gz, b.gz) | bit_diff(a.j2, b.j2)) == 0u;
}

// ---------------------------------------------------------------- solver
//
// Coincident pairs (dx = dy = dz = 0) are common where particles are pushed
// onto walls and corners.  Their gradient terms are signed zeros; the
// accumulators start at +0 and can never become -0 (x + y rounds an exact
// cancellation to +0), so adding a signed zero leaves them bit-for-bit
// unchanged and those additions are skipped.  Only finite products are
// skipped: a non-finite factor times zero is NaN and is added as the
// reference adds it.  A tile whose candidates all sit at one point gives
// every lane one fixed pair geometry, computed once for the tile.

// lambda pass (core.cpp:281-329).  Writes lambda into pred[s].w; with
// sq_out (a recorded substep) also the gradient sum S (xyz) and the
// denominator sum Q (w), lambda = -C / (Q + eps).
__global__ __launch_bounds__(1024) void k_lambda(StepConsts c, float4* __restrict__ P,
                                                 NbrView nv, CellView cv,
                                                 float* __restrict__ rho_out, int write_rho,
                                                 float4* __restrict__ sq_out,
                                                 const unsigned* __restrict__ flags) {
  if (fatal(flags) || past_lanes(c, cv)) return;
  const LaneCell L =
      lane_cell(c, cv, map_lane(), nv.count,
                nv.nw);
  const int s = L.s;
  const float3 p = ld3(P, s);
  float rho = 0.0f, gsx = 0.0f, gsy = 0.0f, gsz = 0.0f, sg2 = 0.0f;
  const float ngs = -c.grad_scale;
  // what a candidate at q adds for a particle at ps
  auto term = [&](float3 ps, float3 q) -> LTerm {
    const float dx = ps.x - q.x;
    const float dy = ps.y - q.y;
    const float dz = ps.z - q.z;
    const float r2 = dx * dx + dy * dy + dz * dz;
    // coincident: poly6(0), and the f * (+-0) terms leave the sums unchanged
    const bool same = (dx == 0.0f) & (dy == 0.0f) & (dz == 0.0f);
    const bool grad = !same & (r2 < c.h2);
    const float f = grad ? pair_spiky(r2, c) : 0.0f;
    const float gx = f * dx, gy = f * dy, gz = f * dz;
    const float jx = ngs * gx, jy = ngs * gy, jz = ngs * gz;
    return LTerm{same ? c.w0 : poly6(r2, c), gx, gy, gz, jx * jx + jy * jy + jz * jz,
                 grad ? 1 : 0};
  };
  // Identical terms of consecutive neighbors are counted and added in one
  // go (exact_sum.h) before anything else is added.
  LTerm run{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0};
  int run_n = 0;
  auto flush = [&]() {
    if (run_n == 0) return;
    if (run_n <= kAddInPlace) {
      // one loop for every accumulator (each still receives its own
      // additions in order), unrolled
      if (run.grad) {
#pragma unroll 4
        for (int i = 0; i < run_n; ++i) {
          rho += run.w;
          gsx += run.gx;
          gsy += run.gy;
          gsz += run.gz;
          sg2 += run.j2;
        }
      } else {
#pragma unroll 4
        for (int i = 0; i < run_n; ++i) rho += run.w;
      }
    } else {
      rho = add_repeat(rho, run.w, run_n);
      if (run.grad) {
        gsx = add_repeat(gsx, run.gx, run_n);
        gsy = add_repeat(gsy, run.gy, run_n);
        gsz = add_repeat(gsz, run.gz, run_n);
        sg2 = add_repeat(sg2, run.j2, run_n);
      }
    }
    run_n = 0;
  };
  auto add = [&](LTerm t) {
    flush();
    rho += t.w;
    if (t.grad) {
      gsx += t.gx;
      gsy += t.gy;
      gsz += t.gz;
      sg2 += t.j2;
    }
  };
  float3 gp = p;      // the group's position
  LTerm tile_t = run;  // the term of the current run at one point
  gather(
      c, nv, cv, L, [&](int t) { return EPos{ld3(P, t)}; },
      pair_ops([&](int leader) { return same_bits(bcast(p, leader), p); },
               [&](int leader) { gp = bcast(p, leader); },
               [&](const EPos& q) { return term(gp, q.p); }, add,
               [&](const EPos& q) { add(term(p, q.p)); }),
      [&](const EPos&, int) { return true; },
      [&](const float3& cp, bool) {
        tile_t = term(p, cp);
        return 2;
      },
      [&](const EPos&) {}, [&](const EPos&, int) {},
      [&](int cnt) {
        // cnt identical neighbors: each accumulator receives the same term
        // cnt times, in order
        if ((run_n > 0) & !same_term(run, tile_t)) flush();
        run = tile_t;
        run_n += cnt;
      });
  flush();
  if (!L.active) return;
  rho += c.w0;
  rho *= c.mass;
  const float C = rho * c.inv_density - 1.0f;
  const float gix = c.grad_scale * gsx, giy = c.grad_scale * gsy, giz = c.grad_scale * gsz;
  sg2 += gix * gix + giy * giy + giz * giz;
  reinterpret_cast<float*>(P + s)[3] = -C / (sg2 + c.eps);
  if (write_rho) rho_out[s] = rho;
  if (sq_out) sq_out[s] = make_float4(gsx, gsy, gsz, sg2);
}

// delta pass with plane projection and apply (core.cpp:334-407).  kLast
// also commits (core.cpp:414-421) into posA/vout; kSolve=false is the
// zero-iteration commit.  kRest fuses plane restitution (core.cpp:573-612).
// A recorded substep also stores the planes that pushed (pmask) and the
// restitution decisions (rmask).  Held to 96 VGPRs (5 wavefronts per SIMD):
// measured 1.4% faster than the compiler's 107 (XSPH below likewise, 0.8%;
// λ would spill to scratch, ω measured slower).
template <bool kSolve, bool kLast, bool kRest>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(5))) void k_delta(
    StepConsts c, const float4* __restrict__ P, float4* __restrict__ Pnext, NbrView nv,
    CellView cv, const ScenePlanes* __restrict__ sp, const float4* __restrict__ posB,
    float4* __restrict__ posA, float4* __restrict__ vout, const float* __restrict__ rho,
    unsigned* __restrict__ pmask, unsigned* __restrict__ rmask,
    const unsigned* __restrict__ flags) {
  if (fatal(flags) || past_lanes(c, cv)) return;
  const LaneCell L =
      lane_cell(c, cv, map_lane(), nv.count,
                nv.nw);
  const int s = L.s;
  const float4 pi = P[s];
  float nx = pi.x, ny = pi.y, nz = pi.z;
  if (kSolve) {
    float ddx = 0.0f, ddy = 0.0f, ddz = 0.0f;
    // uniform-tile geometry
    float hdx = 0.f, hdy = 0.f, hdz = 0.f, hf = 0.f, hsc = 0.f;
    bool hsame = false;
    // what a neighbor with lambda lj adds for a particle with lambda ps.w
    auto dterm = [&](const float4& ps, float dx, float dy, float dz, bool same, float f, float W,
                     float lj) {
      V3Term t{0.0f, 0.0f, 0.0f, 0};
      float sv = ps.w + lj;
      if (c.scorr_on) {
        const float ratio = W * c.scorr_inv_wdq;
        const float scorr = c.scorr_neg_k * pow_ratio(ratio, c.scorr_n);
        sv += scorr;
      }
      const float sf = sv * f;
      if (same && isfinite(sf)) return t;  // (sv f) * (+-0): +-0
      t.x = sf * dx;
      t.y = sf * dy;
      t.z = sf * dz;
      t.on = 1;
      return t;
    };
    auto term = [&](const float4& ps, const EPosL& q) {
      const float dx = ps.x - q.p.x;
      const float dy = ps.y - q.p.y;
      const float dz = ps.z - q.p.z;
      const float r2 = dx * dx + dy * dy + dz * dz;
      if (!(r2 < c.h2)) return V3Term{0.0f, 0.0f, 0.0f, 0};
      const bool same = (dx == 0.0f && dy == 0.0f && dz == 0.0f);
      return dterm(ps, dx, dy, dz, same, same ? c.f_min : pair_spiky(r2, c),
                   same ? c.w0 : poly6(r2, c), q.l);
    };
    auto add = [&](const V3Term& t) {
      if (!t.on) return;
      ddx += t.x;
      ddy += t.y;
      ddz += t.z;
    };
    float4 gpi = pi;  // the group's position and lambda
    gather(
        c, nv, cv, L,
        [&](int t) {
          const float4 q = P[t];
          return EPosL{make_float3(q.x, q.y, q.z), q.w};
        },
        pair_ops([&](int leader) { return same_bits(bcast(pi, leader), pi); },
                 [&](int leader) { gpi = bcast(pi, leader); },
                 [&](const EPosL& q) { return term(gpi, q); }, add,
                 [&](const EPosL& q) { add(term(pi, q)); }),
        // bounded lambdas keep (sv f) finite for coincident neighbors
        [&](const EPosL& tile, int m) {
          const int lane = threadIdx.x & (kWave - 1);
          return tile_all(lane >= m || fabsf(tile.l) < c.lam_safe);
        },
        [&](const float3& cp, bool tf) {
          hdx = pi.x - cp.x;
          hdy = pi.y - cp.y;
          hdz = pi.z - cp.z;
          const float r2 = hdx * hdx + hdy * hdy + hdz * hdz;
          if (!(r2 < c.h2)) return 1;
          hsame = (hdx == 0.0f && hdy == 0.0f && hdz == 0.0f);
          if (hsame && tf && fabsf(pi.w) < c.lam_safe) return 1;
          hf = hsame ? c.f_min : pair_spiky(r2, c);
          hsc = hsame ? c.w0 : poly6(r2, c);  // W, for s_corr
          return 0;
        },
        [&](const EPosL& q) { add(dterm(pi, hdx, hdy, hdz, hsame, hf, hsc, q.l)); },
        [&](const EPosL& q, int k) {
          add_rep3(ddx, ddy, ddz, dterm(pi, hdx, hdy, hdz, hsame, hf, hsc, q.l), k);
        },
        [&](int) {});
    if (!L.active) return;
    ddx *= c.inv_density;
    ddy *= c.inv_density;
    ddz *= c.inv_density;
    const ScenePlanes& sc = sp[scene_of(s, cv.seg, c.n_scenes)];
    unsigned pushed = 0;
    if (sc.n_planes > 0) {
      float qx = pi.x + ddx, qy = pi.y + ddy, qz = pi.z + ddz;
      pushed = project_planes(qx, qy, qz, sc);
      ddx = qx - pi.x;
      ddy = qy - pi.y;
      ddz = qz - pi.z;
    }
    if (pmask) pmask[s] = pushed;
    nx = pi.x + ddx;  // core.cpp:404-406
    ny = pi.y + ddy;
    nz = pi.z + ddz;
  }
  if (!L.active) return;
  if (!kLast) {
    Pnext[s] = make_float4(nx, ny, nz, 0.0f);
    return;
  }
  const float4 x = posB[s];
  float vx = (nx - x.x) / c.dt;  // core.cpp:415-417
  float vy = (ny - x.y) / c.dt;
  float vz = (nz - x.z) / c.dt;
  posA[s] = make_float4(nx, ny, nz, x.w);
  if (kRest) {
    const unsigned rm = restitute(vx, vy, vz, nx, ny, nz, sp[scene_of(s, cv.seg, c.n_scenes)]);
    if (rmask) rmask[s] = rm;
  }
  vout[s] = make_float4(vx, vy, vz, kSolve ? rho[s] : 0.0f);
}

// XSPH (core.cpp:430-465); vin.w carries rho_j of the last lambda pass.
template <bool kRest>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(5))) void k_xsph(StepConsts c, const float4* __restrict__ pos,
                                               const float4* __restrict__ vin,
                                               float4* __restrict__ vout, NbrView nv,
                                               CellView cv, const ScenePlanes* __restrict__ sp,
                                               unsigned* __restrict__ rmask,
                                               const unsigned* __restrict__ flags) {
  if (fatal(flags) || past_lanes(c, cv)) return;
  const LaneCell L =
      lane_cell(c, cv, map_lane(), nv.count,
                nv.nw);
  const int s = L.s;
  const float3 xi = ld3(pos, s);
  const float4 vi = vin[s];
  float ax = 0.0f, ay = 0.0f, az = 0.0f;
  float hW = 0.0f;
  float3 tv = make_float3(0.f, 0.f, 0.f);
  const float3 vi3 = make_float3(vi.x, vi.y, vi.z);
  // what a neighbor adds for a particle moving at vs
  auto xterm = [&](const float3& vs, float W, const EPosV& q) {
    return V3Term{(q.v.x - vs.x) * W * q.irho, (q.v.y - vs.y) * W * q.irho,
                  (q.v.z - vs.z) * W * q.irho, 1};
  };
  auto term = [&](const float3& xs, const float3& vs, const EPosV& q) {
    const float dx = xs.x - q.p.x;
    const float dy = xs.y - q.p.y;
    const float dz = xs.z - q.p.z;
    const float r2 = dx * dx + dy * dy + dz * dz;
    if (!(r2 < c.h2)) return V3Term{0.0f, 0.0f, 0.0f, 0};
    return xterm(vs, poly6(r2, c), q);
  };
  auto add = [&](const V3Term& t) {
    if (!t.on) return;
    ax += t.x;
    ay += t.y;
    az += t.z;
  };
  float3 gx3 = xi, gv3 = vi3;  // the group's position and velocity
  gather(
      c, nv, cv, L,
      [&](int t) {
        const float4 v = vin[t];
        return EPosV{ld3(pos, t), make_float3(v.x, v.y, v.z),
                     (v.w > 0.0f) ? (c.mass / v.w) : 0.0f};
      },
      pair_ops(
          [&](int leader) {
            return same_bits(bcast(xi, leader), xi) && same_bits(bcast(vi3, leader), vi3);
          },
          [&](int leader) {
            gx3 = bcast(xi, leader);
            gv3 = bcast(vi3, leader);
          },
          [&](const EPosV& q) { return term(gx3, gv3, q); }, add,
          [&](const EPosV& q) { add(term(xi, vi3, q)); }),
      // one velocity in the whole tile (bitwise), finite m/rho everywhere
      [&](const EPosV& tile, int m) {
        const int lane = threadIdx.x & (kWave - 1);
        tv = make_float3(bcast(tile.v.x, 0), bcast(tile.v.y, 0), bcast(tile.v.z, 0));
        return tile_all(lane >= m || (__float_as_int(tile.v.x) == __float_as_int(tv.x) &&
                                      __float_as_int(tile.v.y) == __float_as_int(tv.y) &&
                                      __float_as_int(tile.v.z) == __float_as_int(tv.z) &&
                                      isfinite(tile.irho)));
      },
      [&](const float3& cp, bool tf) {
        const float dx = xi.x - cp.x;
        const float dy = xi.y - cp.y;
        const float dz = xi.z - cp.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        if (!(r2 < c.h2)) return 1;
        hW = poly6(r2, c);
        // (vj - vi) = +0 exactly for every neighbor: each term is a zero and
        // leaves the sums unchanged (finite W and m/rho)
        if (tf && vi.x == tv.x && vi.y == tv.y && vi.z == tv.z && isfinite(vi.x) &&
            isfinite(vi.y) && isfinite(vi.z) && isfinite(hW))
          return 1;
        return 0;
      },
      [&](const EPosV& q) { add(xterm(vi3, hW, q)); },
      [&](const EPosV& q, int k) { add_rep3(ax, ay, az, xterm(vi3, hW, q), k); }, [&](int) {});
  if (!L.active) return;
  float vx = vi.x + c.visc_c * ax;
  float vy = vi.y + c.visc_c * ay;
  float vz = vi.z + c.visc_c * az;
  if (kRest) {
    const unsigned rm =
        restitute(vx, vy, vz, xi.x, xi.y, xi.z, sp[scene_of(s, cv.seg, c.n_scenes)]);
    if (rmask) rmask[s] = rm;
  }
  vout[s] = make_float4(vx, vy, vz, 0.0f);
}

// vorticity: omega and |omega| (core.cpp:472-508)
__global__ __launch_bounds__(1024) void k_vort_omega(StepConsts c, const float4* __restrict__ pos,
                                                     const float4* __restrict__ vel,
                                                     float4* __restrict__ omega, NbrView nv,
                                                     CellView cv,
                                                     const unsigned* __restrict__ flags) {
  if (fatal(flags) || past_lanes(c, cv)) return;
  const LaneCell L =
      lane_cell(c, cv, map_lane(), nv.count,
                nv.nw);
  const int s = L.s;
  const float3 xi = ld3(pos, s);
  const float3 vi = ld3(vel, s);
  float ox = 0.0f, oy = 0.0f, oz = 0.0f;
  float hgx = 0.f, hgy = 0.f, hgz = 0.f;
  bool hsame = false;
  // what a neighbor moving at vj adds for a particle moving at vs
  auto wterm = [&](const float3& vs, bool same, float gx, float gy, float gz, const float3& vj) {
    const float dvx = vj.x - vs.x, dvy = vj.y - vs.y, dvz = vj.z - vs.z;
    if (same && isfinite(dvx) && isfinite(dvy) && isfinite(dvz))
      return V3Term{0.0f, 0.0f, 0.0f, 0};  // dv x (f * (+-0)) is a signed zero
    return V3Term{dvy * gz - dvz * gy, dvz * gx - dvx * gz, dvx * gy - dvy * gx, 1};
  };
  auto term = [&](const float3& xs, const float3& vs, const EPosV3& q) {
    const float dx = xs.x - q.p.x;
    const float dy = xs.y - q.p.y;
    const float dz = xs.z - q.p.z;
    const float r2 = dx * dx + dy * dy + dz * dz;
    if (!(r2 < c.h2)) return V3Term{0.0f, 0.0f, 0.0f, 0};
    const bool same = (dx == 0.0f && dy == 0.0f && dz == 0.0f);
    const float f = pair_spiky(r2, c);
    return wterm(vs, same, f * dx, f * dy, f * dz, q.v);
  };
  auto add = [&](const V3Term& t) {
    if (!t.on) return;
    ox += t.x;
    oy += t.y;
    oz += t.z;
  };
  float3 gx3 = xi, gv3 = vi;  // the group's position and velocity
  const bool own_ok = fabsf(vi.x) < kHalfMax && fabsf(vi.y) < kHalfMax && fabsf(vi.z) < kHalfMax;
  gather(
      c, nv, cv, L, [&](int t) { return EPosV3{ld3(pos, t), ld3(vel, t)}; },
      pair_ops(
          [&](int leader) {
            return same_bits(bcast(xi, leader), xi) && same_bits(bcast(vi, leader), vi);
          },
          [&](int leader) {
            gx3 = bcast(xi, leader);
            gv3 = bcast(vi, leader);
          },
          [&](const EPosV3& q) { return term(gx3, gv3, q); }, add,
          [&](const EPosV3& q) { add(term(xi, vi, q)); }),
      // finite velocity differences make coincident terms signed zeros
      [&](const EPosV3& tile, int m) {
        const int lane = threadIdx.x & (kWave - 1);
        return tile_all(lane >= m || (fabsf(tile.v.x) < kHalfMax && fabsf(tile.v.y) < kHalfMax &&
                                      fabsf(tile.v.z) < kHalfMax));
      },
      [&](const float3& cp, bool tf) {
        const float dx = xi.x - cp.x;
        const float dy = xi.y - cp.y;
        const float dz = xi.z - cp.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        if (!(r2 < c.h2)) return 1;
        hsame = (dx == 0.0f && dy == 0.0f && dz == 0.0f);
        if (hsame && tf && own_ok) return 1;
        const float f = pair_spiky(r2, c);
        hgx = f * dx, hgy = f * dy, hgz = f * dz;
        return 0;
      },
      [&](const EPosV3& q) { add(wterm(vi, hsame, hgx, hgy, hgz, q.v)); },
      [&](const EPosV3& q, int k) { add_rep3(ox, oy, oz, wterm(vi, hsame, hgx, hgy, hgz, q.v), k); },
      [&](int) {});
  if (!L.active) return;
  omega[s] = make_float4(ox, oy, oz, sqrtf(ox * ox + oy * oy + oz * oz));
}

// vorticity: eta, confinement force, velocity update (core.cpp:512-570),
// in place on vel (no neighbor reads vel here).
template <bool kRest>
__global__ __launch_bounds__(1024) void k_vort_apply(StepConsts c, const float4* __restrict__ pos,
                                                     const float4* __restrict__ omega,
                                                     float4* __restrict__ vel, NbrView nv,
                                                     CellView cv,
                                                     const ScenePlanes* __restrict__ sp,
                                                     const unsigned* __restrict__ flags) {
  if (fatal(flags) || past_lanes(c, cv)) return;
  const LaneCell L =
      lane_cell(c, cv, map_lane(), nv.count,
                nv.nw);
  const int s = L.s;
  const float3 xi = ld3(pos, s);
  const float4 om = omega[s];
  float ex = 0.0f, ey = 0.0f, ez = 0.0f;
  float hgx = 0.f, hgy = 0.f, hgz = 0.f;
  bool hsame = false;
  // what a neighbor with |omega| = omj adds for a particle with |omega| = ws
  auto eterm = [&](float ws, bool same, float gx, float gy, float gz, float omj) {
    const float co = omj - ws;
    if (same && isfinite(co)) return V3Term{0.0f, 0.0f, 0.0f, 0};
    return V3Term{co * gx, co * gy, co * gz, 1};
  };
  auto term = [&](const float3& xs, float ws, const EPosW& q) {
    const float dx = xs.x - q.p.x;
    const float dy = xs.y - q.p.y;
    const float dz = xs.z - q.p.z;
    const float r2 = dx * dx + dy * dy + dz * dz;
    if (!(r2 < c.h2)) return V3Term{0.0f, 0.0f, 0.0f, 0};
    const bool same = (dx == 0.0f && dy == 0.0f && dz == 0.0f);
    const float f = pair_spiky(r2, c);
    return eterm(ws, same, f * dx, f * dy, f * dz, q.w);
  };
  auto add = [&](const V3Term& t) {
    if (!t.on) return;
    ex += t.x;
    ey += t.y;
    ez += t.z;
  };
  float3 gx3 = xi;  // the group's position and |omega|
  float gw = om.w;
  gather(
      c, nv, cv, L,
      [&](int t) { return EPosW{ld3(pos, t), reinterpret_cast<const float*>(omega + t)[3]}; },
      pair_ops(
          [&](int leader) {
            return same_bits(bcast(xi, leader), xi) && same_bits(bcast(om.w, leader), om.w);
          },
          [&](int leader) {
            gx3 = bcast(xi, leader);
            gw = bcast(om.w, leader);
          },
          [&](const EPosW& q) { return term(gx3, gw, q); }, add,
          [&](const EPosW& q) { add(term(xi, om.w, q)); }),
      [&](const EPosW& tile, int m) {
        const int lane = threadIdx.x & (kWave - 1);
        return tile_all(lane >= m || fabsf(tile.w) < kHalfMax);
      },
      [&](const float3& cp, bool tf) {
        const float dx = xi.x - cp.x;
        const float dy = xi.y - cp.y;
        const float dz = xi.z - cp.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        if (!(r2 < c.h2)) return 1;
        hsame = (dx == 0.0f && dy == 0.0f && dz == 0.0f);
        if (hsame && tf && fabsf(om.w) < kHalfMax) return 1;
        const float f = pair_spiky(r2, c);
        hgx = f * dx, hgy = f * dy, hgz = f * dz;
        return 0;
      },
      [&](const EPosW& q) { add(eterm(om.w, hsame, hgx, hgy, hgz, q.w)); },
      [&](const EPosW& q, int k) { add_rep3(ex, ey, ez, eterm(om.w, hsame, hgx, hgy, hgz, q.w), k); },
      [&](int) {});
  if (!L.active) return;
  const float len = sqrtf(ex * ex + ey * ey + ez * ez);
  float nx = 0.0f, ny = 0.0f, nz = 0.0f;
  if (len > c.vort_norm_eps) {
    const float inv = 1.0f / len;
    nx = ex * inv;
    ny = ey * inv;
    nz = ez * inv;
  }
  const float fx = c.vort_eps * (ny * om.z - nz * om.y);
  const float fy = c.vort_eps * (nz * om.x - nx * om.z);
  const float fz = c.vort_eps * (nx * om.y - ny * om.x);
  const float4 v = vel[s];
  float vx = v.x + c.dt * fx;
  float vy = v.y + c.dt * fy;
  float vz = v.z + c.dt * fz;
  if (kRest) restitute(vx, vy, vz, xi.x, xi.y, xi.z, sp[scene_of(s, cv.seg, c.n_scenes)]);
  vel[s] = make_float4(vx, vy, vz, 0.0f);
}

// ---------------------------------------------------------------- runs
//
// Particles at one position with one velocity (bit for bit, after gravity),
// in one cell, have the same candidates and the same tests: their lists
// differ only in which slot each leaves out as itself.  When every slot
// between two of them is such a particle too, the two lists hold the same
// values in the same order, so every pass computes bit-identical results for
// them.  A maximal run of such consecutive slots is therefore computed by
// its first slot alone (one lane per run), and the others copy its outputs
// after each pass (k_run_fill).

// flag[s] = 1 when slot s starts a run.
// Per slot: run[s] = 1 when s starts a run (identical consecutive slots of
// one cell: bitwise equal position and velocity), cell[s] = 1 when s starts
// a (non-empty) cell.
__global__ __launch_bounds__(1024) void k_slot_flags(StepConsts c, const float4* __restrict__ posB,
                                                     const float4* __restrict__ velB,
                                                     const int* __restrict__ slot_key,
                                                     int* __restrict__ run,
                                                     int* __restrict__ cell,
                                                     const unsigned* __restrict__ flags) {
  if (fatal(flags)) return;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= c.n) return;
  const bool cell_head = s == 0 || slot_key[s] != slot_key[s - 1];
  const bool run_head = cell_head || !same_bits(ld3(posB, s), ld3(posB, s - 1)) ||
                        !same_bits(ld3(velB, s), ld3(velB, s - 1));
  run[s] = run_head ? 1 : 0;
  cell[s] = cell_head ? 1 : 0;
}

// From the scanned flags: the list of run heads and each slot's run number,
// the first slot of each non-empty cell and each slot's cell number (in
// place of the flags).  Slots that are not run heads store no neighbor list.
__global__ __launch_bounds__(1024) void k_slot_lists(StepConsts c, int* __restrict__ run_no,
                                                     const int* __restrict__ run_excl,
                                                     int* __restrict__ rep,
                                                     int* __restrict__ cell_no,
                                                     const int* __restrict__ cell_excl,
                                                     int* __restrict__ first,
                                                     long long* __restrict__ noff,
                                                     int* __restrict__ nwords,
                                                     int* __restrict__ ncount,
                                                     const unsigned* __restrict__ flags) {
  if (fatal(flags)) return;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= c.n) return;
  const int head = run_no[s];
  const int e = run_excl[s];
  if (head) {
    rep[e] = s;
  } else {
    noff[s] = 0;
    nwords[s] = -1;
    ncount[s] = 0;
  }
  run_no[s] = e + head - 1;
  const int ch = cell_no[s];
  const int ce = cell_excl[s];
  if (ch) first[ce] = s;
  cell_no[s] = ce + ch - 1;
}

__device__ __forceinline__ bool aligned_cell(const StepConsts& c, int a, int b) {
  return c.coop && (b - a) >= kDenseCell;
}

// Lanes of non-empty cell j: its runs, packed (lanes1) or rounded up to whole
// waves (lanes2): dense cells, and (up to 2^24 particles, which keeps the
// padded total within int) sparse cells whose 27-cell neighborhood holds at
// least kHeavyList candidates, whose particles walk cooperatively too.
__global__ __launch_bounds__(1024) void k_cell_lanes(StepConsts c, const int* __restrict__ excl,
                                                     const int* __restrict__ first,
                                                     const int* __restrict__ run_no,
                                                     const uint64_t* __restrict__ slot_cell,
                                                     const GridDesc* __restrict__ grid,
                                                     const int* __restrict__ start,
                                                     const int* __restrict__ seg,
                                                     int* __restrict__ lanes1,
                                                     int* __restrict__ lanes2,
                                                     const int* __restrict__ in_tile,
                                                     const unsigned* __restrict__ flags) {
  if (fatal(flags)) return;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int C = excl[c.n];
  if (j >= C) return;
  const int a = first[j];
  const int b = j + 1 < C ? first[j + 1] : c.n;
  if (in_tile && in_tile[a]) {  // the tile kernels compute this cell
    lanes1[j] = 0;
    lanes2[j] = 0;
    return;
  }
  const int R = run_no[b - 1] - run_no[a] + 1;
  bool al = aligned_cell(c, a, b);
  if (!al && c.coop && c.n <= (1 << 24)) {
    int x, y, z;
    unpack_cell(slot_cell[a], x, y, z);
    const GridDesc g = grid[scene_of(a, seg, c.n_scenes)];
    const int rx = x - g.ox, ry = y - g.oy, rz = z - g.oz;
    const int ext = 1 << g.bshift;
    if (rx >= 1 && ry >= 1 && rz >= 1 && rx <= ext * g.nbx - 2 && ry <= ext * g.nby - 2 &&
        rz <= ext * g.nbz - 2) {
      int cand = 0;
      for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dx = -1; dx <= 1; ++dx) {
            const int k = cell_key(rx + dx, ry + dy, rz + dz, g);
            cand += start[k + 1] - start[k];
          }
      al = cand >= kHeavyList;
    }
  }
  lanes1[j] = al ? 0 : R;
  lanes2[j] = al ? (R + kWave - 1) / kWave * kWave : 0;
}

// lane_rep: each run's head slot at its lane (off1/off2: the cells' scanned
// lane counts); padding lanes of aligned cells get -1.  The lanes in use are
// [0, lane_info[0]) and [lane_info[1], lane_info[2]).  Also run_end[s], one
// past the last slot of s's run.
__global__ __launch_bounds__(1024) void k_lane_map(StepConsts c, const int* __restrict__ run_no,
                                                   const int* __restrict__ cell_no,
                                                   const int* __restrict__ first,
                                                   const int* __restrict__ excl,
                                                   const int* __restrict__ lanes1,
                                                   const int* __restrict__ lanes2,
                                                   const int* __restrict__ off1,
                                                   const int* __restrict__ off2,
                                                   const int* __restrict__ runs_total,
                                                   long long lane_cap,
                                                   int* __restrict__ lane_rep,
                                                   int* __restrict__ lane_info,
                                                   const int* __restrict__ rep,
                                                   int* __restrict__ run_end,
                                                   const int* __restrict__ in_tile,
                                                   const unsigned* __restrict__ flags) {
  if (fatal(flags)) return;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= c.n) return;
  {
    // one past the last slot of s's run: the next run's first slot
    const int r = run_no[s] + 1;
    run_end[s] = r < *runs_total ? rep[r] : c.n;
  }
  const int C = excl[c.n];
  // dense and heavy cells first (their waves start first), packed runs after
  const int total2 = off2[C];  // a multiple of 64
  const long long end = static_cast<long long>(total2) + off1[C];
  // beyond the lane capacity: every run at its own number, packed
  const bool fits = end <= lane_cap;
  if (s == 0) {
    const int R = *runs_total;
    lane_info[0] = fits ? total2 : R;
    lane_info[1] = fits ? total2 : R;
    lane_info[2] = fits ? static_cast<int>(end) : R;
  }
  if (s > 0 && run_no[s] == run_no[s - 1]) return;  // not a run head
  if (!fits) {
    // the tile kernels compute tile slots (the eta pass updates vel in
    // place, so no slot may be computed twice): their lanes stay idle
    lane_rep[run_no[s]] = (in_tile && in_tile[s]) ? -1 : s;
    return;
  }
  const int j = cell_no[s];
  const int a = first[j];
  const int b = j + 1 < C ? first[j + 1] : c.n;
  const int rr = run_no[s] - run_no[a];
  if (lanes2[j] == 0) {
    if (lanes1[j] != 0) lane_rep[total2 + off1[j] + rr] = s;  // 0: a tile's cell
    return;
  }
  const int lane0 = off2[j];
  lane_rep[lane0 + rr] = s;
  const int R = run_no[b - 1] - run_no[a] + 1;
  if (rr == R - 1)
    for (int q = R; q < (R + kWave - 1) / kWave * kWave; ++q) lane_rep[lane0 + q] = -1;
}

// Copies each run's first slot's output into the run's other slots.  With
// `ids`, w is each slot's own original id, ids[s].w (the committed position
// carries the id of the slot's input position).
__global__ __launch_bounds__(1024) void k_run_fill(StepConsts c, const int* __restrict__ run_no,
                                                   const int* __restrict__ rep,
                                                   float4* __restrict__ a,
                                                   const float4* __restrict__ ids,
                                                   const int* __restrict__ in_tile,
                                                   const unsigned* __restrict__ flags) {
  if (fatal(flags)) return;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= c.n) return;
  // a tile computes each of its particles (coincident ones included), and
  // its kernel may run beside this one
  if (in_tile && in_tile[s]) return;
  const int h = rep[run_no[s]];
  if (h == s) return;
  float4 v = a[h];
  if (ids) v.w = ids[s].w;
  a[s] = v;
}

// ---------------------------------------------------------------- launching

struct Launcher {
  Ctx& c;
  template <typename F>
  hipError_t operator()(const char* name, F&& f) {
    return on(c.stream, name, f);
  }
  // events on stream `st` (a tile kernel on the side stream: its pass's own
  // events on the main stream span the fork to the join)
  template <typename F>
  hipError_t on(hipStream_t st, const char* name, F&& f) {
    if (c.prof.on) {
      hipEvent_t a, b;
      if (c.prof.pool.size() >= 2) {
        a = c.prof.pool.back();
        c.prof.pool.pop_back();
        b = c.prof.pool.back();
        c.prof.pool.pop_back();
      } else {
        hipError_t e = hipEventCreate(&a);
        if (e != hipSuccess) return e;
        e = hipEventCreate(&b);
        if (e != hipSuccess) return e;
      }
      (void)hipEventRecord(a, st);
      f();
      (void)hipEventRecord(b, st);
      c.prof.pending.push_back({name, a, b, c.prof.step});
    } else {
      f();
    }
    return hipGetLastError();
  }
};

inline int blocks_for(long long n, int bs) { return static_cast<int>((n + bs - 1) / bs); }

}  // namespace

hipError_t enqueue_count_pairs(Ctx& c) {
  k_reset_total<<<1, 64, 0, c.stream>>>(c.pair_total);
  if (c.n == 0) return hipGetLastError();
  StepConsts k = c.consts;
  k.count_only = 1;
  k.runs = 0;
  const CellView cv{c.slot_cell, c.slot_key, c.grid, c.start, c.seg, nullptr, nullptr, nullptr,
                    c.pred[0]};
  k_neighbors<<<blocks_for(c.n, c.block), c.block, 0, c.stream>>>(
      k, c.pred[0], cv, c.noff, c.nwords, c.ncount, c.nbr, c.nbr_cap, c.pair_total, c.flags);
  return hipGetLastError();
}

hipError_t enqueue_reset_flags(Ctx& c) {
  k_reset_flags<<<1, 64, 0, c.stream>>>(c.flags, c.steps_started);
  return hipGetLastError();
}

hipError_t enqueue_init_grid(Ctx& c) {
  k_init_bbox<<<blocks_for(c.n_scenes, 256), 256, 0, c.stream>>>(c.n_scenes, c.bbox);
  return hipGetLastError();
}

hipError_t enqueue_upload(Ctx& c, const float* d_pos, const float* d_vel) {
  if (c.n == 0) return hipSuccess;
  k_upload<<<blocks_for(c.n, 256), 256, 0, c.stream>>>(c.n, d_pos, d_vel, c.posA, c.velA);
  return hipGetLastError();
}

hipError_t enqueue_download(Ctx& c, float* d_pos, float* d_vel) {
  if (c.n == 0) return hipSuccess;
  k_download<<<blocks_for(c.n, 256), 256, 0, c.stream>>>(c.n, c.posA, c.velA, d_pos, d_vel);
  return hipGetLastError();
}

hipError_t enqueue_substep(Ctx& c, const SubstepRec* rec) {
  if (c.n == 0) return hipSuccess;
  if (rec && c.iterations < 1) return hipErrorInvalidValue;
  // iterate k of the constraint loop: p^0 stays in its own array for the
  // whole substep and later iterates alternate between two others, or the
  // recording's own array per iterate
  auto PK = [&](int k) { return rec ? rec->P[k] : iterate_buffer(c, k); };
  Launcher L{c};
  StepConsts k = c.consts;
  // a recorded substep keeps every particle's own list (the adjoint reads them)
  if (rec) {
    k.runs = 0;
    k.full_lists = 1;
  }
  const bool runs = k.runs != 0;
  // sparse bricks through the LDS-tiled passes: the lane map of runs leaves
  // their slots out, or without runs (a recorded substep) the lanes skip them
  const bool tiles = k.tiles != 0 && c.tile_list != nullptr;
  if (!tiles) k.tiles = 0;
  const TileArgs ta{c.slot_cell, c.tile_count, static_cast<int>(c.tile_cap), c.tile_rec,
                    c.tile_seg,  c.tile_src,   c.tile_list, c.tile_list_cap, c.tile_cnt};
  const CellView cv{c.slot_cell, c.slot_key, c.grid,      c.start,
                    c.seg,       c.lane_rep, c.lane_info, runs ? c.run_end : nullptr,
                    PK(0),       (tiles && !runs) ? c.in_tile : nullptr};
  c.lists_partial = !rec && k.coop && !k.full_lists;
  c.last_runs = runs;
  auto fill = [&](float4* a, const float4* ids) {
    if (runs)
      k_run_fill<<<blocks_for(c.n, 256), 256, 0, c.stream>>>(
          k, c.run_no, c.rep_list, a, ids, tiles ? c.in_tile : nullptr, c.flags);
  };
  const int bs = c.block;
  const int nb = blocks_for(c.n, bs);
  // neighbor build and passes: one lane per slot, or with runs the lane map
  const int nl = runs ? blocks_for(c.lane_cap, bs) : nb;
  hipStream_t st = c.stream;
  hipError_t e;

#define FH_TRY(x)                      \
  do {                                 \
    e = (x);                           \
    if (e != hipSuccess) return e;     \
  } while (0)

  FH_TRY(L("predict_bbox", [&] {
    k_predict_bbox<<<nb, bs, 0, st>>>(k, c.posA, c.velA, c.planes, c.seg, c.cellc, c.bbox,
                                      c.tile_bbox, c.flags);
  }));
  FH_TRY(L("grid_setup", [&] {
    k_grid_setup<<<1, kGridSetupThreads, 0, st>>>(c.n_scenes, c.seg, c.tile_bbox, bs, c.bbox, c.grid,
                                    c.total_cells, c.cell_cap, c.brick_shift,
                                    c.flags, c.nbr_total, c.need_cells, c.steps_started,
                                    tiles ? c.tile_count : nullptr);
  }));
  FH_TRY(L("cell_keys", [&] {
    k_cell_keys<<<nb, bs, 0, st>>>(k, c.cellc, c.seg, c.grid, c.key, c.flags);
  }));
  // keys lie below the cell table's capacity: 8 bits per pass
  int key_bits = 1;
  while (key_bits < 31 && (1ll << key_bits) < c.cell_cap) ++key_bits;
  const int passes = (key_bits + 7) / 8;
  FH_TRY(L("sort", [&] {
    k_sort_keys<<<blocks_for(c.n, 256), 256, 0, st>>>(k, c.posA, c.key, c.sort_k[0],
                                                       c.sort_v[0], c.flags);
    for (int q = 0; q < passes; ++q) {
      const int i = q & 1, o = i ^ 1;
      k_radix_hist<<<c.radix_blocks, 256, 0, st>>>(c.n, c.sort_k[i], 8 * q, c.radix_hist,
                                                   c.total_cells, c.flags);
      k_scan_reduce<<<c.scan_blocks, 256, 0, st>>>(c.radix_hist, nullptr, c.radix_len,
                                                   c.scan_partial, c.flags);
      k_scan_down<<<c.scan_blocks, 256, 0, st>>>(c.radix_hist, nullptr, c.radix_len,
                                                 c.scan_partial, c.radix_off, nullptr, 0,
                                                 c.flags);
      k_radix_scatter<<<c.radix_blocks, 256, 0, st>>>(c.n, c.sort_k[i], c.sort_v[i], c.sort_k[o],
                                                      c.sort_v[o], 8 * q, c.radix_off,
                                                      c.total_cells, c.flags);
    }
  }));
  FH_TRY(L("cell_count", [&] {
    // the histogram of the sorted keys and its scan: each cell's first slot
    k_cell_runs<<<blocks_for(c.n, 1024), 1024, 0, st>>>(c.n, c.sort_k[passes & 1], c.count,
                                                        c.flags);
    k_scan_reduce<<<c.scan_blocks, 256, 0, st>>>(c.count, nullptr, c.total_cells,
                                                 c.scan_partial, c.flags);
    k_scan_down<<<c.scan_blocks, 256, 0, st>>>(c.count, nullptr, c.total_cells, c.scan_partial,
                                               c.start, nullptr, 1, c.flags);
  }));
  FH_TRY(L("reorder", [&] {
    k_reorder<<<nb, bs, 0, st>>>(k, c.sort_k[passes & 1], c.sort_v[passes & 1], c.seg, c.planes,
                                 c.posA, c.velA, c.posB, c.velB, PK(0), c.cellc, c.slot_cell,
                                 c.slot_key, c.flags);
  }));
  // Each pass's tile kernel runs on the side stream beside the per-lane
  // kernel and its run fill on the main stream: they write disjoint slots
  // (tile bricks hold no run of coincident particles, and the lane map leaves
  // tile slots out) and read only what earlier passes finished.  fork() makes
  // the side stream wait for the main stream's work so far, join() the main
  // stream for the side stream's.  Both are edges of a captured graph too.
  // An event record and a cross-stream wait leave the device idle for tens
  // of microseconds (about 30 us per fork measured at 65,536 particles), so
  // small batches, whose tile kernels take about as long, stay on one stream.
  const bool forked = tiles && c.side != nullptr && c.n >= kForkMinParticles;
  hipStream_t tst = forked ? c.side : st;
  hipError_t fj = hipSuccess;
  auto fork = [&] {
    if (!forked || fj != hipSuccess) return;
    fj = hipEventRecord(c.ev_fork, st);
    if (fj == hipSuccess) fj = hipStreamWaitEvent(c.side, c.ev_fork, 0);
  };
  auto join = [&] {
    if (!forked || fj != hipSuccess) return;
    fj = hipEventRecord(c.ev_join, c.side);
    if (fj == hipSuccess) fj = hipStreamWaitEvent(st, c.ev_join, 0);
  };
  // The tile lists come before the lane map when a tile can go back to the
  // lanes (a particle with more neighbors than its list holds).  A large
  // batch's lists hold every possible neighbor, so its tiles are final after
  // classify: the lane map is built first (its short kernels would wait for
  // CUs behind the tile list build), then the tile lists are built on the
  // side stream while the main stream builds the lanes' lists (which write
  // only non-tile slots: every tile slot heads its own run).
  const bool early_lanes = forked && c.tile_list_cap >= kTileMaxCand - 1;
  auto tile_lists = [&](hipStream_t s) {
    (void)L.on(s, "tile_neighbors", [&] {
      (void)enqueue_tile_neighbors(c, k, ta, PK(0), rec != nullptr, s);
    });
  };
  if (tiles)
    FH_TRY(L("tiles", [&] {
      (void)enqueue_tile_classify(c, k);
      if (!early_lanes) tile_lists(st);
    }));
  if (runs)
    FH_TRY(L("runs", [&] {
      // runs and non-empty cells: flags, their scans (one dual scan), the
      // lists of run heads and of cells' first slots
      k_slot_flags<<<blocks_for(c.n, 256), 256, 0, st>>>(k, c.posB, c.velB, c.slot_key,
                                                          c.run_no, c.cell_no, c.flags);
      k_scan_reduce<<<c.scan_blocks, 256, 0, st>>>(c.run_no, c.cell_no, c.n_dev, c.scan_partial,
                                                   c.flags);
      k_scan_down<<<c.scan_blocks, 256, 0, st>>>(c.run_no, c.cell_no, c.n_dev, c.scan_partial,
                                                 c.run_excl, c.cell_excl, 0, c.flags);
      k_slot_lists<<<blocks_for(c.n, 256), 256, 0, st>>>(k, c.run_no, c.run_excl, c.rep_list,
                                                          c.cell_no, c.cell_excl, c.cell_first,
                                                          c.noff, c.nwords, c.ncount, c.flags);
      // lanes: the cells' lane counts, their offsets (one dual scan), the lane
      // map (which also writes each slot's run end)
      k_cell_lanes<<<blocks_for(c.n, 256), 256, 0, st>>>(k, c.cell_excl, c.cell_first, c.run_no,
                                                          c.slot_cell, c.grid, c.start, c.seg,
                                                          c.lanes1, c.lanes2,
                                                          tiles ? c.in_tile : nullptr, c.flags);
      k_scan_reduce<<<c.scan_blocks, 256, 0, st>>>(c.lanes1, c.lanes2, c.cell_excl + c.n,
                                                   c.scan_partial, c.flags);
      k_scan_down<<<c.scan_blocks, 256, 0, st>>>(c.lanes1, c.lanes2, c.cell_excl + c.n,
                                                 c.scan_partial, c.off1, c.off2, 0, c.flags);
      k_lane_map<<<blocks_for(c.n, 256), 256, 0, st>>>(
          k, c.run_no, c.cell_no, c.cell_first, c.cell_excl, c.lanes1, c.lanes2, c.off1, c.off2,
          c.run_excl + c.n, c.lane_cap, c.lane_rep, c.lane_info, c.rep_list, c.run_end,
          tiles ? c.in_tile : nullptr, c.flags);
    }));
  FH_TRY(L("neighbors", [&] {
    if (early_lanes) {
      fork();
      tile_lists(tst);
    }
    k_neighbors<<<nl, bs, 0, st>>>(k, PK(0), cv, c.noff, c.nwords,
                                   c.ncount, c.nbr, c.nbr_cap, c.nbr_total, c.flags);
    if (early_lanes) join();
  }));
  FH_TRY(fj);

  // velocity buffer that the commit writes: XSPH needs a second buffer to
  // stay Jacobi, so the commit goes to velB and XSPH back to velA.
  const NbrView nv{c.noff, c.nwords, c.nbr, c.ncount};
  float4* v1 = c.xsph ? c.velB : c.velA;
  const bool rest_in_delta = c.restitution_any && !c.xsph && !c.vort;
  const int I = c.iterations;
  for (int it = 0; it < I; ++it) {
    const bool last = (it + 1 == I);
    float4* P = PK(it);
    float4* Pn = last ? nullptr : PK(it + 1);
    float4* sq = rec ? rec->SQ[it] : nullptr;
    unsigned* pm = rec ? rec->M[it] : nullptr;
    unsigned* rm = rec ? rec->RM : nullptr;
    FH_TRY(L("lambda", [&] {
      if (tiles) {
        fork();
        (void)L.on(tst, "tile_lambda",
                   [&] { (void)enqueue_tile_lambda(c, k, ta, P, last ? 1 : 0, sq, tst); });
      }
      k_lambda<<<nl, bs, 0, st>>>(k, P, nv, cv, c.rho, last ? 1 : 0, sq, c.flags);
      fill(P, nullptr);
      if (tiles) join();
    }));
    FH_TRY(fj);
    FH_TRY(L("delta", [&] {
      if (tiles) {
        fork();
        (void)L.on(tst, "tile_delta", [&] {
          (void)enqueue_tile_delta(c, k, ta, P, Pn, last, rest_in_delta, v1, pm,
                                   last && rest_in_delta ? rm : nullptr, tst);
        });
      }
      if (!last)
        k_delta<true, false, false><<<nl, bs, 0, st>>>(k, P, Pn, nv, cv,
                                                       c.planes, c.posB, c.posA, v1, c.rho,
                                                       pm, nullptr, c.flags);
      else if (rest_in_delta)
        k_delta<true, true, true><<<nl, bs, 0, st>>>(k, P, Pn, nv, cv,
                                                     c.planes, c.posB, c.posA, v1, c.rho,
                                                     pm, rm, c.flags);
      else
        k_delta<true, true, false><<<nl, bs, 0, st>>>(k, P, Pn, nv, cv,
                                                      c.planes, c.posB, c.posA, v1, c.rho,
                                                      pm, nullptr, c.flags);
      if (!last) {
        fill(Pn, nullptr);
      } else {
        fill(c.posA, c.posB);  // each slot keeps its own id
        fill(v1, nullptr);
      }
      if (tiles) join();
    }));
    FH_TRY(fj);
  }
  if (I == 0) {
    FH_TRY(L("delta", [&] {
      if (rest_in_delta)
        k_delta<false, true, true><<<nl, bs, 0, st>>>(k, c.pred[0], c.pred[1], nv, cv,
                                                      c.planes, c.posB, c.posA, v1, c.rho,
                                                      nullptr, nullptr, c.flags);
      else
        k_delta<false, true, false><<<nl, bs, 0, st>>>(k, c.pred[0], c.pred[1], nv, cv,
                                                       c.planes, c.posB, c.posA, v1, c.rho,
                                                       nullptr, nullptr, c.flags);
      fill(c.posA, c.posB);
      fill(v1, nullptr);
    }));
  }
  if (c.xsph) {
    const bool rest = c.restitution_any && !c.vort;
    FH_TRY(L("xsph", [&] {
      if (tiles) {
        fork();
        (void)L.on(tst, "tile_xsph", [&] {
          (void)enqueue_tile_xsph(c, k, ta, rest, rest && rec ? rec->RM : nullptr, tst);
        });
      }
      if (rest)
        k_xsph<true><<<nl, bs, 0, st>>>(k, c.posA, c.velB, c.velA, nv, cv,
                                        c.planes, rec ? rec->RM : nullptr, c.flags);
      else
        k_xsph<false><<<nl, bs, 0, st>>>(k, c.posA, c.velB, c.velA, nv, cv,
                                         c.planes, nullptr, c.flags);
      fill(c.velA, nullptr);
      if (tiles) join();
    }));
    FH_TRY(fj);
  }
  if (c.vort) {
    FH_TRY(L("vorticity_omega", [&] {
      if (tiles) {
        fork();
        (void)L.on(tst, "tile_vorticity_omega",
                   [&] { (void)enqueue_tile_vort_omega(c, k, ta, tst); });
      }
      k_vort_omega<<<nl, bs, 0, st>>>(k, c.posA, c.velA, c.omega, nv, cv, c.flags);
      fill(c.omega, nullptr);
      if (tiles) join();
    }));
    FH_TRY(fj);
    FH_TRY(L("vorticity_apply", [&] {
      if (tiles) {
        fork();
        (void)L.on(tst, "tile_vorticity_apply",
                   [&] { (void)enqueue_tile_vort_apply(c, k, ta, c.restitution_any, tst); });
      }
      if (c.restitution_any)
        k_vort_apply<true><<<nl, bs, 0, st>>>(k, c.posA, c.omega, c.velA, nv, cv,
                                              c.planes, c.flags);
      else
        k_vort_apply<false><<<nl, bs, 0, st>>>(k, c.posA, c.omega, c.velA, nv, cv,
                                               c.planes, c.flags);
      fill(c.velA, nullptr);
      if (tiles) join();
    }));
    FH_TRY(fj);
  }
#undef FH_TRY
  return hipSuccess;
}

}  // namespace fluidhip
