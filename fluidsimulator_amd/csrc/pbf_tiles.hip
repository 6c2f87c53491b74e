// Brick-tiled gather passes: the sparse part of a scene through LDS.
//
// The cell table orders cells in 4x4x4-cell bricks (brick_shift 2), so a
// brick's particles are one contiguous slot range and the 6x6x6 cells around
// it (the brick and a one-cell halo) hold every candidate of those particles.
// A brick of enough particles (kTileMinParticles in large batches) is a tile
// when its whole halo holds at most kTileMaxHalo particles, one of its cells
// is good and, found by the neighbor kernel, none of its particles has more
// neighbors than a tile list holds (kTileList, or kTileListSmall in small
// batches).  A cell is good when its 27-cell neighborhood holds fewer than
// kTileMaxCand candidates and, in large batches, no two of its particles
// coincide (the per-lane passes compute a run of coincident particles once).
// The tile computes the particles of its good cells; the other cells of the
// brick go to the per-lane passes, their particles staged in the halo like
// any other.  For a tile:
//
//   * classify (once per substep): records the tile, the LDS offset of each
//     halo cell (the cells' particle counts, scanned in halo order z, y, x)
//     and each halo cell's first slot, and marks the good cells' slots so
//     that the lane map of the per-lane passes leaves them out;
//   * neighbors (once per substep; in small batches before the lane map, in
//     large ones, whose lists cannot overflow, beside the lanes' lists):
//     stages the halo's p^0 in LDS and writes each particle's neighbor list
//     as LDS records of the halo: the candidates are the 27 cells in
//     (dz, dy, dx) order, each
//     by ascending slot (= ascending id), and candidate j is a neighbor when
//     it is not the particle itself and |p^0_i - p^0_j|^2 < h^2 (near0, the
//     reference's list predicate, core.cpp:205-247).  A tile with a particle
//     of more neighbors than its list holds goes back to the per-lane passes;
//   * each gather pass (lambda, delta, XSPH, omega, eta): stages the halo's
//     records of the pass in LDS and visits every particle's list in order,
//     i.e. the reference's list order, with the per-lane term arithmetic of
//     pbf_kernels.hip, so every sum is bit-identical.
//
// In halo order the three cells (dx = -1, 0, 1) of one (dz, dy) row are
// consecutive, so a particle's 27 cells are 9 contiguous LDS segments.  A
// list entry is 16 bits (an LDS record); lists are chunks of 8 entries, chunk
// q of every particle together, so the loads of one pass coalesce and one
// 16-byte load serves 8 visits.
//
// Every tile kernel is persistent: a fixed grid loops over the tiles the
// classify kernel found (a device count), so no launch depends on host-side
// knowledge of the tile count.

#include <climits>

#include "pbf_ctx.h"
#include "pbf_device.h"

namespace fluidhip {

namespace {

constexpr int kEdge = 4;           // cells per tile edge (brick_shift 2)
constexpr int kHalo = 6;           // cells per halo edge
constexpr int kHaloCells = 216;
constexpr int kThreads = 256;      // workgroup of the classify kernel
constexpr int kPass = 512;         // workgroup of the neighbor and gather kernels
constexpr int kPerThread = kTileMaxHalo / kPass;  // halo records staged per thread
// In batches of kForkMinParticles particles or more, bricks with fewer
// particles take the per-lane passes: a tile's fixed cost (layout, halo
// staging, a 512-thread workgroup) is not worth it for a few particles of
// sparse spray, of which such batches can hold 10^5 bricks.
constexpr int kTileMinParticles = 32;

__device__ __forceinline__ bool tile_fatal(const unsigned* flags) {
  return (*flags & kFlagFatal) != 0u;
}

// Bitwise equal positions.
__device__ __forceinline__ bool coincide(const float3& a, const float3& b) {
  return ((__float_as_uint(a.x) ^ __float_as_uint(b.x)) | (__float_as_uint(a.y) ^ __float_as_uint(b.y)) |
          (__float_as_uint(a.z) ^ __float_as_uint(b.z))) == 0u;
}

// Pair terms of the gather passes for a pair with r2 < h2.  Tiles run only
// when min_r2 >= 2^-96 and sqrt(max(pred(h2), min_r2)) <= h (tile_terms_exact
// in pbf_capi.hip), so for such a pair the clamped argument of the square
// root is a finite float >= 2^-96, where sqrt_core is the correctly rounded
// square root, and r <= h, where spiky's r > h test is false: tile_poly6 and
// tile_spiky return what poly6 and pair_spiky return (core.cpp:35-57 and the
// clamp of core.cpp:303), bit for bit, without their range tests.
__device__ __forceinline__ float tile_poly6(float r2, const StepConsts& c) {
  const float diff = c.h2 - r2;
  const float diff3 = diff * diff * diff;
  return c.w_coeff * diff3;
}

__device__ __forceinline__ float tile_spiky(float r2, const StepConsts& c) {
  // r2 < min_r2 ? min_r2 : r2 for 0 <= r2 < h2 and min_r2 > 0, where the
  // bit patterns order as the floats do (one integer max)
  const float r = sqrt_core(__int_as_float(max(__float_as_int(r2), __float_as_int(c.min_r2))));
  const float diff = c.h - r;
  return c.s_coeff * diff * diff;
}

// Scene of the brick whose first cell has key `k0` (scenes own whole bricks).
__device__ __forceinline__ int scene_of_key(int k0, const GridDesc* __restrict__ grid,
                                            int n_scenes) {
  int lo = 0, hi = n_scenes - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (grid[mid].cell_off <= k0)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Table coordinates of the first cell of brick b (of scene g).
__device__ __forceinline__ void brick_origin(int b, const GridDesc& g, int& rx, int& ry,
                                             int& rz) {
  const int local = b - (g.cell_off >> 6);
  rx = (local % g.nbx) * kEdge;
  ry = ((local / g.nbx) % g.nby) * kEdge;
  rz = (local / (g.nbx * g.nby)) * kEdge;
}

// Exclusive scan of v over a workgroup of kN threads (lds: kN/64 + 1 ints).
template <int kN = kThreads>
__device__ __forceinline__ int tile_excl_scan(int v, int* lds, int* total) {
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  int inc = v;
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(inc, o, 64);
    if (lane >= o) inc += u;
  }
  if (lane == 63) lds[w] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int i = 0; i < kN / 64; ++i) {
      const int t = lds[i];
      lds[i] = run;
      run += t;
    }
    lds[kN / 64] = run;
  }
  __syncthreads();
  const int r = lds[w] + inc - v;
  *total = lds[kN / 64];
  __syncthreads();
  return r;
}

// ------------------------------------------------------------ classify

// Non-empty bricks of at least min_particles particles, appended to
// tile_cand (tile_count[1] of them; one atomic per wavefront); the slots of
// other bricks, and of candidates beyond tile_cap, get in_tile = 0 here.
__global__ __launch_bounds__(kThreads) void k_tile_bricks(
    const int* __restrict__ start, const int* __restrict__ total_cells, int* __restrict__ in_tile,
    int* __restrict__ tile_count, int* __restrict__ tile_cand, int tile_cap, int min_particles,
    const unsigned* __restrict__ flags) {
  if (tile_fatal(flags)) return;
  const int nbricks = *total_cells >> 6;
  const int lane = threadIdx.x & 63;
  // whole wavefronts iterate together (the ballot needs every lane)
  for (int b0 = (blockIdx.x * blockDim.x + threadIdx.x) - lane; b0 < nbricks;
       b0 += gridDim.x * blockDim.x) {
    const int b = b0 + lane;
    bool big = false;
    if (b < nbricks) {
      const int a = start[b << 6];
      const int e = start[(b << 6) + 64];
      big = e > a && e - a >= min_particles;
      if (!big)
        for (int s = a; s < e; ++s) in_tile[s] = 0;
    }
    const unsigned long long m = __ballot(big);
    if (m == 0ull) continue;
    int base = 0;
    if (lane == 0) base = atomicAdd(tile_count + 1, __popcll(m));
    base = __shfl(base, 0, 64);
    const int at = base + __popcll(m & ((1ull << lane) - 1ull));
    if (big && at < tile_cap) {
      tile_cand[at] = b;
    } else if (big) {  // beyond the capacity: the per-lane passes
      for (int s = start[b << 6]; s < start[(b << 6) + 64]; ++s) in_tile[s] = 0;
    }
  }
}

// Examines the candidate bricks (one workgroup at a time, persistent grid):
// a tile gets a record, its halo layout and in_tile[s] = 1 + its index for
// the slots of its good cells; every other slot of a candidate gets
// in_tile = 0.
__global__ __launch_bounds__(kThreads) void k_tile_classify(
    StepConsts c, const GridDesc* __restrict__ grid, const int* __restrict__ start,
    const float4* __restrict__ posB, const int* __restrict__ tile_cand,
    int* __restrict__ in_tile, int* __restrict__ tile_count, int4* __restrict__ tile_rec,
    int* __restrict__ tile_seg, int* __restrict__ tile_src, int tile_cap, int reject_runs,
    const unsigned* __restrict__ flags) {
  __shared__ int hcnt[kHaloCells];
  __shared__ int hsrc[kHaloCells];
  __shared__ int scan_lds[kThreads / 64 + 1];
  __shared__ int cst[65];   // the brick's cell starts (slots)
  __shared__ int cbad[64];  // per brick cell: left to the per-lane passes
  __shared__ int idx_sh;
  if (tile_fatal(flags)) return;
  const int ncand = min(tile_count[1], tile_cap);
  const int t = threadIdx.x;
  for (int j = blockIdx.x; j < ncand; j += gridDim.x) {
    const int b = tile_cand[j];
    const int a = start[b << 6];
    const int e = start[(b << 6) + 64];
    const int sc = scene_of_key(b << 6, grid, c.n_scenes);
    const GridDesc g = grid[sc];
    int ox, oy, oz;
    brick_origin(b, g, ox, oy, oz);
    int cnt = 0, src = 0;
    if (t < kHaloCells) {
      const int rx = ox - 1 + t % kHalo, ry = oy - 1 + (t / kHalo) % kHalo,
                rz = oz - 1 + t / (kHalo * kHalo);
      if (rx >= 0 && ry >= 0 && rz >= 0 && rx < kEdge * g.nbx && ry < kEdge * g.nby &&
          rz < kEdge * g.nbz) {
        const int k = cell_key(rx, ry, rz, g);
        src = start[k];
        cnt = start[k + 1] - src;
      }
      hcnt[t] = cnt;
      hsrc[t] = src;
    }
    if (t <= 64) cst[t] = start[(b << 6) + t];
    __syncthreads();
    int total;
    const int off = tile_excl_scan(cnt, scan_lds, &total);
    // a cell whose 27-cell neighborhood holds kTileMaxCand candidates or more
    // (brick cell t is the cell of key (b << 6) + t: Z-order, bits
    // x0 y0 z0 x1 y1 z1)
    if (t < 64) {
      const int hx = ((t & 1) | ((t >> 2) & 2)) + 1;
      const int hy = (((t >> 1) & 1) | ((t >> 3) & 2)) + 1;
      const int hz = (((t >> 2) & 1) | ((t >> 4) & 2)) + 1;
      int cand = 0;
      for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dx = -1; dx <= 1; ++dx)
            cand += hcnt[((hz + dz) * kHalo + hy + dy) * kHalo + hx + dx];
      cbad[t] = cand >= kTileMaxCand;
    }
    __syncthreads();
    // coincident particles (consecutive slots at one position, so in one
    // cell): in large batches the per-lane passes, which compute a run of
    // them once, take their cell.  Only a brick that can still be a tile
    // (at most kTileMaxHalo slots) is scanned; each thread keeps the cell of
    // its slot (cells lie in slot order) and jumps over excluded cells.
    const bool maybe =
        __syncthreads_or(t < 64 && cst[t + 1] > cst[t] && !cbad[t]) && total <= kTileMaxHalo;
    if (reject_runs && maybe) {
      int q = 0;
      for (int s = a + 1 + t; s < e; s += kThreads) {
        while (cst[q + 1] <= s) ++q;
        if (cbad[q]) {  // left to the per-lane passes: on to the next cell
          s += (cst[q + 1] - s - 1) / kThreads * kThreads;
          continue;
        }
        if (coincide(ld3(posB, s), ld3(posB, s - 1))) cbad[q] = 1;
      }
    }
    __syncthreads();
    const bool good = t < 64 && cst[t + 1] > cst[t] && !cbad[t];
    const bool ok = __syncthreads_or(good) && maybe;
    if (ok && t == 0) idx_sh = atomicAdd(tile_count, 1);
    __syncthreads();
    const bool tile = ok && idx_sh < tile_cap;
    // a tile with excluded cells marks its slots cell by cell
    const bool mixed = __syncthreads_or(tile && t < 64 && cst[t + 1] > cst[t] && cbad[t]);
    if (tile) {
      const int i = idx_sh;
      if (t == 0) {
        tile_rec[2 * i] = make_int4(a, e, total, sc);
        tile_rec[2 * i + 1] = make_int4(g.ox + ox, g.oy + oy, g.oz + oz, b);
      }
      if (t < kHaloCells) {
        tile_seg[i * (kHaloCells + 1) + t] = off;
        tile_src[i * kHaloCells + t] = hsrc[t];
      }
      if (t == 0) tile_seg[i * (kHaloCells + 1) + kHaloCells] = total;
    }
    if (mixed) {
      int q = 0;
      for (int s = a + t; s < e; s += kThreads) {
        while (cst[q + 1] <= s) ++q;
        in_tile[s] = cbad[q] ? 0 : idx_sh + 1;
      }
    } else {
      for (int s = a + t; s < e; s += kThreads) in_tile[s] = tile ? idx_sh + 1 : 0;
    }
    __syncthreads();  // LDS is reused by the next brick
  }
}

// ------------------------------------------------------------ shared walk

// A tile's halo layout in LDS: seg[c] = first LDS record of halo cell c
// (seg[216] = records, seg[217..255] = INT_MAX so that a binary search over
// 256 entries needs no bounds test), src[c] = its first slot.
struct TileLayout {
  int seg[256];
  int src[kHaloCells];
};

__device__ __forceinline__ void load_layout(TileLayout& L, const int* __restrict__ tile_seg,
                                            const int* __restrict__ tile_src, int i) {
  const int t = threadIdx.x;
  if (t < 256) {
    const int seg = t <= kHaloCells ? tile_seg[i * (kHaloCells + 1) + t] : INT_MAX;
    const int src = t < kHaloCells ? tile_src[i * kHaloCells + t] : 0;
    L.seg[t] = seg;
    if (t < kHaloCells) L.src[t] = src;
  }
}

// Cell of slot s inside its tile: (lx, ly, lz) in [0, 4)^3 (cx0..cz0: the
// tile's first cell).
__device__ __forceinline__ void local_cell(uint64_t cell, int cx0, int cy0, int cz0, int& lx,
                                           int& ly, int& lz) {
  int x, y, z;
  unpack_cell(cell, x, y, z);
  lx = x - cx0;
  ly = y - cy0;
  lz = z - cz0;
}

// Stages the tile's halo in LDS: store(r, fetch(slot)) for every record
// r < nrec.  The cell searches of a thread's records run interleaved, then
// every fetch is issued before the first LDS store, so one thread's records
// are one round of loads.
template <typename Fetch, typename Store>
__device__ __forceinline__ void stage_halo(const TileLayout& L, int nrec, Fetch&& fetch,
                                           Store&& store) {
  int cell[kPerThread];
#pragma unroll
  for (int k = 0; k < kPerThread; ++k) cell[k] = 0;
  const int r0 = static_cast<int>(threadIdx.x);
  // the last cell c with seg[c] <= r (seg is non-decreasing, seg[0] = 0)
#pragma unroll
  for (int step = 128; step > 0; step >>= 1) {
#pragma unroll
    for (int k = 0; k < kPerThread; ++k) {
      const int mid = cell[k] + step;
      cell[k] = L.seg[mid] <= r0 + k * kPass ? mid : cell[k];
    }
  }
  int slot[kPerThread];
#pragma unroll
  for (int k = 0; k < kPerThread; ++k) {
    const int r = r0 + k * kPass;
    // records past nrec fetch slot 0 (some particle of the tile) and are not stored
    slot[k] = r < nrec ? L.src[cell[k]] + (r - L.seg[cell[k]]) : 0;
  }
  decltype(fetch(0)) v[kPerThread];
#pragma unroll
  for (int k = 0; k < kPerThread; ++k) v[k] = fetch(slot[k]);
#pragma unroll
  for (int k = 0; k < kPerThread; ++k)
    if (r0 + k * kPass < nrec) store(r0 + k * kPass, v[k]);
}

// Calls visit(r) for the LDS records r of a tile particle's cnt neighbors,
// in list (= candidate) order.  A list is chunks of 8 16-bit entries (one
// uint4, chunk-major: chunk q of every particle together, so one load of a
// pass coalesces); each chunk is loaded one chunk ahead, and the 8 visits of
// a chunk are unrolled (a visit that no lane of the wavefront needs is
// skipped as a whole).
template <typename Visit>
__device__ __forceinline__ void tile_walk(const uint4* __restrict__ list, int n, int s, int cnt,
                                          uint4 first, Visit&& visit) {
  uint4 q = first;
  for (int k = 0; k < cnt; k += 8) {
    const uint4 nq = k + 8 < cnt ? list[static_cast<size_t>((k >> 3) + 1) * n + s]
                                 : make_uint4(0u, 0u, 0u, 0u);
    const unsigned w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int e = 0; e < 8; ++e)
      if (k + e < cnt) visit(static_cast<int>((w[e >> 1] >> (16 * (e & 1))) & 0xffffu));
    q = nq;
  }
}

// A tile's slot range [a, e), halo records, scene and first cell; dead: the
// tile went back to the per-lane passes (a particle with more neighbors than
// a tile list holds).
struct TileInfo {
  int a, e, nrec, sc;
  int cx0, cy0, cz0;
  bool dead;
};

__device__ __forceinline__ TileInfo tile_info(const int4* __restrict__ tile_rec, int i) {
  const int4 r0 = tile_rec[2 * i], r1 = tile_rec[2 * i + 1];
  return TileInfo{r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w < 0};
}

__device__ __forceinline__ int tiles_of(const int* tile_count, int tile_cap) {
  const int t = *tile_count;
  return t < tile_cap ? t : tile_cap;
}

// A tile particle's neighbor count (-1 for a slot of the tile's brick that
// the per-lane passes compute) and first list chunk.
struct Nbrs {
  int cnt;
  uint4 first;  // the list's first chunk
};

__device__ __forceinline__ Nbrs load_nbrs(const TileArgs& ta, int n, int s) {
  return Nbrs{ta.tile_cnt[s], ta.tile_list[s]};
}

// The tile loop of every gather kernel.  Each workgroup takes tiles in turn
// (with xcd set, consecutive tiles, which share halo cells, go to workgroups
// of one XCD); per tile it
//   1. loads the tile record and the halo layout (one round of loads),
//   2. issues the halo gather (stage_halo's fetch and store) together with the
//      loads of its first particle: load(s), the neighbor count and the first
//      list chunk (a second round),
//   3. runs body(s, tile, nbrs, x) for each particle s of the tile.
template <typename X, typename Load, typename Fetch, typename Store, typename Body>
__device__ __forceinline__ void tile_loop(const TileArgs& ta, int n, int xcd, TileLayout& L,
                                          Load&& load, Fetch&& fetch, Store&& store,
                                          Body&& body) {
  const int nt = tiles_of(ta.tile_count, ta.tile_cap);
  const int per = (nt + 7) >> 3;
  const int qend = xcd ? per << 3 : nt;
  for (int q = blockIdx.x; q < qend; q += gridDim.x) {
    const int i = xcd ? (q & 7) * per + (q >> 3) : q;
    if (i >= nt) continue;  // uniform over the workgroup
    const TileInfo T = tile_info(ta.tile_rec, i);
    if (T.dead) continue;
    load_layout(L, ta.tile_seg, ta.tile_src, i);
    int s = T.a + static_cast<int>(threadIdx.x);
    // a tile holds at least one particle: lanes past its end load its last
    // particle's data (in bounds) and do not use it
    const int s0 = s < T.e ? s : T.e - 1;
    Nbrs nb = load_nbrs(ta, n, s0);
    X x = load(s0);
    __syncthreads();
    stage_halo(L, T.nrec, fetch, store);
    __syncthreads();
    while (s < T.e) {
      if (nb.cnt >= 0) body(s, T, nb, x);  // count -1: a cell of the per-lane passes
      s += kPass;
      if (s < T.e) {
        nb = load_nbrs(ta, n, s);
        x = load(s);
      }
    }
    __syncthreads();  // LDS is reused by the next tile
  }
}

// ------------------------------------------------------------ neighbors

// Neighbor lists of the tiles (core.cpp:205-247): each particle's neighbors
// as LDS records of the tile's halo, in the 27 cells' (dz, dy, dx) order,
// each cell by ascending slot (= ascending id); neighbors are the other
// candidates j with near0(p^0_i, p^0_j).  A tile with a particle of more
// neighbors than its list holds is dead: its slots go back to the per-lane passes
// (in_tile = 0) and the passes skip it.  The slots of the brick's other
// cells get the count -1.  With nbr (a recorded substep) the
// live tiles' lists are also stored there as slots (noff, nwords), as the
// per-lane neighbor kernel stores its lists.
// A halo record's p^0 and slot.
struct RecSlot {
  float3 p;
  int t;
};

template <bool kStore>
__global__ __launch_bounds__(kPass) void k_tile_neighbors(
    StepConsts c, TileArgs ta, const float4* __restrict__ p0, int4* __restrict__ tile_rec,
    int* __restrict__ in_tile, int* __restrict__ ncount, long long* __restrict__ noff,
    int* __restrict__ nwords, int* __restrict__ nbr, long long nbr_cap,
    unsigned long long* __restrict__ nbr_total, unsigned* __restrict__ flags) {
  __shared__ TileLayout L;
  __shared__ float X[kTileMaxHalo], Y[kTileMaxHalo], Z[kTileMaxHalo];
  __shared__ int Sl[kStore ? kTileMaxHalo : 1];  // each record's slot (stored lists)
  __shared__ int scan_lds[kPass / 64 + 1];
  __shared__ long long base_sh;
  if (tile_fatal(flags)) return;
  const int nt = tiles_of(ta.tile_count, ta.tile_cap);
  const int per = (nt + 7) >> 3;
  const int qend = c.xcd_remap ? per << 3 : nt;
  uint4* list = const_cast<uint4*>(ta.tile_list);
  // nbr_total counts this substep's stored words from zero (grid setup)
  int* cnt_out = const_cast<int*>(ta.tile_cnt);
  for (int q = blockIdx.x; q < qend; q += gridDim.x) {
    const int i = c.xcd_remap ? (q & 7) * per + (q >> 3) : q;
    if (i >= nt) continue;
    const TileInfo T = tile_info(tile_rec, i);
    load_layout(L, ta.tile_seg, ta.tile_src, i);
    __syncthreads();
    stage_halo(L, T.nrec, [&](int t) { return RecSlot{ld3(p0, t), t}; },
               [&](int r, const RecSlot& v) {
                 X[r] = v.p.x;
                 Y[r] = v.p.y;
                 Z[r] = v.p.z;
                 if (kStore) Sl[r] = v.t;
               });
    __syncthreads();
    int over = 0;
    for (int s = T.a + static_cast<int>(threadIdx.x); s < T.e; s += kPass) {
      if (in_tile[s] == 0) {  // a cell of the per-lane passes: they write its lists
        cnt_out[s] = -1;
        continue;
      }
      int lx, ly, lz;
      local_cell(ta.slot_cell[s], T.cx0, T.cy0, T.cz0, lx, ly, lz);
      const int own = ((lz + 1) * kHalo + ly + 1) * kHalo + lx + 1;
      const int self = L.seg[own] + (s - L.src[own]);
      const float3 me = make_float3(X[self], Y[self], Z[self]);
      // entries enter a 128-bit shift register w3:w2:w1:w0 at the top, 16
      // bits at a time, so after 8 of them entry 0 is w0's low half; a chunk
      // goes out every 8 entries
      int k = 0;
      unsigned w0 = 0u, w1 = 0u, w2 = 0u, w3 = 0u;
      for (int row = 0; row < 9; ++row) {
        const int hc = ((lz + row / 3) * kHalo + ly + row % 3) * kHalo + lx;
        const int lo = L.seg[hc], hi = L.seg[hc + 3];
        // 32 candidates at a time: first the neighbor tests as a bit mask,
        // then the mask's bits in ascending order (the candidates' order);
        // the insertion runs once per neighbor, not once per candidate
        for (int b = lo; b < hi; b += 32) {
          const int nb = min(32, hi - b);
          unsigned m = 0u;
          for (int j = 0; j < nb; ++j) {
            const int r = b + j;
            if (near0(me, make_float3(X[r], Y[r], Z[r]), c)) m |= 1u << j;
          }
          if (self >= b && self < b + 32) m &= ~(1u << (self - b));
          while (m != 0u) {
            const int r = b + __builtin_ctz(m);
            m &= m - 1u;
            w0 = __builtin_amdgcn_alignbit(w1, w0, 16);
            w1 = __builtin_amdgcn_alignbit(w2, w1, 16);
            w2 = __builtin_amdgcn_alignbit(w3, w2, 16);
            w3 = (w3 >> 16) | (static_cast<unsigned>(r) << 16);
            if ((k & 7) == 7 && k < ta.list_cap)
              list[static_cast<size_t>(k >> 3) * c.n + s] = make_uint4(w0, w1, w2, w3);
            ++k;
          }
        }
      }
      if ((k & 7) != 0 && k < ta.list_cap) {
        // the last m < 8 entries sit in the top m halves: shift them down
        for (int m = k & 7; m < 8; ++m) {
          w0 = __builtin_amdgcn_alignbit(w1, w0, 16);
          w1 = __builtin_amdgcn_alignbit(w2, w1, 16);
          w2 = __builtin_amdgcn_alignbit(w3, w2, 16);
          w3 >>= 16;
        }
        list[static_cast<size_t>(k >> 3) * c.n + s] = make_uint4(w0, w1, w2, w3);
      }
      cnt_out[s] = k;
      ncount[s] = k;
      nwords[s] = -1;  // no stored list: the tile passes read their own
      noff[s] = 0;
      over |= k > ta.list_cap;
    }
    if (__syncthreads_or(over)) {
      for (int s = T.a + static_cast<int>(threadIdx.x); s < T.e; s += kPass) in_tile[s] = 0;
      if (threadIdx.x == 0) tile_rec[2 * i + 1].w = -1;
    } else if (kStore) {
      // a recorded substep: the lists also go to the shared list buffer as
      // plain slot words, in the same order, for the adjoint (each list read
      // back from its chunks, each record mapped to its slot); one
      // allocation per tile
      int mine = 0;
      for (int s = T.a + static_cast<int>(threadIdx.x); s < T.e; s += kPass)
        mine += max(cnt_out[s], 0);
      int total;
      const int pre = tile_excl_scan<kPass>(mine, scan_lds, &total);
      if (threadIdx.x == 0)
        base_sh = static_cast<long long>(
            atomicAdd(nbr_total, static_cast<unsigned long long>(total)));
      __syncthreads();
      long long off = base_sh + pre;
      if (base_sh + total > nbr_cap) {
        if (threadIdx.x == 0) atomicOr(flags, kFlagNeighbors);
      } else {
        for (int s = T.a + static_cast<int>(threadIdx.x); s < T.e; s += kPass) {
          const int cnt = cnt_out[s];
          if (cnt < 0) continue;  // not the tile's
          for (int q8 = 0; q8 < cnt; q8 += 8) {
            const uint4 ch = list[static_cast<size_t>(q8 >> 3) * c.n + s];
            const unsigned w[4] = {ch.x, ch.y, ch.z, ch.w};
#pragma unroll
            for (int e = 0; e < 8; ++e)
              if (q8 + e < cnt) nbr[off + q8 + e] = Sl[(w[e >> 1] >> (16 * (e & 1))) & 0xffffu];
          }
          noff[s] = off;
          nwords[s] = cnt;
          off += cnt;
        }
      }
    }
    __syncthreads();  // LDS is reused by the next tile
  }
}

// ------------------------------------------------------------ lambda

// lambda pass (core.cpp:281-329) of the tiles; the per-lane term and its
// accumulation order of k_lambda.  A recorded substep also stores the
// gradient sum S and the denominator sum Q (sq_out).
__global__ __launch_bounds__(kPass) void k_tile_lambda(StepConsts c, TileArgs ta,
                                                       float4* __restrict__ P,
                                                       float* __restrict__ rho_out,
                                                       int write_rho,
                                                       float4* __restrict__ sq_out,
                                                       const unsigned* __restrict__ flags) {
  __shared__ TileLayout L;
  __shared__ float X[kTileMaxHalo], Y[kTileMaxHalo], Z[kTileMaxHalo];
  if (tile_fatal(flags)) return;
  const float ngs = -c.grad_scale;
  tile_loop<float3>(
      ta, c.n, c.xcd_remap, L, [&](int s) { return ld3(P, s); },
      [&](int t) { return ld3(P, t); },
      [&](int r, const float3& q) {
        X[r] = q.x;
        Y[r] = q.y;
        Z[r] = q.z;
      },
      [&](int s, const TileInfo&, const Nbrs& nb, const float3& p) {
        float rho = 0.0f, gsx = 0.0f, gsy = 0.0f, gsz = 0.0f, sg2 = 0.0f;
        tile_walk(ta.tile_list, c.n, s, nb.cnt, nb.first, [&](int r) {
          const float dx = p.x - X[r];
          const float dy = p.y - Y[r];
          const float dz = p.z - Z[r];
          const float r2 = dx * dx + dy * dy + dz * dz;
          // A coincident neighbor (r2 = 0) adds gradient terms of +-0, which
          // leave the sums as they are: none of them is ever -0 (each starts
          // at +0 and a rounded sum is -0 only of two -0).
          if (r2 < c.h2) {
            rho += tile_poly6(r2, c);
            const float f = tile_spiky(r2, c);
            const float gx = f * dx, gy = f * dy, gz = f * dz;
            const float jx = ngs * gx, jy = ngs * gy, jz = ngs * gz;
            gsx += gx;
            gsy += gy;
            gsz += gz;
            sg2 += jx * jx + jy * jy + jz * jz;
          } else {
            rho += r2 > c.h2 ? 0.0f : tile_poly6(r2, c);  // poly6: +0, or NaN
          }
        });
        rho += c.w0;
        rho *= c.mass;
        const float C = rho * c.inv_density - 1.0f;
        const float gix = c.grad_scale * gsx, giy = c.grad_scale * gsy,
                    giz = c.grad_scale * gsz;
        sg2 += gix * gix + giy * giy + giz * giz;
        reinterpret_cast<float*>(P + s)[3] = -C / (sg2 + c.eps);
        if (write_rho) rho_out[s] = rho;
        if (sq_out) sq_out[s] = make_float4(gsx, gsy, gsz, sg2);
      });
}

// ------------------------------------------------------------ delta

// delta pass with plane projection and apply (core.cpp:334-407) of the
// tiles; kLast also commits (core.cpp:414-421), kRest fuses restitution.  A
// recorded substep also stores the planes that pushed (pmask) and the
// restitution decisions (rmask).
template <bool kLast, bool kRest>
__global__ __launch_bounds__(kPass) void k_tile_delta(
    StepConsts c, TileArgs ta, const float4* __restrict__ P, float4* __restrict__ Pnext,
    const ScenePlanes* __restrict__ sp,
    const float4* __restrict__ posB, float4* __restrict__ posA, float4* __restrict__ vout,
    const float* __restrict__ rho, unsigned* __restrict__ pmask, unsigned* __restrict__ rmask,
    const unsigned* __restrict__ flags) {
  __shared__ TileLayout L;
  __shared__ float X[kTileMaxHalo], Y[kTileMaxHalo], Z[kTileMaxHalo], W[kTileMaxHalo];
  if (tile_fatal(flags)) return;
  tile_loop<float4>(
      ta, c.n, c.xcd_remap, L, [&](int s) { return P[s]; },
      [&](int t) { return P[t]; },
      [&](int r, const float4& q) {
        X[r] = q.x;
        Y[r] = q.y;
        Z[r] = q.z;
        W[r] = q.w;
      },
      [&](int s, const TileInfo& T, const Nbrs& nb, const float4& pi) {
        float ddx = 0.0f, ddy = 0.0f, ddz = 0.0f;
        tile_walk(ta.tile_list, c.n, s, nb.cnt, nb.first, [&](int r) {
          const float dx = pi.x - X[r];
          const float dy = pi.y - Y[r];
          const float dz = pi.z - Z[r];
          const float r2 = dx * dx + dy * dy + dz * dz;
          if (!(r2 < c.h2)) return;
          const float f = tile_spiky(r2, c);
          float sv = pi.w + W[r];
          if (c.scorr_on) {
            const float Wr = tile_poly6(r2, c);
            const float ratio = Wr * c.scorr_inv_wdq;
            sv += c.scorr_neg_k * pow_ratio(ratio, c.scorr_n);
          }
          const float sf = sv * f;
          ddx += sf * dx;
          ddy += sf * dy;
          ddz += sf * dz;
        });
        ddx *= c.inv_density;
        ddy *= c.inv_density;
        ddz *= c.inv_density;
        const ScenePlanes& plane = sp[T.sc];
        unsigned pushed = 0;
        if (plane.n_planes > 0) {
          float qx = pi.x + ddx, qy = pi.y + ddy, qz = pi.z + ddz;
          pushed = project_planes(qx, qy, qz, plane);
          ddx = qx - pi.x;
          ddy = qy - pi.y;
          ddz = qz - pi.z;
        }
        if (pmask) pmask[s] = pushed;
        const float nx = pi.x + ddx, ny = pi.y + ddy, nz = pi.z + ddz;  // core.cpp:404-406
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
          const unsigned rm = restitute(vx, vy, vz, nx, ny, nz, plane);
          if (rmask) rmask[s] = rm;
        }
        vout[s] = make_float4(vx, vy, vz, rho[s]);
      });
}

// ------------------------------------------------------------ XSPH

struct PosVel {
  float3 x;
  float4 v;
};

// XSPH (core.cpp:430-465) of the tiles; vin.w carries rho_j.
template <bool kRest>
__global__ __launch_bounds__(kPass) void k_tile_xsph(
    StepConsts c, TileArgs ta, const float4* __restrict__ pos, const float4* __restrict__ vin,
    float4* __restrict__ vout, const ScenePlanes* __restrict__ sp, unsigned* __restrict__ rmask,
    const unsigned* __restrict__ flags) {
  __shared__ TileLayout L;
  __shared__ float X[kTileMaxHalo], Y[kTileMaxHalo], Z[kTileMaxHalo];
  __shared__ float U[kTileMaxHalo], V[kTileMaxHalo], Q[kTileMaxHalo], M[kTileMaxHalo];
  if (tile_fatal(flags)) return;
  tile_loop<PosVel>(
      ta, c.n, c.xcd_remap, L, [&](int s) { return PosVel{ld3(pos, s), vin[s]}; },
      [&](int t) { return PosVel{ld3(pos, t), vin[t]}; },
      [&](int r, const PosVel& pv) {
        const float3 q = pv.x;
        const float4 v = pv.v;
        X[r] = q.x;
        Y[r] = q.y;
        Z[r] = q.z;
        U[r] = v.x;
        V[r] = v.y;
        Q[r] = v.z;
        M[r] = (v.w > 0.0f) ? (c.mass / v.w) : 0.0f;
      },
      [&](int s, const TileInfo& T, const Nbrs& nb, const PosVel& me) {
        const float3 xi = me.x;
        const float4 vi = me.v;
        float ax = 0.0f, ay = 0.0f, az = 0.0f;
        tile_walk(ta.tile_list, c.n, s, nb.cnt, nb.first, [&](int r) {
          const float dx = xi.x - X[r];
          const float dy = xi.y - Y[r];
          const float dz = xi.z - Z[r];
          const float r2 = dx * dx + dy * dy + dz * dz;
          if (!(r2 < c.h2)) return;
          const float Wr = tile_poly6(r2, c);
          ax += (U[r] - vi.x) * Wr * M[r];
          ay += (V[r] - vi.y) * Wr * M[r];
          az += (Q[r] - vi.z) * Wr * M[r];
        });
        float vx = vi.x + c.visc_c * ax;
        float vy = vi.y + c.visc_c * ay;
        float vz = vi.z + c.visc_c * az;
        if (kRest) {
          const unsigned rm = restitute(vx, vy, vz, xi.x, xi.y, xi.z, sp[T.sc]);
          if (rmask) rmask[s] = rm;
        }
        vout[s] = make_float4(vx, vy, vz, 0.0f);
      });
}

// ------------------------------------------------------------ vorticity

struct PosVel3 {
  float3 x, v;
};

// omega and |omega| (core.cpp:472-508) of the tiles.
__global__ __launch_bounds__(kPass) void k_tile_vort_omega(
    StepConsts c, TileArgs ta, const float4* __restrict__ pos, const float4* __restrict__ vel,
    float4* __restrict__ omega, const unsigned* __restrict__ flags) {
  __shared__ TileLayout L;
  __shared__ float X[kTileMaxHalo], Y[kTileMaxHalo], Z[kTileMaxHalo];
  __shared__ float U[kTileMaxHalo], V[kTileMaxHalo], Q[kTileMaxHalo];
  if (tile_fatal(flags)) return;
  tile_loop<PosVel3>(
      ta, c.n, c.xcd_remap, L, [&](int s) { return PosVel3{ld3(pos, s), ld3(vel, s)}; },
      [&](int t) { return PosVel3{ld3(pos, t), ld3(vel, t)}; },
      [&](int r, const PosVel3& pv) {
        const float3 q = pv.x, v = pv.v;
        X[r] = q.x;
        Y[r] = q.y;
        Z[r] = q.z;
        U[r] = v.x;
        V[r] = v.y;
        Q[r] = v.z;
      },
      [&](int s, const TileInfo&, const Nbrs& nb, const PosVel3& me) {
        const float3 xi = me.x, vi = me.v;
        float ox = 0.0f, oy = 0.0f, oz = 0.0f;
        tile_walk(ta.tile_list, c.n, s, nb.cnt, nb.first, [&](int r) {
          const float dx = xi.x - X[r];
          const float dy = xi.y - Y[r];
          const float dz = xi.z - Z[r];
          const float r2 = dx * dx + dy * dy + dz * dz;
          if (!(r2 < c.h2)) return;
          const float f = tile_spiky(r2, c);
          const float gx = f * dx, gy = f * dy, gz = f * dz;
          const float dvx = U[r] - vi.x, dvy = V[r] - vi.y, dvz = Q[r] - vi.z;
          ox += dvy * gz - dvz * gy;
          oy += dvz * gx - dvx * gz;
          oz += dvx * gy - dvy * gx;
        });
        omega[s] = make_float4(ox, oy, oz, sqrtf(ox * ox + oy * oy + oz * oz));
      });
}

struct PosOmega {
  float3 x;
  float4 om;
};

// eta, confinement force and velocity update (core.cpp:512-570) of the
// tiles, in place on vel.
template <bool kRest>
__global__ __launch_bounds__(kPass) void k_tile_vort_apply(
    StepConsts c, TileArgs ta, const float4* __restrict__ pos, const float4* __restrict__ omega,
    float4* __restrict__ vel, const ScenePlanes* __restrict__ sp,
    const unsigned* __restrict__ flags) {
  __shared__ TileLayout L;
  __shared__ float X[kTileMaxHalo], Y[kTileMaxHalo], Z[kTileMaxHalo], O[kTileMaxHalo];
  if (tile_fatal(flags)) return;
  tile_loop<PosOmega>(
      ta, c.n, c.xcd_remap, L, [&](int s) { return PosOmega{ld3(pos, s), omega[s]}; },
      [&](int t) {
        return make_float4(ld3(pos, t).x, ld3(pos, t).y, ld3(pos, t).z,
                           reinterpret_cast<const float*>(omega + t)[3]);
      },
      [&](int r, const float4& q) {
        X[r] = q.x;
        Y[r] = q.y;
        Z[r] = q.z;
        O[r] = q.w;
      },
      [&](int s, const TileInfo& T, const Nbrs& nb, const PosOmega& me) {
        const float3 xi = me.x;
        const float4 om = me.om;
        float ex = 0.0f, ey = 0.0f, ez = 0.0f;
        tile_walk(ta.tile_list, c.n, s, nb.cnt, nb.first, [&](int r) {
          const float dx = xi.x - X[r];
          const float dy = xi.y - Y[r];
          const float dz = xi.z - Z[r];
          const float r2 = dx * dx + dy * dy + dz * dz;
          if (!(r2 < c.h2)) return;
          const float f = tile_spiky(r2, c);
          const float gx = f * dx, gy = f * dy, gz = f * dz;
          const float co = O[r] - om.w;
          ex += co * gx;
          ey += co * gy;
          ez += co * gz;
        });
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
        if (kRest) restitute(vx, vy, vz, xi.x, xi.y, xi.z, sp[T.sc]);
        vel[s] = make_float4(vx, vy, vz, 0.0f);
      });
}

}  // namespace

// ------------------------------------------------------------ launching

hipError_t enqueue_tile_classify(Ctx& c, const StepConsts& k) {
  // small batches tile every brick they can (coincident particles
  // included): their per-lane kernels then have no lanes, whose latency
  // costs each pass about 20 us for any number of them, and a few small
  // tiles cost little; large ones leave bricks of sparse spray and runs of
  // coincident particles to the per-lane passes
  const bool large = c.n >= kForkMinParticles;
  const int min_particles = large ? kTileMinParticles : 1;
  k_tile_bricks<<<1024, kThreads, 0, c.stream>>>(c.start, c.total_cells, c.in_tile, c.tile_count,
                                                c.tile_cand, static_cast<int>(c.tile_cap),
                                                min_particles, c.flags);
  k_tile_classify<<<kTileGrid, kThreads, 0, c.stream>>>(
      k, c.grid, c.start, c.posB, c.tile_cand, c.in_tile, c.tile_count, c.tile_rec, c.tile_seg,
      c.tile_src, static_cast<int>(c.tile_cap), large ? 1 : 0, c.flags);
  return hipGetLastError();
}

hipError_t enqueue_tile_neighbors(Ctx& c, const StepConsts& k, const TileArgs& ta,
                                  const float4* p0, bool store, hipStream_t st) {
  if (store)
    k_tile_neighbors<true><<<kTileGrid, kPass, 0, st>>>(k, ta, p0, c.tile_rec, c.in_tile,
                                                        c.ncount, c.noff, c.nwords, c.nbr,
                                                        c.nbr_cap, c.nbr_total, c.flags);
  else
    k_tile_neighbors<false><<<kTileGrid, kPass, 0, st>>>(k, ta, p0, c.tile_rec, c.in_tile,
                                                         c.ncount, c.noff, c.nwords, nullptr,
                                                         c.nbr_cap, c.nbr_total, c.flags);
  return hipGetLastError();
}

hipError_t enqueue_tile_lambda(Ctx& c, const StepConsts& k, const TileArgs& ta, float4* P,
                               int write_rho, float4* sq, hipStream_t st) {
  k_tile_lambda<<<kTileGrid, kPass, 0, st>>>(k, ta, P, c.rho, write_rho, sq, c.flags);
  return hipGetLastError();
}

hipError_t enqueue_tile_delta(Ctx& c, const StepConsts& k, const TileArgs& ta, const float4* P,
                              float4* Pn, bool last, bool rest, float4* vout, unsigned* pm,
                              unsigned* rm, hipStream_t st) {
  if (!last)
    k_tile_delta<false, false><<<kTileGrid, kPass, 0, st>>>(
        k, ta, P, Pn, c.planes, c.posB, c.posA, vout, c.rho, pm, rm, c.flags);
  else if (rest)
    k_tile_delta<true, true><<<kTileGrid, kPass, 0, st>>>(
        k, ta, P, Pn, c.planes, c.posB, c.posA, vout, c.rho, pm, rm, c.flags);
  else
    k_tile_delta<true, false><<<kTileGrid, kPass, 0, st>>>(
        k, ta, P, Pn, c.planes, c.posB, c.posA, vout, c.rho, pm, rm, c.flags);
  return hipGetLastError();
}

hipError_t enqueue_tile_xsph(Ctx& c, const StepConsts& k, const TileArgs& ta, bool rest,
                             unsigned* rm, hipStream_t st) {
  if (rest)
    k_tile_xsph<true><<<kTileGrid, kPass, 0, st>>>(k, ta, c.posA, c.velB, c.velA, c.planes, rm,
                                                   c.flags);
  else
    k_tile_xsph<false><<<kTileGrid, kPass, 0, st>>>(k, ta, c.posA, c.velB, c.velA, c.planes,
                                                    nullptr, c.flags);
  return hipGetLastError();
}

hipError_t enqueue_tile_vort_omega(Ctx& c, const StepConsts& k, const TileArgs& ta,
                                   hipStream_t st) {
  k_tile_vort_omega<<<kTileGrid, kPass, 0, st>>>(k, ta, c.posA, c.velA, c.omega,
                                                         c.flags);
  return hipGetLastError();
}

hipError_t enqueue_tile_vort_apply(Ctx& c, const StepConsts& k, const TileArgs& ta, bool rest,
                                   hipStream_t st) {
  if (rest)
    k_tile_vort_apply<true><<<kTileGrid, kPass, 0, st>>>(k, ta, c.posA, c.omega, c.velA,
                                                                 c.planes, c.flags);
  else
    k_tile_vort_apply<false><<<kTileGrid, kPass, 0, st>>>(
        k, ta, c.posA, c.omega, c.velA, c.planes, c.flags);
  return hipGetLastError();
}

}  // namespace fluidhip
