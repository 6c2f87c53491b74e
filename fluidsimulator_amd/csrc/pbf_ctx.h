// Host-side context of the gfx950 PBF backend: device buffers, per-scene
// tables and the launch plan of one substep.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/fluid_hip.h"
#include "pbf_device.h"

namespace fluidhip {

// Brick tiles of the LDS-tiled gather passes (pbf_tiles.hip).
constexpr int kTileMaxHalo = 2048;  // particles of a tile's 6^3-cell halo held in LDS
// A brick with a cell of this many candidates (particles in its 27 cells) or
// more is no tile.
constexpr int kTileMaxCand = 512;
// neighbors of a tile particle, at most (else its tile goes per-lane).  In
// batches of kForkMinParticles particles or more a list holds every neighbor
// a tile particle can have (fewer than kTileMaxCand), 1 KiB per particle, so
// no tile goes back and the lane map is built before the tile lists; smaller
// batches, which tile their piles of coincident particles too, keep
// kTileListSmall entries and hand back the tiles that overflow.
constexpr int kTileList = 512;
constexpr int kTileListSmall = 256;
static_assert(kTileList >= kTileMaxCand - 1, "large-batch tiles must never overflow");
// large batches: tile criteria and list capacity, and the tile kernels on the
// side stream (at 65,536 particles overlapping the two kernels of a pass
// saves about what the fork and join cost)
constexpr int kForkMinParticles = 1 << 18;
constexpr int kTileAutoMinParticles = 1 << 15;  // default tile mode: tiles from this size
constexpr int kTileGrid = 2048;     // workgroups of the persistent tile kernels

// What the tile kernels read besides their pass's arrays.
struct TileArgs {
  const uint64_t* slot_cell;
  const int* tile_count;   // tiles found this substep (device)
  int tile_cap;
  const int4* tile_rec;    // per tile: {a, e, halo records, scene}, {first cell x, y, z, brick}
  const int* tile_seg;     // per tile: 217 LDS offsets of the halo cells
  const int* tile_src;     // per tile: 216 first slots of the halo cells
  const uint4* tile_list;  // list_cap / 8 x n chunks of 8 neighbors' LDS records
  int list_cap;            // entries of a list
  const int* tile_cnt;     // n: neighbors of a tile particle
};

struct Prof {
  struct Pending {
    const char* name;
    hipEvent_t a, b;
    int step;  // substep of the queue the launch belongs to
  };
  bool on = false;
  int step = 0;  // substep of the queue being enqueued
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  std::unordered_map<std::string, std::pair<double, long long>> acc;
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // a second stream of the context (default priority) on which each pass's
  // LDS-tile kernel runs beside the per-lane kernel, forked from and joined
  // back into `stream` by the two events; null: the tile kernels run on
  // `stream` (contexts built for one call, e.g. the allocation-free entry)
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  bool own_stream = false;

  // problem shape
  int n = 0;
  int n_scenes = 0;
  std::vector<int> seg_h;  // n_scenes + 1 slot offsets
  std::vector<fluid_hip_params> params;
  StepConsts consts{};
  int iterations = 0;
  bool xsph = false, vort = false, restitution_any = false;

  // launch knobs
  int block = 256;
  int xcd_remap = 1;
  int scan_blocks = 1024;  // workgroups of the device scans (the partials scan takes up to 1024)
  bool record_diag = false;
  int wave_mode = 2;
  int brick_shift = 2;  // log2 of the cells per brick edge of the cell table
  int tile_gather = 0;  // 0 default (on), 1 on, 2 off: LDS-tiled passes for sparse bricks

  // state (A = committed state between substeps)
  float4* posA = nullptr;  // xyz, w = original id bits
  float4* velA = nullptr;  // xyz, w = rho of the last lambda pass (scratch)
  float4* posB = nullptr;
  float4* velB = nullptr;
  // constraint iterates, xyz, w = lambda: pred[0] holds p^0 (the predicted
  // positions the neighbor lists are built from) for the whole substep,
  // later iterates alternate between pred[1] and pred[2]
  float4* pred[3] = {nullptr, nullptr, nullptr};
  float4* omega = nullptr;               // xyz = curl, w = |curl|
  float* rho = nullptr;

  // grid
  uint64_t* cellc = nullptr;  // packed cell of pred0, previous order
  int* key = nullptr;  // per slot of the previous order: cell key
  // stable radix sort of (cell key) over particles in id order; values are
  // previous slots, so the sorted values list the new (cell, id) order
  int* sort_k[2] = {nullptr, nullptr};
  int* sort_v[2] = {nullptr, nullptr};
  int* radix_hist = nullptr;  // 256 digits x radix_blocks
  int* radix_off = nullptr;   // its exclusive scan (+1)
  int* radix_len = nullptr;   // 256 * radix_blocks, on the device (scan length)
  int radix_blocks = 0;
  int* count = nullptr;  // cell_cap
  int* start = nullptr;  // cell_cap + 1
  long long cell_cap = 0;
  int* bbox = nullptr;   // n_scenes * 6
  int* tile_bbox = nullptr;  // 6 per particle tile (workgroup)
  GridDesc* grid = nullptr;
  int* total_cells = nullptr;
  int* scan_partial = nullptr;  // 2 scan_blocks + 2: range sums of one or two scanned arrays
  uint64_t* slot_cell = nullptr;  // packed cell of pred0 per sorted slot
  int* slot_key = nullptr;  // table key of each sorted slot's cell
  // runs of identical consecutive slots (one lane computes each run)
  int* run_no = nullptr;     // per slot: head flag, then the slot's run number
  int* run_excl = nullptr;   // n + 1: exclusive scan of the head flags; [n] = runs
  int* rep_list = nullptr;   // per run: its first slot
  int* run_end = nullptr;    // per slot: one past the last slot of its run
  int* n_dev = nullptr;      // the particle count, on the device (scan length)
  // lanes of the passes: runs of sparse cells packed, then each dense cell's
  // runs from a wave boundary (one cell per wave)
  int* cell_no = nullptr;    // per slot: cell head flag, then the slot's non-empty cell number
  int* cell_excl = nullptr;  // n + 1: exclusive scan of the cell head flags; [n] = cells
  int* cell_first = nullptr; // per non-empty cell: its first slot
  int* lanes1 = nullptr;     // per non-empty cell: packed lanes
  int* lanes2 = nullptr;     // per non-empty cell: aligned lanes (whole waves)
  int* off1 = nullptr;       // n + 1: exclusive scan of lanes1; [cells] = total
  int* off2 = nullptr;       // n + 1: exclusive scan of lanes2; [cells] = total
  int* lane_rep = nullptr;   // lane_cap: the run head slot of each lane, -1 for padding
  int* lane_info = nullptr;  // 3: packed lanes, first aligned lane, end of aligned lanes
  long long lane_cap = 0;    // lanes the passes launch with runs: 3n + 64

  // brick tiles (sparse bricks whose passes run through LDS)
  int* in_tile = nullptr;      // per slot: 1 + its tile's index this substep, 0 outside tiles
  int* tile_count = nullptr;   // [0] tiles found this substep, [1] candidate bricks
  int* tile_cand = nullptr;    // tile_cap: bricks of at least 32 particles (classify input)
  int4* tile_rec = nullptr;    // 2 tile_cap: tile records
  int* tile_seg = nullptr;     // tile_cap x 217: LDS offset of each halo cell
  int* tile_src = nullptr;     // tile_cap x 216: first slot of each halo cell
  uint4* tile_list = nullptr;  // tile_list_cap / 8 x n: chunks of 8 neighbors' LDS records (16 bit)
  int tile_list_cap = 0;       // kTileList, or kTileListSmall for small batches
  int* tile_cnt = nullptr;        // n: neighbors of a tile particle
  long long tile_cap = 0;

  // neighbors
  long long* noff = nullptr;  // first word of each slot's list
  int* nwords = nullptr;      // encoded length (words)
  int* ncount = nullptr;      // neighbors (decoded length)
  int* nbr = nullptr;         // run-length-encoded lists, nbr_cap words
  long long nbr_cap = 0;
  unsigned long long* nbr_total = nullptr;
  unsigned long long* need_cells = nullptr;  // cells a failed substep needed
  unsigned long long* pair_total = nullptr;  // neighbor pairs, counted on request

  // scenes
  int* seg = nullptr;
  ScenePlanes* planes = nullptr;

  // status
  unsigned* flags = nullptr;
  int* steps_started = nullptr;
  float time = 0.0f;
  bool uploaded = false;
  bool stepped = false;
  long long grow_events = 0;  // capacity growths (each replays the failed substep)
  // the last substep stored no lists for cooperative particles (their
  // neighbor counts need a counting walk), and its p^0 is in pred[0]
  bool lists_partial = false;
  // the last substep computed one lane per run of identical slots (the
  // run's other slots copied its outputs; rho was written by the heads only)
  bool last_runs = false;

  // staging for transfers
  float* stage = nullptr;  // 6 * n
  hipEvent_t dl_event = nullptr;  // end of the last asynchronous download
  float dl_time = 0.0f;           // time of the state it downloads
  bool dl_pending = false;

  // graph of one substep
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;

  Prof prof;

  // differentiable rollout (created by the first recording)
  struct AdjRecord* adj = nullptr;
};

// Where a recorded substep writes what its adjoint needs (device arrays in
// that substep's sorted slot order, one entry per constraint iteration).
struct SubstepRec {
  float4* const* P;    // iterate p^k of the constraint loop, w = lambda^k
  float4* const* SQ;   // gradient sum S (xyz) and denominator sum Q (w)
  unsigned* const* M;  // planes that pushed p^k + delta
  unsigned* RM;        // restitution: planes in contact | approached << 8
};

// Array holding iterate k of the constraint loop (unrecorded substeps).
inline float4* iterate_buffer(Ctx& c, int k) {
  return k == 0 ? c.pred[0] : c.pred[1 + ((k - 1) & 1)];
}

// Queues one substep on ctx.stream (kernels in pbf_kernels.hip); with `rec`
// the iterates and branch decisions are also written to the recording.
hipError_t enqueue_substep(Ctx& c, const SubstepRec* rec = nullptr);

// What the adjoint of one substep reads (device arrays, that substep's slot
// order; P, SQ and M hold `iterations` consecutive blocks of n entries).
struct SubTape {
  float4* X = nullptr;    // input positions, w = original id
  float4* PI = nullptr;   // committed positions p^I
  float4* VC = nullptr;   // commit velocity (xyz) and rho of the last lambda pass (w)
  float4* P = nullptr;    // p^k, w = lambda^k
  float4* SQ = nullptr;   // gradient sum S (xyz), Q (w)
  unsigned* M = nullptr;  // planes that pushed p^k + delta
  unsigned* RM = nullptr; // restitution decisions
  long long* noff = nullptr;
  int* nwords = nullptr;
  int* ncount = nullptr;
  int* nbr = nullptr;     // the substep's run-length-encoded lists
  long long nbr_cap = 0;
  std::vector<float4*> Pk, SQk;
  std::vector<unsigned*> Mk;
};

// Adjoint state carried between substeps: position and velocity adjoints in
// original particle order, per-slot work arrays, per-scene g-bar (double).
struct AdjBuffers {
  float4* GX = nullptr;
  float4* GV = nullptr;
  float4* U = nullptr;   // adjoint of the XSPH output velocity
  float4* PB = nullptr;  // adjoint of the current iterate
  float4* XB = nullptr;  // commit's share of the input-position adjoint
  float4* W = nullptr;   // delta-bar / rho0
  float4* LB = nullptr;  // (Q-bar S, Q-bar)
  float* RB = nullptr;   // rho-bar
  double* gacc = nullptr;
  float* stage = nullptr;  // 6n floats: SoA transfers
  unsigned long long* tape_need = nullptr;  // list words a tape lacked room for (0: none)
};

// A recorded rollout: the state before every segment of `every` substeps,
// the final state, and the recordings of one segment's substeps.
struct AdjRecord {
  bool valid = false;
  int nsteps = 0, every = 0, nseg = 0;
  int tape_seg = -1;  // segment the tapes currently hold
  std::vector<float4*> ck_pos, ck_vel;
  std::vector<float> ck_time;
  float4* fin_pos = nullptr;
  float4* fin_vel = nullptr;
  float fin_time = 0.0f;
  std::vector<SubTape> tapes;
  AdjBuffers buf;
  int recomputed = 0;  // substeps re-run by the last backward
};

hipError_t enqueue_adjoint_seed(Ctx& c, const float* d_dpos, const float* d_dvel,
                                const AdjBuffers& b);
hipError_t enqueue_adjoint_substep(Ctx& c, const SubTape& t, const AdjBuffers& b);
hipError_t enqueue_adjoint_result(Ctx& c, float* d_dpos, float* d_dvel, const AdjBuffers& b);
// After a recorded substep (skipped when it failed): its input positions,
// committed state, commit velocity and neighbor lists into the tape; lists
// longer than the tape's room only raise *need to their length.
hipError_t enqueue_tape_copy(Ctx& c, SubTape& t, unsigned long long* need);
hipError_t enqueue_upload(Ctx& c, const float* d_pos, const float* d_vel);
hipError_t enqueue_download(Ctx& c, float* d_pos, float* d_vel);
hipError_t enqueue_reset_flags(Ctx& c);
// Counts the neighbor pairs of the last substep's grid (lists of
// cooperative particles are not stored) into c.pair_total.
hipError_t enqueue_count_pairs(Ctx& c);
hipError_t enqueue_init_grid(Ctx& c);

// The tile kernels (pbf_tiles.hip).
hipError_t enqueue_tile_classify(Ctx& c, const StepConsts& k);
hipError_t enqueue_tile_neighbors(Ctx& c, const StepConsts& k, const TileArgs& ta,
                                  const float4* p0, bool store, hipStream_t st);
hipError_t enqueue_tile_lambda(Ctx& c, const StepConsts& k, const TileArgs& ta, float4* P,
                               int write_rho, float4* sq, hipStream_t st);
hipError_t enqueue_tile_delta(Ctx& c, const StepConsts& k, const TileArgs& ta, const float4* P,
                              float4* Pn, bool last, bool rest, float4* vout, unsigned* pm,
                              unsigned* rm, hipStream_t st);
hipError_t enqueue_tile_xsph(Ctx& c, const StepConsts& k, const TileArgs& ta, bool rest,
                             unsigned* rm, hipStream_t st);
hipError_t enqueue_tile_vort_omega(Ctx& c, const StepConsts& k, const TileArgs& ta,
                                   hipStream_t st);
hipError_t enqueue_tile_vort_apply(Ctx& c, const StepConsts& k, const TileArgs& ta, bool rest,
                                   hipStream_t st);

}  // namespace fluidhip
