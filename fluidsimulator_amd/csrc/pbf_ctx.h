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

struct Prof {
  struct Pending {
    const char* name;
    hipEvent_t a, b;
  };
  bool on = false;
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  std::unordered_map<std::string, std::pair<double, long long>> acc;
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
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
  int scan_blocks = 512;
  bool record_diag = false;
  int wave_mode = 2;

  // state (A = committed state between substeps)
  float4* posA = nullptr;  // xyz, w = original id bits
  float4* velA = nullptr;  // xyz, w = rho of the last lambda pass (scratch)
  float4* posB = nullptr;
  float4* velB = nullptr;
  float4* pred[2] = {nullptr, nullptr};  // xyz, w = lambda
  float4* omega = nullptr;               // xyz = curl, w = |curl|
  float* rho = nullptr;

  // grid
  uint64_t* cellc = nullptr;  // packed cell of pred0, previous order
  int* key = nullptr;
  int* rank = nullptr;
  int* perm = nullptr;
  int* sid = nullptr;
  int* skey = nullptr;
  int* count = nullptr;  // cell_cap
  int* start = nullptr;  // cell_cap + 1
  long long cell_cap = 0;
  int* bbox = nullptr;   // n_scenes * 6
  int* tile_bbox = nullptr;  // 6 per particle tile (workgroup)
  GridDesc* grid = nullptr;
  int* total_cells = nullptr;
  int* scan_partial = nullptr;  // scan_blocks + 1
  uint64_t* slot_cell = nullptr;  // packed cell of pred0 per sorted slot
  int* slot_key = nullptr;  // table key of each sorted slot's cell

  // neighbors
  long long* noff = nullptr;  // first word of each slot's list
  int* nwords = nullptr;      // encoded length (words)
  int* ncount = nullptr;      // neighbors (decoded length)
  int* nbr = nullptr;         // run-length-encoded lists, nbr_cap words
  long long nbr_cap = 0;
  unsigned long long* nbr_total = nullptr;
  unsigned long long* need_cells = nullptr;  // cells a failed substep needed

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

  // staging for transfers
  float* stage = nullptr;  // 6 * n

  // graph of one substep
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;

  Prof prof;
};

// Queues one substep on ctx.stream (kernels in pbf_kernels.hip).
hipError_t enqueue_substep(Ctx& c);
hipError_t enqueue_upload(Ctx& c, const float* d_pos, const float* d_vel);
hipError_t enqueue_download(Ctx& c, float* d_pos, float* d_vel);
hipError_t enqueue_reset_flags(Ctx& c);
hipError_t enqueue_init_grid(Ctx& c);

}  // namespace fluidhip
