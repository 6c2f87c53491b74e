/*
 * fluid_hip.h — C ABI of the MI355X (gfx950) PBF backend.
 *
 * This is the drop-in boundary that replaces the reference's CUDA backend
 * (cuda/include/fluid/cuda.h:7-9) behind the same solver contract
 * (`fluid::Params`, `fluid::State`, core/include/fluid/core.h:11-132).  Plain
 * C types only: no C++ exceptions cross it, every call returns an int status
 * (0 = ok, negative = error) and fluid_hip_last_error() gives the message of
 * the last failing call on the calling thread.
 *
 * Memory: the library owns all device memory of a context; the caller owns
 * every host buffer it passes.  Positions and velocities cross the boundary as
 * SoA blocks of 3*n floats (all x, then all y, then all z) in ORIGINAL
 * particle order; on the device they stay resident, in the solver's own
 * order, across steps.
 *
 * Threading: one host thread per context at a time.  Work is queued on the
 * context's stream (default: a stream the context creates); calls that
 * return data synchronize that stream.
 */
#ifndef FLUID_HIP_H
#define FLUID_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FLUID_HIP_ABI_VERSION 2
#define FLUID_HIP_MAX_PLANES 8

/* Status codes. */
#define FLUID_HIP_OK 0
#define FLUID_HIP_EINVAL -1       /* bad argument or parameter */
#define FLUID_HIP_ENODEV -2       /* no usable HIP device */
#define FLUID_HIP_EHIP -3         /* a HIP runtime call failed */
#define FLUID_HIP_ECAPACITY -4    /* cell table or neighbor list overflowed */
#define FLUID_HIP_ERANGE -5       /* non-finite or out-of-range position */
#define FLUID_HIP_EUNSUPPORTED -6 /* valid in the reference, not offered here */
#define FLUID_HIP_ESTATE -7       /* call out of order (e.g. no data uploaded) */

/*
 * Per-scene solver parameters: a flattened `fluid::Params`
 * (core/include/fluid/core.h:11-79) with the plane list inline.  Fields the
 * reference cannot set from JSON/CLI (solver_iterations, scorr_dq_coeff,
 * vort_epsilon, vort_norm_eps) are ordinary fields here.  In a batch, every
 * scene must agree on all fields except gravity, the planes, restitution and
 * friction, which may differ per scene.
 */
typedef struct fluid_hip_params {
  float dt;
  float density;
  float particle_mass;
  float h;
  float particle_radius; /* carried for completeness; the step never reads it */
  float epsilon;
  int32_t solver_iterations;
  int32_t use_uniform_grid; /* must be 1: the O(N^2) path is CPU-only */
  int32_t enable_scorr;
  int32_t enable_xsph;
  int32_t enable_vorticity;
  float scorr_k;
  int32_t scorr_n;
  float scorr_dq_coeff;
  float visc_c;
  float plane_restitution;
  float plane_friction;
  float vort_epsilon;
  float vort_norm_eps;
  float gravity[3];          /* Params::external_forces */
  int32_t n_planes;          /* <= FLUID_HIP_MAX_PLANES */
  float planes[FLUID_HIP_MAX_PLANES][4]; /* (nx, ny, nz, d): n.p - d >= 0 inside */
  /* Extension with no counterpart in the reference, whose restitution and
   * friction hold for every plane (core/src/core.cpp:573-612): with
   * plane_props = 1, plane p bounces with plane_restitution_of[p] and damps
   * with plane_friction_of[p] (the scene's contact pass runs when any plane
   * has either above zero); with 0 the two globals above hold for every
   * plane, which is the reference's behaviour. */
  int32_t plane_props;
  float plane_restitution_of[FLUID_HIP_MAX_PLANES];
  float plane_friction_of[FLUID_HIP_MAX_PLANES];
} fluid_hip_params;

/* Capacity knobs for fluid_hip_create; zero fields take defaults. */
typedef struct fluid_hip_config {
  int64_t cell_capacity;      /* cells of the dense grid table, all scenes */
  int64_t neighbor_capacity;  /* neighbor-list entries, all scenes */
  int32_t device;             /* HIP device ordinal */
  int32_t block_size;         /* threads per workgroup of the particle kernels (64..1024) */
  int32_t xcd_remap;          /* 1: map consecutive particle tiles to one XCD
                                 (the tile, bounding-box and key kernels; 0, the
                                 default, measured as fast or faster) */
  int32_t record_diag;        /* 1: keep per-step grid diagnostics (slower) */
  int32_t wave_mode;          /* gathers of wavefronts that share one cell: 0 default (2),
                                 1 per-lane lists only, 2 shared candidate tiles,
                                 3 shared tiles without the coincident-tile shortcut */
  int32_t brick_edge;         /* cells per brick edge of the cell table's Z-order:
                                 0 default (4), 1, 2, 4 or 8 */
  int64_t lane_capacity;      /* lanes of the gather passes (runs of identical
                                 particles, each dense cell from a wave boundary):
                                 0 default (3n + 64), otherwise clamped to
                                 [n, 3n + 64]; a layout that does not fit
                                 packs one lane per run */
  int32_t tile_gather;        /* passes of sparse 4x4x4-cell bricks through LDS tiles:
                                 0 default (on from 32,768 particles), 1 on, 2 off; used
                                 with brick edge 4 and gather mode 2 (results are the
                                 same either way) */
} fluid_hip_config;

typedef struct fluid_hip_ctx fluid_hip_ctx;

/* ---- query (cuda/include/fluid/cuda.h:7-8) ---------------------------- */

/* Replaces `int cuda_version()` (cuda/include/fluid/cuda.h:7). */
int fluid_hip_version(void);

/* Replaces `bool cuda_device_available(int*, const char**)`
 * (cuda/include/fluid/cuda.h:8, cuda/src/cuda_stub.cu:740-762): returns 1
 * when at least one device is usable, else 0 with *error (static string)
 * set when the runtime reported an error. */
int fluid_hip_device_available(int* count, const char** error);

/* Message of the last failing call on this thread ("" if none). */
const char* fluid_hip_last_error(void);

/* ---- context ------------------------------------------------------------ */

/*
 * Creates a context for `n_scenes` independent scenes with
 * scene_counts[s] particles each (scenes are concatenated in that order in
 * every array).  One scene is the plain backend; more than one is the
 * segmented batch, stepped by the same kernels in one launch each.
 * `params` holds n_scenes entries.  No counterpart in the reference, whose
 * CUDA step allocates and frees everything per call (cuda_stub.cu:791-816).
 */
int fluid_hip_create(fluid_hip_ctx** out, int32_t n_scenes, const int64_t* scene_counts,
                     const fluid_hip_params* params, const fluid_hip_config* config);
int fluid_hip_destroy(fluid_hip_ctx* ctx);

/* Replace the per-scene parameters (same validation as create). */
int fluid_hip_set_params(fluid_hip_ctx* ctx, const fluid_hip_params* params);

/* Queue work on this stream (a hipStream_t) instead of the context's own. */
int fluid_hip_set_stream(fluid_hip_ctx* ctx, void* stream);

/* ---- state transfer (the H2D/D2H of cuda_stub.cu:791-816, :1092-1097) --- */

/* pos, vel: host SoA blocks of 3*n floats, original particle order. */
int fluid_hip_upload(fluid_hip_ctx* ctx, const float* pos, const float* vel, float time);
int fluid_hip_download(fluid_hip_ctx* ctx, float* pos, float* vel, float* time);

/* Asynchronous download for output while the device keeps stepping (the
 * reference's app downloads nothing: its state lives on the host).  Queues
 * the committed state's conversion to original order and its copy into
 * pos/vel (3*n floats each; pinned memory from fluid_hip_host_alloc makes
 * the copy asynchronous) on the context's stream and returns at once;
 * fluid_hip_download_wait() blocks until that copy has landed and returns
 * the time of the downloaded state.  Steps queued in between run after the
 * copy, in stream order. */
int fluid_hip_download_async(fluid_hip_ctx* ctx, float* pos, float* vel);
int fluid_hip_download_wait(fluid_hip_ctx* ctx, float* time);

/* Page-locked host memory (for fluid_hip_download_async), and its release. */
int fluid_hip_host_alloc(void** out, int64_t bytes);
int fluid_hip_host_free(void* ptr);

/* ---- stepping ------------------------------------------------------------ */

/*
 * Advances every scene by `nsteps` substeps of `fluid::step`
 * (core/src/core.cpp:119-615; the reference's GPU form is `cuda_step`,
 * cuda/src/cuda_stub.cu:764-1099), device-resident.  Synchronizes once at
 * the end to check the device error flags.  use_graph=1 replays a captured
 * hipGraph of one substep.
 */
int fluid_hip_step(fluid_hip_ctx* ctx, int32_t nsteps, int32_t use_graph);

/* Queue `nsteps` substeps without synchronizing or checking errors (for
 * timing loops); fluid_hip_sync() checks afterwards. */
int fluid_hip_step_async(fluid_hip_ctx* ctx, int32_t nsteps, int32_t use_graph);
int fluid_hip_sync(fluid_hip_ctx* ctx);

/* ---- diagnostics of the last step (original particle order) -------------- */

/*
 * cell[3*n]: integer cell (x,y,z) of the predicted position pred0, as
 *   floor(p * (1/h)) (core/src/core.cpp:28-34).  Needs record_diag=1.
 * sorted_ids[n]: original ids in the solver's sorted order (scene, Z-order
 *   cell key, original id).  Needs record_diag=1 for the cell keys, not ids.
 * nbr_count[n]: neighbor count of each particle (core/src/core.cpp:205-247);
 *   when the last substep kept no list for some particles (dense cells, runs
 *   of identical particles), a counting walk of its grid supplies them.
 * rho[n], lambda[n]: density and lambda of the last constraint iteration.
 * Any pointer may be NULL.
 */
int fluid_hip_get_diag(fluid_hip_ctx* ctx, int32_t* cell, int32_t* sorted_ids,
                       int32_t* nbr_count, float* rho, float* lambda);

/* Neighbor lists of the last step as original ids, particle by particle in
 * original order, each list in the reference's order (dz, dy, dx cells,
 * ascending id).  `cap` entries of `ids`; returns the total in *total.
 * Without record_diag, particles of dense cells store no list (their passes
 * retest the neighbor predicate): then FLUID_HIP_ESTATE. */
int fluid_hip_get_neighbors(fluid_hip_ctx* ctx, int32_t* ids, int64_t cap, int64_t* total);

/* Sorted cell key (scene offset + brick-Z-order index) of each sorted slot and
 * the scene grid descriptors, for the bit-exact sort tests.  desc receives 8
 * int32 per scene: origin x,y,z, bricks x,y,z, cell offset, cell count. */
int fluid_hip_get_grid(fluid_hip_ctx* ctx, int32_t* slot_keys, int32_t* desc);

/* Sizes of the last step: out[0] neighbor pairs (every particle's, counted
 * by a walk of the grid when dense cells stored no lists), out[1] stored
 * list words, out[2] grid cells (all scenes), out[3] neighbor-list capacity
 * (words), out[4] cell-table capacity, out[5] capacity growths since the
 * context was created, out[6] device bytes held by the rollout recording
 * (checkpoints and tapes), out[7] substeps the last backward pass
 * recomputed, out[8] bricks the last substep ran through the LDS-tiled
 * passes, out[9] particles those tiles computed.  `n` entries are written
 * (<= 10). */
int fluid_hip_get_stats(fluid_hip_ctx* ctx, int64_t* out, int32_t n);

/* ---- autotuning ---------------------------------------------------------- */

/* No counterpart in the reference.  Picks the launch configuration of this
 * context by timing, one knob at a time over two passes, among the knobs in
 * `knobs` (a mask of FLUID_HIP_TUNE_*; 0 selects FLUID_HIP_TUNE_DEFAULT):
 * the workgroup size (64..1024), the cell table's brick edge (1, 2, 4, 8),
 * the XCD-aware tile order, the LDS-tiled passes of sparse bricks (a
 * kernel-structure choice: one persistent tile kernel per pass, or every
 * particle through the per-lane kernels) and, on request, the gather mode of wavefronts
 * sharing a cell (wave_mode 1..3).  The gather mode is left out by default:
 * which mode is fastest depends on how far particles have piled up, which a
 * short trial at one state does not show.  Every candidate runs
 * `trial_steps` substeps from the current state between two events; its
 * final state must equal the starting configuration's bit for bit, or it is
 * rejected.  The fastest candidate (by at least 3%) is kept; the state and
 * time are restored.  `best` receives the chosen configuration, report[0..3]
 * the starting and chosen ms per substep, the candidates timed and the
 * candidates rejected (either may be NULL).  A recorded rollout is
 * discarded. */
#define FLUID_HIP_TUNE_WAVE_MODE 1
#define FLUID_HIP_TUNE_BLOCK 2
#define FLUID_HIP_TUNE_BRICK 4
#define FLUID_HIP_TUNE_XCD 8
#define FLUID_HIP_TUNE_TILES 16
#define FLUID_HIP_TUNE_DEFAULT \
  (FLUID_HIP_TUNE_BLOCK | FLUID_HIP_TUNE_BRICK | FLUID_HIP_TUNE_XCD | FLUID_HIP_TUNE_TILES)
int fluid_hip_autotune(fluid_hip_ctx* ctx, int32_t trial_steps, int32_t knobs,
                       fluid_hip_config* best, double* report);

/* ---- differentiable rollout ---------------------------------------------- */

/* No counterpart in the reference (its solver has no derivatives).  For
 * inverse problems: gradients of a loss on the final state with respect to
 * the initial positions, velocities and per-scene gravity.
 *
 * Runs `nsteps` substeps from the current state, as fluid_hip_step does, and
 * keeps what the reverse pass needs: the state before every segment of
 * `checkpoint_every` substeps, and the recording (iterates, lambdas, plane
 * and restitution decisions, neighbor lists) of the last segment.  Leaves the
 * context at the final state.  Memory: 32 B per particle per checkpoint, and
 * per recorded substep about (96 + 36 * iterations) B per particle plus its
 * neighbor lists.  Any later upload, step or set_params discards the record.
 * Vorticity confinement and zero constraint iterations are refused
 * (FLUID_HIP_EUNSUPPORTED). */
int fluid_hip_forward_record(fluid_hip_ctx* ctx, int32_t nsteps, int32_t checkpoint_every);

/* Reverse pass through the recorded rollout: the hand-written adjoint of
 * every substep (predict, each lambda/delta iteration with s_corr and plane
 * projection, commit, XSPH, restitution/friction) with each substep's
 * neighbor lists frozen as the forward built them.  Segments other than the
 * recorded one are recomputed from their checkpoints.
 *   dpos, dvel    dL/dx_T and dL/dv_T, 3n floats SoA in original order (NULL:
 *                 zero)
 *   dpos0, dvel0  dL/dx_0, dL/dv_0 (3n SoA; NULL: not returned)
 *   dgravity      dL/dg, 3 floats per scene (NULL: not returned)
 * The context is left at the recorded final state; the record stays valid, so
 * the pass can be repeated with other loss gradients. */
int fluid_hip_backward(fluid_hip_ctx* ctx, const float* dpos, const float* dvel, float* dpos0,
                       float* dvel0, float* dgravity);

/* ---- stepping without device allocation ---------------------------------- */

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc): a stateless
 * variant of create + upload + fluid_hip_step + download + destroy for one
 * scene that allocates nothing.  Every buffer of the substep lives in one
 * static device array of FLUID_HIP_NOALLOC_SCRATCH_BYTES that each GPU the
 * library runs on holds until the process ends; the cell table and the
 * neighbor lists get fixed shares of it and never grow (a scene that needs
 * more fails with FLUID_HIP_ECAPACITY, leaving pos/vel unchanged).
 *   params  one scene's parameters (validated as by fluid_hip_create)
 *   n       particles, 0 <= n <= FLUID_HIP_NOALLOC_MAX_PARTICLES (else EINVAL)
 *   pos,vel DEVICE SoA blocks of 3*n floats in original order, advanced in
 *           place by `nsteps` substeps
 *   time    host float advanced by nsteps * dt (may be NULL)
 *   stream  hipStream_t to queue on (NULL: the null stream); synchronized
 *           once at the end to check the device flags
 * Results are bitwise those of fluid_hip_step.  Every call on one GPU shares
 * that GPU's array: calls that overlap on one GPU (other threads, other
 * streams) race and must be serialized by the caller. */
#define FLUID_HIP_NOALLOC_SCRATCH_BYTES (64ll << 20)
#define FLUID_HIP_NOALLOC_MAX_PARTICLES 32768
int fluid_hip_step_noalloc(const fluid_hip_params* params, int64_t n, float* pos, float* vel,
                           float* time, int32_t nsteps, void* stream);

/* ---- timing -------------------------------------------------------------- */

/* Per-kernel device time (HIP events on the context's stream) of the steps
 * queued while enabled.  Names: "predict_bbox", "grid_setup", "cell_keys",
 * "sort", "cell_count", "reorder", "tiles", "runs", "neighbors", "lambda", "delta", "xsph",
 * "vorticity_omega", "vorticity_apply"; with LDS tiles also "tile_neighbors",
 * "tile_lambda", "tile_delta", "tile_xsph", "tile_vorticity_omega" and
 * "tile_vorticity_apply", the tile kernels alone, timed on the stream they run
 * on (a pass's own name spans its per-lane and tile kernels).  Launches of a substep that
 * overflowed a capacity, and of the substeps queued behind it, are not
 * counted (fluid_hip_step runs them again after growing the tables). */
int fluid_hip_profile(fluid_hip_ctx* ctx, int32_t enable);
int fluid_hip_profile_read(fluid_hip_ctx* ctx, const char* kernel, double* total_ms,
                           int64_t* launches);

#ifdef __cplusplus
}
#endif

#endif /* FLUID_HIP_H */
