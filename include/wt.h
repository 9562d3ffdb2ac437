/*
 * wt.h — C ABI of the MI355X lattice-Boltzmann wind-tunnel library (libwt.so).
 *
 * This is the drop-in boundary for the reference's solver surface: every entry point replaces one
 * member of `SolverGPU` (src/solver/SolverGPU.ts), of the `LbmSolver` contract
 * (src/backend/types.ts:7-18) or of the voxelizer host (src/voxelize/VoxelizerGPU.ts), cited
 * per function below. Hosts bind it through N-API (Node), ctypes (Python) or plain C.
 *
 * Conventions
 *  - Every call returns an int status: WT_OK (0) or a negative WT_E_* code; wt_last_error() gives a
 *    thread-local message for the last failing call on this thread.
 *  - The caller owns host arrays; the library owns all device memory.
 *  - One HIP stream per solver handle. Handles are independent; one handle is not re-entrant.
 *  - wt_step is asynchronous (it enqueues work and returns); wt_step_and_sample, the read calls and
 *    the upload calls synchronise the handle's stream.
 *  - Populations are exchanged with the host unpacked, as full values f_i (not deviations), in the
 *    direction-major SoA order f[i * n + cell], cell = x + nx * (y + ny * z).
 *  - Physics failure (NaN, over-speed) is reported through wt_stats.unstable, not as an error,
 *    as in SolverGPU.ts:277.
 */
#ifndef WT_H
#define WT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WT_OK 0
#define WT_E_ARG -1      /* invalid argument (sizes, enum values, null pointers) */
#define WT_E_HIP -2      /* a HIP runtime call failed (message has the HIP error string) */
#define WT_E_NOMEM -3    /* device allocation failed */
#define WT_E_STATE -4    /* call not valid in the handle's current state */
#define WT_E_NODEV -5    /* no usable GPU */

typedef struct wt_solver wt_solver;

/* GridDims, src/solver/types.ts:56-60. nz == 1 selects D2Q9, nz > 1 D3Q19 (SolverGPU.ts:63-64). */
typedef struct {
  int32_t nx, ny, nz;
} wt_dims;

enum { WT_GROUND_FREESLIP = 0, WT_GROUND_NOSLIP = 1, WT_GROUND_MOVING = 2 };
enum { WT_SIDES_FREESLIP = 0, WT_SIDES_PERIODIC = 1 };
/* Population storage. F16/BF16 store deviations f_i - w_i; like SolverGPU.ts:69 a half-width format
 * needs an even nx and otherwise falls back to F32 (wt_info reports the format in use). */
enum { WT_PREC_F32 = 0, WT_PREC_F16 = 1, WT_PREC_BF16 = 2 };
enum { WT_COLL_REGULARIZED = 0, WT_COLL_BGK = 1 };
/* Sub-grid closure: Smagorinsky (wgslLbm.ts:214) or the learned per-cell MLP eddy viscosity. */
enum { WT_CLOSURE_SMAGORINSKY = 0, WT_CLOSURE_MLP = 1 };

/* SolverParams, src/solver/types.ts:3-29 (same fields and meaning; optional fields take the
 * reference defaults when the caller passes them: sponge_in 0, tau_wall 0.5, interp 1,
 * collision regularized). */
typedef struct {
  double u;            /* lattice inlet velocity */
  double nu;           /* molecular lattice viscosity; tau0 = 3 nu + 1/2 */
  double cs;           /* Smagorinsky constant */
  int32_t ground;      /* WT_GROUND_* */
  int32_t sides;       /* WT_SIDES_* (z boundaries in 3D) */
  double sponge_nu;    /* extra viscosity in the outlet sponge */
  double sponge_start; /* fraction of nx where the outlet sponge starts */
  double sponge_in;    /* width (fraction of nx) of the inlet absorbing layer */
  double tau_wall;     /* minimum relaxation time on floor-adjacent cells */
  int32_t precision;   /* WT_PREC_*, fixed at creation */
  int32_t interp;      /* 1: Bouzidi interpolated bounce-back, 0: halfway */
  int32_t collision;   /* WT_COLL_* */
  int32_t closure;     /* WT_CLOSURE_* */
  double ema_alpha;    /* EMA weight of the time averages at each sample */
} wt_params;

/* Field mask for wt_set_params (Partial<SolverParams>, SolverGPU.ts:155-158). */
#define WT_P_U (1u << 0)
#define WT_P_NU (1u << 1)
#define WT_P_CS (1u << 2)
#define WT_P_GROUND (1u << 3)
#define WT_P_SIDES (1u << 4)
#define WT_P_SPONGE_NU (1u << 5)
#define WT_P_SPONGE_START (1u << 6)
#define WT_P_SPONGE_IN (1u << 7)
#define WT_P_TAU_WALL (1u << 8)
#define WT_P_INTERP (1u << 9)
#define WT_P_COLLISION (1u << 10)
#define WT_P_CLOSURE (1u << 11)
#define WT_P_EMA_ALPHA (1u << 12)
#define WT_P_ALL 0x1fffu

/* SolverStats, src/solver/types.ts:31-42, decoded as SolverGPU.ts:302-308, in FP64. */
typedef struct {
  double fx, fy, fz;   /* momentum-exchange force on the body (lattice units) */
  double max_u;        /* max |u| over fluid cells */
  double rho_ref;      /* mean density on the reference plane x = min(4, nx-1) */
  int32_t unstable;    /* NaN flag, non-finite fx or max u^2 > 0.25 */
  int32_t pad;
  int64_t step;        /* solver step at which the sample was taken */
} wt_stats;

/* Read-only facts about a handle. */
typedef struct {
  wt_dims dims;
  int64_t n;           /* cells */
  int32_t q;           /* 9 or 19 */
  int32_t precision;   /* storage actually used (WT_PREC_*) */
  int64_t step;        /* stepCount */
  int64_t wall_cells;  /* cells next to the body, updated by the wall kernel */
  int64_t device_bytes;   /* device memory held by the handle */
} wt_info;

/* VoxelInfo, src/voxelize/VoxelizerGPU.ts (frontal cells, solid cells, bbox min / max+1). */
typedef struct {
  int64_t frontal;
  int64_t solid_cells;
  int32_t min[3];
  int32_t max[3];
} wt_voxel_info;

/* ---- lifecycle -------------------------------------------------------------------------- */

/* new SolverGPU(device, dims, params), SolverGPU.ts:61-144: allocates everything on the current
 * HIP device, uploads all-fluid flags and unknown distances, and leaves an init pending
 * (reset(), :190-194). */
int wt_create(const wt_dims* dims, const wt_params* params, wt_solver** out);
/* destroy(), SolverGPU.ts:432-437. Accepts NULL. */
int wt_destroy(wt_solver* s);
/* static bytesFor(dims), SolverGPU.ts:439-443 (the reference's FP32 estimate, in bytes). */
int64_t wt_bytes_for(const wt_dims* dims);
int wt_info_get(const wt_solver* s, wt_info* out);
/* Thread-local message of the last failing call ("" if none). */
const char* wt_last_error(void);

/* ---- parameters / reset ----------------------------------------------------------------- */

/* setParams(Partial<SolverParams>), SolverGPU.ts:155-187: copies the fields selected by mask. */
int wt_set_params(wt_solver* s, const wt_params* p, uint32_t mask);
int wt_get_params(const wt_solver* s, wt_params* out);
/* tau0 getter, SolverGPU.ts:151-153. */
double wt_tau0(const wt_solver* s);
/* reset(), SolverGPU.ts:190-194: re-init both buffers at the next step, stepCount = 0, clear stats. */
int wt_reset(wt_solver* s);
/* clearStats(), SolverGPU.ts:196-201: meanB = 0, meanA = (0,0,0,1). */
int wt_clear_stats(wt_solver* s);

/* ---- geometry --------------------------------------------------------------------------- */

/* uploadFlags(Uint32Array), SolverGPU.ts:342-344: 0 fluid, 1 solid, 1 | (k+1) << 8 wheel k. */
int wt_upload_flags(wt_solver* s, const uint32_t* flags);
/* readFlags(), SolverGPU.ts:347-356. */
int wt_read_flags(wt_solver* s, uint32_t* out);
/* uploadSdf(Float32Array), SolverGPU.ts:312-322: wall distance in cells, < 0 (or NaN) = unknown. */
int wt_upload_sdf(wt_solver* s, const float* dist);
/* No reference counterpart (a debug readback beside readFlags): the stored wall distances, decoded,
 * -1 where unknown. */
int wt_read_sdf(wt_solver* s, float* dist);
/* setWheels(list), SolverGPU.ts:325-335: 8 floats per wheel (cx, cy, cz, ax, ay, az, r, hw),
 * grid space, at most 16 (extra wheels are ignored, as in the reference). */
int wt_set_wheels(wt_solver* s, const float* wheels, int32_t count);

/* Analytic-SDF body for the voxelizer: a union of primitives in grid space. */
enum { WT_PRIM_ROUNDED_BOX = 1, WT_PRIM_CYLINDER = 2, WT_PRIM_SPHERE = 3, WT_PRIM_WEDGE = 4 };
typedef struct {
  int32_t kind;        /* WT_PRIM_* */
  int32_t flags;       /* reserved, 0 */
  float center[3];     /* grid-space centre */
  float rot[9];        /* row-major rotation R; local = R^T (p - center) */
  float half[3];       /* box / wedge: half extents; cylinder: (radius, half width, 0); sphere: (r,0,0) */
  float radius;        /* box / wedge corner radius */
  float param;         /* wedge: slope of the top face (height lost per unit length), else 0 */
} wt_prim;

/* Voxelize an analytic body into the solver (no reference counterpart; the benchmark workloads' procedural
 * bodies). Writes flags (solid iff the cell centre is inside; cells inside a wheel cylinder of the
 * solver's wheel list are tagged as VoxelizerGPU's wheel pass does, wgslVoxel.ts:268-290) and
 * distances |sdf| < 2.5 at cell centres (the roundFlags convention, validation.ts:35-53). */
int wt_voxelize_sdf(wt_solver* s, const wt_prim* prims, int32_t count, wt_voxel_info* info);

/* VoxelizerGPU.voxelize(solver, mesh, matrix, wheels), VoxelizerGPU.ts:48-151: triangle mesh
 * (positions xyz per vertex, 3 indices per triangle), column-major 4x4 mesh->grid matrix M, and
 * the wheel list (8 floats per wheel, as wt_set_wheels; may be NULL/0). */
int wt_voxelize_mesh(wt_solver* s, const float* positions, int32_t n_vertices, const uint32_t* indices,
                     int32_t n_triangles, const float* matrix16, const float* wheels, int32_t n_wheels,
                     wt_voxel_info* info);

/* ---- stepping --------------------------------------------------------------------------- */

/* encodeSteps(enc, steps), SolverGPU.ts:204-222: runs a pending init, then `steps` fused
 * pull-stream + collide steps. Asynchronous. */
int wt_step(wt_solver* s, int32_t steps);
/* stepAndSample(steps), SolverGPU.ts:292-309: steps, then macro + momentum exchange + reduction,
 * then a synchronous read of the decoded statistics. steps may be 0. */
int wt_step_and_sample(wt_solver* s, int32_t steps, wt_stats* out);
/* stepAndSample(steps) for hosts with an event loop (SolverGPU.ts:292-309: encode and submit, then
 * await the mapped result): wt_step_and_sample_async enqueues the steps, the macro pass, the
 * reduction and a copy of it into a pinned slot, and returns a ticket without waiting, so the order
 * of work on the handle is the order of the calls. wt_sample_wait(ticket) blocks until that copy is
 * complete and decodes it (the step is the one at the async call); it touches only the ticket's slot
 * and may run on another thread while other calls use the handle (not wt_destroy). Every ticket
 * must be waited for once; the slots grow with the samples in flight. The async path does not
 * change what wt_sample_raw returns. */
int wt_step_and_sample_async(wt_solver* s, int32_t steps, int64_t* ticket);
int wt_sample_wait(wt_solver* s, int64_t ticket, wt_stats* out);
/* encodePost(enc, derived), SolverGPU.ts:225-261: enqueues the macro pass and the force / statistics
 * reduction on the current populations and, with derived != 0, the derived pass (wt_derive) and an
 * asynchronous copy of the reduction into a free slot of a ring of four pinned read slots; returns
 * without waiting. *queued (may be NULL) is 1 when a read was queued, 0 when derived == 0 or every
 * slot was still busy (that post is then not read back, as in the reference). */
int wt_encode_post(wt_solver* s, int32_t derived, int32_t* queued);
/* flushReads(cb), SolverGPU.ts:264-287, as a non-blocking poll: the statistics of every queued read
 * whose copy has completed, decoded as stepAndSample decodes them, with the step of its post, in
 * posting order (at most cap of them, into out); their slots are freed. *pending (may be NULL): reads
 * still in flight. The caller keeps the latest by step and polls again for the pending ones. */
int wt_flush_reads(wt_solver* s, wt_stats* out, int32_t cap, int32_t* count, int32_t* pending);
/* The undecoded 8-value reduction of the last sample as the reference's result buffer holds it
 * (Fx, Fy, Fz, max u^2, sum rho_ref, n_ref, nan count, solid count), wgslLbm.ts:491-521. */
int wt_sample_raw(wt_solver* s, double* out8);
/* Block until the handle's stream is idle. */
int wt_sync(wt_solver* s);
/* Queue fences, for the reference's frame throttle (backend/webgpu.ts:92-94, :120: frames in flight
 * counted with device.queue.onSubmittedWorkDone()). The handle's stream is its queue: wt_fence
 * marks the point after all work queued on the handle so far and returns its ticket (> 0);
 * wt_fence_done sets *done to 1 once the GPU has passed that point (non-blocking; it may report 0
 * a little longer for a ticket whose record was reused after 16 newer fences); wt_fence_wait
 * blocks until it has. Tickets from another handle, or not yet issued, are WT_E_ARG. */
int wt_fence(wt_solver* s, int64_t* ticket);
int wt_fence_done(wt_solver* s, int64_t ticket, int32_t* done);
int wt_fence_wait(wt_solver* s, int64_t ticket);

/* ---- readback / upload ------------------------------------------------------------------ */

/* readPopulations(), SolverGPU.ts:359-369, but unpacked to full f_i (q * n floats). */
int wt_read_populations(wt_solver* s, float* out);
/* readPopulations() restricted to the planes z0 .. z0 + count - 1 (no reference counterpart: a grid
 * of 10^8 cells does not fit one host array): full f_i, direction-major within the slab,
 * out[i * m + c], m = nx * ny * count, c = x + nx * (y + ny * (z - z0)). */
int wt_read_population_planes(wt_solver* s, int32_t z0, int32_t count, float* out);
/* No reference counterpart: overwrite the current populations (full f_i, q * n floats) after any
 * pending init; rounds to the storage format. Seeded initial conditions and checkpoints. */
int wt_upload_populations(wt_solver* s, const float* f);
/* Macro fields of the last sample: velocity texture (ux, uy, uz, rho per cell, SolverGPU velTex),
 * EMA means meanA (4 per cell) and meanB (1 per cell). Any pointer may be NULL. */
int wt_read_macro(wt_solver* s, float* vel4, float* mean_a4, float* mean_b);

/* ---- derived fields and export ----------------------------------------------------------------
 * Derived pass of encodePost(derived = true) (SolverGPU.ts:242-246, derivedWGSL wgslLbm.ts:527-592)
 * over the fields of the last sample (wt_step_and_sample / wt_sample_raw): vorticity and the
 * Q-criterion (faded out within two cells of a solid) from central differences of the velocity
 * record, turbulence intensity sqrt(max(meanB - |meanA.xyz|^2, 0) / 3) / max(uRef, 1e-6) and a copy
 * of the EMA mean, each stored as four IEEE halves per cell (the reference's rgba16float textures).
 * The buffers (24 bytes per cell) are allocated, zeroed, on first use. */
int wt_derive(wt_solver* s);
/* Full-resolution readback of the derived fields, 4 floats per cell each: (wx, wy, wz, Q),
 * (TI, solid, 0, 0), mean (ux, uy, uz, rho). Any pointer may be NULL. */
int wt_read_derived(wt_solver* s, float* vort4, float* stat4, float* mean4);
/* SolverGPU.exportFields (SolverGPU.ts:374-425): output dims ceil(n / stride) per axis, point g
 * samples cell min(g * stride, n - 1); vel4 (ux, uy, uz, rho), mean4 (EMA mean), vort4 (wx, wy, wz, Q),
 * 4 floats per point, x fastest (FieldExport, solver/types.ts:45-55). Reads the fields as the last
 * sample and derived pass left them. With all three pointers NULL only out_dims is written.
 * wt_export_dims: the output dims alone. stride < 1 is WT_E_ARG. */
int wt_export_dims(const wt_solver* s, int32_t stride, int32_t* out_dims);
int wt_export_fields(wt_solver* s, int32_t stride, float* vel4, float* mean4, float* vort4, int32_t* out_dims);

/* ---- probes, streamlines and particle tracers (the compute passes of render/tracers.ts) --------
 * They sample the velocity record of the last sample (and, for the mean, the derived pass's mean
 * record) like the reference's linear, clamp-to-edge textures: texel centres at cell centres,
 * trilinear (here in FP32), the single layer in 2D. Points are in cell units. */
typedef struct wt_rake {  /* seeding line a -> b, parallel lines offset by d (xyz used, w ignored) */
  float a[4], b[4], d[4];
} wt_rake;
/* PROBES_WGSL (tracers.ts:109-125): out8 per point = field (ux, uy, uz, rho), mean (ux, uy, uz, rho);
 * z components 0 in 2D. pos4: 4 floats per point (xyz used). */
int wt_probe(wt_solver* s, const float* pos4, int32_t count, float* out8);
/* STREAMLINES_WGSL (tracers.ts:152-200): `seeds` lines seeded per_line per rake line; `points`
 * points each, step h along the normalised velocity (midpoint rule) of the field (use_mean: the
 * mean field); out4 = (x, y, z, speed), speed -1 once the line has left the domain, entered a
 * solid or slowed below 1e-3 U. */
int wt_streamlines(wt_solver* s, const wt_rake* rake, int32_t seeds, int32_t per_line, int32_t points, float h, int32_t use_mean,
                   float* out4);
/* PARTICLES_WGSL (tracers.ts:3-106): a particle system on the solver's velocity field, state
 * (x, y, z, age) per particle and a ring buffer of `trail` past positions. The handle must be
 * destroyed before its solver. */
typedef struct wt_particles wt_particles;
typedef struct wt_particle_opts {
  int32_t count, trail;
  int32_t emitter;   /* 0: inlet plane x in [1, 3), 1: rake nozzles */
  int32_t nozzles, lines;
  float max_age;     /* steps of life before a forced respawn: 4 x max_age frames */
  wt_rake rake;
} wt_particle_opts;
int wt_particles_create(wt_solver* s, const wt_particle_opts* o, wt_particles** out);
int wt_particles_destroy(wt_particles* p);
/* One frame (RendererGPU.ts:505-527): fill = fresh positions (and trail); else with advance the
 * trail head moves on and every particle advances by `steps` lattice steps. `seed` feeds the
 * per-particle hash generator (the renderer passes a frame counter). */
int wt_particles_update(wt_particles* p, float steps, uint32_t seed, int32_t fill, int32_t advance);
/* state4: count x 4 floats; hist4: trail x count x 4 floats (slot-major); head: current slot.
 * Any pointer may be NULL. */
int wt_particles_read(wt_particles* p, float* state4, float* hist4, int32_t* head);

/* ---- profiling -------------------------------------------------------------------------- */

/* Step-kernel timing with HIP events on the handle's stream (one launch per step: wall cells and
 * bulk cells). WT_PROFILE_LAUNCH binds a start / stop event pair to every launch
 * (hipExtLaunchKernelGGL): exact per-launch durations, but each bound pair holds the queue for a
 * few microseconds between launches. WT_PROFILE_SPAN records one pair around the launches of each
 * wt_step call: the summed time is the stream's busy span, i.e. the launches plus the (sub-
 * microsecond) gaps between back-to-back launches, with nothing added to the queue between them.
 * wt_profile_read synchronises and returns the summed time (ms), the number of launches it covers,
 * and 0 in the third slot (kept for hosts that read it); then clears the accumulators. */
enum { WT_PROFILE_OFF = 0, WT_PROFILE_LAUNCH = 1, WT_PROFILE_SPAN = 2 };
int wt_profile_enable(wt_solver* s, int32_t on);
int wt_profile_read(wt_solver* s, double* bulk_ms, int64_t* launches, double* boundary_ms);

/* Number of fluid cells the bulk part of the step kernel updates and the number of cells its wall
 * part updates (cells next to the body). */
int wt_work_split(const wt_solver* s, int64_t* bulk_cells, int64_t* boundary_cells);

/* HBM calibration (no reference counterpart): copies `bytes` between two fresh device buffers
 * `iters` times with a plain streaming kernel and returns read + write GB/s. */
int wt_copy_probe(int64_t bytes, int32_t iters, double* gbs);

/* Launch tuning knobs (no reference counterpart; set by the autotuner).
 * WT_TUNE_ITEM_CHUNK: the bulk items (64 or 128-cell chunks of x-rows) are cut into runs of this
 * many consecutive items, dealt round-robin to the 8 XCDs; 4 (default) is the plain order, 0 gives
 * every XCD one contiguous eighth of the grid.
 * WT_TUNE_CELLS_PER_THREAD: 2 (default; pairs along x, needs an even nx) or 1. FP16 / BF16
 * storage always uses 2. Changing it rebuilds the wall-cell list. */
enum { WT_TUNE_ITEM_CHUNK = 0, WT_TUNE_CELLS_PER_THREAD = 1 };
int wt_set_tuning(wt_solver* s, int32_t key, int32_t value);
int wt_get_tuning(const wt_solver* s, int32_t key, int32_t* value);

/* Autotuner (no reference counterpart): times `steps` steps for every candidate launch
 * configuration (cells per thread x item-run length), keeps the fastest on the handle and reports
 * all trials. The populations, step counter and profiling state are restored afterwards. */
#define WT_TUNE_MAX_TRIALS 16
typedef struct wt_tune_result {
  int32_t chunk;             /* chosen WT_TUNE_ITEM_CHUNK */
  int32_t cells_per_thread;  /* chosen WT_TUNE_CELLS_PER_THREAD */
  double ms_per_step;        /* its time per step */
  int32_t trials;
  int32_t pad;
  int32_t trial_chunk[WT_TUNE_MAX_TRIALS];
  int32_t trial_cpt[WT_TUNE_MAX_TRIALS];
  double trial_ms[WT_TUNE_MAX_TRIALS];
} wt_tune_result;
int wt_autotune(wt_solver* s, int32_t steps, wt_tune_result* out);

/* ---- batched sweeps (no reference counterpart: the reference runs a study's cases one after
 * another, studies.ts:223-252) ------------------------------------------------------------------
 * A batch steps many independent solver handles (e.g. a yaw / ride-height sweep of uneven grids)
 * with ONE kernel launch per step: a device descriptor table gives every block its domain's
 * parameters, layout and buffers. Steps run as replays of a hipGraph-captured chunk (default 100
 * steps) plus direct launches for the remainder. The members must share lattice, storage
 * precision, cells per thread and collision, must have been stepped equally often, and stay
 * owned by the caller (destroy the batch first). Geometry, parameter and item-order changes on a
 * member are picked up at the next wt_batch_run; a change that breaks the shared kernel variant
 * (cells per thread, collision operator, a loaded network with the learned closure, which the
 * batched step does not run) makes wt_batch_run fail with WT_E_STATE and step nothing. Per-member
 * statistics are read with wt_batch_stats (one wt_stats per member, in member order). */
typedef struct wt_batch wt_batch;
int wt_batch_create(wt_solver* const* members, int32_t count, wt_batch** out);
int wt_batch_destroy(wt_batch* b);
/* chunk_steps >= 2 (rounded up to even); use_graph = 0 launches every step directly. */
int wt_batch_set_chunk(wt_batch* b, int32_t chunk_steps, int32_t use_graph);
int wt_batch_run(wt_batch* b, int32_t steps);
int wt_batch_sync(wt_batch* b);
int wt_batch_stats(wt_batch* b, wt_stats* out);

/* ---- learned subgrid closure (no reference counterpart; it replaces the Smagorinsky tau of
 * wgslLbm.ts:214 when params.closure = WT_CLOSURE_MLP) --------------------------------------------
 * A per-cell MLP 8-32-32-1 (ReLU hidden layers, softplus output) predicts the eddy viscosity nu_t;
 * the cell relaxes with tau = t0 + 3 nu_t (t0: tau0 + outlet sponge, ground-wall floor). Inputs:
 * six rotation invariants of the strain rate S = -3 Pi^neq / (2 rho tau_smag) (tr S, the second
 * and third principal invariants, tr S^2, tr S^3, |S|; each compressed as sign(x) log(1 + |x| /
 * 1e-3^degree)), the wall distance (cells to the nearest solid cell centre, capped at 4) and the
 * local Reynolds number |u| max(d, 1/2) / nu. Inference runs in the step kernel on BF16 MFMA
 * (FP32 accumulation) for bulk cells. Smagorinsky guards it: nu_t is at least smag_floor times
 * the Smagorinsky eddy viscosity of the cell, cells next to the body and non-finite outputs keep
 * Smagorinsky, which is also the closure when no network is loaded. D3Q19 only. */
#define WT_MLP_NPARAM 1377
typedef struct wt_mlp {
  float params[WT_MLP_NPARAM]; /* W1[32][8], b1[32], W2[32][32], b2[32], w3[32], b3 (out x in, row-major) */
  float mu[8];                 /* input normalisation: (x - mu) / sigma */
  float sigma[8];
  float nu_scale;              /* nu_t = nu_scale * min(softplus(z), 10) */
  float smag_floor;            /* nu_t >= smag_floor * Smagorinsky's nu_t */
} wt_mlp;
int wt_mlp_load(wt_solver* s, const wt_mlp* net);
/* Wall distance per cell as bytes (32 per cell, capped at 128). */
int wt_read_wall_distance(wt_solver* s, uint8_t* out);
/* nu_t of n cells from their raw (unnormalised) inputs (n x 8), on the step kernel's MFMA path. */
int wt_mlp_infer(const wt_mlp* net, const float* raw, int64_t n, float* nu);
/* No reference counterpart. wt_mlp_infer for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, hosts that forbid device allocation): the same arguments, results
 * and kernels, but the network and the staged inputs and results live in a static device array of
 * the library (2.3 MB per GPU, held until the process ends) instead of memory allocated per call,
 * and the work runs on the current device's null stream. No hipMalloc, hipFree or stream creation.
 * At most WT_MLP_NOALLOC_MAX cells per call (more is WT_E_ARG). Calls on one GPU share that
 * array: calls that overlap in time on one GPU (other threads) race; each call synchronises
 * before it returns. */
#define WT_MLP_NOALLOC_MAX 65536
int wt_mlp_infer_noalloc(const wt_mlp* net, const float* raw, int64_t n, float* nu);

/* Training data: records of 10 floats (8 raw inputs, target nu_t, coarse cell index as bits) in
 * device memory. wt_dataset_add box-filters a fine D3Q19 run 2x onto the grid of `coarse` (same
 * geometry at half resolution; it supplies the wall distances and nu): per coarse cell the block
 * mean velocity, the subgrid stress tau_ij = <u_i u_j> - <u_i><u_j>, the filtered strain rate by
 * central differences and the least-squares eddy viscosity of the 3x3x3 neighbourhood,
 * nu_t = max(0, sum(-tau_dev:S) / sum(2 S:S));
 * cells touching a solid or a domain face are skipped. Records are appended in cell order. */
typedef struct wt_dataset wt_dataset;
int wt_dataset_create(int64_t capacity, wt_dataset** out);
int wt_dataset_destroy(wt_dataset* d);
int wt_dataset_add(wt_dataset* d, wt_solver* fine, wt_solver* coarse, int64_t* added);
int wt_dataset_count(const wt_dataset* d, int64_t* n);
int wt_dataset_read(const wt_dataset* d, int64_t first, int64_t count, float* out);
int wt_dataset_append(wt_dataset* d, const float* records, int64_t count);

/* GPU training with hand-written kernels (FP32 forward / backward, FP64 gradient reduction, Adam):
 * loss = mean (softplus(z) - min(nu_t / nu_scale, 10))^2 over mini-batches of `batch` records in
 * a fresh random order every epoch (`max_samples` > 0 limits the records drawn per epoch).
 * net->params is the initial state on entry and the trained state on return; mu, sigma and
 * nu_scale are used as given. epoch_loss (epochs entries) may be NULL. */
typedef struct wt_train_opts {
  int32_t epochs;
  int32_t batch;
  double lr, beta1, beta2, eps;
  uint64_t seed;
  int64_t max_samples;
} wt_train_opts;
int wt_mlp_train(wt_dataset* d, wt_mlp* net, const wt_train_opts* o, double* epoch_loss);
/* Loss and its gradient over records [first, first + count) (one mini-batch, for checking). */
int wt_mlp_grad(wt_dataset* d, const wt_mlp* net, int64_t first, int64_t count, float* grad, double* loss);

#ifdef __cplusplus
}
#endif

#endif /* WT_H */
