/*
 * lf_abi.h -- C ABI of libfluid_hip.so, the MI355X (gfx950) WCSPH timestep.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes.
 * Every function returns LF_OK (0) or a negative LF_ERR_* code; the message
 * of the last failure on the calling thread is returned by lf_last_error().
 *
 * The reference (NCCA/libfluid) has no C ABI: its boundary is the C++ class
 * FluidSystem (libfluid/include/fluidsystem.h:16-126).  Each entry point
 * below names the reference member it replaces; include/fluidsystem.h
 * re-creates that class on top of these functions.
 */
#ifndef LF_ABI_H
#define LF_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LF_ABI_VERSION 1

/* Cell key of a particle outside the grid (NULL_HASH,
 * common/include/fluiddefines.h:12). */
#define LF_NULL_HASH 0xFFFFFFFFu

enum lf_status {
    LF_OK = 0,
    LF_ERR_INVALID = -1,  /* bad argument */
    LF_ERR_HIP = -2,      /* HIP runtime failure */
    LF_ERR_ALLOC = -3,    /* device or host allocation failed */
    LF_ERR_STATE = -4,    /* call not valid in the scene's current state */
    LF_ERR_OVERFLOW = -5, /* kept for source compatibility; never returned (cells of any occupancy are sorted) */
    LF_ERR_FULL = -6      /* a batch has no room for another scene */
};

/* Particle order for lf_download. */
enum lf_order {
    LF_ORDER_CELL = 0, /* current cell-sorted order, as exportToData copies it (fluidsystem.cu:168-172) */
    LF_ORDER_ID = 1    /* particle-id order (the order particles were uploaded in) */
};

/*
 * Scene parameters.  Same quantities as FluidParamData
 * (libfluid/cudasrc/fluidparams.cuh:16-69) with the unit cube generalised:
 * the fluid lives in the box [0,box], the uniform grid has grid[a] cells of
 * size 1/inv_cell[a] along axis a, and the smoothing length h is independent
 * of the grid (the reference ties h = 1/res, fluidparams.cu:62).  Cells must
 * not be smaller than h.  Derived constants (h powers, kernel coefficients)
 * are computed by the library.
 */
typedef struct lf_params {
    float rest_density;    /* rho0 */
    float gas_constant;    /* K: Tait stiffness */
    float gamma;           /* Tait exponent */
    float mass;            /* per-particle mass */
    float particle_size;   /* boundary clamp margin (particle radius) */
    float xsph_scale;      /* XSPH blend */
    float surface_tension; /* sigma */
    float viscosity;       /* nu */
    float adhesion;        /* alpha */
    float vel_limit;
    float acc_limit;
    float vdamp;           /* boundary velocity damping */
    float gravity[3];
    float h;               /* smoothing length */
    float inv_cell[3];     /* 1 / cell size per axis (the reference's res) */
    uint32_t grid[3];      /* cells per axis */
    float box[3];          /* box extents; boundary SDF box is [0,box] */
} lf_params;

typedef struct lf_scene lf_scene;

/* ---- library -------------------------------------------------------- */
int lf_version(void);
const char *lf_last_error(void);
/* Number of visible HIP devices (0 when no GPU; never initialises a context). */
int lf_device_count(int *count);

/* ---- parameters ------------------------------------------------------ */
/* Defaults of FluidSystem::init (fluidsystem.cu:196-235) on the unit cube with
 * a res^3 grid and h = 1/res (FluidParams::setRes, fluidparams.cu:51-89). */
int lf_params_default(lf_params *p, uint32_t res);
/* Validates p (positive h and cells, cells >= h, finite coefficients). */
int lf_params_check(const lf_params *p);

/* ---- scene lifetime (FluidSystem::init / clear / ~FluidSystem,
 *      fluidsystem.cu:18-52,196-265) ------------------------------------ */
/* num_particles must be in [1, 2^28). */
int lf_create(const lf_params *p, uint32_t num_particles, lf_scene **out);
int lf_destroy(lf_scene *s);
int lf_num_particles(const lf_scene *s, uint32_t *n);

/* Setters of FluidSystem (fluidsystem.cu:412-475) collapse into one call;
 * the grid (inv_cell, grid, box) cannot change after lf_create. */
int lf_set_params(lf_scene *s, const lf_params *p);
int lf_get_params(const lf_scene *s, lf_params *p);

/* ---- data in / out -------------------------------------------------- */
/* Writes particles with ids [first, first+count) from host float3 arrays
 * (x,y,z interleaved).  vel/acc may be NULL (zero).  Replaces the per-point
 * writes of createFluidBlock (fluidsystem.cu:121) and the copies of
 * setupFromData (fluidsystem.cu:187-188). */
int lf_upload(lf_scene *s, uint32_t first, uint32_t count, const float *pos_xyz, const float *vel_xyz,
              const float *acc_xyz);
/* Copies the state to host float3 arrays (any pointer may be NULL), in the
 * given lf_order.  colour is DensityColourOperator of the last density
 * (fluidutil.cu:67-80).  Replaces exportToData (fluidsystem.cu:154-173). */
int lf_download(lf_scene *s, int order, float *pos_xyz, float *colour_rgb, float *vel_xyz, float *acc_xyz,
                uint32_t *ids);

/* ---- stepping (FluidSystem::advance, fluidsystem.cu:273-399) ---------- */
/* Enqueues `substeps` substeps of dt/substeps on the scene's HIP stream and
 * returns without waiting. */
int lf_step(lf_scene *s, float dt, uint32_t substeps);
/* Waits for the scene's stream and reports any deferred device error. */
int lf_sync(lf_scene *s);
/* Number of substeps run so far. */
int lf_step_count(const lf_scene *s, uint64_t *count);

/* ---- inspection of the last substep (parity tests, tooling) ---------- */
enum lf_field {
    LF_FIELD_KEYS = 0,        /* uint32[n]: cell key of each particle, previous order (NULL_HASH outside) */
    LF_FIELD_START = 1,       /* uint32[ncells+2]: cell start table; [ncells] = first out-of-grid slot */
    LF_FIELD_SORTED_IDS = 2,  /* uint32[n]: particle id in each slot of the new cell order */
    LF_FIELD_DENSITY = 3,     /* float4[n]: (density, pressure, sound speed, 1/density), cell order */
    LF_FIELD_NORMAL = 4,      /* float4[n]: (normal xyz, 0), cell order; computed when read
                                 (from the last substep's data) for scenes without surface
                                 tension, which skip them while stepping */
    LF_FIELD_FORCES = 5,      /* float4[4n]: pressure, viscosity, tension, adhesion per particle
                                 (written only with LF_DEBUG_FORCES), cell order */
    LF_FIELD_VEL_HALF = 6,    /* float4[n]: velocity after integration, before XSPH, cell order */
    LF_FIELD_SORTED_POS = 7   /* float4[n]: pre-integration positions in cell order */
};
/* Copies `bytes` bytes of a raw device field into dst (synchronous). */
int lf_get_field(lf_scene *s, int field, void *dst, size_t bytes);
/* Debug flags: LF_DEBUG_FORCES also writes the four force components
 * (LF_FIELD_FORCES) each substep; LF_DEBUG_SCAN makes the normals, forces and
 * XSPH passes test every slot of the 27-cell stencil instead of walking the
 * neighbour masks the density pass records; LF_DEBUG_NORMALS_PASS runs the
 * normals pass every substep also when surface tension is zero (otherwise
 * the normals, which feed only the tension force, are computed when read).
 * Results are bitwise the same either way. */
#define LF_DEBUG_FORCES 1
#define LF_DEBUG_SCAN 2
#define LF_DEBUG_NORMALS_PASS 4
int lf_set_debug(lf_scene *s, int flags);
/* Neighbour sets of the density pass for the current state (same FP32
 * predicate as the kernels), by particle id: counts[n], and when list is not
 * NULL the ids of each set, ascending, concatenated in id order.
 * *total receives the number of pairs; cap bounds `list`. */
int lf_neighbours(lf_scene *s, uint32_t *counts, uint32_t *list, uint64_t cap, uint64_t *total);

/* Host-side check of the flattened neighbour walk's table (no counterpart
 * in the reference): for one particle's nine stencil rows, masks[r] (high
 * word zero for a row of at most 32 slots) and first_slot[r], builds the
 * table the normals, forces and XSPH passes walk with the kernels' own code:
 * the non-empty 32-slot pseudo-rows in row order, (entry_mask[k],
 * entry_base[k]) for k < min(*count, 16); *count is the number of non-empty
 * pseudo-rows (above 16 the kernels walk row by row instead).  The entry
 * arrays hold 16 elements. */
int lf_debug_flat_table(const uint64_t masks[9], const uint32_t first_slot[9], uint32_t *entry_mask,
                        uint32_t *entry_base, uint32_t *count);
/* Test hooks for the allocation-failure paths (no counterpart in the
 * reference): after lf_debug_fail_alloc(k) the library's (k+1)-th device
 * allocation from now fails with LF_ERR_ALLOC (k < 0: none fails);
 * lf_debug_live_allocs reports how many device allocations the library
 * holds, so a test can check that a failed create releases each exactly
 * once. */
int lf_debug_fail_alloc(int64_t nth);
int lf_debug_live_allocs(int64_t *count);

/* ---- stepping without allocating -------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory: work recorded under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  Runs `substeps` substeps of dt/substeps
 * of one scene with parameters p on device arrays the caller owns: pos_xyz,
 * vel_xyz and acc_xyz hold num_particles float3 (x,y,z interleaved) in
 * particle-id order; they are read first and overwritten with the new state.
 * The result is bitwise that of lf_create + lf_upload + lf_step(dt,
 * substeps) + lf_download(LF_ORDER_ID) with the default launch shapes.  The
 * call allocates nothing (no device, pinned or heap memory, no stream, event
 * or graph): every working buffer lies in a fixed 32 MiB device array of the
 * library, one per device, held from the first call on that device until the
 * process ends.  That fits about 350 bytes per particle (the count rounded
 * up to 512) plus 8 bytes per grid cell: e.g. 65,536 particles on up to
 * 1,000,000 cells; larger scenes are rejected with LF_ERR_INVALID.  Work is
 * only enqueued on `stream` (a hipStream_t; NULL: the default stream);
 * completion is the stream's.  Calls share the device's array, so calls that
 * overlap on one device (other streams or threads) race. */
int lf_step_noalloc(const lf_params *p, uint32_t num_particles, float dt, uint32_t substeps, float *pos_xyz,
                    float *vel_xyz, float *acc_xyz, void *stream);

/* ---- timing ------------------------------------------------------------ */
/* Kernel classes timed by lf_profile_*. */
enum lf_kernel {
    LF_K_HASH = 0, LF_K_SCAN = 1, LF_K_SCATTER = 2, LF_K_GATHER = 3,
    LF_K_DENSITY = 4, LF_K_NORMALS = 5, LF_K_FORCES = 6, LF_K_XSPH = 7,
    LF_K_COUNT = 8
};
/* While on, every kernel launch of lf_step is bracketed by HIP events on the
 * scene's stream.  lf_profile_read waits, then reports per kernel class the
 * launch count and the summed event time in milliseconds, and resets. */
int lf_profile_enable(lf_scene *s, int on);
int lf_profile_read(lf_scene *s, uint64_t *launches, double *ms);
/* Plain device-to-device copy kernel over `bytes` bytes (multiple of 16) of
 * scene-owned scratch; used to calibrate hardware byte counters.  Returns
 * the event-timed duration in ms. */
int lf_copy_probe(lf_scene *s, size_t bytes, int repeats, double *ms);

/* ---- formats either side of the stepping path ------------------------ */
/* Checkpoint: an exact binary snapshot (parameters, step count, last substep
 * length, and per particle in id order the double-float position, velocity
 * and acceleration) with a checksum.  A scene loaded from it continues
 * bitwise as the saved scene would have.  The reference has no checkpoint;
 * its round trip exportToData -> setupFromData (fluidsystem.cu:154-189) keeps
 * FP32 positions and velocities only, and no acceleration. */
int lf_checkpoint_save(lf_scene *s, const char *path);
int lf_checkpoint_load(const char *path, lf_scene **out);
/* Device-side setup: positions of particles [first, ...) generated on the
 * GPU, the same values as lf_jittered_lattice / lf_fluid_block produce on the
 * host (createFluidBlock, fluidsystem.cu:66-127, whose N single-point
 * host-to-device copies, :121, this replaces).  Velocities and accelerations
 * are left as they are. */
int lf_setup_jittered_lattice(lf_scene *s, uint32_t first, uint32_t nx, uint32_t ny, uint32_t nz, float dx,
                              const float origin[3], float jitter, uint64_t seed);
int lf_setup_fluid_block(lf_scene *s, uint32_t first, uint32_t n, const float mn[3], const float mx[3],
                         uint32_t *written, float *mass, float *particle_size);
/* Asynchronous .geo export (the reference driver's dumpToGeo after
 * exportToData, test/src/main.cpp:10-74,137-150, done synchronously there):
 * the current state is snapshotted on the scene's stream, copied to host
 * memory on a second stream and written by a background thread, in the given
 * lf_order, while stepping goes on.  The files are byte-identical to
 * lf_download + lf_write_geo at the same state.  lf_export_wait waits for
 * every pending file and reports the first error since the last wait. */
int lf_export_geo_async(lf_scene *s, const char *filebase, uint32_t cnt, int order);
int lf_export_wait(lf_scene *s);

/* ---- launch shapes (no counterpart in the reference, whose launch shape is
 *      derived from the largest cell occupancy every substep,
 *      fluidsystem.cu:319-323) ------------------------------------------- */
/* Workgroup sizes of the density, normals, forces and XSPH passes, each 128,
 * 256 or 512 (default 256).  Results do not depend on them. */
int lf_set_tuning(lf_scene *s, const int block_sizes[4]);
int lf_get_tuning(const lf_scene *s, int block_sizes[4]);
/* Walk modes of the normals, forces and XSPH passes: 1 walks the density
 * pass's neighbour masks as one flattened stream of set bits (in waves whose
 * stencil rows all fit 64-slot masks), 0 walks them row by row, 2 flattens
 * only the waves in which a row holds more than 32 slots.  Default {1, 1, 2}
 * (measured fastest: the force walk's arithmetic-heavy steps gain from the
 * flattened stream everywhere, the XSPH walk's gather-heavy ones only where
 * rows are long); lf_autotune picks per pass.  Results do not depend on
 * them. */
int lf_set_walks(lf_scene *s, const int flat[3]);
int lf_get_walks(const lf_scene *s, int flat[3]);
/* Times each block size on the scene's current state (`reps` substeps each,
 * state restored) and keeps the fastest per pass; then, at those block
 * sizes, the three walk modes, keeping the fastest per pass.  ms (may be NULL)
 * receives the 12 block-size averages: pass-major, block sizes 128, 256, 512.
 * Needs one lf_step first (the substep length). */
int lf_autotune(lf_scene *s, uint32_t reps, double *ms);
/* On (default): lf_step replays one captured hipGraph per substep; off: every
 * kernel is launched individually.  Profiling (lf_profile_enable) always
 * launches individually.  Results are the same either way. */
int lf_set_graphs(lf_scene *s, int on);

/* ---- batches of independent scenes ------------------------------------
 * The reference steps one FluidSystem at a time; a batch keeps many scenes
 * (own parameters, box, grid, gravity, particle count and substep length)
 * resident in one pool of device memory and advances all active ones with
 * one launch per stage.  Scenes can be added and removed between steps.
 * Each scene's results are bitwise those of the same scene in its own
 * lf_scene.  Scene handles are small integers, reused after removal. */
typedef struct lf_batch lf_batch;
/* Device bytes a batch of the given capacity allocates. */
int lf_batch_footprint(uint32_t max_particles, uint32_t max_cells, uint32_t max_scenes, uint64_t *bytes);
/* max_particles: total particle slots (each scene takes its count rounded up
 * to 512); max_cells: total grid cells plus one per scene; max_scenes:
 * simultaneous scenes. */
int lf_batch_create(uint32_t max_particles, uint32_t max_cells, uint32_t max_scenes, lf_batch **out);
int lf_batch_destroy(lf_batch *b);
/* Admits a scene (state like lf_create: all particles at the origin until
 * uploaded) with substep length dt; LF_ERR_FULL when it does not fit. */
int lf_batch_add(lf_batch *b, const lf_params *p, uint32_t num_particles, float dt, uint32_t *scene);
int lf_batch_remove(lf_batch *b, uint32_t scene);
int lf_batch_set_params(lf_batch *b, uint32_t scene, const lf_params *p);
int lf_batch_set_dt(lf_batch *b, uint32_t scene, float dt);
/* Inactive scenes keep their state but are not stepped. */
int lf_batch_set_active(lf_batch *b, uint32_t scene, int active);
/* As lf_upload / lf_download, for one scene. */
int lf_batch_upload(lf_batch *b, uint32_t scene, uint32_t first, uint32_t count, const float *pos_xyz,
                    const float *vel_xyz, const float *acc_xyz);
int lf_batch_download(lf_batch *b, uint32_t scene, int order, float *pos_xyz, float *colour_rgb, float *vel_xyz,
                      float *acc_xyz, uint32_t *ids);
/* Enqueues `substeps` substeps of every active scene (each its own dt). */
int lf_batch_step(lf_batch *b, uint32_t substeps);
int lf_batch_sync(lf_batch *b);
int lf_batch_scene_info(const lf_batch *b, uint32_t scene, uint32_t *num_particles, uint64_t *steps, int *active);
int lf_batch_usage(const lf_batch *b, uint32_t *scenes, uint32_t *slots_used, uint32_t *slot_cap,
                   uint32_t *buckets_used, uint32_t *bucket_cap);
/* The formats above, for one scene of a batch (checkpoint_load admits it). */
int lf_batch_checkpoint_save(lf_batch *b, uint32_t scene, const char *path);
int lf_batch_checkpoint_load(lf_batch *b, const char *path, uint32_t *scene);
int lf_batch_setup_jittered_lattice(lf_batch *b, uint32_t scene, uint32_t first, uint32_t nx, uint32_t ny,
                                    uint32_t nz, float dx, const float origin[3], float jitter, uint64_t seed);
int lf_batch_setup_fluid_block(lf_batch *b, uint32_t scene, uint32_t first, uint32_t n, const float mn[3],
                               const float mx[3], uint32_t *written, float *mass, float *particle_size);
int lf_batch_export_geo_async(lf_batch *b, uint32_t scene, const char *filebase, uint32_t cnt, int order);
int lf_batch_export_wait(lf_batch *b);
int lf_batch_set_tuning(lf_batch *b, const int block_sizes[4]);
int lf_batch_get_tuning(const lf_batch *b, int block_sizes[4]);
int lf_batch_set_walks(lf_batch *b, const int flat[3]);
int lf_batch_get_walks(const lf_batch *b, int flat[3]);
int lf_batch_autotune(lf_batch *b, uint32_t reps, double *ms);
int lf_batch_set_graphs(lf_batch *b, int on);
int lf_batch_profile_enable(lf_batch *b, int mask);
int lf_batch_profile_read(lf_batch *b, uint64_t *launches, double *ms);

/* ---- serving: continuous batching of scene jobs -----------------------
 * No counterpart in the reference.  A job is a scene (parameters, initial
 * state) plus a number of substeps of length dt.  lf_server_tick admits
 * queued jobs into free pool memory in arrival order (LF_SERVE_BACKFILL lets
 * a later job pass one that does not fit yet; LF_SERVE_FIFO does not),
 * advances every running job by one substep in one batched step, and retires
 * jobs that reached their step count: their final state is kept in host
 * memory until lf_server_job_result and their pool memory is reused at
 * once.  Ticks do not wait for the device (at most two batched steps are
 * queued ahead of it); a job is reported finished once its final state has
 * reached host memory.  A job's result is bitwise what the same scene gives
 * alone in an lf_scene.  Single-threaded: one caller drives a server. */
typedef struct lf_server lf_server;
enum lf_serve_policy { LF_SERVE_FIFO = 0, LF_SERVE_BACKFILL = 1 };
enum lf_job_state { LF_JOB_QUEUED = 0, LF_JOB_RUNNING = 1, LF_JOB_DONE = 2 };
typedef struct lf_server_stats {
    uint64_t submitted, admitted, finished, queued, running;
    uint64_t ticks;              /* lf_server_tick calls */
    uint64_t step_ticks;         /* ticks that stepped at least one job */
    uint64_t particle_steps;     /* sum over stepping ticks of the particles stepped */
    uint32_t resident_particles; /* particles of running jobs now */
    uint32_t slot_cap;           /* particle slots of the pool */
    double busy_seconds;         /* wall time spent inside lf_server_tick */
    double step_p50_ms, step_p99_ms; /* device time of one batched step (all running jobs) */
    double job_p50_ms, job_p99_ms;   /* submit to finish, finished jobs */
    double wait_p50_ms, wait_p99_ms; /* submit to admission, admitted jobs */
    double mean_occupancy;       /* particles stepped / slot_cap, mean over stepping ticks */
} lf_server_stats;
/* Capacity as for lf_batch_create (see lf_batch_footprint for the bytes). */
int lf_server_create(uint32_t max_particles, uint32_t max_cells, uint32_t max_scenes, int policy, lf_server **out);
int lf_server_destroy(lf_server *s);
/* Copies the initial state (vel may be NULL); LF_ERR_FULL if the job could
 * never fit this server. */
int lf_server_submit(lf_server *s, const lf_params *p, uint32_t num_particles, const float *pos_xyz,
                     const float *vel_xyz, float dt, uint32_t steps, uint64_t *job);
/* One serving iteration; *finished = jobs whose results arrived during it. */
int lf_server_tick(lf_server *s, uint32_t *finished);
/* run_ahead: batched steps a tick may queue ahead of the device (default 2;
 * 0 makes every tick wait for its step and retirements); graphs: replay
 * captured substeps (default on).  Results do not depend on either. */
int lf_server_configure(lf_server *s, uint32_t run_ahead, int graphs);
/* Moves up to cap ids of jobs finished since the last call into jobs[]. */
int lf_server_take_finished(lf_server *s, uint64_t *jobs, uint32_t cap, uint32_t *count);
int lf_server_job_status(const lf_server *s, uint64_t job, int *state, uint32_t *steps_done);
/* Final state of a finished job in particle-id order (any pointer may be
 * NULL) and its submit-to-finish latency; forgets the job. */
int lf_server_job_result(lf_server *s, uint64_t job, float *pos_xyz, float *colour_rgb, float *vel_xyz,
                         float *acc_xyz, double *latency_ms);
int lf_server_get_stats(const lf_server *s, lf_server_stats *out);

/* ---- host-side helpers (no device needed) ---------------------------- */
/* createFluidBlock (fluidsystem.cu:66-127): lattice of up to n points in the
 * box [mn, mx], x outermost and z innermost, FP32 float accumulators as in
 * the reference.  Writes the points, returns the count written, the mass
 * (particle volume x rho0) and the particle size (half the spacing). */
int lf_fluid_block(uint32_t n, const float mn[3], const float mx[3], float rest_density, float *pos_xyz,
                   uint32_t *written, float *mass, float *particle_size);
/* Jittered lattice: nx*ny*nz points at origin + (i+0.5)*dx (x outermost,
 * z innermost), each coordinate displaced by U(-jitter, jitter) drawn from
 * SplitMix64(seed) in point order. */
int lf_jittered_lattice(uint32_t nx, uint32_t ny, uint32_t nz, float dx, const float origin[3], float jitter,
                        uint64_t seed, float *pos_xyz);
/* n points uniform in [lo, hi) from SplitMix64(seed). */
int lf_uniform_points(uint32_t n, const float lo[3], const float hi[3], uint64_t seed, float *pos_xyz);
/* Houdini PGEOMETRY V5 ASCII writer, byte-compatible with the reference
 * driver's dumpToGeo (test/src/main.cpp:10-74); writes "<base>.<cnt:04d>.geo". */
int lf_write_geo(const char *filebase, uint32_t cnt, const float *pos_xyz, const float *colour_rgb,
                 const float *vel_xyz, uint32_t n, float pscale);

#ifdef __cplusplus
}
#endif
#endif /* LF_ABI_H */
