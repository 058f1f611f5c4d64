/*
 * mp_abi.h - C ABI of the MI355X particle mover (push + tet walk + nodal scatter).
 *
 * This is the drop-in boundary for mini-pic's MoveParticles hot path.  Every entry
 * point takes plain pointers and sizes, returns an int status (MP_OK == 0) and never
 * throws.  A context owns its device buffers and one HIP stream; all work is
 * stream-ordered on that stream.  One context per host thread, or serialise calls.
 *
 * Reference interfaces replaced (AD1641/mini-fem-pic-main):
 *   src/mini-pic.cpp:39-53    struct DeviceData (ctor uploads Tetra table, owns buffers)
 *   src/mini-pic.cpp:268-320  MoveParticles (copy in, kernel, copy out, erase_if, normalise)
 *   src/mini-pic.cpp:287-308  the device kernel: push, XtoLtet walk, FP64 atomic scatter
 *   src/mini-pic.cpp:348-355  DeviceData::reserveParticles (grow to 2n)
 *   src/meshes.h:75-100       XtoLtet barycentric walk
 *   src/particles.h:20-25     struct Particle (88 B layout kept: mp_particle)
 *   src/meshes.h:33-43        struct Tetra    (168 B layout kept: mp_tetra)
 */
#ifndef MP_ABI_H
#define MP_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------------- */
#define MP_OK            0
#define MP_ERR_ARG       1   /* bad argument (null pointer, size out of range, ...)  */
#define MP_ERR_HIP       2   /* HIP runtime error                                     */
#define MP_ERR_ALLOC     3   /* device or host allocation failed                      */
#define MP_ERR_STATE     4   /* call out of order (no field uploaded, no particles)   */
#define MP_ERR_NODEV     5   /* no usable gfx950 device                               */

/* ---- modes -------------------------------------------------------------------------- */
#define MP_SCATTER_ATOMIC        0  /* FP64 atomics, order nondeterministic (reference)  */
#define MP_SCATTER_DETERMINISTIC 1  /* sort (node,pid) then serial fold per node          */

#define MP_WALL_ABSORB   0  /* walk exits through a boundary face: particle is lost (reference) */
#define MP_WALL_REFLECT  1  /* specular reflection about the boundary face plane (extension)    */

/* ---- records with the reference byte layout ----------------------------------------- */
/* Tetra, src/meshes.h:33-43: con@0 volume@16 alpha@24 beta@56 gamma@88 delta@120 cell_con@152 */
typedef struct mp_tetra {
    int32_t con[4];       /* node ids                                       */
    double  volume;       /* tet volume (positive)                          */
    double  alpha[4];     /* precomputed 3x3 determinants for barycentrics  */
    double  beta[4];
    double  gamma[4];
    double  delta[4];
    int32_t cell_con[4];  /* neighbour across the face opposite node i, -1 = boundary */
} mp_tetra;

/* Particle, src/particles.h:20-25: pos@0 vel@24 lc@48 cell_index@80, size 88 */
typedef struct mp_particle {
    double  pos[3];
    double  vel[3];
    double  lc[4];        /* barycentric weights, written by the walk       */
    int32_t cell_index;   /* containing tet, -1 = left the domain           */
    int32_t _pad;
} mp_particle;

/* Tuning knobs of the device-resident path (mp_step). */
typedef struct mp_options {
    int32_t sort_interval;         /* cell-sort + compact every k steps (0 = never)    */
    int32_t wall_mode;             /* MP_WALL_ABSORB / MP_WALL_REFLECT                 */
    int32_t block_size;            /* threads per workgroup of the push kernel (64, 128, 256;
                                      the placing push of det_algo 2 runs its own: 512, or 256
                                      with dense nodes cut into pid slabs) */
    int32_t particles_per_thread;  /* 1, 2, 4 or 8: rounds of block_size particles per thread */
    int32_t walk_cap;              /* max face moves per particle per step             */
    int32_t scatter_agg;           /* atomic-mode aggregation: 0 one atomic per weight,
                                      2 per workgroup by node (LDS hash)                */
    int32_t sort_shift;            /* cell sort key = cell >> sort_shift (0: exact cell) */
    int32_t keep_pid_order;        /* 1: the periodic pass is a stable compaction that keeps the
                                      particles in pid order instead of a cell sort; the
                                      deterministic records are then written contiguously by
                                      slot.  0 (default): cell sort */
    int32_t gather_sort;           /* mp_step_many with the cell sort: 1 the sort's rank pass
                                      writes only a permutation and the next multi-step
                                      launch reads each particle through it into the other
                                      buffer; 0 (default): the sort copies the records */
    int32_t sort_order;            /* cell sort key: 0 (default) cell >> sort_shift; 1 the
                                      tet's rank along a Morton curve through the tet
                                      centroids, >> sort_shift (buckets compact in space:
                                      fewer distinct nodes per workgroup where particles are
                                      sparse, a few per tet) */
    int32_t det_algo;              /* deterministic scatter: 2 (default) the push places each
                                      particle's records into slab segments sized from the last
                                      step's counts (a node's one slab, or a dense node's pid
                                      slabs as 1 cuts them), and each slab is put in pid order
                                      on chip and folded (a step whose records do not fit, or
                                      with a dense node's slab over 512 records, is done as
                                      with 1); 1 the push
                                      counts each node's records, which are rebuilt from the
                                      pushed state and grouped into node segments (dense nodes
                                      cut into pid slabs); 0 the records are written in pid
                                      order and a device radix sort on the node bits groups
                                      them */
    int32_t sort_pid_bits;         /* cell sort (sort_now) key: 0 (default) the bucket; b > 0
                                      (up to 12) the bucket refined by the top b bits of a pid
                                      digit, so the particles of a dense bucket lie in pid
                                      ranges: the deterministic scatter's pid slabs of a dense
                                      node then take few keys per tile */
} mp_options;

/* Counters accumulated on the device since mp_set_particles / mp_reset_counters. */
typedef struct mp_counters {
    int64_t particle_steps;   /* sum over steps of particles active at step start   */
    int64_t lost;             /* particles that left the domain                     */
    int64_t crossed;          /* particle-steps that moved to at least one other tet */
    int64_t moves;            /* total face moves (+ reflections)                   */
    int64_t cap_hits;         /* particles stopped by walk_cap (treated as lost)    */
    int64_t reflections;      /* wall reflections (MP_WALL_REFLECT)                 */
} mp_counters;

typedef struct mp_ctx mp_ctx;

/* ---- lifetime ---------------------------------------------------------------------- */
/* DeviceData::DeviceData (src/mini-pic.cpp:322-338): checks FP64 atomics, uploads the
 * Tetra table once.  node_volume[n] is Node::volume (src/meshes.h:29). */
int  mp_create(const mp_tetra* tets, int32_t n_tets, const double* node_volume,
               int32_t n_nodes, mp_ctx** out);
/* DeviceData::~DeviceData (src/mini-pic.cpp:340-345). */
void mp_destroy(mp_ctx* ctx);
/* Message of the last failed call on ctx (ctx == NULL: last failed mp_create). */
const char* mp_last_error(const mp_ctx* ctx);

int  mp_get_options(const mp_ctx* ctx, mp_options* opt);
int  mp_set_options(mp_ctx* ctx, const mp_options* opt);
/* Shorthand for the sort interval alone: cell-sort (or compact) every k steps, 0 = never. */
int  mp_set_sort_interval(mp_ctx* ctx, int32_t k);

/* FESolver::ef (src/FESolver.h:49): per-element E, ef[n_tets][3], copied to the device.
 * MoveParticles copies it every step (src/mini-pic.cpp:276). */
int  mp_upload_field(mp_ctx* ctx, const double* ef, int32_t n_tets);

/* ---- drop-in step over the reference's particle array -------------------------------- */
/* MoveParticles (src/mini-pic.cpp:268-320) on an AoS Particle array in host memory:
 * uploads n particles, pushes, walks and scatters on the device, downloads, removes the
 * particles with cell_index < 0 keeping the order of the others (std::erase_if), and
 * writes den_out[n_nodes] = raw * (spwt / node_volume).  *n_inout becomes the number of
 * particles kept; parts[0..*n_inout) holds them. */
int  mp_step_aos(mp_ctx* ctx, mp_particle* parts, size_t* n_inout, double q_over_m,
                 double dt, double spwt, double* den_out, int32_t mode);

/* ---- device-resident path (structure of arrays in HBM) ------------------------------- */
/* Replace the resident particle set.  pos, vel: [n][3]; cell: [n].  Particle ids (pid)
 * are 0..n-1 in array order; pid order is the order a serial MoveParticles sees. */
int  mp_set_particles(mp_ctx* ctx, const double* pos, const double* vel,
                      const int32_t* cell, int64_t n);
/* Append particles (injection); they get the next pids.  A batch larger than an eighth of
 * the resident set is cell-sorted in before the next step; smaller ones join the next
 * periodic sort. */
int  mp_add_particles(mp_ctx* ctx, const double* pos, const double* vel,
                      const int32_t* cell, int64_t n);
/* One MoveParticles step on the resident set: push + walk + scatter, then the nodal
 * density normalised by spwt / node_volume.  Stream-ordered, does not block the host. */
int  mp_step(mp_ctx* ctx, double q_over_m, double dt, double spwt, int32_t mode);
/* n_steps consecutive mp_step calls under the field last uploaded (no reference counterpart:
 * for runs whose field stays frozen).  Atomic mode pushes every stretch of steps between two
 * periodic sorts in one launch, the particles kept in registers from step to step, so the
 * state crosses HBM once per launch; aggregation 0 or 2, 256-thread
 * workgroups, one particle per thread.  Deterministic mode runs the steps one by one.
 * Particles, counters and sorts are those of n_steps mp_step calls; the normalised density
 * of every step is kept for mp_get_density_step, and the last one for mp_get_density. */
int  mp_step_many(mp_ctx* ctx, double q_over_m, double dt, double spwt, int32_t mode, int32_t n_steps);
/* Normalised nodal density of step k (0-based) of the last mp_step_many, den[n_nodes]
 * (blocks). */
int  mp_get_density_step(mp_ctx* ctx, int32_t k, double* den);
/* Block until all queued work on the context's stream is done. */
int  mp_sync(mp_ctx* ctx);
/* Number of resident particles still in the domain (blocks). */
int  mp_num_active(mp_ctx* ctx, int64_t* n_active);
/* Diagnostics (no reference counterpart): the resident slots in their current order, which
 * is bucket order after a cell sort, lost slots included until the next sort or compaction:
 * cell[i] (-1 = lost) and pid[i] of slot i.  Either pointer may be NULL.  *n_inout: capacity
 * in, number of slots out; MP_ERR_ARG (with the count) when the capacity is short (blocks). */
int  mp_get_slots(mp_ctx* ctx, int32_t* cell, int32_t* pid, int64_t* n_inout);
/* Diagnostics (no reference counterpart): rank[e] = position of tet e along the Morton curve
 * of sort_order 1 (a permutation of 0..n_tets-1; blocks). */
int  mp_get_tet_order(mp_ctx* ctx, int32_t* rank);
/* Download the particles still in the domain in ascending pid order (blocks).
 * Any output pointer may be NULL.  *n_inout: capacity in, count out. */
int  mp_get_particles(mp_ctx* ctx, double* pos, double* vel, int32_t* cell, int64_t* pid,
                      int64_t* n_inout);
/* Normalised nodal density of the last step, den[n_nodes] (blocks). */
int  mp_get_density(mp_ctx* ctx, double* den);
int  mp_get_counters(mp_ctx* ctx, mp_counters* out);
int  mp_reset_counters(mp_ctx* ctx);
/* Deterministic steps since mp_set_particles / mp_reset_counters by how their records were
 * grouped: *placed by the push into slab segments sized from the last step (det_algo 2),
 * *rebuilt from the pushed state (det_algo 1, and det_algo 2 steps whose records did not fit
 * or that had a dense node's slab over 512 records).  Either pointer may be NULL. */
int  mp_get_det_paths(mp_ctx* ctx, int64_t* placed, int64_t* rebuilt);
/* Sort the resident particles by cell (or, with keep_pid_order, compact them in pid order)
 * and drop the lost ones now. */
int  mp_sort_now(mp_ctx* ctx);
/* The context's hipStream_t, for callers that capture or time its work. */
void* mp_stream(mp_ctx* ctx);

/* Push-kernel timing with HIP events recorded on the context's stream around every push
 * launch while enabled.  mp_set_timing(ctx, 1) clears and starts recording, 0 stops.
 * mp_get_timing blocks and returns the summed push-kernel time (ms) and launch count. */
int  mp_set_timing(mp_ctx* ctx, int32_t enable);
int  mp_get_timing(mp_ctx* ctx, double* push_ms, int64_t* push_launches);

/* ---- autotuner (no reference counterpart; the SYCL reference has no tunables) ----------- */
/* One timed candidate: the options and the measured time per step (ms). */
typedef struct mp_tune_trial {
    mp_options opt;
    double     ms_per_step;
} mp_tune_trial;
/* Choose particles_per_thread, block_size, scatter_agg (atomic mode), sort_shift,
 * sort_order, sort_interval (and det_algo, keep_pid_order, sort_pid_bits in deterministic mode) for the resident particles and the given step parameters:
 * coordinate descent over the knobs, each candidate timed (HIP events) over a window of
 * max(interval, min_steps) steps that starts from the same saved state and holds one cell
 * sort per sort interval, as a long run does; a candidate replaces the incumbent only when
 * more than 1% faster, and the winner is timed twice more against the starting options,
 * keeping the faster.  Results do not depend on the options, so
 * only time is compared.  The particle state, counters and density are restored before
 * returning; the chosen options are set on ctx and written to *best (may be NULL) with their
 * time per step.  trials[0..max_trials) receives the timed candidates in order (may be NULL
 * when max_trials is 0); *n_trials their number.  Needs one extra particle-state copy of
 * device memory while it runs. */
int  mp_autotune(mp_ctx* ctx, double q_over_m, double dt, double spwt, int32_t mode, int32_t min_steps,
                 mp_options* best, double* best_ms_per_step, mp_tune_trial* trials, int32_t max_trials,
                 int32_t* n_trials);
/* OR into mp_autotune's mode: time the candidates through mp_step_many (frozen-field runs)
 * and sweep only the knobs its kernel takes (sort shift, sort interval, aggregation 0 or 2,
 * gather_sort). */
#define MP_TUNE_MULTI_STEP 0x100

/* ---- batched ensemble: many independent runs on one mesh (no reference counterpart) -- */
/* An ensemble of n_runs independent particle sets pushed through the context's mesh and
 * field.  Each run has its own particles, q/m, dt, spwt and nodal density.  One step of
 * every run is one push launch over all particles (captured once in a hipGraph and
 * replayed when graphs are enabled); particle counts live in device memory so the graph
 * stays valid as particles are lost.  The batch uses the context's stream and options. */
typedef struct mp_batch mp_batch;
int  mp_batch_create(mp_ctx* ctx, int32_t n_runs, mp_batch** out);
void mp_batch_destroy(mp_batch* b);
/* Set run `run`'s particles (pos, vel [n][3], cell [n]) and parameters; every run must be
 * set before the first step.  Local pids are 0..n-1 in array order. */
int  mp_batch_set_run(mp_batch* b, int32_t run, const double* pos, const double* vel,
                      const int32_t* cell, int64_t n, double q_over_m, double dt, double spwt);
/* One step of every run in one launch (graph replay when enabled, the default). */
int  mp_batch_step(mp_batch* b, int32_t mode);
/* The same step as n_runs separate per-run launches (the sequential baseline). */
int  mp_batch_step_sequential(mp_batch* b, int32_t mode);
/* n_steps mp_batch_step calls under the frozen field; atomic mode pushes every stretch of
 * steps between two periodic sorts in one launch with the particles kept in registers (as
 * mp_step_many does), deterministic mode steps one by one.  Every step's densities are kept
 * for mp_batch_get_density_step, the last step's for mp_batch_get_density.
 * n_runs * n_nodes must stay below 2^32. */
int  mp_batch_step_many(mp_batch* b, int32_t mode, int32_t n_steps);
/* Normalised density of run `run` at step k (0-based) of the last mp_batch_step_many,
 * den[n_nodes] (blocks). */
int  mp_batch_get_density_step(mp_batch* b, int32_t k, int32_t run, double* den);
int  mp_batch_set_graph(mp_batch* b, int32_t enable);
int  mp_batch_sync(mp_batch* b);
/* Normalised density of run `run` after the last step, den[n_nodes] (blocks). */
int  mp_batch_get_density(mp_batch* b, int32_t run, double* den);
/* Run `run`'s particles still in the domain, ascending local pid (blocks). */
int  mp_batch_get_particles(mp_batch* b, int32_t run, double* pos, double* vel, int32_t* cell,
                            int64_t* pid, int64_t* n_inout);
int  mp_batch_get_counters(mp_batch* b, mp_counters* out);
int  mp_batch_reset_counters(mp_batch* b);

/* ---- adjoint of K frozen-field steps (no reference counterpart) ------------------------ */
/* Runs K MoveParticles steps from the given initial particles under the context's field and
 * wall mode (absorbing or reflecting walls), with L = sum_n (den_K[n] - den_target[n])^2 (den_K
 * normalised by spwt / node volume), and returns L, dL/dE [n_tets][3] and dL/d(drift) where
 * the initial velocities are u + drift * drift_dir.  Derivatives are those of the piece of the
 * piecewise-smooth loss the inputs sit in (every particle's tet path and reflections held
 * fixed).  The forward pass stores the particle state every checkpoint_every steps; the
 * backward pass replays each segment, recording the tets crossed (and, with reflecting walls,
 * each step's start state, 48 B per particle and step), and sweeps it in reverse.  A step that
 * reflects one particle more than 4 times is reported as MP_ERR_STATE.  Replaces the context's
 * resident particle set (it holds the state after K steps on return). */
int  mp_adjoint(mp_ctx* ctx, const double* pos0, const double* vel0, const int32_t* cell0, int64_t n,
                double q_over_m, double dt, double spwt, int32_t steps, const double* den_target,
                const double drift_dir[3], int32_t checkpoint_every, double* loss, double* dL_dE,
                double* dL_ddrift);
/* mp_adjoint with a scatter mode.  MP_SCATTER_ATOMIC is mp_adjoint (FP64 atomics: the sums
 * of dL/dE vary in the last bits from run to run).  MP_SCATTER_DETERMINISTIC gives the same
 * bits on every run: the last step's density is the ascending-pid serial fold, and dL/dE is
 * summed exactly in 128-bit fixed point (scaled by a power of two chosen from a bound on every
 * contribution) and rounded to FP64 once; loss and dL/d(drift) are summed in a fixed order. */
int  mp_adjoint_ex(mp_ctx* ctx, const double* pos0, const double* vel0, const int32_t* cell0, int64_t n,
                   double q_over_m, double dt, double spwt, int32_t steps, const double* den_target,
                   const double drift_dir[3], int32_t checkpoint_every, double* loss, double* dL_dE,
                   double* dL_ddrift, int32_t mode);

/* ---- drop-in step with no allocation (no reference counterpart) ------------------------ */
/* For callers that cannot allocate device memory: code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc.  The MoveParticles step of mp_step_aos, atomic
 * scatter only, over Particle records that the caller already holds in DEVICE memory:
 * parts[0..n) are pushed, walked and scattered in place (lc written, cell_index = -1 for the
 * lost), the survivors are copied in order to kept[0..*n_kept) (device, capacity n, must not
 * overlap parts), and den_out[n_nodes] (device) = raw * (spwt / node_volume).  The step's
 * scratch (keep flags, prefix sums, tallies) is one fixed device array per GPU of 32.4 MiB,
 * held until the process ends; n must not exceed MP_NOALLOC_MAX_PARTICLES (MP_ERR_ARG beyond).
 * Blocks until done.  Calls that overlap on one GPU share that array and race: serialise
 * them across host threads and contexts. */
#define MP_NOALLOC_MAX_PARTICLES (1 << 22)
int  mp_step_aos_noalloc(mp_ctx* ctx, mp_particle* parts, int64_t n, mp_particle* kept, int64_t* n_kept,
                         double q_over_m, double dt, double spwt, double* den_out, int32_t mode);

#ifdef __cplusplus
}
#endif
#endif /* MP_ABI_H */
