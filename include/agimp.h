/* C-ABI of the MI355X adaptive-GIMP substep library (libagimp.so).
 *
 * Drop-in boundary for the simulator API of g1n0st/adaptive-gimp (adaptive_gimp.py): the
 * Python facade AdaptiveGIMP binds these through ctypes.  Plain pointers and sizes only.
 *
 * Conventions
 *  - Every call returns 0 on success or a negative AGIMP_E* code; agimp_last_error(h) then
 *    holds a message.  Nothing aborts the process.
 *  - Host arrays are borrowed for the duration of the call; device memory is owned by the
 *    handle and freed by agimp_destroy.
 *  - Particle arrays use the reference's field layout and ORIGINAL particle order: x, v
 *    [n][dim], F, C [n][dim][dim] row-major, m, Jp [n] float32; g, fl [n] int32.
 *  - Grid readbacks are dense over the reference extents e_l = 4*align_size(n_l/4+1, 4)
 *    (adaptive_gimp.py:36-37): cell_mask / node_mask int32 [e]^dim, m float32 [e]^dim,
 *    v float32 [e]^dim[dim]; inactive elements read as 0.
 *  - A handle owns one HIP stream; calls are stream-ordered and not thread-safe; the
 *    agimp_get_* calls synchronise.
 */
#ifndef AGIMP_H
#define AGIMP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AGIMP_OK 0
#define AGIMP_EINVAL (-1)   /* bad argument */
#define AGIMP_EHIP (-2)     /* HIP runtime error */
#define AGIMP_ESTATE (-3)   /* call not valid in the current state */
#define AGIMP_ENOMEM (-4)   /* device allocation failed */

typedef struct agimp_handle agimp_handle;
typedef struct agimp_batch agimp_batch;

#define AGIMP_MIN_CHUNK 64    /* particles per P2G work chunk: smallest */
#define AGIMP_MAX_CHUNK 2048  /* ... and largest */

/* Simulator configuration.  Defaults (agimp_config_init) are the reference constants:
 * adaptive_gimp.py:18 (boundary_gap 4), :48-53 (radius 0.5, E 5e3, nu 0.2, gravity -100 on
 * axis 1), an implicit particle volume of 1 (the stress scale at :240 has none), slip walls
 * (:284-287) and no plasticity. */
typedef struct agimp_config {
  int32_t dim;             /* 2 or 3 */
  int32_t level;           /* octree levels, 1..6 (adaptive_gimp.py:14) */
  int32_t coarsest_size;   /* cells per axis at level 0 (adaptive_gimp.py:15) */
  int32_t boundary_gap;    /* finest cells from a wall where walls act (adaptive_gimp.py:18) */
  int64_t n_particles;     /* adaptive_gimp.py:51 */
  float E, nu;             /* adaptive_gimp.py:52 */
  float gravity[3];        /* adaptive_gimp.py:50 */
  float p_vol;             /* particle volume multiplying the stress term (reference: 1) */
  float radius;            /* GIMP half width in cells (adaptive_gimp.py:48); must be 0.5 */
  int32_t wall_mode;       /* 0 per-axis normal zeroing (reference), 1 sticky */
  int32_t plastic;         /* 0 none (reference), 1 singular-value clamp */
  float theta_c, theta_s;  /* clamp to [1-theta_c, 1+theta_s] */
  float hardening;         /* xi in exp(xi (1 - Jp)) scaling mu, lambda */
  int32_t slab_enable;     /* sticky slab collider on grid nodes */
  int32_t slab_axis;
  float slab_lo, slab_hi;
  int32_t particles_per_chunk; /* P2G work chunk (particles of one leaf block), AGIMP_MIN/MAX_CHUNK;
                                 agimp_config_init: 1024 */
  int32_t device;          /* HIP device ordinal */
} agimp_config;

/* Fill cfg with the reference defaults for (dim, level, coarsest_size, n_particles). */
int agimp_config_init(agimp_config* cfg, int32_t dim, int32_t level, int32_t coarsest_size, int64_t n_particles);

/* adaptive_gimp.py:6-66 AdaptiveGIMP.__init__: allocate state on the device. */
int agimp_create(const agimp_config* cfg, agimp_handle** out);
int agimp_destroy(agimp_handle* h);

/* adaptive_gimp.py:64 particle_initializer(self): upload particle fields (original order);
 * Jp may be NULL (all 1).  Recomputes the stored Kirchhoff stress. */
int agimp_set_particles(agimp_handle* h, const float* x, const float* v, const float* F, const float* C,
                        const float* m, const int32_t* g, const float* Jp);

/* adaptive_gimp.py:65-66, :72-85 static adaptivity: the grid initializer evaluated over the
 * padded finest cell box [-4, e_{L-1}+4)^dim, row-major int8: ret_level[I] is what
 * grid_initializer(sim, I) returns, act_level[I] the level it passes to activate_cell
 * (adaptive_gimp.py:68-70) or -1.  Passing NULLs returns to dynamic adaptivity. */
int agimp_set_static_levels(agimp_handle* h, const int8_t* ret_level, const int8_t* act_level);

/* Topology rebuild of adaptive_gimp.py:324-335 (clear, activation, culling, node/T/ghost
 * marks) from the current particles; the physics is not advanced. */
int agimp_refine(agimp_handle* h);

/* adaptive_gimp.py:323-352 substep(dt0): refine, level mapping, P2G, restriction, grid
 * update, prolongation, G2P. */
int agimp_substep(agimp_handle* h, float dt);

/* substep without the topology rebuild: level mapping + physics on the current octree
 * (used when a predicted octree is reused; not in the reference). */
int agimp_substep_fixed_topology(agimp_handle* h, float dt);

/* n substeps back to back on the device, no host round trips: one substep is captured in a
 * hipGraph on first use and replayed n times (eager launches while timing is on). */
int agimp_advance(agimp_handle* h, float dt, int32_t n);

/* Particles per P2G work chunk (the autotuned parameter), AGIMP_MIN_CHUNK..AGIMP_MAX_CHUNK. */
int agimp_set_particles_per_chunk(agimp_handle* h, int32_t ch);

/* Read particle fields (adaptive_gimp.py:55-62) in original order; any pointer may be NULL. */
int agimp_get_particles(agimp_handle* h, float* x, float* v, float* F, float* C, float* Jp, int32_t* fl,
                        int32_t* g);

/* agimp_get_particles without any device allocation, for callers that cannot allocate
 * (allocation-free serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  Same arguments and results; the device staging goes through one file-scope
 * scratch array of 64 MiB per GPU instead of a per-call allocation, which limits it to scenes
 * of at most 64 MiB / (4 (2d + 2d^2 + 1) + 8) particles: 621,378 in 3D, 1,118,481 in 2D; larger
 * scenes get AGIMP_EINVAL.  All calls on one GPU share the array: calls that overlap on one GPU
 * (from several threads) race. */
int agimp_get_particles_noalloc(agimp_handle* h, float* x, float* v, float* F, float* C, float* Jp, int32_t* fl,
                                int32_t* g);

/* Dense cell_mask[l], ad_grid_m[l], ad_grid_v[l] (adaptive_gimp.py:39-42); any may be NULL. */
int agimp_get_level(agimp_handle* h, int32_t l, int32_t* cell_mask, float* m, float* v);

/* Dense node_mask (adaptive_gimp.py:23-26), finest extent. */
int agimp_get_node_mask(agimp_handle* h, int32_t* node_mask);

/* Counters accumulated since creation (or the last reset): [0] out-of-box accesses,
 * [1] frozen particle-substeps (fl_p = -1), [2] transferred particle-substeps,
 * [3] P2G chunks, [4] non-finite particles, [5] particles that failed the refinement
 * criterion check (only counted once agimp_learned_enable has switched the check on),
 * [6], [7] bookkeeping of the learned windows. */
int agimp_get_counters(agimp_handle* h, int64_t* out, int32_t n);
int agimp_reset_counters(agimp_handle* h);

/* Phase timing with HIP events recorded on the handle's stream around each phase of every
 * substep (the Taichi kernel profiler of run_3d.py:12, :43 is the reference's analogue).
 * agimp_get_timings synchronises, returns total milliseconds and call counts per phase
 * (AGIMP_PHASE_*) accumulated since the last call, and resets them. */
#define AGIMP_PHASE_REFINE 0  /* clear + activation + culling + node/T/ghost marks */
#define AGIMP_PHASE_BUCKET 1  /* level mapping + counting sort + block lists */
#define AGIMP_PHASE_P2G 2     /* the P2G kernel */
#define AGIMP_PHASE_GRID 3    /* restriction + grid update + prolongation */
#define AGIMP_PHASE_G2P 4     /* the G2P kernel (incl. plasticity and stress) */
#define AGIMP_NPHASES 5
int agimp_set_timing(agimp_handle* h, int32_t on);
int agimp_get_timings(agimp_handle* h, double* ms, int64_t* calls, int32_t n);
/* Per-kernel GPU time while timing is on (the Taichi kernel profiler's per-kernel table,
 * ti.profiler.print_kernel_profiler_info at run_3d.py:43): HIP events recorded on the handle's
 * stream around every kernel launch of the eager path.  Synchronises, returns what accumulated
 * since the last call and resets it.  Entry k: names + k*name_len (NUL-terminated), total ms,
 * launches, shortest and longest launch in ms; at most cap entries are written and *count
 * receives the number of kernels.  Any output pointer but count may be NULL. */
int agimp_get_kernel_timings(agimp_handle* h, char* names, int32_t name_len, double* total_ms, int64_t* calls,
                             double* min_ms, double* max_ms, int32_t cap, int32_t* count);

/* Reference extent e_l of level l. */
int agimp_level_extent(agimp_handle* h, int32_t l);

int agimp_sync(agimp_handle* h);
const char* agimp_last_error(agimp_handle* h);

/* Batches (not in the reference): many independent simulators of one dimension advanced
 * together.  Every kernel of a substep is launched once for all scenes (one workgroup row
 * per scene), and the whole substep is captured in one hipGraph and replayed.  The handles
 * keep ownership of their state and stay usable on their own; do not use them concurrently
 * with the batch.  dt[i] is scene i's time step; agimp_batch_advance returns when done. */
int agimp_batch_create(agimp_handle* const* scenes, int32_t n, agimp_batch** out);
int agimp_batch_destroy(agimp_batch* b);
int agimp_batch_set_particles_per_chunk(agimp_batch* b, const int32_t* ch);
int agimp_batch_advance(agimp_batch* b, const float* dt, int32_t steps);
const char* agimp_batch_last_error(agimp_batch* b);

/* Learned refinement prediction (not in the reference, whose octree is rebuilt from the
 * current particles every substep, adaptive_gimp.py:323-337).
 *
 * Units are the cells of level max(0, level-3), i.e. blocks of 4^dim finest cells.  For each
 * unit within one unit of a particle-holding one ("candidates"), 16 features are computed
 * from the particles (mass fraction, |v|, |grad v| = |C|, det F, level tags, neighbour mass
 * and levels, mean displacement over the window, upwind level).  A 16-64-64-4 network
 * (agimp_mlp_*) classifies each candidate by the finest level the plain criterion will
 * request in it during the next `window` substeps: 0 finest, 1 one coarser, 2 two coarser,
 * 3 none.  The predicted octree is the plain one of the current particles plus two parts
 * (agimp_learned_set_parts): the network's (in every unit of class c < 3, all cells of level
 * level-1-c if that is not coarser than the unit) and the swept requests (each particle's
 * request cells at g_p between its base cell now and the base cell of x + v * window dt).
 *
 * agimp_learned_advance reuses one predicted octree for up to `window` substeps (captured
 * bucketing + physics replays).  Every substep checks the plain criterion in the level
 * mapping: each particle's 2^dim request cells at level g_p (adaptive_gimp.py:87-96) must be
 * covered at level >= g_p.  If a check fails, the window is rolled back to its start (a
 * device snapshot of the particles), the substeps before the failing one are replayed, and
 * the octree is rebuilt there: no substep runs on an octree coarser than the criterion
 * requires. */
#define AGIMP_MLP_NPARAMS 5508
#define AGIMP_LSTAT_REBUILDS 0    /* predicted octrees built */
#define AGIMP_LSTAT_FORCED 1      /* windows cut short by a failed criterion check */
#define AGIMP_LSTAT_DISCARDED 2   /* substeps run and then rolled back */
#define AGIMP_LSTAT_REPLAYED 3    /* substeps replayed after a rollback */
#define AGIMP_LSTAT_SUBSTEPS 4    /* substeps advanced by agimp_learned_advance */
#define AGIMP_LSTAT_VIOLATIONS 5  /* particle checks that failed (before the rollbacks) */
#define AGIMP_LSTAT_CANDIDATES 6  /* candidate units of the last prediction */
#define AGIMP_LSTAT_REFINED 7     /* units refined by the last prediction */
#define AGIMP_NLSTATS 8

/* Allocate the prediction state and switch the criterion check on; tau in (0, 1] is the
 * decision threshold (class c = first whose cumulative probability reaches tau). */
int agimp_learned_enable(agimp_handle* h, int32_t window, float tau);
/* Network parameters (host, AGIMP_MLP_NPARAMS floats, layout of agimp_mlp_forward). */
int agimp_learned_set_model(agimp_handle* h, const float* params, int32_t n);
/* Candidates of the current particles: features [count][16] and unit ids (block-major cell
 * index of the unit level); copies min(count, cap) rows, returns the total in *count. */
int agimp_learned_features(agimp_handle* h, float dt, float* feat, int32_t* units, int64_t cap, int64_t* count);
/* Window labels for training: reset, then accumulate the plain requests of the current
 * particles before each of the window's substeps, then read the class of every candidate of
 * the last agimp_learned_features call (0 finest .. 3 none). */
int agimp_learned_labels_reset(agimp_handle* h);
int agimp_learned_labels_accumulate(agimp_handle* h);
int agimp_learned_labels(agimp_handle* h, int32_t* cls, int64_t cap);
/* Features + network + decision on the current particles: classes of the candidates and,
 * if level != NULL, the predicted level of every unit (int8, -1 none). */
int agimp_learned_predict(agimp_handle* h, float dt, int32_t* cls, int8_t* level, int64_t cap, int64_t* count);
/* n substeps with predicted octrees (see above). */
int agimp_learned_advance(agimp_handle* h, float dt, int32_t n);
#define AGIMP_PREDICT_NETWORK 1
#define AGIMP_PREDICT_SWEEP 2
/* Which parts a predicted octree holds besides the current requests (default both). */
int agimp_learned_set_parts(agimp_handle* h, int32_t parts);
int agimp_learned_stats(agimp_handle* h, int64_t* out, int32_t n);
int agimp_learned_reset_stats(agimp_handle* h);

/* The network on its own (device pointers, ordered on `stream`, NULL = default stream).
 * Parameters: W1[64][16] b1[64] W2[64][64] b2[64] W3[4][64] b3[4] (row-major, FP32).
 * Products run on BF16 matrix cores with FP32 accumulation: inputs, weights and hidden
 * activations are rounded to BF16 before each product. */
int agimp_mlp_forward(const float* X, int64_t rows, const float* params, float* logits, void* stream);
/* loss = scale * sum_r w[y_r] * CE(softmax(z_r), y_r) and its gradient [5508]; w may be NULL
 * (all 1).  grad and loss are overwritten. */
int agimp_mlp_loss_grad(const float* X, const int32_t* y, int64_t rows, const float* params,
                        const float* class_weight, float scale, float* grad, float* loss, void* stream);
/* One Adam step (bias-corrected; decoupled weight decay on the weight matrices). */
int agimp_mlp_adam(float* params, const float* grad, float* m, float* v, int32_t step, float lr, float beta1,
                   float beta2, float eps, float weight_decay, void* stream);

/* Device-resident training of the network (no other GPU library needed): the trainer owns
 * the dataset (capacity rows of 16 features and a class 0..3), the parameters, the Adam
 * moments and scratch, all on `device`.  Host arrays are borrowed for the call. */
typedef struct agimp_trainer agimp_trainer;
int agimp_trainer_create(int64_t capacity, int32_t device, agimp_trainer** out);
int agimp_trainer_destroy(agimp_trainer* t);
int agimp_trainer_set_data(agimp_trainer* t, const float* X, const int32_t* y, int64_t rows);
/* set the parameters and reset the Adam state */
int agimp_trainer_set_params(agimp_trainer* t, const float* params);
int agimp_trainer_get_params(agimp_trainer* t, float* params);
/* One pass over the dataset in the order perm[0..rows) (a permutation of the rows): for every
 * full batch (and a single short one when rows < batch), the gradient of
 * scale * sum cw[y] CE over the batch and one Adam step (beta 0.9/0.999, eps 1e-8).
 * mean_loss receives the mean of the batch losses. */
int agimp_trainer_epoch(agimp_trainer* t, const int32_t* perm, int64_t batch, const float* class_weight, float scale,
                        float lr, float weight_decay, double* mean_loss);
/* Logits [rows][4] of host rows X [rows][16] (overwrites the dataset). */
int agimp_trainer_logits(agimp_trainer* t, const float* X, int64_t rows, float* Z);
/* Loss and gradient [5508] of host rows (overwrites the dataset); class_weight may be NULL. */
int agimp_trainer_grad(agimp_trainer* t, const float* X, const int32_t* y, int64_t rows, const float* class_weight,
                       float scale, float* grad, float* loss);
/* One Adam step with a given gradient (host). */
int agimp_trainer_apply(agimp_trainer* t, const float* grad, float lr, float weight_decay);

/* Host-side FP64 maths used by the kernels, exported for CPU checks: the rotation-variant
 * SVD (dim 2 or 3, row-major) and the constitutive update (clamp plasticity if enabled in
 * cfg, then Kirchhoff stress, utils.py:11-16). */
int agimp_host_svd(int32_t dim, const double* F, double* U, double* s, double* V);
int agimp_host_constitutive(const agimp_config* cfg, double* F, double* Jp, double* tau, int32_t clamp);

#ifdef __cplusplus
}
#endif

#endif /* AGIMP_H */
