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
  int32_t particles_per_chunk; /* P2G work chunk (particles of one leaf block), AGIMP_MIN/MAX_CHUNK */
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

/* Dense cell_mask[l], ad_grid_m[l], ad_grid_v[l] (adaptive_gimp.py:39-42); any may be NULL. */
int agimp_get_level(agimp_handle* h, int32_t l, int32_t* cell_mask, float* m, float* v);

/* Dense node_mask (adaptive_gimp.py:23-26), finest extent. */
int agimp_get_node_mask(agimp_handle* h, int32_t* node_mask);

/* Counters accumulated since creation (or the last reset): [0] out-of-box accesses,
 * [1] frozen particle-substeps (fl_p = -1), [2] transferred particle-substeps,
 * [3] P2G chunks, [4] non-finite particles. */
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

/* Host-side FP64 maths used by the kernels, exported for CPU checks: the rotation-variant
 * SVD (dim 2 or 3, row-major) and the constitutive update (clamp plasticity if enabled in
 * cfg, then Kirchhoff stress, utils.py:11-16). */
int agimp_host_svd(int32_t dim, const double* F, double* U, double* s, double* V);
int agimp_host_constitutive(const agimp_config* cfg, double* F, double* Jp, double* tau, int32_t clamp);

#ifdef __cplusplus
}
#endif

#endif /* AGIMP_H */
