/*
 * dem.h -- C ABI of the MI355X (gfx950) DEM packing step, libdem.so.
 *
 * Three layers, one set of HIP kernels:
 *
 *  1. Reference-named entry points (section "drop-in").  Same names, argument order and
 *     meaning as the reference's extern "C" operator API in particleSystem.cu, plus the
 *     SDK RadixSort it links, so a dempacking.cu-shaped driver links against this library
 *     unchanged.  Buffers are caller-owned device pointers in particle-index order, AoS
 *     double3 as in the reference.  Launches go to the null stream; host scalars are
 *     returned after a synchronising copy exactly where the reference copies them.
 *     Errors: the reference's gpuErrchk prints and exit()s (particleSystem.cu:15-23); a
 *     library must not end its host process, so these functions print the same message,
 *     record the error and return; dem_last_error() reads (and clears) it.
 *
 *  2. Context API (dem_*).  Owns every buffer (no per-step allocation), keeps particles
 *     in cell-sorted slots, builds CSR pair lists with a transposed index, fuses the
 *     per-particle gather, wall force, gravity and integration into one kernel, and runs
 *     one or many independent instances (batched form) per launch.  Functions return a
 *     dem_status.
 *
 *  3. Adjoint form (dem_adj_*).  Reverse mode of the context step with checkpoints and
 *     segment recompute, for gradients of smooth packing observables with respect to the
 *     material parameters and the initial state.
 *
 * All arithmetic is IEEE FP64.  Units are the reference's reduced units: length = particle
 * diameter d, mass = rho*pi*d^3/6, force = m*g, time = sqrt(d/g) (dempacking.cu:146-153,
 * neiglist.cu:272).
 */
#ifndef DEM_H
#define DEM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef unsigned int uint;

/* ----------------------------------------------------------------------------------- */
/* Shared types                                                                         */
/* ----------------------------------------------------------------------------------- */

/* Reference Matproperties / Wallproperties (neiglist.cu:32-47), same member order. */
typedef struct Matproperties {
    double estarp, ufricp, vpois, dampnp, damptp, upp;
} Matproperties;

typedef struct Wallproperties {
    double estarw, ufricw, vpoisw, dampnw, damptw, upw;
} Wallproperties;

/* SDK KeyValuePair used by RadixSort (dempacking.cu:553-555): (cell hash, index). */
typedef struct KeyValuePair {
    uint key;
    uint value;
} KeyValuePair;

typedef struct dem_material {
    Matproperties p;     /* particle-particle */
    Wallproperties w;    /* particle-wall */
} dem_material;

enum { DEM_WALLS_BOX = 0, DEM_WALLS_CYLINDER = 1 };

/* Container walls.  Box: side walls x = +-size[0]/2, y = +-size[1]/2 and floor z = 0 (the
 * reference's worldSize box, dempacking.cu:333-339).  Cylinder: vertical wall of radius
 * `radius` about the z axis, floor z = 0.  No lid in either case. */
typedef struct dem_walls {
    int kind;
    int pad;
    double size[3];
    double radius;
} dem_walls;

/* One independent packing.  Grid as in dempacking.cu:329-344 (cell, origin, gridSize). */
typedef struct dem_instance {
    uint n;              /* particles */
    uint gsize[3];       /* cells per axis */
    double cell[3];      /* cell edge per axis; must be >= 2*max radius + skin */
    double origin[3];    /* grid origin */
    dem_walls walls;
    dem_material mat;
    double dt;           /* reduced time step */
    double skin;         /* list skin = cub - cua (neiglist.cu:27-28): 0.55 */
} dem_instance;

typedef enum dem_status {
    DEM_OK = 0,
    DEM_ERR_INVALID = 1,      /* bad argument or size */
    DEM_ERR_HIP = 2,          /* a HIP runtime call failed */
    DEM_ERR_CAPACITY = 3,     /* a fixed capacity (pairs, neighbour row, no-alloc cap) exceeded */
    DEM_ERR_NO_DEVICE = 4,    /* no usable gfx950 device */
    DEM_ERR_IO = 5            /* deck / restart file error */
} dem_status;

/* Last recorded error of the drop-in layer (0 if none); reading clears it. */
int dem_last_error(void);
/* Human-readable text for a dem_status. */
const char *dem_status_string(int status);

/* ----------------------------------------------------------------------------------- */
/* 1. Drop-in: reference-named operator API (particleSystem.cu:13-424).                 */
/* ----------------------------------------------------------------------------------- */

/* Replaces cudaMemcpyToSymbol(Par) / (ParW) (dempacking.cu:355-356). */
void setMaterialProperties(const Matproperties *par, const Wallproperties *parw);
/* Selects the wall set used by calculateforceiw (default: the worldSize box). */
void setWallGeometry(const dem_walls *walls);

/* particleSystem.cu:27 -- particleHash[i] = (cell of Pos[i], i). */
void calcHash(double *Pos, uint *particleHash, uint gridSize[3], double cellSize[3],
              double worldOrigin[3], int numBodies);

/* SDK radix sort (call dempacking.cu:553-555): stable sort of pData0 by key over the low
 * `bits` key bits, pData1 is scratch of the same size, result in pData0. */
void RadixSort(KeyValuePair *pData0, KeyValuePair *pData1, uint elements, uint bits);

/* RadixSort without allocation -- no counterpart in the reference.  For callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same stable sort, with RadixSort's scratch (digit counts, their
 * scan) in a fixed device array of 8 MiB that each device holds until the process ends.
 * At most 8,388,608 (2^23) elements: larger counts return DEM_ERR_CAPACITY untouched.
 * Returns a dem_status.  Calls on one device share that array: calls that overlap on one
 * device (several host threads or streams) race. */
int RadixSortNoAlloc(KeyValuePair *pData0, KeyValuePair *pData1, uint elements, uint bits);

/* particleSystem.cu:49 -- cellStart[c] = first sorted slot of cell c, 0xffffffff if empty. */
void findCellStart(uint *particleHash, uint *cellStart, int numBodies, uint numGridCells);

/* particleSystem.cu:67 -- Anei[i*maxParticlesPerCell + k]: candidates of i, |xj-xi| < cub
 * (1.55), 27-cell stencil order; njgi / njli = number with j > i / j < i. */
void neighborarray(double *pos, uint *particleHash, uint *cellStart, uint gridSize[3],
                   double cellSize[3], double worldOrigin[3], int maxParticlesPerCell,
                   int numBodies, uint *Anei, uint *njgi, uint *njli);

/* particleSystem.cu:126 -- oldsjgi <- sjgi; sjgi = inclusive scan of njgi;
 * *totalcontacts (host) = sjgi[numBodies-1]. */
void prefixsum(uint numBodies, uint *njgi, uint *sjgi, uint *oldsjgi, uint *totalcontacts);

/* particleSystem.cu:143 -- pair list Lp/Ln, reference table Aref, tangential history
 * remapped from (oldLn, olddispt, oldfricp) by pair identity. */
void pairlist(uint *sjgi, uint *oldsjgi, uint *oldLn, uint *Anei, uint *njgi, uint *njli,
              double *olddispt, double *oldfricp, int maxParticlesPerCell,
              int maxCnPerParticle, uint numBodies, uint *Aref, uint *Lp, uint *Ln,
              double *dispt, double *fricp);

/* particleSystem.cu:194 -- per-pair contact force and torques (Hertz-Mindlin, damping,
 * Coulomb limit, rolling friction). */
void calculateforcepair(double *pos, double *pdis, double *angv, double *rad, double *rmass,
                        uint *Lp, uint *Ln, uint *oldLn, double *dispt, double *fricp,
                        double *olddispt, double *oldfricp, double dt, double *forcepair,
                        double *torqpairi, double *torqpairj, uint *contactpair,
                        uint totalcontacts);

/* particleSystem.cu:243 -- per-particle sum of pair forces/torques through Aref. */
void sumallforcei(int maxParticlesPerCell, int numBodies, uint *njgi, uint *njli,
                  double *forcepair, double *torqpairi, double *torqpairj, uint *contactpair,
                  uint *Aref, double *forcei, double *torquei, uint *contacti);

/* particleSystem.cu:279 -- adds gravity and particle-wall forces (one history slot). */
void calculateforceiw(double worldSize[3], double *pos, double *pdis, double *angv,
                      double *rad, double *rmass, double dt, uint totalcontacts,
                      int numBodies, double *disptw, double *fricpw, double *forcei,
                      double *torquei);

/* particleSystem.cu:315 -- integrator; *iall (host) = rebuild flag, *maxhz (host) = max z. */
void updateposition(double worldSize[3], int numBodies, double *pos, double *pdis,
                    double *angv, double *rmass, double *inert, double *forcei,
                    double *torquei, double *qi, double dt, double cdamp, uint *iall,
                    double *maxhz);

/* particleSystem.cu:386 -- *pcntot (host) = sum of contacti. */
void outputporositycn(uint numBodies, uint *contacti, double maxhz, double worldSize[3],
                      uint *pcntot);

/* ----------------------------------------------------------------------------------- */
/* 2. Context API                                                                       */
/* ----------------------------------------------------------------------------------- */

typedef struct dem_ctx dem_ctx;

typedef struct dem_stats {
    unsigned long long steps;      /* steps run */
    unsigned long long rebuilds;   /* list rebuilds run */
    uint n;                        /* particles (all instances) */
    uint npairs;                   /* candidate pairs in the current list */
    uint pair_capacity;
    uint ninstances;
    uint last_rebuild_flag;        /* displacement trigger raised by the last step */
    uint overflow;
    double maxhz;                  /* instance 0: max(0, max particle z) after the last step */
    uint pcntot;                   /* instance 0: sum of per-particle contact counts */
    uint step_path;                /* step path in use: 1 fused, 0 split */
} dem_stats;

/* Options for dem_set_option. */
enum {
    DEM_OPT_RECORD_FORCES = 1,     /* 1: keep per-particle force/torque totals for dem_get_forces */
    DEM_OPT_REBUILD_INTERVAL = 2,  /* 0: displacement trigger (reference); K>0: every K steps */
    DEM_OPT_SKIN = 4,              /* list skin (cell edge follows); requests a rebuild */
    DEM_OPT_PAIR_BLOCK = 5,        /* threads per block of the pair kernel: 64, 128, 256 */
    DEM_OPT_PAIR_GRID = 6,         /* pair kernel blocks per CU (grid-stride) */
    DEM_OPT_PARTICLE_BLOCK = 7,    /* threads per block of the per-particle kernels: 64, 128
                                      (default) or 256 */
    DEM_OPT_ADJ_ROLL_REG = 8,      /* adjoint: floor of |w_rel| in the rolling moment's
                                      linearisation, in units of the one-step chatter scale
                                      mu_r fcn dt^2 (1/I_a + 1/I_b); 0 = exact, default 1 */
    DEM_OPT_FUSED_STEP = 9,        /* 1: fused per-particle step, each pair law evaluated
                                      from both sides, no per-pair output buffer; 0: pair
                                      kernel + gather; 2 (default): auto -- dem_step times one
                                      step of each after the first and every 8th rebuild and
                                      keeps the faster.  Same results bit for bit */
    DEM_OPT_TIME_STEPS = 10,       /* 1: dem_step records HIP events around its kernels
                                      (read with dem_get_step_timing) */
    DEM_OPT_PIPELINE = 11,         /* steps dem_step keeps in flight under the displacement
                                      trigger (1..16, default 16; 1 = wait for each step) */
    DEM_OPT_FUSED_FAST = 12,       /* fused step: a particle with at most this many candidates
                                      on each side (own pairs, pairs where it is second) tests
                                      both sides in one pass and runs one law loop (0..64,
                                      default 64); others go through 32-candidate chunks, own
                                      side first.  Same results bit for bit */
    DEM_OPT_XCD_CHUNK = 13,        /* step kernels: the 8 XCDs take chunks of this many
                                      blocks in turn (default 8; 0: one contiguous eighth of
                                      the particles or pairs each).  Same results bit for bit */
    DEM_OPT_PIPELINE_BATCH = 14,   /* queued steps issued together behind one HIP event
                                      (1..16, at most DEM_OPT_PIPELINE; default 8): an event
                                      between two kernels adds ~4 us to the gap between them.
                                      Same results bit for bit */
    DEM_OPT_ROLL_FLOOR = 15,       /* differentiable form: every instance's rolling moment
                                      floors |w_rel| at k mu_r fcn dt^2 (1/I_a + 1/I_b) (one
                                      inertia for a wall), i.e. it is viscous, -w_rel / (k dt^2
                                      (1/I_a + 1/I_b)), where the reference's moment would flip
                                      sign every step; the adjoint then reverses that forward
                                      exactly.  k >= 0; 0 (default) = the reference's moment */
    DEM_OPT_FLOAT_LENGTH = 16      /* 1: the reference's float length() -- sqrtf of the FP64 dot
                                      product rounded to float (cutil_math.h:1086-1089) -- for
                                      vprojec's magnitudes (neiglist.cu:140-142) and the rebuild
                                      trigger rmov = length(q) > 0.5 skin (neiglist.cu:359,370),
                                      in every instance's pair and wall laws; 0 (default): FP64.
                                      Not supported by dem_adj_forward (DEM_ERR_INVALID) */
};

/* Result of dem_autotune. */
typedef struct dem_tune_result {
    double skin;                   /* chosen list skin */
    double cell;                   /* resulting cell edge of instance 0 */
    int interval;                  /* safe fixed rebuild interval for dem_step_graph */
    uint pair_block, pair_grid_per_cu, particle_block;
    double ms_per_step;            /* force-step time at the chosen shapes */
    double score_ms;               /* force step + rebuild / interval at the chosen skin */
    int trials;
    int step_path;                 /* 1 fused, 0 split */
} dem_tune_result;

/* Create a context for `ninstances` independent packings (1 = plain form).  pair_capacity
 * 0 picks 12 pairs per particle; the plain path grows it on demand. */
int dem_create(dem_ctx **ctx, int ninstances, const dem_instance *inst, uint pair_capacity);
int dem_destroy(dem_ctx *ctx);
int dem_set_option(dem_ctx *ctx, int option, double value);

/* Particle state in index order (instances concatenated), AoS double3 for pos/pdis/angv.
 * pdis / angv are per-step displacements (reference convention).  `on_device` = 1 means
 * the pointers are device pointers.  Setting the state clears all contact history and
 * requests a list rebuild before the next step. */
int dem_set_state(dem_ctx *ctx, const double *pos, const double *pdis, const double *angv,
                  const double *rad, const double *rmass, const double *inert, int on_device);
int dem_get_state(dem_ctx *ctx, double *pos, double *pdis, double *angv, double *qi,
                  int on_device);
/* Per-particle wall history (disptw AoS double3, fricpw) in index order. */
int dem_get_wall_history(dem_ctx *ctx, double *disptw, double *fricpw);

/* Full-state checkpoint, the binary, lossless replacement of the reference's preflow.dat
 * (dempacking.cu:805-823, which keeps pos/pdis/angv at 6 decimals and no contact history):
 * records of every particle (active or waiting), wall histories, the live pair list with
 * its tangential histories, the instance table (grids, materials), step counters, the
 * rebuild state and the rebuild interval.  Steps after dem_load_checkpoint repeat the
 * original run's steps bit for bit.  Save writes `path`.part and renames it over `path`.
 * Load needs a context created for the same instances (same particle counts); it replaces
 * the state, the grids and materials, and drops any adjoint recording.  Returns DEM_ERR_IO
 * for a missing, truncated or corrupted file (checksum), DEM_ERR_INVALID for a context of
 * other sizes. */
int dem_save_checkpoint(dem_ctx *ctx, const char *path);
int dem_load_checkpoint(dem_ctx *ctx, const char *path);
/* Layout the checkpoint reader checks (host only, no device): out[0..3] = bytes of a particle
 * record, a history, an instance entry, the control block; out[4..5] = offsets of the
 * record's instance and index words; out[6..9] = offsets of the instance entry's particle
 * offset, count, cell offset and grid size; out[10..13] = offsets of the control block's
 * pair count, overflow, wall-candidate count and zero word (out holds 14 uints). */
int dem_checkpoint_layout(uint *out);

/* Single-instance contexts: make particles [0, n) active; the rest of the state set by
 * dem_set_state waits, as the reference's not-yet-inserted layers do (dempacking.cu:378-462).
 * Any n before the first step after dem_set_state, only growth afterwards.  The next step
 * rebuilds the lists. */
int dem_set_active(dem_ctx *ctx, uint n);
uint dem_active_count(dem_ctx *ctx);
/* Run `nsteps` steps in the reference order: [rebuild if requested] -> pair forces ->
 * gather + walls + gravity + integration -> rebuild trigger. */
int dem_step(dem_ctx *ctx, int nsteps);
/* Run `nsteps` steps as replays of a captured hipGraph (fixed rebuild interval K, set with
 * DEM_OPT_REBUILD_INTERVAL; nsteps must be a multiple of K). */
int dem_step_graph(dem_ctx *ctx, int nsteps);
/* Project autotuner: picks the list skin (hence the cell size), the launch shapes and a safe
 * fixed rebuild interval by timing trials of `steps` steps each.  Trials run the reference
 * trigger, so they are ordinary exact steps of the simulation (they advance it).  Applies the
 * skin and shapes to the context; the caller sets DEM_OPT_REBUILD_INTERVAL to
 * result->interval before dem_step_graph. */
int dem_autotune(dem_ctx *ctx, int steps, dem_tune_result *result);
/* Request a list rebuild before the next step. */
int dem_request_rebuild(dem_ctx *ctx);
/* Rebuild now (no force step). */
int dem_rebuild(dem_ctx *ctx);
int dem_get_stats(dem_ctx *ctx, dem_stats *out);
/* Per-instance scalars after the last step: maxhz[ninst], pcntot[ninst] (either may be NULL). */
int dem_get_instance_stats(dem_ctx *ctx, double *maxhz, uint *pcntot);
/* Synchronise the context's stream. */
int dem_sync(dem_ctx *ctx);
/* Device time (ms) of the last dem_step / dem_step_graph call, from HIP events. */
double dem_last_ms(dem_ctx *ctx);
/* dem_step with HIP events around each kernel (synchronises every step): out_ms[0] mean
 * pair-force kernel time, out_ms[1] mean per-particle kernel time, out_ms[2] mean rebuild
 * time, out_ms[3] total; *nrebuild = rebuilds run. */
int dem_time_kernels(dem_ctx *ctx, int nsteps, double *out_ms, int *nrebuild);
/* With DEM_OPT_TIME_STEPS: kernel times accumulated by dem_step since the last call (ms),
 * then reset.  out[10] = split path: pair kernel + walls total, per-particle kernel total,
 * steps, sum of pair counts; fused path: walls total, step kernel total, steps, sum of pair
 * counts; rebuild total, rebuilds. */
int dem_get_step_timing(dem_ctx *ctx, double *out);
/* One list rebuild now, timed stage by stage with HIP events: out[0..8] = ms of hash (cell
 * counts), sort (count scan and scatter), cell ranges (0: given by the scan), reorder,
 * neighbour count, scans, pair fill, history remap, transpose; out[9] = radix passes (0: the
 * rebuild sorts by counting), out[10] = new pairs, out[11] = old pairs, out[12] = cells,
 * out[13] = particles, out[14] = ms of a second rebuild timed at its two phases' ends only
 * (no per-stage events; the host's read of the pair total between the phases excluded)
 * (out holds 15 doubles).  The trajectory is unchanged. */
int dem_time_rebuild(dem_ctx *ctx, double *out);

/* Introspection (host buffers), for parity checks against the reference layout. */
/* calcHash output of the last rebuild: 2n uint (cell, index), index order. */
int dem_get_hash(dem_ctx *ctx, uint *kv);
/* Sorted (cell, index) of the last rebuild, 2n uint. */
int dem_get_sorted(dem_ctx *ctx, uint *kv);
/* cellStart over all instances' cells. */
int dem_get_cell_start(dem_ctx *ctx, uint *cs, uint ncells);
/* Pairs in list order: ids[2p] = (index of first, index of second); hist[4p] =
 * (dispt xyz, fricp) oriented second-relative-to-first; out[10p] = (F on first xyz,
 * torque first xyz, torque second xyz, contact flag) of the last force step.  Any pointer
 * may be NULL.  Returns DEM_ERR_CAPACITY if cap < npairs. */
int dem_get_pairs(dem_ctx *ctx, uint *ids, double *hist, double *out, uint cap);
/* Last step's per-particle contact counts (contacti) in index order. */
int dem_get_contacts(dem_ctx *ctx, uint *contacti);
/* Last step's per-particle totals (pairs + walls + gravity) in index order; needs
 * DEM_OPT_RECORD_FORCES. */
int dem_get_forces(dem_ctx *ctx, double *force, double *torque, uint *contacti);

/* Current device: CU count and gfx architecture name. */
int dem_device_info(int *cu_count, char *name, int name_len);
/* Device memory for callers without a HIP or torch binding (e.g. a ctypes or FFI driver of
 * the drop-in entry points).  dem_memcpy kind: 0 H2H, 1 H2D, 2 D2H, 3 D2D (synchronous). */
void *dem_dev_alloc(size_t bytes);
int dem_dev_free(void *ptr);
int dem_memcpy(void *dst, const void *src, size_t bytes, int kind);
int dem_memzero(void *dst, size_t bytes);
int dem_device_sync(void);

/* ----------------------------------------------------------------------------------- */
/* Contact laws on the host (the kernels' own inline code) and their adjoints            */
/* ----------------------------------------------------------------------------------- */

/* mat = (E*, mu, nu, gamma_n, gamma_t, mu_r).  Pair: in = xa, xb, va, vb, wa, wb (18), ra, rb,
 * ma, mb, hin (3), dt (26 values); out = F on a, Ta, Tb, h_out, |Ft| (13).  Returns 1 on
 * contact.  vjp: outbar = adjoints of (F, Ta, Tb, h_out) (12); inbar = adjoints of xa, xb, va,
 * vb, wa, wb, hin (21); matbar = adjoints of (E*, mu, gamma_n, gamma_t, mu_r). */
int dem_law_pair_eval(const double *in, const double *mat, double *out);
/* The step kernels' FP32 contact filter on the host: in = count x (xa (3), ra, xb (3), rb);
 * keep[i] = 0 only if the spheres are certainly apart (the FP64 law then finds no contact). */
int dem_contact_filter(const double *in, int count, int *keep);
int dem_law_pair_vjp(const double *in, const double *mat, const double *outbar, double *inbar,
                     double *matbar);
/* Walls: in = x, v, w (9), r, m, hin (3), dt (15); out = F, T, h_out, |Ft| (10), returns the
 * walls hit; vjp outbar = (F, T, h_out) (9), inbar = x, v, w, hin (12). */
int dem_law_walls_eval(const dem_walls *walls, const double *in, const double *mat, double *out);
int dem_law_walls_vjp(const dem_walls *walls, const double *in, const double *mat,
                      const double *outbar, double *inbar, double *matbar);
/* The same four with the rolling floor of DEM_OPT_ROLL_FLOOR: roll_c = k dt^2 (1/I_a + 1/I_b)
 * (pair) or k dt^2 / I (walls); the vjp is the exact reverse of that forward.  roll_c = 0
 * gives the functions above. */
int dem_law_pair_eval_floor(const double *in, const double *mat, double roll_c, double *out);
int dem_law_pair_vjp_floor(const double *in, const double *mat, double roll_c,
                           const double *outbar, double *inbar, double *matbar);
int dem_law_walls_eval_floor(const dem_walls *walls, const double *in, const double *mat,
                             double roll_c, double *out);
int dem_law_walls_vjp_floor(const dem_walls *walls, const double *in, const double *mat,
                            double roll_c, const double *outbar, double *inbar, double *matbar);

/* ----------------------------------------------------------------------------------- */
/* 3. Adjoint (differentiable) form -- no counterpart in the reference                  */
/* ----------------------------------------------------------------------------------- */

/* Smooth packing observables, per instance.  Region: side walls moved in by `margin`,
 * zlo < z < zhi, with logistic edges of width `eps`:
 *   phi = sum_i w(x_i) V_i / V_region,   CN = sum_i w_i c_i / sum_i w_i,
 *   c_i = sum_j S(overlap_ij / eps_c), S the C1 smoothstep (0 at no overlap, 1 past eps_c). */
typedef struct dem_obs {
    double margin, zlo, zhi, eps, eps_c;
} dem_obs;

/* Replace the material of one instance (E*, mu, damping, rolling friction of the pair and
 * wall laws) without touching the state. */
int dem_set_material(dem_ctx *ctx, int instance, const dem_material *mat);

/* phi[ninst], cn[ninst] of the current state (builds the list first if there is none). */
int dem_observe(dem_ctx *ctx, const dem_obs *obs, double *phi, double *cn);

/* Run nsteps steps (dem_step's schedule) keeping a full-state checkpoint every
 * `checkpoint_every` steps, then evaluate the observables at the end.  Requires every
 * particle active. */
int dem_adj_forward(dem_ctx *ctx, int nsteps, int checkpoint_every, const dem_obs *obs,
                    double *phi, double *cn);

/* Reverse sweep of the last dem_adj_forward: for each segment, last first, restore its
 * checkpoint, recompute it with tapes, and run the adjoint kernels back over it.  gphi,
 * gcn[ninst] are the adjoints of the observables; theta_bar[10 * ninst] receives, per
 * instance, d(sum gphi phi + gcn CN) / d(E*, mu, gamma_n, gamma_t, mu_r) of the pair law and
 * then of the wall law.  The context is left at the forward's end state. */
int dem_adj_backward(dem_ctx *ctx, const double *gphi, const double *gcn, double *theta_bar);

/* Adjoints of the initial positions, displacements and rotations (per particle index,
 * 3 per particle each; any may be null) from the last dem_adj_backward. */
int dem_adj_state_grad(dem_ctx *ctx, double *xbar, double *vbar, double *wbar);

/* out[6] = forward ms, backward ms, segments, largest pair list, checkpoint bytes, tape
 * bytes of the last forward/backward. */
int dem_adj_info(dem_ctx *ctx, double *out);

/* Free the checkpoints and tapes. */
int dem_adj_release(dem_ctx *ctx);

/* ----------------------------------------------------------------------------------- */
/* Host utilities: deck, units, restart file, workload generators                       */
/* ----------------------------------------------------------------------------------- */

/* hopp3d.inp, read positionally as dempacking.cu:112-125. */
typedef struct dem_deck {
    int iforcemode;
    double emod, vpois, denp;
    double diam;
    int ntot;
    double dt;
    uint tstop;
    double dmt1, dmt2, thick;
    double zht1, zht2;
    double vel0;
    int nrings;
    double ufricp, ufricw;
    double dampnp, damptp, dampnw, damptw, cdamp;
    double stiffness, dampcoeffn, dampcoefft;
    double upp, upw;
} dem_deck;

int dem_read_deck(const char *path, dem_deck *deck);

/* Deck side-car: the knobs the 12-line deck has no line for, in a file of its own so that
 * hopp3d.inp keeps the reference's layout.  One "key values" per line, '#' starts a
 * comment; lengths in units of the deck's particle diameter:
 *   walls box X Y Z | walls cylinder R      container (default: the deck's box)
 *   height H                                container height (default: from the deck)
 *   particles N                             particle count (default: the deck's ntot)
 *   diameters mono D | lognormal SIGMA DMIN DMAX | bimodal DBIG DSMALL FRACTION_BIG
 *   insertion layers | lattice SPACING JITTER Z0 | random FILL
 *                                           layers: the reference's layer pour (default);
 *                                           lattice / random: every particle placed at the
 *                                           start, at rest (random: seeded non-overlapping
 *                                           insertion up to the height giving volume fraction
 *                                           FILL, at most the container height)
 *   seed S                                  diameters use S, positions S + 1 (lattice: S)
 *   restitution E                           damping of both laws from e (the deck's
 *                                           damping coefficients otherwise)
 *   rebuild trigger | rebuild interval K    list rebuild policy (default: the trigger)
 * Unknown keys and bad values are DEM_ERR_IO. */
typedef struct dem_sidecar {
    int walls;                   /* -1: the deck's box; DEM_WALLS_BOX; DEM_WALLS_CYLINDER */
    int particles;               /* 0: the deck's ntot */
    double box[3];
    double radius;
    double height;               /* 0: from the deck */
    int diameters;               /* dem_gen_diameters kind: 0 mono, 1 lognormal, 2 bimodal */
    int insertion;               /* 0 layers, 1 lattice, 2 random */
    double dpar[3];              /* diameter parameters (mono: dpar[0] = D, default 1) */
    double ipar[3];              /* lattice: spacing, jitter, z0; random: fill */
    unsigned long long seed;     /* default 1 */
    double restitution;          /* < 0: keep the deck's damping coefficients */
    int rebuild_interval;        /* 0: displacement trigger */
    int pad;
} dem_sidecar;

int dem_read_sidecar(const char *path, dem_sidecar *side);
/* Reduced material from the deck (dempacking.cu:141-149, 224-225). */
int dem_deck_material(const dem_deck *deck, dem_material *mat);
/* Writes restart.dat exactly as dempacking.cu:908-913 (40 bytes). */
int dem_write_restart(const char *path);
/* Reads restart.dat as the reference's disabled reader does (fscanf "%d %d %d",
 * dempacking.cu:361-366): nrestart 0 means "resume from the checkpoint".  DEM_ERR_IO if the
 * file is missing, DEM_ERR_INVALID if it does not start with three integers (the end-of-run
 * file written by dem_write_restart), in which case a run starts fresh. */
int dem_read_restart(const char *path, int *nrestart, int *npregas, int *nparticledat);
/* Writes restart.dat in that format ("%8d %8d %8d", the write left commented out at
 * dempacking.cu:829), so that the next run resumes when nrestart is 0. */
int dem_write_restart_resume(const char *path, int nrestart, int npregas, int nparticledat);
/* Normal damping coefficient giving restitution e for the Hertz + dashpot law (bisection
 * on the head-on collision ODE).  Returns the coefficient, or -1 on bad input. */
double dem_damping_for_restitution(double e);
/* Restitution produced by damping coefficient g (same ODE). */
double dem_restitution_for_damping(double g);

/* Diameter generators (reduced units, mean-d = 1): kind 0 monodisperse (p0 = d), kind 1
 * lognormal (p0 = sigma, clipped to [p1, p2]), kind 2 bimodal (p0 = d_big, p1 = d_small,
 * p2 = volume fraction of big).  Deterministic for a seed. */
int dem_gen_diameters(int kind, uint n, double p0, double p1, double p2,
                      unsigned long long seed, double *diam);
/* Seeded random sequential insertion without overlap inside `walls` between z0 and z1
 * (grid-accelerated rejection).  Returns DEM_OK or DEM_ERR_CAPACITY if it cannot place
 * every sphere within the attempt budget. */
int dem_gen_insert(const dem_walls *walls, double z0, double z1, uint n, const double *diam,
                   unsigned long long seed, double *pos);
/* Jittered lattice (config 0): spacing a, jitter amplitude j (in d), box walls. */
int dem_gen_lattice(const dem_walls *walls, double z0, double a, double jitter, uint n,
                    unsigned long long seed, double *pos);
/* Fill rad/rmass/inert from diameters (rmass = 8 r^3, inert = 0.4 m r^2; dempacking.cu:310-312). */
int dem_particle_props(uint n, const double *diam, double *rad, double *rmass, double *inert);
/* Grid covering `walls` up to height h with cell edge c (dempacking.cu:329-344 pattern). */
int dem_make_grid(const dem_walls *walls, double h, double c, uint gsize[3], double origin[3]);

#ifdef __cplusplus
}
#endif

#endif /* DEM_H */
