/*
 * mcs.h -- C ABI of libmcs, the MI355X (gfx950) implementation of the
 * per-step particle-mesh loop of aluque/mcstreamer.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, and returns
 * an int32_t status (MCS_OK = 0).  The library owns every device buffer; host
 * pointers are borrowed for the duration of a call only.  One context owns one
 * HIP stream; calls on a context are not re-entrant.
 *
 * Each declaration names the reference function (file:line, paths relative to
 * the reference tree) whose behaviour it replaces.  A Julia binding calls these
 * with `ccall`; the Python mirror in mcstreamer_amd/ calls them with ctypes.
 */
#ifndef MCS_H
#define MCS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------
 * MCS_E_CAPACITY     : JuMC/src/population.jl:90 "Maximum number of particles reached"
 * MCS_E_FIELD        : MCStreamer/src/gridfields.jl:178-187 "Too high electric field" (|E| > 1e8 V/m)
 * MCS_E_TABLE_RANGE  : JuMC/src/util.jl:27 (energy beyond the last table node; BoundsError)
 * MCS_E_NONCONVERGED : Multigrid/src/Multigrid.jl:458-464 (logged there, returned here)
 * MCS_E_NAN          : JuMC/src/util.jl:27 with a NaN energy (floor(Int, NaN) throws an
 *                      InexactError): a particle whose velocity holds a NaN collides
 */
#define MCS_OK 0
#define MCS_E_ARG 1
#define MCS_E_HIP 2
#define MCS_E_CAPACITY 3
#define MCS_E_FIELD 4
#define MCS_E_TABLE_RANGE 5
#define MCS_E_NONCONVERGED 6
#define MCS_E_NAN 7

/* Collision process kinds (JuMC/src/electron.jl:70-84, collisions.jl:80). */
#define MCS_PROC_NULL 0
#define MCS_PROC_ELASTIC 1
#define MCS_PROC_EXCITATION 2
#define MCS_PROC_IONIZATION 3
#define MCS_PROC_ATTACHMENT 4
#define MCS_PROC_PHOTOEMISSION 5 /* electron.jl:88-105; parameters by mcs_table_set_photoemission */
#define MCS_MAX_PROC 32

/* Energy grid of a collision table. */
#define MCS_EGRID_UNIFORM 0 /* the reference's LinRange (util.jl:23-30) */
#define MCS_EGRID_LOG 1     /* log-spaced nodes, rate linear in energy between nodes */

/* Field descriptor kinds: Swarm.FixedField (Swarm/src/Swarm.jl:28-32) and
 * MCStreamer.FieldInterp (MCStreamer/src/gridfields.jl:86-190). */
#define MCS_FIELD_UNIFORM 0
#define MCS_FIELD_MESH 1

/* Mesh arrays.  Cell-centred arrays carry one ghost layer: (nx+2)(ny+2)(nz+2),
 * x fastest.  Face arrays: EX (nx+1)(ny+2)(nz+2), EY (nx+2)(ny+1)(nz+2),
 * EZ (nx+2)(ny+2)(nz+1). */
#define MCS_ARR_QFIXED 0
#define MCS_ARR_QPART 1
#define MCS_ARR_Q 2
#define MCS_ARR_U 3
#define MCS_ARR_EX 4
#define MCS_ARR_EY 5
#define MCS_ARR_EZ 6

#define MCS_DEPOSIT_NGP 0 /* MCStreamer/src/poisson.jl:65-84 */
#define MCS_DEPOSIT_CIC 1 /* trilinear cloud-in-cell */

/* Multigrid boundary conditions (Multigrid/src/boundaries.jl:100-108). */
#define MCS_BC_DIRICHLET 0
#define MCS_BC_NEUMANN 1

#define MCS_PREC_F64 0
#define MCS_PREC_F32 1

typedef struct mcs_ctx mcs_ctx;
typedef struct mcs_table mcs_table;
typedef struct mcs_pop mcs_pop;
typedef struct mcs_mesh mcs_mesh;
typedef struct mcs_mg mcs_mg;
typedef struct mcs_cgrid mcs_cgrid;

typedef struct mcs_field {
    int32_t kind;      /* MCS_FIELD_UNIFORM or MCS_FIELD_MESH */
    double e[3];       /* uniform field (V/m) */
    mcs_mesh *mesh;    /* face fields for MCS_FIELD_MESH */
} mcs_field;

typedef struct mcs_step_stats {
    int64_t n_in;                   /* population size when the step started */
    int64_t n_out;                  /* after births were appended */
    int64_t births;
    int64_t deaths;
    int64_t flights;                /* free flights (collision or end of step) */
    int64_t events[MCS_MAX_PROC];   /* collisions selected per table row (null included) */
    int32_t flags;                  /* bit 0 table range, bit 1 field limit, bit 2 capacity,
                                       bit 3 NaN energy */
} mcs_step_stats;

typedef struct mcs_moments {
    int64_t nparticles;             /* population.jl:63 */
    int64_t nactive;                /* population.jl:130 */
    double weight;                  /* population.jl:114-124 */
    double wenergy;                 /* sum w*eps over actives (meanenergy numerator, :136-150) */
    double maxenergy;               /* over all slots, active or not (:156-158) */
    double sum_x[3];                /* unweighted sums over actives (Swarm.jl:126-127) */
    double sum_x2[3];
    double sum_v[3];                /* Swarm.jl:152 */
    double sum_energy;
    double collrate[MCS_MAX_PROC];  /* sum of per-process rates (Swarm.collrate, Swarm.jl:242-248) */
} mcs_moments;

typedef struct mcs_resample_stats {
    int64_t n_in;
    int64_t n_out;
    int64_t cells_merged;           /* cells with p > nmax */
    int64_t clones;                 /* particles created by roulette */
    int64_t splits;                 /* particles created by splitting */
    int64_t removed;
} mcs_resample_stats;

typedef struct mcs_mg_config {
    int32_t n[3];       /* interior cells per dimension at the finest level */
    int32_t bc[6];      /* x-, x+, y-, y+, z-, z+ ; applied in this order */
    double s;           /* right-hand-side factor (MGConfig.s, Multigrid.jl:48-49) */
    int32_t levels;     /* coarsenings (MGConfig.levels) */
    int32_t smooth1;
    int32_t smooth2;
    double tolerance;   /* converged when ||r||/sqrt(N) < s*tolerance (Multigrid.jl:391-398);
                           a negative value -t means t times the rms of b (relative mode) */
    int32_t maxiter;
    int32_t precision;  /* MCS_PREC_F64, or MCS_PREC_F32 smoother inside FP64 refinement */
    int32_t ndim;       /* 3 (also 0), or 2: an (n[0], n[1]) grid, n[2] == 1, bc[0..3], FP64,
                           arrays (n[0]+2)(n[1]+2) -- Multigrid.jl with N = 2 */
    int32_t cyl;        /* 0: CartesianConnector; 1 + d: CylindricalConnector{d+1} on dimension d
                           (Multigrid.jl:31-39; 2D only) */
} mcs_mg_config;

/* ---- context ------------------------------------------------------------- */
int32_t mcs_ctx_create(int32_t device, mcs_ctx **out);
int32_t mcs_ctx_destroy(mcs_ctx *ctx);
const char *mcs_last_error(const mcs_ctx *ctx);
int32_t mcs_ctx_synchronize(mcs_ctx *ctx);
/* The hipStream_t the context launches on (for callers that time or order work). */
int32_t mcs_ctx_stream(mcs_ctx *ctx, void **stream);
/* Per-kernel HIP-event timing of the hot kernels; names: "advance", "track" (the ion
 * tracker's ordered sums), "sort_count", "sort_scatter", "sort_gather", "deposit", "faces",
 * "resample", "mg_solve" (a whole mcs_mg_solve_mesh), "mg_rbgs", "mg_residual",
 * "mg_restrict", "mg_prolong". */
int32_t mcs_ctx_profile(mcs_ctx *ctx, int32_t enable);
int32_t mcs_ctx_profile_read(mcs_ctx *ctx, const char *kernel, double *total_ms, int64_t *launches);
int32_t mcs_ctx_profile_reset(mcs_ctx *ctx);

/* ---- collision table: JuMC.CollisionTable (JuMC/src/collisions.jl:86-99) ---
 * rate is energy-major: rate[k*nproc + j] is process j at node k (the reference's
 * column-major (nproc x nE) matrix, collisions.jl:94-95, lxcat.jl:99).  energy holds
 * the ne node energies in J.  param[j] is the threshold (J) of excitation,
 * ionization and attachment rows and the mass ratio of elastic rows. */
int32_t mcs_table_create(mcs_ctx *ctx, int32_t nproc, int32_t ne, int32_t egrid_kind,
                         const double *energy, const double *rate, const int32_t *kind,
                         const double *param, double maxrate, mcs_table **out);
int32_t mcs_table_destroy(mcs_table *t);
/* The fields of PhotoEmission (JuMC/src/electron.jl:88-105) for row `row`, created
 * with kind MCS_PROC_PHOTOEMISSION (its param is unused): log_numin and log_numax
 * as the struct stores them (the logs its constructor takes), photon_multiplier,
 * photon_weight.  The row's collision (electron.jl:141-144, collisions.jl:125-141)
 * leaves the electron as it is and emits Poisson((w / photon_multiplier) /
 * photon_weight) photons at its position, each of weight photon_weight, velocity
 * c randsphere() and absorption rate nu = exp(log_numin + (log_numax - log_numin) U).
 * Every photoemission row must be set before a step uses the table; such a table
 * needs mcs_advance_mixed. */
int32_t mcs_table_set_photoemission(mcs_table *t, int32_t row, double log_numin, double log_numax,
                                    double photon_multiplier, double photon_weight);

/* ---- particle population: JuMC.Population (JuMC/src/population.jl:4-36) -----
 * Host layout of x and v is n x 3 row-major (Julia Vector{SVector{3,Float64}}). */
int32_t mcs_pop_create(mcs_ctx *ctx, int64_t capacity, mcs_pop **out);
/* Split-precision population (precision mode) on the grid of mesh m: positions
 * stored as packed cell coordinates (i, j, k < 1024, 1024, 2048) plus an FP32
 * offset in the cell, FP32 velocity, weight and collision clock.  Upload and
 * download take and give the same FP64 host layout as an FP64 population;
 * mcs_advance runs the FP32 step, mcs_sort_by_cell / mcs_deposit /
 * mcs_resample_cells / mcs_pop_moments accept it; repack and swarm_resample do
 * not.  No counterpart in the reference (its populations are FP64). */
int32_t mcs_pop_create_split(mcs_ctx *ctx, int64_t capacity, const mcs_mesh *m, mcs_pop **out);
int32_t mcs_pop_destroy(mcs_pop *p);
int32_t mcs_pop_upload(mcs_pop *p, int64_t n, const double *x, const double *v,
                       const double *w, const double *s, const uint8_t *active,
                       const uint64_t *id, uint64_t next_id);
int32_t mcs_pop_download(mcs_pop *p, int64_t nmax, int64_t *n, double *x, double *v,
                         double *w, double *s, uint8_t *active, uint64_t *id,
                         uint64_t *next_id);
/* nparticles (population.jl:63): slots in use, active or not. */
int32_t mcs_pop_size(mcs_pop *p, int64_t *n);
/* weight/actives/meanenergy/maxenergy/spread (population.jl:74-183) and the swarm
 * tallies of Swarm.swarm (Swarm.jl:126-156) in one device reduction. */
int32_t mcs_pop_moments(mcs_pop *p, const mcs_table *t, mcs_moments *out);
/* The same reductions, bitwise equal to mcs_pop_moments, with no device or host
 * allocation: the partial results live in a fixed 176.5 KiB device array of the
 * library (one per GPU, held until the process ends), enough for the reduction's
 * largest grid, so every population size fits.  No counterpart in the reference:
 * for callers that cannot allocate (tight serving loops, code between HIP graph
 * captures, harnesses that forbid hipMalloc).  Calls on one GPU share the array:
 * calls that overlap on one GPU (several contexts or threads) race. */
int32_t mcs_pop_moments_noalloc(mcs_pop *p, const mcs_table *t, mcs_moments *out);

/* Energies (J, energy(p) of JuMC/src/electron.jl:27) of the active particles in slot
 * order into energy[0..n-1] (n = the active count, returned even when it exceeds nmax,
 * which is then MCS_E_ARG): the sample of the swarm's Kolmogorov-Smirnov stop
 * (Swarm/src/Swarm.jl:288-306). */
int32_t mcs_pop_energies(mcs_pop *p, int64_t nmax, int64_t *n, double *energy);

/* ---- the fused step: JuMC.advance! (JuMC/src/mixed_population.jl:24-54) ---
 * Yoshida push (electron.jl:51-62), null-collision sampling (collisions.jl:159-192),
 * collision kinematics (electron.jl:108-139) and births/deaths (collisions.jl:173-192,
 * population.jl:88-106) in one kernel.  tracker = NULL is VoidCollisionTracker; a mesh
 * adds +-w/dV ion charge to its QFIXED array (gridfields.jl:193-224): the step logs
 * its deposits and adds each cell's in particle order (then event order within a
 * particle) to the cell's value -- the reference's sequential order, deterministic,
 * no floating-point atomics. */
int32_t mcs_advance(mcs_pop *p, const mcs_table *t, const mcs_field *field, double dt,
                    mcs_mesh *tracker, uint64_t seed, uint32_t instance, uint32_t step,
                    mcs_step_stats *stats);

/* ---- photons: a population of PhotonState (JuMC/src/photon.jl:11-18) ----------
 * x, v (|v| = c), nu (absorption rate, 1/s), w, s, active, id; host layout as
 * mcs_pop_upload, nu one double per photon.  mcs_pop_size, mcs_repack,
 * mcs_pop_destroy and mcs_pop_moments take a photon population (the moments'
 * nparticles, nactive, weight and position/velocity sums; its energy fields and
 * collrate are 0: a photon has no energy in the reference, photon.jl:22, and t
 * may be NULL); mcs_pop_upload/download, the sorts, deposits and resampling do not. */
int32_t mcs_pop_create_photons(mcs_ctx *ctx, int64_t capacity, mcs_pop **out);
int32_t mcs_photons_upload(mcs_pop *p, int64_t n, const double *x, const double *v, const double *nu,
                           const double *w, const double *s, const uint8_t *active,
                           const uint64_t *id, uint64_t next_id);
int32_t mcs_photons_download(mcs_pop *p, int64_t nmax, int64_t *n, double *x, double *v, double *nu,
                             double *w, double *s, uint8_t *active, uint64_t *id, uint64_t *next_id);

/* advance!(mpopl, efield, dt, tracker) over MultiPopulation(:electron => electrons,
 * :photon => photons) (JuMC/src/mixed_population.jl:24-54; the populations in the
 * order main.jl:105-122 builds them):
 *  1. the electron loop of mcs_advance over the electrons present at the call; a
 *     PhotoEmission collision appends its photons to `photons`, in the order of
 *     the emitting electrons' slots, then of their collisions;
 *  2. the photon loop of ZhelezniakCollisions(numax) (photon.jl:27-56) over every
 *     photon present after the emission, the new ones included: free flights
 *     x += t v, a collision every s/numax, absorbed when nu > numax U
 *     (PhotoIonization): an electron at rest with the photon's weight and active
 *     flag is appended to `electrons` (in photon order), the photon is removed,
 *     and the tracker adds +w/dV at the photon's position (gridfields.jl:219-222).
 * A removed particle keeps flying until dt is used up, as in the reference.  The
 * tracker's additions of both loops reach each cell in that order.  Draws: the
 * electron streams as in mcs_advance (the photon count from the emitting
 * electron's stream), each new photon's direction, nu and s from a stream of its
 * own id, each photon's flights from another.  estats describes loop 1 (births =
 * ionization children); pstats loop 2: n_in photons before the call, n_out after
 * the emission, births = electrons appended, deaths = photons removed, flights,
 * events[0] = absorptions.  photons may be NULL when t has no photoemission row.
 * FP64 electron populations only. */
int32_t mcs_advance_mixed(mcs_pop *electrons, mcs_pop *photons, const mcs_table *t, double numax,
                          const mcs_field *field, double dt, mcs_mesh *tracker, uint64_t seed,
                          uint32_t instance, uint32_t step, mcs_step_stats *estats,
                          mcs_step_stats *pstats);

/* repack! (population.jl:191-215): last active particle into the first hole. */
int32_t mcs_repack(mcs_pop *p);

/* Swarm.resample! (Swarm/src/Swarm.jl:193-223): global roulette / Poisson cloning
 * towards n0 when the active count is off by at least tol. did = 1 if it acted. */
int32_t mcs_swarm_resample(mcs_pop *p, int64_t n0, double tol, uint64_t seed,
                           uint32_t instance, uint32_t step, int32_t *did);

/* ---- mesh: 3D Cartesian generalisation of MCStreamer.Grid/GridFields
 * (MCStreamer/src/grid.jl:4-100, gridfields.jl:1-84); origin at 0, spacing h. */
int32_t mcs_mesh_create(mcs_ctx *ctx, const int32_t n[3], double h, mcs_mesh **out);
int32_t mcs_mesh_destroy(mcs_mesh *m);
int32_t mcs_mesh_set(mcs_mesh *m, int32_t which, const double *host);
int32_t mcs_mesh_get(mcs_mesh *m, int32_t which, double *host);
int32_t mcs_mesh_zero(mcs_mesh *m, int32_t which);

/* Canonical cell order: drop inactive particles and sort the rest by (cell, id);
 * particles outside the mesh go last.  Builds the per-cell start offsets used by
 * deposit and resample.  Replaces repack! where the order is free (see DESIGN). */
int32_t mcs_sort_by_cell(mcs_pop *p, mcs_mesh *m);

/* density! (MCStreamer/src/poisson.jl:65-84): add w*val/dV of every active particle
 * to array `which` (QFIXED or QPART) with NGP or CIC weights.  Needs the population in
 * canonical cell order (mcs_sort_by_cell); accumulation is a per-cell gather, no atomics. */
int32_t mcs_deposit(mcs_pop *p, mcs_mesh *m, int32_t which, int32_t scheme, double val);

/* q = e*(qfixed + qpart) (poisson.jl:19-26 with ne = 0 and no denoiser). */
int32_t mcs_assemble_charge(mcs_mesh *m);

/* Face fields from the potential (poisson.jl:47-57): E = -grad u on faces plus eb
 * on z, with the tangential ghost copies. */
int32_t mcs_efield_faces(mcs_mesh *m, double eb);

/* MCStreamer.resample! (MCStreamer/src/resample.jl:7-84) per sorted cell: systematic
 * roulette to ntarget equal weights where a cell holds more than nmax particles,
 * pairwise splitting where it holds fewer than ntarget and w > 2.  Output stays in
 * canonical cell order and is compacted. */
int32_t mcs_resample_cells(mcs_pop *p, mcs_mesh *m, int32_t ntarget, int32_t nmax,
                           uint64_t seed, uint32_t instance, uint32_t step,
                           mcs_resample_stats *stats);

/* ---- multigrid: Multigrid.allocate / solve (Multigrid/src/Multigrid.jl:302-324,
 * 449-467), the fmg!/mgv! cycle (:334-434) with red-black Gauss-Seidel, residual,
 * restriction, trilinear prolongation and the coarse direct solve. */
int32_t mcs_mg_create(mcs_ctx *ctx, const mcs_mg_config *cfg, mcs_mg **out);
int32_t mcs_mg_destroy(mcs_mg *mg);
/* Solve L u + s b = 0 on host arrays with ghosts ((nx+2)(ny+2)(nz+2), x fastest;
 * (nx+2)(ny+2) for ndim 2); u is the initial guess on entry and the solution on
 * return. */
int32_t mcs_mg_solve_host(mcs_mg *mg, double *u, const double *b, int32_t *iters,
                          double *resid);
/* Same on device arrays of the context's device (ghost layout as above). */
int32_t mcs_mg_solve_device(mcs_mg *mg, double *u, const double *b, int32_t *iters, double *resid);
/* Same, on a mesh: u = mesh U, b = mesh Q (device resident). */
int32_t mcs_mg_solve_mesh(mcs_mg *mg, mcs_mesh *m, int32_t *iters, double *resid);
/* setinhom! (boundaries.jl:141-153): b += (2/s) val on the first valid layer of face
 * `face` (0..5); val has the face's interior size, first index fastest. */
int32_t mcs_mg_setinhom_mesh(mcs_mg *mg, mcs_mesh *m, int32_t face, const double *val);

/* ---- batched swarms: many Swarm.swarm instances (Swarm/src/Swarm.jl:102-142) in
 * one set of launches.  Each instance ("slot") owns a slab of a shared particle
 * store, its own collision table, uniform field, time step and resample target,
 * and a Philox key (seed, instance) so that its result does not depend on the
 * other slots.  One step of a slot = advance! + the tallies of Swarm.jl:126-130
 * (fused into the step kernel) + Swarm.resample! (Swarm.jl:193-223), and every
 * `compact_every` steps a stable compaction of the slab (the reference never
 * repacks a swarm; compaction only reorders, it changes no particle). */
#define MCS_TALLY_NQ 4   /* v_z, energy, ionization rate, attachment rate */
#define MCS_DIFFUSION_WINDOW 100   /* steps of the diffusion fit (the l of Swarm.jl:140) */

typedef struct mcs_swarm_batch mcs_swarm_batch;

typedef struct mcs_swarm_params {
    double e[3];          /* uniform field (V/m) */
    double dt;
    int64_t n0;           /* resample target */
    double tol;           /* resample tolerance (0.1 in Swarm.jl:193) */
    int64_t capacity;     /* slab size */
    uint32_t instance;    /* second Philox key word: the request id */
    int32_t relax_steps;  /* steps before the tallies start */
    int32_t block_steps;  /* steps per block mean (statistical error) */
    int32_t pad;
    double init_vth;      /* x == NULL in mcs_swarm_batch_add: n electrons at the origin,
                             Maxwellian velocity components init_vth * N(0,1), w = 1,
                             drawn on the device from phase INIT of the slot's key */
} mcs_swarm_params;

typedef struct mcs_swarm_tally {
    int64_t steps;        /* steps done */
    int64_t n;            /* slab slots in use (active or not) */
    int64_t nactive;      /* actives after the last step's advance */
    int64_t nrecorded;    /* steps entering the time averages */
    int64_t nblocks;      /* complete blocks */
    int32_t flags;        /* bit 0 table range, bit 1 field, bit 2 capacity, bit 3 NaN energy */
    int32_t pad;
    double sum[MCS_TALLY_NQ];    /* sum over recorded steps of the per-step means */
    double bm_s1[MCS_TALLY_NQ];  /* sum of block means */
    double bm_s2[MCS_TALLY_NQ];  /* sum of squared block means */
    double last[MCS_TALLY_NQ];   /* per-step means of the last step */
    double diffusion[3];  /* D_x, D_y, D_z: half the least-squares slope of sigma^2(t) =
                             <x_d^2> - <x_d>^2 over the actives after each step, fitted over
                             the last `window` steps (Swarm.jl:126-130,144-150) */
    int64_t window;       /* steps in that fit: min(steps, MCS_DIFFUSION_WINDOW) */
} mcs_swarm_tally;

/* precision MCS_PREC_F64: the FP64 store and step (bit-exact with a lone FP64
 * population); MCS_PREC_F32: split store (packed cell of a 1e-5 m virtual grid +
 * FP32 offset, FP32 velocity/weight/clock) with the FP32 step; tallies are FP64
 * compensated sums in both. */
int32_t mcs_swarm_batch_create(mcs_ctx *ctx, int64_t total_capacity, int32_t max_slots,
                               uint64_t seed, int32_t precision, mcs_swarm_batch **out);
int32_t mcs_swarm_batch_destroy(mcs_swarm_batch *b);
/* Admit an instance: copies its n initial particles (ids 0..n-1) into a free slab
 * (first fit), or draws them on the device when x is NULL (see init_vth).
 * MCS_E_CAPACITY when no slab of params->capacity is free. */
int32_t mcs_swarm_batch_add(mcs_swarm_batch *b, const mcs_table *t, const mcs_swarm_params *params,
                            int64_t n, const double *x, const double *v, const double *w,
                            const double *s, const uint8_t *active, int32_t *slot);
int32_t mcs_swarm_batch_remove(mcs_swarm_batch *b, int32_t slot);
/* nsteps steps of every slot; the step sequence is replayed from a HIP graph. */
int32_t mcs_swarm_batch_step(mcs_swarm_batch *b, int32_t nsteps);
int32_t mcs_swarm_batch_tally(mcs_swarm_batch *b, int32_t slot, mcs_swarm_tally *out);
int32_t mcs_swarm_batch_tally_all(mcs_swarm_batch *b, int32_t nslots, mcs_swarm_tally *out);
int32_t mcs_swarm_batch_download(mcs_swarm_batch *b, int32_t slot, int64_t nmax, int64_t *n,
                                 double *x, double *v, double *w, double *s, uint8_t *active,
                                 uint64_t *id);
/* A single population of n thermal electrons at the origin, drawn exactly as the
 * x == NULL form of mcs_swarm_batch_add (ids 0..n-1), so the plain per-instance
 * path and a batch slot with the same (seed, instance) start identically. */
int32_t mcs_pop_init_thermal(mcs_pop *p, int64_t n, uint64_t seed, uint32_t instance, double vth);
/* Launch shape and compaction cadence (workgroup size 128 or 256). */
int32_t mcs_swarm_batch_config(mcs_swarm_batch *b, int32_t block, int32_t compact_every);
/* Majorant of the null-collision sampler (collisions.jl:159-192 samples the next
 * collision at the table's one maxrate, mixed_population.jl:35-44).
 * MCS_MAJORANT_GLOBAL: that maxrate (the default; the only mode of an FP64 batch,
 * which stays bit-exact with a lone FP64 population).  MCS_MAJORANT_BANDED (FP32
 * batches): the flight clock runs at the majorant of an energy band [0, cap] that
 * holds the particle's energy with a quarter of headroom (mcs_table_majorants); a
 * flight whose parabola would leave the band stops where it reaches the cap and
 * the remaining exponential clock continues at the next band's majorant.  The
 * process of real collisions is the same Poisson process (the null-collision
 * method is exact for any majorant above the real rate), with far fewer null
 * events at low energy; the particle's draws are no longer those of the global
 * form, so results agree with it statistically, not draw for draw.  Rows of kind
 * MCS_PROC_NULL are not sampled in this mode (a null event is the rest of the band
 * majorant).  Capture of the step graph is redone at the next step. */
#define MCS_MAJORANT_GLOBAL 0
#define MCS_MAJORANT_BANDED 1
int32_t mcs_swarm_batch_majorant(mcs_swarm_batch *b, int32_t mode);
/* The energy bands of MCS_MAJORANT_BANDED: *nbands bands (at most nmax are
 * written), band k holding energies [0, ecap[k]] (J) with majorant rate[k] (1/s):
 * at least the summed rate of the table's non-null rows at every node up to the
 * first node above ecap[k] (so at every energy of the band, the rates being
 * linear between nodes), raised by 1e-5 for the FP32 rounding of the rows.  The
 * caps double from band to band up to the last node's energy; the band above
 * (ecap = +inf) has the table's maxrate. */
int32_t mcs_table_majorants(const mcs_table *t, int32_t nmax, int32_t *nbands, double *ecap, double *rate);

/* ---- axisymmetric (r, z) grid: MCStreamer.Grid / GridFields as the reference
 * builds them (MCStreamer/src/grid.jl:4-90, gridfields.jl:1-84): M cells in r over
 * [0, R], N in z over [0, L].  Cell-centred arrays QFIXED, QPART, Q, U are
 * (M+2)(N+2) with one ghost layer (index (i, j), i in 0..M+1, j in 0..N+1, r
 * fastest, the layout of the two-dimensional multigrid with ndim 2, cyl 1);
 * face arrays ER, EZ are (M+3)(N+3) (calloc_faces, i in 0..M+2, j in 0..N+2).
 * Spacings follow the reference's LinRanges: step(rc), step(zc) for FieldInterp,
 * faces and dV; step(rf), step(zf) for cellindex. */
#define MCS_CARR_QFIXED 0
#define MCS_CARR_QPART 1
#define MCS_CARR_Q 2
#define MCS_CARR_U 3
#define MCS_CARR_ER 4
#define MCS_CARR_EZ 5
/* Grid(R, L, M, N) (grid.jl:33-42) with its arrays zeroed (M, N >= 2). */
int32_t mcs_cgrid_create(mcs_ctx *ctx, double R, double L, int32_t M, int32_t N, mcs_cgrid **out);
int32_t mcs_cgrid_destroy(mcs_cgrid *g);
/* steps[0..3] = dr(grid), dz(grid) (grid.jl:74-75: steps of rc, zc), step(rf), step(zf). */
int32_t mcs_cgrid_steps(const mcs_cgrid *g, double steps[4]);
int32_t mcs_cgrid_set(mcs_cgrid *g, int32_t which, const double *host);
int32_t mcs_cgrid_get(mcs_cgrid *g, int32_t which, double *host);
int32_t mcs_cgrid_zero(mcs_cgrid *g, int32_t which);
/* density! (poisson.jl:65-84) on the axisymmetric grid: every active particle in
 * the NGP cell I = cellindex(grid, x) (grid.jl:84-90) that the ghosted array and
 * rc both hold adds (w val)/dV(grid, I[1]), dV = 2 pi rc[i] dr dz (grid.jl:76).
 * Per cell the terms are added in particle order; no particle order is needed. */
int32_t mcs_cyl_density(mcs_pop *p, mcs_cgrid *g, int32_t which, double val);
/* poisson! (poisson.jl:6-58): q = e (qfixed + qpart) rounded to FP32 as the Float32
 * denoiser input (poisson.jl:31-33, NullDenoiser, no fluid electrons), then
 * Multigrid.solve of u from the previous u with `mg` (ndim 2, n = (M, N), cyl 1),
 * and with freebc != 0 the rest of solve_freebc! (freebc.jl:4-36): the radial
 * gradient at r = R, its DST-II over f_n = L k_n (I1/I0 + K1/K0)(k_n R) / 2
 * (k_n = n pi / L), the DST-III back into the boundary value, setinhom! at r = R,
 * a second solve and the value into the ghost row; then the face fields with eb.
 * The transforms are direct sums (sequential in the index) over a table of
 * sin(pi m / 2N).  iters = the solves' iterations together.  MCS_E_NONCONVERGED
 * still leaves the faces computed. */
int32_t mcs_cyl_poisson(mcs_cgrid *g, mcs_mg *mg, double eb, int32_t freebc, int32_t *iters, double *resid);
/* The free boundary's f_n, n = 1..N (host computation: libstdc++ Bessel functions
 * below x = 50, large-argument expansions above). */
int32_t mcs_cgrid_freebc_coeffs(const mcs_cgrid *g, double *f);
/* Face fields from U (poisson.jl:50-58): er = (u[i-1,j] - u[i,j])/dr and
 * ez = eb + (u[i,j-1] - u[i,j])/dz for i in 1..M+1, j in 1..N+1, then the ghost
 * copies ez[0,:] = ez[1,:], ez[M+1,:] = ez[M,:], er[:,0] = er[:,1], er[:,N+1] = er[:,N].
 * U's ghosts must hold the boundary values (the solve leaves them so). */
int32_t mcs_cyl_efield_faces(mcs_cgrid *g, double eb);
/* Canonical cell order on the axisymmetric grid: inactive particles dropped, the
 * rest sorted by (cell, id) over resample!'s cells (cellindex within the ghosted
 * per-cell arrays, i in 1..M+1, j in 0..N+1), particles outside last.  The bits
 * of (M+2)(N+2) + 1 and of the next id must fit 64 together. */
int32_t mcs_cyl_sort_by_cell(mcs_pop *p, mcs_cgrid *g);
/* MCStreamer.resample! (resample.jl:7-84) per cell of the axisymmetric grid, on the
 * canonical order of mcs_cyl_sort_by_cell (called right before; the rules and
 * streams are those of mcs_resample_cells, keyed by the flat cell index). */
int32_t mcs_cyl_resample(mcs_pop *p, mcs_cgrid *g, int32_t ntarget, int32_t nmax, uint64_t seed,
                         uint32_t instance, uint32_t step, mcs_resample_stats *stats);
/* GridFields.p and .wtotal (gridfields.jl:41-48) as the last mcs_cyl_resample
 * left them (resample.jl:16-30, 63-72): per cell of the ghosted (M+2)(N+2)
 * layout the count of its particles plus one per split, and the sum of their
 * weights; zeros before the first resample.  Read by the HDF5 export
 * (plot.jl:210-240) and SavedGridFields (io.jl:9-44). */
int32_t mcs_cgrid_resample_fields(mcs_cgrid *g, int64_t *p, double *wtotal);
/* advance! (mixed_population.jl:24-54) in the axisymmetric FieldInterp
 * (gridfields.jl:115-190: bilinear er, ez on the staggered faces, er rotated by
 * x/r, y/r; zero outside 0 < r < R, 0 < z < L) with, when track != 0, the
 * CollisionTracker of the same grid (gridfields.jl:203-218: +-w/dV into QFIXED at
 * the NGP cell, added per cell in particle order, as in mcs_advance).
 * FP64 populations only; otherwise as mcs_advance. */
int32_t mcs_advance_cyl(mcs_pop *p, const mcs_table *t, mcs_cgrid *g, double dt, int32_t track,
                        uint64_t seed, uint32_t instance, uint32_t step, mcs_step_stats *stats);
/* mcs_advance_mixed in the axisymmetric FieldInterp of g, tracking into g's QFIXED
 * when track != 0. */
int32_t mcs_advance_cyl_mixed(mcs_pop *electrons, mcs_pop *photons, const mcs_table *t, double numax,
                              mcs_cgrid *g, double dt, int32_t track, uint64_t seed, uint32_t instance,
                              uint32_t step, mcs_step_stats *estats, mcs_step_stats *pstats);

#ifdef __cplusplus
}
#endif

#endif /* MCS_H */
