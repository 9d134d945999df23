/*
 * refcpu.h -- the ORACLE: a single-threaded FP64 C restatement of the
 * reference's per-step particle-mesh loop, used only by the tests, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, as the checker.
 * Nothing in the product (mcstreamer_amd/, include/) links or calls it.
 *
 * It follows the reference's operation order line by line (file:line cited at
 * each function in refcpu.c), with two stated substitutions that the GPU uses
 * too: random numbers come from the per-particle Philox streams of
 * mcstreamer_amd/csrc/mcs_rng.h (the reference calls the global rand()), and
 * the 3D Cartesian mesh generalises the reference's 2D cylindrical grid (which
 * the ref_cgrid / ref_cyl_* functions restate as it is).
 *
 * Pins (tests/): Philox known-answer vectors; elementary functions against
 * libm; table builder against the closed-form cross sections; Yoshida push
 * against the exact parabola; kinematics invariants; null-collision statistics;
 * swarm drift against the analytic constant-frequency model; repack against
 * the rank-matching formulation; multigrid against a dense solve; resampling
 * weight conservation and systematic-sampling bounds.  Axisymmetric grid
 * (test_oracle_axisym.py): ref_fld against exact rational floors; the
 * LinRange geometry and dV against closed forms; the cylindrical field against
 * linear fields it must reproduce; density against brute-force exact cells;
 * faces against a linear potential; the charge's FP32 rounding; the push in
 * uniform axial and radial fields against the parabola; the free boundary's
 * transforms against scipy's DST-II/III and its solution against a point
 * charge between grounded planes; the (r, z) resample against its per-cell
 * rules.  The free boundary's Bessel factors are an input (the product's f_n,
 * itself checked against scipy).  Mesh operators (test_oracle_mesh_pins.py):
 * the 3D and (r, z) canonical sorts against numpy's lexsort of (cell, id) over
 * independently computed cells, and their cell starts against a cumulative
 * bincount; the CIC deposit against the tent assignment function, charge and
 * first-moment conservation, its special points and dropped outside shares;
 * the NGP deposit against cell binning; the reciprocal in normdivrem against
 * the true quotient (it moves cell indices only where the interpolant is
 * continuous).  Each of these checks fails on a transposed cell key or a
 * dropped half-cell stagger.  See each test's header.
 */
#ifndef REFCPU_H
#define REFCPU_H

#include <stdint.h>

#define REF_MAX_PROC 32

int32_t ref_threads(void);

typedef struct ref_pop {
    int64_t n;
    int64_t cap;
    uint64_t next_id;
    double *x;       /* n x 3 */
    double *v;       /* n x 3 */
    double *w;
    double *s;
    uint8_t *active;
    uint64_t *id;
    double *nu;      /* photons: absorption rate (NULL for electrons) */
} ref_pop;

typedef struct ref_table {
    int32_t nproc;
    int32_t ne;
    int32_t egrid;       /* 0 uniform, 1 log */
    int32_t pad;
    const double *energy;
    const double *rate;  /* rate[k*nproc + j] */
    const int32_t *kind;
    const double *param;
    double maxrate;
    const double *photo; /* per row: log numin, log numax, photon_multiplier, photon_weight
                            (PhotoEmission rows, kind 5; NULL without such rows) */
} ref_table;

typedef struct ref_mesh {
    int32_t n[3];
    int32_t pad;
    double h;
    double *qfixed;      /* (nx+2)(ny+2)(nz+2) */
    double *qpart;
    double *q;
    double *u;
    double *ex;          /* (nx+1)(ny+2)(nz+2) */
    double *ey;          /* (nx+2)(ny+1)(nz+2) */
    double *ez;          /* (nx+2)(ny+2)(nz+1) */
    int64_t *cell_start; /* ncells + 2 */
} ref_mesh;

/* Axisymmetric grid (MCStreamer/src/grid.jl:4-90): M x N cells over [0,R] x [0,L]. */
typedef struct ref_cgrid {
    int32_t M, N;
    double R, L;
    double dr, dz;       /* step(rc), step(zc) */
    double drf, dzf;     /* step(rf), step(zf) */
    double *dv;          /* dV(grid, i), i = 1..M */
    double *qfixed;      /* (M+2)(N+2), r fastest */
    double *qpart;
    double *q;
    double *u;
    double *er;          /* (M+3)(N+3) */
    double *ez;
} ref_cgrid;

typedef struct ref_field {
    int32_t kind;        /* 0 uniform, 1 mesh, 2 axisymmetric grid */
    int32_t pad;
    double e[3];
    const ref_mesh *mesh;
    const ref_cgrid *cgrid;
} ref_field;

typedef struct ref_stats {
    int64_t n_in;
    int64_t n_out;
    int64_t births;
    int64_t deaths;
    int64_t flights;
    int64_t events[REF_MAX_PROC];
    int32_t flags;
    int32_t pad;
} ref_stats;

typedef struct ref_moments {
    int64_t nparticles;
    int64_t nactive;
    double weight;
    double wenergy;
    double maxenergy;
    double sum_x[3];
    double sum_x2[3];
    double sum_v[3];
    double sum_energy;
    double collrate[REF_MAX_PROC];
} ref_moments;

typedef struct ref_resample_stats {
    int64_t n_in;
    int64_t n_out;
    int64_t cells_merged;
    int64_t clones;
    int64_t splits;
    int64_t removed;
} ref_resample_stats;

typedef struct ref_mg {
    int32_t n[3];
    int32_t bc[6];       /* 0 Dirichlet, 1 Neumann; x-,x+,y-,y+,z-,z+ */
    double s;
    int32_t levels;
    int32_t smooth1;
    int32_t smooth2;
    int32_t maxiter;
    double tolerance;
    int32_t ndim;        /* 3 (also 0) or 2 (n[2] == 1; bc[0..3]) */
    int32_t cyl;         /* 0 CartesianConnector; 1 + d: CylindricalConnector{d + 1} */
} ref_mg;

/* 2D pieces (Multigrid.jl with N = 2), exported for the tests; cyl as in ref_mg */
void ref_mg2_apply_bc(const int32_t n[2], const int32_t bc[4], double *a);
void ref_mg2_gauss_seidel(const int32_t n[2], int32_t cyl, double *u, const double *b);
void ref_mg2_residual(const int32_t n[2], int32_t cyl, double *r, const double *u, const double *b, double s);
void ref_mg2_restrict(const int32_t nf[2], double *rh, const double *r);
void ref_mg2_interpolate(const int32_t nf[2], double *r, const double *rh, int32_t add);
int32_t ref_mg2_coarse_inverse(const int32_t n[2], const int32_t bc[4], double *ainv);

/* elementary pieces, exported for the tests */
void ref_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void ref_stream_u01(uint64_t seed, uint32_t instance, uint32_t phase, uint64_t id,
                    uint32_t step, int64_t count, double *out);
void ref_draws(uint64_t seed, uint32_t instance, uint32_t phase, uint64_t id0, uint32_t step,
               int32_t kind, double param, int64_t count, double *out);
void ref_math(int32_t which, int64_t count, const double *in, double *out, double *out2);
int32_t ref_indweight(const ref_table *t, double eng, int64_t *k, double *w);
void ref_efield(const ref_field *f, const double x[3], double e[3], int32_t *flags);
void ref_yoshida(const ref_field *f, double x[3], double v[3], double dt, int32_t *flags);

/* JuMC.advance! */
int32_t ref_advance(ref_pop *p, const ref_table *t, const ref_field *f, double dt,
                    ref_mesh *tracker, uint64_t seed, uint32_t instance, uint32_t step,
                    ref_stats *st);
/* advance! over MultiPopulation(:electron => e, :photon => ph): the electron loop
 * (photons of PhotoEmission collisions appended to ph), then the photon loop of
 * ZhelezniakCollisions(numax) (photoelectrons appended to e).  est/pst as the
 * product's mcs_advance_mixed. */
int32_t ref_advance_mixed(ref_pop *e, ref_pop *ph, const ref_table *t, double numax, const ref_field *f,
                          double dt, ref_mesh *tracker, uint64_t seed, uint32_t instance, uint32_t step,
                          ref_stats *est, ref_stats *pst);
int32_t ref_advance_cyl_mixed(ref_pop *e, ref_pop *ph, const ref_table *t, double numax, ref_cgrid *g,
                              double dt, int32_t track, uint64_t seed, uint32_t instance, uint32_t step,
                              ref_stats *est, ref_stats *pst);
/* JuMC.repack! */
void ref_repack(ref_pop *p);
/* population reductions and Swarm tallies */
void ref_moments_compute(const ref_pop *p, const ref_table *t, ref_moments *m);
/* Swarm.resample! */
void ref_swarm_init(ref_pop *p, int64_t n, uint64_t seed, uint32_t instance, double vth);
int32_t ref_swarm_resample(ref_pop *p, int64_t n0, double tol, uint64_t seed,
                           uint32_t instance, uint32_t step, int32_t *did);

/* mesh pieces */
void ref_cgrid_geometry(double R, double L, int32_t M, int32_t N, double steps[4], double *dv);
double ref_fld(double x, double y);
int32_t ref_advance_cyl(ref_pop *p, const ref_table *t, ref_cgrid *g, double dt, int32_t track,
                        uint64_t seed, uint32_t instance, uint32_t step, ref_stats *st);
void ref_cyl_density(const ref_pop *p, ref_cgrid *g, double *arr, double val);
void ref_cyl_charge(ref_cgrid *g);
void ref_cyl_faces(ref_cgrid *g, double eb);
void ref_cyl_freebc_bcond(const ref_cgrid *g, const double *f, double *bcond);
void ref_cyl_freebc_inhom(ref_cgrid *g, double s, const double *bcond);
void ref_cyl_freebc_ghost(ref_cgrid *g, const double *bcond);
int32_t ref_cyl_sort_by_cell(ref_pop *p, const ref_cgrid *g, int64_t *cell_start, ref_pop *scr);
int64_t ref_cell_key(const ref_mesh *m, const double x[3]);
int32_t ref_sort_by_cell(ref_pop *p, ref_mesh *m, ref_pop *scratch);
void ref_deposit(const ref_pop *p, ref_mesh *m, int32_t which, int32_t scheme, double val);
void ref_assemble_charge(ref_mesh *m);
void ref_efield_faces(ref_mesh *m, double eb);
/* resample!'s p and wtotal per cell of a sorted population (cell_start of
 * ncells + 1 entries): the count plus one per split, and the weight sum. */
void ref_resample_counts(const ref_pop *p, const int64_t *cell_start, int64_t ncells, int32_t ntarget,
                         int32_t nmax, int64_t *pcell, double *wcell);
int32_t ref_resample_cells(ref_pop *p, ref_mesh *m, int32_t ntarget, int32_t nmax,
                           uint64_t seed, uint32_t instance, uint32_t step,
                           ref_pop *scratch, ref_resample_stats *st);

/* multigrid */
void ref_mg_apply_bc(const int32_t n[3], const int32_t bc[6], double *a);
void ref_mg_gauss_seidel(const int32_t n[3], double *u, const double *b);
void ref_mg_residual(const int32_t n[3], double *r, const double *u, const double *b, double s);
void ref_mg_restrict(const int32_t nf[3], double *rh, const double *r);
void ref_mg_interpolate(const int32_t nf[3], double *r, const double *rh, int32_t add);
int32_t ref_mg_coarse_inverse(const int32_t n[3], const int32_t bc[6], double *ainv);
int32_t ref_mg_solve(const ref_mg *cfg, double *u, const double *b, int32_t *iters,
                     double *resid);

#endif
