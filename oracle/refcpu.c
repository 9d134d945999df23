/*
 * refcpu.c -- ORACLE (test infrastructure only; see refcpu.h).
 *
 * Single-threaded FP64 restatement of the reference's per-step loop.  Built
 * with -O2 -ffp-contract=off so that every a*b+c rounds twice, as Julia does
 * without @fastmath/muladd.  Every function cites the reference lines it
 * restates (paths relative to the reference tree).
 */
#include "refcpu.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../mcstreamer_amd/csrc/mcs_rng.h"
#include "../mcstreamer_amd/csrc/mcs_phys.h"

#define FLAG_TABLE_RANGE 1
#define FLAG_FIELD 2
#define FLAG_CAPACITY 4
#define FLAG_NAN 8

/* ---------------------------------------------------------------- helpers */

void ref_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    mcs_philox4x32_10(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1], out);
}

void ref_stream_u01(uint64_t seed, uint32_t instance, uint32_t phase, uint64_t id,
                    uint32_t step, int64_t count, double *out)
{
    mcs_stream st;
    mcs_stream_init(&st, seed, instance, phase, id, step);
    for (int64_t i = 0; i < count; ++i) out[i] = mcs_u01(&st);
}

/* Draws of the reference's random primitives, one stream per sample (id0 + i):
 * kind 0 rand(), 1 nextcoll() (util.jl:17), 2 Poisson(param) (Swarm.jl:216),
 * 3 samplemin(param) (resample.jl:84), 4 randsphere() (util.jl:4-12, 3 values). */
void ref_draws(uint64_t seed, uint32_t instance, uint32_t phase, uint64_t id0, uint32_t step,
               int32_t kind, double param, int64_t count, double *out)
{
    for (int64_t i = 0; i < count; ++i) {
        mcs_stream st;
        mcs_stream_init(&st, seed, instance, phase, id0 + (uint64_t)i, step);
        switch (kind) {
        case 0: out[i] = mcs_u01(&st); break;
        case 1: out[i] = mcs_nextcoll(&st); break;
        case 2: out[i] = (double)mcs_poisson(&st, param); break;
        case 3: out[i] = mcs_samplemin(&st, param); break;
        default: mcs_randsphere(&st, &out[3 * i]); break;
        }
    }
}

void ref_math(int32_t which, int64_t count, const double *in, double *out, double *out2)
{
    for (int64_t i = 0; i < count; ++i) {
        switch (which) {
        case 0: out[i] = mcs_log(in[i]); break;
        case 1: out[i] = mcs_exp(in[i]); break;
        case 2: mcs_sincos2pi(in[i], &out[i], &out2[i]); break;
        default: out[i] = sqrt(in[i]); break;
        }
    }
}

/* energy(p) = 0.5 * mass(p) * (v1^2 + v2^2 + v3^2)   (JuMC/src/electron.jl:27) */
static inline double energy3(const double *v)
{
    return MCS_HALF_ME * ((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
}

/* indweight (JuMC/src/util.jl:23-30): index of the node before eng and the
 * weight of that node.  Uniform grid: k = floor((eng - e0)/step),
 * w = (e[k+1] - eng)/step.  Log grid: k located exactly by node comparisons,
 * w = (e[k+1] - eng)/(e[k+1] - e[k]).  Energies past the last node flag a range
 * error (the reference throws a BoundsError) and use the last interval; a NaN
 * energy flags NaN (floor(Int, NaN) throws an InexactError there). */
int32_t ref_indweight(const ref_table *t, double eng, int64_t *kout, double *wout)
{
    const double *e = t->energy;
    int64_t ne = t->ne;
    int32_t flag = 0;
    int64_t k;
    double w;
    if (!(eng == eng)) {
        *kout = 0;
        *wout = 0.0;
        return FLAG_NAN;
    }
    if (eng >= e[ne - 1]) {
        *kout = ne - 2;
        *wout = 0.0;
        return FLAG_TABLE_RANGE;
    }
    if (t->egrid == 0) {
        double step = (e[ne - 1] - e[0]) / (double)(ne - 1);
        double kf = floor((eng - e[0]) / step);
        if (kf < 0.0) kf = 0.0;
        k = (int64_t)kf;
        if (k > ne - 2) k = ne - 2;
        w = (e[k + 1] - eng) / step;
    } else {
        if (eng <= e[0]) {
            *kout = 0;
            *wout = 1.0;
            return 0;
        }
        double inv_dlog = (double)(ne - 1) / mcs_log(e[ne - 1] / e[0]);
        double kf = floor(mcs_log(eng / e[0]) * inv_dlog);
        if (kf < 0.0) kf = 0.0;
        if (kf > (double)(ne - 2)) kf = (double)(ne - 2);
        k = (int64_t)kf;
        while (k > 0 && e[k] > eng) k -= 1;
        while (k < ne - 2 && e[k + 1] <= eng) k += 1;
        w = (e[k + 1] - eng) / (e[k + 1] - e[k]);
    }
    if (w < 0.0) w = 0.0;
    if (w > 1.0) w = 1.0;
    *kout = k;
    *wout = w;
    return flag;
}

/* rate(c, j, pre) = w*R[j,k] + (1-w)*R[j,k+1]   (JuMC/src/collisions.jl:112-116) */
static inline double rate_j(const ref_table *t, int j, int64_t k, double w)
{
    return w * t->rate[k * t->nproc + j] + (1.0 - w) * t->rate[(k + 1) * t->nproc + j];
}

/* ------------------------------------------------------------- fields */

static inline int64_t cidx(const int32_t n[3], int64_t i, int64_t j, int64_t k)
{
    return i + (int64_t)(n[0] + 2) * (j + (int64_t)(n[1] + 2) * k);
}

/* normdivrem + staggered (MCStreamer/src/gridfields.jl:94-109) for one axis:
 * face index m with offset d, and ghost-inclusive centre index m1 with offset d1.
 * normdivrem is @fastmath in the reference, which licenses x * (1/h) for x / h;
 * this restatement fixes that choice. */
static inline void axis_index(double xc, double inv_h, int32_t n, int64_t *m, double *d,
                              int64_t *m1, double *d1)
{
    double r = xc * inv_h;
    int64_t mi = (int64_t)trunc(r);
    if (mi > n - 1) mi = n - 1;
    double di = r - (double)mi;
    *m = mi;
    *d = di;
    if (di < 0.5) {
        *m1 = mi;
        *d1 = di + 0.5;
    } else {
        *m1 = mi + 1;
        *d1 = di - 0.5;
    }
}

static inline double trilerp(const double *a, int64_t sx, int64_t sxy, int64_t base,
                             double dx, double dy, double dz)
{
    double c00 = a[base] * (1.0 - dx) + a[base + 1] * dx;
    double c10 = a[base + sx] * (1.0 - dx) + a[base + sx + 1] * dx;
    double c01 = a[base + sxy] * (1.0 - dx) + a[base + sxy + 1] * dx;
    double c11 = a[base + sxy + sx] * (1.0 - dx) + a[base + sxy + sx + 1] * dx;
    double c0 = c00 * (1.0 - dy) + c10 * dy;
    double c1 = c01 * (1.0 - dy) + c11 * dy;
    return c0 * (1.0 - dz) + c1 * dz;
}

/* FieldInterp (MCStreamer/src/gridfields.jl:115-190), 3D Cartesian: zero outside
 * the open box (inside(), grid.jl:45-49); each component interpolated trilinearly
 * from its staggered face array; |E| > 1e8 V/m raises the field flag. */
/* ---------------------------------------------------- axisymmetric grid */

/* Base.lerpi: element j of LinRange(a, b, d + 1) is (1 - t) a + t b, t = j / d. */
static double lerpi(int64_t j, int64_t d, double a, double b)
{
    double t = (double)j / (double)d;
    return (1.0 - t) * a + t * b;
}

/* Grid(R, L, M, N) (grid.jl:33-42, 74-76): rf = LinRange(0, R, M + 1) and
 * rc = 0.5 * (rf[2:end] + rf[1:end-1]), a LinRange from 0.5 (rf[2] + rf[1]) to
 * 0.5 (rf[M+1] + rf[M]) of M points; dr = step(rc) = (last - first) / (M - 1);
 * likewise in z; dV(i) = 2 pi rc[i] dr dz.  steps = (dr, dz, step(rf), step(zf)). */
void ref_cgrid_geometry(double R, double L, int32_t M, int32_t N, double steps[4], double *dv)
{
    double r0 = 0.5 * (lerpi(1, M, 0.0, R) + lerpi(0, M, 0.0, R));
    double r1 = 0.5 * (lerpi(M, M, 0.0, R) + lerpi(M - 1, M, 0.0, R));
    double z0 = 0.5 * (lerpi(1, N, 0.0, L) + lerpi(0, N, 0.0, L));
    double z1 = 0.5 * (lerpi(N, N, 0.0, L) + lerpi(N - 1, N, 0.0, L));
    double dr = (r1 - r0) / (double)(M - 1), dz = (z1 - z0) / (double)(N - 1);
    steps[0] = dr;
    steps[1] = dz;
    steps[2] = (R - 0.0) / (double)M;
    steps[3] = (L - 0.0) / (double)N;
    double two_pi = 2.0 * 3.141592653589793;
    for (int32_t i = 0; i < M; ++i) dv[i] = two_pi * lerpi(i, M - 1, r0, r1) * dr * dz;
}

/* fld(x, y), y > 0 (Julia Base: the largest integer q with q <= x / y, the exact
 * quotient): floor of the rounded quotient, lowered by one when q y exceeds x. */
double ref_fld(double x, double y)
{
    double q = floor(x / y);
    if (fma(-q, y, x) < 0.0) q -= 1.0;
    else if (fma(-(q + 1.0), y, x) >= 0.0) q += 1.0;
    return q;
}

/* cellindex (grid.jl:84-90) with the bounds tests of density! and track1 (the
 * ghosted array holds i in 0..M+1, j in 0..N+1; rc holds i in 1..M): flat index
 * i + (M+2) j, or -1; *ci = i. */
static int64_t cyl_cell(const ref_cgrid *g, const double x[3], int64_t *ci)
{
    double rho = sqrt(x[0] * x[0] + x[1] * x[1]);
    double fi = ref_fld(rho - 0.0, g->drf), fj = ref_fld(x[2] - 0.0, g->dzf);
    if (!(fi >= 0.0 && fi < (double)g->M && fj >= -1.0 && fj <= (double)g->N)) return -1;
    *ci = (int64_t)fi + 1;
    return *ci + (int64_t)(g->M + 2) * ((int64_t)fj + 1);
}

/* FieldInterp (gridfields.jl:115-190) on the axisymmetric grid: zero unless
 * 0 < r < R and 0 < z < L (inside, grid.jl:45-49); normdivrem (@fastmath x / y,
 * restated as x * (1/y) like the 3D gather; trunc) and staggered per axis (:94-109); er bilinear on (i + 1, j1), ez on (i1, j + 1)
 * with the 0-based origin of the offset arrays; er rotated by x / r, y / r. */
static void cyl_efield(const ref_cgrid *g, const double x[3], double e[3], int32_t *flags)
{
    double r = sqrt(x[0] * x[0] + x[1] * x[1]);
    if (!(0.0 < r && r < g->R && 0.0 < x[2] && x[2] < g->L)) {
        e[0] = 0.0;
        e[1] = 0.0;
        e[2] = 0.0;
        return;
    }
    double qr = r * (1.0 / g->dr), qz = x[2] * (1.0 / g->dz);
    int64_t i = (int64_t)trunc(qr), j = (int64_t)trunc(qz);
    if (i > g->M) i = g->M;   /* never taken: r < R bounds qr by M */
    if (j > g->N) j = g->N;
    double dr = qr - (double)i, dz = qz - (double)j;
    int64_t i1, j1;
    double dr1, dz1;
    if (dr < 0.5) { i1 = i; dr1 = dr + 0.5; } else { i1 = i + 1; dr1 = dr - 0.5; }
    if (dz < 0.5) { j1 = j; dz1 = dz + 0.5; } else { j1 = j + 1; dz1 = dz - 0.5; }
    int64_t S = g->M + 3;
    const double *er = g->er, *ez = g->ez;
    int64_t I = i + 1, J = j + 1;
    double er1 = er[I + S * j1] * (1.0 - dr) * (1.0 - dz1) + er[I + 1 + S * j1] * dr * (1.0 - dz1) +
                 er[I + S * (j1 + 1)] * (1.0 - dr) * dz1 + er[I + 1 + S * (j1 + 1)] * dr * dz1;
    double ez1 = ez[i1 + S * J] * (1.0 - dr1) * (1.0 - dz) + ez[i1 + 1 + S * J] * dr1 * (1.0 - dz) +
                 ez[i1 + S * (J + 1)] * (1.0 - dr1) * dz + ez[i1 + 1 + S * (J + 1)] * dr1 * dz;
    e[0] = er1 * x[0] / r;
    e[1] = er1 * x[1] / r;
    e[2] = ez1;
    double e2 = (e[0] * e[0] + e[1] * e[1]) + e[2] * e[2];
    if (e2 > MCS_EFIELD_LIMIT2) *flags |= FLAG_FIELD;
}

/* density! (poisson.jl:65-84) on the axisymmetric grid, particles in order. */
void ref_cyl_density(const ref_pop *p, ref_cgrid *g, double *arr, double val)
{
    for (int64_t k = 0; k < p->n; ++k) {
        if (!p->active[k]) continue;
        int64_t i;
        int64_t c = cyl_cell(g, p->x + 3 * k, &i);
        if (c < 0) continue;
        arr[c] += p->w[k] * val / g->dv[i - 1];
    }
}

/* poisson.jl:19-33: q = Float64(Float32(e (qfixed + qpart))) on the valid cells
 * (one thread slice, no fluid electrons, NullDenoiser). */
void ref_cyl_charge(ref_cgrid *g)
{
    int64_t su = g->M + 2;
    for (int64_t j = 1; j <= g->N; ++j)
        for (int64_t i = 1; i <= g->M; ++i) {
            int64_t c = i + su * j;
            double q0 = g->qfixed[c];
            q0 += g->qpart[c];
            q0 *= MCS_ELEMENTARY_CHARGE;
            g->q[c] = (double)(float)q0;
        }
}

/* poisson.jl:50-58: face fields from u, then the ghost copies. */
void ref_cyl_faces(ref_cgrid *g, double eb)
{
    int64_t M = g->M, N = g->N, su = M + 2, sf = M + 3;
    for (int64_t i = 1; i <= M + 1; ++i)
        for (int64_t j = 1; j <= N + 1; ++j) {
            g->er[i + sf * j] = (g->u[(i - 1) + su * j] - g->u[i + su * j]) / g->dr;
            g->ez[i + sf * j] = eb + (g->u[i + su * (j - 1)] - g->u[i + su * j]) / g->dz;
        }
    for (int64_t j = 0; j <= N + 2; ++j) g->ez[0 + sf * j] = g->ez[1 + sf * j];
    for (int64_t j = 0; j <= N + 2; ++j) g->ez[(M + 1) + sf * j] = g->ez[M + sf * j];
    for (int64_t i = 0; i <= M + 2; ++i) g->er[i] = g->er[i + sf];
    for (int64_t i = 0; i <= M + 2; ++i) g->er[i + sf * (N + 1)] = g->er[i + sf * N];
}

/* solve_freebc! (MCStreamer/src/freebc.jl:4-36) after its first solve:
 * grad = (u[ghost row] - u[last row]) / dr; am = (-(dr/2) RODFT10(grad)) ./ f;
 * bcond = RODFT01(am) / 2 + RODFT01(am1), am1 = am's last element alone.  FFTW's
 * RODFT10 is Y_k = 2 sum_j X_j sin(pi (j + 1/2)(k + 1) / N) and RODFT01 is
 * Y_k = (-1)^k X_{N-1} + 2 sum_{j < N-1} X_j sin(pi (j + 1)(k + 1/2) / N); here as
 * direct sums over S[m] = sin(pi m / 2N) in a fixed order: 64 partial sums (the
 * terms j = l, l + 64, ... in sequence for l < 64), then the xor butterfly over
 * the partials (offsets 32, ..., 1), read at partial 0.  f is the caller's
 * f_n = L k_n (I1/I0 + K1/K0)(k_n R) / 2. */
static double tree64(double *p)
{
    double q[64];
    for (int off = 32; off >= 1; off >>= 1) {
        for (int l = 0; l < 64; ++l) q[l] = p[l] + p[l ^ off];
        for (int l = 0; l < 64; ++l) p[l] = q[l];
    }
    return p[0];
}

void ref_cyl_freebc_bcond(const ref_cgrid *g, const double *f, double *bcond)
{
    int64_t M = g->M, N = g->N, su = M + 2, four = 4 * N;
    double *S = malloc(sizeof(double) * four);
    double *grad = malloc(sizeof(double) * N);
    double *am = malloc(sizeof(double) * N);
    for (int64_t m = 0; m < four; ++m) S[m] = sin(3.141592653589793 * (double)m / (2.0 * (double)N));
    for (int64_t j = 0; j < N; ++j) grad[j] = (g->u[(M + 1) + su * (j + 1)] - g->u[M + su * (j + 1)]) / g->dr;
    double part[64];
    for (int64_t k = 0; k < N; ++k) {
        for (int l = 0; l < 64; ++l) {
            part[l] = 0.0;
            for (int64_t j = l; j < N; j += 64) part[l] += grad[j] * S[((2 * j + 1) * (k + 1)) % four];
        }
        double acc = tree64(part);
        double bm = -(g->dr / 2.0) * (2.0 * acc);
        am[k] = bm / f[k];
    }
    for (int64_t k = 0; k < N; ++k) {
        for (int l = 0; l < 64; ++l) {
            part[l] = 0.0;
            for (int64_t j = l; j + 1 < N; j += 64) part[l] += am[j] * S[((j + 1) * (2 * k + 1)) % four];
        }
        double acc = tree64(part);
        double sg = (k & 1) ? -am[N - 1] : am[N - 1];
        double t = sg + 2.0 * acc;
        double t1 = sg + 2.0 * 0.0;
        bcond[k] = t / 2.0 + t1;
    }
    free(S);
    free(grad);
    free(am);
}

/* setinhom!(mg, q, TopBnd(), bcond) (Multigrid/src/boundaries.jl:141-153) with the
 * cylindrical connector: q[last row] += (2/s) (1 + 1/(2M - 1)) bcond. */
void ref_cyl_freebc_inhom(ref_cgrid *g, double s, const double *bcond)
{
    int64_t M = g->M, su = M + 2;
    double c = (2.0 / s) * (1.0 + 1.0 / (double)(2 * M - 1));
    for (int64_t j = 0; j < g->N; ++j) g->q[M + su * (j + 1)] += c * bcond[j];
}

/* u[end, :] .+= 2 .* bcond (freebc.jl:35). */
void ref_cyl_freebc_ghost(ref_cgrid *g, const double *bcond)
{
    int64_t M = g->M, su = M + 2;
    for (int64_t j = 0; j < g->N; ++j) g->u[(M + 1) + su * (j + 1)] += 2.0 * bcond[j];
}

void ref_efield(const ref_field *f, const double x[3], double e[3], int32_t *flags)
{
    if (f->kind == 2) {
        cyl_efield(f->cgrid, x, e, flags);
        return;
    }
    if (f->kind == 0) {
        e[0] = f->e[0];
        e[1] = f->e[1];
        e[2] = f->e[2];
        return;
    }
    const ref_mesh *m = f->mesh;
    const int32_t *n = m->n;
    double h = m->h;
    if (!(0.0 < x[0] && x[0] < (double)n[0] * h && 0.0 < x[1] && x[1] < (double)n[1] * h &&
          0.0 < x[2] && x[2] < (double)n[2] * h)) {
        e[0] = 0.0;
        e[1] = 0.0;
        e[2] = 0.0;
        return;
    }
    int64_t mx, my, mz, m1x, m1y, m1z;
    double dx, dy, dz, d1x, d1y, d1z;
    double inv_h = 1.0 / h;
    axis_index(x[0], inv_h, n[0], &mx, &dx, &m1x, &d1x);
    axis_index(x[1], inv_h, n[1], &my, &dy, &m1y, &d1y);
    axis_index(x[2], inv_h, n[2], &mz, &dz, &m1z, &d1z);
    int64_t nx = n[0], ny = n[1];
    /* EX: (nx+1)(ny+2)(nz+2) */
    {
        int64_t sx = nx + 1, sxy = (nx + 1) * (ny + 2);
        e[0] = trilerp(m->ex, sx, sxy, mx + sx * m1y + sxy * m1z, dx, d1y, d1z);
    }
    /* EY: (nx+2)(ny+1)(nz+2) */
    {
        int64_t sx = nx + 2, sxy = (nx + 2) * (ny + 1);
        e[1] = trilerp(m->ey, sx, sxy, m1x + sx * my + sxy * m1z, d1x, dy, d1z);
    }
    /* EZ: (nx+2)(ny+2)(nz+1) */
    {
        int64_t sx = nx + 2, sxy = (nx + 2) * (ny + 2);
        e[2] = trilerp(m->ez, sx, sxy, m1x + sx * m1y + sxy * mz, d1x, d1y, dz);
    }
    double e2 = (e[0] * e[0] + e[1] * e[1]) + e[2] * e[2];
    if (e2 > MCS_EFIELD_LIMIT2) *flags |= FLAG_FIELD;
}

/* advance_free_yoshida (JuMC/src/electron.jl:51-62):
 *   x1 = x + (c1 dt) v;  v1 = v + ((d1 dt) q/m) E(x1);  ...;  x4 = x3 + (c4 dt) v3. */
void ref_yoshida(const ref_field *f, double x[3], double v[3], double dt, int32_t *flags)
{
    double e[3];
    double c1 = MCS_YOSHIDA_C1 * dt, c2 = MCS_YOSHIDA_C2 * dt;
    double c3 = MCS_YOSHIDA_C3 * dt, c4 = MCS_YOSHIDA_C4 * dt;
    double f1 = (MCS_YOSHIDA_D1 * dt) * MCS_QM;
    double f2 = (MCS_YOSHIDA_D2 * dt) * MCS_QM;
    double f3 = (MCS_YOSHIDA_D3 * dt) * MCS_QM;
    for (int d = 0; d < 3; ++d) x[d] = x[d] + c1 * v[d];
    ref_efield(f, x, e, flags);
    for (int d = 0; d < 3; ++d) v[d] = v[d] + f1 * e[d];
    for (int d = 0; d < 3; ++d) x[d] = x[d] + c2 * v[d];
    ref_efield(f, x, e, flags);
    for (int d = 0; d < 3; ++d) v[d] = v[d] + f2 * e[d];
    for (int d = 0; d < 3; ++d) x[d] = x[d] + c3 * v[d];
    ref_efield(f, x, e, flags);
    for (int d = 0; d < 3; ++d) v[d] = v[d] + f3 * e[d];
    for (int d = 0; d < 3; ++d) x[d] = x[d] + c4 * v[d];
}

/* track1 (MCStreamer/src/gridfields.jl:203-213): the NGP cell of x in the fixed
 * charge array and the amount val/dV it receives (NULL outside the mesh). */
static double *track1(ref_mesh *m, const double x[3], double val, double *amount)
{
    if (!m) return NULL;
    int64_t c[3];
    for (int d = 0; d < 3; ++d) {
        double f = floor(x[d] / m->h);
        if (!(f >= 0.0 && f < (double)m->n[d])) return NULL;
        c[d] = (int64_t)f;
    }
    double dV = (m->h * m->h) * m->h;
    *amount = val / dV;
    return &m->qfixed[cidx(m->n, c[0] + 1, c[1] + 1, c[2] + 1)];
}

/* track1 on the axisymmetric grid (gridfields.jl:203-213). */
static double *track1_cyl(ref_cgrid *g, const double x[3], double val, double *amount)
{
    if (!g) return NULL;
    int64_t i;
    int64_t c = cyl_cell(g, x, &i);
    if (c < 0) return NULL;
    *amount = val / g->dv[i - 1];
    return &g->qfixed[c];
}

/* add_particle! (JuMC/src/population.jl:88-97) */
static int append(ref_pop *p, const double x[3], const double v[3], double w, double s,
                  uint8_t active)
{
    if (p->n >= p->cap) return 0;
    int64_t i = p->n;
    for (int d = 0; d < 3; ++d) {
        p->x[3 * i + d] = x[d];
        p->v[3 * i + d] = v[d];
    }
    p->w[i] = w;
    p->s[i] = s;
    p->active[i] = active;
    p->id[i] = p->next_id;
    p->next_id += 1;
    p->n += 1;
    return 1;
}

/* advance! (JuMC/src/mixed_population.jl:24-54) with do_one_collision!
 * (collisions.jl:159-192), collide (electron.jl:108-139), apply!
 * (collisions.jl:173-192) and the tracker (gridfields.jl:215-218).
 * The particle range is fixed when the loop starts, so children are not
 * advanced in the step they are born; a particle removed by attachment keeps
 * flying and colliding until its time step is used up. */
static int32_t advance_impl(ref_pop *p, const ref_table *t, const ref_field *f, double dt,
                            ref_mesh *tracker, ref_cgrid *ctracker, ref_pop *ph, uint64_t seed,
                            uint32_t instance, uint32_t step, ref_stats *st);

int32_t ref_advance(ref_pop *p, const ref_table *t, const ref_field *f, double dt,
                    ref_mesh *tracker, uint64_t seed, uint32_t instance, uint32_t step,
                    ref_stats *st)
{
    return advance_impl(p, t, f, dt, tracker, NULL, NULL, seed, instance, step, st);
}

/* advance! in the axisymmetric FieldInterp of g, tracking into g->qfixed. */
int32_t ref_advance_cyl(ref_pop *p, const ref_table *t, ref_cgrid *g, double dt, int32_t track,
                        uint64_t seed, uint32_t instance, uint32_t step, ref_stats *st)
{
    ref_field f;
    memset(&f, 0, sizeof(f));
    f.kind = 2;
    f.cgrid = g;
    return advance_impl(p, t, &f, dt, NULL, track ? g : NULL, NULL, seed, instance, step, st);
}

/* What one chunk of the particle loop produces, in particle order: the children
 * apply! appends and the tracker's deposits.  The loop's chunks are consecutive
 * particle ranges; their outputs are applied chunk after chunk, which is the
 * order of the reference's sequential loop (children are not advanced in the
 * step they are born, and the fixed-charge array is not read by the step). */
typedef struct {
    double x[3], v[3], w, s;
    uint8_t act;
} birth_rec;

typedef struct {
    double *q;
    double amount;
} dep_rec;

/* a PhotoEmission collision's photons (count n > 0) */
typedef struct {
    double x[3];
    int64_t n;
    int32_t row;
    uint8_t act;
} burst_rec;

typedef struct {
    birth_rec *b;
    int64_t nb, cb;
    dep_rec *d;
    int64_t nd, cd;
    burst_rec *u;
    int64_t nu, cu;
    ref_stats st;
} chunk_out;

static void push_burst(chunk_out *o, int32_t row, const double x[3], uint8_t act, int64_t n)
{
    if (o->nu == o->cu) {
        o->cu = o->cu ? 2 * o->cu : 64;
        o->u = (burst_rec *)realloc(o->u, sizeof(burst_rec) * (size_t)o->cu);
    }
    burst_rec *r = &o->u[o->nu++];
    for (int d = 0; d < 3; ++d) r->x[d] = x[d];
    r->n = n;
    r->row = row;
    r->act = act;
}

static void push_birth(chunk_out *o, const double x[3], const double v[3], double w, double s, uint8_t act)
{
    if (o->nb == o->cb) {
        o->cb = o->cb ? 2 * o->cb : 256;
        o->b = (birth_rec *)realloc(o->b, sizeof(birth_rec) * (size_t)o->cb);
    }
    birth_rec *r = &o->b[o->nb++];
    for (int d = 0; d < 3; ++d) {
        r->x[d] = x[d];
        r->v[d] = v[d];
    }
    r->w = w;
    r->s = s;
    r->act = act;
}

static void push_deposit(chunk_out *o, double *q, double amount)
{
    if (!q) return;
    if (o->nd == o->cd) {
        o->cd = o->cd ? 2 * o->cd : 256;
        o->d = (dep_rec *)realloc(o->d, sizeof(dep_rec) * (size_t)o->cd);
    }
    o->d[o->nd].q = q;
    o->d[o->nd].amount = amount;
    o->nd++;
}

static void advance_range(ref_pop *p, const ref_table *t, const ref_field *f, double dt, ref_mesh *tracker,
                          ref_cgrid *ctracker, uint64_t seed, uint32_t instance, uint32_t step, int64_t i0,
                          int64_t i1, chunk_out *o)
{
    ref_stats *st = &o->st;
    double maxrate = t->maxrate;
    for (int64_t i = i0; i < i1; ++i) {
        if (!p->active[i]) continue;
        mcs_stream rs;
        mcs_stream_init(&rs, seed, instance, MCS_PHASE_ADVANCE, p->id[i], step);
        double x[3] = {p->x[3 * i], p->x[3 * i + 1], p->x[3 * i + 2]};
        double v[3] = {p->v[3 * i], p->v[3 * i + 1], p->v[3 * i + 2]};
        double w = p->w[i];
        double s = p->s[i];
        uint8_t act = p->active[i];
        double trem = dt;
        while (trem > 0.0) {
            double tnext = s / maxrate;
            double tt;
            int collides;
            if (trem > tnext) {
                tt = tnext;
                collides = 1;
            } else {
                tt = trem;
                collides = 0;
                s = s - tt * maxrate;
            }
            ref_yoshida(f, x, v, tt, &st->flags);
            st->flights += 1;
            if (collides) {
                double eng = energy3(v);
                int64_t k;
                double wgt;
                st->flags |= ref_indweight(t, eng, &k, &wgt);
                double xi = mcs_u01(&rs) * maxrate;
                for (int j = 0; j < t->nproc; ++j) {
                    double nu = rate_j(t, j, k, wgt);
                    if (nu > xi) {
                        st->events[j] += 1;
                        int kind = t->kind[j];
                        if (kind == 1) { /* Elastic, electron.jl:133-139 */
                            double mr = t->param[j];
                            double c = mr / (1.0 + mr);
                            double vcm[3] = {c * v[0], c * v[1], c * v[2]};
                            double dv[3] = {v[0] - vcm[0], v[1] - vcm[1], v[2] - vcm[2]};
                            double nrm = sqrt((dv[0] * dv[0] + dv[1] * dv[1]) + dv[2] * dv[2]);
                            double om[3];
                            mcs_randsphere(&rs, om);
                            for (int d = 0; d < 3; ++d) v[d] = om[d] * nrm + vcm[d];
                        } else if (kind == 2) { /* Excitation, electron.jl:108-114 */
                            double e1 = eng - t->param[j];
                            if (e1 < 0.0) e1 = 0.0;
                            double vabs = sqrt((2.0 * e1) / MCS_ELECTRON_MASS);
                            double om[3];
                            mcs_randsphere(&rs, om);
                            for (int d = 0; d < 3; ++d) v[d] = om[d] * vabs;
                        } else if (kind == 3) { /* Ionization, electron.jl:116-127 */
                            double e1 = eng - t->param[j];
                            if (e1 < 0.0) e1 = 0.0;
                            e1 = e1 / 2.0;
                            double vabs = sqrt((2.0 * e1) / MCS_ELECTRON_MASS);
                            double om1[3], om2[3], vc[3];
                            mcs_randsphere(&rs, om1);
                            mcs_randsphere(&rs, om2);
                            for (int d = 0; d < 3; ++d) {
                                v[d] = om1[d] * vabs;
                                vc[d] = om2[d] * vabs;
                            }
                            double sc = mcs_nextcoll(&rs);
                            double amount = 0.0;
                            double *q = track1(tracker, x, w, &amount);
                            push_deposit(o, q, amount);
                            q = track1_cyl(ctracker, x, w, &amount);
                            push_deposit(o, q, amount);
                            /* apply!: parent state written, child appended */
                            push_birth(o, x, vc, w, sc, act);
                        } else if (kind == 4) { /* Attachment, electron.jl:129-131 */
                            double amount = 0.0;
                            double *q = track1(tracker, x, -w, &amount);
                            push_deposit(o, q, amount);
                            q = track1_cyl(ctracker, x, -w, &amount);
                            push_deposit(o, q, amount);
                            if (act) st->deaths += 1;
                            act = 0;
                        } else if (kind == 5 && t->photo) {
                            /* PhotoEmission, electron.jl:141-144: nphot =
                             * rand(Poisson(w / multiplier / photon_weight)); the
                             * electron is unchanged, its photons are appended by
                             * apply! (collisions.jl:125-141) */
                            const double *pr = &t->photo[4 * j];
                            int64_t np = mcs_poisson(&rs, (w / pr[2]) / pr[3]);
                            if (np > 0) push_burst(o, j, x, act, np);
                        }
                        break;
                    } else {
                        xi = xi - nu;
                    }
                }
                s = mcs_nextcoll(&rs);
            }
            trem = trem - tt;
        }
        for (int d = 0; d < 3; ++d) {
            p->x[3 * i + d] = x[d];
            p->v[3 * i + d] = v[d];
        }
        p->s[i] = s;
        p->active[i] = act;
    }
}

#ifdef _OPENMP
#include <omp.h>
#define REF_THREADS omp_get_max_threads()
#else
#define REF_THREADS 1
#endif

/* Threads of the particle loop and the grid loops (1 in the single-threaded build). */
int32_t ref_threads(void) { return REF_THREADS; }

/* add_particle!(photons, PhotonState(x, c randsphere(), nu, photon_weight,
 * nextcoll(), act)) (collisions.jl:134-139), the draws from the new photon's id. */
static int append_photon(ref_pop *ph, const double x[3], double lmin, double lmax, double pw, uint8_t act,
                         uint64_t seed, uint32_t instance, uint32_t step)
{
    if (ph->n >= ph->cap) return 0;
    const int64_t i = ph->n;
    double v[3], nu, s;
    mcs_photon_emit(seed, instance, ph->next_id, step, lmin, lmax, v, &nu, &s);
    for (int d = 0; d < 3; ++d) {
        ph->x[3 * i + d] = x[d];
        ph->v[3 * i + d] = v[d];
    }
    ph->nu[i] = nu;
    ph->w[i] = pw;
    ph->s[i] = s;
    ph->active[i] = act;
    ph->id[i] = ph->next_id;
    ph->next_id += 1;
    ph->n += 1;
    return 1;
}

static int32_t advance_impl(ref_pop *p, const ref_table *t, const ref_field *f, double dt,
                            ref_mesh *tracker, ref_cgrid *ctracker, ref_pop *ph, uint64_t seed,
                            uint32_t instance, uint32_t step, ref_stats *st)
{
    memset(st, 0, sizeof(*st));
    const int64_t n0 = p->n;
    st->n_in = n0;
    const int nch = REF_THREADS;
    chunk_out *out = (chunk_out *)calloc((size_t)nch, sizeof(chunk_out));
#pragma omp parallel for schedule(static, 1)
    for (int c = 0; c < nch; ++c) {
        const int64_t i0 = n0 * c / nch, i1 = n0 * (c + 1) / nch;
        advance_range(p, t, f, dt, tracker, ctracker, seed, instance, step, i0, i1, &out[c]);
    }
    for (int c = 0; c < nch; ++c) {
        chunk_out *o = &out[c];
        st->flags |= o->st.flags;
        st->flights += o->st.flights;
        st->deaths += o->st.deaths;
        for (int j = 0; j < REF_MAX_PROC; ++j) st->events[j] += o->st.events[j];
        for (int64_t k = 0; k < o->nd; ++k) *o->d[k].q += o->d[k].amount;
        for (int64_t k = 0; k < o->nb; ++k) {
            const birth_rec *r = &o->b[k];
            if (append(p, r->x, r->v, r->w, r->s, r->act))
                st->births += 1;
            else
                st->flags |= FLAG_CAPACITY;
        }
        for (int64_t k = 0; k < o->nu; ++k) {
            const burst_rec *r = &o->u[k];
            const double *pr = &t->photo[4 * r->row];
            for (int64_t m = 0; m < r->n; ++m)
                if (!ph || !append_photon(ph, r->x, pr[0], pr[1], pr[3], r->act, seed, instance, step))
                    st->flags |= FLAG_CAPACITY;
        }
        free(o->b);
        free(o->d);
        free(o->u);
    }
    free(out);
    st->n_out = p->n;
    return st->flags;
}

/* The photon loop of advance! (mixed_population.jl:27-51) with
 * ZhelezniakCollisions(numax) (photon.jl:27-56) over photons [i0, i1): straight
 * flights (advance_free, photon.jl:54-56), a collision draws xi = numax rand()
 * (collisions.jl:165) and absorbs when nu > xi: collide(PhotoIonization)
 * (photon.jl:45-51) makes an electron at rest with s = nextcoll(), the tracker
 * adds +w at the photon (gridfields.jl:219-222), apply! removes the photon and
 * appends the electron (collisions.jl:117-123); then s = nextcoll(). */
static void photon_range(ref_pop *ph, double numax, double dt, ref_mesh *tracker, ref_cgrid *ctracker,
                         uint64_t seed, uint32_t instance, uint32_t step, int64_t i0, int64_t i1, chunk_out *o)
{
    ref_stats *st = &o->st;
    const double zero[3] = {0.0, 0.0, 0.0};
    for (int64_t i = i0; i < i1; ++i) {
        if (!ph->active[i]) continue;
        mcs_stream rs;
        mcs_stream_init(&rs, seed, instance, MCS_PHASE_PHOTON, ph->id[i], step);
        double x[3] = {ph->x[3 * i], ph->x[3 * i + 1], ph->x[3 * i + 2]};
        const double v[3] = {ph->v[3 * i], ph->v[3 * i + 1], ph->v[3 * i + 2]};
        const double nu = ph->nu[i], w = ph->w[i];
        double s = ph->s[i];
        uint8_t act = 1;
        double trem = dt;
        while (trem > 0.0) {
            const double tnext = s / numax;
            double tt;
            int collides;
            if (trem > tnext) {
                tt = tnext;
                collides = 1;
            } else {
                tt = trem;
                collides = 0;
                s = s - tt * numax;
            }
            for (int d = 0; d < 3; ++d) x[d] = x[d] + tt * v[d];
            st->flights += 1;
            if (collides) {
                const double xi = mcs_u01(&rs) * numax;
                if (nu > xi) {
                    st->events[0] += 1;
                    const double sc = mcs_nextcoll(&rs);
                    double amount = 0.0;
                    double *q = track1(tracker, x, w, &amount);
                    push_deposit(o, q, amount);
                    q = track1_cyl(ctracker, x, w, &amount);
                    push_deposit(o, q, amount);
                    push_birth(o, x, zero, w, sc, act);
                    if (act) st->deaths += 1;
                    act = 0;
                }
                s = mcs_nextcoll(&rs);
            }
            trem = trem - tt;
        }
        for (int d = 0; d < 3; ++d) ph->x[3 * i + d] = x[d];
        ph->s[i] = s;
        ph->active[i] = act;
    }
}

static int32_t photon_impl(ref_pop *ph, ref_pop *e, double numax, double dt, ref_mesh *tracker,
                           ref_cgrid *ctracker, uint64_t seed, uint32_t instance, uint32_t step, ref_stats *st)
{
    const int64_t n0 = ph->n;
    const int nch = REF_THREADS;
    chunk_out *out = (chunk_out *)calloc((size_t)nch, sizeof(chunk_out));
#pragma omp parallel for schedule(static, 1)
    for (int c = 0; c < nch; ++c) {
        const int64_t i0 = n0 * c / nch, i1 = n0 * (c + 1) / nch;
        photon_range(ph, numax, dt, tracker, ctracker, seed, instance, step, i0, i1, &out[c]);
    }
    for (int c = 0; c < nch; ++c) {
        chunk_out *o = &out[c];
        st->flags |= o->st.flags;
        st->flights += o->st.flights;
        st->deaths += o->st.deaths;
        st->events[0] += o->st.events[0];
        for (int64_t k = 0; k < o->nd; ++k) *o->d[k].q += o->d[k].amount;
        for (int64_t k = 0; k < o->nb; ++k) {
            const birth_rec *r = &o->b[k];
            if (append(e, r->x, r->v, r->w, r->s, r->act))
                st->births += 1;
            else
                st->flags |= FLAG_CAPACITY;
        }
        free(o->b);
        free(o->d);
        free(o->u);
    }
    free(out);
    st->n_out = ph->n;
    return st->flags;
}

int32_t ref_advance_mixed(ref_pop *e, ref_pop *ph, const ref_table *t, double numax, const ref_field *f,
                          double dt, ref_mesh *tracker, uint64_t seed, uint32_t instance, uint32_t step,
                          ref_stats *est, ref_stats *pst)
{
    memset(pst, 0, sizeof(*pst));
    pst->n_in = ph ? ph->n : 0;
    int32_t fl = advance_impl(e, t, f, dt, tracker, NULL, ph, seed, instance, step, est);
    if (!ph) return fl;
    return fl | photon_impl(ph, e, numax, dt, tracker, NULL, seed, instance, step, pst);
}

int32_t ref_advance_cyl_mixed(ref_pop *e, ref_pop *ph, const ref_table *t, double numax, ref_cgrid *g,
                              double dt, int32_t track, uint64_t seed, uint32_t instance, uint32_t step,
                              ref_stats *est, ref_stats *pst)
{
    ref_field f;
    memset(&f, 0, sizeof(f));
    f.kind = 2;
    f.cgrid = g;
    memset(pst, 0, sizeof(*pst));
    pst->n_in = ph ? ph->n : 0;
    int32_t fl = advance_impl(e, t, &f, dt, NULL, track ? g : NULL, ph, seed, instance, step, est);
    if (!ph) return fl;
    return fl | photon_impl(ph, e, numax, dt, NULL, track ? g : NULL, seed, instance, step, pst);
}

static void copy_slot(ref_pop *dst, int64_t di, const ref_pop *src, int64_t si)
{
    for (int d = 0; d < 3; ++d) {
        dst->x[3 * di + d] = src->x[3 * si + d];
        dst->v[3 * di + d] = src->v[3 * si + d];
    }
    dst->w[di] = src->w[si];
    dst->s[di] = src->s[si];
    dst->active[di] = src->active[si];
    dst->id[di] = src->id[si];
    if (src->nu) dst->nu[di] = src->nu[si];
}

/* repack! (JuMC/src/population.jl:191-215): two pointers, the last active
 * particle moves into the first hole; n becomes the number of actives. */
void ref_repack(ref_pop *p)
{
    int64_t l = p->n - 1;
    if (p->n == 0) return;
    while (l >= 0 && !p->active[l]) l -= 1;
    int64_t i = 0;
    while (i <= l) {
        if (!p->active[i]) {
            copy_slot(p, i, p, l);
            l -= 1;
            while (l >= 0 && !p->active[l]) l -= 1;
        }
        i += 1;
    }
    p->n = l + 1;
}

/* weight, actives, meanenergy, maxenergy (population.jl:114-158) and the swarm
 * tallies (Swarm.jl:126-127,152,156 with collrate :242-248), summed in order. */
void ref_moments_compute(const ref_pop *p, const ref_table *t, ref_moments *m)
{
    memset(m, 0, sizeof(*m));
    m->nparticles = p->n;
    double emax = -HUGE_VAL;
    for (int64_t i = 0; i < p->n; ++i) {
        const double *v = &p->v[3 * i];
        const double *x = &p->x[3 * i];
        double eng = energy3(v);
        if (eng > emax) emax = eng;
        if (!p->active[i]) continue;
        m->nactive += 1;
        m->weight += p->w[i];
        m->wenergy += p->w[i] * eng;
        m->sum_energy += eng;
        for (int d = 0; d < 3; ++d) {
            m->sum_x[d] += x[d];
            m->sum_x2[d] += x[d] * x[d];
            m->sum_v[d] += v[d];
        }
        if (t) {
            int64_t k;
            double w;
            ref_indweight(t, eng, &k, &w);
            for (int j = 0; j < t->nproc; ++j) m->collrate[j] += rate_j(t, j, k, w);
        }
    }
    m->maxenergy = p->n > 0 ? emax : 0.0;
}

/* Swarm.resample! (Swarm/src/Swarm.jl:193-223): one draw per active particle;
 * drop with probability (ncur-n)/ncur, or append Poisson((n-ncur)/ncur) copies.
 * The loop range is fixed at its start, so copies are not cloned again. */
int32_t ref_swarm_resample(ref_pop *p, int64_t n0, double tol, uint64_t seed,
                           uint32_t instance, uint32_t step, int32_t *did)
{
    int64_t ncur = 0;
    *did = 0;
    for (int64_t i = 0; i < p->n; ++i) ncur += p->active[i] ? 1 : 0;
    if (fabs((double)(ncur - n0) / (double)n0) < tol) return 0;
    if (ncur == 0) return 0;
    *did = 1;
    int64_t nloop = p->n;
    if (ncur > n0) {
        double pdrop = (double)(ncur - n0) / (double)ncur;
        for (int64_t i = 0; i < nloop; ++i) {
            if (!p->active[i]) continue;
            mcs_stream rs;
            mcs_stream_init(&rs, seed, instance, MCS_PHASE_SWARM_RESAMPLE, p->id[i], step);
            if (mcs_u01(&rs) < pdrop) p->active[i] = 0;
        }
    } else {
        double lam = (double)(n0 - ncur) / (double)ncur;
        for (int64_t i = 0; i < nloop; ++i) {
            if (!p->active[i]) continue;
            mcs_stream rs;
            mcs_stream_init(&rs, seed, instance, MCS_PHASE_SWARM_RESAMPLE, p->id[i], step);
            int64_t nnew = mcs_poisson(&rs, lam);
            for (int64_t j = 0; j < nnew; ++j) {
                double x[3] = {p->x[3 * i], p->x[3 * i + 1], p->x[3 * i + 2]};
                double v[3] = {p->v[3 * i], p->v[3 * i + 1], p->v[3 * i + 2]};
                if (!append(p, x, v, p->w[i], p->s[i], p->active[i])) return FLAG_CAPACITY;
            }
        }
    }
    return 0;
}

/* Initial swarm population: n electrons at the origin, Maxwellian velocities,
 * w = 1, ids 0..n-1 (Swarm/src/Swarm.jl:104-106 with a thermal start), from the
 * shared mcs_maxwell_init draw sequence. */
void ref_swarm_init(ref_pop *p, int64_t n, uint64_t seed, uint32_t instance, double vth)
{
    for (int64_t i = 0; i < n; ++i) {
        double v[3], s;
        mcs_maxwell_init(seed, instance, (uint64_t)i, vth, v, &s);
        for (int d = 0; d < 3; ++d) {
            p->x[3 * i + d] = 0.0;
            p->v[3 * i + d] = v[d];
        }
        p->w[i] = 1.0;
        p->s[i] = s;
        p->active[i] = 1;
        p->id[i] = (uint64_t)i;
    }
    p->n = n;
    p->next_id = (uint64_t)n;
}

/* ------------------------------------------------------------------ mesh */

/* cellindex (MCStreamer/src/grid.jl:84-90), 3D: floor(x/h) per axis; particles
 * outside [0, n h) get key ncells. */
int64_t ref_cell_key(const ref_mesh *m, const double x[3])
{
    int64_t c[3];
    int64_t ncells = (int64_t)m->n[0] * m->n[1] * m->n[2];
    for (int d = 0; d < 3; ++d) {
        double f = floor(x[d] / m->h);
        if (!(f >= 0.0 && f < (double)m->n[d])) return ncells;
        c[d] = (int64_t)f;
    }
    return c[0] + (int64_t)m->n[0] * (c[1] + (int64_t)m->n[1] * c[2]);
}

typedef struct { int64_t key; uint64_t id; int64_t src; } sort_item;

static int cmp_item(const void *a, const void *b)
{
    const sort_item *x = (const sort_item *)a, *y = (const sort_item *)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    if (x->id != y->id) return x->id < y->id ? -1 : 1;
    return 0;
}

/* Canonical cell order: actives sorted by (cell, id), outside particles last. */
int32_t ref_sort_by_cell(ref_pop *p, ref_mesh *m, ref_pop *scr)
{
    int64_t ncells = (int64_t)m->n[0] * m->n[1] * m->n[2];
    sort_item *items = (sort_item *)malloc(sizeof(sort_item) * (size_t)(p->n + 1));
    int64_t na = 0;
    for (int64_t i = 0; i < p->n; ++i) {
        if (!p->active[i]) continue;
        items[na].key = ref_cell_key(m, &p->x[3 * i]);
        items[na].id = p->id[i];
        items[na].src = i;
        na++;
    }
    qsort(items, (size_t)na, sizeof(sort_item), cmp_item);
    for (int64_t c = 0; c <= ncells + 1; ++c) m->cell_start[c] = 0;
    for (int64_t a = 0; a < na; ++a) {
        copy_slot(scr, a, p, items[a].src);
        m->cell_start[items[a].key + 1] += 1;
    }
    for (int64_t c = 1; c <= ncells + 1; ++c) m->cell_start[c] += m->cell_start[c - 1];
    for (int64_t a = 0; a < na; ++a) copy_slot(p, a, scr, a);
    p->n = na;
    free(items);
    return 0;
}

/* Canonical order on the axisymmetric grid: actives by (cell, id) over resample!'s
 * cells (cellindex within calloc_centers' bounds: i in 0..M+1, j in 0..N+1, flat
 * i + (M+2) j; resample.jl:21-23), outside particles last; cell_start has
 * (M+2)(N+2) + 2 entries. */
int32_t ref_cyl_sort_by_cell(ref_pop *p, const ref_cgrid *g, int64_t *cell_start, ref_pop *scr)
{
    int64_t ncg = (int64_t)(g->M + 2) * (g->N + 2);
    sort_item *items = (sort_item *)malloc(sizeof(sort_item) * (size_t)(p->n + 1));
    int64_t na = 0;
    for (int64_t i = 0; i < p->n; ++i) {
        if (!p->active[i]) continue;
        const double *x = &p->x[3 * i];
        double rho = sqrt(x[0] * x[0] + x[1] * x[1]);
        double fi = ref_fld(rho - 0.0, g->drf), fj = ref_fld(x[2] - 0.0, g->dzf);
        int64_t key = ncg;
        if (fi >= -1.0 && fi <= (double)g->M && fj >= -1.0 && fj <= (double)g->N)
            key = ((int64_t)fi + 1) + (int64_t)(g->M + 2) * ((int64_t)fj + 1);
        items[na].key = key;
        items[na].id = p->id[i];
        items[na].src = i;
        na++;
    }
    qsort(items, (size_t)na, sizeof(sort_item), cmp_item);
    for (int64_t c = 0; c <= ncg + 1; ++c) cell_start[c] = 0;
    for (int64_t a = 0; a < na; ++a) {
        copy_slot(scr, a, p, items[a].src);
        cell_start[items[a].key + 1] += 1;
    }
    for (int64_t c = 1; c <= ncg + 1; ++c) cell_start[c] += cell_start[c - 1];
    for (int64_t a = 0; a < na; ++a) copy_slot(p, a, scr, a);
    p->n = na;
    free(items);
    return 0;
}

/* density! (MCStreamer/src/poisson.jl:65-84): arr[cell] += w*val/dV per active
 * particle in population order (NGP), or the same amount spread with trilinear
 * weights over the 8 nearest cell centres (CIC); shares outside the mesh drop. */
#ifdef _OPENMP
/* The same sums as a gather per cell, for the OpenMP build (the CPU baseline):
 * cell C takes the shares of the particles of its 27 neighbour cells, walked in
 * increasing cell order -- increasing particle order in the canonical order of
 * the last ref_sort_by_cell, which must describe p -- so every cell adds its
 * terms in the order of the sequential loop. */
static void deposit_gather(const ref_pop *p, ref_mesh *m, double *arr, int32_t scheme, double val)
{
    const int64_t nx = m->n[0], ny = m->n[1], nz = m->n[2];
    const double h = m->h, dV = (h * h) * h;
    const int64_t *cs = m->cell_start;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t cz = 0; cz < nz; ++cz)
        for (int64_t cy = 0; cy < ny; ++cy)
            for (int64_t cx = 0; cx < nx; ++cx) {
                double acc = arr[cidx(m->n, cx + 1, cy + 1, cz + 1)];
                const int r = scheme == 0 ? 0 : 1;
                for (int64_t dz = -r; dz <= r; ++dz)
                    for (int64_t dy = -r; dy <= r; ++dy)
                        for (int64_t dx = -r; dx <= r; ++dx) {
                            const int64_t kx = cx + dx, ky = cy + dy, kz = cz + dz;
                            if (kx < 0 || kx >= nx || ky < 0 || ky >= ny || kz < 0 || kz >= nz) continue;
                            const int64_t key = kx + nx * (ky + ny * kz);
                            for (int64_t i = cs[key]; i < cs[key + 1]; ++i) {
                                if (!p->active[i]) continue;
                                const double base = (p->w[i] * val) / dV;
                                if (scheme == 0) {
                                    acc += base;
                                    continue;
                                }
                                const double *x = &p->x[3 * i];
                                double wgt[3];
                                int hit = 1;
                                const int64_t tc[3] = {cx, cy, cz};
                                for (int d = 0; d < 3 && hit; ++d) {
                                    double g = x[d] / h - 0.5;
                                    double fl = floor(g);
                                    int64_t i0 = (int64_t)fl;
                                    double fw = g - fl;
                                    if (tc[d] == i0) wgt[d] = 1.0 - fw;
                                    else if (tc[d] == i0 + 1) wgt[d] = fw;
                                    else hit = 0;
                                }
                                if (hit) acc += base * ((wgt[0] * wgt[1]) * wgt[2]);
                            }
                        }
                arr[cidx(m->n, cx + 1, cy + 1, cz + 1)] = acc;
            }
}
#endif

void ref_deposit(const ref_pop *p, ref_mesh *m, int32_t which, int32_t scheme, double val)
{
    double *arr = which == 0 ? m->qfixed : m->qpart;
#ifdef _OPENMP
    if (omp_get_max_threads() > 1) {
        deposit_gather(p, m, arr, scheme, val);
        return;
    }
#endif
    double h = m->h;
    double dV = (h * h) * h;
    int64_t ncells = (int64_t)m->n[0] * m->n[1] * m->n[2];
    for (int64_t i = 0; i < p->n; ++i) {
        if (!p->active[i]) continue;
        const double *x = &p->x[3 * i];
        int64_t key = ref_cell_key(m, x);
        if (key >= ncells) continue;
        double base = (p->w[i] * val) / dV;
        if (scheme == 0) {
            int64_t c0 = key % m->n[0];
            int64_t c1 = (key / m->n[0]) % m->n[1];
            int64_t c2 = key / ((int64_t)m->n[0] * m->n[1]);
            arr[cidx(m->n, c0 + 1, c1 + 1, c2 + 1)] += base;
            continue;
        }
        int64_t i0[3];
        double fw[3];
        for (int d = 0; d < 3; ++d) {
            double g = x[d] / h - 0.5;
            double fl = floor(g);
            i0[d] = (int64_t)fl;
            fw[d] = g - fl;
        }
        for (int dz = 0; dz < 2; ++dz) {
            int64_t kz = i0[2] + dz;
            if (kz < 0 || kz >= m->n[2]) continue;
            double wz = dz ? fw[2] : 1.0 - fw[2];
            for (int dy = 0; dy < 2; ++dy) {
                int64_t ky = i0[1] + dy;
                if (ky < 0 || ky >= m->n[1]) continue;
                double wy = dy ? fw[1] : 1.0 - fw[1];
                for (int dx = 0; dx < 2; ++dx) {
                    int64_t kx = i0[0] + dx;
                    if (kx < 0 || kx >= m->n[0]) continue;
                    double wx = dx ? fw[0] : 1.0 - fw[0];
                    arr[cidx(m->n, kx + 1, ky + 1, kz + 1)] += base * ((wx * wy) * wz);
                }
            }
        }
    }
}

/* poisson! charge assembly (poisson.jl:19-26): q = (qfixed + qpart - ne) e, ne = 0. */
void ref_assemble_charge(ref_mesh *m)
{
    int64_t tot = (int64_t)(m->n[0] + 2) * (m->n[1] + 2) * (m->n[2] + 2);
    for (int64_t c = 0; c < tot; ++c) m->q[c] = 0.0;
#pragma omp parallel for schedule(static)
    for (int64_t k = 1; k <= m->n[2]; ++k)
        for (int64_t j = 1; j <= m->n[1]; ++j)
            for (int64_t i = 1; i <= m->n[0]; ++i) {
                int64_t c = cidx(m->n, i, j, k);
                m->q[c] = (m->qfixed[c] + m->qpart[c]) * MCS_ELEMENTARY_CHARGE;
            }
}

/* Face fields (poisson.jl:47-57): E_d = (u[c-1] - u[c])/h on faces, plus eb on
 * z; tangential ghost layers copy their neighbours, y (or x) first, then z. */
void ref_efield_faces(ref_mesh *m, double eb)
{
    int64_t nx = m->n[0], ny = m->n[1], nz = m->n[2];
    double h = m->h;
    const double *u = m->u;
#define U(i, j, k) u[cidx(m->n, (i), (j), (k))]
    /* EX (nx+1)(ny+2)(nz+2) */
#define EX(i, j, k) m->ex[(i) + (nx + 1) * ((j) + (ny + 2) * (k))]
#pragma omp parallel for schedule(static)
    for (int64_t k = 1; k <= nz; ++k)
        for (int64_t j = 1; j <= ny; ++j)
            for (int64_t i = 0; i <= nx; ++i) EX(i, j, k) = (U(i, j, k) - U(i + 1, j, k)) / h;
    for (int64_t k = 1; k <= nz; ++k)
        for (int64_t i = 0; i <= nx; ++i) {
            EX(i, 0, k) = EX(i, 1, k);
            EX(i, ny + 1, k) = EX(i, ny, k);
        }
    for (int64_t j = 0; j <= ny + 1; ++j)
        for (int64_t i = 0; i <= nx; ++i) {
            EX(i, j, 0) = EX(i, j, 1);
            EX(i, j, nz + 1) = EX(i, j, nz);
        }
#undef EX
#define EY(i, j, k) m->ey[(i) + (nx + 2) * ((j) + (ny + 1) * (k))]
#pragma omp parallel for schedule(static)
    for (int64_t k = 1; k <= nz; ++k)
        for (int64_t j = 0; j <= ny; ++j)
            for (int64_t i = 1; i <= nx; ++i) EY(i, j, k) = (U(i, j, k) - U(i, j + 1, k)) / h;
    for (int64_t k = 1; k <= nz; ++k)
        for (int64_t j = 0; j <= ny; ++j) {
            EY(0, j, k) = EY(1, j, k);
            EY(nx + 1, j, k) = EY(nx, j, k);
        }
    for (int64_t j = 0; j <= ny; ++j)
        for (int64_t i = 0; i <= nx + 1; ++i) {
            EY(i, j, 0) = EY(i, j, 1);
            EY(i, j, nz + 1) = EY(i, j, nz);
        }
#undef EY
#define EZ(i, j, k) m->ez[(i) + (nx + 2) * ((j) + (ny + 2) * (k))]
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k <= nz; ++k)
        for (int64_t j = 1; j <= ny; ++j)
            for (int64_t i = 1; i <= nx; ++i)
                EZ(i, j, k) = eb + (U(i, j, k) - U(i, j, k + 1)) / h;
    for (int64_t k = 0; k <= nz; ++k)
        for (int64_t j = 1; j <= ny; ++j) {
            EZ(0, j, k) = EZ(1, j, k);
            EZ(nx + 1, j, k) = EZ(nx, j, k);
        }
    for (int64_t k = 0; k <= nz; ++k)
        for (int64_t i = 0; i <= nx + 1; ++i) {
            EZ(i, 0, k) = EZ(i, 1, k);
            EZ(i, ny + 1, k) = EZ(i, ny, k);
        }
#undef EZ
#undef U
}

/* resample! (MCStreamer/src/resample.jl:7-84) on the canonical cell order.
 * Pass 1 per cell: count, total weight, and the first systematic offset
 * x = samplemin(ntarget) drawn when the count reaches nmax (:16-29).
 * Pass 2 in order (:31-76): a cell with p > nmax emits copies of weight
 * wtotal/ntarget while wtotal*x < wcum + w and advances x by
 * samplemin(ntarget-pk)*(1-x), then drops the original; a cell with
 * p < ntarget splits particles of w > 2 into round(w/2) and w - round(w/2),
 * raising p each time.  Output per cell: surviving originals, then the new
 * particles in creation order (new ids in output order); outside particles last. */
/* GridFields.p and .wtotal as resample! (MCStreamer/src/resample.jl:7-76) leaves
 * them: the first pass (:16-30) counts the cell's active particles into p and
 * adds their weights into wtotal in order; the second pass (:33-76) adds one to
 * p at each split, which happens to a particle with weight > 2 while p is below
 * ntarget (:63-72) -- so later particles of the cell see the grown count; a cell
 * above nmax is merged and its p stays. */
void ref_resample_counts(const ref_pop *p, const int64_t *cell_start, int64_t ncells, int32_t ntarget,
                         int32_t nmax, int64_t *pcell, double *wcell)
{
    for (int64_t c = 0; c < ncells; ++c) {
        const int64_t a = cell_start[c], b = cell_start[c + 1];
        int64_t pc = b - a;
        double wtotal = 0.0;
        for (int64_t i = a; i < b; ++i) wtotal += p->w[i];
        if (pc <= nmax)
            for (int64_t i = a; i < b; ++i)
                if (pc < ntarget && p->w[i] > 2.0) pc += 1;
        pcell[c] = pc;
        wcell[c] = wtotal;
    }
}

int32_t ref_resample_cells(ref_pop *p, ref_mesh *m, int32_t ntarget, int32_t nmax,
                           uint64_t seed, uint32_t instance, uint32_t step, ref_pop *scr,
                           ref_resample_stats *st)
{
    memset(st, 0, sizeof(*st));
    st->n_in = p->n;
    int64_t ncells = (int64_t)m->n[0] * m->n[1] * m->n[2];
    int64_t *cs = m->cell_start;
    int64_t out = 0;
    int64_t *newstart = (int64_t *)malloc(sizeof(int64_t) * (size_t)(ncells + 2));
    /* created particles of one cell, staged before the survivors are known */
    int64_t stage_cap = 1024;
    int64_t *stage_src = (int64_t *)malloc(sizeof(int64_t) * (size_t)stage_cap);
    double *stage_w = (double *)malloc(sizeof(double) * (size_t)stage_cap);
    for (int64_t c = 0; c < ncells; ++c) {
        newstart[c] = out;
        int64_t a = cs[c], b = cs[c + 1];
        int64_t pc = b - a;
        if (pc == 0) continue;
        mcs_stream rs;
        mcs_stream_init(&rs, seed, instance, MCS_PHASE_CELL_RESAMPLE, (uint64_t)c, step);
        double wtotal = 0.0, xx = 0.0, wcum = 0.0;
        int64_t cnt = 0, pk = 0, nstage = 0;
        for (int64_t i = a; i < b; ++i) {
            cnt += 1;
            wtotal += p->w[i];
            if (cnt == nmax) xx = mcs_samplemin(&rs, (double)ntarget);
        }
        int merged = pc > nmax;
        if (merged) st->cells_merged += 1;
        for (int64_t i = a; i < b; ++i) {
            double wi = p->w[i];
            if (pc > nmax) {
                while ((wtotal * xx < wcum + wi) && pk < ntarget) {
                    pk += 1;
                    if (nstage == stage_cap) {
                        stage_cap *= 2;
                        stage_src = (int64_t *)realloc(stage_src, sizeof(int64_t) * (size_t)stage_cap);
                        stage_w = (double *)realloc(stage_w, sizeof(double) * (size_t)stage_cap);
                    }
                    stage_src[nstage] = i;
                    stage_w[nstage] = wtotal / (double)ntarget;
                    nstage++;
                    st->clones += 1;
                    if (pk < ntarget)
                        xx = xx + mcs_samplemin(&rs, (double)(ntarget - pk)) * (1.0 - xx);
                }
                wcum = wcum + wi;
                st->removed += 1;
            } else if (pc < ntarget && wi > 2.0) {
                double wnew = rint(wi / 2.0);
                if (nstage == stage_cap) {
                    stage_cap *= 2;
                    stage_src = (int64_t *)realloc(stage_src, sizeof(int64_t) * (size_t)stage_cap);
                    stage_w = (double *)realloc(stage_w, sizeof(double) * (size_t)stage_cap);
                }
                stage_src[nstage] = i;
                stage_w[nstage] = wnew;
                nstage++;
                st->splits += 1;
                p->w[i] = wi - wnew;
                pc += 1;
            }
        }
        if (out + (merged ? 0 : (b - a)) + nstage > scr->cap) {
            free(newstart);
            free(stage_src);
            free(stage_w);
            return FLAG_CAPACITY;
        }
        if (!merged) {
            for (int64_t i = a; i < b; ++i) copy_slot(scr, out++, p, i);
        }
        for (int64_t q = 0; q < nstage; ++q) {
            copy_slot(scr, out, p, stage_src[q]);
            scr->w[out] = stage_w[q];
            scr->id[out] = p->next_id;
            p->next_id += 1;
            out++;
        }
    }
    newstart[ncells] = out;
    for (int64_t i = cs[ncells]; i < cs[ncells + 1]; ++i) {
        if (out >= scr->cap) {
            free(newstart);
            free(stage_src);
            free(stage_w);
            return FLAG_CAPACITY;
        }
        copy_slot(scr, out++, p, i);
    }
    newstart[ncells + 1] = out;
    if (out > p->cap) {
        free(newstart);
        free(stage_src);
        free(stage_w);
        return FLAG_CAPACITY;
    }
    for (int64_t i = 0; i < out; ++i) copy_slot(p, i, scr, i);
    p->n = out;
    memcpy(cs, newstart, sizeof(int64_t) * (size_t)(ncells + 2));
    st->n_out = out;
    free(newstart);
    free(stage_src);
    free(stage_w);
    return 0;
}

/* ------------------------------------------------------------- multigrid */

static inline int64_t gsz(const int32_t n[3])
{
    return (int64_t)(n[0] + 2) * (n[1] + 2) * (n[2] + 2);
}

/* apply! with setbc! (Multigrid/src/boundaries.jl:110-133): faces in the order
 * x-, x+, y-, y+, z-, z+; each ghost plane (full extent, ghosts included) is
 * coef * (adjacent valid plane), coef = -1 Dirichlet, +1 Neumann. */
void ref_mg_apply_bc(const int32_t n[3], const int32_t bc[6], double *a)
{
    int64_t sx = 1, sy = n[0] + 2, sz = (int64_t)(n[0] + 2) * (n[1] + 2);
    int64_t ext[3] = {n[0] + 2, n[1] + 2, n[2] + 2};
    int64_t str[3] = {sx, sy, sz};
    for (int face = 0; face < 6; ++face) {
        int d = face / 2;
        int hi = face & 1;
        double coef = bc[face] == 0 ? -1.0 : 1.0;
        int64_t g = hi ? n[d] + 1 : 0;
        int64_t vld = hi ? n[d] : 1;
        int d1 = (d + 1) % 3, d2 = (d + 2) % 3;
        for (int64_t b2 = 0; b2 < ext[d2]; ++b2)
            for (int64_t b1 = 0; b1 < ext[d1]; ++b1) {
                int64_t base = b1 * str[d1] + b2 * str[d2];
                a[base + g * str[d]] = coef * a[base + vld * str[d]];
            }
    }
}

/* gauss_seidel! with redblack3 (Multigrid.jl:168-178, redblack.jl:49-61):
 * colour 0 = cells with even (i+j+k) in interior coordinates, then colour 1;
 * u += 1.0*(L u + b)/6 with L u = -6u + u[+x] + u[+y] + u[+z] + u[-x] + u[-y] + u[-z]. */
void ref_mg_gauss_seidel(const int32_t n[3], double *u, const double *b)
{
    int64_t sy = n[0] + 2, sz = (int64_t)(n[0] + 2) * (n[1] + 2);
    for (int par = 0; par < 2; ++par)
#pragma omp parallel for schedule(static) if (n[2] >= 16)
        for (int64_t k = 1; k <= n[2]; ++k)
            for (int64_t j = 1; j <= n[1]; ++j)
                for (int64_t i = 1; i <= n[0]; ++i) {
                    if ((((i - 1) + (j - 1) + (k - 1)) & 1) != par) continue;
                    int64_t c = i + sy * j + sz * k;
                    double l = -6.0 * u[c];
                    l += u[c + 1];
                    l += u[c + sy];
                    l += u[c + sz];
                    l += u[c - 1];
                    l += u[c - sy];
                    l += u[c - sz];
                    u[c] += (1.0 * (l + b[c])) / 6.0;
                }
}

/* residual! (Multigrid.jl:186-193): r = s b + L u on the interior. */
void ref_mg_residual(const int32_t n[3], double *r, const double *u, const double *b, double s)
{
    int64_t sy = n[0] + 2, sz = (int64_t)(n[0] + 2) * (n[1] + 2);
#pragma omp parallel for schedule(static) if (n[2] >= 16)
    for (int64_t k = 1; k <= n[2]; ++k)
        for (int64_t j = 1; j <= n[1]; ++j)
            for (int64_t i = 1; i <= n[0]; ++i) {
                int64_t c = i + sy * j + sz * k;
                double l = -6.0 * u[c];
                l += u[c + 1];
                l += u[c + sy];
                l += u[c + sz];
                l += u[c - 1];
                l += u[c - sy];
                l += u[c - sz];
                r[c] = s * b[c] + l;
            }
}

/* restrict! (Multigrid.jl:213-226): rh = 4 * (sum of the 8 children)/8, children
 * visited with the x offset slowest (cubestencil order, Multigrid.jl:133-144). */
void ref_mg_restrict(const int32_t nf[3], double *rh, const double *r)
{
    int32_t nc[3] = {nf[0] / 2, nf[1] / 2, nf[2] / 2};
    int64_t fy = nf[0] + 2, fz = (int64_t)(nf[0] + 2) * (nf[1] + 2);
    int64_t cy = nc[0] + 2, cz = (int64_t)(nc[0] + 2) * (nc[1] + 2);
#pragma omp parallel for schedule(static) if (nc[2] >= 16)
    for (int64_t k = 1; k <= nc[2]; ++k)
        for (int64_t j = 1; j <= nc[1]; ++j)
            for (int64_t i = 1; i <= nc[0]; ++i) {
                int64_t f = (2 * i - 1) + fy * (2 * j - 1) + fz * (2 * k - 1);
                double s = 0.0;
                for (int dx = 0; dx < 2; ++dx)
                    for (int dy = 0; dy < 2; ++dy)
                        for (int dz = 0; dz < 2; ++dz) s += r[f + dx + fy * dy + fz * dz];
                rh[i + cy * j + cz * k] = (4.0 * s) / 8.0;
            }
}

/* interpolate! (Multigrid.jl:232-258): each fine child of a coarse cell gets
 * sum_sh w(sh) * rh[coarse + delta.*sh], delta = 2s-1, w = prod(3-2 sh_i)/64,
 * sh in cubestencil order; added to (add=1) or written into r. */
void ref_mg_interpolate(const int32_t nf[3], double *r, const double *rh, int32_t add)
{
    int32_t nc[3] = {nf[0] / 2, nf[1] / 2, nf[2] / 2};
    int64_t fy = nf[0] + 2, fz = (int64_t)(nf[0] + 2) * (nf[1] + 2);
    int64_t cy = nc[0] + 2, cz = (int64_t)(nc[0] + 2) * (nc[1] + 2);
#pragma omp parallel for schedule(static) if (nc[2] >= 16)
    for (int64_t k = 1; k <= nc[2]; ++k)
        for (int64_t j = 1; j <= nc[1]; ++j)
            for (int64_t i = 1; i <= nc[0]; ++i) {
                int64_t cc = i + cy * j + cz * k;
                for (int sx = 0; sx < 2; ++sx)
                    for (int sy = 0; sy < 2; ++sy)
                        for (int sz = 0; sz < 2; ++sz) {
                            int64_t fidx = (2 * i - 1 + sx) + fy * (2 * j - 1 + sy) + fz * (2 * k - 1 + sz);
                            int64_t ddx = 2 * sx - 1, ddy = 2 * sy - 1, ddz = 2 * sz - 1;
                            double v = 0.0;
                            for (int hx = 0; hx < 2; ++hx)
                                for (int hy = 0; hy < 2; ++hy)
                                    for (int hz = 0; hz < 2; ++hz) {
                                        double wgt = (double)((3 - 2 * hx) * (3 - 2 * hy) * (3 - 2 * hz)) / 64.0;
                                        v += wgt * rh[cc + ddx * hx + cy * ddy * hy + cz * ddz * hz];
                                    }
                            if (add)
                                r[fidx] += v;
                            else
                                r[fidx] = v;
                        }
            }
}

/* buildmat (Multigrid.jl:261-295): the 7-point operator on the coarsest interior
 * with the ghost rule folded in (a missing neighbour adds coef to the diagonal),
 * inverted by Gauss-Jordan elimination with partial pivoting (the reference
 * factorises with UMFPACK; any exact solve restates it up to rounding). */
int32_t ref_mg_coarse_inverse(const int32_t n[3], const int32_t bc[6], double *ainv)
{
    int64_t N = (int64_t)n[0] * n[1] * n[2];
    double *a = (double *)calloc((size_t)(N * N), sizeof(double));
    for (int64_t k = 0; k < n[2]; ++k)
        for (int64_t j = 0; j < n[1]; ++j)
            for (int64_t i = 0; i < n[0]; ++i) {
                int64_t row = i + n[0] * (j + n[1] * k);
                int64_t c[3] = {i, j, k};
                a[row * N + row] -= 6.0;
                /* lplstencil order: +x, +y, +z, -x, -y, -z */
                for (int q = 0; q < 6; ++q) {
                    int d = q % 3;
                    int sgn = q < 3 ? 1 : -1;
                    int64_t cc[3] = {c[0], c[1], c[2]};
                    cc[d] += sgn;
                    if (cc[d] < 0 || cc[d] >= n[d]) {
                        int face = 2 * d + (sgn > 0 ? 1 : 0);
                        a[row * N + row] += bc[face] == 0 ? -1.0 : 1.0;
                    } else {
                        int64_t col = cc[0] + n[0] * (cc[1] + n[1] * cc[2]);
                        a[row * N + col] += 1.0;
                    }
                }
            }
    for (int64_t r = 0; r < N; ++r)
        for (int64_t c = 0; c < N; ++c) ainv[r * N + c] = r == c ? 1.0 : 0.0;
    for (int64_t col = 0; col < N; ++col) {
        int64_t piv = col;
        double best = fabs(a[col * N + col]);
        for (int64_t r = col + 1; r < N; ++r)
            if (fabs(a[r * N + col]) > best) {
                best = fabs(a[r * N + col]);
                piv = r;
            }
        if (best == 0.0) {
            free(a);
            return 1;
        }
        if (piv != col) {
            for (int64_t c = 0; c < N; ++c) {
                double t = a[col * N + c];
                a[col * N + c] = a[piv * N + c];
                a[piv * N + c] = t;
                t = ainv[col * N + c];
                ainv[col * N + c] = ainv[piv * N + c];
                ainv[piv * N + c] = t;
            }
        }
        double d = a[col * N + col];
        for (int64_t c = 0; c < N; ++c) {
            a[col * N + c] = a[col * N + c] / d;
            ainv[col * N + c] = ainv[col * N + c] / d;
        }
        for (int64_t r = 0; r < N; ++r) {
            if (r == col) continue;
            double fct = a[r * N + col];
            if (fct == 0.0) continue;
            for (int64_t c = 0; c < N; ++c) {
                a[r * N + c] = a[r * N + c] - fct * a[col * N + c];
                ainv[r * N + c] = ainv[r * N + c] - fct * ainv[col * N + c];
            }
        }
    }
    free(a);
    return 0;
}

typedef struct {
    int32_t n[8][3];
    double *res[8], *res1[8], *sol[8];
    double *ainv;
    int64_t ncoarse;
    double tol_abs;
} ref_ws;

/* mgv! (Multigrid.jl:334-377) */
static void ref_mgv(const ref_mg *cfg, ref_ws *ws, double *x, const double *b, int level)
{
    const int32_t *n = ws->n[level];
    if (level == cfg->levels) {
        for (int it = 0; it < 50; ++it) {
            ref_mg_apply_bc(n, cfg->bc, x);
            ref_mg_gauss_seidel(n, x, b);
        }
        return;
    }
    for (int it = 0; it < cfg->smooth1; ++it) {
        ref_mg_apply_bc(n, cfg->bc, x);
        ref_mg_gauss_seidel(n, x, b);
    }
    double *r = ws->res1[level];
    ref_mg_apply_bc(n, cfg->bc, x);
    ref_mg_residual(n, r, x, b, 1.0);
    double *rh = ws->res[level + 1];
    ref_mg_apply_bc(n, cfg->bc, r);
    ref_mg_restrict(n, rh, r);
    double *xh = ws->sol[level + 1];
    memset(xh, 0, sizeof(double) * (size_t)gsz(ws->n[level + 1]));
    ref_mgv(cfg, ws, xh, rh, level + 1);
    ref_mg_interpolate(n, x, xh, 1);
    for (int it = 0; it < cfg->smooth2; ++it) {
        ref_mg_apply_bc(n, cfg->bc, x);
        ref_mg_gauss_seidel(n, x, b);
    }
}

/* fmg! (Multigrid.jl:380-434).  The coarsest correction is ws.mat \ btop, i.e.
 * the operator applied inverse to the restricted residual, as written there. */
static int ref_fmg(const ref_mg *cfg, ref_ws *ws, double *x, const double *b, double *eps_out)
{
    int L = cfg->levels;
    const int32_t *n = ws->n[0];
    ref_mg_apply_bc(n, cfg->bc, x);
    memset(ws->res[0], 0, sizeof(double) * (size_t)gsz(n));
    ref_mg_residual(n, ws->res[0], x, b, cfg->s);
    double ss = 0.0;
    int64_t tot = gsz(n);
    for (int64_t c = 0; c < tot; ++c) ss += ws->res[0][c] * ws->res[0][c];
    double eps = sqrt(ss) / sqrt((double)n[0] * (double)n[1] * (double)n[2]);
    *eps_out = eps;
    if (cfg->s * ws->tol_abs > eps) return 0;
    for (int k = 0; k < L; ++k) ref_mg_restrict(ws->n[k], ws->res[k + 1], ws->res[k]);
    /* coarse direct solve */
    {
        const int32_t *nc = ws->n[L];
        double *z = ws->sol[L];
        int64_t N = ws->ncoarse;
        double *bt = (double *)malloc(sizeof(double) * (size_t)N);
        for (int64_t k = 0; k < nc[2]; ++k)
            for (int64_t j = 0; j < nc[1]; ++j)
                for (int64_t i = 0; i < nc[0]; ++i)
                    bt[i + nc[0] * (j + nc[1] * k)] = ws->res[L][cidx(nc, i + 1, j + 1, k + 1)];
        for (int64_t k = 0; k < nc[2]; ++k)
            for (int64_t j = 0; j < nc[1]; ++j)
                for (int64_t i = 0; i < nc[0]; ++i) {
                    int64_t row = i + nc[0] * (j + nc[1] * k);
                    double acc = 0.0;
                    for (int64_t c = 0; c < N; ++c) acc += ws->ainv[row * N + c] * bt[c];
                    z[cidx(nc, i + 1, j + 1, k + 1)] = acc;
                }
        free(bt);
    }
    double *z = ws->sol[L];
    for (int k = 1; k <= L; ++k) {
        int lev = L - k;
        double *z1 = ws->sol[lev];
        memset(z1, 0, sizeof(double) * (size_t)gsz(ws->n[lev]));
        ref_mg_apply_bc(ws->n[lev + 1], cfg->bc, z);
        ref_mg_interpolate(ws->n[lev], z1, z, 0);
        ref_mgv(cfg, ws, z1, ws->res[lev], lev);
        z = z1;
    }
#pragma omp parallel for schedule(static) if (tot >= 65536)
    for (int64_t c = 0; c < tot; ++c) x[c] += z[c];
    ref_mg_apply_bc(n, cfg->bc, x);
    return 1;
}

/* ----------------------------------------------- multigrid in two dimensions
 * Multigrid.jl with N = 2: lplstencil (+x, +y, -x, -y), redblack2, cubestencil
 * children (x offset slowest), interpolation weights prod(3 - 2 sh) / 16, and the
 * connector (Multigrid.jl:31-39) on every neighbour term of the laplacian and the
 * residual.  The coarse operator of buildmat stays Cartesian (Multigrid.jl:261-295
 * takes no connector). */

static inline int64_t gsz2(const int32_t n[2])
{
    return (int64_t)(n[0] + 2) * (n[1] + 2);
}

/* apply! in 2D: faces x-, x+, y-, y+; each ghost line (full extent, ghosts
 * included) is coef * (adjacent valid line). */
void ref_mg2_apply_bc(const int32_t n[2], const int32_t bc[4], double *a)
{
    int64_t ext[2] = {n[0] + 2, n[1] + 2};
    int64_t str[2] = {1, n[0] + 2};
    for (int face = 0; face < 4; ++face) {
        int d = face / 2, d1 = 1 - d;
        double coef = bc[face] == 0 ? -1.0 : 1.0;
        int64_t g = (face & 1) ? n[d] + 1 : 0;
        int64_t vld = (face & 1) ? n[d] : 1;
        for (int64_t b1 = 0; b1 < ext[d1]; ++b1)
            a[b1 * str[d1] + g * str[d]] = coef * a[b1 * str[d1] + vld * str[d]];
    }
}

/* CylindricalConnector{D} (Multigrid.jl:37-39): the neighbour at offset sgn along
 * dimension d of the cell with interior index i (0-based) along d enters as
 * x * (1.0 + sgn / (2 i + 1)); with the Cartesian connector (and along the other
 * dimension) the factor is 1.0, which leaves x unchanged. */
static inline double conn2(int32_t cyl, int d, int sgn, int64_t i)
{
    if (cyl != d + 1) return 1.0;
    return 1.0 + (double)sgn / (double)(2 * i + 1);
}

static inline double lap2(const int32_t n[2], int32_t cyl, const double *u, int64_t i, int64_t j)
{
    int64_t sy = n[0] + 2, c = (i + 1) + sy * (j + 1);
    double l = -4.0 * u[c];
    l += u[c + 1] * conn2(cyl, 0, 1, i);
    l += u[c + sy] * conn2(cyl, 1, 1, j);
    l += u[c - 1] * conn2(cyl, 0, -1, i);
    l += u[c - sy] * conn2(cyl, 1, -1, j);
    return l;
}

/* gauss_seidel! with redblack2 (redblack.jl:38-45): colour 0 = cells with even
 * i + j in interior coordinates, then colour 1; u += 1.0 (L u + b) / 4. */
void ref_mg2_gauss_seidel(const int32_t n[2], int32_t cyl, double *u, const double *b)
{
    int64_t sy = n[0] + 2;
    for (int par = 0; par < 2; ++par)
        for (int64_t j = 0; j < n[1]; ++j)
            for (int64_t i = 0; i < n[0]; ++i) {
                if (((i + j) & 1) != par) continue;
                int64_t c = (i + 1) + sy * (j + 1);
                double l = lap2(n, cyl, u, i, j);
                u[c] += (1.0 * (l + b[c])) / 4.0;
            }
}

void ref_mg2_residual(const int32_t n[2], int32_t cyl, double *r, const double *u, const double *b, double s)
{
    int64_t sy = n[0] + 2;
    for (int64_t j = 0; j < n[1]; ++j)
        for (int64_t i = 0; i < n[0]; ++i) {
            int64_t c = (i + 1) + sy * (j + 1);
            r[c] = s * b[c] + lap2(n, cyl, u, i, j);
        }
}

/* restrict! in 2D: rh = 4 * (sum of the 4 children) / 4. */
void ref_mg2_restrict(const int32_t nf[2], double *rh, const double *r)
{
    int32_t nc[2] = {nf[0] / 2, nf[1] / 2};
    int64_t fy = nf[0] + 2, cy = nc[0] + 2;
    for (int64_t j = 1; j <= nc[1]; ++j)
        for (int64_t i = 1; i <= nc[0]; ++i) {
            int64_t f = (2 * i - 1) + fy * (2 * j - 1);
            double s = 0.0;
            for (int dx = 0; dx < 2; ++dx)
                for (int dy = 0; dy < 2; ++dy) s += r[f + dx + fy * dy];
            rh[i + cy * j] = (4.0 * s) / 4.0;
        }
}

/* interpolate! in 2D: weights prod(3 - 2 sh) / 16. */
void ref_mg2_interpolate(const int32_t nf[2], double *r, const double *rh, int32_t add)
{
    int32_t nc[2] = {nf[0] / 2, nf[1] / 2};
    int64_t fy = nf[0] + 2, cy = nc[0] + 2;
    for (int64_t j = 1; j <= nc[1]; ++j)
        for (int64_t i = 1; i <= nc[0]; ++i) {
            int64_t cc = i + cy * j;
            for (int sx = 0; sx < 2; ++sx)
                for (int sy = 0; sy < 2; ++sy) {
                    int64_t fidx = (2 * i - 1 + sx) + fy * (2 * j - 1 + sy);
                    int64_t ddx = 2 * sx - 1, ddy = 2 * sy - 1;
                    double v = 0.0;
                    for (int hx = 0; hx < 2; ++hx)
                        for (int hy = 0; hy < 2; ++hy) {
                            double wgt = (double)((3 - 2 * hx) * (3 - 2 * hy)) / 16.0;
                            v += wgt * rh[cc + ddx * hx + cy * ddy * hy];
                        }
                    if (add)
                        r[fidx] += v;
                    else
                        r[fidx] = v;
                }
        }
}

/* buildmat in 2D (Cartesian 5-point operator, ghost rule folded), Gauss-Jordan. */
int32_t ref_mg2_coarse_inverse(const int32_t n[2], const int32_t bc[4], double *ainv)
{
    int64_t N = (int64_t)n[0] * n[1];
    double *a = (double *)calloc((size_t)(N * N), sizeof(double));
    for (int64_t j = 0; j < n[1]; ++j)
        for (int64_t i = 0; i < n[0]; ++i) {
            int64_t row = i + n[0] * j;
            int64_t c[2] = {i, j};
            a[row * N + row] -= 4.0;
            for (int q = 0; q < 4; ++q) {   /* +x, +y, -x, -y */
                int d = q % 2;
                int sgn = q < 2 ? 1 : -1;
                int64_t cc[2] = {c[0], c[1]};
                cc[d] += sgn;
                if (cc[d] < 0 || cc[d] >= n[d]) {
                    int face = 2 * d + (sgn > 0 ? 1 : 0);
                    a[row * N + row] += bc[face] == 0 ? -1.0 : 1.0;
                } else {
                    a[row * N + cc[0] + n[0] * cc[1]] += 1.0;
                }
            }
        }
    for (int64_t r = 0; r < N; ++r)
        for (int64_t c = 0; c < N; ++c) ainv[r * N + c] = r == c ? 1.0 : 0.0;
    for (int64_t col = 0; col < N; ++col) {
        int64_t piv = col;
        double best = fabs(a[col * N + col]);
        for (int64_t r = col + 1; r < N; ++r)
            if (fabs(a[r * N + col]) > best) {
                best = fabs(a[r * N + col]);
                piv = r;
            }
        if (best == 0.0) {
            free(a);
            return 1;
        }
        if (piv != col) {
            for (int64_t c = 0; c < N; ++c) {
                double t = a[col * N + c];
                a[col * N + c] = a[piv * N + c];
                a[piv * N + c] = t;
                t = ainv[col * N + c];
                ainv[col * N + c] = ainv[piv * N + c];
                ainv[piv * N + c] = t;
            }
        }
        double d = a[col * N + col];
        for (int64_t c = 0; c < N; ++c) {
            a[col * N + c] = a[col * N + c] / d;
            ainv[col * N + c] = ainv[col * N + c] / d;
        }
        for (int64_t r = 0; r < N; ++r) {
            if (r == col) continue;
            double fct = a[r * N + col];
            if (fct == 0.0) continue;
            for (int64_t c = 0; c < N; ++c) {
                a[r * N + c] = a[r * N + c] - fct * a[col * N + c];
                ainv[r * N + c] = ainv[r * N + c] - fct * ainv[col * N + c];
            }
        }
    }
    free(a);
    return 0;
}

typedef struct {
    int32_t n[16][2];
    double *res[16], *res1[16], *sol[16];
    double *ainv;
    int64_t ncoarse;
    double tol_abs;
} ref_ws2;

static void ref_mgv2(const ref_mg *cfg, ref_ws2 *ws, double *x, const double *b, int level)
{
    const int32_t *n = ws->n[level];
    if (level == cfg->levels) {
        for (int it = 0; it < 50; ++it) {
            ref_mg2_apply_bc(n, cfg->bc, x);
            ref_mg2_gauss_seidel(n, cfg->cyl, x, b);
        }
        return;
    }
    for (int it = 0; it < cfg->smooth1; ++it) {
        ref_mg2_apply_bc(n, cfg->bc, x);
        ref_mg2_gauss_seidel(n, cfg->cyl, x, b);
    }
    double *r = ws->res1[level];
    ref_mg2_apply_bc(n, cfg->bc, x);
    ref_mg2_residual(n, cfg->cyl, r, x, b, 1.0);
    double *rh = ws->res[level + 1];
    ref_mg2_apply_bc(n, cfg->bc, r);
    ref_mg2_restrict(n, rh, r);
    double *xh = ws->sol[level + 1];
    memset(xh, 0, sizeof(double) * (size_t)gsz2(ws->n[level + 1]));
    ref_mgv2(cfg, ws, xh, rh, level + 1);
    ref_mg2_interpolate(n, x, xh, 1);
    for (int it = 0; it < cfg->smooth2; ++it) {
        ref_mg2_apply_bc(n, cfg->bc, x);
        ref_mg2_gauss_seidel(n, cfg->cyl, x, b);
    }
}

static int ref_fmg2(const ref_mg *cfg, ref_ws2 *ws, double *x, const double *b, double *eps_out)
{
    int L = cfg->levels;
    const int32_t *n = ws->n[0];
    ref_mg2_apply_bc(n, cfg->bc, x);
    memset(ws->res[0], 0, sizeof(double) * (size_t)gsz2(n));
    ref_mg2_residual(n, cfg->cyl, ws->res[0], x, b, cfg->s);
    double ss = 0.0;
    int64_t tot = gsz2(n);
    for (int64_t c = 0; c < tot; ++c) ss += ws->res[0][c] * ws->res[0][c];
    double eps = sqrt(ss) / sqrt((double)n[0] * (double)n[1]);
    *eps_out = eps;
    if (cfg->s * ws->tol_abs > eps) return 0;
    for (int k = 0; k < L; ++k) ref_mg2_restrict(ws->n[k], ws->res[k + 1], ws->res[k]);
    {
        const int32_t *nc = ws->n[L];
        double *z = ws->sol[L];
        int64_t N = ws->ncoarse, cy = nc[0] + 2;
        double *bt = (double *)malloc(sizeof(double) * (size_t)N);
        for (int64_t j = 0; j < nc[1]; ++j)
            for (int64_t i = 0; i < nc[0]; ++i) bt[i + nc[0] * j] = ws->res[L][(i + 1) + cy * (j + 1)];
        for (int64_t j = 0; j < nc[1]; ++j)
            for (int64_t i = 0; i < nc[0]; ++i) {
                int64_t row = i + nc[0] * j;
                double acc = 0.0;
                for (int64_t c = 0; c < N; ++c) acc += ws->ainv[row * N + c] * bt[c];
                z[(i + 1) + cy * (j + 1)] = acc;
            }
        free(bt);
    }
    double *z = ws->sol[L];
    for (int k = 1; k <= L; ++k) {
        int lev = L - k;
        double *z1 = ws->sol[lev];
        memset(z1, 0, sizeof(double) * (size_t)gsz2(ws->n[lev]));
        ref_mg2_apply_bc(ws->n[lev + 1], cfg->bc, z);
        ref_mg2_interpolate(ws->n[lev], z1, z, 0);
        ref_mgv2(cfg, ws, z1, ws->res[lev], lev);
        z = z1;
    }
    for (int64_t c = 0; c < tot; ++c) x[c] += z[c];
    ref_mg2_apply_bc(n, cfg->bc, x);
    return 1;
}

static int32_t ref_mg2_solve(const ref_mg *cfg, double *u, const double *b, int32_t *iters, double *resid)
{
    ref_ws2 ws;
    memset(&ws, 0, sizeof(ws));
    if (cfg->levels < 0 || cfg->levels > 15 || cfg->n[2] != 1 || cfg->cyl < 0 || cfg->cyl > 2) return 1;
    for (int l = 0; l <= cfg->levels; ++l) {
        for (int d = 0; d < 2; ++d) {
            if (cfg->n[d] % (1 << l)) return 1;
            ws.n[l][d] = cfg->n[d] >> l;
        }
        int64_t sz = gsz2(ws.n[l]);
        ws.res[l] = (double *)calloc((size_t)sz, sizeof(double));
        ws.res1[l] = (double *)calloc((size_t)sz, sizeof(double));
        ws.sol[l] = (double *)calloc((size_t)sz, sizeof(double));
    }
    const int32_t *nc = ws.n[cfg->levels];
    ws.ncoarse = (int64_t)nc[0] * nc[1];
    ws.ainv = (double *)malloc(sizeof(double) * (size_t)(ws.ncoarse * ws.ncoarse));
    int32_t rc = ref_mg2_coarse_inverse(nc, cfg->bc, ws.ainv);
    ws.tol_abs = cfg->tolerance;
    if (cfg->tolerance < 0.0) {
        const int32_t *n0 = ws.n[0];
        double ss = 0.0;
        for (int64_t j = 1; j <= n0[1]; ++j)
            for (int64_t i = 1; i <= n0[0]; ++i) {
                double v = b[i + (int64_t)(n0[0] + 2) * j];
                ss += v * v;
            }
        double brms = sqrt(ss) / sqrt((double)n0[0] * (double)n0[1]);
        ws.tol_abs = -cfg->tolerance * brms;
        if (!(ws.tol_abs > 0.0)) ws.tol_abs = 2.2250738585072014e-308;
    }
    int iter = 0;
    int cont = 1;
    double eps = 0.0;
    while (rc == 0 && cont && iter <= cfg->maxiter) {
        cont = ref_fmg2(cfg, &ws, u, b, &eps);
        iter += 1;
    }
    *iters = iter;
    *resid = eps;
    for (int l = 0; l <= cfg->levels; ++l) {
        free(ws.res[l]);
        free(ws.res1[l]);
        free(ws.sol[l]);
    }
    free(ws.ainv);
    if (rc) return 1;
    return iter > cfg->maxiter ? 6 : 0;
}

/* solve (Multigrid.jl:449-467): fmg! until converged or iter > maxiter. */
int32_t ref_mg_solve(const ref_mg *cfg, double *u, const double *b, int32_t *iters,
                     double *resid)
{
    if (cfg->ndim == 2) return ref_mg2_solve(cfg, u, b, iters, resid);
    ref_ws ws;
    memset(&ws, 0, sizeof(ws));
    if (cfg->levels < 0 || cfg->levels > 7) return 1;
    for (int l = 0; l <= cfg->levels; ++l) {
        for (int d = 0; d < 3; ++d) {
            if (cfg->n[d] % (1 << l)) return 1;
            ws.n[l][d] = cfg->n[d] >> l;
        }
        int64_t sz = gsz(ws.n[l]);
        ws.res[l] = (double *)calloc((size_t)sz, sizeof(double));
        ws.res1[l] = (double *)calloc((size_t)sz, sizeof(double));
        ws.sol[l] = (double *)calloc((size_t)sz, sizeof(double));
    }
    const int32_t *nc = ws.n[cfg->levels];
    ws.ncoarse = (int64_t)nc[0] * nc[1] * nc[2];
    ws.ainv = (double *)malloc(sizeof(double) * (size_t)(ws.ncoarse * ws.ncoarse));
    int32_t rc = ref_mg_coarse_inverse(nc, cfg->bc, ws.ainv);
    ws.tol_abs = cfg->tolerance;
    if (cfg->tolerance < 0.0) {
        const int32_t *n0 = ws.n[0];
        double ss = 0.0;
        for (int64_t k = 1; k <= n0[2]; ++k)
            for (int64_t j = 1; j <= n0[1]; ++j)
                for (int64_t i = 1; i <= n0[0]; ++i) {
                    double v = b[cidx(n0, i, j, k)];
                    ss += v * v;
                }
        double brms = sqrt(ss) / sqrt((double)n0[0] * (double)n0[1] * (double)n0[2]);
        ws.tol_abs = -cfg->tolerance * brms;
        if (!(ws.tol_abs > 0.0)) ws.tol_abs = 2.2250738585072014e-308;
    }
    int iter = 0;
    int cont = 1;
    double eps = 0.0;
    while (rc == 0 && cont && iter <= cfg->maxiter) {
        cont = ref_fmg(cfg, &ws, u, b, &eps);
        iter += 1;
    }
    *iters = iter;
    *resid = eps;
    for (int l = 0; l <= cfg->levels; ++l) {
        free(ws.res[l]);
        free(ws.res1[l]);
        free(ws.sol[l]);
    }
    free(ws.ainv);
    if (rc) return 1;
    return iter > cfg->maxiter ? 6 : 0;
}
