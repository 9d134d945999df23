// mcs_device.h -- device-side building blocks shared by the libmcs kernels.
#pragma once

#include "mcs_internal.h"

// Loads through the global address space (table pointers that arrive inside
// structs read from memory would otherwise be loaded with flat instructions).
template <typename T>
__device__ __forceinline__ T gld(const T *p)
{
    return *(const __attribute__((address_space(1))) T *)p;
}

struct TableView {
    const double *energy;
    const double *rate;
    const int32_t *kind;
    const double *param;
    const double *photo;   // 4 per row: log numin, log numax, photon_multiplier, photon_weight
    int32_t nproc, ne, egrid;
    double step, inv_dlog, maxrate;
};

// The axisymmetric grid as FieldInterp and track1 read it.
struct CylView {
    int32_t M, N;
    double R, L;
    double dr, dz, drf, dzf;
    double inv_dr, inv_dz;   // 1/dr, 1/dz: normdivrem's quotients
    const double *er, *ez;   // (M+3)(N+3)
    double *qfixed;          // (M+2)(N+2)
    const double *dv;        // M
};

struct MeshView {
    int32_t n[3];
    double h, inv_h;
    const double *ex, *ey, *ez;
    double *qfixed;
    CylView C;
};

static inline TableView table_view(const mcs_table *t)
{
    TableView v;
    v.energy = t->energy;
    v.rate = t->rate;
    v.kind = t->kind;
    v.param = t->param;
    v.photo = t->photo;
    v.nproc = t->nproc;
    v.ne = t->ne;
    v.egrid = t->egrid;
    v.step = t->step;
    v.inv_dlog = t->inv_dlog;
    v.maxrate = t->maxrate;
    return v;
}

static inline MeshView mesh_view(const mcs_mesh *m)
{
    MeshView v;
    for (int d = 0; d < 3; ++d) v.n[d] = m ? m->n[d] : 0;
    v.h = m ? m->h : 0.0;
    v.inv_h = m ? 1.0 / m->h : 0.0;
    v.ex = m ? m->arr[MCS_ARR_EX] : nullptr;
    v.ey = m ? m->arr[MCS_ARR_EY] : nullptr;
    v.ez = m ? m->arr[MCS_ARR_EZ] : nullptr;
    v.qfixed = m ? m->arr[MCS_ARR_QFIXED] : nullptr;
    memset(&v.C, 0, sizeof(v.C));
    return v;
}

static inline MeshView cyl_view(const mcs_cgrid *g)
{
    MeshView v;
    memset(&v, 0, sizeof(v));
    v.C.M = g->M;
    v.C.N = g->N;
    v.C.R = g->R;
    v.C.L = g->L;
    v.C.dr = g->dr;
    v.C.dz = g->dz;
    v.C.drf = g->drf;
    v.C.dzf = g->dzf;
    v.C.inv_dr = 1.0 / g->dr;
    v.C.inv_dz = 1.0 / g->dz;
    v.C.er = g->arr[MCS_CARR_ER];
    v.C.ez = g->arr[MCS_CARR_EZ];
    v.C.qfixed = g->arr[MCS_CARR_QFIXED];
    v.C.dv = g->dv;
    return v;
}

// fld(x, y) for y > 0: the largest integer q with q y <= x, exactly (the quotient
// is rounded, so floor(x / y) can be one too high; x - q y has the sign of its
// fused evaluation).
__device__ __forceinline__ double fld_dev(double x, double y)
{
    double q = floor(x / y);
    if (fma(-q, y, x) < 0.0) q -= 1.0;
    else if (fma(-(q + 1.0), y, x) >= 0.0) q += 1.0;
    return q;
}

// cellindex (MCStreamer/src/grid.jl:84-90) with the bounds tests of density! and
// track1 (the ghosted array: i in 0..M+1, j in 0..N+1; rc: i in 1..M): the flat
// index i + (M+2) j of the cell, or -1.
__device__ __forceinline__ int64_t cyl_cell_dev(const CylView &C, const double x[3], int64_t &i)
{
    const double rho = sqrt(x[0] * x[0] + x[1] * x[1]);
    const double fi = fld_dev(rho - 0.0, C.drf), fj = fld_dev(x[2] - 0.0, C.dzf);
    if (!(fi >= 0.0 && fi < (double)C.M && fj >= -1.0 && fj <= (double)C.N)) return -1;
    i = (int64_t)fi + 1;
    return i + (int64_t)(C.M + 2) * ((int64_t)fj + 1);
}

// FieldInterp (MCStreamer/src/gridfields.jl:115-190) on the axisymmetric grid:
// normdivrem and staggered (:94-109) per axis, bilinear er (faces in r, staggered
// in z) and ez (staggered in r, faces in z), er rotated by x/r and y/r.
// normdivrem is @fastmath, which licenses x * (1/y) for its x / y (the 3D gather
// reads it the same way); the rotation's divisions by r are exact divisions.
__device__ __forceinline__ void efield_cyl_dev(const CylView &C, const double x[3], double e[3],
                                               int32_t &flags)
{
    const double r = sqrt(x[0] * x[0] + x[1] * x[1]);
    if (!(0.0 < r && r < C.R && 0.0 < x[2] && x[2] < C.L)) {
        e[0] = 0.0;
        e[1] = 0.0;
        e[2] = 0.0;
        return;
    }
    const double qr = r * C.inv_dr, qz = x[2] * C.inv_dz;
    // trunc; qr < R / dr and qz < L / dz keep these within M and N (the min only
    // guards the arrays)
    const int32_t i = min((int32_t)qr, C.M), j = min((int32_t)qz, C.N);
    const double dr = qr - (double)i, dz = qz - (double)j;
    int32_t i1, j1;
    double dr1, dz1;
    if (dr < 0.5) {
        i1 = i;
        dr1 = dr + 0.5;
    } else {
        i1 = i + 1;
        dr1 = dr - 0.5;
    }
    if (dz < 0.5) {
        j1 = j;
        dz1 = dz + 0.5;
    } else {
        j1 = j + 1;
        dz1 = dz - 0.5;
    }
    const int64_t S = C.M + 3;
    const double *er = C.er + (i + 1) + S * j1;   // er[i + 1, j1] (1-based i)
    const double *ez = C.ez + i1 + S * (j + 1);   // ez[i1, j + 1]
    const double er00 = gld(er), er10 = gld(er + 1), er01 = gld(er + S), er11 = gld(er + S + 1);
    const double ez00 = gld(ez), ez10 = gld(ez + 1), ez01 = gld(ez + S), ez11 = gld(ez + S + 1);
    const double er1 = ((er00 * (1.0 - dr) * (1.0 - dz1) + er10 * dr * (1.0 - dz1)) + er01 * (1.0 - dr) * dz1) +
                       er11 * dr * dz1;
    const double ez1 = ((ez00 * (1.0 - dr1) * (1.0 - dz) + ez10 * dr1 * (1.0 - dz)) + ez01 * (1.0 - dr1) * dz) +
                       ez11 * dr1 * dz;
    e[0] = er1 * x[0] / r;
    e[1] = er1 * x[1] / r;
    e[2] = ez1;
    const double e2 = (e[0] * e[0] + e[1] * e[1]) + e[2] * e[2];
    if (e2 > MCS_EFIELD_LIMIT2) flags |= MCS_FLAG_FIELD;
}

// energy(p) = 0.5*m*(v1^2 + v2^2 + v3^2)  (JuMC/src/electron.jl:27)
__device__ __forceinline__ double energy_dev(double v0, double v1, double v2)
{
    return MCS_HALF_ME * ((v0 * v0 + v1 * v1) + v2 * v2);
}

// indweight (JuMC/src/util.jl:23-30) on a uniform or log grid; see the header of
// mcs_table_create for the grid definitions.
__device__ __forceinline__ void indweight_dev(const double *energy, int32_t ne, int32_t egrid,
                                              double step, double inv_dlog, double eng,
                                              int64_t &k, double &w, int32_t &flags)
{
    if (!(eng == eng)) {   // a NaN velocity: the reference's floor(Int, NaN) throws
        k = 0;
        w = 0.0;
        flags |= MCS_FLAG_NAN;
        return;
    }
    if (eng >= gld(energy + (ne - 1))) {
        k = ne - 2;
        w = 0.0;
        flags |= MCS_FLAG_TABLE_RANGE;
        return;
    }
    if (egrid == MCS_EGRID_UNIFORM) {
        double kf = floor((eng - gld(energy)) / step);
        if (kf < 0.0) kf = 0.0;
        int64_t kk = (int64_t)kf;
        if (kk > ne - 2) kk = ne - 2;
        double ww = (gld(energy + kk + 1) - eng) / step;
        k = kk;
        w = fmin(fmax(ww, 0.0), 1.0);
    } else {
        if (eng <= gld(energy)) {
            k = 0;
            w = 1.0;
            return;
        }
        double kf = floor(mcs_log(eng / gld(energy)) * inv_dlog);
        if (kf < 0.0) kf = 0.0;
        if (kf > (double)(ne - 2)) kf = (double)(ne - 2);
        int64_t kk = (int64_t)kf;
        while (kk > 0 && gld(energy + kk) > eng) kk -= 1;
        while (kk < ne - 2 && gld(energy + kk + 1) <= eng) kk += 1;
        const double e1 = gld(energy + kk + 1), e0 = gld(energy + kk);
        double ww = (e1 - eng) / (e1 - e0);
        k = kk;
        w = fmin(fmax(ww, 0.0), 1.0);
    }
}

// normdivrem + staggered (MCStreamer/src/gridfields.jl:94-109) for one axis;
// normdivrem is @fastmath there, x * (1/h) stands for x / h.  Indices are 32-bit:
// every mesh array has fewer than 2^31 elements (checked at creation), and xc is
// inside (0, n h) here.
__device__ __forceinline__ void axis_index_dev(double xc, double inv_h, int32_t n, uint32_t &m,
                                               double &d, uint32_t &m1, double &d1)
{
    double r = xc * inv_h;
    int32_t mi = (int32_t)trunc(r);
    if (mi > n - 1) mi = n - 1;
    double di = r - (double)mi;
    m = (uint32_t)mi;
    d = di;
    if (di < 0.5) {
        m1 = (uint32_t)mi;
        d1 = di + 0.5;
    } else {
        m1 = (uint32_t)mi + 1u;
        d1 = di - 0.5;
    }
}

// (offsets from one base pointer, so that the +1 neighbours are load offsets)
__device__ __forceinline__ double trilerp_dev(const double *__restrict__ a, int64_t sx,
                                              int64_t sxy, int64_t base, double dx, double dy,
                                              double dz)
{
    const double *__restrict__ p = a + base;
    double c00 = p[0] * (1.0 - dx) + p[1] * dx;
    double c10 = p[sx] * (1.0 - dx) + p[sx + 1] * dx;
    double c01 = p[sxy] * (1.0 - dx) + p[sxy + 1] * dx;
    double c11 = p[sxy + sx] * (1.0 - dx) + p[sxy + sx + 1] * dx;
    double c0 = c00 * (1.0 - dy) + c10 * dy;
    double c1 = c01 * (1.0 - dy) + c11 * dy;
    return c0 * (1.0 - dz) + c1 * dz;
}

// FieldInterp (MCStreamer/src/gridfields.jl:115-190) in 3D Cartesian form: zero
// outside the open box, staggered trilinear per component, |E| > 1e8 flagged.
__device__ __forceinline__ void efield_mesh_dev(const MeshView &M, const double x[3],
                                                double e[3], int32_t &flags)
{
    double h = M.h;
    if (!(0.0 < x[0] && x[0] < (double)M.n[0] * h && 0.0 < x[1] && x[1] < (double)M.n[1] * h &&
          0.0 < x[2] && x[2] < (double)M.n[2] * h)) {
        e[0] = 0.0;
        e[1] = 0.0;
        e[2] = 0.0;
        return;
    }
    uint32_t mx, my, mz, m1x, m1y, m1z;
    double dx, dy, dz, d1x, d1y, d1z;
    axis_index_dev(x[0], M.inv_h, M.n[0], mx, dx, m1x, d1x);
    axis_index_dev(x[1], M.inv_h, M.n[1], my, dy, m1y, d1y);
    axis_index_dev(x[2], M.inv_h, M.n[2], mz, dz, m1z, d1z);
    const uint32_t nx = (uint32_t)M.n[0], ny = (uint32_t)M.n[1];
    {
        const uint32_t sx = nx + 1, sxy = (nx + 1) * (ny + 2);
        e[0] = trilerp_dev(M.ex, sx, sxy, (int64_t)(mx + sx * m1y + sxy * m1z), dx, d1y, d1z);
    }
    {
        const uint32_t sx = nx + 2, sxy = (nx + 2) * (ny + 1);
        e[1] = trilerp_dev(M.ey, sx, sxy, (int64_t)(m1x + sx * my + sxy * m1z), d1x, dy, d1z);
    }
    {
        const uint32_t sx = nx + 2, sxy = (nx + 2) * (ny + 2);
        e[2] = trilerp_dev(M.ez, sx, sxy, (int64_t)(m1x + sx * m1y + sxy * mz), d1x, d1y, dz);
    }
    double e2 = (e[0] * e[0] + e[1] * e[1]) + e[2] * e[2];
    if (e2 > MCS_EFIELD_LIMIT2) flags |= MCS_FLAG_FIELD;
}

template <typename T>
__device__ __forceinline__ T block_excl_scan(T v, T *total)
{
    __shared__ T wsum[MCS_BLOCK / MCS_WAVE];
    int lane = threadIdx.x & (MCS_WAVE - 1), wid = threadIdx.x / MCS_WAVE;
    T x = v;
#pragma unroll
    for (int o = 1; o < MCS_WAVE; o <<= 1) {
        T y = __shfl_up(x, o, MCS_WAVE);
        if (lane >= o) x += y;
    }
    if (lane == MCS_WAVE - 1) wsum[wid] = x;
    __syncthreads();
    T pre = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < MCS_BLOCK / MCS_WAVE; ++k) {
        if (k < wid) pre += wsum[k];
        tot += wsum[k];
    }
    __syncthreads();
    *total = tot;
    return pre + x - v;
}

__device__ __forceinline__ void copy_particle(const PopBuf &src, int64_t si, const PopBuf &dst,
                                              int64_t di)
{
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        dst.x[d][di] = src.x[d][si];
        dst.v[d][di] = src.v[d][si];
    }
    dst.w[di] = src.w[si];
    dst.s[di] = src.s[si];
    dst.active[di] = src.active[si];
    dst.id[di] = src.id[si];
    if (src.nu) dst.nu[di] = src.nu[si];
}

// cell index of a position (grid.jl:84-90 generalised): -1 outside [0, n h)^3
__device__ __forceinline__ int64_t cell_of(const int32_t n[3], double h, double x0, double x1,
                                           double x2)
{
    double f0 = floor(x0 / h), f1 = floor(x1 / h), f2 = floor(x2 / h);
    if (!(f0 >= 0.0 && f0 < (double)n[0] && f1 >= 0.0 && f1 < (double)n[1] && f2 >= 0.0 &&
          f2 < (double)n[2]))
        return -1;
    return (int64_t)f0 + (int64_t)n[0] * ((int64_t)f1 + (int64_t)n[1] * (int64_t)f2);
}
