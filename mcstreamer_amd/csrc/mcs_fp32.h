// mcs_fp32.h -- the split-precision particle step: positions stored as a packed
// integer cell plus an FP32 offset inside the cell (in cell units), velocities,
// weights and collision clocks in FP32, collision sampling in FP32 from 24-bit
// uniforms.  The physics is JuMC.advance! (mixed_population.jl:24-54) exactly as
// in the FP64 body (mcs_advance.h); only the arithmetic precision differs, so the
// results agree with the FP64 path statistically, not bitwise.
#pragma once

#include "mcs_device.h"

// Packed cell coordinates: i in bits 0-9, j in 10-19, k in 20-30.
#define MCS_CX_BITS 10
#define MCS_CY_BITS 10
#define MCS_CZ_BITS 11
#define MCS_CX_MAX (1 << MCS_CX_BITS)
#define MCS_CY_MAX (1 << MCS_CY_BITS)
#define MCS_CZ_MAX (1 << MCS_CZ_BITS)

__host__ __device__ __forceinline__ int32_t pack_cell(int32_t i, int32_t j, int32_t k)
{
    return i | (j << MCS_CX_BITS) | (k << (MCS_CX_BITS + MCS_CY_BITS));
}

__host__ __device__ __forceinline__ void unpack_cell(int32_t c, int32_t &i, int32_t &j, int32_t &k)
{
    i = c & (MCS_CX_MAX - 1);
    j = (c >> MCS_CX_BITS) & (MCS_CY_MAX - 1);
    k = (c >> (MCS_CX_BITS + MCS_CY_BITS)) & (MCS_CZ_MAX - 1);
}

struct TableView32 {
    const float *energy;
    const float *rate;
    const int32_t *kind;
    const float *param;
    int32_t nproc, ne, egrid;
    float e0, emax, inv_e0, inv_step, inv_dlog2, maxrate, inv_maxrate;
    uint64_t kindbits;   // 4 bits per process and bit 63 set (nproc <= 15), else 0: kind[] is read
    const float *band;   // MCS_NBAND + 1 triples (cap, majorant, 1 / majorant), the last unbounded
    float inv_cap0;      // 1 / the first band's cap
};

__device__ __forceinline__ int proc_kind(const TableView32 &T, int j)
{
    return T.kindbits ? (int)((T.kindbits >> (4 * j)) & 15u) : gld(T.kind + j);
}

static inline TableView32 table_view32(const mcs_table *t)
{
    TableView32 v;
    v.energy = t->energy32;
    v.rate = t->rate32;
    v.kind = t->kind;
    v.param = t->param32;
    v.nproc = t->nproc;
    v.ne = t->ne;
    v.egrid = t->egrid;
    v.e0 = (float)t->e0;
    v.emax = (float)t->emax;
    v.inv_e0 = (float)(1.0 / t->e0);
    v.inv_step = (float)(1.0 / t->step);
    v.inv_dlog2 = (float)(t->inv_dlog * 0.6931471805599453);
    v.maxrate = (float)t->maxrate;
    v.inv_maxrate = (float)(1.0 / t->maxrate);
    v.band = t->band32;
    v.inv_cap0 = (float)(1.0 / t->h_band_cap[0]);
    v.kindbits = 0;
    if (t->nproc <= 15) {
        for (int j = 0; j < t->nproc; ++j) v.kindbits |= (uint64_t)(t->h_kind[j] & 15) << (4 * j);
        v.kindbits |= (uint64_t)1 << 63;   // marks the bits valid (kind 0 rows give no bits)
    }
    return v;
}

// Face fields in FP32 on the mesh, cell size h (m).
struct MeshView32 {
    int32_t n[3];
    float h, inv_h;
    const float *ex, *ey, *ez;
    double *qfixed;
    double inv_dv;
};

#define MCS_HALF_ME_F 4.55469178e-31f
#define MCS_QM_F (-175882002360.2124f)
#define MCS_ELECTRON_MASS_F 9.10938356e-31f

__device__ __forceinline__ float energy32(const float v[3])
{
    return MCS_HALF_ME_F * ((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
}

// indweight (JuMC/src/util.jl:23-30) in FP32: node k with energy[k] <= eng <
// energy[k+1] (the float nodes), weight of node k.  The node index comes from
// the grid formula (hardware log2 on the log grid), then is corrected by one
// node against the stored nodes; the general search runs only if that fails.
// Table loads: from global memory (TL = 0), or (TL = 1) from the copy a kernel
// staged in LDS, whose pointers the TableView32 then holds.
template <int TL, typename V>
__device__ __forceinline__ V tld(const V *p)
{
    if constexpr (TL != 0)
        return *(const __attribute__((address_space(3))) V *)p;
    else
        return gld(p);
}

template <int TL = 0>
__device__ __forceinline__ void indweight32(const TableView32 &T, float eng, int32_t &k, float &w,
                                            int32_t &flags)
{
    const float *E = T.energy;
    if (!(eng == eng)) {
        k = 0;
        w = 0.0f;
        flags |= MCS_FLAG_NAN;
        return;
    }
    if (eng >= T.emax) {
        k = T.ne - 2;
        w = 0.0f;
        flags |= MCS_FLAG_TABLE_RANGE;
        return;
    }
    float kf;
    if (T.egrid == MCS_EGRID_UNIFORM) {
        kf = floorf((eng - T.e0) * T.inv_step);
    } else {
        if (eng <= T.e0) {
            k = 0;
            w = 1.0f;
            return;
        }
        kf = floorf(__builtin_amdgcn_logf(eng * T.inv_e0) * T.inv_dlog2);
    }
    int32_t kk = (int32_t)fminf(fmaxf(kf, 0.0f), (float)(T.ne - 2));
    float e0 = tld<TL>(E + kk), e1 = tld<TL>(E + kk + 1);
    if ((e0 > eng) | (e1 <= eng)) {
        kk = e0 > eng ? kk - 1 : kk + 1;
        kk = kk < 0 ? 0 : (kk > T.ne - 2 ? T.ne - 2 : kk);
        while (kk > 0 && tld<TL>(E + kk) > eng) kk -= 1;
        while (kk < T.ne - 2 && tld<TL>(E + kk + 1) <= eng) kk += 1;
        e0 = tld<TL>(E + kk);
        e1 = tld<TL>(E + kk + 1);
    }
    const float ww = (e1 - eng) * __builtin_amdgcn_rcpf(e1 - e0);
    k = kk;
    w = fminf(fmaxf(ww, 0.0f), 1.0f);
}

// nextcoll() = -log(rand()) from a 24-bit uniform; 1 - u >= 2^-24.
__device__ __forceinline__ float nextcoll32(mcs_stream *st)
{
    return -0.6931471805599453f * __builtin_amdgcn_logf(1.0f - mcs_u01f(st));
}

// randsphere() (util.jl:4-12) with phi drawn first; the hardware sine and cosine
// take the angle in turns.
__device__ __forceinline__ void randsphere32(mcs_stream *st, float out[3])
{
    const float uphi = mcs_u01f(st);
    const float sp = __builtin_amdgcn_sinf(uphi), cp = __builtin_amdgcn_cosf(uphi);
    const float u = 2.0f * mcs_u01f(st) - 1.0f;
    const float r = sqrtf(fmaxf(1.0f - u * u, 0.0f));
    out[0] = r * cp;
    out[1] = r * sp;
    out[2] = u;
}

// Staggered trilinear gather of one face array in split coordinates (the FP32
// form of efield_mesh_dev): per axis the integer cell m and fraction d of the
// position, then the half-cell shifted index for the staggered axes.
__device__ __forceinline__ float trilerp32(const float *__restrict__ a, int32_t sx, int32_t sxy,
                                           int32_t base, float dx, float dy, float dz)
{
    float c00 = a[base] * (1.0f - dx) + a[base + 1] * dx;
    float c10 = a[base + sx] * (1.0f - dx) + a[base + sx + 1] * dx;
    float c01 = a[base + sxy] * (1.0f - dx) + a[base + sxy + 1] * dx;
    float c11 = a[base + sxy + sx] * (1.0f - dx) + a[base + sxy + sx + 1] * dx;
    float c0 = c00 * (1.0f - dy) + c10 * dy;
    float c1 = c01 * (1.0f - dy) + c11 * dy;
    return c0 * (1.0f - dz) + c1 * dz;
}

// Returns false (field zero) outside the open box.
__device__ __forceinline__ bool axis32(int32_t c, float off, int32_t n, int32_t &m, float &d,
                                       int32_t &m1, float &d1)
{
    const float fl = floorf(off);
    int32_t mi = c + (int32_t)fl;
    float di = off - fl;
    if (mi < 0 || mi > n - 1 || (mi == 0 && di == 0.0f)) return false;
    m = mi;
    d = di;
    if (di < 0.5f) {
        m1 = mi;
        d1 = di + 0.5f;
    } else {
        m1 = mi + 1;
        d1 = di - 0.5f;
    }
    return true;
}

__device__ __forceinline__ void efield32(const MeshView32 &M, const int32_t c[3], const float off[3],
                                         float e[3], int32_t &flags)
{
    int32_t mx, my, mz, m1x, m1y, m1z;
    float dx, dy, dz, d1x, d1y, d1z;
    if (!axis32(c[0], off[0], M.n[0], mx, dx, m1x, d1x) || !axis32(c[1], off[1], M.n[1], my, dy, m1y, d1y) ||
        !axis32(c[2], off[2], M.n[2], mz, dz, m1z, d1z)) {
        e[0] = 0.0f;
        e[1] = 0.0f;
        e[2] = 0.0f;
        return;
    }
    const int32_t nx = M.n[0], ny = M.n[1];
    {
        int32_t sx = nx + 1, sxy = (nx + 1) * (ny + 2);
        e[0] = trilerp32(M.ex, sx, sxy, mx + sx * m1y + sxy * m1z, dx, d1y, d1z);
    }
    {
        int32_t sx = nx + 2, sxy = (nx + 2) * (ny + 1);
        e[1] = trilerp32(M.ey, sx, sxy, m1x + sx * my + sxy * m1z, d1x, dy, d1z);
    }
    {
        int32_t sx = nx + 2, sxy = (nx + 2) * (ny + 2);
        e[2] = trilerp32(M.ez, sx, sxy, m1x + sx * m1y + sxy * mz, d1x, d1y, dz);
    }
    const float e2 = (e[0] * e[0] + e[1] * e[1]) + e[2] * e[2];
    if (e2 > (float)MCS_EFIELD_LIMIT2) flags |= MCS_FLAG_FIELD;
}

// NGP ion-charge tracker (gridfields.jl:203-213) from split coordinates: the
// cell (flat index into the ghosted array) and the amount val/dV it receives;
// false outside the mesh.
__device__ __forceinline__ bool track_cell32(const MeshView32 &M, const int32_t c[3], const float off[3],
                                             double val, int32_t &cell, double &amount)
{
    int32_t q[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        q[d] = c[d] + (int32_t)floorf(off[d]);
        if (q[d] < 0 || q[d] >= M.n[d]) return false;
    }
    cell = (q[0] + 1) + (M.n[0] + 2) * ((q[1] + 1) + (M.n[1] + 2) * (q[2] + 1));
    amount = val * M.inv_dv;
    return true;
}

// Moves whole cells from the offset into the packed coordinates, clamped to
// [0, nmax) per axis (beyond the clamp the offset carries the distance).
__device__ __forceinline__ void rebase32(int32_t c[3], float off[3], const int32_t nmax[3])
{
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        int32_t nc = c[d] + (int32_t)floorf(off[d]);
        nc = nc < 0 ? 0 : (nc > nmax[d] - 1 ? nmax[d] - 1 : nc);
        off[d] = off[d] - (float)(nc - c[d]);
        c[d] = nc;
    }
}

// One particle through one time step in FP32 (JuMC/src/mixed_population.jl:31-51;
// flights, Yoshida push of electron.jl:51-62 in a mesh field or its exact
// parabola in a uniform one, collisions of collisions.jl:159-192 and
// electron.jl:108-139).  Draws: 24-bit uniforms, one word each, in the order of
// the FP64 body.  `birth(c, off, v, s, active)` receives each ionization child
// at the parent's position, `event(j)` each selected table row, and with TRACK
// `track(c, off, +-w)` each ion-charge deposit of the CollisionTracker.
template <int FIELD_MESH, int TRACK, int TL = 0, typename Birth, typename Event, typename Track>
__device__ __forceinline__ void advance_one32(int32_t c[3], float off[3], float v[3], const float w,
                                              float &s, uint8_t &act, mcs_stream &rs, const float dt,
                                              const TableView32 &T, const MeshView32 &M,
                                              const float *acc_uni, const float inv_h, int32_t &flags,
                                              int &flights, int &deaths, Birth &&birth, Event &&event,
                                              Track &&track)
{
    const float maxrate = T.maxrate;
    float trem = dt;
    const float c1 = (float)MCS_YOSHIDA_C1, c2 = (float)MCS_YOSHIDA_C2;
    const float d1 = (float)MCS_YOSHIDA_D1, d2 = (float)MCS_YOSHIDA_D2;
    while (trem > 0.0f) {
        const float tnext = s * T.inv_maxrate;
        float tt;
        bool collides;
        if (trem > tnext) {
            tt = tnext;
            collides = true;
        } else {
            tt = trem;
            collides = false;
            s = s - tt * maxrate;
        }
        if (FIELD_MESH) {
            float e[3];
            const float k1 = c1 * tt * inv_h, k2 = c2 * tt * inv_h;
            const float f1 = (d1 * tt) * MCS_QM_F, f2 = (d2 * tt) * MCS_QM_F;
#pragma unroll
            for (int d = 0; d < 3; ++d) off[d] = off[d] + k1 * v[d];
            efield32(M, c, off, e, flags);
#pragma unroll
            for (int d = 0; d < 3; ++d) v[d] = v[d] + f1 * e[d];
#pragma unroll
            for (int d = 0; d < 3; ++d) off[d] = off[d] + k2 * v[d];
            efield32(M, c, off, e, flags);
#pragma unroll
            for (int d = 0; d < 3; ++d) v[d] = v[d] + f2 * e[d];
#pragma unroll
            for (int d = 0; d < 3; ++d) off[d] = off[d] + k2 * v[d];
            efield32(M, c, off, e, flags);
#pragma unroll
            for (int d = 0; d < 3; ++d) v[d] = v[d] + f1 * e[d];
#pragma unroll
            for (int d = 0; d < 3; ++d) off[d] = off[d] + k1 * v[d];
        } else {
            // the Yoshida map in a uniform field is the exact parabola
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const float dv = acc_uni[d] * tt;
                off[d] = off[d] + (v[d] + 0.5f * dv) * (tt * inv_h);
                v[d] = v[d] + dv;
            }
        }
        flights += 1;
        if (collides) {
            const float eng = energy32(v);
            int32_t k;
            float wg;
            indweight32<TL>(T, eng, k, wg, flags);
            float xi = mcs_u01f(&rs) * maxrate;
            const float *r0 = T.rate + k * T.nproc;
            const float *r1 = r0 + T.nproc;
            // the selected row first (do_one_collision!), then its outcome
            int sel = -1;
            for (int j = 0; j < T.nproc; ++j) {
                const float nu = wg * tld<TL>(r0 + j) + (1.0f - wg) * tld<TL>(r1 + j);
                if (nu > xi) {
                    sel = j;
                    break;
                }
                xi = xi - nu;
            }
            if (sel >= 0) {
                event(sel);
                const int kind = proc_kind(T, sel);
                if (kind == MCS_PROC_ELASTIC) {
                    const float mr = tld<TL>(T.param + sel);
                    const float cm = mr / (1.0f + mr);
                    const float vc0 = cm * v[0], vc1 = cm * v[1], vc2 = cm * v[2];
                    const float e0 = v[0] - vc0, e1 = v[1] - vc1, e2 = v[2] - vc2;
                    const float nrm = sqrtf((e0 * e0 + e1 * e1) + e2 * e2);
                    float om[3];
                    randsphere32(&rs, om);
                    v[0] = om[0] * nrm + vc0;
                    v[1] = om[1] * nrm + vc1;
                    v[2] = om[2] * nrm + vc2;
                } else if (kind == MCS_PROC_EXCITATION) {
                    const float e1 = fmaxf(eng - tld<TL>(T.param + sel), 0.0f);
                    const float vabs = sqrtf((2.0f * e1) / MCS_ELECTRON_MASS_F);
                    float om[3];
                    randsphere32(&rs, om);
                    v[0] = om[0] * vabs;
                    v[1] = om[1] * vabs;
                    v[2] = om[2] * vabs;
                } else if (kind == MCS_PROC_IONIZATION) {
                    const float e1 = fmaxf(eng - tld<TL>(T.param + sel), 0.0f) * 0.5f;
                    const float vabs = sqrtf((2.0f * e1) / MCS_ELECTRON_MASS_F);
                    float om1[3], om2[3];
                    randsphere32(&rs, om1);
                    randsphere32(&rs, om2);
                    v[0] = om1[0] * vabs;
                    v[1] = om1[1] * vabs;
                    v[2] = om1[2] * vabs;
                    const float sc = nextcoll32(&rs);
                    if (TRACK) track(c, off, (double)w);
                    const float vc[3] = {om2[0] * vabs, om2[1] * vabs, om2[2] * vabs};
                    birth(c, off, vc, sc, act);
                } else if (kind == MCS_PROC_ATTACHMENT) {
                    if (TRACK) track(c, off, -(double)w);
                    if (act) deaths += 1;
                    act = 0;
                }
            }
            s = nextcoll32(&rs);
        }
        trem = trem - tt;
    }
}

// The band of a particle of energy eng under the banded majorant: the first band
// whose cap is at least 1.25 eng (a quarter of headroom, so that few flights run
// into the cap), the unbounded band MCS_NBAND above the last cap or for a NaN.
__device__ __forceinline__ int band_of(const TableView32 &T, float eng)
{
    const float e = 1.25f * eng;
    const float r = e * T.inv_cap0;
    if (!(r > 1.0f)) return r == r ? 0 : MCS_NBAND;
    int b = (int)ceilf(__builtin_amdgcn_logf(r));
    b = b < 0 ? 0 : (b > MCS_NBAND ? MCS_NBAND : b);
    while (b < MCS_NBAND && gld(T.band + 3 * b) < e) b += 1;
    return b;
}

// advance_one32 in a uniform field with the banded majorant (mcs_swarm_batch_majorant):
// each flight's exponential clock runs at the majorant M of the particle's band
// [0, cap]; |v + a t|^2 is convex in t and below the cap at the flight's start, so
// it stays below up to the flight's end when it is below there, and otherwise the
// flight stops where the parabola reaches the cap (the remaining clock s - t M
// carries over to the next band, the exponential being memoryless).  At a
// collision xi = U M selects among the non-null rows in table order (a null event
// when it passes them all), then the outcomes of advance_one32.
template <typename Birth>
__device__ __forceinline__ void advance_one32_banded(int32_t c[3], float off[3], float v[3], const float w,
                                                     float &s, uint8_t &act, mcs_stream &rs, const float dt,
                                                     const TableView32 &T, const float *acc, const float inv_h,
                                                     int32_t &flags, int &flights, int &deaths, Birth &&birth)
{
    const bool field = acc[0] != 0.0f || acc[1] != 0.0f || acc[2] != 0.0f;
    float trem = dt;
    while (trem > 0.0f) {
        const int b = band_of(T, energy32(v));
        const float cap = gld(T.band + 3 * b), M = gld(T.band + 3 * b + 1), invM = gld(T.band + 3 * b + 2);
        const float tnext = s * invM;
        float tt;
        bool collides;
        if (trem > tnext) {
            tt = tnext;
            collides = true;
        } else {
            tt = trem;
            collides = false;
        }
        if (b < MCS_NBAND && field) {
            const float ve[3] = {v[0] + acc[0] * tt, v[1] + acc[1] * tt, v[2] + acc[2] * tt};
            if (energy32(ve) > cap) {
                // positive root of A t^2 + 2 B t + C = 0, |v + a t|^2 = 2 cap / m (C < 0:
                // the energy starts below the cap), in FP64 (A = |a|^2 reaches 1e37);
                // for B > 0 in the form without cancellation, -C / (B + sqrt(B^2 - A C))
                const double a0 = acc[0], a1 = acc[1], a2 = acc[2];
                const double v0 = v[0], v1 = v[1], v2 = v[2];
                const double A = (a0 * a0 + a1 * a1) + a2 * a2;
                const double B = (v0 * a0 + v1 * a1) + v2 * a2;
                const double C = ((v0 * v0 + v1 * v1) + v2 * v2) - (double)cap / (double)MCS_HALF_ME_F;
                const double sq = sqrt(fmax(B * B - A * C, 0.0));
                const double tc = B > 0.0 ? (-C) / (B + sq) : (sq - B) / A;
                tt = fminf((float)tc, tt);
                collides = false;
            }
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const float dv = acc[d] * tt;
            off[d] = off[d] + (v[d] + 0.5f * dv) * (tt * inv_h);
            v[d] = v[d] + dv;
        }
        flights += 1;
        if (!collides) {
            s = fmaxf(s - tt * M, 0.0f);
        } else {
            const float eng = energy32(v);
            int32_t k;
            float wg;
            indweight32(T, eng, k, wg, flags);
            float xi = mcs_u01f(&rs) * M;
            const float *r0 = T.rate + k * T.nproc;
            const float *r1 = r0 + T.nproc;
            int sel = -1, kind = MCS_PROC_NULL;
            for (int j = 0; j < T.nproc; ++j) {
                const int kj = proc_kind(T, j);
                if (kj == MCS_PROC_NULL) continue;
                const float nu = wg * gld(r0 + j) + (1.0f - wg) * gld(r1 + j);
                if (nu > xi) {
                    sel = j;
                    kind = kj;
                    break;
                }
                xi = xi - nu;
            }
            if (sel >= 0) {
                if (kind == MCS_PROC_ELASTIC) {
                    const float mr = gld(T.param + sel);
                    const float cm = mr / (1.0f + mr);
                    const float vc0 = cm * v[0], vc1 = cm * v[1], vc2 = cm * v[2];
                    const float e0 = v[0] - vc0, e1 = v[1] - vc1, e2 = v[2] - vc2;
                    const float nrm = sqrtf((e0 * e0 + e1 * e1) + e2 * e2);
                    float om[3];
                    randsphere32(&rs, om);
                    v[0] = om[0] * nrm + vc0;
                    v[1] = om[1] * nrm + vc1;
                    v[2] = om[2] * nrm + vc2;
                } else if (kind == MCS_PROC_EXCITATION) {
                    const float e1 = fmaxf(eng - gld(T.param + sel), 0.0f);
                    const float vabs = sqrtf((2.0f * e1) / MCS_ELECTRON_MASS_F);
                    float om[3];
                    randsphere32(&rs, om);
                    v[0] = om[0] * vabs;
                    v[1] = om[1] * vabs;
                    v[2] = om[2] * vabs;
                } else if (kind == MCS_PROC_IONIZATION) {
                    const float e1 = fmaxf(eng - gld(T.param + sel), 0.0f) * 0.5f;
                    const float vabs = sqrtf((2.0f * e1) / MCS_ELECTRON_MASS_F);
                    float om1[3], om2[3];
                    randsphere32(&rs, om1);
                    randsphere32(&rs, om2);
                    v[0] = om1[0] * vabs;
                    v[1] = om1[1] * vabs;
                    v[2] = om1[2] * vabs;
                    const float sc = nextcoll32(&rs);
                    const float vc[3] = {om2[0] * vabs, om2[1] * vabs, om2[2] * vabs};
                    birth(c, off, vc, sc, act);
                } else if (kind == MCS_PROC_ATTACHMENT) {
                    if (act) deaths += 1;
                    act = 0;
                }
            }
            s = nextcoll32(&rs);
        }
        trem = trem - tt;
    }
}

__device__ __forceinline__ void copy_rec(const PopBuf &src, int64_t si, const PopBuf &dst, int64_t di)
{
    copy_particle(src, si, dst, di);
}

__device__ __forceinline__ void copy_rec(const SplitBuf &src, int64_t si, const SplitBuf &dst, int64_t di)
{
    dst.cell[di] = src.cell[si];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        dst.off[d][di] = src.off[d][si];
        dst.v[d][di] = src.v[d][si];
    }
    dst.w[di] = src.w[si];
    dst.s[di] = src.s[si];
    dst.active[di] = src.active[si];
    dst.id[di] = src.id[si];
}

// Position (m), velocity and weight of one particle as FP64, for either store;
// h is the cell size of the split coordinates.
__device__ __forceinline__ void load_state(const PopBuf &b, int64_t i, double h, double x[3], double v[3],
                                           double &w)
{
    (void)h;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        x[d] = b.x[d][i];
        v[d] = b.v[d][i];
    }
    w = b.w[i];
}

__device__ __forceinline__ void load_state(const SplitBuf &b, int64_t i, double h, double x[3], double v[3],
                                           double &w)
{
    int32_t c[3];
    unpack_cell(b.cell[i], c[0], c[1], c[2]);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        x[d] = ((double)c[d] + (double)b.off[d][i]) * h;
        v[d] = (double)b.v[d][i];
    }
    w = (double)b.w[i];
}

__device__ __forceinline__ double get_w(const PopBuf &b, int64_t i) { return b.w[i]; }
__device__ __forceinline__ double get_w(const SplitBuf &b, int64_t i) { return (double)b.w[i]; }
__device__ __forceinline__ void set_w(const PopBuf &b, int64_t i, double w) { b.w[i] = w; }
__device__ __forceinline__ void set_w(const SplitBuf &b, int64_t i, double w) { b.w[i] = (float)w; }

// Cell of a particle (floor of the position over h per axis), -1 outside [0, n h).
__device__ __forceinline__ int64_t cell_key(const PopBuf &b, int64_t i, const int32_t n[3], double h)
{
    return cell_of(n, h, b.x[0][i], b.x[1][i], b.x[2][i]);
}

__device__ __forceinline__ int64_t cell_key(const SplitBuf &b, int64_t i, const int32_t n[3], double h)
{
    (void)h;
    int32_t c[3];
    unpack_cell(b.cell[i], c[0], c[1], c[2]);
    int64_t q[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        q[d] = (int64_t)c[d] + (int64_t)floorf(b.off[d][i]);
        if (q[d] < 0 || q[d] >= n[d]) return -1;
    }
    return q[0] + (int64_t)n[0] * (q[1] + (int64_t)n[1] * q[2]);
}

// CIC geometry of a particle: g = x/h - 1/2 per axis, its floor and fraction.
__device__ __forceinline__ void cic_geom(const PopBuf &b, int64_t i, double h, double fl[3], double fr[3])
{
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double gq = b.x[d][i] / h - 0.5;
        fl[d] = floor(gq);
        fr[d] = gq - fl[d];
    }
}

__device__ __forceinline__ void cic_geom(const SplitBuf &b, int64_t i, double h, double fl[3], double fr[3])
{
    (void)h;
    int32_t c[3];
    unpack_cell(b.cell[i], c[0], c[1], c[2]);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const float g = b.off[d][i] - 0.5f;
        const float f = floorf(g);
        fl[d] = (double)c[d] + (double)f;
        fr[d] = (double)(g - f);
    }
}

