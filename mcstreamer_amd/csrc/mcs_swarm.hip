// mcs_swarm.hip -- many Swarm.swarm instances (Swarm/src/Swarm.jl:102-142) in one
// set of launches.
//
// Layout: one particle store; slot s owns the slab [base_s, base_s + cap_s).
// Work is cut into tiles of one workgroup, each tile inside one slab, so every
// per-slot quantity (births, clones, compaction offsets, tallies) is a scan or a
// reduction over the slot's consecutive tiles.  Every kernel runs on a grid that
// depends only on the set of slots, never on particle counts (those live on the
// device), so a step is captured once into a HIP graph and replayed.
//
// The store is either the FP64 SoA of the single population (bit-exact with the
// oracle) or the split FP32 layout of mcs_fp32.h (packed cell of a virtual grid
// plus an FP32 offset; FP32 push and collisions); tallies are FP64 compensated
// sums in both.
#include <algorithm>
#include <cstring>
#include <vector>

#include "mcs_advance.h"
#include "mcs_fp32.h"

#define SW_NQ MCS_TALLY_NQ
#define SW_NT (SW_NQ + 1)     // count + the tallied quantities
#define SW_NX 6               // position moments: sum x_d, sum x_d^2
#define SW_WIN MCS_DIFFUSION_WINDOW
#define SW_XCHUNK 64          // tiles per workgroup of the position moments
#define SW_MODE_NONE 0
#define SW_MODE_DROP 1
#define SW_MODE_CLONE 2

// Virtual grid of the FP32 store: cells of SW_H32 m, x = 0 at the corner of cell
// (SW_OX, SW_OY, SW_OZ); the packed coordinates span the whole 10/10/11-bit range.
#define SW_H32 1e-5
#define SW_OX (MCS_CX_MAX / 2)
#define SW_OY (MCS_CY_MAX / 2)
#define SW_OZ (MCS_CZ_MAX / 2)

struct SlotDev {
    int64_t *base, *cap, *n, *n0;
    int32_t *used, *flags, *mode, *relax, *block;
    double *e, *dt, *tol, *param;
    uint32_t *key, *step;
    uint64_t *next_id;
    TableView *tv;
    TableView32 *tv32;
    int32_t *first_tile, *ntiles;
    // tallies: per slot SW_NQ each
    double *sum, *sumc, *cur, *bm1, *bm2, *last;
    int64_t *steps, *nrec, *nblk, *curcnt, *nact;
    // diffusion: sigma^2 per axis of the last SW_WIN steps (ring by step number)
    // and the least-squares D = slope / 2 over them, 3 per slot
    double *ring, *diff;
};

// Birth logs: one record per ionization child, written in any order and placed
// by (parent's tile offset, offset within the tile, sequence within the parent).
struct SwLog {
    int64_t *parent;
    int32_t *seq;
    double *x[3], *v[3];
    double *s;
    uint8_t *act;
    int64_t cap;
};

struct SwLog32 {
    int64_t *parent;
    int32_t *seq;
    int32_t *cell;
    float *off[3], *v[3];
    float *s;
    uint8_t *act;
    int64_t cap;
};

template <typename Buf, typename Log>
struct SwArgsT {
    Buf b, alt;
    SlotDev S;
    const int32_t *tile_slot;
    const int64_t *tile_start;
    uint64_t seed;
    Log L;
    int64_t *log_count;
    int32_t *loff, *cnt;
    int64_t *tile_births, *tile_bpre, *tile_clones, *tile_cpre, *tile_act, *tile_apre;
    double *tile_part;   // ntiles x SW_NT x 2 (Neumaier sum, compensation)
    double *tile_xpart;  // ntiles x SW_NX x 2: the position moments after the step, at each chunk's first tile
};
using SwArgs64 = SwArgsT<PopBuf, SwLog>;
using SwArgs32 = SwArgsT<SplitBuf, SwLog32>;

struct Neu2 {
    double s, c;
};

__device__ __forceinline__ void nadd(Neu2 &a, double x)
{
    double t = a.s + x;
    if (fabs(a.s) >= fabs(x))
        a.c += (a.s - t) + x;
    else
        a.c += (x - t) + a.s;
    a.s = t;
}

__device__ __forceinline__ void nmerge(Neu2 &a, const Neu2 &b)
{
    double t = a.s + b.s;
    double c = fabs(a.s) >= fabs(b.s) ? (a.s - t) + b.s : (b.s - t) + a.s;
    a.s = t;
    a.c = a.c + b.c + c;
}

template <int BS, typename T>
__device__ __forceinline__ T bscan(T v, T *total)
{
    __shared__ T ws[BS / MCS_WAVE];
    int lane = threadIdx.x & (MCS_WAVE - 1), wid = threadIdx.x / MCS_WAVE;
    T x = v;
#pragma unroll
    for (int o = 1; o < MCS_WAVE; o <<= 1) {
        T y = __shfl_up(x, o, MCS_WAVE);
        if (lane >= o) x += y;
    }
    if (lane == MCS_WAVE - 1) ws[wid] = x;
    __syncthreads();
    T pre = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < BS / MCS_WAVE; ++k) {
        if (k < wid) pre += ws[k];
        tot += ws[k];
    }
    __syncthreads();
    *total = tot;
    return pre + x - v;
}

// Tallies of one particle state (Swarm.jl:126-130,152,156 with collrate :242-248):
// count, v_z, energy, summed rate of the ionization rows, of the attachment rows.
__device__ __forceinline__ void tally_add(Neu2 acc[SW_NT], const double v[3], const TableView &T)
{
    double eng = energy_dev(v[0], v[1], v[2]);
    int64_t k;
    double w;
    int32_t fl = 0;
    indweight_dev(T.energy, T.ne, T.egrid, T.step, T.inv_dlog, eng, k, w, fl);
    const double *r0 = T.rate + k * T.nproc;
    const double *r1 = r0 + T.nproc;
    double ion = 0.0, att = 0.0;
    for (int j = 0; j < T.nproc; ++j) {
        int kd = gld(T.kind + j);
        if (kd == MCS_PROC_IONIZATION) ion += w * gld(r0 + j) + (1.0 - w) * gld(r1 + j);
        else if (kd == MCS_PROC_ATTACHMENT) att += w * gld(r0 + j) + (1.0 - w) * gld(r1 + j);
    }
    nadd(acc[0], 1.0);
    nadd(acc[1], v[2]);
    nadd(acc[2], eng);
    nadd(acc[3], ion);
    nadd(acc[4], att);
}

// The same tallies of an FP32 state: the state and the interpolated rates are
// FP32, the accumulation FP64 (compensated).
__device__ __forceinline__ void tally_add32(Neu2 acc[SW_NT], const float v[3], const TableView32 &T)
{
    const float eng = energy32(v);
    int32_t k;
    float w;
    int32_t fl = 0;
    indweight32(T, eng, k, w, fl);
    const float *r0 = T.rate + k * T.nproc;
    const float *r1 = r0 + T.nproc;
    float ion = 0.0f, att = 0.0f;
    for (int j = 0; j < T.nproc; ++j) {
        const int kd = proc_kind(T, j);
        if (kd == MCS_PROC_IONIZATION) ion += w * gld(r0 + j) + (1.0f - w) * gld(r1 + j);
        else if (kd == MCS_PROC_ATTACHMENT) att += w * gld(r0 + j) + (1.0f - w) * gld(r1 + j);
    }
    const double vd[3] = {(double)v[0], (double)v[1], (double)v[2]};
    nadd(acc[0], 1.0);
    nadd(acc[1], vd[2]);
    nadd(acc[2], energy_dev(vd[0], vd[1], vd[2]));
    nadd(acc[3], (double)ion);
    nadd(acc[4], (double)att);
}

// Block-reduce NQ Neumaier pairs into out (a tile's partials): a butterfly over
// the lanes of each wavefront (nmerge is symmetric, so every lane ends with the
// wave's total), then one barrier and the waves' totals merged in wave order.
// The order is fixed by the tile alone, so a slot's tallies do not depend on
// the other slots.
template <int BS, int NQ = SW_NT>
__device__ __forceinline__ void tile_reduce(Neu2 acc[NQ], double *out)
{
    constexpr int NW = BS / MCS_WAVE;
    __shared__ double sh[NW][NQ][2];
    const int lane = threadIdx.x & (MCS_WAVE - 1), wid = threadIdx.x / MCS_WAVE;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
#pragma unroll
        for (int o = MCS_WAVE / 2; o > 0; o >>= 1) {
            const Neu2 b{__shfl_xor(acc[q].s, o, MCS_WAVE), __shfl_xor(acc[q].c, o, MCS_WAVE)};
            nmerge(acc[q], b);
        }
    }
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            sh[wid][q][0] = acc[q].s;
            sh[wid][q][1] = acc[q].c;
        }
    }
    __syncthreads();
    if (threadIdx.x < NQ) {
        const int q = threadIdx.x;
        Neu2 a{sh[0][q][0], sh[0][q][1]};
#pragma unroll
        for (int w = 1; w < NW; ++w) nmerge(a, Neu2{sh[w][q][0], sh[w][q][1]});
        out[2 * q] = a.s;
        out[2 * q + 1] = a.c;
    }
}

// advance! of every slot's particles (JuMC/src/mixed_population.jl:24-54), births
// to the log, and the post-step tallies of parents and children.  FP64 store; six
// waves per SIMD (a few cold values spill), which hides the loads better than four.
template <int BS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_sw_advance(SwArgs64 a)
{
    const int64_t tile = blockIdx.x;
    const int slot = a.tile_slot[tile];
    const int64_t i = a.tile_start[tile] + threadIdx.x;
    const int64_t lim = a.S.base[slot] + a.S.n[slot];
    const TableView T = a.S.tv[slot];
    Neu2 acc[SW_NT];
#pragma unroll
    for (int q = 0; q < SW_NT; ++q) acc[q] = Neu2{0.0, 0.0};
    int nb = 0;
    int32_t flags = 0;
    if (i < lim && a.b.active[i]) {
        double x[3] = {a.b.x[0][i], a.b.x[1][i], a.b.x[2][i]};
        double v[3] = {a.b.v[0][i], a.b.v[1][i], a.b.v[2][i]};
        const double w = a.b.w[i];
        double s = a.b.s[i];
        uint8_t act = 1;
        mcs_stream rs;
        mcs_stream_init(&rs, a.seed, a.S.key[slot], MCS_PHASE_ADVANCE, a.b.id[i], a.S.step[slot]);
        const double eu[3] = {a.S.e[3 * slot], a.S.e[3 * slot + 1], a.S.e[3 * slot + 2]};
        MeshView M;
        int flights = 0, deaths = 0;
        advance_one<0, 0>(
            x, v, w, s, act, rs, a.S.dt[slot], T, M, eu, flags, flights, deaths,
            [&](const double *bx, const double *bv, double bs, uint8_t bact) {
                unsigned long long r = atomicAdd((unsigned long long *)a.log_count, 1ULL);
                if ((int64_t)r < a.L.cap) {
                    a.L.parent[r] = i;
                    a.L.seq[r] = nb;
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        a.L.x[d][r] = bx[d];
                        a.L.v[d][r] = bv[d];
                    }
                    a.L.s[r] = bs;
                    a.L.act[r] = bact;
                } else {
                    flags |= MCS_FLAG_CAPACITY;
                }
                if (bact) tally_add(acc, bv, T);
                nb += 1;
            },
            [&](int) {}, [&](const double *, double) {});
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            a.b.x[d][i] = x[d];
            a.b.v[d][i] = v[d];
        }
        a.b.s[i] = s;
        if (!act) a.b.active[i] = 0;
        if (act) tally_add(acc, v, T);
    }
    int tot;
    int ex = bscan<BS, int>(nb, &tot);
    if (nb > 0) a.loff[i] = ex;
    if (threadIdx.x == 0) a.tile_births[tile] = tot;
    tile_reduce<BS>(acc, a.tile_part + tile * (2 * SW_NT));
    if (flags) atomicOr(&a.S.flags[slot], flags);
}

// The same step on the FP32 split store, at eight waves per SIMD (64 registers;
// a few cold values spill): more waves hide the table and field loads better.
template <int BS, int BANDED>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_sw_advance32(SwArgs32 a)
{
    const int64_t tile = blockIdx.x;
    const int slot = a.tile_slot[tile];
    const int64_t i = a.tile_start[tile] + threadIdx.x;
    const int64_t lim = a.S.base[slot] + a.S.n[slot];
    const TableView32 T = a.S.tv32[slot];
    Neu2 acc[SW_NT];
#pragma unroll
    for (int q = 0; q < SW_NT; ++q) acc[q] = Neu2{0.0, 0.0};
    int nb = 0;
    int32_t flags = 0;
    if (i < lim && a.b.active[i]) {
        int32_t c[3];
        unpack_cell(a.b.cell[i], c[0], c[1], c[2]);
        float off[3] = {a.b.off[0][i], a.b.off[1][i], a.b.off[2][i]};
        float v[3] = {a.b.v[0][i], a.b.v[1][i], a.b.v[2][i]};
        const float w = a.b.w[i];
        float s = a.b.s[i];
        uint8_t act = 1;
        mcs_stream rs;
        mcs_stream_init(&rs, a.seed, a.S.key[slot], MCS_PHASE_ADVANCE, a.b.id[i], a.S.step[slot]);
        const float acc_u[3] = {(float)(MCS_QM * a.S.e[3 * slot]), (float)(MCS_QM * a.S.e[3 * slot + 1]),
                                (float)(MCS_QM * a.S.e[3 * slot + 2])};
        MeshView32 M;
        int flights = 0, deaths = 0;
        auto birth = [&](const int32_t *bc, const float *boff, const float *bv, float bs, uint8_t bact) {
            unsigned long long r = atomicAdd((unsigned long long *)a.log_count, 1ULL);
            if ((int64_t)r < a.L.cap) {
                a.L.parent[r] = i;
                a.L.seq[r] = nb;
                a.L.cell[r] = pack_cell(bc[0], bc[1], bc[2]);
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    a.L.off[d][r] = boff[d];
                    a.L.v[d][r] = bv[d];
                }
                a.L.s[r] = bs;
                a.L.act[r] = bact;
            } else {
                flags |= MCS_FLAG_CAPACITY;
            }
            if (bact) tally_add32(acc, bv, T);
            nb += 1;
        };
        if (BANDED)
            advance_one32_banded(c, off, v, w, s, act, rs, (float)a.S.dt[slot], T, acc_u, (float)(1.0 / SW_H32),
                                 flags, flights, deaths, birth);
        else
            advance_one32<0, 0>(c, off, v, w, s, act, rs, (float)a.S.dt[slot], T, M, acc_u, (float)(1.0 / SW_H32),
                                flags, flights, deaths, birth, [&](int) {},
                                [&](const int32_t *, const float *, double) {});
        const int32_t nmax[3] = {MCS_CX_MAX, MCS_CY_MAX, MCS_CZ_MAX};
        rebase32(c, off, nmax);
        a.b.cell[i] = pack_cell(c[0], c[1], c[2]);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            a.b.off[d][i] = off[d];
            a.b.v[d][i] = v[d];
        }
        a.b.s[i] = s;
        if (!act) a.b.active[i] = 0;
        if (act) tally_add32(acc, v, T);
    }
    int tot;
    int ex = bscan<BS, int>(nb, &tot);
    if (nb > 0) a.loff[i] = ex;
    if (threadIdx.x == 0) a.tile_births[tile] = tot;
    tile_reduce<BS>(acc, a.tile_part + tile * (2 * SW_NT));
    if (flags) atomicOr(&a.S.flags[slot], flags);
}

// The tile of a particle index: bisection over the (sorted) tile starts.
__device__ __forceinline__ int64_t tile_of(const int64_t *tile_start, int64_t ntiles, int64_t i)
{
    int64_t lo = 0, hi = ntiles - 1;
    while (lo < hi) {
        int64_t mid = (lo + hi + 1) >> 1;
        if (tile_start[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ void put_child(const SwArgs64 &a, int64_t r, int64_t i, int64_t d)
{
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        a.b.x[q][d] = a.L.x[q][r];
        a.b.v[q][d] = a.L.v[q][r];
    }
    a.b.w[d] = a.b.w[i];
    a.b.s[d] = a.L.s[r];
    a.b.active[d] = a.L.act[r];
}

__device__ __forceinline__ void put_child(const SwArgs32 &a, int64_t r, int64_t i, int64_t d)
{
    a.b.cell[d] = a.L.cell[r];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        a.b.off[q][d] = a.L.off[q][r];
        a.b.v[q][d] = a.L.v[q][r];
    }
    a.b.w[d] = a.b.w[i];
    a.b.s[d] = a.L.s[r];
    a.b.active[d] = a.L.act[r];
}

// Children from the birth log to their slabs: rank within the slot in parent
// order (tile offset + offset within the tile + sequence within the parent),
// ids next_id + rank.  Grid-stride over the log with a fixed grid.
template <typename A>
__global__ __launch_bounds__(MCS_BLOCK) void k_sw_births(A a, int64_t logcap, int64_t ntiles)
{
    const int64_t cnt = *a.log_count < logcap ? *a.log_count : logcap;
    for (int64_t r = gidx(); r < cnt; r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = a.L.parent[r];
        const int64_t tile = tile_of(a.tile_start, ntiles, i);
        const int slot = a.tile_slot[tile];
        const int64_t base = a.S.base[slot], cap = a.S.cap[slot], n = a.S.n[slot];
        const int64_t f = a.S.first_tile[slot];
        const int64_t rank = (a.tile_bpre[tile] - a.tile_bpre[f]) + a.loff[i] + a.L.seq[r];
        if (n + rank >= cap) continue;
        const int64_t d = base + n + rank;
        put_child(a, r, i, d);
        a.b.id[d] = a.S.next_id[slot] + (uint64_t)rank;
    }
}

// Position of particle i (FP64 store: x itself; split store: the virtual grid's
// cell and offset, as the download converts them).
__device__ __forceinline__ void sw_pos(const PopBuf &b, int64_t i, double x[3])
{
    x[0] = b.x[0][i];
    x[1] = b.x[1][i];
    x[2] = b.x[2][i];
}

__device__ __forceinline__ void sw_pos(const SplitBuf &b, int64_t i, double x[3])
{
    int32_t c[3];
    unpack_cell(b.cell[i], c[0], c[1], c[2]);
    const int32_t o[3] = {SW_OX, SW_OY, SW_OZ};
#pragma unroll
    for (int q = 0; q < 3; ++q) x[q] = ((double)(c[q] - o[q]) + (double)b.off[q][i]) * SW_H32;
}

// The position moments of Swarm.jl:126-127 after the step (the slot's parents and
// the children placed after them; actives only) as Neumaier pairs: sum x_d,
// sum x_d^2; the count is the tallies' (the same actives).  One workgroup per
// (slot, chunk of SW_XCHUNK tiles): each thread walks the chunk at a stride of the
// workgroup, and the chunk's partials go to the slot of its first tile.  The
// chunks start at fixed offsets from the slab base, so a slot's sums do not
// depend on the other slots.
template <int BS, typename A>
__global__ __launch_bounds__(BS) void k_sw_xmom(A a)
{
    const int slot = blockIdx.x;
    const int64_t chunk = blockIdx.y;
    if (!a.S.used[slot]) return;
    const int64_t f = a.S.first_tile[slot], nt = a.S.ntiles[slot];
    if (chunk * SW_XCHUNK >= nt) return;
    const int64_t births = a.tile_bpre[f + nt] - a.tile_bpre[f];
    int64_t lim = a.S.n[slot] + births;
    if (lim > a.S.cap[slot]) lim = a.S.cap[slot];
    const int64_t base = a.S.base[slot];
    const int64_t i0 = base + chunk * (SW_XCHUNK * BS);
    const int64_t i1 = min(base + lim, i0 + (int64_t)SW_XCHUNK * BS);
    Neu2 acc[SW_NX];
#pragma unroll
    for (int q = 0; q < SW_NX; ++q) acc[q] = Neu2{0.0, 0.0};
#pragma unroll 4
    for (int64_t i = i0 + threadIdx.x; i < i1; i += BS) {
        if (a.b.active[i]) {
            double x[3];
            sw_pos(a.b, i, x);
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                nadd(acc[d], x[d]);
                nadd(acc[3 + d], x[d] * x[d]);
            }
        }
    }
    tile_reduce<BS, SW_NX>(acc, a.tile_xpart + (f + chunk * SW_XCHUNK) * (2 * SW_NX));
}

// Sum of NQ Neumaier pairs at the tiles f, f + stride, ... (n of them) by one
// workgroup; the totals (sum + compensation) to tot.
template <int NQ>
__device__ __forceinline__ void slot_merge(const double *part, int64_t f, int64_t n, int64_t stride, double *tot)
{
    Neu2 acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = Neu2{0.0, 0.0};
    for (int64_t t = threadIdx.x; t < n; t += MCS_BLOCK) {
        const double *p = part + (f + t * stride) * (2 * NQ);
#pragma unroll
        for (int q = 0; q < NQ; ++q) nmerge(acc[q], Neu2{p[2 * q], p[2 * q + 1]});
    }
    __shared__ double sh[MCS_BLOCK], shc[MCS_BLOCK];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        sh[threadIdx.x] = acc[q].s;
        shc[threadIdx.x] = acc[q].c;
        __syncthreads();
        for (int o = MCS_BLOCK / 2; o > 0; o >>= 1) {
            if (threadIdx.x < o) {
                Neu2 x{sh[threadIdx.x], shc[threadIdx.x]};
                nmerge(x, Neu2{sh[threadIdx.x + o], shc[threadIdx.x + o]});
                sh[threadIdx.x] = x.s;
                shc[threadIdx.x] = x.c;
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) tot[q] = sh[0] + shc[0];
        __syncthreads();
    }
}

// sigma^2 of this step into the slot's ring, and D = slope / 2 of the least-squares
// line through (t, sigma^2) over the last min(steps, SW_WIN) steps (Swarm.jl:144-150:
// [ones(l) t] \ sigma2 over the last l steps), per axis.
__device__ void slot_diffusion(const SlotDev &S, int slot, int64_t st, int64_t nact, const double *xm)
{
    double *ring = S.ring + (int64_t)slot * SW_WIN * 3;
    const int64_t idx = st % SW_WIN;
    for (int d = 0; d < 3; ++d) {
        double s2 = 0.0;
        if (nact > 0) {
            const double X = xm[d] / (double)nact, X2 = xm[3 + d] / (double)nact;
            s2 = X2 - X * X;
        }
        ring[3 * idx + d] = s2;
    }
    const int64_t m = st + 1 < SW_WIN ? st + 1 : SW_WIN;
    const double dt = S.dt[slot];
    for (int d = 0; d < 3; ++d) {
        double D = 0.0;
        if (m >= 2) {
            // steps st - j, j = 0..m-1; centred at their mean
            const double jbar = 0.5 * (double)(m - 1);
            double ybar = 0.0;
            for (int64_t j = 0; j < m; ++j) ybar += ring[3 * ((st - j) % SW_WIN) + d];
            ybar /= (double)m;
            double sxy = 0.0, sxx = 0.0;
            for (int64_t j = 0; j < m; ++j) {
                const double u = jbar - (double)j;   // (step - mean step)
                sxy += u * (ring[3 * ((st - j) % SW_WIN) + d] - ybar);
                sxx += u * u;
            }
            D = sxy / (sxx * dt) / 2.0;
        }
        S.diff[3 * slot + d] = D;
    }
}

// Per slot: births -> n, next_id; tallies of this step; resample decision of
// Swarm.resample! (Swarm.jl:194-199).  One workgroup per slot.
template <typename A>
__global__ __launch_bounds__(MCS_BLOCK) void k_sw_slot_step(A a)
{
    const int slot = blockIdx.x;
    if (!a.S.used[slot]) return;
    const int64_t f = a.S.first_tile[slot], nt = a.S.ntiles[slot];
    Neu2 acc[SW_NT];
#pragma unroll
    for (int q = 0; q < SW_NT; ++q) acc[q] = Neu2{0.0, 0.0};
    for (int64_t t = threadIdx.x; t < nt; t += blockDim.x) {
        const double *p = a.tile_part + (f + t) * (2 * SW_NT);
#pragma unroll
        for (int q = 0; q < SW_NT; ++q) nmerge(acc[q], Neu2{p[2 * q], p[2 * q + 1]});
    }
    __shared__ double sh[MCS_BLOCK], shc[MCS_BLOCK];
    __shared__ double tot[SW_NT];
#pragma unroll
    for (int q = 0; q < SW_NT; ++q) {
        sh[threadIdx.x] = acc[q].s;
        shc[threadIdx.x] = acc[q].c;
        __syncthreads();
        for (int o = MCS_BLOCK / 2; o > 0; o >>= 1) {
            if (threadIdx.x < o) {
                Neu2 x{sh[threadIdx.x], shc[threadIdx.x]};
                nmerge(x, Neu2{sh[threadIdx.x + o], shc[threadIdx.x + o]});
                sh[threadIdx.x] = x.s;
                shc[threadIdx.x] = x.c;
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) tot[q] = sh[0] + shc[0];
        __syncthreads();
    }
    __shared__ double xm[SW_NX];
    slot_merge<SW_NX>(a.tile_xpart, f, (nt + SW_XCHUNK - 1) / SW_XCHUNK, SW_XCHUNK, xm);
    if (threadIdx.x != 0) return;
    // births
    int64_t births = a.tile_bpre[f + nt] - a.tile_bpre[f];
    int64_t n = a.S.n[slot], cap = a.S.cap[slot];
    if (n + births > cap) {
        a.S.flags[slot] |= MCS_FLAG_CAPACITY;
        births = cap - n;
    }
    a.S.n[slot] = n + births;
    a.S.next_id[slot] += (uint64_t)births;
    // tallies
    const int64_t nact = (int64_t)tot[0];
    a.S.nact[slot] = nact;
    const int64_t st = a.S.steps[slot];
    slot_diffusion(a.S, slot, st, nact, xm);
    double m[SW_NQ];
    for (int q = 0; q < SW_NQ; ++q) {
        m[q] = nact > 0 ? tot[q + 1] / (double)nact : 0.0;
        a.S.last[SW_NQ * slot + q] = m[q];
    }
    if (st >= a.S.relax[slot] && nact > 0) {
        for (int q = 0; q < SW_NQ; ++q) {
            Neu2 sacc{a.S.sum[SW_NQ * slot + q], a.S.sumc[SW_NQ * slot + q]};
            nadd(sacc, m[q]);
            a.S.sum[SW_NQ * slot + q] = sacc.s;
            a.S.sumc[SW_NQ * slot + q] = sacc.c;
            a.S.cur[SW_NQ * slot + q] += m[q];
        }
        a.S.nrec[slot] += 1;
        a.S.curcnt[slot] += 1;
        if (a.S.curcnt[slot] == a.S.block[slot]) {
            for (int q = 0; q < SW_NQ; ++q) {
                double bm = a.S.cur[SW_NQ * slot + q] / (double)a.S.block[slot];
                a.S.bm1[SW_NQ * slot + q] += bm;
                a.S.bm2[SW_NQ * slot + q] += bm * bm;
                a.S.cur[SW_NQ * slot + q] = 0.0;
            }
            a.S.nblk[slot] += 1;
            a.S.curcnt[slot] = 0;
        }
    }
    // Swarm.resample!: nothing if the count is within tol of n0
    const int64_t n0 = a.S.n0[slot];
    int mode = SW_MODE_NONE;
    double par = 0.0;
    if (!(fabs((double)(nact - n0) / (double)n0) < a.S.tol[slot]) && nact > 0) {
        if (nact > n0) {
            mode = SW_MODE_DROP;
            par = (double)(nact - n0) / (double)nact;
        } else {
            mode = SW_MODE_CLONE;
            par = (double)(n0 - nact) / (double)nact;
        }
    }
    a.S.mode[slot] = mode;
    a.S.param[slot] = par;
}

// Per-slot clone counts / drops of Swarm.resample! (Swarm.jl:200-221): one draw
// stream per particle (phase swarm-resample), as in mcs_swarm_resample.
template <int BS, typename A>
__global__ __launch_bounds__(BS) void k_sw_resample(A a)
{
    const int64_t tile = blockIdx.x;
    const int slot = a.tile_slot[tile];
    const int64_t i = a.tile_start[tile] + threadIdx.x;
    const int64_t lim = a.S.base[slot] + a.S.n[slot];
    const int mode = a.S.mode[slot];
    // slots that need no clones (none, or a drop) write only their zero tile total;
    // k_sw_clone_place skips them
    if (mode != SW_MODE_CLONE) {
        if (mode == SW_MODE_DROP && i < lim && a.b.active[i]) {
            mcs_stream rs;
            mcs_stream_init(&rs, a.seed, a.S.key[slot], MCS_PHASE_SWARM_RESAMPLE, a.b.id[i], a.S.step[slot]);
            if (mcs_u01(&rs) < a.S.param[slot]) a.b.active[i] = 0;
        }
        if (threadIdx.x == 0) a.tile_clones[tile] = 0;
        return;
    }
    int c = 0;
    if (i < lim && a.b.active[i]) {
        mcs_stream rs;
        mcs_stream_init(&rs, a.seed, a.S.key[slot], MCS_PHASE_SWARM_RESAMPLE, a.b.id[i], a.S.step[slot]);
        c = (int)mcs_poisson(&rs, a.S.param[slot]);
    }
    a.cnt[i] = c;
    int tot;
    int ex = bscan<BS, int>(c, &tot);
    a.loff[i] = ex;
    if (threadIdx.x == 0) a.tile_clones[tile] = tot;
}

template <int BS, typename A>
__global__ __launch_bounds__(BS) void k_sw_clone_place(A a)
{
    const int64_t tile = blockIdx.x;
    const int slot = a.tile_slot[tile];
    if (a.S.mode[slot] != SW_MODE_CLONE) return;
    const int64_t i = a.tile_start[tile] + threadIdx.x;
    const int c = a.cnt[i];
    if (c == 0) return;
    const int64_t base = a.S.base[slot], cap = a.S.cap[slot], n = a.S.n[slot];
    const int64_t f = a.S.first_tile[slot];
    const int64_t off = (a.tile_cpre[tile] - a.tile_cpre[f]) + a.loff[i];
    const uint64_t nid = a.S.next_id[slot];
    for (int j = 0; j < c; ++j) {
        int64_t rank = off + j;
        if (n + rank >= cap) break;
        int64_t d = base + n + rank;
        copy_rec(a.b, i, a.b, d);
        a.b.id[d] = nid + (uint64_t)rank;
    }
}

template <typename A>
__global__ void k_sw_slot_step2(A a, int nslots)
{
    const int slot = gidx();
    if (slot >= nslots || !a.S.used[slot]) return;
    const int64_t f = a.S.first_tile[slot], nt = a.S.ntiles[slot];
    int64_t clones = a.tile_cpre[f + nt] - a.tile_cpre[f];
    int64_t n = a.S.n[slot], cap = a.S.cap[slot];
    if (n + clones > cap) {
        a.S.flags[slot] |= MCS_FLAG_CAPACITY;
        clones = cap - n;
    }
    a.S.n[slot] = n + clones;
    a.S.next_id[slot] += (uint64_t)clones;
    a.S.step[slot] += 1;
    a.S.steps[slot] += 1;
}

// Stable compaction of every slab: actives keep their order.
template <int BS, typename A>
__global__ __launch_bounds__(BS) void k_sw_compact_count(A a)
{
    const int64_t tile = blockIdx.x;
    const int slot = a.tile_slot[tile];
    const int64_t i = a.tile_start[tile] + threadIdx.x;
    const int64_t lim = a.S.base[slot] + a.S.n[slot];
    int c = (i < lim && a.b.active[i]) ? 1 : 0;
    int tot;
    int ex = bscan<BS, int>(c, &tot);
    a.loff[i] = ex;
    if (threadIdx.x == 0) a.tile_act[tile] = tot;
}

template <int BS, typename A>
__global__ __launch_bounds__(BS) void k_sw_compact_write(A a)
{
    const int64_t tile = blockIdx.x;
    const int slot = a.tile_slot[tile];
    const int64_t i = a.tile_start[tile] + threadIdx.x;
    const int64_t lim = a.S.base[slot] + a.S.n[slot];
    if (!(i < lim && a.b.active[i])) return;
    const int64_t f = a.S.first_tile[slot];
    int64_t d = a.S.base[slot] + (a.tile_apre[tile] - a.tile_apre[f]) + a.loff[i];
    copy_rec(a.b, i, a.alt, d);
}

template <typename A>
__global__ void k_sw_compact_slot(A a, int nslots)
{
    const int slot = gidx();
    if (slot >= nslots || !a.S.used[slot]) return;
    const int64_t f = a.S.first_tile[slot], nt = a.S.ntiles[slot];
    a.S.n[slot] = a.tile_apre[f + nt] - a.tile_apre[f];
}

template <int BS, typename A>
__global__ __launch_bounds__(BS) void k_sw_copy_back(A a)
{
    const int64_t tile = blockIdx.x;
    const int slot = a.tile_slot[tile];
    const int64_t i = a.tile_start[tile] + threadIdx.x;
    if (i < a.S.base[slot] + a.S.n[slot]) copy_rec(a.alt, i, a.b, i);
}

// n electrons at the origin with Maxwellian velocities (mcs_maxwell_init).
__global__ __launch_bounds__(MCS_BLOCK) void k_sw_init(PopBuf b, int64_t base, int64_t n, uint64_t seed,
                                                       uint32_t key, double vth)
{
    int64_t i = gidx();
    if (i >= n) return;
    double v[3], s;
    mcs_maxwell_init(seed, key, (uint64_t)i, vth, v, &s);
    const int64_t d = base + i;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        b.x[q][d] = 0.0;
        b.v[q][d] = v[q];
    }
    b.w[d] = 1.0;
    b.s[d] = s;
    b.active[d] = 1;
    b.id[d] = (uint64_t)i;
}

// The same electrons in the FP32 store (draws in FP64, stored rounded).
__global__ __launch_bounds__(MCS_BLOCK) void k_sw_init32(SplitBuf b, int64_t base, int64_t n, uint64_t seed,
                                                         uint32_t key, double vth)
{
    int64_t i = gidx();
    if (i >= n) return;
    double v[3], s;
    mcs_maxwell_init(seed, key, (uint64_t)i, vth, v, &s);
    const int64_t d = base + i;
    b.cell[d] = pack_cell(SW_OX, SW_OY, SW_OZ);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        b.off[q][d] = 0.0f;
        b.v[q][d] = (float)v[q];
    }
    b.w[d] = 1.0f;
    b.s[d] = (float)s;
    b.active[d] = 1;
    b.id[d] = (uint64_t)i;
}

// Host particles (x, v as n x 3 doubles) into the FP32 store.
__global__ __launch_bounds__(MCS_BLOCK) void k_sw_from64(SplitBuf b, int64_t base, int64_t n, const double *x,
                                                         const double *v, const double *w, const double *s,
                                                         const uint8_t *act)
{
    int64_t i = gidx();
    if (i >= n) return;
    const int64_t d = base + i;
    const int32_t o[3] = {SW_OX, SW_OY, SW_OZ}, nmax[3] = {MCS_CX_MAX, MCS_CY_MAX, MCS_CZ_MAX};
    int32_t c[3];
    float off[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const double r = x[3 * i + q] * (1.0 / SW_H32);
        const double fl = floor(r);
        c[q] = o[q] + (int32_t)fmax(fmin(fl, 1e9), -1e9);
        off[q] = (float)(r - fl);
        b.v[q][d] = (float)v[3 * i + q];
    }
    rebase32(c, off, nmax);
    b.cell[d] = pack_cell(c[0], c[1], c[2]);
#pragma unroll
    for (int q = 0; q < 3; ++q) b.off[q][d] = off[q];
    b.w[d] = (float)w[i];
    b.s[d] = (float)s[i];
    b.active[d] = act[i];
    b.id[d] = (uint64_t)i;
}

// The FP32 store back to FP64 positions (m) and velocities, into n x 3 arrays.
__global__ __launch_bounds__(MCS_BLOCK) void k_sw_to64(SplitBuf b, int64_t base, int64_t n, double *x, double *v,
                                                       double *w, double *s)
{
    int64_t i = gidx();
    if (i >= n) return;
    const int64_t d = base + i;
    int32_t c[3];
    unpack_cell(b.cell[d], c[0], c[1], c[2]);
    const int32_t o[3] = {SW_OX, SW_OY, SW_OZ};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        x[3 * i + q] = ((double)(c[q] - o[q]) + (double)b.off[q][d]) * SW_H32;
        v[3 * i + q] = (double)b.v[q][d];
    }
    w[i] = (double)b.w[d];
    s[i] = (double)b.s[d];
}

__global__ void k_sw_zero(int64_t *p, int n)
{
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

// One slot's parameters and zeroed tallies, written in one launch.
struct SlotInit {
    int64_t base, cap, n, n0;
    int32_t relax, block;
    double e[3], dt, tol;
    uint32_t key;
    uint64_t next_id;
    TableView tv;
    TableView32 tv32;
};

__global__ void k_sw_slot_init(SlotDev S, int slot, SlotInit p)
{
    if (threadIdx.x != 0) return;
    S.base[slot] = p.base;
    S.cap[slot] = p.cap;
    S.n[slot] = p.n;
    S.n0[slot] = p.n0;
    S.used[slot] = 1;
    S.flags[slot] = 0;
    S.mode[slot] = SW_MODE_NONE;
    S.relax[slot] = p.relax;
    S.block[slot] = p.block;
    for (int d = 0; d < 3; ++d) S.e[3 * slot + d] = p.e[d];
    S.dt[slot] = p.dt;
    S.tol[slot] = p.tol;
    S.param[slot] = 0.0;
    S.key[slot] = p.key;
    S.step[slot] = 0;
    S.next_id[slot] = p.next_id;
    S.tv[slot] = p.tv;
    S.tv32[slot] = p.tv32;
    for (int q = 0; q < SW_NQ; ++q) {
        const int k = SW_NQ * slot + q;
        S.sum[k] = 0.0;
        S.sumc[k] = 0.0;
        S.cur[k] = 0.0;
        S.bm1[k] = 0.0;
        S.bm2[k] = 0.0;
        S.last[k] = 0.0;
    }
    for (int d = 0; d < 3; ++d) S.diff[3 * slot + d] = 0.0;
    S.steps[slot] = 0;
    S.nrec[slot] = 0;
    S.nblk[slot] = 0;
    S.curcnt[slot] = 0;
    S.nact[slot] = 0;
}

__global__ void k_sw_slot_free(SlotDev S, int slot)
{
    if (threadIdx.x == 0) S.used[slot] = 0;
}

// ---------------------------------------------------------------- host side

struct HostSlot {
    bool used = false;
    int64_t base = 0, cap = 0;
    const mcs_table *table = nullptr;
    mcs_swarm_params params;
};

struct mcs_swarm_batch {
    mcs_ctx *ctx;
    int precision;
    int64_t total_cap;
    int max_slots;
    uint64_t seed;
    PopBuf buf, alt;
    SplitBuf buf32, alt32;
    SlotDev S;
    SwLog L;
    SwLog32 L32;
    int64_t *log_count = nullptr;
    int32_t *loff = nullptr, *cnt = nullptr;
    int64_t max_tiles = 0;
    int32_t *tile_slot = nullptr;
    int64_t *tile_start = nullptr;
    int64_t *tile_births = nullptr, *tile_bpre = nullptr, *tile_clones = nullptr, *tile_cpre = nullptr;
    int64_t *tile_act = nullptr, *tile_apre = nullptr;
    int64_t *scan_scratch = nullptr;
    double *tile_part = nullptr;
    double *tile_xpart = nullptr;
    double *stage = nullptr;       // host-exchange staging (FP32 store), 8 doubles per particle
    int64_t stage_cap = 0;
    std::vector<HostSlot> slots;
    std::vector<void *> owned;     // every device allocation, freed at destroy
    int64_t ntiles = 0;
    int block = MCS_BLOCK;
    int compact_every = 8;
    int majorant = MCS_MAJORANT_GLOBAL;
    int64_t step_count = 0;
    bool dirty = true;
    hipGraphExec_t g_step = nullptr, g_compact = nullptr;
};

template <typename T>
static int32_t dalloc(mcs_swarm_batch *b, T **p, int64_t n)
{
    mcs_ctx *ctx = b->ctx;
    const size_t bytes = sizeof(T) * (size_t)(n > 0 ? n : 1);
    MCS_HIP(ctx, hipMalloc((void **)p, bytes));
    b->owned.push_back((void *)*p);
    MCS_HIP(ctx, mcs_zero(ctx, *p, bytes));
    return MCS_OK;
}

static int32_t alloc_store(mcs_swarm_batch *b, PopBuf &p, int64_t cap)
{
    int32_t rc = 0;
    for (int d = 0; d < 3; ++d) rc |= dalloc(b, &p.x[d], cap) | dalloc(b, &p.v[d], cap);
    rc |= dalloc(b, &p.w, cap) | dalloc(b, &p.s, cap) | dalloc(b, &p.active, cap) | dalloc(b, &p.id, cap);
    return rc ? MCS_E_HIP : MCS_OK;
}

static int32_t alloc_store(mcs_swarm_batch *b, SplitBuf &p, int64_t cap)
{
    int32_t rc = dalloc(b, &p.cell, cap);
    for (int d = 0; d < 3; ++d) rc |= dalloc(b, &p.off[d], cap) | dalloc(b, &p.v[d], cap);
    rc |= dalloc(b, &p.w, cap) | dalloc(b, &p.s, cap) | dalloc(b, &p.active, cap) | dalloc(b, &p.id, cap);
    return rc ? MCS_E_HIP : MCS_OK;
}

static void free_all(mcs_swarm_batch *b)
{
    for (void *p : b->owned) hipFree(p);
    b->owned.clear();
}

extern "C" int32_t mcs_swarm_batch_create(mcs_ctx *ctx, int64_t total_capacity, int32_t max_slots,
                                          uint64_t seed, int32_t precision, mcs_swarm_batch **out)
{
    if (!ctx || !out) return MCS_E_ARG;
    if (total_capacity < 1 || total_capacity >= ((int64_t)1 << 31) || max_slots < 1 || max_slots > 65536)
        return mcs_fail(ctx, MCS_E_ARG, "swarm batch: capacity must be in [1, 2^31), slots in [1, 65536]");
    if (precision != MCS_PREC_F64 && precision != MCS_PREC_F32)
        return mcs_fail(ctx, MCS_E_ARG, "swarm batch: unknown precision %d", precision);
    hipSetDevice(ctx->device);
    mcs_swarm_batch *b = new mcs_swarm_batch();
    b->ctx = ctx;
    b->precision = precision;
    b->total_cap = total_capacity;
    b->max_slots = max_slots;
    b->seed = seed;
    b->slots.resize(max_slots);
    int32_t rc = 0;
    if (precision == MCS_PREC_F64) {
        rc |= alloc_store(b, b->buf, total_capacity) | alloc_store(b, b->alt, total_capacity);
        SwLog &L = b->L;
        L.cap = total_capacity;
        rc |= dalloc(b, &L.parent, L.cap) | dalloc(b, &L.seq, L.cap) | dalloc(b, &L.s, L.cap) |
              dalloc(b, &L.act, L.cap);
        for (int d = 0; d < 3; ++d) rc |= dalloc(b, &L.x[d], L.cap) | dalloc(b, &L.v[d], L.cap);
    } else {
        rc |= alloc_store(b, b->buf32, total_capacity) | alloc_store(b, b->alt32, total_capacity);
        SwLog32 &L = b->L32;
        L.cap = total_capacity;
        rc |= dalloc(b, &L.parent, L.cap) | dalloc(b, &L.seq, L.cap) | dalloc(b, &L.cell, L.cap) |
              dalloc(b, &L.s, L.cap) | dalloc(b, &L.act, L.cap);
        for (int d = 0; d < 3; ++d) rc |= dalloc(b, &L.off[d], L.cap) | dalloc(b, &L.v[d], L.cap);
    }
    SlotDev &S = b->S;
    int64_t ms = max_slots;
    rc |= dalloc(b, &S.base, ms) | dalloc(b, &S.cap, ms) | dalloc(b, &S.n, ms) | dalloc(b, &S.n0, ms);
    rc |= dalloc(b, &S.used, ms) | dalloc(b, &S.flags, ms) | dalloc(b, &S.mode, ms) | dalloc(b, &S.relax, ms) |
          dalloc(b, &S.block, ms);
    rc |= dalloc(b, &S.e, 3 * ms) | dalloc(b, &S.dt, ms) | dalloc(b, &S.tol, ms) | dalloc(b, &S.param, ms);
    rc |= dalloc(b, &S.key, ms) | dalloc(b, &S.step, ms) | dalloc(b, &S.next_id, ms);
    rc |= dalloc(b, &S.tv, ms) | dalloc(b, &S.tv32, ms) | dalloc(b, &S.first_tile, ms) | dalloc(b, &S.ntiles, ms);
    rc |= dalloc(b, &S.sum, SW_NQ * ms) | dalloc(b, &S.sumc, SW_NQ * ms) | dalloc(b, &S.cur, SW_NQ * ms) |
          dalloc(b, &S.bm1, SW_NQ * ms) | dalloc(b, &S.bm2, SW_NQ * ms) | dalloc(b, &S.last, SW_NQ * ms);
    rc |= dalloc(b, &S.steps, ms) | dalloc(b, &S.nrec, ms) | dalloc(b, &S.nblk, ms) | dalloc(b, &S.curcnt, ms) |
          dalloc(b, &S.nact, ms);
    rc |= dalloc(b, &S.ring, (int64_t)SW_WIN * 3 * ms) | dalloc(b, &S.diff, 3 * ms);
    rc |= dalloc(b, &b->log_count, 8);
    rc |= dalloc(b, &b->loff, total_capacity + 512) | dalloc(b, &b->cnt, total_capacity + 512);
    b->max_tiles = total_capacity / 128 + 2 * (int64_t)max_slots + 2;
    rc |= dalloc(b, &b->tile_slot, b->max_tiles) | dalloc(b, &b->tile_start, b->max_tiles);
    rc |= dalloc(b, &b->tile_births, b->max_tiles) | dalloc(b, &b->tile_bpre, b->max_tiles + 1);
    rc |= dalloc(b, &b->tile_clones, b->max_tiles) | dalloc(b, &b->tile_cpre, b->max_tiles + 1);
    rc |= dalloc(b, &b->tile_act, b->max_tiles) | dalloc(b, &b->tile_apre, b->max_tiles + 1);
    rc |= dalloc(b, &b->scan_scratch, b->max_tiles / MCS_SCAN_TILE + 8);
    rc |= dalloc(b, &b->tile_part, b->max_tiles * 2 * SW_NT);
    rc |= dalloc(b, &b->tile_xpart, b->max_tiles * 2 * SW_NX);
    if (rc) {
        free_all(b);
        delete b;
        return mcs_fail(ctx, MCS_E_HIP, "swarm batch: device allocation failed");
    }
    *out = b;
    return MCS_OK;
}

static void drop_graphs(mcs_swarm_batch *b)
{
    if (b->g_step) hipGraphExecDestroy(b->g_step);
    if (b->g_compact) hipGraphExecDestroy(b->g_compact);
    b->g_step = nullptr;
    b->g_compact = nullptr;
}

extern "C" int32_t mcs_swarm_batch_destroy(mcs_swarm_batch *b)
{
    if (!b) return MCS_OK;
    hipSetDevice(b->ctx->device);
    hipStreamSynchronize(b->ctx->stream);
    drop_graphs(b);
    free_all(b);
    delete b;
    return MCS_OK;
}

// Tiles of `block` particles over every used slab, in increasing slab base, so
// that tile_start is sorted (k_sw_births finds a parent's tile by bisection) and
// each slot's tiles are consecutive.
static int32_t rebuild_tiles(mcs_swarm_batch *b)
{
    mcs_ctx *ctx = b->ctx;
    std::vector<int32_t> ts;
    std::vector<int64_t> tst;
    std::vector<int32_t> first(b->max_slots, 0), nt(b->max_slots, 0);
    std::vector<int> order;
    for (int s = 0; s < b->max_slots; ++s)
        if (b->slots[s].used) order.push_back(s);
    std::sort(order.begin(), order.end(),
              [&](int p, int q) { return b->slots[p].base < b->slots[q].base; });
    for (int s : order) {
        const HostSlot &h = b->slots[s];
        first[s] = (int32_t)ts.size();
        for (int64_t o = 0; o < h.cap; o += b->block) {
            ts.push_back(s);
            tst.push_back(h.base + o);
        }
        nt[s] = (int32_t)ts.size() - first[s];
    }
    b->ntiles = (int64_t)ts.size();
    if (b->ntiles > b->max_tiles) return mcs_fail(ctx, MCS_E_ARG, "swarm batch: too many tiles");
    if (b->ntiles) {
        MCS_HIP(ctx, mcs_h2d(ctx, b->tile_slot, ts.data(), sizeof(int32_t) * ts.size()));
        MCS_HIP(ctx, mcs_h2d(ctx, b->tile_start, tst.data(), sizeof(int64_t) * tst.size()));
    }
    MCS_HIP(ctx, mcs_h2d(ctx, b->S.first_tile, first.data(), sizeof(int32_t) * first.size()));
    MCS_HIP(ctx, mcs_h2d(ctx, b->S.ntiles, nt.data(), sizeof(int32_t) * nt.size()));
    return MCS_OK;
}

static int32_t ensure_stage(mcs_swarm_batch *b, int64_t n)
{
    if (n <= b->stage_cap) return MCS_OK;
    mcs_ctx *ctx = b->ctx;
    if (b->stage) {
        hipFree(b->stage);
        b->owned.erase(std::find(b->owned.begin(), b->owned.end(), (void *)b->stage));
        b->stage = nullptr;
    }
    int64_t cap = std::max<int64_t>(n, 1 << 16);
    MCS_HIP(ctx, hipMalloc((void **)&b->stage, sizeof(double) * 9 * (size_t)cap));
    b->owned.push_back(b->stage);
    b->stage_cap = cap;
    return MCS_OK;
}

extern "C" int32_t mcs_swarm_batch_add(mcs_swarm_batch *b, const mcs_table *t,
                                       const mcs_swarm_params *prm, int64_t n, const double *x,
                                       const double *v, const double *w, const double *s,
                                       const uint8_t *active, int32_t *slot_out)
{
    if (!b || !t || !prm || !slot_out) return MCS_E_ARG;
    mcs_ctx *ctx = b->ctx;
    hipSetDevice(ctx->device);
    if (prm->capacity < 1 || n < 0 || n > prm->capacity || prm->n0 < 1 || !(prm->dt >= 0.0) ||
        prm->block_steps < 1 || prm->relax_steps < 0)
        return mcs_fail(ctx, MCS_E_ARG, "swarm batch: bad instance parameters");
    if (t->nphoto > 0) return mcs_fail(ctx, MCS_E_ARG, "swarm batch: tables with photoemission rows need photons");
    const bool dev_init = (x == nullptr);
    if (n > 0 && !dev_init && (!v || !w || !s || !active))
        return mcs_fail(ctx, MCS_E_ARG, "swarm batch: null array");
    if (dev_init && !(prm->init_vth >= 0.0)) return mcs_fail(ctx, MCS_E_ARG, "swarm batch: bad init_vth");
    int slot = -1;
    for (int k = 0; k < b->max_slots; ++k)
        if (!b->slots[k].used) {
            slot = k;
            break;
        }
    if (slot < 0) return mcs_fail(ctx, MCS_E_CAPACITY, "swarm batch: no free slot");
    // first fit: slabs are rounded to whole tiles
    int64_t cap = cdiv(prm->capacity, 512) * 512;
    std::vector<std::pair<int64_t, int64_t>> used;
    for (auto &h : b->slots)
        if (h.used) used.emplace_back(h.base, h.base + h.cap);
    std::sort(used.begin(), used.end());
    int64_t base = -1, at = 0;
    for (auto &u : used) {
        if (u.first - at >= cap) {
            base = at;
            break;
        }
        at = std::max(at, u.second);
    }
    if (base < 0 && b->total_cap - at >= cap) base = at;
    if (base < 0) return mcs_fail(ctx, MCS_E_CAPACITY, "swarm batch: no free slab of %lld particles", (long long)cap);
    const unsigned grid = (unsigned)std::max<int64_t>(cdiv(n, MCS_BLOCK), 1);
    // particles
    if (dev_init) {
        if (n > 0) {
            if (b->precision == MCS_PREC_F64)
                hipLaunchKernelGGL(k_sw_init, dim3(grid), dim3(MCS_BLOCK), 0, ctx->stream, b->buf, base, n,
                                   b->seed, prm->instance, prm->init_vth);
            else
                hipLaunchKernelGGL(k_sw_init32, dim3(grid), dim3(MCS_BLOCK), 0, ctx->stream, b->buf32, base, n,
                                   b->seed, prm->instance, prm->init_vth);
            MCS_LAUNCH_CHECK(ctx);
        }
    } else if (b->precision == MCS_PREC_F64) {
        MCS_HIP(ctx, hipStreamSynchronize(ctx->stream));
        std::vector<double> tmp((size_t)n);
        for (int d = 0; d < 3; ++d) {
            for (int64_t i = 0; i < n; ++i) tmp[i] = x[3 * i + d];
            MCS_HIP(ctx, mcs_h2d(ctx, b->buf.x[d] + base, tmp.data(), sizeof(double) * n));
            for (int64_t i = 0; i < n; ++i) tmp[i] = v[3 * i + d];
            MCS_HIP(ctx, mcs_h2d(ctx, b->buf.v[d] + base, tmp.data(), sizeof(double) * n));
        }
        MCS_HIP(ctx, mcs_h2d(ctx, b->buf.w + base, w, sizeof(double) * n));
        MCS_HIP(ctx, mcs_h2d(ctx, b->buf.s + base, s, sizeof(double) * n));
        MCS_HIP(ctx, mcs_h2d(ctx, b->buf.active + base, active, n));
        std::vector<uint64_t> ids((size_t)n);
        for (int64_t i = 0; i < n; ++i) ids[i] = (uint64_t)i;
        MCS_HIP(ctx, mcs_h2d(ctx, b->buf.id + base, ids.data(), sizeof(uint64_t) * n));
    } else if (n > 0) {
        MCS_HIP(ctx, hipStreamSynchronize(ctx->stream));
        int32_t rc = ensure_stage(b, n);
        if (rc) return rc;
        double *dx = b->stage, *dv = dx + 3 * n, *dw = dv + 3 * n, *ds = dw + n;
        uint8_t *da = (uint8_t *)(ds + n);
        MCS_HIP(ctx, mcs_h2d(ctx, dx, x, sizeof(double) * 3 * n));
        MCS_HIP(ctx, mcs_h2d(ctx, dv, v, sizeof(double) * 3 * n));
        MCS_HIP(ctx, mcs_h2d(ctx, dw, w, sizeof(double) * n));
        MCS_HIP(ctx, mcs_h2d(ctx, ds, s, sizeof(double) * n));
        MCS_HIP(ctx, mcs_h2d(ctx, da, active, n));
        hipLaunchKernelGGL(k_sw_from64, dim3(grid), dim3(MCS_BLOCK), 0, ctx->stream, b->buf32, base, n,
                           (const double *)dx, (const double *)dv, (const double *)dw, (const double *)ds,
                           (const uint8_t *)da);
        MCS_LAUNCH_CHECK(ctx);
        MCS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    // slot state
    SlotInit p;
    p.base = base;
    p.cap = cap;
    p.n = n;
    p.n0 = prm->n0;
    p.relax = prm->relax_steps;
    p.block = prm->block_steps;
    for (int d = 0; d < 3; ++d) p.e[d] = prm->e[d];
    p.dt = prm->dt;
    p.tol = prm->tol;
    p.key = prm->instance;
    p.next_id = (uint64_t)n;
    p.tv = table_view(t);
    p.tv32 = table_view32(t);
    hipLaunchKernelGGL(k_sw_slot_init, dim3(1), dim3(64), 0, ctx->stream, b->S, slot, p);
    MCS_LAUNCH_CHECK(ctx);
    HostSlot &h = b->slots[slot];
    h.used = true;
    h.base = base;
    h.cap = cap;
    h.table = t;
    h.params = *prm;
    b->dirty = true;
    *slot_out = slot;
    return MCS_OK;
}

extern "C" int32_t mcs_swarm_batch_remove(mcs_swarm_batch *b, int32_t slot)
{
    if (!b || slot < 0 || slot >= b->max_slots || !b->slots[slot].used) return MCS_E_ARG;
    mcs_ctx *ctx = b->ctx;
    hipSetDevice(ctx->device);
    hipLaunchKernelGGL(k_sw_slot_free, dim3(1), dim3(64), 0, ctx->stream, b->S, slot);
    MCS_LAUNCH_CHECK(ctx);
    b->slots[slot].used = false;
    b->dirty = true;
    return MCS_OK;
}

extern "C" int32_t mcs_swarm_batch_config(mcs_swarm_batch *b, int32_t block, int32_t compact_every)
{
    if (!b || (block != 128 && block != 256) || compact_every < 0) return MCS_E_ARG;
    b->block = block;
    b->compact_every = compact_every;
    b->dirty = true;
    return MCS_OK;
}

extern "C" int32_t mcs_swarm_batch_majorant(mcs_swarm_batch *b, int32_t mode)
{
    if (!b || (mode != MCS_MAJORANT_GLOBAL && mode != MCS_MAJORANT_BANDED)) return MCS_E_ARG;
    if (mode == MCS_MAJORANT_BANDED && b->precision != MCS_PREC_F32)
        return mcs_fail(b->ctx, MCS_E_ARG, "swarm batch: the banded majorant is for FP32 batches (an FP64 batch "
                                           "stays draw for draw with a lone FP64 population)");
    b->majorant = mode;
    b->dirty = true;
    return MCS_OK;
}

template <typename A>
static void fill_common(mcs_swarm_batch *b, A &a)
{
    a.S = b->S;
    a.tile_slot = b->tile_slot;
    a.tile_start = b->tile_start;
    a.seed = b->seed;
    a.log_count = b->log_count;
    a.loff = b->loff;
    a.cnt = b->cnt;
    a.tile_births = b->tile_births;
    a.tile_bpre = b->tile_bpre;
    a.tile_clones = b->tile_clones;
    a.tile_cpre = b->tile_cpre;
    a.tile_act = b->tile_act;
    a.tile_apre = b->tile_apre;
    a.tile_part = b->tile_part;
    a.tile_xpart = b->tile_xpart;
}

static SwArgs64 args64(mcs_swarm_batch *b)
{
    SwArgs64 a;
    a.b = b->buf;
    a.alt = b->alt;
    a.L = b->L;
    fill_common(b, a);
    return a;
}

static SwArgs32 args32(mcs_swarm_batch *b)
{
    SwArgs32 a;
    a.b = b->buf32;
    a.alt = b->alt32;
    a.L = b->L32;
    fill_common(b, a);
    return a;
}

template <int BS>
static void launch_advance(mcs_swarm_batch *b, const SwArgs64 &a, dim3 g)
{
    hipLaunchKernelGGL(k_sw_advance<BS>, g, dim3(BS), 0, b->ctx->stream, a);
}

template <int BS>
static void launch_advance(mcs_swarm_batch *b, const SwArgs32 &a, dim3 g)
{
    if (b->majorant == MCS_MAJORANT_BANDED)
        hipLaunchKernelGGL((k_sw_advance32<BS, 1>), g, dim3(BS), 0, b->ctx->stream, a);
    else
        hipLaunchKernelGGL((k_sw_advance32<BS, 0>), g, dim3(BS), 0, b->ctx->stream, a);
}

template <int BS, typename A>
static int32_t launch_step_t(mcs_swarm_batch *b, const A &a)
{
    mcs_ctx *ctx = b->ctx;
    const dim3 g((unsigned)b->ntiles);
    hipLaunchKernelGGL(k_sw_zero, dim3(1), dim3(64), 0, ctx->stream, b->log_count, 8);
    launch_advance<BS>(b, a, g);
    int32_t rc = mcs_scan_i64_buf(ctx, b->tile_births, b->tile_bpre, b->ntiles, b->scan_scratch);
    if (rc) return rc;
    const int64_t logcap = b->precision == MCS_PREC_F64 ? b->L.cap : b->L32.cap;
    hipLaunchKernelGGL(k_sw_births<A>, dim3((unsigned)std::min<int64_t>(cdiv(logcap, MCS_BLOCK), 2048)),
                       dim3(MCS_BLOCK), 0, ctx->stream, a, logcap, b->ntiles);
    int64_t maxnt = 1;
    for (const HostSlot &h : b->slots)
        if (h.used) maxnt = std::max<int64_t>(maxnt, cdiv(h.cap, (int64_t)BS));
    const int64_t nchunk = cdiv(maxnt, (int64_t)SW_XCHUNK);
    if (nchunk > 65535) return mcs_fail(ctx, MCS_E_ARG, "swarm batch: a slab of %lld particles is too large",
                                        (long long)(maxnt * BS));
    hipLaunchKernelGGL((k_sw_xmom<BS, A>), dim3((unsigned)b->max_slots, (unsigned)nchunk), dim3(BS), 0,
                       ctx->stream, a);
    hipLaunchKernelGGL(k_sw_slot_step<A>, dim3((unsigned)b->max_slots), dim3(MCS_BLOCK), 0, ctx->stream, a);
    hipLaunchKernelGGL((k_sw_resample<BS, A>), g, dim3(BS), 0, ctx->stream, a);
    rc = mcs_scan_i64_buf(ctx, b->tile_clones, b->tile_cpre, b->ntiles, b->scan_scratch);
    if (rc) return rc;
    hipLaunchKernelGGL((k_sw_clone_place<BS, A>), g, dim3(BS), 0, ctx->stream, a);
    hipLaunchKernelGGL(k_sw_slot_step2<A>, dim3((unsigned)cdiv(b->max_slots, 64)), dim3(64), 0, ctx->stream, a,
                       b->max_slots);
    MCS_LAUNCH_CHECK(ctx);
    return MCS_OK;
}

template <int BS, typename A>
static int32_t launch_compact_t(mcs_swarm_batch *b, const A &a)
{
    mcs_ctx *ctx = b->ctx;
    const dim3 g((unsigned)b->ntiles);
    hipLaunchKernelGGL((k_sw_compact_count<BS, A>), g, dim3(BS), 0, ctx->stream, a);
    int32_t rc = mcs_scan_i64_buf(ctx, b->tile_act, b->tile_apre, b->ntiles, b->scan_scratch);
    if (rc) return rc;
    hipLaunchKernelGGL((k_sw_compact_write<BS, A>), g, dim3(BS), 0, ctx->stream, a);
    hipLaunchKernelGGL(k_sw_compact_slot<A>, dim3((unsigned)cdiv(b->max_slots, 64)), dim3(64), 0, ctx->stream,
                       a, b->max_slots);
    hipLaunchKernelGGL((k_sw_copy_back<BS, A>), g, dim3(BS), 0, ctx->stream, a);
    MCS_LAUNCH_CHECK(ctx);
    return MCS_OK;
}

static int32_t launch_step(mcs_swarm_batch *b)
{
    if (b->precision == MCS_PREC_F64)
        return b->block == 128 ? launch_step_t<128>(b, args64(b)) : launch_step_t<256>(b, args64(b));
    return b->block == 128 ? launch_step_t<128>(b, args32(b)) : launch_step_t<256>(b, args32(b));
}

static int32_t launch_compact(mcs_swarm_batch *b)
{
    if (b->precision == MCS_PREC_F64)
        return b->block == 128 ? launch_compact_t<128>(b, args64(b)) : launch_compact_t<256>(b, args64(b));
    return b->block == 128 ? launch_compact_t<128>(b, args32(b)) : launch_compact_t<256>(b, args32(b));
}

static int32_t capture(mcs_swarm_batch *b, int32_t (*fn)(mcs_swarm_batch *), hipGraphExec_t *out)
{
    mcs_ctx *ctx = b->ctx;
    hipGraph_t g = nullptr;
    MCS_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
    int32_t rc = fn(b);
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc) return rc;
    if (e != hipSuccess) return mcs_fail(ctx, MCS_E_HIP, "swarm batch: graph capture failed");
    MCS_HIP(ctx, hipGraphInstantiate(out, g, nullptr, nullptr, 0));
    hipGraphDestroy(g);
    return MCS_OK;
}

extern "C" int32_t mcs_swarm_batch_step(mcs_swarm_batch *b, int32_t nsteps)
{
    if (!b || nsteps < 0) return MCS_E_ARG;
    mcs_ctx *ctx = b->ctx;
    hipSetDevice(ctx->device);
    if (b->dirty) {
        MCS_HIP(ctx, hipStreamSynchronize(ctx->stream));
        drop_graphs(b);
        int32_t rc = rebuild_tiles(b);
        if (rc) return rc;
        b->dirty = false;
        if (b->ntiles > 0) {
            rc = capture(b, launch_step, &b->g_step);
            if (rc) return rc;
            rc = capture(b, launch_compact, &b->g_compact);
            if (rc) return rc;
        }
    }
    if (b->ntiles == 0) return MCS_OK;
    ProfScope ps(ctx, "swarm_batch");
    for (int k = 0; k < nsteps; ++k) {
        MCS_HIP(ctx, hipGraphLaunch(b->g_step, ctx->stream));
        b->step_count += 1;
        if (b->compact_every > 0 && b->step_count % b->compact_every == 0)
            MCS_HIP(ctx, hipGraphLaunch(b->g_compact, ctx->stream));
    }
    return MCS_OK;
}

extern "C" int32_t mcs_swarm_batch_tally(mcs_swarm_batch *b, int32_t slot, mcs_swarm_tally *out)
{
    if (!b || !out || slot < 0 || slot >= b->max_slots) return MCS_E_ARG;
    std::vector<mcs_swarm_tally> all((size_t)slot + 1);
    int32_t rc = mcs_swarm_batch_tally_all(b, slot + 1, all.data());
    if (rc) return rc;
    *out = all[slot];
    return MCS_OK;
}

extern "C" int32_t mcs_swarm_batch_tally_all(mcs_swarm_batch *b, int32_t nslots, mcs_swarm_tally *out)
{
    if (!b || !out || nslots < 0 || nslots > b->max_slots) return MCS_E_ARG;
    mcs_ctx *ctx = b->ctx;
    hipSetDevice(ctx->device);
    MCS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (nslots == 0) return MCS_OK;
    const SlotDev &S = b->S;
    std::vector<int64_t> steps(nslots), n(nslots), nact(nslots), nrec(nslots), nblk(nslots);
    std::vector<int32_t> flags(nslots);
    std::vector<double> sum(SW_NQ * nslots), sumc(SW_NQ * nslots), bm1(SW_NQ * nslots),
        bm2(SW_NQ * nslots), last(SW_NQ * nslots);
    MCS_HIP(ctx, hipMemcpy(steps.data(), S.steps, sizeof(int64_t) * nslots, hipMemcpyDeviceToHost));
    MCS_HIP(ctx, hipMemcpy(n.data(), S.n, sizeof(int64_t) * nslots, hipMemcpyDeviceToHost));
    MCS_HIP(ctx, hipMemcpy(nact.data(), S.nact, sizeof(int64_t) * nslots, hipMemcpyDeviceToHost));
    MCS_HIP(ctx, hipMemcpy(nrec.data(), S.nrec, sizeof(int64_t) * nslots, hipMemcpyDeviceToHost));
    MCS_HIP(ctx, hipMemcpy(nblk.data(), S.nblk, sizeof(int64_t) * nslots, hipMemcpyDeviceToHost));
    MCS_HIP(ctx, hipMemcpy(flags.data(), S.flags, sizeof(int32_t) * nslots, hipMemcpyDeviceToHost));
    MCS_HIP(ctx, hipMemcpy(sum.data(), S.sum, sizeof(double) * SW_NQ * nslots, hipMemcpyDeviceToHost));
    MCS_HIP(ctx, hipMemcpy(sumc.data(), S.sumc, sizeof(double) * SW_NQ * nslots, hipMemcpyDeviceToHost));
    MCS_HIP(ctx, hipMemcpy(bm1.data(), S.bm1, sizeof(double) * SW_NQ * nslots, hipMemcpyDeviceToHost));
    MCS_HIP(ctx, hipMemcpy(bm2.data(), S.bm2, sizeof(double) * SW_NQ * nslots, hipMemcpyDeviceToHost));
    MCS_HIP(ctx, hipMemcpy(last.data(), S.last, sizeof(double) * SW_NQ * nslots, hipMemcpyDeviceToHost));
    std::vector<double> diff(3 * (size_t)nslots);
    MCS_HIP(ctx, hipMemcpy(diff.data(), S.diff, sizeof(double) * 3 * nslots, hipMemcpyDeviceToHost));
    for (int k = 0; k < nslots; ++k) {
        mcs_swarm_tally &o = out[k];
        memset(&o, 0, sizeof(o));
        o.steps = steps[k];
        o.n = n[k];
        o.nactive = nact[k];
        o.nrecorded = nrec[k];
        o.nblocks = nblk[k];
        o.flags = flags[k];
        for (int q = 0; q < SW_NQ; ++q) {
            o.sum[q] = sum[SW_NQ * k + q] + sumc[SW_NQ * k + q];
            o.bm_s1[q] = bm1[SW_NQ * k + q];
            o.bm_s2[q] = bm2[SW_NQ * k + q];
            o.last[q] = last[SW_NQ * k + q];
        }
        for (int d = 0; d < 3; ++d) o.diffusion[d] = diff[3 * k + d];
        o.window = std::min<int64_t>(steps[k], SW_WIN);
    }
    return MCS_OK;
}

extern "C" int32_t mcs_swarm_batch_download(mcs_swarm_batch *b, int32_t slot, int64_t nmax,
                                            int64_t *nout, double *x, double *v, double *w, double *s,
                                            uint8_t *active, uint64_t *id)
{
    if (!b || slot < 0 || slot >= b->max_slots || !b->slots[slot].used) return MCS_E_ARG;
    mcs_ctx *ctx = b->ctx;
    hipSetDevice(ctx->device);
    MCS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    int64_t n;
    MCS_HIP(ctx, hipMemcpy(&n, b->S.n + slot, sizeof(int64_t), hipMemcpyDeviceToHost));
    if (nout) *nout = n;
    int64_t m = n < nmax ? n : nmax;
    if (m <= 0) return MCS_OK;
    const int64_t base = b->slots[slot].base;
    if (b->precision == MCS_PREC_F32) {
        int32_t rc = ensure_stage(b, m);
        if (rc) return rc;
        double *dx = b->stage, *dv = dx + 3 * m, *dw = dv + 3 * m, *ds = dw + m;
        hipLaunchKernelGGL(k_sw_to64, dim3((unsigned)cdiv(m, MCS_BLOCK)), dim3(MCS_BLOCK), 0, ctx->stream,
                           b->buf32, base, m, dx, dv, dw, ds);
        MCS_LAUNCH_CHECK(ctx);
        MCS_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (x) MCS_HIP(ctx, hipMemcpy(x, dx, sizeof(double) * 3 * m, hipMemcpyDeviceToHost));
        if (v) MCS_HIP(ctx, hipMemcpy(v, dv, sizeof(double) * 3 * m, hipMemcpyDeviceToHost));
        if (w) MCS_HIP(ctx, hipMemcpy(w, dw, sizeof(double) * m, hipMemcpyDeviceToHost));
        if (s) MCS_HIP(ctx, hipMemcpy(s, ds, sizeof(double) * m, hipMemcpyDeviceToHost));
        if (active) MCS_HIP(ctx, hipMemcpy(active, b->buf32.active + base, m, hipMemcpyDeviceToHost));
        if (id) MCS_HIP(ctx, hipMemcpy(id, b->buf32.id + base, sizeof(uint64_t) * m, hipMemcpyDeviceToHost));
        return MCS_OK;
    }
    std::vector<double> tmp((size_t)m);
    for (int d = 0; d < 3; ++d) {
        if (x) {
            MCS_HIP(ctx, hipMemcpy(tmp.data(), b->buf.x[d] + base, sizeof(double) * m, hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < m; ++i) x[3 * i + d] = tmp[i];
        }
        if (v) {
            MCS_HIP(ctx, hipMemcpy(tmp.data(), b->buf.v[d] + base, sizeof(double) * m, hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < m; ++i) v[3 * i + d] = tmp[i];
        }
    }
    if (w) MCS_HIP(ctx, hipMemcpy(w, b->buf.w + base, sizeof(double) * m, hipMemcpyDeviceToHost));
    if (s) MCS_HIP(ctx, hipMemcpy(s, b->buf.s + base, sizeof(double) * m, hipMemcpyDeviceToHost));
    if (active) MCS_HIP(ctx, hipMemcpy(active, b->buf.active + base, m, hipMemcpyDeviceToHost));
    if (id) MCS_HIP(ctx, hipMemcpy(id, b->buf.id + base, sizeof(uint64_t) * m, hipMemcpyDeviceToHost));
    return MCS_OK;
}

extern "C" int32_t mcs_pop_init_thermal(mcs_pop *p, int64_t n, uint64_t seed, uint32_t instance, double vth)
{
    if (!p) return MCS_E_ARG;
    MCS_ELECTRONS_ONLY(p, "pop_init_thermal");
    mcs_ctx *ctx = p->ctx;
    hipSetDevice(ctx->device);
    if (n < 0 || n > p->cap || !(vth >= 0.0))
        return mcs_fail(ctx, MCS_E_ARG, "pop_init_thermal: n must be in [0, capacity], vth >= 0");
    mcs_pop_touch(p);
    if (n > 0) {
        hipLaunchKernelGGL(k_sw_init, dim3((unsigned)cdiv(n, MCS_BLOCK)), dim3(MCS_BLOCK), 0, ctx->stream,
                           p->buf[0], (int64_t)0, n, seed, instance, vth);
        MCS_LAUNCH_CHECK(ctx);
    }
    p->n = n;
    p->next_id = (uint64_t)n;
    return MCS_OK;
}
