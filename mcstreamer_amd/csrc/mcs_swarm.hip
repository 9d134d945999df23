// mcs_swarm.hip -- many Swarm.swarm instances (Swarm/src/Swarm.jl:102-142) in one
// set of launches.
//
// Layout: one particle store; slot s owns the slab [base_s, base_s + cap_s).
// Work is cut into tiles of one workgroup, each tile inside one slab, so every
// per-slot quantity (births, clones, compaction offsets, tallies) is a scan or a
// reduction over the slot's consecutive tiles.  Every kernel runs on a grid that
// depends only on the set of slots, never on particle counts (those live on the
// device), so a step is captured once into a HIP graph and replayed.
#include <algorithm>
#include <cstring>
#include <vector>

#include "mcs_advance.h"

#define SW_NQ MCS_TALLY_NQ
#define SW_NT (SW_NQ + 1)     // count + the tallied quantities
#define SW_MODE_NONE 0
#define SW_MODE_DROP 1
#define SW_MODE_CLONE 2

struct SlotDev {
    int64_t *base, *cap, *n, *n0;
    int32_t *used, *flags, *mode, *relax, *block;
    double *e, *dt, *tol, *param;
    uint32_t *key, *step;
    uint64_t *next_id;
    TableView *tv;
    int32_t *first_tile, *ntiles;
    // tallies: per slot SW_NQ each
    double *sum, *sumc, *cur, *bm1, *bm2, *last;
    int64_t *steps, *nrec, *nblk, *curcnt, *nact;
};

struct SwLog {
    int64_t *parent;
    int32_t *seq;
    double *x[3], *v[3];
    double *s;
    uint8_t *act;
    int64_t cap;
};

struct SwArgs {
    PopBuf b, alt;
    SlotDev S;
    const int32_t *tile_slot;
    const int64_t *tile_start;
    uint64_t seed;
    SwLog L;
    int64_t *log_count;
    int32_t *loff, *cnt;
    int64_t *tile_births, *tile_bpre, *tile_clones, *tile_cpre, *tile_act, *tile_apre;
    double *tile_part;   // ntiles x SW_NT x 2 (Neumaier sum, compensation)
};

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
        int kd = T.kind[j];
        if (kd == MCS_PROC_IONIZATION) ion += w * r0[j] + (1.0 - w) * r1[j];
        else if (kd == MCS_PROC_ATTACHMENT) att += w * r0[j] + (1.0 - w) * r1[j];
    }
    nadd(acc[0], 1.0);
    nadd(acc[1], v[2]);
    nadd(acc[2], eng);
    nadd(acc[3], ion);
    nadd(acc[4], att);
}

// Block-reduce SW_NT Neumaier pairs into tile_part[tile].
template <int BS>
__device__ __forceinline__ void tile_reduce(Neu2 acc[SW_NT], double *out)
{
    __shared__ double sh[BS], shc[BS];
#pragma unroll
    for (int q = 0; q < SW_NT; ++q) {
        sh[threadIdx.x] = acc[q].s;
        shc[threadIdx.x] = acc[q].c;
        __syncthreads();
        for (int o = BS / 2; o > 0; o >>= 1) {
            if (threadIdx.x < o) {
                Neu2 a{sh[threadIdx.x], shc[threadIdx.x]};
                Neu2 b{sh[threadIdx.x + o], shc[threadIdx.x + o]};
                nmerge(a, b);
                sh[threadIdx.x] = a.s;
                shc[threadIdx.x] = a.c;
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            out[2 * q] = sh[0];
            out[2 * q + 1] = shc[0];
        }
        __syncthreads();
    }
}

// advance! of every slot's particles (JuMC/src/mixed_population.jl:24-54), births
// to the log, and the post-step tallies of parents and children.
template <int BS>
__global__ __launch_bounds__(BS) void k_sw_advance(SwArgs a)
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
            [&](int) {});
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

// Per slot: births -> n, next_id; tallies of this step; resample decision of
// Swarm.resample! (Swarm.jl:194-199).  One workgroup per slot.
__global__ __launch_bounds__(MCS_BLOCK) void k_sw_slot_step(SwArgs a)
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
template <int BS>
__global__ __launch_bounds__(BS) void k_sw_resample(SwArgs a)
{
    const int64_t tile = blockIdx.x;
    const int slot = a.tile_slot[tile];
    const int64_t i = a.tile_start[tile] + threadIdx.x;
    const int64_t lim = a.S.base[slot] + a.S.n[slot];
    const int mode = a.S.mode[slot];
    int c = 0;
    if (mode != SW_MODE_NONE && i < lim && a.b.active[i]) {
        mcs_stream rs;
        mcs_stream_init(&rs, a.seed, a.S.key[slot], MCS_PHASE_SWARM_RESAMPLE, a.b.id[i], a.S.step[slot]);
        if (mode == SW_MODE_DROP) {
            if (mcs_u01(&rs) < a.S.param[slot]) a.b.active[i] = 0;
        } else {
            c = (int)mcs_poisson(&rs, a.S.param[slot]);
        }
    }
    a.cnt[i] = c;
    int tot;
    int ex = bscan<BS, int>(c, &tot);
    a.loff[i] = ex;
    if (threadIdx.x == 0) a.tile_clones[tile] = tot;
}

template <int BS>
__global__ __launch_bounds__(BS) void k_sw_clone_place(SwArgs a)
{
    const int64_t tile = blockIdx.x;
    const int slot = a.tile_slot[tile];
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
        copy_particle(a.b, i, a.b, d);
        a.b.id[d] = nid + (uint64_t)rank;
    }
}

__global__ void k_sw_slot_step2(SwArgs a, int nslots)
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
template <int BS>
__global__ __launch_bounds__(BS) void k_sw_compact_count(SwArgs a)
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

template <int BS>
__global__ __launch_bounds__(BS) void k_sw_compact_write(SwArgs a)
{
    const int64_t tile = blockIdx.x;
    const int slot = a.tile_slot[tile];
    const int64_t i = a.tile_start[tile] + threadIdx.x;
    const int64_t lim = a.S.base[slot] + a.S.n[slot];
    if (!(i < lim && a.b.active[i])) return;
    const int64_t f = a.S.first_tile[slot];
    int64_t d = a.S.base[slot] + (a.tile_apre[tile] - a.tile_apre[f]) + a.loff[i];
    copy_particle(a.b, i, a.alt, d);
}

__global__ void k_sw_compact_slot(SwArgs a, int nslots)
{
    const int slot = gidx();
    if (slot >= nslots || !a.S.used[slot]) return;
    const int64_t f = a.S.first_tile[slot], nt = a.S.ntiles[slot];
    a.S.n[slot] = a.tile_apre[f + nt] - a.tile_apre[f];
}

template <int BS>
__global__ __launch_bounds__(BS) void k_sw_copy_back(SwArgs a)
{
    const int64_t tile = blockIdx.x;
    const int slot = a.tile_slot[tile];
    const int64_t i = a.tile_start[tile] + threadIdx.x;
    if (i < a.S.base[slot] + a.S.n[slot]) copy_particle(a.alt, i, a.b, i);
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

__global__ void k_sw_zero(int64_t *p, int n)
{
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
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
    int64_t total_cap;
    int max_slots;
    uint64_t seed;
    PopBuf buf, alt;
    SlotDev S;
    SwLog L;
    int64_t *log_count = nullptr;
    int32_t *loff = nullptr, *cnt = nullptr;
    int64_t max_tiles = 0;
    int32_t *tile_slot = nullptr;
    int64_t *tile_start = nullptr;
    int64_t *tile_births = nullptr, *tile_bpre = nullptr, *tile_clones = nullptr, *tile_cpre = nullptr;
    int64_t *tile_act = nullptr, *tile_apre = nullptr;
    int64_t *scan_scratch = nullptr;
    double *tile_part = nullptr;
    std::vector<HostSlot> slots;
    int64_t ntiles = 0;
    int block = MCS_BLOCK;
    int compact_every = 8;
    int64_t step_count = 0;
    bool dirty = true;
    hipGraphExec_t g_step = nullptr, g_compact = nullptr;
};

template <typename T>
static int32_t dalloc(mcs_ctx *ctx, T **p, int64_t n)
{
    MCS_HIP(ctx, hipMalloc((void **)p, sizeof(T) * (size_t)(n > 0 ? n : 1)));
    MCS_HIP(ctx, hipMemset(*p, 0, sizeof(T) * (size_t)(n > 0 ? n : 1)));
    return MCS_OK;
}

static int32_t alloc_popbuf(mcs_ctx *ctx, PopBuf &b, int64_t cap)
{
    int32_t rc = 0;
    for (int d = 0; d < 3 && !rc; ++d) {
        rc |= dalloc(ctx, &b.x[d], cap);
        rc |= dalloc(ctx, &b.v[d], cap);
    }
    rc |= dalloc(ctx, &b.w, cap);
    rc |= dalloc(ctx, &b.s, cap);
    rc |= dalloc(ctx, &b.active, cap);
    rc |= dalloc(ctx, &b.id, cap);
    return rc ? MCS_E_HIP : MCS_OK;
}

static void free_popbuf(PopBuf &b)
{
    for (int d = 0; d < 3; ++d) {
        hipFree(b.x[d]);
        hipFree(b.v[d]);
    }
    hipFree(b.w);
    hipFree(b.s);
    hipFree(b.active);
    hipFree(b.id);
}

extern "C" int32_t mcs_swarm_batch_create(mcs_ctx *ctx, int64_t total_capacity, int32_t max_slots,
                                          uint64_t seed, mcs_swarm_batch **out)
{
    if (!ctx || !out) return MCS_E_ARG;
    if (total_capacity < 1 || total_capacity >= ((int64_t)1 << 31) || max_slots < 1 || max_slots > 65536)
        return mcs_fail(ctx, MCS_E_ARG, "swarm batch: capacity must be in [1, 2^31), slots in [1, 65536]");
    hipSetDevice(ctx->device);
    mcs_swarm_batch *b = new mcs_swarm_batch();
    b->ctx = ctx;
    b->total_cap = total_capacity;
    b->max_slots = max_slots;
    b->seed = seed;
    b->slots.resize(max_slots);
    int32_t rc = alloc_popbuf(ctx, b->buf, total_capacity);
    rc |= alloc_popbuf(ctx, b->alt, total_capacity);
    SlotDev &S = b->S;
    int64_t ms = max_slots;
    rc |= dalloc(ctx, &S.base, ms) | dalloc(ctx, &S.cap, ms) | dalloc(ctx, &S.n, ms) | dalloc(ctx, &S.n0, ms);
    rc |= dalloc(ctx, &S.used, ms) | dalloc(ctx, &S.flags, ms) | dalloc(ctx, &S.mode, ms) |
          dalloc(ctx, &S.relax, ms) | dalloc(ctx, &S.block, ms);
    rc |= dalloc(ctx, &S.e, 3 * ms) | dalloc(ctx, &S.dt, ms) | dalloc(ctx, &S.tol, ms) | dalloc(ctx, &S.param, ms);
    rc |= dalloc(ctx, &S.key, ms) | dalloc(ctx, &S.step, ms) | dalloc(ctx, &S.next_id, ms);
    rc |= dalloc(ctx, &S.tv, ms) | dalloc(ctx, &S.first_tile, ms) | dalloc(ctx, &S.ntiles, ms);
    rc |= dalloc(ctx, &S.sum, SW_NQ * ms) | dalloc(ctx, &S.sumc, SW_NQ * ms) | dalloc(ctx, &S.cur, SW_NQ * ms) |
          dalloc(ctx, &S.bm1, SW_NQ * ms) | dalloc(ctx, &S.bm2, SW_NQ * ms) | dalloc(ctx, &S.last, SW_NQ * ms);
    rc |= dalloc(ctx, &S.steps, ms) | dalloc(ctx, &S.nrec, ms) | dalloc(ctx, &S.nblk, ms) |
          dalloc(ctx, &S.curcnt, ms) | dalloc(ctx, &S.nact, ms);
    SwLog &L = b->L;
    L.cap = total_capacity;
    rc |= dalloc(ctx, &L.parent, L.cap) | dalloc(ctx, &L.seq, L.cap) | dalloc(ctx, &L.s, L.cap) |
          dalloc(ctx, &L.act, L.cap);
    for (int d = 0; d < 3; ++d) rc |= dalloc(ctx, &L.x[d], L.cap) | dalloc(ctx, &L.v[d], L.cap);
    rc |= dalloc(ctx, &b->log_count, 8);
    rc |= dalloc(ctx, &b->loff, total_capacity + 512) | dalloc(ctx, &b->cnt, total_capacity + 512);
    b->max_tiles = total_capacity / 128 + 2 * (int64_t)max_slots + 2;
    rc |= dalloc(ctx, &b->tile_slot, b->max_tiles) | dalloc(ctx, &b->tile_start, b->max_tiles);
    rc |= dalloc(ctx, &b->tile_births, b->max_tiles) | dalloc(ctx, &b->tile_bpre, b->max_tiles + 1);
    rc |= dalloc(ctx, &b->tile_clones, b->max_tiles) | dalloc(ctx, &b->tile_cpre, b->max_tiles + 1);
    rc |= dalloc(ctx, &b->tile_act, b->max_tiles) | dalloc(ctx, &b->tile_apre, b->max_tiles + 1);
    rc |= dalloc(ctx, &b->scan_scratch, b->max_tiles / MCS_SCAN_TILE + 8);
    rc |= dalloc(ctx, &b->tile_part, b->max_tiles * 2 * SW_NT);
    if (rc) {
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
    free_popbuf(b->buf);
    free_popbuf(b->alt);
    SlotDev &S = b->S;
    void *ptrs[] = {S.base, S.cap, S.n, S.n0, S.used, S.flags, S.mode, S.relax, S.block, S.e, S.dt,
                    S.tol, S.param, S.key, S.step, S.next_id, S.tv, S.first_tile, S.ntiles, S.sum,
                    S.sumc, S.cur, S.bm1, S.bm2, S.last, S.steps, S.nrec, S.nblk, S.curcnt, S.nact,
                    b->L.parent, b->L.seq, b->L.s, b->L.act, b->L.x[0], b->L.x[1], b->L.x[2],
                    b->L.v[0], b->L.v[1], b->L.v[2], b->log_count, b->loff, b->cnt, b->tile_slot,
                    b->tile_start, b->tile_births, b->tile_bpre, b->tile_clones, b->tile_cpre,
                    b->tile_act, b->tile_apre, b->scan_scratch, b->tile_part};
    for (void *p : ptrs) hipFree(p);
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
        MCS_HIP(ctx, hipMemcpy(b->tile_slot, ts.data(), sizeof(int32_t) * ts.size(), hipMemcpyHostToDevice));
        MCS_HIP(ctx, hipMemcpy(b->tile_start, tst.data(), sizeof(int64_t) * tst.size(), hipMemcpyHostToDevice));
    }
    MCS_HIP(ctx, hipMemcpy(b->S.first_tile, first.data(), sizeof(int32_t) * first.size(), hipMemcpyHostToDevice));
    MCS_HIP(ctx, hipMemcpy(b->S.ntiles, nt.data(), sizeof(int32_t) * nt.size(), hipMemcpyHostToDevice));
    return MCS_OK;
}

template <typename T>
static int32_t put1(mcs_ctx *ctx, T *dev, int slot, const T &v)
{
    MCS_HIP(ctx, hipMemcpy(dev + slot, &v, sizeof(T), hipMemcpyHostToDevice));
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
    MCS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    // particles
    if (dev_init && n > 0) {
        hipLaunchKernelGGL(k_sw_init, dim3((unsigned)cdiv(n, MCS_BLOCK)), dim3(MCS_BLOCK), 0, ctx->stream,
                           b->buf, base, n, b->seed, prm->instance, prm->init_vth);
        MCS_LAUNCH_CHECK(ctx);
        MCS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    std::vector<double> tmp((size_t)(dev_init ? 0 : n));
    for (int d = 0; d < 3 && !dev_init; ++d) {
        for (int64_t i = 0; i < n; ++i) tmp[i] = x[3 * i + d];
        MCS_HIP(ctx, hipMemcpy(b->buf.x[d] + base, tmp.data(), sizeof(double) * n, hipMemcpyHostToDevice));
        for (int64_t i = 0; i < n; ++i) tmp[i] = v[3 * i + d];
        MCS_HIP(ctx, hipMemcpy(b->buf.v[d] + base, tmp.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    }
    if (!dev_init) {
        MCS_HIP(ctx, hipMemcpy(b->buf.w + base, w, sizeof(double) * n, hipMemcpyHostToDevice));
        MCS_HIP(ctx, hipMemcpy(b->buf.s + base, s, sizeof(double) * n, hipMemcpyHostToDevice));
        MCS_HIP(ctx, hipMemcpy(b->buf.active + base, active, n, hipMemcpyHostToDevice));
        std::vector<uint64_t> ids((size_t)n);
        for (int64_t i = 0; i < n; ++i) ids[i] = (uint64_t)i;
        MCS_HIP(ctx, hipMemcpy(b->buf.id + base, ids.data(), sizeof(uint64_t) * n, hipMemcpyHostToDevice));
    }
    // slot state
    SlotDev &S = b->S;
    int32_t rc = 0;
    rc |= put1(ctx, S.base, slot, base) | put1(ctx, S.cap, slot, cap) | put1(ctx, S.n, slot, n) |
          put1(ctx, S.n0, slot, prm->n0);
    rc |= put1(ctx, S.used, slot, (int32_t)1) | put1(ctx, S.flags, slot, (int32_t)0) |
          put1(ctx, S.mode, slot, (int32_t)0) | put1(ctx, S.relax, slot, (int32_t)prm->relax_steps) |
          put1(ctx, S.block, slot, (int32_t)prm->block_steps);
    for (int d = 0; d < 3; ++d) rc |= put1(ctx, S.e, 3 * slot + d, prm->e[d]);
    rc |= put1(ctx, S.dt, slot, prm->dt) | put1(ctx, S.tol, slot, prm->tol) | put1(ctx, S.param, slot, 0.0);
    rc |= put1(ctx, S.key, slot, prm->instance) | put1(ctx, S.step, slot, (uint32_t)0) |
          put1(ctx, S.next_id, slot, (uint64_t)n);
    rc |= put1(ctx, S.tv, slot, table_view(t));
    for (int q = 0; q < SW_NQ; ++q) {
        int k = SW_NQ * slot + q;
        rc |= put1(ctx, S.sum, k, 0.0) | put1(ctx, S.sumc, k, 0.0) | put1(ctx, S.cur, k, 0.0) |
              put1(ctx, S.bm1, k, 0.0) | put1(ctx, S.bm2, k, 0.0) | put1(ctx, S.last, k, 0.0);
    }
    rc |= put1(ctx, S.steps, slot, (int64_t)0) | put1(ctx, S.nrec, slot, (int64_t)0) |
          put1(ctx, S.nblk, slot, (int64_t)0) | put1(ctx, S.curcnt, slot, (int64_t)0) |
          put1(ctx, S.nact, slot, (int64_t)0);
    if (rc) return rc;
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
    MCS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    int32_t rc = put1(ctx, b->S.used, slot, (int32_t)0);
    if (rc) return rc;
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

static SwArgs args_of(mcs_swarm_batch *b)
{
    SwArgs a;
    a.b = b->buf;
    a.alt = b->alt;
    a.S = b->S;
    a.tile_slot = b->tile_slot;
    a.tile_start = b->tile_start;
    a.seed = b->seed;
    a.L = b->L;
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
    return a;
}

// The slot of a particle index: binary search over tile starts.
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

// Children from the birth log to their slabs: rank within the slot in parent
// order (tile offset + offset within the tile + sequence within the parent),
// ids next_id + rank.  Grid-stride over the log with a fixed grid.
__global__ __launch_bounds__(MCS_BLOCK) void k_sw_births(SwArgs a, int64_t logcap, int64_t ntiles)
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
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            a.b.x[q][d] = a.L.x[q][r];
            a.b.v[q][d] = a.L.v[q][r];
        }
        a.b.w[d] = a.b.w[i];
        a.b.s[d] = a.L.s[r];
        a.b.active[d] = a.L.act[r];
        a.b.id[d] = a.S.next_id[slot] + (uint64_t)rank;
    }
}

static int32_t launch_step(mcs_swarm_batch *b)
{
    mcs_ctx *ctx = b->ctx;
    SwArgs a = args_of(b);
    const dim3 g((unsigned)b->ntiles);
    hipLaunchKernelGGL(k_sw_zero, dim3(1), dim3(64), 0, ctx->stream, b->log_count, 8);
    if (b->block == 128)
        hipLaunchKernelGGL(k_sw_advance<128>, g, dim3(128), 0, ctx->stream, a);
    else
        hipLaunchKernelGGL(k_sw_advance<256>, g, dim3(256), 0, ctx->stream, a);
    int32_t rc = mcs_scan_i64_buf(ctx, b->tile_births, b->tile_bpre, b->ntiles, b->scan_scratch);
    if (rc) return rc;
    hipLaunchKernelGGL(k_sw_births, dim3((unsigned)std::min<int64_t>(cdiv(b->L.cap, MCS_BLOCK), 2048)), dim3(MCS_BLOCK), 0,
                       ctx->stream, a, b->L.cap, b->ntiles);
    hipLaunchKernelGGL(k_sw_slot_step, dim3((unsigned)b->max_slots), dim3(MCS_BLOCK), 0, ctx->stream, a);
    if (b->block == 128)
        hipLaunchKernelGGL(k_sw_resample<128>, g, dim3(128), 0, ctx->stream, a);
    else
        hipLaunchKernelGGL(k_sw_resample<256>, g, dim3(256), 0, ctx->stream, a);
    rc = mcs_scan_i64_buf(ctx, b->tile_clones, b->tile_cpre, b->ntiles, b->scan_scratch);
    if (rc) return rc;
    if (b->block == 128)
        hipLaunchKernelGGL(k_sw_clone_place<128>, g, dim3(128), 0, ctx->stream, a);
    else
        hipLaunchKernelGGL(k_sw_clone_place<256>, g, dim3(256), 0, ctx->stream, a);
    hipLaunchKernelGGL(k_sw_slot_step2, dim3((unsigned)cdiv(b->max_slots, 64)), dim3(64), 0, ctx->stream,
                       a, b->max_slots);
    MCS_LAUNCH_CHECK(ctx);
    return MCS_OK;
}

static int32_t launch_compact(mcs_swarm_batch *b)
{
    mcs_ctx *ctx = b->ctx;
    SwArgs a = args_of(b);
    const dim3 g((unsigned)b->ntiles);
    if (b->block == 128)
        hipLaunchKernelGGL(k_sw_compact_count<128>, g, dim3(128), 0, ctx->stream, a);
    else
        hipLaunchKernelGGL(k_sw_compact_count<256>, g, dim3(256), 0, ctx->stream, a);
    int32_t rc = mcs_scan_i64_buf(ctx, b->tile_act, b->tile_apre, b->ntiles, b->scan_scratch);
    if (rc) return rc;
    if (b->block == 128)
        hipLaunchKernelGGL(k_sw_compact_write<128>, g, dim3(128), 0, ctx->stream, a);
    else
        hipLaunchKernelGGL(k_sw_compact_write<256>, g, dim3(256), 0, ctx->stream, a);
    hipLaunchKernelGGL(k_sw_compact_slot, dim3((unsigned)cdiv(b->max_slots, 64)), dim3(64), 0, ctx->stream,
                       a, b->max_slots);
    if (b->block == 128)
        hipLaunchKernelGGL(k_sw_copy_back<128>, g, dim3(128), 0, ctx->stream, a);
    else
        hipLaunchKernelGGL(k_sw_copy_back<256>, g, dim3(256), 0, ctx->stream, a);
    MCS_LAUNCH_CHECK(ctx);
    return MCS_OK;
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

static int32_t read_tally(mcs_swarm_batch *b, int32_t slot, mcs_swarm_tally *o)
{
    mcs_ctx *ctx = b->ctx;
    const SlotDev &S = b->S;
    memset(o, 0, sizeof(*o));
    MCS_HIP(ctx, hipMemcpy(&o->steps, S.steps + slot, sizeof(int64_t), hipMemcpyDeviceToHost));
    MCS_HIP(ctx, hipMemcpy(&o->n, S.n + slot, sizeof(int64_t), hipMemcpyDeviceToHost));
    MCS_HIP(ctx, hipMemcpy(&o->nactive, S.nact + slot, sizeof(int64_t), hipMemcpyDeviceToHost));
    MCS_HIP(ctx, hipMemcpy(&o->nrecorded, S.nrec + slot, sizeof(int64_t), hipMemcpyDeviceToHost));
    MCS_HIP(ctx, hipMemcpy(&o->nblocks, S.nblk + slot, sizeof(int64_t), hipMemcpyDeviceToHost));
    MCS_HIP(ctx, hipMemcpy(&o->flags, S.flags + slot, sizeof(int32_t), hipMemcpyDeviceToHost));
    double sum[SW_NQ], sumc[SW_NQ];
    MCS_HIP(ctx, hipMemcpy(sum, S.sum + SW_NQ * slot, sizeof(sum), hipMemcpyDeviceToHost));
    MCS_HIP(ctx, hipMemcpy(sumc, S.sumc + SW_NQ * slot, sizeof(sumc), hipMemcpyDeviceToHost));
    for (int q = 0; q < SW_NQ; ++q) o->sum[q] = sum[q] + sumc[q];
    MCS_HIP(ctx, hipMemcpy(o->bm_s1, S.bm1 + SW_NQ * slot, sizeof(o->bm_s1), hipMemcpyDeviceToHost));
    MCS_HIP(ctx, hipMemcpy(o->bm_s2, S.bm2 + SW_NQ * slot, sizeof(o->bm_s2), hipMemcpyDeviceToHost));
    MCS_HIP(ctx, hipMemcpy(o->last, S.last + SW_NQ * slot, sizeof(o->last), hipMemcpyDeviceToHost));
    return MCS_OK;
}

extern "C" int32_t mcs_swarm_batch_tally(mcs_swarm_batch *b, int32_t slot, mcs_swarm_tally *out)
{
    if (!b || !out || slot < 0 || slot >= b->max_slots) return MCS_E_ARG;
    hipSetDevice(b->ctx->device);
    MCS_HIP(b->ctx, hipStreamSynchronize(b->ctx->stream));
    return read_tally(b, slot, out);
}

extern "C" int32_t mcs_swarm_batch_tally_all(mcs_swarm_batch *b, int32_t nslots, mcs_swarm_tally *out)
{
    if (!b || !out || nslots < 0 || nslots > b->max_slots) return MCS_E_ARG;
    mcs_ctx *ctx = b->ctx;
    hipSetDevice(ctx->device);
    MCS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    // one bulk copy per array
    const SlotDev &S = b->S;
    std::vector<int64_t> steps(nslots), n(nslots), nact(nslots), nrec(nslots), nblk(nslots);
    std::vector<int32_t> flags(nslots);
    std::vector<double> sum(SW_NQ * nslots), sumc(SW_NQ * nslots), bm1(SW_NQ * nslots),
        bm2(SW_NQ * nslots), last(SW_NQ * nslots);
    if (nslots == 0) return MCS_OK;
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
    mcs_ctx *ctx = p->ctx;
    hipSetDevice(ctx->device);
    if (n < 0 || n > p->cap || !(vth >= 0.0))
        return mcs_fail(ctx, MCS_E_ARG, "pop_init_thermal: n must be in [0, capacity], vth >= 0");
    if (n > 0) {
        hipLaunchKernelGGL(k_sw_init, dim3((unsigned)cdiv(n, MCS_BLOCK)), dim3(MCS_BLOCK), 0, ctx->stream,
                           p->buf[0], (int64_t)0, n, seed, instance, vth);
        MCS_LAUNCH_CHECK(ctx);
    }
    p->n = n;
    p->next_id = (uint64_t)n;
    return MCS_OK;
}
