// mcs_step.hip -- the fused particle step (JuMC.advance!), birth placement,
// repack! and the swarm's global resample.
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "mcs_advance.h"
#include "mcs_fp32.h"

// counters block layout in ctx->d_small
#define C_LOG 0
#define C_FLIGHTS 1
#define C_DEATHS 2
#define C_FLAGS 3
#define C_TLOG 4     // tracker deposits logged
#define C_BIRTHS 5   // births (the births' scan total)
#define C_TTOT 6     // tracker deposits (their scan total)
#define C_EVENTS 8
#define C_BLOG (C_EVENTS + MCS_MAX_PROC)   // photon bursts logged
#define C_PTOT (C_BLOG + 1)                // photons emitted (their scan total)

struct BirthLog {
    int32_t *parent, *seq;
    double *x[3], *v[3];
    double *w, *s;
    uint8_t *act;
    int64_t cap;
};

static inline BirthLog birth_log(mcs_pop *p)
{
    BirthLog L;
    L.parent = p->log_parent;
    L.seq = p->log_seq;
    for (int d = 0; d < 3; ++d) {
        L.x[d] = p->log_x[d];
        L.v[d] = p->log_v[d];
    }
    L.w = p->log_w;
    L.s = p->log_s;
    L.act = p->log_act;
    L.cap = p->logcap;
    return L;
}

// The ion tracker's deposits of one step (CollisionTracker, gridfields.jl:203-218):
// each (cell, amount) is logged under (parent slot, ordinal within the parent) at
// a slot taken from a counter; per-thread counts and their block prefix
// (toff, tbtot) give every deposit its place in parent order afterwards, where
// the deposits are sorted stably by cell and each cell's run is added to its
// value in that order -- the order of the reference's sequential loop, with no
// floating-point atomics.
struct TrackLog {
    int32_t *parent, *seq, *cell;
    double *val;
    int64_t cap;
    int32_t *toff;
    int64_t *tbtot;
};

static inline TrackLog track_log(mcs_pop *p)
{
    TrackLog T;
    T.parent = p->trk_parent;
    T.seq = p->trk_seq;
    T.cell = p->trk_cell;
    T.val = p->trk_val;
    T.cap = p->logcap;
    T.toff = p->i32b;
    T.tbtot = p->tblk;
    return T;
}

__device__ __forceinline__ void track_push(const TrackLog &TL, int64_t *ctr, int64_t parent, int &nt,
                                           int32_t cell, double amount, int32_t &flags)
{
    const unsigned long long slot = atomicAdd((unsigned long long *)&ctr[C_TLOG], 1ULL);
    if ((int64_t)slot < TL.cap) {
        TL.parent[slot] = (int32_t)parent;
        TL.seq[slot] = nt;
        TL.cell[slot] = cell;
        TL.val[slot] = amount;
    } else {
        flags |= MCS_FLAG_CAPACITY;
    }
    nt += 1;
}

// The per-thread deposit counts of a block: offsets of each thread's deposits
// within the block, the block's total at its logical index.
__device__ __forceinline__ void track_block_counts(const TrackLog &TL, int64_t i, int nt)
{
    int ttot;
    const int tex = block_excl_scan<int>(nt, &ttot);
    if (nt > 0) TL.toff[i] = tex;
    if (threadIdx.x == 0) TL.tbtot[xcd_block()] = ttot;
}

// The photon bursts of a step's PhotoEmission collisions: each burst with at
// least one photon is logged under its parent slot with the parent's photons
// before it (base); per-thread photon totals give, after a block scan (poff, per
// slot) and a scan of the block totals (pbtot), every burst its first photon's
// place in parent order.
struct BurstLog {
    int32_t *parent, *row;
    int64_t *base, *n;
    uint8_t *act;
    double *x[3];
    int64_t cap;
    int64_t *poff, *pbtot;
};

static inline BurstLog burst_log(mcs_pop *p)
{
    BurstLog B;
    B.parent = p->bur_parent;
    B.row = p->bur_row;
    B.base = p->bur_base;
    B.n = p->bur_n;
    B.act = p->bur_act;
    for (int d = 0; d < 3; ++d) B.x[d] = p->bur_x[d];
    B.cap = p->logcap;
    B.poff = p->poff;
    B.pbtot = p->pblk;
    return B;
}

struct AdvArgs {
    PopBuf b;
    int64_t n0;
    TableView T;
    MeshView M;
    double e[3];
    double dt;
    uint64_t seed;
    uint32_t instance, step;
    BirthLog L;
    int32_t *loff;
    int64_t *btot;
    int64_t *ctr;
    TrackLog TL;
};

// The fused step: one thread per particle slot that exists when the step starts
// (JuMC/src/mixed_population.jl:27 fixes the range, so children are not advanced).
// Per particle, free flights to the next (null) collision until dt is used up
// (:31-51), the Yoshida push of electron.jl:51-62 for each flight, the sampler of
// collisions.jl:159-192 and the outcomes of electron.jl:108-139.  Children go to a
// birth log tagged (parent, ordinal); k_place writes them after the parents in
// parent order, which is the append order of the single-threaded reference.
// With PHOTO, PhotoEmission collisions log photon bursts (BurstLog).  k_advance
// below keeps its own copy of this body without the photon sink: instantiated
// through this function it needs 120 bytes of scratch per lane instead of 56
// (the compiler's resource-usage report), which its performance cannot spare.
template <int FIELD_MESH, int TRACK, int PHOTO>
__device__ __forceinline__ void advance_body(const AdvArgs &a, const BurstLog &BL)
{
    __shared__ int s_ev[MCS_MAX_PROC];
    __shared__ int s_flights, s_deaths, s_flags;
    if (threadIdx.x < MCS_MAX_PROC) s_ev[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
        s_flights = 0;
        s_deaths = 0;
        s_flags = 0;
    }
    __syncthreads();

    const int64_t i = xcd_gidx();
    int nb = 0, nt = 0;
    int64_t nph = 0;
    int32_t flags = 0;
    int flights = 0, deaths = 0;
    if (i < a.n0 && a.b.active[i]) {
        double x[3] = {a.b.x[0][i], a.b.x[1][i], a.b.x[2][i]};
        double v[3] = {a.b.v[0][i], a.b.v[1][i], a.b.v[2][i]};
        const double w = a.b.w[i];
        double s = a.b.s[i];
        uint8_t act = 1;
        mcs_stream rs;
        mcs_stream_init(&rs, a.seed, a.instance, MCS_PHASE_ADVANCE, a.b.id[i], a.step);
        auto birth = [&](const double *bx, const double *bv, double bs, uint8_t bact) {
            unsigned long long slot = atomicAdd((unsigned long long *)&a.ctr[C_LOG], 1ULL);
            if ((int64_t)slot < a.L.cap) {
                a.L.parent[slot] = (int32_t)i;
                a.L.seq[slot] = nb;
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    a.L.x[d][slot] = bx[d];
                    a.L.v[d][slot] = bv[d];
                }
                a.L.w[slot] = w;
                a.L.s[slot] = bs;
                a.L.act[slot] = bact;
            } else {
                flags |= MCS_FLAG_CAPACITY;
            }
            nb += 1;
        };
        auto event = [&](int j) { atomicAdd(&s_ev[j], 1); };
        auto track = [&](const double *tx, double val) {
            int32_t cell;
            double amount;
            if (track_cell<FIELD_MESH>(a.M, tx, val, cell, amount))
                track_push(a.TL, a.ctr, i, nt, cell, amount, flags);
        };
        if constexpr (PHOTO) {
            auto photo = [&](int row, const double *px, uint8_t pact, int64_t np) {
                if (np <= 0) return;
                const unsigned long long slot = atomicAdd((unsigned long long *)&a.ctr[C_BLOG], 1ULL);
                if ((int64_t)slot < BL.cap) {
                    BL.parent[slot] = (int32_t)i;
                    BL.row[slot] = row;
                    BL.base[slot] = nph;
                    BL.n[slot] = np;
                    BL.act[slot] = pact;
#pragma unroll
                    for (int d = 0; d < 3; ++d) BL.x[d][slot] = px[d];
                } else {
                    flags |= MCS_FLAG_CAPACITY;
                }
                nph += np;
            };
            advance_one<FIELD_MESH, TRACK>(x, v, w, s, act, rs, a.dt, a.T, a.M, a.e, flags, flights, deaths,
                                           birth, event, track, photo);
        } else {
            advance_one<FIELD_MESH, TRACK>(x, v, w, s, act, rs, a.dt, a.T, a.M, a.e, flags, flights, deaths,
                                           birth, event, track);
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            a.b.x[d][i] = x[d];
            a.b.v[d][i] = v[d];
        }
        a.b.s[i] = s;
        if (!act) a.b.active[i] = 0;
    }
    int tot;
    int ex = block_excl_scan<int>(nb, &tot);
    if (nb > 0) a.loff[i] = ex;
    if (threadIdx.x == 0) a.btot[xcd_block()] = tot;
    if (TRACK) track_block_counts(a.TL, i, nt);
    if constexpr (PHOTO) {
        int64_t ptot;
        const int64_t pex = block_excl_scan<int64_t>(nph, &ptot);
        if (nph > 0) BL.poff[i] = pex;
        if (threadIdx.x == 0) BL.pbtot[xcd_block()] = ptot;
    }
    if (flights) atomicAdd(&s_flights, flights);
    if (deaths) atomicAdd(&s_deaths, deaths);
    if (flags) atomicOr(&s_flags, flags);
    __syncthreads();
    if (threadIdx.x < a.T.nproc && s_ev[threadIdx.x])
        atomicAdd((unsigned long long *)&a.ctr[C_EVENTS + threadIdx.x],
                  (unsigned long long)s_ev[threadIdx.x]);
    if (threadIdx.x == 0) {
        if (s_flights) atomicAdd((unsigned long long *)&a.ctr[C_FLIGHTS], (unsigned long long)s_flights);
        if (s_deaths) atomicAdd((unsigned long long *)&a.ctr[C_DEATHS], (unsigned long long)s_deaths);
        if (s_flags) atomicOr((unsigned long long *)&a.ctr[C_FLAGS], (unsigned long long)s_flags);
    }
}

// Five waves per SIMD (96 registers; a few cold values spill): the field gathers'
// latency is hidden better than at the four that 106 registers allow.
template <int FIELD_MESH, int TRACK>
__global__ __launch_bounds__(MCS_BLOCK) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_advance(AdvArgs a)
{
    __shared__ int s_ev[MCS_MAX_PROC];
    __shared__ int s_flights, s_deaths, s_flags;
    if (threadIdx.x < MCS_MAX_PROC) s_ev[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
        s_flights = 0;
        s_deaths = 0;
        s_flags = 0;
    }
    __syncthreads();

    const int64_t i = xcd_gidx();
    int nb = 0, nt = 0;
    int32_t flags = 0;
    int flights = 0, deaths = 0;
    if (i < a.n0 && a.b.active[i]) {
        double x[3] = {a.b.x[0][i], a.b.x[1][i], a.b.x[2][i]};
        double v[3] = {a.b.v[0][i], a.b.v[1][i], a.b.v[2][i]};
        const double w = a.b.w[i];
        double s = a.b.s[i];
        uint8_t act = 1;
        mcs_stream rs;
        mcs_stream_init(&rs, a.seed, a.instance, MCS_PHASE_ADVANCE, a.b.id[i], a.step);
        advance_one<FIELD_MESH, TRACK>(
            x, v, w, s, act, rs, a.dt, a.T, a.M, a.e, flags, flights, deaths,
            [&](const double *bx, const double *bv, double bs, uint8_t bact) {
                unsigned long long slot = atomicAdd((unsigned long long *)&a.ctr[C_LOG], 1ULL);
                if ((int64_t)slot < a.L.cap) {
                    a.L.parent[slot] = (int32_t)i;
                    a.L.seq[slot] = nb;
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        a.L.x[d][slot] = bx[d];
                        a.L.v[d][slot] = bv[d];
                    }
                    a.L.w[slot] = w;
                    a.L.s[slot] = bs;
                    a.L.act[slot] = bact;
                } else {
                    flags |= MCS_FLAG_CAPACITY;
                }
                nb += 1;
            },
            [&](int j) { atomicAdd(&s_ev[j], 1); },
            [&](const double *tx, double val) {
                int32_t cell;
                double amount;
                if (track_cell<FIELD_MESH>(a.M, tx, val, cell, amount))
                    track_push(a.TL, a.ctr, i, nt, cell, amount, flags);
            });
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            a.b.x[d][i] = x[d];
            a.b.v[d][i] = v[d];
        }
        a.b.s[i] = s;
        if (!act) a.b.active[i] = 0;
    }
    int tot;
    int ex = block_excl_scan<int>(nb, &tot);
    if (nb > 0) a.loff[i] = ex;
    if (threadIdx.x == 0) a.btot[xcd_block()] = tot;
    if (TRACK) track_block_counts(a.TL, i, nt);
    if (flights) atomicAdd(&s_flights, flights);
    if (deaths) atomicAdd(&s_deaths, deaths);
    if (flags) atomicOr(&s_flags, flags);
    __syncthreads();
    if (threadIdx.x < a.T.nproc && s_ev[threadIdx.x])
        atomicAdd((unsigned long long *)&a.ctr[C_EVENTS + threadIdx.x],
                  (unsigned long long)s_ev[threadIdx.x]);
    if (threadIdx.x == 0) {
        if (s_flights) atomicAdd((unsigned long long *)&a.ctr[C_FLIGHTS], (unsigned long long)s_flights);
        if (s_deaths) atomicAdd((unsigned long long *)&a.ctr[C_DEATHS], (unsigned long long)s_deaths);
        if (s_flags) atomicOr((unsigned long long *)&a.ctr[C_FLAGS], (unsigned long long)s_flags);
    }
}

// The same step for a table with PhotoEmission rows.
template <int FIELD_MESH, int TRACK>
__global__ __launch_bounds__(MCS_BLOCK) void k_advance_photo(AdvArgs a, BurstLog BL)
{
    advance_body<FIELD_MESH, TRACK, 1>(a, BL);
}

// Children land at n0 + (births of earlier blocks) + (births of earlier threads in
// the block) + ordinal, with ids continuing next_id in that order.
__global__ __launch_bounds__(MCS_BLOCK) void k_place_births(BirthLog L, int64_t count,
                                                            const int32_t *loff,
                                                            const int64_t *bpre, int64_t n0,
                                                            uint64_t next_id, PopBuf b)
{
    int64_t r = gidx();
    if (r >= count) return;
    int32_t par = L.parent[r];
    int64_t dst = n0 + bpre[par / MCS_BLOCK] + loff[par] + L.seq[r];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        b.x[d][dst] = L.x[d][r];
        b.v[d][dst] = L.v[d][r];
    }
    b.w[dst] = L.w[r];
    b.s[dst] = L.s[r];
    b.active[dst] = L.act[r];
    b.id[dst] = next_id + (uint64_t)(dst - n0);
}

__global__ void k_fill_i64(int64_t *a, int64_t n, int64_t v)
{
    int64_t i = gidx();
    if (i < n) a[i] = v;
}

// ---- the same step on the split store (precision mode) ----------------------

struct BirthLog32 {
    int32_t *parent, *seq, *cell;
    float *off[3], *v[3];
    float *w, *s;
    uint8_t *act;
    int64_t cap;
};

struct AdvArgs32 {
    SplitBuf b;
    int64_t n0;
    TableView32 T;
    MeshView32 M;
    float acc[3];      // q/m E of a uniform field
    float dt, inv_h;
    int32_t nmax[3];   // grid extent for rebasing
    uint64_t seed;
    uint32_t instance, step;
    BirthLog32 L;
    int32_t *loff;
    int64_t *btot;
    int64_t *ctr;
    TrackLog TL;
};

template <int FIELD_MESH, int TRACK>
__global__ __launch_bounds__(MCS_BLOCK) void k_advance32(AdvArgs32 a)
{
    __shared__ int s_ev[MCS_MAX_PROC];
    __shared__ int s_flights, s_deaths, s_flags;
    if (threadIdx.x < MCS_MAX_PROC) s_ev[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
        s_flights = 0;
        s_deaths = 0;
        s_flags = 0;
    }
    __syncthreads();
    const int64_t i = xcd_gidx();
    int nb = 0, nt = 0;
    int32_t flags = 0;
    int flights = 0, deaths = 0;
    if (i < a.n0 && a.b.active[i]) {
        int32_t c[3];
        unpack_cell(a.b.cell[i], c[0], c[1], c[2]);
        float off[3] = {a.b.off[0][i], a.b.off[1][i], a.b.off[2][i]};
        float v[3] = {a.b.v[0][i], a.b.v[1][i], a.b.v[2][i]};
        const float w = a.b.w[i];
        float s = a.b.s[i];
        uint8_t act = 1;
        mcs_stream rs;
        mcs_stream_init(&rs, a.seed, a.instance, MCS_PHASE_ADVANCE, a.b.id[i], a.step);
        advance_one32<FIELD_MESH, TRACK>(
            c, off, v, w, s, act, rs, a.dt, a.T, a.M, a.acc, a.inv_h, flags, flights, deaths,
            [&](const int32_t *bc, const float *boff, const float *bv, float bs, uint8_t bact) {
                unsigned long long slot = atomicAdd((unsigned long long *)&a.ctr[C_LOG], 1ULL);
                if ((int64_t)slot < a.L.cap) {
                    a.L.parent[slot] = (int32_t)i;
                    a.L.seq[slot] = nb;
                    a.L.cell[slot] = pack_cell(bc[0], bc[1], bc[2]);
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        a.L.off[d][slot] = boff[d];
                        a.L.v[d][slot] = bv[d];
                    }
                    a.L.w[slot] = w;
                    a.L.s[slot] = bs;
                    a.L.act[slot] = bact;
                } else {
                    flags |= MCS_FLAG_CAPACITY;
                }
                nb += 1;
            },
            [&](int j) { atomicAdd(&s_ev[j], 1); },
            [&](const int32_t *tc, const float *toff, double val) {
                int32_t cell;
                double amount;
                if (track_cell32(a.M, tc, toff, val, cell, amount))
                    track_push(a.TL, a.ctr, i, nt, cell, amount, flags);
            });
        rebase32(c, off, a.nmax);
        a.b.cell[i] = pack_cell(c[0], c[1], c[2]);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            a.b.off[d][i] = off[d];
            a.b.v[d][i] = v[d];
        }
        a.b.s[i] = s;
        if (!act) a.b.active[i] = 0;
    }
    int tot;
    int ex = block_excl_scan<int>(nb, &tot);
    if (nb > 0) a.loff[i] = ex;
    if (threadIdx.x == 0) a.btot[xcd_block()] = tot;
    if (TRACK) track_block_counts(a.TL, i, nt);
    if (flights) atomicAdd(&s_flights, flights);
    if (deaths) atomicAdd(&s_deaths, deaths);
    if (flags) atomicOr(&s_flags, flags);
    __syncthreads();
    if (threadIdx.x < a.T.nproc && s_ev[threadIdx.x])
        atomicAdd((unsigned long long *)&a.ctr[C_EVENTS + threadIdx.x], (unsigned long long)s_ev[threadIdx.x]);
    if (threadIdx.x == 0) {
        if (s_flights) atomicAdd((unsigned long long *)&a.ctr[C_FLIGHTS], (unsigned long long)s_flights);
        if (s_deaths) atomicAdd((unsigned long long *)&a.ctr[C_DEATHS], (unsigned long long)s_deaths);
        if (s_flags) atomicOr((unsigned long long *)&a.ctr[C_FLAGS], (unsigned long long)s_flags);
    }
}

// The FP32 step with the collision table staged in LDS: one workgroup of 1024
// threads per CU copies the FP32 rates, energy nodes and parameters (ne (nproc
// + 1) + nproc floats, 128 KB for the 7-row, 4000-node gas) into LDS and then
// runs k_advance32's body over chunks of 1024 particle slots, each chunk as
// four logical blocks of MCS_BLOCK slots with their own birth and deposit
// scans, so that the placement kernels and every result are those of
// k_advance32 (the table's values are the same; only where they are read from
// differs).
#define ADV_LDS_THREADS 1024
#define ADV_LDS_GROUPS (ADV_LDS_THREADS / MCS_BLOCK)

// Exclusive scan over the MCS_BLOCK threads of this thread's group.
__device__ __forceinline__ int group_excl_scan(int v, int *total, int *wsum)
{
    const int lane = threadIdx.x & (MCS_WAVE - 1), wid = threadIdx.x / MCS_WAVE;
    const int g0 = (wid / (MCS_BLOCK / MCS_WAVE)) * (MCS_BLOCK / MCS_WAVE);
    int x = v;
#pragma unroll
    for (int o = 1; o < MCS_WAVE; o <<= 1) {
        const int y = __shfl_up(x, o, MCS_WAVE);
        if (lane >= o) x += y;
    }
    if (lane == MCS_WAVE - 1) wsum[wid] = x;
    __syncthreads();
    int pre = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < MCS_BLOCK / MCS_WAVE; ++k) {
        const int sk = wsum[g0 + k];
        if (g0 + k < wid) pre += sk;
        tot += sk;
    }
    __syncthreads();
    *total = tot;
    return pre + x - v;
}

template <int FIELD_MESH, int TRACK>
__global__ __launch_bounds__(ADV_LDS_THREADS) void k_advance32_lds(AdvArgs32 a, int64_t nchunks, int64_t nblk)
{
    extern __shared__ float tab[];
    __shared__ int s_ev[MCS_MAX_PROC];
    __shared__ int s_flights, s_deaths, s_flags;
    __shared__ int s_wsum[ADV_LDS_THREADS / MCS_WAVE];
    const int nr = a.T.ne * a.T.nproc;
    float *lrate = tab, *lenergy = tab + nr, *lparam = lenergy + a.T.ne;
    for (int q = threadIdx.x; q < nr; q += ADV_LDS_THREADS) lrate[q] = gld(a.T.rate + q);
    for (int q = threadIdx.x; q < a.T.ne; q += ADV_LDS_THREADS) lenergy[q] = gld(a.T.energy + q);
    if (threadIdx.x < a.T.nproc) lparam[threadIdx.x] = gld(a.T.param + threadIdx.x);
    if (threadIdx.x < MCS_MAX_PROC) s_ev[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
        s_flights = 0;
        s_deaths = 0;
        s_flags = 0;
    }
    __syncthreads();
    TableView32 T = a.T;
    T.rate = lrate;
    T.energy = lenergy;
    T.param = lparam;
    const int grp = threadIdx.x / MCS_BLOCK, gt = threadIdx.x % MCS_BLOCK;
    int32_t flags = 0;
    int flights = 0, deaths = 0;
    for (int64_t ch = xcd_block(); ch < nchunks; ch += gridDim.x) {
        const int64_t lb = ch * ADV_LDS_GROUPS + grp;
        const int64_t i = lb * MCS_BLOCK + gt;
        int nb = 0, nt = 0;
        if (i < a.n0 && a.b.active[i]) {
            int32_t c[3];
            unpack_cell(a.b.cell[i], c[0], c[1], c[2]);
            float off[3] = {a.b.off[0][i], a.b.off[1][i], a.b.off[2][i]};
            float v[3] = {a.b.v[0][i], a.b.v[1][i], a.b.v[2][i]};
            const float w = a.b.w[i];
            float s = a.b.s[i];
            uint8_t act = 1;
            mcs_stream rs;
            mcs_stream_init(&rs, a.seed, a.instance, MCS_PHASE_ADVANCE, a.b.id[i], a.step);
            advance_one32<FIELD_MESH, TRACK, 1>(
                c, off, v, w, s, act, rs, a.dt, T, a.M, a.acc, a.inv_h, flags, flights, deaths,
                [&](const int32_t *bc, const float *boff, const float *bv, float bs, uint8_t bact) {
                    unsigned long long slot = atomicAdd((unsigned long long *)&a.ctr[C_LOG], 1ULL);
                    if ((int64_t)slot < a.L.cap) {
                        a.L.parent[slot] = (int32_t)i;
                        a.L.seq[slot] = nb;
                        a.L.cell[slot] = pack_cell(bc[0], bc[1], bc[2]);
#pragma unroll
                        for (int d = 0; d < 3; ++d) {
                            a.L.off[d][slot] = boff[d];
                            a.L.v[d][slot] = bv[d];
                        }
                        a.L.w[slot] = w;
                        a.L.s[slot] = bs;
                        a.L.act[slot] = bact;
                    } else {
                        flags |= MCS_FLAG_CAPACITY;
                    }
                    nb += 1;
                },
                [&](int j) { atomicAdd(&s_ev[j], 1); },
                [&](const int32_t *tc, const float *toff, double val) {
                    int32_t cell;
                    double amount;
                    if (track_cell32(a.M, tc, toff, val, cell, amount))
                        track_push(a.TL, a.ctr, i, nt, cell, amount, flags);
                });
            rebase32(c, off, a.nmax);
            a.b.cell[i] = pack_cell(c[0], c[1], c[2]);
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                a.b.off[d][i] = off[d];
                a.b.v[d][i] = v[d];
            }
            a.b.s[i] = s;
            if (!act) a.b.active[i] = 0;
        }
        int tot;
        const int ex = group_excl_scan(nb, &tot, s_wsum);
        if (nb > 0) a.loff[i] = ex;
        if (gt == 0 && lb < nblk) a.btot[lb] = tot;
        if (TRACK) {
            int ttot;
            const int tex = group_excl_scan(nt, &ttot, s_wsum);
            if (nt > 0) a.TL.toff[i] = tex;
            if (gt == 0 && lb < nblk) a.TL.tbtot[lb] = ttot;
        }
    }
    if (flights) atomicAdd(&s_flights, flights);
    if (deaths) atomicAdd(&s_deaths, deaths);
    if (flags) atomicOr(&s_flags, flags);
    __syncthreads();
    if (threadIdx.x < a.T.nproc && s_ev[threadIdx.x])
        atomicAdd((unsigned long long *)&a.ctr[C_EVENTS + threadIdx.x], (unsigned long long)s_ev[threadIdx.x]);
    if (threadIdx.x == 0) {
        if (s_flights) atomicAdd((unsigned long long *)&a.ctr[C_FLIGHTS], (unsigned long long)s_flights);
        if (s_deaths) atomicAdd((unsigned long long *)&a.ctr[C_DEATHS], (unsigned long long)s_deaths);
        if (s_flags) atomicOr((unsigned long long *)&a.ctr[C_FLAGS], (unsigned long long)s_flags);
    }
}

__global__ __launch_bounds__(MCS_BLOCK) void k_place_births32(BirthLog32 L, int64_t count, const int32_t *loff,
                                                              const int64_t *bpre, int64_t n0, uint64_t next_id,
                                                              SplitBuf b, int32_t n0x, int32_t n1x, int32_t n2x)
{
    int64_t r = gidx();
    if (r >= count) return;
    int32_t par = L.parent[r];
    int64_t dst = n0 + bpre[par / MCS_BLOCK] + loff[par] + L.seq[r];
    int32_t c[3];
    unpack_cell(L.cell[r], c[0], c[1], c[2]);
    float off[3] = {L.off[0][r], L.off[1][r], L.off[2][r]};
    const int32_t nmax[3] = {n0x, n1x, n2x};
    rebase32(c, off, nmax);
    b.cell[dst] = pack_cell(c[0], c[1], c[2]);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        b.off[d][dst] = off[d];
        b.v[d][dst] = L.v[d][r];
    }
    b.w[dst] = L.w[r];
    b.s[dst] = L.s[r];
    b.active[dst] = L.act[r];
    b.id[dst] = next_id + (uint64_t)(dst - n0);
}

static BirthLog32 birth_log32(mcs_pop *p)
{
    BirthLog32 L;
    L.parent = p->log_parent;
    L.seq = p->log_seq;
    L.cell = p->log_cell;
    for (int d = 0; d < 3; ++d) {
        L.off[d] = p->log_off[d];
        L.v[d] = p->log_v32[d];
    }
    L.w = p->log_w32;
    L.s = p->log_s32;
    L.act = p->log_act;
    L.cap = p->logcap;
    return L;
}

// The device's compute units (one workgroup of k_advance32_lds each).
static int64_t adv_lds_cus(mcs_ctx *ctx)
{
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess || cus < 1)
        cus = 256;
    return cus;
}

// Whether k_advance32_lds may take tab_bytes of dynamic LDS (the attribute set
// once per process for each instantiation).
static bool adv_lds_ready(mcs_ctx *ctx, size_t tab_bytes)
{
    static int ok = -1;
    if (ok < 0) {
        int lds_max = 0;
        if (hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device) != hipSuccess)
            lds_max = 64 * 1024;
        const int want = 144 * 1024;
        ok = lds_max >= want + 4096 &&
             hipFuncSetAttribute((const void *)k_advance32_lds<1, 1>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 want) == hipSuccess &&
             hipFuncSetAttribute((const void *)k_advance32_lds<1, 0>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 want) == hipSuccess &&
             hipFuncSetAttribute((const void *)k_advance32_lds<0, 1>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 want) == hipSuccess &&
             hipFuncSetAttribute((const void *)k_advance32_lds<0, 0>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 want) == hipSuccess;
        if (!ok) (void)hipGetLastError();
    }
    return ok && tab_bytes <= 144 * 1024;
}

static int32_t launch_advance32(mcs_pop *p, const mcs_table *t, const mcs_field *field, double dt,
                                mcs_mesh *tracker, uint64_t seed, uint32_t instance, uint32_t step,
                                int64_t n0, int64_t nblk, int64_t *btot, int64_t *ctr)
{
    mcs_ctx *ctx = p->ctx;
    mcs_mesh *m = field->kind == MCS_FIELD_MESH ? field->mesh : tracker;
    if (m)
        for (int d = 0; d < 3; ++d)
            if (m->n[d] != p->gn[d] || m->h != p->gh)
                return mcs_fail(ctx, MCS_E_ARG, "advance: the mesh differs from the population's grid");
    AdvArgs32 a;
    a.b = p->sbuf[0];
    a.n0 = n0;
    a.T = table_view32(t);
    memset(&a.M, 0, sizeof(a.M));
    if (m) {
        if (field->kind == MCS_FIELD_MESH) {
            int32_t rc = mcs_mesh_faces32(m);
            if (rc) return rc;
        }
        for (int d = 0; d < 3; ++d) a.M.n[d] = m->n[d];
        a.M.h = (float)m->h;
        a.M.inv_h = (float)(1.0 / m->h);
        a.M.ex = m->face32[0];
        a.M.ey = m->face32[1];
        a.M.ez = m->face32[2];
        a.M.qfixed = m->arr[MCS_ARR_QFIXED];
        a.M.inv_dv = 1.0 / ((m->h * m->h) * m->h);
    }
    for (int d = 0; d < 3; ++d) {
        a.acc[d] = (float)(MCS_QM * field->e[d]);
        a.nmax[d] = p->gn[d];
    }
    a.dt = (float)dt;
    a.inv_h = (float)(1.0 / p->gh);
    a.seed = seed;
    a.instance = instance;
    a.step = step;
    a.L = birth_log32(p);
    a.loff = p->i32a;
    a.btot = btot;
    a.ctr = ctr;
    a.TL = track_log(p);
    ProfScope ps(ctx, "advance");
    const int fm = field->kind == MCS_FIELD_MESH, tr = tracker != nullptr;
    // MCS_ADV_LDS=1: the table in LDS (k_advance32_lds) when it fits beside the
    // kernel's own and the process kinds travel as bits.  Same results; on the
    // 256x256x512 workload 12.9 ms per launch against 11.2 ms reading the table
    // from global memory, where it stays in L2, so the default reads it there.
    const size_t tab_bytes = sizeof(float) * ((size_t)t->ne * (t->nproc + 1) + (size_t)t->nproc);
    const char *le = getenv("MCS_ADV_LDS");
    if (le && le[0] == '1' && a.T.kindbits != 0 && tab_bytes <= 144 * 1024 && adv_lds_ready(ctx, tab_bytes)) {
        const int64_t nchunks = cdiv(n0, (int64_t)ADV_LDS_THREADS);
        const dim3 gl((unsigned)std::min<int64_t>(nchunks, adv_lds_cus(ctx))), bll(ADV_LDS_THREADS);
        if (fm && tr)
            hipLaunchKernelGGL((k_advance32_lds<1, 1>), gl, bll, tab_bytes, ctx->stream, a, nchunks, nblk);
        else if (fm)
            hipLaunchKernelGGL((k_advance32_lds<1, 0>), gl, bll, tab_bytes, ctx->stream, a, nchunks, nblk);
        else if (tr)
            hipLaunchKernelGGL((k_advance32_lds<0, 1>), gl, bll, tab_bytes, ctx->stream, a, nchunks, nblk);
        else
            hipLaunchKernelGGL((k_advance32_lds<0, 0>), gl, bll, tab_bytes, ctx->stream, a, nchunks, nblk);
        MCS_LAUNCH_CHECK(ctx);
        return MCS_OK;
    }
    dim3 g((unsigned)nblk), bl(MCS_BLOCK);
    if (fm && tr)
        hipLaunchKernelGGL((k_advance32<1, 1>), g, bl, 0, ctx->stream, a);
    else if (fm)
        hipLaunchKernelGGL((k_advance32<1, 0>), g, bl, 0, ctx->stream, a);
    else if (tr)
        hipLaunchKernelGGL((k_advance32<0, 1>), g, bl, 0, ctx->stream, a);
    else
        hipLaunchKernelGGL((k_advance32<0, 0>), g, bl, 0, ctx->stream, a);
    MCS_LAUNCH_CHECK(ctx);
    return MCS_OK;
}

// Deposits placed in parent order: the deposit logged at slot r, the seq-th of
// its parent, goes to (deposits of earlier blocks) + (of earlier threads in the
// block) + seq.
__global__ __launch_bounds__(MCS_BLOCK) void k_track_place(TrackLog TL, int64_t count, const int64_t *tpre,
                                                           uint32_t *keys, double *vals)
{
    const int64_t r = gidx();
    if (r >= count) return;
    const int32_t par = TL.parent[r];
    const int64_t dst = tpre[par / MCS_BLOCK] + TL.toff[par] + TL.seq[r];
    keys[dst] = (uint32_t)TL.cell[r];
    vals[dst] = TL.val[r];
}

// The deposits sorted stably by cell: the first thread of each cell's run adds
// the run to the cell's value in order, ((q + d1) + d2) + ..., as the
// reference's sequential loop does.
__global__ __launch_bounds__(MCS_BLOCK) void k_track_sum(const uint32_t *keys, const double *vals, int64_t count,
                                                         double *qfixed)
{
    const int64_t r = gidx();
    if (r >= count) return;
    const uint32_t c = keys[r];
    if (r > 0 && keys[r - 1] == c) return;
    double q = qfixed[c];
    for (int64_t j = r; j < count && keys[j] == c; ++j) q = q + vals[j];
    qfixed[c] = q;
}

// The tracker's array and its cell count (the ghosted array's length), or none.
struct TrackTarget {
    double *qfixed = nullptr;
    int64_t ncells = 0;
};

// The step's tracker deposits (count of them, logged in p's track log) added to
// the target array in parent order per cell.
static int32_t track_apply(mcs_pop *p, const TrackTarget &tt, int64_t count, const int64_t *tpre)
{
    mcs_ctx *ctx = p->ctx;
    if (count == 0) return MCS_OK;
    uint32_t *keys = (uint32_t *)p->i32c, *keys_out = (uint32_t *)p->trk_cell;
    double *vals = (double *)p->u64a, *vals_out = p->trk_val;
    ProfScope ps(ctx, "track");
    hipLaunchKernelGGL(k_track_place, dim3((unsigned)cdiv(count, MCS_BLOCK)), dim3(MCS_BLOCK), 0, ctx->stream,
                       track_log(p), count, tpre, keys, vals);
    MCS_LAUNCH_CHECK(ctx);
    unsigned bits = 1;
    while (bits < 32 && ((uint64_t)(tt.ncells - 1) >> bits) != 0) ++bits;
    size_t tb = 0;
    MCS_HIP(ctx, rocprim::radix_sort_pairs(nullptr, tb, keys, keys_out, vals, vals_out, (size_t)count, 0u, bits,
                                           ctx->stream));
    void *tmp = nullptr;
    int32_t rc = mcs_scratch(ctx, tb, &tmp);
    if (rc) return rc;
    MCS_HIP(ctx, rocprim::radix_sort_pairs(tmp, tb, keys, keys_out, vals, vals_out, (size_t)count, 0u, bits,
                                           ctx->stream));
    hipLaunchKernelGGL(k_track_sum, dim3((unsigned)cdiv(count, MCS_BLOCK)), dim3(MCS_BLOCK), 0, ctx->stream,
                       (const uint32_t *)keys_out, (const double *)vals_out, count, tt.qfixed);
    MCS_LAUNCH_CHECK(ctx);
    return MCS_OK;
}

// Photon emission of an electron step: the photons of the logged bursts go to
// `photons`, drawn from streams of their own ids.
struct EmitArgs {
    mcs_pop *photons = nullptr;
    const double *photo = nullptr;   // the table's photoemission parameters, 4 per row
    uint64_t seed = 0;
    uint32_t instance = 0, step = 0;
};

__global__ void k_zero_u64(uint64_t *a, int64_t n)
{
    const int64_t i = gidx();
    if (i < n) a[i] = 0;
}

// Each logged burst (all hold at least one photon, so their places differ) marks
// the place of its first photon among the step's photons with (place << 32) | slot.
__global__ __launch_bounds__(MCS_BLOCK) void k_burst_mark(BurstLog BL, int64_t count, const int64_t *ppre,
                                                          uint64_t *marks)
{
    const int64_t r = gidx();
    if (r >= count) return;
    const int32_t par = BL.parent[r];
    const int64_t start = ppre[par / MCS_BLOCK] + BL.poff[par] + BL.base[r];
    marks[start] = ((uint64_t)start << 32) | (uint64_t)r;
}

// Photon q of the step, whose burst the running maximum of the marks names, at
// slot n0 + q with id next_id + q: c randsphere(), nu and s from the stream of its
// id (collisions.jl:134-139), weight photon_weight, the burst's position and
// active flag.
__global__ __launch_bounds__(MCS_BLOCK) void k_place_photons(const uint64_t *owner, int64_t total, BurstLog BL,
                                                             const double *photo, int64_t n0, uint64_t next_id,
                                                             PopBuf b, uint64_t seed, uint32_t instance,
                                                             uint32_t step)
{
    const int64_t q = gidx();
    if (q >= total) return;
    const uint32_t r = (uint32_t)(owner[q] & 0xffffffffull);
    const int32_t row = BL.row[r];
    const double lmin = photo[4 * row], lmax = photo[4 * row + 1], pw = photo[4 * row + 3];
    const int64_t dst = n0 + q;
    const uint64_t id = next_id + (uint64_t)q;
    double v[3], nu, s;
    mcs_photon_emit(seed, instance, id, step, lmin, lmax, v, &nu, &s);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        b.x[d][dst] = BL.x[d][r];
        b.v[d][dst] = v[d];
    }
    b.nu[dst] = nu;
    b.w[dst] = pw;
    b.s[dst] = s;
    b.active[dst] = BL.act[r];
    b.id[dst] = id;
}

// The step's `total` photons from its `nbur` logged bursts, appended to the
// photon population in parent order (apply! of MultiplePhotonOutcome,
// collisions.jl:125-141).
static int32_t emit_photons(mcs_pop *p, const EmitArgs &em, int64_t nbur, int64_t total, const int64_t *ppre)
{
    mcs_ctx *ctx = p->ctx;
    mcs_pop *ph = em.photons;
    ProfScope ps(ctx, "photon_emit");
    size_t tb = 0;
    uint64_t *none = nullptr;
    MCS_HIP(ctx, rocprim::inclusive_scan(nullptr, tb, none, none, (size_t)total, rocprim::maximum<uint64_t>(),
                                         ctx->stream));
    void *scr = nullptr;
    int32_t rc = mcs_scratch(ctx, 2 * sizeof(uint64_t) * (size_t)total + tb + 256, &scr);
    if (rc) return rc;
    uint64_t *marks = (uint64_t *)scr, *owner = marks + total;
    void *tmp = (void *)(((uintptr_t)(owner + total) + 255) & ~(uintptr_t)255);
    hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)cdiv(total, MCS_BLOCK)), dim3(MCS_BLOCK), 0, ctx->stream, marks,
                       total);
    hipLaunchKernelGGL(k_burst_mark, dim3((unsigned)cdiv(nbur, MCS_BLOCK)), dim3(MCS_BLOCK), 0, ctx->stream,
                       burst_log(p), nbur, ppre, marks);
    MCS_LAUNCH_CHECK(ctx);
    MCS_HIP(ctx, rocprim::inclusive_scan(tmp, tb, marks, owner, (size_t)total, rocprim::maximum<uint64_t>(),
                                         ctx->stream));
    hipLaunchKernelGGL(k_place_photons, dim3((unsigned)cdiv(total, MCS_BLOCK)), dim3(MCS_BLOCK), 0, ctx->stream,
                       (const uint64_t *)owner, total, burst_log(p), em.photo, ph->n, ph->next_id, ph->buf[0],
                       em.seed, em.instance, em.step);
    MCS_LAUNCH_CHECK(ctx);
    ph->n += total;
    ph->next_id += (uint64_t)total;
    mcs_pop_touch(ph);
    return MCS_OK;
}

// After the step kernel of population p: the births' prefix over blocks (and the
// tracker deposits', and the photons'), the counters and flags in one read-back,
// then the children placed after the particles of `dst` (p itself, or the
// electrons for the photon loop), the photons emitted and the deposits added to
// the tracker's array.  nev table rows are counted.
static int32_t advance_finish(mcs_pop *p, mcs_pop *dst, int nev, int64_t nblk, int64_t *btot, int64_t *bpre,
                              int64_t *ctr, const TrackTarget &tt, mcs_step_stats *stats,
                              const EmitArgs *em = nullptr)
{
    mcs_ctx *ctx = p->ctx;
    int32_t rc = mcs_scan_i64_to(ctx, btot, bpre, nblk, ctr + C_BIRTHS);
    if (rc) return rc;
    int64_t *tpre = p->tblk + (nblk + 1);
    if (tt.qfixed) {
        rc = mcs_scan_i64_to(ctx, p->tblk, tpre, nblk, ctr + C_TTOT);
        if (rc) return rc;
    }
    int64_t *ppre = em ? p->pblk + (nblk + 1) : nullptr;
    if (em) {
        rc = mcs_scan_i64_to(ctx, p->pblk, ppre, nblk, ctr + C_PTOT);
        if (rc) return rc;
    }
    int64_t h[C_PTOT + 1];
    rc = mcs_readback(ctx, ctr, h, em ? C_PTOT + 1 : C_EVENTS + nev);
    if (rc) return rc;
    const int64_t logged = h[C_LOG], births = h[C_BIRTHS];
    const int64_t tlogged = tt.qfixed ? h[C_TLOG] : 0, tcount = tt.qfixed ? h[C_TTOT] : 0;
    const int64_t blogged = em ? h[C_BLOG] : 0, ptot = em ? h[C_PTOT] : 0;
    int32_t flags = (int32_t)h[C_FLAGS];
    if (stats) {
        stats->births = births;
        stats->deaths = h[C_DEATHS];
        stats->flights = h[C_FLIGHTS];
        for (int j = 0; j < nev; ++j) stats->events[j] = h[C_EVENTS + j];
        stats->flags = flags;
    }
    const int64_t nd = dst->n;
    if (logged > p->logcap || nd + births > dst->cap || tlogged > p->logcap) {
        if (stats) stats->flags |= MCS_FLAG_CAPACITY;
        return mcs_fail(ctx, MCS_E_CAPACITY, "Maximum number of particles reached (%lld + %lld > %lld)",
                        (long long)nd, (long long)births, (long long)dst->cap);
    }
    if (em && (blogged > p->logcap || em->photons->n + ptot > em->photons->cap)) {
        if (stats) stats->flags |= MCS_FLAG_CAPACITY;
        return mcs_fail(ctx, MCS_E_CAPACITY, "Maximum number of photons reached (%lld + %lld > %lld)",
                        (long long)em->photons->n, (long long)ptot, (long long)em->photons->cap);
    }
    if (tcount != tlogged)
        return mcs_fail(ctx, MCS_E_HIP, "advance: %lld tracker deposits logged, %lld counted", (long long)tlogged,
                        (long long)tcount);
    if (births > 0) {
        if (dst != p) mcs_pop_touch(dst);
        if (dst->prec == MCS_PREC_F32)
            hipLaunchKernelGGL(k_place_births32, dim3((unsigned)cdiv(births, MCS_BLOCK)), dim3(MCS_BLOCK), 0,
                               ctx->stream, birth_log32(p), births, (const int32_t *)p->i32a,
                               (const int64_t *)bpre, nd, dst->next_id, dst->sbuf[0], dst->gn[0], dst->gn[1],
                               dst->gn[2]);
        else
            hipLaunchKernelGGL(k_place_births, dim3((unsigned)cdiv(births, MCS_BLOCK)), dim3(MCS_BLOCK), 0,
                               ctx->stream, birth_log(p), births, (const int32_t *)p->i32a,
                               (const int64_t *)bpre, nd, dst->next_id, dst->buf[0]);
        MCS_LAUNCH_CHECK(ctx);
        dst->n = nd + births;
        dst->next_id += (uint64_t)births;
    }
    if (em && ptot > 0) {
        rc = emit_photons(p, *em, blogged, ptot, ppre);
        if (rc) return rc;
    }
    if (tt.qfixed) {
        rc = track_apply(p, tt, tcount, tpre);
        if (rc) return rc;
    }
    if (stats) stats->n_out = p->n;
    if (flags & MCS_FLAG_FIELD) return mcs_fail(ctx, MCS_E_FIELD, "Too high electric field");
    if (flags & MCS_FLAG_NAN) return mcs_fail(ctx, MCS_E_NAN, "NaN particle energy (a NaN velocity)");
    if (flags & MCS_FLAG_TABLE_RANGE)
        return mcs_fail(ctx, MCS_E_TABLE_RANGE, "particle energy beyond the collision table");
    return MCS_OK;
}

// The electron loop of advance! (mcs_advance), emitting photons when em is set.
static int32_t electron_step(mcs_pop *p, const mcs_table *t, const mcs_field *field, double dt, mcs_mesh *tracker,
                             uint64_t seed, uint32_t instance, uint32_t step, mcs_step_stats *stats,
                             const EmitArgs *em)
{
    mcs_ctx *ctx = p->ctx;
    hipSetDevice(ctx->device);
    if (!(dt >= 0.0)) return mcs_fail(ctx, MCS_E_ARG, "advance: dt must be >= 0");
    if (field->kind == MCS_FIELD_MESH && !field->mesh)
        return mcs_fail(ctx, MCS_E_ARG, "advance: mesh field without a mesh");
    if (field->kind == MCS_FIELD_MESH && tracker && tracker != field->mesh)
        return mcs_fail(ctx, MCS_E_ARG, "advance: tracker mesh must be the field mesh");
    mcs_pop_touch(p);
    if (stats) memset(stats, 0, sizeof(*stats));
    const int64_t n0 = p->n;
    if (stats) {
        stats->n_in = n0;
        stats->n_out = n0;
    }
    if (n0 == 0) return MCS_OK;
    int64_t nblk = cdiv(n0, MCS_BLOCK);
    int64_t *btot = p->blk;
    int64_t *bpre = p->blk + (nblk + 1);
    int64_t *ctr = ctx->d_small;
    TrackTarget tt;
    if (tracker) {
        int32_t rc = mcs_pop_ensure_track(p);
        if (rc) return rc;
        tt.qfixed = tracker->arr[MCS_ARR_QFIXED];
        tt.ncells = tracker->ncg;
    }
    if (em) {
        int32_t rc = mcs_pop_ensure_bursts(p);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_fill_i64, dim3(1), dim3(64), 0, ctx->stream, ctr, (int64_t)64, (int64_t)0);
    if (p->prec == MCS_PREC_F32) {
        int32_t rc = launch_advance32(p, t, field, dt, tracker, seed, instance, step, n0, nblk, btot, ctr);
        if (rc) return rc;
    } else {
        AdvArgs a;
        a.b = p->buf[0];
        a.n0 = n0;
        a.T = table_view(t);
        a.M = mesh_view(field->kind == MCS_FIELD_MESH ? field->mesh : tracker);
        for (int d = 0; d < 3; ++d) a.e[d] = field->e[d];
        a.dt = dt;
        a.seed = seed;
        a.instance = instance;
        a.step = step;
        a.L = birth_log(p);
        a.loff = p->i32a;
        a.btot = btot;
        a.ctr = ctr;
        a.TL = track_log(p);
        ProfScope ps(ctx, "advance");
        const int fm = field->kind == MCS_FIELD_MESH;
        const int tr = tracker != nullptr;
        dim3 g((unsigned)nblk), bl(MCS_BLOCK);
        if (em) {
            if (fm && tr)
                hipLaunchKernelGGL((k_advance_photo<1, 1>), g, bl, 0, ctx->stream, a, burst_log(p));
            else if (fm)
                hipLaunchKernelGGL((k_advance_photo<1, 0>), g, bl, 0, ctx->stream, a, burst_log(p));
            else if (tr)
                hipLaunchKernelGGL((k_advance_photo<0, 1>), g, bl, 0, ctx->stream, a, burst_log(p));
            else
                hipLaunchKernelGGL((k_advance_photo<0, 0>), g, bl, 0, ctx->stream, a, burst_log(p));
        } else if (fm && tr) {
            hipLaunchKernelGGL((k_advance<1, 1>), g, bl, 0, ctx->stream, a);
        } else if (fm) {
            hipLaunchKernelGGL((k_advance<1, 0>), g, bl, 0, ctx->stream, a);
        } else if (tr) {
            hipLaunchKernelGGL((k_advance<0, 1>), g, bl, 0, ctx->stream, a);
        } else {
            hipLaunchKernelGGL((k_advance<0, 0>), g, bl, 0, ctx->stream, a);
        }
    }
    MCS_LAUNCH_CHECK(ctx);
    return advance_finish(p, p, t->nproc, nblk, btot, bpre, ctr, tt, stats, em);
}

// The electron loop of advance! in the axisymmetric field of grid g, tracking into
// g's QFIXED when track != 0, emitting photons when em is set.
static int32_t electron_step_cyl(mcs_pop *p, const mcs_table *t, mcs_cgrid *g, double dt, int32_t track,
                                 uint64_t seed, uint32_t instance, uint32_t step, mcs_step_stats *stats,
                                 const EmitArgs *em)
{
    mcs_ctx *ctx = p->ctx;
    hipSetDevice(ctx->device);
    if (!(dt >= 0.0)) return mcs_fail(ctx, MCS_E_ARG, "advance: dt must be >= 0");
    if (p->prec != MCS_PREC_F64) return mcs_fail(ctx, MCS_E_ARG, "advance_cyl: FP64 populations only");
    if (g->ctx != ctx) return mcs_fail(ctx, MCS_E_ARG, "advance_cyl: grid of another context");
    mcs_pop_touch(p);
    if (stats) memset(stats, 0, sizeof(*stats));
    const int64_t n0 = p->n;
    if (stats) {
        stats->n_in = n0;
        stats->n_out = n0;
    }
    if (n0 == 0) return MCS_OK;
    int64_t nblk = cdiv(n0, MCS_BLOCK);
    int64_t *btot = p->blk;
    int64_t *bpre = p->blk + (nblk + 1);
    int64_t *ctr = ctx->d_small;
    TrackTarget tt;
    if (track) {
        int32_t rc = mcs_pop_ensure_track(p);
        if (rc) return rc;
        tt.qfixed = g->arr[MCS_CARR_QFIXED];
        tt.ncells = (int64_t)(g->M + 2) * (g->N + 2);
    }
    if (em) {
        int32_t rc = mcs_pop_ensure_bursts(p);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_fill_i64, dim3(1), dim3(64), 0, ctx->stream, ctr, (int64_t)64, (int64_t)0);
    AdvArgs a;
    a.b = p->buf[0];
    a.n0 = n0;
    a.T = table_view(t);
    a.M = cyl_view(g);
    for (int d = 0; d < 3; ++d) a.e[d] = 0.0;
    a.dt = dt;
    a.seed = seed;
    a.instance = instance;
    a.step = step;
    a.L = birth_log(p);
    a.loff = p->i32a;
    a.btot = btot;
    a.ctr = ctr;
    a.TL = track_log(p);
    {
        ProfScope ps(ctx, "advance");
        dim3 gr((unsigned)nblk), bl(MCS_BLOCK);
        if (em && track)
            hipLaunchKernelGGL((k_advance_photo<2, 1>), gr, bl, 0, ctx->stream, a, burst_log(p));
        else if (em)
            hipLaunchKernelGGL((k_advance_photo<2, 0>), gr, bl, 0, ctx->stream, a, burst_log(p));
        else if (track)
            hipLaunchKernelGGL((k_advance<2, 1>), gr, bl, 0, ctx->stream, a);
        else
            hipLaunchKernelGGL((k_advance<2, 0>), gr, bl, 0, ctx->stream, a);
    }
    MCS_LAUNCH_CHECK(ctx);
    return advance_finish(p, p, t->nproc, nblk, btot, bpre, ctr, tt, stats, em);
}

static int32_t single_population_checks(mcs_pop *p, const mcs_table *t)
{
    if (p->species == MCS_SPECIES_PHOTON)
        return mcs_fail(p->ctx, MCS_E_ARG, "advance: photons move in mcs_advance_mixed");
    if (t->nphoto > 0)
        return mcs_fail(p->ctx, MCS_E_ARG, "advance: a table with photoemission rows needs mcs_advance_mixed");
    return MCS_OK;
}

extern "C" int32_t mcs_advance(mcs_pop *p, const mcs_table *t, const mcs_field *field, double dt,
                               mcs_mesh *tracker, uint64_t seed, uint32_t instance, uint32_t step,
                               mcs_step_stats *stats)
{
    if (!p || !t || !field) return MCS_E_ARG;
    int32_t rc = single_population_checks(p, t);
    if (rc) return rc;
    return electron_step(p, t, field, dt, tracker, seed, instance, step, stats, nullptr);
}

extern "C" int32_t mcs_advance_cyl(mcs_pop *p, const mcs_table *t, mcs_cgrid *g, double dt, int32_t track,
                                   uint64_t seed, uint32_t instance, uint32_t step, mcs_step_stats *stats)
{
    if (!p || !t || !g) return MCS_E_ARG;
    int32_t rc = single_population_checks(p, t);
    if (rc) return rc;
    return electron_step_cyl(p, t, g, dt, track, seed, instance, step, stats, nullptr);
}

// ---------------------------------------------------------------- photons

struct PhotonArgs {
    PopBuf b;
    int64_t n0;
    double numax, dt;
    uint64_t seed;
    uint32_t instance, step;
    BirthLog L;
    int32_t *loff;
    int64_t *btot;
    int64_t *ctr;
    TrackLog TL;
    MeshView M;
};

// The photon loop of advance! (mixed_population.jl:27-51) with
// ZhelezniakCollisions(numax) (photon.jl:27-56): one thread per photon slot that
// exists when the loop starts.  Free flights x += t v to the next (null) collision
// at s/numax; a collision draws xi = numax U and absorbs when nu > xi
// (PhotoIonization): the new electron (at x, at rest, weight w, s = nextcoll(),
// the photon's active flag) goes to the birth log, the tracker (TK 1: 3D mesh,
// 2: axisymmetric grid, 0: none) gets +w at x, the photon is removed; then the
// photon's s = nextcoll().  A removed photon keeps flying, as in the reference.
template <int TK>
__global__ __launch_bounds__(MCS_BLOCK) void k_photon_advance(PhotonArgs a)
{
    __shared__ int s_flights, s_deaths, s_abs, s_flags;
    if (threadIdx.x == 0) {
        s_flights = 0;
        s_deaths = 0;
        s_abs = 0;
        s_flags = 0;
    }
    __syncthreads();
    const int64_t i = xcd_gidx();
    int nb = 0, nt = 0, flights = 0, deaths = 0, absorbed = 0;
    int32_t flags = 0;
    if (i < a.n0 && a.b.active[i]) {
        double x[3] = {a.b.x[0][i], a.b.x[1][i], a.b.x[2][i]};
        const double v[3] = {a.b.v[0][i], a.b.v[1][i], a.b.v[2][i]};
        const double nu = a.b.nu[i], w = a.b.w[i];
        double s = a.b.s[i];
        uint8_t act = 1;
        mcs_stream rs;
        mcs_stream_init(&rs, a.seed, a.instance, MCS_PHASE_PHOTON, a.b.id[i], a.step);
        double trem = a.dt;
        while (trem > 0.0) {
            const double tnext = s / a.numax;
            double tt;
            bool collides;
            if (trem > tnext) {
                tt = tnext;
                collides = true;
            } else {
                tt = trem;
                collides = false;
                s = s - tt * a.numax;
            }
#pragma unroll
            for (int d = 0; d < 3; ++d) x[d] = x[d] + tt * v[d];
            flights += 1;
            if (collides) {
                const double xi = mcs_u01(&rs) * a.numax;
                if (nu > xi) {
                    absorbed += 1;
                    const double sc = mcs_nextcoll(&rs);
                    if (TK) {
                        int32_t cell;
                        double amount;
                        if (track_cell<TK == 2 ? 2 : 1>(a.M, x, w, cell, amount))
                            track_push(a.TL, a.ctr, i, nt, cell, amount, flags);
                    }
                    const unsigned long long slot = atomicAdd((unsigned long long *)&a.ctr[C_LOG], 1ULL);
                    if ((int64_t)slot < a.L.cap) {
                        a.L.parent[slot] = (int32_t)i;
                        a.L.seq[slot] = nb;
#pragma unroll
                        for (int d = 0; d < 3; ++d) {
                            a.L.x[d][slot] = x[d];
                            a.L.v[d][slot] = 0.0;
                        }
                        a.L.w[slot] = w;
                        a.L.s[slot] = sc;
                        a.L.act[slot] = act;
                    } else {
                        flags |= MCS_FLAG_CAPACITY;
                    }
                    nb += 1;
                    if (act) deaths += 1;
                    act = 0;
                }
                s = mcs_nextcoll(&rs);
            }
            trem = trem - tt;
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) a.b.x[d][i] = x[d];
        a.b.s[i] = s;
        if (!act) a.b.active[i] = 0;
    }
    int tot;
    const int ex = block_excl_scan<int>(nb, &tot);
    if (nb > 0) a.loff[i] = ex;
    if (threadIdx.x == 0) a.btot[xcd_block()] = tot;
    if (TK) track_block_counts(a.TL, i, nt);
    if (flights) atomicAdd(&s_flights, flights);
    if (deaths) atomicAdd(&s_deaths, deaths);
    if (absorbed) atomicAdd(&s_abs, absorbed);
    if (flags) atomicOr(&s_flags, flags);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (s_flights) atomicAdd((unsigned long long *)&a.ctr[C_FLIGHTS], (unsigned long long)s_flights);
        if (s_deaths) atomicAdd((unsigned long long *)&a.ctr[C_DEATHS], (unsigned long long)s_deaths);
        if (s_abs) atomicAdd((unsigned long long *)&a.ctr[C_EVENTS], (unsigned long long)s_abs);
        if (s_flags) atomicOr((unsigned long long *)&a.ctr[C_FLAGS], (unsigned long long)s_flags);
    }
}

// The photon loop over the photons of ph, appending photoelectrons to el; tk and
// M select the tracker (1: the 3D mesh of M, 2: the axisymmetric grid of M).
static int32_t photon_step(mcs_pop *ph, mcs_pop *el, double numax, double dt, const MeshView &M, int tk,
                           const TrackTarget &tt, uint64_t seed, uint32_t instance, uint32_t step,
                           mcs_step_stats *stats)
{
    mcs_ctx *ctx = ph->ctx;
    mcs_pop_touch(ph);
    const int64_t n0 = ph->n;
    if (stats) stats->n_out = n0;
    if (n0 == 0) return MCS_OK;
    const int64_t nblk = cdiv(n0, MCS_BLOCK);
    int64_t *btot = ph->blk;
    int64_t *bpre = ph->blk + (nblk + 1);
    int64_t *ctr = ctx->d_small;
    if (tk) {
        int32_t rc = mcs_pop_ensure_track(ph);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_fill_i64, dim3(1), dim3(64), 0, ctx->stream, ctr, (int64_t)64, (int64_t)0);
    PhotonArgs a;
    a.b = ph->buf[0];
    a.n0 = n0;
    a.numax = numax;
    a.dt = dt;
    a.seed = seed;
    a.instance = instance;
    a.step = step;
    a.L = birth_log(ph);
    a.loff = ph->i32a;
    a.btot = btot;
    a.ctr = ctr;
    a.TL = track_log(ph);
    a.M = M;
    {
        ProfScope ps(ctx, "photon_advance");
        dim3 g((unsigned)nblk), bl(MCS_BLOCK);
        if (tk == 2)
            hipLaunchKernelGGL(k_photon_advance<2>, g, bl, 0, ctx->stream, a);
        else if (tk == 1)
            hipLaunchKernelGGL(k_photon_advance<1>, g, bl, 0, ctx->stream, a);
        else
            hipLaunchKernelGGL(k_photon_advance<0>, g, bl, 0, ctx->stream, a);
    }
    MCS_LAUNCH_CHECK(ctx);
    return advance_finish(ph, el, 1, nblk, btot, bpre, ctr, tt, stats, nullptr);
}

static int32_t mixed_checks(mcs_pop *e, mcs_pop *ph, const mcs_table *t, double numax)
{
    mcs_ctx *ctx = e->ctx;
    if (e->species != MCS_SPECIES_ELECTRON)
        return mcs_fail(ctx, MCS_E_ARG, "advance_mixed: the first population must hold electrons");
    if (t->ctx != ctx) return mcs_fail(ctx, MCS_E_ARG, "advance_mixed: table of another context");
    if (!ph) {
        if (t->nphoto > 0)
            return mcs_fail(ctx, MCS_E_ARG, "advance_mixed: a table with photoemission rows needs photons");
        return MCS_OK;
    }
    if (ph->species != MCS_SPECIES_PHOTON)
        return mcs_fail(ctx, MCS_E_ARG, "advance_mixed: the second population must hold photons");
    if (ph->ctx != ctx) return mcs_fail(ctx, MCS_E_ARG, "advance_mixed: photons of another context");
    if (e->prec != MCS_PREC_F64)
        return mcs_fail(ctx, MCS_E_ARG, "advance_mixed: FP64 electron populations only");
    if (!(numax > 0.0)) return mcs_fail(ctx, MCS_E_ARG, "advance_mixed: numax must be positive");
    for (int j = 0; j < t->nproc; ++j)
        if (t->h_kind[j] == MCS_PROC_PHOTOEMISSION && !t->photo_set[j])
            return mcs_fail(ctx, MCS_E_ARG, "advance_mixed: photoemission row %d has no parameters", j);
    return MCS_OK;
}

extern "C" int32_t mcs_advance_mixed(mcs_pop *e, mcs_pop *ph, const mcs_table *t, double numax,
                                     const mcs_field *field, double dt, mcs_mesh *tracker, uint64_t seed,
                                     uint32_t instance, uint32_t step, mcs_step_stats *estats,
                                     mcs_step_stats *pstats)
{
    if (!e || !t || !field) return MCS_E_ARG;
    int32_t rc = mixed_checks(e, ph, t, numax);
    if (rc) return rc;
    if (pstats) memset(pstats, 0, sizeof(*pstats));
    if (!ph) return electron_step(e, t, field, dt, tracker, seed, instance, step, estats, nullptr);
    if (pstats) pstats->n_in = ph->n;
    EmitArgs em;
    em.photons = ph;
    em.photo = t->photo;
    em.seed = seed;
    em.instance = instance;
    em.step = step;
    rc = electron_step(e, t, field, dt, tracker, seed, instance, step, estats, t->nphoto > 0 ? &em : nullptr);
    if (rc) return rc;
    TrackTarget tt;
    if (tracker) {
        tt.qfixed = tracker->arr[MCS_ARR_QFIXED];
        tt.ncells = tracker->ncg;
    }
    return photon_step(ph, e, numax, dt, mesh_view(tracker), tracker ? 1 : 0, tt, seed, instance, step, pstats);
}

extern "C" int32_t mcs_advance_cyl_mixed(mcs_pop *e, mcs_pop *ph, const mcs_table *t, double numax, mcs_cgrid *g,
                                         double dt, int32_t track, uint64_t seed, uint32_t instance,
                                         uint32_t step, mcs_step_stats *estats, mcs_step_stats *pstats)
{
    if (!e || !t || !g) return MCS_E_ARG;
    int32_t rc = mixed_checks(e, ph, t, numax);
    if (rc) return rc;
    if (pstats) memset(pstats, 0, sizeof(*pstats));
    if (!ph) return electron_step_cyl(e, t, g, dt, track, seed, instance, step, estats, nullptr);
    if (pstats) pstats->n_in = ph->n;
    EmitArgs em;
    em.photons = ph;
    em.photo = t->photo;
    em.seed = seed;
    em.instance = instance;
    em.step = step;
    rc = electron_step_cyl(e, t, g, dt, track, seed, instance, step, estats, t->nphoto > 0 ? &em : nullptr);
    if (rc) return rc;
    TrackTarget tt;
    if (track) {
        tt.qfixed = g->arr[MCS_CARR_QFIXED];
        tt.ncells = (int64_t)(g->M + 2) * (g->N + 2);
    }
    return photon_step(ph, e, numax, dt, cyl_view(g), track ? 2 : 0, tt, seed, instance, step, pstats);
}

// ---------------------------------------------------------------- repack!

__global__ __launch_bounds__(MCS_BLOCK) void k_count_active(const uint8_t *act, int64_t n,
                                                            int64_t *btot)
{
    int64_t i = gidx();
    int a = (i < n && act[i]) ? 1 : 0;
    int tot;
    block_excl_scan<int>(a, &tot);
    if (threadIdx.x == 0) btot[blockIdx.x] = tot;
}

// Rank matching equivalent of the two-pointer loop of population.jl:191-215: the
// k-th hole below A (the number of actives) receives the k-th active from the top.
__global__ __launch_bounds__(MCS_BLOCK) void k_repack_ranks(const uint8_t *act, int64_t n,
                                                            const int64_t *bpre, int64_t A,
                                                            int32_t *holes, int32_t *srcs)
{
    int64_t i = gidx();
    int a = (i < n && act[i]) ? 1 : 0;
    int tot;
    int ex = block_excl_scan<int>(a, &tot);
    if (i >= n) return;
    int64_t before = bpre[blockIdx.x] + ex;
    if (i < A && !a) holes[i - before] = (int32_t)i;
    if (i >= A && a) srcs[A - before - 1] = (int32_t)i;
}

__global__ __launch_bounds__(MCS_BLOCK) void k_repack_move(PopBuf b, const int32_t *holes,
                                                           const int32_t *srcs, int64_t h)
{
    int64_t r = gidx();
    if (r >= h) return;
    copy_particle(b, srcs[r], b, holes[r]);
}

static int32_t count_actives(mcs_pop *p, int64_t *A, int64_t **bpre_out)
{
    mcs_ctx *ctx = p->ctx;
    int64_t n = p->n;
    int64_t nblk = cdiv(n, MCS_BLOCK);
    int64_t *btot = p->blk;
    int64_t *bpre = p->blk + (nblk + 1);
    const uint8_t *act = p->prec == MCS_PREC_F32 ? p->sbuf[0].active : p->buf[0].active;
    hipLaunchKernelGGL(k_count_active, dim3((unsigned)nblk), dim3(MCS_BLOCK), 0, ctx->stream,
                       act, n, btot);
    MCS_LAUNCH_CHECK(ctx);
    int32_t rc = mcs_scan_i64(ctx, btot, bpre, nblk);
    if (rc) return rc;
    rc = mcs_readback(ctx, bpre + nblk, A, 1);
    *bpre_out = bpre;
    return rc;
}

// Energies (J) of the active particles in slot order (energy, JuMC/src/electron.jl:27),
// the sample of the swarm's Kolmogorov-Smirnov test (Swarm/src/Swarm.jl:288-306).
template <typename Buf>
__global__ __launch_bounds__(MCS_BLOCK) void k_active_energy(Buf b, int64_t n, const int64_t *bpre, double *out)
{
    const int64_t i = gidx();
    const int a = (i < n && b.active[i]) ? 1 : 0;
    int tot;
    const int ex = block_excl_scan<int>(a, &tot);
    if (!a) return;
    const double v0 = b.v[0][i], v1 = b.v[1][i], v2 = b.v[2][i];
    out[bpre[blockIdx.x] + ex] = energy_dev(v0, v1, v2);
}

extern "C" int32_t mcs_pop_energies(mcs_pop *p, int64_t nmax, int64_t *nout, double *energy)
{
    if (!p || !nout) return MCS_E_ARG;
    MCS_ELECTRONS_ONLY(p, "pop_energies");
    mcs_ctx *ctx = p->ctx;
    hipSetDevice(ctx->device);
    *nout = 0;
    const int64_t n = p->n;
    if (n == 0) return MCS_OK;
    int64_t A;
    int64_t *bpre;
    int32_t rc = count_actives(p, &A, &bpre);
    if (rc) return rc;
    *nout = A;
    if (A > nmax) return mcs_fail(ctx, MCS_E_ARG, "pop_energies: %lld actives, room for %lld", (long long)A,
                                  (long long)nmax);
    if (A == 0) return MCS_OK;
    if (!energy) return mcs_fail(ctx, MCS_E_ARG, "pop_energies: null output");
    void *scr;
    rc = mcs_scratch(ctx, sizeof(double) * (size_t)A, &scr);
    if (rc) return rc;
    const dim3 g((unsigned)cdiv(n, MCS_BLOCK)), bl(MCS_BLOCK);
    if (p->prec == MCS_PREC_F32)
        hipLaunchKernelGGL(k_active_energy<SplitBuf>, g, bl, 0, ctx->stream, p->sbuf[0], n, (const int64_t *)bpre,
                           (double *)scr);
    else
        hipLaunchKernelGGL(k_active_energy<PopBuf>, g, bl, 0, ctx->stream, p->buf[0], n, (const int64_t *)bpre,
                           (double *)scr);
    MCS_LAUNCH_CHECK(ctx);
    MCS_HIP(ctx, hipMemcpyAsync(energy, scr, sizeof(double) * (size_t)A, hipMemcpyDeviceToHost, ctx->stream));
    MCS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return MCS_OK;
}

extern "C" int32_t mcs_repack(mcs_pop *p)
{
    if (!p) return MCS_E_ARG;
    mcs_ctx *ctx = p->ctx;
    if (p->prec == MCS_PREC_F32)
        return mcs_fail(ctx, MCS_E_ARG, "repack: not available on a split-precision population");
    mcs_pop_touch(p);
    hipSetDevice(ctx->device);
    int64_t n = p->n;
    if (n == 0) return MCS_OK;
    int64_t A;
    int64_t *bpre;
    int32_t rc = count_actives(p, &A, &bpre);
    if (rc) return rc;
    // holes below A = actives at or above A
    int64_t h = 0;
    if (A < n) {
        int64_t nblk = cdiv(n, MCS_BLOCK);
        hipLaunchKernelGGL(k_repack_ranks, dim3((unsigned)nblk), dim3(MCS_BLOCK), 0, ctx->stream,
                           p->buf[0].active, n, (const int64_t *)bpre, A, p->i32a, p->i32b);
        MCS_LAUNCH_CHECK(ctx);
        // h = A - (actives in [0, A)); read the prefix at A's block and recount inside it
        int64_t blockA = A / MCS_BLOCK;
        int64_t pre;
        rc = mcs_readback(ctx, bpre + blockA, &pre, 1);
        if (rc) return rc;
        std::vector<uint8_t> tail((size_t)(A - blockA * MCS_BLOCK));
        if (!tail.empty()) {
            MCS_HIP(ctx, hipMemcpy(tail.data(), p->buf[0].active + blockA * MCS_BLOCK, tail.size(),
                                   hipMemcpyDeviceToHost));
            for (uint8_t c : tail) pre += c ? 1 : 0;
        }
        h = A - pre;
    }
    if (h > 0) {
        hipLaunchKernelGGL(k_repack_move, dim3((unsigned)cdiv(h, MCS_BLOCK)), dim3(MCS_BLOCK), 0,
                           ctx->stream, p->buf[0], (const int32_t *)p->i32a,
                           (const int32_t *)p->i32b, h);
        MCS_LAUNCH_CHECK(ctx);
    }
    p->n = A;
    return MCS_OK;
}

// ---------------------------------------------------------------- Swarm.resample!

__global__ __launch_bounds__(MCS_BLOCK) void k_swarm_drop(PopBuf b, int64_t n, double pdrop,
                                                          uint64_t seed, uint32_t instance,
                                                          uint32_t step)
{
    int64_t i = gidx();
    if (i >= n || !b.active[i]) return;
    mcs_stream rs;
    mcs_stream_init(&rs, seed, instance, MCS_PHASE_SWARM_RESAMPLE, b.id[i], step);
    if (mcs_u01(&rs) < pdrop) b.active[i] = 0;
}

__global__ __launch_bounds__(MCS_BLOCK) void k_swarm_count(PopBuf b, int64_t n, double lam,
                                                           uint64_t seed, uint32_t instance,
                                                           uint32_t step, int32_t *cnt,
                                                           int64_t *btot)
{
    int64_t i = gidx();
    int c = 0;
    if (i < n && b.active[i]) {
        mcs_stream rs;
        mcs_stream_init(&rs, seed, instance, MCS_PHASE_SWARM_RESAMPLE, b.id[i], step);
        c = (int)mcs_poisson(&rs, lam);
    }
    if (i < n) cnt[i] = c;
    int tot;
    block_excl_scan<int>(c, &tot);
    if (threadIdx.x == 0) btot[blockIdx.x] = tot;
}

__global__ __launch_bounds__(MCS_BLOCK) void k_swarm_place(PopBuf b, int64_t n, const int32_t *cnt,
                                                           const int64_t *bpre, uint64_t next_id)
{
    int64_t i = gidx();
    int c = i < n ? cnt[i] : 0;
    int tot;
    int ex = block_excl_scan<int>(c, &tot);
    if (c == 0) return;
    int64_t base = n + bpre[blockIdx.x] + ex;
    for (int j = 0; j < c; ++j) {
        copy_particle(b, i, b, base + j);
        b.id[base + j] = next_id + (uint64_t)(base + j - n);
    }
}

extern "C" int32_t mcs_swarm_resample(mcs_pop *p, int64_t n0, double tol, uint64_t seed,
                                      uint32_t instance, uint32_t step, int32_t *did)
{
    if (!p || n0 <= 0) return MCS_E_ARG;
    MCS_ELECTRONS_ONLY(p, "swarm_resample");
    mcs_ctx *ctx = p->ctx;
    if (p->prec == MCS_PREC_F32)
        return mcs_fail(ctx, MCS_E_ARG, "swarm_resample: not available on a split-precision population");
    mcs_pop_touch(p);
    hipSetDevice(ctx->device);
    if (did) *did = 0;
    int64_t n = p->n;
    if (n == 0) return MCS_OK;
    int64_t ncur;
    int64_t *bpre;
    int32_t rc = count_actives(p, &ncur, &bpre);
    if (rc) return rc;
    if (fabs((double)(ncur - n0) / (double)n0) < tol) return MCS_OK;
    if (ncur == 0) return MCS_OK;
    if (did) *did = 1;
    int64_t nblk = cdiv(n, MCS_BLOCK);
    if (ncur > n0) {
        double pdrop = (double)(ncur - n0) / (double)ncur;
        hipLaunchKernelGGL(k_swarm_drop, dim3((unsigned)nblk), dim3(MCS_BLOCK), 0, ctx->stream,
                           p->buf[0], n, pdrop, seed, instance, step);
        MCS_LAUNCH_CHECK(ctx);
        return MCS_OK;
    }
    double lam = (double)(n0 - ncur) / (double)ncur;
    int64_t *btot = p->blk;
    bpre = p->blk + (nblk + 1);
    hipLaunchKernelGGL(k_swarm_count, dim3((unsigned)nblk), dim3(MCS_BLOCK), 0, ctx->stream,
                       p->buf[0], n, lam, seed, instance, step, p->i32a, btot);
    MCS_LAUNCH_CHECK(ctx);
    rc = mcs_scan_i64(ctx, btot, bpre, nblk);
    if (rc) return rc;
    int64_t added;
    rc = mcs_readback(ctx, bpre + nblk, &added, 1);
    if (rc) return rc;
    if (n + added > p->cap)
        return mcs_fail(ctx, MCS_E_CAPACITY, "Maximum number of particles reached (%lld + %lld > %lld)",
                        (long long)n, (long long)added, (long long)p->cap);
    if (added > 0) {
        hipLaunchKernelGGL(k_swarm_place, dim3((unsigned)nblk), dim3(MCS_BLOCK), 0, ctx->stream,
                           p->buf[0], n, (const int32_t *)p->i32a, (const int64_t *)bpre,
                           p->next_id);
        MCS_LAUNCH_CHECK(ctx);
        p->n = n + added;
        p->next_id += (uint64_t)added;
    }
    return MCS_OK;
}
