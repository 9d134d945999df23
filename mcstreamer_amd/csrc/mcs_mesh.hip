// mcs_mesh.hip -- the mesh side of the step: canonical cell sort, charge
// deposition, charge assembly, face fields and per-cell resampling.
#include <cstring>

#include <vector>

#include <rocprim/device/device_radix_sort.hpp>

#include "mcs_device.h"
#include "mcs_fp32.h"

#define MESH_BOX_COPIES 64   // copies of the occupied-cell box (blocks spread their atomics)
#define MESH_BOX_FINAL (6 * MESH_BOX_COPIES)   // the reduced box after the copies
#define MESH_BOX_WORDS (MESH_BOX_FINAL + 6)   // the box copies and the reduced box after the cell counts
#define M_SHORT 0
#define M_BIG 1
#define M_TOT 2    // records in the mesh's segments (cell_start[nseg])
#define M_HUGE 8   // count, then (start, length) pairs

// ---------------------------------------------------------------- mesh object

static int64_t arr_len(const int32_t n[3], int which)
{
    int64_t nx = n[0], ny = n[1], nz = n[2];
    switch (which) {
    case MCS_ARR_EX: return (nx + 1) * (ny + 2) * (nz + 2);
    case MCS_ARR_EY: return (nx + 2) * (ny + 1) * (nz + 2);
    case MCS_ARR_EZ: return (nx + 2) * (ny + 2) * (nz + 1);
    default: return (nx + 2) * (ny + 2) * (nz + 2);
    }
}

extern "C" int32_t mcs_mesh_create(mcs_ctx *ctx, const int32_t n[3], double h, mcs_mesh **out)
{
    if (!ctx || !n || !out) return MCS_E_ARG;
    if (n[0] < 1 || n[1] < 1 || n[2] < 1 || !(h > 0.0))
        return mcs_fail(ctx, MCS_E_ARG, "mesh: sizes must be >= 1 and h > 0");
    int64_t ncells = (int64_t)n[0] * n[1] * n[2];
    // every array (ghosts included) is indexed with 32-bit integers on the device
    if ((int64_t)(n[0] + 2) * (n[1] + 2) * (n[2] + 2) >= ((int64_t)1 << 31) - 2)
        return mcs_fail(ctx, MCS_E_ARG, "mesh too large");
    hipSetDevice(ctx->device);
    mcs_mesh *m = new mcs_mesh();
    m->ctx = ctx;
    for (int d = 0; d < 3; ++d) m->n[d] = n[d];
    m->h = h;
    m->ncells = ncells;
    m->ncg = arr_len(n, MCS_ARR_Q);
    for (int a = 0; a < 7; ++a) {
        m->arr_len[a] = arr_len(n, a);
        if (hipMalloc(&m->arr[a], sizeof(double) * m->arr_len[a]) != hipSuccess)
            return mcs_fail(ctx, MCS_E_HIP, "mesh: device allocation failed");
        MCS_HIP(ctx, mcs_zero(ctx, m->arr[a], sizeof(double) * m->arr_len[a]));
    }
    if (hipMalloc(&m->cell_start, sizeof(int32_t) * (ncells + 2)) != hipSuccess ||
        hipMalloc(&m->cell_cnt, sizeof(int32_t) * (ncells + 2 + MESH_BOX_WORDS)) != hipSuccess ||
        hipMalloc(&m->cnt64, sizeof(int64_t) * (ncells + 2)) != hipSuccess ||
        hipMalloc(&m->pre64, sizeof(int64_t) * (ncells + 2)) != hipSuccess)
        return mcs_fail(ctx, MCS_E_HIP, "mesh: device allocation failed");
    MCS_HIP(ctx, mcs_zero(ctx, m->cell_start, sizeof(int32_t) * (ncells + 2)));
    *out = m;
    return MCS_OK;
}

extern "C" int32_t mcs_mesh_destroy(mcs_mesh *m)
{
    if (!m) return MCS_OK;
    hipSetDevice(m->ctx->device);
    hipStreamSynchronize(m->ctx->stream);
    for (int a = 0; a < 7; ++a) hipFree(m->arr[a]);
    for (int d = 0; d < 3; ++d)
        if (m->face32[d]) hipFree(m->face32[d]);
    hipFree(m->cell_start);
    hipFree(m->cell_cnt);
    hipFree(m->cnt64);
    hipFree(m->pre64);
    if (m->dep_list) hipFree(m->dep_list);
    if (m->dep_host) hipHostFree(m->dep_host);
    delete m;
    return MCS_OK;
}

extern "C" int32_t mcs_mesh_set(mcs_mesh *m, int32_t which, const double *host)
{
    if (!m || !host || which < 0 || which > 6) return MCS_E_ARG;
    mcs_ctx *ctx = m->ctx;
    hipSetDevice(ctx->device);
    MCS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    MCS_HIP(ctx, mcs_h2d(ctx, m->arr[which], host, sizeof(double) * m->arr_len[which]));
    return MCS_OK;
}

extern "C" int32_t mcs_mesh_get(mcs_mesh *m, int32_t which, double *host)
{
    if (!m || !host || which < 0 || which > 6) return MCS_E_ARG;
    mcs_ctx *ctx = m->ctx;
    hipSetDevice(ctx->device);
    MCS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    MCS_HIP(ctx, hipMemcpy(host, m->arr[which], sizeof(double) * m->arr_len[which],
                           hipMemcpyDeviceToHost));
    return MCS_OK;
}

__global__ void k_zero_f64(double *a, int64_t n)
{
    int64_t i = gidx();
    if (i < n) a[i] = 0.0;
}

__global__ void k_zero_i32(int32_t *a, int64_t n)
{
    int64_t i = gidx();
    if (i < n) a[i] = 0;
}

extern "C" int32_t mcs_mesh_zero(mcs_mesh *m, int32_t which)
{
    if (!m || which < 0 || which > 6) return MCS_E_ARG;
    mcs_ctx *ctx = m->ctx;
    hipSetDevice(ctx->device);
    hipLaunchKernelGGL(k_zero_f64, dim3((unsigned)cdiv(m->arr_len[which], MCS_BLOCK)), dim3(MCS_BLOCK),
                       0, ctx->stream, m->arr[which], m->arr_len[which]);
    MCS_LAUNCH_CHECK(ctx);
    return MCS_OK;
}

// ---------------------------------------------------------------- sort by cell

struct MeshGeom {
    int32_t n[3];
    double h;
    int64_t ncells;
};

static inline int32_t *mesh_box(const mcs_mesh *m)
{
    return m->cell_cnt + m->ncells + 2;
}

static inline MeshGeom geom(const mcs_mesh *m)
{
    MeshGeom g;
    for (int d = 0; d < 3; ++d) g.n[d] = m->n[d];
    g.h = m->h;
    g.ncells = m->ncells;
    return g;
}

// Counting pass.  Lanes of a wave that share a cell take consecutive ranks from a
// single atomic (ballot over equal keys), so a mostly sorted population costs one
// atomic per cell run rather than one per particle.
// The box of the cells that hold particles, kept after the cell counts (at
// cell_cnt + ncells + 2, zeroed with them) in MESH_BOX_COPIES copies of six
// words: d = max over particles of n[d] - i_d and 3 + d = max of i_d + 1, so
// that zero means no particle.  The segment listing reduces the copies and
// stores the six words at MESH_BOX_FINAL for the later walks.

struct CellBox {
    int lo[3], ext[3];   // first cell and extent per axis; ext[0] == 0: empty
};

// The six box words, the maxima over the copies (every thread of the block calls).
__device__ __forceinline__ void box_words(const int32_t *box, int32_t bw[6])
{
    __shared__ int32_t sw[6];
    if (threadIdx.x < 6) {
        int v = 0;
        for (int c = 0; c < MESH_BOX_COPIES; ++c) v = max(v, box[6 * c + threadIdx.x]);
        sw[threadIdx.x] = v;
    }
    __syncthreads();
#pragma unroll
    for (int w = 0; w < 6; ++w) bw[w] = sw[w];
}

// The box grown by g cells per side, clamped to the mesh.
__device__ __forceinline__ CellBox cell_box(const int32_t *bw, const int32_t n[3], int g)
{
    CellBox B;
    if (bw[3] == 0) {
        B.lo[0] = B.lo[1] = B.lo[2] = 0;
        B.ext[0] = B.ext[1] = B.ext[2] = 0;
        return B;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int lo = max(n[d] - bw[d] - g, 0), hi = min(bw[3 + d] - 1 + g, n[d] - 1);
        B.lo[d] = lo;
        B.ext[d] = hi - lo + 1;
    }
    return B;
}

// Cell index (linear, x fastest) of position t of the box, in increasing order.
__device__ __forceinline__ int64_t box_cell(const CellBox &B, const int32_t n[3], int64_t t, int64_t &i,
                                            int64_t &j, int64_t &k)
{
    const int64_t q = t / B.ext[0];
    i = B.lo[0] + (t - q * B.ext[0]);
    j = B.lo[1] + q % B.ext[1];
    k = B.lo[2] + q / B.ext[1];
    return i + (int64_t)n[0] * (j + (int64_t)n[1] * k);
}

template <typename Buf>
__global__ __launch_bounds__(MCS_BLOCK) void k_sort_count(Buf b, int64_t n, MeshGeom g, int32_t *cnt,
                                                          int32_t *key_out, int32_t *rank_out, int32_t *box)
{
    int64_t i = gidx();
    int lane = threadIdx.x & (MCS_WAVE - 1);
    int key = -1;
    // box words of this particle (0: none)
    int bw[6] = {0, 0, 0, 0, 0, 0};
    if (i < n && b.active[i]) {
        int64_t c = cell_key(b, i, g.n, g.h);
        key = c < 0 ? (int)g.ncells : (int)c;
        if (c >= 0) {
            const uint32_t cu = (uint32_t)c, nx = (uint32_t)g.n[0], ny = (uint32_t)g.n[1];
            const uint32_t q = cu / nx, ci = cu - q * nx, cj = q % ny, ck = q / ny;
            bw[0] = g.n[0] - (int)ci;
            bw[1] = g.n[1] - (int)cj;
            bw[2] = g.n[2] - (int)ck;
            bw[3] = (int)ci + 1;
            bw[4] = (int)cj + 1;
            bw[5] = (int)ck + 1;
        }
    }
    {
        // block maxima, then one atomic per word into the block's copy of the box
        __shared__ int sbw[MCS_BLOCK / MCS_WAVE][6];
#pragma unroll
        for (int w = 0; w < 6; ++w) {
            const int v = __reduce_max_sync(~0ULL, bw[w]);
            if (lane == 0) sbw[threadIdx.x / MCS_WAVE][w] = v;
        }
        __syncthreads();
        if (threadIdx.x < 6) {
            int v = 0;
            for (int w = 0; w < MCS_BLOCK / MCS_WAVE; ++w) v = max(v, sbw[w][threadIdx.x]);
            if (v > 0) atomicMax(box + 6 * (blockIdx.x % MESH_BOX_COPIES) + threadIdx.x, v);
        }
    }
    unsigned long long todo = __ballot(key >= 0);
    int rank = 0;
    while (todo) {
        int leader = __ffsll((unsigned long long)todo) - 1;
        int lk = __shfl(key, leader, MCS_WAVE);
        unsigned long long peers = __ballot(key == lk) & todo;
        int base = 0;
        if (lane == leader) base = atomicAdd(&cnt[lk], __popcll(peers));
        base = __shfl(base, leader, MCS_WAVE);
        if ((peers >> lane) & 1ULL) rank = base + __popcll(peers & ((1ULL << lane) - 1ULL));
        todo &= ~peers;
    }
    if (i < n) {
        key_out[i] = key;
        rank_out[i] = rank;
    }
}

template <typename Buf>
__global__ __launch_bounds__(MCS_BLOCK) void k_sort_scatter(Buf b, int64_t n,
                                                            const int32_t *key,
                                                            const int32_t *rank,
                                                            const int32_t *cs, int32_t *perm,
                                                            uint64_t *sid)
{
    int64_t i = gidx();
    if (i >= n) return;
    int k = key[i];
    if (k < 0) return;
    int64_t pos = (int64_t)cs[k] + rank[i];
    perm[pos] = (int32_t)i;
    sid[pos] = b.id[i];
}

#define SEG_SMALL 64

// Within each cell, order by particle id.  Thread per cell: segments of 2..64
// particles go to a list sorted one wave per segment (k_segsort_wave), longer
// ones to a list sorted one block per segment (k_segsort_big).  The lists share
// one buffer: short ones fill it from the front, long ones from the back.
#define SEG_HUGE 4096
#define HUGE_MAX 200   // (start, length) pairs recorded after the counters in ctx->d_small

__global__ __launch_bounds__(MCS_BLOCK) void k_segsort_list(const int32_t *cs, int64_t nseg, MeshGeom g,
                                                            int32_t *box, int32_t *lists, int64_t *nshort,
                                                            int64_t *nbig, int64_t *huge)
{
    if (gidx() == 0) huge[M_TOT - M_HUGE] = cs[nseg];
    // the segments that can hold two records or more: the occupied box's cells,
    // then the bucket of particles outside the mesh (segment nseg - 1)
    int32_t bw[6];
    box_words(box, bw);
    // the reduced box for the later walks (every block writes the same words)
    if (threadIdx.x < 6) box[MESH_BOX_FINAL + threadIdx.x] = bw[threadIdx.x];
    const CellBox B = cell_box(bw, g.n, 0);
    const int64_t nbox = (int64_t)B.ext[0] * B.ext[1] * B.ext[2];
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base <= nbox; base += (int64_t)gridDim.x * blockDim.x) {
        const int64_t t = base + threadIdx.x;
        int64_t c = -1;
        if (t < nbox) {
            int64_t i, j, k;
            c = box_cell(B, g.n, t, i, j, k);
        } else if (t == nbox) {
            c = nseg - 1;
        }
        const int64_t len = c >= 0 ? (int64_t)cs[c + 1] - cs[c] : 0;
        if (len > SEG_HUGE) {
            unsigned long long k = atomicAdd((unsigned long long *)huge, 1ULL);
            if (k < HUGE_MAX) {
                huge[1 + 2 * k] = cs[c];
                huge[2 + 2 * k] = len;
            }
        }
        // one atomic per block and list (the counters are single addresses shared
        // by the whole grid): threads take consecutive entries in block order
        const int fs = (len > 1 && len <= SEG_SMALL) ? 1 : 0, fb = len > SEG_SMALL ? 1 : 0;
        int ts, tb;
        const int es = block_excl_scan<int>(fs, &ts);
        const int eb = block_excl_scan<int>(fb, &tb);
        __shared__ unsigned long long base_s, base_b;
        if (threadIdx.x == 0) {
            base_s = ts ? atomicAdd((unsigned long long *)nshort, (unsigned long long)ts) : 0ULL;
            base_b = tb ? atomicAdd((unsigned long long *)nbig, (unsigned long long)tb) : 0ULL;
        }
        __syncthreads();
        if (fs) lists[base_s + es] = (int32_t)c;
        if (fb) lists[nseg - 1 - (int64_t)(base_b + eb)] = (int32_t)c;
        __syncthreads();
    }
}

// One wave per short segment: lane l holds element l; its rank is the number of
// smaller ids in the segment (ids are unique).
__global__ __launch_bounds__(MCS_BLOCK) void k_segsort_wave(const int32_t *cs, int32_t *perm,
                                                            uint64_t *sid, const int32_t *lists,
                                                            const int64_t *nshort)
{
    const int lane = threadIdx.x & (MCS_WAVE - 1);
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / MCS_WAVE;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / MCS_WAVE;
    const int64_t ns = *nshort;
    for (int64_t q = wave; q < ns; q += nwaves) {
        int32_t c = lists[q];
        int64_t a = cs[c];
        int len = (int)((int64_t)cs[c + 1] - a);
        uint64_t key = lane < len ? sid[a + lane] : ~0ULL;
        int32_t pv = lane < len ? perm[a + lane] : 0;
        int rank = 0;
        for (int m = 0; m < len; ++m) {
            uint64_t other = __shfl(key, m, MCS_WAVE);
            rank += other < key ? 1 : 0;
        }
        if (lane < len) {
            sid[a + rank] = key;
            perm[a + rank] = pv;
        }
    }
}


// One block per segment of 65..4096 particles (longer ones take the grid-wide path).
__global__ __launch_bounds__(MCS_BLOCK) void k_segsort_big(const int32_t *cs, int32_t *perm,
                                                           uint64_t *sid, const int32_t *lists,
                                                           int64_t nseg, const int64_t *nbig,
                                                           int32_t *tmp_perm, uint64_t *tmp_sid)
{
    int64_t nb = *nbig;
    for (int64_t q = blockIdx.x; q < nb; q += gridDim.x) {
        int32_t c = lists[nseg - 1 - q];
        int64_t a = cs[c], e = cs[c + 1];
        if (e - a > SEG_HUGE) continue;
        for (int64_t t = a + threadIdx.x; t < e; t += blockDim.x) {
            uint64_t key = sid[t];
            int64_t rank = 0;
            for (int64_t u = a; u < e; ++u) rank += sid[u] < key ? 1 : 0;
            tmp_perm[a + rank] = perm[t];
            tmp_sid[a + rank] = key;
        }
        __syncthreads();
        for (int64_t t = a + threadIdx.x; t < e; t += blockDim.x) {
            perm[t] = tmp_perm[t];
            sid[t] = tmp_sid[t];
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(MCS_BLOCK) void k_copy_back_seg(int32_t *perm, uint64_t *sid, int64_t a, int64_t len,
                                                             const int32_t *tmp_perm, const uint64_t *tmp_sid)
{
    const int64_t t = gidx();
    if (t >= len) return;
    perm[a + t] = tmp_perm[a + t];
    sid[a + t] = tmp_sid[a + t];
}

template <typename Buf>
__global__ __launch_bounds__(MCS_BLOCK) void k_sort_gather(Buf src, Buf dst, const int32_t *perm, int64_t na)
{
    int64_t pos = gidx();
    if (pos >= na) return;
    copy_rec(src, perm[pos], dst, pos);
}

template <typename Buf>
static int32_t sort_impl(mcs_pop *p, mcs_mesh *m, const Buf &src, const Buf &dst)
{
    mcs_ctx *ctx = p->ctx;
    int32_t rc;
    int64_t n = p->n;
    int64_t nseg = m->ncells + 1;
    // the counts and, after them, the box of the occupied cells
    hipLaunchKernelGGL(k_zero_i32, dim3((unsigned)cdiv(nseg + 1 + MESH_BOX_WORDS, MCS_BLOCK)), dim3(MCS_BLOCK), 0,
                       ctx->stream, m->cell_cnt, nseg + 1 + MESH_BOX_WORDS);
    int64_t *ctr = ctx->d_small;
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, ctx->stream, (int32_t *)ctr, (int64_t)(2 * (M_HUGE + 1)));
    if (n > 0) {
        ProfScope ps(ctx, "sort_count");
        hipLaunchKernelGGL(k_sort_count<Buf>, dim3((unsigned)cdiv(n, MCS_BLOCK)), dim3(MCS_BLOCK), 0,
                           ctx->stream, src, n, geom(m), m->cell_cnt, p->i32a, p->i32b, mesh_box(m));
    }
    MCS_LAUNCH_CHECK(ctx);
    rc = mcs_scan_i32(ctx, m->cell_cnt, m->cell_start, nseg);
    if (rc) return rc;
    int64_t tot = 0;   // no particles: every segment is empty
    if (n > 0) {
        ProfScope ps(ctx, "sort_scatter");
        hipLaunchKernelGGL(k_sort_scatter<Buf>, dim3((unsigned)cdiv(n, MCS_BLOCK)), dim3(MCS_BLOCK), 0,
                           ctx->stream, src, n, (const int32_t *)p->i32a,
                           (const int32_t *)p->i32b, (const int32_t *)m->cell_start, p->i32c,
                           p->u64a);
        // the count buffer is free again: it holds the lists of segments to sort
        hipLaunchKernelGGL(k_segsort_list, dim3((unsigned)std::min<int64_t>(cdiv(nseg, MCS_BLOCK), 2048)),
                           dim3(MCS_BLOCK), 0, ctx->stream, (const int32_t *)m->cell_start, nseg, geom(m),
                           mesh_box(m), m->cell_cnt, ctr + M_SHORT, ctr + M_BIG, ctr + M_HUGE);
        hipLaunchKernelGGL(k_segsort_wave, dim3(2048), dim3(MCS_BLOCK), 0, ctx->stream,
                           (const int32_t *)m->cell_start, p->i32c, p->u64a,
                           (const int32_t *)m->cell_cnt, (const int64_t *)(ctr + M_SHORT));
        hipLaunchKernelGGL(k_segsort_big, dim3(512), dim3(MCS_BLOCK), 0, ctx->stream,
                           (const int32_t *)m->cell_start, p->i32c, p->u64a,
                           (const int32_t *)m->cell_cnt, nseg, (const int64_t *)(ctr + M_BIG), p->i32b,
                           (uint64_t *)dst.id);
        // segments longer than SEG_HUGE (in practice the bucket of particles outside
        // the mesh): their (start, length) pairs come back in one read; each is
        // ordered by id with a radix sort of (id, position) pairs (ids are unique)
        // (with the record count in front)
        int64_t hb[2 + 2 * HUGE_MAX];
        rc = mcs_readback2(ctx, ctr + M_TOT, 1, ctr + M_HUGE, 1 + 2 * HUGE_MAX, hb);
        if (rc) return rc;
        tot = hb[0];
        const int64_t *hv = hb + 1;
        if (hv[0] > HUGE_MAX)
            return mcs_fail(ctx, MCS_E_CAPACITY, "sort: more than %d cells hold over %d particles", HUGE_MAX,
                            SEG_HUGE);
        for (int64_t q = 0; q < hv[0]; ++q) {
            const int64_t a = hv[1 + 2 * q], len = hv[2 + 2 * q];
            const uint64_t *kin = (const uint64_t *)p->u64a + a;
            uint64_t *kout = (uint64_t *)dst.id + a;
            const int32_t *vin = (const int32_t *)p->i32c + a;
            int32_t *vout = p->i32b + a;
            size_t tb = 0;
            MCS_HIP(ctx, rocprim::radix_sort_pairs(nullptr, tb, kin, kout, vin, vout, (size_t)len, 0u, 64u,
                                                   ctx->stream));
            void *tmp = nullptr;
            rc = mcs_scratch(ctx, tb, &tmp);
            if (rc) return rc;
            MCS_HIP(ctx, rocprim::radix_sort_pairs(tmp, tb, kin, kout, vin, vout, (size_t)len, 0u, 64u,
                                                   ctx->stream));
            hipLaunchKernelGGL(k_copy_back_seg, dim3((unsigned)cdiv(len, MCS_BLOCK)), dim3(MCS_BLOCK), 0,
                               ctx->stream, p->i32c, p->u64a, a, len, (const int32_t *)p->i32b,
                               (const uint64_t *)dst.id);
        }
    }
    MCS_LAUNCH_CHECK(ctx);
    const int64_t na = tot;
    if (na > 0) {
        ProfScope ps(ctx, "sort_gather");
        hipLaunchKernelGGL(k_sort_gather<Buf>, dim3((unsigned)cdiv(na, MCS_BLOCK)), dim3(MCS_BLOCK), 0,
                           ctx->stream, src, dst, (const int32_t *)p->i32c, na);
    }
    MCS_LAUNCH_CHECK(ctx);
    mcs_pop_swap(p);
    p->n = na;
    mcs_pop_touch(p);
    m->sorted_gen = p->gen;
    return MCS_OK;
}

extern "C" int32_t mcs_sort_by_cell(mcs_pop *p, mcs_mesh *m)
{
    if (!p || !m) return MCS_E_ARG;
    MCS_ELECTRONS_ONLY(p, "sort_by_cell");
    mcs_ctx *ctx = p->ctx;
    hipSetDevice(ctx->device);
    int32_t rc = mcs_pop_ensure_alt(p);
    if (rc) return rc;
    if (p->prec == MCS_PREC_F32) {
        for (int d = 0; d < 3; ++d)
            if (p->gn[d] != m->n[d]) return mcs_fail(ctx, MCS_E_ARG, "sort: mesh differs from the population's grid");
        return sort_impl(p, m, p->sbuf[0], p->sbuf[1]);
    }
    return sort_impl(p, m, p->buf[0], p->buf[1]);
}

// ---------------------------------------------------------------- deposit

template <typename Buf>
struct DepArgs {
    Buf b;
    const int32_t *cs;
    MeshGeom g;
    double *arr;
    double val;
    // per-particle CIC record, computed once: the lower cell index per axis packed
    // in 21-bit fields (offset by one), and the shares of the 8 cells it feeds
    uint64_t *pk;
    double *pc;       // 8 per particle, corner (dx, dy, dz) at 4 dx + 2 dy + dz
    int64_t n;
    int32_t *dense;   // [0] count, then cells with more than dense_min records to visit
    int32_t dense_min;
    // cells with more than cand_min records (the listing threshold), counted for
    // the next call's choice of gather: dense[cand_at] counts them when they are
    // not listed, and k_deposit_dense writes the count to *count_out (host memory)
    int32_t cand_min;
    int64_t cand_at;
    int32_t *count_out;
    const int32_t *box;   // the occupied-cell box of the last sort
};

#define DEP_MASK 0x1FFFFFULL
#define DEP_PF 8       // records loaded ahead of the sequential accumulation
#define DEP_DENSE 64   // default: cells visiting more particle records take a wavefront each
#define DEP_LANE_LISTED (1 << 20)   // default: above this many such cells, one lane per cell

// Per particle: g = x/h - 1/2, i0 = floor(g), fr = g - i0, base = (w val)/dV and
// the share base * ((wx wy) wz) of each of the 8 cells i0 + (dx, dy, dz), with
// w = 1 - fr towards i0 and fr towards i0 + 1: exactly the products a
// sequential CIC scatter adds.
template <typename Buf>
__global__ __launch_bounds__(MCS_BLOCK) void k_deposit_prep(DepArgs<Buf> a)
{
    int64_t q = gidx();
    if (q == 0) {
        a.dense[0] = 0;
        a.dense[a.cand_at] = 0;
    }
    if (q >= a.n) return;
    const double h = a.g.h;
    const double dV = (h * h) * h;
    double fl[3], fr[3];
    cic_geom(a.b, q, h, fl, fr);
    uint64_t key = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        // particles outside [-1, n] along an axis are never gathered: clamp the field
        double fc = fmin(fmax(fl[d] + 1.0, 0.0), (double)DEP_MASK);
        key |= ((uint64_t)fc & DEP_MASK) << (21 * d);
    }
    a.pk[q] = key;
    const double base = (get_w(a.b, q) * a.val) / dV;
    double sh[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const double wx = (c & 4) ? fr[0] : 1.0 - fr[0];
        const double wy = (c & 2) ? fr[1] : 1.0 - fr[1];
        const double wz = (c & 1) ? fr[2] : 1.0 - fr[2];
        sh[c] = base * ((wx * wy) * wz);
    }
    double4 *dst = (double4 *)(a.pc + 8 * q);
    dst[0] = make_double4(sh[0], sh[1], sh[2], sh[3]);
    dst[1] = make_double4(sh[4], sh[5], sh[6], sh[7]);
}

// Loads that stay inside the arrays whatever the predicate (a record past the
// range reads record 0 and is discarded).
template <typename Buf>
__device__ __forceinline__ uint64_t dep_key(const DepArgs<Buf> &a, int64_t q, int64_t s1)
{
    const uint64_t v = a.pk[q < s1 ? q : 0];
    return q < s1 ? v : ~0ULL;
}

template <typename Buf>
__device__ __forceinline__ double dep_share(const DepArgs<Buf> &a, bool on, int64_t q, int corner)
{
    return a.pc[on ? 8 * q + corner : 0];
}

// Whether a record (key) feeds cell (i, j, k), and at which corner.
__device__ __forceinline__ bool dep_corner(uint64_t key, int i, int j, int k, int &corner)
{
    const int dx = i + 1 - (int)(key & DEP_MASK);
    const int dy = j + 1 - (int)((key >> 21) & DEP_MASK);
    const int dz = k + 1 - (int)((key >> 42) & DEP_MASK);
    corner = 4 * dx + 2 * dy + dz;
    return (uint32_t)dx <= 1u && (uint32_t)dy <= 1u && (uint32_t)dz <= 1u;
}

// The records a cell visits, as one sequence: the ranges of its nine neighbour
// rows (three consecutive cells along x each) in (dk, dj) order, which is the
// order of the sequential scatter over the sorted population.  Rows outside the
// mesh are empty.
struct DepRows {
    int32_t start[9];
    int32_t pre[10];   // prefix of the range lengths (record counts are int32, as cell starts)
};

__device__ __forceinline__ void dep_rows(const int32_t *cs, int64_t i, int64_t j, int64_t k, int64_t nx, int64_t ny,
                                         int64_t nz, DepRows &R)
{
    const int64_t ilo = i > 0 ? i - 1 : 0, ihi = i + 1 < nx ? i + 1 : nx - 1;
    R.pre[0] = 0;
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        const int64_t kk = k + r / 3 - 1, jj = j + r % 3 - 1;
        int32_t s0 = 0, len = 0;
        if (kk >= 0 && kk < nz && jj >= 0 && jj < ny) {
            const int64_t row = nx * (jj + ny * kk);
            s0 = cs[row + ilo];
            len = cs[row + ihi + 1] - s0;
        }
        R.start[r] = s0;
        R.pre[r + 1] = R.pre[r] + len;
    }
}

// Record index of position v (< pre[9]) of the sequence.  (Each candidate is
// computed before the select: a select among the row starts themselves becomes
// an indexed load from private memory.)
__device__ __forceinline__ int32_t dep_at(const DepRows &R, int32_t v)
{
    int32_t q = R.start[0] + v;
#pragma unroll
    for (int r = 1; r < 9; ++r)
        if (v >= R.pre[r]) q = R.start[r] + (v - R.pre[r]);
    return q;
}

// density! as a per-cell gather over the canonical order: cell c sums the shares of
// the particles of its 27 neighbours in (cell, id) order, which is the order in
// which a sequential scatter over the sorted population adds them.  No atomics.
// Keys run DEP_PF records ahead along the cell's sequence; a share is loaded only
// for a record that feeds the cell.  Cells with many records go to
// k_deposit_dense.
template <int CIC, typename Buf>
__device__ __forceinline__ void deposit_cell(const DepArgs<Buf> &a, const CellBox &B, int64_t t)
{
    int64_t i, j, k;
    const int64_t c = box_cell(B, a.g.n, t, i, j, k);
    const int64_t nx = a.g.n[0], ny = a.g.n[1], nz = a.g.n[2];
    int64_t idx = (i + 1) + (nx + 2) * ((j + 1) + (ny + 2) * (k + 1));
    const double h = a.g.h;
    const double dV = (h * h) * h;
    if (!CIC) {
        double acc = a.arr[idx];
        int64_t s0 = a.cs[c], s1 = a.cs[c + 1];
        for (int64_t q = s0; q < s1; ++q) acc += (get_w(a.b, q) * a.val) / dV;
        a.arr[idx] = acc;
        return;
    }
    DepRows R;
    dep_rows(a.cs, i, j, k, nx, ny, nz, R);
    const int32_t tot = R.pre[9];
    if (a.dense_min > a.cand_min) {
        const uint64_t cm = __ballot(tot > a.cand_min);
        if (cm && (threadIdx.x & (MCS_WAVE - 1)) == (uint32_t)(__ffsll((unsigned long long)cm) - 1))
            atomicAdd(a.dense + a.cand_at, __popcll(cm));
    }
    // listed cells: one atomic per wavefront, slots in lane order
    const bool listed = tot > a.dense_min;
    const uint64_t lm = __ballot(listed);
    if (listed) {
        const int below = __builtin_amdgcn_mbcnt_hi((uint32_t)(lm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)lm, 0));
        int base = 0;
        if (below == 0) base = atomicAdd(a.dense, __popcll(lm));
        base = __shfl(base, __ffsll((unsigned long long)lm) - 1);
        a.dense[1 + base + below] = (int32_t)c;
        return;
    }
    if (tot == 0) return;
    double acc = a.arr[idx];
    const int ii = (int)i, jj = (int)j, kk = (int)k;
    int32_t q8[DEP_PF];
    uint64_t kk8[DEP_PF];
#pragma unroll
    for (int u = 0; u < DEP_PF; ++u) {
        q8[u] = u < tot ? dep_at(R, u) : 0;
        kk8[u] = dep_key(a, q8[u], u < tot ? q8[u] + 1 : 0);
    }
    for (int32_t v = 0; v < tot; v += DEP_PF) {
        // shares of these records while the keys of the next ones load
        double t8[DEP_PF];
        bool on8[DEP_PF];
#pragma unroll
        for (int u = 0; u < DEP_PF; ++u) {
            int corner;
            on8[u] = dep_corner(kk8[u], ii, jj, kk, corner);
            t8[u] = dep_share(a, on8[u], q8[u], corner);
        }
#pragma unroll
        for (int u = 0; u < DEP_PF; ++u) {
            const int32_t vn = v + DEP_PF + u;
            q8[u] = vn < tot ? dep_at(R, vn) : 0;
            kk8[u] = dep_key(a, q8[u], vn < tot ? q8[u] + 1 : 0);
        }
#pragma unroll
        for (int u = 0; u < DEP_PF; ++u)
            if (on8[u]) acc += t8[u];
    }
    a.arr[idx] = acc;
}

template <int CIC, typename Buf>
__global__ __launch_bounds__(MCS_BLOCK) void k_deposit(DepArgs<Buf> a)
{
    // the cells that particles can reach: the occupied box (reduced by the sort's
    // segment listing), grown by one for CIC
    const CellBox B = cell_box(a.box + MESH_BOX_FINAL, a.g.n, CIC ? 1 : 0);
    const int64_t nbox = (int64_t)B.ext[0] * B.ext[1] * B.ext[2];
    for (int64_t t = gidx(); t < nbox; t += (int64_t)gridDim.x * blockDim.x)
        deposit_cell<CIC>(a, B, t);
}

#define DEP_SLOTS (MCS_WAVE + 8)   // per wavefront: 64 compacted shares and 8 pad slots
#define DEP_UNROLL 1               // sequence positions per round: DEP_UNROLL x 64

// The CIC gather of the listed cells, one wavefront per cell.  The lanes take
// DEP_UNROLL x 64 consecutive positions of the cell's record sequence at once
// (the keys of the next ones are in flight while their shares load); per 64
// positions, the lanes whose record feeds the cell write their shares to LDS in
// sequence order (compacted by a ballot prefix), and lane 0 adds them to the
// cell's sum one by one, in that order: the additions of k_deposit's sequential
// loop, so the sum is bitwise the same.  Slots past the last share hold -0.0,
// which leaves any sum unchanged under round-to-nearest, so lane 0 adds whole
// groups of eight.
template <typename Buf>
__global__ __launch_bounds__(MCS_BLOCK) void k_deposit_dense(DepArgs<Buf> a)
{
    __shared__ __align__(16) double stage[(MCS_BLOCK / MCS_WAVE) * DEP_SLOTS];
    const int lane = threadIdx.x & (MCS_WAVE - 1);
    const int wpb = blockDim.x / MCS_WAVE;
    const int nw = gridDim.x * wpb;
    double *st = stage + (threadIdx.x / MCS_WAVE) * DEP_SLOTS;
    const int ndense = a.dense[0];
    if (blockIdx.x == 0 && threadIdx.x == 0)
        *(volatile int32_t *)a.count_out = a.dense_min > a.cand_min ? a.dense[a.cand_at] : ndense;
    const int64_t nx = a.g.n[0], ny = a.g.n[1], nz = a.g.n[2];
    const int e0 = __builtin_amdgcn_readfirstlane(blockIdx.x * wpb + threadIdx.x / MCS_WAVE);
    for (int e = e0; e < ndense; e += nw) {
        const int64_t c = a.dense[1 + e];
        const uint32_t cu = (uint32_t)c, qy = cu / (uint32_t)nx;
        const int64_t i = cu - qy * (uint32_t)nx, j = qy % (uint32_t)ny, k = qy / (uint32_t)ny;
        const int64_t idx = (i + 1) + (nx + 2) * ((j + 1) + (ny + 2) * (k + 1));
        DepRows R;
        dep_rows(a.cs, i, j, k, nx, ny, nz, R);
        const int32_t tot = R.pre[9];
        const int ii = (int)i, jj = (int)j, kk = (int)k;
        double acc = a.arr[idx];
        int32_t qc[DEP_UNROLL];
        uint64_t kc[DEP_UNROLL];
#pragma unroll
        for (int u = 0; u < DEP_UNROLL; ++u) {
            const int32_t vu = u * MCS_WAVE + lane;
            qc[u] = vu < tot ? dep_at(R, vu) : 0;
            kc[u] = dep_key(a, qc[u], vu < tot ? qc[u] + 1 : 0);
        }
        for (int32_t v = 0; v < tot; v += DEP_UNROLL * MCS_WAVE) {
            // shares of these positions, keys of the next ones
            double t[DEP_UNROLL];
            bool on[DEP_UNROLL];
#pragma unroll
            for (int u = 0; u < DEP_UNROLL; ++u) {
                int corner;
                on[u] = dep_corner(kc[u], ii, jj, kk, corner);
                t[u] = dep_share(a, on[u], qc[u], corner);
            }
#pragma unroll
            for (int u = 0; u < DEP_UNROLL; ++u) {
                const int32_t vn = v + (DEP_UNROLL + u) * MCS_WAVE + lane;
                qc[u] = vn < tot ? dep_at(R, vn) : 0;
                kc[u] = dep_key(a, qc[u], vn < tot ? qc[u] + 1 : 0);
            }
#pragma unroll
            for (int u = 0; u < DEP_UNROLL; ++u) {
                const uint64_t m = __ballot(on[u]);
                if (m == 0) continue;
                const int cnt = __popcll(m);
                const int below =
                    __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
                // on lanes: their share at their rank (below cnt); lanes cnt..cnt+7: pad
                if (on[u]) st[below] = t[u];
                if (lane >= cnt && lane < cnt + 8) st[lane] = -0.0;
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                if (lane == 0) {
                    for (int g = 0; g < cnt; g += 8) {
                        const double2 p0 = *(const double2 *)(st + g), p1 = *(const double2 *)(st + g + 2);
                        const double2 p2 = *(const double2 *)(st + g + 4), p3 = *(const double2 *)(st + g + 6);
                        acc += p0.x;
                        acc += p0.y;
                        acc += p1.x;
                        acc += p1.y;
                        acc += p2.x;
                        acc += p2.y;
                        acc += p3.x;
                        acc += p3.y;
                    }
                }
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            }
        }
        if (lane == 0) a.arr[idx] = acc;
    }
}

template <typename Buf>
static int32_t deposit_impl(mcs_pop *p, mcs_mesh *m, const Buf &src, int32_t which, int32_t scheme,
                            double val);

extern "C" int32_t mcs_deposit(mcs_pop *p, mcs_mesh *m, int32_t which, int32_t scheme, double val)
{
    if (!p || !m) return MCS_E_ARG;
    MCS_ELECTRONS_ONLY(p, "deposit");
    mcs_ctx *ctx = p->ctx;
    if (which != MCS_ARR_QFIXED && which != MCS_ARR_QPART)
        return mcs_fail(ctx, MCS_E_ARG, "deposit: target must be QFIXED or QPART");
    if (scheme != MCS_DEPOSIT_NGP && scheme != MCS_DEPOSIT_CIC)
        return mcs_fail(ctx, MCS_E_ARG, "deposit: unknown scheme %d", scheme);
    if (!mcs_pop_sorted(p, m->sorted_gen))
        return mcs_fail(ctx, MCS_E_ARG,
                        "deposit: the population is not in this mesh's cell order (call mcs_sort_by_cell)");
    hipSetDevice(ctx->device);
    if (p->prec == MCS_PREC_F32) return deposit_impl(p, m, p->sbuf[0], which, scheme, val);
    return deposit_impl(p, m, p->buf[0], which, scheme, val);
}

template <typename Buf>
static int32_t deposit_impl(mcs_pop *p, mcs_mesh *m, const Buf &src, int32_t which, int32_t scheme,
                            double val)
{
    mcs_ctx *ctx = p->ctx;
    DepArgs<Buf> a;
    a.b = src;
    a.cs = m->cell_start;
    a.g = geom(m);
    a.arr = m->arr[which];
    a.val = val;
    if (scheme == MCS_DEPOSIT_CIC && !p->dep_w) {
        MCS_HIP(ctx, hipMalloc(&p->dep_w, sizeof(double) * 8 * p->cap));
        MCS_HIP(ctx, hipMalloc(&p->dep_k, sizeof(uint64_t) * p->cap));
    }
    a.pk = p->dep_k;
    a.pc = (double *)p->dep_w;
    a.n = p->n;
    if (scheme == MCS_DEPOSIT_CIC && !m->dep_list) {
        MCS_HIP(ctx, hipMalloc(&m->dep_list, sizeof(int32_t) * (m->ncells + 2)));
        MCS_HIP(ctx, hipHostMalloc((void **)&m->dep_host, 64, hipHostMallocMapped | hipHostMallocCoherent));
        *(volatile int32_t *)m->dep_host = 0;
        MCS_HIP(ctx, hipHostGetDevicePointer((void **)&m->dep_host_dev, m->dep_host, 0));
        const char *e = getenv("MCS_DEP_DENSE");
        m->dep_dense = e ? atoi(e) : DEP_DENSE;
        if (m->dep_dense < 1) m->dep_dense = DEP_DENSE;
        const char *f = getenv("MCS_DEP_LANE_LISTED");
        m->dep_lane_listed = f ? atoll(f) : DEP_LANE_LISTED;
    }
    a.dense = m->dep_list;
    a.box = mesh_box(m);
    a.cand_min = m->dep_dense;
    a.cand_at = m->ncells + 1;
    a.count_out = m->dep_host_dev;
    // The previous call's count of cells above the listing threshold (written by
    // the device into host memory; any value gives the same sums): many such
    // cells fill the device with one lane per cell, few do not and take a
    // wavefront each.
    a.dense_min = scheme == MCS_DEPOSIT_CIC && *(volatile int32_t *)m->dep_host > m->dep_lane_listed
                      ? INT32_MAX
                      : m->dep_dense;
    {
        ProfScope ps(ctx, "deposit");
        // one thread per cell of the occupied box (the grid covers the mesh; the
        // threads past the box end at once)
        dim3 g((unsigned)cdiv(m->ncells, MCS_BLOCK)), bl(MCS_BLOCK);
        if (scheme == MCS_DEPOSIT_CIC) {
            // the prep kernel also clears the list count (one block at least)
            hipLaunchKernelGGL(k_deposit_prep<Buf>, dim3((unsigned)cdiv(p->n > 0 ? p->n : 1, MCS_BLOCK)), bl, 0,
                               ctx->stream, a);
            hipLaunchKernelGGL((k_deposit<1, Buf>), g, bl, 0, ctx->stream, a);
            hipLaunchKernelGGL(k_deposit_dense<Buf>, dim3(2048), bl, 0, ctx->stream, a);
        } else {
            hipLaunchKernelGGL((k_deposit<0, Buf>), g, bl, 0, ctx->stream, a);
        }
    }
    MCS_LAUNCH_CHECK(ctx);
    return MCS_OK;
}

// ---------------------------------------------------------------- charge, faces

__global__ __launch_bounds__(MCS_BLOCK) void k_assemble(MeshGeom g, const double *qf,
                                                        const double *qp, double *q)
{
    int64_t c = gidx();
    int64_t sx = g.n[0] + 2, sy = g.n[1] + 2;
    int64_t tot = sx * sy * (g.n[2] + 2);
    if (c >= tot) return;
    int64_t i = c % sx, j = (c / sx) % sy, k = c / (sx * sy);
    bool interior = i >= 1 && i <= g.n[0] && j >= 1 && j <= g.n[1] && k >= 1 && k <= g.n[2];
    q[c] = interior ? (qf[c] + qp[c]) * MCS_ELEMENTARY_CHARGE : 0.0;
}

extern "C" int32_t mcs_assemble_charge(mcs_mesh *m)
{
    if (!m) return MCS_E_ARG;
    mcs_ctx *ctx = m->ctx;
    hipSetDevice(ctx->device);
    hipLaunchKernelGGL(k_assemble, dim3((unsigned)cdiv(m->ncg, MCS_BLOCK)), dim3(MCS_BLOCK), 0,
                       ctx->stream, geom(m), (const double *)m->arr[MCS_ARR_QFIXED],
                       (const double *)m->arr[MCS_ARR_QPART], m->arr[MCS_ARR_Q]);
    MCS_LAUNCH_CHECK(ctx);
    return MCS_OK;
}

__device__ __forceinline__ int64_t clampi(int64_t v, int64_t lo, int64_t hi)
{
    return v < lo ? lo : (v > hi ? hi : v);
}

// Face fields (MCStreamer/src/poisson.jl:47-57).  Ghost faces take the value of
// the nearest interior face, which is what the reference's ghost copies produce.
__global__ __launch_bounds__(MCS_BLOCK) void k_faces(MeshGeom g, const double *u, double *ex,
                                                     double *ey, double *ez, double eb, float *fx,
                                                     float *fy, float *fz)
{
    int64_t t = xcd_gidx();
    const int64_t nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const int64_t ux = nx + 2, uxy = (nx + 2) * (ny + 2);
    const double h = g.h;
    int64_t lx = (nx + 1) * (ny + 2) * (nz + 2);
    int64_t ly = (nx + 2) * (ny + 1) * (nz + 2);
    int64_t lz = (nx + 2) * (ny + 2) * (nz + 1);
    if (t < lx) {
        int64_t i = t % (nx + 1), j = (t / (nx + 1)) % (ny + 2), k = t / ((nx + 1) * (ny + 2));
        int64_t jj = clampi(j, 1, ny), kk = clampi(k, 1, nz);
        int64_t c = i + ux * jj + uxy * kk;
        ex[t] = (u[c] - u[c + 1]) / h;
        if (fx) fx[t] = (float)ex[t];
        return;
    }
    t -= lx;
    if (t < ly) {
        int64_t i = t % (nx + 2), j = (t / (nx + 2)) % (ny + 1), k = t / ((nx + 2) * (ny + 1));
        int64_t ii = clampi(i, 1, nx), kk = clampi(k, 1, nz);
        int64_t c = ii + ux * j + uxy * kk;
        ey[t] = (u[c] - u[c + ux]) / h;
        if (fy) fy[t] = (float)ey[t];
        return;
    }
    t -= ly;
    if (t < lz) {
        int64_t i = t % (nx + 2), j = (t / (nx + 2)) % (ny + 2), k = t / ((nx + 2) * (ny + 2));
        int64_t ii = clampi(i, 1, nx), jj = clampi(j, 1, ny);
        int64_t c = ii + ux * jj + uxy * k;
        ez[t] = eb + (u[c] - u[c + uxy]) / h;
        if (fz) fz[t] = (float)ez[t];
    }
}

extern "C" int32_t mcs_efield_faces(mcs_mesh *m, double eb)
{
    if (!m) return MCS_E_ARG;
    mcs_ctx *ctx = m->ctx;
    hipSetDevice(ctx->device);
    int64_t tot = m->arr_len[MCS_ARR_EX] + m->arr_len[MCS_ARR_EY] + m->arr_len[MCS_ARR_EZ];
    {
        ProfScope ps(ctx, "faces");
        hipLaunchKernelGGL(k_faces, dim3((unsigned)cdiv(tot, MCS_BLOCK)), dim3(MCS_BLOCK), 0,
                           ctx->stream, geom(m), (const double *)m->arr[MCS_ARR_U],
                           m->arr[MCS_ARR_EX], m->arr[MCS_ARR_EY], m->arr[MCS_ARR_EZ], eb, m->face32[0],
                           m->face32[1], m->face32[2]);
    }
    MCS_LAUNCH_CHECK(ctx);
    return MCS_OK;
}

__global__ void k_to_f32(float *dst, const double *src, int64_t n)
{
    int64_t i = gidx();
    if (i < n) dst[i] = (float)src[i];
}

int32_t mcs_mesh_faces32(mcs_mesh *m)
{
    if (m->face32[0]) return MCS_OK;
    mcs_ctx *ctx = m->ctx;
    for (int d = 0; d < 3; ++d) {
        const int a = MCS_ARR_EX + d;
        MCS_HIP(ctx, hipMalloc(&m->face32[d], sizeof(float) * m->arr_len[a]));
        hipLaunchKernelGGL(k_to_f32, dim3((unsigned)cdiv(m->arr_len[a], MCS_BLOCK)), dim3(MCS_BLOCK), 0,
                           ctx->stream, m->face32[d], (const double *)m->arr[a], m->arr_len[a]);
    }
    MCS_LAUNCH_CHECK(ctx);
    return MCS_OK;
}

// ---------------------------------------------------------------- resample!

#define R_MERGED 2
#define R_CLONES 3
#define R_SPLITS 4
#define R_REMOVED 5

// One cell of MCStreamer/src/resample.jl:16-76 on the canonical order (the
// oracle restatement documents the rules).  WRITE = false counts the outputs;
// WRITE = true replays the same draws and writes them: surviving originals first,
// then the created particles, ids next_id + (created before this one).
#define RS_PF 8   // weights loaded ahead in resample_cell

// pfinal and wfinal receive the cell's p and wtotal as resample! leaves them
// (resample.jl:16-30, 63-72): the count of its particles plus one per split, and
// the sum of their weights in order.
template <bool WRITE, typename Buf>
__device__ void resample_cell(int64_t c, int64_t a, int64_t e, int32_t ntarget, int32_t nmax,
                              uint64_t seed, uint32_t instance, uint32_t step, const Buf &src,
                              const Buf &dst, int64_t base, uint64_t idbase, int64_t &nout,
                              int64_t &ncreated, int64_t *ctr, int64_t &pfinal, double &wfinal)
{
    int64_t pc = e - a;
    nout = 0;
    ncreated = 0;
    pfinal = 0;
    wfinal = 0.0;
    if (pc == 0) return;
    mcs_stream rs;
    mcs_stream_init(&rs, seed, instance, MCS_PHASE_CELL_RESAMPLE, (uint64_t)c, step);
    double wtotal = 0.0, xx = 0.0, wcum = 0.0;
    int64_t cnt = 0, pk = 0;
    // weights read RS_PF at a time (the loads in flight together), used in order
    for (int64_t i0 = a; i0 < e; i0 += RS_PF) {
        double wv[RS_PF];
#pragma unroll
        for (int u = 0; u < RS_PF; ++u) wv[u] = i0 + u < e ? get_w(src, i0 + u) : 0.0;
#pragma unroll
        for (int u = 0; u < RS_PF; ++u) {
            if (i0 + u >= e) break;
            cnt += 1;
            wtotal += wv[u];
            if (cnt == nmax) xx = mcs_samplemin(&rs, (double)ntarget);
        }
    }
    const bool merged = pc > nmax;
    const int64_t created_at = merged ? 0 : pc;
    int64_t splits = 0, removed = 0;
    for (int64_t i0 = a; i0 < e; i0 += RS_PF) {
        double wv[RS_PF];
#pragma unroll
        for (int u = 0; u < RS_PF; ++u) wv[u] = i0 + u < e ? get_w(src, i0 + u) : 0.0;
#pragma unroll
        for (int u = 0; u < RS_PF; ++u) {
            const int64_t i = i0 + u;
            if (i >= e) break;
            const double wi = wv[u];
            if (pc > nmax) {
                while ((wtotal * xx < wcum + wi) && pk < ntarget) {
                    pk += 1;
                    if (WRITE) {
                        int64_t d = base + created_at + ncreated;
                        copy_rec(src, i, dst, d);
                        set_w(dst, d, wtotal / (double)ntarget);
                        dst.id[d] = idbase + (uint64_t)ncreated;
                    }
                    ncreated += 1;
                    if (pk < ntarget) xx = xx + mcs_samplemin(&rs, (double)(ntarget - pk)) * (1.0 - xx);
                }
                wcum = wcum + wi;
                removed += 1;
            } else {
                double wkeep = wi;
                if (pc < ntarget && wi > 2.0) {
                    double wnew = rint(wi / 2.0);
                    if (WRITE) {
                        int64_t d = base + created_at + ncreated;
                        copy_rec(src, i, dst, d);
                        set_w(dst, d, wnew);
                        dst.id[d] = idbase + (uint64_t)ncreated;
                    }
                    ncreated += 1;
                    splits += 1;
                    wkeep = wi - wnew;
                    pc += 1;
                }
                if (WRITE && wkeep != wi) set_w(dst, base + (i - a), wkeep);  // survivors copied by k_resample_copy
            }
        }
    }
    nout = (merged ? 0 : (e - a)) + ncreated;
    pfinal = pc;
    wfinal = wtotal;
    if (!WRITE && ctr) {
        if (merged) {
            atomicAdd((unsigned long long *)&ctr[R_MERGED], 1ULL);
            atomicAdd((unsigned long long *)&ctr[R_CLONES], (unsigned long long)ncreated);
            atomicAdd((unsigned long long *)&ctr[R_REMOVED], (unsigned long long)removed);
        } else if (splits) {
            atomicAdd((unsigned long long *)&ctr[R_SPLITS], (unsigned long long)splits);
        }
    }
}

template <typename Buf>
struct ResArgs {
    Buf src, dst;
    const int32_t *cs;
    int64_t ncells;
    int32_t ntarget, nmax;
    uint64_t seed;
    uint32_t instance, step;
    int64_t *packed;   // ncells + 1 counts: out | created << 32
    const int64_t *prefix;
    uint64_t next_id;
    int32_t *new_cs;
    int64_t *ctr;
    int64_t *pcell;    // per cell p and wtotal as resample! leaves them, or null
    double *wcell;
};

template <typename Buf>
__global__ __launch_bounds__(MCS_BLOCK) void k_resample_count(ResArgs<Buf> r)
{
    int64_t c = gidx();
    if (c > r.ncells) return;
    int64_t a = r.cs[c], e = r.cs[c + 1];
    if (c == r.ncells) {  // particles outside the mesh are kept as they are
        r.packed[c] = e - a;
        return;
    }
    int64_t nout, ncr, pf;
    double wf;
    resample_cell<false, Buf>(c, a, e, r.ntarget, r.nmax, r.seed, r.instance, r.step, r.src, r.dst, 0,
                         0, nout, ncr, r.ctr, pf, wf);
    r.packed[c] = nout | (ncr << 32);
    if (r.pcell) {
        r.pcell[c] = pf;
        r.wcell[c] = wf;
    }
}

template <typename Buf>
__global__ __launch_bounds__(MCS_BLOCK) void k_resample_write(ResArgs<Buf> r)
{
    int64_t c = gidx();
    if (c > r.ncells) return;
    int64_t a = r.cs[c], e = r.cs[c + 1];
    int64_t pre = r.prefix[c];
    int64_t base = pre & 0xffffffffLL;
    int64_t crbase = pre >> 32;
    r.new_cs[c] = (int32_t)base;
    if (c == r.ncells) {
        int64_t tot = r.prefix[c + 1];
        r.new_cs[c + 1] = (int32_t)(tot & 0xffffffffLL);
        return;
    }
    int64_t nout, ncr, pf;
    double wf;
    resample_cell<true, Buf>(c, a, e, r.ntarget, r.nmax, r.seed, r.instance, r.step, r.src, r.dst, base,
                        r.next_id + (uint64_t)crbase, nout, ncr, nullptr, pf, wf);
}

// Survivors (particles of cells that are not merged, and those outside the mesh)
// move as whole records, one thread per particle; their slot is the cell's output
// start plus their rank in the cell.
template <typename Buf>
__global__ __launch_bounds__(MCS_BLOCK) void k_resample_copy(ResArgs<Buf> r, MeshGeom g, int64_t n)
{
    int64_t i = gidx();
    if (i >= n) return;
    int64_t c = cell_key(r.src, i, g.n, g.h);
    if (c < 0) c = r.ncells;
    int64_t a = r.cs[c];
    if (c < r.ncells && r.cs[c + 1] - a > r.nmax) return;
    int64_t base = r.prefix[c] & 0xffffffffLL;
    copy_rec(r.src, i, r.dst, base + (i - a));
}

template <typename Buf>
static int32_t resample_impl(mcs_pop *p, mcs_mesh *m, const Buf &src, const Buf &dst, int32_t ntarget,
                             int32_t nmax, uint64_t seed, uint32_t instance, uint32_t step,
                             mcs_resample_stats *stats);

extern "C" int32_t mcs_resample_cells(mcs_pop *p, mcs_mesh *m, int32_t ntarget, int32_t nmax,
                                      uint64_t seed, uint32_t instance, uint32_t step,
                                      mcs_resample_stats *stats)
{
    if (!p || !m) return MCS_E_ARG;
    MCS_ELECTRONS_ONLY(p, "resample_cells");
    mcs_ctx *ctx = p->ctx;
    if (ntarget < 1 || nmax < ntarget)
        return mcs_fail(ctx, MCS_E_ARG, "resample: need 1 <= ntarget <= nmax");
    if (!mcs_pop_sorted(p, m->sorted_gen))
        return mcs_fail(ctx, MCS_E_ARG,
                        "resample: the population is not in this mesh's cell order (call mcs_sort_by_cell)");
    hipSetDevice(ctx->device);
    int32_t rc = mcs_pop_ensure_alt(p);
    if (rc) return rc;
    mcs_pop_touch(p);
    rc = p->prec == MCS_PREC_F32
             ? resample_impl(p, m, p->sbuf[0], p->sbuf[1], ntarget, nmax, seed, instance, step, stats)
             : resample_impl(p, m, p->buf[0], p->buf[1], ntarget, nmax, seed, instance, step, stats);
    // the output is in canonical order and the cell starts were rewritten for it
    if (rc == MCS_OK) m->sorted_gen = p->gen;
    return rc;
}

// The cell lists a resample works on: starts (ncells + 2), per-cell packed counts
// and their prefix (ncells + 2 each), and the new starts.
struct CellLists {
    int32_t *cs;
    int64_t ncells;
    int64_t *cnt64, *pre64;
    int32_t *new_cs;
    int64_t *pcell = nullptr;   // resample!'s p and wtotal per cell, when kept
    double *wcell = nullptr;
};

template <typename Buf, typename Copy>
static int32_t resample_run(mcs_pop *p, const CellLists &L, const Buf &src, const Buf &dst, int32_t ntarget,
                            int32_t nmax, uint64_t seed, uint32_t instance, uint32_t step,
                            mcs_resample_stats *stats, Copy &&copy_survivors);

template <typename Buf>
static int32_t resample_impl(mcs_pop *p, mcs_mesh *m, const Buf &src, const Buf &dst, int32_t ntarget,
                             int32_t nmax, uint64_t seed, uint32_t instance, uint32_t step,
                             mcs_resample_stats *stats)
{
    const CellLists L = {m->cell_start, m->ncells, m->cnt64, m->pre64, m->cell_cnt};
    const MeshGeom g = geom(m);
    return resample_run(p, L, src, dst, ntarget, nmax, seed, instance, step, stats,
                        [&](const ResArgs<Buf> &r, int64_t n) {
                            hipLaunchKernelGGL(k_resample_copy<Buf>, dim3((unsigned)cdiv(n, MCS_BLOCK)),
                                               dim3(MCS_BLOCK), 0, p->ctx->stream, r, g, n);
                        });
}

template <typename Buf, typename Copy>
static int32_t resample_run(mcs_pop *p, const CellLists &L, const Buf &src, const Buf &dst, int32_t ntarget,
                            int32_t nmax, uint64_t seed, uint32_t instance, uint32_t step,
                            mcs_resample_stats *stats, Copy &&copy_survivors)
{
    mcs_ctx *ctx = p->ctx;
    int32_t rc;
    if (stats) memset(stats, 0, sizeof(*stats));
    int64_t nseg = L.ncells + 1;
    int64_t *ctr = ctx->d_small;
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, ctx->stream, (int32_t *)ctr, (int64_t)16);
    ResArgs<Buf> r;
    r.src = src;
    r.dst = dst;
    r.cs = L.cs;
    r.ncells = L.ncells;
    r.ntarget = ntarget;
    r.nmax = nmax;
    r.seed = seed;
    r.instance = instance;
    r.step = step;
    r.packed = L.cnt64;
    r.prefix = L.pre64;
    r.next_id = p->next_id;
    r.new_cs = L.new_cs;
    r.ctr = ctr;
    r.pcell = L.pcell;
    r.wcell = L.wcell;
    {
        ProfScope ps(ctx, "resample");
        hipLaunchKernelGGL(k_resample_count<Buf>, dim3((unsigned)cdiv(nseg, MCS_BLOCK)), dim3(MCS_BLOCK), 0,
                           ctx->stream, r);
    }
    MCS_LAUNCH_CHECK(ctx);
    rc = mcs_scan_i64(ctx, L.cnt64, L.pre64, nseg);
    if (rc) return rc;
    int64_t hb[9];
    rc = mcs_readback2(ctx, L.pre64 + nseg, 1, ctr, 8, hb);
    if (rc) return rc;
    int64_t tot = hb[0];
    int64_t h[8];
    memcpy(h, hb + 1, sizeof(h));
    int64_t nout = tot & 0xffffffffLL;
    int64_t ncreated = tot >> 32;
    if (nout > p->cap)
        return mcs_fail(ctx, MCS_E_CAPACITY, "Maximum number of particles reached (%lld > %lld)",
                        (long long)nout, (long long)p->cap);
    {
        ProfScope ps(ctx, "resample");
        if (p->n > 0) copy_survivors(r, p->n);
        hipLaunchKernelGGL(k_resample_write<Buf>, dim3((unsigned)cdiv(nseg, MCS_BLOCK)), dim3(MCS_BLOCK), 0,
                           ctx->stream, r);
    }
    MCS_LAUNCH_CHECK(ctx);
    MCS_HIP(ctx, hipMemcpyAsync(L.cs, L.new_cs, sizeof(int32_t) * (nseg + 1), hipMemcpyDeviceToDevice,
                                ctx->stream));
    mcs_pop_swap(p);
    if (stats) {
        stats->n_in = p->n;
        stats->n_out = nout;
        stats->cells_merged = h[R_MERGED];
        stats->clones = h[R_CLONES];
        stats->splits = h[R_SPLITS];
        stats->removed = h[R_REMOVED];
    }
    p->n = nout;
    p->next_id += (uint64_t)ncreated;
    return MCS_OK;
}

// ------------------------------------------- the axisymmetric grid: cell order, resample!

// resample!'s cell of x on the axisymmetric grid (resample.jl:21-23): cellindex
// with the bounds of its per-cell arrays (calloc_centers: i in 0..M+1, j in
// 0..N+1, no test against rc), flat i + (M+2) j; -1 outside.
__device__ __forceinline__ int64_t cyl_rs_cell(const CylView &C, const double x[3])
{
    const double rho = sqrt(x[0] * x[0] + x[1] * x[1]);
    const double fi = fld_dev(rho - 0.0, C.drf), fj = fld_dev(x[2] - 0.0, C.dzf);
    if (!(fi >= -1.0 && fi <= (double)C.M && fj >= -1.0 && fj <= (double)C.N)) return -1;
    return ((int64_t)fi + 1) + (int64_t)(C.M + 2) * ((int64_t)fj + 1);
}

// (cell, id) keys of the canonical order: cell << idbits | id (ids are below
// 2^idbits), the cells outside as cell ncg, inactive particles as cell ncg + 1
// (dropped: they sort after every kept one).
__global__ __launch_bounds__(MCS_BLOCK) void kr_keys(PopBuf b, int64_t n, CylView C, uint64_t ncg, int idbits,
                                                     uint64_t *key, int32_t *val)
{
    const int64_t p = gidx();
    if (p >= n) return;
    uint64_t k = (ncg + 1) << idbits;
    if (b.active[p]) {
        const double x[3] = {b.x[0][p], b.x[1][p], b.x[2][p]};
        const int64_t c = cyl_rs_cell(C, x);
        k = ((c < 0 ? ncg : (uint64_t)c) << idbits) | b.id[p];
    }
    key[p] = k;
    val[p] = (int32_t)p;
}

// cs[c] = the first sorted key of cell c or later (c = 0..ncg + 1); the last is
// the number of kept particles, also stored as an 8-byte word at *na.
__global__ __launch_bounds__(MCS_BLOCK) void kr_starts(const uint64_t *key, int64_t n, int64_t ncg, int idbits,
                                                       int32_t *cs, int64_t *na)
{
    const int64_t c = gidx();
    if (c > ncg + 1) return;
    const uint64_t t = (uint64_t)c << idbits;
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (key[mid] < t) lo = mid + 1;
        else hi = mid;
    }
    cs[c] = (int32_t)lo;
    if (c == ncg + 1) *na = lo;
}

static int32_t cgrid_lists(mcs_cgrid *g)
{
    if (g->cs) return MCS_OK;
    const int64_t ncg = (int64_t)(g->M + 2) * (g->N + 2);
    if (hipMalloc(&g->cs, sizeof(int32_t) * (ncg + 2)) != hipSuccess ||
        hipMalloc(&g->cs_new, sizeof(int32_t) * (ncg + 2)) != hipSuccess ||
        hipMalloc(&g->cnt64, sizeof(int64_t) * (ncg + 2)) != hipSuccess ||
        hipMalloc(&g->pre64, sizeof(int64_t) * (ncg + 2)) != hipSuccess)
        return mcs_fail(g->ctx, MCS_E_HIP, "cgrid: device allocation failed");
    return mcs_cgrid_resample_arrays(g);
}

// resample!'s per-cell p and wtotal, zero until the first resample.
int32_t mcs_cgrid_resample_arrays(mcs_cgrid *g)
{
    if (g->rs_p) return MCS_OK;
    const int64_t ncg = (int64_t)(g->M + 2) * (g->N + 2);
    if (hipMalloc(&g->rs_p, sizeof(int64_t) * ncg) != hipSuccess ||
        hipMalloc(&g->rs_w, sizeof(double) * ncg) != hipSuccess)
        return mcs_fail(g->ctx, MCS_E_HIP, "cgrid: device allocation failed");
    MCS_HIP(g->ctx, mcs_zero(g->ctx, g->rs_p, sizeof(int64_t) * ncg));
    MCS_HIP(g->ctx, mcs_zero(g->ctx, g->rs_w, sizeof(double) * ncg));
    return MCS_OK;
}

extern "C" int32_t mcs_cyl_sort_by_cell(mcs_pop *p, mcs_cgrid *g)
{
    if (!p || !g) return MCS_E_ARG;
    MCS_ELECTRONS_ONLY(p, "cyl_sort_by_cell");
    mcs_ctx *ctx = p->ctx;
    if (g->ctx != ctx) return mcs_fail(ctx, MCS_E_ARG, "cyl_sort: grid of another context");
    if (p->prec != MCS_PREC_F64) return mcs_fail(ctx, MCS_E_ARG, "cyl_sort: FP64 populations only");
    const int64_t ncg = (int64_t)(g->M + 2) * (g->N + 2);
    // the key holds the id in its low idbits bits and the cell above; the sort
    // looks at those bits only
    int idbits = 1, cbits = 1;
    while (idbits < 63 && (p->next_id >> idbits) != 0) ++idbits;
    while (cbits < 63 && ((uint64_t)(ncg + 1) >> cbits) != 0) ++cbits;
    if (idbits + cbits > 64) return mcs_fail(ctx, MCS_E_ARG, "cyl_sort: cell and id do not fit 64 bits");
    const unsigned endbit = (unsigned)(idbits + cbits);
    hipSetDevice(ctx->device);
    int32_t rc = mcs_pop_ensure_alt(p);
    if (rc) return rc;
    rc = cgrid_lists(g);
    if (rc) return rc;
    const int64_t n = p->n;
    ProfScope ps(ctx, "cyl_sort");
    int64_t na = 0;
    if (n > 0) {
        uint64_t *kin = (uint64_t *)p->u64a;
        int32_t *vin = p->i32a, *vout = p->i32b;
        size_t tb = 0;
        MCS_HIP(ctx, rocprim::radix_sort_pairs(nullptr, tb, kin, kin, vin, vout, (size_t)n, 0u, endbit, ctx->stream));
        const size_t kbytes = ((sizeof(uint64_t) * (size_t)n + 255) / 256) * 256;
        void *tmp = nullptr;
        rc = mcs_scratch(ctx, kbytes + tb, &tmp);
        if (rc) return rc;
        uint64_t *kout = (uint64_t *)tmp;
        const MeshView mv = cyl_view(g);
        hipLaunchKernelGGL(kr_keys, dim3((unsigned)cdiv(n, MCS_BLOCK)), dim3(MCS_BLOCK), 0, ctx->stream,
                           p->buf[0], n, mv.C, (uint64_t)ncg, idbits, kin, vin);
        MCS_HIP(ctx, rocprim::radix_sort_pairs((char *)tmp + kbytes, tb, kin, kout, vin, vout, (size_t)n, 0u, endbit,
                                               ctx->stream));
        hipLaunchKernelGGL(kr_starts, dim3((unsigned)cdiv(ncg + 2, MCS_BLOCK)), dim3(MCS_BLOCK), 0, ctx->stream,
                           (const uint64_t *)kout, n, ncg, idbits, g->cs, g->cnt64 + ncg + 1);
        MCS_LAUNCH_CHECK(ctx);
        rc = mcs_readback(ctx, g->cnt64 + ncg + 1, &na, 1);
        if (rc) return rc;
        if (na > 0)
            hipLaunchKernelGGL(k_sort_gather<PopBuf>, dim3((unsigned)cdiv(na, MCS_BLOCK)), dim3(MCS_BLOCK), 0,
                               ctx->stream, p->buf[0], p->buf[1], (const int32_t *)vout, na);
        MCS_LAUNCH_CHECK(ctx);
        mcs_pop_swap(p);
    } else {
        hipLaunchKernelGGL(k_zero_i32, dim3((unsigned)cdiv(ncg + 2, MCS_BLOCK)), dim3(MCS_BLOCK), 0, ctx->stream,
                           g->cs, ncg + 2);
        MCS_LAUNCH_CHECK(ctx);
    }
    p->n = na;
    mcs_pop_touch(p);
    g->sorted_gen = p->gen;
    return MCS_OK;
}

// Survivors of the (r, z) resample, as k_resample_copy with the axisymmetric cell.
__global__ __launch_bounds__(MCS_BLOCK) void k_resample_copy_cyl(ResArgs<PopBuf> r, CylView C, int64_t n)
{
    int64_t i = gidx();
    if (i >= n) return;
    const double x[3] = {r.src.x[0][i], r.src.x[1][i], r.src.x[2][i]};
    int64_t c = cyl_rs_cell(C, x);
    if (c < 0) c = r.ncells;
    int64_t a = r.cs[c];
    if (c < r.ncells && r.cs[c + 1] - a > r.nmax) return;
    int64_t base = r.prefix[c] & 0xffffffffLL;
    copy_rec(r.src, i, r.dst, base + (i - a));
}

extern "C" int32_t mcs_cyl_resample(mcs_pop *p, mcs_cgrid *g, int32_t ntarget, int32_t nmax, uint64_t seed,
                                    uint32_t instance, uint32_t step, mcs_resample_stats *stats)
{
    if (!p || !g) return MCS_E_ARG;
    MCS_ELECTRONS_ONLY(p, "cyl_resample");
    mcs_ctx *ctx = p->ctx;
    if (ntarget < 1 || nmax < ntarget) return mcs_fail(ctx, MCS_E_ARG, "resample: need 1 <= ntarget <= nmax");
    if (!mcs_pop_sorted(p, g->sorted_gen))
        return mcs_fail(ctx, MCS_E_ARG,
                        "resample: the population is not in this grid's cell order (call mcs_cyl_sort_by_cell)");
    hipSetDevice(ctx->device);
    int32_t rc = mcs_pop_ensure_alt(p);
    if (rc) return rc;
    const int64_t ncg = (int64_t)(g->M + 2) * (g->N + 2);
    const CellLists L = {g->cs, ncg, g->cnt64, g->pre64, g->cs_new, g->rs_p, g->rs_w};
    const CylView C = cyl_view(g).C;
    mcs_pop_touch(p);
    rc = resample_run(p, L, p->buf[0], p->buf[1], ntarget, nmax, seed, instance, step, stats,
                      [&](const ResArgs<PopBuf> &r, int64_t n) {
                          hipLaunchKernelGGL(k_resample_copy_cyl, dim3((unsigned)cdiv(n, MCS_BLOCK)),
                                             dim3(MCS_BLOCK), 0, ctx->stream, r, C, n);
                      });
    if (rc == MCS_OK) g->sorted_gen = p->gen;
    return rc;
}
