// mcs_api.hip -- context, collision table, particle population and the
// population reductions of libmcs.
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <map>
#include <mutex>

#include "mcs_device.h"
#include "mcs_fp32.h"

// ---------------------------------------------------------------- errors

int32_t mcs_fail(mcs_ctx *ctx, int32_t code, const char *fmt, ...)
{
    if (ctx) {
        char buf[1024];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return code;
}

extern "C" const char *mcs_last_error(const mcs_ctx *ctx)
{
    return ctx ? ctx->err.c_str() : "null context";
}

// ---------------------------------------------------------------- context

static int32_t noalloc_setup(int device);

extern "C" int32_t mcs_ctx_create(int32_t device, mcs_ctx **out)
{
    if (!out) return MCS_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return MCS_E_HIP;
    if (device < 0 || device >= ndev) return MCS_E_ARG;
    if (hipSetDevice(device) != hipSuccess) return MCS_E_HIP;
    if (noalloc_setup(device) != MCS_OK) return MCS_E_HIP;
    mcs_ctx *c = new mcs_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return MCS_E_HIP;
    }
    if (hipHostMalloc((void **)&c->h_small, 512 * sizeof(int64_t), hipHostMallocDefault) !=
            hipSuccess ||
        hipMalloc((void **)&c->d_small, 512 * sizeof(int64_t)) != hipSuccess ||
        hipHostMalloc((void **)&c->h_pub, (MCS_PUB_WORDS + 1) * sizeof(uint64_t),
                      hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
        hipHostGetDevicePointer((void **)&c->d_pub, c->h_pub, 0) != hipSuccess) {
        delete c;
        return MCS_E_HIP;
    }
    __atomic_store_n(c->h_pub, (uint64_t)0, __ATOMIC_RELEASE);
    *out = c;
    return MCS_OK;
}

extern "C" int32_t mcs_ctx_destroy(mcs_ctx *ctx)
{
    if (!ctx) return MCS_OK;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    for (auto &kv : ctx->profile)
        for (auto &pr : kv.second.pending) {
            hipEventDestroy(pr.first);
            hipEventDestroy(pr.second);
        }
    if (ctx->scratch) hipFree(ctx->scratch);
    if (ctx->d_small) hipFree(ctx->d_small);
    if (ctx->h_small) hipHostFree(ctx->h_small);
    if (ctx->h_pub) hipHostFree(ctx->h_pub);
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return MCS_OK;
}

extern "C" int32_t mcs_ctx_synchronize(mcs_ctx *ctx)
{
    if (!ctx) return MCS_E_ARG;
    MCS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return MCS_OK;
}

extern "C" int32_t mcs_ctx_stream(mcs_ctx *ctx, void **stream)
{
    if (!ctx || !stream) return MCS_E_ARG;
    *stream = (void *)ctx->stream;
    return MCS_OK;
}

extern "C" int32_t mcs_ctx_profile(mcs_ctx *ctx, int32_t enable)
{
    if (!ctx) return MCS_E_ARG;
    ctx->prof = enable != 0;
    return MCS_OK;
}

static int32_t settle_profile(mcs_ctx *ctx)
{
    MCS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (auto &kv : ctx->profile) {
        for (auto &pr : kv.second.pending) {
            float ms = 0.f;
            hipEventElapsedTime(&ms, pr.first, pr.second);
            kv.second.total_ms += ms;
            kv.second.launches += 1;
            hipEventDestroy(pr.first);
            hipEventDestroy(pr.second);
        }
        kv.second.pending.clear();
    }
    return MCS_OK;
}

extern "C" int32_t mcs_ctx_profile_read(mcs_ctx *ctx, const char *kernel, double *total_ms,
                                        int64_t *launches)
{
    if (!ctx || !kernel) return MCS_E_ARG;
    int32_t rc = settle_profile(ctx);
    if (rc) return rc;
    auto it = ctx->profile.find(kernel);
    if (it == ctx->profile.end()) {
        if (total_ms) *total_ms = 0.0;
        if (launches) *launches = 0;
        return MCS_OK;
    }
    if (total_ms) *total_ms = it->second.total_ms;
    if (launches) *launches = it->second.launches;
    return MCS_OK;
}

extern "C" int32_t mcs_ctx_profile_reset(mcs_ctx *ctx)
{
    if (!ctx) return MCS_E_ARG;
    int32_t rc = settle_profile(ctx);
    if (rc) return rc;
    ctx->profile.clear();
    return MCS_OK;
}

int32_t mcs_scratch(mcs_ctx *ctx, size_t bytes, void **out)
{
    if (bytes > ctx->scratch_bytes) {
        if (ctx->scratch) {
            MCS_HIP(ctx, hipStreamSynchronize(ctx->stream));
            MCS_HIP(ctx, hipFree(ctx->scratch));
            ctx->scratch = nullptr;
        }
        size_t nb = bytes < (1u << 20) ? (1u << 20) : bytes + bytes / 4;
        MCS_HIP(ctx, hipMalloc(&ctx->scratch, nb));
        ctx->scratch_bytes = nb;
    }
    *out = ctx->scratch;
    return MCS_OK;
}

__global__ void k_publish(const uint64_t *src, int n, const uint64_t *src2, int n2, uint64_t *pub, uint64_t seq)
{
    for (int t = threadIdx.x; t < n + n2; t += blockDim.x) pub[1 + t] = t < n ? src[t] : src2[t - n];
    __threadfence_system();
    __syncthreads();
    if (threadIdx.x == 0) mcs_pub_raise(pub, seq);
}

uint64_t mcs_pub_next(mcs_ctx *ctx)
{
    return ++ctx->pub_seq;
}

int32_t mcs_pub_wait(mcs_ctx *ctx, uint64_t seq, void *host, int n)
{
    // spin on the sequence number; now and then ask the stream, so that a failed
    // or lost publication ends in an error instead of a hang
    for (uint64_t it = 1;; ++it) {
        if (__atomic_load_n(ctx->h_pub, __ATOMIC_ACQUIRE) == seq) break;
        if ((it & 1023) == 0) {
            const hipError_t q = hipStreamQuery(ctx->stream);
            if (q == hipErrorNotReady) continue;
            if (__atomic_load_n(ctx->h_pub, __ATOMIC_ACQUIRE) == seq) break;
            if (q != hipSuccess)
                return mcs_fail(ctx, MCS_E_HIP, "read-back: stream failed: %s", hipGetErrorString(q));
            return mcs_fail(ctx, MCS_E_HIP, "read-back: sequence %llu not published by an idle stream",
                            (unsigned long long)seq);
        }
    }
    if (n > 0) memcpy(host, (const void *)(ctx->h_pub + 1), sizeof(uint64_t) * n);
    return MCS_OK;
}

int32_t mcs_readback2(mcs_ctx *ctx, const void *dev, int n, const void *dev2, int n2, void *host)
{
    if (n < 0 || n2 < 0 || n + n2 > MCS_PUB_WORDS)
        return mcs_fail(ctx, MCS_E_ARG, "read-back of %d words", n + n2);
    const uint64_t seq = mcs_pub_next(ctx);
    hipLaunchKernelGGL(k_publish, dim3(1), dim3(MCS_WAVE), 0, ctx->stream, (const uint64_t *)dev, n,
                       (const uint64_t *)dev2, n2, ctx->d_pub, seq);
    MCS_LAUNCH_CHECK(ctx);
    return mcs_pub_wait(ctx, seq, host, n + n2);
}

int32_t mcs_readback(mcs_ctx *ctx, const int64_t *dev, int64_t *host, int n)
{
    return mcs_readback2(ctx, dev, n, nullptr, 0, host);
}

// ---------------------------------------------------------------- scans

#define SCAN_ITEMS 8
#define SCAN_TILE (MCS_BLOCK * SCAN_ITEMS)

template <typename TI, typename TO>
__global__ __launch_bounds__(MCS_BLOCK) void k_scan_reduce(const TI *in, int64_t len, TO *bsum)
{
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    TO acc = 0;
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        int64_t i = base + (int64_t)it * MCS_BLOCK + threadIdx.x;
        if (i < len) acc += (TO)in[i];
    }
    TO tot;
    block_excl_scan<TO>(acc, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

template <typename TO>
__global__ __launch_bounds__(MCS_BLOCK) void k_scan_top(TO *bsum, int64_t nb, TO *total_out)
{
    TO carry = 0;
    for (int64_t base = 0; base < nb; base += MCS_BLOCK) {
        int64_t i = base + threadIdx.x;
        TO v = i < nb ? bsum[i] : 0;
        TO tot;
        TO ex = block_excl_scan<TO>(v, &tot);
        if (i < nb) bsum[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) *total_out = carry;
}

template <typename TI, typename TO>
__global__ __launch_bounds__(MCS_BLOCK) void k_scan_down(const TI *in, int64_t len,
                                                         const TO *bsum, TO *out)
{
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    TO vals[SCAN_ITEMS];
    TO acc = 0;
#pragma unroll
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        int64_t i = base + it;
        vals[it] = i < len ? (TO)in[i] : 0;
        acc += vals[it];
    }
    TO tot;
    TO ex = block_excl_scan<TO>(acc, &tot) + bsum[blockIdx.x];
#pragma unroll
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        int64_t i = base + it;
        if (i < len) out[i] = ex;
        ex += vals[it];
    }
}

template <typename TI, typename TO>
static int32_t scan_impl(mcs_ctx *ctx, const TI *in, TO *out, int64_t len, TO *total = nullptr)
{
    int64_t nb = cdiv(len, SCAN_TILE);
    if (nb == 0) nb = 1;
    void *scr;
    int32_t rc = mcs_scratch(ctx, sizeof(TO) * (size_t)(nb + 1), &scr);
    if (rc) return rc;
    TO *bsum = (TO *)scr;
    hipLaunchKernelGGL((k_scan_reduce<TI, TO>), dim3((unsigned)nb), dim3(MCS_BLOCK), 0,
                       ctx->stream, in, len, bsum);
    hipLaunchKernelGGL((k_scan_top<TO>), dim3(1), dim3(MCS_BLOCK), 0, ctx->stream, bsum, nb,
                       total ? total : out + len);
    hipLaunchKernelGGL((k_scan_down<TI, TO>), dim3((unsigned)nb), dim3(MCS_BLOCK), 0,
                       ctx->stream, in, len, bsum, out);
    MCS_LAUNCH_CHECK(ctx);
    return MCS_OK;
}

int32_t mcs_scan_i64_to(mcs_ctx *ctx, const int64_t *in, int64_t *out, int64_t len, int64_t *total)
{
    return scan_impl<int64_t, int64_t>(ctx, in, out, len, total);
}

// Same scan with a caller-owned block-sum buffer (cdiv(len, 2048) + 1 entries),
// for launch sequences that are captured into graphs.
int32_t mcs_scan_i64_buf(mcs_ctx *ctx, const int64_t *in, int64_t *out, int64_t len, int64_t *bsum)
{
    int64_t nb = cdiv(len, SCAN_TILE);
    if (nb == 0) nb = 1;
    hipLaunchKernelGGL((k_scan_reduce<int64_t, int64_t>), dim3((unsigned)nb), dim3(MCS_BLOCK), 0,
                       ctx->stream, in, len, bsum);
    hipLaunchKernelGGL((k_scan_top<int64_t>), dim3(1), dim3(MCS_BLOCK), 0, ctx->stream, bsum, nb,
                       out + len);
    hipLaunchKernelGGL((k_scan_down<int64_t, int64_t>), dim3((unsigned)nb), dim3(MCS_BLOCK), 0,
                       ctx->stream, in, len, (const int64_t *)bsum, out);
    MCS_LAUNCH_CHECK(ctx);
    return MCS_OK;
}

int32_t mcs_scan_i32(mcs_ctx *ctx, const int32_t *in, int32_t *out, int64_t len)
{
    return scan_impl<int32_t, int32_t>(ctx, in, out, len);
}

int32_t mcs_scan_i64(mcs_ctx *ctx, const int64_t *in, int64_t *out, int64_t len)
{
    return scan_impl<int64_t, int64_t>(ctx, in, out, len);
}

// ---------------------------------------------------------------- table

// Energy bands of the banded majorant: caps emax 2^(k - (MCS_NBAND - 1)), k = 0 ..
// MCS_NBAND - 1, each with the largest summed rate of the non-null rows over the
// nodes up to the first node above the cap (the rates are linear between nodes,
// so that bounds the band), raised by 1e-5 for the FP32 rows; then the unbounded
// band at maxrate.  FP32 triples (cap, majorant, 1 / majorant) for the device.
static void table_bands(mcs_table *t, const double *energy, const double *rate, const int32_t *kind,
                        std::vector<float> &b32)
{
    const int32_t ne = t->ne, nproc = t->nproc;
    std::vector<double> tot((size_t)ne, 0.0);
    for (int32_t k = 0; k < ne; ++k)
        for (int32_t j = 0; j < nproc; ++j)
            if (kind[j] != MCS_PROC_NULL) tot[k] += rate[(size_t)k * nproc + j];
    t->h_band_cap.clear();
    t->h_band_rate.clear();
    b32.clear();
    for (int b = 0; b < MCS_NBAND; ++b) {
        const double cap = t->emax * std::ldexp(1.0, b - (MCS_NBAND - 1));
        double m = 0.0;
        for (int32_t k = 0; k < ne; ++k) {
            m = std::max(m, tot[k]);
            if (energy[k] > cap) break;
        }
        m = std::min(m * (1.0 + 1e-5), t->maxrate);
        if (!(m > 0.0)) m = t->maxrate;
        t->h_band_cap.push_back(cap);
        t->h_band_rate.push_back(m);
        b32.push_back((float)cap);
        b32.push_back((float)m);
        b32.push_back((float)(1.0 / m));
    }
    t->h_band_cap.push_back(HUGE_VAL);
    t->h_band_rate.push_back(t->maxrate);
    b32.push_back(HUGE_VALF);
    b32.push_back((float)t->maxrate);
    b32.push_back((float)(1.0 / t->maxrate));
    t->nband = MCS_NBAND;
}

extern "C" int32_t mcs_table_create(mcs_ctx *ctx, int32_t nproc, int32_t ne, int32_t egrid_kind,
                                    const double *energy, const double *rate,
                                    const int32_t *kind, const double *param, double maxrate,
                                    mcs_table **out)
{
    if (!ctx || !out || !energy || !rate || !kind || !param) return MCS_E_ARG;
    if (nproc < 1 || nproc > MCS_MAX_PROC || ne < 2)
        return mcs_fail(ctx, MCS_E_ARG, "table: nproc must be 1..%d and ne >= 2", MCS_MAX_PROC);
    if (egrid_kind != MCS_EGRID_UNIFORM && egrid_kind != MCS_EGRID_LOG)
        return mcs_fail(ctx, MCS_E_ARG, "table: unknown energy grid kind %d", egrid_kind);
    if (!(maxrate > 0.0)) return mcs_fail(ctx, MCS_E_ARG, "table: maxrate must be positive");
    for (int k = 1; k < ne; ++k)
        if (!(energy[k] > energy[k - 1]))
            return mcs_fail(ctx, MCS_E_ARG, "table: energy nodes must increase");
    if (egrid_kind == MCS_EGRID_LOG && !(energy[0] > 0.0))
        return mcs_fail(ctx, MCS_E_ARG, "table: log grid needs a positive first node");
    for (int j = 0; j < nproc; ++j)
        if (kind[j] < MCS_PROC_NULL || kind[j] > MCS_PROC_PHOTOEMISSION)
            return mcs_fail(ctx, MCS_E_ARG, "table: process %d has unknown kind %d", j, kind[j]);
    hipSetDevice(ctx->device);
    mcs_table *t = new mcs_table();
    t->ctx = ctx;
    t->nproc = nproc;
    t->ne = ne;
    t->egrid = egrid_kind;
    t->maxrate = maxrate;
    t->e0 = energy[0];
    t->emax = energy[ne - 1];
    t->step = (energy[ne - 1] - energy[0]) / (double)(ne - 1);
    t->inv_dlog = egrid_kind == MCS_EGRID_LOG
                      ? (double)(ne - 1) / mcs_log(energy[ne - 1] / energy[0])
                      : 0.0;
    t->h_kind.assign(kind, kind + nproc);
    t->h_photo.assign((size_t)nproc * 4, 0.0);
    t->photo_set.assign((size_t)nproc, 0);
    for (int j = 0; j < nproc; ++j) t->nphoto += kind[j] == MCS_PROC_PHOTOEMISSION;
    if (hipMalloc(&t->energy, sizeof(double) * ne) != hipSuccess ||
        hipMalloc(&t->photo, sizeof(double) * 4 * (size_t)nproc) != hipSuccess ||
        hipMalloc(&t->rate, sizeof(double) * (size_t)ne * nproc) != hipSuccess ||
        hipMalloc(&t->kind, sizeof(int32_t) * nproc) != hipSuccess ||
        hipMalloc(&t->param, sizeof(double) * nproc) != hipSuccess) {
        delete t;
        return mcs_fail(ctx, MCS_E_HIP, "table: device allocation failed");
    }
    if (mcs_h2d(ctx, t->energy, energy, sizeof(double) * ne) != hipSuccess ||
        mcs_h2d(ctx, t->rate, rate, sizeof(double) * (size_t)ne * nproc) != hipSuccess ||
        mcs_h2d(ctx, t->kind, kind, sizeof(int32_t) * nproc) != hipSuccess ||
        mcs_h2d(ctx, t->param, param, sizeof(double) * nproc) != hipSuccess ||
        mcs_h2d(ctx, t->photo, t->h_photo.data(), sizeof(double) * 4 * (size_t)nproc) != hipSuccess) {
        mcs_table_destroy(t);
        return mcs_fail(ctx, MCS_E_HIP, "table: copy to the device failed");
    }
    {
        std::vector<float> e32(energy, energy + ne), r32(rate, rate + (size_t)ne * nproc),
            p32(param, param + nproc);
        if (hipMalloc(&t->energy32, sizeof(float) * ne) != hipSuccess ||
            hipMalloc(&t->rate32, sizeof(float) * (size_t)ne * nproc) != hipSuccess ||
            hipMalloc(&t->param32, sizeof(float) * nproc) != hipSuccess) {
            mcs_table_destroy(t);
            return mcs_fail(ctx, MCS_E_HIP, "table: device allocation failed");
        }
        if (mcs_h2d(ctx, t->energy32, e32.data(), sizeof(float) * ne) != hipSuccess ||
            mcs_h2d(ctx, t->rate32, r32.data(), sizeof(float) * r32.size()) != hipSuccess ||
            mcs_h2d(ctx, t->param32, p32.data(), sizeof(float) * nproc) != hipSuccess) {
            mcs_table_destroy(t);
            return mcs_fail(ctx, MCS_E_HIP, "table: copy to the device failed");
        }
    }
    {
        std::vector<float> b32;
        table_bands(t, energy, rate, kind, b32);
        if (hipMalloc(&t->band32, sizeof(float) * b32.size()) != hipSuccess) {
            mcs_table_destroy(t);
            return mcs_fail(ctx, MCS_E_HIP, "table: device allocation failed");
        }
        if (mcs_h2d(ctx, t->band32, b32.data(), sizeof(float) * b32.size()) != hipSuccess) {
            mcs_table_destroy(t);
            return mcs_fail(ctx, MCS_E_HIP, "table: copy to the device failed");
        }
    }
    *out = t;
    return MCS_OK;
}

extern "C" int32_t mcs_table_majorants(const mcs_table *t, int32_t nmax, int32_t *nbands, double *ecap, double *rate)
{
    if (!t || !nbands || nmax < 0 || (nmax > 0 && (!ecap || !rate))) return MCS_E_ARG;
    const int32_t nb = (int32_t)t->h_band_cap.size();
    *nbands = nb;
    for (int32_t k = 0; k < nb && k < nmax; ++k) {
        ecap[k] = t->h_band_cap[k];
        rate[k] = t->h_band_rate[k];
    }
    return MCS_OK;
}

extern "C" int32_t mcs_table_set_photoemission(mcs_table *t, int32_t row, double log_numin, double log_numax,
                                               double photon_multiplier, double photon_weight)
{
    if (!t) return MCS_E_ARG;
    mcs_ctx *ctx = t->ctx;
    if (row < 0 || row >= t->nproc || t->h_kind[row] != MCS_PROC_PHOTOEMISSION)
        return mcs_fail(ctx, MCS_E_ARG, "table: row %d is not a photoemission row", row);
    if (!(log_numax >= log_numin) || !std::isfinite(log_numin) || !std::isfinite(log_numax) ||
        !(photon_multiplier > 0.0) || !(photon_weight > 0.0))
        return mcs_fail(ctx, MCS_E_ARG,
                        "table: photoemission needs finite log_numin <= log_numax and positive multiplier and weight");
    double *r = &t->h_photo[(size_t)row * 4];
    r[0] = log_numin;
    r[1] = log_numax;
    r[2] = photon_multiplier;
    r[3] = photon_weight;
    hipSetDevice(ctx->device);
    MCS_HIP(ctx, mcs_h2d(ctx, t->photo + (size_t)row * 4, r, sizeof(double) * 4));
    t->photo_set[row] = 1;
    return MCS_OK;
}

extern "C" int32_t mcs_table_destroy(mcs_table *t)
{
    if (!t) return MCS_OK;
    hipSetDevice(t->ctx->device);
    hipFree(t->photo);
    hipFree(t->energy);
    hipFree(t->rate);
    hipFree(t->kind);
    hipFree(t->param);
    hipFree(t->energy32);
    hipFree(t->rate32);
    hipFree(t->param32);
    hipFree(t->band32);
    delete t;
    return MCS_OK;
}

// ---------------------------------------------------------------- population

static int32_t alloc_buf(mcs_ctx *ctx, PopBuf &b, int64_t cap, bool photon = false)
{
    if (photon) MCS_HIP(ctx, hipMalloc(&b.nu, sizeof(double) * cap));
    for (int d = 0; d < 3; ++d) {
        MCS_HIP(ctx, hipMalloc(&b.x[d], sizeof(double) * cap));
        MCS_HIP(ctx, hipMalloc(&b.v[d], sizeof(double) * cap));
    }
    MCS_HIP(ctx, hipMalloc(&b.w, sizeof(double) * cap));
    MCS_HIP(ctx, hipMalloc(&b.s, sizeof(double) * cap));
    MCS_HIP(ctx, hipMalloc(&b.active, cap));
    MCS_HIP(ctx, hipMalloc(&b.id, sizeof(uint64_t) * cap));
    return MCS_OK;
}

static void free_buf(PopBuf &b)
{
    hipFree(b.nu);
    for (int d = 0; d < 3; ++d) {
        hipFree(b.x[d]);
        hipFree(b.v[d]);
    }
    hipFree(b.w);
    hipFree(b.s);
    hipFree(b.active);
    hipFree(b.id);
    b = PopBuf();
}

static int32_t alloc_split(mcs_ctx *ctx, SplitBuf &b, int64_t cap)
{
    MCS_HIP(ctx, hipMalloc(&b.cell, sizeof(int32_t) * cap));
    for (int d = 0; d < 3; ++d) {
        MCS_HIP(ctx, hipMalloc(&b.off[d], sizeof(float) * cap));
        MCS_HIP(ctx, hipMalloc(&b.v[d], sizeof(float) * cap));
    }
    MCS_HIP(ctx, hipMalloc(&b.w, sizeof(float) * cap));
    MCS_HIP(ctx, hipMalloc(&b.s, sizeof(float) * cap));
    MCS_HIP(ctx, hipMalloc(&b.active, cap));
    MCS_HIP(ctx, hipMalloc(&b.id, sizeof(uint64_t) * cap));
    return MCS_OK;
}

static void free_split(SplitBuf &b)
{
    hipFree(b.cell);
    for (int d = 0; d < 3; ++d) {
        hipFree(b.off[d]);
        hipFree(b.v[d]);
    }
    hipFree(b.w);
    hipFree(b.s);
    hipFree(b.active);
    hipFree(b.id);
    b = SplitBuf();
}

// Host particles (x, v as n x 3, FP64) into the split store of a population.
__global__ __launch_bounds__(MCS_BLOCK) void k_split_in(SplitBuf b, int64_t n, const double *x, const double *v,
                                                        const double *w, const double *s, const uint8_t *act,
                                                        double inv_h, int32_t n0, int32_t n1, int32_t n2)
{
    int64_t i = gidx();
    if (i >= n) return;
    const int32_t nmax[3] = {n0, n1, n2};
    int32_t c[3];
    float off[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const double r = x[3 * i + q] * inv_h;
        const double fl = floor(r);
        c[q] = (int32_t)fmax(fmin(fl, 1e9), -1e9);
        off[q] = (float)(r - fl);
        b.v[q][i] = (float)v[3 * i + q];
    }
    rebase32(c, off, nmax);
    b.cell[i] = pack_cell(c[0], c[1], c[2]);
#pragma unroll
    for (int q = 0; q < 3; ++q) b.off[q][i] = off[q];
    b.w[i] = (float)w[i];
    b.s[i] = (float)s[i];
    b.active[i] = act[i];
}

// The split store as FP64 host layout (x, v as n x 3).
__global__ __launch_bounds__(MCS_BLOCK) void k_split_out(SplitBuf b, int64_t n, double h, double *x, double *v,
                                                         double *w, double *s)
{
    int64_t i = gidx();
    if (i >= n) return;
    double xs[3], vs[3], ws;
    load_state(b, i, h, xs, vs, ws);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        x[3 * i + q] = xs[q];
        v[3 * i + q] = vs[q];
    }
    w[i] = ws;
    s[i] = (double)b.s[i];
}

static int32_t alloc_pop_common(mcs_ctx *ctx, mcs_pop *p, int64_t capacity)
{
    int64_t nblk = cdiv(capacity, MCS_BLOCK) + 2;
    if (hipMalloc(&p->log_parent, sizeof(int32_t) * p->logcap) != hipSuccess ||
        hipMalloc(&p->log_seq, sizeof(int32_t) * p->logcap) != hipSuccess ||
        hipMalloc(&p->log_act, p->logcap) != hipSuccess ||
        hipMalloc(&p->i32a, sizeof(int32_t) * (capacity + 1)) != hipSuccess ||
        hipMalloc(&p->i32b, sizeof(int32_t) * (capacity + 1)) != hipSuccess ||
        hipMalloc(&p->i32c, sizeof(int32_t) * (capacity + 1)) != hipSuccess ||
        hipMalloc(&p->u64a, sizeof(uint64_t) * (capacity + 1)) != hipSuccess ||
        hipMalloc(&p->blk, sizeof(int64_t) * (2 * nblk + 4)) != hipSuccess)
        return mcs_fail(ctx, MCS_E_HIP, "population: device allocation failed");
    return MCS_OK;
}

extern "C" int32_t mcs_pop_create_split(mcs_ctx *ctx, int64_t capacity, const mcs_mesh *m, mcs_pop **out)
{
    if (!ctx || !out || !m) return MCS_E_ARG;
    if (capacity < 1 || capacity >= (int64_t)1 << 31)
        return mcs_fail(ctx, MCS_E_ARG, "population capacity must be in [1, 2^31)");
    if (m->n[0] > MCS_CX_MAX || m->n[1] > MCS_CY_MAX || m->n[2] > MCS_CZ_MAX)
        return mcs_fail(ctx, MCS_E_ARG, "split population: mesh exceeds %d x %d x %d cells", MCS_CX_MAX,
                        MCS_CY_MAX, MCS_CZ_MAX);
    hipSetDevice(ctx->device);
    mcs_pop *p = new mcs_pop();
    p->ctx = ctx;
    p->cap = capacity;
    p->prec = MCS_PREC_F32;
    for (int d = 0; d < 3; ++d) p->gn[d] = m->n[d];
    p->gh = m->h;
    p->logcap = capacity;
    int32_t rc = alloc_split(ctx, p->sbuf[0], capacity);
    if (!rc) rc = alloc_pop_common(ctx, p, capacity);
    if (!rc) {
        if (hipMalloc(&p->log_cell, sizeof(int32_t) * p->logcap) != hipSuccess ||
            hipMalloc(&p->log_w32, sizeof(float) * p->logcap) != hipSuccess ||
            hipMalloc(&p->log_s32, sizeof(float) * p->logcap) != hipSuccess)
            rc = mcs_fail(ctx, MCS_E_HIP, "population: device allocation failed");
        for (int d = 0; d < 3 && !rc; ++d)
            if (hipMalloc(&p->log_off[d], sizeof(float) * p->logcap) != hipSuccess ||
                hipMalloc(&p->log_v32[d], sizeof(float) * p->logcap) != hipSuccess)
                rc = mcs_fail(ctx, MCS_E_HIP, "population: device allocation failed");
    }
    if (rc) {
        mcs_pop_destroy(p);
        return rc;
    }
    *out = p;
    return MCS_OK;
}

static int32_t pop_create_fp64(mcs_ctx *ctx, int64_t capacity, int species, mcs_pop **out)
{
    if (!ctx || !out) return MCS_E_ARG;
    if (capacity < 1 || capacity >= (int64_t)1 << 31)
        return mcs_fail(ctx, MCS_E_ARG, "population capacity must be in [1, 2^31)");
    hipSetDevice(ctx->device);
    mcs_pop *p = new mcs_pop();
    p->ctx = ctx;
    p->cap = capacity;
    p->n = 0;
    p->next_id = 0;
    p->species = species;
    int32_t rc = alloc_buf(ctx, p->buf[0], capacity, species == MCS_SPECIES_PHOTON);
    if (rc) {
        free_buf(p->buf[0]);
        delete p;
        return rc;
    }
    p->logcap = capacity;
    int64_t nblk = cdiv(capacity, MCS_BLOCK) + 2;
    if (hipMalloc(&p->log_parent, sizeof(int32_t) * p->logcap) != hipSuccess ||
        hipMalloc(&p->log_seq, sizeof(int32_t) * p->logcap) != hipSuccess ||
        hipMalloc(&p->log_w, sizeof(double) * p->logcap) != hipSuccess ||
        hipMalloc(&p->log_s, sizeof(double) * p->logcap) != hipSuccess ||
        hipMalloc(&p->log_act, p->logcap) != hipSuccess ||
        hipMalloc(&p->i32a, sizeof(int32_t) * (capacity + 1)) != hipSuccess ||
        hipMalloc(&p->i32b, sizeof(int32_t) * (capacity + 1)) != hipSuccess ||
        hipMalloc(&p->i32c, sizeof(int32_t) * (capacity + 1)) != hipSuccess ||
        hipMalloc(&p->u64a, sizeof(uint64_t) * (capacity + 1)) != hipSuccess ||
        hipMalloc(&p->blk, sizeof(int64_t) * (2 * nblk + 4)) != hipSuccess)
        return mcs_fail(ctx, MCS_E_HIP, "population: device allocation failed");
    for (int d = 0; d < 3; ++d) {
        if (hipMalloc(&p->log_x[d], sizeof(double) * p->logcap) != hipSuccess ||
            hipMalloc(&p->log_v[d], sizeof(double) * p->logcap) != hipSuccess)
            return mcs_fail(ctx, MCS_E_HIP, "population: device allocation failed");
    }
    *out = p;
    return MCS_OK;
}

extern "C" int32_t mcs_pop_create(mcs_ctx *ctx, int64_t capacity, mcs_pop **out)
{
    return pop_create_fp64(ctx, capacity, MCS_SPECIES_ELECTRON, out);
}

extern "C" int32_t mcs_pop_create_photons(mcs_ctx *ctx, int64_t capacity, mcs_pop **out)
{
    return pop_create_fp64(ctx, capacity, MCS_SPECIES_PHOTON, out);
}

int32_t mcs_pop_ensure_bursts(mcs_pop *p)
{
    if (p->bur_parent) return MCS_OK;
    const int64_t nblk = cdiv(p->cap, MCS_BLOCK) + 2;
    bool ok = hipMalloc(&p->bur_parent, sizeof(int32_t) * p->logcap) == hipSuccess &&
              hipMalloc(&p->bur_row, sizeof(int32_t) * p->logcap) == hipSuccess &&
              hipMalloc(&p->bur_base, sizeof(int64_t) * p->logcap) == hipSuccess &&
              hipMalloc(&p->bur_n, sizeof(int64_t) * p->logcap) == hipSuccess &&
              hipMalloc(&p->bur_act, p->logcap) == hipSuccess &&
              hipMalloc(&p->poff, sizeof(int64_t) * (p->cap + 1)) == hipSuccess &&
              hipMalloc(&p->pblk, sizeof(int64_t) * (2 * nblk + 4)) == hipSuccess;
    for (int d = 0; d < 3 && ok; ++d) ok = hipMalloc(&p->bur_x[d], sizeof(double) * p->logcap) == hipSuccess;
    if (!ok) {
        hipFree(p->bur_parent);
        hipFree(p->bur_row);
        hipFree(p->bur_base);
        hipFree(p->bur_n);
        hipFree(p->bur_act);
        hipFree(p->poff);
        hipFree(p->pblk);
        for (int d = 0; d < 3; ++d) hipFree(p->bur_x[d]);
        p->bur_parent = p->bur_row = nullptr;
        p->bur_base = p->bur_n = p->poff = p->pblk = nullptr;
        p->bur_act = nullptr;
        for (int d = 0; d < 3; ++d) p->bur_x[d] = nullptr;
        return mcs_fail(p->ctx, MCS_E_HIP, "population: photon burst log allocation failed");
    }
    return MCS_OK;
}

int32_t mcs_pop_ensure_track(mcs_pop *p)
{
    if (p->trk_parent) return MCS_OK;
    const int64_t nblk = cdiv(p->cap, MCS_BLOCK) + 2;
    if (hipMalloc(&p->trk_parent, sizeof(int32_t) * p->logcap) != hipSuccess ||
        hipMalloc(&p->trk_seq, sizeof(int32_t) * p->logcap) != hipSuccess ||
        hipMalloc(&p->trk_cell, sizeof(int32_t) * p->logcap) != hipSuccess ||
        hipMalloc(&p->trk_val, sizeof(double) * p->logcap) != hipSuccess ||
        hipMalloc(&p->tblk, sizeof(int64_t) * (2 * nblk + 4)) != hipSuccess) {
        hipFree(p->trk_parent);
        hipFree(p->trk_seq);
        hipFree(p->trk_cell);
        hipFree(p->trk_val);
        hipFree(p->tblk);
        p->trk_parent = p->trk_seq = p->trk_cell = nullptr;
        p->trk_val = nullptr;
        p->tblk = nullptr;
        return mcs_fail(p->ctx, MCS_E_HIP, "population: tracker log allocation failed");
    }
    return MCS_OK;
}

int32_t mcs_pop_ensure_alt(mcs_pop *p)
{
    if (p->has_alt) return MCS_OK;
    int32_t rc = p->prec == MCS_PREC_F32 ? alloc_split(p->ctx, p->sbuf[1], p->cap)
                                         : alloc_buf(p->ctx, p->buf[1], p->cap, p->species == MCS_SPECIES_PHOTON);
    if (rc) return rc;
    p->has_alt = true;
    return MCS_OK;
}

void mcs_pop_touch(mcs_pop *p)
{
    static std::atomic<uint64_t> serial{0};
    p->gen = serial.fetch_add(1, std::memory_order_relaxed) + 1;
}

void mcs_pop_swap(mcs_pop *p)
{
    PopBuf t = p->buf[0];
    p->buf[0] = p->buf[1];
    p->buf[1] = t;
    SplitBuf u = p->sbuf[0];
    p->sbuf[0] = p->sbuf[1];
    p->sbuf[1] = u;
}

extern "C" int32_t mcs_pop_destroy(mcs_pop *p)
{
    if (!p) return MCS_OK;
    hipSetDevice(p->ctx->device);
    hipStreamSynchronize(p->ctx->stream);
    if (p->prec == MCS_PREC_F32) {
        free_split(p->sbuf[0]);
        if (p->has_alt) free_split(p->sbuf[1]);
        hipFree(p->log_cell);
        hipFree(p->log_w32);
        hipFree(p->log_s32);
        for (int d = 0; d < 3; ++d) {
            hipFree(p->log_off[d]);
            hipFree(p->log_v32[d]);
        }
    } else {
        free_buf(p->buf[0]);
        if (p->has_alt) free_buf(p->buf[1]);
    }
    hipFree(p->log_parent);
    hipFree(p->log_seq);
    for (int d = 0; d < 3; ++d) {
        hipFree(p->log_x[d]);
        hipFree(p->log_v[d]);
    }
    hipFree(p->log_w);
    hipFree(p->log_s);
    hipFree(p->log_act);
    hipFree(p->i32a);
    hipFree(p->i32b);
    hipFree(p->i32c);
    hipFree(p->u64a);
    hipFree(p->blk);
    if (p->dep_w) hipFree(p->dep_w);
    if (p->dep_k) hipFree(p->dep_k);
    hipFree(p->trk_parent);
    hipFree(p->trk_seq);
    hipFree(p->trk_cell);
    hipFree(p->trk_val);
    hipFree(p->tblk);
    hipFree(p->bur_parent);
    hipFree(p->bur_row);
    hipFree(p->bur_base);
    hipFree(p->bur_n);
    hipFree(p->bur_act);
    for (int d = 0; d < 3; ++d) hipFree(p->bur_x[d]);
    hipFree(p->poff);
    hipFree(p->pblk);
    delete p;
    return MCS_OK;
}

static int32_t upload_impl(mcs_pop *p, int64_t n, const double *x, const double *v, const double *w,
                           const double *s, const uint8_t *active, const uint64_t *id, uint64_t next_id);

extern "C" int32_t mcs_pop_upload(mcs_pop *p, int64_t n, const double *x, const double *v,
                                  const double *w, const double *s, const uint8_t *active,
                                  const uint64_t *id, uint64_t next_id)
{
    if (!p) return MCS_E_ARG;
    if (p->species == MCS_SPECIES_PHOTON)
        return mcs_fail(p->ctx, MCS_E_ARG, "pop_upload: a photon population takes mcs_photons_upload");
    return upload_impl(p, n, x, v, w, s, active, id, next_id);
}

extern "C" int32_t mcs_photons_upload(mcs_pop *p, int64_t n, const double *x, const double *v, const double *nu,
                                      const double *w, const double *s, const uint8_t *active,
                                      const uint64_t *id, uint64_t next_id)
{
    if (!p) return MCS_E_ARG;
    mcs_ctx *ctx = p->ctx;
    if (p->species != MCS_SPECIES_PHOTON)
        return mcs_fail(ctx, MCS_E_ARG, "photons_upload: not a photon population");
    if (n > 0 && !nu) return mcs_fail(ctx, MCS_E_ARG, "photons_upload: null array");
    int32_t rc = upload_impl(p, n, x, v, w, s, active, id, next_id);
    if (rc) return rc;
    if (n > 0) MCS_HIP(ctx, mcs_h2d(ctx, p->buf[0].nu, nu, sizeof(double) * n));
    return MCS_OK;
}

static int32_t upload_impl(mcs_pop *p, int64_t n, const double *x, const double *v, const double *w,
                           const double *s, const uint8_t *active, const uint64_t *id, uint64_t next_id)
{
    mcs_ctx *ctx = p->ctx;
    if (n < 0 || n > p->cap)
        return mcs_fail(ctx, MCS_E_CAPACITY, "Maximum number of particles reached (%lld > %lld)",
                        (long long)n, (long long)p->cap);
    if (n > 0 && (!x || !v || !w || !s || !active))
        return mcs_fail(ctx, MCS_E_ARG, "pop_upload: null array");
    mcs_pop_touch(p);
    hipSetDevice(ctx->device);
    MCS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (p->prec == MCS_PREC_F32) {
        void *scr;
        int32_t rc = mcs_scratch(ctx, (sizeof(double) * 8 + 1) * (size_t)(n > 0 ? n : 1), &scr);
        if (rc) return rc;
        double *dx = (double *)scr, *dv = dx + 3 * n, *dw = dv + 3 * n, *ds = dw + n;
        uint8_t *da = (uint8_t *)(ds + n);
        if (n > 0) {
            MCS_HIP(ctx, mcs_h2d(ctx, dx, x, sizeof(double) * 3 * n));
            MCS_HIP(ctx, mcs_h2d(ctx, dv, v, sizeof(double) * 3 * n));
            MCS_HIP(ctx, mcs_h2d(ctx, dw, w, sizeof(double) * n));
            MCS_HIP(ctx, mcs_h2d(ctx, ds, s, sizeof(double) * n));
            MCS_HIP(ctx, mcs_h2d(ctx, da, active, n));
            hipLaunchKernelGGL(k_split_in, dim3((unsigned)cdiv(n, MCS_BLOCK)), dim3(MCS_BLOCK), 0, ctx->stream,
                               p->sbuf[0], n, (const double *)dx, (const double *)dv, (const double *)dw,
                               (const double *)ds, (const uint8_t *)da, 1.0 / p->gh, p->gn[0], p->gn[1], p->gn[2]);
            MCS_LAUNCH_CHECK(ctx);
            MCS_HIP(ctx, hipStreamSynchronize(ctx->stream));
        }
    }
    std::vector<double> tmp((size_t)(p->prec == MCS_PREC_F32 ? 0 : n));
    PopBuf &b = p->buf[0];
    for (int d = 0; d < 3 && p->prec != MCS_PREC_F32; ++d) {
        for (int64_t i = 0; i < n; ++i) tmp[i] = x[3 * i + d];
        MCS_HIP(ctx, mcs_h2d(ctx, b.x[d], tmp.data(), sizeof(double) * n));
        for (int64_t i = 0; i < n; ++i) tmp[i] = v[3 * i + d];
        MCS_HIP(ctx, mcs_h2d(ctx, b.v[d], tmp.data(), sizeof(double) * n));
    }
    if (p->prec != MCS_PREC_F32) {
        MCS_HIP(ctx, mcs_h2d(ctx, b.w, w, sizeof(double) * n));
        MCS_HIP(ctx, mcs_h2d(ctx, b.s, s, sizeof(double) * n));
        MCS_HIP(ctx, mcs_h2d(ctx, b.active, active, n));
    }
    uint64_t *dst_id = p->prec == MCS_PREC_F32 ? p->sbuf[0].id : b.id;
    uint64_t nid = next_id;
    if (id) {
        MCS_HIP(ctx, mcs_h2d(ctx, dst_id, id, sizeof(uint64_t) * n));
        for (int64_t i = 0; i < n; ++i)
            if (id[i] + 1 > nid) nid = id[i] + 1;
    } else {
        std::vector<uint64_t> ids((size_t)n);
        for (int64_t i = 0; i < n; ++i) ids[i] = next_id + (uint64_t)i;
        MCS_HIP(ctx, mcs_h2d(ctx, dst_id, ids.data(), sizeof(uint64_t) * n));
        nid = next_id + (uint64_t)n;
    }
    p->n = n;
    p->next_id = nid;
    return MCS_OK;
}

static int32_t download_impl(mcs_pop *p, int64_t nmax, int64_t *nout, double *x, double *v, double *w,
                             double *s, uint8_t *active, uint64_t *id, uint64_t *next_id);

extern "C" int32_t mcs_pop_download(mcs_pop *p, int64_t nmax, int64_t *nout, double *x,
                                    double *v, double *w, double *s, uint8_t *active,
                                    uint64_t *id, uint64_t *next_id)
{
    if (!p) return MCS_E_ARG;
    if (p->species == MCS_SPECIES_PHOTON)
        return mcs_fail(p->ctx, MCS_E_ARG, "pop_download: a photon population takes mcs_photons_download");
    return download_impl(p, nmax, nout, x, v, w, s, active, id, next_id);
}

extern "C" int32_t mcs_photons_download(mcs_pop *p, int64_t nmax, int64_t *nout, double *x, double *v,
                                        double *nu, double *w, double *s, uint8_t *active, uint64_t *id,
                                        uint64_t *next_id)
{
    if (!p) return MCS_E_ARG;
    mcs_ctx *ctx = p->ctx;
    if (p->species != MCS_SPECIES_PHOTON)
        return mcs_fail(ctx, MCS_E_ARG, "photons_download: not a photon population");
    int32_t rc = download_impl(p, nmax, nout, x, v, w, s, active, id, next_id);
    if (rc) return rc;
    const int64_t n = p->n < nmax ? p->n : nmax;
    if (n > 0 && nu) MCS_HIP(ctx, hipMemcpy(nu, p->buf[0].nu, sizeof(double) * n, hipMemcpyDeviceToHost));
    return MCS_OK;
}

static int32_t download_impl(mcs_pop *p, int64_t nmax, int64_t *nout, double *x, double *v, double *w,
                             double *s, uint8_t *active, uint64_t *id, uint64_t *next_id)
{
    mcs_ctx *ctx = p->ctx;
    hipSetDevice(ctx->device);
    MCS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (nout) *nout = p->n;
    if (next_id) *next_id = p->next_id;
    int64_t n = p->n < nmax ? p->n : nmax;
    if (n <= 0) return MCS_OK;
    if (p->prec == MCS_PREC_F32) {
        void *scr;
        int32_t rc = mcs_scratch(ctx, sizeof(double) * 8 * (size_t)n, &scr);
        if (rc) return rc;
        double *dx = (double *)scr, *dv = dx + 3 * n, *dw = dv + 3 * n, *ds = dw + n;
        hipLaunchKernelGGL(k_split_out, dim3((unsigned)cdiv(n, MCS_BLOCK)), dim3(MCS_BLOCK), 0, ctx->stream,
                           p->sbuf[0], n, p->gh, dx, dv, dw, ds);
        MCS_LAUNCH_CHECK(ctx);
        MCS_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (x) MCS_HIP(ctx, hipMemcpy(x, dx, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
        if (v) MCS_HIP(ctx, hipMemcpy(v, dv, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
        if (w) MCS_HIP(ctx, hipMemcpy(w, dw, sizeof(double) * n, hipMemcpyDeviceToHost));
        if (s) MCS_HIP(ctx, hipMemcpy(s, ds, sizeof(double) * n, hipMemcpyDeviceToHost));
        if (active) MCS_HIP(ctx, hipMemcpy(active, p->sbuf[0].active, n, hipMemcpyDeviceToHost));
        if (id) MCS_HIP(ctx, hipMemcpy(id, p->sbuf[0].id, sizeof(uint64_t) * n, hipMemcpyDeviceToHost));
        return MCS_OK;
    }
    PopBuf &b = p->buf[0];
    std::vector<double> tmp((size_t)n);
    for (int d = 0; d < 3; ++d) {
        if (x) {
            MCS_HIP(ctx, hipMemcpy(tmp.data(), b.x[d], sizeof(double) * n, hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < n; ++i) x[3 * i + d] = tmp[i];
        }
        if (v) {
            MCS_HIP(ctx, hipMemcpy(tmp.data(), b.v[d], sizeof(double) * n, hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < n; ++i) v[3 * i + d] = tmp[i];
        }
    }
    if (w) MCS_HIP(ctx, hipMemcpy(w, b.w, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (s) MCS_HIP(ctx, hipMemcpy(s, b.s, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (active) MCS_HIP(ctx, hipMemcpy(active, b.active, n, hipMemcpyDeviceToHost));
    if (id) MCS_HIP(ctx, hipMemcpy(id, b.id, sizeof(uint64_t) * n, hipMemcpyDeviceToHost));
    return MCS_OK;
}

extern "C" int32_t mcs_pop_size(mcs_pop *p, int64_t *n)
{
    if (!p || !n) return MCS_E_ARG;
    *n = p->n;
    return MCS_OK;
}

// ---------------------------------------------------------------- moments

struct Neu {
    double s, c;
};

__device__ __forceinline__ void neu_add(Neu &a, double x)
{
    double t = a.s + x;
    if (fabs(a.s) >= fabs(x))
        a.c += (a.s - t) + x;
    else
        a.c += (x - t) + a.s;
    a.s = t;
}

__device__ __forceinline__ void neu_merge(Neu &a, const Neu &b)
{
    double t = a.s + b.s;
    double c;
    if (fabs(a.s) >= fabs(b.s))
        c = (a.s - t) + b.s;
    else
        c = (b.s - t) + a.s;
    a.s = t;
    a.c = a.c + b.c + c;
}

#define MOM_NV 13   // weight, wenergy, sum_energy, x[3], x2[3], v[3]
#define MOM_REC (2 * MOM_NV + 2)  // + maxenergy + nactive
#define MOM_BLOCKS 512

template <typename Buf>
__global__ __launch_bounds__(MCS_BLOCK) void k_moments(Buf b, int64_t n, double gh, double *partial)
{
    Neu acc[MOM_NV];
#pragma unroll
    for (int q = 0; q < MOM_NV; ++q) acc[q] = Neu{0.0, 0.0};
    double emax = -HUGE_VAL;
    double nact = 0.0;
    for (int64_t i = gidx(); i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double xs[3], vs[3], w;
        load_state(b, i, gh, xs, vs, w);
        double v0 = vs[0], v1 = vs[1], v2 = vs[2];
        double eng = energy_dev(v0, v1, v2);
        emax = fmax(emax, eng);
        if (!b.active[i]) continue;
        double x0 = xs[0], x1 = xs[1], x2 = xs[2];
        nact += 1.0;
        neu_add(acc[0], w);
        neu_add(acc[1], w * eng);
        neu_add(acc[2], eng);
        neu_add(acc[3], x0);
        neu_add(acc[4], x1);
        neu_add(acc[5], x2);
        neu_add(acc[6], x0 * x0);
        neu_add(acc[7], x1 * x1);
        neu_add(acc[8], x2 * x2);
        neu_add(acc[9], v0);
        neu_add(acc[10], v1);
        neu_add(acc[11], v2);
    }
    __shared__ double sh[MCS_BLOCK];
    double *rec = partial + (int64_t)blockIdx.x * MOM_REC;
#pragma unroll
    for (int q = 0; q < MOM_NV; ++q) {
        // tree over the block for (s, c)
        __shared__ double shc[MCS_BLOCK];
        sh[threadIdx.x] = acc[q].s;
        shc[threadIdx.x] = acc[q].c;
        __syncthreads();
        for (int o = MCS_BLOCK / 2; o > 0; o >>= 1) {
            if (threadIdx.x < o) {
                Neu a{sh[threadIdx.x], shc[threadIdx.x]};
                Neu bb{sh[threadIdx.x + o], shc[threadIdx.x + o]};
                neu_merge(a, bb);
                sh[threadIdx.x] = a.s;
                shc[threadIdx.x] = a.c;
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            rec[2 * q] = sh[0];
            rec[2 * q + 1] = shc[0];
        }
        __syncthreads();
    }
    sh[threadIdx.x] = emax;
    __syncthreads();
    for (int o = MCS_BLOCK / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) rec[2 * MOM_NV] = sh[0];
    __syncthreads();
    sh[threadIdx.x] = nact;
    __syncthreads();
    for (int o = MCS_BLOCK / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) rec[2 * MOM_NV + 1] = sh[0];
}

// Combine the block records with a fixed pairwise tree: block q of the grid owns
// quantity q (deterministic for a given grid size).
__global__ __launch_bounds__(MCS_BLOCK) void k_moments_final(const double *partial, int nb, double *out)
{
    int q = blockIdx.x;
    __shared__ double sh[MCS_BLOCK], shc[MCS_BLOCK];
    Neu a{0.0, 0.0};
    double m = -HUGE_VAL;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        const double *rec = partial + (int64_t)b * MOM_REC;
        if (q < MOM_NV) {
            Neu r{rec[2 * q], rec[2 * q + 1]};
            neu_merge(a, r);
        } else if (q == MOM_NV) {
            m = fmax(m, rec[2 * MOM_NV]);
        } else {
            a.s += rec[2 * MOM_NV + 1];
        }
    }
    sh[threadIdx.x] = q == MOM_NV ? m : a.s;
    shc[threadIdx.x] = a.c;
    __syncthreads();
    for (int o = MCS_BLOCK / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            if (q < MOM_NV) {
                Neu x{sh[threadIdx.x], shc[threadIdx.x]};
                Neu y{sh[threadIdx.x + o], shc[threadIdx.x + o]};
                neu_merge(x, y);
                sh[threadIdx.x] = x.s;
                shc[threadIdx.x] = x.c;
            } else if (q == MOM_NV) {
                sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
            } else {
                sh[threadIdx.x] += sh[threadIdx.x + o];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = q < MOM_NV ? sh[0] + shc[0] : sh[0];
}

#define CR_CHUNK 8

template <typename Buf>
__global__ __launch_bounds__(MCS_BLOCK) void k_collrate(Buf b, int64_t n, const double *energy,
                                                        const double *rate, int32_t nproc,
                                                        int32_t ne, int32_t egrid, double step,
                                                        double inv_dlog, int j0, double *partial)
{
    Neu acc[CR_CHUNK];
#pragma unroll
    for (int q = 0; q < CR_CHUNK; ++q) acc[q] = Neu{0.0, 0.0};
    int32_t flags = 0;
    for (int64_t i = gidx(); i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (!b.active[i]) continue;
        double xs[3], vs[3], wq;
        load_state(b, i, 0.0, xs, vs, wq);
        double eng = energy_dev(vs[0], vs[1], vs[2]);
        int64_t k;
        double w;
        indweight_dev(energy, ne, egrid, step, inv_dlog, eng, k, w, flags);
#pragma unroll
        for (int q = 0; q < CR_CHUNK; ++q) {
            int j = j0 + q;
            if (j < nproc)
                neu_add(acc[q], w * rate[k * nproc + j] + (1.0 - w) * rate[(k + 1) * nproc + j]);
        }
    }
    __shared__ double sh[MCS_BLOCK], shc[MCS_BLOCK];
#pragma unroll
    for (int q = 0; q < CR_CHUNK; ++q) {
        sh[threadIdx.x] = acc[q].s;
        shc[threadIdx.x] = acc[q].c;
        __syncthreads();
        for (int o = MCS_BLOCK / 2; o > 0; o >>= 1) {
            if (threadIdx.x < o) {
                Neu a{sh[threadIdx.x], shc[threadIdx.x]};
                Neu bb{sh[threadIdx.x + o], shc[threadIdx.x + o]};
                neu_merge(a, bb);
                sh[threadIdx.x] = a.s;
                shc[threadIdx.x] = a.c;
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            partial[((int64_t)blockIdx.x * CR_CHUNK + q) * 2] = sh[0];
            partial[((int64_t)blockIdx.x * CR_CHUNK + q) * 2 + 1] = shc[0];
        }
        __syncthreads();
    }
}

__global__ void k_collrate_final(const double *partial, int nb, double *out)
{
    int q = threadIdx.x;
    if (q >= CR_CHUNK) return;
    Neu a{0.0, 0.0};
    for (int b = 0; b < nb; ++b) {
        Neu r{partial[((int64_t)b * CR_CHUNK + q) * 2], partial[((int64_t)b * CR_CHUNK + q) * 2 + 1]};
        neu_merge(a, r);
    }
    out[q] = a.s + a.c;
}

// The reductions of mcs_pop_moments on a caller-chosen scratch of at least
// moments_scratch_doubles(nb) doubles: per-block partial records, then the
// final values.
static size_t moments_scratch_doubles(int nb) { return (size_t)nb * (MOM_REC + 2 * CR_CHUNK) + 64; }

static int moments_blocks(int64_t n)
{
    int64_t nb = cdiv(n, MCS_BLOCK);
    return (int)(nb > MOM_BLOCKS ? MOM_BLOCKS : nb);
}

static int32_t moments_run(mcs_pop *p, const mcs_table *t, mcs_moments *out, int nb, double *scr)
{
    mcs_ctx *ctx = p->ctx;
    double *partial = scr;
    double *res = partial + (size_t)nb * (MOM_REC + 2 * CR_CHUNK);
    if (p->prec == MCS_PREC_F32)
        hipLaunchKernelGGL(k_moments<SplitBuf>, dim3(nb), dim3(MCS_BLOCK), 0, ctx->stream, p->sbuf[0], p->n,
                           p->gh, partial);
    else
        hipLaunchKernelGGL(k_moments<PopBuf>, dim3(nb), dim3(MCS_BLOCK), 0, ctx->stream, p->buf[0], p->n,
                           0.0, partial);
    hipLaunchKernelGGL(k_moments_final, dim3(MOM_NV + 2), dim3(MCS_BLOCK), 0, ctx->stream, partial, nb, res);
    MCS_LAUNCH_CHECK(ctx);
    double h[MOM_NV + 2];
    MCS_HIP(ctx, hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    MCS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    out->weight = h[0];
    out->wenergy = h[1];
    out->sum_energy = h[2];
    for (int d = 0; d < 3; ++d) {
        out->sum_x[d] = h[3 + d];
        out->sum_x2[d] = h[6 + d];
        out->sum_v[d] = h[9 + d];
    }
    out->maxenergy = h[MOM_NV];
    out->nactive = (int64_t)h[MOM_NV + 1];
    if (p->species == MCS_SPECIES_PHOTON) {
        // photons carry no energy (photon.jl:22) and no collision table rates
        out->wenergy = out->sum_energy = out->maxenergy = 0.0;
        return MCS_OK;
    }
    if (t) {
        for (int j0 = 0; j0 < t->nproc; j0 += CR_CHUNK) {
            if (p->prec == MCS_PREC_F32)
                hipLaunchKernelGGL(k_collrate<SplitBuf>, dim3(nb), dim3(MCS_BLOCK), 0, ctx->stream, p->sbuf[0],
                                   p->n, t->energy, t->rate, t->nproc, t->ne, t->egrid, t->step,
                                   t->inv_dlog, j0, partial);
            else
                hipLaunchKernelGGL(k_collrate<PopBuf>, dim3(nb), dim3(MCS_BLOCK), 0, ctx->stream, p->buf[0],
                                   p->n, t->energy, t->rate, t->nproc, t->ne, t->egrid, t->step,
                                   t->inv_dlog, j0, partial);
            hipLaunchKernelGGL(k_collrate_final, dim3(1), dim3(64), 0, ctx->stream, partial, nb,
                               res);
            MCS_LAUNCH_CHECK(ctx);
            double cr[CR_CHUNK];
            MCS_HIP(ctx, hipMemcpyAsync(cr, res, sizeof(cr), hipMemcpyDeviceToHost, ctx->stream));
            MCS_HIP(ctx, hipStreamSynchronize(ctx->stream));
            for (int q = 0; q < CR_CHUNK && j0 + q < t->nproc; ++q) out->collrate[j0 + q] = cr[q];
        }
    }
    return MCS_OK;
}

extern "C" int32_t mcs_pop_moments(mcs_pop *p, const mcs_table *t, mcs_moments *out)
{
    if (!p || !out) return MCS_E_ARG;
    mcs_ctx *ctx = p->ctx;
    hipSetDevice(ctx->device);
    memset(out, 0, sizeof(*out));
    out->nparticles = p->n;
    if (p->n == 0) return MCS_OK;
    const int nb = moments_blocks(p->n);
    void *scr;
    int32_t rc = mcs_scratch(ctx, sizeof(double) * moments_scratch_doubles(nb), &scr);
    if (rc) return rc;
    return moments_run(p, t, out, nb, (double *)scr);
}

// ---------------------------------------------------- allocation-free moments

// Scratch of mcs_pop_moments_noalloc: room for the partial records of the
// reduction's largest grid (MOM_BLOCKS blocks) and the final values, 176.5 KiB.
// Each device holds its own copy from the library's load until the process ends;
// host code reaches it only through the address hipGetSymbolAddress gives for
// that device (noalloc_scratch_ptr), never through the name.
#define MCS_NOALLOC_DOUBLES (MOM_BLOCKS * (MOM_REC + 2 * CR_CHUNK) + 64)
static __device__ __align__(256) double scratch[MCS_NOALLOC_DOUBLES];

static std::mutex g_noalloc_mu;
static std::map<int, double *> g_noalloc_ptr;   // device -> address of its scratch

// Looks up the device address of `scratch` on the current device (once per
// device); called by mcs_ctx_create with `device` current.
static int32_t noalloc_setup(int device)
{
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_noalloc_ptr.count(device)) return MCS_OK;
    double *ptr = nullptr;
    if (hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return MCS_E_HIP;
    g_noalloc_ptr[device] = ptr;
    return MCS_OK;
}

static double *noalloc_scratch_ptr(int device)
{
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    auto it = g_noalloc_ptr.find(device);
    return it == g_noalloc_ptr.end() ? nullptr : it->second;
}

extern "C" int32_t mcs_pop_moments_noalloc(mcs_pop *p, const mcs_table *t, mcs_moments *out)
{
    if (!p || !out) return MCS_E_ARG;
    mcs_ctx *ctx = p->ctx;
    hipSetDevice(ctx->device);
    memset(out, 0, sizeof(*out));
    out->nparticles = p->n;
    if (p->n == 0) return MCS_OK;
    const int nb = moments_blocks(p->n);
    if (moments_scratch_doubles(nb) > (size_t)MCS_NOALLOC_DOUBLES)
        return mcs_fail(ctx, MCS_E_ARG, "moments_noalloc: %d blocks need %zu doubles of scratch, the maximum is %d",
                        nb, moments_scratch_doubles(nb), MCS_NOALLOC_DOUBLES);
    double *scr = noalloc_scratch_ptr(ctx->device);
    if (!scr) return mcs_fail(ctx, MCS_E_HIP, "moments_noalloc: no scratch address for device %d", ctx->device);
    return moments_run(p, t, out, nb, scr);
}
