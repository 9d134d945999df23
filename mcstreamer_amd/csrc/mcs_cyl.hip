// mcs_cyl.hip -- the axisymmetric (r, z) grid of MCStreamer: Grid and GridFields
// (MCStreamer/src/grid.jl:4-90, gridfields.jl:1-84), the cylindrical density!
// (poisson.jl:65-84) and poisson!'s charge assembly and face fields
// (poisson.jl:6-58).  FieldInterp and track1 on this grid are device functions
// of the particle step (mcs_device.h); mcs_advance_cyl is in mcs_step.hip.
#include <rocprim/device/device_radix_sort.hpp>

#include <cmath>

#include "mcs_device.h"

namespace {

// Element j (0-based) of LinRange(a, b, d + 1): Base.lerpi, (1 - t) a + t b with
// t = j / d.
double lerpi(int64_t j, int64_t d, double a, double b)
{
    const double t = (double)j / (double)d;
    return (1.0 - t) * a + t * b;
}

// Centres of LinRange(0, X, n + 1) as grid.jl:37-40 builds them: the LinRange
// 0.5 * (f[2:end] + f[1:end-1]) of n points; its first and last element.
void centres(double X, int64_t n, double &c0, double &c1)
{
    const double f1 = lerpi(0, n, 0.0, X), f2 = lerpi(1, n, 0.0, X);
    const double fn = lerpi(n - 1, n, 0.0, X), fn1 = lerpi(n, n, 0.0, X);
    c0 = 0.5 * (f2 + f1);
    c1 = 0.5 * (fn1 + fn);
}

int64_t centre_len(const mcs_cgrid *g) { return (int64_t)(g->M + 2) * (g->N + 2); }
int64_t face_len(const mcs_cgrid *g) { return (int64_t)(g->M + 3) * (g->N + 3); }

__global__ void kc_zero(double *a, int64_t n)
{
    const int64_t i = gidx();
    if (i < n) a[i] = 0.0;
}

// Key of each particle slot: its cell, or `none` when the particle is inactive or
// outside what density! adds to; the value is the slot.
__global__ __launch_bounds__(MCS_BLOCK) void kc_keys(PopBuf b, int64_t n, CylView C, uint32_t none,
                                                     uint32_t *key, uint32_t *val, int32_t *longs)
{
    const int64_t p = gidx();
    if (p == 0) longs[0] = 0;   // the long-run list, filled by kc_runs
    if (p >= n) return;
    uint32_t k = none;
    if (b.active[p]) {
        const double x[3] = {b.x[0][p], b.x[1][p], b.x[2][p]};
        int64_t i;
        const int64_t c = cyl_cell_dev(C, x, i);
        if (c >= 0) k = (uint32_t)c;
    }
    key[p] = k;
    val[p] = (uint32_t)p;
}

// Run bounds of the stably sorted (key, slot) pairs: the first and one past the
// last position of each cell's run, stored at the cell (only cells that have a
// run are written, and only those are read).
__global__ __launch_bounds__(MCS_BLOCK) void kc_bounds(const uint32_t *key, int64_t n, uint32_t none,
                                                       int32_t *first, int32_t *last)
{
    const int64_t t = gidx();
    if (t >= n) return;
    const uint32_t k = key[t];
    if (k == none) return;
    if (t == 0 || key[t - 1] != k) first[k] = (int32_t)t;
    if (t == n - 1 || key[t + 1] != k) last[k] = (int32_t)(t + 1);
}

#define CYL_LONG 256    // runs longer than this are summed by a wavefront each

// The term (w val)/dV of every sorted pair, at its sorted position (parallel; the
// runs below only add them up in order).
__global__ __launch_bounds__(MCS_BLOCK) void kc_terms(const uint32_t *key, const uint32_t *val, int64_t n,
                                                      uint32_t none, const double *w, const double *dv, int32_t M,
                                                      double v, double *term)
{
    const int64_t t = gidx();
    if (t >= n) return;
    const uint32_t k = key[t];
    if (k == none) return;
    term[t] = (w[val[t]] * v) / dv[k % (uint32_t)(M + 2) - 1];
}

// One thread per run of at most CYL_LONG: the cell's terms added to its value in
// slot order.  Longer runs are listed for kc_runs_long.
__global__ __launch_bounds__(MCS_BLOCK) void kc_runs(const uint32_t *key, int64_t n, uint32_t none,
                                                     const int32_t *first, const int32_t *last, const double *term,
                                                     double *arr, int32_t *longs)
{
    const int64_t t = gidx();
    if (t >= n) return;
    const uint32_t k = key[t];
    if (k == none || (t > 0 && key[t - 1] == k)) return;
    const int32_t a = first[k], b = last[k];
    if (b - a > CYL_LONG) {
        longs[1 + atomicAdd(&longs[0], 1)] = (int32_t)k;
        return;
    }
    double acc = arr[k];
    int32_t q = a;
    for (; q + 8 <= b; q += 8) {
        double c[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) c[u] = term[q + u];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc = acc + c[u];
    }
    for (; q < b; ++q) acc = acc + term[q];
    arr[k] = acc;
}

// Lane u's value of v, as a wave-uniform value (u uniform).
__device__ __forceinline__ double lane_value(double v, int u)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), u);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), u);
    return __hiloint2double(hi, lo);
}

// One wavefront per listed run: the run's terms come in chunks of 64, one per
// lane (three chunks in flight), and every lane adds a chunk's terms in slot
// order, reading them lane by lane, so the sum is the same sequence of additions
// as one thread's.
__global__ __launch_bounds__(MCS_BLOCK) void kc_runs_long(const int32_t *first, const int32_t *last,
                                                          const double *term, double *arr, const int32_t *longs)
{
    const int lane = threadIdx.x & (MCS_WAVE - 1);
    const int64_t wave = (int64_t)blockIdx.x * (MCS_BLOCK / MCS_WAVE) + threadIdx.x / MCS_WAVE;
    const int64_t nwaves = (int64_t)gridDim.x * (MCS_BLOCK / MCS_WAVE);
    const int32_t count = longs[0];
    for (int64_t r = wave; r < count; r += nwaves) {
        const uint32_t k = (uint32_t)longs[1 + r];
        const int32_t a = first[k], b = last[k];
        double acc = arr[k];
        double c0 = a + lane < b ? term[a + lane] : 0.0;
        double c1 = a + MCS_WAVE + lane < b ? term[a + MCS_WAVE + lane] : 0.0;
        double c2 = a + 2 * MCS_WAVE + lane < b ? term[a + 2 * MCS_WAVE + lane] : 0.0;
        for (int32_t q = a; q < b; q += MCS_WAVE) {
            const int32_t qn = q + 3 * MCS_WAVE + lane;
            const double c3 = qn < b ? term[qn] : 0.0;
            const int m = b - q < MCS_WAVE ? b - q : MCS_WAVE;
            for (int u = 0; u < m; ++u) acc = acc + lane_value(c0, u);
            c0 = c1;
            c1 = c2;
            c2 = c3;
        }
        if (lane == 0) arr[k] = acc;
    }
}

// q = Float64(Float32(e (qfixed + qpart))) on the valid cells (poisson.jl:19-33).
__global__ __launch_bounds__(MCS_BLOCK) void kc_charge(const double *qf, const double *qp, double *q, int32_t M,
                                                       int32_t N)
{
    const int64_t t = gidx();
    if (t >= (int64_t)M * N) return;
    const int64_t i = t % M + 1, j = t / M + 1, c = i + (int64_t)(M + 2) * j;
    double q0 = qf[c];
    q0 = q0 + qp[c];
    q0 = q0 * MCS_ELEMENTARY_CHARGE;
    q[c] = (double)(float)q0;
}

// er, ez on i in 1..M+1, j in 1..N+1 (poisson.jl:50-55).
__global__ __launch_bounds__(MCS_BLOCK) void kc_faces(const double *u, double *er, double *ez, int32_t M, int32_t N,
                                                      double dr, double dz, double eb)
{
    const int64_t t = gidx();
    if (t >= (int64_t)(M + 1) * (N + 1)) return;
    const int64_t i = t % (M + 1) + 1, j = t / (M + 1) + 1;
    const int64_t su = M + 2, sf = M + 3;
    er[i + sf * j] = (u[(i - 1) + su * j] - u[i + su * j]) / dr;
    ez[i + sf * j] = eb + (u[i + su * (j - 1)] - u[i + su * j]) / dz;
}

// The ghost copies of poisson.jl:55-58; the rows and columns written are not
// among those read.
__global__ void kc_face_ghosts(double *er, double *ez, int32_t M, int32_t N)
{
    const int64_t t = gidx();
    const int64_t sf = M + 3;
    if (t < N + 3) {
        ez[0 + sf * t] = ez[1 + sf * t];
        ez[(M + 1) + sf * t] = ez[M + sf * t];
    }
    if (t < M + 3) {
        er[t] = er[t + sf];
        er[t + sf * (N + 1)] = er[t + sf * N];
    }
}

// ---- the free radial boundary (MCStreamer/src/freebc.jl:4-36)

// grad = (u[ghost row, j] - u[last row, j]) / dr over the valid columns.
__global__ void kf_grad(const double *u, double *grad, int32_t M, int32_t N, double dr)
{
    const int64_t j = gidx();
    if (j >= N) return;
    const int64_t su = M + 2;
    grad[j] = (u[(M + 1) + su * (j + 1)] - u[M + su * (j + 1)]) / dr;
}

// The transforms' sums are formed in a fixed order: lane l of a wavefront adds the
// terms j = l, l + 64, ... in sequence, then the 64 partial sums meet in the xor
// butterfly (offsets 32, 16, ..., 1), which leaves the same value in every lane.
__device__ __forceinline__ double wave_tree(double acc)
{
#pragma unroll
    for (int off = MCS_WAVE / 2; off >= 1; off >>= 1) acc = acc + __shfl_xor(acc, off, MCS_WAVE);
    return acc;
}

// am = (-(dr/2) RODFT10(grad)) ./ f with RODFT10 (DST-II) Y_k = 2 sum_j X_j sin(pi (2j+1)(k+1) / 2N);
// S[m] = sin(pi m / 2N) for m in [0, 4N), the phase index reduced mod 4N.  One
// wavefront per k.
__global__ __launch_bounds__(MCS_BLOCK) void kf_dst2(const double *grad, const double *S, const double *f,
                                                     int32_t N, double dr, double *am)
{
    const int lane = threadIdx.x & (MCS_WAVE - 1);
    const int64_t k = (int64_t)blockIdx.x * (MCS_BLOCK / MCS_WAVE) + threadIdx.x / MCS_WAVE;
    if (k >= N) return;
    const int64_t four = 4 * (int64_t)N, inc = (2 * MCS_WAVE * (k + 1)) % four;
    int64_t m = ((2 * lane + 1) * (k + 1)) % four;
    double acc = 0.0;
    for (int32_t j = lane; j < N; j += MCS_WAVE) {
        acc = acc + grad[j] * S[m];
        m += inc;
        if (m >= four) m -= four;
    }
    acc = wave_tree(acc);
    if (lane == 0) {
        const double y = 2.0 * acc;
        const double bm = -(dr / 2.0) * y;
        am[k] = bm / f[k];
    }
}

// bcond = RODFT01(am) / 2 + RODFT01(am1), am1 = am's last element alone; RODFT01
// (DST-III) Y_k = (-1)^k X_{N-1} + 2 sum_{j < N-1} X_j sin(pi (j+1)(2k+1) / 2N).
__global__ __launch_bounds__(MCS_BLOCK) void kf_dst3(const double *am, const double *S, int32_t N, double *bcond)
{
    const int lane = threadIdx.x & (MCS_WAVE - 1);
    const int64_t k = (int64_t)blockIdx.x * (MCS_BLOCK / MCS_WAVE) + threadIdx.x / MCS_WAVE;
    if (k >= N) return;
    const int64_t four = 4 * (int64_t)N, inc = (MCS_WAVE * (2 * k + 1)) % four;
    int64_t m = ((lane + 1) * (2 * k + 1)) % four;
    double acc = 0.0;
    for (int32_t j = lane; j + 1 < N; j += MCS_WAVE) {
        acc = acc + am[j] * S[m];
        m += inc;
        if (m >= four) m -= four;
    }
    acc = wave_tree(acc);
    if (lane == 0) {
        const double last = am[N - 1];
        const double sg = (k & 1) ? -last : last;
        const double t = sg + 2.0 * acc;
        const double t1 = sg + 2.0 * 0.0;
        bcond[k] = t / 2.0 + t1;
    }
}

// setinhom!(mg, b, TopBnd(), bcond) (boundaries.jl:141-153): b += c bcond on the
// last valid row, c = (2/s) (1 + 1/(2M - 1)) (the cylindrical connector's factor).
__global__ void kf_inhom(double *b, const double *bcond, int32_t M, int32_t N, double c)
{
    const int64_t j = gidx();
    if (j >= N) return;
    double *p = b + M + (int64_t)(M + 2) * (j + 1);
    *p = *p + c * bcond[j];
}

// u[end, :] .+= 2 .* bcond on the ghost row (freebc.jl:35).
__global__ void kf_ghost(double *u, const double *bcond, int32_t M, int32_t N)
{
    const int64_t j = gidx();
    if (j >= N) return;
    double *p = u + (M + 1) + (int64_t)(M + 2) * (j + 1);
    *p = *p + 2.0 * bcond[j];
}

// I1/I0 + K1/K0 at x > 0 (the scaled forms of freebc.jl:17-18 have the same
// ratios): libstdc++'s cyl_bessel_i/k below x = 50, above it the large-argument
// expansions (DLMF 10.40.1-2: terms a_k(nu)/x^k, a_k = prod_{m<=k} (4 nu^2 -
// (2m-1)^2) / (8 m)), whose e^{-2x} remainders are below double precision there.
double bessel_ratio_sum(double x)
{
    if (x < 50.0)
        return std::cyl_bessel_i(1.0, x) / std::cyl_bessel_i(0.0, x) +
               std::cyl_bessel_k(1.0, x) / std::cyl_bessel_k(0.0, x);
    double t0 = 1.0, t1 = 1.0, i0 = 1.0, i1 = 1.0, k0 = 1.0, k1 = 1.0;
    for (int m = 1; m < 200; ++m) {
        const double q = (2.0 * m - 1.0) * (2.0 * m - 1.0);
        t0 *= (0.0 - q) / (8.0 * x * m);
        t1 *= (4.0 - q) / (8.0 * x * m);
        const double sg = (m & 1) ? -1.0 : 1.0;
        i0 += sg * t0;
        i1 += sg * t1;
        k0 += t0;
        k1 += t1;
        if (std::fabs(t0) < 1e-20 && std::fabs(t1) < 1e-20) break;
    }
    return i1 / i0 + k1 / k0;
}

}  // namespace

extern "C" int32_t mcs_cgrid_freebc_coeffs(const mcs_cgrid *g, double *f)
{
    if (!g || !f) return MCS_E_ARG;
    const double c = 3.141592653589793 / g->L;
    for (int32_t n = 1; n <= g->N; ++n) {
        const double k = c * (double)n;
        const double x = k * g->R;
        f[n - 1] = g->L * k * bessel_ratio_sum(x) / 2.0;
    }
    return MCS_OK;
}

static int32_t freebc_setup(mcs_cgrid *g)
{
    mcs_ctx *ctx = g->ctx;
    if (g->fb_S) return MCS_OK;
    const int64_t N = g->N;
    std::vector<double> S(4 * N), f(N);
    for (int64_t m = 0; m < 4 * N; ++m) S[m] = std::sin(3.141592653589793 * (double)m / (2.0 * (double)N));
    mcs_cgrid_freebc_coeffs(g, f.data());
    if (hipMalloc(&g->fb_S, sizeof(double) * 4 * N) != hipSuccess ||
        hipMalloc(&g->fb_f, sizeof(double) * 4 * N) != hipSuccess)
        return mcs_fail(ctx, MCS_E_HIP, "freebc: device allocation failed");
    MCS_HIP(ctx, mcs_h2d(ctx, g->fb_S, S.data(), sizeof(double) * 4 * N));
    MCS_HIP(ctx, mcs_h2d(ctx, g->fb_f, f.data(), sizeof(double) * N));
    return MCS_OK;
}

extern "C" int32_t mcs_cgrid_create(mcs_ctx *ctx, double R, double L, int32_t M, int32_t N, mcs_cgrid **out)
{
    if (!ctx || !out) return MCS_E_ARG;
    if (M < 2 || N < 2 || !(R > 0.0) || !(L > 0.0))
        return mcs_fail(ctx, MCS_E_ARG, "cgrid: M, N must be >= 2 and R, L > 0");
    if ((int64_t)(M + 3) * (N + 3) >= ((int64_t)1 << 31))
        return mcs_fail(ctx, MCS_E_ARG, "cgrid too large");
    hipSetDevice(ctx->device);
    mcs_cgrid *g = new mcs_cgrid();
    g->ctx = ctx;
    g->M = M;
    g->N = N;
    g->R = R;
    g->L = L;
    double r0, r1, z0, z1;
    centres(R, M, r0, r1);
    centres(L, N, z0, z1);
    g->dr = (r1 - r0) / (double)(M - 1);   // step(rc)
    g->dz = (z1 - z0) / (double)(N - 1);   // step(zc)
    g->drf = (R - 0.0) / (double)M;        // step(rf)
    g->dzf = (L - 0.0) / (double)N;
    std::vector<double> dv(M);
    const double two_pi = 2.0 * 3.141592653589793;
    for (int32_t i = 0; i < M; ++i) dv[i] = two_pi * lerpi(i, M - 1, r0, r1) * g->dr * g->dz;
    for (int a = 0; a < 6; ++a) {
        g->arr_len[a] = a < 4 ? centre_len(g) : face_len(g);
        if (hipMalloc(&g->arr[a], sizeof(double) * g->arr_len[a]) != hipSuccess) {
            mcs_cgrid_destroy(g);
            return mcs_fail(ctx, MCS_E_HIP, "cgrid: device allocation failed");
        }
        if (hipMemsetAsync(g->arr[a], 0, sizeof(double) * g->arr_len[a], ctx->stream) != hipSuccess) {
            mcs_cgrid_destroy(g);
            return mcs_fail(ctx, MCS_E_HIP, "cgrid: clearing the arrays failed");
        }
    }
    if (hipMalloc(&g->dv, sizeof(double) * M) != hipSuccess) {
        mcs_cgrid_destroy(g);
        return mcs_fail(ctx, MCS_E_HIP, "cgrid: device allocation failed");
    }
    MCS_HIP(ctx, hipMemcpyAsync(g->dv, dv.data(), sizeof(double) * M, hipMemcpyHostToDevice, ctx->stream));
    MCS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *out = g;
    return MCS_OK;
}

extern "C" int32_t mcs_cgrid_destroy(mcs_cgrid *g)
{
    if (!g) return MCS_OK;
    hipSetDevice(g->ctx->device);
    hipStreamSynchronize(g->ctx->stream);
    for (int a = 0; a < 6; ++a)
        if (g->arr[a]) hipFree(g->arr[a]);
    if (g->dv) hipFree(g->dv);
    if (g->run_first) hipFree(g->run_first);
    if (g->run_last) hipFree(g->run_last);
    if (g->run_long) hipFree(g->run_long);
    if (g->fb_S) hipFree(g->fb_S);
    if (g->fb_f) hipFree(g->fb_f);
    if (g->cs) hipFree(g->cs);
    if (g->cs_new) hipFree(g->cs_new);
    if (g->cnt64) hipFree(g->cnt64);
    if (g->pre64) hipFree(g->pre64);
    if (g->rs_p) hipFree(g->rs_p);
    if (g->rs_w) hipFree(g->rs_w);
    delete g;
    return MCS_OK;
}

extern "C" int32_t mcs_cgrid_resample_fields(mcs_cgrid *g, int64_t *p, double *wtotal)
{
    if (!g || !p || !wtotal) return MCS_E_ARG;
    mcs_ctx *ctx = g->ctx;
    const int64_t ncg = (int64_t)(g->M + 2) * (g->N + 2);
    hipSetDevice(ctx->device);
    if (!g->rs_p) {
        memset(p, 0, sizeof(int64_t) * ncg);
        memset(wtotal, 0, sizeof(double) * ncg);
        return MCS_OK;
    }
    MCS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    MCS_HIP(ctx, hipMemcpy(p, g->rs_p, sizeof(int64_t) * ncg, hipMemcpyDeviceToHost));
    MCS_HIP(ctx, hipMemcpy(wtotal, g->rs_w, sizeof(double) * ncg, hipMemcpyDeviceToHost));
    return MCS_OK;
}

extern "C" int32_t mcs_cgrid_steps(const mcs_cgrid *g, double steps[4])
{
    if (!g || !steps) return MCS_E_ARG;
    steps[0] = g->dr;
    steps[1] = g->dz;
    steps[2] = g->drf;
    steps[3] = g->dzf;
    return MCS_OK;
}

extern "C" int32_t mcs_cgrid_set(mcs_cgrid *g, int32_t which, const double *host)
{
    if (!g || !host || which < 0 || which > 5) return MCS_E_ARG;
    mcs_ctx *ctx = g->ctx;
    hipSetDevice(ctx->device);
    MCS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    MCS_HIP(ctx, mcs_h2d(ctx, g->arr[which], host, sizeof(double) * g->arr_len[which]));
    return MCS_OK;
}

extern "C" int32_t mcs_cgrid_get(mcs_cgrid *g, int32_t which, double *host)
{
    if (!g || !host || which < 0 || which > 5) return MCS_E_ARG;
    mcs_ctx *ctx = g->ctx;
    hipSetDevice(ctx->device);
    MCS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    MCS_HIP(ctx, hipMemcpy(host, g->arr[which], sizeof(double) * g->arr_len[which], hipMemcpyDeviceToHost));
    return MCS_OK;
}

extern "C" int32_t mcs_cgrid_zero(mcs_cgrid *g, int32_t which)
{
    if (!g || which < 0 || which > 5) return MCS_E_ARG;
    mcs_ctx *ctx = g->ctx;
    hipSetDevice(ctx->device);
    hipLaunchKernelGGL(kc_zero, dim3((unsigned)cdiv(g->arr_len[which], MCS_BLOCK)), dim3(MCS_BLOCK), 0, ctx->stream,
                       g->arr[which], g->arr_len[which]);
    MCS_LAUNCH_CHECK(ctx);
    return MCS_OK;
}

// density!: stable radix sort of (cell, slot) pairs, then one thread per cell run
// (a wavefront per run longer than CYL_LONG).
extern "C" int32_t mcs_cyl_density(mcs_pop *p, mcs_cgrid *g, int32_t which, double val)
{
    if (!p || !g || (which != MCS_CARR_QFIXED && which != MCS_CARR_QPART)) return MCS_E_ARG;
    MCS_ELECTRONS_ONLY(p, "cyl_density");
    mcs_ctx *ctx = p->ctx;
    if (g->ctx != ctx) return mcs_fail(ctx, MCS_E_ARG, "cyl_density: grid of another context");
    if (p->prec != MCS_PREC_F64) return mcs_fail(ctx, MCS_E_ARG, "cyl_density: FP64 populations only");
    hipSetDevice(ctx->device);
    const int64_t n = p->n;
    if (n == 0) return MCS_OK;
    const uint32_t none = (uint32_t)centre_len(g);
    unsigned bits = 1;
    while (bits < 32 && (none >> bits) != 0) ++bits;
    uint32_t *kin = (uint32_t *)p->i32a, *kout = (uint32_t *)p->i32b;
    uint32_t *vin = (uint32_t *)p->i32c, *vout = (uint32_t *)p->u64a;
    ProfScope ps(ctx, "cyl_density");
    const MeshView mv = cyl_view(g);
    const dim3 grid((unsigned)cdiv(n, MCS_BLOCK)), blk(MCS_BLOCK);
    if (!g->run_first) {
        const int64_t nc = centre_len(g);
        if (hipMalloc(&g->run_first, sizeof(int32_t) * nc) != hipSuccess ||
            hipMalloc(&g->run_last, sizeof(int32_t) * nc) != hipSuccess ||
            hipMalloc(&g->run_long, sizeof(int32_t) * (nc + 1)) != hipSuccess)
            return mcs_fail(ctx, MCS_E_HIP, "cyl_density: device allocation failed");
    }
    hipLaunchKernelGGL(kc_keys, grid, blk, 0, ctx->stream, p->buf[0], n, mv.C, none, kin, vin, g->run_long);
    MCS_LAUNCH_CHECK(ctx);
    size_t tb = 0;
    MCS_HIP(ctx, rocprim::radix_sort_pairs(nullptr, tb, kin, kout, vin, vout, (size_t)n, 0u, bits, ctx->stream));
    // one scratch buffer: the sort's temporary storage, then the terms
    const size_t need = tb > sizeof(double) * (size_t)n ? tb : sizeof(double) * (size_t)n;
    void *tmp = nullptr;
    int32_t rc = mcs_scratch(ctx, need, &tmp);
    if (rc) return rc;
    MCS_HIP(ctx, rocprim::radix_sort_pairs(tmp, tb, kin, kout, vin, vout, (size_t)n, 0u, bits, ctx->stream));
    double *term = (double *)tmp;
    hipLaunchKernelGGL(kc_bounds, grid, blk, 0, ctx->stream, (const uint32_t *)kout, n, none, g->run_first,
                       g->run_last);
    hipLaunchKernelGGL(kc_terms, grid, blk, 0, ctx->stream, (const uint32_t *)kout, (const uint32_t *)vout, n, none,
                       (const double *)p->buf[0].w, (const double *)g->dv, g->M, val, term);
    hipLaunchKernelGGL(kc_runs, grid, blk, 0, ctx->stream, (const uint32_t *)kout, n, none,
                       (const int32_t *)g->run_first, (const int32_t *)g->run_last, (const double *)term,
                       g->arr[which], g->run_long);
    hipLaunchKernelGGL(kc_runs_long, dim3(512), blk, 0, ctx->stream, (const int32_t *)g->run_first,
                       (const int32_t *)g->run_last, (const double *)term, g->arr[which],
                       (const int32_t *)g->run_long);
    MCS_LAUNCH_CHECK(ctx);
    return MCS_OK;
}

extern "C" int32_t mcs_cyl_efield_faces(mcs_cgrid *g, double eb)
{
    if (!g) return MCS_E_ARG;
    mcs_ctx *ctx = g->ctx;
    hipSetDevice(ctx->device);
    ProfScope ps(ctx, "cyl_faces");
    const int64_t nf = (int64_t)(g->M + 1) * (g->N + 1);
    hipLaunchKernelGGL(kc_faces, dim3((unsigned)cdiv(nf, MCS_BLOCK)), dim3(MCS_BLOCK), 0, ctx->stream,
                       (const double *)g->arr[MCS_CARR_U], g->arr[MCS_CARR_ER], g->arr[MCS_CARR_EZ], g->M, g->N,
                       g->dr, g->dz, eb);
    const int64_t ng = (g->M > g->N ? g->M : g->N) + 3;
    hipLaunchKernelGGL(kc_face_ghosts, dim3((unsigned)cdiv(ng, MCS_BLOCK)), dim3(MCS_BLOCK), 0, ctx->stream,
                       g->arr[MCS_CARR_ER], g->arr[MCS_CARR_EZ], g->M, g->N);
    MCS_LAUNCH_CHECK(ctx);
    return MCS_OK;
}

extern "C" int32_t mcs_cyl_poisson(mcs_cgrid *g, mcs_mg *mg, double eb, int32_t freebc, int32_t *iters,
                                   double *resid)
{
    if (!g || !mg) return MCS_E_ARG;
    mcs_ctx *ctx = g->ctx;
    if (mg->ctx != ctx) return mcs_fail(ctx, MCS_E_ARG, "cyl_poisson: solver of another context");
    if (mg->cfg.ndim != 2 || mg->cfg.n[0] != g->M || mg->cfg.n[1] != g->N || mg->cfg.cyl != 1 ||
        mg->cfg.precision != MCS_PREC_F64)
        return mcs_fail(ctx, MCS_E_ARG, "cyl_poisson: the solver must be FP64, ndim 2, n = (M, N), cyl 1");
    if (freebc && mg->cfg.bc[1] != MCS_BC_DIRICHLET)
        return mcs_fail(ctx, MCS_E_ARG, "cyl_poisson: the free boundary needs Dirichlet at r = R");
    hipSetDevice(ctx->device);
    if (freebc) {
        int32_t rs = freebc_setup(g);
        if (rs) return rs;
    }
    const int64_t nc = (int64_t)g->M * g->N;
    double *u = g->arr[MCS_CARR_U], *q = g->arr[MCS_CARR_Q];
    hipLaunchKernelGGL(kc_charge, dim3((unsigned)cdiv(nc, MCS_BLOCK)), dim3(MCS_BLOCK), 0, ctx->stream,
                       (const double *)g->arr[MCS_CARR_QFIXED], (const double *)g->arr[MCS_CARR_QPART], q, g->M,
                       g->N);
    MCS_LAUNCH_CHECK(ctx);
    int32_t it1 = 0, it2 = 0;
    int32_t rc = mcs_mg2_solve_dev(mg, u, q, &it1, resid);
    if (rc != MCS_OK && rc != MCS_E_NONCONVERGED) return rc;
    if (freebc) {
        // solve_freebc! (freebc.jl:4-36): the homogeneous solve's radial gradient at
        // r = R, its sine transform over the Bessel factors back into a boundary
        // value, that value into b, a second solve, and the value into the ghost row
        ProfScope ps(ctx, "cyl_freebc");
        const int32_t M = g->M, N = g->N;
        const dim3 gb((unsigned)cdiv(N, MCS_BLOCK)), bl(MCS_BLOCK);
        double *grad = g->fb_f + N, *am = g->fb_f + 2 * N, *bcond = g->fb_f + 3 * N;
        hipLaunchKernelGGL(kf_grad, gb, bl, 0, ctx->stream, (const double *)u, grad, M, N, g->dr);
        const dim3 gw((unsigned)cdiv(N, MCS_BLOCK / MCS_WAVE));   // a wavefront per output
        hipLaunchKernelGGL(kf_dst2, gw, bl, 0, ctx->stream, (const double *)grad, (const double *)g->fb_S,
                           (const double *)g->fb_f, N, g->dr, am);
        hipLaunchKernelGGL(kf_dst3, gw, bl, 0, ctx->stream, (const double *)am, (const double *)g->fb_S, N, bcond);
        const double c = (2.0 / mg->cfg.s) * (1.0 + 1.0 / (double)(2 * M - 1));
        hipLaunchKernelGGL(kf_inhom, gb, bl, 0, ctx->stream, q, (const double *)bcond, M, N, c);
        MCS_LAUNCH_CHECK(ctx);
        int32_t rc2 = mcs_mg2_solve_dev(mg, u, q, &it2, resid);
        if (rc2 != MCS_OK && rc2 != MCS_E_NONCONVERGED) return rc2;
        if (rc2) rc = rc2;
        hipLaunchKernelGGL(kf_ghost, gb, bl, 0, ctx->stream, u, (const double *)bcond, M, N);
        MCS_LAUNCH_CHECK(ctx);
    }
    if (iters) *iters = it1 + it2;
    int32_t rf = mcs_cyl_efield_faces(g, eb);
    return rf ? rf : rc;
}
