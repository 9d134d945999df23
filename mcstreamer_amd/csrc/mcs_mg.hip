// mcs_mg.hip -- geometric multigrid for L u + s b = 0 on a cell-centred 3D
// mesh with one ghost layer (Multigrid/src/Multigrid.jl, redblack.jl,
// boundaries.jl).  The host drives the same sequence of operations as
// fmg!/mgv!; each operation is a coalesced stencil kernel over one level.
#include <cmath>
#include <cstring>

#include "mcs_device.h"

struct Lvl {
    int32_t n[3];
};

__device__ __forceinline__ int64_t g_index(const Lvl &L, int64_t i, int64_t j, int64_t k)
{
    return i + (int64_t)(L.n[0] + 2) * (j + (int64_t)(L.n[1] + 2) * k);
}

// setbc! for both faces of one dimension (boundaries.jl:171-178): the ghost
// plane, over the full extent of the other two dimensions, is coef times the
// adjacent valid plane.
__global__ __launch_bounds__(MCS_BLOCK) void k_mg_bc(double *a, Lvl L, int d, double clo, double chi)
{
    int64_t ext[3] = {L.n[0] + 2, L.n[1] + 2, L.n[2] + 2};
    int64_t str[3] = {1, ext[0], ext[0] * ext[1]};
    int d1 = (d + 1) % 3, d2 = (d + 2) % 3;
    int64_t t = gidx();
    if (t >= ext[d1] * ext[d2]) return;
    int64_t b1 = t % ext[d1], b2 = t / ext[d1];
    int64_t base = b1 * str[d1] + b2 * str[d2];
    a[base] = clo * a[base + str[d]];
    a[base + (L.n[d] + 1) * str[d]] = chi * a[base + L.n[d] * str[d]];
}

// One colour of gauss_seidel! (Multigrid.jl:168-178 via redblack3): cells with
// (i+j+k) % 2 == par in interior coordinates; each row holds nx/2 of them.
__global__ __launch_bounds__(MCS_BLOCK) void k_mg_rbgs(double *u, const double *b, Lvl L, int par)
{
    int64_t hx = L.n[0] / 2;
    int64_t t = gidx();
    if (t >= hx * L.n[1] * L.n[2]) return;
    int64_t ti = t % hx, j = (t / hx) % L.n[1], k = t / (hx * L.n[1]);
    int64_t i = 2 * ti + ((j + k + par) & 1);
    const int64_t sy = L.n[0] + 2, sz = (int64_t)(L.n[0] + 2) * (L.n[1] + 2);
    int64_t c = (i + 1) + sy * (j + 1) + sz * (k + 1);
    double l = -6.0 * u[c];
    l += u[c + 1];
    l += u[c + sy];
    l += u[c + sz];
    l += u[c - 1];
    l += u[c - sy];
    l += u[c - sz];
    u[c] += (1.0 * (l + b[c])) / 6.0;
}

// residual! (Multigrid.jl:186-193): r = s b + L u; optional per-block sum of r^2.
__global__ __launch_bounds__(MCS_BLOCK) void k_mg_residual(double *r, const double *u, const double *b,
                                                           double s, Lvl L, double *partial)
{
    int64_t t = gidx();
    int64_t N = (int64_t)L.n[0] * L.n[1] * L.n[2];
    double sq = 0.0;
    if (t < N) {
        int64_t i = t % L.n[0], j = (t / L.n[0]) % L.n[1], k = t / ((int64_t)L.n[0] * L.n[1]);
        const int64_t sy = L.n[0] + 2, sz = (int64_t)(L.n[0] + 2) * (L.n[1] + 2);
        int64_t c = (i + 1) + sy * (j + 1) + sz * (k + 1);
        double l = -6.0 * u[c];
        l += u[c + 1];
        l += u[c + sy];
        l += u[c + sz];
        l += u[c - 1];
        l += u[c - sy];
        l += u[c - sz];
        double rv = s * b[c] + l;
        r[c] = rv;
        sq = rv * rv;
    }
    if (!partial) return;
    __shared__ double sh[MCS_BLOCK];
    sh[threadIdx.x] = sq;
    __syncthreads();
    for (int o = MCS_BLOCK / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(MCS_BLOCK) void k_mg_sumsq(const double *b, Lvl L, double *partial)
{
    int64_t t = gidx();
    int64_t N = (int64_t)L.n[0] * L.n[1] * L.n[2];
    double sq = 0.0;
    if (t < N) {
        int64_t i = t % L.n[0], j = (t / L.n[0]) % L.n[1], k = t / ((int64_t)L.n[0] * L.n[1]);
        double v = b[g_index(L, i + 1, j + 1, k + 1)];
        sq = v * v;
    }
    __shared__ double sh[MCS_BLOCK];
    sh[threadIdx.x] = sq;
    __syncthreads();
    for (int o = MCS_BLOCK / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ void k_sum_partials(const double *partial, int64_t nb, double *out)
{
    __shared__ double sh[MCS_BLOCK];
    double acc = 0.0;
    for (int64_t q = threadIdx.x; q < nb; q += blockDim.x) acc += partial[q];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int o = MCS_BLOCK / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// restrict! (Multigrid.jl:213-226): 4*(sum of 8 children)/8, x offset slowest.
__global__ __launch_bounds__(MCS_BLOCK) void k_mg_restrict(double *rh, const double *r, Lvl F)
{
    int32_t nc0 = F.n[0] / 2, nc1 = F.n[1] / 2, nc2 = F.n[2] / 2;
    int64_t t = gidx();
    if (t >= (int64_t)nc0 * nc1 * nc2) return;
    int64_t i = t % nc0, j = (t / nc0) % nc1, k = t / ((int64_t)nc0 * nc1);
    const int64_t fy = F.n[0] + 2, fz = (int64_t)(F.n[0] + 2) * (F.n[1] + 2);
    const int64_t cy = nc0 + 2, cz = (int64_t)(nc0 + 2) * (nc1 + 2);
    int64_t f = (2 * i + 1) + fy * (2 * j + 1) + fz * (2 * k + 1);
    double s = 0.0;
    s += r[f];
    s += r[f + fz];
    s += r[f + fy];
    s += r[f + fy + fz];
    s += r[f + 1];
    s += r[f + 1 + fz];
    s += r[f + 1 + fy];
    s += r[f + 1 + fy + fz];
    rh[(i + 1) + cy * (j + 1) + cz * (k + 1)] = (4.0 * s) / 8.0;
}

// interpolate! (Multigrid.jl:232-258): trilinear weights 27/64 ... 1/64 over the
// coarse neighbours in the direction of each fine child; add or overwrite.
__global__ __launch_bounds__(MCS_BLOCK) void k_mg_interp(double *r, const double *rh, Lvl F, int add)
{
    int32_t nc0 = F.n[0] / 2, nc1 = F.n[1] / 2, nc2 = F.n[2] / 2;
    int64_t t = gidx();
    if (t >= (int64_t)nc0 * nc1 * nc2) return;
    int64_t i = t % nc0, j = (t / nc0) % nc1, k = t / ((int64_t)nc0 * nc1);
    const int64_t fy = F.n[0] + 2, fz = (int64_t)(F.n[0] + 2) * (F.n[1] + 2);
    const int64_t cy = nc0 + 2, cz = (int64_t)(nc0 + 2) * (nc1 + 2);
    int64_t cc = (i + 1) + cy * (j + 1) + cz * (k + 1);
    const double wt[8] = {27.0 / 64.0, 9.0 / 64.0, 9.0 / 64.0, 3.0 / 64.0,
                          9.0 / 64.0,  3.0 / 64.0, 3.0 / 64.0, 1.0 / 64.0};
#pragma unroll
    for (int sx = 0; sx < 2; ++sx)
#pragma unroll
        for (int sy = 0; sy < 2; ++sy)
#pragma unroll
            for (int sz = 0; sz < 2; ++sz) {
                int64_t fidx = (2 * i + 1 + sx) + fy * (2 * j + 1 + sy) + fz * (2 * k + 1 + sz);
                int64_t ox = 2 * sx - 1, oy = (2 * sy - 1) * cy, oz = (2 * sz - 1) * cz;
                double v = 0.0;
                v += wt[0] * rh[cc];
                v += wt[1] * rh[cc + oz];
                v += wt[2] * rh[cc + oy];
                v += wt[3] * rh[cc + oy + oz];
                v += wt[4] * rh[cc + ox];
                v += wt[5] * rh[cc + ox + oz];
                v += wt[6] * rh[cc + ox + oy];
                v += wt[7] * rh[cc + ox + oy + oz];
                if (add)
                    r[fidx] += v;
                else
                    r[fidx] = v;
            }
}

// Coarse direct solve of fmg! (Multigrid.jl:404-410): z = A^-1 btop, one row per
// thread, columns summed in order.
__global__ __launch_bounds__(MCS_BLOCK) void k_mg_coarse(const double *ainv, const double *res,
                                                         double *z, Lvl C)
{
    int64_t N = (int64_t)C.n[0] * C.n[1] * C.n[2];
    for (int64_t row = threadIdx.x; row < N; row += blockDim.x) {
        double acc = 0.0;
        for (int64_t col = 0; col < N; ++col) {
            int64_t i = col % C.n[0], j = (col / C.n[0]) % C.n[1], k = col / ((int64_t)C.n[0] * C.n[1]);
            acc += ainv[row * N + col] * res[g_index(C, i + 1, j + 1, k + 1)];
        }
        int64_t i = row % C.n[0], j = (row / C.n[0]) % C.n[1], k = row / ((int64_t)C.n[0] * C.n[1]);
        z[g_index(C, i + 1, j + 1, k + 1)] = acc;
    }
}

// The 50 (apply!, gauss_seidel!) rounds at the coarsest level of mgv!
// (Multigrid.jl:340-346) in one workgroup, phases separated by barriers.
__global__ __launch_bounds__(1024) void k_mg_coarsest(double *x, const double *b, Lvl L,
                                                      const int32_t *bc_coef, int rounds)
{
    int64_t ext[3] = {L.n[0] + 2, L.n[1] + 2, L.n[2] + 2};
    int64_t str[3] = {1, ext[0], ext[0] * ext[1]};
    const int64_t N = (int64_t)L.n[0] * L.n[1] * L.n[2];
    for (int it = 0; it < rounds; ++it) {
        for (int d = 0; d < 3; ++d) {
            int d1 = (d + 1) % 3, d2 = (d + 2) % 3;
            double clo = bc_coef[2 * d] == MCS_BC_DIRICHLET ? -1.0 : 1.0;
            double chi = bc_coef[2 * d + 1] == MCS_BC_DIRICHLET ? -1.0 : 1.0;
            for (int64_t t = threadIdx.x; t < ext[d1] * ext[d2]; t += blockDim.x) {
                int64_t b1 = t % ext[d1], b2 = t / ext[d1];
                int64_t base = b1 * str[d1] + b2 * str[d2];
                x[base] = clo * x[base + str[d]];
                x[base + (L.n[d] + 1) * str[d]] = chi * x[base + L.n[d] * str[d]];
            }
            __syncthreads();
        }
        for (int par = 0; par < 2; ++par) {
            for (int64_t t = threadIdx.x; t < N; t += blockDim.x) {
                int64_t i = t % L.n[0], j = (t / L.n[0]) % L.n[1], k = t / ((int64_t)L.n[0] * L.n[1]);
                if (((i + j + k) & 1) != par) continue;
                int64_t c = (i + 1) + str[1] * (j + 1) + str[2] * (k + 1);
                double l = -6.0 * x[c];
                l += x[c + 1];
                l += x[c + str[1]];
                l += x[c + str[2]];
                l += x[c - 1];
                l += x[c - str[1]];
                l += x[c - str[2]];
                x[c] += (1.0 * (l + b[c])) / 6.0;
            }
            __syncthreads();
        }
    }
}

__global__ void k_mg_fill0(double *a, int64_t n)
{
    int64_t i = gidx();
    if (i < n) a[i] = 0.0;
}

__global__ void k_mg_axpy(double *x, const double *z, int64_t n)
{
    int64_t i = gidx();
    if (i < n) x[i] += z[i];
}

// setinhom! (boundaries.jl:202-214): b += (2/s) val on the first valid layer.
__global__ void k_mg_inhom(double *b, Lvl L, int face, double c, const double *val)
{
    int d = face / 2, hi = face & 1;
    int d1 = (d + 1) % 3, d2 = (d + 2) % 3;
    int64_t n1 = L.n[d1], n2 = L.n[d2];
    int64_t t = gidx();
    if (t >= n1 * n2) return;
    int64_t a1 = t % n1, a2 = t / n1;
    int64_t idx[3];
    idx[d] = hi ? L.n[d] : 1;
    idx[d1] = a1 + 1;
    idx[d2] = a2 + 1;
    int64_t cidx = g_index(L, idx[0], idx[1], idx[2]);
    b[cidx] = b[cidx] + c * val[t];
}

// ---------------------------------------------------------------- host side

static int64_t glen(const int32_t n[3])
{
    return (int64_t)(n[0] + 2) * (n[1] + 2) * (n[2] + 2);
}

// Gauss-Jordan inverse of the coarsest operator with the ghost rule folded in
// (buildmat, Multigrid.jl:261-295).  Same elimination order as the CPU restatement.
static int coarse_inverse(const int32_t n[3], const int32_t bc[6], std::vector<double> &ainv)
{
    int64_t N = (int64_t)n[0] * n[1] * n[2];
    std::vector<double> a((size_t)(N * N), 0.0);
    ainv.assign((size_t)(N * N), 0.0);
    for (int64_t k = 0; k < n[2]; ++k)
        for (int64_t j = 0; j < n[1]; ++j)
            for (int64_t i = 0; i < n[0]; ++i) {
                int64_t row = i + n[0] * (j + n[1] * k);
                int64_t c[3] = {i, j, k};
                a[row * N + row] -= 6.0;
                for (int q = 0; q < 6; ++q) {
                    int d = q % 3;
                    int sgn = q < 3 ? 1 : -1;
                    int64_t cc[3] = {c[0], c[1], c[2]};
                    cc[d] += sgn;
                    if (cc[d] < 0 || cc[d] >= n[d]) {
                        int face = 2 * d + (sgn > 0 ? 1 : 0);
                        a[row * N + row] += bc[face] == MCS_BC_DIRICHLET ? -1.0 : 1.0;
                    } else {
                        int64_t col = cc[0] + n[0] * (cc[1] + n[1] * cc[2]);
                        a[row * N + col] += 1.0;
                    }
                }
            }
    for (int64_t r = 0; r < N; ++r) ainv[r * N + r] = 1.0;
    for (int64_t col = 0; col < N; ++col) {
        int64_t piv = col;
        double best = fabs(a[col * N + col]);
        for (int64_t r = col + 1; r < N; ++r)
            if (fabs(a[r * N + col]) > best) {
                best = fabs(a[r * N + col]);
                piv = r;
            }
        if (best == 0.0) return 1;
        if (piv != col) {
            for (int64_t c = 0; c < N; ++c) {
                std::swap(a[col * N + c], a[piv * N + c]);
                std::swap(ainv[col * N + c], ainv[piv * N + c]);
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
    return 0;
}

extern "C" int32_t mcs_mg_create(mcs_ctx *ctx, const mcs_mg_config *cfg, mcs_mg **out)
{
    if (!ctx || !cfg || !out) return MCS_E_ARG;
    const mcs_mg_config &c = *cfg;
    if (c.levels < 0 || c.levels > 12 || c.smooth1 < 0 || c.smooth2 < 0 || c.maxiter < 0)
        return mcs_fail(ctx, MCS_E_ARG, "mg: bad levels/smoothing/maxiter");
    for (int d = 0; d < 3; ++d) {
        if (c.n[d] < 1 || c.n[d] % (1 << c.levels))
            return mcs_fail(ctx, MCS_E_ARG, "mg: n[%d]=%d is not divisible by 2^levels", d, c.n[d]);
    }
    if ((c.n[0] >> c.levels) % 2)
        return mcs_fail(ctx, MCS_E_ARG, "mg: red-black needs an even x size on every level (redblack.jl:13)");
    for (int f = 0; f < 6; ++f)
        if (c.bc[f] != MCS_BC_DIRICHLET && c.bc[f] != MCS_BC_NEUMANN)
            return mcs_fail(ctx, MCS_E_ARG, "mg: unknown boundary condition on face %d", f);
    if (c.precision != MCS_PREC_F64)
        return mcs_fail(ctx, MCS_E_ARG, "mg: only MCS_PREC_F64 is implemented");
    hipSetDevice(ctx->device);
    mcs_mg *mg = new mcs_mg();
    mg->ctx = ctx;
    mg->cfg = c;
    for (int l = 0; l <= c.levels; ++l) {
        MgLevel L;
        for (int d = 0; d < 3; ++d) L.n[d] = c.n[d] >> l;
        L.len = glen(L.n);
        L.fres = nullptr;
        L.fsol = nullptr;
        if (hipMalloc(&L.res, sizeof(double) * L.len) != hipSuccess ||
            hipMalloc(&L.res1, sizeof(double) * L.len) != hipSuccess ||
            hipMalloc(&L.sol, sizeof(double) * L.len) != hipSuccess)
            return mcs_fail(ctx, MCS_E_HIP, "mg: device allocation failed");
        hipMemset(L.res, 0, sizeof(double) * L.len);
        hipMemset(L.res1, 0, sizeof(double) * L.len);
        hipMemset(L.sol, 0, sizeof(double) * L.len);
        mg->lv.push_back(L);
    }
    const int32_t *nc = mg->lv.back().n;
    mg->ncoarse = (int64_t)nc[0] * nc[1] * nc[2];
    if (mg->ncoarse > 4096)
        return mcs_fail(ctx, MCS_E_ARG, "mg: coarsest level has %lld cells; use more levels",
                        (long long)mg->ncoarse);
    std::vector<double> ainv;
    if (coarse_inverse(nc, c.bc, ainv))
        return mcs_fail(ctx, MCS_E_ARG, "mg: coarsest operator is singular (all-Neumann?)");
    if (hipMalloc(&mg->ainv, sizeof(double) * ainv.size()) != hipSuccess)
        return mcs_fail(ctx, MCS_E_HIP, "mg: device allocation failed");
    hipMemcpy(mg->ainv, ainv.data(), sizeof(double) * ainv.size(), hipMemcpyHostToDevice);
    int64_t N0 = (int64_t)c.n[0] * c.n[1] * c.n[2];
    mg->npartial = cdiv(N0, MCS_BLOCK);
    if (hipMalloc(&mg->partial, sizeof(double) * (mg->npartial + 8)) != hipSuccess)
        return mcs_fail(ctx, MCS_E_HIP, "mg: device allocation failed");
    *out = mg;
    return MCS_OK;
}

extern "C" int32_t mcs_mg_destroy(mcs_mg *mg)
{
    if (!mg) return MCS_OK;
    hipSetDevice(mg->ctx->device);
    hipStreamSynchronize(mg->ctx->stream);
    for (auto &L : mg->lv) {
        hipFree(L.res);
        hipFree(L.res1);
        hipFree(L.sol);
    }
    hipFree(mg->ainv);
    hipFree(mg->partial);
    if (mg->u_tmp) hipFree(mg->u_tmp);
    if (mg->b_tmp) hipFree(mg->b_tmp);
    if (mg->bc_dev) hipFree(mg->bc_dev);
    delete mg;
    return MCS_OK;
}

static inline Lvl lvl(const mcs_mg *mg, int l)
{
    Lvl L;
    for (int d = 0; d < 3; ++d) L.n[d] = mg->lv[l].n[d];
    return L;
}

static void mg_bc(mcs_mg *mg, int l, double *a)
{
    Lvl L = lvl(mg, l);
    hipStream_t st = mg->ctx->stream;
    for (int d = 0; d < 3; ++d) {
        int d1 = (d + 1) % 3, d2 = (d + 2) % 3;
        int64_t cnt = (int64_t)(L.n[d1] + 2) * (L.n[d2] + 2);
        double clo = mg->cfg.bc[2 * d] == MCS_BC_DIRICHLET ? -1.0 : 1.0;
        double chi = mg->cfg.bc[2 * d + 1] == MCS_BC_DIRICHLET ? -1.0 : 1.0;
        hipLaunchKernelGGL(k_mg_bc, dim3((unsigned)cdiv(cnt, MCS_BLOCK)), dim3(MCS_BLOCK), 0, st, a, L,
                           d, clo, chi);
    }
}

static void mg_gs(mcs_mg *mg, int l, double *u, const double *b)
{
    Lvl L = lvl(mg, l);
    int64_t cnt = (int64_t)(L.n[0] / 2) * L.n[1] * L.n[2];
    ProfScope ps(mg->ctx, "mg_rbgs");
    for (int par = 0; par < 2; ++par)
        hipLaunchKernelGGL(k_mg_rbgs, dim3((unsigned)cdiv(cnt, MCS_BLOCK)), dim3(MCS_BLOCK), 0,
                           mg->ctx->stream, u, b, L, par);
}

static void mg_residual(mcs_mg *mg, int l, double *r, const double *u, const double *b, double s,
                        double *partial)
{
    Lvl L = lvl(mg, l);
    int64_t N = (int64_t)L.n[0] * L.n[1] * L.n[2];
    ProfScope ps(mg->ctx, "mg_residual");
    hipLaunchKernelGGL(k_mg_residual, dim3((unsigned)cdiv(N, MCS_BLOCK)), dim3(MCS_BLOCK), 0,
                       mg->ctx->stream, r, u, b, s, L, partial);
}

static void mg_restrict(mcs_mg *mg, int lf, double *rh, const double *r)
{
    Lvl F = lvl(mg, lf);
    int64_t Nc = (int64_t)(F.n[0] / 2) * (F.n[1] / 2) * (F.n[2] / 2);
    ProfScope ps(mg->ctx, "mg_restrict");
    hipLaunchKernelGGL(k_mg_restrict, dim3((unsigned)cdiv(Nc, MCS_BLOCK)), dim3(MCS_BLOCK), 0,
                       mg->ctx->stream, rh, r, F);
}

static void mg_interp(mcs_mg *mg, int lf, double *r, const double *rh, int add)
{
    Lvl F = lvl(mg, lf);
    int64_t Nc = (int64_t)(F.n[0] / 2) * (F.n[1] / 2) * (F.n[2] / 2);
    ProfScope ps(mg->ctx, "mg_prolong");
    hipLaunchKernelGGL(k_mg_interp, dim3((unsigned)cdiv(Nc, MCS_BLOCK)), dim3(MCS_BLOCK), 0,
                       mg->ctx->stream, r, rh, F, add);
}

static void mg_fill0(mcs_mg *mg, double *a, int64_t n)
{
    hipLaunchKernelGGL(k_mg_fill0, dim3((unsigned)cdiv(n, MCS_BLOCK)), dim3(MCS_BLOCK), 0,
                       mg->ctx->stream, a, n);
}

static int32_t ensure_bc_dev(mcs_mg *mg)
{
    if (mg->bc_dev) return MCS_OK;
    MCS_HIP(mg->ctx, hipMalloc(&mg->bc_dev, sizeof(int32_t) * 6));
    MCS_HIP(mg->ctx, hipMemcpy(mg->bc_dev, mg->cfg.bc, sizeof(int32_t) * 6, hipMemcpyHostToDevice));
    return MCS_OK;
}

// mgv! (Multigrid.jl:334-377)
static void mgv(mcs_mg *mg, double *x, const double *b, int level)
{
    const int L = mg->cfg.levels;
    if (level == L) {
        Lvl C = lvl(mg, level);
        if (mg->lv[level].len <= 65536) {
            hipLaunchKernelGGL(k_mg_coarsest, dim3(1), dim3(1024), 0, mg->ctx->stream, x, b, C,
                               (const int32_t *)mg->bc_dev, 50);
        } else {
            for (int it = 0; it < 50; ++it) {
                mg_bc(mg, level, x);
                mg_gs(mg, level, x, b);
            }
        }
        return;
    }
    for (int it = 0; it < mg->cfg.smooth1; ++it) {
        mg_bc(mg, level, x);
        mg_gs(mg, level, x, b);
    }
    double *r = mg->lv[level].res1;
    mg_bc(mg, level, x);
    mg_residual(mg, level, r, x, b, 1.0, nullptr);
    double *rh = mg->lv[level + 1].res;
    mg_bc(mg, level, r);
    mg_restrict(mg, level, rh, r);
    double *xh = mg->lv[level + 1].sol;
    mg_fill0(mg, xh, mg->lv[level + 1].len);
    mgv(mg, xh, rh, level + 1);
    mg_interp(mg, level, x, xh, 1);
    for (int it = 0; it < mg->cfg.smooth2; ++it) {
        mg_bc(mg, level, x);
        mg_gs(mg, level, x, b);
    }
}

// fmg! (Multigrid.jl:380-434).  Returns 1 when a correction was applied, 0 when
// the residual already meets the tolerance, negative on error.
static int fmg(mcs_mg *mg, double *x, const double *b, double *eps_out)
{
    mcs_ctx *ctx = mg->ctx;
    const int L = mg->cfg.levels;
    const MgLevel &F = mg->lv[0];
    mg_bc(mg, 0, x);
    mg_fill0(mg, F.res, F.len);
    mg_residual(mg, 0, F.res, x, b, mg->cfg.s, mg->partial);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(MCS_BLOCK), 0, ctx->stream,
                       (const double *)mg->partial, mg->npartial, mg->partial + mg->npartial);
    double ss = 0.0;
    if (hipMemcpyAsync(ctx->h_small, mg->partial + mg->npartial, sizeof(double),
                       hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        return -1;
    memcpy(&ss, ctx->h_small, sizeof(double));
    int64_t N = (int64_t)F.n[0] * F.n[1] * F.n[2];
    double eps = sqrt(ss) / sqrt((double)F.n[0] * (double)F.n[1] * (double)F.n[2]);
    (void)N;
    *eps_out = eps;
    if (mg->cfg.s * mg->tol_abs > eps) return 0;
    for (int k = 0; k < L; ++k) mg_restrict(mg, k, mg->lv[k + 1].res, mg->lv[k].res);
    {
        Lvl C = lvl(mg, L);
        hipLaunchKernelGGL(k_mg_coarse, dim3(1), dim3(MCS_BLOCK), 0, ctx->stream,
                           (const double *)mg->ainv, (const double *)mg->lv[L].res, mg->lv[L].sol, C);
    }
    double *z = mg->lv[L].sol;
    for (int k = 1; k <= L; ++k) {
        int lev = L - k;
        double *z1 = mg->lv[lev].sol;
        mg_fill0(mg, z1, mg->lv[lev].len);
        mg_bc(mg, lev + 1, z);
        mg_interp(mg, lev, z1, z, 0);
        mgv(mg, z1, mg->lv[lev].res, lev);
        z = z1;
    }
    hipLaunchKernelGGL(k_mg_axpy, dim3((unsigned)cdiv(F.len, MCS_BLOCK)), dim3(MCS_BLOCK), 0,
                       ctx->stream, x, (const double *)z, F.len);
    mg_bc(mg, 0, x);
    return 1;
}

// solve (Multigrid.jl:449-467)
static int32_t mg_solve_dev(mcs_mg *mg, double *u, const double *b, int32_t *iters, double *resid)
{
    mcs_ctx *ctx = mg->ctx;
    int32_t rc = ensure_bc_dev(mg);
    if (rc) return rc;
    mg->tol_abs = mg->cfg.tolerance;
    if (mg->cfg.tolerance < 0.0) {
        // relative mode: -tolerance times the rms of b over the interior
        Lvl F = lvl(mg, 0);
        int64_t N = (int64_t)F.n[0] * F.n[1] * F.n[2];
        hipLaunchKernelGGL(k_mg_sumsq, dim3((unsigned)cdiv(N, MCS_BLOCK)), dim3(MCS_BLOCK), 0,
                           ctx->stream, b, F, mg->partial);
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(MCS_BLOCK), 0, ctx->stream,
                           (const double *)mg->partial, mg->npartial, mg->partial + mg->npartial);
        double ss = 0.0;
        MCS_HIP(ctx, hipMemcpyAsync(ctx->h_small, mg->partial + mg->npartial, sizeof(double),
                                    hipMemcpyDeviceToHost, ctx->stream));
        MCS_HIP(ctx, hipStreamSynchronize(ctx->stream));
        memcpy(&ss, ctx->h_small, sizeof(double));
        double brms = sqrt(ss) / sqrt((double)F.n[0] * (double)F.n[1] * (double)F.n[2]);
        mg->tol_abs = -mg->cfg.tolerance * brms;
        if (!(mg->tol_abs > 0.0)) mg->tol_abs = 2.2250738585072014e-308;
    }
    int iter = 0;
    int cont = 1;
    double eps = 0.0;
    while (cont && iter <= mg->cfg.maxiter) {
        cont = fmg(mg, u, b, &eps);
        if (cont < 0) return mcs_fail(ctx, MCS_E_HIP, "mg: device error during the cycle");
        iter += 1;
    }
    MCS_LAUNCH_CHECK(ctx);
    if (iters) *iters = iter;
    if (resid) *resid = eps;
    if (iter > mg->cfg.maxiter)
        return mcs_fail(ctx, MCS_E_NONCONVERGED, "Convergence failed with [maxiter=] %d iterations",
                        mg->cfg.maxiter);
    return MCS_OK;
}

extern "C" int32_t mcs_mg_solve_host(mcs_mg *mg, double *u, const double *b, int32_t *iters,
                                     double *resid)
{
    if (!mg || !u || !b) return MCS_E_ARG;
    mcs_ctx *ctx = mg->ctx;
    hipSetDevice(ctx->device);
    int64_t len = mg->lv[0].len;
    if (!mg->u_tmp) {
        MCS_HIP(ctx, hipMalloc(&mg->u_tmp, sizeof(double) * len));
        MCS_HIP(ctx, hipMalloc(&mg->b_tmp, sizeof(double) * len));
    }
    MCS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    MCS_HIP(ctx, hipMemcpy(mg->u_tmp, u, sizeof(double) * len, hipMemcpyHostToDevice));
    MCS_HIP(ctx, hipMemcpy(mg->b_tmp, b, sizeof(double) * len, hipMemcpyHostToDevice));
    int32_t rc = mg_solve_dev(mg, mg->u_tmp, mg->b_tmp, iters, resid);
    MCS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    MCS_HIP(ctx, hipMemcpy(u, mg->u_tmp, sizeof(double) * len, hipMemcpyDeviceToHost));
    return rc;
}

extern "C" int32_t mcs_mg_solve_mesh(mcs_mg *mg, mcs_mesh *m, int32_t *iters, double *resid)
{
    if (!mg || !m) return MCS_E_ARG;
    mcs_ctx *ctx = mg->ctx;
    for (int d = 0; d < 3; ++d)
        if (m->n[d] != mg->cfg.n[d]) return mcs_fail(ctx, MCS_E_ARG, "mg: mesh size differs from the solver's");
    hipSetDevice(ctx->device);
    return mg_solve_dev(mg, m->arr[MCS_ARR_U], m->arr[MCS_ARR_Q], iters, resid);
}

extern "C" int32_t mcs_mg_setinhom_mesh(mcs_mg *mg, mcs_mesh *m, int32_t face, const double *val)
{
    if (!mg || !m || !val || face < 0 || face > 5) return MCS_E_ARG;
    mcs_ctx *ctx = mg->ctx;
    hipSetDevice(ctx->device);
    Lvl L = lvl(mg, 0);
    int d = face / 2;
    int64_t cnt = (int64_t)L.n[(d + 1) % 3] * L.n[(d + 2) % 3];
    double *dval;
    void *scr;
    int32_t rc = mcs_scratch(ctx, sizeof(double) * cnt, &scr);
    if (rc) return rc;
    dval = (double *)scr;
    MCS_HIP(ctx, hipMemcpyAsync(dval, val, sizeof(double) * cnt, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_mg_inhom, dim3((unsigned)cdiv(cnt, MCS_BLOCK)), dim3(MCS_BLOCK), 0, ctx->stream,
                       m->arr[MCS_ARR_Q], L, (int)face, 2.0 / mg->cfg.s, (const double *)dval);
    MCS_LAUNCH_CHECK(ctx);
    MCS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return MCS_OK;
}
