// Multigrid in two dimensions (Multigrid/src/Multigrid.jl with N = 2) with the
// Cartesian or the cylindrical connector (Multigrid.jl:31-39), the reference's
// axisymmetric (r, z) solver (MCStreamer/src/main.jl:139 passes
// CylindricalConnector{1}).  One launch per phase; the phases and their
// arithmetic are those of the three-dimensional solver in mcs_mg.hip with the
// 5-point stencil: lplstencil +x, +y, -x, -y; redblack2 colours; 4 children per
// coarse cell (x offset slowest); prolongation weights prod(3 - 2 sh) / 16.  The
// coarse direct solve uses the Cartesian operator, as buildmat does.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "mcs_internal.h"

namespace {

struct Lv2 {
    int32_t n[2];
};

__device__ __forceinline__ int64_t g2(const Lv2 &L, int64_t i, int64_t j)   // interior (i, j), 0-based
{
    return (i + 1) + (int64_t)(L.n[0] + 2) * (j + 1);
}

// Connector factor of the neighbour at offset sgn along d of the cell with
// interior index i along d: 1 + sgn / (2 i + 1) on the cylindrical dimension
// (cyl = 1 + d), 1 elsewhere (x * 1.0 is x).
__device__ __forceinline__ double conn2(int32_t cyl, int d, int sgn, int64_t i)
{
    if (cyl != d + 1) return 1.0;
    return 1.0 + (double)sgn / (double)(2 * i + 1);
}

// The laplacian with apply!'s ghost values folded in: a neighbour beyond face f
// is the cell itself times the face's factor, which is what the ghost holds
// when apply! ran since the cell last changed (the cell is not updated between
// apply! and its own update, nor before the residual).
__device__ __forceinline__ double lap2(const double *u, const Lv2 &L, int32_t cyl, uint32_t neg, int64_t i,
                                       int64_t j, int64_t c)
{
    const int64_t sy = L.n[0] + 2;
    const double uc = u[c];
    const double vxp = i + 1 < L.n[0] ? u[c + 1] : ((neg >> 1) & 1u ? -uc : uc);
    const double vyp = j + 1 < L.n[1] ? u[c + sy] : ((neg >> 3) & 1u ? -uc : uc);
    const double vxm = i > 0 ? u[c - 1] : (neg & 1u ? -uc : uc);
    const double vym = j > 0 ? u[c - sy] : ((neg >> 2) & 1u ? -uc : uc);
    double l = -4.0 * uc;
    l += vxp * conn2(cyl, 0, 1, i);
    l += vyp * conn2(cyl, 1, 1, j);
    l += vxm * conn2(cyl, 0, -1, i);
    l += vym * conn2(cyl, 1, -1, j);
    return l;
}

// apply! in 2D: a ghost cell takes the nearest interior value times the factor
// (+1 Neumann, -1 Dirichlet) of every face it lies beyond, which is what the
// face-by-face copies x-, x+, y-, y+ leave (corners included).
__global__ void k2_bc(double *a, Lv2 L, uint32_t neg)
{
    const int64_t t = gidx();
    const int64_t sx = L.n[0] + 2, tot = sx * (L.n[1] + 2);
    if (t >= tot) return;
    const int64_t i = t % sx, j = t / sx;
    const bool xl = i == 0, xh = i == L.n[0] + 1, yl = j == 0, yh = j == L.n[1] + 1;
    if (!(xl || xh || yl || yh)) return;
    const int64_t vi = xl ? 1 : (xh ? L.n[0] : i), vj = yl ? 1 : (yh ? L.n[1] : j);
    uint32_t flip = 0;
    if (xl) flip ^= neg & 1u;
    if (xh) flip ^= (neg >> 1) & 1u;
    if (yl) flip ^= (neg >> 2) & 1u;
    if (yh) flip ^= (neg >> 3) & 1u;
    const double v = a[vi + sx * vj];
    a[t] = flip ? -v : v;
}

// One colour of gauss_seidel! (redblack2: colour par = cells with i + j % 2 == par).
__global__ void k2_gs(double *u, const double *b, Lv2 L, int32_t cyl, uint32_t neg, int par)
{
    const int64_t t = gidx();
    const int64_t nxh = L.n[0] / 2;
    if (t >= nxh * L.n[1]) return;
    const int64_t j = t / nxh, i = 2 * (t - j * nxh) + ((j + par) & 1);
    const int64_t c = g2(L, i, j);
    const double l = lap2(u, L, cyl, neg, i, j, c);
    u[c] += (1.0 * (l + b[c])) / 4.0;
}

// residual! (r = s b + L u on the interior) with per-block sums of r^2 when
// partial is given (thread order, then a fixed tree).
__global__ __launch_bounds__(MCS_BLOCK) void k2_resid(double *r, const double *u, const double *b, double s, Lv2 L,
                                                      int32_t cyl, uint32_t neg, double *partial)
{
    __shared__ double sh[MCS_BLOCK];
    const int64_t t = gidx();
    double q = 0.0;
    if (t < (int64_t)L.n[0] * L.n[1]) {
        const int64_t j = t / L.n[0], i = t - j * L.n[0];
        const int64_t c = g2(L, i, j);
        const double v = s * b[c] + lap2(u, L, cyl, neg, i, j, c);
        r[c] = v;
        q = v * v;
    }
    if (!partial) return;
    sh[threadIdx.x] = q;
    __syncthreads();
    for (int o = MCS_BLOCK / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// Per-block sums of b^2 over the interior (the relative tolerance's rms of b).
__global__ __launch_bounds__(MCS_BLOCK) void k2_sumsq(const double *b, Lv2 L, double *partial)
{
    __shared__ double sh[MCS_BLOCK];
    const int64_t t = gidx();
    double q = 0.0;
    if (t < (int64_t)L.n[0] * L.n[1]) {
        const int64_t j = t / L.n[0], i = t - j * L.n[0];
        const double v = b[g2(L, i, j)];
        q = v * v;
    }
    sh[threadIdx.x] = q;
    __syncthreads();
    for (int o = MCS_BLOCK / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// The partials summed in order by one block, the total published to the host.
__global__ __launch_bounds__(MCS_BLOCK) void k2_sum(const double *partial, int64_t nb, uint64_t *pub, uint64_t seq)
{
    __shared__ double sh[MCS_BLOCK];
    double acc = 0.0;
    const int64_t bd = MCS_BLOCK;
    int64_t q = threadIdx.x;
    // eight loads in flight, added in the order of the plain loop
    for (; q + 7 * bd < nb; q += 8 * bd) {
        double v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = partial[q + k * bd];
#pragma unroll
        for (int k = 0; k < 8; ++k) acc += v[k];
    }
    for (; q < nb; q += bd) acc += partial[q];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int o = MCS_BLOCK / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        pub[1] = (uint64_t)__double_as_longlong(sh[0]);
        mcs_pub_raise(pub, seq);
    }
}

// restrict!: rh = 4 (sum of the 4 children) / 4, children (0,0), (0,1), (1,0), (1,1).
__global__ void k2_restrict(double *rh, const double *r, Lv2 F)
{
    const int64_t t = gidx();
    const int64_t nc0 = F.n[0] / 2, nc1 = F.n[1] / 2;
    if (t >= nc0 * nc1) return;
    const int64_t jc = t / nc0, ic = t - jc * nc0;
    const int64_t fy = F.n[0] + 2;
    const int64_t f = (2 * ic + 1) + fy * (2 * jc + 1);
    double s = 0.0;
    s += r[f];
    s += r[f + fy];
    s += r[f + 1];
    s += r[f + 1 + fy];
    rh[(ic + 1) + (nc0 + 2) * (jc + 1)] = (4.0 * s) / 4.0;
}

// interpolate! into fine cell (i, j): the four coarse values towards the cell,
// weights prod(3 - 2 sh) / 16 in cubestencil order; added or written.
__global__ void k2_interp(double *r, const double *rh, Lv2 F, int add)
{
    const int64_t t = gidx();
    if (t >= (int64_t)F.n[0] * F.n[1]) return;
    const int64_t j = t / F.n[0], i = t - j * F.n[0];
    const int64_t nc0 = F.n[0] / 2, cy = nc0 + 2;
    const int64_t ic = i / 2, jc = j / 2, sx = i & 1, sy = j & 1;
    const int64_t cc = (ic + 1) + cy * (jc + 1);
    const int64_t ddx = 2 * sx - 1, ddy = 2 * sy - 1;
    double v = 0.0;
#pragma unroll
    for (int hx = 0; hx < 2; ++hx)
#pragma unroll
        for (int hy = 0; hy < 2; ++hy) {
            const double w = (double)((3 - 2 * hx) * (3 - 2 * hy)) / 16.0;
            v += w * rh[cc + ddx * hx + cy * ddy * hy];
        }
    const int64_t c = g2(F, i, j);
    r[c] = add ? r[c] + v : v;
}

// z = A^-1 btop on the coarsest interior (at most 4096 cells), one row per
// thread of one block, columns in order; btop staged in LDS.
__global__ __launch_bounds__(MCS_BLOCK) void k2_coarse(const double *ainv, const double *res, double *z, Lv2 C)
{
    __shared__ double bt[4096];
    const int N = C.n[0] * C.n[1];
    for (int col = threadIdx.x; col < N; col += MCS_BLOCK) {
        const int jc = col / C.n[0], ic = col - jc * C.n[0];
        bt[col] = res[g2(C, ic, jc)];
    }
    __syncthreads();
    for (int row = threadIdx.x; row < N; row += MCS_BLOCK) {
        const double *a = ainv + (int64_t)row * N;
        double acc = 0.0;
        int col = 0;
        for (; col + 8 <= N; col += 8) {
            double av[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) av[k] = a[col + k];
#pragma unroll
            for (int k = 0; k < 8; ++k) acc += av[k] * bt[col + k];
        }
        for (; col < N; ++col) acc += a[col] * bt[col];
        const int j = row / C.n[0], i = row - j * C.n[0];
        z[g2(C, i, j)] = acc;
    }
}

__global__ void k2_fill0(double *a, int64_t n)
{
    const int64_t t = gidx();
    if (t < n) a[t] = 0.0;
}

__global__ void k2_axpy(double *x, const double *z, int64_t n)
{
    const int64_t t = gidx();
    if (t < n) x[t] += z[t];
}

// ---- the coarse levels in one workgroup ---------------------------------------

#ifndef MG2_TAIL_THREADS
#define MG2_TAIL_THREADS 1024
#endif
#define MG2_MAXL 16

struct Tail2 {
    Lv2 n[MG2_MAXL];
    int64_t len[MG2_MAXL];
    int64_t off_x[MG2_MAXL], off_b[MG2_MAXL];   // LDS offsets (doubles) of the levels' x and b
    int64_t off_r;                              // fine residual scratch (start level, ghost layout)
    int L, smooth1, smooth2, cyl;
    uint32_t neg;
};

#define TLOOP(t, N) for (int64_t t = threadIdx.x; t < (N); t += MG2_TAIL_THREADS)

__device__ __forceinline__ Lv2 t2lvl(const Tail2 &T, int l)
{
    Lv2 L;
    L.n[0] = T.n[l].n[0];
    L.n[1] = T.n[l].n[1];
    return L;
}

__device__ __forceinline__ void t2_gs(double *u, const double *b, const Lv2 &L, const Tail2 &T, int par)
{
    const int64_t nxh = L.n[0] / 2;
    TLOOP(t, nxh * L.n[1])
    {
        const int64_t j = t / nxh, i = 2 * (t - j * nxh) + ((j + par) & 1);
        const int64_t c = g2(L, i, j);
        const double l = lap2(u, L, T.cyl, T.neg, i, j, c);
        u[c] += (1.0 * (l + b[c])) / 4.0;
    }
    __syncthreads();
}

__device__ __forceinline__ void t2_bc(double *a, const Lv2 &L, uint32_t neg)
{
    const int64_t sx = L.n[0] + 2, tot = sx * (L.n[1] + 2);
    TLOOP(t, tot)
    {
        const int64_t i = t % sx, j = t / sx;
        const bool xl = i == 0, xh = i == L.n[0] + 1, yl = j == 0, yh = j == L.n[1] + 1;
        if (xl || xh || yl || yh) {
            const int64_t vi = xl ? 1 : (xh ? L.n[0] : i), vj = yl ? 1 : (yh ? L.n[1] : j);
            uint32_t flip = 0;
            if (xl) flip ^= neg & 1u;
            if (xh) flip ^= (neg >> 1) & 1u;
            if (yl) flip ^= (neg >> 2) & 1u;
            if (yh) flip ^= (neg >> 3) & 1u;
            const double v = a[vi + sx * vj];
            a[t] = flip ? -v : v;
        }
    }
    __syncthreads();
}

// mgv! from level `start` (x and b of that level in LDS), as mgv2 does it.
__device__ void t2_vcycle(const Tail2 &T, double *lds, int start)
{
    double *r = lds + T.off_r;
    for (int lev = start; lev < T.L; ++lev) {
        double *x = lds + T.off_x[lev];
        const double *b = lds + T.off_b[lev];
        const Lv2 L = t2lvl(T, lev);
        for (int it = 0; it < T.smooth1; ++it) {
            t2_gs(x, b, L, T, 0);
            t2_gs(x, b, L, T, 1);
        }
        if (T.smooth2 == 0) t2_bc(x, L, T.neg);
        TLOOP(t, (int64_t)L.n[0] * L.n[1])
        {
            const int64_t j = t / L.n[0], i = t - j * L.n[0];
            const int64_t c = g2(L, i, j);
            r[c] = 1.0 * b[c] + lap2(x, L, T.cyl, T.neg, i, j, c);
        }
        __syncthreads();
        const Lv2 C = t2lvl(T, lev + 1);
        double *bh = lds + T.off_b[lev + 1], *xh = lds + T.off_x[lev + 1];
        TLOOP(t, T.len[lev + 1]) xh[t] = 0.0;
        TLOOP(t, (int64_t)C.n[0] * C.n[1])
        {
            const int64_t jc = t / C.n[0], ic = t - jc * C.n[0];
            const int64_t fy = L.n[0] + 2;
            const int64_t f = (2 * ic + 1) + fy * (2 * jc + 1);
            double sum = 0.0;
            sum += r[f];
            sum += r[f + fy];
            sum += r[f + 1];
            sum += r[f + 1 + fy];
            bh[g2(C, ic, jc)] = (4.0 * sum) / 4.0;
        }
        __syncthreads();
    }
    {
        double *x = lds + T.off_x[T.L];
        const double *b = lds + T.off_b[T.L];
        const Lv2 L = t2lvl(T, T.L);
        for (int it = 0; it < 50; ++it) {
            if (it == 49) t2_bc(x, L, T.neg);
            t2_gs(x, b, L, T, 0);
            t2_gs(x, b, L, T, 1);
        }
    }
    for (int lev = T.L - 1; lev >= start; --lev) {
        double *x = lds + T.off_x[lev];
        const double *b = lds + T.off_b[lev];
        const double *xh = lds + T.off_x[lev + 1];
        const Lv2 L = t2lvl(T, lev);
        TLOOP(t, (int64_t)L.n[0] * L.n[1])
        {
            const int64_t j = t / L.n[0], i = t - j * L.n[0];
            const int64_t cy = L.n[0] / 2 + 2;
            const int64_t ic = i / 2, jc = j / 2, sx = i & 1, sy = j & 1;
            const int64_t cc = (ic + 1) + cy * (jc + 1);
            const int64_t ddx = 2 * sx - 1, ddy = 2 * sy - 1;
            double v = 0.0;
#pragma unroll
            for (int hx = 0; hx < 2; ++hx)
#pragma unroll
                for (int hy = 0; hy < 2; ++hy) {
                    const double w = (double)((3 - 2 * hx) * (3 - 2 * hy)) / 16.0;
                    v += w * xh[cc + ddx * hx + cy * ddy * hy];
                }
            const int64_t c = g2(L, i, j);
            x[c] = x[c] + v;
        }
        __syncthreads();
        for (int it = 0; it < T.smooth2; ++it) {
            if (it == T.smooth2 - 1) t2_bc(x, L, T.neg);
            t2_gs(x, b, L, T, 0);
            t2_gs(x, b, L, T, 1);
        }
    }
}

// mgv! of levels start..L in one workgroup: b0 and x0 (or zero) loaded, x0 (ghost
// shell included, as the cycle leaves it) written back.
__device__ __forceinline__ void tail2_body(const Tail2 &T, int start, double *x0, const double *b0, int zero,
                                           double *lds)
{
    double *x = lds + T.off_x[start], *b = lds + T.off_b[start];
    TLOOP(t, T.len[start])
    {
        b[t] = b0[t];
        x[t] = zero ? 0.0 : x0[t];
    }
    __syncthreads();
    t2_vcycle(T, lds, start);
    TLOOP(t, T.len[start]) x0[t] = x[t];
}

__global__ __launch_bounds__(MG2_TAIL_THREADS) void k2_tail(Tail2 T, int start, double *x0, const double *b0,
                                                            int zero)
{
    extern __shared__ double lds[];
    tail2_body(T, start, x0, b0, zero, lds);
}

// Tabulation of the tail: block q runs mgv! from x = 0 on its own pair of
// start-level arrays, q lengths into x0 and b0.
__global__ __launch_bounds__(MG2_TAIL_THREADS) void k2_tail_batch(Tail2 T, int start, double *x0, const double *b0)
{
    extern __shared__ double lds[];
    const int64_t off = (int64_t)blockIdx.x * T.len[start];
    tail2_body(T, start, x0 + off, b0 + off, 1, lds);
}

// ---- tabulated tail ----------------------------------------------------------
//
// As in three dimensions (mcs_mg.hip): below the tail's start level S the cycle
// from x = 0 is a fixed linear map of S's right-hand side, tabulated once by
// running the tail on unit right-hand sides and then applied as one wide
// matrix-vector product.  Same operator; the product rounds differently from
// the sweeps.

// Unit inputs: array q (len apart) zero except at valid cell q of level L.
__global__ void k2_units(double *a, int64_t len, Lv2 L, int64_t ncols)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= len * ncols) return;
    const int64_t q = t / len, p = t - q * len;
    a[t] = p == g2(L, q % L.n[0], q / L.n[0]) ? 1.0 : 0.0;
}

// a = 0 except 1 at valid cell q of level L.
__global__ void k2_unit1(double *a, int64_t len, Lv2 L, int64_t q)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < len) a[p] = p == g2(L, q % L.n[0], q / L.n[0]) ? 1.0 : 0.0;
}

__global__ void k2_copy(double *dst, const double *src, int64_t n)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) dst[p] = src[p];
}

// M[r][q] = X[q][r].
__global__ void k2_transpose(double *M, const double *X, int64_t rows, int64_t ncols)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= rows * ncols) return;
    const int64_t r = t / ncols, q = t - r * ncols;
    M[t] = X[q * rows + r];
}

#define MG2_TM_WAVES 4
// x = M b on level L: b's valid cells staged in LDS in cell order, a wavefront per
// output position (ghosts included), fused multiply-adds and a fixed butterfly.
__global__ __launch_bounds__(MG2_TM_WAVES * MCS_WAVE) void k2_tailmat(const double *__restrict__ M,
                                                                     const double *__restrict__ b,
                                                                     double *__restrict__ x, Lv2 L, int64_t rows)
{
    extern __shared__ double bs[];
    const int C = L.n[0] * L.n[1];
    for (int t = threadIdx.x; t < C; t += MG2_TM_WAVES * MCS_WAVE) bs[t] = b[g2(L, t % L.n[0], t / L.n[0])];
    __syncthreads();
    const int lane = threadIdx.x & (MCS_WAVE - 1);
    const int64_t r = (int64_t)blockIdx.x * MG2_TM_WAVES + (threadIdx.x / MCS_WAVE);
    if (r >= rows) return;
    const double *row = M + r * C;
    double acc = 0.0;
    for (int q = lane; q < C; q += MCS_WAVE) acc = __builtin_fma(row[q], bs[q], acc);
#pragma unroll
    for (int o = MCS_WAVE / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, MCS_WAVE);
    if (lane == 0) x[r] = acc;
}

inline unsigned nblk(int64_t n)
{
    return (unsigned)((n + MCS_BLOCK - 1) / MCS_BLOCK);
}

Lv2 lv2(const mcs_mg *mg, int l)
{
    Lv2 L;
    L.n[0] = mg->lv[l].n[0];
    L.n[1] = mg->lv[l].n[1];
    return L;
}

uint32_t neg2(const mcs_mg *mg)
{
    uint32_t m = 0;
    for (int f = 0; f < 4; ++f)
        if (mg->cfg.bc[f] == MCS_BC_DIRICHLET) m |= 1u << f;
    return m;
}

// buildmat in 2D (Cartesian 5-point operator, ghost rule folded), inverted by the
// Gauss-Jordan elimination of the CPU restatement.
int coarse_inverse2(const int32_t n[2], const int32_t bc[4], std::vector<double> &ainv)
{
    const int64_t N = (int64_t)n[0] * n[1];
    std::vector<double> a((size_t)(N * N), 0.0);
    ainv.assign((size_t)(N * N), 0.0);
    for (int64_t j = 0; j < n[1]; ++j)
        for (int64_t i = 0; i < n[0]; ++i) {
            const int64_t row = i + n[0] * j;
            const int64_t c[2] = {i, j};
            a[row * N + row] -= 4.0;
            for (int q = 0; q < 4; ++q) {
                const int d = q % 2, sgn = q < 2 ? 1 : -1;
                int64_t cc[2] = {c[0], c[1]};
                cc[d] += sgn;
                if (cc[d] < 0 || cc[d] >= n[d]) {
                    const int face = 2 * d + (sgn > 0 ? 1 : 0);
                    a[row * N + row] += bc[face] == MCS_BC_DIRICHLET ? -1.0 : 1.0;
                } else {
                    a[row * N + cc[0] + n[0] * cc[1]] += 1.0;
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
        if (piv != col)
            for (int64_t c = 0; c < N; ++c) {
                std::swap(a[col * N + c], a[piv * N + c]);
                std::swap(ainv[col * N + c], ainv[piv * N + c]);
            }
        const double d = a[col * N + col];
        for (int64_t c = 0; c < N; ++c) {
            a[col * N + c] = a[col * N + c] / d;
            ainv[col * N + c] = ainv[col * N + c] / d;
        }
        for (int64_t r = 0; r < N; ++r) {
            if (r == col) continue;
            const double fct = a[r * N + col];
            if (fct == 0.0) continue;
            for (int64_t c = 0; c < N; ++c) {
                a[r * N + c] = a[r * N + c] - fct * a[col * N + c];
                ainv[r * N + c] = ainv[r * N + c] - fct * ainv[col * N + c];
            }
        }
    }
    return 0;
}

void bc2(mcs_mg *mg, int l, double *a)
{
    const Lv2 L = lv2(mg, l);
    hipLaunchKernelGGL(k2_bc, dim3(nblk(mg->lv[l].len)), dim3(MCS_BLOCK), 0, mg->ctx->stream, a, L, neg2(mg));
}

void gs2(mcs_mg *mg, int l, double *u, const double *b)
{
    const Lv2 L = lv2(mg, l);
    const int64_t cnt = (int64_t)(L.n[0] / 2) * L.n[1];
    // the finest level's sweeps are timed on their own (one sweep per record)
    ProfScope ps(mg->ctx, l == 0 ? "mg2_sweep0" : "mg2_sweep");
    for (int par = 0; par < 2; ++par)
        hipLaunchKernelGGL(k2_gs, dim3(nblk(cnt)), dim3(MCS_BLOCK), 0, mg->ctx->stream, u, b, L, mg->cfg.cyl,
                           neg2(mg), par);
}

void fill0(mcs_mg *mg, double *a, int64_t n)
{
    hipLaunchKernelGGL(k2_fill0, dim3(nblk(n)), dim3(MCS_BLOCK), 0, mg->ctx->stream, a, n);
}

int64_t tail2_setup(const mcs_mg *mg, int start, Tail2 &T)
{
    memset(&T, 0, sizeof(T));
    T.L = mg->cfg.levels;
    T.smooth1 = mg->cfg.smooth1;
    T.smooth2 = mg->cfg.smooth2;
    T.cyl = mg->cfg.cyl;
    T.neg = neg2(mg);
    int64_t off = 0;
    for (int l = 0; l <= T.L; ++l) {
        T.n[l] = lv2(mg, l);
        T.len[l] = mg->lv[l].len;
        if (l >= start) {
            T.off_x[l] = off;
            off += T.len[l];
            T.off_b[l] = off;
            off += T.len[l];
        }
    }
    T.off_r = off;
    off += mg->lv[start].len;
    return off;   // doubles
}

// The first level from which a V-cycle runs in one workgroup (its levels fit the
// LDS budget), or levels + 1 when none does.
int tail2_start(const mcs_mg *mg)
{
    if (mg->tail_lds == 0) return mg->cfg.levels + 1;
    int S = mg->cfg.levels + 1;
    for (int l = mg->cfg.levels; l >= 0; --l) {
        Tail2 T;
        if ((size_t)tail2_setup(mg, l, T) * sizeof(double) > mg->tail_lds) break;
        S = l;
    }
    return S;
}

void tail2(mcs_mg *mg, int level, double *x, const double *b, bool zero)
{
    if (zero && level == mg->tm_level) {
        const int64_t rows = mg->lv[level].len;
        const Lv2 L = lv2(mg, level);
        hipLaunchKernelGGL(k2_tailmat, dim3((unsigned)((rows + MG2_TM_WAVES - 1) / MG2_TM_WAVES)),
                           dim3(MG2_TM_WAVES * MCS_WAVE), sizeof(double) * (size_t)L.n[0] * L.n[1], mg->ctx->stream,
                           (const double *)mg->tm_z, b, x, L, rows);
        return;
    }
    Tail2 T;
    const int64_t words = tail2_setup(mg, level, T);
    hipLaunchKernelGGL(k2_tail, dim3(1), dim3(MG2_TAIL_THREADS), (size_t)words * sizeof(double), mg->ctx->stream, T,
                       level, x, b, zero ? 1 : 0);
}

// mgv! (Multigrid.jl:334-377).  The sweeps and the residual fold apply! in
// (lap2); the ghost shell is written where a later step reads it: before the
// last sweep (the coarsest level's 50th, the last post-smoothing one), whose
// values the caller's interpolation reads, as the reference leaves them; without
// post-smoothing, before the residual.  apply!(r) before restrict! is left out:
// the restriction reads interior cells only.
void mgv2(mcs_mg *mg, double *x, const double *b, int level, bool zero = false)
{
    hipStream_t st = mg->ctx->stream;
    if (level >= mg->tail2) {
        tail2(mg, level, x, b, zero);
        return;
    }
    if (zero) fill0(mg, x, mg->lv[level].len);
    if (level == mg->cfg.levels) {
        for (int it = 0; it < 50; ++it) {
            if (it == 49) bc2(mg, level, x);
            gs2(mg, level, x, b);
        }
        return;
    }
    for (int it = 0; it < mg->cfg.smooth1; ++it) gs2(mg, level, x, b);
    const Lv2 L = lv2(mg, level);
    double *r = mg->lv[level].res1;
    if (mg->cfg.smooth2 == 0) bc2(mg, level, x);
    hipLaunchKernelGGL(k2_resid, dim3(nblk((int64_t)L.n[0] * L.n[1])), dim3(MCS_BLOCK), 0, st, r, (const double *)x,
                       b, 1.0, L, mg->cfg.cyl, neg2(mg), (double *)nullptr);
    double *rh = mg->lv[level + 1].res;
    hipLaunchKernelGGL(k2_restrict, dim3(nblk((int64_t)(L.n[0] / 2) * (L.n[1] / 2))), dim3(MCS_BLOCK), 0, st, rh,
                       (const double *)r, L);
    double *xh = mg->lv[level + 1].sol;
    mgv2(mg, xh, rh, level + 1, true);
    hipLaunchKernelGGL(k2_interp, dim3(nblk((int64_t)L.n[0] * L.n[1])), dim3(MCS_BLOCK), 0, st, x,
                       (const double *)xh, L, 1);
    for (int it = 0; it < mg->cfg.smooth2; ++it) {
        if (it == mg->cfg.smooth2 - 1) bc2(mg, level, x);
        gs2(mg, level, x, b);
    }
}

void fmg2_body(mcs_mg *mg, double *x, const double *b);
double *fmg2_bottom(mcs_mg *mg, int S);

// Tabulates mgv! from x = 0 at the tail's start level (mcs_mg::tm_level and
// tm_z) and fmg!'s coarse part from that level's right-hand side (tm_f); off with
// MCS_MG_TAIL_MATRIX=0, and when the start is the finest level or holds more
// than 4096 cells.
int32_t tabulate_tail2(mcs_mg *mg)
{
    mcs_ctx *ctx = mg->ctx;
    mg->tm_level = -1;
    const char *e = getenv("MCS_MG_TAIL_MATRIX");
    if (e && e[0] == '0') return MCS_OK;
    const int S = mg->tail2;
    if (S < 1 || S > mg->cfg.levels) return MCS_OK;
    const Lv2 L = lv2(mg, S);
    const int64_t C = (int64_t)L.n[0] * L.n[1], R = mg->lv[S].len;
    if (C > 4096) return MCS_OK;
    Tail2 T;
    const size_t lds = (size_t)tail2_setup(mg, S, T) * sizeof(double);
    if (lds > 64 * 1024 &&
        hipFuncSetAttribute((const void *)k2_tail_batch, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)mg->tail_lds) != hipSuccess)
        return MCS_OK;
    double *X = nullptr, *B = nullptr, *M = nullptr;
    if (hipMalloc(&X, sizeof(double) * C * R) != hipSuccess || hipMalloc(&B, sizeof(double) * C * R) != hipSuccess ||
        hipMalloc(&M, sizeof(double) * R * C) != hipSuccess) {
        if (X) hipFree(X);
        if (B) hipFree(B);
        if (M) hipFree(M);
        return mcs_fail(ctx, MCS_E_HIP, "mg: device allocation of the tail table failed");
    }
    hipLaunchKernelGGL(k2_units, dim3(nblk(C * R)), dim3(MCS_BLOCK), 0, ctx->stream, B, R, L, C);
    hipLaunchKernelGGL(k2_tail_batch, dim3((unsigned)C), dim3(MG2_TAIL_THREADS), lds, ctx->stream, T, S, X,
                       (const double *)B);
    hipLaunchKernelGGL(k2_transpose, dim3(nblk(C * R)), dim3(MCS_BLOCK), 0, ctx->stream, M, (const double *)X, R, C);
    const hipError_t le = hipGetLastError();
    const hipError_t se = hipStreamSynchronize(ctx->stream);
    hipFree(X);
    hipFree(B);
    if (le != hipSuccess || se != hipSuccess) {
        hipFree(M);
        return mcs_fail(ctx, MCS_E_HIP, "mg: tabulating the tail failed (%s)",
                        hipGetErrorString(le != hipSuccess ? le : se));
    }
    mg->tm_level = S;
    mg->tm_z = M;
    // fmg!'s coarse part from S's right-hand side: one unit right-hand side at a
    // time through the same launches, its solution a column of the table
    double *Xf = nullptr, *Mf = nullptr;
    if (hipMalloc(&Xf, sizeof(double) * C * R) != hipSuccess || hipMalloc(&Mf, sizeof(double) * R * C) != hipSuccess) {
        if (Xf) hipFree(Xf);
        return mcs_fail(ctx, MCS_E_HIP, "mg: device allocation of the tail table failed");
    }
    for (int64_t q = 0; q < C; ++q) {
        hipLaunchKernelGGL(k2_unit1, dim3(nblk(R)), dim3(MCS_BLOCK), 0, ctx->stream, mg->lv[S].res, R, L, q);
        const double *zs = fmg2_bottom(mg, S);
        hipLaunchKernelGGL(k2_copy, dim3(nblk(R)), dim3(MCS_BLOCK), 0, ctx->stream, Xf + q * R, zs, R);
    }
    hipLaunchKernelGGL(k2_transpose, dim3(nblk(C * R)), dim3(MCS_BLOCK), 0, ctx->stream, Mf, (const double *)Xf, R,
                       C);
    const hipError_t fe = hipGetLastError();
    const hipError_t fs = hipStreamSynchronize(ctx->stream);
    hipFree(Xf);
    if (fe != hipSuccess || fs != hipSuccess) {
        hipFree(Mf);
        return mcs_fail(ctx, MCS_E_HIP, "mg: tabulating fmg!'s coarse part failed (%s)",
                        hipGetErrorString(fe != hipSuccess ? fe : fs));
    }
    mg->tm_f = Mf;
    return MCS_OK;
}

// fmg! (Multigrid.jl:380-434).  Returns 1 after a correction, 0 when converged,
// negative on a device error.
int fmg2(mcs_mg *mg, double *x, const double *b, double *eps_out)
{
    mcs_ctx *ctx = mg->ctx;
    hipStream_t st = ctx->stream;
    const int L = mg->cfg.levels;
    const Lv2 F = lv2(mg, 0);
    const int64_t N0 = (int64_t)F.n[0] * F.n[1];
    // apply!(x) folded into the residual; x's ghost shell is already what apply!
    // writes (the previous correction, or the solve's start, left it)
    fill0(mg, mg->lv[0].res, mg->lv[0].len);
    hipLaunchKernelGGL(k2_resid, dim3(nblk(N0)), dim3(MCS_BLOCK), 0, st, mg->lv[0].res, (const double *)x, b,
                       mg->cfg.s, F, mg->cfg.cyl, neg2(mg), mg->partial);
    const uint64_t seq = mcs_pub_next(ctx);
    hipLaunchKernelGGL(k2_sum, dim3(1), dim3(MCS_BLOCK), 0, st, (const double *)mg->partial, (int64_t)nblk(N0),
                       ctx->d_pub, seq);
    double ss = 0.0;
    if (hipGetLastError() != hipSuccess || mcs_pub_wait(ctx, seq, &ss, 1) != MCS_OK) return -1;
    const double eps = sqrt(ss) / sqrt((double)F.n[0] * (double)F.n[1]);
    *eps_out = eps;
    if (mg->cfg.s * mg->tol_abs > eps) return 0;
    if (mg->use_graph) {
        // the correction is the same launch sequence for a given (x, b): captured
        // once (profiling records left out) and replayed
        if (!mg->graph_exec || mg->graph_x != x || mg->graph_b != b) {
            if (mg->graph_exec) {
                hipGraphExecDestroy(mg->graph_exec);
                mg->graph_exec = nullptr;
            }
            const bool prof = ctx->prof;
            ctx->prof = false;
            hipGraph_t g = nullptr;
            if (hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed) != hipSuccess) {
                ctx->prof = prof;
                return -1;
            }
            fmg2_body(mg, x, b);
            const hipError_t e = hipStreamEndCapture(st, &g);
            ctx->prof = prof;
            if (e != hipSuccess || hipGraphInstantiate(&mg->graph_exec, g, nullptr, nullptr, 0) != hipSuccess) {
                if (g) hipGraphDestroy(g);
                return -1;
            }
            hipGraphDestroy(g);
            mg->graph_x = x;
            mg->graph_b = b;
        }
        ProfScope ps(ctx, "mg2_cycle");
        if (hipGraphLaunch(mg->graph_exec, st) != hipSuccess) return -1;
        return hipGetLastError() == hipSuccess ? 1 : -1;
    }
    {
        ProfScope ps(ctx, "mg2_cycle");
        fmg2_body(mg, x, b);
    }
    return hipGetLastError() == hipSuccess ? 1 : -1;
}

// fmg!'s restrictions of the right-hand sides from level `from` down to the
// coarsest (Multigrid.jl:400-402).
void fmg2_restrict(mcs_mg *mg, int from, int to)
{
    for (int k = from; k < to; ++k) {
        const Lv2 Lk = lv2(mg, k);
        hipLaunchKernelGGL(k2_restrict, dim3(nblk((int64_t)(Lk.n[0] / 2) * (Lk.n[1] / 2))), dim3(MCS_BLOCK), 0,
                           mg->ctx->stream, mg->lv[k + 1].res, (const double *)mg->lv[k].res, Lk);
    }
}

// fmg!'s interpolate! + mgv! sweep (Multigrid.jl:418-429) from the solution z of
// level top + 1 up to level `to`; returns the solution of `to`.
double *fmg2_sweep(mcs_mg *mg, double *z, int top, int to)
{
    hipStream_t st = mg->ctx->stream;
    for (int lev = top; lev >= to; --lev) {
        double *z1 = mg->lv[lev].sol;
        fill0(mg, z1, mg->lv[lev].len);
        bc2(mg, lev + 1, z);
        const Lv2 Ll = lv2(mg, lev);
        hipLaunchKernelGGL(k2_interp, dim3(nblk((int64_t)Ll.n[0] * Ll.n[1])), dim3(MCS_BLOCK), 0, st, z1,
                           (const double *)z, Ll, 0);
        mgv2(mg, z1, mg->lv[lev].res, lev);
        z = z1;
    }
    return z;
}

// fmg!'s coarse part from level S's right-hand side in res[S]: the restrictions
// below it, the coarse solve and the sweep up to S; the result in sol[S].
double *fmg2_bottom(mcs_mg *mg, int S)
{
    const int L = mg->cfg.levels;
    fmg2_restrict(mg, S, L);
    hipLaunchKernelGGL(k2_coarse, dim3(1), dim3(MCS_BLOCK), 0, mg->ctx->stream, (const double *)mg->ainv,
                       (const double *)mg->lv[L].res, mg->lv[L].sol, lv2(mg, L));
    return fmg2_sweep(mg, mg->lv[L].sol, L - 1, S);
}

// The correction of fmg! after the convergence test: restrictions, the coarse
// solve, the interpolate! + mgv! sweep up the levels, x += z and apply!(x).  With
// the coarse part tabulated (tm_f), level tm_level's solution is its product
// with the restricted right-hand side.
void fmg2_body(mcs_mg *mg, double *x, const double *b)
{
    hipStream_t st = mg->ctx->stream;
    double *z;
    if (mg->tm_f) {
        const int S = mg->tm_level;
        fmg2_restrict(mg, 0, S);
        const int64_t rows = mg->lv[S].len;
        const Lv2 LS = lv2(mg, S);
        hipLaunchKernelGGL(k2_tailmat, dim3((unsigned)((rows + MG2_TM_WAVES - 1) / MG2_TM_WAVES)),
                           dim3(MG2_TM_WAVES * MCS_WAVE), sizeof(double) * (size_t)LS.n[0] * LS.n[1], st,
                           (const double *)mg->tm_f, (const double *)mg->lv[S].res, mg->lv[S].sol, LS, rows);
        z = fmg2_sweep(mg, mg->lv[S].sol, S - 1, 0);
    } else {
        z = fmg2_bottom(mg, 0);
    }
    hipLaunchKernelGGL(k2_axpy, dim3(nblk(mg->lv[0].len)), dim3(MCS_BLOCK), 0, st, x, (const double *)z,
                       mg->lv[0].len);
    bc2(mg, 0, x);
}

}  // namespace

int32_t mcs_mg2_create(mcs_ctx *ctx, const mcs_mg_config *cfg, mcs_mg **out)
{
    const mcs_mg_config &c = *cfg;
    if (c.n[2] != 1) return mcs_fail(ctx, MCS_E_ARG, "mg: a 2D solve has n[2] == 1");
    if (c.levels < 0 || c.levels >= 16 || c.smooth1 < 0 || c.smooth2 < 0 || c.maxiter < 0)
        return mcs_fail(ctx, MCS_E_ARG, "mg: bad levels/smoothing/maxiter");
    if (c.cyl < 0 || c.cyl > 2) return mcs_fail(ctx, MCS_E_ARG, "mg: cyl is 0 (Cartesian), 1 or 2");
    if (c.precision != MCS_PREC_F64) return mcs_fail(ctx, MCS_E_ARG, "mg: the 2D solve is FP64");
    for (int d = 0; d < 2; ++d)
        if (c.n[d] < 1 || c.n[d] % (1 << c.levels))
            return mcs_fail(ctx, MCS_E_ARG, "mg: n[%d]=%d is not divisible by 2^levels", d, c.n[d]);
    if ((c.n[0] >> c.levels) % 2)
        return mcs_fail(ctx, MCS_E_ARG, "mg: red-black needs an even x size on every level (redblack.jl:13)");
    const int64_t nc = (int64_t)(c.n[0] >> c.levels) * (c.n[1] >> c.levels);
    if (nc > 4096) return mcs_fail(ctx, MCS_E_ARG, "mg: coarsest level above 4096 cells; add levels");
    hipSetDevice(ctx->device);
    mcs_mg *mg = new mcs_mg();
    mg->ctx = ctx;
    mg->cfg = c;
    const char *ge = getenv("MCS_MG_GRAPH");
    mg->use_graph = !(ge && ge[0] == '0');
    for (int l = 0; l <= c.levels; ++l) {
        MgLevel L;
        memset(&L, 0, sizeof(L));
        L.n[0] = c.n[0] >> l;
        L.n[1] = c.n[1] >> l;
        L.n[2] = 1;
        L.len = (int64_t)(L.n[0] + 2) * (L.n[1] + 2);
        if (hipMalloc(&L.res, sizeof(double) * L.len) != hipSuccess ||
            hipMalloc(&L.res1, sizeof(double) * L.len) != hipSuccess ||
            hipMalloc(&L.sol, sizeof(double) * L.len) != hipSuccess) {
            mg->lv.push_back(L);
            mcs_mg_destroy(mg);
            return mcs_fail(ctx, MCS_E_HIP, "mg: device allocation failed");
        }
        mg->lv.push_back(L);
        if (hipMemsetAsync(L.res, 0, sizeof(double) * L.len, ctx->stream) != hipSuccess ||
            hipMemsetAsync(L.res1, 0, sizeof(double) * L.len, ctx->stream) != hipSuccess ||
            hipMemsetAsync(L.sol, 0, sizeof(double) * L.len, ctx->stream) != hipSuccess) {
            mcs_mg_destroy(mg);
            return mcs_fail(ctx, MCS_E_HIP, "mg: clearing the level arrays failed");
        }
    }
    const int32_t ncn[2] = {c.n[0] >> c.levels, c.n[1] >> c.levels};
    std::vector<double> ainv;
    if (coarse_inverse2(ncn, c.bc, ainv)) {
        mcs_mg_destroy(mg);
        return mcs_fail(ctx, MCS_E_ARG, "mg: singular coarse operator (all-Neumann boundaries?)");
    }
    mg->ncoarse = nc;
    if (hipMalloc(&mg->ainv, sizeof(double) * ainv.size()) != hipSuccess) {
        mcs_mg_destroy(mg);
        return mcs_fail(ctx, MCS_E_HIP, "mg: device allocation failed");
    }
    if (mcs_h2d(ctx, mg->ainv, ainv.data(), sizeof(double) * ainv.size()) != hipSuccess) {
        mcs_mg_destroy(mg);
        return mcs_fail(ctx, MCS_E_HIP, "mg: copy to the device failed");
    }
    mg->npartial = nblk((int64_t)c.n[0] * c.n[1]);
    if (hipMalloc(&mg->partial, sizeof(double) * (mg->npartial + 8)) != hipSuccess) {
        mcs_mg_destroy(mg);
        return mcs_fail(ctx, MCS_E_HIP, "mg: device allocation failed");
    }
    int lds_max = 0;
    if (hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device) != hipSuccess)
        lds_max = 64 * 1024;
    mg->tail_lds = (size_t)lds_max;
    if (mg->tail_lds > 64 * 1024 &&
        hipFuncSetAttribute((const void *)k2_tail, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max) != hipSuccess)
        mg->tail_lds = 64 * 1024;
    const char *te = getenv("MCS_MG2_TAIL");
    if (te && atoi(te) == 0) mg->tail_lds = 0;
    mg->tail2 = tail2_start(mg);
    MCS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const int32_t trc = tabulate_tail2(mg);
    if (trc) {
        mcs_mg_destroy(mg);
        return trc;
    }
    *out = mg;
    return MCS_OK;
}

// solve (Multigrid.jl:449-467) on device arrays with ghosts ((n0+2)(n1+2)).
int32_t mcs_mg2_solve_dev(mcs_mg *mg, double *u, const double *b, int32_t *iters, double *resid)
{
    mcs_ctx *ctx = mg->ctx;
    const Lv2 F = lv2(mg, 0);
    mg->tol_abs = mg->cfg.tolerance;
    if (mg->cfg.tolerance < 0.0) {
        const int64_t N0 = (int64_t)F.n[0] * F.n[1];
        hipLaunchKernelGGL(k2_sumsq, dim3(nblk(N0)), dim3(MCS_BLOCK), 0, ctx->stream, b, F, mg->partial);
        const uint64_t seq = mcs_pub_next(ctx);
        hipLaunchKernelGGL(k2_sum, dim3(1), dim3(MCS_BLOCK), 0, ctx->stream, (const double *)mg->partial,
                           (int64_t)nblk(N0), ctx->d_pub, seq);
        MCS_LAUNCH_CHECK(ctx);
        double ss = 0.0;
        int32_t rc = mcs_pub_wait(ctx, seq, &ss, 1);
        if (rc) return rc;
        const double brms = sqrt(ss) / sqrt((double)F.n[0] * (double)F.n[1]);
        mg->tol_abs = -mg->cfg.tolerance * brms;
        if (!(mg->tol_abs > 0.0)) mg->tol_abs = 2.2250738585072014e-308;
    }
    int iter = 0, cont = 1;
    double eps = 0.0;
    bc2(mg, 0, u);   // fmg!'s first apply!(x) (the later ones follow each x += z)
    while (cont && iter <= mg->cfg.maxiter) {
        cont = fmg2(mg, u, b, &eps);
        if (cont < 0) return mcs_fail(ctx, MCS_E_HIP, "mg: device error during the cycle");
        iter += 1;
    }
    if (iters) *iters = iter;
    if (resid) *resid = eps;
    if (iter > mg->cfg.maxiter)
        return mcs_fail(ctx, MCS_E_NONCONVERGED, "Convergence failed with [maxiter=] %d iterations",
                        mg->cfg.maxiter);
    return MCS_OK;
}
