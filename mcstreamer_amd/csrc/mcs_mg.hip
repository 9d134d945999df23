// mcs_mg.hip -- geometric multigrid for L u + s b = 0 on a cell-centred 3D
// mesh with one ghost layer (Multigrid/src/Multigrid.jl, redblack.jl,
// boundaries.jl).  The host drives the operation sequence of fmg!/mgv!; each
// operation is a coalesced stencil kernel over one level, and the levels below
// a few thousand cells run as one single-workgroup kernel.
//
// Ghost handling.  apply! (boundaries.jl:180-194) sets every ghost plane to
// +-1 times the adjacent valid plane, x first, then y, then z; a ghost cell
// therefore equals the product of the +-1 factors of its ghost directions
// times the nearest interior cell.  For the 7-point stencil the ghost neighbour
// of a boundary cell is that cell itself times the factor, so smoothing and
// residual kernels read it from the cell directly ("folded") -- the same value
// the materialised ghost would hold at that moment.  Ghosts are written to
// memory only where a later operation reads them: before the last smoothing
// sweep of a level (the prolongation into the finer level reads them), before
// the coarse-to-fine prolongation of fmg!, and at the start and end of fmg!.
//
// Precision mode (MCS_PREC_F32): the finest residual and the update x += z stay
// FP64; the correction equation L z + r = 0 of each fmg! pass (restrictions,
// coarse solve, V-cycles) runs on an FP32 copy of the hierarchy.  Every stencil
// body is a template on the element type, so the FP64 path is unchanged.
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "mcs_device.h"

#define MG_MAXL 13
#define MG_TAIL_LDS (160 * 1024)
#define MG_TAIL_THREADS 512

struct Lvl {
    int32_t n[3];
};

// Boundary factors of the ghost rule, each exactly +1 (Neumann) or -1
// (Dirichlet), as a sign mask: bit d for the low face of axis d, bit 3 + d for
// the high face.  Applying a factor is a conditional negation, which equals the
// product by +-1 exactly.
struct BcCoef {
    uint32_t neg;
};

template <typename T>
__device__ __forceinline__ T bcf(const BcCoef &bc, int bit, T v)
{
    return ((bc.neg >> bit) & 1u) ? -v : v;
}

// x / 6 and (4 sum) / 8: rounded divisions in FP64 (the restatement's order),
// multiplications by the exact reciprocal in FP32.
__device__ __forceinline__ double div6(double x) { return x / 6.0; }
__device__ __forceinline__ float div6(float x) { return x * (1.0f / 6.0f); }

__device__ __forceinline__ int64_t g_index(const Lvl &L, int64_t i, int64_t j, int64_t k)
{
    return i + (int64_t)(L.n[0] + 2) * (j + (int64_t)(L.n[1] + 2) * k);
}

// t -> (i, j, k) with i fastest, in 32-bit arithmetic (levels hold < 2^31 cells).
__device__ __forceinline__ void split3(int64_t t64, uint32_t nx, uint32_t ny, int64_t &i, int64_t &j,
                                       int64_t &k)
{
    uint32_t t = (uint32_t)t64;
    uint32_t q, r;
    if ((nx & (nx - 1)) == 0) {
        q = t >> __builtin_ctz(nx);
        i = t & (nx - 1);
    } else {
        q = t / nx;
        i = t - q * nx;
    }
    if ((ny & (ny - 1)) == 0) {
        r = q >> __builtin_ctz(ny);
        j = q & (ny - 1);
    } else {
        r = q / ny;
        j = q - r * ny;
    }
    k = r;
}

// L u at interior cell (i, j, k) with folded ghosts, in lplstencil order
// (+x, +y, +z, -x, -y, -z; Multigrid.jl:124-130,152-158).
template <typename T>
__device__ __forceinline__ T lap_fold(const T *__restrict__ u, const Lvl &L, const BcCoef &bc, int64_t i,
                                     int64_t j, int64_t k, int64_t c, T uc)
{
    const int64_t sy = L.n[0] + 2, sz = (int64_t)(L.n[0] + 2) * (L.n[1] + 2);
    T l = T(-6) * uc;
    l += (i + 1 < L.n[0]) ? u[c + 1] : bcf(bc, 3, uc);
    l += (j + 1 < L.n[1]) ? u[c + sy] : bcf(bc, 4, uc);
    l += (k + 1 < L.n[2]) ? u[c + sz] : bcf(bc, 5, uc);
    l += (i > 0) ? u[c - 1] : bcf(bc, 0, uc);
    l += (j > 0) ? u[c - sy] : bcf(bc, 1, uc);
    l += (k > 0) ? u[c - sz] : bcf(bc, 2, uc);
    return l;
}

// ---- per-item bodies shared by the grid kernels and the tail kernel ----------

// One cell of one colour of gauss_seidel! (Multigrid.jl:168-178, redblack.jl:49-61):
// colour par holds cells with (i+j+k) % 2 == par; each row has nx/2 of them.
template <typename T>
__device__ __forceinline__ void gs_item(T *u, const T *b, const Lvl &L, const BcCoef &bc, int par, int64_t t)
{
    int64_t ti, j, k;
    split3(t, (uint32_t)(L.n[0] / 2), (uint32_t)L.n[1], ti, j, k);
    int64_t i = 2 * ti + ((j + k + par) & 1);
    int64_t c = g_index(L, i + 1, j + 1, k + 1);
    T uc = u[c];
    T l = lap_fold(u, L, bc, i, j, k, c, uc);
    u[c] = uc + div6(T(1) * (l + b[c]));
}

// Two cells of one colour at once: both cells' operands are loaded before either
// is stored (within a colour every neighbour read belongs to the other colour).
template <typename T>
__device__ __forceinline__ void gs_item2(T *u, const T *b, const Lvl &L, const BcCoef &bc, int par, int64_t t0,
                                         int64_t t1, bool has1)
{
    int64_t ti0, j0, k0, ti1, j1, k1;
    split3(t0, (uint32_t)(L.n[0] / 2), (uint32_t)L.n[1], ti0, j0, k0);
    split3(has1 ? t1 : t0, (uint32_t)(L.n[0] / 2), (uint32_t)L.n[1], ti1, j1, k1);
    int64_t i0 = 2 * ti0 + ((j0 + k0 + par) & 1), i1 = 2 * ti1 + ((j1 + k1 + par) & 1);
    int64_t c0 = g_index(L, i0 + 1, j0 + 1, k0 + 1), c1 = g_index(L, i1 + 1, j1 + 1, k1 + 1);
    T u0 = u[c0], u1 = u[c1], b0 = b[c0], b1 = b[c1];
    T l0 = lap_fold(u, L, bc, i0, j0, k0, c0, u0);
    T l1 = lap_fold(u, L, bc, i1, j1, k1, c1, u1);
    u[c0] = u0 + div6(T(1) * (l0 + b0));
    if (has1) u[c1] = u1 + div6(T(1) * (l1 + b1));
}

// Ghost shell item: x faces (full y-z extent), then y faces (interior x, full z),
// then z faces (interior x, y).
template <typename T>
__device__ __forceinline__ void bc_item(T *a, const Lvl &L, const BcCoef &bc, int64_t t)
{
    const int64_t nx = L.n[0], ny = L.n[1], nz = L.n[2];
    int64_t i, j, k;
    int64_t nA = 2 * (ny + 2) * (nz + 2), nB = 2 * nx * (nz + 2);
    int64_t ua, ub, side;
    if (t < nA) {
        split3(t, (uint32_t)(ny + 2), (uint32_t)(nz + 2), ua, ub, side);
        i = side ? nx + 1 : 0;
        j = ua;
        k = ub;
    } else if (t < nA + nB) {
        split3(t - nA, (uint32_t)nx, (uint32_t)(nz + 2), ua, ub, side);
        j = side ? ny + 1 : 0;
        i = 1 + ua;
        k = ub;
    } else {
        split3(t - nA - nB, (uint32_t)nx, (uint32_t)ny, ua, ub, side);
        k = side ? nz + 1 : 0;
        i = 1 + ua;
        j = 1 + ub;
    }
    uint32_t flip = 0;
    int64_t ci = i, cj = j, ck = k;
    if (i == 0) { flip ^= bc.neg & 1u; ci = 1; } else if (i == nx + 1) { flip ^= (bc.neg >> 3) & 1u; ci = nx; }
    if (j == 0) { flip ^= (bc.neg >> 1) & 1u; cj = 1; } else if (j == ny + 1) { flip ^= (bc.neg >> 4) & 1u; cj = ny; }
    if (k == 0) { flip ^= (bc.neg >> 2) & 1u; ck = 1; } else if (k == nz + 1) { flip ^= (bc.neg >> 5) & 1u; ck = nz; }
    const T v = a[g_index(L, ci, cj, ck)];
    a[g_index(L, i, j, k)] = flip ? -v : v;
}

__device__ __forceinline__ int64_t bc_count(const Lvl &L)
{
    const int64_t nx = L.n[0], ny = L.n[1], nz = L.n[2];
    return 2 * (ny + 2) * (nz + 2) + 2 * nx * (nz + 2) + 2 * nx * ny;
}

// residual! (Multigrid.jl:186-193) at the 8 children of coarse cell t, summed in
// cubestencil order (x offset slowest) and restricted as 4*sum/8 (restrict!,
// Multigrid.jl:213-226).  The fine residual itself is not stored.
template <typename T>
__device__ __forceinline__ void resid_restrict_item(T *rh, const T *u, const T *b, T s, const Lvl &F,
                                                    const BcCoef &bc, int64_t t)
{
    int64_t ic, jc, kc;
    split3(t, (uint32_t)(F.n[0] / 2), (uint32_t)(F.n[1] / 2), ic, jc, kc);
    T sum = T(0);
#pragma unroll
    for (int dx = 0; dx < 2; ++dx)
#pragma unroll
        for (int dy = 0; dy < 2; ++dy)
#pragma unroll
            for (int dz = 0; dz < 2; ++dz) {
                int64_t i = 2 * ic + dx, j = 2 * jc + dy, k = 2 * kc + dz;
                int64_t c = g_index(F, i + 1, j + 1, k + 1);
                T uc = u[c];
                T l = lap_fold(u, F, bc, i, j, k, c, uc);
                sum += s * b[c] + l;
            }
    Lvl C;
    C.n[0] = F.n[0] / 2;
    C.n[1] = F.n[1] / 2;
    C.n[2] = F.n[2] / 2;
    rh[g_index(C, ic + 1, jc + 1, kc + 1)] = (T(4) * sum) / T(8);
}

// restrict! of a stored fine array.
template <typename T>
__device__ __forceinline__ void restrict_item(T *rh, const T *r, const Lvl &F, int64_t t)
{
    const int64_t nc0 = F.n[0] / 2, nc1 = F.n[1] / 2;
    int64_t i, j, k;
    split3(t, (uint32_t)nc0, (uint32_t)nc1, i, j, k);
    const int64_t fy = F.n[0] + 2, fz = (int64_t)(F.n[0] + 2) * (F.n[1] + 2);
    const int64_t cy = nc0 + 2, cz = (int64_t)(nc0 + 2) * (nc1 + 2);
    int64_t f = (2 * i + 1) + fy * (2 * j + 1) + fz * (2 * k + 1);
    T s = T(0);
    s += r[f];
    s += r[f + fz];
    s += r[f + fy];
    s += r[f + fy + fz];
    s += r[f + 1];
    s += r[f + 1 + fz];
    s += r[f + 1 + fy];
    s += r[f + 1 + fy + fz];
    rh[(i + 1) + cy * (j + 1) + cz * (k + 1)] = (T(4) * s) / T(8);
}

// interpolate! (Multigrid.jl:232-258): weights prod(3 - 2 sh_i)/64 over the coarse
// neighbours towards each fine child, sh in cubestencil order; add or overwrite.
template <typename T>
__device__ __forceinline__ void interp_item(T *r, const T *rh, const Lvl &F, int add, int64_t t)
{
    const int64_t nc0 = F.n[0] / 2, nc1 = F.n[1] / 2;
    int64_t i, j, k;
    split3(t, (uint32_t)nc0, (uint32_t)nc1, i, j, k);
    const int64_t fy = F.n[0] + 2, fz = (int64_t)(F.n[0] + 2) * (F.n[1] + 2);
    const int64_t cy = nc0 + 2, cz = (int64_t)(nc0 + 2) * (nc1 + 2);
    int64_t cc = (i + 1) + cy * (j + 1) + cz * (k + 1);
    const T w0 = T(27) / T(64), w1 = T(9) / T(64), w2 = T(3) / T(64), w3 = T(1) / T(64);
#pragma unroll
    for (int sx = 0; sx < 2; ++sx)
#pragma unroll
        for (int sy = 0; sy < 2; ++sy)
#pragma unroll
            for (int sz = 0; sz < 2; ++sz) {
                int64_t fidx = (2 * i + 1 + sx) + fy * (2 * j + 1 + sy) + fz * (2 * k + 1 + sz);
                int64_t ox = 2 * sx - 1, oy = (2 * sy - 1) * cy, oz = (2 * sz - 1) * cz;
                T v = T(0);
                v += w0 * rh[cc];
                v += w1 * rh[cc + oz];
                v += w1 * rh[cc + oy];
                v += w2 * rh[cc + oy + oz];
                v += w1 * rh[cc + ox];
                v += w2 * rh[cc + ox + oz];
                v += w2 * rh[cc + ox + oy];
                v += w3 * rh[cc + ox + oy + oz];
                if (add)
                    r[fidx] += v;
                else
                    r[fidx] = v;
            }
}

// ---- grid kernels ------------------------------------------------------------

template <typename T>
__global__ __launch_bounds__(MCS_BLOCK) void k_mg_bc(T *a, Lvl L, BcCoef bc)
{
    int64_t t = gidx();
    if (t < bc_count(L)) bc_item(a, L, bc, t);
}

template <typename T>
__global__ __launch_bounds__(MCS_BLOCK) void k_mg_rbgs(T *u, const T *b, Lvl L, BcCoef bc, int par)
{
    int64_t t = xcd_gidx();
    if (t < (int64_t)(L.n[0] / 2) * L.n[1] * L.n[2]) gs_item(u, b, L, bc, par, t);
}

// Also zeroes the interior of the coarse solution xh that the next mgv! starts
// from (its ghosts are always materialised before they are read).
template <typename T>
__global__ __launch_bounds__(MCS_BLOCK) void k_mg_resid_restrict(T *rh, const T *u, const T *b, Lvl F, BcCoef bc,
                                                                 T *xh)
{
    int64_t t = xcd_gidx();
    if (t < (int64_t)(F.n[0] / 2) * (F.n[1] / 2) * (F.n[2] / 2)) {
        resid_restrict_item(rh, u, b, T(1), F, bc, t);
        int64_t ic, jc, kc;
        split3(t, (uint32_t)(F.n[0] / 2), (uint32_t)(F.n[1] / 2), ic, jc, kc);
        Lvl C;
        C.n[0] = F.n[0] / 2;
        C.n[1] = F.n[1] / 2;
        C.n[2] = F.n[2] / 2;
        xh[g_index(C, ic + 1, jc + 1, kc + 1)] = T(0);
    }
}

// residual! stored on the interior, with per-block sums of r^2 (fmg!'s norm).
__global__ __launch_bounds__(MCS_BLOCK) void k_mg_residual(double *r, const double *u, const double *b,
                                                           double s, Lvl L, BcCoef bc, double *partial)
{
    int64_t t = xcd_gidx();
    int64_t N = (int64_t)L.n[0] * L.n[1] * L.n[2];
    double sq = 0.0;
    if (t < N) {
        int64_t i, j, k;
        split3(t, (uint32_t)L.n[0], (uint32_t)L.n[1], i, j, k);
        int64_t c = g_index(L, i + 1, j + 1, k + 1);
        double uc = u[c];
        double rv = s * b[c] + lap_fold(u, L, bc, i, j, k, c, uc);
        r[c] = rv;
        sq = rv * rv;
    }
    __shared__ double sh[MCS_BLOCK];
    sh[threadIdx.x] = sq;
    __syncthreads();
    for (int o = MCS_BLOCK / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[xcd_block()] = sh[0];
}

__global__ __launch_bounds__(MCS_BLOCK) void k_mg_sumsq(const double *b, Lvl L, double *partial)
{
    int64_t t = gidx();
    int64_t N = (int64_t)L.n[0] * L.n[1] * L.n[2];
    double sq = 0.0;
    if (t < N) {
        int64_t i, j, k;
        split3(t, (uint32_t)L.n[0], (uint32_t)L.n[1], i, j, k);
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

template <typename T>
__global__ __launch_bounds__(MCS_BLOCK) void k_mg_restrict(T *rh, const T *r, Lvl F)
{
    int64_t t = gidx();
    if (t < (int64_t)(F.n[0] / 2) * (F.n[1] / 2) * (F.n[2] / 2)) restrict_item(rh, r, F, t);
}

template <typename T>
__global__ __launch_bounds__(MCS_BLOCK) void k_mg_interp(T *r, const T *rh, Lvl F, int add)
{
    int64_t t = xcd_gidx();
    if (t < (int64_t)(F.n[0] / 2) * (F.n[1] / 2) * (F.n[2] / 2)) interp_item(r, rh, F, add, t);
}

// Coarse direct solve of fmg! (Multigrid.jl:404-410): z = A^-1 btop, one row per
// thread, columns summed in order.  btop is staged in LDS (at most 4096 cells).
template <typename T>
__global__ __launch_bounds__(MCS_BLOCK) void k_mg_coarse(const T *ainv, const T *res, T *z, Lvl C)
{
    __shared__ T bt[4096];
    const int64_t N = (int64_t)C.n[0] * C.n[1] * C.n[2];
    for (int64_t col = threadIdx.x; col < N; col += blockDim.x) {
        int64_t i, j, k;
        split3(col, (uint32_t)C.n[0], (uint32_t)C.n[1], i, j, k);
        bt[col] = res[g_index(C, i + 1, j + 1, k + 1)];
    }
    __syncthreads();
    for (int64_t row = threadIdx.x; row < N; row += blockDim.x) {
        T acc = T(0);
        for (int64_t col = 0; col < N; ++col) acc += ainv[row * N + col] * bt[col];
        int64_t i, j, k;
        split3(row, (uint32_t)C.n[0], (uint32_t)C.n[1], i, j, k);
        z[g_index(C, i + 1, j + 1, k + 1)] = acc;
    }
}

template <typename T>
__global__ void k_mg_fill0(T *a, int64_t n)
{
    int64_t i = gidx();
    if (i < n) a[i] = T(0);
}

// FP64 residual -> FP32 right-hand side of the correction equation (whole array).
__global__ void k_mg_to32(float *dst, const double *src, int64_t n)
{
    int64_t i = gidx();
    if (i < n) dst[i] = (float)src[i];
}

// x += z on the interior with an FP32 correction.
__global__ __launch_bounds__(MCS_BLOCK) void k_mg_axpy32(double *x, const float *z, Lvl L)
{
    int64_t t = gidx();
    if (t >= (int64_t)L.n[0] * L.n[1] * L.n[2]) return;
    int64_t i, j, k;
    split3(t, (uint32_t)L.n[0], (uint32_t)L.n[1], i, j, k);
    int64_t c = g_index(L, i + 1, j + 1, k + 1);
    x[c] += (double)z[c];
}

// x += z on the interior (the ghosts are rewritten by the apply! that follows).
__global__ __launch_bounds__(MCS_BLOCK) void k_mg_axpy(double *x, const double *z, Lvl L)
{
    int64_t t = gidx();
    if (t >= (int64_t)L.n[0] * L.n[1] * L.n[2]) return;
    int64_t i, j, k;
        split3(t, (uint32_t)L.n[0], (uint32_t)L.n[1], i, j, k);
    int64_t c = g_index(L, i + 1, j + 1, k + 1);
    x[c] += z[c];
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

// ---- single-workgroup tail: mgv! from a small level down to the coarsest -----

struct Tail {
    Lvl n[MG_MAXL];
    int64_t len[MG_MAXL];
    int64_t off_x[MG_MAXL];   // LDS offsets (elements) of the level solutions
    int64_t off_b[MG_MAXL];   // and of their right-hand sides
    int L;
    int smooth1, smooth2;
    BcCoef bc;
};

#define BLOOP(t, N) for (int64_t t = threadIdx.x; t < (N); t += blockDim.x)

// The 50 (apply!, gauss_seidel!) rounds of the coarsest level (Multigrid.jl:340-346)
// with one cell per lane of the first wave, neighbours exchanged by shuffles; the
// arithmetic is that of gs_item.  The ghosts are materialised from the values at
// the start of the last round, as the last apply! leaves them.  Called by wave 0
// only; LDS accesses of one wave complete in order, the fences keep the compiler
// from reordering them.
template <typename T>
__device__ void coarsest_in_registers(T *x, const T *b, const Lvl &Lv, const BcCoef &bc)
{
    const int lane = threadIdx.x;
    const int nx = Lv.n[0], ny = Lv.n[1], nz = Lv.n[2];
    const int N = nx * ny * nz;
    const bool own = lane < N;
    int i = 0, j = 0, k = 0;
    if (own) {
        i = lane % nx;
        j = (lane / nx) % ny;
        k = lane / (nx * ny);
    }
    const int64_t c = g_index(Lv, i + 1, j + 1, k + 1);
    T u = own ? x[c] : T(0);
    const T bb = own ? b[c] : T(0);
    const int par_me = (i + j + k) & 1;
    const int sxp = (i + 1 < nx) ? lane + 1 : lane, sxm = (i > 0) ? lane - 1 : lane;
    const int syp = (j + 1 < ny) ? lane + nx : lane, sym = (j > 0) ? lane - nx : lane;
    const int szp = (k + 1 < nz) ? lane + nx * ny : lane, szm = (k > 0) ? lane - nx * ny : lane;
    for (int it = 0; it < 50; ++it) {
        if (it == 49) {
            if (own) x[c] = u;
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            for (int64_t t = lane; t < bc_count(Lv); t += MCS_WAVE) bc_item(x, Lv, bc, t);
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
        for (int par = 0; par < 2; ++par) {
            T vxp = __shfl(u, sxp, MCS_WAVE), vyp = __shfl(u, syp, MCS_WAVE);
            T vzp = __shfl(u, szp, MCS_WAVE), vxm = __shfl(u, sxm, MCS_WAVE);
            T vym = __shfl(u, sym, MCS_WAVE), vzm = __shfl(u, szm, MCS_WAVE);
            if (own && par_me == par) {
                T l = T(-6) * u;
                l += (i + 1 < nx) ? vxp : bcf(bc, 3, u);
                l += (j + 1 < ny) ? vyp : bcf(bc, 4, u);
                l += (k + 1 < nz) ? vzp : bcf(bc, 5, u);
                l += (i > 0) ? vxm : bcf(bc, 0, u);
                l += (j > 0) ? vym : bcf(bc, 1, u);
                l += (k > 0) ? vzm : bcf(bc, 2, u);
                u = u + div6(T(1) * (l + bb));
            }
        }
    }
    if (own) x[c] = u;
}

// Block-wide copy with four independent loads in flight per thread.
template <typename E>
__device__ __forceinline__ void tail_copy(E *dst, const E *src, int64_t n)
{
    const int64_t bd = blockDim.x;
    int64_t t = threadIdx.x;
    for (; t + 3 * bd < n; t += 4 * bd) {
        const E a0 = src[t], a1 = src[t + bd], a2 = src[t + 2 * bd], a3 = src[t + 3 * bd];
        dst[t] = a0;
        dst[t + bd] = a1;
        dst[t + 2 * bd] = a2;
        dst[t + 3 * bd] = a3;
    }
    for (; t < n; t += bd) dst[t] = src[t];
}

// The coarsest rounds when the level has at most 16 cells and power-of-two
// strides: neighbours come from DPP row shifts within lanes 0-15 (row_shl:k gives
// lane l the value of lane l + k, row_shr:k that of lane l - k) instead of LDS
// permutes; the arithmetic is that of coarsest_in_registers.  SZ = 0: nz = 1.
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v)
{
    const uint64_t w = (uint64_t)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)w, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(w >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

template <int CTRL>
__device__ __forceinline__ float dpp_move(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}

template <typename T, int SY, int SZ>
__device__ void coarsest_dpp(T *x, const T *b, const Lvl &Lv, const BcCoef &bc)
{
    constexpr int ZP = SZ ? 0x100 + SZ : 0x101, ZM = SZ ? 0x110 + SZ : 0x111;
    const int lane = threadIdx.x;
    const int nx = Lv.n[0], ny = Lv.n[1], nz = Lv.n[2];
    const int N = nx * ny * nz;
    const bool own = lane < N;
    int i = 0, j = 0, k = 0;
    if (own) {
        i = lane % nx;
        j = (lane / nx) % ny;
        k = lane / (nx * ny);
    }
    const int64_t c = g_index(Lv, i + 1, j + 1, k + 1);
    T u = own ? x[c] : T(0);
    const T bb = own ? b[c] : T(0);
    const int par_me = (i + j + k) & 1;
    for (int it = 0; it < 50; ++it) {
        if (it == 49) {
            if (own) x[c] = u;
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            for (int64_t t = lane; t < bc_count(Lv); t += MCS_WAVE) bc_item(x, Lv, bc, t);
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
        for (int par = 0; par < 2; ++par) {
            const T vxp = dpp_move<0x101>(u), vxm = dpp_move<0x111>(u);
            const T vyp = dpp_move<0x100 + SY>(u), vym = dpp_move<0x110 + SY>(u);
            const T vzp = dpp_move<ZP>(u), vzm = dpp_move<ZM>(u);
            if (own && par_me == par) {
                T l = T(-6) * u;
                l += (i + 1 < nx) ? vxp : bcf(bc, 3, u);
                l += (j + 1 < ny) ? vyp : bcf(bc, 4, u);
                l += (k + 1 < nz) ? vzp : bcf(bc, 5, u);
                l += (i > 0) ? vxm : bcf(bc, 0, u);
                l += (j > 0) ? vym : bcf(bc, 1, u);
                l += (k > 0) ? vzm : bcf(bc, 2, u);
                u = u + div6(T(1) * (l + bb));
            }
        }
    }
    if (own) x[c] = u;
}

// Picks the DPP form for levels of at most 16 cells with power-of-two x and y
// strides; returns false when it does not apply.
template <typename T>
__device__ bool coarsest_try_dpp(T *x, const T *b, const Lvl &Lv, const BcCoef &bc)
{
    const int nx = Lv.n[0], ny = Lv.n[1], nz = Lv.n[2];
    if (nx * ny * nz > 16) return false;
    const int sz = nz > 1 ? nx * ny : 0;
    if (nx == 2 && sz == 0) { coarsest_dpp<T, 2, 0>(x, b, Lv, bc); return true; }
    if (nx == 2 && sz == 2) { coarsest_dpp<T, 2, 2>(x, b, Lv, bc); return true; }
    if (nx == 2 && sz == 4) { coarsest_dpp<T, 2, 4>(x, b, Lv, bc); return true; }
    if (nx == 2 && sz == 8) { coarsest_dpp<T, 2, 8>(x, b, Lv, bc); return true; }
    if (nx == 4 && sz == 0) { coarsest_dpp<T, 4, 0>(x, b, Lv, bc); return true; }
    if (nx == 4 && sz == 4) { coarsest_dpp<T, 4, 4>(x, b, Lv, bc); return true; }
    if (nx == 4 && sz == 8) { coarsest_dpp<T, 4, 8>(x, b, Lv, bc); return true; }
    if (nx == 8 && sz == 0) { coarsest_dpp<T, 8, 0>(x, b, Lv, bc); return true; }
    if (nx == 8 && sz == 8) { coarsest_dpp<T, 8, 8>(x, b, Lv, bc); return true; }
    return false;
}

// mgv! (Multigrid.jl:334-377) for levels start..L inside one workgroup: the level
// solutions and the deeper right-hand sides live in LDS, the start level's
// right-hand side is read from global memory; x0 is loaded and written back.
// Phases are separated by barriers and follow the grid path's order exactly.
template <typename E>
__global__ __launch_bounds__(MG_TAIL_THREADS) void k_mg_tail(Tail T, int start, E *x0, const E *b0)
{
    extern __shared__ double lds_raw[];
    E *lds = (E *)lds_raw;
    const BcCoef bc = T.bc;
    tail_copy(lds + T.off_x[start], x0, T.len[start]);
    __syncthreads();
    for (int lev = start; lev < T.L; ++lev) {
        E *x = lds + T.off_x[lev];
        const E *b = lev == start ? b0 : lds + T.off_b[lev];
        const Lvl Lv = T.n[lev];
        const int64_t ncol = (int64_t)(Lv.n[0] / 2) * Lv.n[1] * Lv.n[2];
        for (int it = 0; it < T.smooth1; ++it) {
            for (int par = 0; par < 2; ++par) {
                for (int64_t t = threadIdx.x; t < ncol; t += 2 * (int64_t)blockDim.x)
                    gs_item2(x, b, Lv, bc, par, t, t + blockDim.x, t + blockDim.x < ncol);
                __syncthreads();
            }
        }
        if (T.smooth2 == 0) {
            BLOOP(t, bc_count(Lv)) bc_item(x, Lv, bc, t);
            __syncthreads();
        }
        const int64_t nc = (int64_t)(Lv.n[0] / 2) * (Lv.n[1] / 2) * (Lv.n[2] / 2);
        BLOOP(t, nc) resid_restrict_item(lds + T.off_b[lev + 1], x, b, E(1), Lv, bc, t);
        BLOOP(t, T.len[lev + 1]) lds[T.off_x[lev + 1] + t] = E(0);
        __syncthreads();
    }
    {
        const int lev = T.L;
        E *x = lds + T.off_x[lev];
        const E *b = lev == start ? b0 : lds + T.off_b[lev];
        const Lvl Lv = T.n[lev];
        const int64_t N = (int64_t)Lv.n[0] * Lv.n[1] * Lv.n[2];
        if (N <= MCS_WAVE) {
            if (threadIdx.x < MCS_WAVE && !coarsest_try_dpp(x, b, Lv, bc)) coarsest_in_registers(x, b, Lv, bc);
            __syncthreads();
        } else {
            const int64_t ncol = (int64_t)(Lv.n[0] / 2) * Lv.n[1] * Lv.n[2];
            for (int it = 0; it < 50; ++it) {
                if (it == 49) {
                    BLOOP(t, bc_count(Lv)) bc_item(x, Lv, bc, t);
                    __syncthreads();
                }
                for (int par = 0; par < 2; ++par) {
                    BLOOP(t, ncol) gs_item(x, b, Lv, bc, par, t);
                    __syncthreads();
                }
            }
        }
    }
    for (int lev = T.L - 1; lev >= start; --lev) {
        E *x = lds + T.off_x[lev];
        const E *b = lev == start ? b0 : lds + T.off_b[lev];
        const Lvl Lv = T.n[lev];
        const int64_t nc = (int64_t)(Lv.n[0] / 2) * (Lv.n[1] / 2) * (Lv.n[2] / 2);
        BLOOP(t, nc) interp_item(x, lds + T.off_x[lev + 1], Lv, 1, t);
        __syncthreads();
        const int64_t ncol = (int64_t)(Lv.n[0] / 2) * Lv.n[1] * Lv.n[2];
        for (int it = 0; it < T.smooth2; ++it) {
            if (it == T.smooth2 - 1) {
                BLOOP(t, bc_count(Lv)) bc_item(x, Lv, bc, t);
                __syncthreads();
            }
            for (int par = 0; par < 2; ++par) {
                for (int64_t t = threadIdx.x; t < ncol; t += 2 * (int64_t)blockDim.x)
                    gs_item2(x, b, Lv, bc, par, t, t + blockDim.x, t + blockDim.x < ncol);
                __syncthreads();
            }
        }
    }
    tail_copy(x0, lds + T.off_x[start], T.len[start]);
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
    int lds_max = 0;
    if (hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device) !=
        hipSuccess)
        lds_max = 64 * 1024;
    size_t tail_lds = (size_t)lds_max < (size_t)MG_TAIL_LDS ? (size_t)lds_max : (size_t)MG_TAIL_LDS;
    if (tail_lds > 64 * 1024 &&
        (hipFuncSetAttribute((const void *)k_mg_tail<double>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)tail_lds) != hipSuccess ||
         hipFuncSetAttribute((const void *)k_mg_tail<float>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)tail_lds) != hipSuccess))
        tail_lds = 64 * 1024;
    if (!ctx || !cfg || !out) return MCS_E_ARG;
    const mcs_mg_config &c = *cfg;
    if (c.levels < 0 || c.levels >= MG_MAXL || c.smooth1 < 0 || c.smooth2 < 0 || c.maxiter < 0)
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
    if (c.precision != MCS_PREC_F64 && c.precision != MCS_PREC_F32)
        return mcs_fail(ctx, MCS_E_ARG, "mg: unknown precision %d", c.precision);
    const bool f32 = c.precision == MCS_PREC_F32;
    hipSetDevice(ctx->device);
    mcs_mg *mg = new mcs_mg();
    mg->ctx = ctx;
    mg->cfg = c;
    mg->tail_lds = tail_lds;
    const char *ge = getenv("MCS_MG_GRAPH");
    mg->use_graph = !(ge && ge[0] == '0');
    for (int l = 0; l <= c.levels; ++l) {
        MgLevel L;
        for (int d = 0; d < 3; ++d) L.n[d] = c.n[d] >> l;
        L.len = glen(L.n);
        L.fres = nullptr;
        L.fsol = nullptr;
        L.res1 = nullptr;
        if (hipMalloc(&L.res, sizeof(double) * L.len) != hipSuccess ||
            hipMalloc(&L.sol, sizeof(double) * L.len) != hipSuccess)
            return mcs_fail(ctx, MCS_E_HIP, "mg: device allocation failed");
        hipMemset(L.res, 0, sizeof(double) * L.len);
        hipMemset(L.sol, 0, sizeof(double) * L.len);
        if (f32) {
            if (hipMalloc(&L.fres, sizeof(float) * L.len) != hipSuccess ||
                hipMalloc(&L.fsol, sizeof(float) * L.len) != hipSuccess)
                return mcs_fail(ctx, MCS_E_HIP, "mg: device allocation failed");
            hipMemset(L.fres, 0, sizeof(float) * L.len);
            hipMemset(L.fsol, 0, sizeof(float) * L.len);
        }
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
    if (f32) {
        std::vector<float> a32(ainv.begin(), ainv.end());
        if (hipMalloc(&mg->ainv32, sizeof(float) * a32.size()) != hipSuccess)
            return mcs_fail(ctx, MCS_E_HIP, "mg: device allocation failed");
        hipMemcpy(mg->ainv32, a32.data(), sizeof(float) * a32.size(), hipMemcpyHostToDevice);
    }
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
        hipFree(L.sol);
        if (L.fres) hipFree(L.fres);
        if (L.fsol) hipFree(L.fsol);
    }
    hipFree(mg->ainv);
    if (mg->ainv32) hipFree(mg->ainv32);
    hipFree(mg->partial);
    if (mg->u_tmp) hipFree(mg->u_tmp);
    if (mg->b_tmp) hipFree(mg->b_tmp);
    if (mg->bc_dev) hipFree(mg->bc_dev);
    if (mg->graph_exec) hipGraphExecDestroy(mg->graph_exec);
    delete mg;
    return MCS_OK;
}

static inline Lvl lvl(const mcs_mg *mg, int l)
{
    Lvl L;
    for (int d = 0; d < 3; ++d) L.n[d] = mg->lv[l].n[d];
    return L;
}

static inline BcCoef coefs(const mcs_mg *mg)
{
    BcCoef c;
    c.neg = 0;
    for (int d = 0; d < 3; ++d) {
        if (mg->cfg.bc[2 * d] == MCS_BC_DIRICHLET) c.neg |= 1u << d;
        if (mg->cfg.bc[2 * d + 1] == MCS_BC_DIRICHLET) c.neg |= 1u << (3 + d);
    }
    return c;
}

static inline int64_t ncells(const mcs_mg *mg, int l)
{
    return (int64_t)mg->lv[l].n[0] * mg->lv[l].n[1] * mg->lv[l].n[2];
}

static inline int64_t ncoarse_cells(const mcs_mg *mg, int lf)
{
    return (int64_t)(mg->lv[lf].n[0] / 2) * (mg->lv[lf].n[1] / 2) * (mg->lv[lf].n[2] / 2);
}

static inline dim3 grid_for(int64_t n) { return dim3((unsigned)cdiv(n > 0 ? n : 1, MCS_BLOCK)); }

// The level arrays of one precision.
template <typename T> static T *lres(mcs_mg *mg, int l);
template <typename T> static T *lsol(mcs_mg *mg, int l);
template <typename T> static const T *coarse_inv(mcs_mg *mg);
template <> double *lres<double>(mcs_mg *mg, int l) { return mg->lv[l].res; }
template <> double *lsol<double>(mcs_mg *mg, int l) { return mg->lv[l].sol; }
template <> const double *coarse_inv<double>(mcs_mg *mg) { return mg->ainv; }
template <> float *lres<float>(mcs_mg *mg, int l) { return mg->lv[l].fres; }
template <> float *lsol<float>(mcs_mg *mg, int l) { return mg->lv[l].fsol; }
template <> const float *coarse_inv<float>(mcs_mg *mg) { return mg->ainv32; }

template <typename T>
static void mg_bc(mcs_mg *mg, int l, T *a)
{
    Lvl L = lvl(mg, l);
    int64_t nx = L.n[0], ny = L.n[1], nz = L.n[2];
    int64_t cnt = 2 * (ny + 2) * (nz + 2) + 2 * nx * (nz + 2) + 2 * nx * ny;
    hipLaunchKernelGGL(k_mg_bc<T>, grid_for(cnt), dim3(MCS_BLOCK), 0, mg->ctx->stream, a, L, coefs(mg));
}

template <typename T>
static void mg_gs(mcs_mg *mg, int l, T *u, const T *b)
{
    Lvl L = lvl(mg, l);
    int64_t cnt = (int64_t)(L.n[0] / 2) * L.n[1] * L.n[2];
    ProfScope ps(mg->ctx, "mg_rbgs");
    for (int par = 0; par < 2; ++par)
        hipLaunchKernelGGL(k_mg_rbgs<T>, grid_for(cnt), dim3(MCS_BLOCK), 0, mg->ctx->stream, u, b, L, coefs(mg),
                           par);
}

template <typename T>
static void mg_fill0(mcs_mg *mg, T *a, int64_t n)
{
    hipLaunchKernelGGL(k_mg_fill0<T>, grid_for(n), dim3(MCS_BLOCK), 0, mg->ctx->stream, a, n);
}

template <typename T>
static void mg_interp(mcs_mg *mg, int lf, T *r, const T *rh, int add)
{
    ProfScope ps(mg->ctx, "mg_prolong");
    hipLaunchKernelGGL(k_mg_interp<T>, grid_for(ncoarse_cells(mg, lf)), dim3(MCS_BLOCK), 0, mg->ctx->stream, r,
                       rh, lvl(mg, lf), add);
}

static int32_t ensure_bc_dev(mcs_mg *mg)
{
    if (mg->bc_dev) return MCS_OK;
    MCS_HIP(mg->ctx, hipMalloc(&mg->bc_dev, sizeof(int32_t) * 6));
    MCS_HIP(mg->ctx, hipMemcpy(mg->bc_dev, mg->cfg.bc, sizeof(int32_t) * 6, hipMemcpyHostToDevice));
    return MCS_OK;
}

// LDS bytes of a tail from `level`: the solutions of levels level..L and the
// right-hand sides of levels level+1..L.
template <typename T>
static size_t tail_bytes(const mcs_mg *mg, int level)
{
    size_t tot = (size_t)mg->lv[level].len * sizeof(T);
    for (int l = level + 1; l <= mg->cfg.levels; ++l) tot += 2 * (size_t)mg->lv[l].len * sizeof(T);
    return tot;
}

template <typename T>
static bool in_tail(const mcs_mg *mg, int level)
{
    return tail_bytes<T>(mg, level) <= mg->tail_lds;
}

template <typename T>
static void mg_tail(mcs_mg *mg, int level, T *x, const T *b)
{
    Tail Tl;
    memset(&Tl, 0, sizeof(Tl));
    Tl.L = mg->cfg.levels;
    Tl.smooth1 = mg->cfg.smooth1;
    Tl.smooth2 = mg->cfg.smooth2;
    Tl.bc = coefs(mg);
    int64_t off = 0;
    for (int l = 0; l <= Tl.L; ++l) {
        Tl.n[l] = lvl(mg, l);
        Tl.len[l] = mg->lv[l].len;
        if (l >= level) {
            Tl.off_x[l] = off;
            off += Tl.len[l];
            Tl.off_b[l] = off;
            if (l > level) off += Tl.len[l];
        }
    }
    ProfScope ps(mg->ctx, "mg_tail");
    hipLaunchKernelGGL(k_mg_tail<T>, dim3(1), dim3(MG_TAIL_THREADS), (size_t)off * sizeof(T), mg->ctx->stream, Tl,
                       level, x, b);
}

// mgv! (Multigrid.jl:334-377).  apply! before each smoothing sweep and before the
// residual is folded into the kernels; the ghosts are materialised before the
// last post-smoothing sweep (they are what the caller's prolongation reads).
template <typename T>
static void mgv(mcs_mg *mg, T *x, const T *b, int level)
{
    if (in_tail<T>(mg, level)) {
        mg_tail<T>(mg, level, x, b);
        return;
    }
    for (int it = 0; it < mg->cfg.smooth1; ++it) mg_gs<T>(mg, level, x, b);
    if (mg->cfg.smooth2 == 0) mg_bc<T>(mg, level, x);
    {
        ProfScope ps(mg->ctx, "mg_residual");
        hipLaunchKernelGGL(k_mg_resid_restrict<T>, grid_for(ncoarse_cells(mg, level)), dim3(MCS_BLOCK), 0,
                           mg->ctx->stream, lres<T>(mg, level + 1), (const T *)x, b, lvl(mg, level), coefs(mg),
                           lsol<T>(mg, level + 1));
    }
    T *xh = lsol<T>(mg, level + 1);
    mgv<T>(mg, xh, lres<T>(mg, level + 1), level + 1);
    mg_interp<T>(mg, level, x, xh, 1);
    for (int it = 0; it < mg->cfg.smooth2; ++it) {
        if (it == mg->cfg.smooth2 - 1) mg_bc<T>(mg, level, x);
        mg_gs<T>(mg, level, x, b);
    }
}

// The correction part of fmg! (Multigrid.jl:396-426) on the hierarchy of one
// precision, from the finest residual in lres<T>(0): restrictions, coarse solve
// and the coarse-to-fine sweep of V-cycles.  Returns the finest correction.
template <typename T>
static T *fmg_correction(mcs_mg *mg)
{
    mcs_ctx *ctx = mg->ctx;
    const int L = mg->cfg.levels;
    for (int k = 0; k < L; ++k) {
        ProfScope ps(ctx, "mg_restrict");
        hipLaunchKernelGGL(k_mg_restrict<T>, grid_for(ncoarse_cells(mg, k)), dim3(MCS_BLOCK), 0, ctx->stream,
                           lres<T>(mg, k + 1), (const T *)lres<T>(mg, k), lvl(mg, k));
    }
    hipLaunchKernelGGL(k_mg_coarse<T>, dim3(1), dim3(MCS_BLOCK), 0, ctx->stream, coarse_inv<T>(mg),
                       (const T *)lres<T>(mg, L), lsol<T>(mg, L), lvl(mg, L));
    T *z = lsol<T>(mg, L);
    for (int k = 1; k <= L; ++k) {
        int lev = L - k;
        T *z1 = lsol<T>(mg, lev);
        mg_bc<T>(mg, lev + 1, z);
        mg_interp<T>(mg, lev, z1, z, 0);
        mgv<T>(mg, z1, lres<T>(mg, lev), lev);
        z = z1;
    }
    return z;
}

// The part of fmg! after the convergence test, ending with x += z and apply!.
// No host interaction, so it is captured once into a HIP graph per (x, b) pair
// and replayed.  In precision mode the FP64 residual is rounded to FP32 first
// and the FP32 correction is added to the FP64 solution.
static void fmg_body(mcs_mg *mg, double *x, const double *b)
{
    mcs_ctx *ctx = mg->ctx;
    (void)b;
    if (mg->cfg.precision == MCS_PREC_F32) {
        hipLaunchKernelGGL(k_mg_to32, grid_for(mg->lv[0].len), dim3(MCS_BLOCK), 0, ctx->stream,
                           mg->lv[0].fres, (const double *)mg->lv[0].res, mg->lv[0].len);
        float *z = fmg_correction<float>(mg);
        hipLaunchKernelGGL(k_mg_axpy32, grid_for(ncells(mg, 0)), dim3(MCS_BLOCK), 0, ctx->stream, x,
                           (const float *)z, lvl(mg, 0));
    } else {
        double *z = fmg_correction<double>(mg);
        hipLaunchKernelGGL(k_mg_axpy, grid_for(ncells(mg, 0)), dim3(MCS_BLOCK), 0, ctx->stream, x,
                           (const double *)z, lvl(mg, 0));
    }
    mg_bc<double>(mg, 0, x);
}

// fmg! (Multigrid.jl:380-434).  Returns 1 when a correction was applied, 0 when
// the residual already meets the tolerance, negative on error.
static int fmg(mcs_mg *mg, double *x, const double *b, double *eps_out)
{
    mcs_ctx *ctx = mg->ctx;
    const MgLevel &F = mg->lv[0];
    mg_bc<double>(mg, 0, x);
    {
        ProfScope ps(ctx, "mg_residual");
        hipLaunchKernelGGL(k_mg_residual, grid_for(ncells(mg, 0)), dim3(MCS_BLOCK), 0, ctx->stream,
                           F.res, (const double *)x, b, mg->cfg.s, lvl(mg, 0), coefs(mg), mg->partial);
    }
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(MCS_BLOCK), 0, ctx->stream,
                       (const double *)mg->partial, mg->npartial, mg->partial + mg->npartial);
    double ss = 0.0;
    if (hipMemcpyAsync(ctx->h_small, mg->partial + mg->npartial, sizeof(double),
                       hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        return -1;
    memcpy(&ss, ctx->h_small, sizeof(double));
    double eps = sqrt(ss) / sqrt((double)F.n[0] * (double)F.n[1] * (double)F.n[2]);
    *eps_out = eps;
    if (mg->cfg.s * mg->tol_abs > eps) return 0;
    if (mg->use_graph) {
        if (!mg->graph_exec || mg->graph_x != x || mg->graph_b != b) {
            if (mg->graph_exec) {
                hipGraphExecDestroy(mg->graph_exec);
                mg->graph_exec = nullptr;
            }
            bool prof = ctx->prof;
            ctx->prof = false;
            hipGraph_t g = nullptr;
            if (hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed) != hipSuccess) return -1;
            fmg_body(mg, x, b);
            hipError_t e = hipStreamEndCapture(ctx->stream, &g);
            ctx->prof = prof;
            if (e != hipSuccess || hipGraphInstantiate(&mg->graph_exec, g, nullptr, nullptr, 0) != hipSuccess)
                return -1;
            hipGraphDestroy(g);
            mg->graph_x = x;
            mg->graph_b = b;
        }
        ProfScope ps(ctx, "mg_cycle");
        if (hipGraphLaunch(mg->graph_exec, ctx->stream) != hipSuccess) return -1;
        return 1;
    }
    ProfScope ps(ctx, "mg_cycle");
    fmg_body(mg, x, b);
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
