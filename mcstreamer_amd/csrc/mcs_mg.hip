// mcs_mg.hip -- geometric multigrid for L u + s b = 0 on a cell-centred 3D
// mesh with one ghost layer (Multigrid/src/Multigrid.jl, redblack.jl,
// boundaries.jl).  The host drives the operation sequence of fmg!/mgv!; each
// operation is a coalesced stencil kernel over one level, and the levels below
// a few thousand cells run as one single-workgroup kernel.
//
// Ghost handling.  apply! (boundaries.jl:110-133) sets every ghost plane to
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
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "mcs_device.h"

#define MG_MAXL 13
#define MG_TAIL_LDS (160 * 1024)
#define MG_TAIL_THREADS 512
// Measured on the 128x128x256 and 256x256x512 streamer solves: the
// single-workgroup tail pays off from levels of up to ~2^10 cells (it also needs
// the levels in LDS), the tiled smoother on levels of up to ~2^18 cells; the
// finer levels run one launch per pass.
#define MG_TAIL_MAXCELLS 1024
#define MG_SM_MAXCELLS ((int64_t)1 << 18)

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
// FP64 x / 6 is formed as q = x*c, r = fma(-q, 6, x), q + r*c with c = RN(1/6):
// by Markstein's theorem (q within 1 ulp of x/6, c within 1/2 ulp of 1/6) the
// result is the correctly rounded quotient, i.e. equal to x / 6.0, while r does
// not underflow; zeros, tiny and huge |x| and non-finite x take the division.
__device__ __forceinline__ double div6(double x)
{
    const double a = fabs(x);
    if (__builtin_expect(!(a > 0x1p-960 && a < 0x1p960), 0)) return x / 6.0;
    const double c = 1.0 / 6.0;
    const double q = x * c;
    return __builtin_fma(__builtin_fma(-q, 6.0, x), c, q);
}
__device__ __forceinline__ float div6(float x) { return x * (1.0f / 6.0f); }

__device__ __forceinline__ int64_t g_index(const Lvl &L, int64_t i, int64_t j, int64_t k)
{
    return i + (int64_t)(L.n[0] + 2) * (j + (int64_t)(L.n[1] + 2) * k);
}

// Colour-split index (the level-0 correction, see k_mg_rbgs_s): cell (i, j, k) of
// colour p = (i+j+k) % 2 at p*N/2 + (i>>1) + nx/2*(j + ny*k).
__device__ __forceinline__ int64_t s_index(const Lvl &L, int64_t i, int64_t j, int64_t k)
{
    const int64_t nxh = L.n[0] / 2;
    const int64_t half = nxh * L.n[1] * L.n[2];
    return ((i + j + k) & 1) * half + (i >> 1) + nxh * (j + (int64_t)L.n[1] * k);
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

// q = t / d and r = t - q d for 0 <= t < 2^24 and 1 <= d < 2^24, with rd = 1/d
// rounded to float: the float quotient is within one of the true one (t is
// exact in float and the two roundings are below 2^-23 relative), and one
// correction step makes it exact.
__device__ __forceinline__ int fdivmod(int t, int d, float rd, int &r)
{
    int q = (int)((float)t * rd);
    r = t - q * d;
    if (r < 0) {
        q -= 1;
        r += d;
    } else if (r >= d) {
        q += 1;
        r -= d;
    }
    return q;
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
__device__ __forceinline__ void interp_item(T *__restrict__ r, const T *__restrict__ rh, const Lvl &F, int add,
                                            int64_t t)
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

// bc_item for every ghost cell whose nearest interior cell is (i, j, k), which
// holds v: up to 7 ghost cells of a corner cell, none of an inner cell.
template <typename T>
__device__ __forceinline__ void ghost_write(T *a, const Lvl &L, const BcCoef &bc, int64_t i, int64_t j, int64_t k,
                                            T v)
{
    const int n0 = L.n[0], n1 = L.n[1], n2 = L.n[2];
    if (i > 0 && i + 1 < n0 && j > 0 && j + 1 < n1 && k > 0 && k + 1 < n2) return;
    // per axis: the cell itself, its low ghost, its high ghost (with their flips)
    const int xs[3] = {(int)i, -1, n0}, ys[3] = {(int)j, -1, n1}, zs[3] = {(int)k, -1, n2};
    const bool xo[3] = {true, i == 0, i + 1 == n0}, yo[3] = {true, j == 0, j + 1 == n1},
               zo[3] = {true, k == 0, k + 1 == n2};
    const uint32_t neg = bc.neg;
    const uint32_t fx[3] = {0u, neg & 1u, (neg >> 3) & 1u}, fy[3] = {0u, (neg >> 1) & 1u, (neg >> 4) & 1u},
                   fz[3] = {0u, (neg >> 2) & 1u, (neg >> 5) & 1u};
#pragma unroll
    for (int ax = 0; ax < 3; ++ax)
#pragma unroll
        for (int ay = 0; ay < 3; ++ay)
#pragma unroll
            for (int az = 0; az < 3; ++az) {
                if ((ax | ay | az) == 0 || !xo[ax] || !yo[ay] || !zo[az]) continue;
                const uint32_t flip = fx[ax] ^ fy[ay] ^ fz[az];
                a[g_index(L, (int64_t)xs[ax] + 1, (int64_t)ys[ay] + 1, (int64_t)zs[az] + 1)] = flip ? -v : v;
            }
}

// ---- grid kernels ------------------------------------------------------------

template <typename T>
__global__ __launch_bounds__(MCS_BLOCK) void k_mg_bc(T *a, Lvl L, BcCoef bc)
{
    int64_t t = gidx();
    if (t < bc_count(L)) bc_item(a, L, bc, t);
}

// ghosts: after the update, apply! of the level as it stands then (the pass's
// cell with its new value and its pair cell along x, of the other colour, which
// this pass does not change), folded into the pass instead of a launch.
template <typename T>
__global__ __launch_bounds__(MCS_BLOCK) void k_mg_rbgs(T *u, const T *b, Lvl L, BcCoef bc, int par, int ghosts)
{
    int64_t t = xcd_gidx();
    if (t >= (int64_t)(L.n[0] / 2) * L.n[1] * L.n[2]) return;
    if (!ghosts) {
        gs_item(u, b, L, bc, par, t);
        return;
    }
    int64_t ti, j, k;
    split3(t, (uint32_t)(L.n[0] / 2), (uint32_t)L.n[1], ti, j, k);
    const int64_t i = 2 * ti + ((j + k + par) & 1);
    const int64_t c = g_index(L, i + 1, j + 1, k + 1);
    const T uc = u[c];
    const T l = lap_fold(u, L, bc, i, j, k, c, uc);
    const T un = uc + div6(T(1) * (l + b[c]));
    u[c] = un;
    ghost_write(u, L, bc, i, j, k, un);
    const int64_t ip = i ^ 1;
    ghost_write(u, L, bc, ip, j, k, u[c + (ip - i)]);
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
// split: r is stored in the colour-split layout of the level-0 correction.
__global__ __launch_bounds__(MCS_BLOCK) void k_mg_residual(double *r, const double *u, const double *b,
                                                           double s, Lvl L, BcCoef bc, double *partial, int split)
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
        r[split ? s_index(L, i, j, k) : c] = rv;
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

// The sum of the partials to *out and, when pub is given, published to the host
// with sequence number seq.
__global__ void k_sum_partials(const double *partial, int64_t nb, double *out, uint64_t *pub, uint64_t seq)
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
        *out = sh[0];
        if (pub) {
            pub[1] = (uint64_t)__double_as_longlong(sh[0]);
            mcs_pub_raise(pub, seq);
        }
    }
}

template <typename T>
__global__ __launch_bounds__(MCS_BLOCK) void k_mg_restrict(T *rh, const T *r, Lvl F)
{
    int64_t t = gidx();
    if (t < (int64_t)(F.n[0] / 2) * (F.n[1] / 2) * (F.n[2] / 2)) restrict_item(rh, r, F, t);
}

template <typename T>
__global__ __launch_bounds__(MCS_BLOCK) void k_mg_interp(T *__restrict__ r, const T *__restrict__ rh, Lvl F, int add)
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
    for (int64_t col = threadIdx.x; col < N; col += MCS_BLOCK) {
        int64_t i, j, k;
        split3(col, (uint32_t)C.n[0], (uint32_t)C.n[1], i, j, k);
        bt[col] = res[g_index(C, i + 1, j + 1, k + 1)];
    }
    __syncthreads();
    for (int64_t row = threadIdx.x; row < N; row += MCS_BLOCK) {
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
// x .+= z; apply!(x) at the end of fmg!: a boundary cell also writes the ghost
// cells whose nearest interior cell it is, as bc_item would from the new value.
template <typename Z>
__global__ __launch_bounds__(MCS_BLOCK) void k_mg_axpy_bc(double *x, const Z *z, Lvl L, BcCoef bc, int split)
{
    int64_t t = xcd_gidx();
    if (t >= (int64_t)L.n[0] * L.n[1] * L.n[2]) return;
    int64_t i, j, k;
    split3(t, (uint32_t)L.n[0], (uint32_t)L.n[1], i, j, k);
    const int64_t c = g_index(L, i + 1, j + 1, k + 1);
    const double v = x[c] + (double)z[split ? s_index(L, i, j, k) : c];
    x[c] = v;
    ghost_write(x, L, bc, i, j, k, v);
}

// setinhom! (boundaries.jl:141-153): b += (2/s) val on the first valid layer.
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

// ---- colour-split layout of the level-0 correction ----------------------------
//
// The finest level's correction and right-hand side (private to the solver, their
// ghosts never read) are stored by colour: cell (i, j, k) of colour
// p = (i+j+k) % 2 at p*N/2 + (i>>1) + nx/2*(j + ny*k).  A red-black pass then reads
// and writes only its colour's half and the other colour's neighbours, instead
// of every line of both arrays; the arithmetic is unchanged.

// lap_fold on the split layout: o is the other colour's half, off the cell's
// offset within a half.  Out-of-level neighbours read the cell's own offset (a
// valid index) and are replaced by the ghost rule.
template <typename T>
__device__ __forceinline__ T lap_split(const T *__restrict__ o, const Lvl &L, const BcCoef &bc, int64_t i,
                                      int64_t j, int64_t k, int64_t off, T uc)
{
    const int64_t sy = L.n[0] / 2, sz = (int64_t)(L.n[0] / 2) * L.n[1];
    const bool xp = i + 1 < L.n[0], yp = j + 1 < L.n[1], zp = k + 1 < L.n[2], xm = i > 0, ym = j > 0, zm = k > 0;
    const int64_t odd = i & 1;
    const T vxp = o[xp ? off + odd : off], vyp = o[yp ? off + sy : off], vzp = o[zp ? off + sz : off];
    const T vxm = o[xm ? off - 1 + odd : off], vym = o[ym ? off - sy : off], vzm = o[zm ? off - sz : off];
    T l = T(-6) * uc;
    l += xp ? vxp : bcf(bc, 3, uc);
    l += yp ? vyp : bcf(bc, 4, uc);
    l += zp ? vzp : bcf(bc, 5, uc);
    l += xm ? vxm : bcf(bc, 0, uc);
    l += ym ? vym : bcf(bc, 1, uc);
    l += zm ? vzm : bcf(bc, 2, uc);
    return l;
}

// gs_item on the split layout: thread t is the t-th cell of colour par.
template <typename T>
__global__ __launch_bounds__(MCS_BLOCK) void k_mg_rbgs_s(T *u, const T *b, Lvl L, BcCoef bc, int par)
{
    const int64_t t = xcd_gidx();
    const int64_t half = (int64_t)(L.n[0] / 2) * L.n[1] * L.n[2];
    if (t >= half) return;
    int64_t ti, j, k;
    split3(t, (uint32_t)(L.n[0] / 2), (uint32_t)L.n[1], ti, j, k);
    const int64_t i = 2 * ti + ((j + k + par) & 1);
    T *up = u + par * half;
    const T uc = up[t];
    const T l = lap_split(u + (1 - par) * half, L, bc, i, j, k, t, uc);
    up[t] = uc + div6(T(1) * (l + b[par * half + t]));
}

// The same pass with two consecutive cells of a row per thread (a half row has
// an even count): the pair's own values, right-hand sides and y/z neighbours
// are aligned pairs loaded and stored together, the three x neighbours they
// share loaded once; each cell's arithmetic is lap_split's and gs_item's.
template <typename T>
struct Pair2;
template <>
struct Pair2<double> {
    using V = double2;
};
template <>
struct Pair2<float> {
    using V = float2;
};

template <typename T>
__global__ __launch_bounds__(MCS_BLOCK) void k_mg_rbgs_s2(T *u, const T *b, Lvl L, BcCoef bc, int par)
{
    using V = typename Pair2<T>::V;
    const int64_t half = (int64_t)(L.n[0] / 2) * L.n[1] * L.n[2];
    const int64_t t = 2 * xcd_gidx();
    if (t >= half) return;
    int64_t ti, j, k;
    split3(t, (uint32_t)(L.n[0] / 2), (uint32_t)L.n[1], ti, j, k);
    const int64_t odd = (j + k + par) & 1;
    const int64_t i0 = 2 * ti + odd, i1 = i0 + 2;
    const int64_t sy = L.n[0] / 2, sz = (int64_t)(L.n[0] / 2) * L.n[1];
    T *up = u + par * half;
    const T *o = u + (1 - par) * half;
    const V uc = *(const V *)(up + t);
    const V bb = *(const V *)(b + par * half + t);
    const bool yp = j + 1 < L.n[1], zp = k + 1 < L.n[2], ym = j > 0, zm = k > 0;
    const V vyp = *(const V *)(o + (yp ? t + sy : t)), vzp = *(const V *)(o + (zp ? t + sz : t));
    const V vym = *(const V *)(o + (ym ? t - sy : t)), vzm = *(const V *)(o + (zm ? t - sz : t));
    // x neighbours: cell 0 at offset t, cell 1 at t + 1 (lap_split's indices)
    const bool xp0 = i0 + 1 < L.n[0], xm0 = i0 > 0, xp1 = i1 + 1 < L.n[0], xm1 = i1 > 0;
    const T vxp0 = o[xp0 ? t + odd : t], vxm0 = o[xm0 ? t - 1 + odd : t];
    const T vxp1 = o[xp1 ? t + 1 + odd : t + 1], vxm1 = o[xm1 ? t + odd : t + 1];
    T l0 = T(-6) * uc.x;
    l0 += xp0 ? vxp0 : bcf(bc, 3, uc.x);
    l0 += yp ? vyp.x : bcf(bc, 4, uc.x);
    l0 += zp ? vzp.x : bcf(bc, 5, uc.x);
    l0 += xm0 ? vxm0 : bcf(bc, 0, uc.x);
    l0 += ym ? vym.x : bcf(bc, 1, uc.x);
    l0 += zm ? vzm.x : bcf(bc, 2, uc.x);
    T l1 = T(-6) * uc.y;
    l1 += xp1 ? vxp1 : bcf(bc, 3, uc.y);
    l1 += yp ? vyp.y : bcf(bc, 4, uc.y);
    l1 += zp ? vzp.y : bcf(bc, 5, uc.y);
    l1 += xm1 ? vxm1 : bcf(bc, 0, uc.y);
    l1 += ym ? vym.y : bcf(bc, 1, uc.y);
    l1 += zm ? vzm.y : bcf(bc, 2, uc.y);
    V r;
    r.x = uc.x + div6(T(1) * (l0 + bb.x));
    r.y = uc.y + div6(T(1) * (l1 + bb.y));
    *(V *)(up + t) = r;
}

// interp_item into a split fine array for two consecutive coarse cells of a row
// per thread (a coarse row has an even count): the two cells' children of one
// colour, row and plane sit side by side, so each of the eight (colour, y, z)
// child pairs is one paired store (and load when adding); each child's value is
// interp_item's sum, term by term.
template <typename T>
__global__ __launch_bounds__(MCS_BLOCK) void k_mg_interp_s2(T *__restrict__ r, const T *__restrict__ rh, Lvl F,
                                                            int add)
{
    using V = typename Pair2<T>::V;
    const int64_t nc0 = F.n[0] / 2, nc1 = F.n[1] / 2;
    const int64_t t = 2 * xcd_gidx();
    if (t >= nc0 * nc1 * (F.n[2] / 2)) return;
    int64_t i, j, k;
    split3(t, (uint32_t)nc0, (uint32_t)nc1, i, j, k);
    const int64_t cy = nc0 + 2, cz = (int64_t)(nc0 + 2) * (nc1 + 2);
    const int64_t cc = (i + 1) + cy * (j + 1) + cz * (k + 1);
    const T w0 = T(27) / T(64), w1 = T(9) / T(64), w2 = T(3) / T(64), w3 = T(1) / T(64);
#pragma unroll
    for (int sx = 0; sx < 2; ++sx)
#pragma unroll
        for (int sy = 0; sy < 2; ++sy)
#pragma unroll
            for (int sz = 0; sz < 2; ++sz) {
                const int64_t f = s_index(F, 2 * i + sx, 2 * j + sy, 2 * k + sz);
                const int64_t ox = 2 * sx - 1, oy = (2 * sy - 1) * cy, oz = (2 * sz - 1) * cz;
                T v[2];
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    const int64_t c = cc + e;
                    T a = T(0);
                    a += w0 * rh[c];
                    a += w1 * rh[c + oz];
                    a += w1 * rh[c + oy];
                    a += w2 * rh[c + oy + oz];
                    a += w1 * rh[c + ox];
                    a += w2 * rh[c + ox + oz];
                    a += w2 * rh[c + ox + oy];
                    a += w3 * rh[c + ox + oy + oz];
                    v[e] = a;
                }
                V out;
                if (add) {
                    const V old = *(const V *)(r + f);
                    out.x = old.x + v[0];
                    out.y = old.y + v[1];
                } else {
                    out.x = v[0];
                    out.y = v[1];
                }
                *(V *)(r + f) = out;
            }
}

// resid_restrict_item with the fine arrays split (the coarse ones keep ghosts).
template <typename T>
__global__ __launch_bounds__(MCS_BLOCK) void k_mg_resid_restrict_s(T *rh, const T *u, const T *b, Lvl F, BcCoef bc,
                                                                   T *xh)
{
    const int64_t t = xcd_gidx();
    const int64_t half = (int64_t)(F.n[0] / 2) * F.n[1] * F.n[2];
    if (t >= (int64_t)(F.n[0] / 2) * (F.n[1] / 2) * (F.n[2] / 2)) return;
    int64_t ic, jc, kc;
    split3(t, (uint32_t)(F.n[0] / 2), (uint32_t)(F.n[1] / 2), ic, jc, kc);
    T sum = T(0);
#pragma unroll
    for (int dx = 0; dx < 2; ++dx)
#pragma unroll
        for (int dy = 0; dy < 2; ++dy)
#pragma unroll
            for (int dz = 0; dz < 2; ++dz) {
                const int64_t i = 2 * ic + dx, j = 2 * jc + dy, k = 2 * kc + dz;
                const int64_t p = (i + j + k) & 1, off = (i >> 1) + (int64_t)(F.n[0] / 2) * (j + (int64_t)F.n[1] * k);
                const T uc = u[p * half + off];
                const T l = lap_split(u + (1 - p) * half, F, bc, i, j, k, off, uc);
                sum += T(1) * b[p * half + off] + l;
            }
    Lvl C;
    C.n[0] = F.n[0] / 2;
    C.n[1] = F.n[1] / 2;
    C.n[2] = F.n[2] / 2;
    const int64_t cidx = g_index(C, ic + 1, jc + 1, kc + 1);
    rh[cidx] = (T(4) * sum) / T(8);
    xh[cidx] = T(0);
}

// restrict_item of a split fine array.
template <typename T>
__global__ __launch_bounds__(MCS_BLOCK) void k_mg_restrict_s(T *rh, const T *r, Lvl F)
{
    const int64_t t = gidx();
    if (t >= (int64_t)(F.n[0] / 2) * (F.n[1] / 2) * (F.n[2] / 2)) return;
    int64_t i, j, k;
    split3(t, (uint32_t)(F.n[0] / 2), (uint32_t)(F.n[1] / 2), i, j, k);
    const int64_t cy = F.n[0] / 2 + 2, cz = (int64_t)(F.n[0] / 2 + 2) * (F.n[1] / 2 + 2);
    const int64_t x0 = 2 * i, y0 = 2 * j, z0 = 2 * k;
    T s = T(0);
    s += r[s_index(F, x0, y0, z0)];
    s += r[s_index(F, x0, y0, z0 + 1)];
    s += r[s_index(F, x0, y0 + 1, z0)];
    s += r[s_index(F, x0, y0 + 1, z0 + 1)];
    s += r[s_index(F, x0 + 1, y0, z0)];
    s += r[s_index(F, x0 + 1, y0, z0 + 1)];
    s += r[s_index(F, x0 + 1, y0 + 1, z0)];
    s += r[s_index(F, x0 + 1, y0 + 1, z0 + 1)];
    rh[(i + 1) + cy * (j + 1) + cz * (k + 1)] = (T(4) * s) / T(8);
}

// interp_item into a split fine array (add or overwrite).
template <typename T>
__global__ __launch_bounds__(MCS_BLOCK) void k_mg_interp_s(T *__restrict__ r, const T *__restrict__ rh, Lvl F,
                                                           int add)
{
    const int64_t t = xcd_gidx();
    if (t >= (int64_t)(F.n[0] / 2) * (F.n[1] / 2) * (F.n[2] / 2)) return;
    const int64_t nc0 = F.n[0] / 2, nc1 = F.n[1] / 2;
    int64_t i, j, k;
    split3(t, (uint32_t)nc0, (uint32_t)nc1, i, j, k);
    const int64_t cy = nc0 + 2, cz = (int64_t)(nc0 + 2) * (nc1 + 2);
    const int64_t cc = (i + 1) + cy * (j + 1) + cz * (k + 1);
    const T w0 = T(27) / T(64), w1 = T(9) / T(64), w2 = T(3) / T(64), w3 = T(1) / T(64);
#pragma unroll
    for (int sx = 0; sx < 2; ++sx)
#pragma unroll
        for (int sy = 0; sy < 2; ++sy)
#pragma unroll
            for (int sz = 0; sz < 2; ++sz) {
                const int64_t f = s_index(F, 2 * i + sx, 2 * j + sy, 2 * k + sz);
                const int64_t ox = 2 * sx - 1, oy = (2 * sy - 1) * cy, oz = (2 * sz - 1) * cz;
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
                    r[f] += v;
                else
                    r[f] = v;
            }
}

// ---- tiled smoother: one visit of mgv! to a grid level in two launches --------
//
// A workgroup owns a tile of the level and holds the tile grown by H cells (cut
// at the level's edges) in LDS.  Pass p of the P red-black passes updates its
// colour on the tile grown by H-1-p cells: by induction every update reads
// neighbours that are as the launch-per-pass path leaves them, and the outer layer
// is only read, so after the passes the tile grown by H-P cells holds exactly the
// global result.  The per-cell arithmetic is that of gs_item, resid_restrict_item,
// interp_item and bc_item, so results are bitwise those of one launch per pass.
// Input and output are distinct arrays: neighbouring tiles read each other's
// halos from the input while the outputs are written.

#define MG_SM_THREADS 1024
#define SM_LD 4
#define SM_GS 4

struct SmoothArgs {
    Lvl L;
    BcCoef bc;
    int32_t tile[3];     // tile size: even, divides L.n
    int32_t ntile[3];    // tiles per axis
    int32_t passes;      // red-black passes, colour 0 first (2 per sweep)
    int32_t halo;        // passes + resid
    int32_t ghost_pass;  // ghost shell of u_out from the state after this many passes; -1: none
    int32_t resid;       // after the passes: residual restricted into rh
    int32_t interp;      // start: 0 u_in, 1 u_in + interpolate!(xc), 2 interpolate!(xc), 3 zero
};

// Cells t, t + step, ... of a box (nx, ny, nz) in x-fastest order as a carried
// mixed-radix counter (no divisions inside the loops).
struct BoxWalk {
    int x, y, z, sx, sy, sz, nx, ny;
    __device__ __forceinline__ BoxWalk(int nx_, int ny_, int t, int step) : nx(nx_), ny(ny_)
    {
        const float rx = 1.0f / (float)nx, ry = 1.0f / (float)ny;
        int q = fdivmod(t, nx, rx, x);
        z = fdivmod(q, ny, ry, y);
        q = fdivmod(step, nx, rx, sx);
        sz = fdivmod(q, ny, ry, sy);
    }
    __device__ __forceinline__ void next()
    {
        x += sx;
        y += sy;
        z += sz;
        if (x >= nx) { x -= nx; ++y; }
        if (y >= ny) { y -= ny; ++z; }
    }
};

__device__ __forceinline__ int64_t g_at(const Lvl &L, int x, int y, int z)
{
    return g_index(L, (int64_t)x + 1, (int64_t)y + 1, (int64_t)z + 1);
}

// lap_fold on the LDS box: a neighbour outside the level reads the cell itself,
// so every LDS read stays inside the box, and the ghost rule is applied to it.
template <typename T>
__device__ __forceinline__ T lap_box(const T *s, int li, int sy, int sz, int x, int y, int z, const Lvl &L,
                                    const BcCoef &bc, T uc)
{
    const bool xp = x + 1 < L.n[0], yp = y + 1 < L.n[1], zp = z + 1 < L.n[2];
    const bool xm = x > 0, ym = y > 0, zm = z > 0;
    const T vxp = s[xp ? li + 1 : li], vyp = s[yp ? li + sy : li], vzp = s[zp ? li + sz : li];
    const T vxm = s[xm ? li - 1 : li], vym = s[ym ? li - sy : li], vzm = s[zm ? li - sz : li];
    T l = T(-6) * uc;
    l += xp ? vxp : bcf(bc, 3, vxp);
    l += yp ? vyp : bcf(bc, 4, vyp);
    l += zp ? vzp : bcf(bc, 5, vzp);
    l += xm ? vxm : bcf(bc, 0, vxm);
    l += ym ? vym : bcf(bc, 1, vym);
    l += zm ? vzm : bcf(bc, 2, vzm);
    return l;
}

// interp_item's value for fine cell (x, y, z) of level F from the coarse array rh.
template <typename T>
__device__ __forceinline__ T prolong_at(const T *rh, const Lvl &F, int x, int y, int z)
{
    const int64_t cy = F.n[0] / 2 + 2, cz = (int64_t)(F.n[0] / 2 + 2) * (F.n[1] / 2 + 2);
    const int64_t cc = (int64_t)((x >> 1) + 1) + cy * ((y >> 1) + 1) + cz * ((z >> 1) + 1);
    const int64_t ox = 2 * (x & 1) - 1, oy = (2 * (y & 1) - 1) * cy, oz = (2 * (z & 1) - 1) * cz;
    const T w0 = T(27) / T(64), w1 = T(9) / T(64), w2 = T(3) / T(64), w3 = T(1) / T(64);
    T v = T(0);
    v += w0 * rh[cc];
    v += w1 * rh[cc + oz];
    v += w1 * rh[cc + oy];
    v += w2 * rh[cc + oy + oz];
    v += w1 * rh[cc + ox];
    v += w2 * rh[cc + ox + oz];
    v += w2 * rh[cc + ox + oy];
    v += w3 * rh[cc + ox + oy + oz];
    return v;
}

struct SmBox {
    int t0[3], t1[3], a0[3], R[3];   // tile, and the box of u: the tile grown by H
    int b0[3], Rb[3];                // box of b: the tile grown by H - 1
};

// bc_item for the ghost cells whose nearest interior cell lies in this tile, from
// the LDS state.
template <typename T>
__device__ void sm_ghosts(const SmoothArgs &A, const SmBox &B, const T *s, T *u_out)
{
    const Lvl &L = A.L;
    int e0[3], E[3];
    bool touch = false;
    for (int d = 0; d < 3; ++d) {
        const bool lo = B.t0[d] == 0, hi = B.t1[d] == L.n[d];
        touch |= lo || hi;
        e0[d] = B.t0[d] - (lo ? 1 : 0);
        E[d] = B.t1[d] + (hi ? 1 : 0) - e0[d];
    }
    if (!touch) return;
    const int sy = B.R[0], sz = B.R[0] * B.R[1];
    const uint32_t neg = A.bc.neg;
    for (BoxWalk w(E[0], E[1], threadIdx.x, MG_SM_THREADS); w.z < E[2]; w.next()) {
        const int x = e0[0] + w.x, y = e0[1] + w.y, z = e0[2] + w.z;
        uint32_t flip = 0;
        int cx = x, cy = y, cz = z;
        bool shell = false;
        if (x < 0) { flip ^= neg & 1u; cx = 0; shell = true; }
        else if (x >= L.n[0]) { flip ^= (neg >> 3) & 1u; cx = L.n[0] - 1; shell = true; }
        if (y < 0) { flip ^= (neg >> 1) & 1u; cy = 0; shell = true; }
        else if (y >= L.n[1]) { flip ^= (neg >> 4) & 1u; cy = L.n[1] - 1; shell = true; }
        if (z < 0) { flip ^= (neg >> 2) & 1u; cz = 0; shell = true; }
        else if (z >= L.n[2]) { flip ^= (neg >> 5) & 1u; cz = L.n[2] - 1; shell = true; }
        if (!shell) continue;
        const T v = s[(cx - B.a0[0]) + sy * (cy - B.a0[1]) + sz * (cz - B.a0[2])];
        u_out[g_at(L, x, y, z)] = flip ? -v : v;
    }
}

template <typename T>
__global__ __launch_bounds__(MG_SM_THREADS) void k_mg_smooth(SmoothArgs A, const T *u_in, const T *b, T *u_out,
                                                             const T *xc, T *rh)
{
    extern __shared__ double lds_raw[];
    T *s = (T *)lds_raw;
    const Lvl L = A.L;
    const BcCoef bc = A.bc;
    SmBox B;
    {
        uint32_t q = xcd_block();
        for (int d = 0; d < 3; ++d) {
            int td;
            if (d < 2) {
                td = (int)(q % (uint32_t)A.ntile[d]);
                q /= (uint32_t)A.ntile[d];
            } else {
                td = (int)q;
            }
            B.t0[d] = td * A.tile[d];
            B.t1[d] = B.t0[d] + A.tile[d];
            B.a0[d] = max(B.t0[d] - A.halo, 0);
            B.R[d] = min(B.t1[d] + A.halo, L.n[d]) - B.a0[d];
            B.b0[d] = max(B.t0[d] - (A.halo - 1), 0);
            B.Rb[d] = min(B.t1[d] + A.halo - 1, L.n[d]) - B.b0[d];
        }
    }
    const int sy = B.R[0], sz = B.R[0] * B.R[1];
    const int by = B.Rb[0], bz = B.Rb[0] * B.Rb[1];
    T *sb = s + sz * B.R[2];
    // both boxes, SM_LD cells per thread at a time so that the global loads overlap
    for (BoxWalk w(B.R[0], B.R[1], threadIdx.x, MG_SM_THREADS); w.z < B.R[2];) {
        T v[SM_LD], vb[SM_LD];
        int li[SM_LD], lb[SM_LD];
#pragma unroll
        for (int q = 0; q < SM_LD; ++q) {
            li[q] = -1;
            lb[q] = -1;
            if (w.z < B.R[2]) {
                const int x = B.a0[0] + w.x, y = B.a0[1] + w.y, z = B.a0[2] + w.z;
                li[q] = w.x + sy * w.y + sz * w.z;
                if (A.interp == 3) {
                    v[q] = T(0);
                } else if (A.interp == 2) {
                    v[q] = prolong_at(xc, L, x, y, z);
                } else {
                    v[q] = u_in[g_at(L, x, y, z)];
                    if (A.interp) v[q] += prolong_at(xc, L, x, y, z);
                }
                const int xb = x - B.b0[0], yb = y - B.b0[1], zb = z - B.b0[2];
                if ((unsigned)xb < (unsigned)B.Rb[0] && (unsigned)yb < (unsigned)B.Rb[1] &&
                    (unsigned)zb < (unsigned)B.Rb[2]) {
                    lb[q] = xb + by * yb + bz * zb;
                    vb[q] = b[g_at(L, x, y, z)];
                }
            }
            w.next();
        }
#pragma unroll
        for (int q = 0; q < SM_LD; ++q) {
            if (li[q] >= 0) s[li[q]] = v[q];
            if (lb[q] >= 0) sb[lb[q]] = vb[q];
        }
    }
    __syncthreads();
    if (A.ghost_pass == 0) {
        sm_ghosts(A, B, s, u_out);
        __syncthreads();
    }
    for (int p = 0; p < A.passes; ++p) {
        const int e = A.halo - 1 - p, par = p & 1;
        int lo[3], W[3];
        for (int d = 0; d < 3; ++d) {
            lo[d] = max(B.t0[d] - e, 0);
            W[d] = min(B.t1[d] + e, L.n[d]) - lo[d];
        }
        const int half = (W[0] + 1) / 2;
        // SM_GS cells of one colour per thread at a time: their right-hand sides are
        // loaded together, and none of them reads another
        for (BoxWalk w(half, W[1], threadIdx.x, MG_SM_THREADS); w.z < W[2];) {
            int li[SM_GS], cx[SM_GS], cy[SM_GS], cz[SM_GS];
            T bv[SM_GS];
#pragma unroll
            for (int q = 0; q < SM_GS; ++q) {
                li[q] = -1;
                if (w.z < W[2]) {
                    const int y = lo[1] + w.y, z = lo[2] + w.z;
                    const int x = lo[0] + 2 * w.x + ((lo[0] + y + z + par) & 1);
                    if (x < lo[0] + W[0]) {
                        li[q] = (x - B.a0[0]) + sy * (y - B.a0[1]) + sz * (z - B.a0[2]);
                        cx[q] = x;
                        cy[q] = y;
                        cz[q] = z;
                        bv[q] = sb[(x - B.b0[0]) + by * (y - B.b0[1]) + bz * (z - B.b0[2])];
                    }
                }
                w.next();
            }
#pragma unroll
            for (int q = 0; q < SM_GS; ++q)
                if (li[q] >= 0) {
                    const T uc = s[li[q]];
                    const T l = lap_box(s, li[q], sy, sz, cx[q], cy[q], cz[q], L, bc, uc);
                    s[li[q]] = uc + div6(T(1) * (l + bv[q]));
                }
        }
        __syncthreads();
        if (A.ghost_pass == p + 1) {
            sm_ghosts(A, B, s, u_out);
            __syncthreads();
        }
    }
    {
        const int ox = B.t0[0] - B.a0[0], oy = B.t0[1] - B.a0[1], oz = B.t0[2] - B.a0[2];
        for (BoxWalk w(A.tile[0], A.tile[1], threadIdx.x, MG_SM_THREADS); w.z < A.tile[2]; w.next())
            u_out[g_at(L, B.t0[0] + w.x, B.t0[1] + w.y, B.t0[2] + w.z)] =
                s[(ox + w.x) + sy * (oy + w.y) + sz * (oz + w.z)];
    }
    if (A.resid) {
        Lvl C;
        C.n[0] = L.n[0] / 2;
        C.n[1] = L.n[1] / 2;
        C.n[2] = L.n[2] / 2;
        for (BoxWalk w(A.tile[0] / 2, A.tile[1] / 2, threadIdx.x, MG_SM_THREADS); w.z < A.tile[2] / 2; w.next()) {
            const int ic = B.t0[0] / 2 + w.x, jc = B.t0[1] / 2 + w.y, kc = B.t0[2] / 2 + w.z;
            T sum = T(0);
#pragma unroll
            for (int dx = 0; dx < 2; ++dx)
#pragma unroll
                for (int dy = 0; dy < 2; ++dy)
#pragma unroll
                    for (int dz = 0; dz < 2; ++dz) {
                        const int x = 2 * ic + dx, y = 2 * jc + dy, z = 2 * kc + dz;
                        const int li = (x - B.a0[0]) + sy * (y - B.a0[1]) + sz * (z - B.a0[2]);
                        const T uc = s[li];
                        const T l = lap_box(s, li, sy, sz, x, y, z, L, bc, uc);
                        sum += T(1) * sb[(x - B.b0[0]) + by * (y - B.b0[1]) + bz * (z - B.b0[2])] + l;
                    }
            rh[g_at(C, ic, jc, kc)] = (T(4) * sum) / T(8);
        }
    }
}

// ---- single-workgroup tail: mgv! from a small level down to the coarsest -----

struct Tail {
    Lvl n[MG_MAXL];
    int64_t len[MG_MAXL];
    int64_t off_x[MG_MAXL];   // LDS offsets (elements) of the level solutions
    int64_t off_b[MG_MAXL];   // and of their right-hand sides
    int64_t off_t;            // LDS offset of the fine-residual scratch (start level's cells)
    int L;
    int smooth1, smooth2;
    BcCoef bc;
};

#define BLOOP(t, N) for (int64_t t = threadIdx.x; t < (N); t += MG_TAIL_THREADS)

// Level lev's extents, read element by element (a copy of the indexed struct
// goes through private memory).
__device__ __forceinline__ Lvl tail_lvl(const Tail &T, int lev)
{
    Lvl L;
    L.n[0] = T.n[lev].n[0];
    L.n[1] = T.n[lev].n[1];
    L.n[2] = T.n[lev].n[2];
    return L;
}

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
    const int64_t bd = MG_TAIL_THREADS;
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

// x / 6 without a branch: the correction of div6 with the sign of x kept (so
// zeros come out as x / 6.0 does); `bad` records an x for which the correction
// is not proven (tiny, huge, non-finite), whose caller then recomputes with
// the division.  FP32 multiplies by the reciprocal, as div6 does.
__device__ __forceinline__ double div6_flag(double x, bool &bad)
{
    const double a = fabs(x);
    bad |= !(a < 0x1p960) || (a != 0.0 && !(a > 0x1p-960));
    const double c = 1.0 / 6.0;
    const double q = x * c;
    return copysign(__builtin_fma(__builtin_fma(-q, 6.0, x), c, q), x);
}
__device__ __forceinline__ float div6_flag(float x, bool &) { return x * (1.0f / 6.0f); }

// The 50 rounds on the lanes' registers; EXACT divides, otherwise div6_flag.
// Every lane computes both colours' updates and keeps its own: no divergence.
// Returns whether any update of this lane needed the division.
template <typename T, int SY, int SZ, bool EXACT>
__device__ __forceinline__ bool coarsest_dpp_rounds(T *x, const Lvl &Lv, const BcCoef &bc, T u, const T bb,
                                                    int64_t c, bool own, int i, int j, int k)
{
    constexpr int ZP = SZ ? 0x100 + SZ : 0x101, ZM = SZ ? 0x110 + SZ : 0x111;
    const int lane = threadIdx.x;
    const int nx = Lv.n[0], ny = Lv.n[1], nz = Lv.n[2];
    const bool up[2] = {own && ((i + j + k) & 1) == 0, own && ((i + j + k) & 1) == 1};
    const bool ixp = i + 1 < nx, iyp = j + 1 < ny, izp = k + 1 < nz, ixm = i > 0, iym = j > 0, izm = k > 0;
    bool bad = false;
    for (int it = 0; it < 50; ++it) {
        if (it == 49) {
            if (own) x[c] = u;
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            for (int64_t t = lane; t < bc_count(Lv); t += MCS_WAVE) bc_item(x, Lv, bc, t);
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
#pragma unroll
        for (int par = 0; par < 2; ++par) {
            const T vxp = dpp_move<0x101>(u), vxm = dpp_move<0x111>(u);
            const T vyp = dpp_move<0x100 + SY>(u), vym = dpp_move<0x110 + SY>(u);
            const T vzp = dpp_move<ZP>(u), vzm = dpp_move<ZM>(u);
            T l = T(-6) * u;
            l += ixp ? vxp : bcf(bc, 3, u);
            l += iyp ? vyp : bcf(bc, 4, u);
            l += izp ? vzp : bcf(bc, 5, u);
            l += ixm ? vxm : bcf(bc, 0, u);
            l += iym ? vym : bcf(bc, 1, u);
            l += izm ? vzm : bcf(bc, 2, u);
            const T y = T(1) * (l + bb);
            bool bd = false;
            const T q = EXACT ? y / T(6) : div6_flag(y, bd);
            bad |= bd && up[par];
            const T un = u + q;
            u = up[par] ? un : u;
        }
    }
    if (own) x[c] = u;
    return bad;
}

template <typename T, int SY, int SZ>
__device__ void coarsest_dpp(T *x, const T *b, const Lvl &Lv, const BcCoef &bc)
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
    const T u0 = own ? x[c] : T(0);
    const T bb = own ? b[c] : T(0);
    const bool bad = coarsest_dpp_rounds<T, SY, SZ, false>(x, Lv, bc, u0, bb, c, own, i, j, k);
    if (__builtin_amdgcn_ballot_w64(bad)) {
        // an update the correction does not cover: redo all rounds dividing
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        coarsest_dpp_rounds<T, SY, SZ, true>(x, Lv, bc, u0, bb, c, own, i, j, k);
    }
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

// ---- per-item bodies of the tail, in 32-bit arithmetic (levels < 2^31 cells) --

// lap_fold at (i, j, k), index c of the ghost layout.
template <typename T>
__device__ __forceinline__ T tail_lap(const T *u, const Lvl &L, const BcCoef &bc, int i, int j, int k, int c, T uc)
{
    const int sy = L.n[0] + 2, sz = (L.n[0] + 2) * (L.n[1] + 2);
    T l = T(-6) * uc;
    l += (i + 1 < L.n[0]) ? u[c + 1] : bcf(bc, 3, uc);
    l += (j + 1 < L.n[1]) ? u[c + sy] : bcf(bc, 4, uc);
    l += (k + 1 < L.n[2]) ? u[c + sz] : bcf(bc, 5, uc);
    l += (i > 0) ? u[c - 1] : bcf(bc, 0, uc);
    l += (j > 0) ? u[c - sy] : bcf(bc, 1, uc);
    l += (k > 0) ? u[c - sz] : bcf(bc, 2, uc);
    return l;
}

__device__ __forceinline__ int tail_cell(const Lvl &L, int t, int &i, int &j, int &k)
{
    const int q = fdivmod(t, L.n[0], 1.0f / (float)L.n[0], i);
    k = fdivmod(q, L.n[1], 1.0f / (float)L.n[1], j);
    return (i + 1) + (L.n[0] + 2) * ((j + 1) + (L.n[1] + 2) * (k + 1));
}

// One colour pass of gauss_seidel! (gs_item's arithmetic).
template <typename T>
__device__ __forceinline__ void tail_gs(T *u, const T *b, const Lvl &L, const BcCoef &bc, int par)
{
    const int nxh = L.n[0] / 2, ny = L.n[1];
    const int ncol = nxh * ny * L.n[2];
    const float rx = 1.0f / (float)nxh, ry = 1.0f / (float)ny;
    for (int t = threadIdx.x; t < ncol; t += MG_TAIL_THREADS) {
        int ti, j;
        const int q = fdivmod(t, nxh, rx, ti);
        const int k = fdivmod(q, ny, ry, j);
        const int i = 2 * ti + ((j + k + par) & 1);
        const int c = (i + 1) + (L.n[0] + 2) * ((j + 1) + (ny + 2) * (k + 1));
        const T uc = u[c];
        const T l = tail_lap(u, L, bc, i, j, k, c, uc);
        u[c] = uc + div6(T(1) * (l + b[c]));
    }
}

// resid_restrict_item in two phases: every fine cell's s b + L u into rt, then
// the sums of the 8 children in cubestencil order (the same additions).
template <typename T>
__device__ __forceinline__ void tail_resid(T *rt, const T *u, const T *b, const Lvl &L, const BcCoef &bc)
{
    const int N = L.n[0] * L.n[1] * L.n[2];
    for (int t = threadIdx.x; t < N; t += MG_TAIL_THREADS) {
        int i, j, k;
        const int c = tail_cell(L, t, i, j, k);
        const T uc = u[c];
        rt[t] = T(1) * b[c] + tail_lap(u, L, bc, i, j, k, c, uc);
    }
}

template <typename T>
__device__ __forceinline__ void tail_restrict(T *rh, T *xh, const T *rt, const Lvl &F)
{
    Lvl C;
    C.n[0] = F.n[0] / 2;
    C.n[1] = F.n[1] / 2;
    C.n[2] = F.n[2] / 2;
    const int NC = C.n[0] * C.n[1] * C.n[2];
    const int sy = F.n[0], sz = F.n[0] * F.n[1];
    for (int t = threadIdx.x; t < NC; t += MG_TAIL_THREADS) {
        int ic, jc, kc;
        const int cc = tail_cell(C, t, ic, jc, kc);
        const int f = 2 * ic + sy * (2 * jc) + sz * (2 * kc);
        T sum = T(0);
        sum += rt[f];
        sum += rt[f + sz];
        sum += rt[f + sy];
        sum += rt[f + sy + sz];
        sum += rt[f + 1];
        sum += rt[f + 1 + sz];
        sum += rt[f + 1 + sy];
        sum += rt[f + 1 + sy + sz];
        rh[cc] = (T(4) * sum) / T(8);
        xh[cc] = T(0);
    }
}

// interpolate! into level L one fine cell per thread (interp_item's value).
template <typename T>
__device__ __forceinline__ void tail_interp(T *x, const T *rh, const Lvl &L, bool add)
{
    const int N = L.n[0] * L.n[1] * L.n[2];
    for (int t = threadIdx.x; t < N; t += MG_TAIL_THREADS) {
        int i, j, k;
        const int c = tail_cell(L, t, i, j, k);
        const T v = prolong_at(rh, L, i, j, k);
        x[c] = add ? x[c] + v : v;
    }
}

// mgv! (Multigrid.jl:334-377) for levels start..L inside one workgroup, on the
// level solutions and right-hand sides in LDS (x and b of `start` loaded).
// Phases are separated by barriers and follow the grid path's order exactly.
template <typename E>
__device__ void tail_vcycle(const Tail &T, E *lds, int start)
{
    const BcCoef bc = T.bc;
    E *rt = lds + T.off_t;   // fine residuals of the level being restricted
    for (int lev = start; lev < T.L; ++lev) {
        E *x = lds + T.off_x[lev];
        const E *b = lds + T.off_b[lev];
        const Lvl Lv = tail_lvl(T, lev);
        for (int it = 0; it < T.smooth1; ++it) {
            for (int par = 0; par < 2; ++par) {
                tail_gs(x, b, Lv, bc, par);
                __syncthreads();
            }
        }
        if (T.smooth2 == 0) {
            BLOOP(t, bc_count(Lv)) bc_item(x, Lv, bc, t);
            __syncthreads();
        }
        tail_resid(rt, x, b, Lv, bc);
        __syncthreads();
        // restricted residual and the zeroed interior of the next solution (its
        // ghosts are materialised before they are read)
        tail_restrict(lds + T.off_b[lev + 1], lds + T.off_x[lev + 1], rt, Lv);
        __syncthreads();
    }
    {
        const int lev = T.L;
        E *x = lds + T.off_x[lev];
        const E *b = lds + T.off_b[lev];
        const Lvl Lv = tail_lvl(T, lev);
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
        const E *b = lds + T.off_b[lev];
        const Lvl Lv = tail_lvl(T, lev);
        tail_interp(x, lds + T.off_x[lev + 1], Lv, true);
        __syncthreads();
        for (int it = 0; it < T.smooth2; ++it) {
            if (it == T.smooth2 - 1) {
                BLOOP(t, bc_count(Lv)) bc_item(x, Lv, bc, t);
                __syncthreads();
            }
            for (int par = 0; par < 2; ++par) {
                tail_gs(x, b, Lv, bc, par);
                __syncthreads();
            }
        }
    }
}

// mgv! for levels start..L in one workgroup: x0 and b0 are loaded (x0 unless the
// start is zero or a prolongation), x0 is written back.
template <typename E>
__device__ __forceinline__ void tail_body(const Tail &T, int start, E *x0, const E *b0, const E *zc, int zero,
                                          int fresh, E *lds)
{
    const BcCoef bc = T.bc;
    if (zc) {
        // the start level's interior is interpolate!(x, zc) (overwrite); its
        // ghosts are materialised before they are read
        tail_interp(lds + T.off_x[start], zc, tail_lvl(T, start), false);
    } else if (zero) {
        BLOOP(t, T.len[start]) lds[T.off_x[start] + t] = E(0);
    } else {
        tail_copy(lds + T.off_x[start], x0, T.len[start]);
    }
    tail_copy(lds + T.off_b[start], b0, T.len[start]);
    __syncthreads();
    tail_vcycle(T, lds, start);
    if (fresh) {
        // ghosts of the final state (the apply! the caller would issue next)
        __syncthreads();
        const Lvl Ls = tail_lvl(T, start);
        BLOOP(t, bc_count(Ls)) bc_item(lds + T.off_x[start], Ls, bc, t);
    }
    __syncthreads();
    tail_copy(x0, lds + T.off_x[start], T.len[start]);
}

template <typename E>
__global__ __launch_bounds__(MG_TAIL_THREADS) void k_mg_tail(Tail T, int start, E *x0, const E *b0, const E *zc,
                                                            int zero, int fresh)
{
    extern __shared__ double lds_raw[];
    tail_body(T, start, x0, b0, zc, zero, fresh, (E *)lds_raw);
}

// Tabulation of the tail (tabulate_tails): block q runs mgv! from x = 0 on its
// own pair of start-level arrays, q lengths into x0 and b0.
template <typename E>
__global__ __launch_bounds__(MG_TAIL_THREADS) void k_mg_tail_batch(Tail T, int start, E *x0, const E *b0)
{
    extern __shared__ double lds_raw[];
    const int64_t off = (int64_t)blockIdx.x * T.len[start];
    tail_body(T, start, x0 + off, b0 + off, (const E *)nullptr, 1, 0, (E *)lds_raw);
}

// The coarse part of fmg!'s correction (Multigrid.jl:396-426) in one workgroup:
// the right-hand sides of levels start..L restricted in LDS from rfine (the
// residual of level start - 1, ghost layout; restrict_item's arithmetic), the
// direct solve of the coarsest level (k_mg_coarse's), its apply!, then for each
// level L-1..start the interpolation of the coarser solution, mgv! and apply!
// (the fresh shell the next interpolation reads).  The start level's solution,
// shell included, is written to x0.
template <typename E>
__device__ __forceinline__ void tail_fmg_body(const Tail &T, int start, E *x0, const E *rfine, const E *ainv,
                                              E *lds)
{
    const BcCoef bc = T.bc;
    {
        const Lvl F = tail_lvl(T, start - 1), C = tail_lvl(T, start);
        E *rh = lds + T.off_b[start];
        BLOOP(t, T.len[start]) rh[t] = E(0);
        __syncthreads();
        BLOOP(t, (int64_t)C.n[0] * C.n[1] * C.n[2]) restrict_item(rh, rfine, F, t);
        __syncthreads();
    }
    for (int lev = start; lev < T.L; ++lev) {
        const Lvl F = tail_lvl(T, lev), C = tail_lvl(T, lev + 1);
        E *rh = lds + T.off_b[lev + 1];
        BLOOP(t, T.len[lev + 1]) rh[t] = E(0);
        __syncthreads();
        BLOOP(t, (int64_t)C.n[0] * C.n[1] * C.n[2]) restrict_item(rh, (const E *)(lds + T.off_b[lev]), F, t);
        __syncthreads();
    }
    {
        const Lvl C = tail_lvl(T, T.L);
        const int64_t N = (int64_t)C.n[0] * C.n[1] * C.n[2];
        E *z = lds + T.off_x[T.L];
        const E *bt = lds + T.off_b[T.L];
        BLOOP(t, T.len[T.L]) z[t] = E(0);
        __syncthreads();
        BLOOP(row, N) {
            E acc = E(0);
            for (int64_t col = 0; col < N; ++col) {
                int64_t i, j, k;
                split3(col, (uint32_t)C.n[0], (uint32_t)C.n[1], i, j, k);
                acc += ainv[row * N + col] * bt[g_index(C, i + 1, j + 1, k + 1)];
            }
            int64_t i, j, k;
            split3(row, (uint32_t)C.n[0], (uint32_t)C.n[1], i, j, k);
            z[g_index(C, i + 1, j + 1, k + 1)] = acc;
        }
        __syncthreads();
        BLOOP(t, bc_count(C)) bc_item(z, C, bc, t);
        __syncthreads();
    }
    for (int lev = T.L - 1; lev >= start; --lev) {
        const Lvl Lv = tail_lvl(T, lev);
        E *x = lds + T.off_x[lev];
        tail_interp(x, lds + T.off_x[lev + 1], Lv, false);
        __syncthreads();
        tail_vcycle(T, lds, lev);
        __syncthreads();
        BLOOP(t, bc_count(Lv)) bc_item(x, Lv, bc, t);
        __syncthreads();
    }
    tail_copy(x0, lds + T.off_x[start], T.len[start]);
}

template <typename E>
__global__ __launch_bounds__(MG_TAIL_THREADS) void k_mg_tail_fmg(Tail T, int start, E *x0, const E *rfine,
                                                                const E *ainv)
{
    extern __shared__ double lds_raw[];
    tail_fmg_body(T, start, x0, rfine, ainv, (E *)lds_raw);
}

// Tabulation of fmg!'s coarse part: block q on its own output and fine residual
// (rlen apart).
template <typename E>
__global__ __launch_bounds__(MG_TAIL_THREADS) void k_mg_tail_fmg_batch(Tail T, int start, E *x0, const E *rfine,
                                                                      const E *ainv, int64_t rlen)
{
    extern __shared__ double lds_raw[];
    tail_fmg_body(T, start, x0 + (int64_t)blockIdx.x * T.len[start], rfine + (int64_t)blockIdx.x * rlen, ainv,
                  (E *)lds_raw);
}

// ---- tabulated tails ---------------------------------------------------------
//
// Below the tail's start level S the cycle is linear in S's right-hand side:
// mgv! starts from x = 0 (the caller's xh .= 0), the coarse levels' boundary
// rules are the homogeneous ghost = +-valid, and every operation (sweeps,
// residual, restriction, the direct coarse solve, prolongation) is linear.  The
// whole ghosted output of S is therefore x = M b over b's valid cells.  M is
// tabulated once per solver by running the tail itself on unit right-hand sides
// and applied as one wide matrix-vector product in place of the single-workgroup
// chain.  Same operator; the rounding differs from the sweeps' at the level of
// the last bits of each entry.

// Unit inputs: array q (len apart) is zero except at one valid cell, which holds
// v: cell q of level L (i fastest), or with child = 1 the first child
// (2i, 2j, 2k) of coarse cell q on the finer level L (whose restriction
// 4 v / 8 then puts v / 2 in cell q).
template <typename E>
__global__ void k_tm_units(E *a, int64_t len, Lvl L, int64_t ncols, int child, E v)
{
    const int64_t t = gidx();
    if (t >= len * ncols) return;
    const int64_t q = t / len, p = t - q * len;
    const uint32_t cx = child ? L.n[0] / 2 : L.n[0], cy = child ? L.n[1] / 2 : L.n[1];
    int64_t i, j, k;
    split3(q, cx, cy, i, j, k);
    const int64_t pos = child ? g_index(L, 2 * i + 1, 2 * j + 1, 2 * k + 1) : g_index(L, i + 1, j + 1, k + 1);
    a[t] = p == pos ? v : E(0);
}

// M[r][q] = X[q][r] (row-major rows x ncols from ncols arrays of rows).
template <typename E>
__global__ void k_tm_transpose(E *M, const E *X, int64_t rows, int64_t ncols)
{
    const int64_t t = gidx();
    if (t >= rows * ncols) return;
    const int64_t r = t / ncols, q = t - r * ncols;
    M[t] = X[q * rows + r];
}

#define MG_TM_WAVES 4
// x = M b on level L: b's valid cells staged in LDS in cell order, a wavefront per
// output position (ghosts included), 64 columns per step in FP64 with fused
// multiply-adds, the lane sums combined by a fixed butterfly: the same sum on
// every run.
template <typename E>
__global__ __launch_bounds__(MG_TM_WAVES * MCS_WAVE) void k_mg_tailmat(const E *__restrict__ M,
                                                                      const E *__restrict__ b, E *__restrict__ x,
                                                                      Lvl L, int64_t rows)
{
    __shared__ E bs[MG_TAIL_MAXCELLS];
    const int C = L.n[0] * L.n[1] * L.n[2];
    for (int t = threadIdx.x; t < C; t += MG_TM_WAVES * MCS_WAVE) {
        int64_t i, j, k;
        split3(t, (uint32_t)L.n[0], (uint32_t)L.n[1], i, j, k);
        bs[t] = b[g_index(L, i + 1, j + 1, k + 1)];
    }
    __syncthreads();
    const int lane = threadIdx.x & (MCS_WAVE - 1);
    const int64_t r = (int64_t)blockIdx.x * MG_TM_WAVES + (threadIdx.x / MCS_WAVE);
    if (r >= rows) return;
    const E *row = M + r * C;
    double acc = 0.0;
    for (int q = lane; q < C; q += MCS_WAVE) acc = __builtin_fma((double)row[q], (double)bs[q], acc);
#pragma unroll
    for (int o = MCS_WAVE / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, MCS_WAVE);
    if (lane == 0) x[r] = (E)acc;
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

template <typename T> static bool in_tail(const mcs_mg *mg, int level);
template <typename T> static bool smooth_fits(mcs_mg *mg, int level);
template <typename T> static int32_t tabulate_tails(mcs_mg *mg);

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
                             (int)tail_lds) != hipSuccess ||
         hipFuncSetAttribute((const void *)k_mg_tail_fmg<double>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)tail_lds) != hipSuccess ||
         hipFuncSetAttribute((const void *)k_mg_tail_fmg<float>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)tail_lds) != hipSuccess))
        tail_lds = 64 * 1024;
    if (!ctx || !cfg || !out) return MCS_E_ARG;
    if (cfg->ndim == 2) return mcs_mg2_create(ctx, cfg, out);
    if (cfg->ndim != 0 && cfg->ndim != 3) return mcs_fail(ctx, MCS_E_ARG, "mg: ndim is 2 or 3");
    if (cfg->cyl != 0) return mcs_fail(ctx, MCS_E_ARG, "mg: the cylindrical connector is for 2D grids");
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
    size_t sm_lds = (size_t)lds_max < (size_t)MG_TAIL_LDS ? (size_t)lds_max : (size_t)MG_TAIL_LDS;
    if (hipFuncSetAttribute((const void *)k_mg_smooth<double>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)sm_lds) != hipSuccess ||
        hipFuncSetAttribute((const void *)k_mg_smooth<float>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)sm_lds) != hipSuccess)
        sm_lds = 64 * 1024;
    // MCS_MG_FUSED=0 selects one launch per pass on the grid levels (same results);
    // MCS_MG_TAIL_LEVELS caps the levels of the single-workgroup tail (at least 1).
    const char *fe = getenv("MCS_MG_FUSED");
    if ((fe && fe[0] == '0') || c.smooth2 < 1 || c.smooth1 > 2 || c.smooth2 > 2) sm_lds = 0;
    // MCS_MG_SM_TILE: largest tile edge of the tiled smoother; MCS_MG_SM_MAXCELLS:
    // levels of more cells take one launch per pass (the halos would cost more
    // traffic than the launches they save); MCS_MG_SM_BIG: levels of at least this
    // many cells take one sweep per launch instead of two
    const char *se = getenv("MCS_MG_SM_TILE");
    int sm_tile = se ? atoi(se) : 8;
    if (sm_tile < 2) sm_tile = 2;
    // MCS_MG_SM_MINBLOCKS: smaller tiles on a level until it has this many (64:
    // 16x16x32 takes 4^3 tiles, 128 workgroups instead of 16; 0: the largest tile)
    const char *mbe = getenv("MCS_MG_SM_MINBLOCKS");
    const unsigned sm_minblocks = mbe ? (unsigned)atoi(mbe) : 64u;
    const char *me = getenv("MCS_MG_SM_MAXCELLS");
    int64_t sm_max = me ? atoll(me) : MG_SM_MAXCELLS;
    const char *be = getenv("MCS_MG_SM_BIG");
    int64_t sm_big = be ? atoll(be) : ((int64_t)1 << 62);
    // MCS_MG_TAIL_MAXCELLS: largest level the tail starts from
    const char *tm = getenv("MCS_MG_TAIL_MAXCELLS");
    const int64_t tail_max = tm ? atoll(tm) : MG_TAIL_MAXCELLS;
    const char *fte = getenv("MCS_MG_FMG_TAIL");
    const int fmg_tail = fte ? atoi(fte) : 1;
    const char *te = getenv("MCS_MG_TAIL_LEVELS");
    int tail_levels = te ? atoi(te) : MG_MAXL;
    if (tail_levels < 1) tail_levels = 1;
    mcs_mg *mg = new mcs_mg();
    mg->ctx = ctx;
    mg->cfg = c;
    mg->tail_lds = tail_lds;
    mg->fmg_tail = fmg_tail;
    mg->tail_levels = tail_levels;
    mg->tail_max = tail_max;
    mg->sm_lds = sm_lds;
    mg->sm_minblocks = sm_minblocks;
    mg->sm_tile = sm_tile;
    mg->sm_big = sm_big;
    mg->sm_max = sm_max;
    const char *ge = getenv("MCS_MG_GRAPH");
    mg->use_graph = !(ge && ge[0] == '0');
    for (int l = 0; l <= c.levels; ++l) {
        MgLevel L;
        for (int d = 0; d < 3; ++d) L.n[d] = c.n[d] >> l;
        L.len = glen(L.n);
        L.fres = nullptr;
        L.fsol = nullptr;
        L.res1 = nullptr;
        L.alt = nullptr;
        L.falt = nullptr;
        if (hipMalloc(&L.res, sizeof(double) * L.len) != hipSuccess ||
            hipMalloc(&L.sol, sizeof(double) * L.len) != hipSuccess)
            return mcs_fail(ctx, MCS_E_HIP, "mg: device allocation failed");
        MCS_HIP(ctx, mcs_zero(ctx, L.res, sizeof(double) * L.len));
        MCS_HIP(ctx, mcs_zero(ctx, L.sol, sizeof(double) * L.len));
        if (f32) {
            if (hipMalloc(&L.fres, sizeof(float) * L.len) != hipSuccess ||
                hipMalloc(&L.fsol, sizeof(float) * L.len) != hipSuccess)
                return mcs_fail(ctx, MCS_E_HIP, "mg: device allocation failed");
            MCS_HIP(ctx, mcs_zero(ctx, L.fres, sizeof(float) * L.len));
            MCS_HIP(ctx, mcs_zero(ctx, L.fsol, sizeof(float) * L.len));
        }
        mg->lv.push_back(L);
    }
    // second solution arrays of the tiled smoother, on the levels outside the tail
    if (sm_lds) {
        for (int l = 0; l < c.levels; ++l) {
            MgLevel &L = mg->lv[l];
            if ((in_tail<double>(mg, l) && in_tail<float>(mg, l)) ||
                (int64_t)L.n[0] * L.n[1] * L.n[2] > sm_max)
                continue;
            if (f32) {
                if (hipMalloc(&L.falt, sizeof(float) * L.len) != hipSuccess)
                    return mcs_fail(ctx, MCS_E_HIP, "mg: device allocation failed");
                MCS_HIP(ctx, mcs_zero(ctx, L.falt, sizeof(float) * L.len));
            } else {
                if (hipMalloc(&L.alt, sizeof(double) * L.len) != hipSuccess)
                    return mcs_fail(ctx, MCS_E_HIP, "mg: device allocation failed");
                MCS_HIP(ctx, mcs_zero(ctx, L.alt, sizeof(double) * L.len));
            }
        }
    }
    // the level-0 correction in the colour-split layout when level 0 runs one
    // launch per pass (MCS_MG_SPLIT0=0: ghost layout, same results)
    {
        const char *ze = getenv("MCS_MG_SPLIT0");
        const bool per_pass = f32 ? (!in_tail<float>(mg, 0) && !smooth_fits<float>(mg, 0))
                                  : (!in_tail<double>(mg, 0) && !smooth_fits<double>(mg, 0));
        mg->split0 = c.levels > 0 && per_pass && !(ze && ze[0] == '0');
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
    MCS_HIP(ctx, mcs_h2d(ctx, mg->ainv, ainv.data(), sizeof(double) * ainv.size()));
    if (f32) {
        std::vector<float> a32(ainv.begin(), ainv.end());
        if (hipMalloc(&mg->ainv32, sizeof(float) * a32.size()) != hipSuccess)
            return mcs_fail(ctx, MCS_E_HIP, "mg: device allocation failed");
        MCS_HIP(ctx, mcs_h2d(ctx, mg->ainv32, a32.data(), sizeof(float) * a32.size()));
    }
    int64_t N0 = (int64_t)c.n[0] * c.n[1] * c.n[2];
    mg->npartial = cdiv(N0, MCS_BLOCK);
    if (hipMalloc(&mg->partial, sizeof(double) * (mg->npartial + 8)) != hipSuccess)
        return mcs_fail(ctx, MCS_E_HIP, "mg: device allocation failed");
    const int32_t trc = f32 ? tabulate_tails<float>(mg) : tabulate_tails<double>(mg);
    if (trc) return trc;
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
        if (L.res1) hipFree(L.res1);
        if (L.fres) hipFree(L.fres);
        if (L.fsol) hipFree(L.fsol);
        if (L.alt) hipFree(L.alt);
        if (L.falt) hipFree(L.falt);
    }
    hipFree(mg->ainv);
    if (mg->ainv32) hipFree(mg->ainv32);
    hipFree(mg->partial);
    if (mg->u_tmp) hipFree(mg->u_tmp);
    if (mg->b_tmp) hipFree(mg->b_tmp);
    if (mg->bc_dev) hipFree(mg->bc_dev);
    if (mg->tm_z) hipFree(mg->tm_z);
    if (mg->tm_f) hipFree(mg->tm_f);
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
template <typename T> static T *lalt(mcs_mg *mg, int l);
template <> double *lalt<double>(mcs_mg *mg, int l) { return mg->lv[l].alt; }
template <> float *lalt<float>(mcs_mg *mg, int l) { return mg->lv[l].falt; }

template <typename T>
static void mg_bc(mcs_mg *mg, int l, T *a)
{
    Lvl L = lvl(mg, l);
    int64_t nx = L.n[0], ny = L.n[1], nz = L.n[2];
    int64_t cnt = 2 * (ny + 2) * (nz + 2) + 2 * nx * (nz + 2) + 2 * nx * ny;
    hipLaunchKernelGGL(k_mg_bc<T>, grid_for(cnt), dim3(MCS_BLOCK), 0, mg->ctx->stream, a, L, coefs(mg));
}

// MCS_MG_RBGS_SINGLE=1: the colour-split passes and prolongation take one cell
// per thread (k_mg_rbgs_s, k_mg_interp_s) instead of two (same results).
static bool mg_rbgs_single()
{
    const char *e = getenv("MCS_MG_RBGS_SINGLE");
    return e && e[0] == '1';
}

// One sweep (colour 0, then 1); ghosts: the colour-1 pass also writes the ghost
// shell of the state the sweep ends with (ghost layout only).
template <typename T>
static void mg_gs(mcs_mg *mg, int l, T *u, const T *b, bool split = false, bool ghosts = false)
{
    Lvl L = lvl(mg, l);
    int64_t cnt = (int64_t)(L.n[0] / 2) * L.n[1] * L.n[2];
    ProfScope ps(mg->ctx, "mg_rbgs");
    for (int par = 0; par < 2; ++par) {
        if (split) {
            if ((L.n[0] / 2) % 2 == 0 && !mg_rbgs_single())
                hipLaunchKernelGGL(k_mg_rbgs_s2<T>, grid_for(cnt / 2), dim3(MCS_BLOCK), 0, mg->ctx->stream, u, b,
                                   L, coefs(mg), par);
            else
                hipLaunchKernelGGL(k_mg_rbgs_s<T>, grid_for(cnt), dim3(MCS_BLOCK), 0, mg->ctx->stream, u, b, L,
                                   coefs(mg), par);
        } else {
            hipLaunchKernelGGL(k_mg_rbgs<T>, grid_for(cnt), dim3(MCS_BLOCK), 0, mg->ctx->stream, u, b, L,
                               coefs(mg), par, ghosts && par == 1 ? 1 : 0);
        }
    }
}

template <typename T>
static void mg_fill0(mcs_mg *mg, T *a, int64_t n)
{
    hipLaunchKernelGGL(k_mg_fill0<T>, grid_for(n), dim3(MCS_BLOCK), 0, mg->ctx->stream, a, n);
}

template <typename T>
static void mg_interp(mcs_mg *mg, int lf, T *r, const T *rh, int add, bool split = false)
{
    ProfScope ps(mg->ctx, "mg_prolong");
    if (split && (mg->lv[lf].n[0] / 2) % 2 == 0 && !mg_rbgs_single())
        hipLaunchKernelGGL(k_mg_interp_s2<T>, grid_for(ncoarse_cells(mg, lf) / 2), dim3(MCS_BLOCK), 0,
                           mg->ctx->stream, r, rh, lvl(mg, lf), add);
    else if (split)
        hipLaunchKernelGGL(k_mg_interp_s<T>, grid_for(ncoarse_cells(mg, lf)), dim3(MCS_BLOCK), 0, mg->ctx->stream,
                           r, rh, lvl(mg, lf), add);
    else
        hipLaunchKernelGGL(k_mg_interp<T>, grid_for(ncoarse_cells(mg, lf)), dim3(MCS_BLOCK), 0, mg->ctx->stream,
                           r, rh, lvl(mg, lf), add);
}

static int32_t ensure_bc_dev(mcs_mg *mg)
{
    if (mg->bc_dev) return MCS_OK;
    MCS_HIP(mg->ctx, hipMalloc(&mg->bc_dev, sizeof(int32_t) * 6));
    MCS_HIP(mg->ctx, mcs_h2d(mg->ctx, mg->bc_dev, mg->cfg.bc, sizeof(int32_t) * 6));
    return MCS_OK;
}

// LDS bytes of a tail from `level`: the solutions and right-hand sides of
// levels level..L, and the fine residuals of the start level.
template <typename T>
static size_t tail_bytes(const mcs_mg *mg, int level)
{
    size_t tot = 0;
    for (int l = level; l <= mg->cfg.levels; ++l) tot += 2 * (size_t)mg->lv[l].len * sizeof(T);
    const int32_t *n = mg->lv[level].n;
    return tot + (size_t)n[0] * n[1] * n[2] * sizeof(T);
}

template <typename T>
static bool in_tail(const mcs_mg *mg, int level)
{
    const int32_t *n = mg->lv[level].n;
    return tail_bytes<T>(mg, level) <= mg->tail_lds && mg->cfg.levels - level < mg->tail_levels &&
           (level == mg->cfg.levels || (int64_t)n[0] * n[1] * n[2] <= mg->tail_max);
}

template <typename T>
static int64_t tail_setup(mcs_mg *mg, int level, Tail &Tl)
{
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
            off += Tl.len[l];
        }
    }
    Tl.off_t = off;
    off += (int64_t)mg->lv[level].n[0] * mg->lv[level].n[1] * mg->lv[level].n[2];
    return off;
}

// The first level of fmg!'s coarse-to-fine sweep that k_mg_tail_fmg takes on
// (with all coarser ones), or -1: a level whose mgv! is a tail, below the finest
// (its right-hand side is restricted from the level above, in the ghost layout).
template <typename T>
static int fmg_tail_start(const mcs_mg *mg)
{
    if (!mg->fmg_tail) return -1;
    const int L = mg->cfg.levels;
    int S = -1;
    for (int l = L - 1; l >= 1; --l) {
        if (!in_tail<T>(mg, l)) break;
        S = l;
    }
    if (S == 1 && mg->split0) return -1;
    return S;
}

// x = M b on `level` from a tabulated tail.
template <typename T>
static void tail_matvec(mcs_mg *mg, const void *M, const T *b, T *x, int level)
{
    ProfScope ps(mg->ctx, "mg_tail");
    const int64_t rows = mg->lv[level].len;
    hipLaunchKernelGGL(k_mg_tailmat<T>, dim3((unsigned)cdiv(rows, MG_TM_WAVES)), dim3(MG_TM_WAVES * MCS_WAVE), 0,
                       mg->ctx->stream, (const T *)M, b, x, lvl(mg, level), rows);
}

// Tabulates the tails of the working precision T (the mcs_mg fields tm_*): M_z
// from mgv! on unit right-hand sides of the tail's start level S, and, when
// fmg!'s coarse part starts at S too, M_f from unit residuals of level S - 1
// whose restrictions are the unit right-hand sides.  Off with
// MCS_MG_TAIL_MATRIX=0, or when S has more cells than the matrix-vector kernel
// stages.
template <typename T>
static int32_t tabulate_tails(mcs_mg *mg)
{
    mcs_ctx *ctx = mg->ctx;
    mg->tm_level = -1;
    const char *e = getenv("MCS_MG_TAIL_MATRIX");
    if (e && e[0] == '0') return MCS_OK;
    int S = -1;
    for (int l = 1; l <= mg->cfg.levels; ++l)
        if (in_tail<T>(mg, l)) {
            S = l;
            break;
        }
    if (S < 1) return MCS_OK;
    const int64_t C = ncells(mg, S), R = mg->lv[S].len;
    if (C > MG_TAIL_MAXCELLS) return MCS_OK;
    Tail Tl;
    const int64_t words = tail_setup<T>(mg, S, Tl);
    const size_t lds = (size_t)words * sizeof(T);
    if (lds > 64 * 1024 &&
        (hipFuncSetAttribute((const void *)k_mg_tail_batch<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)mg->tail_lds) != hipSuccess ||
         hipFuncSetAttribute((const void *)k_mg_tail_fmg_batch<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)mg->tail_lds) != hipSuccess))
        return MCS_OK;
    const int Sf = fmg_tail_start<T>(mg);
    const int64_t Rf = Sf == S ? mg->lv[S - 1].len : 0;
    T *X = nullptr, *B = nullptr, *Mz = nullptr, *Mf = nullptr;
    auto release = [&]() {
        if (X) hipFree(X);
        if (B) hipFree(B);
        if (Mz) hipFree(Mz);
        if (Mf) hipFree(Mf);
    };
    if (hipMalloc(&X, sizeof(T) * C * R) != hipSuccess ||
        hipMalloc(&B, sizeof(T) * C * std::max(R, Rf)) != hipSuccess ||
        hipMalloc(&Mz, sizeof(T) * R * C) != hipSuccess || (Rf && hipMalloc(&Mf, sizeof(T) * R * C) != hipSuccess)) {
        release();
        return mcs_fail(ctx, MCS_E_HIP, "mg: device allocation of the tail tables failed");
    }
    hipLaunchKernelGGL(k_tm_units<T>, grid_for(C * R), dim3(MCS_BLOCK), 0, ctx->stream, B, R, lvl(mg, S), C, 0,
                       T(1));
    hipLaunchKernelGGL(k_mg_tail_batch<T>, dim3((unsigned)C), dim3(MG_TAIL_THREADS), lds, ctx->stream, Tl, S, X,
                       (const T *)B);
    hipLaunchKernelGGL(k_tm_transpose<T>, grid_for(C * R), dim3(MCS_BLOCK), 0, ctx->stream, Mz, (const T *)X, R, C);
    if (Rf) {
        hipLaunchKernelGGL(k_tm_units<T>, grid_for(C * Rf), dim3(MCS_BLOCK), 0, ctx->stream, B, Rf,
                           lvl(mg, S - 1), C, 1, T(2));
        hipLaunchKernelGGL(k_mg_tail_fmg_batch<T>, dim3((unsigned)C), dim3(MG_TAIL_THREADS), lds, ctx->stream, Tl,
                           S, X, (const T *)B, coarse_inv<T>(mg), Rf);
        hipLaunchKernelGGL(k_tm_transpose<T>, grid_for(C * R), dim3(MCS_BLOCK), 0, ctx->stream, Mf, (const T *)X,
                           R, C);
    }
    const hipError_t le = hipGetLastError();
    const hipError_t se = hipStreamSynchronize(ctx->stream);
    hipFree(X);
    hipFree(B);
    X = B = nullptr;
    if (le != hipSuccess || se != hipSuccess) {
        release();
        return mcs_fail(ctx, MCS_E_HIP, "mg: tabulating the tail failed (%s)",
                        hipGetErrorString(le != hipSuccess ? le : se));
    }
    mg->tm_level = S;
    mg->tm_z = Mz;
    mg->tm_f = Mf;
    return MCS_OK;
}

template <typename T>
static void mg_tail(mcs_mg *mg, int level, T *x, const T *b, const T *zc, bool zero, bool fresh)
{
    if (level == mg->tm_level && zero && !zc && !fresh) {
        tail_matvec<T>(mg, mg->tm_z, b, x, level);
        return;
    }
    Tail Tl;
    const int64_t off = tail_setup<T>(mg, level, Tl);
    ProfScope ps(mg->ctx, "mg_tail");
    hipLaunchKernelGGL(k_mg_tail<T>, dim3(1), dim3(MG_TAIL_THREADS), (size_t)off * sizeof(T), mg->ctx->stream, Tl,
                       level, x, b, zc, zero ? 1 : 0, fresh ? 1 : 0);
}

// Tile of the tiled smoother for `level` with halo H: per axis the largest even
// divisor of the level size up to sm_tile, shrunk until the boxes of u (grown by
// H) and b (grown by H - 1) fit in LDS and the level has at least sm_minblocks
// tiles (the tile does not change the results).
template <typename T>
static bool smooth_plan(const mcs_mg *mg, int level, int halo, SmoothArgs &A, size_t &lds, unsigned &nblocks)
{
    const int32_t *n = mg->lv[level].n;
    for (int target = mg->sm_tile; target >= 2; target /= 2) {
        size_t box = sizeof(T), boxb = sizeof(T);
        nblocks = 1;
        for (int d = 0; d < 3; ++d) {
            int t = 2;
            for (int c = 2; c <= target && c <= n[d]; c += 2)
                if (n[d] % c == 0) t = c;
            if (n[d] % t) return false;
            A.tile[d] = t;
            A.ntile[d] = n[d] / t;
            nblocks *= (unsigned)A.ntile[d];
            box *= (size_t)std::min(t + 2 * halo, n[d]);
            boxb *= (size_t)std::min(t + 2 * halo - 2, n[d]);
        }
        if (box + boxb <= mg->sm_lds && (nblocks >= mg->sm_minblocks || target <= 2)) {
            lds = box + boxb;
            return true;
        }
    }
    return false;
}

template <typename T>
static bool smooth_launch(mcs_mg *mg, int level, const T *u_in, const T *b, T *u_out, int passes, int ghost_pass,
                          int resid, const T *xc, T *rh, int interp)
{
    SmoothArgs A;
    memset(&A, 0, sizeof(A));
    A.L = lvl(mg, level);
    A.bc = coefs(mg);
    A.passes = passes;
    A.halo = passes + resid;
    A.ghost_pass = ghost_pass;
    A.resid = resid;
    A.interp = interp;
    size_t lds = 0;
    unsigned nb = 0;
    if (!smooth_plan<T>(mg, level, A.halo, A, lds, nb)) return false;
    ProfScope ps(mg->ctx, "mg_rbgs");
    hipLaunchKernelGGL(k_mg_smooth<T>, dim3(nb), dim3(MG_SM_THREADS), lds, mg->ctx->stream, A, u_in, b, u_out,
                       xc, rh);
    return true;
}

template <typename T>
static bool smooth_fits(mcs_mg *mg, int level)
{
    if (!mg->sm_lds || !lalt<T>(mg, level)) return false;
    if ((int64_t)mg->lv[level].n[0] * mg->lv[level].n[1] * mg->lv[level].n[2] > mg->sm_max) return false;
    SmoothArgs A;
    size_t lds;
    unsigned nb;
    return smooth_plan<T>(mg, level, 2 * mg->cfg.smooth1 + 1, A, lds, nb) &&
           smooth_plan<T>(mg, level, 2 * mg->cfg.smooth2, A, lds, nb);
}

// mgv! (Multigrid.jl:334-377).  apply! before each smoothing sweep and before the
// residual is folded into the kernels; the ghosts are materialised before the
// last post-smoothing sweep (they are what the caller's prolongation reads).
// The call also takes the steps around it: zero starts from x = 0 (the xh .= 0
// of the caller's mgv!); for the coarse-to-fine sweep of fmg!, zc != nullptr
// starts from x = interpolate!(zc) (zc's ghosts materialised) and fresh leaves
// the ghosts of the final x (fmg!'s apply! before its next interpolate!).
// x itself is only written, so it may not alias zc.
template <typename T>
static void mgv(mcs_mg *mg, T *x, const T *b, int level, bool zero, const T *zc = nullptr, bool fresh = false)
{
    if (in_tail<T>(mg, level)) {
        mg_tail<T>(mg, level, x, b, zc, zero, fresh);
        return;
    }
    if (smooth_fits<T>(mg, level)) {
        // pre-smoothing ending with the restricted residual; the coarse cycle;
        // post-smoothing starting with the added prolongation, the ghosts as the
        // last apply! leaves them (before the last sweep).  Each launch reads one
        // of x and y and writes the other; the launch count is kept even so that
        // the result ends in x.
        T *xh = lsol<T>(mg, level + 1);
        T *rh = lres<T>(mg, level + 1);
        const int s1 = mg->cfg.smooth1, s2 = mg->cfg.smooth2;
        int spl = (int64_t)mg->lv[level].n[0] * mg->lv[level].n[1] * mg->lv[level].n[2] >= mg->sm_big ? 1 : 2;
        int k1 = s1 > 0 ? (s1 + spl - 1) / spl : 1, k2 = (s2 + spl - 1) / spl;
        if ((k1 + k2) % 2) {
            spl = 2;
            k1 = 1;
            k2 = 1;
        }
        T *cur = x, *oth = lalt<T>(mg, level);
        int left = s1;
        for (int i = 0; i < k1; ++i) {
            const int sw = std::min(spl, left);
            left -= sw;
            const bool last = i == k1 - 1;
            smooth_launch<T>(mg, level, cur, b, oth, 2 * sw, -1, last ? 1 : 0, i == 0 ? zc : (const T *)nullptr,
                             last ? rh : (T *)nullptr, i > 0 ? 0 : zc ? 2 : zero ? 3 : 0);
            std::swap(cur, oth);
        }
        mgv<T>(mg, xh, rh, level + 1, true);
        left = s2;
        for (int i = 0; i < k2; ++i) {
            const int sw = std::min(spl, left);
            left -= sw;
            const bool last = i == k2 - 1;
            smooth_launch<T>(mg, level, cur, b, oth, 2 * sw, last ? (fresh ? 2 * sw : 2 * sw - 2) : -1, 0,
                             i == 0 ? (const T *)xh : (const T *)nullptr, (T *)nullptr, i == 0 ? 1 : 0);
            std::swap(cur, oth);
        }
        return;
    }
    const bool sp = level == 0 && mg->split0;
    if (zc) mg_interp<T>(mg, level, x, zc, 0, sp);
    else if (zero) mg_fill0<T>(mg, x, mg->lv[level].len);
    for (int it = 0; it < mg->cfg.smooth1; ++it) mg_gs<T>(mg, level, x, b, sp);
    // the correction of level 0 is only added to the solution (interior): its
    // ghost shell is never read, so its apply! launches are left out
    const bool ghosts = level > 0;
    if (mg->cfg.smooth2 == 0 && ghosts) mg_bc<T>(mg, level, x);
    {
        ProfScope ps(mg->ctx, "mg_residual");
        if (sp)
            hipLaunchKernelGGL(k_mg_resid_restrict_s<T>, grid_for(ncoarse_cells(mg, level)), dim3(MCS_BLOCK), 0,
                               mg->ctx->stream, lres<T>(mg, level + 1), (const T *)x, b, lvl(mg, level),
                               coefs(mg), lsol<T>(mg, level + 1));
        else
            hipLaunchKernelGGL(k_mg_resid_restrict<T>, grid_for(ncoarse_cells(mg, level)), dim3(MCS_BLOCK), 0,
                               mg->ctx->stream, lres<T>(mg, level + 1), (const T *)x, b, lvl(mg, level),
                               coefs(mg), lsol<T>(mg, level + 1));
    }
    // xh's interior was zeroed with the residual; a tail reads no ghost before it
    // writes it, so for a tail this start is x = 0 (which a tabulated tail needs)
    T *xh = lsol<T>(mg, level + 1);
    mgv<T>(mg, xh, lres<T>(mg, level + 1), level + 1, in_tail<T>(mg, level + 1));
    mg_interp<T>(mg, level, x, xh, 1, sp);
    // the ghost shell the caller reads: before the last sweep (the second-to-last
    // sweep's colour-1 pass writes it), or after it (fresh: the last pass writes
    // it); a single sweep takes the launch before it
    const int s2 = mg->cfg.smooth2;
    if (ghosts && !fresh && s2 == 1) mg_bc<T>(mg, level, x);
    for (int it = 0; it < s2; ++it) {
        const bool gh = ghosts && (fresh ? it == s2 - 1 : it == s2 - 2);
        mg_gs<T>(mg, level, x, b, sp, gh);
    }
    if (fresh && ghosts && s2 == 0) mg_bc<T>(mg, level, x);
}

// The correction part of fmg! (Multigrid.jl:396-426) on the hierarchy of one
// precision, from the finest residual in lres<T>(0): restrictions, coarse solve
// and the coarse-to-fine sweep of V-cycles.  Returns the finest correction.
template <typename T>
static T *fmg_correction(mcs_mg *mg)
{
    mcs_ctx *ctx = mg->ctx;
    const int L = mg->cfg.levels;
    const int S = fmg_tail_start<T>(mg);
    const int kend = S > 0 ? S - 1 : L;   // restrictions on the grid path
    for (int k = 0; k < kend; ++k) {
        ProfScope ps(ctx, "mg_restrict");
        if (k == 0 && mg->split0)
            hipLaunchKernelGGL(k_mg_restrict_s<T>, grid_for(ncoarse_cells(mg, k)), dim3(MCS_BLOCK), 0, ctx->stream,
                               lres<T>(mg, k + 1), (const T *)lres<T>(mg, k), lvl(mg, k));
        else
            hipLaunchKernelGGL(k_mg_restrict<T>, grid_for(ncoarse_cells(mg, k)), dim3(MCS_BLOCK), 0, ctx->stream,
                               lres<T>(mg, k + 1), (const T *)lres<T>(mg, k), lvl(mg, k));
    }
    int top = L;   // the finest level solved so far
    T *z = lsol<T>(mg, L);
    if (S > 0 && S == mg->tm_level && mg->tm_f) {
        // the restricted right-hand side of S (restrict_item's arithmetic, as the
        // tail forms it), then the tabulated coarse part
        {
            ProfScope ps(ctx, "mg_restrict");
            hipLaunchKernelGGL(k_mg_restrict<T>, grid_for(ncoarse_cells(mg, S - 1)), dim3(MCS_BLOCK), 0, ctx->stream,
                               lres<T>(mg, S), (const T *)lres<T>(mg, S - 1), lvl(mg, S - 1));
        }
        tail_matvec<T>(mg, mg->tm_f, (const T *)lres<T>(mg, S), lsol<T>(mg, S), S);
        top = S;
        z = lsol<T>(mg, S);
    } else if (S > 0) {
        Tail Tl;
        const int64_t words = tail_setup<T>(mg, S, Tl);
        ProfScope ps(ctx, "mg_tail");
        hipLaunchKernelGGL(k_mg_tail_fmg<T>, dim3(1), dim3(MG_TAIL_THREADS), (size_t)words * sizeof(T), ctx->stream,
                           Tl, S, lsol<T>(mg, S), (const T *)lres<T>(mg, S - 1), coarse_inv<T>(mg));
        top = S;
        z = lsol<T>(mg, S);
    } else {
        hipLaunchKernelGGL(k_mg_coarse<T>, dim3(1), dim3(MCS_BLOCK), 0, ctx->stream, coarse_inv<T>(mg),
                           (const T *)lres<T>(mg, L), lsol<T>(mg, L), lvl(mg, L));
        // apply!(z) of the coarse solution; the later ones are left by each mgv!
        // (fresh), which also starts from interpolate!(z1, z)
        mg_bc<T>(mg, L, z);
    }
    for (int k = L - top + 1; k <= L; ++k) {
        int lev = L - k;
        T *z1 = lsol<T>(mg, lev);
        mgv<T>(mg, z1, lres<T>(mg, lev), lev, false, (const T *)z, true);
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
        hipLaunchKernelGGL(k_mg_axpy_bc<float>, grid_for(ncells(mg, 0)), dim3(MCS_BLOCK), 0, ctx->stream, x,
                           (const float *)z, lvl(mg, 0), coefs(mg), mg->split0 ? 1 : 0);
    } else {
        double *z = fmg_correction<double>(mg);
        hipLaunchKernelGGL(k_mg_axpy_bc<double>, grid_for(ncells(mg, 0)), dim3(MCS_BLOCK), 0, ctx->stream, x,
                           (const double *)z, lvl(mg, 0), coefs(mg), mg->split0 ? 1 : 0);
    }
}

// fmg! (Multigrid.jl:380-434).  Returns 1 when a correction was applied, 0 when
// the residual already meets the tolerance, negative on error.  ghosts_fresh:
// x's ghost shell is already what apply!(x) writes (the previous correction of
// this solve ended with it), so the first apply! is skipped.
static int fmg(mcs_mg *mg, double *x, const double *b, double *eps_out, bool ghosts_fresh)
{
    mcs_ctx *ctx = mg->ctx;
    const MgLevel &F = mg->lv[0];
    if (!ghosts_fresh) mg_bc<double>(mg, 0, x);
    {
        ProfScope ps(ctx, "mg_residual");
        hipLaunchKernelGGL(k_mg_residual, grid_for(ncells(mg, 0)), dim3(MCS_BLOCK), 0, ctx->stream,
                           F.res, (const double *)x, b, mg->cfg.s, lvl(mg, 0), coefs(mg), mg->partial,
                           mg->split0 ? 1 : 0);
    }
    const uint64_t seq = mcs_pub_next(ctx);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(MCS_BLOCK), 0, ctx->stream,
                       (const double *)mg->partial, mg->npartial, mg->partial + mg->npartial, ctx->d_pub, seq);
    double ss = 0.0;
    if (hipGetLastError() != hipSuccess || mcs_pub_wait(ctx, seq, &ss, 1) != MCS_OK) return -1;
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
        const uint64_t seq = mcs_pub_next(ctx);
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(MCS_BLOCK), 0, ctx->stream,
                           (const double *)mg->partial, mg->npartial, mg->partial + mg->npartial, ctx->d_pub, seq);
        MCS_LAUNCH_CHECK(ctx);
        double ss = 0.0;
        int32_t rc = mcs_pub_wait(ctx, seq, &ss, 1);
        if (rc) return rc;
        double brms = sqrt(ss) / sqrt((double)F.n[0] * (double)F.n[1] * (double)F.n[2]);
        mg->tol_abs = -mg->cfg.tolerance * brms;
        if (!(mg->tol_abs > 0.0)) mg->tol_abs = 2.2250738585072014e-308;
    }
    int iter = 0;
    int cont = 1;
    double eps = 0.0;
    while (cont && iter <= mg->cfg.maxiter) {
        cont = fmg(mg, u, b, &eps, iter > 0);
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
    MCS_HIP(ctx, mcs_h2d(ctx, mg->u_tmp, u, sizeof(double) * len));
    MCS_HIP(ctx, mcs_h2d(ctx, mg->b_tmp, b, sizeof(double) * len));
    int32_t rc = mg->cfg.ndim == 2 ? mcs_mg2_solve_dev(mg, mg->u_tmp, mg->b_tmp, iters, resid)
                                   : mg_solve_dev(mg, mg->u_tmp, mg->b_tmp, iters, resid);
    MCS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    MCS_HIP(ctx, hipMemcpy(u, mg->u_tmp, sizeof(double) * len, hipMemcpyDeviceToHost));
    return rc;
}

extern "C" int32_t mcs_mg_solve_device(mcs_mg *mg, double *u, const double *b, int32_t *iters, double *resid)
{
    if (!mg || !u || !b) return MCS_E_ARG;
    hipSetDevice(mg->ctx->device);
    return mg->cfg.ndim == 2 ? mcs_mg2_solve_dev(mg, u, b, iters, resid) : mg_solve_dev(mg, u, b, iters, resid);
}

extern "C" int32_t mcs_mg_solve_mesh(mcs_mg *mg, mcs_mesh *m, int32_t *iters, double *resid)
{
    if (!mg || !m) return MCS_E_ARG;
    mcs_ctx *ctx = mg->ctx;
    if (mg->cfg.ndim == 2) return mcs_fail(ctx, MCS_E_ARG, "mg: a 2D solver has no 3D mesh; use the host solve");
    for (int d = 0; d < 3; ++d)
        if (m->n[d] != mg->cfg.n[d]) return mcs_fail(ctx, MCS_E_ARG, "mg: mesh size differs from the solver's");
    hipSetDevice(ctx->device);
    ProfScope ps(ctx, "mg_solve");
    return mg_solve_dev(mg, m->arr[MCS_ARR_U], m->arr[MCS_ARR_Q], iters, resid);
}

extern "C" int32_t mcs_mg_setinhom_mesh(mcs_mg *mg, mcs_mesh *m, int32_t face, const double *val)
{
    if (!mg || !m || !val || face < 0 || face > 5) return MCS_E_ARG;
    mcs_ctx *ctx = mg->ctx;
    if (mg->cfg.ndim == 2) return mcs_fail(ctx, MCS_E_ARG, "mg: a 2D solver has no 3D mesh");
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
