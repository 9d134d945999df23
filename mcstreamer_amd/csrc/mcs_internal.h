// mcs_internal.h -- shared definitions of libmcs (HIP, gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/mcs.h"
#include "mcs_phys.h"
#include "mcs_rng.h"

#pragma clang fp contract(off)

#define MCS_BLOCK 256
#define MCS_WAVE 64

#define MCS_FLAG_TABLE_RANGE 1
#define MCS_FLAG_FIELD 2
#define MCS_FLAG_CAPACITY 4
#define MCS_FLAG_NAN 8

// energy bands of the banded majorant below the unbounded one (caps doubling up
// to the table's last node)
#define MCS_NBAND 20

struct ProfEntry {
    double total_ms = 0.0;
    int64_t launches = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
};

struct mcs_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    bool prof = false;
    std::map<std::string, ProfEntry> profile;
    // small pinned host buffer for read-backs
    int64_t *h_small = nullptr;
    int64_t *d_small = nullptr;   // device counters block (512 int64)
    // growable device scratch used by scans and reductions
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    // host memory (mapped, coherent) the device publishes small results to:
    // [0] the sequence number of the last publication, then its words
    uint64_t *h_pub = nullptr;
    uint64_t *d_pub = nullptr;   // its device address
    uint64_t pub_seq = 0;
};

#define MCS_PUB_WORDS 511

// set error text and return the code
int32_t mcs_fail(mcs_ctx *ctx, int32_t code, const char *fmt, ...);

#define MCS_HIP(ctx, call)                                                                  \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            return mcs_fail((ctx), MCS_E_HIP, "%s failed: %s (%s:%d)", #call,               \
                            hipGetErrorString(e_), __FILE__, __LINE__);                     \
    } while (0)

#define MCS_LAUNCH_CHECK(ctx)                                                               \
    do {                                                                                    \
        hipError_t e_ = hipGetLastError();                                                  \
        if (e_ != hipSuccess)                                                               \
            return mcs_fail((ctx), MCS_E_HIP, "kernel launch failed: %s (%s:%d)",          \
                            hipGetErrorString(e_), __FILE__, __LINE__);                     \
    } while (0)

// Every transfer and fill of the library runs on the context's stream.  That
// stream is non-blocking, so it does not wait for work on the null stream: a
// plain hipMemcpy or hipMemset there is not ordered with the kernels queued on
// ctx->stream, and a fill still running when a later kernel has written its
// buffer wipes what the kernel wrote.
//
// Host-to-device copy queued on ctx->stream and waited for, so that kernels
// queued afterwards read the new data and the host buffer is free on return.
static inline hipError_t mcs_h2d(mcs_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(ctx->stream);
}

// Zero fill queued on ctx->stream (ordered before every kernel queued after it).
// For set-up code only: a fill recorded into a captured graph is not used here.
static inline hipError_t mcs_zero(mcs_ctx *ctx, void *dst, size_t bytes)
{
    return hipMemsetAsync(dst, 0, bytes, ctx->stream);
}

// Records a HIP event pair around the enclosed launches when profiling is on.
struct ProfScope {
    mcs_ctx *ctx;
    const char *name;
    hipEvent_t a = nullptr, b = nullptr;
    ProfScope(mcs_ctx *c, const char *n) : ctx(c), name(n)
    {
        if (ctx->prof) {
            hipEventCreate(&a);
            hipEventCreate(&b);
            hipEventRecord(a, ctx->stream);
        }
    }
    ~ProfScope()
    {
        if (ctx->prof) {
            hipEventRecord(b, ctx->stream);
            ctx->profile[name].pending.emplace_back(a, b);
        }
    }
};

// scratch of at least `bytes`, reused across calls on the context's stream
int32_t mcs_scratch(mcs_ctx *ctx, size_t bytes, void **out);
// Read-back of n (<= MCS_PUB_WORDS) int64 values from device memory: a kernel
// copies them to ctx->h_pub and raises the sequence number, which the host
// waits for (no stream synchronisation, no copy engine).
int32_t mcs_readback(mcs_ctx *ctx, const int64_t *dev, int64_t *host, int n);
// the same for n 8-byte words at dev followed by n2 at dev2
int32_t mcs_readback2(mcs_ctx *ctx, const void *dev, int n, const void *dev2, int n2, void *host);
// For kernels that publish by themselves: the sequence number the next
// publication is to carry, and the wait for it (then n words are copied out).
uint64_t mcs_pub_next(mcs_ctx *ctx);
int32_t mcs_pub_wait(mcs_ctx *ctx, uint64_t seq, void *host, int n);

// Device side of a publication by one thread, after the words are stored to
// pub[1..]: the words are made visible to the host before the sequence number.
__device__ __forceinline__ void mcs_pub_raise(uint64_t *pub, uint64_t seq)
{
    __threadfence_system();
    __hip_atomic_store(pub, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

struct mcs_table {
    mcs_ctx *ctx;
    int32_t nproc, ne, egrid;
    double maxrate;
    double e0;        // first energy node
    double emax;      // last energy node
    double inv_dlog;
    double step;
    double *energy;   // ne
    double *rate;     // ne * nproc
    int32_t *kind;    // nproc
    double *param;    // nproc
    // FP32 copies for the split-precision kernels
    float *energy32 = nullptr, *rate32 = nullptr, *param32 = nullptr;
    // energy bands of the banded majorant (mcs_table_majorants): per band
    // (cap, majorant, 1 / majorant) in FP32, nband bands + the unbounded one
    float *band32 = nullptr;
    int32_t nband = 0;
    std::vector<double> h_band_cap, h_band_rate;
    std::vector<int32_t> h_kind;
    // PhotoEmission rows: (log numin, log numax, photon_multiplier, photon_weight)
    // per row (nproc x 4, device and host), which rows are set, how many rows
    // are photoemission rows
    double *photo = nullptr;
    std::vector<double> h_photo;
    std::vector<char> photo_set;
    int32_t nphoto = 0;
};

// Particle SoA in HBM.  Buffer set 0 is live; set 1 receives sort/resample output.
struct PopBuf {
    double *x[3] = {nullptr, nullptr, nullptr};
    double *v[3] = {nullptr, nullptr, nullptr};
    double *w = nullptr, *s = nullptr;
    uint8_t *active = nullptr;
    uint64_t *id = nullptr;
    double *nu = nullptr;   // photon populations: absorption rate
};

// Split-precision SoA (precision mode): packed cell coordinates (i in bits 0-9,
// j in 10-19, k in 20-30) plus the position inside the cell in cell units, FP32.
struct SplitBuf {
    int32_t *cell = nullptr;
    float *off[3] = {nullptr, nullptr, nullptr};
    float *v[3] = {nullptr, nullptr, nullptr};
    float *w = nullptr, *s = nullptr;
    uint8_t *active = nullptr;
    uint64_t *id = nullptr;
};

struct mcs_pop {
    mcs_ctx *ctx;
    int64_t cap;
    int64_t n;
    uint64_t next_id;
    // Serial of the population's current order and count: every call that moves,
    // adds, removes or reorders particles takes a fresh one (mcs_pop_touch), so a
    // mesh or grid that recorded the serial at its sort knows whether its cell
    // starts still describe this population.
    uint64_t gen = 0;
    PopBuf buf[2];
    bool has_alt = false;
    // precision mode: the split store on a mesh grid (gn cells of gh metres)
    int prec = 0;
    SplitBuf sbuf[2];
    int32_t gn[3] = {0, 0, 0};
    double gh = 0.0;
    int32_t *log_cell = nullptr;
    float *log_off[3] = {nullptr, nullptr, nullptr};
    float *log_v32[3] = {nullptr, nullptr, nullptr};
    float *log_w32 = nullptr, *log_s32 = nullptr;
    // birth log (structure of arrays)
    int64_t logcap = 0;
    int32_t *log_parent = nullptr;
    int32_t *log_seq = nullptr;
    double *log_x[3] = {nullptr, nullptr, nullptr};
    double *log_v[3] = {nullptr, nullptr, nullptr};
    double *log_w = nullptr, *log_s = nullptr;
    uint8_t *log_act = nullptr;
    // per-particle scratch (int32 x cap): local offsets, keys, ranks, permutation
    int32_t *i32a = nullptr, *i32b = nullptr, *i32c = nullptr;
    uint64_t *u64a = nullptr;
    int64_t *blk = nullptr;   // block totals, then their prefix (2 * (cap/256 + 2) + 4)
    void *dep_w = nullptr;     // CIC deposit shares (8 doubles x cap), allocated on first use
    uint64_t *dep_k = nullptr;
    // ion-tracker deposits of a step (logcap each; then their cell-sorted copy),
    // and per-block deposit counts and their prefix; allocated on the first
    // tracking step
    int32_t *trk_parent = nullptr, *trk_seq = nullptr, *trk_cell = nullptr;
    double *trk_val = nullptr;
    int64_t *tblk = nullptr;
    // MCS_SPECIES_PHOTON: a photon population (buf[].nu allocated)
    int species = 0;
    // photon bursts of a step's PhotoEmission collisions (logcap each): parent
    // slot, the parent's photons before the burst, the burst's photon count, table
    // row, active flag and position; per-slot photon offsets within the block and
    // per-block photon totals and their prefix.  Allocated on the first step
    // that emits.
    int32_t *bur_parent = nullptr, *bur_row = nullptr;
    int64_t *bur_base = nullptr, *bur_n = nullptr;
    uint8_t *bur_act = nullptr;
    double *bur_x[3] = {nullptr, nullptr, nullptr};
    int64_t *poff = nullptr, *pblk = nullptr;
};

#define MCS_SPECIES_ELECTRON 0
#define MCS_SPECIES_PHOTON 1

// Entry points of electron populations refuse a photon population.
#define MCS_ELECTRONS_ONLY(p, fn)                                                            \
    do {                                                                                    \
        if ((p)->species == MCS_SPECIES_PHOTON)                                             \
            return mcs_fail((p)->ctx, MCS_E_ARG, fn ": not for a photon population");       \
    } while (0)

struct mcs_mesh {
    mcs_ctx *ctx;
    int32_t n[3];
    double h;
    int64_t ncells;
    int64_t ncg;      // cells with ghosts
    double *arr[7];   // QFIXED..EZ
    int64_t arr_len[7];
    int32_t *cell_start = nullptr;  // ncells + 2
    int32_t *cell_cnt = nullptr;    // ncells + 2 (counts, then the long-segment list)
    int64_t *cnt64 = nullptr;       // ncells + 2 (per-cell resample output counts)
    int64_t *pre64 = nullptr;       // ncells + 2 (their prefix)
    uint64_t sorted_gen = 0;        // mcs_pop::gen of the population the cell starts describe
    float *face32[3] = {nullptr, nullptr, nullptr};  // FP32 face fields (precision mode)
    int32_t *dep_list = nullptr;    // ncells + 2: count, the cells the CIC gather gives a wavefront, a counter
    int32_t *dep_host = nullptr;    // host memory the device writes the count of such cells to
    int32_t *dep_host_dev = nullptr;
    int32_t dep_dense = 0;          // records a cell visits above which it takes a wavefront
    int64_t dep_lane_listed = 0;    // above this many such cells (last call), one lane per cell instead
};

// Axisymmetric grid (mcs_cyl.hip).
struct mcs_cgrid {
    mcs_ctx *ctx;
    int32_t M, N;
    double R, L;
    double dr, dz;     // step(rc), step(zc)
    double drf, dzf;   // step(rf), step(zf)
    double *arr[6];    // QFIXED, QPART, Q, U: (M+2)(N+2); ER, EZ: (M+3)(N+3)
    int64_t arr_len[6];
    double *dv;        // dV(grid, i) for i = 1..M
    int32_t *run_first = nullptr, *run_last = nullptr;   // per cell: its run in the density's sorted pairs
    int32_t *run_long = nullptr;                         // count, then the cells whose run takes a wavefront
    double *fb_S = nullptr;   // free boundary: sin(pi m / 2N), m < 4N
    double *fb_f = nullptr;   // free boundary: f (N), then grad, am, bcond (N each)
    // canonical (cell, id) order over resample!'s cells, (M+2)(N+2) + 2 entries each
    int32_t *cs = nullptr, *cs_new = nullptr;
    int64_t *cnt64 = nullptr, *pre64 = nullptr;
    uint64_t sorted_gen = 0;   // mcs_pop::gen of the population the cell starts describe
    // resample!'s p and wtotal per cell ((M+2)(N+2) each), as the last resample left them
    int64_t *rs_p = nullptr;
    double *rs_w = nullptr;
};

// Allocates (zeroed) the grid's resample p and wtotal arrays if it has none.
int32_t mcs_cgrid_resample_arrays(mcs_cgrid *g);

// The two-dimensional solver (mcs_mg2.hip): creation for cfg->ndim == 2, and the
// solve on device arrays with ghosts.
int32_t mcs_mg2_create(mcs_ctx *ctx, const mcs_mg_config *cfg, mcs_mg **out);
int32_t mcs_mg2_solve_dev(mcs_mg *mg, double *u, const double *b, int32_t *iters, double *resid);

// FP32 copies of the face fields, allocated and filled on first use; from then on
// mcs_efield_faces writes both precisions.
int32_t mcs_mesh_faces32(mcs_mesh *m);

struct MgLevel {
    int32_t n[3];
    int64_t len;
    double *res, *res1, *sol;
    float *fres, *fsol;   // FP32 smoother copies (precision mode)
    double *alt;          // second solution array of the tiled smoother (grid levels)
    float *falt;
};

struct mcs_mg {
    mcs_ctx *ctx;
    mcs_mg_config cfg;
    std::vector<MgLevel> lv;
    double *ainv = nullptr;   // coarse inverse, ncoarse^2
    float *ainv32 = nullptr;  // its FP32 copy (precision mode)
    int64_t ncoarse = 0;
    double *u_tmp = nullptr, *b_tmp = nullptr;  // host-solve staging (finest size)
    double *partial = nullptr;                   // norm partial sums
    int64_t npartial = 0;
    int32_t *bc_dev = nullptr;                   // boundary kinds on the device
    double tol_abs = 0.0;                        // tolerance in force for the current solve
    size_t tail_lds = 0;                         // LDS budget of the single-workgroup tail
    int tail_levels = 0;                         // most levels the tail may hold
    int64_t tail_max = 0;                        // largest level (cells) a tail starts from
    int fmg_tail = 1;                            // fmg!'s tail levels in one launch (k_mg_tail_fmg)
    int tail2 = 0;                               // 2D solver: first level of the one-workgroup V-cycle
    size_t sm_lds = 0;                           // LDS budget of the tiled smoother; 0: one launch per pass
    int sm_tile = 16;                            // its largest tile edge
    unsigned sm_minblocks = 0;                   // smaller tiles until a level has this many blocks
    int64_t sm_big = 0;                          // levels of this many cells: one sweep per launch
    int64_t sm_max = 0;                          // levels of more cells: one launch per pass
    bool split0 = false;                         // level-0 correction in the colour-split layout
    bool use_graph = true;                       // replay the cycle body as a HIP graph
    // Tabulated tails: below the tail's start level the cycle is a fixed linear
    // map of that level's right-hand side (x starts at zero, the coarse boundary
    // rules are homogeneous), kept as a (ghosted length x cells) row-major matrix
    // of the working precision: tm_z for mgv! from x = 0, tm_f for fmg!'s coarse
    // part from the restricted right-hand side.  tm_level -1: none.
    int tm_level = -1;
    void *tm_z = nullptr, *tm_f = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    double *graph_x = nullptr;
    const double *graph_b = nullptr;
};

// ---- device helpers --------------------------------------------------------

__device__ __forceinline__ int64_t gidx() { return (int64_t)blockIdx.x * blockDim.x + threadIdx.x; }

// XCD-aware block order.  gfx950 dispatches the workgroups of a 1D grid
// round-robin over its 8 XCDs (block b on XCD b % 8), each with its own L2.
// Renumbering so that XCD x runs one contiguous range of logical blocks keeps
// neighbouring rows and planes of a stencil, and the cells of neighbouring
// particles, in the same L2.  A bijection on [0, gridDim.x).
#define MCS_NXCD 8u
__device__ __forceinline__ uint32_t xcd_block()
{
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    const uint32_t per = nb / MCS_NXCD, rem = nb % MCS_NXCD;
    const uint32_t x = b % MCS_NXCD, k = b / MCS_NXCD;
    return x * per + (x < rem ? x : rem) + k;
}

__device__ __forceinline__ int64_t xcd_gidx() { return (int64_t)xcd_block() * blockDim.x + threadIdx.x; }

static inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// exclusive scan of int32 counts (len entries) into out (len+1 entries, out[len] = total).
int32_t mcs_scan_i32(mcs_ctx *ctx, const int32_t *in, int32_t *out, int64_t len);
// exclusive scan of int64 values (len entries) into out (len+1 entries).
int32_t mcs_scan_i64(mcs_ctx *ctx, const int64_t *in, int64_t *out, int64_t len);
int32_t mcs_scan_i64_buf(mcs_ctx *ctx, const int64_t *in, int64_t *out, int64_t len, int64_t *bsum);
// exclusive scan of len int64 values into out (len entries) with the total at *total.
int32_t mcs_scan_i64_to(mcs_ctx *ctx, const int64_t *in, int64_t *out, int64_t len, int64_t *total);
#define MCS_SCAN_TILE 2048

int32_t mcs_pop_ensure_alt(mcs_pop *p);
// the tracker's deposit log (allocated once)
int32_t mcs_pop_ensure_track(mcs_pop *p);
// the photon-burst log of an emitting electron population (allocated once)
int32_t mcs_pop_ensure_bursts(mcs_pop *p);
// Gives p a fresh order serial (unique over all populations of the process).
void mcs_pop_touch(mcs_pop *p);
// Whether cell starts recorded at serial `gen` describe p's current order.
static inline bool mcs_pop_sorted(const mcs_pop *p, uint64_t gen) { return gen != 0 && gen == p->gen; }
void mcs_pop_swap(mcs_pop *p);
