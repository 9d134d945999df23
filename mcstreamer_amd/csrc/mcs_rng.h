/*
 * mcs_rng.h -- counter-based random streams and the elementary functions the
 * particle step needs, written once for both the HIP kernels and host C code.
 *
 * Why this header exists: the reference draws every random number from Julia's
 * global task-local RNG (`rand()` in JuMC/src/util.jl:4-17, collisions.jl:166,
 * electron.jl:111-125).  Here every draw comes from a Philox4x32-10 stream keyed
 * per particle, so the GPU and any CPU restatement that includes this header see
 * the same numbers in the same order, whatever the thread schedule.
 *
 * Counter layout of one stream (all words uint32):
 *   key     = (lo32(seed), hi32(seed) + instance)
 *   counter = (block, step, lo32(id), (hi32(id) & 0xFFFFFF) | phase << 24)
 * A stream hands out doubles two per Philox block (each double uses two words),
 * or floats four per block.  `phase` separates the consumers of one particle id
 * in one step (advance, global resample, per-cell resample, initialisation).
 *
 * Elementary functions (log, exp, sin/cos of 2*pi*u) are evaluated with plain
 * IEEE add/mul/div only (compile with contraction off), so host and device
 * results are bitwise identical.  They are series evaluations derived from the
 * mathematics (atanh series of log, Taylor series of exp/sin/cos after exact
 * argument reduction); accuracy is within a few ulp of the correctly rounded
 * value, which the tests check against libm.
 */
#ifndef MCS_RNG_H
#define MCS_RNG_H

#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MCS_FN __host__ __device__ static inline
#else
#define MCS_FN static inline
#endif

#if defined(__clang__)
#pragma clang fp contract(off)
#endif

/* Philox4x32 multipliers and Weyl key increments (Salmon et al., SC'11). */
#define MCS_PHILOX_M0 0xD2511F53u
#define MCS_PHILOX_M1 0xCD9E8D57u
#define MCS_PHILOX_W0 0x9E3779B9u
#define MCS_PHILOX_W1 0xBB67AE85u

#define MCS_PHASE_ADVANCE 0u
#define MCS_PHASE_SWARM_RESAMPLE 1u
#define MCS_PHASE_CELL_RESAMPLE 2u
#define MCS_PHASE_INIT 3u
#define MCS_PHASE_PHOTON 4u   /* a photon's flights (ZhelezniakCollisions) */
#define MCS_PHASE_EMIT 5u     /* a new photon's direction, nu and s */

MCS_FN void mcs_philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                              uint32_t k0, uint32_t k1, uint32_t out[4])
{
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += MCS_PHILOX_W0;
            k1 += MCS_PHILOX_W1;
        }
        uint64_t p0 = (uint64_t)MCS_PHILOX_M0 * (uint64_t)c0;
        uint64_t p1 = (uint64_t)MCS_PHILOX_M1 * (uint64_t)c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

typedef struct mcs_stream {
    uint32_t k0, k1;       /* key */
    uint32_t c1, c2, c3;   /* fixed counter words */
    uint32_t blk;          /* next block index (counter word 0) */
    uint32_t w0, w1, w2, w3; /* current block; named words keep them in registers */
    int pos;               /* next unused word of the block; 4 = empty */
} mcs_stream;

MCS_FN void mcs_stream_init(mcs_stream *st, uint64_t seed, uint32_t instance,
                            uint32_t phase, uint64_t id, uint32_t step)
{
    st->k0 = (uint32_t)seed;
    st->k1 = (uint32_t)(seed >> 32) + instance;
    st->c1 = step;
    st->c2 = (uint32_t)id;
    st->c3 = ((uint32_t)(id >> 32) & 0x00FFFFFFu) | (phase << 24);
    st->blk = 0;
    st->pos = 4;
    st->w0 = st->w1 = st->w2 = st->w3 = 0;
}

MCS_FN void mcs_stream_refill(mcs_stream *st)
{
    uint32_t o[4];
    mcs_philox4x32_10(st->blk, st->c1, st->c2, st->c3, st->k0, st->k1, o);
    st->w0 = o[0];
    st->w1 = o[1];
    st->w2 = o[2];
    st->w3 = o[3];
    st->blk += 1;
    st->pos = 0;
}

/* Uniform double on [0, 1) with 53 random bits (two words: (w0, w1) then (w2, w3)). */
MCS_FN double mcs_u01(mcs_stream *st)
{
    if (st->pos > 2) mcs_stream_refill(st);
    uint32_t a = st->pos == 0 ? st->w0 : st->w2;
    uint32_t b = st->pos == 0 ? st->w1 : st->w3;
    st->pos += 2;
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

/* Uniform float on [0, 1) with 24 random bits (one word). */
MCS_FN float mcs_u01f(mcs_stream *st)
{
    if (st->pos > 3) mcs_stream_refill(st);
    uint32_t a = st->pos == 0 ? st->w0 : (st->pos == 1 ? st->w1 : (st->pos == 2 ? st->w2 : st->w3));
    st->pos += 1;
    return (float)(a >> 8) * (1.0f / 16777216.0f);
}

/* ---- elementary functions ------------------------------------------------ */

typedef union { double d; uint64_t u; } mcs_dbits;

#define MCS_LN2_HI 0.6931471805592082      /* ln 2 rounded to 40 significant bits */
#define MCS_LN2_LO 7.371002565167799e-13   /* ln 2 - MCS_LN2_HI */
#define MCS_INV_LN2 1.4426950408889634
#define MCS_SQRT2 1.4142135623730951
#define MCS_PI_HALF 1.5707963267948966

/* Natural log for positive finite normal x.
 * x = 2^e m, m in (sqrt(2)/2, sqrt(2)], f = m - 1 (exact), s = f/(2+f),
 * log(1+f) = 2 atanh(s) = 2 s (1 + s^2/3 + s^4/5 + ...), |s| < 0.1716. */
MCS_FN double mcs_log(double x)
{
    mcs_dbits b;
    b.d = x;
    int e = (int)((b.u >> 52) & 0x7ffu) - 1023;
    b.u = (b.u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;
    double m = b.d;
    if (m > MCS_SQRT2) {
        m = m * 0.5;
        e += 1;
    }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double p = 0.08;                     /* 2/25 */
    p = 0.08695652173913043 + z * p;     /* 2/23 */
    p = 0.09523809523809523 + z * p;     /* 2/21 */
    p = 0.10526315789473684 + z * p;     /* 2/19 */
    p = 0.11764705882352941 + z * p;     /* 2/17 */
    p = 0.13333333333333333 + z * p;     /* 2/15 */
    p = 0.15384615384615385 + z * p;     /* 2/13 */
    p = 0.18181818181818182 + z * p;     /* 2/11 */
    p = 0.2222222222222222 + z * p;      /* 2/9 */
    p = 0.2857142857142857 + z * p;      /* 2/7 */
    p = 0.4 + z * p;                     /* 2/5 */
    p = 0.6666666666666666 + z * p;      /* 2/3 */
    double r = s * (2.0 + z * p);
    double de = (double)e;
    return de * MCS_LN2_HI + (r + de * MCS_LN2_LO);
}

/* exp(x): x = k ln2 + r with |r| <= ln2/2, Taylor series of e^r to degree 13,
 * then exact scaling by 2^k. */
MCS_FN double mcs_exp(double x)
{
    if (x < -745.0) return 0.0;
    if (x > 709.0) return HUGE_VAL;
    double k = rint(x * MCS_INV_LN2);
    double r = (x - k * MCS_LN2_HI) - k * MCS_LN2_LO;
    double q = 1.6059043836821613e-10;       /* 1/13! */
    q = 2.08767569878681e-09 + r * q;        /* 1/12! */
    q = 2.505210838544172e-08 + r * q;       /* 1/11! */
    q = 2.755731922398589e-07 + r * q;       /* 1/10! */
    q = 2.7557319223985893e-06 + r * q;      /* 1/9! */
    q = 2.48015873015873e-05 + r * q;        /* 1/8! */
    q = 0.0001984126984126984 + r * q;       /* 1/7! */
    q = 0.001388888888888889 + r * q;        /* 1/6! */
    q = 0.008333333333333333 + r * q;        /* 1/5! */
    q = 0.041666666666666664 + r * q;        /* 1/4! */
    q = 0.16666666666666666 + r * q;         /* 1/3! */
    q = 0.5 + r * q;                         /* 1/2! */
    q = 1.0 + r * q;                         /* 1/1! */
    double y = 1.0 + r * q;
    int ki = (int)k;
    if (ki < -1000) {
        y = y * 8.673617379884035e-19;       /* 2^-60 */
        ki += 60;
    }
    mcs_dbits sc;
    sc.u = (uint64_t)(ki + 1023) << 52;
    return y * sc.d;
}

/* sin(2 pi u) and cos(2 pi u).  4u = q + t with q integer and |t| <= 1/2 is
 * exact; theta = t pi/2 lies in [-pi/4, pi/4] where Taylor series to degree
 * 19/20 are accurate to below 1e-20; the quadrant q selects the signs. */
MCS_FN void mcs_sincos2pi(double u, double *sn, double *cs)
{
    double y = 4.0 * u;
    double q = rint(y);
    double t = (y - q) * MCS_PI_HALF;
    double t2 = t * t;
    double ps = -8.22063524662433e-18;       /* -1/19! */
    ps = 2.8114572543455206e-15 + t2 * ps;   /*  1/17! */
    ps = -7.647163731819816e-13 + t2 * ps;   /* -1/15! */
    ps = 1.6059043836821613e-10 + t2 * ps;   /*  1/13! */
    ps = -2.505210838544172e-08 + t2 * ps;   /* -1/11! */
    ps = 2.7557319223985893e-06 + t2 * ps;   /*  1/9!  */
    ps = -0.0001984126984126984 + t2 * ps;   /* -1/7!  */
    ps = 0.008333333333333333 + t2 * ps;     /*  1/5!  */
    ps = -0.16666666666666666 + t2 * ps;     /* -1/3!  */
    double s = t + t * (t2 * ps);
    double pc = 4.110317623312165e-19;       /*  1/20! */
    pc = -1.5619206968586225e-16 + t2 * pc;  /* -1/18! */
    pc = 4.779477332387385e-14 + t2 * pc;    /*  1/16! */
    pc = -1.1470745597729725e-11 + t2 * pc;  /* -1/14! */
    pc = 2.08767569878681e-09 + t2 * pc;     /*  1/12! */
    pc = -2.755731922398589e-07 + t2 * pc;   /* -1/10! */
    pc = 2.48015873015873e-05 + t2 * pc;     /*  1/8!  */
    pc = -0.001388888888888889 + t2 * pc;    /* -1/6!  */
    pc = 0.041666666666666664 + t2 * pc;     /*  1/4!  */
    pc = -0.5 + t2 * pc;                     /* -1/2!  */
    double c = 1.0 + t2 * pc;
    int qi = ((int)q) & 3;
    if (qi == 0) { *sn = s;  *cs = c;  }
    else if (qi == 1) { *sn = c;  *cs = -s; }
    else if (qi == 2) { *sn = -s; *cs = -c; }
    else { *sn = -c; *cs = s; }
}

/* ---- the reference's random primitives on top of a stream --------------- */

/* nextcoll() = -log(rand())  (JuMC/src/util.jl:17).  1 - u lies in (0, 1], so
 * the log is finite; the subtraction is exact for 53-bit u. */
MCS_FN double mcs_nextcoll(mcs_stream *st)
{
    double u = mcs_u01(st);
    return -mcs_log(1.0 - u);
}

/* randsphere() (JuMC/src/util.jl:4-12): phi = 2 pi rand(), then u = 2 rand() - 1,
 * returns (sqrt(1-u^2) cos phi, sqrt(1-u^2) sin phi, u).  Two draws, phi first. */
MCS_FN void mcs_randsphere(mcs_stream *st, double out[3])
{
    double uphi = mcs_u01(st);
    double sp, cp;
    mcs_sincos2pi(uphi, &sp, &cp);
    double u = 2.0 * mcs_u01(st) - 1.0;
    double v = sqrt(1.0 - u * u);
    out[0] = v * cp;
    out[1] = v * sp;
    out[2] = u;
}

/* Poisson(lambda) by sequential inversion of one uniform per chunk of mean at
 * most 500 (exp(-500) is a normal double).  Stands in for Distributions.jl's
 * Poisson sampler used by Swarm.resample! (Swarm/src/Swarm.jl:216). */
MCS_FN int64_t mcs_poisson(mcs_stream *st, double lambda)
{
    int64_t total = 0;
    while (lambda > 0.0) {
        double lam = lambda > 500.0 ? 500.0 : lambda;
        lambda = lambda - lam;
        double u = mcs_u01(st);
        double p = mcs_exp(-lam);
        double F = p;
        int64_t k = 0;
        while (u > F && k < 100000) {
            k += 1;
            p = p * lam / (double)k;
            F = F + p;
            if (p == 0.0 && (double)k > lam) break;
        }
        total += k;
    }
    return total;
}

/* samplemin(m) = 1 - (1 - rand())^(1/m) (MCStreamer/src/resample.jl:84). */
MCS_FN double mcs_samplemin(mcs_stream *st, double m)
{
    double u = mcs_u01(st);
    return 1.0 - mcs_exp(mcs_log(1.0 - u) / m);
}

/* Initial swarm electron (ElectronState(x0, v0) of Swarm/src/Swarm.jl:104-106 with
 * a Maxwellian instead of v0 = 0): velocity components vth * N(0,1) by two
 * Box-Muller pairs (u1,u2) -> (n0,n1), (u3,u4) -> n2, then s = nextcoll()
 * (electron.jl:17).  Drawn from phase INIT of particle `id`, step 0. */
MCS_FN void mcs_maxwell_init(uint64_t seed, uint32_t instance, uint64_t id, double vth,
                             double v[3], double *s)
{
    mcs_stream st;
    mcs_stream_init(&st, seed, instance, MCS_PHASE_INIT, id, 0u);
    double u1 = mcs_u01(&st), u2 = mcs_u01(&st);
    double r = sqrt(-2.0 * mcs_log(1.0 - u1));
    double sn, cs;
    mcs_sincos2pi(u2, &sn, &cs);
    v[0] = vth * (r * cs);
    v[1] = vth * (r * sn);
    double u3 = mcs_u01(&st), u4 = mcs_u01(&st);
    r = sqrt(-2.0 * mcs_log(1.0 - u3));
    mcs_sincos2pi(u4, &sn, &cs);
    v[2] = vth * (r * cs);
    *s = mcs_nextcoll(&st);
}

/* One photon of a PhotoEmission collision (JuMC/src/collisions.jl:134-139):
 * v = c randsphere(), nu = exp(lmin + (lmax - lmin) rand()), then s = nextcoll(),
 * drawn in that order from phase EMIT of the photon's own id.  c is the speed of
 * light (299792458 m/s). */
MCS_FN void mcs_photon_emit(uint64_t seed, uint32_t instance, uint64_t id, uint32_t step, double lmin,
                            double lmax, double v[3], double *nu, double *s)
{
    mcs_stream st;
    mcs_stream_init(&st, seed, instance, MCS_PHASE_EMIT, id, step);
    double om[3];
    mcs_randsphere(&st, om);
    v[0] = om[0] * 299792458.0;
    v[1] = om[1] * 299792458.0;
    v[2] = om[2] * 299792458.0;
    double u = mcs_u01(&st);
    *nu = mcs_exp(lmin + (lmax - lmin) * u);
    *s = mcs_nextcoll(&st);
}

#endif /* MCS_RNG_H */
