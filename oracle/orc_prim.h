/*
 * orc_prim.h - oracle primitives: Philox-4x32-10 streams and deterministic
 * FP64 elementary functions.  TEST INFRASTRUCTURE ONLY (the CPU oracle).
 *
 * Philox-4x32-10: Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11; checked against the Random123 known-answer
 * vectors in tests/golden/philox_kat.json.  It replaces OpenFOAM's
 * sequential Random stream (uniGasCloud.C:490, one clock-seeded stream per
 * rank), which cannot be reproduced; parity with the reference is therefore
 * statistical, and bit-exactness is defined against this restatement.
 *
 * det_log / det_exp / det_sincos2pi: Cody-Waite argument reduction and
 * truncated Taylor/atanh series using only +, -, *, / and floor, so that a
 * second implementation with the same operation order (the HIP kernels)
 * rounds identically.  Checked against libm to a few ulp in
 * tests/test_oracle_prims.py.  Build with -ffp-contract=off.
 */
#ifndef ORC_PRIM_H
#define ORC_PRIM_H

#include <math.h>
#include <stdint.h>
#include <string.h>

/* ---------------------------------------------------------------- Philox */

static inline void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Stream kinds (top 8 bits of counter word 3). */
enum {
    ORC_RNG_MOVE = 1,        /* id = particle index in case, per step   */
    ORC_RNG_NTC = 2,         /* id = cell, per step                     */
    ORC_RNG_BGK = 3,         /* id = cell, per step                     */
    ORC_RNG_INFLOW_N = 4,    /* id = inflow face-species, per step      */
    ORC_RNG_INFLOW_P = 5,    /* id = inflow face-species, sub = index   */
    ORC_RNG_FILL = 6,        /* id = cell, sub = species                */
    ORC_RNG_BGK_V = 7,       /* id = cell, sub = selection index        */
    ORC_RNG_WEIGHT = 8       /* id = slot (inflow: 2^31 | index), per step */
};

typedef struct {
    uint32_t key[2];
    uint32_t ctr[4];
    uint32_t buf[4];
    int half;                /* 0: next pair is buf[0..1], 1: buf[2..3], 2: empty */
} orc_rng;

static inline void orc_rng_init(orc_rng* r, uint64_t seed, uint32_t kind,
                                uint32_t sub, uint32_t id, uint32_t step)
{
    r->key[0] = (uint32_t)seed;
    r->key[1] = (uint32_t)(seed >> 32);
    r->ctr[0] = 0u;
    r->ctr[1] = id;
    r->ctr[2] = step;
    r->ctr[3] = (kind << 24) | (sub & 0xFFFFFFu);
    r->half = 2;
}

static inline uint64_t orc_rng_u64(orc_rng* r)
{
    if (r->half == 2) {
        orc_philox(r->ctr, r->key, r->buf);
        r->ctr[0] += 1u;
        r->half = 0;
    }
    uint64_t v;
    if (r->half == 0) {
        v = ((uint64_t)r->buf[0] << 32) | r->buf[1];
        r->half = 1;
    } else {
        v = ((uint64_t)r->buf[2] << 32) | r->buf[3];
        r->half = 2;
    }
    return v;
}

/* Uniform on (0, 1]: 53 random bits. */
static inline double orc_uniform(orc_rng* r)
{
    return (double)((orc_rng_u64(r) >> 11) + 1u) * 0x1.0p-53;
}

/* Uniform integer in [0, n): multiply-shift on the top 32 bits. */
static inline uint32_t orc_index(orc_rng* r, uint32_t n)
{
    return (uint32_t)(((orc_rng_u64(r) >> 32) * (uint64_t)n) >> 32);
}

/* ------------------------------------------------------ deterministic math */

#define ORC_LN2_HI 6.93147180369123816490e-01
#define ORC_LN2_LO 1.90821492927058770002e-10
#define ORC_INV_LN2 1.44269504088896338700e+00
#define ORC_TWO_PI 6.28318530717958647692e+00

static inline double orc_as_double(uint64_t b) { double d; memcpy(&d, &b, 8); return d; }
static inline uint64_t orc_as_bits(double d) { uint64_t b; memcpy(&b, &d, 8); return b; }

/* Natural log for finite x > 0: x = 2^e m, m in [sqrt(1/2), sqrt(2)],
 * log m = 2 atanh(s), s = (m-1)/(m+1), series to s^23. */
static inline double orc_log(double x)
{
    if (x < 2.2250738585072014e-308) x = 2.2250738585072014e-308;
    uint64_t b = orc_as_bits(x);
    int e = (int)((b >> 52) & 0x7FF) - 1023;
    double m = orc_as_double((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
    if (m > 1.4142135623730951) { m = m * 0.5; e = e + 1; }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double p = 2.0 / 23.0;
    p = p * z + 2.0 / 21.0;
    p = p * z + 2.0 / 19.0;
    p = p * z + 2.0 / 17.0;
    p = p * z + 2.0 / 15.0;
    p = p * z + 2.0 / 13.0;
    p = p * z + 2.0 / 11.0;
    p = p * z + 2.0 / 9.0;
    p = p * z + 2.0 / 7.0;
    p = p * z + 2.0 / 5.0;
    p = p * z + 2.0 / 3.0;
    double lm = 2.0 * s + s * (z * p);
    double de = (double)e;
    return de * ORC_LN2_HI + (de * ORC_LN2_LO + lm);
}

/* exp(x): x = k ln2 + r, |r| <= ln2/2, Taylor to r^13. */
static inline double orc_exp(double x)
{
    if (x < -745.0) return 0.0;
    if (x > 709.0) x = 709.0;
    double kd = floor(x * ORC_INV_LN2 + 0.5);
    double r = (x - kd * ORC_LN2_HI) - kd * ORC_LN2_LO;
    double p = 1.0 / 6227020800.0;           /* 1/13! */
    p = p * r + 1.0 / 479001600.0;            /* 1/12! */
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    int k = (int)kd;
    if (k < -1020) {
        p = p * orc_as_double((uint64_t)(k + 1000 + 1023) << 52);
        return p * orc_as_double((uint64_t)(1023 - 1000) << 52);
    }
    return p * orc_as_double((uint64_t)(k + 1023) << 52);
}

static inline double orc_pow(double x, double y) { return orc_exp(y * orc_log(x)); }

/* sin and cos of 2*pi*u: quarter-turn reduction (exact), Taylor to t^20. */
static inline void orc_sincos2pi(double u, double* s, double* c)
{
    double qd = floor(u * 4.0 + 0.5);
    double r = u - qd * 0.25;
    int q = ((int)qd) & 3;
    double t = r * ORC_TWO_PI;
    double t2 = t * t;
    double ps = -1.0 / 121645100408832000.0;  /* -1/19! */
    ps = ps * t2 + 1.0 / 355687428096000.0;   /*  1/17! */
    ps = ps * t2 - 1.0 / 1307674368000.0;     /* -1/15! */
    ps = ps * t2 + 1.0 / 6227020800.0;        /*  1/13! */
    ps = ps * t2 - 1.0 / 39916800.0;          /* -1/11! */
    ps = ps * t2 + 1.0 / 362880.0;            /*  1/9!  */
    ps = ps * t2 - 1.0 / 5040.0;              /* -1/7!  */
    ps = ps * t2 + 1.0 / 120.0;               /*  1/5!  */
    ps = ps * t2 - 1.0 / 6.0;                 /* -1/3!  */
    double sn = t + t * (t2 * ps);
    double pc = 1.0 / 2432902008176640000.0;  /*  1/20! */
    pc = pc * t2 - 1.0 / 6402373705728000.0;  /* -1/18! */
    pc = pc * t2 + 1.0 / 20922789888000.0;    /*  1/16! */
    pc = pc * t2 - 1.0 / 87178291200.0;       /* -1/14! */
    pc = pc * t2 + 1.0 / 479001600.0;         /*  1/12! */
    pc = pc * t2 - 1.0 / 3628800.0;           /* -1/10! */
    pc = pc * t2 + 1.0 / 40320.0;             /*  1/8!  */
    pc = pc * t2 - 1.0 / 720.0;               /* -1/6!  */
    pc = pc * t2 + 1.0 / 24.0;                /*  1/4!  */
    pc = pc * t2 - 0.5;                       /* -1/2!  */
    double cs = 1.0 + t2 * pc;
    switch (q) {
    case 0: *s = sn;  *c = cs;  break;
    case 1: *s = cs;  *c = -sn; break;
    case 2: *s = -sn; *c = -cs; break;
    default: *s = -cs; *c = sn; break;
    }
}

/* Two standard normals by Box-Muller from two uniforms. */
static inline void orc_gauss2(orc_rng* r, double* g0, double* g1)
{
    double u1 = orc_uniform(r);
    double u2 = orc_uniform(r);
    double rad = sqrt(-2.0 * orc_log(u1));
    double sn, cs;
    orc_sincos2pi(u2, &sn, &cs);
    *g0 = rad * cs;
    *g1 = rad * sn;
}

/* Three standard normals (two Box-Muller pairs, the fourth value unused). */
static inline void orc_gauss3(orc_rng* r, double g[3])
{
    double a, b, c, d;
    orc_gauss2(r, &a, &b);
    orc_gauss2(r, &c, &d);
    (void)d;
    g[0] = a; g[1] = b; g[2] = c;
}

#endif /* ORC_PRIM_H */
