// ugx_dev.hpp - device primitives and per-simulator physics of the particle
// time step, for gfx950.  Compiled with -ffp-contract=off: every FP64
// expression rounds exactly as written, which makes the FP64 path
// bit-identical to a CPU restatement with the same operation order.
//
// Reference routines restated here (paths relative to src/lagrangian/uniGas/):
//   uniGasParcel::move / hitWallPatch        parcels/uniGasParcel.C:35-145
//   specular / diffuse reflection            boundaries/basic/uniGasPatchBoundary/uniGasPatchBoundary.C:305-403
//   freestream insertion (Bird 12.5)         boundaries/basic/uniGasGeneralBoundary/uniGasGeneralBoundary.C:537-761
//   equipartition / LB rotational helpers    clouds/uniGasCloud.C:975-1017, 1129-1189
//   VHS / VSS / Larsen-Borgnakke collide      dsmcCollisions/derived/*/*.C
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ugx.h"

namespace ugx {

constexpr double kB = UGX_KB;
constexpr double kPi = 3.14159265358979323846;
constexpr double kVSmall = 1.0e-300;
constexpr double kSmall = 1.0e-15;
constexpr int kMaxFaces = UGX_MAX_CELL_FACES;
constexpr int32_t kNbrEmpty = INT32_MIN;
// Cell-face kinds for the hybrid decomposition's face interpolation.
constexpr int8_t kFaceSkip = 0;       // padded slot or empty patch: no face value
constexpr int8_t kFaceInternal = 1;
constexpr int8_t kFaceCyclic = 2;
constexpr int8_t kFaceBoundary = 3;   // zeroGradient: face value = cell value
constexpr int8_t kFaceSymmetry = 4;   // symmetryPlane: vectors lose their normal component
constexpr uint32_t kDead = 0xFFFFFFFFu;
constexpr int kMaxCrossings = 1000;
constexpr int kNS = UGX_MAX_SPECIES;
constexpr int kNSamp = UGX_S_COUNT;
// wall-face geometry record: unit outward normal (3), area, tangent t1 (3), pad
constexpr int kWGeo = 8;

enum RngKind : uint32_t {
    kRngMove = 1, kRngNtc = 2, kRngBgk = 3, kRngInflowN = 4, kRngInflowP = 5, kRngFill = 6,
    kRngWeight = 8
};

// ------------------------------------------------------------ Philox RNG
// Philox-4x32-10 (Salmon et al., SC'11).  Counter words: {draw block, entity,
// step, kind<<24 | sub}; key = case seed.

struct Rng {
    uint32_t k0, k1;
    uint32_t c0, c1, c2, c3;
    uint32_t b0, b1, b2, b3;
    int half;
};

__host__ __device__ inline void philox10(uint32_t& c0, uint32_t& c1, uint32_t& c2, uint32_t& c3,
                                         uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

__host__ __device__ inline void rng_init(Rng& r, uint32_t seed_lo, uint32_t seed_hi, uint32_t kind,
                                         uint32_t sub, uint32_t id, uint32_t step)
{
    r.k0 = seed_lo; r.k1 = seed_hi;
    r.c0 = 0u; r.c1 = id; r.c2 = step; r.c3 = (kind << 24) | (sub & 0xFFFFFFu);
    r.half = 2;
}

__host__ __device__ inline uint64_t rng_u64(Rng& r)
{
    if (r.half == 2) {
        uint32_t a = r.c0, b = r.c1, c = r.c2, d = r.c3;
        philox10(a, b, c, d, r.k0, r.k1);
        r.b0 = a; r.b1 = b; r.b2 = c; r.b3 = d;
        r.c0 += 1u;
        r.half = 0;
    }
    uint64_t v;
    if (r.half == 0) { v = ((uint64_t)r.b0 << 32) | r.b1; r.half = 1; }
    else { v = ((uint64_t)r.b2 << 32) | r.b3; r.half = 2; }
    return v;
}

// Uniform on (0, 1].
__host__ __device__ inline double uniform(Rng& r)
{
    return (double)((rng_u64(r) >> 11) + 1u) * 0x1.0p-53;
}

// Uniform integer in [0, n).
__host__ __device__ inline uint32_t uindex(Rng& r, uint32_t n)
{
    return (uint32_t)(((rng_u64(r) >> 32) * (uint64_t)n) >> 32);
}

// ----------------------------------------------------- deterministic math
// Cody-Waite reduction + truncated series; only +, -, *, /, floor.

constexpr double kLn2Hi = 6.93147180369123816490e-01;
constexpr double kLn2Lo = 1.90821492927058770002e-10;
constexpr double kInvLn2 = 1.44269504088896338700e+00;
constexpr double kTwoPi = 6.28318530717958647692e+00;
constexpr double kSqrtHalf = 7.07106781186547524401e-01;   // 1/sqrt(2): the BGK-family draws' N(0, 1/2)
constexpr double kThird = 1.0 / 3.0;

__host__ __device__ inline double bits_to_double(uint64_t b)
{
    union { uint64_t u; double d; } x;
    x.u = b;
    return x.d;
}

__host__ __device__ inline uint64_t double_to_bits(double d)
{
    union { uint64_t u; double d; } x;
    x.d = d;
    return x.u;
}

__host__ __device__ inline double dlog(double x)
{
    if (x < 2.2250738585072014e-308) x = 2.2250738585072014e-308;
    uint64_t b = double_to_bits(x);
    int e = (int)((b >> 52) & 0x7FF) - 1023;
    double m = bits_to_double((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
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
    return de * kLn2Hi + (de * kLn2Lo + lm);
}

__host__ __device__ inline double dexp(double x)
{
    if (x < -745.0) return 0.0;
    if (x > 709.0) x = 709.0;
    double kd = floor(x * kInvLn2 + 0.5);
    double r = (x - kd * kLn2Hi) - kd * kLn2Lo;
    double p = 1.0 / 6227020800.0;
    p = p * r + 1.0 / 479001600.0;
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
        p = p * bits_to_double((uint64_t)(k + 1000 + 1023) << 52);
        return p * bits_to_double((uint64_t)(1023 - 1000) << 52);
    }
    return p * bits_to_double((uint64_t)(k + 1023) << 52);
}

__host__ __device__ inline double dpow(double x, double y) { return dexp(y * dlog(x)); }

__host__ __device__ inline void dsincos2pi(double u, double& s, double& c)
{
    double qd = floor(u * 4.0 + 0.5);
    double r = u - qd * 0.25;
    int q = ((int)qd) & 3;
    double t = r * kTwoPi;
    double t2 = t * t;
    double ps = -1.0 / 121645100408832000.0;
    ps = ps * t2 + 1.0 / 355687428096000.0;
    ps = ps * t2 - 1.0 / 1307674368000.0;
    ps = ps * t2 + 1.0 / 6227020800.0;
    ps = ps * t2 - 1.0 / 39916800.0;
    ps = ps * t2 + 1.0 / 362880.0;
    ps = ps * t2 - 1.0 / 5040.0;
    ps = ps * t2 + 1.0 / 120.0;
    ps = ps * t2 - 1.0 / 6.0;
    double sn = t + t * (t2 * ps);
    double pc = 1.0 / 2432902008176640000.0;
    pc = pc * t2 - 1.0 / 6402373705728000.0;
    pc = pc * t2 + 1.0 / 20922789888000.0;
    pc = pc * t2 - 1.0 / 87178291200.0;
    pc = pc * t2 + 1.0 / 479001600.0;
    pc = pc * t2 - 1.0 / 3628800.0;
    pc = pc * t2 + 1.0 / 40320.0;
    pc = pc * t2 - 1.0 / 720.0;
    pc = pc * t2 + 1.0 / 24.0;
    pc = pc * t2 - 0.5;
    double cs = 1.0 + t2 * pc;
    if (q == 0) { s = sn; c = cs; }
    else if (q == 1) { s = cs; c = -sn; }
    else if (q == 2) { s = -sn; c = -cs; }
    else { s = -cs; c = sn; }
}

__host__ __device__ inline void gauss2(Rng& r, double& g0, double& g1)
{
    double u1 = uniform(r);
    double u2 = uniform(r);
    double rad = sqrt(-2.0 * dlog(u1));
    double sn, cs;
    dsincos2pi(u2, sn, cs);
    g0 = rad * cs;
    g1 = rad * sn;
}

__host__ __device__ inline void gauss3(Rng& r, double& a, double& b, double& c)
{
    double d;
    gauss2(r, a, b);
    gauss2(r, c, d);
}

// Single-precision draws for the FP32-state path, where the drawn
// velocities are stored in FP32: hardware log2 and sin/cos (argument in
// revolutions), 24-bit uniforms.  The FP64 path uses the deterministic
// functions above, which the oracle restates bit for bit.
// Uniform on (0, 1] from the top 24 bits of a word, and on [0, 1) from 24 bits.
__device__ inline float uniform24_open0(uint32_t w) { return (float)((w >> 8) + 1u) * 0x1.0p-24f; }
__device__ inline float uniform24(uint32_t w) { return (float)(w >> 8) * 0x1.0p-24f; }

__device__ inline void gauss2f(Rng& r, float& g0, float& g1)
{
    const uint64_t v = rng_u64(r);
    const float u1 = uniform24_open0((uint32_t)(v >> 32));
    const float u2 = uniform24((uint32_t)v);
    // -2 ln u1 = -2 ln2 log2 u1
    const float rad = __builtin_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u1));
    g0 = rad * __builtin_amdgcn_cosf(u2);
    g1 = rad * __builtin_amdgcn_sinf(u2);
}

__device__ inline void gauss3f(Rng& r, float& a, float& b, float& c)
{
    float d;
    gauss2f(r, a, b);
    gauss2f(r, c, d);
}

__device__ inline float uniformf(Rng& r) { return uniform24_open0((uint32_t)(rng_u64(r) >> 32)); }

// ------------------------------------------------------------ world view

template <class R>
struct alignas(4 * sizeof(R)) V4 {
    R x, y, z, w;
};

// Per-case parameters (device copy; step advanced on the device).
struct CaseDev {
    double dt, fn, inv_zrot, theta, tref;
    double mask[3];
    uint32_t seed_lo, seed_hi;
    uint32_t step;
    int32_t sample_interval;
    int32_t coll_model, dsmc_model, bgk_model;
    int32_t ns, sp_base, pair_base, bm_base;
    int32_t cell_begin, cell_end;
    int32_t inflow_begin, inflow_end;
    uint32_t particle_begin;     // first slot of the case in the sorted order (set each step)
    int32_t active;
    int32_t decomposition_interval, smoothing_passes;
    double breakdown_max, kn_theta;
    double kn_time;
    int32_t sampler_full;        // sample rows 8-19 as well
    int32_t sum_level;           // per-cell sums needed: 0 rows 0-7, 1 +8-13, 2 +14-19
    int32_t reset_at_decomposition;
    int32_t weighted;            // cell-weighted case (clones and deletions on weight changes)
    double reset_until_time;     // <= 0: no limit
    double time;                 // simulated time at the start of the step
    int32_t adaptive;            // adaptive simulation: the adapter's time sums are kept every step
    int32_t macro;               // macroInterpolation: BGK-family draws take cellPoint-interpolated fields
};

// Hybrid decomposition: time-sum rows per cell (N, M, M u, M v, M w, M |u|^2,
// two unused, then the per-species counts).
constexpr int kKnRows = UGX_KN_ROWS;

// Per cell-face data of the decomposition's face interpolation, FP64 for
// either particle precision: area, linear weight of the cell's own value,
// distance between the cell reference points (internal faces), face normal.
struct KnFace {
    double area, w, dist, nx, ny, nz;
};

struct BModel {
    int32_t kind;
    int32_t pad;
    double T, U[3];
};

struct InflowFS {
    int32_t cell;        // global cell
    int32_t species;     // global species
    int32_t ccase;       // case index
    int32_t local_id;    // face-species index within the case
    int32_t nv;
    int32_t pad;
    double nbar;         // expected inserts per step at cell weight 1: nbar_num / nbar_den
    double nbar_num;     // (uniGasGeneralBoundary.C:154-165) numerator and F_N part of the
    double nbar_den;     // denominator; the count is nbar_num / (nbar_den W_c)
    double vloc[8][3];
    double cum[6];
    double n[3], t1[3], t2[3];
    double T, rotT, U[3];
};

template <class R>
struct Parts {
    R *x, *y, *z, *vx, *vy, *vz, *erot;
    uint32_t *cell, *sp;
};

template <class R>
struct Geom {
    const V4<R>* plane;     // [nc*6] outward plane (n, d) in the cell frame
    const V4<R>* delta;     // [nc*6] frame shift to the neighbour
    const int32_t* nbr;     // [nc*6] neighbour cell, -1-bmodel, or kNbrEmpty
};

struct Species {
    const double* mass;
    const double* diam;
    const int32_t* rotdof;
    const double* omega;
    const double* muref;
    const double* prterm;
    const int32_t* local;   // case-local index
};

struct Pairs {
    const double *mr, *C, *e, *chi, *alpha;
};

// ------------------------------------------------------- kinetic helpers

__host__ __device__ inline double equip_erot(Rng& r, double T, int dof)
{
    if (dof <= 0) return 0.0;
    if (dof == 2) return -dlog(uniform(r)) * kB * T;
    double a = 0.5 * (double)dof - 1.0;
    double eps = uniform(r);
    double x = 0.0, P = -1.0;
    for (int it = 0; it < 100000; ++it) {
        x = 10.0 * uniform(r);
        P = dpow(x / a, a) * dexp(a - x);
        if (!(P < eps)) break;
    }
    return x * kB * T;
}

__host__ __device__ inline double post_rot_ratio(Rng& r, int dof, double chiB)
{
    if (dof == 2) return 1.0 - dpow(uniform(r), 1.0 / chiB);
    double chiA = 0.5 * (double)dof;
    double am1 = chiA - 1.0, bm1 = chiB - 1.0;
    if (am1 < kSmall && bm1 < kSmall) return uniform(r);
    double eps = uniform(r);
    double ratio = 0.0;
    for (int it = 0; it < 100000; ++it) {
        ratio = uniform(r);
        double P;
        if (am1 < kSmall) P = dpow(1.0 - ratio, bm1);
        else if (bm1 < kSmall) P = dpow(1.0 - ratio, am1);
        else P = dpow((am1 + bm1) * ratio / am1, am1) * dpow((am1 + bm1) * (1.0 - ratio) / bm1, bm1);
        if (!(P < eps)) break;
    }
    return ratio;
}

// uniGasPatchBoundary::diffuseReflection (uniGasPatchBoundary.C:305-388)
__host__ __device__ inline void diffuse_reflect(Rng& r, const double nw[3], double U[3], double& erot,
                                                double T, const double Uw[3], double mass, int rotdof)
{
    double udn = U[0] * nw[0] + U[1] * nw[1] + U[2] * nw[2];
    double Ut[3];
    for (int d = 0; d < 3; ++d) Ut[d] = U[d] - udn * nw[d];
    double mUt = sqrt(Ut[0] * Ut[0] + Ut[1] * Ut[1] + Ut[2] * Ut[2]);
    for (int it = 0; it < 1000 && mUt < kSmall; ++it) {
        for (int d = 0; d < 3; ++d) U[d] = U[d] * (0.8 + 0.2 * uniform(r));
        udn = U[0] * nw[0] + U[1] * nw[1] + U[2] * nw[2];
        for (int d = 0; d < 3; ++d) Ut[d] = U[d] - udn * nw[d];
        mUt = sqrt(Ut[0] * Ut[0] + Ut[1] * Ut[1] + Ut[2] * Ut[2]);
    }
    double t1[3], t2[3];
    for (int d = 0; d < 3; ++d) t1[d] = Ut[d] / mUt;
    t2[0] = nw[1] * t1[2] - nw[2] * t1[1];
    t2[1] = nw[2] * t1[0] - nw[0] * t1[2];
    t2[2] = nw[0] * t1[1] - nw[1] * t1[0];
    double sig = sqrt(kB * T / mass);
    double g1, g2;
    gauss2(r, g1, g2);
    double un = sqrt(-2.0 * dlog(uniform(r)));
    for (int d = 0; d < 3; ++d) U[d] = sig * ((g1 * t1[d] + g2 * t2[d]) - un * nw[d]) + Uw[d];
    erot = equip_erot(r, T, rotdof);
}

struct TrackOut {
    int alive;
    int hits;
    int errs;
    int queued;      // stopped at a diffuse wall (deferred): face and remaining time below
    int qface;
    int qiter;
};

// Face tracking through convex planar-faced cells (uniGasParcel::move,
// uniGasParcel.C:35-108) with the wall dispatch of hitWallPatch (:131-145).
// Positions and planes in R; wall physics in FP64.
// Species data of a simulator, looked up only when a wall model needs it:
// MassOf carries known values (new simulators), a world view loads
// species id and E_rot of slot t lazily (existing simulators).
struct MassOf {
    double mass;
    int rotdof;
    uint32_t sp;      // global species id
};

__device__ inline void wall_species(const MassOf& m, uint32_t, double& erot, double& mass, int& rotdof, uint32_t& sp,
                                    bool&)
{
    mass = m.mass;
    rotdof = m.rotdof;
    sp = m.sp;
    (void)erot;
}

template <class WV>
__device__ inline void wall_species(const WV& w, uint32_t t, double& erot, double& mass, int& rotdof, uint32_t& sp,
                                    bool& loaded)
{
    if (!loaded) {
        const uint32_t g = w.a.sp[t];
        mass = w.sp.mass[g];
        rotdof = w.sp.rotdof[g];
        erot = (double)w.a.erot[t];
        sp = g;
        loaded = true;
    }
}

// A wall-hit recorder for track_t: measured(c, f) tells whether face f of
// cell c is a measured wall face, hit(...) records one interaction (cell,
// face, slot, crossing number, velocity and E_rot before and after, global
// species).  NoSink records nothing.
struct NoSink {
    __device__ bool measured(uint32_t, int) const { return false; }
    __device__ void hit(uint32_t, int, uint32_t, int, const double*, const double*, double, double, uint32_t) const {}
};

// Defer = true: stop at the first diffuse wall (position on the wall, the
// remaining time in tau_io) and report it as queued; the caller finishes
// the track with the wall model later.  Nothing random is drawn before a
// diffuse wall, so the continuation uses a freshly initialised stream.
template <bool Defer, class R, class Src, class V = double, class Sink = NoSink>
__device__ inline TrackOut track_t(const Geom<R>& g, const BModel* bm, const double mask[3],
                                   uint32_t& cell, R x[3], V U[3], double& erot,
                                   const Src& src, uint32_t slot, R& tau, Rng& rng, int it0,
                                   const Sink& sink = Sink{})
{
    double mass = 0.0;
    int rotdof = 0;
    uint32_t spg = 0u;
    bool loaded = false;
    TrackOut out{1, 0, 0, 0, 0, 0};
    uint32_t c = cell;
    R vt[3];
    for (int d = 0; d < 3; ++d) vt[d] = (R)(U[d] * (V)mask[d]);
    for (int it = it0; it < kMaxCrossings; ++it) {
        const size_t base = (size_t)c * kMaxFaces;
        // every slot of a cell exists (padded ones are (0,0,0,1) / kNbrEmpty), so
        // all face data is loaded before any decision: one memory round trip per cell
        V4<R> q[kMaxFaces];
        int32_t nb[kMaxFaces];
#pragma unroll
        for (int f = 0; f < kMaxFaces; ++f) { q[f] = g.plane[base + f]; nb[f] = g.nbr[base + f]; }
        R tmin = tau;
        int fmin = -1;
        int32_t code = 0;
        R qx = (R)0, qy = (R)0, qz = (R)0;
        // every face is evaluated and the nearest one selected without
        // branches: a face counts when it exists and the simulator moves
        // towards it; ties keep the first face
#pragma unroll
        for (int f = 0; f < kMaxFaces; ++f) {
            const R dn = q[f].x * vt[0] + q[f].y * vt[1] + q[f].z * vt[2];
            R t = (q[f].w - (q[f].x * x[0] + q[f].y * x[1] + q[f].z * x[2])) / dn;
            t = (t < (R)0) ? (R)0 : t;
            const bool take = (nb[f] != kNbrEmpty) && (dn > (R)0) && (t < tmin);
            tmin = take ? t : tmin;
            fmin = take ? f : fmin;
            code = take ? nb[f] : code;
            qx = take ? q[f].x : qx;
            qy = take ? q[f].y : qy;
            qz = take ? q[f].z : qz;
        }
        if (fmin < 0) {
            for (int d = 0; d < 3; ++d) x[d] = x[d] + vt[d] * tau;
            cell = c;
            return out;
        }
        for (int d = 0; d < 3; ++d) x[d] = x[d] + vt[d] * tmin;
        tau = tau - tmin;
        if (code >= 0) {
            V4<R> dl = g.delta[base + fmin];
            x[0] = x[0] + dl.x;
            x[1] = x[1] + dl.y;
            x[2] = x[2] + dl.z;
            c = (uint32_t)code;
            continue;
        }
        const BModel b = bm[-1 - code];
        double nw[3] = {(double)qx, (double)qy, (double)qz};
        if (b.kind == UGX_BND_DELETION) { cell = kDead; out.alive = 0; return out; }
        if (Defer && b.kind != UGX_BND_SPECULAR) {
            out.queued = 1;
            out.qface = fmin;
            out.qiter = it;
            cell = c;
            return out;
        }
        out.hits += 1;
        const bool meas = sink.measured(c, fmin);
        const double pre[3] = {(double)U[0], (double)U[1], (double)U[2]};
        if (b.kind == UGX_BND_SPECULAR) {
            // mirror; E_rot and species are needed only to measure the hit
            const V n0 = (V)qx, n1 = (V)qy, n2 = (V)qz;
            V udn = U[0] * n0 + U[1] * n1 + U[2] * n2;
            if (udn > (V)0) {
                U[0] = U[0] - (V)2 * udn * n0;
                U[1] = U[1] - (V)2 * udn * n1;
                U[2] = U[2] - (V)2 * udn * n2;
            }
            if (meas) {
                wall_species(src, slot, erot, mass, rotdof, spg, loaded);
                const double post[3] = {(double)U[0], (double)U[1], (double)U[2]};
                sink.hit(c, fmin, slot, it, pre, post, erot, erot, spg);
            }
        } else if constexpr (!Defer) {
            wall_species(src, slot, erot, mass, rotdof, spg, loaded);
            const double e0 = erot;
            double Ud[3] = {(double)U[0], (double)U[1], (double)U[2]};
            diffuse_reflect(rng, nw, Ud, erot, b.T, b.U, mass, rotdof);
            U[0] = (V)Ud[0]; U[1] = (V)Ud[1]; U[2] = (V)Ud[2];
            if (meas) {
                const double post[3] = {(double)U[0], (double)U[1], (double)U[2]};
                sink.hit(c, fmin, slot, it, pre, post, e0, erot, spg);
            }
        }
        for (int d = 0; d < 3; ++d) vt[d] = (R)(U[d] * (V)mask[d]);
    }
    out.errs = 1;
    cell = c;
    return out;
}

template <class R, class Src, class Sink = NoSink>
__device__ inline TrackOut track(const Geom<R>& g, const BModel* bm, const double mask[3],
                                 uint32_t& cell, R x[3], double U[3], double& erot,
                                 const Src& src, uint32_t slot, double tau_d, Rng& rng, const Sink& sink = Sink{})
{
    R tau = (R)tau_d;
    return track_t<false, R, Src, double, Sink>(g, bm, mask, cell, x, U, erot, src, slot, tau, rng, 0, sink);
}

// Diffuse reflection at face f of cell c, then the rest of the track:
// the continuation of a deferred track_t<true>.
template <class R, class Src, class Sink = NoSink>
__device__ inline TrackOut finish_diffuse(const Geom<R>& g, const BModel* bm, const double mask[3],
                                          uint32_t& cell, int f, int it, R x[3], double U[3], double& erot,
                                          const Src& src, uint32_t slot, R tau, Rng& rng, const Sink& sink = Sink{})
{
    double mass = 0.0;
    int rotdof = 0;
    uint32_t spg = 0u;
    bool loaded = false;
    wall_species(src, slot, erot, mass, rotdof, spg, loaded);
    const size_t k = (size_t)cell * kMaxFaces + f;
    const BModel b = bm[-1 - g.nbr[k]];
    V4<R> q = g.plane[k];
    double nw[3] = {(double)q.x, (double)q.y, (double)q.z};
    const double pre[3] = {U[0], U[1], U[2]};
    const double e0 = erot;
    diffuse_reflect(rng, nw, U, erot, b.T, b.U, mass, rotdof);
    if (sink.measured(cell, f)) sink.hit(cell, f, slot, it, pre, U, e0, erot, spg);
    // the rest of the track carries this E_rot (not the stored one)
    TrackOut o = track_t<false, R, MassOf, double, Sink>(g, bm, mask, cell, x, U, erot, MassOf{mass, rotdof, spg}, slot,
                                                         tau, rng, it + 1, sink);
    o.hits += 1;
    return o;
}

// ------------------------------------------------------------ collisions

__device__ inline void scatter(Rng& r, int vss, double alpha, double cR, const double crc[3],
                               double mP, double mQ, double UP[3], double UQ[3])
{
    double Ucm[3];
    for (int d = 0; d < 3; ++d) Ucm[d] = (mP * UP[d] + mQ * UQ[d]) / (mP + mQ);
    double cosT = vss ? 2.0 * dpow(uniform(r), 1.0 / alpha) - 1.0 : 2.0 * uniform(r) - 1.0;
    double sinT = sqrt(1.0 - cosT * cosT);
    double sp, cp;
    dsincos2pi(uniform(r), sp, cp);
    double c[3];
    double D = sqrt(crc[1] * crc[1] + crc[2] * crc[2]);
    if (vss && D > 0.0) {
        c[0] = cosT * crc[0] + sinT * sp * D;
        c[1] = cosT * crc[1] + sinT * (cR * crc[2] * cp - crc[0] * crc[1] * sp) / D;
        c[2] = cosT * crc[2] - sinT * (cR * crc[1] * cp + crc[0] * crc[2] * sp) / D;
    } else {
        c[0] = cR * cosT;
        c[1] = cR * (sinT * cp);
        c[2] = cR * (sinT * sp);
    }
    for (int d = 0; d < 3; ++d) {
        UP[d] = Ucm[d] + c[d] * mQ / (mP + mQ);
        UQ[d] = Ucm[d] - c[d] * mP / (mP + mQ);
    }
}

// dsmcCollisionModel::collide: VHS, VSS, serial Larsen-Borgnakke.
__device__ inline void binary_collide(Rng& r, int model, double inv_zrot, int rdP, int rdQ,
                                      double mP, double mQ, double mR, double chiB, double alpha,
                                      double UP[3], double UQ[3], double& eP, double& eQ)
{
    double crc[3] = {UP[0] - UQ[0], UP[1] - UQ[1], UP[2] - UQ[2]};
    double cR2 = crc[0] * crc[0] + crc[1] * crc[1] + crc[2] * crc[2];
    double cR = sqrt(cR2);
    int vss = (model == UGX_DSMC_VSS || model == UGX_DSMC_LB_VSS);
    if (model == UGX_DSMC_LB_VHS || model == UGX_DSMC_LB_VSS) {
        double Et = 0.5 * mR * cR2;
        double preP = eP, preQ = eQ;
        if (rdP > 0 && inv_zrot > uniform(r)) {
            double Ec = Et + preP;
            double ratio = post_rot_ratio(r, rdP, chiB);
            eP = ratio * Ec;
            Et = Ec - eP;
        }
        if (rdQ > 0 && inv_zrot > uniform(r)) {
            double Ec = Et + preQ;
            double ratio = post_rot_ratio(r, rdQ, chiB);
            eQ = ratio * Ec;
            Et = Ec - eQ;
        }
        double cRn = sqrt(2.0 * Et / mR);
        if (vss && cR > 0.0)
            for (int d = 0; d < 3; ++d) crc[d] = crc[d] * (cRn / cR);
        cR = cRn;
    }
    scatter(r, vss, alpha, cR, crc, mP, mQ, UP, UQ);
}

}  // namespace ugx
