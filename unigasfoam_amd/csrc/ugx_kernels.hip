// ugx_kernels.hip - the particle time step of uniGasCloud::evolve
// (src/lagrangian/uniGas/clouds/uniGasCloud.C:821-869) as gfx950 kernels.
//
// Per step, on one HIP stream, for every case of a world at once:
//   begin      per-case slot base, insert counts (Bird 4.22, stochastic round)
//   inflow     create + first move of new simulators     (uniGasGeneralBoundary.C:537-761)
//   move       face tracking + wall models + histogram   (uniGasParcel.C:35-145)
//   scan       cell offsets                              (CloudWithModels.C:109-138)
//   scatter    counting-sort slots (wave-aggregated atomics)
//   segsort    per-tile LDS bitonic sort of each cell's slot list by previous
//              slot -> the stable order, independent of atomic timing
//   gather     SoA records into the sorted buffers
//   sample     per-cell sums -> Delta-t weighted FP64 sampler (cellMeasurements.C:408-513,
//              uniGasVolFields.C:723-837)
//   bgk        BGK / ES-BGK / USP relaxation per cell    (bgkCollisions/derived/*)
//   ntc        No-Time-Counter + VHS/VSS/LB per cell     (noTimeCounter.C:66-343)
//   end        step counters
#include <hip/hip_runtime.h>

#include "ugx_kernels.hpp"

namespace ugx {

constexpr int kBlock = 256;
constexpr int kScanBlock = 1024;      // elements per scan block (256 threads x 4)

__device__ inline uint32_t wave_agg_inc(uint32_t* ctr, uint32_t key, bool active)
{
    const unsigned lane = __lane_id();
    uint32_t ret = 0;
    unsigned long long pending = __ballot(active);
    while (pending) {
        int leader = __ffsll(pending) - 1;
        uint32_t lkey = __shfl(key, leader);
        unsigned long long grp = __ballot(active && key == lkey) & pending;
        uint32_t base = 0;
        if ((int)lane == leader) base = atomicAdd(&ctr[lkey], (uint32_t)__popcll(grp));
        base = __shfl(base, leader);
        if ((grp >> lane) & 1ull) ret = base + (uint32_t)__popcll(grp & ((1ull << lane) - 1ull));
        pending &= ~grp;
    }
    return ret;
}

// Adds v to counters[ccase][idx]; one atomic per wave when the wave's
// non-zero contributions share a case.  Called by every lane of the wave.
__device__ inline void case_add(unsigned long long* counters, int ccase, int idx, unsigned long long v)
{
    unsigned long long nz = __ballot(v != 0ull);
    if (nz == 0ull) return;
    int leader = __ffsll(nz) - 1;
    int lc = __shfl(ccase, leader);
    unsigned long long other = __ballot(v != 0ull && ccase != lc);
    if (other == 0ull) {
        unsigned long long s = v;
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if ((int)__lane_id() == leader) atomicAdd(&counters[(size_t)lc * kCntNum + idx], s);
    } else if (v != 0ull) {
        atomicAdd(&counters[(size_t)ccase * kCntNum + idx], v);
    }
}


// Per-thread counters of one case, reduced over the block at kernel end so
// that a kernel issues one global atomic per counter per block instead of
// one per wave.  Slots: 0 deleted, 1 wall hits, 2 track errors, 3 inserted.
struct TCount {
    int ccase;
    unsigned long long v[4];
};

__device__ inline void tcount_flush_global(const TCount& t, unsigned long long* counters)
{
    const int idx[4] = {kCntDeleted, kCntWallHits, kCntTrackErrors, kCntInserted};
    if (t.ccase < 0) return;
    for (int k = 0; k < 4; ++k)
        if (t.v[k]) atomicAdd(&counters[(size_t)t.ccase * kCntNum + idx[k]], t.v[k]);
}

__device__ inline void tcount_add(TCount& t, unsigned long long* counters, int ccase, int k, unsigned long long v)
{
    if (!v) return;
    if (t.ccase != ccase) {
        tcount_flush_global(t, counters);
        t.ccase = ccase;
        t.v[0] = t.v[1] = t.v[2] = t.v[3] = 0ull;
    }
    t.v[k] += v;
}

// Called by every thread of the block.
__device__ inline void tcount_block_flush(const TCount& t, unsigned long long* counters)
{
    __shared__ int s_case[kBlock];
    __shared__ unsigned long long s_v[kBlock][4];
    const uint32_t tid = threadIdx.x;
    s_case[tid] = t.ccase;
    for (int k = 0; k < 4; ++k) s_v[tid][k] = t.v[k];
    __syncthreads();
    if (tid == 0) {
        TCount sum{-1, {0ull, 0ull, 0ull, 0ull}};
        for (uint32_t i = 0; i < kBlock; ++i) {
            if (s_case[i] < 0) continue;
            if (sum.ccase < 0) sum.ccase = s_case[i];
            if (s_case[i] == sum.ccase) {
                for (int k = 0; k < 4; ++k) sum.v[k] += s_v[i][k];
            } else {
                TCount o{s_case[i], {s_v[i][0], s_v[i][1], s_v[i][2], s_v[i][3]}};
                tcount_flush_global(o, counters);
            }
        }
        tcount_flush_global(sum, counters);
    }
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

template <class R>
__global__ void k_begin_step(WorldView<R> w)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < w.ncases) w.cases[c].particle_begin = w.off[w.cases[c].cell_begin];
    if (c == 0) { w.st->n_ins = 0u; w.st->nwall = 0u; }
}

// ------------------------------------------------------------- inflow

// One block: insert counts per face-species (uniGasGeneralBoundary.C:602-609)
// and their exclusive scan.
template <class R>
__global__ void k_inflow_count(WorldView<R> w)
{
    __shared__ uint32_t sh[1024];
    const int tid = threadIdx.x;
    uint32_t carry = 0;
    for (int base = 0; base < w.n_inflow; base += 1024) {
        int k = base + tid;
        uint32_t cnt = 0;
        if (k < w.n_inflow) {
            const InflowFS& e = w.inflow[k];
            const CaseDev& cs = w.cases[e.ccase];
            if (cs.active) {
                Rng r;
                rng_init(r, cs.seed_lo, cs.seed_hi, kRngInflowN, 0u, (uint32_t)e.local_id, cs.step);
                double nb = e.nbar;
                double fl = floor(nb);
                long long c = (long long)fl;
                if (c < 0) c = 0;
                if (uniform(r) < nb - fl) c += 1;
                cnt = (uint32_t)c;
            }
        }
        sh[tid] = cnt;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            uint32_t v = (tid >= o) ? sh[tid - o] : 0u;
            __syncthreads();
            sh[tid] += v;
            __syncthreads();
        }
        if (k < w.n_inflow) w.ins_off[k] = carry + sh[tid] - cnt;
        uint32_t tot = sh[1023];
        __syncthreads();
        carry += tot;
    }
    if (tid == 0) {
        w.ins_off[w.n_inflow] = carry;
        if ((unsigned long long)w.st->n + carry > (unsigned long long)w.cap) {
            w.st->overflow = 1u;
            carry = 0u;
        }
        w.st->n_ins = carry;
    }
}

// Bird eq. 12.5 insertion + first move over (1 - R) deltaT.
template <class R>
__device__ inline int make_inflow(const WorldView<R>& w, const InflowFS& e, const CaseDev& cs, Rng& r,
                                  R x[3], double U[3], double& erot, uint32_t& cell, TrackOut& to)
{
    double us = uniform(r);
    int tri = 0;
    for (int t = 0; t < e.nv - 2; ++t) { tri = t; if (e.cum[t] >= us) break; }
    double s = sqrt(uniform(r));
    double tt = uniform(r);
    for (int d = 0; d < 3; ++d)
        x[d] = (R)(e.vloc[0][d] * (1.0 - s) + e.vloc[tri + 1][d] * (s * (1.0 - tt)) + e.vloc[tri + 2][d] * (s * tt));
    const int g = e.species;
    double m = w.sp.mass[g];
    double mps = sqrt(2.0 * kB * e.T / m);
    double udn = e.U[0] * e.n[0] + e.U[1] * e.n[1] + e.U[2] * e.n[2];
    double sCos = udn / mps;
    double uNormal;
    if (fabs(udn) > kVSmall) {
        double cA = sCos + sqrt(sCos * sCos + 2.0);
        double cB = 0.5 * (1.0 + sCos * (sCos - sqrt(sCos * sCos + 2.0)));
        double scale = 3.0;
        if (sCos < -3.0) scale = fabs(sCos) + 1.0;
        uNormal = 0.0;
        for (int it = 0; it < 100000; ++it) {
            double th = scale * (2.0 * uniform(r) - 1.0);
            uNormal = th + sCos;
            double P = (uNormal < 0.0) ? -1.0 : 2.0 * uNormal / cA * dexp(cB - th * th);
            if (!(P < uniform(r))) break;
        }
    } else {
        uNormal = sqrt(-dlog(uniform(r)));
    }
    double g1, g2;
    gauss2(r, g1, g2);
    double sig = sqrt(kB * e.T / m);
    double ut1 = e.t1[0] * e.U[0] + e.t1[1] * e.U[1] + e.t1[2] * e.U[2];
    double ut2 = e.t2[0] * e.U[0] + e.t2[1] * e.U[1] + e.t2[2] * e.U[2];
    for (int d = 0; d < 3; ++d)
        U[d] = ((sig * (g1 * e.t1[d] + g2 * e.t2[d]) + ut1 * e.t1[d]) + ut2 * e.t2[d]) + mps * uNormal * e.n[d];
    const int rd = w.sp.rotdof[g];
    erot = equip_erot(r, e.rotT, rd);
    cell = (uint32_t)e.cell;
    double frac = uniform(r);
    to = track<R>(w.geom, w.bm, cs.mask, cell, x, U, erot, MassOf{m, rd}, 0u, (1.0 - frac) * cs.dt, r);
    return to.alive;
}

template <class R>
__global__ void k_inflow_gen(WorldView<R> w)
{
    TCount tc{-1, {0ull, 0ull, 0ull, 0ull}};
    const uint32_t total = w.st->n_ins;
    const uint32_t n0 = w.st->n;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < w.max_ins; base += stride) {
        uint32_t t = base + threadIdx.x;
        bool active = t < total;
        uint32_t cell = kDead;
        int ccase = 0;
        TrackOut to{1, 0, 0};
        if (active) {
            int lo = 0, hi = w.n_inflow - 1;
            while (lo < hi) {
                int mid = (lo + hi + 1) >> 1;
                if (w.ins_off[mid] <= t) lo = mid; else hi = mid - 1;
            }
            const InflowFS& e = w.inflow[lo];
            const CaseDev& cs = w.cases[e.ccase];
            ccase = e.ccase;
            uint32_t i = t - w.ins_off[lo];
            Rng r;
            rng_init(r, cs.seed_lo, cs.seed_hi, kRngInflowP, i, (uint32_t)e.local_id, cs.step);
            R x[3];
            double U[3], erot;
            make_inflow<R>(w, e, cs, r, x, U, erot, cell, to);
            uint32_t j = n0 + t;
            w.a.x[j] = x[0]; w.a.y[j] = x[1]; w.a.z[j] = x[2];
            w.a.vx[j] = (R)U[0]; w.a.vy[j] = (R)U[1]; w.a.vz[j] = (R)U[2];
            w.a.erot[j] = (R)erot;
            w.a.cell[j] = cell;
            w.a.sp[j] = (uint32_t)e.species;
        }
        bool alive = active && cell != kDead;
        wave_agg_inc(w.cnt, alive ? cell : 0u, alive);
        if (active) {
            tcount_add(tc, w.counters, ccase, 3, 1ull);
            tcount_add(tc, w.counters, ccase, 0, alive ? 0ull : 1ull);
            tcount_add(tc, w.counters, ccase, 1, (unsigned long long)to.hits);
            tcount_add(tc, w.counters, ccase, 2, (unsigned long long)to.errs);
        }
    }
    tcount_block_flush(tc, w.counters);
}

// --------------------------------------------------------------- move

// Case of an existing simulator from its slot: the cases own consecutive slot
// ranges starting at particle_begin (set by k_begin_step from the offsets).
template <class R>
__device__ inline int case_of_slot(const WorldView<R>& w, uint32_t t)
{
    int lo = 0, hi = w.ncases - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (w.cases[mid].particle_begin <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Ballistic move of the existing simulators.  Specular walls, cyclic and
// deletion faces are handled in line; a simulator reaching a diffuse wall
// stops there and is queued for k_wall, which keeps this kernel free of the
// FP64 transcendental code of the wall model.
template <class R>
__global__ void __launch_bounds__(256) k_move(WorldView<R> w)
{
    TCount tc{-1, {0ull, 0ull, 0ull, 0ull}};
    const uint32_t n0 = w.st->n;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < w.cap; base += stride) {
        if (base >= n0) break;
        uint32_t t = base + threadIdx.x;
        bool active = t < n0;
        uint32_t cell = kDead;
        int ccase = 0;
        TrackOut to{1, 0, 0, 0, 0, 0};
        if (active) {
            cell = w.a.cell[t];
            ccase = case_of_slot(w, t);
            const CaseDev& cs = w.cases[ccase];
            R x[3] = {w.a.x[t], w.a.y[t], w.a.z[t]};
            R U[3] = {w.a.vx[t], w.a.vy[t], w.a.vz[t]};
            double erot = 0.0;
            Rng r;
            rng_init(r, cs.seed_lo, cs.seed_hi, kRngMove, 0u, t - cs.particle_begin, cs.step);
            R tau = (R)cs.dt;
            to = track_t<true, R, WorldView<R>, R>(w.geom, w.bm, cs.mask, cell, x, U, erot, w, t, tau, r, 0);
            w.a.x[t] = x[0]; w.a.y[t] = x[1]; w.a.z[t] = x[2];
            w.a.cell[t] = cell;
            if (to.hits) {
                w.a.vx[t] = U[0]; w.a.vy[t] = U[1]; w.a.vz[t] = U[2];
            }
            if (to.queued) {
                const uint32_t q = atomicAdd(&w.st->nwall, 1u);
                if (q < w.wq_cap) {
                    w.wq_slot[q] = t;
                    w.wq_info[q] = (uint32_t)to.qface | ((uint32_t)to.qiter << 3);
                    w.wq_tau[q] = (double)tau;
                } else {
                    w.st->overflow = 1u;
                }
            }
        }
        bool done = active && cell != kDead && !to.queued;
        wave_agg_inc(w.cnt, done ? cell : 0u, done);
        if (active) {
            tcount_add(tc, w.counters, ccase, 0, (cell == kDead) ? 1ull : 0ull);
            tcount_add(tc, w.counters, ccase, 1, (unsigned long long)to.hits);
            tcount_add(tc, w.counters, ccase, 2, (unsigned long long)to.errs);
        }
    }
    tcount_block_flush(tc, w.counters);
}

// Diffuse-wall continuation of the queued tracks.
template <class R>
__global__ void k_wall(WorldView<R> w)
{
    TCount tc{-1, {0ull, 0ull, 0ull, 0ull}};
    const uint32_t nq = min(w.st->nwall, w.wq_cap);
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < nq; base += stride) {
        uint32_t q = base + threadIdx.x;
        bool active = q < nq;
        uint32_t cell = kDead;
        int ccase = 0;
        TrackOut to{1, 0, 0, 0, 0, 0};
        if (active) {
            const uint32_t t = w.wq_slot[q];
            const uint32_t info = w.wq_info[q];
            cell = w.a.cell[t];
            ccase = w.cell_case[cell];
            const CaseDev& cs = w.cases[ccase];
            R x[3] = {w.a.x[t], w.a.y[t], w.a.z[t]};
            double U[3] = {(double)w.a.vx[t], (double)w.a.vy[t], (double)w.a.vz[t]};
            double erot = 0.0;
            Rng r;
            rng_init(r, cs.seed_lo, cs.seed_hi, kRngMove, 0u, t - cs.particle_begin, cs.step);
            to = finish_diffuse<R>(w.geom, w.bm, cs.mask, cell, (int)(info & 7u), (int)(info >> 3), x, U, erot,
                                   w, t, (R)w.wq_tau[q], r);
            w.a.x[t] = x[0]; w.a.y[t] = x[1]; w.a.z[t] = x[2];
            w.a.cell[t] = cell;
            w.a.vx[t] = (R)U[0]; w.a.vy[t] = (R)U[1]; w.a.vz[t] = (R)U[2];
            w.a.erot[t] = (R)erot;
        }
        bool done = active && cell != kDead;
        wave_agg_inc(w.cnt, done ? cell : 0u, done);
        if (active) {
            tcount_add(tc, w.counters, ccase, 0, (cell == kDead) ? 1ull : 0ull);
            tcount_add(tc, w.counters, ccase, 1, (unsigned long long)to.hits);
            tcount_add(tc, w.counters, ccase, 2, (unsigned long long)to.errs);
        }
    }
    tcount_block_flush(tc, w.counters);
}

// Histogram only (sort after fill / upload).
template <class R>
__global__ void k_hist(WorldView<R> w)
{
    const uint32_t n_all = w.st->n + w.st->n_ins;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < w.cap; base += stride) {
        if (base >= n_all) break;
        uint32_t t = base + threadIdx.x;
        bool active = t < n_all;
        uint32_t c = active ? w.a.cell[t] : kDead;
        bool alive = active && c != kDead;
        wave_agg_inc(w.cnt, alive ? c : 0u, alive);
    }
}

// --------------------------------------------------------------- scan

__global__ void k_scan_reduce(const uint32_t* in, uint32_t n, uint32_t* bsum)
{
    __shared__ uint32_t sh[kBlock];
    uint32_t b = blockIdx.x * kScanBlock;
    uint32_t s = 0;
    for (int k = 0; k < 4; ++k) {
        uint32_t i = b + threadIdx.x * 4 + k;
        if (i < n) s += in[i];
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = sh[0];
}

__global__ void k_scan_bsum(uint32_t* bsum, uint32_t nb)
{
    __shared__ uint32_t sh[1024];
    const int tid = threadIdx.x;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nb; base += 1024) {
        uint32_t k = base + tid;
        uint32_t v0 = (k < nb) ? bsum[k] : 0u;
        sh[tid] = v0;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            uint32_t v = (tid >= o) ? sh[tid - o] : 0u;
            __syncthreads();
            sh[tid] += v;
            __syncthreads();
        }
        if (k < nb) bsum[k] = carry + sh[tid] - v0;
        uint32_t tot = sh[1023];
        __syncthreads();
        carry += tot;
    }
}

// Exclusive scan of in[0..n) into out; out[n-1] is the total when in[n-1]==0.
__global__ void k_scan_apply(const uint32_t* in, uint32_t n, const uint32_t* bsum, uint32_t* out, uint32_t* total_out)
{
    __shared__ uint32_t sh[kBlock];
    uint32_t b = blockIdx.x * kScanBlock + threadIdx.x * 4;
    uint32_t v[4], s = 0;
    for (int k = 0; k < 4; ++k) { v[k] = (b + k < n) ? in[b + k] : 0u; s += v[k]; }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 1; o < kBlock; o <<= 1) {
        uint32_t t = ((int)threadIdx.x >= o) ? sh[threadIdx.x - o] : 0u;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    uint32_t run = bsum[blockIdx.x] + sh[threadIdx.x] - s;
    for (int k = 0; k < 4; ++k) {
        if (b + k < n) {
            out[b + k] = run;
            if (b + k == n - 1 && total_out) *total_out = run;
        }
        run += v[k];
    }
}

static void scan_u32(const uint32_t* in, uint32_t n, uint32_t* out, uint32_t* bsum, uint32_t* total, hipStream_t s)
{
    uint32_t nb = (n + kScanBlock - 1) / kScanBlock;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kBlock), 0, s, in, n, bsum);
    hipLaunchKernelGGL(k_scan_bsum, dim3(1), dim3(1024), 0, s, bsum, nb);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(kBlock), 0, s, in, n, bsum, out, total);
}

// ------------------------------------------------------------ scatter

template <class R>
__global__ void k_scatter(WorldView<R> w)
{
    const uint32_t n_all = w.st->n + w.st->n_ins;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < w.cap; base += stride) {
        if (base >= n_all) break;
        uint32_t t = base + threadIdx.x;
        bool active = t < n_all;
        uint32_t c = active ? w.a.cell[t] : kDead;
        bool alive = active && c != kDead;
        uint32_t rank = wave_agg_inc(w.cursor, alive ? c : 0u, alive);
        if (alive) w.perm[w.off[c] + rank] = t;
    }
}

__device__ inline uint32_t lower_bound_u32(const uint32_t* a, uint32_t n, uint32_t v)
{
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// ------------------------------------------------------- cell accessors
// The per-cell physics reads and writes simulators through an accessor:
// LdsAcc for a tile staged in LDS, GlobAcc for the sorted buffers in HBM
// (tiles too large for LDS).  Both address simulators by sorted slot.

template <class R>
struct LdsAcc {
    R *vx, *vy, *vz, *er;
    uint8_t* sl;
    uint16_t* idx;
    uint32_t p0;
    __device__ void load(uint32_t i, double& a, double& b, double& c) const
    {
        const uint32_t k = i - p0;
        a = (double)vx[k]; b = (double)vy[k]; c = (double)vz[k];
    }
    __device__ void store(uint32_t i, double a, double b, double c) const
    {
        const uint32_t k = i - p0;
        vx[k] = (R)a; vy[k] = (R)b; vz[k] = (R)c;
    }
    __device__ double erot(uint32_t i) const { return (double)er[i - p0]; }
    __device__ void set_erot(uint32_t i, double e) const { er[i - p0] = (R)e; }
    __device__ int local_species(uint32_t i) const { return sl[i - p0]; }
    __device__ uint32_t idx_get(uint32_t b, uint32_t k) const { return idx[b - p0 + k]; }
    __device__ void idx_set(uint32_t b, uint32_t k, uint32_t v) const { idx[b - p0 + k] = (uint16_t)v; }
};

template <class R>
struct GlobAcc {
    Parts<R> p;
    const int32_t* local;
    uint32_t* idx;
    __device__ void load(uint32_t i, double& a, double& b, double& c) const
    {
        a = (double)p.vx[i]; b = (double)p.vy[i]; c = (double)p.vz[i];
    }
    __device__ void store(uint32_t i, double a, double b, double c) const
    {
        p.vx[i] = (R)a; p.vy[i] = (R)b; p.vz[i] = (R)c;
    }
    __device__ double erot(uint32_t i) const { return (double)p.erot[i]; }
    __device__ void set_erot(uint32_t i, double e) const { p.erot[i] = (R)e; }
    __device__ int local_species(uint32_t i) const { return local[p.sp[i]]; }
    __device__ uint32_t idx_get(uint32_t b, uint32_t k) const { return idx[b + k]; }
    __device__ void idx_set(uint32_t b, uint32_t k, uint32_t v) const { idx[b + k] = v; }
};

// Canonical per-cell sums, wave-cooperative: the cell's simulators are taken
// 64 at a time (lane l holds simulator chunk + l, missing lanes hold +0.0),
// every value is canonicalised (v + 0.0) and reduced by the xor butterfly
// 32, 16, 8, 4, 2, 1; the chunk totals are added in order to 0.0.  Every
// lane of the wave ends with the totals.  Called by all 64 lanes of a wave
// with the same [b, e).
__device__ inline double wave_bfly(double v)
{
    v = v + 0.0;
    v = v + __shfl_xor(v, 32);
    v = v + __shfl_xor(v, 16);
    v = v + __shfl_xor(v, 8);
    v = v + __shfl_xor(v, 4);
    v = v + __shfl_xor(v, 2);
    v = v + __shfl_xor(v, 1);
    return v;
}

// Rows of cellMeasurements::calculateFields (cellMeasurements.C:408-513).
// L selects the rows needed: 0 rows 0-7, 1 adds 8-13, 2 adds 14-19; the
// per-species rows follow (a single-species case copies rows N and MCC,
// which are the same sums).
template <int L, class R, class Acc>
__device__ inline void wave_cell_sums_l(const WorldView<R>& w, const CaseDev& cs, const Acc& A, uint32_t b,
                                        uint32_t e, double S[kNSamp])
{
#pragma unroll
    for (int k = 0; k < kNSamp; ++k) S[k] = 0.0;
    const uint32_t lane = __lane_id();
    const bool multi = cs.ns > 1;
    for (uint32_t cb = b; cb < e; cb += 64u) {
        const uint32_t i = cb + lane;
        const bool on = i < e;
        double m = 0.0, vx = 0.0, vy = 0.0, vz = 0.0, er = 0.0, rd = 0.0, one = 0.0;
        int s = -1;
        if (on) {
            s = A.local_species(i);
            const int g = cs.sp_base + s;
            m = w.sp.mass[g];
            rd = (double)w.sp.rotdof[g];
            A.load(i, vx, vy, vz);
            er = A.erot(i);
            one = 1.0;
        }
        const double mcc = m * (vx * vx + vy * vy + vz * vz);
        S[UGX_S_N] = S[UGX_S_N] + wave_bfly(one);
        S[UGX_S_M] = S[UGX_S_M] + wave_bfly(m);
        S[UGX_S_MU] = S[UGX_S_MU] + wave_bfly(m * vx);
        S[UGX_S_MV] = S[UGX_S_MV] + wave_bfly(m * vy);
        S[UGX_S_MW] = S[UGX_S_MW] + wave_bfly(m * vz);
        S[UGX_S_MCC] = S[UGX_S_MCC] + wave_bfly(mcc);
        S[UGX_S_EROT] = S[UGX_S_EROT] + wave_bfly(er);
        S[UGX_S_ROTDOF] = S[UGX_S_ROTDOF] + wave_bfly(rd);
        if (L >= 1) {
            S[UGX_S_MUU] = S[UGX_S_MUU] + wave_bfly(m * (vx * vx));
            S[UGX_S_MUV] = S[UGX_S_MUV] + wave_bfly(m * (vx * vy));
            S[UGX_S_MUW] = S[UGX_S_MUW] + wave_bfly(m * (vx * vz));
            S[UGX_S_MVV] = S[UGX_S_MVV] + wave_bfly(m * (vy * vy));
            S[UGX_S_MVW] = S[UGX_S_MVW] + wave_bfly(m * (vy * vz));
            S[UGX_S_MWW] = S[UGX_S_MWW] + wave_bfly(m * (vz * vz));
        }
        if (L >= 2) {
            S[UGX_S_MCCU] = S[UGX_S_MCCU] + wave_bfly(mcc * vx);
            S[UGX_S_MCCV] = S[UGX_S_MCCV] + wave_bfly(mcc * vy);
            S[UGX_S_MCCW] = S[UGX_S_MCCW] + wave_bfly(mcc * vz);
            S[UGX_S_EROTU] = S[UGX_S_EROTU] + wave_bfly(er * vx);
            S[UGX_S_EROTV] = S[UGX_S_EROTV] + wave_bfly(er * vy);
            S[UGX_S_EROTW] = S[UGX_S_EROTW] + wave_bfly(er * vz);
        }
        if (multi) {
#pragma unroll
            for (int q = 0; q < kNS; ++q) {
                if (q < cs.ns) {
                    S[UGX_S_SPECIES_N + q] = S[UGX_S_SPECIES_N + q] + wave_bfly((s == q) ? 1.0 : 0.0);
                    S[UGX_S_SPECIES_MCC + q] = S[UGX_S_SPECIES_MCC + q] + wave_bfly((s == q) ? mcc : 0.0);
                }
            }
        }
    }
    if (!multi) {
        S[UGX_S_SPECIES_N] = S[UGX_S_N];
        S[UGX_S_SPECIES_MCC] = S[UGX_S_MCC];
    }
}

template <class R, class Acc>
__device__ inline void wave_cell_sums(const WorldView<R>& w, const CaseDev& cs, const Acc& A, uint32_t b, uint32_t e,
                                      double S[kNSamp])
{
    if (cs.sum_level >= 2) wave_cell_sums_l<2, R, Acc>(w, cs, A, b, e, S);
    else if (cs.sum_level == 1) wave_cell_sums_l<1, R, Acc>(w, cs, A, b, e, S);
    else wave_cell_sums_l<0, R, Acc>(w, cs, A, b, e, S);
}

// noTimeCounter::collide for one cell (noTimeCounter.C:83-312), subcell level 1.
template <class R, class Acc>
__device__ inline void ntc_cell(const WorldView<R>& w, const CaseDev& cs, uint32_t c, uint32_t b, uint32_t nC,
                                const Acc& A, unsigned long long& cand, unsigned long long& coll)
{
    Rng r;
    rng_init(r, cs.seed_lo, cs.seed_hi, kRngNtc, 0u, c - (uint32_t)cs.cell_begin, cs.step);
    const double smax0 = w.smax[c];
    double smax = smax0;
    const double sel = 0.5 * (double)nC * (double)(nC - 1) * cs.fn * smax0 * cs.dt / w.vol[c];
    const double fl = floor(sel);
    long long ncand = (long long)fl;
    if (uniform(r) < sel - fl) ncand += 1;
    cand += (unsigned long long)ncand;
    for (long long k = 0; k < ncand; ++k) {
        uint32_t P = uindex(r, nC);
        uint32_t Q;
        do { Q = uindex(r, nC); } while (Q == P);
        const uint32_t ip = b + P, iq = b + Q;
        const int lp = A.local_species(ip), lq = A.local_species(iq);
        const int gp = cs.sp_base + lp, gq = cs.sp_base + lq;
        const int pk = cs.pair_base + lp * cs.ns + lq;
        double UP[3], UQ[3];
        A.load(ip, UP[0], UP[1], UP[2]);
        A.load(iq, UQ[0], UQ[1], UQ[2]);
        double d0 = UP[0] - UQ[0], d1 = UP[1] - UQ[1], d2 = UP[2] - UQ[2];
        double cR2 = d0 * d0 + d1 * d1 + d2 * d2;
        double scr = (cR2 < kVSmall) ? 0.0 : w.pr.C[pk] * dexp(w.pr.e[pk] * dlog(cR2));
        if (scr > smax) smax = scr;
        if (scr / smax0 > uniform(r)) {
            double eP = A.erot(ip), eQ = A.erot(iq);
            binary_collide(r, cs.dsmc_model, cs.inv_zrot, w.sp.rotdof[gp], w.sp.rotdof[gq], w.sp.mass[gp],
                           w.sp.mass[gq], w.pr.mr[pk], w.pr.chi[pk], w.pr.alpha[pk], UP, UQ, eP, eQ);
            A.store(ip, UP[0], UP[1], UP[2]);
            A.store(iq, UQ[0], UQ[1], UQ[2]);
            A.set_erot(ip, eP);
            A.set_erot(iq, eQ);
            coll += 1;
        }
    }
    w.smax[c] = smax;
}

// BGK-family relaxation (stochasticParticleBGK.C:287-796,
// stochasticParticleESBGK.C:311-906, unifiedStochasticParticleSBGK.C:379-1096)
// in three phases so that the velocity draws run one simulator per lane:
//  A  (lane per cell)  moments, Pr, nu, USP q/sigma time relaxation, the
//     number of relaxed simulators and their selection (partial Fisher-Yates
//     on the cell stream); parameters go to a per-cell table;
//  B  (lane per selected simulator) new velocity from its own stream
//     (kRngBgkV, sub = selection index); USP draws test against the cell's
//     maxProb at the start of the step and record the largest raise;
//  C  (lane per cell)  conserveMomentumAndEnergy, maxProb update.
constexpr uint32_t kSelNone = 0xFFFFFFFFu;
constexpr uint32_t kRound = 32;       // cells per phase round in a tile
constexpr uint32_t kRngBgkV = 7;

struct BgkP {
    double U[3];
    double s2kT, p, cQ, cS, m0, T;
    double q[3];
    double sm[9];      // ES-BGK S matrix (row major) or USP stress (xx xy xz yy yz zz)
    int32_t perform;
    int32_t model;
};

template <class R, class Acc>
__device__ inline void bgk_prepare(const WorldView<R>& w, const CaseDev& cs, uint32_t c, uint32_t b, uint32_t nC,
                                   const double* S, const Acc& A, uint32_t* sel, uint32_t sel_base,
                                   uint32_t slot, BgkP& P, unsigned long long& nrel_out)
{
    const int model = cs.bgk_model;
    const double FN = cs.fn, dt = cs.dt, V = w.vol[c];
    const double N = S[UGX_S_N];
    const bool usp = model == UGX_BGK_USP;
    bool perform = true;
    double T = 0.0, p = 0.0, U[3] = {0.0, 0.0, 0.0};
    double Pt[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double q[3], sh[6];
    for (int d = 0; d < 3; ++d) q[d] = w.uspq[3 * (size_t)c + d];
    for (int d = 0; d < 6; ++d) sh[d] = w.usps[6 * (size_t)c + d];
    double Pr = 0.0, nu = 0.0;
    if (N > kVSmall) {
        const double rhoN = FN * N / V;
        for (int d = 0; d < 3; ++d) U[d] = S[UGX_S_MU + d] / S[UGX_S_M];
        const double uu = U[0] * U[0] + U[1] * U[1] + U[2] * U[2];
        T = (S[UGX_S_MCC] - S[UGX_S_M] * uu) / (3.0 * kB * N);
        p = rhoN * kB * T;
        const int ia[6] = {0, 0, 0, 1, 1, 2};
        const int ib[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
        for (int k = 0; k < 6; ++k)
            Pt[k] = rhoN * (S[UGX_S_MUU + k] / N - (S[UGX_S_M] / N) * U[ia[k]] * U[ib[k]]);
        if (usp) {
            const double tr = (Pt[0] + Pt[3] + Pt[5]) / 3.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) sh[k] = Pt[k];
            sh[0] = sh[0] - tr; sh[3] = sh[3] - tr; sh[5] = sh[5] - tr;
            const int r0[3] = {0, 1, 2}, r1[3] = {1, 3, 4}, r2[3] = {2, 4, 5};
#pragma unroll
            for (int d = 0; d < 3; ++d)
                q[d] = rhoN * (0.5 * (S[UGX_S_MCCU + d] / N) - 0.5 * (S[UGX_S_MCC] / N) * U[d]
                               + S[UGX_S_EROTU + d] / N - (S[UGX_S_EROT] / N) * U[d])
                       - Pt[r0[d]] * U[0] - Pt[r1[d]] * U[1] - Pt[r2[d]] * U[2];
        }
        if (N > (usp ? 2.0 : 1.0)) {
            const double f1 = N / (N - 1.0);
            p = f1 * p;
            T = f1 * T;
#pragma unroll
            for (int k = 0; k < 6; ++k) Pt[k] = f1 * Pt[k];
            if (usp) {
                const double f2 = N * N / (N - 1.0) / (N - 2.0);
                for (int d = 0; d < 3; ++d) q[d] = f2 * q[d];
#pragma unroll
                for (int k = 0; k < 6; ++k) sh[k] = f1 * sh[k];
            }
        } else {
            perform = false;
        }
        if (T > kVSmall) {
            double visc = 0.0;
            for (int s = 0; s < cs.ns; ++s) {
                const int g = cs.sp_base + s;
                double nsv = 0.0;
#pragma unroll
                for (int qq = 0; qq < kNS; ++qq) if (qq == s) nsv = S[UGX_S_SPECIES_N + qq];
                visc = visc + nsv * w.sp.muref[g] * dpow(T / cs.tref, w.sp.omega[g]);
                Pr = Pr + nsv * w.sp.prterm[g];
            }
            visc = visc / N;
            Pr = Pr / N;
            nu = (model == UGX_BGK_ESBGK) ? Pr * p / visc : p / visc;
        } else {
            perform = false;
        }
    } else {
        perform = false;
    }
    if (usp) {
        const double th = cs.theta;
        for (int d = 0; d < 3; ++d)
            q[d] = th * q[d] / (1.0 + 0.5 * Pr * nu * dt) + (1.0 - th) * w.uspq[3 * (size_t)c + d];
        for (int k = 0; k < 6; ++k)
            sh[k] = th * sh[k] / (1.0 + 0.5 * nu * dt) + (1.0 - th) * w.usps[6 * (size_t)c + k];
        for (int d = 0; d < 3; ++d) w.uspq[3 * (size_t)c + d] = q[d];
        for (int k = 0; k < 6; ++k) w.usps[6 * (size_t)c + k] = sh[k];
    }
    P.model = model;
    P.m0 = w.uspmax[c];
    P.perform = 0;
    if (perform && w.cmodel[c] == 0) {
        P.perform = 1;
        Rng r;
        rng_init(r, cs.seed_lo, cs.seed_hi, kRngBgk, 0u, c - (uint32_t)cs.cell_begin, cs.step);
        const double pc = N * (1.0 - dexp(-nu * dt));
        const double fl = floor(pc);
        long long nrel = (long long)fl;
        if (uniform(r) < pc - fl) nrel += 1;
        if (nrel > (long long)nC) nrel = nC;
        nrel_out += (unsigned long long)nrel;
        for (uint32_t i = 0; i < nC; ++i) A.idx_set(b, i, i);
        for (long long i = 0; i < nrel; ++i) {
            const uint32_t j = (uint32_t)i + uindex(r, (uint32_t)((long long)nC - i));
            const uint32_t ti = A.idx_get(b, (uint32_t)i), tj = A.idx_get(b, j);
            A.idx_set(b, (uint32_t)i, tj);
            A.idx_set(b, j, ti);
            sel[b - sel_base + tj] = ((uint32_t)i << 6) | slot;
        }
        for (int d = 0; d < 3; ++d) { P.U[d] = U[d]; P.q[d] = q[d]; }
        P.T = T;
        P.p = p;
        P.s2kT = sqrt(2.0 * kB * T);
        const double tau = 0.5 * nu * dt;
        const double ex = 1.0 + 2.0 / (dexp(2.0 * tau) - 1.0);
        P.cQ = 2.0 * (1.0 - Pr * tau * ex);
        P.cS = 1.0 - tau * ex;
        if (model == UGX_BGK_ESBGK) {
            const double a = 0.5 * (1.0 - Pr) / Pr;
            const int full[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                const double dl = (k == 0 || k == 4 || k == 8) ? 1.0 : 0.0;
                P.sm[k] = dl - a * (Pt[full[k]] / p - dl);
            }
        } else {
#pragma unroll
            for (int k = 0; k < 6; ++k) P.sm[k] = sh[k];
        }
    }
}

// Phase B: the new velocity of one selected simulator (global slot i).
template <class R, class Acc>
__device__ inline void bgk_sample(const WorldView<R>& w, const CaseDev& cs, uint32_t c, uint32_t i, uint32_t k,
                                  const BgkP& P, unsigned long long* raise, const Acc& A)
{
    Rng rv;
    rng_init(rv, cs.seed_lo, cs.seed_hi, kRngBgkV, k, c - (uint32_t)cs.cell_begin, cs.step);
    const double u0 = P.s2kT / sqrt(w.sp.mass[cs.sp_base + A.local_species(i)]);
    double xi[3];
    if (P.model == UGX_BGK_USP) {
        const double m0 = P.m0, p = P.p, cQ = P.cQ, cS = P.cS;
        for (int it = 0; it < 100000; ++it) {
            double g0, g1, g2;
            gauss3(rv, g0, g1, g2);
            xi[0] = g0 / sqrt(2.0); xi[1] = g1 / sqrt(2.0); xi[2] = g2 / sqrt(2.0);
            const double vSq = xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2];
            const double vTr = vSq / 3.0;
            const double* sh = P.sm;
            const double prob = 1.0 + cQ / (p * u0) * (P.q[0] * xi[0] + P.q[1] * xi[1] + P.q[2] * xi[2]) * (vSq / 2.5 - 1.0)
                                + cS / p * (sh[0] * (xi[0] * xi[0] - vTr) + sh[3] * (xi[1] * xi[1] - vTr)
                                            + sh[5] * (xi[2] * xi[2] - vTr) + 2.0 * sh[1] * xi[0] * xi[1]
                                            + 2.0 * sh[2] * xi[0] * xi[2] + 2.0 * sh[4] * xi[1] * xi[2]);
            if (prob > m0 && prob < 10.0) {
                atomicMax(raise, (unsigned long long)double_to_bits(prob));
                break;
            }
            if (uniform(rv) < prob / m0) break;
        }
    } else {
        double g0, g1, g2;
        gauss3(rv, g0, g1, g2);
        xi[0] = g0 / sqrt(2.0); xi[1] = g1 / sqrt(2.0); xi[2] = g2 / sqrt(2.0);
        if (P.model == UGX_BGK_ESBGK) {
            double y[3];
            for (int d = 0; d < 3; ++d) y[d] = P.sm[3 * d] * xi[0] + P.sm[3 * d + 1] * xi[1] + P.sm[3 * d + 2] * xi[2];
            xi[0] = y[0]; xi[1] = y[1]; xi[2] = y[2];
        }
    }
    A.store(i, P.U[0] + u0 * xi[0], P.U[1] + u0 * xi[1], P.U[2] + u0 * xi[2]);
}

// Phase C: conserveMomentumAndEnergy (canonical wave sums) and the USP
// maxProb update; called by all lanes of one wave for one cell.
template <class R, class Acc>
__device__ inline void bgk_finish(const WorldView<R>& w, const CaseDev& cs, uint32_t c, uint32_t b, uint32_t nC,
                                  const double* S, const BgkP& P, unsigned long long raise_bits, const Acc& A)
{
    const uint32_t lane = __lane_id();
    if (P.perform) {
        const double N = S[UGX_S_N];
        double ke = 0.0, mom0 = 0.0, mom1 = 0.0, mom2 = 0.0;
        for (uint32_t cb = b; cb < b + nC; cb += 64u) {
            const uint32_t i = cb + lane;
            double m = 0.0, vx = 0.0, vy = 0.0, vz = 0.0;
            if (i < b + nC) {
                m = w.sp.mass[cs.sp_base + A.local_species(i)];
                A.load(i, vx, vy, vz);
            }
            ke = ke + wave_bfly(m * (vx * vx + vy * vy + vz * vz));
            mom0 = mom0 + wave_bfly(m * vx);
            mom1 = mom1 + wave_bfly(m * vy);
            mom2 = mom2 + wave_bfly(m * vz);
        }
        const double pU0 = mom0 / S[UGX_S_M], pU1 = mom1 / S[UGX_S_M], pU2 = mom2 / S[UGX_S_M];
        const double pT = (ke - S[UGX_S_M] * (pU0 * pU0 + pU1 * pU1 + pU2 * pU2)) / (3.0 * kB * (N - 1.0));
        if (pT > kVSmall) {
            const double sc = sqrt(P.T / pT);
            for (uint32_t i = b + lane; i < b + nC; i += 64u) {
                double vx, vy, vz;
                A.load(i, vx, vy, vz);
                A.store(i, P.U[0] + (vx - pU0) * sc, P.U[1] + (vy - pU1) * sc, P.U[2] + (vz - pU2) * sc);
            }
        }
    }
    if (P.model == UGX_BGK_USP && lane == 0) {
        const double raise = bits_to_double(raise_bits);
        w.uspmax[c] = (raise > 0.0) ? raise : P.m0 * 0.999;
    }
}

// Round of up to kRound cells [cb, ce) of a tile:
//  sums    wave per cell: canonical sums into S_tab, sampler and hybrid sums
//  A       lane per cell: BGK moments, selection (wave 0)
//  B       lane per selected simulator: new velocities (all waves)
//  C       wave per cell: conservation, USP maxProb
//  NTC     lane per cell (wave 0)
// sel is indexed by slot - sel_base.
template <class R, class Acc>
__device__ inline void cell_round(const WorldView<R>& w, uint32_t cb, uint32_t ce, const Acc& A, uint32_t* sel,
                                  uint32_t sel_base, BgkP* bp, unsigned long long* raise, double (*S_tab)[kNSamp],
                                  unsigned long long (*s_cnt)[3], int case0)
{
    const uint32_t tid = threadIdx.x;
    const uint32_t wv = tid >> 6, lane = tid & 63u;
    const uint32_t ncell = ce - cb;
    const uint32_t rb = w.off[cb], re = w.off[ce];
    for (uint32_t i = rb + tid; i < re; i += kBlock) sel[i - sel_base] = kSelNone;
    if (tid < kRound) { raise[tid] = 0ull; bp[tid].perform = 0; bp[tid].model = UGX_BGK_NONE; }
    // sums
    for (uint32_t k = wv; k < ncell; k += kBlock / 64u) {
        const uint32_t c = cb + k;
        const CaseDev& cs = w.cases[w.cell_case[c]];
        if (!cs.active) continue;
        const bool sample = ((cs.step + 1u) % (uint32_t)cs.sample_interval) == 0u;
        const bool hybrid = cs.coll_model == UGX_COLL_HYBRID;
        const bool bgk = (cs.coll_model == UGX_COLL_BGK || hybrid) && cs.bgk_model != UGX_BGK_NONE;
        if (!(sample || hybrid || bgk)) continue;
        double S[kNSamp];
        wave_cell_sums<R>(w, cs, A, w.off[c], w.off[c + 1], S);
        if (lane == 0) {
            const double dt = cs.dt;
#pragma unroll
            for (int q = 0; q < kNSamp; ++q) S_tab[k][q] = S[q];
            if (sample) {
#pragma unroll
                for (int q = 0; q < kNSamp; ++q)
                    if (cs.sampler_full || q < UGX_S_MUU || q >= UGX_S_BASE)
                        w.samp[(size_t)q * w.nc + c] += dt * S[q];
            }
            if (hybrid) {
#pragma unroll
                for (int q = 0; q < 8; ++q) w.knacc[(size_t)8 * c + q] += dt * S[q];
            }
        }
    }
    __syncthreads();
    unsigned long long cnt[3] = {0ull, 0ull, 0ull};
    if (tid < ncell) {
        const uint32_t c = cb + tid;
        const CaseDev& cs = w.cases[w.cell_case[c]];
        const bool hybrid = cs.coll_model == UGX_COLL_HYBRID;
        const bool bgk = (cs.coll_model == UGX_COLL_BGK || hybrid) && cs.bgk_model != UGX_BGK_NONE;
        if (cs.active && bgk) {
            const uint32_t b = w.off[c];
            bgk_prepare<R>(w, cs, c, b, w.off[c + 1] - b, S_tab[tid], A, sel, sel_base, tid, bp[tid], cnt[2]);
        }
    }
    __syncthreads();
    for (uint32_t i = rb + tid; i < re; i += kBlock) {
        const uint32_t mark = sel[i - sel_base];
        if (mark == kSelNone) continue;
        const uint32_t sl = mark & 63u;
        const uint32_t cc = cb + sl;
        bgk_sample<R>(w, w.cases[w.cell_case[cc]], cc, i, mark >> 6, bp[sl], &raise[sl], A);
    }
    __syncthreads();
    for (uint32_t k = wv; k < ncell; k += kBlock / 64u) {
        const uint32_t c = cb + k;
        const CaseDev& cs = w.cases[w.cell_case[c]];
        const bool hybrid = cs.coll_model == UGX_COLL_HYBRID;
        const bool bgk = (cs.coll_model == UGX_COLL_BGK || hybrid) && cs.bgk_model != UGX_BGK_NONE;
        if (cs.active && bgk) {
            const uint32_t b = w.off[c];
            bgk_finish<R>(w, cs, c, b, w.off[c + 1] - b, S_tab[k], bp[k], raise[k], A);
        }
    }
    __syncthreads();
    if (tid < ncell) {
        const uint32_t c = cb + tid;
        const CaseDev& cs = w.cases[w.cell_case[c]];
        const bool dsmc = (cs.coll_model == UGX_COLL_DSMC || cs.coll_model == UGX_COLL_HYBRID)
                          && cs.dsmc_model != UGX_DSMC_NONE;
        const uint32_t b = w.off[c], nC = w.off[c + 1] - b;
        if (cs.active && dsmc && w.cmodel[c] == 1 && nC >= 2) ntc_cell<R>(w, cs, c, b, nC, A, cnt[0], cnt[1]);
        const int k = w.cell_case[c] - case0;
        for (int q = 0; q < 3; ++q) {
            if (!cnt[q]) continue;
            if (k >= 0 && k < 4) atomicAdd(&s_cnt[k][q], cnt[q]);
            else atomicAdd(&w.counters[(size_t)w.cell_case[c] * kCntNum + (q == 0 ? kCntCandidates : q == 1 ? kCntCollisions : kCntBgk)], cnt[q]);
        }
    }
    __syncthreads();
}

constexpr uint32_t kTileN = 512;     // nominal simulators per cell tile
constexpr uint32_t kTileMax = 1024;  // LDS capacity of a tile

// One tile = the cells whose first sorted slot lies in [t*kTileN, (t+1)*kTileN)
// (the last tile also owns the trailing empty cells).  Steps:
//  1. each cell's slot list (filled by k_scatter in atomic order) is put in
//     ascending previous-slot order: an entry's rank in its cell is the
//     number of entries of that cell with a smaller slot;
//  2. records are gathered into the sorted buffers (x, cell, species to HBM,
//     velocity, E_rot and species to LDS);
//  3. lane per cell: sampler, BGK, NTC on the LDS copy;
//  4. velocities and E_rot are written back.
// SortOnly skips 3.  Tiles larger than kTileMax run the same steps from HBM.
template <class R, bool SortOnly>
__global__ void __launch_bounds__(kBlock) k_tile(WorldView<R> w)
{
    __shared__ uint32_t s_in[kTileMax];
    __shared__ uint32_t s_perm[kTileMax];
    __shared__ R s_vx[SortOnly ? 1 : kTileMax], s_vy[SortOnly ? 1 : kTileMax], s_vz[SortOnly ? 1 : kTileMax];
    __shared__ R s_er[SortOnly ? 1 : kTileMax];
    __shared__ uint16_t s_idx[SortOnly ? 1 : kTileMax];
    __shared__ uint8_t s_sl[SortOnly ? 1 : kTileMax];
    __shared__ unsigned long long s_cnt[4][3];
    __shared__ BgkP s_bp[SortOnly ? 1 : kRound];
    __shared__ unsigned long long s_raise[SortOnly ? 1 : kRound];
    __shared__ double s_S[SortOnly ? 1 : kRound][kNSamp];
    const uint32_t nn = w.st->n_new;
    const uint32_t ntiles = nn / kTileN + 1u;
    if (blockIdx.x >= ntiles) return;
    const uint32_t tid = threadIdx.x;
    const uint32_t ncp = (uint32_t)w.nc + 1u;
    const uint32_t start = blockIdx.x * kTileN;
    uint32_t c0 = lower_bound_u32(w.off, ncp, start);
    uint32_t c1 = (blockIdx.x == ntiles - 1u) ? (uint32_t)w.nc : lower_bound_u32(w.off, ncp, start + kTileN);
    if (c0 > (uint32_t)w.nc) c0 = (uint32_t)w.nc;
    if (c1 > (uint32_t)w.nc) c1 = (uint32_t)w.nc;
    const uint32_t p0 = w.off[c0], p1 = w.off[c1];
    const uint32_t size = p1 - p0;
    const int case0 = (c0 < (uint32_t)w.nc) ? w.cell_case[c0] : 0;
    if (tid < 12) s_cnt[tid / 3][tid % 3] = 0ull;
    if (size <= kTileMax) {
        for (uint32_t i = tid; i < size; i += kBlock) s_in[i] = w.perm[p0 + i];
        __syncthreads();
        for (uint32_t i = tid; i < size; i += kBlock) {
            const uint32_t slot = s_in[i];
            const uint32_t c = w.a.cell[slot];
            const uint32_t b = w.off[c] - p0, e = w.off[c + 1] - p0;
            uint32_t r = 0;
            for (uint32_t j = b; j < e; ++j) r += (s_in[j] < slot) ? 1u : 0u;
            s_perm[b + r] = slot;
        }
        __syncthreads();
        for (uint32_t i = tid; i < size; i += kBlock) {
            const uint32_t slot = s_perm[i];
            const uint32_t j = p0 + i;
            w.b.x[j] = w.a.x[slot]; w.b.y[j] = w.a.y[slot]; w.b.z[j] = w.a.z[slot];
            const uint32_t g = w.a.sp[slot];
            w.b.cell[j] = w.a.cell[slot];
            w.b.sp[j] = g;
            if (SortOnly) {
                w.b.vx[j] = w.a.vx[slot]; w.b.vy[j] = w.a.vy[slot]; w.b.vz[j] = w.a.vz[slot];
                w.b.erot[j] = w.a.erot[slot];
            } else {
                s_vx[i] = w.a.vx[slot]; s_vy[i] = w.a.vy[slot]; s_vz[i] = w.a.vz[slot];
                s_er[i] = w.a.erot[slot];
                s_sl[i] = (uint8_t)w.sp.local[g];
            }
        }
        if (SortOnly) return;
        __syncthreads();
        LdsAcc<R> A{s_vx, s_vy, s_vz, s_er, s_sl, s_idx, p0};
        for (uint32_t cb = c0; cb < c1; cb += kRound)
            cell_round<R>(w, cb, min(cb + kRound, c1), A, s_in, p0, s_bp, s_raise, s_S, s_cnt, case0);
        for (uint32_t i = tid; i < size; i += kBlock) {
            const uint32_t j = p0 + i;
            w.b.vx[j] = s_vx[i]; w.b.vy[j] = s_vy[i]; w.b.vz[j] = s_vz[i];
            w.b.erot[j] = s_er[i];
        }
    } else {
        // Tile larger than LDS: the same steps on HBM, slot lists ordered into scratch.
        if (tid == 0) atomicAdd(&w.st->big_tiles, 1u);
        for (uint32_t i = tid; i < size; i += kBlock) {
            const uint32_t slot = w.perm[p0 + i];
            const uint32_t c = w.a.cell[slot];
            const uint32_t b = w.off[c], e = w.off[c + 1];
            uint32_t r = 0;
            for (uint32_t j = b; j < e; ++j) r += (w.perm[j] < slot) ? 1u : 0u;
            w.scratch[b + r] = slot;
        }
        __syncthreads();
        for (uint32_t i = tid; i < size; i += kBlock) {
            const uint32_t slot = w.scratch[p0 + i];
            const uint32_t j = p0 + i;
            w.b.x[j] = w.a.x[slot]; w.b.y[j] = w.a.y[slot]; w.b.z[j] = w.a.z[slot];
            w.b.vx[j] = w.a.vx[slot]; w.b.vy[j] = w.a.vy[slot]; w.b.vz[j] = w.a.vz[slot];
            w.b.erot[j] = w.a.erot[slot];
            w.b.cell[j] = w.a.cell[slot];
            w.b.sp[j] = w.a.sp[slot];
        }
        if (SortOnly) return;
        __syncthreads();
        GlobAcc<R> A{w.b, w.sp.local, w.perm};
        for (uint32_t cb = c0; cb < c1; cb += kRound)
            cell_round<R>(w, cb, min(cb + kRound, c1), A, w.scratch, 0u, s_bp, s_raise, s_S, s_cnt, case0);
    }
    __syncthreads();
    if (tid < 12) {
        const int k = tid / 3, q = tid % 3;
        const unsigned long long v = s_cnt[k][q];
        if (v && case0 + k < w.ncases)
            atomicAdd(&w.counters[(size_t)(case0 + k) * kCntNum + (q == 0 ? kCntCandidates : q == 1 ? kCntCollisions : kCntBgk)], v);
    }
}

// ------------------------------------------------------------ end step

template <class R>
__global__ void k_end_step(WorldView<R> w)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < w.ncases) {
        CaseDev& cs = w.cases[c];
        if (cs.active) {
            if (((cs.step + 1u) % (uint32_t)cs.sample_interval) == 0u) w.counters[(size_t)c * kCntNum + kCntSampled] += 1ull;
            w.counters[(size_t)c * kCntNum + kCntParticleSteps] +=
                (unsigned long long)(w.off[cs.cell_end] - w.off[cs.cell_begin]);
            if (cs.coll_model == UGX_COLL_HYBRID) cs.kn_time += cs.dt;
            cs.step += 1u;
        }
    }
    if (c == 0) w.st->n = w.st->n_new;
}

template <class R>
__global__ void k_finish_sort(WorldView<R> w)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) { w.st->n = w.st->n_new; w.st->n_ins = 0u; }
}

// --------------------------------------------------------------- fill

template <class R>
__global__ void k_fill_count(WorldView<R> w, int ccase, const double* nd, int exact, uint32_t* fcnt, int ns, int ncl)
{
    const CaseDev& cs = w.cases[ccase];
    double ntot = 0.0;
    for (int s = 0; s < ns; ++s) ntot += nd[s];
    const uint32_t total = (uint32_t)ncl * (uint32_t)ns;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t <= total; t += gridDim.x * blockDim.x) {
        if (t == total) { fcnt[t] = 0u; continue; }
        const uint32_t cl = t / (uint32_t)ns, s = t % (uint32_t)ns;
        const uint32_t c = (uint32_t)cs.cell_begin + cl;
        long long k;
        if (exact > 0) {
            k = (ntot > 0.0) ? (long long)floor((double)exact * nd[s] / ntot + 0.5) : 0;
        } else {
            Rng r;
            rng_init(r, cs.seed_lo, cs.seed_hi, kRngFill, s, cl, 0u);
            const double req = nd[s] / cs.fn * w.vol[c];
            const double f0 = floor(req);
            k = (long long)f0;
            if (uniform(r) < req - f0) k += 1;
        }
        fcnt[t] = (uint32_t)k;
    }
}

template <class R>
__global__ void k_fill_gen(WorldView<R> w, int ccase, const double* nd, const double* Tax, double rotT,
                           const double* Uf, int exact, const double* bblo, const double* bbhi,
                           const uint32_t* foff, int ns, int ncl)
{
    const CaseDev& cs = w.cases[ccase];
    const uint32_t total = (uint32_t)ncl * (uint32_t)ns;
    const uint32_t nadd = foff[total];
    const uint32_t n0 = w.st->n;
    if ((unsigned long long)n0 + nadd > (unsigned long long)w.cap) {
        if (blockIdx.x == 0 && threadIdx.x == 0) w.st->overflow = 1u;
        return;
    }
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        const uint32_t cl = t / (uint32_t)ns, s = t % (uint32_t)ns;
        const uint32_t c = (uint32_t)cs.cell_begin + cl;
        const uint32_t g = (uint32_t)cs.sp_base + s;
        Rng r;
        rng_init(r, cs.seed_lo, cs.seed_hi, kRngFill, s, cl, 0u);
        if (exact <= 0) (void)uniform(r);
        const double m = w.sp.mass[g];
        double sig[3];
        for (int d = 0; d < 3; ++d) sig[d] = sqrt(kB * Tax[d] / m);
        const uint32_t cnt = foff[t + 1] - foff[t];
        const double* lo = bblo + 3 * (size_t)cl;
        const double* hi = bbhi + 3 * (size_t)cl;
        for (uint32_t i = 0; i < cnt; ++i) {
            const uint32_t j = n0 + foff[t] + i;
            double x[3] = {0.0, 0.0, 0.0};
            for (int it = 0; it < 1000; ++it) {
                double y[3];
                for (int d = 0; d < 3; ++d) {
                    const double u = uniform(r);
                    y[d] = lo[d] + u * (hi[d] - lo[d]);
                }
                bool inside = true;
                for (int f = 0; f < kMaxFaces; ++f) {
                    const V4<R> q = w.geom.plane[(size_t)c * kMaxFaces + f];
                    if ((double)q.x * y[0] + (double)q.y * y[1] + (double)q.z * y[2] > (double)q.w) { inside = false; break; }
                }
                if (inside) { x[0] = y[0]; x[1] = y[1]; x[2] = y[2]; break; }
            }
            double g0, g1, g2;
            gauss3(r, g0, g1, g2);
            w.a.x[j] = (R)(x[0] * cs.mask[0] + 0.0);
            w.a.y[j] = (R)(x[1] * cs.mask[1] + 0.0);
            w.a.z[j] = (R)(x[2] * cs.mask[2] + 0.0);
            w.a.vx[j] = (R)(sig[0] * g0 + Uf[0]);
            w.a.vy[j] = (R)(sig[1] * g1 + Uf[1]);
            w.a.vz[j] = (R)(sig[2] * g2 + Uf[2]);
            w.a.erot[j] = (R)equip_erot(r, rotT, w.sp.rotdof[g]);
            w.a.cell[j] = c;
            w.a.sp[j] = g;
        }
    }
}

template <class R>
__global__ void k_fill_finish(WorldView<R> w, const uint32_t* foff, uint32_t total)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (!w.st->overflow) w.st->n = w.st->n + foff[total];
        w.st->n_ins = 0u;
    }
}

// ------------------------------------------------------------ launchers

static inline unsigned grid_for(uint64_t n, unsigned cap_blocks = 8192)
{
    uint64_t g = (n + kBlock - 1) / kBlock;
    if (g < 1) g = 1;
    if (g > cap_blocks) g = cap_blocks;
    return (unsigned)g;
}

static inline unsigned tile_grid(uint32_t cap) { return (unsigned)(cap / kTileN) + 2u; }

template <class R>
void launch_sort(const WorldView<R>& w, hipStream_t s)
{
    const uint32_t ncp = (uint32_t)w.nc + 1u;
    hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(ncp)), dim3(kBlock), 0, s, w.cnt, ncp);
    hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(ncp)), dim3(kBlock), 0, s, w.cursor, ncp);
    hipLaunchKernelGGL(k_hist<R>, dim3(grid_for(w.cap)), dim3(kBlock), 0, s, w);
    scan_u32(w.cnt, ncp, w.off, w.blocksum, &w.st->n_new, s);
    hipLaunchKernelGGL(k_scatter<R>, dim3(grid_for(w.cap)), dim3(kBlock), 0, s, w);
    hipLaunchKernelGGL((k_tile<R, true>), dim3(tile_grid(w.cap)), dim3(kBlock), 0, s, w);
    hipLaunchKernelGGL(k_finish_sort<R>, dim3(1), dim3(1), 0, s, w);
}

template <class R>
void launch_step(const WorldView<R>& w0, hipStream_t s, hipEvent_t* marks)
{
    WorldView<R> w = w0;
    const uint32_t ncp = (uint32_t)w.nc + 1u;
    const unsigned gc = grid_for((uint64_t)w.ncases);
    auto mark = [&](int i) { if (marks) (void)hipEventRecord(marks[i], s); };
    mark(0);
    hipLaunchKernelGGL(k_begin_step<R>, dim3(gc), dim3(kBlock), 0, s, w);
    hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(ncp)), dim3(kBlock), 0, s, w.cnt, ncp);
    hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(ncp)), dim3(kBlock), 0, s, w.cursor, ncp);
    mark(1);
    if (w.n_inflow > 0) {
        hipLaunchKernelGGL(k_inflow_count<R>, dim3(1), dim3(1024), 0, s, w);
        hipLaunchKernelGGL(k_inflow_gen<R>, dim3(grid_for(w.max_ins)), dim3(kBlock), 0, s, w);
    }
    mark(2);
    hipLaunchKernelGGL(k_move<R>, dim3(grid_for(w.cap)), dim3(kBlock), 0, s, w);
    mark(3);
    hipLaunchKernelGGL(k_wall<R>, dim3(grid_for(w.wq_cap, 2048)), dim3(kBlock), 0, s, w);
    mark(4);
    scan_u32(w.cnt, ncp, w.off, w.blocksum, &w.st->n_new, s);
    mark(5);
    hipLaunchKernelGGL(k_scatter<R>, dim3(grid_for(w.cap)), dim3(kBlock), 0, s, w);
    mark(6);
    hipLaunchKernelGGL((k_tile<R, false>), dim3(tile_grid(w.cap)), dim3(kBlock), 0, s, w);
    mark(7);
    hipLaunchKernelGGL(k_end_step<R>, dim3(gc), dim3(kBlock), 0, s, w);
    mark(8);
}

template <class R>
void launch_fill(const WorldView<R>& w, hipStream_t s, int ccase, const double* nd, const double* T, double rotT,
                 const double* U, int exact, const double* bblo, const double* bbhi, uint32_t* fcnt,
                 uint32_t* foff, int ns, int ncl)
{
    const uint32_t total = (uint32_t)ncl * (uint32_t)ns;
    hipLaunchKernelGGL(k_fill_count<R>, dim3(grid_for(total + 1)), dim3(kBlock), 0, s, w, ccase, nd, exact, fcnt, ns, ncl);
    scan_u32(fcnt, total + 1, foff, w.blocksum, nullptr, s);
    hipLaunchKernelGGL(k_fill_gen<R>, dim3(grid_for(total)), dim3(kBlock), 0, s, w, ccase, nd, T, rotT, U, exact,
                       bblo, bbhi, foff, ns, ncl);
    hipLaunchKernelGGL(k_fill_finish<R>, dim3(1), dim3(1), 0, s, w, foff, total);
}

__global__ void k_copy_probe(const float4* __restrict__ in, float4* __restrict__ out, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) out[i] = in[i];
}

void launch_copy_probe(const void* in, void* out, size_t bytes, hipStream_t s)
{
    size_t n = bytes / sizeof(float4);
    hipLaunchKernelGGL(k_copy_probe, dim3(grid_for(n, 65536)), dim3(kBlock), 0, s, (const float4*)in, (float4*)out, n);
}

template void launch_step<float>(const WorldView<float>&, hipStream_t, hipEvent_t*);
template void launch_step<double>(const WorldView<double>&, hipStream_t, hipEvent_t*);
template void launch_sort<float>(const WorldView<float>&, hipStream_t);
template void launch_sort<double>(const WorldView<double>&, hipStream_t);
template void launch_fill<float>(const WorldView<float>&, hipStream_t, int, const double*, const double*, double,
                                 const double*, int, const double*, const double*, uint32_t*, uint32_t*, int, int);
template void launch_fill<double>(const WorldView<double>&, hipStream_t, int, const double*, const double*, double,
                                  const double*, int, const double*, const double*, uint32_t*, uint32_t*, int, int);

}  // namespace ugx
