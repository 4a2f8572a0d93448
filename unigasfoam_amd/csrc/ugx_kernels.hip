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

#include <algorithm>
#include <type_traits>
#include <utility>

#include "ugx_kernels.hpp"

namespace ugx {

constexpr int kBlock = 256;
constexpr int kScanBlock = 1024;      // elements per scan block (256 threads x 4)

// Adds one to ctr[key[u]] for every active lane and key set u and returns
// each lane's rank among the lanes of the wave that added to the same
// counter, offset by the counter's previous value.  One atomic per distinct
// key of the wave and key set; every atomic of all U sets is issued before
// any result is used, so the wave waits for one round trip in total.
// Called by every lane of the wave.
template <int U>
__device__ inline void wave_agg_inc(uint32_t* ctr, const uint32_t (&key)[U], const bool (&active)[U],
                                    uint32_t (&rank)[U])
{
    const unsigned lane = __lane_id();
    uint32_t lane_base[U], my_leader[U], my_rank[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        lane_base[u] = 0u;
        my_leader[u] = 0u;
        my_rank[u] = 0u;
        unsigned long long pending = __ballot(active[u]);
        while (pending) {
            const int leader = __ffsll(pending) - 1;
            const uint32_t lkey = __shfl(key[u], leader);
            const unsigned long long grp = __ballot(active[u] && key[u] == lkey) & pending;
            if ((int)lane == leader) lane_base[u] = atomicAdd(&ctr[lkey], (uint32_t)__popcll(grp));
            if ((grp >> lane) & 1ull) {
                my_leader[u] = (uint32_t)leader;
                my_rank[u] = (uint32_t)__popcll(grp & ((1ull << lane) - 1ull));
            }
            pending &= ~grp;
        }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) rank[u] = __shfl(lane_base[u], (int)my_leader[u]) + my_rank[u];
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
// T is the per-thread count type; 32-bit counts go to the global counters
// before one of them could pass 2^31.
// Slot maps: M = 0 the tracking counters above, M = 1 the cell weighting's
// (0 clones, 1 removals).
template <class T, int M = 0>
struct TCount {
    int ccase;
    T v[4];
};

template <int M>
__device__ inline int tcount_idx(int k)
{
    constexpr int kMap0[4] = {kCntDeleted, kCntWallHits, kCntTrackErrors, kCntInserted};
    constexpr int kMap1[4] = {kCntCloned, kCntWeightDeleted, kCntCloned, kCntCloned};
    return M == 0 ? kMap0[k] : kMap1[k];
}

template <int M>
__device__ inline void tcount_flush_one(int ccase, const unsigned long long (&v)[4], unsigned long long* counters)
{
    if (ccase < 0) return;
    for (int k = 0; k < 4; ++k)
        if (v[k]) atomicAdd(&counters[(size_t)ccase * kCntNum + tcount_idx<M>(k)], v[k]);
}

template <class T, int M>
__device__ inline void tcount_add(TCount<T, M>& t, unsigned long long* counters, int ccase, int k, uint32_t v)
{
    if (!v) return;
    bool flush = t.ccase != ccase;
    if constexpr (sizeof(T) == sizeof(uint32_t)) flush = flush || t.v[k] >= 0x80000000u - v;
    if (flush) {
        const unsigned long long o[4] = {t.v[0], t.v[1], t.v[2], t.v[3]};
        tcount_flush_one<M>(t.ccase, o, counters);
        t.ccase = ccase;
        t.v[0] = t.v[1] = t.v[2] = t.v[3] = T(0);
    }
    t.v[k] += v;
}

// Called by every thread of the block.
template <class T, int M>
__device__ inline void tcount_block_flush(const TCount<T, M>& t, unsigned long long* counters)
{
    __shared__ int s_case[kBlock];
    __shared__ unsigned long long s_v[kBlock][4];
    const uint32_t tid = threadIdx.x;
    s_case[tid] = t.ccase;
    for (int k = 0; k < 4; ++k) s_v[tid][k] = t.v[k];
    __syncthreads();
    if (tid == 0) {
        int sc = -1;
        unsigned long long sum[4] = {0ull, 0ull, 0ull, 0ull};
        for (uint32_t i = 0; i < kBlock; ++i) {
            if (s_case[i] < 0) continue;
            if (sc < 0) sc = s_case[i];
            if (s_case[i] == sc) {
                for (int k = 0; k < 4; ++k) sum[k] += s_v[i][k];
            } else {
                const unsigned long long o[4] = {s_v[i][0], s_v[i][1], s_v[i][2], s_v[i][3]};
                tcount_flush_one<M>(s_case[i], o, counters);
            }
        }
        tcount_flush_one<M>(sc, sum, counters);
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
    if (c == 0) {
        w.st->n_ins = 0u; w.st->nwall = 0u; w.st->nsel = 0u;
        w.st->nev_peak = max(w.st->nev_peak, w.st->nev);
        w.st->nev = 0u;
    }
}

// ------------------------------------------------------------- inflow

// One block: insert counts per face-species (uniGasGeneralBoundary.C:602-609)
// and their exclusive scan.
template <class R>
__global__ void __launch_bounds__(1024) k_inflow_count(WorldView<R> w)
{
    __shared__ uint32_t sh[16];
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
                // divided by F_N W_c of the face's cell (W_c = 1 unweighted: the same as nbar)
                double nb = e.nbar_num / (e.nbar_den * w.cw[e.cell]);
                double fl = floor(nb);
                long long c = (long long)fl;
                if (c < 0) c = 0;
                if (uniform(r) < nb - fl) c += 1;
                cnt = (uint32_t)c;
            }
        }
        // block scan: shuffles within each wave, then the 16 wave totals
        uint32_t inc = cnt;
        const int lane = tid & 63;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t u = __shfl_up(inc, o);
            if (lane >= o) inc += u;
        }
        if (lane == 63) sh[tid >> 6] = inc;
        __syncthreads();
        uint32_t below = 0u, tot = 0u;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const uint32_t v = sh[q];
            if (q < (tid >> 6)) below += v;
            tot += v;
        }
        if (k < w.n_inflow) w.ins_off[k] = carry + below + inc - cnt;
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

// The same counts with a thread per face-species, written into ins_off for
// an exclusive scan in place (ins_off[n_inflow] = 0 ends it).
template <class R>
__global__ void k_inflow_face_count(WorldView<R> w)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k <= w.n_inflow; k += gridDim.x * blockDim.x) {
    uint32_t cnt = 0;
    if (k < w.n_inflow) {
        const InflowFS& e = w.inflow[k];
        const CaseDev& cs = w.cases[e.ccase];
        if (cs.active) {
            Rng r;
            rng_init(r, cs.seed_lo, cs.seed_hi, kRngInflowN, 0u, (uint32_t)e.local_id, cs.step);
            const double nb = e.nbar_num / (e.nbar_den * w.cw[e.cell]);
            const double fl = floor(nb);
            long long c = (long long)fl;
            if (c < 0) c = 0;
            if (uniform(r) < nb - fl) c += 1;
            cnt = (uint32_t)c;
        }
    }
    w.ins_off[k] = cnt;
    }
}

// After the scan: the step's inflow total, or none when it would exceed the
// capacity (the overflow is reported).
template <class R>
__global__ void k_inflow_finish(WorldView<R> w)
{
    uint32_t carry = w.ins_off[w.n_inflow];
    if ((unsigned long long)w.st->n + carry > (unsigned long long)w.cap) {
        w.st->overflow = 1u;
        carry = 0u;
    }
    w.st->n_ins = carry;
}

// Bird eq. 12.5 insertion + first move over (1 - R) deltaT.
template <class R>
__device__ inline int make_inflow(const WorldView<R>& w, const InflowFS& e, const CaseDev& cs, Rng& r,
                                  R x[3], double U[3], double& erot, uint32_t& cell, TrackOut& to, uint32_t slot)
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
    to = track<R>(w.geom, w.bm, cs.mask, cell, x, U, erot, MassOf{m, rd, (uint32_t)g}, slot, (1.0 - frac) * cs.dt, r,
                  wall_sink(w));
    return to.alive;
}

template <class R>
__global__ void __launch_bounds__(kBlock) k_inflow_gen(WorldView<R> w)
{
    TCount<unsigned long long> tc{-1, {0ull, 0ull, 0ull, 0ull}};
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
            uint32_t j = n0 + t;
            make_inflow<R>(w, e, cs, r, x, U, erot, cell, to, j);
            w.a.x[j] = x[0]; w.a.y[j] = x[1]; w.a.z[j] = x[2];
            w.a.vx[j] = (R)U[0]; w.a.vy[j] = (R)U[1]; w.a.vz[j] = (R)U[2];
            w.a.erot[j] = (R)erot;
            w.a.cell[j] = cell;
            w.a.sp[j] = (uint32_t)e.species;
            // a new simulator carries the weight of its insertion cell (uniGasGeneralBoundary.C:738-748)
            if (w.has_weighting) w.cstart[j] = (uint32_t)e.cell;
        }
        if (active) {
            tcount_add(tc, w.counters, ccase, 3, 1u);
            tcount_add(tc, w.counters, ccase, 0, (cell != kDead) ? 0u : 1u);
            tcount_add(tc, w.counters, ccase, 1, (uint32_t)to.hits);
            tcount_add(tc, w.counters, ccase, 2, (uint32_t)to.errs);
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

// Case fields the move needs: the time step in the state precision and the
// moving directions as a bit set (bit d set when direction d moves).
template <class R>
struct MoveCase {
    R dt;
    uint32_t mb;
    int32_t active;
};

template <class R>
__device__ inline MoveCase<R> move_case(const WorldView<R>& w, int c)
{
    const CaseDev& cs = w.cases[c];
    const uint32_t mb = (cs.mask[0] != 0.0 ? 1u : 0u) | (cs.mask[1] != 0.0 ? 2u : 0u) | (cs.mask[2] != 0.0 ? 4u : 0u);
    return MoveCase<R>{(R)cs.dt, mb, cs.active};
}

// Slots of one refill chunk: one per lane of a wave.
constexpr uint32_t kMoveChunk = 64;

// One chunk of simulator records, lane k holding slot base + k.
template <class R>
struct MoveRec {
    R x, y, z, u, v, w;
    uint32_t cell;
};

template <class R>
__device__ inline void load_chunk(const WorldView<R>& w, uint64_t base, uint32_t n0, uint32_t lane, MoveRec<R>& r)
{
    const uint64_t s = base + lane;
    if (s < n0) {
        r.cell = w.a.cell[s];
        r.x = w.a.x[s]; r.y = w.a.y[s]; r.z = w.a.z[s];
        r.u = w.a.vx[s]; r.v = w.a.vy[s]; r.w = w.a.vz[s];
    }
}

template <class R>
__device__ inline MoveRec<R> shfl_rec(const MoveRec<R>& r, int src)
{
    MoveRec<R> g;
    g.x = __shfl(r.x, src); g.y = __shfl(r.y, src); g.z = __shfl(r.z, src);
    g.u = __shfl(r.u, src); g.v = __shfl(r.v, src); g.w = __shfl(r.w, src);
    g.cell = __shfl(r.cell, src);
    return g;
}

__device__ inline uint32_t chunk_valid(uint64_t base, uint32_t n0)
{
    return (base < n0) ? (uint32_t)min<uint64_t>(kMoveChunk, n0 - base) : 0u;
}

// Retired move results of a wave's most recent chunks, staged in LDS so that
// each chunk is stored to HBM as whole lines once all of its simulators have
// finished.  Entry e of ring slot q % kMoveRing holds slot base(q) + e; a
// cell of kMoveUnset marks an entry whose simulator has not finished.
constexpr uint32_t kMoveRing = 4;
constexpr uint32_t kMoveUnset = 0xFFFFFFFEu;
constexpr int kMoveWaves = kBlock / 64;  // waves of a move block

template <class R>
struct MoveStage {
    R x[kMoveWaves][kMoveRing][kMoveChunk];
    R y[kMoveWaves][kMoveRing][kMoveChunk];
    R z[kMoveWaves][kMoveRing][kMoveChunk];
    uint32_t cell[kMoveWaves][kMoveRing][kMoveChunk];
};

// Orders a wave's LDS stores before its later LDS loads from other lanes.
__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Stores the finished entries of ring chunk q (slots base.., valid of them)
// to HBM, lane k storing entry k.
template <class R>
__device__ inline void move_flush(const WorldView<R>& w, MoveStage<R>& st, int wv, uint32_t q, uint64_t base,
                                  uint32_t valid, uint32_t lane)
{
    const uint32_t r = q % kMoveRing;
    const uint32_t c = st.cell[wv][r][lane];
    if (lane < valid && c != kMoveUnset) {
        const uint64_t s = base + lane;
        w.a.x[s] = st.x[wv][r][lane];
        w.a.y[s] = st.y[wv][r][lane];
        w.a.z[s] = st.z[wv][r][lane];
        w.a.cell[s] = c;
    }
}

// True when every entry of ring chunk q has finished.
template <class R>
__device__ inline bool move_complete(const MoveStage<R>& st, int wv, uint32_t q, uint32_t valid, uint32_t lane)
{
    const uint32_t c = st.cell[wv][q % kMoveRing][lane];
    return __ballot(lane < valid && c == kMoveUnset) == 0ull;
}

// Ballistic move of the existing simulators (the face-crossing loop of
// track_t<true> restated lane by lane).  Specular walls, cyclic and deletion
// faces are handled in line; a simulator reaching a diffuse wall stops there
// and is queued for k_wall, which keeps this kernel free of the FP64
// transcendental code of the wall model.
//
// One loop pass advances every lane by one face crossing.  A lane whose
// simulator has finished takes the next slot of its wave at once, so a wave
// is not held by its slowest simulator.  A wave works through chunks of
// kMoveChunk slots, interleaved over the waves of the grid so that regions of
// fast flow are spread over all of them.  The records of the current chunk
// and of the next are held in registers (lane k holds slot base + k), so a
// refill is a cross-lane exchange rather than a memory round trip; the next
// chunk is loaded when the current one is used up.  A finished simulator's
// position and cell go to the wave's LDS ring (MoveStage) and a chunk is
// stored as whole lines when its last simulator finishes; a chunk still
// unfinished when the ring needs its place is stored without the stragglers,
// which then store their own results.  The arithmetic of each simulator is
// exactly that of track_t<true>.
template <class R>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_move(WorldView<R> w)
{
    TCount<uint32_t> tc{-1, {0u, 0u, 0u, 0u}};
    const uint32_t n0 = w.st->n;
    const uint32_t lane = __lane_id();
    const unsigned long long below = (1ull << lane) - 1ull;
    const uint64_t cstride = (uint64_t)gridDim.x * (blockDim.x >> 6) * kMoveChunk;
    // chunk C (slots cbase.., cvalid of them, pos taken) and the next chunk N
    uint64_t cbase = ((uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * kMoveChunk;
    uint64_t nbase = cbase + cstride;
    uint32_t cvalid = chunk_valid(cbase, n0), nvalid = chunk_valid(nbase, n0);
    uint32_t pos = 0u;
    MoveRec<R> C{}, N{};
    load_chunk(w, cbase, n0, lane, C);
    load_chunk(w, nbase, n0, lane, N);
    // chunk indices: C is chunk cq, N chunk cq + 1; chunks oldest..cq + 1 are in the ring
    __shared__ MoveStage<R> stage;
    const int wv = (int)(threadIdx.x >> 6);
    const uint64_t base0 = cbase;
    uint32_t cq = 0u, oldest = 0u;
    // slots are 32-bit: the sum wraps to the slot exactly
    auto slot_of = [&](uint32_t v) { return (uint32_t)base0 + (v >> 6) * (uint32_t)cstride + (v & 63u); };
    stage.cell[wv][0][lane] = kMoveUnset;
    stage.cell[wv][1][lane] = kMoveUnset;
    // case cache: slots [wb, we) belong to case wc
    int wc = 0;
    uint64_t wb = 1u, we = 0u;
    MoveCase<R> wcs{};
    // lane state
    bool has = false;
    // tqo: the chunk index of the lane's slot times 64 plus its entry in the chunk
    uint32_t tqo = 0u, c = 0u, mb = 0u;
    int ccase = 0, it = 0, hits = 0;
    R x[3] = {(R)0, (R)0, (R)0}, U[3] = {(R)0, (R)0, (R)0};
    R tau = (R)0;
    for (;;) {
        // refill: lanes without a simulator take the next slots in lane order
        unsigned long long need = __ballot(!has);
        while (need != 0ull && pos < cvalid) {
            const uint32_t availc = cvalid - pos;
            const uint32_t ntake = min((uint32_t)__popcll(need), availc + nvalid);
            const uint32_t k = (uint32_t)__popcll(need & below);
            const bool taker = ((need >> lane) & 1ull) && k < ntake;
            const bool fromn = k >= availc;
            const int src = (int)(fromn ? k - availc : pos + k);
            MoveRec<R> g = shfl_rec(C, src);
            if (ntake > availc) {
                const MoveRec<R> h = shfl_rec(N, src);
                if (fromn) g = h;
            }
            // the slots taken lie in [first, last]; a case is a contiguous slot range
            const uint64_t first = cbase + pos;
            const uint64_t last = (ntake > availc) ? nbase + (ntake - availc) - 1u : cbase + pos + ntake - 1u;
            if (!(first >= wb && last < we)) {
                wc = case_of_slot(w, (uint32_t)first);
                wb = w.cases[wc].particle_begin;
                we = (wc + 1 < w.ncases) ? (uint64_t)w.cases[wc + 1].particle_begin : ~0ull;
                wcs = move_case(w, wc);
            }
            if (taker) {
                const uint32_t sl = (uint32_t)(fromn ? nbase + (k - availc) : cbase + pos + k);
                int cc = wc;
                MoveCase<R> cs = wcs;
                if (last >= we) {
                    cc = case_of_slot(w, sl);
                    cs = move_case(w, cc);
                }
                const uint32_t q = fromn ? cq + 1u : cq;
                const uint32_t e = fromn ? k - availc : pos + k;
                // simulators of a paused case (a retired batch case) stay where they
                // are (their entry keeps the record); their lane takes another slot
                if (!cs.active) {
                    const uint32_t r = q % kMoveRing;
                    stage.x[wv][r][e] = g.x; stage.y[wv][r][e] = g.y; stage.z[wv][r][e] = g.z;
                    stage.cell[wv][r][e] = g.cell;
                } else {
                    tqo = (q << 6) | e;
                    has = true;
                    ccase = cc;
                    c = g.cell;
                    if (w.has_weighting) w.cstart[sl] = g.cell;
                    x[0] = g.x; x[1] = g.y; x[2] = g.z;
                    U[0] = g.u; U[1] = g.v; U[2] = g.w;
                    tau = cs.dt;
                    it = 0;
                    hits = 0;
                    mb = cs.mb;
                }
            }
            pos += ntake;
            if (pos >= cvalid) {
                pos -= cvalid;
                C = N;
                cbase = nbase;
                cvalid = nvalid;
                nbase += cstride;
                nvalid = chunk_valid(nbase, n0);
                load_chunk(w, nbase, n0, lane, N);
                cq += 1u;
                // the new N takes ring place cq + 1: store the oldest chunk first if it
                // holds that place (stragglers of a chunk stored early store themselves)
                wave_sync();
                while (cq + 1u - oldest >= kMoveRing) {
                    const uint64_t ob = base0 + (uint64_t)oldest * cstride;
                    move_flush(w, stage, wv, oldest, ob, chunk_valid(ob, n0), lane);
                    oldest += 1u;
                }
                wave_sync();
                stage.cell[wv][(cq + 1u) % kMoveRing][lane] = kMoveUnset;
                wave_sync();
            }
            need = __ballot(!has);
        }
        if (__ballot(has) == 0ull) break;
        bool fin = false, queued = false, err = false;
        int qface = 0;
        if (has) {
            // one crossing: all face data of the cell in one memory round trip;
            // vt is the velocity along the moving directions
            const size_t base = (size_t)c * kMaxFaces;
            R vt[3];
            for (int d = 0; d < 3; ++d) vt[d] = U[d] * (((mb >> d) & 1u) ? (R)1 : (R)0);
            V4<R> q[kMaxFaces];
            int32_t nb[kMaxFaces];
#pragma unroll
            for (int f = 0; f < kMaxFaces; ++f) { q[f] = w.geom.plane[base + f]; nb[f] = w.geom.nbr[base + f]; }
            R tmin = tau;
            int fmin = -1;
            int32_t code = 0;
            R qx = (R)0, qy = (R)0, qz = (R)0;
#pragma unroll
            for (int f = 0; f < kMaxFaces; ++f) {
                const R dn = q[f].x * vt[0] + q[f].y * vt[1] + q[f].z * vt[2];
                R tf = (q[f].w - (q[f].x * x[0] + q[f].y * x[1] + q[f].z * x[2])) / dn;
                tf = (tf < (R)0) ? (R)0 : tf;
                const bool take = (nb[f] != kNbrEmpty) && (dn > (R)0) && (tf < tmin);
                tmin = take ? tf : tmin;
                fmin = take ? f : fmin;
                code = take ? nb[f] : code;
                qx = take ? q[f].x : qx;
                qy = take ? q[f].y : qy;
                qz = take ? q[f].z : qz;
            }
            if (fmin < 0) {
                for (int d = 0; d < 3; ++d) x[d] = x[d] + vt[d] * tau;
                fin = true;
            } else {
                for (int d = 0; d < 3; ++d) x[d] = x[d] + vt[d] * tmin;
                tau = tau - tmin;
                bool next = true;
                if (code >= 0) {
                    const V4<R> dl = w.geom.delta[base + fmin];
                    x[0] = x[0] + dl.x;
                    x[1] = x[1] + dl.y;
                    x[2] = x[2] + dl.z;
                    c = (uint32_t)code;
                } else {
                    const int kind = w.bm[-1 - code].kind;
                    if (kind == UGX_BND_DELETION) {
                        c = kDead;
                        fin = true;
                        next = false;
                    } else if (kind != UGX_BND_SPECULAR) {
                        queued = true;
                        qface = fmin;
                        fin = true;
                        next = false;
                    } else {
                        hits += 1;
                        const double pre[3] = {(double)U[0], (double)U[1], (double)U[2]};
                        const R udn = U[0] * qx + U[1] * qy + U[2] * qz;
                        if (udn > (R)0) {
                            U[0] = U[0] - (R)2 * udn * qx;
                            U[1] = U[1] - (R)2 * udn * qy;
                            U[2] = U[2] - (R)2 * udn * qz;
                        }
                        const WallSink sk = wall_sink(w);
                        if (sk.measured(c, fmin)) {
                            const double post[3] = {(double)U[0], (double)U[1], (double)U[2]};
                            const uint32_t t = slot_of(tqo);
                            const double er = (double)w.a.erot[t];
                            sk.hit(c, fmin, t, it, pre, post, er, er, w.a.sp[t]);
                        }
                    }
                }
                if (next) {
                    it += 1;
                    if (it >= kMaxCrossings) { err = true; fin = true; }
                }
            }
        }
        if (fin) {
            has = false;
            const uint32_t t = slot_of(tqo);
            if ((tqo >> 6) >= oldest) {
                const uint32_t r = (tqo >> 6) % kMoveRing, e = tqo & 63u;
                stage.x[wv][r][e] = x[0]; stage.y[wv][r][e] = x[1]; stage.z[wv][r][e] = x[2];
                stage.cell[wv][r][e] = c;
            } else {
                w.a.x[t] = x[0]; w.a.y[t] = x[1]; w.a.z[t] = x[2];
                w.a.cell[t] = c;
            }
            if (hits) {
                w.a.vx[t] = U[0]; w.a.vy[t] = U[1]; w.a.vz[t] = U[2];
            }
            if (queued) {
                const uint32_t qi = atomicAdd(&w.st->nwall, 1u);
                if (qi < w.wq_cap) {
                    w.wq_slot[qi] = t;
                    w.wq_info[qi] = (uint32_t)qface | ((uint32_t)it << 3);
                    w.wq_tau[qi] = (double)tau;
                } else {
                    w.st->overflow = 1u;
                }
            }
            tcount_add(tc, w.counters, ccase, 0, (c == kDead) ? 1u : 0u);
            tcount_add(tc, w.counters, ccase, 1, (uint32_t)hits);
            tcount_add(tc, w.counters, ccase, 2, err ? 1u : 0u);
        }
        // store the taken chunks whose simulators have all finished
        wave_sync();
        while (oldest < cq) {
            const uint64_t ob = base0 + (uint64_t)oldest * cstride;
            const uint32_t ov = chunk_valid(ob, n0);
            if (!move_complete(stage, wv, oldest, ov, lane)) break;
            move_flush(w, stage, wv, oldest, ob, ov, lane);
            oldest += 1u;
        }
        wave_sync();
    }
    // every simulator has finished: store what the ring still holds
    wave_sync();
    for (; oldest <= cq; ++oldest) {
        const uint64_t ob = base0 + (uint64_t)oldest * cstride;
        move_flush(w, stage, wv, oldest, ob, chunk_valid(ob, n0), lane);
    }
    tcount_block_flush(tc, w.counters);
}

// Diffuse-wall continuation of the queued tracks.
template <class R>
__global__ void __launch_bounds__(kBlock) k_wall(WorldView<R> w)
{
    TCount<unsigned long long> tc{-1, {0ull, 0ull, 0ull, 0ull}};
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
                                   w, t, (R)w.wq_tau[q], r, wall_sink(w));
            w.a.x[t] = x[0]; w.a.y[t] = x[1]; w.a.z[t] = x[2];
            w.a.cell[t] = cell;
            w.a.vx[t] = (R)U[0]; w.a.vy[t] = (R)U[1]; w.a.vz[t] = (R)U[2];
            w.a.erot[t] = (R)erot;
        }
        if (active) {
            tcount_add(tc, w.counters, ccase, 0, (cell == kDead) ? 1u : 0u);
            tcount_add(tc, w.counters, ccase, 1, (uint32_t)to.hits);
            tcount_add(tc, w.counters, ccase, 2, (uint32_t)to.errs);
        }
    }
    tcount_block_flush(tc, w.counters);
}

// ---------------------------------------------------------- cell weighting
// uniGasCloud::cellWeighting (uniGasCloud.C:1353-1425) after the move: a
// simulator whose cell weight went down from W_o to W_n (W_o > W_n) gets
// floor-like clones for W_o/W_n - 1 (one per whole unit, one more with the
// probability of the rest), one whose weight went up is removed with
// probability 1 - W_o/W_n.  W_o is the weight of the cell the simulator
// started the step in (a new inflow simulator: its insertion cell).  Each
// simulator draws from its own stream: existing ones keyed by their slot in
// the case, new ones by their index in the case's inflow of the step.  The
// clones are copies appended after the step's inflow in the order of their
// parents' slots.
template <class R>
__global__ void k_weight(WorldView<R> w)
{
    TCount<uint32_t, 1> tc{-1, {0u, 0u, 0u, 0u}};
    const uint32_t n0 = w.st->n;
    const uint32_t n_all = n0 + w.st->n_ins;
    // every entry up to cap is written (zero past the simulators): the scan
    // runs over a host-known length
    const bool one_case = w.ncases == 1;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t <= w.cap; t += gridDim.x * blockDim.x) {
        uint32_t k = 0u;
        bool del = false;
        int ccase = 0;
        if (t < n_all) {
            const uint32_t c = w.a.cell[t];
            if (c != kDead) {
                ccase = one_case ? 0 : w.cell_case[c];
                const CaseDev& cs = w.cases[ccase];
                if (cs.active && cs.weighted) {
                    const double wo = w.cw[w.cstart[t]], wn = w.cw[c];
                    if (wo != wn) {
                        const uint32_t id = (t < n0) ? t - cs.particle_begin
                                                     : (0x80000000u | (t - n0 - w.ins_off[cs.inflow_begin]));
                        Rng r;
                        rng_init(r, cs.seed_lo, cs.seed_hi, kRngWeight, 0u, id, cs.step);
                        if (wo > wn) {
                            double prob = wo / wn - 1.0;
                            while (prob > 1.0) { k += 1u; prob = prob - 1.0; }
                            if (prob > uniform(r)) k += 1u;
                        } else if (wo / wn < uniform(r)) {
                            del = true;
                            w.a.cell[t] = kDead;
                        }
                    }
                }
            }
            tcount_add(tc, w.counters, ccase, 0, k);
            tcount_add(tc, w.counters, ccase, 1, del ? 1u : 0u);
        }
        w.wclone[t] = k;
    }
    tcount_block_flush(tc, w.counters);
}

// The clones: copies of slot t at n_all + wcoff[t] + i.  Nothing is written
// when they would exceed the capacity (the overflow is reported).
template <class R>
__global__ void k_clone(WorldView<R> w)
{
    const uint32_t n_all = w.st->n + w.st->n_ins;
    const uint32_t total = w.st->nclone;
    if ((unsigned long long)n_all + total > (unsigned long long)w.cap) {
        if (blockIdx.x == 0 && threadIdx.x == 0) w.st->overflow = 1u;
        return;
    }
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n_all; t += gridDim.x * blockDim.x) {
        const uint32_t k = w.wclone[t];
        if (!k) continue;
        const uint32_t d0 = n_all + w.wcoff[t];
        const R x = w.a.x[t], y = w.a.y[t], z = w.a.z[t], vx = w.a.vx[t], vy = w.a.vy[t], vz = w.a.vz[t];
        const R er = w.a.erot[t];
        const uint32_t c = w.a.cell[t], g = w.a.sp[t];
        for (uint32_t i = 0; i < k; ++i) {
            const uint32_t j = d0 + i;
            w.a.x[j] = x; w.a.y[j] = y; w.a.z[j] = z;
            w.a.vx[j] = vx; w.a.vy[j] = vy; w.a.vz[j] = vz;
            w.a.erot[j] = er;
            w.a.cell[j] = c;
            w.a.sp[j] = g;
        }
    }
}

// The re-weighting after a cell weight adaptation (uniGasDynamicAdapter.C:
// 661-672, then uniGasCloud.C:1353-1425): every simulator of case ccase
// goes from its cell's previous weight cw_old to the current one, by the
// rule of k_weight, on its own stream (sub 1, keyed by its slot in the case
// and the step count after the step).  Run between steps on the sorted
// simulators; the clones follow them and a sort compacts.
template <class R>
__global__ void k_reweight_begin(WorldView<R> w)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) w.st->n_ins = 0u;
}

template <class R>
__global__ void k_reweight(WorldView<R> w, const double* cw_old, int ccase)
{
    TCount<uint32_t, 1> tc{-1, {0u, 0u, 0u, 0u}};
    const uint32_t n = w.st->n;
    const CaseDev& cs = w.cases[ccase];
    const uint32_t first = w.off[cs.cell_begin];   // the case's first slot in the sorted order
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t <= w.cap; t += gridDim.x * blockDim.x) {
        uint32_t k = 0u;
        bool del = false;
        if (t < n) {
            const uint32_t c = w.a.cell[t];
            if (c != kDead && w.cell_case[c] == ccase) {
                const double wo = cw_old[c], wn = w.cw[c];
                if (wo != wn) {
                    Rng r;
                    rng_init(r, cs.seed_lo, cs.seed_hi, kRngWeight, 1u, t - first, cs.step);
                    if (wo > wn) {
                        double prob = wo / wn - 1.0;
                        while (prob > 1.0) { k += 1u; prob = prob - 1.0; }
                        if (prob > uniform(r)) k += 1u;
                    } else if (wo / wn < uniform(r)) {
                        del = true;
                        w.a.cell[t] = kDead;
                    }
                }
            }
            tcount_add(tc, w.counters, ccase, 0, k);
            tcount_add(tc, w.counters, ccase, 1, del ? 1u : 0u);
        }
        w.wclone[t] = k;
    }
    tcount_block_flush(tc, w.counters);
}

template <class R>
__global__ void k_clone_finish(WorldView<R> w)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const uint32_t n_all = w.st->n + w.st->n_ins;
        if ((unsigned long long)n_all + w.st->nclone <= (unsigned long long)w.cap) w.st->n_ins += w.st->nclone;
    }
}

// ------------------------------------------------------ wall measurements
// The step's wall hits (events) go to per-face buckets; each face sorts its
// bucket by (species, slot, crossing) and sums it in that order, so the sums
// do not depend on the order in which the hits were recorded
// (uniGasPatchBoundary.C:130-302 measures in the order of the cloud's list).

template <class R>
__global__ void k_wall_count(WorldView<R> w)
{
    const uint32_t n = min(w.st->nev, w.ev_cap);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        atomicAdd(&w.wcnt[w.ev[i].wf], 1u);
}

template <class R>
__global__ void k_wall_bucket(WorldView<R> w)
{
    const uint32_t n = min(w.st->nev, w.ev_cap);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t wf = w.ev[i].wf;
        const uint32_t pos = w.woff[wf] + atomicAdd(&w.wcnt[wf], 1u);
        w.bkey[pos] = w.ev_key[i];
        w.bidx[pos] = i;
    }
}

// In-place heap sort of n keys (ascending), carrying their values.
__device__ inline void heap_sort(unsigned long long* k, uint32_t* v, uint32_t n)
{
    auto sift = [&](uint32_t root, uint32_t end) {
        for (;;) {
            uint32_t ch = 2u * root + 1u;
            if (ch >= end) break;
            if (ch + 1u < end && k[ch] < k[ch + 1u]) ch += 1u;
            if (!(k[root] < k[ch])) break;
            const unsigned long long tk = k[root]; k[root] = k[ch]; k[ch] = tk;
            const uint32_t tv = v[root]; v[root] = v[ch]; v[ch] = tv;
            root = ch;
        }
    };
    for (uint32_t s = n / 2u; s-- > 0u;) sift(s, n);
    for (uint32_t e = n; e-- > 1u;) {
        const unsigned long long tk = k[0]; k[0] = k[e]; k[e] = tk;
        const uint32_t tv = v[0]; v[0] = v[e]; v[e] = tv;
        sift(0u, e);
    }
}

// measurePropertiesBeforeControl / AfterControl terms of one velocity and
// E_rot (uniGasPatchBoundary.C:147-184, 223-254) added to a[0..9].
__device__ inline void wall_terms(const double U[3], double e, const double n[3], double fA, double m, int rotdof,
                                  double a[kWRows])
{
    const double Un = U[0] * n[0] + U[1] * n[1] + U[2] * n[2];
    const double inv = 1.0 / fmax(fabs(Un) * fA, kVSmall);
    const double UU = U[0] * U[0] + U[1] * U[1] + U[2] * U[2];
    a[UGX_W_RHON] += inv;
    if (rotdof > 0) a[UGX_W_RHON_INT] += inv;
    a[UGX_W_RHOM] += m * inv;
    a[UGX_W_LINEAR_KE] += 0.5 * m * UU * inv;
    a[UGX_W_MCC] += m * UU * inv;
    a[UGX_W_MU] += m * U[0] * inv;
    a[UGX_W_MV] += m * U[1] * inv;
    a[UGX_W_MW] += m * U[2] * inv;
    a[UGX_W_EROT] += e * inv;
    a[UGX_W_ROTDOF] += (double)rotdof * inv;
}

// One thread per wall face: sort its bucket, sum each species' hits in key
// order, and add deltaT times the sums at sample steps.
template <class R>
__global__ void __launch_bounds__(kBlock) k_wall_reduce(WorldView<R> w)
{
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= (uint32_t)w.nwf) return;
    const uint32_t b = w.woff[f], e = w.woff[f + 1];
    if (e == b) return;
    const CaseDev& cs = w.cases[w.cell_case[w.wcell[f]]];
    if (((cs.step + 1u) % (uint32_t)cs.sample_interval) != 0u) return;
    heap_sort(w.bkey + b, w.bidx + b, e - b);
    const double* gq = &w.wgeo[(size_t)f * kWGeo];
    const double n[3] = {gq[0], gq[1], gq[2]};
    const double fA = gq[3];
    // nParticle F_N W_c of the face's cell (uniGasPatchBoundary.C:292-294)
    const double FN = cs.fn * w.cw[w.wcell[f]], dt = cs.dt;
    uint32_t i = b;
    while (i < e) {
        const uint32_t ls = (uint32_t)(w.bkey[i] >> 42);
        const uint32_t g = (uint32_t)cs.sp_base + ls;
        const double m = w.sp.mass[g];
        const int rotdof = w.sp.rotdof[g];
        double a[kWRows];
        for (int r = 0; r < kWRows; ++r) a[r] = 0.0;
        for (; i < e && (uint32_t)(w.bkey[i] >> 42) == ls; ++i) {
            const WallEv ev = w.ev[w.bidx[i]];
            wall_terms(ev.u0, ev.e0, n, fA, m, rotdof, a);
            const double preIE = 0.5 * m * (ev.u0[0] * ev.u0[0] + ev.u0[1] * ev.u0[1] + ev.u0[2] * ev.u0[2]) + ev.e0;
            wall_terms(ev.u1, ev.e1, n, fA, m, rotdof, a);
            const double postIE = 0.5 * m * (ev.u1[0] * ev.u1[0] + ev.u1[1] * ev.u1[1] + ev.u1[2] * ev.u1[2]) + ev.e1;
            a[UGX_W_Q] += FN * (preIE - postIE) / (dt * fA);
            for (int d = 0; d < 3; ++d) a[UGX_W_FX + d] += FN * (m * ev.u0[d] - m * ev.u1[d]) / (dt * fA);
        }
        double* dst = &w.wsamp[((size_t)f * kNS + ls) * kWRows];
        for (int r = 0; r < kWRows; ++r) dst[r] += dt * a[r];
    }
}

// ------------------------------------------------------------ occupancy
// The sort after the move needs each cell's new population and, for every
// simulator, a place within its cell.  Most simulators stay in their cell or
// cross into a face neighbour, so they are counted per old cell in LDS, over
// the tiles of the previous sort (which partition the cells and the slots as
// they stand at the start of the step): class 0 stays, class 1 + f left
// through face f.  The new cell d gathers its arrivals from its distinct
// neighbours; a mover from c is placed in d's range of the first face of d
// that points back to c.  Simulators ending anywhere else, and the inflow,
// are counted by device atomics (class 7).  Each simulator's word in the
// per-slot scratch array is (class of the new cell << 28) | rank within that
// class.  The order within a cell is arbitrary: the cell tiles order each cell
// by slot.
constexpr uint32_t kRankBits = 28;
constexpr uint32_t kRankMask = (1u << kRankBits) - 1u;
constexpr uint32_t kClassOther = 7;
constexpr int kCountUnroll = 4;
constexpr uint32_t kMaxTileCells = 512u / kCellW + 2u;   // cells of the largest tile variant
constexpr int kCountTiles = 4;                           // tiles per block of k_count_tiles
// Largest grid of the occupancy count (blocks stride over the step's tile groups).
constexpr uint64_t kCountGrid = 8192;

// Class-7 counting of the slots [begin, n + n_ins): the inflow of a step, or
// every slot for the sort after an upload.
template <class R>
__global__ void k_count(WorldView<R> w, int all)
{
    constexpr int U = kCountUnroll;
    const uint32_t n_all = w.st->n + w.st->n_ins;
    const uint64_t begin = all ? 0u : w.st->n;
    const uint64_t span = (uint64_t)gridDim.x * blockDim.x * U;
    for (uint64_t base = begin + (uint64_t)blockIdx.x * blockDim.x * U; base < n_all; base += span) {
        uint32_t t[U], c[U], r[U];
        bool alive[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t tu = base + (uint64_t)u * blockDim.x + threadIdx.x;
            t[u] = (uint32_t)tu;
            c[u] = (tu < n_all) ? w.a.cell[tu] : kDead;
            alive[u] = c[u] != kDead;
        }
        wave_agg_inc<U>(w.ocnt, c, alive, r);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (!alive[u]) continue;
            if (r[u] > kRankMask) w.st->overflow = 1u;
            w.scratch[t[u]] = (kClassOther << kRankBits) | (r[u] & kRankMask);
        }
    }
}

// Classes of the simulators that were in the cells of TPB consecutive tiles of
// the previous sort (one block per TPB tiles), in windows of kCountWindow
// slots.  The rank within (old cell, class) follows the order of the slots:
// within each 64-slot chunk of the window, ballots per class give a
// simulator's rank among the same cell's and class's earlier slots of the
// chunk; a cell that spans chunks carries its counts from chunk to chunk, and
// from window to window.  Class 7 takes its rank from the device atomics.
// 512-slot windows (two entries per thread) measured faster than 256, 1024
// (1.18 against 1.21 and 1.28 ms on 1e8 simulators); 2048 spills.
constexpr uint32_t kCountWindow = 512;

template <class R, int Block, int TPB>
__global__ void __launch_bounds__(Block) __attribute__((amdgpu_waves_per_eu(7))) k_count_tiles(WorldView<R> w)
{
    constexpr int U = kCountWindow / Block;
    constexpr uint32_t kChunks = kCountWindow / 64;
    static_assert(U >= 1 && kCountWindow % Block == 0, "whole entries per thread");
    static_assert(kChunks * 7u <= (uint32_t)Block, "one thread per (chunk, class) carry");
    constexpr uint32_t kCells = TPB * kMaxTileCells;
    __shared__ uint32_t s_off[kCells + 1];
    __shared__ int32_t s_nb[kCells * kMaxFaces];
    __shared__ int8_t s_rev[kCells * kMaxFaces];
    __shared__ uint32_t s_cnt[kCells * 7];
    __shared__ uint32_t s_first[kChunks], s_last[kChunks], s_whole[kChunks];
    __shared__ uint32_t s_lastn[kChunks * 7], s_carry[kChunks * 7];
    const uint32_t ntiles = w.st->ntiles;
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    const unsigned long long below = (1ull << lane) - 1ull;
    // the grid is sized for the largest tile table; blocks stride over the
    // groups of this step's tiles
    for (uint32_t t0 = blockIdx.x * (uint32_t)TPB; t0 < ntiles; t0 += gridDim.x * (uint32_t)TPB) {
    const uint32_t t1 = min(t0 + (uint32_t)TPB, ntiles);
    const uint32_t c0 = w.tile_first[t0];
    const uint32_t nct = w.tile_first[t1] - c0;
    const uint32_t p0 = w.tile_p0[t0];
    const uint32_t size = w.tile_p0[t1] - p0;
    if (nct == 0u) continue;
    for (uint32_t k = tid; k <= nct; k += Block) s_off[k] = w.off[c0 + k] - p0;
    for (uint32_t k = tid; k < nct * kMaxFaces; k += Block) {
        s_nb[k] = w.geom.nbr[(size_t)c0 * kMaxFaces + k];
        s_rev[k] = w.rev[(size_t)c0 * kMaxFaces + k];
    }
    for (uint32_t k = tid; k < nct * 7u; k += Block) s_cnt[k] = 0u;
    __syncthreads();
    for (uint32_t w0 = 0; w0 < size; w0 += kCountWindow) {
        const uint32_t wn = min(kCountWindow, size - w0);
        uint32_t d[U], cls[U], lc[U], ro[U], rin[U];
        bool other[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t i = (uint32_t)u * Block + tid;
            d[u] = (i < wn) ? w.a.cell[p0 + w0 + i] : kDead;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t i = (uint32_t)u * Block + tid;
            cls[u] = kClassOther;
            lc[u] = 0u;
            if (i < wn) {
                // old cell of entry i: the last k with s_off[k] <= w0 + i
                uint32_t lo = 0, hi = nct - 1u;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi + 1u) >> 1;
                    if (s_off[mid] <= w0 + i) lo = mid; else hi = mid - 1u;
                }
                lc[u] = lo;
                if (d[u] == c0 + lo) {
                    cls[u] = 0u;
                } else if (d[u] != kDead) {
                    for (int f = 0; f < kMaxFaces; ++f)
                        if (s_nb[lo * kMaxFaces + f] == (int32_t)d[u] && s_rev[lo * kMaxFaces + f] >= 0) {
                            cls[u] = 1u + (uint32_t)f;
                            break;
                        }
                }
            }
            other[u] = d[u] != kDead && cls[u] == kClassOther;
        }
        wave_agg_inc<U>(w.ocnt, d, other, ro);
        // ranks within the chunk (entries u * Block + tid: chunk u * Block / 64 + wave)
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t i = (uint32_t)u * Block + tid;
            const uint32_t chunk = i >> 6, cs0 = w0 + (chunk << 6);
            const bool ranked = i < wn && d[u] != kDead && cls[u] < kClassOther;
            unsigned long long m[7];
#pragma unroll
            for (int c = 0; c < 7; ++c) m[c] = __ballot(ranked && cls[u] == (uint32_t)c);
            // lanes of this lane's cell in the chunk start at lane fl
            const uint32_t fl = (i < wn && s_off[lc[u]] > cs0) ? s_off[lc[u]] - cs0 : 0u;
            const unsigned long long seg = ~((1ull << fl) - 1ull);
            unsigned long long mine = 0ull;
#pragma unroll
            for (int c = 0; c < 7; ++c) mine = (cls[u] == (uint32_t)c) ? m[c] : mine;
            rin[u] = (uint32_t)__popcll(mine & below & seg);
            // the chunk's last cell: its counts per class, and whether it fills the chunk
            const unsigned long long valid = __ballot(i < wn);
            if (valid != 0ull) {
                const int ll = 63 - __clzll(valid);
                const uint32_t lcl = (uint32_t)__shfl((int)lc[u], ll);
                const uint32_t fll = (uint32_t)__shfl((int)fl, ll);
                const unsigned long long segl = ~((1ull << fll) - 1ull);
                if (lane == 0u) {
                    s_first[chunk] = lc[u];
                    s_last[chunk] = lcl;
                    s_whole[chunk] = (fll == 0u) ? 1u : 0u;
#pragma unroll
                    for (int c = 0; c < 7; ++c) s_lastn[chunk * 7u + c] = (uint32_t)__popcll(m[c] & segl);
                }
            } else if (lane == 0u) {
                s_first[chunk] = 0xFFFFFFFFu;
                s_last[chunk] = 0xFFFFFFFFu;
                s_whole[chunk] = 0u;
            }
        }
        __syncthreads();
        // carry into chunk k's first cell (class c): the counts of the same cell in
        // the chunks before, walking back through chunks that the cell fills
        if (tid < kChunks * 7u) {
            const uint32_t k = tid / 7u, c = tid % 7u;
            uint32_t in = 0u;
            for (uint32_t j = k; j > 0u; --j) {
                if (s_first[j] == 0xFFFFFFFFu || s_first[j] != s_last[j - 1u]) break;
                in += s_lastn[(j - 1u) * 7u + c];
                if (!s_whole[j - 1u]) break;
            }
            s_carry[tid] = in;
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t i = (uint32_t)u * Block + tid;
            if (i >= wn || d[u] == kDead) continue;
            uint32_t word;
            if (other[u]) {
                if (ro[u] > kRankMask) w.st->overflow = 1u;
                word = (kClassOther << kRankBits) | (ro[u] & kRankMask);
            } else {
                const uint32_t chunk = i >> 6;
                const uint32_t r = s_cnt[lc[u] * 7u + cls[u]] +
                                   ((lc[u] == s_first[chunk]) ? s_carry[chunk * 7u + cls[u]] : 0u) + rin[u];
                const uint32_t kc = (cls[u] == 0u) ? 0u : 1u + (uint32_t)s_rev[lc[u] * kMaxFaces + (cls[u] - 1u)];
                if (r > kRankMask) w.st->overflow = 1u;
                word = (kc << kRankBits) | (r & kRankMask);
            }
            w.scratch[p0 + w0 + i] = word;
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t i = (uint32_t)u * Block + tid;
            if (i < wn && d[u] != kDead && cls[u] < kClassOther) atomicAdd(&s_cnt[lc[u] * 7u + cls[u]], 1u);
        }
        __syncthreads();
    }
    for (uint32_t k = tid; k < nct * 7u; k += Block) w.out7[(size_t)c0 * 7u + k] = s_cnt[k];
    __syncthreads();
    }
}

// Per cell: population (stayers, arrivals through each distinct neighbour,
// class 7) and the start of each class within the cell.
template <class R>
__global__ void k_gather(WorldView<R> w)
{
    const uint32_t nc = (uint32_t)w.nc;
    for (uint32_t d = blockIdx.x * blockDim.x + threadIdx.x; d < nc; d += gridDim.x * blockDim.x) {
        const size_t fb = (size_t)d * kMaxFaces;
        int32_t nb[kMaxFaces];
#pragma unroll
        for (int g = 0; g < kMaxFaces; ++g) nb[g] = w.geom.nbr[fb + g];
        uint32_t in[kMaxFaces];
#pragma unroll
        for (int g = 0; g < kMaxFaces; ++g) {
            bool first = nb[g] >= 0 && nb[g] != (int32_t)d;
            for (int h = 0; h < g; ++h) first = first && nb[h] != nb[g];
            const int f = w.rev[fb + g];
            in[g] = (first && f >= 0) ? w.out7[(size_t)nb[g] * 7u + 1u + (uint32_t)f] : 0u;
        }
        // class k's source: the cell itself (k = 0) or the neighbour behind face
        // k - 1.  A class starts after the classes of smaller source cells, so the
        // simulators of classes 0-6 follow the order of their slots; class 7 ends
        // the cell
        uint32_t src[7], num[7];
        src[0] = d;
        num[0] = w.out7[(size_t)d * 7u];
#pragma unroll
        for (int g = 0; g < kMaxFaces; ++g) {
            src[1 + g] = (uint32_t)nb[g];
            num[1 + g] = in[g];
        }
        uint32_t* pre = w.pre + (size_t)d * 8u;
        uint32_t total = 0u;
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            uint32_t p = 0u;
#pragma unroll
            for (int q = 0; q < 7; ++q) p += (num[q] != 0u && src[q] < src[k]) ? num[q] : 0u;
            pre[k] = p;
            total += num[k];
        }
        pre[7] = total;
        w.cnt[d] = total + w.ocnt[d];
    }
}

// --------------------------------------------------------------- scan

__device__ inline uint32_t wave_incl_scan_u32(uint32_t v)
{
    const int lane = (int)__lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(v, o);
        if (lane >= o) v += u;
    }
    return v;
}

// Block sums of kScanBlock-entry blocks (256 threads, four entries each).
__global__ void k_scan_reduce(const uint32_t* in, uint32_t n, uint32_t* bsum)
{
    __shared__ uint32_t sh[kBlock / 64];
    const uint32_t b = blockIdx.x * kScanBlock + threadIdx.x * 4;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (b + k < n) s += in[b + k];
    s = wave_incl_scan_u32(s);
    if ((threadIdx.x & 63u) == 63u) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
#pragma unroll
        for (int q = 0; q < kBlock / 64; ++q) t += sh[q];
        bsum[blockIdx.x] = t;
    }
}

// Exclusive scan of the nb block sums in place, one block of 1024 threads:
// each thread sums a contiguous segment, the segment totals are scanned
// across the block, and each thread writes its segment's prefixes.
__global__ void k_scan_bsum(uint32_t* bsum, uint32_t nb)
{
    __shared__ uint32_t sh[16];
    const uint32_t tid = threadIdx.x, wv = tid >> 6;
    const uint32_t seg = (nb + 1023u) / 1024u;
    const uint32_t lo = min(nb, tid * seg), hi = min(nb, lo + seg);
    uint32_t s = 0;
    for (uint32_t i = lo; i < hi; ++i) s += bsum[i];
    const uint32_t inc = wave_incl_scan_u32(s);
    if ((tid & 63u) == 63u) sh[wv] = inc;
    __syncthreads();
    uint32_t run = inc - s;
    for (uint32_t q = 0; q < wv; ++q) run += sh[q];
    for (uint32_t i = lo; i < hi; ++i) {
        const uint32_t v = bsum[i];
        bsum[i] = run;
        run += v;
    }
}

// Exclusive scan of in[0..n) into out; out[n-1] is the total when in[n-1]==0.
__global__ void k_scan_apply(const uint32_t* in, uint32_t n, const uint32_t* bsum, uint32_t* out, uint32_t* total_out)
{
    __shared__ uint32_t sh[kBlock / 64];
    const uint32_t wv = threadIdx.x >> 6;
    const uint32_t b = blockIdx.x * kScanBlock + threadIdx.x * 4;
    uint32_t v[4], s = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) { v[k] = (b + k < n) ? in[b + k] : 0u; s += v[k]; }
    const uint32_t inc = wave_incl_scan_u32(s);
    if ((threadIdx.x & 63u) == 63u) sh[wv] = inc;
    __syncthreads();
    uint32_t run = bsum[blockIdx.x] + inc - s;
    for (uint32_t q = 0; q < wv; ++q) run += sh[q];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (b + k < n) {
            out[b + k] = run;
            if (b + k == n - 1 && total_out) *total_out = run;
        }
        run += v[k];
    }
}

// Exclusive scan of in[0..n) into out.  bsum holds scan_bsum_words(n) words:
// the nb block sums, and past them, when nb exceeds kScanDirect, the sums of
// the block sums (scanned in blocks themselves, as a single block scanning
// many sums is slow).
constexpr uint32_t kScanDirect = 4096;

static void scan_u32(const uint32_t* in, uint32_t n, uint32_t* out, uint32_t* bsum, uint32_t* total, hipStream_t s)
{
    const uint32_t nb = (n + kScanBlock - 1) / kScanBlock;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kBlock), 0, s, in, n, bsum);
    if (nb > kScanDirect) {
        uint32_t* b2 = bsum + nb + 1;
        const uint32_t nb2 = (nb + kScanBlock - 1) / kScanBlock;
        hipLaunchKernelGGL(k_scan_reduce, dim3(nb2), dim3(kBlock), 0, s, bsum, nb, b2);
        hipLaunchKernelGGL(k_scan_bsum, dim3(1), dim3(1024), 0, s, b2, nb2);
        hipLaunchKernelGGL(k_scan_apply, dim3(nb2), dim3(kBlock), 0, s, bsum, nb, b2, bsum, nullptr);
    } else {
        hipLaunchKernelGGL(k_scan_bsum, dim3(1), dim3(1024), 0, s, bsum, nb);
    }
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(kBlock), 0, s, in, n, bsum, out, total);
}

// ------------------------------------------------------------ scatter

// Slot list in cell order: perm[off[d] + start of the class + rank] = slot.
template <class R>
__global__ void k_scatter(WorldView<R> w)
{
    constexpr int U = kCountUnroll;
    const uint32_t n_all = w.st->n + w.st->n_ins;
    const uint64_t span = (uint64_t)gridDim.x * blockDim.x * U;
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x * U; base < n_all; base += span) {
        uint32_t t[U], c[U], r[U], o[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t tu = base + (uint64_t)u * blockDim.x + threadIdx.x;
            t[u] = (uint32_t)tu;
            c[u] = kDead;
            r[u] = 0u;
            if (tu < n_all) {
                c[u] = w.a.cell[tu];
                r[u] = w.scratch[tu];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            o[u] = (c[u] != kDead) ? w.off[c[u]] + w.pre[(size_t)c[u] * 8u + (r[u] >> kRankBits)] : 0u;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (c[u] == kDead) continue;
            const uint32_t p = o[u] + (r[u] & kRankMask);
            if (p < w.cap) w.perm[p] = t[u];
            else w.st->overflow = 1u;
        }
    }
}

// ------------------------------------------------------- cell accessors
// The per-cell physics reads simulators through an accessor addressed by
// sorted slot: LdsAcc for a tile staged in LDS, GlobAcc for the sorted
// buffers in HBM.

template <class R>
struct LdsAcc {
    static constexpr int kInFlight = 1;
    const R *vx, *vy, *vz, *er;
    const uint16_t* g;               // species id of each entry
    const double* spm;               // species mass, rotational DoF and case-local
    const uint8_t *spr, *spl;        // index, staged in LDS when the table fits
    bool lds;
    uint32_t p0;
    __device__ double mass(uint32_t i, const Species& sp) const
    {
        const uint32_t s = g[i - p0];
        return lds ? spm[s] : sp.mass[s];
    }
    __device__ double rotdof(uint32_t i, const Species& sp) const
    {
        const uint32_t s = g[i - p0];
        return lds ? (double)spr[s] : (double)sp.rotdof[s];
    }
    __device__ void load(uint32_t i, double& a, double& b, double& c) const
    {
        const uint32_t k = i - p0;
        a = (double)vx[k]; b = (double)vy[k]; c = (double)vz[k];
    }
    __device__ double erot(uint32_t i) const { return (double)er[i - p0]; }
    __device__ int local_species(uint32_t i, const Species& sp) const
    {
        const uint32_t s = g[i - p0];
        return lds ? (int)spl[s] : sp.local[s];
    }
};

template <class R>
struct GlobAcc {
    static constexpr int kInFlight = 2;   // cell sums keep this many records in flight per lane
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
    __device__ int local_species(uint32_t i, const Species&) const { return local[p.sp[i]]; }
    __device__ double mass(uint32_t i, const Species& sp) const { return sp.mass[p.sp[i]]; }
    __device__ double rotdof(uint32_t i, const Species& sp) const { return (double)sp.rotdof[p.sp[i]]; }
    __device__ uint32_t idx_get(uint32_t b, uint32_t k) const { return idx[b + k]; }
    __device__ void idx_set(uint32_t b, uint32_t k, uint32_t v) const { idx[b + k] = v; }
};

// ----------------------------------------------------- canonical cell sums
// Every per-cell sum (sampler rows, BGK moments, the conservation sums) is
// taken in one fixed order: eight partial sums, partial j adding the
// simulators j, j+8, j+16, ... of the cell in slot order to +0.0, then the
// tree ((P0+P1)+(P2+P3)) + ((P4+P5)+(P6+P7)).  On the GPU a group of eight
// lanes holds the partials and the tree is three DPP exchanges (xor 1, xor 2
// by quad permutes, xor 4 by the half-row mirror, which pairs the two
// uniform halves of an eight-lane row).

template <int Ctrl>
__device__ inline double dpp_f64(double v)
{
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), Ctrl, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), Ctrl, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// Called with all 64 lanes active; returns the group total in each lane.
__device__ inline double group8_total(double p)
{
    p = p + dpp_f64<0xB1>(p);    // quad_perm [1,0,3,2]
    p = p + dpp_f64<0x4E>(p);    // quad_perm [2,3,0,1]
    p = p + dpp_f64<0x141>(p);   // row_half_mirror
    return p;
}

// Partial sums of cellMeasurements::calculateFields rows
// (cellMeasurements.C:408-513) for the simulators [i0, e) step 8 of one
// cell.  L: 0 rows 0-7, 1 adds 8-13, 2 adds 14-19; species rows for
// q < nsw.  S must start at +0.0.
// Accumulators of the per-cell sums: the base rows of sum level L (rows
// 0-7, 0-13 or 0-19), then, for worlds with NS > 1 species slots (a case
// with up to NS species), the per-species counts and M|v|^2 of every slot.
template <int L, int NS>
struct SumRows {
    static constexpr int kBase = (L == 0) ? 8 : (L == 1) ? 14 : 20;
    static constexpr int kN = kBase + (NS > 1 ? 2 * NS : 0);
};

// One simulator's terms added to the partial sums S, in row order.
template <int L, int NS, int N>
__device__ inline void partial_add(double m, double vx, double vy, double vz, double er, double rd, int s,
                                   double (&S)[N])
{
    static_assert(N == SumRows<L, NS>::kN, "sum rows");
    constexpr int B = SumRows<L, NS>::kBase;
    const double mcc = m * (vx * vx + vy * vy + vz * vz);
    S[UGX_S_N] = S[UGX_S_N] + 1.0;
    S[UGX_S_M] = S[UGX_S_M] + m;
    S[UGX_S_MU] = S[UGX_S_MU] + m * vx;
    S[UGX_S_MV] = S[UGX_S_MV] + m * vy;
    S[UGX_S_MW] = S[UGX_S_MW] + m * vz;
    S[UGX_S_MCC] = S[UGX_S_MCC] + mcc;
    S[UGX_S_EROT] = S[UGX_S_EROT] + er;
    S[UGX_S_ROTDOF] = S[UGX_S_ROTDOF] + rd;
    if constexpr (L >= 1) {
        S[UGX_S_MUU] = S[UGX_S_MUU] + m * (vx * vx);
        S[UGX_S_MUV] = S[UGX_S_MUV] + m * (vx * vy);
        S[UGX_S_MUW] = S[UGX_S_MUW] + m * (vx * vz);
        S[UGX_S_MVV] = S[UGX_S_MVV] + m * (vy * vy);
        S[UGX_S_MVW] = S[UGX_S_MVW] + m * (vy * vz);
        S[UGX_S_MWW] = S[UGX_S_MWW] + m * (vz * vz);
    }
    if constexpr (L >= 2) {
        S[UGX_S_MCCU] = S[UGX_S_MCCU] + mcc * vx;
        S[UGX_S_MCCV] = S[UGX_S_MCCV] + mcc * vy;
        S[UGX_S_MCCW] = S[UGX_S_MCCW] + mcc * vz;
        S[UGX_S_EROTU] = S[UGX_S_EROTU] + er * vx;
        S[UGX_S_EROTV] = S[UGX_S_EROTV] + er * vy;
        S[UGX_S_EROTW] = S[UGX_S_EROTW] + er * vz;
    }
    if constexpr (NS > 1) {
#pragma unroll
        for (int q = 0; q < NS; ++q) {
            S[B + q] = S[B + q] + ((s == q) ? 1.0 : 0.0);
            S[B + NS + q] = S[B + NS + q] + ((s == q) ? mcc : 0.0);
        }
    }
}

template <int L, int NS, class R, class Acc>
__device__ inline void cell_partials(const WorldView<R>& w, const Acc& A, uint32_t i0, uint32_t e,
                                     double (&S)[SumRows<L, NS>::kN])
{
    auto rec = [&](uint32_t i, double& m, double& vx, double& vy, double& vz, double& er, double& rd, int& s) {
        m = A.mass(i, w.sp);
        A.load(i, vx, vy, vz);
        er = A.erot(i);
        rd = A.rotdof(i, w.sp);
        s = (NS > 1) ? A.local_species(i, w.sp) : 0;
    };
    uint32_t i = i0;
    if constexpr (Acc::kInFlight > 1) {
        // records in HBM: several simulators' reads in flight, added in slot order
        constexpr int F = Acc::kInFlight;
        for (; i + 8u * (F - 1) < e; i += 8u * F) {
            double m[F], x[F], y[F], z[F], er[F], rd[F];
            int sp[F];
#pragma unroll
            for (int u = 0; u < F; ++u) rec(i + 8u * u, m[u], x[u], y[u], z[u], er[u], rd[u], sp[u]);
#pragma unroll
            for (int u = 0; u < F; ++u) partial_add<L, NS>(m[u], x[u], y[u], z[u], er[u], rd[u], sp[u], S);
        }
    }
    for (; i < e; i += 8u) {
        double m, vx, vy, vz, er, rd;
        int s;
        rec(i, m, vx, vy, vz, er, rd, s);
        partial_add<L, NS>(m, vx, vy, vz, er, rd, s, S);
    }
}

// Fields of a case that the per-cell sums need.
struct TileCase {
    double dt;
    uint32_t step;
    int32_t sample_interval, coll_model, bgk_model, active, ns, sampler_full, adaptive;
};

__device__ inline TileCase tile_case_of(const CaseDev& c)
{
    return TileCase{c.dt, c.step, c.sample_interval, c.coll_model, c.bgk_model, c.active, c.ns, c.sampler_full,
                    c.adaptive};
}

// Sums of the cells of one tile, eight cells per wave (group g of lanes
// 8g..8g+7 takes cell k), then the sampler (uniGasVolFields.C:761-797), the
// hybrid time sums and the BGK moments.  All 64 lanes of every wave call it
// with the same arguments.  Single-species worlds take the species rows
// from rows N and M|v|^2 (the same sums).
// Sampler rows of one cell that lane j of the cell's 8-lane group updates
// (slot i of five): base rows j, j + 8, j + 16 below the sum level's row
// count (rows 8-19 only for the full sampler), then the species rows of
// species j (several species) or, single-species, the species-N row (lane 0)
// and the species-MCC row (lane 1).  -1: no row.
template <int L, int NS>
__device__ inline void samp_lane_rows(uint32_t j, int ns, bool full, int (&rows)[5])
{
    constexpr bool MS = NS > 1;
    constexpr int B = SumRows<L, NS>::kBase;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int q = (int)j + 8 * i;
        rows[i] = (q < B && (full || q < UGX_S_MUU)) ? q : -1;
    }
    if constexpr (MS) {
        rows[3] = ((int)j < ns) ? UGX_S_SPECIES_N + (int)j : -1;
        rows[4] = ((int)j < ns) ? UGX_S_SPECIES_MCC + (int)j : -1;
    } else {
        rows[3] = (j == 0u) ? UGX_S_SPECIES_N : -1;
        rows[4] = (j == 1u) ? UGX_S_SPECIES_MCC : -1;
    }
}

// The totals of slots i of samp_lane_rows for lane j.
template <int L, int NS, int NR>
__device__ inline void samp_lane_values(uint32_t j, const double (&S)[NR], double (&v)[5])
{
    constexpr bool MS = NS > 1;
    constexpr int B = SumRows<L, NS>::kBase;
#pragma unroll
    for (int i = 0; i < 5; ++i) v[i] = 0.0;
#pragma unroll
    for (int q = 0; q < B; ++q)
        if ((uint32_t)(q & 7) == j) v[q >> 3] = S[q];
    if constexpr (MS) {
#pragma unroll
        for (int q = 0; q < NS; ++q)
            if ((uint32_t)q == j) { v[3] = S[B + q]; v[4] = S[B + NS + q]; }
    } else {
        v[3] = S[UGX_S_N];
        v[4] = S[UGX_S_MCC];
    }
}

// Sampler rows of the first cell pass of a tile, loaded ahead of the sums.
struct SampPre {
    double v[5];
    bool valid;
};

__device__ inline bool samples_now(const TileCase& cs)
{
    return cs.active && ((cs.step + 1u) % (uint32_t)cs.sample_interval) == 0u;
}

template <int L, int NS, int Block, class R, class Acc>
__device__ inline void tile_sums(const WorldView<R>& w, uint32_t c0, uint32_t nct, const uint32_t* s_off,
                                 const Acc& A, const int32_t* s_cc, const TileCase* s_tc, int case0,
                                 const SampPre& pre)
{
    constexpr bool MS = NS > 1;
    constexpr int B = SumRows<L, NS>::kBase, NR = SumRows<L, NS>::kN;
    const uint32_t tid = threadIdx.x;
    const uint32_t wv = tid >> 6, lane = tid & 63u, g = lane >> 3, j = lane & 7u;
    for (uint32_t kb = 0; kb < nct; kb += (Block / 64) * 8) {
        const uint32_t k = kb + wv * 8u + g;
        uint32_t b = 0, e = 0;
        bool sample = false, hybrid = false, bgk = false, adapt = false;
        double dt = 0.0;
        const uint32_t c = c0 + k;
        TileCase cs{};
        if (k < nct) {
            const int ci = s_cc[k];
            cs = (ci == case0) ? *s_tc : tile_case_of(w.cases[ci]);
            if (cs.active) {
                sample = ((cs.step + 1u) % (uint32_t)cs.sample_interval) == 0u;
                hybrid = cs.coll_model == UGX_COLL_HYBRID;
                bgk = (cs.coll_model == UGX_COLL_BGK || hybrid) && cs.bgk_model != UGX_BGK_NONE;
                adapt = cs.adaptive != 0;
                if (sample || hybrid || bgk || adapt) { b = s_off[k]; e = s_off[k + 1]; }
                dt = cs.dt;
            }
        }
        const bool need = sample || hybrid || bgk || adapt;
        if (__ballot(need) == 0ull) continue;
        double S[NR];
#pragma unroll
        for (int q = 0; q < NR; ++q) S[q] = 0.0;
        cell_partials<L, NS, R, Acc>(w, A, b + j, e, S);
#pragma unroll
        for (int q = 0; q < NR; ++q) S[q] = group8_total(S[q]);
        if (need && sample) {
            // every lane of the group holds the totals; lane j updates its rows
            int rows[5];
            double v[5];
            samp_lane_rows<L, NS>(j, cs.ns, cs.sampler_full != 0, rows);
            samp_lane_values<L, NS>(j, S, v);
            double* sc = &w.samp[(size_t)c * kNSamp];
            const bool have = pre.valid && kb == 0u;
#pragma unroll
            for (int i = 0; i < 5; ++i)
                if (rows[i] >= 0) sc[rows[i]] = (have ? pre.v[i] : sc[rows[i]]) + dt * v[i];
        }
        // lane j of the group writes rows j, j + 8, j + 16 and species j
        double v[3] = {0.0, 0.0, 0.0}, vsp = 0.0;
#pragma unroll
        for (int q = 0; q < B; ++q)
            if ((uint32_t)(q & 7) == j) v[q >> 3] = S[q];
        if constexpr (MS) {
#pragma unroll
            for (int q = 0; q < NS; ++q)
                if ((uint32_t)q == j) vsp = S[B + q];
        } else {
            vsp = (j == 0u) ? S[UGX_S_N] : 0.0;
        }
        if (need && hybrid) {
            double* ka = &w.knacc[(size_t)kKnRows * c];
            ka[j] += dt * v[0];
            if (MS || j == 0u) ka[8 + j] += dt * vsp;
        }
        if (need && adapt) {
            // uniGasDynamicAdapter::adapt (uniGasDynamicAdapter.C:526-547): the
            // cell measurements of every step, weighted by the time step
            double* aa = &w.adacc[(size_t)kKnRows * c];
            aa[j] += dt * v[0];
            if (MS || j == 0u) aa[8 + j] += dt * vsp;
        }
        if (need && bgk) {
            double* mo = &w.mom[(size_t)c * kNMom];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const uint32_t q = j + 8u * (uint32_t)i;
                if (q < (uint32_t)UGX_S_BASE) mo[q] = v[i];
            }
            mo[UGX_S_SPECIES_N + j] = vsp;
        }
    }
}

// ------------------------------------------------------------ tile table

// tile_first[t] = first cell whose key off[c] + kCellW * c is >= t * TileN;
// tile_first[ntiles] = nc.  Each cell writes the tiles that start at it.
template <class R, uint32_t TileN>
__global__ void k_tile_table(WorldView<R> w)
{
    const uint32_t nc = (uint32_t)w.nc;
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c <= nc; c += gridDim.x * blockDim.x) {
        const uint64_t key = (uint64_t)w.off[c] + (uint64_t)kCellW * c;
        const uint64_t lo = (c == 0u) ? 0ull : ((uint64_t)w.off[c - 1] + (uint64_t)kCellW * (c - 1u)) / TileN + 1ull;
        uint64_t hi = key / TileN;
        if (c == nc) {
            hi += 1ull;
            w.st->ntiles = (uint32_t)hi;
        }
        const uint32_t p = w.off[c];
        const int32_t ci = (c < nc) ? w.cell_case[c] : 0;
        for (uint64_t t = lo; t <= hi; ++t) {
            w.tile_first[t] = c;
            w.tile_p0[t] = p;
            w.tile_case[t] = ci;
        }
    }
}

// ------------------------------------------------------------ cell tiles

// One block per tile (k_tile_table).  Steps:
//  1. each cell's slot list (written by k_scatter in atomic order) is put in
//     ascending previous-slot order: an entry's rank in its cell is the
//     number of entries of that cell with a smaller slot;
//  2. records are gathered into the sorted buffers B (velocity, E_rot and
//     local species also into LDS);
//  3. per-cell sums from LDS: sampler, hybrid accumulators, BGK moments.
// SortOnly stops after 2.  Tiles larger than TileMax run the same steps
// from HBM.
template <class R, bool SortOnly, int L, int NS, int Block, uint32_t TileN, uint32_t TileMax>
__global__ void __launch_bounds__(Block) __attribute__((amdgpu_waves_per_eu((NS == 1 || (NS == 2 && L == 0)) ? 5 : 1))) k_tile(WorldView<R> w)
{
    constexpr uint32_t kM = SortOnly ? 1u : TileMax;
    __shared__ uint32_t s_off[TileN / kCellW + 2];
    __shared__ uint32_t s_in[TileMax];
    __shared__ uint32_t s_perm[TileMax];
    __shared__ uint8_t s_k[TileMax];
    // FP32 x-velocities take the slot list's place once the ranks are known
    constexpr bool kVxInSlots = sizeof(R) == sizeof(uint32_t);
    __shared__ R s_vx_own[(SortOnly || kVxInSlots) ? 1 : kM];
    __shared__ R s_vy[kM], s_vz[kM], s_er[kM];
    __shared__ uint16_t s_g[kM];
    __shared__ int32_t s_cc[SortOnly ? 1 : TileN / kCellW + 2];
    __shared__ uint32_t s_oc[TileN / kCellW + 2];
    __shared__ TileCase s_tc;
    __shared__ double s_spm[SortOnly ? 1 : kLdsSpecies];
    __shared__ uint8_t s_spr[SortOnly ? 1 : kLdsSpecies], s_spl[SortOnly ? 1 : kLdsSpecies];
    R* const s_vx = kVxInSlots ? reinterpret_cast<R*>(s_in) : s_vx_own;
    const uint32_t t = blockIdx.x;
    if (t >= w.st->ntiles) return;
    const uint32_t tid = threadIdx.x;
    // the tile table gives cells, slots and the first case in one round trip ...
    const uint32_t c0 = w.tile_first[t];
    const uint32_t nct = w.tile_first[t + 1] - c0;
    const uint32_t p0 = w.tile_p0[t];
    const uint32_t size = w.tile_p0[t + 1] - p0;
    const int case0 = w.tile_case[t];
    const bool lds_species = w.nsp <= (int)kLdsSpecies;
    // ... and everything else a tile reads before its gather comes in one more
    for (uint32_t k = tid; k <= nct; k += Block) {
        s_off[k] = w.off[c0 + k];
        if (k < nct) s_oc[k] = w.ocnt[c0 + k];
        if (!SortOnly && k < nct) s_cc[k] = w.cell_case[c0 + k];
    }
    if (!SortOnly) {
        if (tid == 0) s_tc = tile_case_of(w.cases[case0]);
        if (lds_species)
            for (int g = (int)tid; g < w.nsp; g += Block) {
                s_spm[g] = w.sp.mass[g];
                s_spr[g] = (uint8_t)w.sp.rotdof[g];
                s_spl[g] = (uint8_t)w.sp.local[g];
            }
    }
    if (size <= TileMax) {
        for (uint32_t i = tid; i < size; i += Block) s_in[i] = w.perm[p0 + i];
        __syncthreads();
        // sampler rows of the first cell pass, loaded while the tile is sorted
        // (not at sum level 2, where the ten registers they hold through the
        // sort cost more than the round trip they save)
        SampPre spre{{0.0, 0.0, 0.0, 0.0, 0.0}, false};
        if constexpr (!SortOnly && L < 2) {
            const uint32_t k = (tid >> 6) * 8u + ((tid & 63u) >> 3);
            if (k < nct) {
                const int ci = s_cc[k];
                const TileCase cs = (ci == case0) ? s_tc : tile_case_of(w.cases[ci]);
                if (samples_now(cs)) {
                    int rows[5];
                    samp_lane_rows<L, NS>(tid & 7u, cs.ns, cs.sampler_full != 0, rows);
                    const double* sc = &w.samp[(size_t)(c0 + k) * kNSamp];
#pragma unroll
                    for (int i = 0; i < 5; ++i) spre.v[i] = (rows[i] >= 0) ? sc[rows[i]] : 0.0;
                    spre.valid = true;
                }
            }
        }
        for (uint32_t i = tid; i < size; i += Block) {
            const uint32_t slot = s_in[i];
            // cell of entry i: the last k with s_off[k] <= p0 + i
            uint32_t lo = 0, hi = nct - 1u;
            while (lo < hi) {
                const uint32_t mid = (lo + hi + 1u) >> 1;
                if (s_off[mid] <= p0 + i) lo = mid; else hi = mid - 1u;
            }
            // the cell's simulators up to eo are in slot order already; those from
            // eo on (multi-hop, inflow) are merged in by slot
            const uint32_t b = s_off[lo] - p0, e = s_off[lo + 1] - p0;
            const uint32_t eo = e - s_oc[lo];
            uint32_t r;
            if (i < eo) {
                r = i - b;
            } else {
                uint32_t l2 = b, h2 = eo;
                while (l2 < h2) {
                    const uint32_t mid = (l2 + h2) >> 1;
                    if (s_in[mid] < slot) l2 = mid + 1u; else h2 = mid;
                }
                r = l2 - b;
            }
            for (uint32_t q = eo; q < e; ++q) r += (s_in[q] < slot) ? 1u : 0u;
            s_perm[b + r] = slot;
            s_k[b + r] = (uint8_t)lo;
        }
        __syncthreads();
        for (uint32_t i = tid; i < size; i += Block) {
            const uint32_t slot = s_perm[i];
            const uint32_t j = p0 + i;
            const R vx = w.a.vx[slot], vy = w.a.vy[slot], vz = w.a.vz[slot], er = w.a.erot[slot];
            const uint32_t g = w.a.sp[slot];
            w.b.x[j] = w.a.x[slot]; w.b.y[j] = w.a.y[slot]; w.b.z[j] = w.a.z[slot];
            w.b.vx[j] = vx; w.b.vy[j] = vy; w.b.vz[j] = vz;
            w.b.erot[j] = er;
            w.b.cell[j] = c0 + s_k[i];
            w.b.sp[j] = g;
            if (!SortOnly) {
                s_vx[i] = vx; s_vy[i] = vy; s_vz[i] = vz; s_er[i] = er;
                s_g[i] = (uint16_t)g;
            }
        }
        if (SortOnly) return;
        __syncthreads();
        LdsAcc<R> A{s_vx, s_vy, s_vz, s_er, s_g, s_spm, s_spr, s_spl, lds_species, p0};
        tile_sums<L, NS, Block, R>(w, c0, nct, s_off, A, s_cc, &s_tc, case0, spre);
    } else {
        __syncthreads();
        // Tile larger than LDS: the same steps in HBM.  As in LDS, a cell's
        // slot list is in slot order up to eo (the entries laid out by the
        // occupancy count) and the few entries from eo on (multi-hop, inflow)
        // are merged in by slot; the ordered slots go to scratch.
        if (tid == 0) atomicAdd(&w.st->big_tiles, 1u);
        for (uint32_t k = 0; k < nct; ++k) {
            const uint32_t b = s_off[k], e = s_off[k + 1];
            const uint32_t eo = e - s_oc[k];
            for (uint32_t i = b + tid; i < e; i += Block) {
                const uint32_t slot = w.perm[i];
                uint32_t r;
                if (i < eo) {
                    r = i - b;
                } else {
                    uint32_t l2 = b, h2 = eo;
                    while (l2 < h2) {
                        const uint32_t mid = (l2 + h2) >> 1;
                        if (w.perm[mid] < slot) l2 = mid + 1u; else h2 = mid;
                    }
                    r = l2 - b;
                }
                for (uint32_t q = eo; q < e; ++q) r += (w.perm[q] < slot) ? 1u : 0u;
                w.scratch[b + r] = slot;
            }
        }
        __syncthreads();
        for (uint32_t i = tid; i < size; i += Block) {
            const uint32_t slot = w.scratch[p0 + i];
            const uint32_t j = p0 + i;
            w.b.x[j] = w.a.x[slot]; w.b.y[j] = w.a.y[slot]; w.b.z[j] = w.a.z[slot];
            w.b.vx[j] = w.a.vx[slot]; w.b.vy[j] = w.a.vy[slot]; w.b.vz[j] = w.a.vz[slot];
            w.b.erot[j] = w.a.erot[slot];
            w.b.cell[j] = w.a.cell[slot];
            w.b.sp[j] = w.a.sp[slot];
        }
        if (SortOnly) return;
        __threadfence_block();
        __syncthreads();
        GlobAcc<R> A{w.b, w.sp.local, w.perm};
        tile_sums<L, NS, Block, R>(w, c0, nct, s_off, A, s_cc, &s_tc, case0, SampPre{{0.0, 0.0, 0.0, 0.0, 0.0}, false});
    }
}

// ------------------------------------------------------------- collisions

// Subcell of a simulator along one axis (noTimeCounter.C:141-155): the
// cell's bounding box in lv equal slabs, the index truncated and kept in
// [0, lv - 1] (a simulator on the upper face would index one past the last).
__device__ inline uint32_t subcell_axis(double x, double lo, double len, uint32_t lv)
{
    const double t = (double)lv * (x - lo) / len;
    int k = (int)t;
    k = k < 0 ? 0 : k;
    return (uint32_t)min(k, (int)lv - 1);
}

// Subcell lists of cell c (noTimeCounter.C:133-161), sequential in the
// cell's lane: scid[b + i] the subcell of its i-th simulator (axes weighted
// 1, lx, lx ly), tab[s]..tab[s+1] the range of sclist[b + .] holding subcell
// s's cell-local indices in ascending order; tab[nsub + 1 + s] are cursors.
template <class R>
__device__ inline void subcell_lists(const WorldView<R>& w, uint32_t c, uint32_t b, uint32_t nC, const uint32_t lv[3],
                                     uint32_t nsub, uint32_t* tab)
{
    double lo[3], len[3];
    for (int d = 0; d < 3; ++d) { lo[d] = w.bblo[3 * (size_t)c + d]; len[d] = w.bbhi[3 * (size_t)c + d] - lo[d]; }
    for (uint32_t s = 0; s <= nsub; ++s) tab[s] = 0u;
    for (uint32_t i = 0; i < nC; ++i) {
        const uint32_t j = b + i;
        const uint32_t k0 = subcell_axis((double)w.b.x[j], lo[0], len[0], lv[0]);
        const uint32_t k1 = subcell_axis((double)w.b.y[j], lo[1], len[1], lv[1]);
        const uint32_t k2 = subcell_axis((double)w.b.z[j], lo[2], len[2], lv[2]);
        const uint32_t s = k0 + lv[0] * (k1 + lv[1] * k2);
        w.scid[j] = s;
        tab[s + 1u] += 1u;
    }
    for (uint32_t s = 1; s <= nsub; ++s) tab[s] += tab[s - 1u];
    uint32_t* cur = tab + nsub + 1u;
    for (uint32_t s = 0; s < nsub; ++s) cur[s] = tab[s];
    for (uint32_t i = 0; i < nC; ++i) w.sclist[b + cur[w.scid[b + i]]++] = i;
}

// noTimeCounter::collide for one cell (noTimeCounter.C:83-312).  With more
// than one subcell the second candidate comes from the first's subcell when
// that holds two or more simulators, else from the whole cell (:193-225).
template <class R, class Acc>
__device__ inline void ntc_cell(const WorldView<R>& w, const CaseDev& cs, uint32_t c, uint32_t b, uint32_t nC,
                                const Acc& A, unsigned long long& cand, unsigned long long& coll)
{
    uint32_t lv[3] = {1u, 1u, 1u}, nsub = 1u;
    uint32_t* tab = nullptr;
    if (w.has_subcells) {
        const uint32_t p = w.sclev[c];
        lv[0] = p & 0xffu; lv[1] = (p >> 8) & 0xffu; lv[2] = (p >> 16) & 0xffu;
        nsub = lv[0] * lv[1] * lv[2];
        if (nsub > 1u) {
            tab = w.sctab + w.scoff[c];
            subcell_lists<R>(w, c, b, nC, lv, nsub, tab);
        }
    }
    Rng r;
    rng_init(r, cs.seed_lo, cs.seed_hi, kRngNtc, 0u, c - (uint32_t)cs.cell_begin, cs.step);
    const double smax0 = w.smax[c];
    double smax = smax0;
    const bool lb = cs.dsmc_model == UGX_DSMC_LB_VHS || cs.dsmc_model == UGX_DSMC_LB_VSS;
    // a single-species case has one species and one pair: their constants are
    // read once, not per candidate
    const bool one = cs.ns == 1;
    double oC = 0.0, oE = 0.0, oMass = 0.0, oMr = 0.0, oChi = 0.0, oAlpha = 0.0;
    int oRd = 0;
    if (one) {
        oC = w.pr.C[cs.pair_base]; oE = w.pr.e[cs.pair_base];
        oMr = w.pr.mr[cs.pair_base]; oChi = w.pr.chi[cs.pair_base]; oAlpha = w.pr.alpha[cs.pair_base];
        oMass = w.sp.mass[cs.sp_base]; oRd = w.sp.rotdof[cs.sp_base];
    }
    const double sel = 0.5 * (double)nC * (double)(nC - 1) * cs.fn * w.cw[c] * smax0 * cs.dt / w.vol[c];
    const double fl = floor(sel);
    long long ncand = (long long)fl;
    if (uniform(r) < sel - fl) ncand += 1;
    cand += (unsigned long long)ncand;
    for (long long k = 0; k < ncand; ++k) {
        uint32_t P = uindex(r, nC);
        uint32_t Q, nSC = 0u, s0 = 0u;
        if (nsub > 1u) {
            const uint32_t s = w.scid[b + P];
            s0 = tab[s];
            nSC = tab[s + 1u] - s0;
        }
        if (nSC > 1u) {
            do { Q = w.sclist[b + s0 + uindex(r, nSC)]; } while (Q == P);
        } else {
            do { Q = uindex(r, nC); } while (Q == P);
        }
        const uint32_t ip = b + P, iq = b + Q;
        // a single-species case needs no species lookup
        const int lp = one ? 0 : A.local_species(ip), lq = one ? 0 : A.local_species(iq);
        const int gp = cs.sp_base + lp, gq = cs.sp_base + lq;
        const int pk = cs.pair_base + lp * cs.ns + lq;
        double UP[3], UQ[3];
        A.load(ip, UP[0], UP[1], UP[2]);
        A.load(iq, UQ[0], UQ[1], UQ[2]);
        // E_rot takes part only in Larsen-Borgnakke exchanges; it is read with
        // the velocities, and the other models leave it untouched
        double eP = 0.0, eQ = 0.0;
        if (lb) { eP = A.erot(ip); eQ = A.erot(iq); }
        double d0 = UP[0] - UQ[0], d1 = UP[1] - UQ[1], d2 = UP[2] - UQ[2];
        double cR2 = d0 * d0 + d1 * d1 + d2 * d2;
        const double pC = one ? oC : w.pr.C[pk], pE = one ? oE : w.pr.e[pk];
        double scr = (cR2 < kVSmall) ? 0.0 : pC * dexp(pE * dlog(cR2));
        if (scr > smax) smax = scr;
        if (scr / smax0 > uniform(r)) {
            int rdP = oRd, rdQ = oRd;
            double mP = oMass, mQ = oMass, mr = oMr, chi = oChi, alpha = oAlpha;
            if (!one) {
                rdP = w.sp.rotdof[gp]; rdQ = w.sp.rotdof[gq];
                mP = w.sp.mass[gp]; mQ = w.sp.mass[gq];
                mr = w.pr.mr[pk]; chi = w.pr.chi[pk]; alpha = w.pr.alpha[pk];
            }
            binary_collide(r, cs.dsmc_model, cs.inv_zrot, rdP, rdQ, mP, mQ, mr, chi, alpha, UP, UQ, eP, eQ);
            A.store(ip, UP[0], UP[1], UP[2]);
            A.store(iq, UQ[0], UQ[1], UQ[2]);
            if (lb) {
                A.set_erot(ip, eP);
                A.set_erot(iq, eQ);
            }
            coll += 1;
        }
    }
    w.smax[c] = smax;
}

// BGK-family relaxation (stochasticParticleBGK.C:287-796,
// stochasticParticleESBGK.C:311-906, unifiedStochasticParticleSBGK.C:379-1096)
// in two kernels:
//  k_collide   (lane per cell)  moments, Pr, nu, USP q/sigma time
//              relaxation, the number of relaxed simulators and their
//              selection (partial Fisher-Yates on the cell stream) into a
//              selection list; parameters into a per-cell record;
//  k_bgk_conserve (eight lanes per cell) new velocities of the
//              selections of the wave's eight cells, spread over the 64
//              lanes, each from the selection's own stream (kRngBgkV,
//              sub = selection index), USP draws testing against the cell's maxProb at the
//              start of the step and recording the largest raise; then
//              conserveMomentumAndEnergy and the USP maxProb update.
constexpr uint32_t kRngBgkV = 7;

// Parameters and relaxed count of one cell; S holds the cell's sums.
template <class R>
__device__ inline void bgk_params(const WorldView<R>& w, const CaseDev& cs, uint32_t c, uint32_t nC,
                                  const double* S, BgkP& P, uint32_t& nrel_out)
{
    const int model = cs.bgk_model;
    const double FN = cs.fn * w.cw[c], dt = cs.dt, V = w.vol[c];
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
    if (w.has_macro && cs.macro) {
        // the cell's fields for the cellPoint interpolation, every cell of the case
        double* mf = &w.mfield[(size_t)kMiRows * c];
        mf[0] = T;
        for (int d = 0; d < 3; ++d) mf[1 + d] = U[d];
        mf[4] = Pr;
        mf[5] = nu;
        mf[6] = p;
        for (int d = 0; d < 3; ++d) mf[7 + d] = usp ? q[d] : 0.0;
        for (int k = 0; k < 6; ++k) mf[10 + k] = usp ? sh[k] : Pt[k];
    }
    P.model = model;
    P.m0 = w.uspmax[c];
    P.perform = 0;
    nrel_out = 0u;
    if (perform && w.cmodel[c] == 0) {
        P.perform = 1;
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
        // relaxed count: stochastic round of N (1 - exp(-nu dt)), first draw of the cell stream
        Rng r;
        rng_init(r, cs.seed_lo, cs.seed_hi, kRngBgk, 0u, c - (uint32_t)cs.cell_begin, cs.step);
        const double pc = N * (1.0 - dexp(-nu * dt));
        const double fl = floor(pc);
        long long nrel = (long long)fl;
        if (uniform(r) < pc - fl) nrel += 1;
        if (nrel > (long long)nC) nrel = nC;
        nrel_out = (uint32_t)nrel;
    }
}

// Partial Fisher-Yates selection of nrel of the cell's nC simulators on the
// cell stream (after the relaxed-count draw); entries go to the selection
// list at base.  Up to kSelReg selections run on a register list of the
// displaced entries: after step i only positions above i are read again,
// and each step displaces at most one entry (to position j > i), so the
// list holds the (position, entry) pairs that differ from the identity and
// a lookup falls back to the position itself.  Larger selections in cells
// of up to kSelLds simulators swap in a lane-private index array in LDS
// (perm, entry e of the lane at perm[e * kBlock]); larger cells swap in the
// cell's index array in HBM.
constexpr int kSelReg = 16;
constexpr uint32_t kSelLds = 96;

template <class R>
__device__ inline void bgk_select(const WorldView<R>& w, const CaseDev& cs, uint32_t c, uint32_t b, uint32_t nC,
                                  uint32_t nrel, uint32_t base, const GlobAcc<R>& A, uint16_t* perm)
{
    Rng r;
    rng_init(r, cs.seed_lo, cs.seed_hi, kRngBgk, 0u, c - (uint32_t)cs.cell_begin, cs.step);
    (void)uniform(r);
    if (nrel <= (uint32_t)kSelReg) {
        uint32_t pos[kSelReg], val[kSelReg];
        uint32_t nd = 0u;
        for (uint32_t i = 0; i < nrel; ++i) {
            const uint32_t j = i + uindex(r, nC - i);
            uint32_t ti = i, tj = j;
            bool have_j = false;
#pragma unroll
            for (int q = 0; q < kSelReg; ++q) {
                if ((uint32_t)q < nd) {
                    if (pos[q] == i) ti = val[q];
                    if (pos[q] == j) { tj = val[q]; have_j = true; }
                }
            }
            if (j != i) {
                // position j now holds the entry of position i
#pragma unroll
                for (int q = 0; q < kSelReg; ++q) {
                    if ((uint32_t)q < nd && pos[q] == j) val[q] = ti;
                    if (!have_j && (uint32_t)q == nd) { pos[q] = j; val[q] = ti; }
                }
                if (!have_j) ++nd;
            }
            w.sel_slot[base + i] = b + tj;
        }
        return;
    }
    if (nC <= kSelLds) {
        for (uint32_t i = 0; i < nC; ++i) perm[i * kBlock] = (uint16_t)i;
        for (uint32_t i = 0; i < nrel; ++i) {
            const uint32_t j = i + uindex(r, nC - i);
            const uint32_t ti = perm[i * kBlock], tj = perm[j * kBlock];
            perm[i * kBlock] = (uint16_t)tj;
            perm[j * kBlock] = (uint16_t)ti;
            w.sel_slot[base + i] = b + tj;
        }
        return;
    }
    for (uint32_t i = 0; i < nC; ++i) A.idx_set(b, i, i);
    for (uint32_t i = 0; i < nrel; ++i) {
        const uint32_t j = i + uindex(r, nC - i);
        const uint32_t ti = A.idx_get(b, i), tj = A.idx_get(b, j);
        A.idx_set(b, i, tj);
        A.idx_set(b, j, ti);
        w.sel_slot[base + i] = b + tj;
    }
}

// Exclusive wave prefix of v and one atomic reservation per wave on *ctr;
// called by all 64 lanes.  Returns the lane's base.
__device__ inline uint32_t wave_reserve(uint32_t* ctr, uint32_t v)
{
    const uint32_t lane = __lane_id();
    uint32_t incl = v;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(incl, o);
        if ((int)lane >= o) incl += u;
    }
    const uint32_t total = __shfl(incl, 63);
    uint32_t base = 0;
    if (lane == 63u && total) base = atomicAdd(ctr, total);
    base = __shfl(base, 63);
    return base + incl - v;
}

// Lane per cell: BGK-family parameters and selections, NTC collisions.
template <class R>
__global__ void __launch_bounds__(kBlock) k_collide(WorldView<R> w)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    const bool on = c < (uint32_t)w.nc;
    unsigned long long cand = 0ull, coll = 0ull;
    uint32_t nrel = 0u, b = 0u, nC = 0u;
    int ccase = 0;
    bool bgk = false, dsmc = false;
    // the cell's own rows are loaded along with its case, not after it
    double S[kNMom];
    if (on) {
        const uint32_t cb = w.off[c], ce = w.off[c + 1];
        const int cm = w.cmodel[c];
        if (w.has_bgk) {
#pragma unroll
            for (int q = 0; q < kNMom; ++q) S[q] = w.mom[(size_t)c * kNMom + q];
        }
        ccase = w.cell_case[c];
        const CaseDev& cs = w.cases[ccase];
        if (cs.active) {
            const bool hybrid = cs.coll_model == UGX_COLL_HYBRID;
            bgk = (cs.coll_model == UGX_COLL_BGK || hybrid) && cs.bgk_model != UGX_BGK_NONE;
            dsmc = (cs.coll_model == UGX_COLL_DSMC || hybrid) && cs.dsmc_model != UGX_DSMC_NONE && cm == 1;
            b = cb;
            nC = ce - cb;
        }
    }
    GlobAcc<R> A{w.b, w.sp.local, w.perm};
    if (bgk) {
        const CaseDev& cs = w.cases[ccase];
        BgkP P;
        bgk_params<R>(w, cs, c, nC, S, P, nrel);
        w.bgkp[c] = P;
        w.raise[c] = 0ull;
    }
    const uint32_t base = wave_reserve(&w.st->nsel, nrel);
    if (bgk) {
        w.bgkp[c].sel_base = base;
        w.bgkp[c].nrel = nrel;
    }
    __shared__ uint16_t s_perm[kSelLds * kBlock];
    if (nrel) bgk_select<R>(w, w.cases[ccase], c, b, nC, nrel, base, A, &s_perm[threadIdx.x]);
    if (dsmc && nC >= 2u) ntc_cell<R>(w, w.cases[ccase], c, b, nC, A, cand, coll);
    case_add(w.counters, ccase, kCntCandidates, cand);
    case_add(w.counters, ccase, kCntCollisions, coll);
    case_add(w.counters, ccase, kCntBgk, (unsigned long long)nrel);
}

// New velocity of one selected simulator (slot i of cell c, selection k)
// from the parameters P.
template <class R>
__device__ inline void bgk_draw_core(const WorldView<R>& w, const CaseDev& cs, uint32_t c, uint32_t i, uint32_t k,
                                     const BgkP& P, unsigned long long* raise, const GlobAcc<R>& A)
{
    Rng rv;
    rng_init(rv, cs.seed_lo, cs.seed_hi, kRngBgkV, k, c - (uint32_t)cs.cell_begin, cs.step);
    // a single-species case needs no species lookup
    const int ls = (cs.ns == 1) ? 0 : A.local_species(i);
    const double u0 = P.s2kT / sqrt(w.sp.mass[cs.sp_base + ls]);
    if constexpr (sizeof(R) == sizeof(float)) {
        // FP32 state: the draw in single precision (the stored velocity is FP32)
        float xf[3];
        if (P.model == UGX_BGK_USP) {
            const float m0 = (float)P.m0;
            const float Af = (float)(P.cQ / (P.p * u0)), Bf = (float)(P.cS / P.p), inv_m0 = 1.0f / m0;
            const float q0 = (float)P.q[0], q1 = (float)P.q[1], q2 = (float)P.q[2];
            const float s0 = (float)P.sm[0], s1 = (float)P.sm[1], s2 = (float)P.sm[2], s3 = (float)P.sm[3],
                        s4 = (float)P.sm[4], s5 = (float)P.sm[5];
            for (int it = 0; it < 100000; ++it) {
                float g0, g1, g2;
                gauss3f(rv, g0, g1, g2);
                xf[0] = g0 * (float)kSqrtHalf; xf[1] = g1 * (float)kSqrtHalf; xf[2] = g2 * (float)kSqrtHalf;
                const float vSq = xf[0] * xf[0] + xf[1] * xf[1] + xf[2] * xf[2];
                const float vTr = vSq * (float)kThird;
                const float prob = 1.0f + Af * (q0 * xf[0] + q1 * xf[1] + q2 * xf[2]) * (vSq * 0.4f - 1.0f)
                                   + Bf * (s0 * (xf[0] * xf[0] - vTr) + s3 * (xf[1] * xf[1] - vTr)
                                           + s5 * (xf[2] * xf[2] - vTr) + 2.0f * s1 * xf[0] * xf[1]
                                           + 2.0f * s2 * xf[0] * xf[2] + 2.0f * s4 * xf[1] * xf[2]);
                if (prob > m0 && prob < 10.0f) {
                    atomicMax(raise, (unsigned long long)double_to_bits((double)prob));
                    break;
                }
                if (uniformf(rv) < prob * inv_m0) break;
            }
        } else {
            float g0, g1, g2;
            gauss3f(rv, g0, g1, g2);
            xf[0] = g0 * (float)kSqrtHalf; xf[1] = g1 * (float)kSqrtHalf; xf[2] = g2 * (float)kSqrtHalf;
            if (P.model == UGX_BGK_ESBGK) {
                float y[3];
                for (int d = 0; d < 3; ++d)
                    y[d] = (float)P.sm[3 * d] * xf[0] + (float)P.sm[3 * d + 1] * xf[1] + (float)P.sm[3 * d + 2] * xf[2];
                xf[0] = y[0]; xf[1] = y[1]; xf[2] = y[2];
            }
        }
        A.store(i, P.U[0] + u0 * (double)xf[0], P.U[1] + u0 * (double)xf[1], P.U[2] + u0 * (double)xf[2]);
        return;
    }
    double xi[3];
    if (P.model == UGX_BGK_USP) {
        // the draw's constants are divided out once (usp_draw of the oracle)
        const double m0 = P.m0;
        const double A = P.cQ / (P.p * u0), B = P.cS / P.p, inv_m0 = 1.0 / m0;
        for (int it = 0; it < 100000; ++it) {
            double g0, g1, g2;
            gauss3(rv, g0, g1, g2);
            xi[0] = g0 * kSqrtHalf; xi[1] = g1 * kSqrtHalf; xi[2] = g2 * kSqrtHalf;
            const double vSq = xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2];
            const double vTr = vSq * kThird;
            const double* sh = P.sm;
            const double prob = 1.0 + A * (P.q[0] * xi[0] + P.q[1] * xi[1] + P.q[2] * xi[2]) * (vSq * 0.4 - 1.0)
                                + B * (sh[0] * (xi[0] * xi[0] - vTr) + sh[3] * (xi[1] * xi[1] - vTr)
                                       + sh[5] * (xi[2] * xi[2] - vTr) + 2.0 * sh[1] * xi[0] * xi[1]
                                       + 2.0 * sh[2] * xi[0] * xi[2] + 2.0 * sh[4] * xi[1] * xi[2]);
            if (prob > m0 && prob < 10.0) {
                atomicMax(raise, (unsigned long long)double_to_bits(prob));
                break;
            }
            if (uniform(rv) < prob * inv_m0) break;
        }
    } else {
        double g0, g1, g2;
        gauss3(rv, g0, g1, g2);
        xi[0] = g0 * kSqrtHalf; xi[1] = g1 * kSqrtHalf; xi[2] = g2 * kSqrtHalf;
        if (P.model == UGX_BGK_ESBGK) {
            double y[3];
            for (int d = 0; d < 3; ++d) y[d] = P.sm[3 * d] * xi[0] + P.sm[3 * d + 1] * xi[1] + P.sm[3 * d + 2] * xi[2];
            xi[0] = y[0]; xi[1] = y[1]; xi[2] = y[2];
        }
    }
    A.store(i, P.U[0] + u0 * xi[0], P.U[1] + u0 * xi[1], P.U[2] + u0 * xi[2]);
}

// interpolationCellPoint at position x (cell frame) of cell c: the tet of
// the cell's decomposition (centroid, then per face the fan (p0, p_i,
// p_i+1), faces in ascending order) that holds x, or the one x is least
// outside of, and its barycentric weights; rows 0..nr-1 of the cell's and
// the tet points' fields combined with them.
constexpr double kMiTol = 1.0e-12;

template <class R>
__device__ inline void mi_interp(const WorldView<R>& w, uint32_t c, const double x[3], int nr, double* out)
{
    const double* ref = &w.cc[3 * (size_t)c];
    double a[3], r[3];
    for (int d = 0; d < 3; ++d) { a[d] = w.mi_cent[3 * (size_t)c + d]; r[d] = x[d] - a[d]; }
    double best = -1.0e300, bl[4] = {1.0, 0.0, 0.0, 0.0};
    int32_t bid[3] = {-1, -1, -1};
    bool found = false;
    for (int j = 0; j < kMaxFaces && !found; ++j) {
        const int32_t* fp = &w.mi_fp[((size_t)c * kMaxFaces + j) * kMiFaceVerts];
        if (fp[0] < 0) break;
        int nv = 3;
        while (nv < kMiFaceVerts && fp[nv] >= 0) ++nv;
        double e1[3];
        for (int d = 0; d < 3; ++d) e1[d] = (w.mi_pts[3 * (size_t)fp[0] + d] - ref[d]) - a[d];
        for (int t = 1; t + 1 < nv; ++t) {
            double e2[3], e3[3];
            for (int d = 0; d < 3; ++d) {
                e2[d] = (w.mi_pts[3 * (size_t)fp[t] + d] - ref[d]) - a[d];
                e3[d] = (w.mi_pts[3 * (size_t)fp[t + 1] + d] - ref[d]) - a[d];
            }
            const double n23[3] = {e2[1] * e3[2] - e2[2] * e3[1], e2[2] * e3[0] - e2[0] * e3[2], e2[0] * e3[1] - e2[1] * e3[0]};
            const double det = e1[0] * n23[0] + e1[1] * n23[1] + e1[2] * n23[2];
            if (det == 0.0) continue;
            const double rx3[3] = {r[1] * e3[2] - r[2] * e3[1], r[2] * e3[0] - r[0] * e3[2], r[0] * e3[1] - r[1] * e3[0]};
            const double e2r[3] = {e2[1] * r[2] - e2[2] * r[1], e2[2] * r[0] - e2[0] * r[2], e2[0] * r[1] - e2[1] * r[0]};
            const double l1 = (r[0] * n23[0] + r[1] * n23[1] + r[2] * n23[2]) / det;
            const double l2 = (e1[0] * rx3[0] + e1[1] * rx3[1] + e1[2] * rx3[2]) / det;
            const double l3 = (e1[0] * e2r[0] + e1[1] * e2r[1] + e1[2] * e2r[2]) / det;
            const double l0 = ((1.0 - l1) - l2) - l3;
            const double mn = fmin(fmin(l0, l1), fmin(l2, l3));
            if (mn > best) {
                best = mn;
                bl[0] = l0; bl[1] = l1; bl[2] = l2; bl[3] = l3;
                bid[0] = fp[0]; bid[1] = fp[t]; bid[2] = fp[t + 1];
            }
            if (mn >= -kMiTol) { found = true; break; }
        }
    }
    const double* fc = &w.mfield[(size_t)kMiRows * c];
    for (int q = 0; q < nr; ++q) {
        double v = bl[0] * fc[q];
        for (int k = 0; k < 3; ++k)
            if (bid[k] >= 0) v = v + bl[k + 1] * w.pfield[(size_t)kMiRows * bid[k] + q];
        out[q] = v;
    }
}

// The draw's parameters at the simulator's position (stochasticParticleBGK.C:
// 680-687, stochasticParticleESBGK.C:779-789, unifiedStochasticParticleSBGK.C:
// 933-946 with the samplers at :887-906 and :1047-1096): T and U for the
// BGK model; with Pr, p and the pressure tensor the ES-BGK matrix
// I - (1 - Pr)/(2 Pr) (P/p - I); with Pr, the collision frequency, p, q and
// the stress the USP coefficients cQ, cS.  maxProb stays the cell's.
template <class R>
__device__ inline void macro_params(const WorldView<R>& w, const CaseDev& cs, uint32_t c, uint32_t i, BgkP& Q)
{
    const double x[3] = {(double)w.b.x[i], (double)w.b.y[i], (double)w.b.z[i]};
    double f[kMiRows];
    mi_interp<R>(w, c, x, Q.model == UGX_BGK_BGK ? 4 : kMiRows, f);
    const double T = f[0];
    for (int d = 0; d < 3; ++d) Q.U[d] = f[1 + d];
    Q.T = T;
    Q.s2kT = sqrt(2.0 * kB * T);
    if (Q.model == UGX_BGK_ESBGK) {
        const double Pr = f[4], p = f[6];
        const double a = 0.5 * (1.0 - Pr) / Pr;
        const int full[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const double dl = (k == 0 || k == 4 || k == 8) ? 1.0 : 0.0;
            Q.sm[k] = dl - a * (f[10 + full[k]] / p - dl);
        }
        Q.p = p;
    } else if (Q.model == UGX_BGK_USP) {
        const double Pr = f[4], nu = f[5];
        Q.p = f[6];
        for (int d = 0; d < 3; ++d) Q.q[d] = f[7 + d];
        for (int k = 0; k < 6; ++k) Q.sm[k] = f[10 + k];
        const double tau = 0.5 * nu * cs.dt;
        const double ex = 1.0 + 2.0 / (dexp(2.0 * tau) - 1.0);
        Q.cQ = 2.0 * (1.0 - Pr * tau * ex);
        Q.cS = 1.0 - tau * ex;
    }
}

// Macro: the world has macroInterpolation cases (a separate instance keeps
// the interpolation's registers out of the plain draws).
template <class R, bool Macro>
__device__ inline void bgk_draw(const WorldView<R>& w, const CaseDev& cs, uint32_t c, uint32_t i, uint32_t k,
                                const BgkP& P, unsigned long long* raise, const GlobAcc<R>& A)
{
    if (Macro && cs.macro) {
        BgkP Q = P;
        macro_params<R>(w, cs, c, i, Q);
        bgk_draw_core<R>(w, cs, c, i, k, Q, raise, A);
    } else {
        bgk_draw_core<R>(w, cs, c, i, k, P, raise, A);
    }
}

// One selected simulator's draw held by a lane across its rejection trials
// (the same stream and arithmetic as bgk_draw_core, one trial at a time).
// T is the draw's precision: single for the FP32 state.  e0..e8 hold the USP
// stress (xx xy xz yy yz zz) or the ES-BGK matrix (row major).  Scalar
// members, not arrays, so that the state stays in registers.
template <class R>
struct DrawLane {
    using T = typename std::conditional<sizeof(R) == sizeof(float), float, double>::type;
    Rng rv;
    T m0, inv_m0, a, bq, q0, q1, q2, e0, e1, e2, e3, e4, e5, e6, e7, e8;
    double u0, U0, U1, U2;
    uint32_t i, cc;
    int model, it;
};

// The draw's parameter f of the cell (BgkP's doubles in order: U 0-2, s2kT 3,
// p 4, cQ 5, cS 6, m0 7, T 8, q 9-11, sm 12-20) comes from get(f).
constexpr int kDrawPars = 21;

__device__ __attribute__((always_inline)) inline double bgkp_field(const BgkP& P, int f)
{
    switch (f) {
    case 0: return P.U[0];
    case 1: return P.U[1];
    case 2: return P.U[2];
    case 3: return P.s2kT;
    case 4: return P.p;
    case 5: return P.cQ;
    case 6: return P.cS;
    case 7: return P.m0;
    case 8: return P.T;
    case 9: return P.q[0];
    case 10: return P.q[1];
    case 11: return P.q[2];
    default: return P.sm[f - 12];
    }
}

// pre, when not null, holds the cell's u0 and the USP draw's a, bq and
// inv_m0 as computed below once per cell (single-species cells).
template <class R, class Get>
__device__ __attribute__((always_inline)) inline void draw_setup(const WorldView<R>& w, const CaseDev& cs, uint32_t c,
                                                                 uint32_t i, uint32_t k, int model, Get get,
                                                                 const GlobAcc<R>& A, DrawLane<R>& D,
                                                                 const double* pre = nullptr)
{
    using T = typename DrawLane<R>::T;
    rng_init(D.rv, cs.seed_lo, cs.seed_hi, kRngBgkV, k, c - (uint32_t)cs.cell_begin, cs.step);
    if (pre) {
        D.u0 = pre[0];
    } else {
        const int ls = (cs.ns == 1) ? 0 : A.local_species(i);
        D.u0 = get(3) / sqrt(w.sp.mass[cs.sp_base + ls]);
    }
    D.model = model;
    D.i = i;
    D.cc = c;
    D.it = 0;
    D.U0 = get(0); D.U1 = get(1); D.U2 = get(2);
    if (model == UGX_BGK_USP) {
        D.m0 = (T)get(7);
        if (pre) {
            D.a = (T)pre[1];
            D.bq = (T)pre[2];
            D.inv_m0 = (T)pre[3];
        } else {
            D.a = (T)(get(5) / (get(4) * D.u0));
            D.bq = (T)(get(6) / get(4));
            D.inv_m0 = (T)1 / D.m0;
        }
        D.q0 = (T)get(9); D.q1 = (T)get(10); D.q2 = (T)get(11);
    }
    if (model == UGX_BGK_USP || model == UGX_BGK_ESBGK) {
        D.e0 = (T)get(12); D.e1 = (T)get(13); D.e2 = (T)get(14); D.e3 = (T)get(15); D.e4 = (T)get(16);
        D.e5 = (T)get(17);
    }
    if (model == UGX_BGK_ESBGK) {
        D.e6 = (T)get(18); D.e7 = (T)get(19); D.e8 = (T)get(20);
    }
}

// One trial: true when (x0, x1, x2) holds the drawn thermal velocity
// (accepted, the only draw of the BGK and ES-BGK models, or the last
// allowed USP trial).
template <class R>
__device__ __attribute__((always_inline)) inline bool draw_trial(DrawLane<R>& D, unsigned long long* raise,
                                                                 typename DrawLane<R>::T& x0,
                                                                 typename DrawLane<R>::T& x1,
                                                                 typename DrawLane<R>::T& x2)
{
    using T = typename DrawLane<R>::T;
    T g0, g1, g2;
    if constexpr (sizeof(T) == sizeof(float)) gauss3f(D.rv, g0, g1, g2);
    else gauss3(D.rv, g0, g1, g2);
    x0 = g0 * (T)kSqrtHalf; x1 = g1 * (T)kSqrtHalf; x2 = g2 * (T)kSqrtHalf;
    if (D.model == UGX_BGK_USP) {
        const T vSq = x0 * x0 + x1 * x1 + x2 * x2;
        const T vTr = vSq * (T)kThird;
        const T prob = (T)1 + D.a * (D.q0 * x0 + D.q1 * x1 + D.q2 * x2) * (vSq * (T)0.4 - (T)1)
                       + D.bq * (D.e0 * (x0 * x0 - vTr) + D.e3 * (x1 * x1 - vTr)
                                 + D.e5 * (x2 * x2 - vTr) + (T)2 * D.e1 * x0 * x1
                                 + (T)2 * D.e2 * x0 * x2 + (T)2 * D.e4 * x1 * x2);
        if (prob > D.m0 && prob < (T)10) {
            atomicMax(raise, (unsigned long long)double_to_bits((double)prob));
            return true;
        }
        T u;
        if constexpr (sizeof(T) == sizeof(float)) u = uniformf(D.rv);
        else u = uniform(D.rv);
        if (u < prob * D.inv_m0) return true;
        D.it += 1;
        return D.it >= 100000;
    }
    if (D.model == UGX_BGK_ESBGK) {
        const T y0 = D.e0 * x0 + D.e1 * x1 + D.e2 * x2;
        const T y1 = D.e3 * x0 + D.e4 * x1 + D.e5 * x2;
        const T y2 = D.e6 * x0 + D.e7 * x1 + D.e8 * x2;
        x0 = y0; x1 = y1; x2 = y2;
    }
    return true;
}

// The point values of the step's fields (volPointInterpolation): per point
// the weighted sum of its cells' rows, in the order of its entries.
template <class R>
__global__ void k_point_fields(WorldView<R> w)
{
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < w.npts; p += gridDim.x * blockDim.x) {
        double v[kMiRows];
#pragma unroll
        for (int q = 0; q < kMiRows; ++q) v[q] = 0.0;
        for (uint32_t k = w.mi_poff[p]; k < w.mi_poff[p + 1]; ++k) {
            const double wt = w.mi_pw[k];
            const double* f = &w.mfield[(size_t)kMiRows * w.mi_pcell[k]];
#pragma unroll
            for (int q = 0; q < kMiRows; ++q) v[q] = v[q] + wt * f[q];
        }
#pragma unroll
        for (int q = 0; q < kMiRows; ++q) w.pfield[(size_t)kMiRows * p + q] = v[q];
    }
}

// Selections of a conservation wave staged in LDS at a time.
constexpr uint32_t kDrawWindow = 512;

// conserveMomentumAndEnergy (stochasticParticleBGK.C:731-780) with the
// canonical sums, and the USP maxProb update; eight lanes per cell.
constexpr int kConsRegs = 8;  // simulators per lane kept in registers between the sums and the rescale
template <class R, bool Macro, int kConsReg>
__device__ __attribute__((always_inline)) inline void bgk_conserve_body(const WorldView<R>& w)
{
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u, g = lane >> 3, j = lane & 7u;
    const uint32_t c = (blockIdx.x * blockDim.x + tid) >> 3;
    bool perform = false, usp = false;
    uint32_t b = 0, e = 0, sb = 0, nrel = 0;
    int ccase = 0;
    double T = 0.0, U0 = 0.0, U1 = 0.0, U2 = 0.0, M = 0.0, N = 0.0;
    int one_sp = -1;   // the global species of a single-species case, else -1
    (void)g;
    if (c < (uint32_t)w.nc) {
        // the cell's own rows are loaded along with its case, not after it
        const uint32_t cb = w.off[c], ce = w.off[c + 1];
        const double cM = w.mom[(size_t)c * kNMom + UGX_S_M], cN = w.mom[(size_t)c * kNMom + UGX_S_N];
        ccase = w.cell_case[c];
        const CaseDev& cs = w.cases[ccase];
        if (cs.ns == 1) one_sp = cs.sp_base;
        const bool hybrid = cs.coll_model == UGX_COLL_HYBRID;
        const bool bgk = cs.active && (cs.coll_model == UGX_COLL_BGK || hybrid) && cs.bgk_model != UGX_BGK_NONE;
        if (bgk) {
            const BgkP& P = w.bgkp[c];
            perform = P.perform != 0;
            usp = P.model == UGX_BGK_USP;
            sb = P.sel_base;
            nrel = P.nrel;
            if (perform) {
                b = cb;
                e = ce;
                T = P.T; U0 = P.U[0]; U1 = P.U[1]; U2 = P.U[2];
                M = cM;
                N = cN;
            }
        }
    }
    // the selections of the wave's eight cells get their new velocities
    // first, spread over all 64 lanes: draw t of the wave is selection
    // t - pf[g] of group g (each from its own stream)
    {
        uint32_t pf[9], sbg[8];
        int cg[8];
        pf[0] = 0u;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            pf[q + 1] = pf[q] + (uint32_t)__builtin_amdgcn_readlane((int)nrel, 8 * q);
            sbg[q] = (uint32_t)__builtin_amdgcn_readlane((int)sb, 8 * q);
            cg[q] = __builtin_amdgcn_readlane(ccase, 8 * q);
        }
        const uint32_t cw = (blockIdx.x * blockDim.x + (tid & ~63u)) >> 3;
        GlobAcc<R> A{w.b, w.sp.local, w.perm};
        const uint32_t total = pf[8];
        if (total <= 64u) {
            // one draw per lane: set up from memory and run the trials in place
            if (lane < total) {
                const uint32_t t = lane;
                uint32_t k = 0u, slot_at = 0u, gg = 0u;
                int ct = 0;
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    if (t >= pf[q] && t < pf[q + 1]) { k = t - pf[q]; slot_at = sbg[q] + k; ct = cg[q]; gg = (uint32_t)q; }
                const uint32_t cc = cw + gg;
                const CaseDev& cs = w.cases[ct];
                const uint32_t i = w.sel_slot[slot_at];
                DrawLane<R> D;
                if (Macro && cs.macro) {
                    BgkP Q = w.bgkp[cc];
                    macro_params<R>(w, cs, cc, i, Q);
                    draw_setup<R>(w, cs, cc, i, k, Q.model, [&](int f) { return bgkp_field(Q, f); }, A, D);
                } else {
                    const BgkP& P = w.bgkp[cc];
                    draw_setup<R>(w, cs, cc, i, k, P.model, [&](int f) { return bgkp_field(P, f); }, A, D);
                }
                typename DrawLane<R>::T x0, x1, x2;
                while (!draw_trial<R>(D, &w.raise[cc], x0, x1, x2)) {}
                A.store(i, D.U0 + D.u0 * (double)x0, D.U1 + D.u0 * (double)x1, D.U2 + D.u0 * (double)x2);
            }
        } else {
        // the eight cells' draw parameters and a window of the wave's
        // selections are staged in LDS, so that a draw's set-up makes no
        // memory round trip
        // (and, for single-species cells, the draw constants that depend
        // only on the cell: entries kDrawPars + 1.. as draw_setup's pre,
        // entry kDrawPars + 1 negative when the draws compute them)
        __shared__ double s_par[kBlock / 8][kDrawPars + 5];
        __shared__ uint32_t s_sel[kBlock / 64][kDrawWindow];
        const uint32_t wv = tid >> 6, cl0 = wv * 8u;
        if (c < (uint32_t)w.nc && nrel > 0u) {
            const double* src = reinterpret_cast<const double*>(&w.bgkp[c]);
            for (uint32_t f = j; f < (uint32_t)kDrawPars; f += 8u) s_par[tid >> 3][f] = src[f];
            if (j == 0u) {
                using T = typename DrawLane<R>::T;
                const BgkP& P = w.bgkp[c];
                const CaseDev& cs = w.cases[ccase];
                double* pre = &s_par[tid >> 3][kDrawPars + 1];
                s_par[tid >> 3][kDrawPars] = (double)P.model;
                pre[0] = -1.0;
                if (cs.ns == 1 && !(Macro && cs.macro)) {
                    const double u0 = P.s2kT / sqrt(w.sp.mass[cs.sp_base]);
                    pre[0] = u0;
                    if (P.model == UGX_BGK_USP) {
                        pre[1] = (double)(T)(P.cQ / (P.p * u0));
                        pre[2] = (double)(T)(P.cS / P.p);
                        pre[3] = (double)((T)1 / (T)P.m0);
                    }
                }
            }
        }
        uint32_t w0 = 0u;
        auto stage_sel = [&](uint32_t from) {
#pragma unroll
            for (uint32_t m = 0; m < kDrawWindow / 64u; ++m) {
                const uint32_t t = from + lane + 64u * m;
                if (t < pf[8]) {
                    uint32_t slot_at = 0u;
#pragma unroll
                    for (int q = 0; q < 8; ++q)
                        if (t >= pf[q] && t < pf[q + 1]) slot_at = sbg[q] + (t - pf[q]);
                    s_sel[wv][t - from] = w.sel_slot[slot_at];
                }
            }
        };
        stage_sel(0u);
        wave_sync();
        // a lane whose draw is accepted takes the next draw of the wave at
        // once, so the USP rejection trials of a wave do not wait for its
        // slowest lane; each draw keeps its own stream, whichever lane runs it
        const unsigned long long below = (1ull << lane) - 1ull;
        uint32_t tnext = 0u;
        bool has = false;
        DrawLane<R> D;
        for (;;) {
            const unsigned long long need = __ballot(!has);
            if (need != 0ull && tnext < total) {
                if (tnext + 64u > w0 + kDrawWindow) {
                    // the next draws leave the staged window: stage the one from tnext
                    wave_sync();
                    w0 = tnext;
                    stage_sel(w0);
                    wave_sync();
                }
                const uint32_t t = tnext + (uint32_t)__popcll(need & below);
                if (!has && t < total) {
                    uint32_t k = 0u, gg = 0u;
                    int ct = 0;
#pragma unroll
                    for (int q = 0; q < 8; ++q)
                        if (t >= pf[q] && t < pf[q + 1]) { k = t - pf[q]; ct = cg[q]; gg = (uint32_t)q; }
                    const uint32_t cc = cw + gg;
                    const CaseDev& cs = w.cases[ct];
                    const uint32_t i = s_sel[wv][t - w0];
                    const double* sp = s_par[cl0 + gg];
                    const int model = (int)sp[kDrawPars];
                    if (Macro && cs.macro) {
                        BgkP Q = w.bgkp[cc];
                        macro_params<R>(w, cs, cc, i, Q);
                        draw_setup<R>(w, cs, cc, i, k, Q.model, [&](int f) { return bgkp_field(Q, f); }, A, D);
                    } else {
                        const double* pre = (sp[kDrawPars + 1] >= 0.0) ? &sp[kDrawPars + 1] : nullptr;
                        draw_setup<R>(w, cs, cc, i, k, model, [&](int f) { return sp[f]; }, A, D, pre);
                    }
                    has = true;
                }
                tnext += (uint32_t)__popcll(need);
            }
            if (__ballot(has) == 0ull) break;
            if (has) {
                typename DrawLane<R>::T x0, x1, x2;
                if (draw_trial<R>(D, &w.raise[D.cc], x0, x1, x2)) {
                    A.store(D.i, D.U0 + D.u0 * (double)x0, D.U1 + D.u0 * (double)x1, D.U2 + D.u0 * (double)x2);
                    has = false;
                }
            }
        }
        }
    }
    __threadfence_block();
    __syncthreads();
    if (__ballot(perform) != 0ull) {
        double ke = 0.0, m0 = 0.0, m1 = 0.0, m2 = 0.0;
        // the first kConsReg simulators of each lane (cells up to 8 kConsReg)
        // are loaded once and kept for the rescale; the rest of a larger cell
        // is read again.  Accumulation in slot order.
        R rx[kConsReg], ry[kConsReg], rz[kConsReg];
        double mm[kConsReg];
#pragma unroll
        for (int u = 0; u < kConsReg; ++u) {
            const uint32_t i = b + j + 8u * (uint32_t)u;
            rx[u] = R(0); ry[u] = R(0); rz[u] = R(0); mm[u] = 0.0;
            if (i < e) {
                rx[u] = w.b.vx[i]; ry[u] = w.b.vy[i]; rz[u] = w.b.vz[i];
                mm[u] = w.sp.mass[(one_sp >= 0) ? (uint32_t)one_sp : w.b.sp[i]];
            }
        }
#pragma unroll
        for (int u = 0; u < kConsReg; ++u) {
            if (b + j + 8u * (uint32_t)u < e) {
                const double ux = (double)rx[u], uy = (double)ry[u], uz = (double)rz[u];
                ke = ke + mm[u] * (ux * ux + uy * uy + uz * uz);
                m0 = m0 + mm[u] * ux;
                m1 = m1 + mm[u] * uy;
                m2 = m2 + mm[u] * uz;
            }
        }
        const uint32_t bt = b + 8u * (uint32_t)kConsReg;
        for (uint32_t i0 = bt + j; i0 < e; i0 += 32u) {
            double tm[4], ux[4], uy[4], uz[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const uint32_t i = min(i0 + 8u * u, e - 1u);
                tm[u] = w.sp.mass[(one_sp >= 0) ? (uint32_t)one_sp : w.b.sp[i]];
                ux[u] = (double)w.b.vx[i]; uy[u] = (double)w.b.vy[i]; uz[u] = (double)w.b.vz[i];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (i0 + 8u * u < e) {
                    ke = ke + tm[u] * (ux[u] * ux[u] + uy[u] * uy[u] + uz[u] * uz[u]);
                    m0 = m0 + tm[u] * ux[u];
                    m1 = m1 + tm[u] * uy[u];
                    m2 = m2 + tm[u] * uz[u];
                }
            }
        }
        ke = group8_total(ke);
        m0 = group8_total(m0);
        m1 = group8_total(m1);
        m2 = group8_total(m2);
        if (perform) {
            const double pU0 = m0 / M, pU1 = m1 / M, pU2 = m2 / M;
            const double pT = (ke - M * (pU0 * pU0 + pU1 * pU1 + pU2 * pU2)) / (3.0 * kB * (N - 1.0));
            if (pT > kVSmall) {
                const double sc = sqrt(T / pT);
#pragma unroll
                for (int u = 0; u < kConsReg; ++u) {
                    const uint32_t i = b + j + 8u * (uint32_t)u;
                    if (i < e) {
                        w.b.vx[i] = (R)(U0 + ((double)rx[u] - pU0) * sc);
                        w.b.vy[i] = (R)(U1 + ((double)ry[u] - pU1) * sc);
                        w.b.vz[i] = (R)(U2 + ((double)rz[u] - pU2) * sc);
                    }
                }
                for (uint32_t i = bt + j; i < e; i += 8u) {
                    const double vx = (double)w.b.vx[i], vy = (double)w.b.vy[i], vz = (double)w.b.vz[i];
                    w.b.vx[i] = (R)(U0 + (vx - pU0) * sc);
                    w.b.vy[i] = (R)(U1 + (vy - pU1) * sc);
                    w.b.vz[i] = (R)(U2 + (vz - pU2) * sc);
                }
            }
        }
    }
    if (usp && j == 0u) {
        const double raise = bits_to_double(w.raise[c]);
        w.uspmax[c] = (raise > 0.0) ? raise : w.bgkp[c].m0 * 0.999;
    }
}

template <class R, bool Macro>
__global__ void __launch_bounds__(kBlock) k_bgk_conserve(WorldView<R> w)
{
    bgk_conserve_body<R, Macro, kConsRegs>(w);
}

// The FP32-state instance without macroInterpolation at five waves per SIMD,
// keeping four simulators per lane in registers (the same sums in the same
// order): faster where the USP draws dominate, slower where the
// conservation does; the autotuner picks.
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5))) k_bgk_conserve5(WorldView<float> w)
{
    bgk_conserve_body<float, false, 4>(w);
}

// --------------------------------------------------- hybrid decomposition
// localKnudsen::decompose (localKnudsen.C:216-562) for the cells of every
// hybrid case whose decomposition falls on this step.  Same restatement as
// the CPU oracle: fvc::average(fvc::interpolate()) on the cell-face data,
// Bird eq. 4.76/4.77 mean free path, theta average, smoothing, threshold,
// and the refinement passes with each loop deciding from the model ids at
// its start.

__device__ inline bool kn_due(const CaseDev& cs)
{
    return cs.active && cs.coll_model == UGX_COLL_HYBRID
           && ((cs.step + 1u) % (uint32_t)cs.decomposition_interval) == 0u;
}

template <class R>
__global__ void k_kn_begin(WorldView<R> w)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < w.ncases) {
        CaseDev& cs = w.cases[c];
        if (cs.active && cs.coll_model == UGX_COLL_HYBRID) cs.kn_time += cs.dt;
    }
}

// kw[8c + 0..6] = rhoN, rhoM, p, T, U (localKnudsen.C:243-279)
template <class R>
__global__ void k_kn_fields(WorldView<R> w)
{
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < (uint32_t)w.nc; c += gridDim.x * blockDim.x) {
        const CaseDev& cs = w.cases[w.cell_case[c]];
        if (!kn_due(cs)) continue;
        const double* a = &w.knacc[(size_t)kKnRows * c];
        const double FN = cs.fn * w.cw[c], tav = cs.kn_time, V = w.vol[c];
        double f[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (a[0] > kVSmall) {
            f[0] = FN * a[0] / (tav * V);
            f[1] = FN * a[1] / (tav * V);
            const double rhoMMean = FN * a[1] / (V * tav);
            for (int d = 0; d < 3; ++d) f[4 + d] = FN * a[2 + d] / (rhoMMean * V * tav);
            const double ke = 0.5 * (FN * a[5]) / (V * tav);
            const double rhoNMean = FN * a[0] / (V * tav);
            const double uu = f[4] * f[4] + f[5] * f[5] + f[6] * f[6];
            f[3] = 2.0 / (3.0 * kB * rhoNMean) * (ke - 0.5 * rhoMMean * uu);
            f[2] = f[0] * kB * f[3];
        }
        for (int q = 0; q < 7; ++q) w.kw[8 * (size_t)c + q] = f[q];
    }
}

// One pass of fvc::average(fvc::interpolate(.)) for one cell: the face
// values (linear interpolation on internal and cyclic faces, the cell value
// on the others) averaged with the face areas, faces in order.  All face
// data and neighbour values are loaded before the sums.

// Face data of cell c; nb[j] is the cell whose value face j reads (c itself
// unless the face is linked).
template <class R>
__device__ inline void kn_face_data(const WorldView<R>& w, uint32_t c, int8_t (&kd)[kMaxFaces],
                                    uint32_t (&nb)[kMaxFaces], double2 (&aw)[kMaxFaces])
{
#pragma unroll
    for (int j = 0; j < kMaxFaces; ++j) {
        const size_t k = (size_t)c * kMaxFaces + j;
        kd[j] = w.fkind[k];
        const int32_t n = w.geom.nbr[k];
        aw[j] = w.knaw[k];
        nb[j] = (kd[j] == kFaceInternal || kd[j] == kFaceCyclic) ? (uint32_t)n : c;
    }
}

// KnGLL (stride 1): one scalar per cell.
template <class R>
__device__ inline void kn_smooth_scalar(const WorldView<R>& w, uint32_t c, const double* in, double* out)
{
    int8_t kd[kMaxFaces];
    uint32_t nb[kMaxFaces];
    double2 aw[kMaxFaces];
    kn_face_data<R>(w, c, kd, nb, aw);
    const double fc = in[c];
    double fo[kMaxFaces];
#pragma unroll
    for (int j = 0; j < kMaxFaces; ++j) fo[j] = in[nb[j]];
    double num = 0.0, den = 0.0;
#pragma unroll
    for (int j = 0; j < kMaxFaces; ++j) {
        if (kd[j] == kFaceSkip) continue;
        const double fv = (kd[j] == kFaceInternal || kd[j] == kFaceCyclic) ? aw[j].y * fc + (1.0 - aw[j].y) * fo[j] : fc;
        num = num + aw[j].x * fv;
        den = den + aw[j].x;
    }
    out[c] = (den > 0.0) ? num / den : fc;
}

// Components 1-6 of a cell's decomposition record (rhoM, p, T scalars and
// the U vector) in one pass over its faces: per component the same sums as
// the scalar pass, with the symmetry-plane mirror applied to U only.
template <class R>
__device__ inline void kn_smooth_fields(const WorldView<R>& w, uint32_t c, const double* in, double* out)
{
    int8_t kd[kMaxFaces];
    uint32_t nb[kMaxFaces];
    double2 aw[kMaxFaces];
    kn_face_data<R>(w, c, kd, nb, aw);
    double own[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) own[q] = in[8 * (size_t)c + 1 + q];
    double fo[kMaxFaces][6];
#pragma unroll
    for (int j = 0; j < kMaxFaces; ++j)
#pragma unroll
        for (int q = 0; q < 6; ++q) fo[j][q] = in[8 * (size_t)nb[j] + 1 + q];
    double num[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, den = 0.0;
#pragma unroll
    for (int j = 0; j < kMaxFaces; ++j) {
        if (kd[j] == kFaceSkip) continue;
        double fv[6];
        if (kd[j] == kFaceInternal || kd[j] == kFaceCyclic) {
#pragma unroll
            for (int q = 0; q < 6; ++q) fv[q] = aw[j].y * own[q] + (1.0 - aw[j].y) * fo[j][q];
        } else {
#pragma unroll
            for (int q = 0; q < 6; ++q) fv[q] = own[q];
            if (kd[j] == kFaceSymmetry) {
                const KnFace kf = w.knf[(size_t)c * kMaxFaces + j];
                const double un = kf.nx * own[3] + kf.ny * own[4] + kf.nz * own[5];
                fv[3] = own[3] - un * kf.nx;
                fv[4] = own[4] - un * kf.ny;
                fv[5] = own[5] - un * kf.nz;
            }
        }
#pragma unroll
        for (int q = 0; q < 6; ++q) num[q] = num[q] + aw[j].x * fv[q];
        den = den + aw[j].x;
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) out[8 * (size_t)c + 1 + q] = (den > 0.0) ? num[q] / den : own[q];
}

template <class R>
__global__ void __launch_bounds__(kBlock) k_kn_smooth(WorldView<R> w, const double* in, double* out, int pass)
{
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < (uint32_t)w.nc; c += gridDim.x * blockDim.x) {
        const CaseDev& cs = w.cases[w.cell_case[c]];
        if (!kn_due(cs)) continue;
        if (pass < cs.smoothing_passes) {
            out[8 * (size_t)c] = in[8 * (size_t)c];
            kn_smooth_fields<R>(w, c, in, out);
        } else {
            for (int q = 0; q < 7; ++q) out[8 * (size_t)c + q] = in[8 * (size_t)c + q];
        }
    }
}

// Smoothing pass of KnGLL (stride 1): kn -> out.
template <class R>
__global__ void __launch_bounds__(kBlock) k_kn_smooth_kn(WorldView<R> w, const double* in, double* out, int pass)
{
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < (uint32_t)w.nc; c += gridDim.x * blockDim.x) {
        const CaseDev& cs = w.cases[w.cell_case[c]];
        if (!kn_due(cs)) continue;
        if (pass < cs.smoothing_passes) kn_smooth_scalar<R>(w, c, in, out);
        else out[c] = in[c];
    }
}

// Gradients, mean free path, instantaneous Kn and its theta average
// (localKnudsen.C:299-398); f is the smoothed field table.
template <class R>
__global__ void __launch_bounds__(kBlock) k_kn_compute(WorldView<R> w, const double* f)
{
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < (uint32_t)w.nc; c += gridDim.x * blockDim.x) {
        const CaseDev& cs = w.cases[w.cell_case[c]];
        if (!kn_due(cs)) continue;
        const double* fc = &f[8 * (size_t)c];
        const double mU = sqrt(fc[4] * fc[4] + fc[5] * fc[5] + fc[6] * fc[6]);
        double gr = 0.0, gt = 0.0, gu = 0.0;
        // face data and neighbour values first, then the face loop in order
        bool in_f[kMaxFaces];
        double dist[kMaxFaces], fo1[kMaxFaces], fo3[kMaxFaces], fo4[kMaxFaces], fo5[kMaxFaces], fo6[kMaxFaces];
#pragma unroll
        for (int j = 0; j < kMaxFaces; ++j) {
            const size_t k = (size_t)c * kMaxFaces + j;
            in_f[j] = w.fkind[k] == kFaceInternal;
            const int32_t n = w.geom.nbr[k];
            dist[j] = w.knf[k].dist;
            const double* fo = &f[8 * (size_t)(in_f[j] ? (uint32_t)n : c)];
            fo1[j] = fo[1]; fo3[j] = fo[3]; fo4[j] = fo[4]; fo5[j] = fo[5]; fo6[j] = fo[6];
        }
#pragma unroll
        for (int j = 0; j < kMaxFaces; ++j) {
            if (!in_f[j]) continue;
            const double mUo = sqrt(fo4[j] * fo4[j] + fo5[j] * fo5[j] + fo6[j] * fo6[j]);
            double v;
            v = fabs(fo1[j] - fc[1]) / dist[j]; if (gr < v) gr = v;
            v = fabs(fo3[j] - fc[3]) / dist[j]; if (gt < v) gt = v;
            v = fabs(mUo - mU) / dist[j]; if (gu < v) gu = v;
        }
        const double* a = &w.knacc[(size_t)kKnRows * c];
        const double bmax = cs.breakdown_max;
        double inst;
        if (a[0] > kVSmall && fc[3] > kVSmall) {
            const double V = w.vol[c], FN = cs.fn * w.cw[c], T = fc[3];
            double mfp = 0.0;
            for (int i = 0; i < cs.ns; ++i) {
                const int gi = cs.sp_base + i;
                double smfp = 0.0;
                for (int q = 0; q < cs.ns; ++q) {
                    const int gq = cs.sp_base + q;
                    const double dPQ = 0.5 * (w.sp.diam[gi] + w.sp.diam[gq]);
                    const double omPQ = 0.5 * (w.sp.omega[gi] + w.sp.omega[gq]);
                    const double mr = w.sp.mass[gi] / w.sp.mass[gq];
                    const double nq = FN * a[8 + q];
                    if (nq > kVSmall) {
                        const double nD = nq / V;
                        smfp += kPi * dPQ * dPQ * nD * dpow(cs.tref / T, omPQ - 0.5) * sqrt(1.0 + mr);
                    }
                }
                const double ni = FN * a[8 + i];
                if (ni > kVSmall) {
                    smfp = 1.0 / smfp;
                    mfp += smfp * ni / (fc[0] * V);
                }
            }
            const double u0 = sqrt(2.0 * kB / (fc[1] / fc[0]) * T);
            const double kr = mfp * gr / fc[1];
            const double kt = mfp * gt / T;
            const double ku = mfp * gu / ((mU < u0) ? u0 : mU);
            inst = (kr < kt) ? kt : kr;
            inst = (inst < ku) ? ku : inst;
        } else {
            inst = 2.0 * bmax;
        }
        w.kn[c] = cs.kn_theta * inst + (1.0 - cs.kn_theta) * w.kn[c];
    }
}

// Threshold, then the snapshot for the first refinement loop.
template <class R>
__global__ void k_kn_decide(WorldView<R> w, const double* kn)
{
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < (uint32_t)w.nc; c += gridDim.x * blockDim.x) {
        const CaseDev& cs = w.cases[w.cell_case[c]];
        if (!kn_due(cs)) continue;
        w.kn[c] = kn[c];
        w.cmodel[c] = (kn[c] > cs.breakdown_max) ? 1 : 0;
    }
}

template <class R>
__global__ void k_kn_snapshot(WorldView<R> w)
{
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < (uint32_t)w.nc; c += gridDim.x * blockDim.x)
        w.ids2[c] = w.cmodel[c];
}

constexpr int kNbCap = 192;   // 3 levels of at most 6 faces: 1 + 6 + 30 + 150 cells

// Cells within three face-neighbour steps of c (internal faces only), c
// first: uniGasHybridDecomposition::fetchCellNeighborhood
// (uniGasHybridDecomposition.C:127-174).  The lists depend on the mesh only
// and are built once per world into a CSR table.
template <class R>
__device__ inline int nb_bfs(const WorldView<R>& w, uint32_t c, uint32_t* list)
{
    int n = 0, lo = 0;
    list[n++] = c;
    for (int l = 0; l < 3; ++l) {
        const int hi = n;
        for (int i = lo; i < hi; ++i)
            for (int j = 0; j < kMaxFaces; ++j) {
                const size_t k = (size_t)list[i] * kMaxFaces + j;
                if (w.fkind[k] != kFaceInternal) continue;
                const uint32_t o = (uint32_t)w.geom.nbr[k];
                bool seen = false;
                for (int q = 0; q < n; ++q) if (list[q] == o) { seen = true; break; }
                if (!seen && n < kNbCap) list[n++] = o;
            }
        lo = hi;
    }
    return n;
}

// fill == nullptr: counts into cnt[c]; else the list of c at fill + off[c].
template <class R>
__global__ void __launch_bounds__(64) k_nb_build(WorldView<R> w, uint32_t* cnt, const uint32_t* off, uint32_t* fill)
{
    uint32_t list[kNbCap];
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < (uint32_t)w.nc; c += gridDim.x * blockDim.x) {
        const CaseDev& cs = w.cases[w.cell_case[c]];
        if (cs.coll_model != UGX_COLL_HYBRID) {
            if (!fill) cnt[c] = 0u;
            continue;
        }
        const int n = nb_bfs<R>(w, c, list);
        if (!fill) {
            cnt[c] = (uint32_t)n;
        } else {
            for (int q = 0; q < n; ++q) fill[off[c] + q] = list[q];
        }
    }
}

static_assert(kMaxFaces <= 8, "one lane of the group per face");

// Sum of v over the eight lanes of a group (lanes 8g .. 8g+7); called with
// all 64 lanes active.
__device__ inline uint32_t group8_sum_u32(uint32_t v)
{
    v += __shfl_xor(v, 1);
    v += __shfl_xor(v, 2);
    v += __shfl_xor(v, 4);
    return v;
}

// One refinement loop (uniGasHybridDecomposition.C:73-75; localKnudsen.C:436-545)
// for the cells of model `me` (1 DSMC, 0 BGK), deciding from ids2.  Eight
// lanes per cell: lane j < 6 looks at face j, and lane j counts the
// neighbourhood entries j, j + 8, ... (the eight lanes read consecutive
// entries of the list); every lane of the wave runs every iteration.
template <class R>
__global__ void __launch_bounds__(kBlock) k_kn_refine(WorldView<R> w, int me)
{
    const uint32_t j = threadIdx.x & 7u;
    const uint32_t nc = (uint32_t)w.nc;
    const uint32_t first = (blockIdx.x * blockDim.x + (threadIdx.x & ~63u)) >> 3;
    const uint32_t stride = (gridDim.x * blockDim.x) >> 3;
    for (uint32_t c0 = first; c0 < nc; c0 += stride) {
        const uint32_t c = c0 + ((threadIdx.x & 63u) >> 3);
        bool on = c < nc;
        if (on) {
            const CaseDev& cs = w.cases[w.cell_case[c]];
            on = kn_due(cs) && w.ids2[c] == me;
        }
        uint32_t same = 0u, other = 0u;
        if (on && j < (uint32_t)kMaxFaces) {
            const size_t k = (size_t)c * kMaxFaces + j;
            if (w.fkind[k] == kFaceInternal) {
                if (w.ids2[w.geom.nbr[k]] == me) same = 1u; else other = 1u;
            }
        }
        same = group8_sum_u32(same);
        other = group8_sum_u32(other);
        const bool flip_faces = same == 0u || (same == 1u && other > 1u);
        uint32_t b = 0u, n = 0u, cnt = 0u;
        if (on && !flip_faces) {
            b = w.nb_off[c];
            n = w.nb_off[c + 1] - b;
        }
        for (uint32_t q = j; q < n; q += 32u) {
            int32_t id[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const uint32_t qq = q + 8u * (uint32_t)u;
                id[u] = (qq < n) ? w.ids2[w.nb_list[b + qq]] : -1;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) cnt += (id[u] == me) ? 1u : 0u;
        }
        cnt = group8_sum_u32(cnt);
        if (on && j == 0u) {
            if (flip_faces) w.cmodel[c] = 1 - me;
            else if ((double)cnt < 0.4 * (double)n) w.cmodel[c] = 1 - me;
        }
    }
}

__device__ inline bool kn_reset_now(const CaseDev& cs)
{
    const double time = cs.time + cs.dt;
    const double until = (cs.reset_until_time > 0.0) ? cs.reset_until_time : 1.0e300;
    return cs.reset_at_decomposition && time < until + 0.5 * cs.dt;
}

template <class R>
__global__ void k_kn_reset(WorldView<R> w)
{
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < (uint32_t)w.nc; c += gridDim.x * blockDim.x) {
        const CaseDev& cs = w.cases[w.cell_case[c]];
        if (!kn_due(cs) || !kn_reset_now(cs)) continue;
        for (int q = 0; q < kKnRows; ++q) w.knacc[(size_t)kKnRows * c + q] = 0.0;
    }
}

template <class R>
__global__ void k_kn_end(WorldView<R> w)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < w.ncases) {
        CaseDev& cs = w.cases[c];
        if (kn_due(cs) && kn_reset_now(cs)) cs.kn_time = 0.0;
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
            cs.step += 1u;
            cs.time = cs.time + cs.dt;
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
            const double req = nd[s] / cs.fn * w.vol[c] / w.cw[c];
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

template <class R>
uint32_t move_blocks(int device)
{
    int per_cu = 0, ncu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_move<R>, kBlock, 0) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 0;
    if (per_cu < 1) per_cu = 1;
    return (uint32_t)(per_cu * ncu);
}
template uint32_t move_blocks<float>(int);
template uint32_t move_blocks<double>(int);

static inline unsigned grid_for(uint64_t n, unsigned cap_blocks = 8192)
{
    uint64_t g = (n + kBlock - 1) / kBlock;
    if (g < 1) g = 1;
    if (g > cap_blocks) g = cap_blocks;
    return (unsigned)g;
}

// Blocks of the counting and scatter passes (kCountUnroll slots per thread).
static inline unsigned count_grid(uint64_t cap)
{
    return grid_for((cap + kCountUnroll - 1) / kCountUnroll);
}

template <class R>
static void launch_decompose(const WorldView<R>& w, hipStream_t s)
{
    const unsigned g = grid_for((uint64_t)w.nc);
    hipLaunchKernelGGL(k_kn_fields<R>, dim3(g), dim3(kBlock), 0, s, w);
    double* a = w.kw;
    double* b = w.kw2;
    for (int p = 0; p < w.kn_passes; ++p) {
        hipLaunchKernelGGL(k_kn_smooth<R>, dim3(g), dim3(kBlock), 0, s, w, (const double*)a, b, p);
        std::swap(a, b);
    }
    hipLaunchKernelGGL(k_kn_compute<R>, dim3(g), dim3(kBlock), 0, s, w, (const double*)a);
    // KnGLL smoothing ping-pongs between kn and the first column-free table b
    double* k0 = w.kn;
    double* k1 = b;
    for (int p = 0; p < w.kn_passes; ++p) {
        hipLaunchKernelGGL(k_kn_smooth_kn<R>, dim3(g), dim3(kBlock), 0, s, w, (const double*)k0, k1, p);
        std::swap(k0, k1);
    }
    hipLaunchKernelGGL(k_kn_decide<R>, dim3(g), dim3(kBlock), 0, s, w, (const double*)k0);
    const unsigned g8 = grid_for((uint64_t)w.nc * 8u);
    for (int pass = 0; pass < 3; ++pass)
        for (int me = 1; me >= 0; --me) {
            hipLaunchKernelGGL(k_kn_snapshot<R>, dim3(g), dim3(kBlock), 0, s, w);
            hipLaunchKernelGGL(k_kn_refine<R>, dim3(g8), dim3(kBlock), 0, s, w, me);
        }
    hipLaunchKernelGGL(k_kn_reset<R>, dim3(g), dim3(kBlock), 0, s, w);
    hipLaunchKernelGGL(k_kn_end<R>, dim3(grid_for((uint64_t)w.ncases)), dim3(kBlock), 0, s, w);
}

// Neighbourhood table of the hybrid refinement: counts, exclusive scan into
// off (off[nc] = total, also copied to *total), then the lists.
template <class R>
void launch_nb_count(const WorldView<R>& w, uint32_t* cnt, uint32_t* off, uint32_t* bsum, uint32_t* total, hipStream_t s)
{
    const uint32_t ncp = (uint32_t)w.nc + 1u;
    hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(ncp)), dim3(kBlock), 0, s, cnt, ncp);
    hipLaunchKernelGGL(k_nb_build<R>, dim3(grid_for((uint64_t)w.nc * 4, 65536)), dim3(64), 0, s, w, cnt,
                       (const uint32_t*)nullptr, (uint32_t*)nullptr);
    scan_u32(cnt, ncp, off, bsum, total, s);
}

template <class R>
void launch_nb_fill(const WorldView<R>& w, const uint32_t* off, uint32_t* list, hipStream_t s)
{
    hipLaunchKernelGGL(k_nb_build<R>, dim3(grid_for((uint64_t)w.nc * 4, 65536)), dim3(64), 0, s, w,
                       (uint32_t*)nullptr, off, list);
}

template void launch_nb_count<float>(const WorldView<float>&, uint32_t*, uint32_t*, uint32_t*, uint32_t*, hipStream_t);
template void launch_nb_count<double>(const WorldView<double>&, uint32_t*, uint32_t*, uint32_t*, uint32_t*, hipStream_t);
template void launch_nb_fill<float>(const WorldView<float>&, const uint32_t*, uint32_t*, hipStream_t);
template void launch_nb_fill<double>(const WorldView<double>&, const uint32_t*, uint32_t*, hipStream_t);

template <class R, int Block, uint32_t TileN, uint32_t TileMax>
static void launch_tiles_cfg(const WorldView<R>& w, hipStream_t s, bool sort_only)
{
    const uint32_t ncp = (uint32_t)w.nc + 1u;
    const unsigned nt = (unsigned)((w.cap + (uint64_t)kCellW * (uint64_t)w.nc) / TileN + 2);
    hipLaunchKernelGGL((k_tile_table<R, TileN>), dim3(grid_for(ncp)), dim3(kBlock), 0, s, w);
    // species slots of the sums: 1 (single-species cases), 2, or all kNS
    const int nsl = (w.ns_max <= 1) ? 1 : (w.ns_max <= 2) ? 2 : kNS;
#define UGX_TILE(SO, L, NS) hipLaunchKernelGGL((k_tile<R, SO, L, NS, Block, TileN, TileMax>), dim3(nt), dim3(Block), 0, s, w)
#define UGX_TILE_NS(L) \
    do { if (nsl == 1) UGX_TILE(false, L, 1); else if (nsl == 2) UGX_TILE(false, L, 2); else UGX_TILE(false, L, kNS); } while (0)
    if (sort_only) UGX_TILE(true, 0, 1);
    else if (w.sum_level >= 2) UGX_TILE_NS(2);
    else if (w.sum_level == 1) UGX_TILE_NS(1);
    else UGX_TILE_NS(0);
#undef UGX_TILE_NS
#undef UGX_TILE
}

// Cell-tile configurations (threads, key window, LDS simulators); the
// choice changes the speed only, never a result.
template <class R>
static void launch_tiles(const WorldView<R>& w, hipStream_t s, bool sort_only)
{
    switch (w.tile_variant) {
    case 1: launch_tiles_cfg<R, 256, 512, 1024>(w, s, sort_only); break;
    case 2: launch_tiles_cfg<R, 64, 128, 256>(w, s, sort_only); break;
    default: launch_tiles_cfg<R, 128, 256, 512>(w, s, sort_only); break;
    }
}

template <class R>
void launch_sort(const WorldView<R>& w, hipStream_t s)
{
    const uint32_t ncp = (uint32_t)w.nc + 1u;
    const uint32_t n7 = (uint32_t)w.nc * 7u;
    hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(ncp)), dim3(kBlock), 0, s, w.cnt, ncp);
    hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(ncp)), dim3(kBlock), 0, s, w.ocnt, ncp);
    hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(n7)), dim3(kBlock), 0, s, w.out7, n7);
    hipLaunchKernelGGL(k_count<R>, dim3(count_grid(w.cap)), dim3(kBlock), 0, s, w, 1);
    hipLaunchKernelGGL(k_gather<R>, dim3(grid_for(w.nc)), dim3(kBlock), 0, s, w);
    scan_u32(w.cnt, ncp, w.off, w.blocksum, &w.st->n_new, s);
    hipLaunchKernelGGL(k_scatter<R>, dim3(count_grid(w.cap)), dim3(kBlock), 0, s, w);
    launch_tiles<R>(w, s, true);
    hipLaunchKernelGGL(k_finish_sort<R>, dim3(1), dim3(1), 0, s, w);
}

template <class R>
void launch_reweight(const WorldView<R>& w, const double* cw_old, int ccase, hipStream_t s)
{
    const unsigned gw = count_grid(w.cap);
    hipLaunchKernelGGL(k_reweight_begin<R>, dim3(1), dim3(1), 0, s, w);
    hipLaunchKernelGGL(k_reweight<R>, dim3(gw), dim3(kBlock), 0, s, w, cw_old, ccase);
    scan_u32(w.wclone, w.cap + 1u, w.wcoff, w.wbsum, &w.st->nclone, s);
    hipLaunchKernelGGL(k_clone<R>, dim3(gw), dim3(kBlock), 0, s, w);
    hipLaunchKernelGGL(k_clone_finish<R>, dim3(1), dim3(64), 0, s, w);
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
    hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(ncp)), dim3(kBlock), 0, s, w.ocnt, ncp);
    mark(1);
    if (w.n_inflow > 0) {
        if (w.n_inflow > 4096) {
            // many inflow faces: counted in parallel and scanned in place
            hipLaunchKernelGGL(k_inflow_face_count<R>, dim3(grid_for((uint64_t)w.n_inflow + 1)), dim3(kBlock), 0, s, w);
            scan_u32(w.ins_off, (uint32_t)w.n_inflow + 1u, w.ins_off, w.blocksum, nullptr, s);
            hipLaunchKernelGGL(k_inflow_finish<R>, dim3(1), dim3(1), 0, s, w);
        } else {
            hipLaunchKernelGGL(k_inflow_count<R>, dim3(1), dim3(1024), 0, s, w);
        }
        hipLaunchKernelGGL(k_inflow_gen<R>, dim3(grid_for(w.max_ins)), dim3(kBlock), 0, s, w);
    }
    mark(2);
    hipLaunchKernelGGL(k_move<R>, dim3(w.move_grid), dim3(kBlock), 0, s, w);
    mark(3);
    hipLaunchKernelGGL(k_wall<R>, dim3(grid_for(w.wq_cap, 2048)), dim3(kBlock), 0, s, w);
    if (w.nwf > 0) {
        const uint32_t nw1 = (uint32_t)w.nwf + 1u;
        hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(nw1)), dim3(kBlock), 0, s, w.wcnt, nw1);
        hipLaunchKernelGGL(k_wall_count<R>, dim3(grid_for(w.ev_cap, 1024)), dim3(kBlock), 0, s, w);
        scan_u32(w.wcnt, nw1, w.woff, w.blocksum, nullptr, s);
        hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(nw1)), dim3(kBlock), 0, s, w.wcnt, nw1);
        hipLaunchKernelGGL(k_wall_bucket<R>, dim3(grid_for(w.ev_cap, 1024)), dim3(kBlock), 0, s, w);
        hipLaunchKernelGGL(k_wall_reduce<R>, dim3(grid_for((uint64_t)w.nwf)), dim3(kBlock), 0, s, w);
    }
    if (w.has_weighting) {
        const unsigned gw = count_grid(w.cap);
        hipLaunchKernelGGL(k_weight<R>, dim3(gw), dim3(kBlock), 0, s, w);
        scan_u32(w.wclone, w.cap + 1u, w.wcoff, w.wbsum, &w.st->nclone, s);
        hipLaunchKernelGGL(k_clone<R>, dim3(gw), dim3(kBlock), 0, s, w);
        hipLaunchKernelGGL(k_clone_finish<R>, dim3(1), dim3(64), 0, s, w);
    }
    mark(4);
    hipLaunchKernelGGL((k_count_tiles<R, 256, kCountTiles>),
                       dim3((unsigned)std::min<uint64_t>((max_tiles(w.cap, (uint64_t)w.nc) + kCountTiles - 1) / kCountTiles, kCountGrid)),
                       dim3(256), 0, s, w);
    if (w.has_weighting)
        hipLaunchKernelGGL(k_count<R>, dim3(count_grid(w.cap)), dim3(kBlock), 0, s, w, 0);
    else if (w.n_inflow > 0)
        hipLaunchKernelGGL(k_count<R>, dim3(grid_for(((uint64_t)w.max_ins + kCountUnroll - 1) / kCountUnroll)), dim3(kBlock), 0, s, w, 0);
    hipLaunchKernelGGL(k_gather<R>, dim3(grid_for(w.nc)), dim3(kBlock), 0, s, w);
    scan_u32(w.cnt, ncp, w.off, w.blocksum, &w.st->n_new, s);
    mark(5);
    hipLaunchKernelGGL(k_scatter<R>, dim3(count_grid(w.cap)), dim3(kBlock), 0, s, w);
    mark(6);
    launch_tiles<R>(w, s, false);
    mark(7);
    if (w.has_collide) {
        const unsigned gcell = (unsigned)(((uint64_t)w.nc + kBlock - 1) / kBlock);
        hipLaunchKernelGGL(k_collide<R>, dim3(gcell), dim3(kBlock), 0, s, w);
        if (w.has_macro)
            hipLaunchKernelGGL(k_point_fields<R>, dim3(grid_for(w.npts)), dim3(kBlock), 0, s, w);
        if (w.has_bgk) {
            const unsigned gcons = (unsigned)(((uint64_t)w.nc * 8u + kBlock - 1) / kBlock);
            if (w.has_macro)
                hipLaunchKernelGGL((k_bgk_conserve<R, true>), dim3(gcons), dim3(kBlock), 0, s, w);
            else {
                bool six = false;
                if constexpr (sizeof(R) == sizeof(float)) {
                    six = w.cons_variant == 1;
                    if (six) hipLaunchKernelGGL(k_bgk_conserve5, dim3(gcons), dim3(kBlock), 0, s, w);
                }
                if (!six) hipLaunchKernelGGL((k_bgk_conserve<R, false>), dim3(gcons), dim3(kBlock), 0, s, w);
            }
        }
    }
    if (w.has_hybrid) {
        hipLaunchKernelGGL(k_kn_begin<R>, dim3(gc), dim3(kBlock), 0, s, w);
        if (w.kn_due) launch_decompose<R>(w, s);
    }
    mark(8);
    hipLaunchKernelGGL(k_end_step<R>, dim3(gc), dim3(kBlock), 0, s, w);
    mark(9);
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

// ------------------------------------------------------ state carry-over
// A case moved into a rebuilt world (batched and served forms) keeps its
// particles, per-cell state and counters on the device: these copies shift
// its cell and species ids to the new world's numbering.

template <class R>
__global__ void k_carry_particles(Parts<R> src, Parts<R> dst, uint32_t src_begin, uint32_t n, uint32_t dst_begin,
                                  int32_t cell_delta, int32_t sp_delta)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t a = src_begin + i, b = dst_begin + i;
        dst.x[b] = src.x[a]; dst.y[b] = src.y[a]; dst.z[b] = src.z[a];
        dst.vx[b] = src.vx[a]; dst.vy[b] = src.vy[a]; dst.vz[b] = src.vz[a];
        dst.erot[b] = src.erot[a];
        dst.cell[b] = (uint32_t)((int32_t)src.cell[a] + cell_delta);
        dst.sp[b] = (uint32_t)((int32_t)src.sp[a] + sp_delta);
    }
}

// Per-cell state of ncl cells: old cells [ob, ob+ncl) -> new cells [nb, nb+ncl).
__global__ void k_carry_cells(CellState o, CellState d, uint32_t ob, uint32_t nb, uint32_t ncl)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ncl; i += gridDim.x * blockDim.x) {
        const size_t a = ob + i, b = nb + i;
        d.smax[b] = o.smax[a];
        d.cmodel[b] = o.cmodel[a];
        for (int q = 0; q < kNSamp; ++q) d.samp[b * kNSamp + q] = o.samp[a * kNSamp + q];
        for (int q = 0; q < 3; ++q) d.uspq[3 * b + q] = o.uspq[3 * a + q];
        for (int q = 0; q < 6; ++q) d.usps[6 * b + q] = o.usps[6 * a + q];
        d.uspmax[b] = o.uspmax[a];
        for (int q = 0; q < kKnRows; ++q) d.knacc[(size_t)kKnRows * b + q] = o.knacc[(size_t)kKnRows * a + q];
        d.kn[b] = o.kn[a];
    }
}

template <class R>
void launch_carry_particles(const Parts<R>& src, const Parts<R>& dst, uint32_t src_begin, uint32_t n,
                            uint32_t dst_begin, int32_t cell_delta, int32_t sp_delta, hipStream_t s)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_carry_particles<R>, dim3(grid_for(n)), dim3(kBlock), 0, s, src, dst, src_begin, n, dst_begin,
                       cell_delta, sp_delta);
}

void launch_carry_cells(const CellState& o, const CellState& d, uint32_t ob, uint32_t nb, uint32_t ncl, hipStream_t s)
{
    if (ncl == 0) return;
    hipLaunchKernelGGL(k_carry_cells, dim3(grid_for(ncl)), dim3(kBlock), 0, s, o, d, ob, nb, ncl);
}

template void launch_carry_particles<float>(const Parts<float>&, const Parts<float>&, uint32_t, uint32_t, uint32_t,
                                            int32_t, int32_t, hipStream_t);
template void launch_carry_particles<double>(const Parts<double>&, const Parts<double>&, uint32_t, uint32_t, uint32_t,
                                             int32_t, int32_t, hipStream_t);

__global__ void k_fill_params(FillParams p, double* d)
{
    const int i = threadIdx.x;
    if (i < kNS) d[i] = p.nd[i];
    else if (i < kNS + 3) d[i] = p.T[i - kNS];
    else if (i < kNS + 6) d[i] = p.U[i - kNS - 3];
}

void launch_fill_params(const FillParams& p, double* d, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_params, dim3(1), dim3(64), 0, s, p, d);
}

__global__ void k_set_f64(double* d, double v, uint32_t n)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) d[i] = v;
}

void launch_set_f64(double* d, double v, uint32_t n, hipStream_t s)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_set_f64, dim3(grid_for(n)), dim3(kBlock), 0, s, d, v, n);
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
template void launch_reweight<float>(const WorldView<float>&, const double*, int, hipStream_t);
template void launch_reweight<double>(const WorldView<double>&, const double*, int, hipStream_t);
template void launch_fill<float>(const WorldView<float>&, hipStream_t, int, const double*, const double*, double,
                                 const double*, int, const double*, const double*, uint32_t*, uint32_t*, int, int);
template void launch_fill<double>(const WorldView<double>&, hipStream_t, int, const double*, const double*, double,
                                  const double*, int, const double*, const double*, uint32_t*, uint32_t*, int, int);

}  // namespace ugx
