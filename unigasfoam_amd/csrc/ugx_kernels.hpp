// ugx_kernels.hpp - device-state layout shared by the kernels and the host
// orchestration, and the host-side launch entry points.
#pragma once

#include "ugx_dev.hpp"

namespace ugx {

// Device-resident scalar state of a world.
struct DevState {
    uint32_t n;             // simulators before this step's sort
    uint32_t n_ins;         // inserted this step
    uint32_t n_new;         // after the sort
    uint32_t overflow;      // capacity exceeded
    uint32_t big_tiles;     // cell tiles handled by the global-memory path
    uint32_t nwall;         // queued diffuse-wall tracks this step
    uint32_t pad[2];
};

// Per-case cumulative counters (uint64).
enum CounterIdx {
    kCntCandidates = 0, kCntCollisions, kCntBgk, kCntInserted, kCntDeleted,
    kCntWallHits, kCntTrackErrors, kCntSampled, kCntParticleSteps, kCntNum
};

template <class R>
struct WorldView {
    Parts<R> a;               // current particle buffers
    Parts<R> b;               // sort destination
    Geom<R> geom;
    const BModel* bm;
    const InflowFS* inflow;
    int32_t n_inflow;         // face-species entries (all cases)
    uint32_t max_ins;         // upper bound on inserts per step
    CaseDev* cases;
    int32_t ncases;
    const int32_t* cell_case; // [nc]
    const double* vol;        // [nc]
    double* smax;             // [nc]
    int32_t* cmodel;          // [nc] 1 DSMC, 0 BGK
    double* samp;             // [kNSamp * nc]
    double* uspq;             // [3*nc]
    double* usps;             // [6*nc]
    double* uspmax;           // [nc]
    double* knacc;            // [8*nc]
    double* kn;               // [nc]
    const double* cc;         // [3*nc] reference points (hybrid distances)
    Species sp;
    Pairs pr;
    uint32_t* cnt;            // [nc+1] counts (histogram)
    uint32_t* off;            // [nc+1] offsets
    uint32_t* cursor;         // [nc]
    uint32_t* perm;           // [cap]
    uint32_t* ins_off;        // [n_inflow+1]
    uint32_t* scratch;        // [cap] per-cell scratch (BGK index lists)
    uint32_t* blocksum;       // scan scratch
    unsigned long long* counters; // [ncases*kCntNum]
    DevState* st;
    int32_t nc;
    uint32_t cap;
    // diffuse-wall queue
    uint32_t* wq_slot;
    uint32_t* wq_info;
    double* wq_tau;
    uint32_t wq_cap;
};

// Host launchers (ugx_kernels.hip).
constexpr int kPhases = 8;  // begin, inflow, move, wall, scan, scatter, cells, end
template <class R> void launch_step(const WorldView<R>& w, hipStream_t s, hipEvent_t* marks);
template <class R> void launch_sort(const WorldView<R>& w, hipStream_t s);
template <class R> void launch_fill(const WorldView<R>& w, hipStream_t s, int ccase,
                                    const double* nd, const double* T, double rotT, const double* U,
                                    int exact, const double* bblo, const double* bbhi,
                                    uint32_t* fill_cnt, uint32_t* fill_off, int ns, int ncell_case);
void launch_copy_probe(const void* in, void* out, size_t bytes, hipStream_t s);
template <class R> void launch_hybrid(const WorldView<R>& w, hipStream_t s, double* work);

}  // namespace ugx
