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
    uint32_t ntiles;        // cell tiles of this step's sort
    uint32_t nsel;          // BGK selections this step
    uint32_t nev;           // wall-hit events this step
    uint32_t nev_peak;      // largest nev of a finished step
    uint32_t evover;        // a step had more wall hits than the event buffer holds
    uint32_t nclone;        // clones of this step's cell weighting
};

// One interaction of a simulator with a measured wall face: velocity and
// E_rot before and after it, the world wall face and the global species.
struct WallEv {
    double u0[3], u1[3];
    double e0, e1;
    uint32_t wf;
    uint32_t sp;
};
// Rows of the wall sums per (face, species): ugx_wall_index.
constexpr int kWRows = UGX_W_COUNT;

// BGK-family parameters of one cell for the velocity draws and the
// conservation step (one record per cell of a BGK-active case).
struct BgkP {
    double U[3];
    double s2kT, p, cQ, cS, m0, T;
    double q[3];
    double sm[9];      // ES-BGK S matrix (row major) or USP stress (xx xy xz yy yz zz)
    int32_t perform;
    int32_t model;
    uint32_t sel_base;  // first entry of the cell's selections in the selection list
    uint32_t nrel;      // number of selections
};

// Words of the block-sum buffer of an exclusive scan of n entries (1024
// entries per block): the block sums, then the sums of the block sums.
inline size_t scan_bsum_words(uint64_t n)
{
    const size_t nb = (size_t)((n + 1023) / 1024);
    return nb + 1 + (nb + 1023) / 1024 + 2;
}

// Per-case cumulative counters (uint64).
enum CounterIdx {
    kCntCandidates = 0, kCntCollisions, kCntBgk, kCntInserted, kCntDeleted,
    kCntWallHits, kCntTrackErrors, kCntSampled, kCntParticleSteps, kCntCloned, kCntWeightDeleted, kCntNum
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
    double* samp;             // [nc][kNSamp] (cell-major; downloads transpose)
    double* uspq;             // [3*nc]
    double* usps;             // [6*nc]
    double* uspmax;           // [nc]
    double* knacc;            // [kKnRows*nc]
    double* kn;               // [nc]
    const double* cc;         // [3*nc] reference points (hybrid distances)
    Species sp;
    Pairs pr;
    uint32_t* cnt;            // [nc+1] simulators per cell after the move
    uint32_t* ocnt;           // [nc+1] of them, those counted by atomics (class 7)
    uint32_t* out7;           // [7*nc] per old cell: simulators staying (0) and leaving to the
                              //        neighbour behind face f (1 + f)
    uint32_t* pre;            // [8*nc] per cell: start of each arrival class within the cell
    const int8_t* rev;        // [6*nc] first face of the neighbour behind face f that points back
    uint32_t* off;            // [nc+1] offsets
    uint32_t* perm;           // [cap]
    uint32_t* ins_off;        // [n_inflow+1]
    uint32_t* scratch;        // [cap] rank of each slot within its cell (count -> scatter), then
                              //       the slot order of tiles larger than LDS
    uint32_t* tile_first;     // [max tiles + 2] first cell of each tile
    uint32_t* tile_p0;        //                 its first sorted slot
    int32_t* tile_case;       //                 the case of its first cell
    int32_t nsp;              // species of all cases
    double* mom;              // [nc][kNMom] per-cell sums for the BGK family
    BgkP* bgkp;               // [nc]
    unsigned long long* raise;// [nc] USP maxProb raise (bits of a positive double)
    uint32_t* sel_slot;       // [cap] BGK selections: sorted slot of each (cell-major, in selection order)
    int32_t ns_max;           // largest species count of a case
    int32_t sum_level;        // largest sampler level of a case (0, 1, 2)
    int32_t has_collide;      // some case collides (DSMC or BGK family)
    int32_t has_hybrid;       // some case switches models per cell
    int32_t tile_variant;     // cell-tile configuration (0..kTileVariants-1)
    int32_t cons_variant;     // BGK conservation kernel: 0 four waves per SIMD, 1 five (FP32 state)
    uint32_t move_grid;       // blocks of the move: as many as are resident at once
    int32_t kn_due;           // a hybrid case decomposes at this step
    int32_t kn_passes;        // largest smoothing pass count of the due cases
    const KnFace* knf;        // [nc*kMaxFaces]
    const double2* knaw;      // [nc*kMaxFaces] area and interpolation weight of each face (the smoothing's reads)
    const int8_t* fkind;      // [nc*kMaxFaces]
    double* kw;               // [8*nc] decomposition fields (ping)
    double* kw2;              // [8*nc] (pong)
    int32_t* ids2;            // [nc] model ids at the start of a refinement loop
    const uint32_t* nb_off;   // [nc+1] neighbourhood table of the refinement (CSR)
    const uint32_t* nb_list;
    int32_t has_bgk;          // some case runs a BGK-family model
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
    // wall measurements (boundaryMeasurements): per cell-face the world wall
    // face or -1; per wall face its geometry (kWGeo doubles) and cell; the
    // step's events, their per-face buckets, and the time sums
    int32_t nwf;
    const int32_t* wface;     // [nc*kMaxFaces]
    const double* wgeo;       // [nwf*kWGeo]
    const int32_t* wcell;     // [nwf]
    WallEv* ev;               // [ev_cap]
    unsigned long long* ev_key;   // [ev_cap] local species << 42 | slot << 10 | crossing
    uint32_t ev_cap;
    uint32_t* wcnt;           // [nwf+1]
    uint32_t* woff;           // [nwf+1]
    unsigned long long* bkey; // [ev_cap] keys in face buckets
    uint32_t* bidx;           // [ev_cap]
    double* wsamp;            // [nwf][kNS][kWRows]
    // cell weighting: per cell its weight W_c (1 in unweighted cases); with
    // weighting on, per slot the cell a simulator started the step in, its
    // clone count and their offsets
    const double* cw;         // [nc]
    int32_t has_weighting;
    uint32_t* cstart;         // [cap]
    uint32_t* wclone;         // [cap + 1]
    uint32_t* wcoff;          // [cap + 1]
    uint32_t* wbsum;          // scan scratch for cap entries
    // NTC subcells (noTimeCounter.C:97-161): with any cell split, per cell
    // its packed levels (x | y << 8 | z << 16) and the offset of its subcell
    // table (2 nsub + 1 words: subcell starts, then cursors), per slot of
    // the sorted order its subcell and the cells' subcell lists
    int32_t has_subcells;
    const uint32_t* sclev;    // [nc]
    const uint32_t* scoff;    // [nc]
    uint32_t* sctab;          // [sum over cells of 2 nsub + 1]
    uint32_t* scid;           // [cap]
    uint32_t* sclist;         // [cap]
    const double* bblo;       // [nc*3] cell bounding box in the cell frame
    const double* bbhi;       // [nc*3]
    // adaptive simulation: per cell the adapter's time sums since its last
    // adaptation (rows as the hybrid time sums)
    double* adacc;            // [nc*kKnRows]
    // macroInterpolation (cells of cases with it on): per cell the centroid
    // in the cell frame and the point ids of its faces, the mesh points,
    // per point the cells and weights of its value; per cell and per point
    // the kMiRows field rows of the step
    int32_t has_macro;
    uint32_t npts;
    const double* mi_cent;    // [nc*3]
    const int32_t* mi_fp;     // [nc*kMaxFaces*kMiFaceVerts] world point ids, -1 padded
    const double* mi_pts;     // [npts*3]
    const uint32_t* mi_poff;  // [npts+1]
    const int32_t* mi_pcell;  // world cells
    const double* mi_pw;
    double* mfield;           // [nc*kMiRows]
    double* pfield;           // [npts*kMiRows]
};

// Largest subcell level along an axis (uniGasDynamicAdapter.C:49).
constexpr int kMaxSubcellLevels = 10;
// macroInterpolation: points per face kept for the tets, and the per-cell
// and per-point field rows (T, U, Pr, collision frequency, p, q, then the
// pressure tensor (ES-BGK) or the stress (USP) as xx xy xz yy yz zz).
constexpr int kMiFaceVerts = 8;
constexpr int kMiRows = 16;

// Records wall hits of measured faces into the step's event list.  Only
// the hits of a sampling step of the cell's case are recorded: the
// reference's per-step boundary sums enter its time sums only at those
// steps and are cleared after every step (uniGasVolFields.C:800-830,
// uniGasCloud.C:865).
struct WallSink {
    const int32_t* wface;
    WallEv* ev;
    unsigned long long* key;
    uint32_t cap;
    DevState* st;
    const int32_t* local;
    const int32_t* cell_case;
    const CaseDev* cases;
    __device__ bool measured(uint32_t c, int f) const
    {
        if (wface == nullptr || wface[(size_t)c * kMaxFaces + f] < 0) return false;
        const CaseDev& cs = cases[cell_case[c]];
        return ((cs.step + 1u) % (uint32_t)cs.sample_interval) == 0u;
    }
    __device__ void hit(uint32_t c, int f, uint32_t slot, int ord, const double* up, const double* uq, double e0,
                        double e1, uint32_t sp) const
    {
        const int32_t wf = wface[(size_t)c * kMaxFaces + f];
        const uint32_t i = atomicAdd(&st->nev, 1u);
        if (i >= cap) { st->evover = 1u; return; }
        key[i] = ((unsigned long long)(uint32_t)local[sp] << 42) | ((unsigned long long)slot << 10) |
                 (unsigned long long)(uint32_t)ord;
        WallEv e;
        for (int d = 0; d < 3; ++d) { e.u0[d] = up[d]; e.u1[d] = uq[d]; }
        e.e0 = e0;
        e.e1 = e1;
        e.wf = (uint32_t)wf;
        e.sp = sp;
        ev[i] = e;
    }
};

template <class R>
__device__ inline WallSink wall_sink(const WorldView<R>& w)
{
    return WallSink{w.nwf > 0 ? w.wface : nullptr, w.ev, w.ev_key, w.ev_cap, w.st, w.sp.local, w.cell_case, w.cases};
}

// Rows of the per-cell sums kept for the BGK family: every base row and the
// per-species counts.
constexpr int kNMom = UGX_S_SPECIES_MCC;

// Cell tiles of the sort: tile t holds the cells c whose key
// off[c] + kCellW * c lies in [t * TileN, (t + 1) * TileN), so a tile has
// at most TileN / kCellW cells and fewer than TileN simulators plus those
// of its last cell.
// Tile configurations: (threads, TileN window, LDS simulators) =
// (128, 256, 512), (256, 512, 1024), (64, 128, 256); the table is sized for
// the smallest window.
constexpr uint32_t kCellW = 8;
constexpr uint32_t kMinTileN = 128;
constexpr int kTileVariants = 3;
constexpr int kConsVariants = 2;
constexpr uint32_t kLdsSpecies = 64;   // species tables up to this size are staged in LDS per tile
inline uint64_t max_tiles(uint64_t cap, uint64_t nc) { return (cap + kCellW * nc) / kMinTileN + 2; }

// Host launchers (ugx_kernels.hip).
constexpr int kPhases = 9;  // begin, inflow, move, wall, scan, scatter, cells, collide, end
template <class R> void launch_step(const WorldView<R>& w, hipStream_t s, hipEvent_t* marks);
// Blocks of the move kernel that are resident on `device` at once (0 on error).
template <class R> uint32_t move_blocks(int device);
template <class R> void launch_sort(const WorldView<R>& w, hipStream_t s);
// Every simulator of case ccase from its cell's weight in cw_old to the
// current weight (clones appended, removals marked); a sort follows.
template <class R> void launch_reweight(const WorldView<R>& w, const double* cw_old, int ccase, hipStream_t s);
template <class R> void launch_fill(const WorldView<R>& w, hipStream_t s, int ccase,
                                    const double* nd, const double* T, double rotT, const double* U,
                                    int exact, const double* bblo, const double* bbhi,
                                    uint32_t* fill_cnt, uint32_t* fill_off, int ns, int ncell_case);
void launch_copy_probe(const void* in, void* out, size_t bytes, hipStream_t s);
// Parameters of a meshFill: number densities per species, per-axis
// temperatures and the drift velocity, stored into device memory at d
// (kFillParams doubles: nd[kNS], T[3], U[3]) by a kernel.
constexpr int kFillParams = kNS + 6;
struct FillParams {
    double nd[kNS];
    double T[3];
    double U[3];
};
void launch_fill_params(const FillParams& p, double* d, hipStream_t s);
void launch_set_f64(double* d, double v, uint32_t n, hipStream_t s);
template <class R>
void launch_nb_count(const WorldView<R>& w, uint32_t* cnt, uint32_t* off, uint32_t* bsum, uint32_t* total, hipStream_t s);
template <class R>
void launch_nb_fill(const WorldView<R>& w, const uint32_t* off, uint32_t* list, hipStream_t s);

// Per-cell state of a world, for the carry-over of a case between worlds.
struct CellState {
    double* smax;
    int32_t* cmodel;
    double* samp;
    double* uspq;
    double* usps;
    double* uspmax;
    double* knacc;
    double* kn;
    int32_t nc;
};
template <class R>
void launch_carry_particles(const Parts<R>& src, const Parts<R>& dst, uint32_t src_begin, uint32_t n,
                            uint32_t dst_begin, int32_t cell_delta, int32_t sp_delta, hipStream_t s);
void launch_carry_cells(const CellState& o, const CellState& d, uint32_t ob, uint32_t nb, uint32_t ncl, hipStream_t s);

}  // namespace ugx
