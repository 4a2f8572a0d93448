// ugx_world.hpp - host-side world: device arrays of one or more cases.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "ugx_kernels.hpp"

namespace ugx {

extern thread_local std::string g_err;
int set_err(int code, const char* fmt, ...);

// Geometry of one case's mesh in the storage form of the kernels.
struct MeshGeom {
    int nc = 0;
    std::vector<double> cc, vol, bblo, bbhi, plane, delta;
    std::vector<double> farea;     // per cell-face area (0 for padded and empty faces)
    std::vector<int32_t> ncf, nbr;
    std::vector<int8_t> fkind;     // per cell-face kFace*
    double mask[3] = {1.0, 1.0, 1.0};
    std::vector<BModel> bm;        // per patch
    std::vector<InflowFS> inflow;  // face-species entries (case-local)
    // per entry the parts of its count that do not depend on the time step:
    // face area times number density, most probable speed, and the flux
    // bracket; nbar_num = ((fA n) dt) c_mp bracket
    std::vector<double> inflow_parts;   // [3 * inflow.size()]
    // measured wall faces: the faces of wall patches with a specular or
    // diffuse model (uniGasPatchBoundary measurePropertiesAtWall_), in
    // patch then face order
    int nwf = 0;
    std::vector<int32_t> wface;    // per cell-face: wall face index, or -1
    std::vector<double> wgeo;      // per wall face [kWGeo]: unit outward normal, area, tangent t1
    std::vector<int32_t> wpatch, wlocal, wcell;   // per wall face: patch, face within the patch, cell
    // macroInterpolation (ugx_interp.cpp): per cell its centroid in the
    // cell frame and per cell-face up to kMiFaceVerts point ids (-1 padded;
    // the face's tets are (centroid, p0, p_i, p_i+1)); the mesh points; per
    // point the cells and weights of its value
    std::vector<double> mi_cent;     // [nc*3]
    std::vector<int32_t> mi_fp;      // [nc*kMaxFaces*kMiFaceVerts]
    std::vector<double> mi_pts;      // [npts*3]
    std::vector<uint32_t> mi_poff;   // [npts+1]
    std::vector<int32_t> mi_pcell;
    std::vector<double> mi_pw;
};

// The geometry of macroInterpolation for mesh m into g (g.cc set).
int build_interp(const ugx_mesh_desc* m, MeshGeom& g);

int build_geometry(const ugx_mesh_desc* m, const std::vector<ugx_boundary_desc>& bnd, int ns,
                   const double* mass, const ugx_model_desc& model, MeshGeom& g);

struct CaseSpec {
    MeshGeom geom;
    std::vector<ugx_species_desc> species;
    int ns = 0;
    ugx_model_desc model{};
    uint64_t seed = 0;
    int cell_begin = 0;
};

struct World {
    int precision = UGX_FP64;
    int device = 0;
    hipStream_t stream = nullptr;
    int32_t nc = 0;
    uint32_t cap = 0;
    // upper bound of the simulator count at the start of the next step: the
    // count read at the last synchronisation plus the most inflow since
    // (valid only when n_bound_valid)
    uint64_t n_bound = 0;
    bool n_bound_valid = false;
    int cur = 0;
    void* pbuf[2][9] = {};
    void* d_plane = nullptr;
    void* d_delta = nullptr;
    int32_t* d_nbr = nullptr;
    BModel* d_bm = nullptr;
    InflowFS* d_inflow = nullptr;
    int32_t n_inflow = 0;
    uint32_t max_ins = 0;
    CaseDev* d_cases = nullptr;
    int32_t* d_cell_case = nullptr;
    double* d_vol = nullptr;
    double* d_smax = nullptr;
    int32_t* d_cmodel = nullptr;
    double* d_samp = nullptr;
    double* d_uspq = nullptr;
    double* d_usps = nullptr;
    double* d_uspmax = nullptr;
    double* d_knacc = nullptr;
    double* d_kn = nullptr;
    double* d_cc = nullptr;
    double* d_bblo = nullptr;     // [nc*3] per cell bounding box in the cell frame (fill)
    double* d_bbhi = nullptr;
    double* d_mass = nullptr;
    int32_t* d_rotdof = nullptr;
    double* d_omega = nullptr;
    double* d_diam = nullptr;
    KnFace* d_knf = nullptr;
    double2* d_knaw = nullptr;
    int8_t* d_fkind = nullptr;
    double* d_kw = nullptr;
    double* d_kw2 = nullptr;
    int32_t* d_ids2 = nullptr;
    uint32_t* d_nb_off = nullptr;    // hybrid refinement neighbourhoods (CSR)
    uint32_t* d_nb_list = nullptr;
    int build_neighbourhoods();
    int32_t has_hybrid = 0;
    double* d_muref = nullptr;
    double* d_prterm = nullptr;
    int32_t* d_local = nullptr;
    double *d_pmr = nullptr, *d_pC = nullptr, *d_pe = nullptr, *d_pchi = nullptr, *d_palpha = nullptr;
    uint32_t* d_cnt = nullptr;
    uint32_t* d_ocnt = nullptr;
    uint32_t* d_out7 = nullptr;
    uint32_t* d_pre = nullptr;
    int8_t* d_rev = nullptr;
    uint32_t* d_off = nullptr;
    uint32_t* d_perm = nullptr;
    uint32_t* d_ins_off = nullptr;
    uint32_t* d_scratch = nullptr;
    uint32_t* d_tile_first = nullptr;
    uint32_t* d_tile_p0 = nullptr;
    int32_t* d_tile_case = nullptr;
    int32_t nsp_total = 0;
    double* d_mom = nullptr;
    BgkP* d_bgkp = nullptr;
    unsigned long long* d_raise = nullptr;
    uint32_t* d_sel_slot = nullptr;
    int32_t ns_max = 1, sum_level = 0, has_collide = 0, has_bgk = 0;
    uint32_t* d_blocksum = nullptr;
    unsigned long long* d_counters = nullptr;
    DevState* d_st = nullptr;
    uint32_t* d_wq_slot = nullptr;
    uint32_t* d_wq_info = nullptr;
    double* d_wq_tau = nullptr;
    uint32_t wq_cap = 0;
    // wall measurements: world wall faces (each case's in a contiguous
    // range), the event buffers of a step and the per-face time sums
    int32_t nwf = 0;
    std::vector<int32_t> wf_begin;   // [ncases + 1]
    int32_t* d_wface = nullptr;
    double* d_wgeo = nullptr;
    int32_t* d_wcell = nullptr;
    WallEv* d_ev = nullptr;
    unsigned long long* d_ev_key = nullptr;
    uint32_t ev_cap = 0;
    uint32_t* d_wcnt = nullptr;
    uint32_t* d_woff = nullptr;
    unsigned long long* d_bkey = nullptr;
    uint32_t* d_bidx = nullptr;
    double* d_wsamp = nullptr;
    int alloc_events(uint32_t n);
    // cell weighting: per-cell weights (device and host copies; 1 unless a
    // weighted case's fill or an upload set them), and with any weighted
    // case the per-slot start cells, clone counts and offsets
    int32_t has_weighting = 0;
    double* d_cw = nullptr;
    std::vector<double> cw_host;
    uint32_t* d_cstart = nullptr;
    uint32_t* d_wclone = nullptr;
    uint32_t* d_wcoff = nullptr;
    uint32_t* d_wbsum = nullptr;
    // sets the weights of case ci's cells (ncl values) and the bound of a step's inflow
    int set_cell_weights(int ci, const double* w);
    void update_max_ins();
    // NTC subcells: packed levels per cell (x | y << 8 | z << 16; host and
    // device), and once any cell is split the subcell tables and the per-slot
    // subcell ids and lists
    int32_t has_subcells = 0;
    std::vector<uint32_t> sclev_host;
    uint32_t* d_sclev = nullptr;
    uint32_t* d_scoff = nullptr;
    uint32_t* d_sctab = nullptr;
    uint32_t* d_scid = nullptr;
    uint32_t* d_sclist = nullptr;
    // sets the levels of case ci's cells ([ncl*3], each in 1..10, 1 in empty directions)
    int set_subcell_levels(int ci, const int32_t* lv);
    // the subcell tables and per-slot buffers for sclev_host
    int upload_subcells();
    // adaptive simulation: the adapter's per-cell time sums (device), per
    // cell the size ratio of the last adaptation, per case the time summed
    // into the sums since then; per case the time of the sampled steps
    // since the last sampler reset (every case)
    int32_t has_adaptive = 0;
    double* d_adacc = nullptr;
    std::vector<double> adapt_ratio;    // [nc*3]
    std::vector<double> adapt_time;     // [ncases]
    std::vector<double> tsamp;          // [ncases]
    // the adaptation of case ci at the end of a step
    int adapt_case(int ci);
    // the adapter's start from case ci's fill (before the fill's weights)
    int adapt_fill(int ci, const ugx_fill_desc* f);
    // sets case ci's time step and the inflow counts that depend on it
    int set_time_step(int ci, double dt);
    // applies an adapter's result to case ci
    int apply_adaptation(int ci, const struct AdaptOut& a);
    // every simulator of case ci from its cell's weight in old_w to the
    // current one, then sorted again
    int reweight(int ci, const std::vector<double>& old_w);
    double* d_cw_old = nullptr;         // [nc] the weights before an adaptation
    // macroInterpolation: the cases' interpolation geometry (world-wide ids)
    // and the step's cell and point fields
    int32_t has_macro = 0;
    uint32_t npts = 0;
    double* d_mi_cent = nullptr;
    int32_t* d_mi_fp = nullptr;
    double* d_mi_pts = nullptr;
    uint32_t* d_mi_poff = nullptr;
    int32_t* d_mi_pcell = nullptr;
    double* d_mi_pw = nullptr;
    double* d_mfield = nullptr;
    double* d_pfield = nullptr;

    // per-phase HIP-event timing (off unless enabled)
    bool timing = false;
    // HIP-graph replay of whole steps: one executable graph per buffer parity
    // (the step's kernels and arguments are otherwise fixed for a world)
    bool use_graphs = false;
    // autotuner: each of the first kTileVariants steps times one cell-tile
    // configuration on the live case (results do not depend on it)
    bool autotune = true;
    int tile_variant = 0;
    uint32_t move_grid = 0;       // resident blocks of the move kernel
    int tune_step = 0;
    double tune_ms[kTileVariants] = {};
    // FP32 worlds with BGK-family cells and no macroInterpolation: the
    // conservation kernel's occupancy, re-timed every kConsRetune steps
    static constexpr uint64_t kConsRetune = 256;
    int cons_variant = 0;
    double cons_ms[kConsVariants] = {};
    uint64_t steps_done = 0;
    std::vector<hipEvent_t> tune_ev;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    void drop_graphs();
    std::vector<hipEvent_t> tev;
    int tev_used = 0;
    double phase_ms[kPhases] = {};
    int64_t timed_steps = 0;
    int flush_timing();

    std::vector<CaseDev> cases;          // host mirror (step advanced with the device)
    std::vector<CaseSpec> specs_host;
    std::vector<int64_t> samp_reset_steps;

    ~World();
    void release();
    size_t rsize() const;
    int build(std::vector<CaseSpec>& specs, int precision, int device, uint64_t capacity);
    int alloc_particles(uint32_t newcap);
    int ensure_capacity(uint64_t need);
    template <class R> WorldView<R> view() const;
    int check_state();
    int count(uint32_t& n);
    int sort_now();
    int step(int n);
    int upload(int ci, const ugx_particle_soa* p, const MeshGeom& g);
    // Scratch of a fill supplied by the caller: the fill's parameters
    // (kFillParams doubles), the per (cell, species) counts and their
    // offsets (ncl * ns + 1 words each).  With it the fill allocates
    // nothing and never grows the particle capacity.
    struct FillScratch {
        double* params;
        uint32_t* fcnt;
        uint32_t* foff;
    };
    int fill(int ci, const ugx_fill_desc* f, const MeshGeom& g, bool sort = true, const FillScratch* fixed = nullptr);
    // Copies the state of the cases of `old` into this freshly built world:
    // new case i takes old case old_of_new[i] (or starts empty if < 0).
    int adopt(World& old, const std::vector<int>& old_of_new);
    CellState cell_state();
    // Stops a case in place (its simulators are kept but no longer move,
    // collide or sample); a batch retires cases this way until its next rebuild.
    int pause(int ci);
    int download(int ci, ugx_particle_soa* p, const MeshGeom& g);
    int case_counters(int ci, ugx_counters* out);
    int case_samples(int ci, double* out, int64_t len);
    int case_samples_reset(int ci);
    // wall sums of case ci, [face][kNS][kWRows]: out to `buf`, or in from it
    int case_wall_samples(int ci, double* buf, int64_t len, bool upload);
    // restart state of case ci: out to s, or in from s
    int case_state(int ci, ugx_case_state* s, bool upload);
    int hybrid_update();
};

// Write-time fields of uniGasVolFields from raw sums (ugx_fields.cpp).
// cw: the case's cell weights (F_N W_c molecules per simulator).
int compute_fields(const CaseSpec& spec, const double* samp, int ncl, const ugx_counters& cnt, const double* cw,
                   ugx_fields* f);
// Write-time wall values of uniGasVolFields from the raw wall sums.
int compute_wall_fields(const CaseSpec& spec, const double* wsamp, const ugx_counters& cnt, const double* cw,
                        ugx_wall_fields* f);
// Cell weights of a weighted case's fill (uniGasMeshFill.C:111-168): the
// total number density ntot spread at particles_per_subcell simulators per
// cell, smoothed until neighbour ratios stay below 1 + max_cell_weight_ratio,
// never below minParticlesPerSubCell's weight (ugx_weights.cpp).
// lv: packed subcell levels of the case's cells (nullptr: one subcell each).
int fill_cell_weights(const CaseSpec& spec, double ntot, const uint32_t* lv, std::vector<double>& w);

// Result of the dynamic adapter for one case (ugx_adapt.cpp).
struct AdaptOut {
    double dt = 0.0;                  // the time step from now on
    std::vector<double> ratio;        // [nc*3] smoothed cell size to mean free path ratio
    std::vector<uint32_t> levels;     // packed subcell levels; empty: unchanged
    std::vector<double> cw;           // new cell weights (cell weight adaptation); empty: unchanged
};
// At the fill: the filled gas (number densities nd[ns], temperature T,
// velocity U) with the case's initial model ids.
int adapt_initial(const CaseSpec& spec, const double* nd, double T, const double U[3], const int32_t* cmodel,
                  AdaptOut& out);
// Every adaptation interval: acc the adapter's time sums [nc*kKnRows] over
// the time tav, cw the cell weights, cmodel the model ids, dt the current
// time step, prev [nc*3] the size ratio of the last adaptation.
int adapt_step(const CaseSpec& spec, const double* acc, double tav, const double* cw, const int32_t* cmodel,
               double dt, const double* prev, const uint32_t* levels, AdaptOut& out);
// fvc::smooth(f, coeff): the smallest field >= f with no cell below a face
// neighbour's value over 1 + coeff (internal and cyclic faces).
void smooth_max_ratio(const MeshGeom& g, std::vector<double>& f, double max_ratio);

}  // namespace ugx

// Opaque handle of the C ABI: a world plus the index of the case in it.
struct ugx_case {
    ugx::World* world = nullptr;
    int index = 0;
    bool owns_world = false;
    ugx_batch* batch = nullptr;   // set for the handle of a batch slot
    int slot = -1;
};

namespace ugx {
// Points a batch-slot handle at its current world and case index; fails
// while the slot has not entered the world yet.
int resolve_case(ugx_case* c);
// CaseSpec from the C descriptors (ugx_api.cpp).
int make_spec(const ugx_mesh_desc* mesh, const ugx_species_desc* species, int n_species,
              const ugx_model_desc* model, const ugx_boundary_desc* bnd, int n_bnd, uint64_t seed, CaseSpec& spec);
}  // namespace ugx
