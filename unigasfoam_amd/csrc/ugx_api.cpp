// ugx_api.cpp - the extern "C" entry points of include/ugx.h for the plain
// (one case per world) form.  Every entry point validates its arguments,
// returns a ugx_status and never lets a C++ exception cross the boundary.
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>
#include <vector>

#include "ugx_world.hpp"

using namespace ugx;

#define UGX_GUARD_BEGIN try {
#define UGX_GUARD_END                                                           \
    }                                                                           \
    catch (const std::bad_alloc&) { return set_err(UGX_ERR_DEVICE, "host allocation failed"); } \
    catch (...) { return set_err(UGX_ERR_STATE, "unexpected exception"); }

static int validate_model(const ugx_model_desc* m)
{
    if (!(m->n_equivalent_particles > 0.0)) return set_err(UGX_ERR_ARG, "nEquivalentParticles must be > 0");
    if (!(m->delta_t > 0.0)) return set_err(UGX_ERR_ARG, "deltaT must be > 0");
    if (m->collision_model < UGX_COLL_NONE || m->collision_model > UGX_COLL_HYBRID) return set_err(UGX_ERR_ARG, "unknown collisionModel");
    if (m->dsmc_model < UGX_DSMC_NONE || m->dsmc_model > UGX_DSMC_LB_VSS) return set_err(UGX_ERR_ARG, "unknown dsmcCollisionModel");
    if (m->bgk_model < UGX_BGK_NONE || m->bgk_model > UGX_BGK_USP) return set_err(UGX_ERR_ARG, "unknown bgkCollisionModel");
    if (m->partner_model != UGX_PARTNER_NTC) return set_err(UGX_ERR_ARG, "unknown dsmcCollisionPartnerModel");
    if ((m->dsmc_model != UGX_DSMC_NONE) && !(m->Tref > 0.0)) return set_err(UGX_ERR_ARG, "collisionProperties.Tref must be > 0");
    if ((m->dsmc_model == UGX_DSMC_LB_VHS || m->dsmc_model == UGX_DSMC_LB_VSS) && !(m->Zrot > 0.0))
        return set_err(UGX_ERR_ARG, "rotationalRelaxationCollisionNumber must be > 0");
    if (m->cell_weighted && m->particles_per_subcell <= 0)
        return set_err(UGX_ERR_ARG, "cellWeightedProperties.particlesPerSubCell must be > 0");
    if (m->adaptive) {
        if (m->adaptation_interval < 0) return set_err(UGX_ERR_ARG, "adaptiveProperties.adaptationInterval must be > 0");
        if (m->max_time_step_mct_ratio < 0.0 || m->max_courant_number < 0.0 || m->max_subcell_size_mfp_ratio < 0.0)
            return set_err(UGX_ERR_ARG, "adaptiveProperties ratios must be > 0");
        if (m->cell_weight_adaptation && !m->cell_weighted)
            return set_err(UGX_ERR_ARG, "cellWeightAdaptation needs cellWeightedSimulation");
    }
    return UGX_OK;
}

int ugx::make_spec(const ugx_mesh_desc* mesh, const ugx_species_desc* species, int n_species,
                   const ugx_model_desc* model, const ugx_boundary_desc* bnd, int n_bnd, uint64_t seed,
                   CaseSpec& spec)
{
    if (!mesh || !species || !model) return set_err(UGX_ERR_ARG, "null descriptor");
    if (n_species < 1 || n_species > UGX_MAX_SPECIES) return set_err(UGX_ERR_ARG, "n_species must be in 1..%d", UGX_MAX_SPECIES);
    if (n_bnd < 0 || (n_bnd > 0 && !bnd)) return set_err(UGX_ERR_ARG, "bad boundary list");
    if (int rc = validate_model(model)) return rc;
    for (int s = 0; s < n_species; ++s) {
        const auto& sd = species[s];
        if (!(sd.mass > 0.0) || !(sd.diameter > 0.0) || !(sd.alpha > 0.0) || sd.rotational_dof < 0)
            return set_err(UGX_ERR_ARG, "species %d: invalid moleculeProperties", s);
    }
    spec.species.assign(species, species + n_species);
    spec.ns = n_species;
    spec.model = *model;
    if (spec.model.sample_interval < 1) spec.model.sample_interval = 1;
    if (spec.model.theta == 0.0 && spec.model.bgk_model == UGX_BGK_USP) spec.model.theta = 1.0;
    if (spec.model.adaptive) {
        // adaptiveProperties defaults (uniGasDynamicAdapter.C:198-216)
        ugx_model_desc& a = spec.model;
        if (a.adaptation_interval == 0) a.adaptation_interval = 20;
        if (a.adaptation_smoothing_passes == 0) a.adaptation_smoothing_passes = 25;
        if (a.adaptation_smoothing_passes < 0) a.adaptation_smoothing_passes = 0;
        if (a.max_time_step_mct_ratio == 0.0) a.max_time_step_mct_ratio = 0.2;
        if (a.max_courant_number == 0.0) a.max_courant_number = 0.5;
        if (a.max_subcell_size_mfp_ratio == 0.0) a.max_subcell_size_mfp_ratio = 0.5;
    }
    spec.seed = seed;
    std::vector<ugx_boundary_desc> b(bnd, bnd + n_bnd);
    for (const auto& x : b)
        if (x.patch < 0 || x.patch >= mesh->n_patches) return set_err(UGX_ERR_ARG, "boundary patch %d out of range", x.patch);
    std::vector<double> mass(n_species);
    for (int s = 0; s < n_species; ++s) mass[s] = species[s].mass;
    return build_geometry(mesh, b, n_species, mass.data(), spec.model, spec.geom);
}


extern "C" {

const char* ugx_version(void) { return "ugx 0.1 (gfx950)"; }

const char* ugx_last_error(void) { return g_err.c_str(); }

int ugx_case_create(const ugx_mesh_desc* mesh, const ugx_species_desc* species, int n_species,
                    const ugx_model_desc* model, const ugx_boundary_desc* boundaries, int n_boundaries,
                    const ugx_options* options, ugx_case** out)
{
    UGX_GUARD_BEGIN
    if (!out) return set_err(UGX_ERR_ARG, "out is null");
    *out = nullptr;
    ugx_options opt{};
    if (options) opt = *options;
    if (opt.precision != UGX_FP64 && opt.precision != UGX_FP32_STATE) return set_err(UGX_ERR_ARG, "unknown precision");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_err(UGX_ERR_DEVICE, "no HIP device available");
    if (opt.device < 0 || opt.device >= ndev) return set_err(UGX_ERR_ARG, "device %d out of range", opt.device);
    std::vector<CaseSpec> specs(1);
    if (int rc = make_spec(mesh, species, n_species, model, boundaries, n_boundaries, opt.seed, specs[0])) return rc;
    auto* w = new World();
    w->device = opt.device;
    if (int rc = w->build(specs, opt.precision, opt.device, (uint64_t)opt.particle_capacity)) { delete w; return rc; }
    w->use_graphs = opt.use_graphs != 0;
    auto* c = new ugx_case();
    c->world = w;
    c->index = 0;
    c->owns_world = true;
    *out = c;
    return UGX_OK;
    UGX_GUARD_END
}

int ugx_case_destroy(ugx_case* c)
{
    if (!c) return UGX_OK;
    if (c->batch) return set_err(UGX_ERR_STATE, "a batch case is released by ugx_batch_retire");
    if (c->owns_world) delete c->world;
    delete c;
    return UGX_OK;
}

#define CHECK_CASE(c) \
    if (!(c)) return set_err(UGX_ERR_ARG, "null case"); \
    if ((c)->batch) { if (int rc_ = resolve_case(c)) return rc_; } \
    if (!(c)->world) return set_err(UGX_ERR_ARG, "null case"); \
    if (hipSetDevice((c)->world->device) != hipSuccess) return set_err(UGX_ERR_DEVICE, "hipSetDevice failed")

static int drain(ugx_case* c)
{
    if (hipStreamSynchronize(c->world->stream) != hipSuccess) return set_err(UGX_ERR_DEVICE, "stream synchronize failed");
    return UGX_OK;
}

int ugx_case_fill(ugx_case* c, const ugx_fill_desc* fill)
{
    UGX_GUARD_BEGIN
    CHECK_CASE(c);
    if (c->batch) return set_err(UGX_ERR_STATE, "a batch case is filled at ugx_batch_admit");
    if (!fill) return set_err(UGX_ERR_ARG, "null fill descriptor");
    for (int d = 0; d < 3; ++d) if (!(fill->T[d] > 0.0)) return set_err(UGX_ERR_ARG, "fill temperature must be > 0");
    return c->world->fill(c->index, fill, c->world->specs_host[c->index].geom);
    UGX_GUARD_END
}

int ugx_particles_upload(ugx_case* c, const ugx_particle_soa* p)
{
    UGX_GUARD_BEGIN
    CHECK_CASE(c);
    if (c->batch) return set_err(UGX_ERR_STATE, "particles of a batch case come from its admission fill");
    if (!p || p->n < 0) return set_err(UGX_ERR_ARG, "bad particle arrays");
    if (p->n == 0) return UGX_OK;
    if ((!p->position && !p->local_position) || !p->velocity || !p->cell) return set_err(UGX_ERR_ARG, "position, velocity and cell are required");
    return c->world->upload(c->index, p, c->world->specs_host[c->index].geom);
    UGX_GUARD_END
}

int ugx_particles_download(ugx_case* c, ugx_particle_soa* p)
{
    UGX_GUARD_BEGIN
    CHECK_CASE(c);
    if (!p) return set_err(UGX_ERR_ARG, "null particle arrays");
    return c->world->download(c->index, p, c->world->specs_host[c->index].geom);
    UGX_GUARD_END
}

int ugx_particle_count(ugx_case* c, int64_t* n)
{
    UGX_GUARD_BEGIN
    CHECK_CASE(c);
    if (!n) return set_err(UGX_ERR_ARG, "null output");
    ugx_counters k;
    if (int rc = c->world->case_counters(c->index, &k)) return rc;
    *n = k.n_particles;
    return UGX_OK;
    UGX_GUARD_END
}

int ugx_step(ugx_case* c, int n_steps)
{
    UGX_GUARD_BEGIN
    CHECK_CASE(c);
    if (c->batch) return set_err(UGX_ERR_STATE, "a batch case advances with ugx_batch_step");
    if (n_steps < 0) return set_err(UGX_ERR_ARG, "n_steps must be >= 0");
    return c->world->step(n_steps);
    UGX_GUARD_END
}

int ugx_sync(ugx_case* c)
{
    UGX_GUARD_BEGIN
    CHECK_CASE(c);
    if (hipStreamSynchronize(c->world->stream) != hipSuccess) return set_err(UGX_ERR_DEVICE, "stream synchronize failed");
    return c->world->check_state();
    UGX_GUARD_END
}

int ugx_samples_download(ugx_case* c, double* sums, int64_t len)
{
    UGX_GUARD_BEGIN
    CHECK_CASE(c);
    if (!sums) return set_err(UGX_ERR_ARG, "null output");
    return c->world->case_samples(c->index, sums, len);
    UGX_GUARD_END
}

int ugx_samples_reset(ugx_case* c)
{
    UGX_GUARD_BEGIN
    CHECK_CASE(c);
    return c->world->case_samples_reset(c->index);
    UGX_GUARD_END
}

int ugx_counters_get(ugx_case* c, ugx_counters* out)
{
    UGX_GUARD_BEGIN
    CHECK_CASE(c);
    if (!out) return set_err(UGX_ERR_ARG, "null output");
    return c->world->case_counters(c->index, out);
    UGX_GUARD_END
}

int ugx_fields_compute(ugx_case* c, ugx_fields* f)
{
    UGX_GUARD_BEGIN
    CHECK_CASE(c);
    if (!f) return set_err(UGX_ERR_ARG, "null output");
    const CaseSpec& spec = c->world->specs_host[c->index];
    int ncl = spec.geom.nc;
    std::vector<double> s((size_t)UGX_S_COUNT * ncl);
    if (int rc = c->world->case_samples(c->index, s.data(), (int64_t)s.size())) return rc;
    ugx_counters k;
    if (int rc = c->world->case_counters(c->index, &k)) return rc;
    return compute_fields(spec, s.data(), ncl, k, c->world->cw_host.data() + c->world->cases[c->index].cell_begin, f);
    UGX_GUARD_END
}

static int cell_range(ugx_case* c, int64_t len, int per, int& b, int& n)
{
    const CaseDev& cd = c->world->cases[c->index];
    b = cd.cell_begin;
    n = cd.cell_end - cd.cell_begin;
    if (len < (int64_t)n * per) return set_err(UGX_ERR_ARG, "buffer holds %lld, need %lld", (long long)len, (long long)n * per);
    return UGX_OK;
}

int ugx_sigma_max_download(ugx_case* c, double* out, int64_t len)
{
    UGX_GUARD_BEGIN
    CHECK_CASE(c);
    int b, n;
    if (int rc = cell_range(c, len, 1, b, n)) return rc;
    if (int rc = drain(c)) return rc;
    if (hipMemcpy(out, c->world->d_smax + b, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess)
        return set_err(UGX_ERR_DEVICE, "download failed");
    return UGX_OK;
    UGX_GUARD_END
}

int ugx_sigma_max_upload(ugx_case* c, const double* in, int64_t len)
{
    UGX_GUARD_BEGIN
    CHECK_CASE(c);
    int b, n;
    if (int rc = cell_range(c, len, 1, b, n)) return rc;
    if (int rc = drain(c)) return rc;
    if (hipMemcpy(c->world->d_smax + b, in, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess)
        return set_err(UGX_ERR_DEVICE, "upload failed");
    return UGX_OK;
    UGX_GUARD_END
}

int ugx_state_download(ugx_case* c, ugx_case_state* s)
{
    UGX_GUARD_BEGIN
    CHECK_CASE(c);
    if (!s) return set_err(UGX_ERR_ARG, "null state");
    return c->world->case_state(c->index, s, false);
    UGX_GUARD_END
}

int ugx_state_upload(ugx_case* c, const ugx_case_state* s)
{
    UGX_GUARD_BEGIN
    CHECK_CASE(c);
    if (c->batch) return set_err(UGX_ERR_STATE, "the state of a batch case is carried by the batch");
    if (!s) return set_err(UGX_ERR_ARG, "null state");
    return c->world->case_state(c->index, const_cast<ugx_case_state*>(s), true);
    UGX_GUARD_END
}

int ugx_cell_model_download(ugx_case* c, int32_t* out, int64_t len)
{
    UGX_GUARD_BEGIN
    CHECK_CASE(c);
    int b, n;
    if (int rc = cell_range(c, len, 1, b, n)) return rc;
    if (int rc = drain(c)) return rc;
    if (hipMemcpy(out, c->world->d_cmodel + b, sizeof(int32_t) * n, hipMemcpyDeviceToHost) != hipSuccess)
        return set_err(UGX_ERR_DEVICE, "download failed");
    return UGX_OK;
    UGX_GUARD_END
}

int ugx_kn_download(ugx_case* c, double* out, int64_t len)
{
    UGX_GUARD_BEGIN
    CHECK_CASE(c);
    int b, n;
    if (int rc = cell_range(c, len, 1, b, n)) return rc;
    if (int rc = drain(c)) return rc;
    if (hipMemcpy(out, c->world->d_kn + b, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess)
        return set_err(UGX_ERR_DEVICE, "download failed");
    return UGX_OK;
    UGX_GUARD_END
}

int ugx_cell_volumes(ugx_case* c, double* out, int64_t len)
{
    UGX_GUARD_BEGIN
    CHECK_CASE(c);
    const MeshGeom& g = c->world->specs_host[c->index].geom;
    if (len < g.nc) return set_err(UGX_ERR_ARG, "buffer too small");
    std::memcpy(out, g.vol.data(), sizeof(double) * g.nc);
    return UGX_OK;
    UGX_GUARD_END
}

int ugx_cell_centres(ugx_case* c, double* out, int64_t len)
{
    UGX_GUARD_BEGIN
    CHECK_CASE(c);
    const MeshGeom& g = c->world->specs_host[c->index].geom;
    if (len < 3 * (int64_t)g.nc) return set_err(UGX_ERR_ARG, "buffer too small");
    std::memcpy(out, g.cc.data(), sizeof(double) * 3 * g.nc);
    return UGX_OK;
    UGX_GUARD_END
}

int ugx_cell_weights_download(ugx_case* c, double* out, int64_t len)
{
    UGX_GUARD_BEGIN
    CHECK_CASE(c);
    if (!out) return set_err(UGX_ERR_ARG, "null output");
    int b, n;
    if (int rc = cell_range(c, len, 1, b, n)) return rc;
    std::memcpy(out, c->world->cw_host.data() + b, sizeof(double) * n);
    return UGX_OK;
    UGX_GUARD_END
}

int ugx_cell_weights_upload(ugx_case* c, const double* in, int64_t len)
{
    UGX_GUARD_BEGIN
    CHECK_CASE(c);
    if (c->batch) return set_err(UGX_ERR_STATE, "the weights of a batch case are carried by the batch");
    if (!in) return set_err(UGX_ERR_ARG, "null input");
    int b, n;
    if (int rc = cell_range(c, len, 1, b, n)) return rc;
    return c->world->set_cell_weights(c->index, in);
    UGX_GUARD_END
}

int ugx_time_get(ugx_case* c, double* delta_t, double* time)
{
    UGX_GUARD_BEGIN
    CHECK_CASE(c);
    const CaseDev& cd = c->world->cases[c->index];
    if (delta_t) *delta_t = cd.dt;
    if (time) *time = cd.time;
    return UGX_OK;
    UGX_GUARD_END
}

int ugx_subcell_levels_download(ugx_case* c, int32_t* out, int64_t len)
{
    UGX_GUARD_BEGIN
    CHECK_CASE(c);
    if (!out) return set_err(UGX_ERR_ARG, "null output");
    int b, n;
    if (int rc = cell_range(c, len, 3, b, n)) return rc;
    for (int i = 0; i < n; ++i) {
        const uint32_t p = c->world->sclev_host[(size_t)b + i];
        out[3 * (size_t)i] = (int32_t)(p & 0xffu);
        out[3 * (size_t)i + 1] = (int32_t)((p >> 8) & 0xffu);
        out[3 * (size_t)i + 2] = (int32_t)((p >> 16) & 0xffu);
    }
    return UGX_OK;
    UGX_GUARD_END
}

int ugx_subcell_levels_upload(ugx_case* c, const int32_t* in, int64_t len)
{
    UGX_GUARD_BEGIN
    CHECK_CASE(c);
    if (c->batch) return set_err(UGX_ERR_STATE, "the subcell levels of a batch case are carried by the batch");
    if (!in) return set_err(UGX_ERR_ARG, "null input");
    int b, n;
    if (int rc = cell_range(c, len, 3, b, n)) return rc;
    return c->world->set_subcell_levels(c->index, in);
    UGX_GUARD_END
}

int ugx_wall_face_count(ugx_case* c, int64_t* n)
{
    UGX_GUARD_BEGIN
    CHECK_CASE(c);
    if (!n) return set_err(UGX_ERR_ARG, "null output");
    *n = c->world->specs_host[c->index].geom.nwf;
    return UGX_OK;
    UGX_GUARD_END
}

int ugx_wall_faces(ugx_case* c, int32_t* patch, int32_t* face, int64_t len)
{
    UGX_GUARD_BEGIN
    CHECK_CASE(c);
    const MeshGeom& g = c->world->specs_host[c->index].geom;
    if (len < (int64_t)g.nwf) return set_err(UGX_ERR_ARG, "buffer holds %lld, need %d", (long long)len, g.nwf);
    if (patch && g.nwf) std::memcpy(patch, g.wpatch.data(), sizeof(int32_t) * g.nwf);
    if (face && g.nwf) std::memcpy(face, g.wlocal.data(), sizeof(int32_t) * g.nwf);
    return UGX_OK;
    UGX_GUARD_END
}

int ugx_wall_samples_download(ugx_case* c, double* sums, int64_t len)
{
    UGX_GUARD_BEGIN
    CHECK_CASE(c);
    if (!sums) return set_err(UGX_ERR_ARG, "null output");
    if (int rc = c->world->check_state()) return rc;
    return c->world->case_wall_samples(c->index, sums, len, false);
    UGX_GUARD_END
}

int ugx_wall_fields_compute(ugx_case* c, ugx_wall_fields* f)
{
    UGX_GUARD_BEGIN
    CHECK_CASE(c);
    if (!f) return set_err(UGX_ERR_ARG, "null output");
    const CaseSpec& spec = c->world->specs_host[c->index];
    std::vector<double> s((size_t)spec.geom.nwf * kNS * kWRows);
    if (int rc = c->world->check_state()) return rc;
    if (int rc = c->world->case_wall_samples(c->index, s.data(), (int64_t)s.size(), false)) return rc;
    ugx_counters k;
    if (int rc = c->world->case_counters(c->index, &k)) return rc;
    return compute_wall_fields(spec, s.data(), k, c->world->cw_host.data() + c->world->cases[c->index].cell_begin, f);
    UGX_GUARD_END
}

static const char* kPhaseNames[kPhases] = {"begin", "inflow", "move", "wall", "scan", "scatter", "cells", "collide", "end"};

const char* ugx_phase_name(int i) { return (i >= 0 && i < kPhases) ? kPhaseNames[i] : ""; }

int ugx_timing_enable(ugx_case* c, int on)
{
    UGX_GUARD_BEGIN
    CHECK_CASE(c);
    World* w = c->world;
    if (int rc = w->flush_timing()) return rc;
    w->timing = on != 0;
    for (double& v : w->phase_ms) v = 0.0;
    w->timed_steps = 0;
    return UGX_OK;
    UGX_GUARD_END
}

int ugx_timing_get(ugx_case* c, double* ms_total, int n, int64_t* steps)
{
    UGX_GUARD_BEGIN
    CHECK_CASE(c);
    World* w = c->world;
    if (int rc = w->flush_timing()) return rc;
    for (int i = 0; i < n && i < kPhases; ++i) ms_total[i] = w->phase_ms[i];
    if (steps) *steps = w->timed_steps;
    return UGX_OK;
    UGX_GUARD_END
}

int ugx_tile_config(ugx_case* c, int32_t* variant, int32_t* threads, int32_t* window, int32_t* lds_simulators,
                    int32_t* tuned)
{
    UGX_GUARD_BEGIN
    CHECK_CASE(c);
    const World* w = c->world;
    static const int32_t cfg[kTileVariants][3] = {{128, 256, 512}, {256, 512, 1024}, {64, 128, 256}};
    const int v = w->tile_variant;
    if (variant) *variant = v;
    if (threads) *threads = cfg[v][0];
    if (window) *window = cfg[v][1];
    if (lds_simulators) *lds_simulators = cfg[v][2];
    if (tuned) *tuned = (!w->autotune || w->tune_step >= kTileVariants) ? 1 : 0;
    return UGX_OK;
    UGX_GUARD_END
}

int ugx_probe_copy(int device, int64_t bytes, int reps, double* seconds_per_copy)
{
    UGX_GUARD_BEGIN
    if (bytes <= 0 || reps <= 0 || !seconds_per_copy) return set_err(UGX_ERR_ARG, "bad probe arguments");
    if (hipSetDevice(device) != hipSuccess) return set_err(UGX_ERR_DEVICE, "hipSetDevice failed");
    void *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = UGX_OK;
    bytes = bytes / 16 * 16;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess || hipStreamCreate(&s) != hipSuccess
        || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        rc = set_err(UGX_ERR_DEVICE, "probe allocation failed");
    } else {
        launch_copy_probe(a, b, (size_t)bytes, s);
        (void)hipEventRecord(e0, s);
        for (int i = 0; i < reps; ++i) launch_copy_probe(a, b, (size_t)bytes, s);
        (void)hipEventRecord(e1, s);
        float ms = 0.0f;
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            rc = set_err(UGX_ERR_DEVICE, "probe timing failed");
        else
            *seconds_per_copy = (double)ms * 1e-3 / reps;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
    UGX_GUARD_END
}

// Geometry in kernel storage form, for parity with the CPU oracle
// (no counterpart in the reference; test-facing).
int ugx_geometry(const ugx_mesh_desc* mesh, const ugx_species_desc* species, int n_species,
                 const ugx_model_desc* model, const ugx_boundary_desc* bnd, int n_bnd,
                 double* cc, double* vol, int32_t* nfaces, double* planes, int32_t* nbr, double* delta,
                 double* nbar, int64_t nbar_len, int64_t* n_inflow)
{
    UGX_GUARD_BEGIN
    CaseSpec spec;
    if (int rc = make_spec(mesh, species, n_species, model, bnd, n_bnd, 0, spec)) return rc;
    const MeshGeom& g = spec.geom;
    size_t nc = (size_t)g.nc;
    if (cc) std::memcpy(cc, g.cc.data(), sizeof(double) * 3 * nc);
    if (vol) std::memcpy(vol, g.vol.data(), sizeof(double) * nc);
    if (nfaces) std::memcpy(nfaces, g.ncf.data(), sizeof(int32_t) * nc);
    if (planes) std::memcpy(planes, g.plane.data(), sizeof(double) * 4 * kMaxFaces * nc);
    if (nbr) std::memcpy(nbr, g.nbr.data(), sizeof(int32_t) * kMaxFaces * nc);
    if (delta) std::memcpy(delta, g.delta.data(), sizeof(double) * 3 * kMaxFaces * nc);
    if (n_inflow) *n_inflow = (int64_t)g.inflow.size();
    if (nbar)
        for (size_t k = 0; k < g.inflow.size() && (int64_t)k < nbar_len; ++k) nbar[k] = g.inflow[k].nbar;
    return UGX_OK;
    UGX_GUARD_END
}

}  // extern "C"
