// ugx_batch.cpp - batched and served forms: many independent cases in one
// device world.  A batch keeps its admitted cases in slots; admissions and
// retirements mark the batch dirty, and the next ugx_batch_step builds a new
// world from the live slots, carries every remaining case's device state
// into it (World::adopt), fills the newly admitted cases and sorts once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstring>
#include <memory>
#include <new>
#include <vector>

#include "ugx_world.hpp"

using namespace ugx;

struct ugx_batch {
    struct Slot {
        bool used = false;       // admitted and not retired
        bool live = false;       // present in the current world
        CaseSpec spec;
        ugx_fill_desc fill{};
        bool has_fill = false;
        int64_t expected = 0;
        int64_t bytes = 0;
        int world_index = -1;
        std::unique_ptr<ugx_case> handle;
    };
    ugx_batch_options opt{};
    int64_t budget = 0;
    int64_t used_bytes = 0;
    int64_t paused_particles = 0;   // simulators of retired cases still in the world
    std::vector<Slot> slots;
    World* world = nullptr;
    bool dirty = false;
    int64_t rebuilds = 0;
    ~ugx_batch() { delete world; }
};

namespace ugx {

int resolve_case(ugx_case* c)
{
    ugx_batch* b = c->batch;
    if (c->slot < 0 || c->slot >= (int)b->slots.size() || !b->slots[c->slot].used)
        return set_err(UGX_ERR_STATE, "batch slot %d is not in use", c->slot);
    const auto& s = b->slots[c->slot];
    if (!s.live || !b->world) return set_err(UGX_ERR_STATE, "batch slot %d enters the batch at the next ugx_batch_step", c->slot);
    c->world = b->world;
    c->index = s.world_index;
    return UGX_OK;
}

}  // namespace ugx

static int64_t footprint(int nc, int precision, int64_t particles, bool bgk, bool hybrid)
{
    const int64_t R = precision == UGX_FP32_STATE ? 4 : 8;
    // geometry (planes, shifts, neighbours), per-cell FP64 state, offsets
    int64_t per_cell = kMaxFaces * (8 * R + 4) + 8 * (kNSamp + 1 + 3 + 6 + 1 + kKnRows + 1 + 2 + 3) + 4 * 5;
    if (bgk) per_cell += 8 * kNMom + (int64_t)sizeof(BgkP) + 8;
    if (hybrid) per_cell += kMaxFaces * ((int64_t)sizeof(KnFace) + 1) + 8 * 16 + 4;
    // two particle buffers, permutation, scratch, selections, wall queue; twice the expected count
    const int64_t per_particle = 2 * (7 * R + 8) + 4 + 4 + 12 + 8;
    return (int64_t)nc * per_cell + 2 * std::max<int64_t>(particles, 1024) * per_particle + (1 << 20);
}

static int rebuild(ugx_batch* b)
{
    std::vector<CaseSpec> specs;
    std::vector<int> old_of_new, slot_of_new;
    uint64_t cap = 0;
    for (size_t i = 0; i < b->slots.size(); ++i) {
        auto& s = b->slots[i];
        if (!s.used) continue;
        specs.push_back(s.spec);
        old_of_new.push_back(s.live ? s.world_index : -1);
        slot_of_new.push_back((int)i);
        cap += (uint64_t)(2 * std::max<int64_t>(s.expected, 1024));
    }
    World* nw = nullptr;
    if (!specs.empty()) {
        nw = new World();
        nw->device = b->opt.device;
        if (int rc = nw->build(specs, b->opt.precision, b->opt.device, cap)) { delete nw; return rc; }
        nw->use_graphs = b->opt.use_graphs != 0;
        if (b->world && b->world->tune_step >= 2 * kTileVariants) {
            // keep the tuned cell-tile configuration across rebuilds
            nw->tile_variant = b->world->tile_variant;
            nw->tune_step = b->world->tune_step;
        }
        if (b->world) {
            if (int rc = nw->adopt(*b->world, old_of_new)) { delete nw; return rc; }
        }
        bool filled = false;
        for (size_t i = 0; i < specs.size(); ++i) {
            auto& s = b->slots[slot_of_new[i]];
            if (old_of_new[i] < 0 && s.has_fill) {
                if (int rc = nw->fill((int)i, &s.fill, nw->specs_host[i].geom, false)) { delete nw; return rc; }
                filled = true;
            }
        }
        if (filled || b->world)
            if (int rc = nw->sort_now()) { delete nw; return rc; }
        if (hipStreamSynchronize(nw->stream) != hipSuccess) { delete nw; return set_err(UGX_ERR_DEVICE, "batch rebuild failed"); }
    }
    delete b->world;
    b->world = nw;
    for (auto& s : b->slots) { s.live = false; s.world_index = -1; }
    for (size_t i = 0; i < slot_of_new.size(); ++i) {
        auto& s = b->slots[slot_of_new[i]];
        s.live = true;
        s.world_index = (int)i;
    }
    b->dirty = false;
    b->paused_particles = 0;
    b->rebuilds += 1;
    return UGX_OK;
}

#define UGX_GUARD_BEGIN try {
#define UGX_GUARD_END                                                           \
    }                                                                           \
    catch (const std::bad_alloc&) { return set_err(UGX_ERR_DEVICE, "host allocation failed"); } \
    catch (...) { return set_err(UGX_ERR_STATE, "unexpected exception"); }

extern "C" {

int ugx_batch_create(const ugx_batch_options* options, ugx_batch** out)
{
    UGX_GUARD_BEGIN
    if (!out) return set_err(UGX_ERR_ARG, "out is null");
    *out = nullptr;
    ugx_batch_options opt{};
    if (options) opt = *options;
    if (opt.precision != UGX_FP64 && opt.precision != UGX_FP32_STATE) return set_err(UGX_ERR_ARG, "unknown precision");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_err(UGX_ERR_DEVICE, "no HIP device available");
    if (opt.device < 0 || opt.device >= ndev) return set_err(UGX_ERR_ARG, "device %d out of range", opt.device);
    if (hipSetDevice(opt.device) != hipSuccess) return set_err(UGX_ERR_DEVICE, "hipSetDevice failed");
    auto* b = new ugx_batch();
    b->opt = opt;
    b->budget = opt.hbm_budget;
    if (b->budget <= 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) { delete b; return set_err(UGX_ERR_DEVICE, "hipMemGetInfo failed"); }
        // the rebuild holds the old and the new world at once
        b->budget = (int64_t)(0.3 * (double)fr);
    }
    *out = b;
    return UGX_OK;
    UGX_GUARD_END
}

int ugx_batch_destroy(ugx_batch* b)
{
    if (!b) return UGX_OK;
    (void)hipSetDevice(b->opt.device);
    delete b;
    return UGX_OK;
}

int ugx_case_footprint(const ugx_mesh_desc* mesh, const ugx_model_desc* model, int64_t particles, int precision,
                       int64_t* bytes)
{
    if (!mesh || !model || !bytes) return set_err(UGX_ERR_ARG, "null argument");
    const bool hybrid = model->collision_model == UGX_COLL_HYBRID;
    const bool bgk = (model->collision_model == UGX_COLL_BGK || hybrid) && model->bgk_model != UGX_BGK_NONE;
    *bytes = footprint(mesh->n_cells, precision, particles, bgk, hybrid);
    // cell weighting: start cell, clone count and offset per simulator slot, weight per cell
    if (model->cell_weighted) *bytes += (int64_t)(12.0 * 1.5 * (double)particles) + 8 * (int64_t)mesh->n_cells;
    // adaptive simulation: the adapter's time sums per cell; subcell ids and lists per simulator slot
    if (model->adaptive) *bytes += 8 * UGX_KN_ROWS * (int64_t)mesh->n_cells + (int64_t)(8.0 * 1.5 * (double)particles);
    return UGX_OK;
}

int ugx_batch_admit(ugx_batch* b, const ugx_mesh_desc* mesh, const ugx_species_desc* species, int n_species,
                    const ugx_model_desc* model, const ugx_boundary_desc* boundaries, int n_boundaries,
                    const ugx_fill_desc* fill, uint64_t seed, int64_t expected_particles, int* slot)
{
    UGX_GUARD_BEGIN
    if (!b || !slot) return set_err(UGX_ERR_ARG, "null argument");
    *slot = -1;
    if (expected_particles < 0) return set_err(UGX_ERR_ARG, "expected_particles must be >= 0");
    if (fill)
        for (int d = 0; d < 3; ++d)
            if (!(fill->T[d] > 0.0)) return set_err(UGX_ERR_ARG, "fill temperature must be > 0");
    int64_t bytes = 0;
    if (int rc = ugx_case_footprint(mesh, model, expected_particles, b->opt.precision, &bytes)) return rc;
    int n_live = 0;
    for (const auto& s : b->slots) n_live += s.used ? 1 : 0;
    if (b->opt.max_cases > 0 && n_live >= b->opt.max_cases)
        return set_err(UGX_ERR_CAPACITY, "batch holds its maximum of %d cases", b->opt.max_cases);
    if (b->used_bytes + bytes > b->budget)
        return set_err(UGX_ERR_CAPACITY, "case needs %lld bytes, %lld of %lld left", (long long)bytes,
                       (long long)(b->budget - b->used_bytes), (long long)b->budget);
    ugx_batch::Slot s;
    if (int rc = make_spec(mesh, species, n_species, model, boundaries, n_boundaries, seed, s.spec)) return rc;
    s.used = true;
    s.has_fill = fill != nullptr;
    if (fill) s.fill = *fill;
    s.expected = expected_particles;
    s.bytes = bytes;
    int idx = -1;
    for (size_t i = 0; i < b->slots.size(); ++i)
        if (!b->slots[i].used && !b->slots[i].live) { idx = (int)i; break; }
    if (idx < 0) { idx = (int)b->slots.size(); b->slots.emplace_back(); }
    b->slots[idx] = std::move(s);
    auto h = std::make_unique<ugx_case>();
    h->batch = b;
    h->slot = idx;
    b->slots[idx].handle = std::move(h);
    b->used_bytes += bytes;
    b->dirty = true;
    *slot = idx;
    return UGX_OK;
    UGX_GUARD_END
}

int ugx_batch_retire(ugx_batch* b, int slot)
{
    UGX_GUARD_BEGIN
    if (!b || slot < 0 || slot >= (int)b->slots.size() || !b->slots[slot].used)
        return set_err(UGX_ERR_ARG, "no case in slot %d", slot);
    auto& s = b->slots[slot];
    s.used = false;
    s.handle.reset();
    b->used_bytes -= s.bytes;
    if (s.live && b->world) {
        // stop it in place; its memory is reclaimed at the next rebuild
        ugx_counters k{};
        if (int rc = b->world->case_counters(s.world_index, &k)) return rc;
        b->paused_particles += k.n_particles;
        if (int rc = b->world->pause(s.world_index)) return rc;
    }
    return UGX_OK;
    UGX_GUARD_END
}

int ugx_batch_step(ugx_batch* b, int n_steps)
{
    UGX_GUARD_BEGIN
    if (!b) return set_err(UGX_ERR_ARG, "null batch");
    if (n_steps < 0) return set_err(UGX_ERR_ARG, "n_steps must be >= 0");
    if (hipSetDevice(b->opt.device) != hipSuccess) return set_err(UGX_ERR_DEVICE, "hipSetDevice failed");
    // rebuild for waiting admissions, or when stopped cases hold most of the simulators
    if (!b->dirty && b->world && b->paused_particles > 0) {
        uint32_t n = 0;
        if (int rc = b->world->count(n)) return rc;
        if (2 * b->paused_particles > (int64_t)n) b->dirty = true;
    }
    if (b->dirty)
        if (int rc = rebuild(b)) return rc;
    if (!b->world || n_steps == 0) return UGX_OK;
    return b->world->step(n_steps);
    UGX_GUARD_END
}

int ugx_batch_sync(ugx_batch* b)
{
    if (!b) return set_err(UGX_ERR_ARG, "null batch");
    if (!b->world) return UGX_OK;
    if (hipSetDevice(b->opt.device) != hipSuccess) return set_err(UGX_ERR_DEVICE, "hipSetDevice failed");
    if (hipStreamSynchronize(b->world->stream) != hipSuccess) return set_err(UGX_ERR_DEVICE, "stream synchronize failed");
    return b->world->check_state();
}

int ugx_batch_case(ugx_batch* b, int slot, ugx_case** out)
{
    if (!b || !out) return set_err(UGX_ERR_ARG, "null argument");
    if (slot < 0 || slot >= (int)b->slots.size() || !b->slots[slot].used) return set_err(UGX_ERR_ARG, "no case in slot %d", slot);
    *out = b->slots[slot].handle.get();
    return UGX_OK;
}

int ugx_batch_info(ugx_batch* b, int64_t* n_cases, int64_t* bytes_used, int64_t* n_particles)
{
    if (!b) return set_err(UGX_ERR_ARG, "null batch");
    int64_t k = 0;
    for (const auto& s : b->slots) k += s.used ? 1 : 0;
    if (n_cases) *n_cases = k;
    if (bytes_used) *bytes_used = b->used_bytes;
    if (n_particles) {
        uint32_t n = 0;
        if (b->world)
            if (int rc = b->world->count(n)) return rc;
        *n_particles = n;
    }
    return UGX_OK;
}

}  // extern "C"
