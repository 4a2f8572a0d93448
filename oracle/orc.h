/*
 * orc.h - CPU oracle of the uniGasFoam particle time step (FP64, OpenMP).
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker; never linked into the
 * product library.
 *
 * The oracle takes the same POD descriptors as include/ugx.h and follows the
 * reference functions cited in orc.c, one function per reference routine.
 */
#ifndef ORC_H
#define ORC_H

#include "../include/ugx.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_world orc_world;

const char* orc_last_error(void);

int  orc_create(const ugx_mesh_desc* mesh, const ugx_species_desc* species, int n_species,
                const ugx_model_desc* model, const ugx_boundary_desc* bnd, int n_bnd,
                uint64_t seed, int64_t capacity, orc_world** out);
void orc_destroy(orc_world* w);

int  orc_fill(orc_world* w, const ugx_fill_desc* fill);
int  orc_upload(orc_world* w, const ugx_particle_soa* p);
int  orc_download(orc_world* w, ugx_particle_soa* p);
int64_t orc_count(orc_world* w);

/* One evolve(): inflow, move, sort, sample+collide, hybrid switch. */
int  orc_step(orc_world* w, int n_steps);
/* Phases of one step, for kernel-level parity (phase_mask bits:
 * 1 inflow+move, 2 sort, 4 sample+collide, 8 hybrid switch, 16 advance step). */
int  orc_step_phases(orc_world* w, int phase_mask);

int  orc_samples(orc_world* w, double* out, int64_t len);
int  orc_counters(orc_world* w, ugx_counters* out);
int  orc_sigma_max(orc_world* w, double* out, int64_t len);
int  orc_set_sigma_max(orc_world* w, const double* in, int64_t len);
int  orc_cell_model(orc_world* w, int32_t* out, int64_t len);
int  orc_set_threads(int n);

/* Geometry as the oracle computes it: per cell reference point [n*3],
 * volume [n], face count [n], planes [n*6*4] (nx ny nz d, local frame),
 * face neighbour code [n*6], frame shift [n*6*3]. */
int  orc_geometry(orc_world* w, double* cc, double* vol, int32_t* nfaces,
                  double* planes, int32_t* nbr, double* delta);
/* Inflow face-species expected counts per step (Bird 4.22). */
int  orc_inflow_nbar(orc_world* w, double* out, int64_t len, int64_t* n_out);

/* Sequential per-cell sums instead of the canonical butterfly order (used
 * only to time the CPU baseline; breaks bit parity with the GPU). */
void orc_set_fast_sums(int on);

/* Primitive hooks for the known-answer and accuracy tests. */
void orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_mathfn(int fn, const double* in, const double* in2, double* out, int64_t n);
void orc_stream(uint64_t seed, uint32_t kind, uint32_t sub, uint32_t id, uint32_t step, int what,
                uint32_t m, double* out, int64_t n);

/* Test hook: n USP velocity draws (unifiedStochasticParticleSBGK.C:1047-1096)
 * for the given cQ, cS, p, most probable speed u0, heat flux q[3], stress
 * sh[6] (xx xy xz yy yz zz) and maxProb m0; xi[3n] are the draws in units of
 * u0, *raise the largest accepted P above m0 (0 if none). */
int orc_usp_sample(double cQ, double cS, double p, double u0, const double* q, const double* sh, double m0,
                   int64_t n, uint64_t seed, double* xi, double* raise);

/* Test hook: n collisions of one pair (species p, q) from the same
 * pre-collision state through the case's collision model, each on its own
 * stream; post-collision velocities outP/outQ [3n], E_rot erP/erQ [n]. */
int orc_pair_collide(orc_world* w, int p, int q, const double* UP, const double* UQ, double eP, double eQ,
                     int64_t n, uint64_t seed, double* outP, double* outQ, double* erP, double* erQ);

/* Test hooks: USP per-cell state (heat flux [nc*3], stress [nc*6], maxProb
 * [nc]); NULL skips an array. */
int orc_usp_state(orc_world* w, double* q, double* sh, double* maxp);

/* Wall measurements: the count of measured wall faces and their time sums,
 * layout [face][UGX_MAX_SPECIES][UGX_W_COUNT] as ugx_wall_samples_download. */
int64_t orc_wall_face_count(orc_world* w);
/* Cell weights [n_cells] (1 unless a cell-weighted fill or a set changed them). */
int orc_cell_weights(orc_world* w, double* out, int64_t len);
int orc_set_cell_weights(orc_world* w, const double* in, int64_t len);
/* NTC subcell levels per cell and axis, [nc*3] (noTimeCounter.C:97-161) */
int orc_subcell_levels(orc_world* w, int32_t* out, int64_t len);
int orc_set_subcell_levels(orc_world* w, const int32_t* in, int64_t len);
/* the current time step (adaptive cases change it) and the simulated time */
double orc_delta_t(orc_world* w);
double orc_time(orc_world* w);
/* test hooks: adapter sums [nc*UGX_KN_ROWS] and time, then one adaptation;
 * cell model ids */
int orc_adapt_inject(orc_world* w, const double* acc, double tav);
int orc_set_cell_model(orc_world* w, const int32_t* in, int64_t len);
/* test hook: fvc::smooth(f, max_ratio - 1) in place, [nc] */
int orc_smooth_max_ratio(orc_world* w, double* f, double max_ratio);
/* test hook: cellPoint interpolation of cell fields [nc*16] at n global positions in cells[] */
int orc_mi_probe(orc_world* w, const double* fields, int64_t n, const double* x, const int32_t* cells, double* out);
int orc_wall_samples(orc_world* w, double* out, int64_t len);
int orc_set_usp_state(orc_world* w, const double* q, const double* sh, const double* maxp);

/* Test hook: zero the sampler sums (ugx_samples_reset). */
int orc_samples_reset(orc_world* w);

#ifdef __cplusplus
}
#endif

#endif
