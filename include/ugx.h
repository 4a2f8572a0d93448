/*
 * ugx.h - C ABI of the MI355X-native uniGasFoam particle time step.
 *
 * The library owns device memory; callers own every host buffer they pass.
 * All entry points return 0 (UGX_OK) on success and a ugx_status code
 * otherwise; ugx_last_error() returns a thread-local message.  No C++ types,
 * exceptions or ownership transfers cross this boundary.  One case (or batch)
 * is not thread-safe; distinct cases are.  Results are deterministic given
 * (seed, inputs): every random draw comes from a counter-based Philox-4x32-10
 * stream keyed by (case seed, entity, step).
 *
 * Reference interfaces replaced (paths relative to src/lagrangian/uniGas/ of
 * the uniGasFoam sources):
 *   ugx_case_create      uniGasCloud::uniGasCloud           clouds/uniGasCloud.C:396-781
 *                        (model selection: dsmcCollisionModel::New
 *                         dsmcCollisions/basic/dsmcCollisionModel.C:59-83,
 *                         bgkCollisionModel::New bgkCollisions/basic/bgkCollisionModel.C:55-80,
 *                         dsmcCollisionPartner::New dsmcCollisionPartner/basic/dsmcCollisionPartner.C:57-83,
 *                         uniGasBoundaries ctor boundaries/basic/uniGasBoundaries/uniGasBoundaries.C:493-556)
 *   ugx_case_fill        uniGasMeshFill::setInitialConfiguration
 *                        uniGasInitialisation/derived/uniGasMeshFill/uniGasMeshFill.C:58-297
 *   ugx_step             uniGasCloud::evolve                clouds/uniGasCloud.C:821-869
 *   ugx_counters         uniGasCloud::info / noTimeCounter info lines
 *                        clouds/uniGasCloud.C:872-920, noTimeCounter.C:318-342
 *   ugx_samples_download uniGasVolFields raw time sums
 *                        macroscopicProperties/derived/volumetric/uniGasVolFields/uniGasVolFields.C:723-837
 *   ugx_fields_compute   uniGasVolFields write-time fields  uniGasVolFields.C:839-1430
 *   ugx_particles_*      uniGasParcel state / lagrangian IO parcels/uniGasParcelIO.C:89-181
 *   ugx_state_*          restart of the cloud and of the resumable sampler sums
 *                        clouds/uniGasCloud.C:577-703 (time-step and cloud state read at start-up),
 *                        uniGasVolFields.C:550-669 (readIn / writeOut of the running sums)
 */
#ifndef UGX_H
#define UGX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UGX_MAX_SPECIES    8   /* species per case */
#define UGX_MAX_CELL_FACES 6   /* convex cells with up to six planar faces */

/* Boltzmann constant used by every kernel (CODATA 2014, as OpenFOAM's
 * physicoChemical::k). */
#define UGX_KB 1.38064852e-23

typedef enum {
    UGX_OK = 0,
    UGX_ERR_ARG = 1,       /* invalid argument / descriptor */
    UGX_ERR_DEVICE = 2,    /* HIP runtime failure or no device */
    UGX_ERR_CAPACITY = 3,  /* particle or cell capacity exceeded */
    UGX_ERR_STATE = 4,     /* call not valid in the current state */
    UGX_ERR_MESH = 5,      /* mesh cannot be used (non-convex, >6 faces, ...) */
    UGX_ERR_IO = 6,        /* dictionary / case directory error */
    UGX_ERR_UNSUPPORTED = 7
} ugx_status;

/* OpenFOAM polyPatch types (constant/polyMesh/boundary "type"). */
typedef enum {
    UGX_PATCH_PATCH = 0,     /* generic patch: open boundary */
    UGX_PATCH_WALL = 1,
    UGX_PATCH_SYMMETRY = 2,  /* symmetry and symmetryPlane: mirror */
    UGX_PATCH_EMPTY = 3,     /* reduced-D direction, never crossed */
    UGX_PATCH_CYCLIC = 4     /* translational periodic pair */
} ugx_patch_type;

/* Mesh in OpenFOAM polyMesh form (points, faces, owner, neighbour, boundary).
 * Faces point from owner to neighbour; boundary faces point out of the
 * domain.  Patches are contiguous face ranges after the internal faces. */
typedef struct {
    int32_t n_points;
    int32_t n_faces;
    int32_t n_internal_faces;
    int32_t n_cells;
    const double*  points;          /* [n_points*3] */
    const int32_t* face_offsets;    /* [n_faces+1] into face_points */
    const int32_t* face_points;     /* vertex ids, right-handed about the face normal */
    const int32_t* owner;           /* [n_faces] */
    const int32_t* neighbour;       /* [n_internal_faces] */
    int32_t n_patches;
    int32_t pad0;
    const int32_t* patch_start;     /* [n_patches] first face id */
    const int32_t* patch_size;      /* [n_patches] */
    const int32_t* patch_type;      /* [n_patches] ugx_patch_type */
    const int32_t* patch_neighbour; /* [n_patches] cyclic partner patch or -1; may be NULL */
} ugx_mesh_desc;

/* constant/uniGasProperties moleculeProperties entry
 * (parcels/uniGasParcel.H:68-187, parcels/uniGasParcelI.H:40-141). */
typedef struct {
    char    name[32];
    double  mass;
    double  diameter;
    double  omega;
    double  alpha;
    int32_t rotational_dof;
    int32_t vibrational_modes;  /* accepted, not exchanged (out of scope) */
    int32_t charge;
    int32_t pad0;
} ugx_species_desc;

typedef enum { UGX_COLL_NONE = 0, UGX_COLL_DSMC = 1, UGX_COLL_BGK = 2, UGX_COLL_HYBRID = 3 } ugx_collision_model;
typedef enum {
    UGX_DSMC_NONE = 0,   /* noDSMCCollision */
    UGX_DSMC_VHS = 1,    /* variableHardSphere */
    UGX_DSMC_VSS = 2,    /* variableSoftSphere */
    UGX_DSMC_LB_VHS = 3, /* LarsenBorgnakkeVariableHardSphere */
    UGX_DSMC_LB_VSS = 4  /* LarsenBorgnakkeVariableSoftSphere */
} ugx_dsmc_model;
typedef enum {
    UGX_BGK_NONE = 0,    /* noBGKCollision */
    UGX_BGK_BGK = 1,     /* stochasticParticleBGK */
    UGX_BGK_ESBGK = 2,   /* stochasticParticleESBGK */
    UGX_BGK_USP = 3      /* unifiedStochasticParticleSBGK */
} ugx_bgk_model;
typedef enum { UGX_PARTNER_NTC = 0 } ugx_partner_model;

/* constant/uniGasProperties + system/controlDict + hybridDecompositionDict. */
typedef struct {
    double  n_equivalent_particles;   /* F_N (nEquivalentParticles) */
    double  delta_t;                  /* controlDict deltaT */
    int32_t collision_model;          /* ugx_collision_model */
    int32_t dsmc_model;               /* ugx_dsmc_model */
    int32_t partner_model;            /* ugx_partner_model */
    int32_t bgk_model;                /* ugx_bgk_model */
    double  Tref;                     /* collisionProperties.Tref */
    double  Zrot;                     /* rotationalRelaxationCollisionNumber (LB) */
    double  theta;                    /* USP time-relaxation theta (default 1) */
    int32_t macro_interpolation;      /* must be 0: cell values */
    int32_t sample_interval;          /* fieldPropertiesDict sampleInterval (>=1) */
    /* hybrid (localKnudsen) */
    int32_t decomposition_interval;   /* default 100 */
    int32_t smoothing_passes;         /* default 5 */
    double  breakdown_max;            /* Kn_GLL threshold (default 0.05) */
    double  kn_theta;                 /* EMA weight of Kn (default 0.2) */
    int32_t sampler_full;             /* 1: also sample the second moments and heat-flux
                                         moments (measureHeatFluxShearStress); 0: the
                                         uniGasVolFields set (rows 0-7 and per species) */
    int32_t reset_at_decomposition;   /* hybrid: clear the Kn time sums at each
                                         decomposition (default on) */
    double  reset_until_time;         /* ... while time < this (<= 0: always) */
    /* cell weighting (cellWeightedSimulation; uniGasCloud.C:1353-1425 and
       uniGasMeshFill.C:111-168): the fill sets a weight W_c per cell from its
       number density, a simulator of cell c stands for F_N W_c molecules, and
       a simulator entering a cell of another weight is cloned or deleted */
    int32_t cell_weighted;            /* 0 off */
    int32_t particles_per_subcell;    /* cellWeightedProperties.particlesPerSubCell (> 0 when weighted) */
    int32_t min_particles_per_subcell;/* ... minParticlesPerSubCell (0: particles_per_subcell) */
    int32_t max_smoothing_passes;     /* weight smoothing passes at most (0: 500) */
    double  max_cell_weight_ratio;    /* smoothing stops once neighbour ratios are below 1 + this (0: 0.05) */
    /* adaptive simulation (adaptiveSimulation and adaptiveProperties;
       uniGasDynamicAdapter.C:187-217, 526-700): the fill sets the time step
       and NTC subcell levels from the filled gas, then every
       adaptation_interval steps the time-averaged density, velocity and
       temperature set them again (and the cell weights of a cell-weighted
       case) */
    int32_t adaptive;                 /* 0 off */
    int32_t time_step_adaptation;     /* timeStepAdaptation */
    int32_t subcell_adaptation;       /* subCellAdaptation */
    int32_t cell_weight_adaptation;   /* cellWeightAdaptation (a cell-weighted case only) */
    int32_t adaptation_interval;      /* adaptationInterval (0: 20) */
    int32_t adaptation_smoothing_passes; /* smoothingPasses (0: 25; < 0: none) */
    double  max_time_step_mct_ratio;  /* maxTimeStepMCTRatio (0: 0.2) */
    double  max_courant_number;       /* maxCourantNumber (0: 0.5) */
    double  max_subcell_size_mfp_ratio; /* maxSubCellSizeMFPRatio (0: 0.5) */
} ugx_model_desc;

typedef enum {
    UGX_BND_NONE = 0,
    UGX_BND_SPECULAR = 1,   /* uniGasSpecularWallPatch */
    UGX_BND_DIFFUSE = 2,    /* uniGasDiffuseWallPatch */
    UGX_BND_DELETION = 3,   /* uniGasDeletionPatch */
    UGX_BND_FREESTREAM = 4  /* uniGasFreeStreamInflowPatch (general boundary) */
} ugx_boundary_kind;

/* system/boundariesDict entry. A patch may carry one patch model (specular,
 * diffuse, deletion) and, independently, one freestream inflow. */
typedef struct {
    int32_t patch;
    int32_t kind;                      /* ugx_boundary_kind */
    double  temperature;               /* diffuse wall temperature */
    double  velocity[3];               /* diffuse wall velocity / inflow velocity */
    double  rot_temperature;           /* inflow rotational temperature */
    double  number_density[UGX_MAX_SPECIES]; /* inflow, per case species (0 = none) */
} ugx_boundary_desc;

/* system/uniGasInitialisationDict uniGasMeshFill entry.  T[3] are per-axis
 * translational temperatures (equal values = the reference's isotropic
 * fill); exact_per_cell > 0 places exactly that many simulators per cell,
 * split over species by number density (a test initialiser). */
typedef struct {
    double  number_density[UGX_MAX_SPECIES];
    double  T[3];
    double  rot_temperature;
    double  velocity[3];
    int32_t exact_per_cell;
    int32_t pad0;
} ugx_fill_desc;

typedef enum { UGX_FP64 = 0, UGX_FP32_STATE = 1 } ugx_precision;

typedef struct {
    int32_t  precision;           /* ugx_precision */
    int32_t  device;              /* HIP device ordinal */
    int64_t  particle_capacity;   /* simulators the buffers hold at first (0: sized by the
                                     fill or upload); ugx_step grows them between steps
                                     whenever a step's inflow could exceed them */
    uint64_t seed;
    int32_t  use_graphs;          /* replay each step as a captured HIP graph */
    int32_t  pad0;
} ugx_options;

/* Host-side particle arrays (always FP64 on the host). local_position, when
 * non-NULL, holds positions relative to the cell reference point (the
 * library's storage form); position holds global coordinates. */
typedef struct {
    int64_t  n;                /* in: capacity of the arrays, out: count */
    double*  position;         /* [n*3] or NULL */
    double*  local_position;   /* [n*3] or NULL */
    double*  velocity;         /* [n*3] */
    double*  erot;             /* [n] */
    int32_t* cell;             /* [n] */
    int32_t* species;          /* [n] case species index */
} ugx_particle_soa;

/* Raw sampler sums per cell, sum over steps of deltaT * (per-step sum over
 * the cell's simulators); layout [UGX_S_COUNT][n_cells].  Per-species blocks
 * follow: UGX_S_SPECIES_N + s and UGX_S_SPECIES_MCC + s. */
typedef enum {
    UGX_S_N = 0, UGX_S_M, UGX_S_MU, UGX_S_MV, UGX_S_MW, UGX_S_MCC,
    UGX_S_EROT, UGX_S_ROTDOF,
    UGX_S_MUU, UGX_S_MUV, UGX_S_MUW, UGX_S_MVV, UGX_S_MVW, UGX_S_MWW,
    UGX_S_MCCU, UGX_S_MCCV, UGX_S_MCCW, UGX_S_EROTU, UGX_S_EROTV, UGX_S_EROTW,
    UGX_S_BASE,
    UGX_S_SPECIES_N = UGX_S_BASE,
    UGX_S_SPECIES_MCC = UGX_S_BASE + UGX_MAX_SPECIES,
    UGX_S_COUNT = UGX_S_BASE + 2 * UGX_MAX_SPECIES
} ugx_sample_index;

/* Write-time volFields (uniGasVolFields.C:839-1430); every array has n_cells
 * entries (U_mean 3*n_cells).  Any pointer may be NULL. */
typedef struct {
    double* rhoN;
    double* rhoM;
    double* U_mean;
    double* translational_T;
    double* rotational_T;
    double* overall_T;
    double* p;
    double* Ma;
    double* mfp;
    double* mcr;
    double* density_error;
    double* velocity_error;
    double* temperature_error;
    double* pressure_error;
    double* shear_stress;      /* [6*n_cells] xx xy xz yy yz zz, Pa */
    double* heat_flux;         /* [3*n_cells] W/m^2 */
} ugx_fields;

/* Wall measurements: the boundaryMeasurements of every wall patch with a
 * specular or diffuse model (boundaries/basic/uniGasPatchBoundary/
 * uniGasPatchBoundary.C:130-302), time-summed as uniGasVolFields does
 * (uniGasVolFields.C:800-830).  Wall faces are numbered in patch then face
 * order.  Per face and case species, each row is the sum over sampled steps
 * of deltaT times the step's sum over the wall hits of that species: a hit
 * adds its before- and after-interaction terms weighted by
 * 1/max(|U.n| A, VSMALL), and to Q and F the energy and momentum it leaves
 * on the wall per unit area and time (F_N (before - after)/(deltaT A)).
 * Only the hits of sampling steps are recorded.  The step's hits of a face
 * are summed per species in the order (slot, crossing).
 * Layout [face][UGX_MAX_SPECIES][UGX_W_COUNT]. */
typedef enum {
    UGX_W_RHON = 0, UGX_W_RHON_INT, UGX_W_RHOM, UGX_W_LINEAR_KE, UGX_W_MCC,
    UGX_W_MU, UGX_W_MV, UGX_W_MW, UGX_W_EROT, UGX_W_ROTDOF,
    UGX_W_Q, UGX_W_FX, UGX_W_FY, UGX_W_FZ,
    UGX_W_COUNT
} ugx_wall_index;

/* Write-time wall values of the uniGasVolFields (uniGasVolFields.C:1262-1350),
 * one entry per wall face (U_mean and force_density 3 per face).  Any
 * pointer may be NULL. */
typedef struct {
    double* rhoN;
    double* rhoM;
    double* U_mean;
    double* translational_T;
    double* rotational_T;
    double* overall_T;
    double* heat_flux;         /* q: energy left on the wall, W/m^2 */
    double* force_density;     /* fD: momentum left on the wall, Pa */
    double* pressure;          /* fD . n */
    double* shear;             /* |fD . t1, fD . t2| */
} ugx_wall_fields;

typedef struct {
    int64_t n_particles;
    int64_t steps;
    int64_t candidates;        /* NTC collision candidates, cumulative */
    int64_t collisions;        /* accepted DSMC collisions, cumulative */
    int64_t bgk_collisions;    /* BGK-family relaxed simulators, cumulative */
    int64_t inserted;
    int64_t deleted;
    int64_t wall_hits;
    int64_t track_errors;      /* tracks stopped at the crossing limit */
    int64_t n_sampled_steps;
    double  time_sampled;      /* sum of deltaT over sampled steps */
    int64_t n_dsmc_cells;      /* hybrid: cells using DSMC now */
    int64_t particle_steps;    /* sum over steps of the simulator count after the step */
    int64_t cloned;            /* cell weighting: clones made, cumulative */
    int64_t weight_deleted;    /* cell weighting: simulators removed, cumulative */
} ugx_counters;

typedef struct ugx_case ugx_case;

const char* ugx_version(void);
const char* ugx_last_error(void);

int ugx_case_create(const ugx_mesh_desc* mesh,
                    const ugx_species_desc* species, int n_species,
                    const ugx_model_desc* model,
                    const ugx_boundary_desc* boundaries, int n_boundaries,
                    const ugx_options* options,
                    ugx_case** out);
int ugx_case_destroy(ugx_case* c);

int ugx_case_fill(ugx_case* c, const ugx_fill_desc* fill);
/* The fill of ugx_case_fill without any device allocation.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops).  Its scratch (the
 * fill's parameters and the per (cell, species) counts and offsets) is a
 * fixed 16 MiB device array that each GPU holds until the process ends, so a
 * fill may cover at most 2,097,088 (cell, species) pairs (UGX_ERR_CAPACITY
 * beyond).  The particle capacity is never grown: the case must have been
 * created with room for the fill (ugx_options.particle_capacity), else
 * UGX_ERR_CAPACITY.  Calls that overlap on one GPU race on the shared
 * scratch.  Results equal ugx_case_fill's bit for bit. */
int ugx_case_fill_noalloc(ugx_case* c, const ugx_fill_desc* fill);
int ugx_particles_upload(ugx_case* c, const ugx_particle_soa* p);
int ugx_particles_download(ugx_case* c, ugx_particle_soa* p);
int ugx_particle_count(ugx_case* c, int64_t* n);

/* Enqueue n_steps calls of the evolve() equivalent on the case's stream. */
int ugx_step(ugx_case* c, int n_steps);
int ugx_sync(ugx_case* c);

int ugx_samples_download(ugx_case* c, double* sums, int64_t len);
int ugx_samples_reset(ugx_case* c);
int ugx_fields_compute(ugx_case* c, ugx_fields* f);
int ugx_counters_get(ugx_case* c, ugx_counters* out);
/* The case's current time step (the controlDict deltaT, or the adapter's,
 * uniGasDynamicAdapter.C:316-372) and the simulated time. */
int ugx_time_get(ugx_case* c, double* delta_t, double* time);
int ugx_sigma_max_download(ugx_case* c, double* out, int64_t len);
int ugx_sigma_max_upload(ugx_case* c, const double* in, int64_t len);
int ugx_cell_model_download(ugx_case* c, int32_t* out, int64_t len);
/* Kn_GLL per cell of a hybrid case after the last decomposition: the
 * KnGLL volScalarField of localKnudsen (localKnudsen.C:395-430). */
int ugx_kn_download(ugx_case* c, double* out, int64_t len);
/* Cell weights W_c (1 in a case without cell weighting): set by the fill of
 * a cell-weighted case (uniGasMeshFill.C:111-168) or uploaded (a restart
 * reads them back, as the reference reads its cellWeightFactor field,
 * uniGasCloud.C:612-623); an upload also rescales the inflow counts. */
int ugx_cell_weights_download(ugx_case* c, double* out, int64_t len);
int ugx_cell_weights_upload(ugx_case* c, const double* in, int64_t len);
/* NTC subcell levels per cell and axis, [n_cells][3] (the SubCellLevels
 * field of uniGasCloud.C:461-474 read by noTimeCounter.C:97-161): the
 * second candidate of a pair comes from the first's virtual Cartesian
 * subcell when that holds two or more simulators.  Levels are 1..10 and 1
 * in empty directions; all 1 (the default) is the plain NTC. */
int ugx_subcell_levels_download(ugx_case* c, int32_t* out, int64_t len);
int ugx_subcell_levels_upload(ugx_case* c, const int32_t* in, int64_t len);
int ugx_cell_volumes(ugx_case* c, double* out, int64_t len);
/* Wall faces of the case: their count, and per face its patch and its face
 * index within the patch. */
int ugx_wall_face_count(ugx_case* c, int64_t* n);
int ugx_wall_faces(ugx_case* c, int32_t* patch, int32_t* face, int64_t len);
/* Raw wall sums, layout [face][UGX_MAX_SPECIES][UGX_W_COUNT]. */
int ugx_wall_samples_download(ugx_case* c, double* sums, int64_t len);
int ugx_wall_fields_compute(ugx_case* c, ugx_wall_fields* f);
int ugx_cell_centres(ugx_case* c, double* out, int64_t len);

/* ------------------------------------------------------------ restart
 * Everything a plain case carries from one step to the next besides its
 * simulators: the step count (which drives every random stream and every
 * interval), the hybrid averaging time, the cumulative counters and the
 * per-cell arrays.  A checkpoint is this state plus the simulators from
 * ugx_particles_download with local_position set (the storage form); a
 * restart creates a case with the same descriptors and seed, uploads those
 * simulators, then this state.  The restarted case continues bit for bit.
 * Per-cell arrays are cell-major (row r of cell c at [c * rows + r]).
 * This is the library's own restart form: it is not the reference's
 * <time>/lagrangian/uniGas parcel files (uniGasParcelIO.C:89-181), which
 * store positions as tet-barycentric coordinates, so a state written by
 * uniGasFoam cannot be read here nor the other way round. */
#define UGX_KN_ROWS     16   /* hybrid time sums per cell */
#define UGX_N_COUNTERS  16   /* counter slots (library order; opaque) */

typedef struct {
    int64_t  n_cells;            /* in: the case's cell count */
    int64_t  steps;              /* steps taken */
    double   kn_time;            /* hybrid: time summed into kn_sums since their last reset */
    int64_t  sampled_at_reset;   /* sampled-step counter at the last ugx_samples_reset */
    uint64_t counters[UGX_N_COUNTERS];
    double*  sigma_max;          /* [n_cells] (sigma_T c_r)max */
    int32_t* cell_model;         /* [n_cells] 1 DSMC, 0 BGK family */
    double*  samples;            /* [n_cells * UGX_S_COUNT] sampler time sums */
    double*  usp_heat_flux;      /* [n_cells * 3] USP relaxed heat flux */
    double*  usp_stress;         /* [n_cells * 6] USP relaxed stress (xx xy xz yy yz zz) */
    double*  usp_max_prob;       /* [n_cells] USP maxProb */
    double*  kn_sums;            /* [n_cells * UGX_KN_ROWS] */
    double*  kn;                 /* [n_cells] Kn_GLL */
    int64_t  n_wall_faces;       /* in: the case's wall face count (ugx_wall_face_count) */
    double*  wall_samples;       /* [n_wall_faces * UGX_MAX_SPECIES * UGX_W_COUNT] wall sums;
                                    NULL on upload: the case's wall sums are left as they are */
    double*  cell_weight;        /* [n_cells] cell weights; NULL: not carried */
    /* time and adaptation */
    double   delta_t;            /* the current time step (changed by time-step adaptation) */
    double   time;               /* simulated time */
    double   time_sampled;       /* time steps of the sampled steps summed since the last sampler reset
                                    (NaN on upload: the sampled steps times the time step) */
    double   adapt_time;         /* time summed into adapt_sums since the last adaptation */
    int32_t* subcell_levels;     /* [n_cells * 3]; NULL: not carried */
    double*  adapt_sums;         /* [n_cells * UGX_KN_ROWS] adapter time sums; NULL: not carried */
    double*  adapt_ratio;        /* [n_cells * 3] cell size to mean free path ratio of the last
                                    adaptation; NULL: not carried */
} ugx_case_state;

/* Copy the state out (any case) or in (a plain case; every pointer non-NULL). */
int ugx_state_download(ugx_case* c, ugx_case_state* s);
int ugx_state_upload(ugx_case* c, const ugx_case_state* s);

/* ------------------------------------------------------------ batches
 * Batched and served forms.  No counterpart in the reference, which runs one
 * case per solver process (uniGasFoam.C:83); these entry points are for
 * callers that run many independent cases (parametric sweeps, a request
 * loop).  A batch holds many cases in one device world and advances them
 * with the same kernels.  Admissions and retirements take effect at the next
 * ugx_batch_step: the world is rebuilt and every remaining case keeps its
 * particles, samplers, per-cell state and counters on the device.  Every
 * random stream and every per-cell sum depends only on the case itself, so a
 * case gives the same results alone, in a batch, and across rebuilds. */
typedef struct ugx_batch ugx_batch;

typedef struct {
    int32_t precision;     /* ugx_precision */
    int32_t device;
    int64_t hbm_budget;    /* bytes the batch's cases may take (0: 60% of free HBM) */
    int32_t max_cases;     /* 0: no limit */
    int32_t use_graphs;    /* replay each step as a captured HIP graph */
} ugx_batch_options;

int ugx_batch_create(const ugx_batch_options* options, ugx_batch** out);
int ugx_batch_destroy(ugx_batch* b);
/* Device bytes a case takes in a batch: mesh arrays, per-cell state and
 * particle storage for `particles` simulators. */
int ugx_case_footprint(const ugx_mesh_desc* mesh, const ugx_model_desc* model, int64_t particles,
                       int precision, int64_t* bytes);
/* Queues a case: the descriptors of ugx_case_create plus its meshFill
 * (fill may be NULL: the case starts empty).  Returns UGX_ERR_CAPACITY when
 * its footprint does not fit the remaining budget or slot count. */
int ugx_batch_admit(ugx_batch* b, const ugx_mesh_desc* mesh, const ugx_species_desc* species, int n_species,
                    const ugx_model_desc* model, const ugx_boundary_desc* boundaries, int n_boundaries,
                    const ugx_fill_desc* fill, uint64_t seed, int64_t expected_particles, int* slot);
int ugx_batch_retire(ugx_batch* b, int slot);
int ugx_batch_step(ugx_batch* b, int n_steps);
int ugx_batch_sync(ugx_batch* b);
/* Non-owning handle of an admitted case for the per-case entry points
 * (samples, fields, counters, particles, ...); ugx_step and ugx_case_destroy
 * refuse it.  Valid until the slot is retired. */
int ugx_batch_case(ugx_batch* b, int slot, ugx_case** out);
int ugx_batch_info(ugx_batch* b, int64_t* n_cases, int64_t* bytes_used, int64_t* n_particles);

/* Per-phase timing of ugx_step with HIP events on the case's stream (no
 * counterpart in the reference, whose only timing is the ExecutionTime /
 * ClockTime log line of uniGasFoam.C:97-101).  Phases, in launch order:
 * begin, inflow, move, wall (diffuse-wall continuation), scan, scatter,
 * cells (slot order + gather + per-cell sums + sampler), collide (NTC and
 * the BGK family), end. */
#define UGX_N_PHASES 9
const char* ugx_phase_name(int i);
int ugx_timing_enable(ugx_case* c, int on);       /* also resets the totals */
int ugx_timing_get(ugx_case* c, double* ms_total, int n_phases, int64_t* steps);

/* Cell-tile configuration the case runs (no counterpart in the reference):
 * its index, threads per block, key window and LDS simulators, and whether
 * the autotuner has finished.  The tuner times each of the three
 * configurations on one of the case's first three steps (with or without
 * ugx_timing_enable) and keeps the fastest; results never depend on it. */
int ugx_tile_config(ugx_case* c, int32_t* variant, int32_t* threads, int32_t* window, int32_t* lds_simulators,
                    int32_t* tuned);

/* Device-to-device copy of `bytes` bytes timed over `reps` launches: the
 * stream-copy probe that confirms the HBM peak (no reference counterpart). */
int ugx_probe_copy(int device, int64_t bytes, int reps, double* seconds_per_copy);

/* Host-only geometry in kernel storage form (no device needed): per cell
 * reference point [3n], volume [n], face count [n], outward planes
 * [n*6*4] (nx ny nz d in the cell frame), neighbour codes [n*6] (cell id,
 * -1-patch for boundary models, INT32_MIN for empty), frame shifts [n*6*3],
 * and the inflow face-species expected insert counts per step.  No
 * counterpart in the reference: it exposes what the tracking kernels use so
 * that callers can check mesh conversion. */
int ugx_geometry(const ugx_mesh_desc* mesh, const ugx_species_desc* species, int n_species,
                 const ugx_model_desc* model, const ugx_boundary_desc* boundaries, int n_boundaries,
                 double* cc, double* vol, int32_t* nfaces, double* planes, int32_t* nbr, double* delta,
                 double* nbar, int64_t nbar_len, int64_t* n_inflow);

#ifdef __cplusplus
}
#endif

#endif /* UGX_H */
