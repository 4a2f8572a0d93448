"""ctypes mirror of include/ugx.h (POD descriptors and enums)."""
import ctypes as C

MAX_SPECIES = 8
MAX_CELL_FACES = 6
KB = 1.38064852e-23

# ugx_status
OK, ERR_ARG, ERR_DEVICE, ERR_CAPACITY, ERR_STATE, ERR_MESH, ERR_IO, ERR_UNSUPPORTED = range(8)
# ugx_patch_type
PATCH, WALL, SYMMETRY, EMPTY, CYCLIC = range(5)
# collision / dsmc / bgk models
COLL_NONE, COLL_DSMC, COLL_BGK, COLL_HYBRID = range(4)
DSMC_NONE, DSMC_VHS, DSMC_VSS, DSMC_LB_VHS, DSMC_LB_VSS = range(5)
BGK_NONE, BGK_BGK, BGK_ESBGK, BGK_USP = range(4)
PARTNER_NTC = 0
# boundary kinds
BND_NONE, BND_SPECULAR, BND_DIFFUSE, BND_DELETION, BND_FREESTREAM = range(5)
# precision
FP64, FP32_STATE = 0, 1

# sampler layout (ugx_sample_index)
S_NAMES = ["N", "M", "MU", "MV", "MW", "MCC", "EROT", "ROTDOF",
           "MUU", "MUV", "MUW", "MVV", "MVW", "MWW",
           "MCCU", "MCCV", "MCCW", "EROTU", "EROTV", "EROTW"]
S = {n: i for i, n in enumerate(S_NAMES)}
S_BASE = len(S_NAMES)
S_SPECIES_N = S_BASE
S_SPECIES_MCC = S_BASE + MAX_SPECIES
S_COUNT = S_BASE + 2 * MAX_SPECIES

DBL_P = C.POINTER(C.c_double)
I32_P = C.POINTER(C.c_int32)


class MeshDesc(C.Structure):
    _fields_ = [
        ("n_points", C.c_int32), ("n_faces", C.c_int32),
        ("n_internal_faces", C.c_int32), ("n_cells", C.c_int32),
        ("points", DBL_P), ("face_offsets", I32_P), ("face_points", I32_P),
        ("owner", I32_P), ("neighbour", I32_P),
        ("n_patches", C.c_int32), ("pad0", C.c_int32),
        ("patch_start", I32_P), ("patch_size", I32_P),
        ("patch_type", I32_P), ("patch_neighbour", I32_P),
    ]


class SpeciesDesc(C.Structure):
    _fields_ = [
        ("name", C.c_char * 32), ("mass", C.c_double), ("diameter", C.c_double),
        ("omega", C.c_double), ("alpha", C.c_double),
        ("rotational_dof", C.c_int32), ("vibrational_modes", C.c_int32),
        ("charge", C.c_int32), ("pad0", C.c_int32),
    ]


class ModelDesc(C.Structure):
    _fields_ = [
        ("n_equivalent_particles", C.c_double), ("delta_t", C.c_double),
        ("collision_model", C.c_int32), ("dsmc_model", C.c_int32),
        ("partner_model", C.c_int32), ("bgk_model", C.c_int32),
        ("Tref", C.c_double), ("Zrot", C.c_double), ("theta", C.c_double),
        ("macro_interpolation", C.c_int32), ("sample_interval", C.c_int32),
        ("decomposition_interval", C.c_int32), ("smoothing_passes", C.c_int32),
        ("breakdown_max", C.c_double), ("kn_theta", C.c_double),
        ("sampler_full", C.c_int32), ("reset_at_decomposition", C.c_int32),
        ("reset_until_time", C.c_double),
        ("cell_weighted", C.c_int32), ("particles_per_subcell", C.c_int32),
        ("min_particles_per_subcell", C.c_int32), ("max_smoothing_passes", C.c_int32),
        ("max_cell_weight_ratio", C.c_double),
        ("adaptive", C.c_int32), ("time_step_adaptation", C.c_int32),
        ("subcell_adaptation", C.c_int32), ("cell_weight_adaptation", C.c_int32),
        ("adaptation_interval", C.c_int32), ("adaptation_smoothing_passes", C.c_int32),
        ("max_time_step_mct_ratio", C.c_double), ("max_courant_number", C.c_double),
        ("max_subcell_size_mfp_ratio", C.c_double),
    ]


class BoundaryDesc(C.Structure):
    _fields_ = [
        ("patch", C.c_int32), ("kind", C.c_int32),
        ("temperature", C.c_double), ("velocity", C.c_double * 3),
        ("rot_temperature", C.c_double),
        ("number_density", C.c_double * MAX_SPECIES),
    ]


class FillDesc(C.Structure):
    _fields_ = [
        ("number_density", C.c_double * MAX_SPECIES), ("T", C.c_double * 3),
        ("rot_temperature", C.c_double), ("velocity", C.c_double * 3),
        ("exact_per_cell", C.c_int32), ("pad0", C.c_int32),
    ]


class Options(C.Structure):
    _fields_ = [
        ("precision", C.c_int32), ("device", C.c_int32),
        ("particle_capacity", C.c_int64), ("seed", C.c_uint64),
        ("use_graphs", C.c_int32), ("pad0", C.c_int32),
    ]


class BatchOptions(C.Structure):
    _fields_ = [
        ("precision", C.c_int32), ("device", C.c_int32), ("hbm_budget", C.c_int64),
        ("max_cases", C.c_int32), ("use_graphs", C.c_int32),
    ]


class ParticleSoA(C.Structure):
    _fields_ = [
        ("n", C.c_int64), ("position", DBL_P), ("local_position", DBL_P),
        ("velocity", DBL_P), ("erot", DBL_P), ("cell", I32_P), ("species", I32_P),
    ]


class Fields(C.Structure):
    _fields_ = [(n, DBL_P) for n in (
        "rhoN", "rhoM", "U_mean", "translational_T", "rotational_T", "overall_T",
        "p", "Ma", "mfp", "mcr", "density_error", "velocity_error",
        "temperature_error", "pressure_error", "shear_stress", "heat_flux")]


KN_ROWS = 16          # UGX_KN_ROWS
N_COUNTERS = 16       # UGX_N_COUNTERS


class CaseState(C.Structure):
    """ugx_case_state: restart state of a case besides its simulators."""
    _fields_ = [
        ("n_cells", C.c_int64), ("steps", C.c_int64), ("kn_time", C.c_double),
        ("sampled_at_reset", C.c_int64), ("counters", C.c_uint64 * N_COUNTERS),
        ("sigma_max", DBL_P), ("cell_model", I32_P), ("samples", DBL_P), ("usp_heat_flux", DBL_P),
        ("usp_stress", DBL_P), ("usp_max_prob", DBL_P), ("kn_sums", DBL_P), ("kn", DBL_P),
        ("n_wall_faces", C.c_int64), ("wall_samples", DBL_P), ("cell_weight", DBL_P),
        ("delta_t", C.c_double), ("time", C.c_double), ("time_sampled", C.c_double),
        ("adapt_time", C.c_double), ("subcell_levels", I32_P), ("adapt_sums", DBL_P),
        ("adapt_ratio", DBL_P),
    ]


# per-cell arrays of CaseState: (name, rows per cell, dtype)
STATE_ARRAYS = (("sigma_max", 1, "f8"), ("cell_model", 1, "i4"), ("samples", S_COUNT, "f8"),
                ("usp_heat_flux", 3, "f8"), ("usp_stress", 6, "f8"), ("usp_max_prob", 1, "f8"),
                ("kn_sums", KN_ROWS, "f8"), ("kn", 1, "f8"))


# ugx_wall_index: rows of the wall sums per (face, species)
(W_RHON, W_RHON_INT, W_RHOM, W_LINEAR_KE, W_MCC, W_MU, W_MV, W_MW, W_EROT, W_ROTDOF,
 W_Q, W_FX, W_FY, W_FZ, W_COUNT) = range(15)

WALL_FIELD_NAMES = ("rhoN", "rhoM", "U_mean", "translational_T", "rotational_T", "overall_T",
                    "heat_flux", "force_density", "pressure", "shear")


class WallFields(C.Structure):
    _fields_ = [(n, DBL_P) for n in WALL_FIELD_NAMES]


class Counters(C.Structure):
    _fields_ = [
        ("n_particles", C.c_int64), ("steps", C.c_int64), ("candidates", C.c_int64),
        ("collisions", C.c_int64), ("bgk_collisions", C.c_int64),
        ("inserted", C.c_int64), ("deleted", C.c_int64), ("wall_hits", C.c_int64),
        ("track_errors", C.c_int64), ("n_sampled_steps", C.c_int64),
        ("time_sampled", C.c_double), ("n_dsmc_cells", C.c_int64), ("particle_steps", C.c_int64),
        ("cloned", C.c_int64), ("weight_deleted", C.c_int64),
    ]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


def dptr(a):
    return a.ctypes.data_as(DBL_P) if a is not None else None


def iptr(a):
    return a.ctypes.data_as(I32_P) if a is not None else None
