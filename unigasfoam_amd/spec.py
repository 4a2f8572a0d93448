"""Python-side case description mirroring the OpenFOAM dictionaries.

``CaseSpec`` holds what constant/uniGasProperties, system/boundariesDict,
system/controlDict and the mesh give uniGasCloud: species
(moleculeProperties), the collision models, boundary models per patch, and
deltaT.  ``descriptors()`` turns it into the POD structs of include/ugx.h.
"""
from __future__ import annotations

from dataclasses import dataclass, field

from . import abi
from .mesh import PolyMesh

DSMC_MODELS = {
    "noDSMCCollision": abi.DSMC_NONE,
    "variableHardSphere": abi.DSMC_VHS,
    "variableSoftSphere": abi.DSMC_VSS,
    "LarsenBorgnakkeVariableHardSphere": abi.DSMC_LB_VHS,
    "LarsenBorgnakkeVariableSoftSphere": abi.DSMC_LB_VSS,
}
BGK_MODELS = {
    "noBGKCollision": abi.BGK_NONE,
    "stochasticParticleBGK": abi.BGK_BGK,
    "stochasticParticleESBGK": abi.BGK_ESBGK,
    "unifiedStochasticParticleSBGK": abi.BGK_USP,
}
COLLISION_MODELS = {"none": abi.COLL_NONE, "dsmc": abi.COLL_DSMC, "bgk": abi.COLL_BGK, "hybrid": abi.COLL_HYBRID}


@dataclass
class Species:
    name: str
    mass: float
    diameter: float
    omega: float
    alpha: float = 1.0
    rotational_dof: int = 0
    vibrational_modes: int = 0
    charge: int = 0


@dataclass
class Boundary:
    patch: str
    kind: int                      # abi.BND_*
    temperature: float = 0.0
    velocity: tuple = (0.0, 0.0, 0.0)
    rot_temperature: float = 0.0
    number_density: dict = field(default_factory=dict)   # species name -> m^-3


@dataclass
class CaseSpec:
    mesh: PolyMesh
    species: list
    delta_t: float
    n_equivalent_particles: float
    collision_model: str = "dsmc"
    dsmc_model: str = "variableHardSphere"
    bgk_model: str = "noBGKCollision"
    Tref: float = 273.0
    Zrot: float = 5.0
    theta: float = 1.0
    macro_interpolation: bool = False    # collisionProperties.macroInterpolation (BGK family)
    sample_interval: int = 1
    boundaries: list = field(default_factory=list)
    decomposition_interval: int = 100
    smoothing_passes: int = 5
    breakdown_max: float = 0.05
    kn_theta: float = 0.2
    reset_at_decomposition: bool = True
    reset_until_time: float = 0.0        # <= 0: no limit (the reference's VGREAT)
    sampler_full: bool = False
    seed: int = 12345
    # cell weighting (cellWeightedSimulation / cellWeightedProperties)
    cell_weighted: bool = False
    particles_per_subcell: int = 0
    min_particles_per_subcell: int = 0
    max_smoothing_passes: int = 500
    max_cell_weight_ratio: float = 0.05
    # adaptive simulation (adaptiveSimulation / adaptiveProperties)
    adaptive: bool = False
    time_step_adaptation: bool = False
    subcell_adaptation: bool = False
    cell_weight_adaptation: bool = False
    adaptation_interval: int = 20
    adaptation_smoothing_passes: int = 25
    max_time_step_mct_ratio: float = 0.2
    max_courant_number: float = 0.5
    max_subcell_size_mfp_ratio: float = 0.5

    def species_index(self, name):
        for i, s in enumerate(self.species):
            if s.name == name:
                return i
        raise KeyError(name)

    def descriptors(self):
        """(mesh_desc, species_array, n_species, model_desc, bnd_array, n_bnd)."""
        md = self.mesh.desc()
        ns = len(self.species)
        sp = (abi.SpeciesDesc * ns)()
        for i, s in enumerate(self.species):
            sp[i].name = s.name.encode()[:31]
            sp[i].mass = s.mass
            sp[i].diameter = s.diameter
            sp[i].omega = s.omega
            sp[i].alpha = s.alpha
            sp[i].rotational_dof = s.rotational_dof
            sp[i].vibrational_modes = s.vibrational_modes
            sp[i].charge = s.charge
        m = abi.ModelDesc()
        m.n_equivalent_particles = self.n_equivalent_particles
        m.delta_t = self.delta_t
        m.collision_model = COLLISION_MODELS[self.collision_model]
        m.dsmc_model = DSMC_MODELS[self.dsmc_model]
        m.partner_model = abi.PARTNER_NTC
        m.bgk_model = BGK_MODELS[self.bgk_model]
        m.Tref = self.Tref
        m.Zrot = self.Zrot
        m.theta = self.theta
        m.macro_interpolation = 1 if self.macro_interpolation else 0
        m.sample_interval = self.sample_interval
        m.decomposition_interval = self.decomposition_interval
        m.smoothing_passes = self.smoothing_passes
        m.breakdown_max = self.breakdown_max
        m.kn_theta = self.kn_theta
        m.sampler_full = 1 if self.sampler_full else 0
        m.reset_at_decomposition = 1 if self.reset_at_decomposition else 0
        m.reset_until_time = self.reset_until_time
        m.cell_weighted = 1 if self.cell_weighted else 0
        m.particles_per_subcell = self.particles_per_subcell
        m.min_particles_per_subcell = self.min_particles_per_subcell
        m.max_smoothing_passes = self.max_smoothing_passes
        m.max_cell_weight_ratio = self.max_cell_weight_ratio
        m.adaptive = 1 if self.adaptive else 0
        m.time_step_adaptation = 1 if self.time_step_adaptation else 0
        m.subcell_adaptation = 1 if self.subcell_adaptation else 0
        m.cell_weight_adaptation = 1 if self.cell_weight_adaptation else 0
        m.adaptation_interval = self.adaptation_interval
        m.adaptation_smoothing_passes = self.adaptation_smoothing_passes if self.adaptation_smoothing_passes > 0 else -1
        m.max_time_step_mct_ratio = self.max_time_step_mct_ratio
        m.max_courant_number = self.max_courant_number
        m.max_subcell_size_mfp_ratio = self.max_subcell_size_mfp_ratio
        nb = len(self.boundaries)
        bd = (abi.BoundaryDesc * max(nb, 1))()
        for i, b in enumerate(self.boundaries):
            bd[i].patch = self.mesh.patch_index(b.patch)
            bd[i].kind = b.kind
            bd[i].temperature = b.temperature
            for d in range(3):
                bd[i].velocity[d] = b.velocity[d]
            bd[i].rot_temperature = b.rot_temperature
            for name, nd in b.number_density.items():
                bd[i].number_density[self.species_index(name)] = nd
        self._keep = (md, sp, m, bd)
        return md, sp, ns, m, bd, nb


def fill_desc(spec, number_density, T, rot_T=None, velocity=(0.0, 0.0, 0.0), exact_per_cell=0):
    """ugx_fill_desc; number_density: dict species->m^-3; T scalar or (Tx,Ty,Tz)."""
    f = abi.FillDesc()
    for name, nd in number_density.items():
        f.number_density[spec.species_index(name)] = nd
    Ts = (T, T, T) if isinstance(T, (int, float)) else tuple(T)
    for d in range(3):
        f.T[d] = Ts[d]
        f.velocity[d] = velocity[d]
    f.rot_temperature = (sum(Ts) / 3.0) if rot_T is None else rot_T
    f.exact_per_cell = exact_per_cell
    return f
