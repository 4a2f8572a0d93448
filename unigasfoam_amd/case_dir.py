"""uniGasFoam case directories: the dictionaries of a case read into the
library's case description.

    spec, fills, run = read_case("path/to/case")

reads what uniGasFoam reads at start-up and maps every run-time selection
word onto the ABI's enums:

  constant/uniGasProperties          uniGasCloud ctor (clouds/uniGasCloud.C:396-781):
                                     nEquivalentParticles, typeIdList, moleculeProperties
                                     (parcels/uniGasParcelI.H:40-60), collisionModel,
                                     dsmcCollisionModel, bgkCollisionModel,
                                     dsmcCollisionPartnerModel, collisionProperties
                                     (LarsenBorgnakkeVariableHardSphere.C:59-66,
                                     stochasticParticleBGK.C:53-54,
                                     unifiedStochasticParticleSBGK.C:53-58),
                                     cellWeightedProperties, adaptiveProperties
                                     (dynamicAdaptation/uniGasDynamicAdapter.C:187-217)
  system/controlDict                 deltaT, endTime, writeInterval, nTerminalOutputs
  system/boundariesDict              uniGasBoundaries (boundaries/basic/uniGasBoundaries/
                                     uniGasBoundaries.C:33-556): patch, general and cyclic models
  system/uniGasInitialisationDict    configurations (uniGasMeshFill.C:58-100)
  system/hybridDecompositionDict     uniGasHybridDecomposition.C:58-76, localKnudsen.C:54-56
  system/fieldPropertiesDict         uniGasField.C:53-55, uniGasVolFields.C:54-60
  constant/polyMesh or system/blockMeshDict   the mesh (foam_io.read_polymesh, blockmesh.build)

Defaults are the reference's (``getOrDefault`` values of the cited lines).
A switch or model this build does not run (axisymmetry, chemistry,
charged species, vibrational or electronic exchange, other model names) is
collected, and
``read_case`` raises ``Unsupported`` (code ``abi.ERR_UNSUPPORTED``, the
C ABI's UGX_ERR_UNSUPPORTED) listing every one of them; ``scan_case``
returns the list without raising.

Two keys of uniGasMeshFill entries are extensions (uniGasFoam ignores
them): ``translationalTemperatureComponents (Tx Ty Tz)`` fills with per-axis
temperatures and ``exactParticlesPerCell n`` places exactly n simulators in
each cell; both reproduce the anisotropic heat-bath start.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

from . import abi, blockmesh
from .foam_dict import DictError, FoamDict, Tokens, as_switch
from .foam_dict import read as read_dict
from .spec import BGK_MODELS, DSMC_MODELS, Boundary, CaseSpec, Species


class CaseError(DictError):
    """A case directory that cannot be read (missing file or entry)."""
    code = abi.ERR_IO


class Unsupported(DictError):
    """Switches and models the case selects that this build does not run."""
    code = abi.ERR_UNSUPPORTED

    def __init__(self, items):
        self.items = list(items)
        super().__init__("UGX_ERR_UNSUPPORTED: " + "; ".join(f"{f}: {w}" for f, w in self.items))


@dataclass
class RunControl:
    """system/controlDict and the sampler's output settings."""
    delta_t: float
    end_time: float
    start_time: float = 0.0
    write_interval: float = 0.0
    n_terminal_outputs: int = 1
    reset_at_output: bool = True
    reset_at_output_until_time: float = float("inf")
    field_name: str = "uniGasFields"


@dataclass
class CaseInputs:
    spec: CaseSpec = None
    fills: list = field(default_factory=list)
    run: RunControl = None
    unsupported: list = field(default_factory=list)


PATCH_MODELS = {"uniGasSpecularWallPatch": abi.BND_SPECULAR, "uniGasDiffuseWallPatch": abi.BND_DIFFUSE,
                "uniGasDeletionPatch": abi.BND_DELETION}
GENERAL_MODELS = {"uniGasFreeStreamInflowPatch": abi.BND_FREESTREAM}
COLLISION_WORDS = ("dsmc", "bgk", "hybrid")
_UNSUPPORTED_SWITCHES = (
    ("axisymmetricSimulation", "axisymmetric radial weighting"),
    ("chemicalReactions", "chemistry"),
    ("chargedParticles", "charged species"),
)


def _path(case, *parts):
    return os.path.join(case, *parts)


def _need(case, *parts):
    p = _path(case, *parts)
    if not os.path.exists(p):
        raise CaseError(f"{p} is missing")
    return read_dict(p)


def _species(props, bad):
    ids = props.get("typeIdList")
    if not isinstance(ids, list) or not ids:
        raise CaseError("uniGasProperties: typeIdList missing")
    if len(ids) > abi.MAX_SPECIES:
        bad.append(("uniGasProperties", f"{len(ids)} species (at most {abi.MAX_SPECIES})"))
    mol = props.sub("moleculeProperties")
    out = []
    for name in ids:
        d = mol.sub(name)
        sp = Species(name, mass=d.scalar("mass"), diameter=d.scalar("diameter"), omega=d.scalar("omega"),
                     alpha=d.scalar("alpha"), rotational_dof=d.label("rotationalDegreesOfFreedom"),
                     vibrational_modes=d.label("vibrationalModes", 0), charge=d.label("charge", 0))
        if sp.vibrational_modes > 0:
            bad.append(("uniGasProperties", f"{name}: vibrationalModes {sp.vibrational_modes} (vibrational exchange)"))
        if sp.charge != 0:
            bad.append(("uniGasProperties", f"{name}: charge {sp.charge} (charged species)"))
        if d.label("numberOfElectronicLevels", 1) > 1:
            bad.append(("uniGasProperties", f"{name}: numberOfElectronicLevels > 1 (electronic exchange)"))
        out.append(sp)
    return out


def _model(props, spec_kw, bad):
    cm = props.word("collisionModel")
    if cm not in COLLISION_WORDS:
        raise CaseError(f"uniGasProperties: unknown collisionModel {cm} (valid: {', '.join(COLLISION_WORDS)})")
    spec_kw["collision_model"] = cm
    cp = props.sub_or_empty("collisionProperties")
    if cm in ("dsmc", "hybrid"):
        dm = props.word("dsmcCollisionModel")
        if dm not in DSMC_MODELS:
            bad.append(("uniGasProperties", f"dsmcCollisionModel {dm}"))
            dm = "noDSMCCollision"
        spec_kw["dsmc_model"] = dm
        pm = props.word("dsmcCollisionPartnerModel", "noTimeCounter")
        if pm != "noTimeCounter":
            bad.append(("uniGasProperties", f"dsmcCollisionPartnerModel {pm} (only noTimeCounter)"))
        if dm.startswith("LarsenBorgnakke"):
            spec_kw["Zrot"] = cp.scalar("rotationalRelaxationCollisionNumber")
    else:
        spec_kw["dsmc_model"] = "noDSMCCollision"
    if cm in ("bgk", "hybrid"):
        bm = props.word("bgkCollisionModel")
        if bm not in BGK_MODELS:
            bad.append(("uniGasProperties", f"bgkCollisionModel {bm}"))
            bm = "noBGKCollision"
        spec_kw["bgk_model"] = bm
        if bm == "unifiedStochasticParticleSBGK":
            spec_kw["theta"] = cp.scalar("theta", 1.0)
        # cellPoint interpolation of the fields at each relaxed simulator (stochasticParticleBGK.C:54, 645-688)
        spec_kw["macro_interpolation"] = bm != "noBGKCollision" and as_switch(cp.get("macroInterpolation", False),
                                                                             "macroInterpolation")
    else:
        spec_kw["bgk_model"] = "noBGKCollision"
    if spec_kw["dsmc_model"] != "noDSMCCollision" or spec_kw["bgk_model"] != "noBGKCollision":
        spec_kw["Tref"] = cp.scalar("Tref")
    # cell weighting (uniGasCloud.C:555-560): the weights come from the fill
    if "cellWeightedSimulation" in props and as_switch(props["cellWeightedSimulation"], "cellWeightedSimulation"):
        cw = props.sub("cellWeightedProperties")
        spec_kw["cell_weighted"] = True
        spec_kw["particles_per_subcell"] = cw.label("particlesPerSubCell")
        spec_kw["min_particles_per_subcell"] = cw.label("minParticlesPerSubCell", spec_kw["particles_per_subcell"])
    # dynamic adaptation (uniGasCloud.C:416, uniGasDynamicAdapter.C:187-217)
    if "adaptiveSimulation" in props and as_switch(props["adaptiveSimulation"], "adaptiveSimulation"):
        ad = props.sub_or_empty("adaptiveProperties")
        spec_kw["adaptive"] = True
        for key, kw in (("timeStepAdaptation", "time_step_adaptation"), ("subCellAdaptation", "subcell_adaptation"),
                        ("cellWeightAdaptation", "cell_weight_adaptation")):
            spec_kw[kw] = as_switch(ad.get(key, False), key)
        spec_kw["adaptation_interval"] = ad.label("adaptationInterval", 20)
        spec_kw["adaptation_smoothing_passes"] = ad.label("smoothingPasses", 25)
        spec_kw["max_time_step_mct_ratio"] = ad.scalar("maxTimeStepMCTRatio", 0.2)
        spec_kw["max_courant_number"] = ad.scalar("maxCourantNumber", 0.5)
        spec_kw["max_subcell_size_mfp_ratio"] = ad.scalar("maxSubCellSizeMFPRatio", 0.5)
        if spec_kw["cell_weight_adaptation"] and not spec_kw.get("cell_weighted"):
            bad.append(("uniGasProperties", "adaptiveProperties.cellWeightAdaptation without cellWeightedSimulation"))
    for key, what in _UNSUPPORTED_SWITCHES:
        if key in props and as_switch(props[key], key):
            bad.append(("uniGasProperties", f"{key} ({what})"))


def _boundaries(case, patches, species, bad):
    p = _path(case, "system", "boundariesDict")
    if not os.path.exists(p):
        return []
    d = read_dict(p)
    names = {s.name for s in species}
    out = []
    for kind_key, table, props_key in (("uniGasPatchBoundaries", PATCH_MODELS, "patchBoundaryProperties"),
                                       ("uniGasGeneralBoundaries", GENERAL_MODELS, "generalBoundaryProperties")):
        for item in d.get(kind_key, []):
            if not (isinstance(item, tuple) and item[0] == "boundary"):
                raise CaseError(f"boundariesDict: {kind_key} holds {item!r}")
            bd = item[1]
            patch = bd.sub(props_key).word("patch")
            model = bd.word("boundaryModel")
            if patch not in patches:
                raise CaseError(f"boundariesDict: patch {patch} is not in the mesh")
            if model not in table:
                bad.append(("boundariesDict", f"{model} on {patch}"))
                continue
            md = bd.sub_or_empty(model + "Properties")
            kind = table[model]
            if kind == abi.BND_DIFFUSE:
                out.append(Boundary(patch, kind, temperature=md.scalar("temperature"), velocity=md.vector("velocity")))
            elif kind == abi.BND_FREESTREAM:
                nd = md.sub("numberDensities")
                dens = {}
                for k, v in nd.items():
                    if k not in names:
                        raise CaseError(f"boundariesDict: inflow species {k} is not in typeIdList")
                    dens[k] = float(v)
                out.append(Boundary(patch, kind, temperature=md.scalar("translationalTemperature"),
                                    velocity=md.vector("velocity"), rot_temperature=md.scalar("rotationalTemperature"),
                                    number_density=dens))
            else:
                out.append(Boundary(patch, kind))
    for item in d.get("uniGasCyclicBoundaries", []):
        model = item[1].get("boundaryModel", "?") if isinstance(item, tuple) else item
        bad.append(("boundariesDict", f"cyclic boundary model {model}"))
    return out


def _fills(case, species, bad):
    p = _path(case, "system", "uniGasInitialisationDict")
    if not os.path.exists(p):
        return []
    d = read_dict(p)
    names = {s.name for s in species}
    out = []
    for item in d.get("configurations", []):
        if not (isinstance(item, tuple) and item[0] == "configuration"):
            raise CaseError(f"uniGasInitialisationDict: configurations holds {item!r}")
        c = item[1]
        t = c.word("type")
        if t != "uniGasMeshFill":
            bad.append(("uniGasInitialisationDict", f"configuration type {t}"))
            continue
        dens = {}
        for k, v in c.sub("numberDensities").items():
            if k not in names:
                raise CaseError(f"uniGasInitialisationDict: species {k} is not in typeIdList")
            dens[k] = float(v)
        T = c.scalar("translationalTemperature")
        if "translationalTemperatureComponents" in c:
            T = c.vector("translationalTemperatureComponents")
        out.append(dict(number_density=dens, T=T, rot_T=c.scalar("rotationalTemperature"),
                        velocity=c.vector("velocity"), exact_per_cell=c.label("exactParticlesPerCell", 0)))
    return out


def _hybrid(case, spec_kw, bad):
    p = _path(case, "system", "hybridDecompositionDict")
    if spec_kw.get("collision_model") != "hybrid":
        return
    if not os.path.exists(p):
        raise CaseError(f"{p} is missing (collisionModel hybrid)")
    d = read_dict(p)
    model = d.word("decompositionModel")
    if model != "localKnudsen":
        bad.append(("hybridDecompositionDict", f"decompositionModel {model}"))
        return
    tp = d.sub("timeProperties")
    spec_kw["decomposition_interval"] = tp.label("decompositionInterval", 100)
    spec_kw["reset_at_decomposition"] = tp.switch("resetAtDecomposition", True)
    spec_kw["reset_until_time"] = tp.scalar("resetAtDecompositionUntilTime", 0.0)
    lk = d.sub_or_empty("localKnudsenProperties")
    spec_kw["breakdown_max"] = lk.scalar("breakdownMax", 0.05)
    spec_kw["kn_theta"] = lk.scalar("theta", 1.0)
    spec_kw["smoothing_passes"] = lk.label("smoothingPasses", 0)


def _fields(case, species, spec_kw, run, bad):
    p = _path(case, "system", "fieldPropertiesDict")
    if not os.path.exists(p):
        return
    d = read_dict(p)
    names = [s.name for s in species]
    seen = False
    for item in d.get("uniGasFields", []):
        fd = item[1] if isinstance(item, tuple) else None
        if fd is None:
            raise CaseError(f"fieldPropertiesDict: uniGasFields holds {item!r}")
        model = fd.word("fieldModel")
        if model != "uniGasVolFields":
            bad.append(("fieldPropertiesDict", f"fieldModel {model}"))
            continue
        if seen:
            bad.append(("fieldPropertiesDict", "more than one uniGasVolFields entry"))
            continue
        seen = True
        tp = fd.sub_or_empty("timeProperties")
        spec_kw["sample_interval"] = tp.label("sampleInterval", 1)
        run.reset_at_output = tp.switch("resetAtOutput", True)
        run.reset_at_output_until_time = tp.scalar("resetAtOutputUntilTime", float("inf"))
        vp = fd.sub("uniGasVolFieldsProperties")
        run.field_name = vp.word("field")
        ids = vp.get("typeIds", names)
        if sorted(ids) != sorted(names):
            bad.append(("fieldPropertiesDict", f"uniGasVolFields over typeIds {ids} (this build samples all species)"))
        spec_kw["sampler_full"] = vp.switch("measureHeatFluxShearStress", False)


def scan_case(case, seed=12345) -> CaseInputs:
    """Everything the case selects; inputs.unsupported lists what this build
    does not run (empty when read_case would succeed)."""
    if not os.path.isdir(case):
        raise CaseError(f"{case} is not a directory")
    bad = []
    props = _need(case, "constant", "uniGasProperties")
    ctrl = _need(case, "system", "controlDict")
    patch_names = None
    if os.path.isdir(_path(case, "constant", "polyMesh")):
        from .foam_io import read_polymesh
        mesh = read_polymesh(case)
    elif os.path.exists(_path(case, "system", "blockMeshDict")):
        bm = read_dict(_path(case, "system", "blockMeshDict"), counted=False)
        try:
            mesh = blockmesh.build(bm)
        except blockmesh.Unsupported as e:
            bad.append(("blockMeshDict", str(e)))
            mesh = None
            patch_names = {item[0] for item in bm.get("boundary", []) if isinstance(item, tuple)}
    else:
        raise CaseError(f"{case}: neither constant/polyMesh nor system/blockMeshDict")
    if mesh is not None:
        patch_names = {q.name for q in mesh.patches}
    species = _species(props, bad)
    spec_kw = dict(n_equivalent_particles=props.scalar("nEquivalentParticles"))
    _model(props, spec_kw, bad)
    run = RunControl(delta_t=ctrl.scalar("deltaT"), end_time=ctrl.scalar("endTime"),
                     start_time=ctrl.scalar("startTime", 0.0), write_interval=ctrl.scalar("writeInterval", 0.0),
                     n_terminal_outputs=ctrl.label("nTerminalOutputs", 1))
    if ctrl.word("writeControl", "timeStep") not in ("timeStep", "runTime", "adjustableRunTime"):
        bad.append(("controlDict", f"writeControl {ctrl.word('writeControl')}"))
    if ctrl.word("writeControl", "timeStep") == "timeStep":
        run.write_interval = run.write_interval * run.delta_t
    _hybrid(case, spec_kw, bad)
    _fields(case, species, spec_kw, run, bad)
    boundaries = _boundaries(case, patch_names, species, bad)
    fills = _fills(case, species, bad)
    spec = None
    if mesh is not None:
        spec = CaseSpec(mesh=mesh, species=species, delta_t=run.delta_t, boundaries=boundaries, seed=seed, **spec_kw)
    return CaseInputs(spec=spec, fills=fills, run=run, unsupported=bad)


def read_case(case, seed=12345):
    """(CaseSpec, list of fill kwargs, RunControl) of a case directory;
    raises Unsupported listing every switch this build does not run."""
    inp = scan_case(case, seed)
    if inp.unsupported:
        raise Unsupported(inp.unsupported)
    return inp.spec, inp.fills, inp.run
