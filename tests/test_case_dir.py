"""Case directories in OpenFOAM syntax (CPU): the dictionary reader, the
blockMesh subset, the mapping of the run-time selection words onto the ABI
enums, and the switches reported as unsupported.

The cases under tests/cases/ are written for these tests: heatBath is the
0-D heat bath (config C1), couetteN2 the 1-D N2 Couette flow (config C2),
platePlume a three-block N2/O2 flow over a plate with inflow, deletion and a
diffuse wall (config C3 in small).
"""
import math
import os
import shutil

import numpy as np
import pytest

from oracle import orc
from unigasfoam_amd import abi, blockmesh, case_dir, core, foam_dict, workloads

CASES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cases")
REF_TUT = "/root/reference/tutorials/uniGasFoam"


# ------------------------------------------------------------------ parser

def test_parser_syntax():
    text = r'''
    FoamFile { version 2.0; format ascii; class dictionary; object x; }
    /* block
       comment */ a 1;   // line comment
    b  -2.5e-3;
    w  word;
    s  "quoted string";
    v  (1 2 3);
    nested { x 1; y { z (a b); } }
    numberDensities{Ar 4.2e20; N2 1e21;}
    lst ( boundary { patch p; } boundary { patch q; } );
    counted 3(4 5 6);
    dims [0 1 -1 0 0 0 0];
    ref $a;
    multi hex (0 1 2 3) (8 8 8) simpleGrading (1 1 1);
    '''
    d = foam_dict.parse(text)
    assert "FoamFile" not in d
    assert d["a"] == 1 and d.scalar("b") == -2.5e-3 and d["w"] == "word" and d["s"] == "quoted string"
    assert d.vector("v") == (1.0, 2.0, 3.0)
    assert d.sub("nested").sub("y")["z"] == ["a", "b"]
    assert d.sub("numberDensities") == {"Ar": 4.2e20, "N2": 1e21}
    assert [name for name, _ in d["lst"]] == ["boundary", "boundary"]
    assert d["lst"][1][1].word("patch") == "q"
    assert d["counted"] == [4, 5, 6]
    assert d["dims"] == [0, 1, -1, 0, 0, 0, 0]
    assert d["ref"] == 1
    assert isinstance(d["multi"], foam_dict.Tokens) and d["multi"][0] == "hex"
    for word, val in (("yes", True), ("off", False), ("true", True), ("no", False), ("on", True)):
        assert foam_dict.as_switch(word) is val


@pytest.mark.parametrize("text", ["a 1", "a { b 1;", "#include \"x\"\na 1;", "a (1 2;", "r $missing;"])
def test_parser_refuses_bad_input(text):
    with pytest.raises(foam_dict.DictError):
        foam_dict.parse(text)


# ------------------------------------------------------------------ cases

def _same_spec(a, b, rtol=1e-12):
    assert a.mesh.n_cells == b.mesh.n_cells
    assert np.allclose(a.mesh.points, b.mesh.points, rtol=rtol, atol=1e-18)
    for k in ("face_offsets", "face_points", "owner", "neighbour"):
        assert np.array_equal(getattr(a.mesh, k), getattr(b.mesh, k)), k
    assert [(p.name, p.type, p.start, p.size, p.neighbour) for p in a.mesh.patches] == \
           [(p.name, p.type, p.start, p.size, p.neighbour) for p in b.mesh.patches]
    for k in ("delta_t", "n_equivalent_particles", "Tref", "theta"):
        assert np.isclose(getattr(a, k), getattr(b, k), rtol=rtol), k
    for k in ("collision_model", "dsmc_model", "bgk_model", "sample_interval"):
        assert getattr(a, k) == getattr(b, k), k
    if a.dsmc_model.startswith("LarsenBorgnakke"):
        assert a.Zrot == b.Zrot
    assert [(s.name, s.mass, s.diameter, s.omega, s.alpha, s.rotational_dof) for s in a.species] == \
           [(s.name, s.mass, s.diameter, s.omega, s.alpha, s.rotational_dof) for s in b.species]
    ka = sorted((x.patch, x.kind, x.temperature, tuple(x.velocity)) for x in a.boundaries)
    kb = sorted((x.patch, x.kind, x.temperature, tuple(x.velocity)) for x in b.boundaries)
    assert ka == kb


def test_heat_bath_case_is_config_c1():
    spec, fills, run = case_dir.read_case(os.path.join(CASES, "heatBath"), seed=1)
    ref, fill = workloads.heat_bath()
    _same_spec(spec, ref)
    assert len(fills) == 1
    f = fills[0]
    assert f["number_density"] == fill["number_density"] and f["T"] == fill["T"]
    assert f["exact_per_cell"] == fill["exact_per_cell"]
    assert run.end_time / run.delta_t == pytest.approx(2000)
    # the same case, bit for bit: the mesh points and every model constant agree exactly
    assert np.array_equal(spec.mesh.points, ref.mesh.points)
    a, b = orc.OracleCase(spec, capacity=1 << 16), orc.OracleCase(ref, capacity=1 << 16)
    a.fill(**f)
    b.fill(**fill)
    a.step(5)
    b.step(5)
    pa, pb = a.particles(), b.particles()
    for k in ("local", "velocity", "cell"):
        assert np.array_equal(pa[k], pb[k]), k
    assert a.counters()["collisions"] == b.counters()["collisions"] > 0


def test_couette_case_is_config_c2():
    spec, fills, run = case_dir.read_case(os.path.join(CASES, "couetteN2"), seed=2)
    ref, fill = workloads.couette()
    _same_spec(spec, ref)
    assert spec.sampler_full
    assert fills[0]["exact_per_cell"] == 200 and fills[0]["T"] == 273.0
    assert run.end_time / run.delta_t == pytest.approx(20000)
    assert run.write_interval == pytest.approx(5e-5)          # timeStep control: 10000 steps
    assert run.reset_at_output_until_time == pytest.approx(5e-5)


def test_plate_case_blocks_merge_into_valid_geometry():
    spec, fills, run = case_dir.read_case(os.path.join(CASES, "platePlume"))
    m = spec.mesh
    assert m.n_cells == (10 + 30 + 20) * 40
    sizes = {p.name: (p.type, p.size) for p in m.patches}
    assert sizes == {"inlet": (abi.PATCH, 100), "outlet": (abi.PATCH, 40), "plate": (abi.WALL, 30),
                     "symmetry": (abi.SYMMETRY, 30), "frontAndBack": (abi.EMPTY, 4800)}
    # every internal face is shared by two cells: the blocks' common faces were merged
    assert m.n_internal_faces == 2 * 60 * 40 - 60 - 40
    g = core.geometry(spec)            # host-only: convex, planar, closed cells
    assert np.isclose(g["vol"].sum(), 3e-3 * 2e-3 * 5e-5, rtol=1e-12)
    assert np.all(g["vol"] > 0)
    kinds = sorted((b.patch, b.kind) for b in spec.boundaries)
    assert kinds == [("inlet", abi.BND_DELETION), ("inlet", abi.BND_FREESTREAM), ("outlet", abi.BND_DELETION),
                     ("plate", abi.BND_DIFFUSE)]
    inflow = [b for b in spec.boundaries if b.kind == abi.BND_FREESTREAM][0]
    assert inflow.number_density == {"N2": 7.9e20, "O2": 2.1e20}
    assert len(spec.species) == 2 and spec.Zrot == 5.0


def test_unsupported_switches_are_named(tmp_path):
    case = tmp_path / "c"
    shutil.copytree(os.path.join(CASES, "heatBath"), case)
    p = case / "constant" / "uniGasProperties"
    text = p.read_text()
    text = text.replace("axisymmetricSimulation          false;", "axisymmetricSimulation          true;")
    text = text.replace("adaptiveSimulation              false;",
                        "adaptiveSimulation              on;\nadaptiveProperties { cellWeightAdaptation true; }")
    text = text.replace("dsmcCollisionPartnerModel       noTimeCounter;",
                        "dsmcCollisionPartnerModel       noTimeCounterSubCycled;")
    p.write_text(text)
    with pytest.raises(case_dir.Unsupported) as e:
        case_dir.read_case(str(case))
    assert e.value.code == abi.ERR_UNSUPPORTED == 7
    words = " ".join(w for _, w in e.value.items)
    assert "axisymmetricSimulation" in words and "cellWeightAdaptation without cellWeightedSimulation" in words
    assert "noTimeCounterSubCycled" in words
    assert len(e.value.items) == 3


def test_adaptive_properties_are_read(tmp_path):
    case = tmp_path / "c"
    shutil.copytree(os.path.join(CASES, "heatBath"), case)
    p = case / "constant" / "uniGasProperties"
    text = p.read_text().replace("adaptiveSimulation              false;",
                                 "adaptiveSimulation              true;\nadaptiveProperties\n{\n"
                                 "    timeStepAdaptation true;\n    subCellAdaptation yes;\n    adaptationInterval 50;\n"
                                 "    maxCourantNumber 0.4;\n}")
    p.write_text(text)
    spec, fills, run = case_dir.read_case(str(case))
    assert spec.adaptive and spec.time_step_adaptation and spec.subcell_adaptation
    assert not spec.cell_weight_adaptation
    assert spec.adaptation_interval == 50 and spec.max_courant_number == 0.4
    assert spec.max_time_step_mct_ratio == 0.2 and spec.adaptation_smoothing_passes == 25


def test_cell_weighting_switch_is_read(tmp_path):
    case = tmp_path / "c"
    shutil.copytree(os.path.join(CASES, "heatBath"), case)
    p = case / "constant" / "uniGasProperties"
    text = p.read_text().replace("cellWeightedSimulation          false;",
                                 "cellWeightedSimulation          true;\ncellWeightedProperties\n{\n"
                                 "    minParticlesPerSubCell 10;\n    particlesPerSubCell 20;\n}")
    p.write_text(text)
    spec, fills, run = case_dir.read_case(str(case))
    assert spec.cell_weighted and spec.particles_per_subcell == 20 and spec.min_particles_per_subcell == 10


def test_missing_files_are_io_errors(tmp_path):
    with pytest.raises(case_dir.CaseError) as e:
        case_dir.read_case(str(tmp_path))
    assert e.value.code == abi.ERR_IO


@pytest.mark.skipif(not os.path.isdir(REF_TUT), reason="reference tutorials not present")
def test_reference_tutorials_report_exactly_their_unsupported_switches():
    common = {"uniGasProperties": []}
    expect = {
        "expansionInVacuum": {"uniGasProperties": common["uniGasProperties"] + ["axisymmetricSimulation"],
                              "fieldPropertiesDict": ["fieldModel uniGasMassFluxSurface"],
                              "uniGasInitialisationDict": ["configuration type uniGasMeshFieldFill"]},
        "hypersonicCylinder": {"uniGasProperties": common["uniGasProperties"]},
        "plumeImpingement": {"uniGasProperties": common["uniGasProperties"] + ["axisymmetricSimulation"],
                             "blockMeshDict": ["curved edges ('polyLine'"],
                             "boundariesDict": ["uniGasLiouFangPressureInletPatch on inlet"],
                             "uniGasInitialisationDict": ["configuration type uniGasMeshFieldFill"]},
        "supersonicPlate": {"uniGasProperties": common["uniGasProperties"]},
    }
    for tut, want in expect.items():
        want = {f: items for f, items in want.items() if items}
        inp = case_dir.scan_case(os.path.join(REF_TUT, tut))
        got = {}
        for f, w in inp.unsupported:
            got.setdefault(f, []).append(w)
        assert sorted(got) == sorted(want), (tut, got)
        for f, items in want.items():
            assert len(got[f]) == len(items), (tut, f, got[f])
            for prefix, g in zip(items, got[f]):
                assert g.startswith(prefix), (tut, f, g)
    # the cylinder tutorial's arc-edged half annulus builds: its volume is
    # the half annulus less the chords' segments (100 per quarter arc)
    inp = case_dir.scan_case(os.path.join(REF_TUT, "hypersonicCylinder"))
    assert inp.spec.mesh.n_cells == 20000
    ft = 0.3048
    seg = 1.0 - math.sin(0.5 * math.pi / 100) / (0.5 * math.pi / 100)
    want = 0.5 * math.pi * (4.0 - 0.25) * ft ** 2 * (0.1 * ft) * (1.0 - seg)
    assert core.geometry(inp.spec)["vol"].sum() == pytest.approx(want, rel=1e-9)
    # the plate tutorial's three-block mesh builds and nothing in it is
    # unsupported: its Mach-3 argon case runs as the reference configures it
    inp = case_dir.scan_case(os.path.join(REF_TUT, "supersonicPlate"))
    assert inp.spec.mesh.n_cells == 45000
    assert inp.spec.collision_model == "hybrid" and inp.spec.bgk_model == "unifiedStochasticParticleSBGK"
    # its adaptive properties are read (uniGasProperties:57-66 of the tutorial)
    sp = inp.spec
    assert sp.macro_interpolation
    assert sp.adaptive and sp.time_step_adaptation and sp.subcell_adaptation and sp.cell_weight_adaptation
    assert sp.cell_weighted and sp.adaptation_interval == 10 and sp.max_subcell_size_mfp_ratio == 0.5
    assert np.all(core.geometry(inp.spec)["vol"] > 0)


QUARTER_ANNULUS = """scale 0.5;
vertices ((1 0 0) (3 0 0) (0 3 0) (0 1 0) (1 0 0.2) (3 0 0.2) (0 3 0.2) (0 1 0.2));
blocks ( hex (0 1 2 3 4 5 6 7) (4 12 1) simpleGrading (2 1 1) );
edges ( {edges} );
boundary ( w {{ type wall; faces ((3 7 4 0) (1 2 6 5)); }} f {{ type empty; faces ((0 3 2 1) (4 5 6 7)); }} );"""


@pytest.mark.parametrize("form", ["point", "angle"])
def test_blockmesh_arc_edges(form):
    """A quarter annulus r in [1, 3] (scale 0.5) from one block with arcs on
    both curved sides: every point lies on a ray of the angle its column
    has and at the radius of its row; the volume is the exact annulus area
    less the chords' segments, (N/2) r^2 (theta - sin theta) per arc of N
    segments of angle theta; the 'angle (axis)' form gives the same mesh as
    the form through a point."""
    h = math.sqrt(0.5)
    if form == "point":
        e = f"arc 1 2 ({3 * h} {3 * h} 0) arc 0 3 ({h} {h} 0) arc 5 6 ({3 * h} {3 * h} 0.2) arc 4 7 ({h} {h} 0.2)"
    else:
        e = "arc 1 2 90 (0 0 1) arc 0 3 90 (0 0 1) arc 5 6 90 (0 0 1) arc 4 7 90 (0 0 1)"
    d = foam_dict.parse(QUARTER_ANNULUS.format(edges=e), counted=False)
    m = blockmesh.build(d)
    assert m.n_cells == 48
    P = m.points / 0.5
    r = np.hypot(P[:, 0], P[:, 1])
    # radii of the graded rows (x direction 1 -> 3 with expansion 2 over 4 cells)
    rows = 1.0 + 2.0 * blockmesh._graded(4, 2.0)
    assert np.allclose(np.sort(np.unique(np.round(r, 12))), rows, atol=1e-12)
    ang = np.arctan2(P[:, 1], P[:, 0])
    assert np.allclose(np.sort(np.unique(np.round(ang, 12))), np.linspace(0.0, 0.5 * math.pi, 13), atol=1e-12)
    vol = core.geometry(case_dir_spec(m))["vol"].sum() / 0.5 ** 3
    th = 0.5 * math.pi / 12
    area = sum(0.5 * 12 * (rows[k + 1] ** 2 - rows[k] ** 2) * math.sin(th) for k in range(4))
    assert vol == pytest.approx(area * 0.2, rel=1e-12)
    assert area < 0.25 * math.pi * (9 - 1)


def case_dir_spec(mesh):
    from unigasfoam_amd.spec import Boundary, CaseSpec, Species
    return CaseSpec(mesh=mesh, species=[Species("Ar", mass=6.63e-26, diameter=4.17e-10, omega=0.81)], delta_t=1e-6,
                    n_equivalent_particles=1.0, collision_model="dsmc", boundaries=[Boundary("w", abi.BND_SPECULAR)])


def test_arc_vertex_labels_are_not_list_counts():
    """'arc 2 3 (x y z)': with counted lists off the second label stays a
    label although a three-entry list follows it."""
    d = foam_dict.parse("edges ( arc 2 3 (1 2 3) arc 0 1 (4 5 6) );", counted=False)
    assert d["edges"] == ["arc", 2, 3, [1, 2, 3], "arc", 0, 1, [4, 5, 6]]
    assert foam_dict.parse("v 3 (1 2 3);")["v"] == [1, 2, 3]


def test_blockmesh_refuses_what_it_does_not_build():
    base = """vertices ((0 0 0) (1 0 0) (1 1 0) (0 1 0) (0 0 1) (1 0 1) (1 1 1) (0 1 1));
    blocks ( hex (0 1 2 3 4 5 6 7) (2 2 2) simpleGrading (1 1 1) );
    boundary ( w { type wall; faces ((0 4 7 3)); } );"""
    m = blockmesh.build(foam_dict.parse(base))
    assert [p.name for p in m.patches] == ["w", "defaultFaces"]
    for extra in ["edges ( polyLine 0 1 ((0.5 -0.1 0)) );", "edges ( spline 0 1 ((0.5 -0.1 0)) );",
                  "mergePatchPairs ( (a b) );"]:
        with pytest.raises(blockmesh.Unsupported):
            blockmesh.build(foam_dict.parse(base + extra))
    with pytest.raises(blockmesh.Unsupported):
        blockmesh.build(foam_dict.parse(base.replace("simpleGrading (1 1 1)", "edgeGrading (1 1 1 1 1 1 1 1 1 1 1 1)")))
