"""OpenFOAM case I/O without a GPU: polyMesh write/read round trips for the
workload meshes (graded, mapped, cyclic, empty), a hand-written polyMesh in
OpenFOAM's own layout (comments, header, a cyclic pair), the geometry the
library derives from a read mesh, and the sampled-field files under the
uniGasVolFields names."""
import os
import textwrap

import numpy as np
import pytest

from unigasfoam_amd import abi, foam_io, workloads
from unigasfoam_amd.core import geometry
from unigasfoam_amd.mesh import box_mesh, channel_mesh, graded_box_mesh


def _same_mesh(a, b):
    assert a.n_cells == b.n_cells
    for k in ("points", "face_offsets", "face_points", "owner", "neighbour"):
        x, y = np.asarray(getattr(a, k)), np.asarray(getattr(b, k))
        assert x.shape == y.shape and np.array_equal(x, y), k
    assert [(p.name, p.type, p.start, p.size, p.neighbour) for p in a.patches] == \
           [(p.name, p.type, p.start, p.size, p.neighbour) for p in b.patches]


@pytest.mark.parametrize("name", ["box", "graded", "channel_cyclic", "channel_empty", "nozzle", "cylinder"])
def test_polymesh_round_trip(tmp_path, name):
    if name == "box":
        m = box_mesh((0.0, 0.0, 0.0), (1.0e-3, 2.0e-3, 3.0e-3), (3, 4, 5))
    elif name == "graded":
        m = graded_box_mesh((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (4, 3, 2), grading=(3.0, 0.5, 1.0))
    elif name == "channel_cyclic":
        m = channel_mesh(12, 1.0e-3, 1.0e-4, periodic="cyclic")
    elif name == "channel_empty":
        m = channel_mesh(12, 1.0e-3, 1.0e-4, periodic="empty")
    elif name == "nozzle":
        m = workloads.nozzle(nx=6, ny=4, nz=4)[0].mesh
    else:
        m = workloads.cylinder(n_theta=8, n_r=4, target=1.0e3)[0].mesh
    foam_io.write_polymesh(m, tmp_path)
    _same_mesh(m, foam_io.read_polymesh(tmp_path))


def test_geometry_of_a_read_mesh_is_identical(tmp_path):
    spec, _ = workloads.couette(n_cells=10, per_cell=5)
    foam_io.write_polymesh(spec.mesh, tmp_path)
    g0 = geometry(spec)
    spec.mesh = foam_io.read_polymesh(tmp_path)
    g1 = geometry(spec)
    for k in g0:
        assert np.array_equal(np.asarray(g0[k]), np.asarray(g1[k])), k


HAND = {
    "points": """
        8
        (
        (0 0 0)
        (1 0 0)
        (1 1 0)
        (0 1 0)
        (0 0 1)
        (1 0 1)
        (1 1 1)
        (0 1 1)
        )
    """,
    "faces": """
        // six faces of one unit cube, outward from cell 0
        6
        (
        4(0 3 2 1)   // zmin
        4(4 5 6 7)   /* zmax */
        4(0 4 7 3)
        4(1 2 6 5)
        4(0 1 5 4)
        4(3 7 6 2)
        )
    """,
    "owner": "6\n(\n0\n0\n0\n0\n0\n0\n)\n",
    "neighbour": "0\n(\n)\n",
    "boundary": """
        3
        (
            zwalls
            {
                type            wall;
                inGroups        List<word> 1(wall);
                nFaces          2;
                startFace       0;
            }
            xperiodic0
            {
                type            cyclic;
                inGroups        List<word> 1(cyclic);
                nFaces          1;
                startFace       2;
                matchTolerance  0.0001;
                transform       unknown;
                neighbourPatch  xperiodic1;
            }
            xperiodic1
            {
                type            cyclic;
                nFaces          1;
                startFace       3;
                neighbourPatch  xperiodic0;
            }
        )
    """,
}


def test_reads_a_hand_written_polymesh(tmp_path):
    d = tmp_path / "constant" / "polyMesh"
    d.mkdir(parents=True)
    classes = {"points": "vectorField", "faces": "faceList", "owner": "labelList", "neighbour": "labelList",
               "boundary": "polyBoundaryMesh"}
    for name, body in HAND.items():
        hdr = ("/*--------------------------------*- C++ -*----------------------------------*\\\n"
               "  =========                 |\n\\*---------------------------------------------------------------------------*/\n"
               f"FoamFile\n{{\n    version     2.0;\n    format      ascii;\n    class       {classes[name]};\n"
               f"    object      {name};\n}}\n// * * * //\n")
        (d / name).write_text(hdr + textwrap.dedent(body))
    m = foam_io.read_polymesh(tmp_path)
    assert m.n_cells == 1 and m.n_faces == 6 and m.n_internal_faces == 0
    assert np.array_equal(m.points[6], [1.0, 1.0, 1.0])
    assert list(m.face_points[4:8]) == [4, 5, 6, 7]
    assert [(p.name, p.type, p.start, p.size, p.neighbour) for p in m.patches] == [
        ("zwalls", abi.WALL, 0, 2, -1), ("xperiodic0", abi.CYCLIC, 2, 1, 2), ("xperiodic1", abi.CYCLIC, 3, 1, 1)]


def test_refuses_binary(tmp_path):
    m = box_mesh((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (1, 1, 1))
    foam_io.write_polymesh(m, tmp_path)
    p = tmp_path / "constant" / "polyMesh" / "points"
    p.write_text(p.read_text().replace("format      ascii", "format      binary"))
    with pytest.raises(ValueError, match="ASCII"):
        foam_io.read_polymesh(tmp_path)


def test_field_files_use_the_volfields_names(tmp_path):
    m = channel_mesh(5, 1.0e-3, 1.0e-4, periodic="empty")
    rng = np.random.default_rng(1)
    n = m.n_cells
    fields = {"rhoN": rng.random(n) * 1e20, "translational_T": rng.random(n) * 300.0,
              "U_mean": rng.normal(size=(n, 3)), "shear_stress": rng.normal(size=(n, 6)), "mfp": rng.random(n),
              "heat_flux": rng.normal(size=(n, 3))}
    written = foam_io.write_fields(tmp_path, "0.001", m, fields, field_name="uniGasFields")
    assert sorted(written) == sorted(["rhoN_uniGasFields", "translationalT_uniGasFields", "UMean_uniGasFields",
                                      "pressureTensor_uniGasFields", "variableHardSphereMeanFreePath_uniGasFields",
                                      "heatFlux_uniGasFields",
                                      "surfaceHeatTransfer_uniGasFields", "surfaceShearStress_uniGasFields",
                                      "fD_uniGasFields"])
    for key, arr in fields.items():
        stem = foam_io.FIELD_FILES[key][0]
        back = foam_io.read_internal_field(os.path.join(tmp_path, "0.001", f"{stem}_uniGasFields"))
        assert np.array_equal(back, np.asarray(arr)), key
    text = open(os.path.join(tmp_path, "0.001", "rhoN_uniGasFields")).read()
    for p in m.patches:
        assert p.name in text
    assert "type            empty;" in text and "type            calculated;" in text


def test_boundary_values_follow_volfields(tmp_path):
    """Wall patches carry the wall measurements for T, p and UMean and the
    owner cell's value for rhoN; other patches the owner cell's value; the
    wall-only fields are zero inside (uniGasVolFields.C:1262-1392)."""
    m = channel_mesh(4, 1.0e-3, 1.0e-4)       # cyclic x/z, walls at y
    n = m.n_cells
    rng = np.random.default_rng(2)
    fields = {"rhoN": rng.random(n) * 1e20, "translational_T": 300.0 + rng.random(n), "p": rng.random(n),
              "U_mean": rng.normal(size=(n, 3))}
    lo, up = m.patch_index("lowerWall"), m.patch_index("upperWall")
    wf = {"translational_T": np.array([501.0, 502.0]), "rotational_T": np.array([0.0, 0.0]),
          "overall_T": np.array([503.0, 504.0]), "pressure": np.array([1.5, 2.5]),
          "U_mean": np.array([[1.0, 2.0, 3.0], [4.0, 5.0, 6.0]]), "heat_flux": np.array([-7.0, 8.0]),
          "shear": np.array([0.25, 0.5]), "force_density": np.array([[0.1, -1.5, 0.0], [0.2, 2.5, 0.0]])}
    wall = {"patch": np.array([lo, up]), "face": np.array([0, 0]), "fields": wf}
    foam_io.write_fields(tmp_path, "1", m, fields, wall=wall)
    d = os.path.join(tmp_path, "1")
    T = lambda name, patch: foam_io.read_patch_field(os.path.join(d, f"{name}_uniGasFields"), patch)
    assert T("translationalT", "lowerWall").tolist() == [501.0]
    assert T("translationalT", "upperWall").tolist() == [502.0]
    assert T("p", "upperWall").tolist() == [2.5]
    assert T("UMean", "lowerWall").tolist() == [[1.0, 2.0, 3.0]]
    assert T("surfaceHeatTransfer", "lowerWall").tolist() == [-7.0]
    assert T("surfaceShearStress", "upperWall").tolist() == [0.5]
    assert T("fD", "upperWall").tolist() == [[0.2, 2.5, 0.0]]
    owner_lo = m.owner[m.patches[lo].start]
    assert T("rhoN", "lowerWall").tolist() == [fields["rhoN"][owner_lo]]
    assert np.all(foam_io.read_internal_field(os.path.join(d, "surfaceHeatTransfer_uniGasFields")) == 0.0)
