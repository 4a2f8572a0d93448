"""uniGasFoam's cloud restart layout (uniGasParcelIO.C:89-181,
uniGasCloud.C:577-703): <time>/lagrangian/uniGas with the parcel fields and
positions, the restart cell fields and uniform/time.

Checked on the CPU: a written cloud reads back to the same doubles; a
hand-written file set in OpenFOAM's own syntax (FoamFile header, comments,
the legacy "(x y z) celli" positions) is read; OpenFOAM's barycentric
"coordinates" form gives the Cartesian positions that the tet vertices fix
(the coordinates (1 0 0 0) are the cell centroid, (0 1 0 0) the face's
first point, equal weights the tet's centroid); OpenFOAM's cell centroid of
a box cell is its centre and of a wedge-shaped cell the volume centroid;
excited electronic or vibrational states are refused."""
import os

import numpy as np
import pytest

from oracle.orc import OracleCase
from unigasfoam_amd import abi, foam_io, workloads
from unigasfoam_amd.mesh import box_mesh, mapped_box_mesh

SIDES = ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")
HDR = """/*--------------------------------*- C++ -*----------------------------------*\\
  =========                 |
\\*---------------------------------------------------------------------------*/
FoamFile
{{
    version     2.0;
    format      ascii;
    arch        "LSB;label=32;scalar=64";
    class       {cls};
    location    "0.002/lagrangian/uniGas";
    object      {obj};
}}
// * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * //

"""


def test_written_cloud_reads_back_identically(tmp_path):
    spec, fill = workloads.couette(n_cells=12, per_cell=30, seed=4)
    o = OracleCase(spec, capacity=1 << 12)
    o.fill(**fill)
    o.step(3)
    p = o.particles()
    sm = o.sigma_max()
    foam_io.write_lagrangian(str(tmp_path), "0.001", spec.mesh, p, cell_weights=np.full(spec.mesh.n_cells, 1.5),
                             sigma_max=sm, cell_model=np.ones(spec.mesh.n_cells),
                             subcell_levels=np.tile([1, 2, 1], (spec.mesh.n_cells, 1)), delta_t=spec.delta_t)
    o.close()
    r = foam_io.read_lagrangian(str(tmp_path), "0.001", spec.mesh)
    assert np.array_equal(r["position"], p["position"])
    assert np.array_equal(r["cell"], p["cell"])
    assert np.array_equal(r["velocity"], p["velocity"])
    assert np.array_equal(r["erot"], p["erot"])
    assert np.array_equal(r["species"], p["species"])
    assert (r["cell_weight"] == 1.5).all()
    rf = foam_io.read_restart_fields(str(tmp_path), "0.001")
    assert np.array_equal(rf["sigma_max"], sm)
    assert (rf["cell_model"] == 1.0).all()
    assert np.array_equal(rf["subcell_levels"], np.tile([1.0, 2.0, 1.0], (spec.mesh.n_cells, 1)))
    assert rf["delta_t"] == spec.delta_t
    props = open(tmp_path / "0.001" / "uniform" / "lagrangian" / "uniGas" / "cloudProperties").read()
    assert "geometry        positions;" in props and f"particleCount   {len(p['cell'])};" in props


def write_hand(d, name, cls, body):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, name), "w") as f:
        f.write(HDR.format(cls=cls, obj=name) + body)


def common_fields(d, n):
    write_hand(d, "U", "vectorField", f"{n}\n(\n" + "".join(f"({i} {-i} 0.5)\n" for i in range(n)) + ")\n")
    write_hand(d, "ERot", "scalarField", f"{n}\n(\n" + "".join(f"{1e-21 * i}\n" for i in range(n)) + ")\n")
    write_hand(d, "typeId", "labelField", f"{n}\n(\n" + "".join(f"{i % 2}\n" for i in range(n)) + ")\n")
    write_hand(d, "ELevel", "labelField", f"{n}{{0}}\n")
    write_hand(d, "vibLevel", "labelFieldField", f"{n}\n(\n" + "0()\n" * n + ")\n")


def test_hand_written_legacy_positions(tmp_path):
    mesh = box_mesh((0.0, 0.0, 0.0), (2.0, 1.0, 1.0), (2, 1, 1), {k: abi.WALL for k in SIDES})
    d = str(tmp_path / "0.002" / "lagrangian" / "uniGas")
    write_hand(d, "positions", "Cloud<uniGasParcel>",
               "3\n(\n(0.25 0.5 0.5) 0\n(1.75 0.1 0.9) 1 // a comment\n(1e-1 2.5e-1 7.5e-1) 0\n)\n")
    common_fields(d, 3)
    r = foam_io.read_lagrangian(str(tmp_path), "0.002", mesh)
    assert np.array_equal(r["position"], [[0.25, 0.5, 0.5], [1.75, 0.1, 0.9], [0.1, 0.25, 0.75]])
    assert r["cell"].tolist() == [0, 1, 0]
    assert r["velocity"].tolist() == [[0, 0, 0.5], [1, -1, 0.5], [2, -2, 0.5]]
    assert r["species"].tolist() == [0, 1, 0]
    assert r["erot"][2] == 2e-21


def test_barycentric_coordinates(tmp_path):
    mesh = box_mesh((0.0, 0.0, 0.0), (2.0, 1.0, 1.0), (2, 1, 1), {k: abi.WALL for k in SIDES})
    # the internal face (0) between the cells: owner 0, neighbour 1
    f = 0
    assert mesh.owner[0] == 0 and mesh.neighbour[0] == 1
    ids = mesh.face_points[mesh.face_offsets[f]:mesh.face_offsets[f + 1]]
    P = mesh.points
    d = str(tmp_path / "0.002" / "lagrangian" / "uniGas")
    write_hand(d, "coordinates", "Cloud<uniGasParcel>",
               "4\n(\n(1 0 0 0) 0 0 1\n(0 1 0 0) 0 0 1\n(0.25 0.25 0.25 0.25) 0 0 2\n(0.25 0.25 0.25 0.25) 1 0 2\n)\n")
    common_fields(d, 4)
    r = foam_io.read_lagrangian(str(tmp_path), "0.002", mesh)
    assert np.allclose(r["position"][0], [0.5, 0.5, 0.5], atol=1e-15)       # the centroid of cell 0
    assert np.allclose(r["position"][1], P[ids[0]], atol=1e-15)            # the face's first point
    # the tet (C, p0, p2, p3) seen from the owner and (C', p0, p3, p2) from the neighbour
    t0 = 0.25 * (np.array([0.5, 0.5, 0.5]) + P[ids[0]] + P[ids[2]] + P[ids[3]])
    t1 = 0.25 * (np.array([1.5, 0.5, 0.5]) + P[ids[0]] + P[ids[3]] + P[ids[2]])
    assert np.allclose(r["position"][2], t0, atol=1e-15)
    assert np.allclose(r["position"][3], t1, atol=1e-15)
    assert r["cell"].tolist() == [0, 0, 0, 1]


def test_openfoam_centroids():
    mesh = box_mesh((0.0, 0.0, 0.0), (2.0, 1.0, 1.0), (2, 1, 1), {k: abi.WALL for k in SIDES})
    fc, cc = foam_io.foam_centroids(mesh)
    assert np.allclose(cc, [[0.5, 0.5, 0.5], [1.5, 0.5, 0.5]], atol=1e-15)

    # a cell whose y extent grows linearly with x (a trapezoidal prism):
    # height h(x) = 1 + x on 0 <= x <= 1, so the volume centroid is at
    # x = int x h / int h = (1/2 + 1/3)/(3/2) = 5/9
    def fn(S):
        x, t, u = S[..., 0], S[..., 1], S[..., 2]
        return np.stack([x, t * (1.0 + x), u], axis=-1)

    m2 = mapped_box_mesh((1, 1, 1), fn, {k: abi.WALL for k in SIDES})
    _, c2 = foam_io.foam_centroids(m2)
    assert c2[0][0] == pytest.approx(5.0 / 9.0, rel=1e-12)
    # y centroid: int h^2/2 / int h = (7/6)/(3/2) = 7/9
    assert c2[0][1] == pytest.approx(7.0 / 9.0, rel=1e-12)


def test_excited_states_are_refused(tmp_path):
    mesh = box_mesh((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (1, 1, 1), {k: abi.WALL for k in SIDES})
    d = str(tmp_path / "0.002" / "lagrangian" / "uniGas")
    write_hand(d, "positions", "Cloud<uniGasParcel>", "1\n(\n(0.5 0.5 0.5) 0\n)\n")
    common_fields(d, 1)
    write_hand(d, "vibLevel", "labelFieldField", "1\n(\n2(1 0)\n)\n")
    with pytest.raises(ValueError, match="vibrational"):
        foam_io.read_lagrangian(str(tmp_path), "0.002", mesh)
