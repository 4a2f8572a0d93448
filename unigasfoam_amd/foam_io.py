"""OpenFOAM case I/O in ASCII: constant/polyMesh (points, faces, owner,
neighbour, boundary) read and written, and the sampled fields written under
the names uniGasVolFields gives them (uniGasVolFields.C:66-345, written at
uniGasVolFields.C:1397-1428).

A mesh read here is the same PolyMesh the builders in mesh.py make, so a case
can start from a blockMesh (or any other) polyMesh of convex cells with up to
six planar faces.  Points are written with 17 significant digits, so a write
and a read give back the same doubles.  Binary polyMesh files are refused.
The cloud's restart layout (<time>/lagrangian/uniGas and the restart cell
fields) is read and written as well (write_lagrangian, read_lagrangian).
"""
from __future__ import annotations

import os
import re

import numpy as np

from . import abi
from .mesh import Patch, PolyMesh

# OpenFOAM patch type names <-> library patch types
_TYPE_NAMES = {abi.PATCH: "patch", abi.WALL: "wall", abi.SYMMETRY: "symmetryPlane", abi.EMPTY: "empty",
               abi.CYCLIC: "cyclic"}
_TYPES = {"patch": abi.PATCH, "wall": abi.WALL, "symmetryPlane": abi.SYMMETRY, "symmetry": abi.SYMMETRY,
          "empty": abi.EMPTY, "cyclic": abi.CYCLIC}

# library field -> (uniGasVolFields file name stem, OpenFOAM class, dimensions)
FIELD_FILES = {
    "rhoN": ("rhoN", "volScalarField", "[0 -3 0 0 0 0 0]"),
    "rhoM": ("rhoM", "volScalarField", "[1 -3 0 0 0 0 0]"),
    "p": ("p", "volScalarField", "[1 -1 -2 0 0 0 0]"),
    "translational_T": ("translationalT", "volScalarField", "[0 0 0 1 0 0 0]"),
    "rotational_T": ("rotationalT", "volScalarField", "[0 0 0 1 0 0 0]"),
    "overall_T": ("overallT", "volScalarField", "[0 0 0 1 0 0 0]"),
    "Ma": ("Ma", "volScalarField", "[0 0 0 0 0 0 0]"),
    "U_mean": ("UMean", "volVectorField", "[0 1 -1 0 0 0 0]"),
    "mfp": ("variableHardSphereMeanFreePath", "volScalarField", "[0 1 0 0 0 0 0]"),
    "mcr": ("meanCollisionRate", "volScalarField", "[0 0 -1 0 0 0 0]"),
    "density_error": ("densityError", "volScalarField", "[0 0 0 0 0 0 0]"),
    "velocity_error": ("velocityError", "volScalarField", "[0 0 0 0 0 0 0]"),
    "temperature_error": ("temperatureError", "volScalarField", "[0 0 0 0 0 0 0]"),
    "pressure_error": ("pressureError", "volScalarField", "[0 0 0 0 0 0 0]"),
    # cell heat flux and pressure tensor: uniGasVolFields keeps only their
    # wall values (surfaceHeatTransfer, surfaceShearStress), so these two
    # volume fields take names of their own
    "heat_flux": ("heatFlux", "volVectorField", "[1 0 -3 0 0 0 0]"),
    "shear_stress": ("pressureTensor", "volSymmTensorField", "[1 -1 -2 0 0 0 0]"),
}


def _header(cls, obj, location=None):
    loc = f'    location    "{location}";\n' if location else ""
    return ("FoamFile\n{\n    version     2.0;\n    format      ascii;\n"
            f"    class       {cls};\n{loc}    object      {obj};\n}}\n\n")


def _strip_comments(text):
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return re.sub(r"//[^\n]*", " ", text)


def _body(text, path):
    """The FoamFile header dictionary and what follows it."""
    text = _strip_comments(text)
    m = re.search(r"FoamFile\s*\{(.*?)\}", text, flags=re.S)
    if not m:
        raise ValueError(f"{path}: no FoamFile header")
    hdr = dict(re.findall(r"(\w+)\s+([^;]+);", m.group(1)))
    if hdr.get("format", "ascii").strip() != "ascii":
        raise ValueError(f"{path}: only ASCII polyMesh files are read")
    return text[m.end():]


def _list(body, path):
    """Count and the text between the outer parentheses of an OpenFOAM list."""
    m = re.match(r"\s*(\d+)\s*\(", body)
    if not m:
        raise ValueError(f"{path}: expected 'N ( ... )'")
    end = body.rfind(")")
    return int(m.group(1)), body[m.end():end]


def _read(path):
    with open(path) as f:
        return _body(f.read(), path)


def _numbers(s, dtype):
    return np.array(s.replace("(", " ").replace(")", " ").split(), dtype=dtype)


def write_polymesh(mesh: PolyMesh, case_dir):
    """Write constant/polyMesh/{points,faces,owner,neighbour,boundary}."""
    d = os.path.join(case_dir, "constant", "polyMesh")
    os.makedirs(d, exist_ok=True)
    loc = "constant/polyMesh"
    pts = np.asarray(mesh.points, dtype=np.float64)
    with open(os.path.join(d, "points"), "w") as f:
        f.write(_header("vectorField", "points", loc))
        f.write(f"{len(pts)}\n(\n")
        f.write("".join(f"({x:.17g} {y:.17g} {z:.17g})\n" for x, y, z in pts))
        f.write(")\n")
    off = np.asarray(mesh.face_offsets)
    fp = np.asarray(mesh.face_points)
    with open(os.path.join(d, "faces"), "w") as f:
        f.write(_header("faceList", "faces", loc))
        f.write(f"{len(off) - 1}\n(\n")
        f.write("".join(f"{off[i + 1] - off[i]}({' '.join(map(str, fp[off[i]:off[i + 1]]))})\n"
                        for i in range(len(off) - 1)))
        f.write(")\n")
    note = (f"nPoints:{len(pts)}  nCells:{mesh.n_cells}  nFaces:{mesh.n_faces}  "
            f"nInternalFaces:{mesh.n_internal_faces}")
    for name, arr in (("owner", mesh.owner), ("neighbour", mesh.neighbour)):
        with open(os.path.join(d, name), "w") as f:
            f.write(_header("labelList", name, loc).replace("    object", f'    note        "{note}";\n    object'))
            f.write(f"{len(arr)}\n(\n")
            f.write("".join(f"{v}\n" for v in np.asarray(arr)))
            f.write(")\n")
    with open(os.path.join(d, "boundary"), "w") as f:
        f.write(_header("polyBoundaryMesh", "boundary", loc))
        f.write(f"{len(mesh.patches)}\n(\n")
        for p in mesh.patches:
            f.write(f"    {p.name}\n    {{\n        type            {_TYPE_NAMES[p.type]};\n")
            if p.type == abi.CYCLIC:
                f.write(f"        neighbourPatch  {mesh.patches[p.neighbour].name};\n")
            f.write(f"        nFaces          {p.size};\n        startFace       {p.start};\n    }}\n")
        f.write(")\n")


def read_polymesh(case_dir) -> PolyMesh:
    """Read an ASCII constant/polyMesh into a PolyMesh."""
    d = os.path.join(case_dir, "constant", "polyMesh")
    n, s = _list(_read(os.path.join(d, "points")), "points")
    points = _numbers(s, np.float64).reshape(-1, 3)
    if len(points) != n:
        raise ValueError(f"points: {len(points)} entries, header says {n}")
    n, s = _list(_read(os.path.join(d, "faces")), "faces")
    toks = _numbers(s, np.int64)
    k0 = int(toks[0]) if len(toks) else 0
    if k0 > 0 and len(toks) == n * (k0 + 1) and np.all(toks[::k0 + 1] == k0):
        # every face has k0 points (hexahedral meshes): no walk needed
        face_points = toks.reshape(n, k0 + 1)[:, 1:].reshape(-1).astype(np.int32)
        offsets = list(range(0, n * k0 + 1, k0))
    else:
        offsets = [0]
        pts = []
        i = 0
        while i < len(toks):
            k = int(toks[i])
            pts.append(toks[i + 1:i + 1 + k])
            offsets.append(offsets[-1] + k)
            i += 1 + k
        face_points = np.concatenate(pts).astype(np.int32) if pts else np.zeros(0, np.int32)
    if len(offsets) - 1 != n:
        raise ValueError(f"faces: {len(offsets) - 1} entries, header says {n}")
    lists = {}
    for name in ("owner", "neighbour"):
        n, s = _list(_read(os.path.join(d, name)), name)
        lists[name] = _numbers(s, np.int64).astype(np.int32)
        if len(lists[name]) != n:
            raise ValueError(f"{name}: {len(lists[name])} entries, header says {n}")
    n, s = _list(_read(os.path.join(d, "boundary")), "boundary")
    entries = re.findall(r"(\w+)\s*\{([^{}]*)\}", s)
    if len(entries) != n:
        raise ValueError(f"boundary: {len(entries)} patches, header says {n}")
    patches, nbr_names = [], []
    for name, body in entries:
        kv = dict(re.findall(r"(\w+)\s+([^;]+);", body))
        t = kv["type"].strip()
        if t not in _TYPES:
            raise ValueError(f"boundary: patch {name} has unsupported type {t}")
        patches.append(Patch(name, _TYPES[t], int(kv["startFace"]), int(kv["nFaces"])))
        nbr_names.append(kv.get("neighbourPatch", "").strip())
    names = [p.name for p in patches]
    for p, nb in zip(patches, nbr_names):
        if p.type == abi.CYCLIC:
            if nb not in names:
                raise ValueError(f"boundary: cyclic patch {p.name} has no neighbourPatch")
            p.neighbour = names.index(nb)
    owner, neighbour = lists["owner"], lists["neighbour"]
    n_cells = int(max(owner.max(initial=-1), neighbour.max(initial=-1)) + 1)
    return PolyMesh(points=points, face_offsets=np.asarray(offsets, dtype=np.int32), face_points=face_points,
                    owner=owner, neighbour=neighbour, patches=patches, n_cells=n_cells)


def _values(a, cls):
    if cls == "volScalarField":
        return "".join(f"{v:.17g}\n" for v in a)
    return "".join("(" + " ".join(f"{v:.17g}" for v in row) + ")\n" for row in a)


# wall-only vol fields of uniGasVolFields (zero inside, measured values on
# wall faces): file stem -> (class, dimensions, key of Case.wall_fields())
WALL_FIELD_FILES = {
    "surfaceHeatTransfer": ("volScalarField", "[1 0 -3 0 0 0 0]", "heat_flux"),
    "surfaceShearStress": ("volScalarField", "[1 -1 -2 0 0 0 0]", "shear"),
    "fD": ("volVectorField", "[1 -1 -2 0 0 0 0]", "force_density"),
}
# cell fields whose wall-face values come from the wall measurements
# (uniGasVolFields.C:1262-1350); on other patches every field except these
# wall-only ones takes the owner cell's value (:1354-1392)
_WALL_MEASURED = {"translational_T": "translational_T", "rotational_T": "rotational_T", "overall_T": "overall_T",
                  "p": "pressure", "U_mean": "U_mean"}
_WIDTH = {"volScalarField": 1, "volVectorField": 3, "volSymmTensorField": 6}


def _patch_values(mesh, p, key, a, wall, wall_only):
    """Boundary values of field `key` (cell values `a`) on patch p; for a
    wall-only field `key` names the wall value."""
    owners = mesh.owner[p.start:p.start + p.size]
    if wall_only:
        vals = np.zeros((p.size,) + a.shape[1:])
    elif p.type == abi.WALL and key in _WALL_MEASURED:
        vals = np.zeros((p.size,) + a.shape[1:])
    else:
        return a[owners]
    if p.type == abi.WALL and wall is not None:
        src = wall["fields"][key if wall_only else _WALL_MEASURED[key]]
        for k in np.nonzero(wall["patch"] == mesh.patch_index(p.name))[0]:
            vals[wall["face"][k]] = src[k]
    return vals


def write_fields(case_dir, time_name, mesh: PolyMesh, fields, field_name="uniGasFields", wall=None):
    """Write the sampled fields (Case.fields()) to <case>/<time>/<stem>_<field_name>
    as OpenFOAM vol fields, with the boundary values uniGasVolFields writes
    (uniGasVolFields.C:1262-1392): on wall patches temperatures, pressure
    (fD.n) and UMean from the wall measurements, other fields the owner
    cell's value; and the wall-only fields surfaceHeatTransfer,
    surfaceShearStress and fD.  `wall` is {"patch", "face", "fields"} from
    Case.wall_faces() and Case.wall_fields(); without it wall-measured
    values are written as zero."""
    d = os.path.join(case_dir, str(time_name))
    os.makedirs(d, exist_ok=True)
    written = []
    # (key, file stem, class, dimensions, cell values, wall-only)
    files = [(key, stem, cls, dims, fields.get(key), False) for key, (stem, cls, dims) in FIELD_FILES.items()]
    for stem, (cls, dims, wkey) in WALL_FIELD_FILES.items():
        width = _WIDTH[cls]
        files.append((wkey, stem, cls, dims, np.zeros((mesh.n_cells, width) if width > 1 else mesh.n_cells), True))
    for key, stem, cls, dims, cell_values, wall_only in files:
        if cell_values is None:
            continue
        a = np.asarray(cell_values, dtype=np.float64)
        width = _WIDTH[cls]
        a = a.reshape(mesh.n_cells, width) if width > 1 else a.reshape(mesh.n_cells)
        obj = f"{stem}_{field_name}"
        kind = cls[3:].replace("Field", "").lower()     # scalar, vector, symmtensor
        kind = {"symmtensor": "symmTensor"}.get(kind, kind)
        with open(os.path.join(d, obj), "w") as f:
            f.write(_header(cls, obj, str(time_name)))
            f.write(f"dimensions      {dims};\n\n")
            f.write(f"internalField   nonuniform List<{kind}>\n{mesh.n_cells}\n(\n{_values(a, cls)})\n;\n\n")
            f.write("boundaryField\n{\n")
            for p in mesh.patches:
                if p.type == abi.EMPTY:
                    f.write(f"    {p.name}\n    {{\n        type            empty;\n    }}\n")
                elif p.type == abi.CYCLIC:
                    f.write(f"    {p.name}\n    {{\n        type            cyclic;\n    }}\n")
                else:
                    v = _patch_values(mesh, p, key, a, wall, wall_only)
                    f.write(f"    {p.name}\n    {{\n        type            calculated;\n"
                            f"        value           nonuniform List<{kind}>\n{p.size}\n(\n{_values(v, cls)})\n;\n    }}\n")
            f.write("}\n")
        written.append(obj)
    return written


def read_patch_field(path, patch):
    """Values of one patch's entry in an ASCII vol field file."""
    with open(path) as f:
        body = _body(f.read(), path)
    i = body.index("boundaryField")
    m = re.search(r"\b" + re.escape(patch) + r"\s*\{[^{}]*?value\s+nonuniform\s+List<(\w+)>\s*(\d+)\s*\(", body[i:])
    if not m:
        raise ValueError(f"{path}: no nonuniform value for patch {patch}")
    n = int(m.group(2))
    start = i + m.end()
    end = body.index(";", start)
    a = _numbers(body[start:body.rfind(")", start, end)], np.float64)
    return a.reshape(n, -1) if m.group(1) != "scalar" else a


def read_internal_field(path):
    """internalField of an ASCII vol field file (nonuniform List or uniform)."""
    with open(path) as f:
        body = _body(f.read(), path)
    m = re.search(r"internalField\s+nonuniform\s+List<(\w+)>\s*(\d+)\s*\(", body)
    if not m:
        m2 = re.search(r"internalField\s+uniform\s+([^;]+);", body)
        if not m2:
            raise ValueError(f"{path}: no internalField")
        return _numbers(m2.group(1), np.float64)
    n = int(m.group(2))
    end = body.index("boundaryField")
    s = body[m.end():body.rfind(")", 0, end)]
    a = _numbers(s, np.float64)
    return a.reshape(n, -1) if m.group(1) != "scalar" else a


# ------------------------------------------------------------- lagrangian
# The cloud's restart files (uniGasParcelIO.C:89-181, uniGasCloud.C:577-703):
# <time>/lagrangian/uniGas/ holds one IOField per parcel property (U,
# cellWeight, radialWeight, ERot, ELevel, typeId, newParcel, vibLevel) and
# the parcel positions, either as OpenFOAM's "coordinates" (barycentric in a
# tet of the cell: "(a b c d) celli tetFacei tetPti") or in the legacy
# "positions" form ("(x y z) celli"), which OpenFOAM reads when
# <time>/uniform/lagrangian/uniGas/cloudProperties says "geometry positions".
# The cell fields of a restart are uniGasSigmaTcRMax, uniGasCollisionModelId,
# uniGasCellWeightFactor and uniGasSubCellLevels, and <time>/uniform/time
# holds deltaT.
CLOUD = "uniGas"
LAGRANGIAN_FIELDS = {"U": "vectorField", "cellWeight": "scalarField", "radialWeight": "scalarField",
                     "ERot": "scalarField", "ELevel": "labelField", "typeId": "labelField",
                     "newParcel": "labelField"}


def _write_list_file(path, cls, obj, location, lines, n):
    with open(path, "w") as f:
        f.write(_header(cls, obj, location))
        f.write(f"{n}\n(\n")
        f.writelines(lines)
        f.write(")\n")


def _vol_file(path, cls, obj, location, dims, a, mesh: PolyMesh):
    kind = {"volScalarField": "scalar", "volVectorField": "vector"}[cls]
    with open(path, "w") as f:
        f.write(_header(cls, obj, location))
        f.write(f"dimensions      {dims};\n\n")
        f.write(f"internalField   nonuniform List<{kind}>\n{mesh.n_cells}\n(\n{_values(a, cls)})\n;\n\n")
        f.write("boundaryField\n{\n")
        for p in mesh.patches:
            t = {abi.EMPTY: "empty", abi.CYCLIC: "cyclic"}.get(p.type, "zeroGradient")
            f.write(f"    {p.name}\n    {{\n        type            {t};\n    }}\n")
        f.write("}\n")


def write_lagrangian(case_dir, time_name, mesh: PolyMesh, particles, cell_weights=None, sigma_max=None,
                     cell_model=None, subcell_levels=None, delta_t=None, cloud=CLOUD):
    """Write a cloud in uniGasFoam's restart layout.  `particles` holds
    position (global, [n, 3]), cell, velocity, erot and species (type ids);
    the per-cell arrays, when given, go to the restart cell fields and
    delta_t to uniform/time.  Positions are written in the legacy form with
    17 significant digits (cloudProperties: geometry positions)."""
    t = str(time_name)
    d = os.path.join(case_dir, t, "lagrangian", cloud)
    os.makedirs(d, exist_ok=True)
    loc = f"{t}/lagrangian/{cloud}"
    pos = np.asarray(particles["position"], dtype=np.float64).reshape(-1, 3)
    cell = np.asarray(particles["cell"], dtype=np.int64)
    n = len(cell)
    cw = np.ones(n) if cell_weights is None else np.asarray(cell_weights, dtype=np.float64)[cell]
    _write_list_file(os.path.join(d, "positions"), f"Cloud<{cloud}Parcel>", "positions", loc,
                     [f"({x:.17g} {y:.17g} {z:.17g}) {c}\n" for (x, y, z), c in zip(pos, cell)], n)
    cols = {"U": np.asarray(particles["velocity"], dtype=np.float64).reshape(-1, 3), "cellWeight": cw,
            "radialWeight": np.ones(n), "ERot": np.asarray(particles["erot"], dtype=np.float64),
            "ELevel": np.zeros(n, dtype=np.int64), "typeId": np.asarray(particles["species"], dtype=np.int64),
            "newParcel": np.zeros(n, dtype=np.int64)}
    for name, cls in LAGRANGIAN_FIELDS.items():
        a = cols[name]
        if cls == "vectorField":
            lines = ["(" + " ".join(f"{v:.17g}" for v in row) + ")\n" for row in a]
        elif cls == "scalarField":
            lines = [f"{v:.17g}\n" for v in a]
        else:
            lines = [f"{int(v)}\n" for v in a]
        _write_list_file(os.path.join(d, name), cls, name, loc, lines, n)
    _write_list_file(os.path.join(d, "vibLevel"), "labelFieldField", "vibLevel", loc, ["0()\n"] * n, n)
    u = os.path.join(case_dir, t, "uniform", "lagrangian", cloud)
    os.makedirs(u, exist_ok=True)
    with open(os.path.join(u, "cloudProperties"), "w") as f:
        f.write(_header("dictionary", "cloudProperties", f"{t}/uniform/lagrangian/{cloud}"))
        f.write(f"geometry        positions;\n\nprocessor0\n{{\n    particleCount   {n};\n}}\n")
    tl = os.path.join(case_dir, t)
    if sigma_max is not None:
        _vol_file(os.path.join(tl, f"{cloud}SigmaTcRMax"), "volScalarField", f"{cloud}SigmaTcRMax", t,
                  "[0 3 -1 0 0 0 0]", np.asarray(sigma_max, dtype=np.float64), mesh)
    if cell_model is not None:
        _vol_file(os.path.join(tl, f"{cloud}CollisionModelId"), "volScalarField", f"{cloud}CollisionModelId", t,
                  "[0 0 0 0 0 0 0]", np.asarray(cell_model, dtype=np.float64), mesh)
    if cell_weights is not None:
        _vol_file(os.path.join(tl, f"{cloud}CellWeightFactor"), "volScalarField", f"{cloud}CellWeightFactor", t,
                  "[0 0 0 0 0 0 0]", np.asarray(cell_weights, dtype=np.float64), mesh)
    if subcell_levels is not None:
        _vol_file(os.path.join(tl, f"{cloud}SubCellLevels"), "volVectorField", f"{cloud}SubCellLevels", t,
                  "[0 0 0 0 0 0 0]", np.asarray(subcell_levels, dtype=np.float64).reshape(-1, 3), mesh)
    if delta_t is not None:
        ud = os.path.join(case_dir, t, "uniform")
        with open(os.path.join(ud, "time"), "w") as f:
            f.write(_header("dictionary", "time", f"{t}/uniform"))
            f.write(f"value           {float(time_name):.17g};\nname            \"{t}\";\n"
                    f"index           0;\ndeltaT          {float(delta_t):.17g};\ndeltaT0         {float(delta_t):.17g};\n")
    return d


def foam_centroids(mesh: PolyMesh):
    """OpenFOAM's face centroids and cell centroids of a mesh
    (primitiveMeshFaceCentresAndAreas, primitiveMeshCellCentresAndVols)."""
    P = mesh.points
    nf = len(mesh.face_offsets) - 1
    fc = np.zeros((nf, 3))
    fa = np.zeros((nf, 3))
    for f in range(nf):
        ids = mesh.face_points[mesh.face_offsets[f]:mesh.face_offsets[f + 1]]
        pts = P[ids]
        if len(ids) == 3:
            fc[f] = (1.0 / 3.0) * pts.sum(axis=0)
            fa[f] = 0.5 * np.cross(pts[1] - pts[0], pts[2] - pts[0])
            continue
        mean = pts.mean(axis=0)
        nxt = np.roll(pts, -1, axis=0)
        c = pts + nxt + mean
        nn = np.cross(nxt - pts, mean - pts)
        a = np.linalg.norm(nn, axis=1)
        fc[f] = (1.0 / 3.0) * (a[:, None] * c).sum(axis=0) / a.sum()
        fa[f] = 0.5 * nn.sum(axis=0)
    nc = mesh.n_cells
    ni = len(mesh.neighbour)
    own, nei = np.asarray(mesh.owner), np.asarray(mesh.neighbour)
    cest = np.zeros((nc, 3))
    cnt = np.zeros(nc)
    np.add.at(cest, own, fc)
    np.add.at(cnt, own, 1.0)
    np.add.at(cest, nei, fc[:ni])
    np.add.at(cnt, nei, 1.0)
    cest /= cnt[:, None]
    cctr = np.zeros((nc, 3))
    vol = np.zeros(nc)
    pyr = np.maximum((fa * (fc - cest[own])).sum(axis=1), 1e-300)
    np.add.at(cctr, own, pyr[:, None] * (0.75 * fc + 0.25 * cest[own]))
    np.add.at(vol, own, pyr)
    pyr = np.maximum((fa[:ni] * (cest[nei] - fc[:ni])).sum(axis=1), 1e-300)
    np.add.at(cctr, nei, pyr[:, None] * (0.75 * fc[:ni] + 0.25 * cest[nei]))
    np.add.at(vol, nei, pyr)
    return fc, cctr / vol[:, None]


def _read_list_lines(path):
    n, inner = _list(_read(path), path)
    return n, inner


def read_lagrangian(case_dir, time_name, mesh: PolyMesh, cloud=CLOUD):
    """Read a cloud written by uniGasFoam (or write_lagrangian): positions
    from "positions" (legacy form) or from "coordinates" (barycentric in the
    tet (cell centroid, p0, p_i, p_i+1) of face tetFacei, the last two swapped
    in the face's neighbour cell), and the parcel fields.  Returns position,
    cell, velocity, erot, species and cell_weight arrays; a parcel with
    ELevel or vibLevel set is refused (no electronic or vibrational states)."""
    d = os.path.join(case_dir, str(time_name), "lagrangian", cloud)
    if os.path.exists(os.path.join(d, "positions")):
        n, inner = _read_list_lines(os.path.join(d, "positions"))
        rows = re.findall(r"\(\s*([^()]+?)\s*\)\s*(-?\d+)", inner)
        if len(rows) != n:
            raise ValueError(f"{d}/positions: {len(rows)} entries, header says {n}")
        pos = np.array([r[0].split() for r in rows], dtype=np.float64).reshape(n, 3)
        cell = np.array([r[1] for r in rows], dtype=np.int32)
    elif os.path.exists(os.path.join(d, "coordinates")):
        n, inner = _read_list_lines(os.path.join(d, "coordinates"))
        rows = re.findall(r"\(\s*([^()]+?)\s*\)\s*(-?\d+)\s+(-?\d+)\s+(-?\d+)", inner)
        if len(rows) != n:
            raise ValueError(f"{d}/coordinates: {len(rows)} entries, header says {n}")
        bary = np.array([r[0].split() for r in rows], dtype=np.float64).reshape(n, 4)
        cell = np.array([r[1] for r in rows], dtype=np.int32)
        tf = np.array([r[2] for r in rows], dtype=np.int64)
        tp = np.array([r[3] for r in rows], dtype=np.int64)
        _, cctr = foam_centroids(mesh)
        pos = np.zeros((n, 3))
        P = mesh.points
        for i in range(n):
            ids = mesh.face_points[mesh.face_offsets[tf[i]]:mesh.face_offsets[tf[i] + 1]]
            a, b = ids[tp[i] % len(ids)], ids[(tp[i] + 1) % len(ids)]
            if mesh.owner[tf[i]] != cell[i]:
                a, b = b, a
            pos[i] = bary[i, 0] * cctr[cell[i]] + bary[i, 1] * P[ids[0]] + bary[i, 2] * P[a] + bary[i, 3] * P[b]
    else:
        raise ValueError(f"{d}: neither positions nor coordinates")

    def field(name, dtype, width=1):
        path = os.path.join(d, name)
        body = _read(path)
        u = re.match(r"\s*(\d+)\s*\{([^}]*)\}", body)
        if u:
            # OpenFOAM's compact form of a uniform list: N{value}
            m = int(u.group(1))
            v = _numbers(u.group(2), dtype)
            a = np.tile(v, m)
        else:
            m, inner = _list(body, path)
            a = _numbers(inner, dtype)
        if len(a) != m * width:
            raise ValueError(f"{d}/{name}: {len(a)} values for {m} parcels")
        return a.reshape(m, width) if width > 1 else a

    out = {"position": pos, "cell": cell, "velocity": field("U", np.float64, 3), "erot": field("ERot", np.float64),
           "species": field("typeId", np.int32)}
    if os.path.exists(os.path.join(d, "cellWeight")):
        out["cell_weight"] = field("cellWeight", np.float64)
    if os.path.exists(os.path.join(d, "ELevel")) and field("ELevel", np.int64).any():
        raise ValueError(f"{d}: parcels in excited electronic levels")
    if os.path.exists(os.path.join(d, "vibLevel")):
        body = _read(os.path.join(d, "vibLevel"))
        if not re.match(r"\s*\d+\s*\{\s*0\s*\(\s*\)\s*\}", body):
            _, inner = _list(body, os.path.join(d, "vibLevel"))
            if re.search(r"\b[1-9]\d*\s*\(", inner):
                raise ValueError(f"{d}: parcels with vibrational levels")
    for k in ("velocity", "erot", "species"):
        if len(out[k]) != n:
            raise ValueError(f"{d}: {k} holds {len(out[k])} parcels, positions {n}")
    return out


def read_restart_fields(case_dir, time_name, cloud=CLOUD):
    """The restart cell fields and deltaT that exist under <time>."""
    t = os.path.join(case_dir, str(time_name))
    out = {}
    for key, stem in (("sigma_max", "SigmaTcRMax"), ("cell_model", "CollisionModelId"),
                      ("cell_weights", "CellWeightFactor"), ("subcell_levels", "SubCellLevels")):
        p = os.path.join(t, cloud + stem)
        if os.path.exists(p):
            out[key] = read_internal_field(p)
    p = os.path.join(t, "uniform", "time")
    if os.path.exists(p):
        m = re.search(r"\bdeltaT\s+([^;]+);", _read(p))
        if m:
            out["delta_t"] = float(m.group(1))
    return out


def write_cloud(case_dir, time_name, case, cloud=CLOUD):
    """A running Case's simulators and restart cell fields in uniGasFoam's
    restart layout (write_lagrangian)."""
    case.sync()
    p = case.particles()
    return write_lagrangian(case_dir, time_name, case.spec.mesh, p, cell_weights=case.cell_weights(),
                            sigma_max=case.sigma_max(), cell_model=case.cell_model(),
                            subcell_levels=case.subcell_levels(), delta_t=case.delta_t(), cloud=cloud)


def case_from_cloud(spec, case_dir, time_name, cloud=CLOUD, **case_kw):
    """A new Case of spec holding the cloud of <case_dir>/<time_name> (as
    uniGasCloud.C:577-703 starts from existing lagrangian data): the
    simulators at their positions and cells, then (σ_T c_r)max, the cell
    weights and subcell levels where the files hold them; a deltaT in
    uniform/time replaces the spec's.  The samplers start empty."""
    import dataclasses

    from .core import Case
    rf = read_restart_fields(case_dir, time_name, cloud)
    if "delta_t" in rf:
        spec = dataclasses.replace(spec, delta_t=rf["delta_t"])
    parts = read_lagrangian(case_dir, time_name, spec.mesh, cloud)
    c = Case(spec, **case_kw)
    if "cell_weights" in rf and spec.cell_weighted:
        c.set_cell_weights(np.asarray(rf["cell_weights"], dtype=np.float64).reshape(spec.mesh.n_cells))
    if "subcell_levels" in rf:
        c.set_subcell_levels(np.rint(np.asarray(rf["subcell_levels"]).reshape(-1, 3)).astype(np.int32))
    if len(parts["cell"]):
        c.upload(velocity=parts["velocity"], cell=parts["cell"], species=parts["species"], erot=parts["erot"],
                 position=parts["position"])
    if "sigma_max" in rf:
        c.set_sigma_max(np.asarray(rf["sigma_max"], dtype=np.float64).reshape(spec.mesh.n_cells))
    if "cell_model" in rf and spec.collision_model == "hybrid":
        st = c.state()
        st["cell_model"] = np.rint(np.asarray(rf["cell_model"])).astype(np.int32).reshape(spec.mesh.n_cells, 1)
        c.restore_state(st)
    return c
