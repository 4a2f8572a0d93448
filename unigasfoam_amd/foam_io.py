"""OpenFOAM case I/O in ASCII: constant/polyMesh (points, faces, owner,
neighbour, boundary) read and written, and the sampled fields written under
the names uniGasVolFields gives them (uniGasVolFields.C:66-345, written at
uniGasVolFields.C:1397-1428).

A mesh read here is the same PolyMesh the builders in mesh.py make, so a case
can start from a blockMesh (or any other) polyMesh of convex cells with up to
six planar faces.  Points are written with 17 significant digits, so a write
and a read give back the same doubles.  Binary polyMesh files are refused.
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
