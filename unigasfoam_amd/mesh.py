"""Structured hexahedral meshes in OpenFOAM polyMesh form.

A ``PolyMesh`` holds exactly what constant/polyMesh holds (points, faces,
owner, neighbour, boundary).  ``box_mesh`` builds the axis-aligned blockMesh
case used by the heat-bath and Couette workloads; ``graded_box_mesh`` and
``mapped_box_mesh`` give non-uniform and curvilinear (planar-faced) hexes.
Face ordering follows blockMesh: internal faces sorted by (owner, neighbour),
then boundary faces patch by patch; every face is right-handed about the
normal pointing out of its owner.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import abi


@dataclass
class Patch:
    name: str
    type: int
    start: int
    size: int
    neighbour: int = -1


@dataclass
class PolyMesh:
    points: np.ndarray            # (n_points, 3) float64
    face_offsets: np.ndarray      # (n_faces+1,) int32
    face_points: np.ndarray       # int32
    owner: np.ndarray             # (n_faces,) int32
    neighbour: np.ndarray         # (n_internal,) int32
    patches: list = field(default_factory=list)
    n_cells: int = 0

    @property
    def n_faces(self):
        return len(self.owner)

    @property
    def n_internal_faces(self):
        return len(self.neighbour)

    def patch_index(self, name):
        for i, p in enumerate(self.patches):
            if p.name == name:
                return i
        raise KeyError(name)

    def desc(self):
        """ugx_mesh_desc view; keeps the backing arrays alive on self."""
        self._pst = np.array([p.start for p in self.patches], dtype=np.int32)
        self._psz = np.array([p.size for p in self.patches], dtype=np.int32)
        self._pty = np.array([p.type for p in self.patches], dtype=np.int32)
        self._pnb = np.array([p.neighbour for p in self.patches], dtype=np.int32)
        self._pts = np.ascontiguousarray(self.points, dtype=np.float64)
        for a in ("face_offsets", "face_points", "owner", "neighbour"):
            setattr(self, a, np.ascontiguousarray(getattr(self, a), dtype=np.int32))
        d = abi.MeshDesc()
        d.n_points = len(self._pts)
        d.n_faces = self.n_faces
        d.n_internal_faces = self.n_internal_faces
        d.n_cells = self.n_cells
        d.points = abi.dptr(self._pts)
        d.face_offsets = abi.iptr(self.face_offsets)
        d.face_points = abi.iptr(self.face_points)
        d.owner = abi.iptr(self.owner)
        d.neighbour = abi.iptr(self.neighbour) if len(self.neighbour) else C.cast(None, abi.I32_P)
        d.n_patches = len(self.patches)
        d.patch_start = abi.iptr(self._pst)
        d.patch_size = abi.iptr(self._psz)
        d.patch_type = abi.iptr(self._pty)
        d.patch_neighbour = abi.iptr(self._pnb)
        return d


def _structured(X, nx, ny, nz, patch_types, names):
    """Hex mesh from a structured point grid X[(nx+1),(ny+1),(nz+1),3].

    patch_types: dict face-name ("xmin","xmax","ymin","ymax","zmin","zmax")
    -> ugx patch type; the two faces of an axis set to CYCLIC are paired.
    """
    def pid(i, j, k):
        return (k * (ny + 1) + j) * (nx + 1) + i

    def cid(i, j, k):
        return (k * ny + j) * nx + i

    pts = X.transpose(2, 1, 0, 3).reshape(-1, 3)  # k slowest, i fastest
    # cells in id order (k slowest); their +x, +y, +z faces have increasing
    # neighbour ids, so emitting them cell by cell gives (owner, neighbour) order.
    K, J, I = np.meshgrid(np.arange(nz, dtype=np.int64), np.arange(ny, dtype=np.int64),
                          np.arange(nx, dtype=np.int64), indexing="ij")
    I, J, K = I.ravel(), J.ravel(), K.ravel()
    nc = nx * ny * nz
    F3 = np.empty((nc, 3, 4), dtype=np.int64)
    N3 = np.empty((nc, 3), dtype=np.int64)
    F3[:, 0] = np.stack([pid(I + 1, J, K), pid(I + 1, J + 1, K), pid(I + 1, J + 1, K + 1), pid(I + 1, J, K + 1)], 1)
    F3[:, 1] = np.stack([pid(I, J + 1, K), pid(I, J + 1, K + 1), pid(I + 1, J + 1, K + 1), pid(I + 1, J + 1, K)], 1)
    F3[:, 2] = np.stack([pid(I, J, K + 1), pid(I + 1, J, K + 1), pid(I + 1, J + 1, K + 1), pid(I, J + 1, K + 1)], 1)
    c0 = cid(I, J, K)
    N3[:, 0] = c0 + 1
    N3[:, 1] = c0 + nx
    N3[:, 2] = c0 + nx * ny
    valid = np.stack([I + 1 < nx, J + 1 < ny, K + 1 < nz], 1)
    F = F3[valid]
    nei = N3[valid]
    own = np.broadcast_to(c0[:, None], (nc, 3))[valid]
    # boundary faces: index order (outer loop slowest axis)
    j2, k2 = np.meshgrid(np.arange(ny), np.arange(nz), indexing="xy")   # k slow, j fast
    j2, k2 = j2.ravel(), k2.ravel()
    i3, k3 = np.meshgrid(np.arange(nx), np.arange(nz), indexing="xy")
    i3, k3 = i3.ravel(), k3.ravel()
    i4, j4 = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    i4, j4 = i4.ravel(), j4.ravel()
    z2, zk = np.zeros_like(j2), np.zeros_like(i3)
    zz = np.zeros_like(i4)
    bf = {
        "xmin": (np.stack([pid(z2, j2, k2), pid(z2, j2, k2 + 1), pid(z2, j2 + 1, k2 + 1), pid(z2, j2 + 1, k2)], 1), cid(z2, j2, k2)),
        "xmax": (np.stack([pid(z2 + nx, j2, k2), pid(z2 + nx, j2 + 1, k2), pid(z2 + nx, j2 + 1, k2 + 1), pid(z2 + nx, j2, k2 + 1)], 1), cid(z2 + nx - 1, j2, k2)),
        "ymin": (np.stack([pid(i3, zk, k3), pid(i3 + 1, zk, k3), pid(i3 + 1, zk, k3 + 1), pid(i3, zk, k3 + 1)], 1), cid(i3, zk, k3)),
        "ymax": (np.stack([pid(i3, zk + ny, k3), pid(i3, zk + ny, k3 + 1), pid(i3 + 1, zk + ny, k3 + 1), pid(i3 + 1, zk + ny, k3)], 1), cid(i3, zk + ny - 1, k3)),
        "zmin": (np.stack([pid(i4, j4, zz), pid(i4, j4 + 1, zz), pid(i4 + 1, j4 + 1, zz), pid(i4 + 1, j4, zz)], 1), cid(i4, j4, zz)),
        "zmax": (np.stack([pid(i4, j4, zz + nz), pid(i4 + 1, j4, zz + nz), pid(i4 + 1, j4 + 1, zz + nz), pid(i4, j4 + 1, zz + nz)], 1), cid(i4, j4, zz + nz - 1)),
    }
    patches, index = [], {}
    Fs, Os = [F], [own]
    nf = len(F)
    for key in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax"):
        fv, c = bf[key]
        index[key] = len(patches)
        patches.append(Patch(names.get(key, key), patch_types.get(key, abi.WALL), nf, len(fv)))
        Fs.append(fv); Os.append(c)
        nf += len(fv)
    for a, b in (("xmin", "xmax"), ("ymin", "ymax"), ("zmin", "zmax")):
        if patches[index[a]].type == abi.CYCLIC:
            patches[index[a]].neighbour = index[b]
            patches[index[b]].neighbour = index[a]
    Fa = np.concatenate(Fs)
    off = np.arange(len(Fa) + 1, dtype=np.int32) * 4
    return PolyMesh(points=np.ascontiguousarray(pts, dtype=np.float64), face_offsets=off,
                    face_points=Fa.astype(np.int32).ravel(),
                    owner=np.concatenate(Os).astype(np.int32), neighbour=nei.astype(np.int32),
                    patches=patches, n_cells=nx * ny * nz)


def box_mesh(lo, hi, n, patch_types=None, names=None):
    """Axis-aligned uniform box lo..hi with n=(nx,ny,nz) cells (points at
    lo + (hi - lo) * i / n, as the blockMeshDict reader places them)."""
    return graded_box_mesh(lo, hi, n, (1.0, 1.0, 1.0), patch_types, names)


def graded_box_mesh(lo, hi, n, grading=(1.0, 1.0, 1.0), patch_types=None, names=None):
    """Box with blockMesh simpleGrading (last/first cell size ratio per axis)."""
    xs = []
    for a in range(3):
        m, g = n[a], grading[a]
        if m == 1 or g == 1.0:
            w = np.ones(m)
        else:
            r = g ** (1.0 / (m - 1))
            w = r ** np.arange(m)
        e = np.concatenate([[0.0], np.cumsum(w)]) / np.sum(w)
        xs.append(lo[a] + (hi[a] - lo[a]) * e)
    X = np.stack(np.meshgrid(*xs, indexing="ij"), axis=-1)
    return _structured(X, n[0], n[1], n[2], patch_types or {}, names or {})


def mapped_box_mesh(n, fn, patch_types=None, names=None):
    """Curvilinear hexes: fn maps unit-cube grid coordinates (s,t,u) to points.
    Faces are planar when fn is affine along each face (e.g. a cross-section
    scaled along x)."""
    s = [np.linspace(0.0, 1.0, n[a] + 1) for a in range(3)]
    S = np.stack(np.meshgrid(*s, indexing="ij"), axis=-1)
    X = fn(S)
    return _structured(X, n[0], n[1], n[2], patch_types or {}, names or {})


def channel_mesh(n_cells_y, gap, width, patch_y=abi.WALL, periodic="cyclic"):
    """1-D channel in y (Couette): one cell in x and z.  periodic="cyclic"
    pairs x and z patches; "empty" makes x and z reduced directions."""
    t = abi.CYCLIC if periodic == "cyclic" else abi.EMPTY
    pt = {"xmin": t, "xmax": t, "zmin": t, "zmax": t, "ymin": patch_y, "ymax": patch_y}
    names = {"ymin": "lowerWall", "ymax": "upperWall"}
    return box_mesh((0.0, 0.0, 0.0), (width, gap, width), (1, n_cells_y, 1), pt, names)
