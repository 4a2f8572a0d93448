"""A blockMesh subset: system/blockMeshDict to a PolyMesh.

Covered: ``scale`` / ``convertToMeters``, ``vertices``, ``blocks`` of
``hex (8 vertices) [zone] (nx ny nz) simpleGrading (gx gy gz)`` with scalar
expansion ratios, straight edges and circular ``arc`` edges (through a
point, ``arc v1 v2 (x y z)``, or by angle and axis, ``arc v1 v2 angle
(ax ay az)``), and ``boundary`` (or the older ``patches``) entries of type
patch, wall, symmetry, symmetryPlane, empty and cyclic (``neighbourPatch``).
Blocks are merged where they share faces, as blockMesh does; block faces
that no patch names go to ``defaultFaces`` (type empty, blockMesh's
default).  Anything else (other curved edges, multi-grading, edgeGrading,
mergePatchPairs, projection) raises Unsupported with the entry's name.

Points inside a block: the graded parameters of each direction (cell sizes
in geometric progression, last/first = the expansion ratio) mapped through
the block's vertices: exactly v0 + s (v1 - v0) + t (v3 - v0) + u (v4 - v0)
for a parallelepiped block, the trilinear map otherwise.  A block with arc
edges takes the edge-based transfinite interpolation: the sum over its
twelve edges, each evaluated at its graded parameter (an arc at that
fraction of its angle) and weighted bilinearly by the other two
parameters, less twice the trilinear map, which reproduces every edge and
is the trilinear map when the edges are straight.  Cells are numbered
block by block, x fastest; faces follow OpenFOAM's order (internal faces by
owner then neighbour, then the boundary patch by patch), each right-handed
about the normal out of its owner.
"""
from __future__ import annotations

import numpy as np

from . import abi
from .foam_dict import DictError, FoamDict, Tokens
from .mesh import Patch, PolyMesh

_PATCH_TYPES = {"patch": abi.PATCH, "wall": abi.WALL, "symmetry": abi.SYMMETRY, "symmetryPlane": abi.SYMMETRY,
                "empty": abi.EMPTY, "cyclic": abi.CYCLIC}

# faces of a hex in local corner numbers, outward (blockMesh's hex model)
_HEX_FACES = ((0, 4, 7, 3), (1, 2, 6, 5), (0, 1, 5, 4), (3, 7, 6, 2), (0, 3, 2, 1), (4, 5, 6, 7))


class Unsupported(DictError):
    """A blockMeshDict feature this subset does not build."""


def _graded(n, g):
    if n == 1 or g == 1.0:
        w = np.ones(n)
    else:
        w = (g ** (1.0 / (n - 1))) ** np.arange(n)
    return np.concatenate([[0.0], np.cumsum(w)]) / np.sum(w)


def _parse_blocks(blocks):
    out = []
    i = 0
    while i < len(blocks):
        if blocks[i] != "hex":
            raise Unsupported(f"block shape {blocks[i]!r} (only hex)")
        verts = blocks[i + 1]
        i += 2
        if isinstance(blocks[i], str) and blocks[i] not in ("simpleGrading", "edgeGrading"):
            i += 1                    # zone name
        div = blocks[i]
        i += 1
        kind = blocks[i]
        grad = blocks[i + 1]
        i += 2
        if kind != "simpleGrading":
            raise Unsupported(f"{kind} (only simpleGrading)")
        if not all(isinstance(g, (int, float)) for g in grad):
            raise Unsupported("multi-grading in simpleGrading")
        if len(verts) != 8 or len(div) != 3:
            raise DictError("hex block needs 8 vertices and 3 divisions")
        out.append(([int(v) for v in verts], [int(n) for n in div], [float(g) for g in grad]))
    return out


def _block_points(V, div, grad):
    s, t, u = (_graded(div[a], grad[a]) for a in range(3))
    S, T, Uu = np.meshgrid(s, t, u, indexing="ij")
    v = V
    para = (np.array_equal(v[2], v[1] + v[3] - v[0]) and np.array_equal(v[5], v[1] + v[4] - v[0])
            and np.array_equal(v[7], v[3] + v[4] - v[0]) and np.array_equal(v[6], v[2] + v[4] - v[0]))
    if para:
        X = v[0] + S[..., None] * (v[1] - v[0]) + T[..., None] * (v[3] - v[0]) + Uu[..., None] * (v[4] - v[0])
    else:
        S, T, Uu = S[..., None], T[..., None], Uu[..., None]
        X = ((1 - S) * (1 - T) * (1 - Uu) * v[0] + S * (1 - T) * (1 - Uu) * v[1] + S * T * (1 - Uu) * v[2]
             + (1 - S) * T * (1 - Uu) * v[3] + (1 - S) * (1 - T) * Uu * v[4] + S * (1 - T) * Uu * v[5]
             + S * T * Uu * v[6] + (1 - S) * T * Uu * v[7])
    return X  # [nx+1, ny+1, nz+1, 3]


# the 12 edges of a hex block: (corner a, corner b, direction, the two
# other-direction parameters' sides), in blockMesh's local corner numbers
_HEX_EDGES = (
    (0, 1, 0, (0, 0)), (3, 2, 0, (1, 0)), (4, 5, 0, (0, 1)), (7, 6, 0, (1, 1)),
    (0, 3, 1, (0, 0)), (1, 2, 1, (1, 0)), (4, 7, 1, (0, 1)), (5, 6, 1, (1, 1)),
    (0, 4, 2, (0, 0)), (1, 5, 2, (1, 0)), (3, 7, 2, (0, 1)), (2, 6, 2, (1, 1)),
)


class _Arc:
    """A circular arc from p0 to p1 (arcEdge): by a third point on it or by
    the swept angle (degrees) about an axis; position(f) is the point at
    fraction f of the angle."""

    def __init__(self, p0, p1, mid=None, angle=None, axis=None):
        p0, p1 = np.asarray(p0, float), np.asarray(p1, float)
        if mid is not None:
            pm = np.asarray(mid, float)
            a, b = pm - p0, p1 - p0
            n = np.cross(a, b)
            nn = np.dot(n, n)
            if nn <= 0.0:
                raise DictError("arc through three collinear points")
            # circumcentre of p0, pm, p1
            c = p0 + (np.dot(a, a) * np.cross(b, n) + np.dot(b, b) * np.cross(n, a)) / (2.0 * nn)
            n = n / np.sqrt(nn)
            r0, r1, rm = p0 - c, p1 - c, pm - c
            ang = np.arctan2(np.dot(np.cross(r0, r1), n), np.dot(r0, r1))
            if ang <= 0.0:
                ang += 2.0 * np.pi
            self.c, self.n, self.r0, self.ang = c, n, r0, ang
        else:
            ax = np.asarray(axis, float)
            ax = ax / np.linalg.norm(ax)
            ang = np.radians(float(angle))
            # the centre on the bisector plane of the chord at the distance the angle fixes
            chord = p1 - p0
            half = 0.5 * np.linalg.norm(chord)
            mid = 0.5 * (p0 + p1)
            toward = np.cross(ax, chord)
            toward = toward / np.linalg.norm(toward)
            c = mid + toward * (half / np.tan(0.5 * ang))
            self.c, self.n, self.r0, self.ang = c, ax, p0 - c, ang

    def position(self, f):
        f = np.asarray(f, float)[..., None]
        th = f * self.ang
        r0, n = self.r0, self.n
        return self.c + r0 * np.cos(th) + np.cross(n, r0) * np.sin(th) + n * np.dot(n, r0) * (1.0 - np.cos(th))


def _parse_edges(edges, verts, scale):
    """{(v1, v2): curve} of the arc edges (verts already scaled)."""
    out = {}
    i = 0
    while i < len(edges):
        kind = edges[i]
        if kind != "arc":
            raise Unsupported(f"curved edges ({kind!r} ...)")
        v1, v2 = int(edges[i + 1]), int(edges[i + 2])
        if isinstance(edges[i + 3], list):
            out[(v1, v2)] = _Arc(verts[v1], verts[v2], mid=np.asarray(edges[i + 3], float) * scale)
            i += 4
        else:
            out[(v1, v2)] = _Arc(verts[v1], verts[v2], angle=edges[i + 3], axis=edges[i + 4])
            i += 5
    return out


def _curved_block_points(V, vid, div, grad, curves):
    """Edge-based transfinite interpolation of a hex block with curved edges."""
    s, t, u = (_graded(div[a], grad[a]) for a in range(3))
    par = (s, t, u)
    S, T, Uu = np.meshgrid(s, t, u, indexing="ij")
    S, T, Uu = S[..., None], T[..., None], Uu[..., None]
    P = (S, T, Uu)
    X = -2.0 * ((1 - S) * (1 - T) * (1 - Uu) * V[0] + S * (1 - T) * (1 - Uu) * V[1] + S * T * (1 - Uu) * V[2]
                + (1 - S) * T * (1 - Uu) * V[3] + (1 - S) * (1 - T) * Uu * V[4] + S * (1 - T) * Uu * V[5]
                + S * T * Uu * V[6] + (1 - S) * T * Uu * V[7])
    for a, b, dirn, (sa, sb) in _HEX_EDGES:
        p = par[dirn]
        ga, gb = vid[a], vid[b]
        if (ga, gb) in curves:
            e = curves[(ga, gb)].position(p)
        elif (gb, ga) in curves:
            e = curves[(gb, ga)].position(1.0 - p)
        else:
            e = V[a][None, :] + p[:, None] * (V[b] - V[a])[None, :]
        o1, o2 = [P[k] for k in range(3) if k != dirn]
        w = (o1 if sa else 1 - o1) * (o2 if sb else 1 - o2)
        shape = [1, 1, 1, 3]
        shape[dirn] = len(p)
        X = X + w * e.reshape(shape)
    return X


def build(d: FoamDict) -> PolyMesh:
    scale = float(d.get("scale", d.get("convertToMeters", 1.0)))
    edges = d.get("edges", [])
    verts = np.array(d["vertices"], dtype=np.float64) * scale
    curves = _parse_edges(edges, verts, scale) if edges else {}
    if d.get("mergePatchPairs", []):
        raise Unsupported("mergePatchPairs")
    for key in ("geometry", "faces", "projectVertices"):
        if key in d:
            raise Unsupported(key)
    blocks = _parse_blocks(d["blocks"])
    # points of every block; coincident points of neighbouring blocks merged
    pts, cells, bface = [], [], []
    base = 0
    for b, (vid, div, grad) in enumerate(blocks):
        if any((vid[a], vid[c]) in curves or (vid[c], vid[a]) in curves for a, c, _, _ in _HEX_EDGES):
            X = _curved_block_points(verts[vid], vid, div, grad, curves)
        else:
            X = _block_points(verts[vid], div, grad)
        nx, ny, nz = div
        P = X.transpose(2, 1, 0, 3).reshape(-1, 3)    # k slowest, i fastest
        pts.append(P)
        k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        i, j, k = i.ravel(), j.ravel(), k.ravel()

        def pid(a, bb, c):
            return base + (c * (ny + 1) + bb) * (nx + 1) + a

        corners = np.stack([pid(i, j, k), pid(i + 1, j, k), pid(i + 1, j + 1, k), pid(i, j + 1, k),
                            pid(i, j, k + 1), pid(i + 1, j, k + 1), pid(i + 1, j + 1, k + 1), pid(i, j + 1, k + 1)], 1)
        cells.append(corners)
        # block side of each cell face that lies on the block's boundary
        side = np.full((len(i), 6), -1, dtype=np.int64)
        side[i == 0, 0] = 0
        side[i == nx - 1, 1] = 1
        side[j == 0, 2] = 2
        side[j == ny - 1, 3] = 3
        side[k == 0, 4] = 4
        side[k == nz - 1, 5] = 5
        bface.append(np.where(side >= 0, side + 6 * b, -1))
        base += len(P)
    P = np.concatenate(pts)
    corners = np.concatenate(cells)
    bside = np.concatenate(bface)
    # merge coincident points (blockMesh merges the shared faces of blocks)
    span = float(np.max(P.max(0) - P.min(0))) if len(P) else 1.0
    tol = 1e-7 * max(span, 1e-300)
    from scipy.spatial import cKDTree
    parent = np.arange(len(P))
    pairs = cKDTree(P).query_pairs(tol, output_type="ndarray")
    if len(pairs):
        for a, b in sorted(map(tuple, np.sort(pairs, axis=1))):
            ra, rb = a, b
            while parent[ra] != ra:
                ra = parent[ra]
            while parent[rb] != rb:
                rb = parent[rb]
            if ra != rb:
                parent[max(ra, rb)] = min(ra, rb)
        for q in range(len(parent)):
            r = q
            while parent[r] != r:
                r = parent[r]
            parent[q] = r
    keep = parent == np.arange(len(P))
    newid = np.cumsum(keep) - 1
    corners = newid[parent[corners]]
    P = P[keep]
    nc = len(corners)
    # every cell face, keyed by its sorted point ids
    F = corners[:, np.array(_HEX_FACES)]           # [nc, 6, 4]
    key = np.sort(F, axis=2).reshape(nc * 6, 4)
    order = np.lexsort(key.T[::-1])
    ks = key[order]
    same = np.all(ks[1:] == ks[:-1], axis=1)
    cellof = np.repeat(np.arange(nc), 6)
    faceof = np.tile(np.arange(6), nc)
    first = np.ones(len(ks), dtype=bool)
    first[1:] = ~same
    internal_a = order[:-1][same]
    internal_b = order[1:][same]
    # internal faces: owner = lower cell, points as the owner sees them
    ca, cb = cellof[internal_a], cellof[internal_b]
    own = np.minimum(ca, cb)
    nei = np.maximum(ca, cb)
    src = np.where(ca < cb, internal_a, internal_b)
    ipts = F.reshape(nc * 6, 4)[src]
    io = np.lexsort((nei, own))
    own, nei, ipts = own[io], nei[io], ipts[io]
    # boundary faces
    paired = np.zeros(nc * 6, dtype=bool)
    paired[internal_a] = True
    paired[internal_b] = True
    bidx = np.nonzero(~paired)[0]
    bcell, bfc = cellof[bidx], faceof[bidx]
    bs = bside.reshape(nc * 6)[bidx]
    bpts = F.reshape(nc * 6, 4)[bidx]
    # patches: block sides by vertex set
    bnd = d.get("boundary", d.get("patches"))
    if bnd is None:
        raise DictError("blockMeshDict has no boundary")
    side_patch = {}
    plist = []
    if "boundary" in d:
        for name, pd in bnd:
            ptype = pd.word("type")
            if ptype not in _PATCH_TYPES:
                raise Unsupported(f"patch type {ptype}")
            plist.append((name, _PATCH_TYPES[ptype], pd.get("neighbourPatch")))
            for fv in pd.get("faces", []):
                side_patch[frozenset(int(v) for v in fv)] = len(plist) - 1
    else:
        for q in range(0, len(bnd), 3):
            ptype, name, faces = bnd[q], bnd[q + 1], bnd[q + 2]
            if ptype not in _PATCH_TYPES:
                raise Unsupported(f"patch type {ptype}")
            plist.append((name, _PATCH_TYPES[ptype], None))
            for fv in faces:
                side_patch[frozenset(int(v) for v in fv)] = len(plist) - 1
    default = None
    bpatch = np.empty(len(bidx), dtype=np.int64)
    for q in range(len(bidx)):
        b, sd = divmod(int(bs[q]), 6)
        vid = blocks[b][0]
        fs = frozenset(vid[c] for c in _HEX_FACES[sd])
        pi = side_patch.get(fs)
        if pi is None:
            if default is None:
                plist.append(("defaultFaces", abi.EMPTY, None))
                default = len(plist) - 1
            pi = default
        bpatch[q] = pi
    patches = []
    faces = [ipts]
    owners = [own]
    start = len(ipts)
    for pi, (name, ptype, nbname) in enumerate(plist):
        sel = np.nonzero(bpatch == pi)[0]
        sel = sel[np.lexsort((bfc[sel], bcell[sel]))]
        patches.append(Patch(name, ptype, start, len(sel)))
        faces.append(bpts[sel])
        owners.append(bcell[sel])
        start += len(sel)
    names = [p.name for p in patches]
    for pi, (name, ptype, nbname) in enumerate(plist):
        if ptype == abi.CYCLIC:
            if nbname is None or nbname not in names:
                raise DictError(f"cyclic patch {name} needs neighbourPatch")
            patches[pi].neighbour = names.index(nbname)
    Fa = np.concatenate(faces).astype(np.int32)
    off = (np.arange(len(Fa) + 1) * 4).astype(np.int32)
    return PolyMesh(points=np.ascontiguousarray(P), face_offsets=off, face_points=Fa.ravel(),
                    owner=np.concatenate(owners).astype(np.int32), neighbour=nei.astype(np.int32),
                    patches=patches, n_cells=int(nc))
