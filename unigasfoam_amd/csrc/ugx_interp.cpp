// ugx_interp.cpp - geometry of macroInterpolation: the cellPoint
// interpolation of the BGK-family fields to a simulator's position
// (interpolationCellPoint, used at stochasticParticleBGK.C:645-688,
// stochasticParticleESBGK.C:741-790, unifiedStochasticParticleSBGK.C:893-946).
//
// OpenFOAM's cellPoint interpolation is linear on a decomposition of each
// cell into tetrahedra (the cell centroid and, per face, the triangles
// (p0, p_i, p_i+1) of a fan from the face's first point), with the cell's
// own value at the centroid and point values at the mesh points.  The point
// values follow volPointInterpolation: a point not on a wall, patch or
// symmetry face takes the inverse-distance weighted mean of the cells around
// it (1/|P - C_c|, cell centroids); a point on such a face the inverse-
// distance weighted mean of those faces' values (1/|P - C_f|, face
// centroids; a zeroGradient field's face value is its owner cell's); points
// of empty faces are not boundary points; points joined by cyclic patches
// share one value, summed over every image with one normalisation.
// Centroids are OpenFOAM's: the face centroid and area of the fan of
// triangles about the vertex mean (primitiveMeshFaceCentresAndAreas) and
// the cell centroid of the face pyramids about the mean of the face
// centroids (primitiveMeshCellCentresAndVols).
#include <algorithm>
#include <cmath>
#include <numeric>

#include "ugx_world.hpp"

namespace ugx {

static void cross(const double a[3], const double b[3], double r[3])
{
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

static double dist(const double* a, const double* b)
{
    const double d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
    return std::sqrt(d0 * d0 + d1 * d1 + d2 * d2);
}

static int find_root(std::vector<int>& up, int a)
{
    while (up[a] != a) {
        up[a] = up[up[a]];
        a = up[a];
    }
    return a;
}

int build_interp(const ugx_mesh_desc* m, MeshGeom& g)
{
    const int nc = m->n_cells, nf = m->n_faces, np = m->n_points;
    const double* P = m->points;
    // face centroids and areas
    std::vector<double> fc((size_t)nf * 3), fa((size_t)nf * 3);
    for (int f = 0; f < nf; ++f) {
        const int b = m->face_offsets[f], n = m->face_offsets[f + 1] - b;
        const int32_t* fp = &m->face_points[b];
        double* C = &fc[3 * (size_t)f];
        double* A = &fa[3 * (size_t)f];
        if (n == 3) {
            for (int d = 0; d < 3; ++d) C[d] = (1.0 / 3.0) * (P[3 * fp[0] + d] + P[3 * fp[1] + d] + P[3 * fp[2] + d]);
            double u[3], v[3];
            for (int d = 0; d < 3; ++d) { u[d] = P[3 * fp[1] + d] - P[3 * fp[0] + d]; v[d] = P[3 * fp[2] + d] - P[3 * fp[0] + d]; }
            cross(u, v, A);
            for (int d = 0; d < 3; ++d) A[d] = 0.5 * A[d];
            continue;
        }
        double mean[3] = {0.0, 0.0, 0.0};
        for (int i = 0; i < n; ++i)
            for (int d = 0; d < 3; ++d) mean[d] = mean[d] + P[3 * fp[i] + d];
        for (int d = 0; d < 3; ++d) mean[d] = mean[d] / (double)n;
        double sumN[3] = {0.0, 0.0, 0.0}, sumA = 0.0, sumAc[3] = {0.0, 0.0, 0.0};
        for (int i = 0; i < n; ++i) {
            const double* p0 = &P[3 * fp[i]];
            const double* p1 = &P[3 * fp[(i + 1) % n]];
            double c[3], u[3], v[3], nn[3];
            for (int d = 0; d < 3; ++d) {
                c[d] = p0[d] + p1[d] + mean[d];
                u[d] = p1[d] - p0[d];
                v[d] = mean[d] - p0[d];
            }
            cross(u, v, nn);
            const double a = std::sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
            for (int d = 0; d < 3; ++d) { sumN[d] = sumN[d] + nn[d]; sumAc[d] = sumAc[d] + a * c[d]; }
            sumA = sumA + a;
        }
        if (sumA < 1.0e-150) {
            for (int d = 0; d < 3; ++d) { C[d] = mean[d]; A[d] = 0.0; }
        } else {
            for (int d = 0; d < 3; ++d) { C[d] = (1.0 / 3.0) * sumAc[d] / sumA; A[d] = 0.5 * sumN[d]; }
        }
    }
    // cell centroids: owner faces first, then neighbour faces, as OpenFOAM sums them
    std::vector<double> cest((size_t)nc * 3, 0.0), cctr((size_t)nc * 3, 0.0), cvol((size_t)nc, 0.0);
    std::vector<int> nfc((size_t)nc, 0);
    auto cell_of = [&](int f, int side) { return side == 0 ? m->owner[f] : m->neighbour[f]; };
    for (int side = 0; side < 2; ++side)
        for (int f = 0; f < (side == 0 ? nf : m->n_internal_faces); ++f) {
            const int c = cell_of(f, side);
            for (int d = 0; d < 3; ++d) cest[3 * (size_t)c + d] = cest[3 * (size_t)c + d] + fc[3 * (size_t)f + d];
            nfc[c] += 1;
        }
    for (int c = 0; c < nc; ++c)
        for (int d = 0; d < 3; ++d) cest[3 * (size_t)c + d] = cest[3 * (size_t)c + d] / (double)nfc[c];
    for (int side = 0; side < 2; ++side)
        for (int f = 0; f < (side == 0 ? nf : m->n_internal_faces); ++f) {
            const int c = cell_of(f, side);
            const double* C = &fc[3 * (size_t)f];
            const double* A = &fa[3 * (size_t)f];
            const double* E = &cest[3 * (size_t)c];
            double pyr = side == 0 ? A[0] * (C[0] - E[0]) + A[1] * (C[1] - E[1]) + A[2] * (C[2] - E[2])
                                   : A[0] * (E[0] - C[0]) + A[1] * (E[1] - C[1]) + A[2] * (E[2] - C[2]);
            pyr = std::max(pyr, 1.0e-300);
            for (int d = 0; d < 3; ++d)
                cctr[3 * (size_t)c + d] = cctr[3 * (size_t)c + d] + pyr * (0.75 * C[d] + 0.25 * E[d]);
            cvol[c] = cvol[c] + pyr;
        }
    g.mi_cent.assign((size_t)nc * 3, 0.0);
    for (int c = 0; c < nc; ++c)
        for (int d = 0; d < 3; ++d) {
            const double x = std::fabs(cvol[c]) > 1.0e-300 ? cctr[3 * (size_t)c + d] / cvol[c] : cest[3 * (size_t)c + d];
            cctr[3 * (size_t)c + d] = x;
            g.mi_cent[3 * (size_t)c + d] = x - g.cc[3 * (size_t)c + d];
        }
    // the faces of each cell in ascending face order, their points
    g.mi_fp.assign((size_t)nc * kMaxFaces * kMiFaceVerts, -1);
    std::vector<int> nface((size_t)nc, 0);
    std::vector<std::vector<int>> pcells((size_t)np), pfaces((size_t)np);
    for (int f = 0; f < nf; ++f) {
        const int b = m->face_offsets[f], n = m->face_offsets[f + 1] - b;
        if (n > kMiFaceVerts) return set_err(UGX_ERR_MESH, "macroInterpolation: face %d has more than %d points", f, kMiFaceVerts);
        for (int side = 0; side < (f < m->n_internal_faces ? 2 : 1); ++side) {
            const int c = cell_of(f, side);
            const int j = nface[c]++;
            if (j >= kMaxFaces) return set_err(UGX_ERR_MESH, "cell %d has more than %d faces", c, kMaxFaces);
            for (int i = 0; i < n; ++i)
                g.mi_fp[((size_t)c * kMaxFaces + j) * kMiFaceVerts + i] = m->face_points[b + i];
        }
        for (int i = 0; i < n; ++i) {
            const int p = m->face_points[b + i];
            pfaces[p].push_back(f);
            for (int side = 0; side < (f < m->n_internal_faces ? 2 : 1); ++side) pcells[p].push_back(cell_of(f, side));
        }
    }
    for (auto& v : pcells) {
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
    }
    // boundary faces that carry point values, and the points on them
    std::vector<int8_t> vface((size_t)nf, 0), ppoint((size_t)np, 0);
    std::vector<int> up((size_t)np);
    std::iota(up.begin(), up.end(), 0);
    for (int pa = 0; pa < m->n_patches; ++pa) {
        const int t = m->patch_type[pa];
        if (t == UGX_PATCH_EMPTY) continue;
        if (t == UGX_PATCH_CYCLIC) {
            // join each face's points with the matching points of its partner face
            const int pb = m->patch_neighbour ? m->patch_neighbour[pa] : -1;
            if (pb < pa) continue;
            for (int i = 0; i < m->patch_size[pa]; ++i) {
                const int f0 = m->patch_start[pa] + i, f1 = m->patch_start[pb] + i;
                double T[3];
                for (int d = 0; d < 3; ++d) T[d] = fc[3 * (size_t)f1 + d] - fc[3 * (size_t)f0 + d];
                const double sz = std::sqrt(fa[3 * (size_t)f0] * fa[3 * (size_t)f0] + fa[3 * (size_t)f0 + 1] * fa[3 * (size_t)f0 + 1] +
                                            fa[3 * (size_t)f0 + 2] * fa[3 * (size_t)f0 + 2]);
                for (int a = m->face_offsets[f0]; a < m->face_offsets[f0 + 1]; ++a) {
                    const int p0 = m->face_points[a];
                    int best = -1;
                    double bd = 1.0e300;
                    for (int q = m->face_offsets[f1]; q < m->face_offsets[f1 + 1]; ++q) {
                        const int p1 = m->face_points[q];
                        double x[3];
                        for (int d = 0; d < 3; ++d) x[d] = P[3 * p0 + d] + T[d];
                        const double dd = dist(x, &P[3 * p1]);
                        if (dd < bd) { bd = dd; best = p1; }
                    }
                    if (best < 0 || bd > 1.0e-6 * std::sqrt(sz))
                        return set_err(UGX_ERR_MESH, "cyclic faces %d and %d do not match point for point", f0, f1);
                    const int r0 = find_root(up, p0), r1 = find_root(up, best);
                    if (r0 != r1) up[std::max(r0, r1)] = std::min(r0, r1);
                }
            }
            continue;
        }
        for (int f = m->patch_start[pa]; f < m->patch_start[pa] + m->patch_size[pa]; ++f) {
            vface[f] = 1;
            for (int a = m->face_offsets[f]; a < m->face_offsets[f + 1]; ++a) ppoint[m->face_points[a]] = 1;
        }
    }
    // the weights of each class of joined points, shared by its members
    std::vector<std::vector<int>> members((size_t)np);
    for (int p = 0; p < np; ++p) members[find_root(up, p)].push_back(p);
    g.mi_pts.assign(P, P + 3 * (size_t)np);
    g.mi_poff.assign((size_t)np + 1, 0u);
    std::vector<std::vector<std::pair<int, double>>> ent((size_t)np);
    for (int r = 0; r < np; ++r) {
        const auto& mem = members[r];
        if (mem.empty()) continue;
        bool patch = false;
        for (int p : mem) patch = patch || ppoint[p];
        std::vector<std::pair<int, double>> e;
        for (int p : mem) {
            if (patch) {
                for (int f : pfaces[p])
                    if (vface[f]) e.emplace_back(m->owner[f], 1.0 / dist(&P[3 * p], &fc[3 * (size_t)f]));
            } else {
                for (int c : pcells[p]) e.emplace_back(c, 1.0 / dist(&P[3 * p], &cctr[3 * (size_t)c]));
            }
        }
        double sum = 0.0;
        for (const auto& x : e) sum = sum + x.second;
        for (auto& x : e) x.second = x.second / sum;
        for (int p : mem) ent[p] = e;
    }
    for (int p = 0; p < np; ++p) g.mi_poff[p + 1] = g.mi_poff[p] + (uint32_t)ent[p].size();
    g.mi_pcell.resize(g.mi_poff[np]);
    g.mi_pw.resize(g.mi_poff[np]);
    for (int p = 0; p < np; ++p)
        for (size_t k = 0; k < ent[p].size(); ++k) {
            g.mi_pcell[g.mi_poff[p] + k] = ent[p][k].first;
            g.mi_pw[g.mi_poff[p] + k] = ent[p][k].second;
        }
    return UGX_OK;
}

}  // namespace ugx
