// ugx_world.hip - host side of the C ABI (include/ugx.h): mesh geometry,
// descriptors -> device arrays, per-step launch sequence, downloads and the
// write-time field computation.
//
// A "world" holds one or more independent cases in shared device arrays:
// global cell ids are the concatenation of the cases' cells, species and
// boundary models are per-case blocks.  The plain solver (ugx_case_*) is a
// world with one case; the batched and served forms put many cases in one
// world and advance them with the same kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "ugx_kernels.hpp"
#include "ugx_world.hpp"

namespace ugx {

thread_local std::string g_err;

int set_err(int code, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return set_err(UGX_ERR_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

// ------------------------------------------------------------- geometry
// Same operation order as the CPU oracle: face centre = vertex mean, area
// vector = fan about the centre, cell reference point = mean of unique
// vertices (first appearance over faces in id order), outward planes in the
// cell frame, volume by the divergence theorem.

static void cross3(const double a[3], const double b[3], double r[3])
{
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

static double dot3(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

int build_geometry(const ugx_mesh_desc* m, const std::vector<ugx_boundary_desc>& bnd, int ns,
                   const double* mass, const ugx_model_desc& model, MeshGeom& g)
{
    const int nc = m->n_cells, nf = m->n_faces;
    if (nc <= 0 || nf <= 0 || !m->points || !m->face_offsets || !m->face_points || !m->owner)
        return set_err(UGX_ERR_ARG, "mesh descriptor incomplete");
    if (m->n_internal_faces > 0 && !m->neighbour) return set_err(UGX_ERR_ARG, "mesh neighbour list missing");
    g.nc = nc;
    g.cc.assign((size_t)nc * 3, 0.0);
    g.vol.assign(nc, 0.0);
    g.bblo.assign((size_t)nc * 3, 0.0);
    g.bbhi.assign((size_t)nc * 3, 0.0);
    g.ncf.assign(nc, 0);
    g.plane.assign((size_t)nc * kMaxFaces * 4, 0.0);
    g.nbr.assign((size_t)nc * kMaxFaces, kNbrEmpty);
    g.farea.assign((size_t)nc * kMaxFaces, 0.0);
    g.fkind.assign((size_t)nc * kMaxFaces, kFaceSkip);
    g.delta.assign((size_t)nc * kMaxFaces * 3, 0.0);
    std::vector<int32_t> cf((size_t)nc * kMaxFaces);
    std::vector<double> pc((size_t)nf * 3), A((size_t)nf * 3);
    std::vector<int> fpatch(nf, -1);
    for (int f = 0; f < nf; ++f) {
        int b = m->face_offsets[f], e = m->face_offsets[f + 1], k = e - b;
        if (k < 3) return set_err(UGX_ERR_MESH, "face %d has fewer than 3 vertices", f);
        double s[3] = {0.0, 0.0, 0.0};
        for (int i = b; i < e; ++i)
            for (int d = 0; d < 3; ++d) s[d] = s[d] + m->points[3 * m->face_points[i] + d];
        for (int d = 0; d < 3; ++d) pc[3 * f + d] = s[d] / (double)k;
        double a[3] = {0.0, 0.0, 0.0};
        for (int i = 0; i < k; ++i) {
            const double* p0 = &m->points[3 * m->face_points[b + i]];
            const double* p1 = &m->points[3 * m->face_points[b + (i + 1) % k]];
            double u[3], w[3], c[3];
            for (int d = 0; d < 3; ++d) { u[d] = p0[d] - pc[3 * f + d]; w[d] = p1[d] - pc[3 * f + d]; }
            cross3(u, w, c);
            for (int d = 0; d < 3; ++d) a[d] = a[d] + c[d];
        }
        for (int d = 0; d < 3; ++d) A[3 * f + d] = 0.5 * a[d];
    }
    for (int p = 0; p < m->n_patches; ++p)
        for (int f = m->patch_start[p]; f < m->patch_start[p] + m->patch_size[p]; ++f) {
            if (f < 0 || f >= nf) return set_err(UGX_ERR_MESH, "patch %d face range out of bounds", p);
            fpatch[f] = p;
        }
    for (int f = 0; f < nf; ++f) {
        int o = m->owner[f];
        if (o < 0 || o >= nc) return set_err(UGX_ERR_MESH, "face %d owner out of range", f);
        if (g.ncf[o] >= kMaxFaces) return set_err(UGX_ERR_MESH, "cell %d has more than 6 faces", o);
        cf[(size_t)o * kMaxFaces + g.ncf[o]++] = f;
        if (f < m->n_internal_faces) {
            int nb = m->neighbour[f];
            if (nb < 0 || nb >= nc) return set_err(UGX_ERR_MESH, "face %d neighbour out of range", f);
            if (g.ncf[nb] >= kMaxFaces) return set_err(UGX_ERR_MESH, "cell %d has more than 6 faces", nb);
            cf[(size_t)nb * kMaxFaces + g.ncf[nb]++] = f;
        }
    }
    for (int c = 0; c < nc; ++c) {
        int ids[64], nid = 0;
        for (int j = 0; j < g.ncf[c]; ++j) {
            int f = cf[(size_t)c * kMaxFaces + j];
            for (int i = m->face_offsets[f]; i < m->face_offsets[f + 1]; ++i) {
                int pid = m->face_points[i];
                bool seen = false;
                for (int q = 0; q < nid; ++q) if (ids[q] == pid) { seen = true; break; }
                if (!seen && nid < 64) ids[nid++] = pid;
            }
        }
        double s[3] = {0.0, 0.0, 0.0};
        for (int q = 0; q < nid; ++q)
            for (int d = 0; d < 3; ++d) s[d] = s[d] + m->points[3 * ids[q] + d];
        for (int d = 0; d < 3; ++d) g.cc[3 * (size_t)c + d] = s[d] / (double)nid;
        for (int d = 0; d < 3; ++d) { g.bblo[3 * (size_t)c + d] = 1.7976931348623157e308; g.bbhi[3 * (size_t)c + d] = -1.7976931348623157e308; }
        for (int q = 0; q < nid; ++q)
            for (int d = 0; d < 3; ++d) {
                double r = m->points[3 * ids[q] + d] - g.cc[3 * (size_t)c + d];
                if (r < g.bblo[3 * (size_t)c + d]) g.bblo[3 * (size_t)c + d] = r;
                if (r > g.bbhi[3 * (size_t)c + d]) g.bbhi[3 * (size_t)c + d] = r;
            }
    }
    // boundary models per patch
    g.bm.assign(std::max(m->n_patches, 1), BModel{});
    for (int p = 0; p < m->n_patches; ++p) {
        int t = m->patch_type[p];
        g.bm[p].kind = (t == UGX_PATCH_SYMMETRY) ? UGX_BND_SPECULAR : UGX_BND_DELETION;
        bool found = false;
        for (const auto& b : bnd) {
            if (b.patch != p) continue;
            if (b.kind == UGX_BND_SPECULAR || b.kind == UGX_BND_DIFFUSE || b.kind == UGX_BND_DELETION) {
                g.bm[p].kind = b.kind;
                g.bm[p].T = b.temperature;
                for (int d = 0; d < 3; ++d) g.bm[p].U[d] = b.velocity[d];
                found = true;
            }
        }
        if (t == UGX_PATCH_WALL && !found) return set_err(UGX_ERR_ARG, "wall patch %d has no boundary model", p);
    }
    // measured wall faces, numbered in patch then face order
    std::vector<int32_t> wid(nf, -1);
    g.nwf = 0;
    g.wgeo.clear(); g.wpatch.clear(); g.wlocal.clear(); g.wcell.clear();
    for (int p = 0; p < m->n_patches; ++p) {
        if (m->patch_type[p] != UGX_PATCH_WALL) continue;
        if (g.bm[p].kind != UGX_BND_SPECULAR && g.bm[p].kind != UGX_BND_DIFFUSE) continue;
        for (int f = m->patch_start[p]; f < m->patch_start[p] + m->patch_size[p]; ++f) {
            wid[f] = g.nwf++;
            const double* a = &A[3 * (size_t)f];
            const double mag = std::sqrt(dot3(a, a));
            double n[3], t1[3];
            for (int d = 0; d < 3; ++d) n[d] = a[d] / mag;
            // tangent: face centre minus the face's first vertex (uniGasVolFields.C:687-691)
            const double* v0 = &m->points[3 * (size_t)m->face_points[m->face_offsets[f]]];
            for (int d = 0; d < 3; ++d) t1[d] = pc[3 * (size_t)f + d] - v0[d];
            const double tm = std::sqrt(dot3(t1, t1));
            for (int d = 0; d < 3; ++d) t1[d] = t1[d] / tm;
            const double w[kWGeo] = {n[0], n[1], n[2], mag, t1[0], t1[1], t1[2], 0.0};
            g.wgeo.insert(g.wgeo.end(), w, w + kWGeo);
            g.wpatch.push_back(p);
            g.wlocal.push_back(f - m->patch_start[p]);
            g.wcell.push_back(m->owner[f]);
        }
    }
    g.wface.assign((size_t)nc * kMaxFaces, -1);
    for (int d = 0; d < 3; ++d) g.mask[d] = 1.0;
    for (int p = 0; p < m->n_patches; ++p) {
        if (m->patch_type[p] != UGX_PATCH_EMPTY || m->patch_size[p] == 0) continue;
        const double* a = &A[3 * (size_t)m->patch_start[p]];
        int ax = 0;
        for (int d = 1; d < 3; ++d) if (std::fabs(a[d]) > std::fabs(a[ax])) ax = d;
        g.mask[ax] = 0.0;
    }
    for (int c = 0; c < nc; ++c) {
        double vs = 0.0;
        for (int j = 0; j < kMaxFaces; ++j) {
            double* pl = &g.plane[((size_t)c * kMaxFaces + j) * 4];
            const size_t k = (size_t)c * kMaxFaces + j;
            g.farea[k] = 0.0;
            g.fkind[k] = kFaceSkip;
            if (j >= g.ncf[c]) { pl[0] = pl[1] = pl[2] = 0.0; pl[3] = 1.0; g.nbr[k] = kNbrEmpty; continue; }
            int f = cf[(size_t)c * kMaxFaces + j];
            double sg = (m->owner[f] == c) ? 1.0 : -1.0;
            double a[3] = {A[3 * (size_t)f], A[3 * (size_t)f + 1], A[3 * (size_t)f + 2]};
            double r[3];
            for (int d = 0; d < 3; ++d) r[d] = pc[3 * (size_t)f + d] - g.cc[3 * (size_t)c + d];
            double dt = dot3(a, r);
            vs = vs + sg * dt;
            double mag = std::sqrt(dot3(a, a));
            double nn[3];
            for (int d = 0; d < 3; ++d) nn[d] = sg * (a[d] / mag);
            pl[0] = nn[0]; pl[1] = nn[1]; pl[2] = nn[2];
            pl[3] = dot3(nn, r);
            g.farea[k] = mag;
            double* dl = &g.delta[((size_t)c * kMaxFaces + j) * 3];
            if (f < m->n_internal_faces) {
                int o = (m->owner[f] == c) ? m->neighbour[f] : m->owner[f];
                g.nbr[(size_t)c * kMaxFaces + j] = o;
                g.fkind[k] = kFaceInternal;
                for (int d = 0; d < 3; ++d) dl[d] = g.cc[3 * (size_t)c + d] - g.cc[3 * (size_t)o + d];
            } else {
                int p = fpatch[f];
                if (p < 0) return set_err(UGX_ERR_MESH, "boundary face %d outside every patch", f);
                int t = m->patch_type[p];
                if (t == UGX_PATCH_EMPTY) {
                    g.nbr[(size_t)c * kMaxFaces + j] = kNbrEmpty;
                    g.farea[k] = 0.0;
                } else if (t == UGX_PATCH_CYCLIC) {
                    g.fkind[k] = kFaceCyclic;
                    int q = m->patch_neighbour ? m->patch_neighbour[p] : -1;
                    if (q < 0 || q >= m->n_patches || m->patch_size[q] != m->patch_size[p])
                        return set_err(UGX_ERR_MESH, "cyclic patch %d has no matching partner", p);
                    int gf = m->patch_start[q] + (f - m->patch_start[p]);
                    int o = m->owner[gf];
                    g.nbr[(size_t)c * kMaxFaces + j] = o;
                    for (int d = 0; d < 3; ++d)
                        dl[d] = (g.cc[3 * (size_t)c + d] - g.cc[3 * (size_t)o + d]) + (pc[3 * (size_t)gf + d] - pc[3 * (size_t)f + d]);
                } else {
                    g.nbr[(size_t)c * kMaxFaces + j] = -1 - p;
                    g.fkind[k] = (t == UGX_PATCH_SYMMETRY) ? kFaceSymmetry : kFaceBoundary;
                    g.wface[k] = wid[f];
                }
            }
        }
        g.vol[c] = vs / 3.0;
        if (!(g.vol[c] > 0.0)) return set_err(UGX_ERR_MESH, "cell %d has non-positive volume", c);
    }
    // freestream inflow face-species (uniGasGeneralBoundary.C:115-169)
    g.inflow.clear();
    g.inflow_parts.clear();
    const double sqrtPi = std::sqrt(kPi);
    for (size_t bi = 0; bi < bnd.size(); ++bi) {
        const auto& b = bnd[bi];
        if (b.kind != UGX_BND_FREESTREAM) continue;
        int p = b.patch;
        if (p < 0 || p >= m->n_patches) return set_err(UGX_ERR_ARG, "inflow patch %d out of range", p);
        for (int f = m->patch_start[p]; f < m->patch_start[p] + m->patch_size[p]; ++f) {
            for (int s = 0; s < ns; ++s) {
                double nd = b.number_density[s];
                if (!(nd > 0.0)) continue;
                InflowFS e{};
                int c = m->owner[f];
                e.cell = c; e.species = s; e.ccase = 0; e.local_id = (int)g.inflow.size();
                e.T = b.temperature; e.rotT = b.rot_temperature;
                for (int d = 0; d < 3; ++d) e.U[d] = b.velocity[d];
                const double* sF = &A[3 * (size_t)f];
                double fA = std::sqrt(dot3(sF, sF));
                double mps = std::sqrt(2.0 * kB * e.T / mass[s]);
                double nrm[3] = {-sF[0] / fA, -sF[1] / fA, -sF[2] / fA};
                double sCos = dot3(e.U, nrm) / mps;
                const double fAn = fA * nd;
                const double br = std::exp(-(sCos * sCos)) + sqrtPi * sCos * (1.0 + std::erf(sCos));
                e.nbar_num = fAn * model.delta_t * mps * br;
                e.nbar_den = 2.0 * sqrtPi * model.n_equivalent_particles;
                g.inflow_parts.push_back(fAn);
                g.inflow_parts.push_back(mps);
                g.inflow_parts.push_back(br);
                e.nbar = e.nbar_num / e.nbar_den;
                int fb = m->face_offsets[f], fe = m->face_offsets[f + 1];
                e.nv = fe - fb;
                if (e.nv > 8) return set_err(UGX_ERR_MESH, "inflow face with more than 8 vertices");
                for (int i = 0; i < e.nv; ++i)
                    for (int d = 0; d < 3; ++d)
                        e.vloc[i][d] = m->points[3 * m->face_points[fb + i] + d] - g.cc[3 * (size_t)c + d];
                double tot = 0.0, ar[6];
                for (int t = 0; t < e.nv - 2; ++t) {
                    double u[3], ww[3], cr[3];
                    for (int d = 0; d < 3; ++d) { u[d] = e.vloc[t + 1][d] - e.vloc[0][d]; ww[d] = e.vloc[t + 2][d] - e.vloc[0][d]; }
                    cross3(u, ww, cr);
                    ar[t] = 0.5 * std::sqrt(dot3(cr, cr));
                    tot = tot + ar[t];
                }
                double run = 0.0;
                for (int t = 0; t < e.nv - 2; ++t) { run = run + ar[t]; e.cum[t] = run / tot; }
                e.cum[e.nv - 3] = 1.0;
                for (int d = 0; d < 3; ++d) e.n[d] = nrm[d];
                double t1[3];
                for (int d = 0; d < 3; ++d) t1[d] = pc[3 * (size_t)f + d] - m->points[3 * m->face_points[fb] + d];
                double tm = std::sqrt(dot3(t1, t1));
                for (int d = 0; d < 3; ++d) e.t1[d] = t1[d] / tm;
                double t2[3];
                cross3(e.n, e.t1, t2);
                double t2m = std::sqrt(dot3(t2, t2));
                for (int d = 0; d < 3; ++d) e.t2[d] = t2[d] / t2m;
                g.inflow.push_back(e);
            }
        }
    }
    if (model.macro_interpolation)
        if (int rc = build_interp(m, g)) return rc;
    return UGX_OK;
}

// --------------------------------------------------------- device buffers

template <class T>
static int dalloc(T*& p, size_t n)
{
    p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void**)&p, n * sizeof(T));
    if (e != hipSuccess) return set_err(UGX_ERR_DEVICE, "hipMalloc(%zu bytes) failed: %s", n * sizeof(T), hipGetErrorString(e));
    return UGX_OK;
}

template <class T>
static void dfree(T*& p)
{
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

World::~World() { release(); }

void World::drop_graphs()
{
    for (auto& g : gexec) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
}

void World::release()
{
    drop_graphs();
    for (auto& e : tune_ev) if (e) (void)hipEventDestroy(e);
    tune_ev.clear();
    for (auto& e : tev) if (e) (void)hipEventDestroy(e);
    tev.clear();
    for (int k = 0; k < 2; ++k)
        for (int a = 0; a < 9; ++a) dfree(pbuf[k][a]);
    void** ptrs[] = {(void**)&d_plane, (void**)&d_delta, (void**)&d_nbr, (void**)&d_bm, (void**)&d_inflow,
                     (void**)&d_cases, (void**)&d_cell_case, (void**)&d_vol, (void**)&d_smax, (void**)&d_cmodel,
                     (void**)&d_samp, (void**)&d_uspq, (void**)&d_usps, (void**)&d_uspmax, (void**)&d_knacc,
                     (void**)&d_kn, (void**)&d_cc, (void**)&d_bblo, (void**)&d_bbhi, (void**)&d_mass, (void**)&d_rotdof, (void**)&d_omega,
                     (void**)&d_diam, (void**)&d_knf, (void**)&d_knaw, (void**)&d_fkind, (void**)&d_kw, (void**)&d_kw2, (void**)&d_ids2,
                     (void**)&d_nb_off, (void**)&d_nb_list,
                     (void**)&d_muref, (void**)&d_prterm, (void**)&d_local, (void**)&d_pmr, (void**)&d_pC,
                     (void**)&d_pe, (void**)&d_pchi, (void**)&d_palpha, (void**)&d_cnt, (void**)&d_ocnt, (void**)&d_out7, (void**)&d_pre, (void**)&d_rev, (void**)&d_off,
                     (void**)&d_perm, (void**)&d_ins_off, (void**)&d_scratch,
                     (void**)&d_tile_first, (void**)&d_tile_p0, (void**)&d_tile_case,
                     (void**)&d_mom, (void**)&d_bgkp, (void**)&d_raise,
                     (void**)&d_sel_slot,
                     (void**)&d_blocksum, (void**)&d_counters, (void**)&d_st,
                     (void**)&d_wq_slot, (void**)&d_wq_info, (void**)&d_wq_tau,
                     (void**)&d_wface, (void**)&d_wgeo, (void**)&d_wcell, (void**)&d_ev, (void**)&d_ev_key,
                     (void**)&d_wcnt, (void**)&d_woff, (void**)&d_bkey, (void**)&d_bidx, (void**)&d_wsamp,
                     (void**)&d_cw, (void**)&d_cstart, (void**)&d_wclone, (void**)&d_wcoff, (void**)&d_wbsum,
                     (void**)&d_sclev, (void**)&d_scoff, (void**)&d_sctab, (void**)&d_scid, (void**)&d_sclist,
                     (void**)&d_adacc, (void**)&d_cw_old, (void**)&d_mi_cent, (void**)&d_mi_fp, (void**)&d_mi_pts,
                     (void**)&d_mi_poff, (void**)&d_mi_pcell, (void**)&d_mi_pw, (void**)&d_mfield, (void**)&d_pfield};
    for (auto pp : ptrs) { if (*pp) (void)hipFree(*pp); *pp = nullptr; }
    if (stream) { (void)hipStreamDestroy(stream); stream = nullptr; }
}

size_t World::rsize() const { return precision == UGX_FP32_STATE ? sizeof(float) : sizeof(double); }

int World::alloc_particles(uint32_t newcap)
{
    drop_graphs();
    void* nb[2][9] = {};
    size_t rs = rsize();
    for (int k = 0; k < 2; ++k)
        for (int a = 0; a < 9; ++a) {
            size_t es = (a < 7) ? rs : sizeof(uint32_t);
            hipError_t e = hipMalloc(&nb[k][a], (size_t)newcap * es);
            if (e != hipSuccess) {
                for (int kk = 0; kk < 2; ++kk) for (int aa = 0; aa < 9; ++aa) if (nb[kk][aa]) (void)hipFree(nb[kk][aa]);
                return set_err(UGX_ERR_DEVICE, "particle allocation (%u simulators) failed: %s", newcap, hipGetErrorString(e));
            }
        }
    uint32_t n = 0;
    if (cap > 0) {
        HIP_TRY(hipMemcpyAsync(&n, &d_st->n, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        for (int a = 0; a < 9; ++a) {
            size_t es = (a < 7) ? rs : sizeof(uint32_t);
            if (n) HIP_TRY(hipMemcpyAsync(nb[cur][a], pbuf[cur][a], (size_t)n * es, hipMemcpyDeviceToDevice, stream));
        }
        HIP_TRY(hipStreamSynchronize(stream));
    }
    for (int k = 0; k < 2; ++k)
        for (int a = 0; a < 9; ++a) { if (pbuf[k][a]) (void)hipFree(pbuf[k][a]); pbuf[k][a] = nullptr; }
    for (int a = 0; a < 9; ++a) { pbuf[0][a] = nb[cur][a]; pbuf[1][a] = nb[1 - cur][a]; }
    cur = 0;
    dfree(d_perm);
    dfree(d_scratch);
    // the tile table of the last sort describes the particles just carried
    // over; the occupancy count of the next step reads it
    {
        const size_t otiles = cap > 0 ? (size_t)max_tiles(cap, (uint64_t)nc) : 0;
        const size_t ntiles_new = (size_t)max_tiles(newcap, (uint64_t)nc);
        uint32_t *tf = nullptr, *tp = nullptr;
        int32_t* tc = nullptr;
        if (int rc = dalloc(tf, ntiles_new)) return rc;
        if (int rc = dalloc(tp, ntiles_new)) { dfree(tf); return rc; }
        if (int rc = dalloc(tc, ntiles_new)) { dfree(tf); dfree(tp); return rc; }
        if (otiles > 0) {
            HIP_TRY(hipMemcpyAsync(tf, d_tile_first, otiles * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream));
            HIP_TRY(hipMemcpyAsync(tp, d_tile_p0, otiles * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream));
            HIP_TRY(hipMemcpyAsync(tc, d_tile_case, otiles * sizeof(int32_t), hipMemcpyDeviceToDevice, stream));
            HIP_TRY(hipStreamSynchronize(stream));
        }
        dfree(d_tile_first);
        dfree(d_tile_p0);
        dfree(d_tile_case);
        d_tile_first = tf;
        d_tile_p0 = tp;
        d_tile_case = tc;
    }
    dfree(d_sel_slot);
    dfree(d_wq_slot);
    dfree(d_wq_info);
    dfree(d_wq_tau);
    if (int rc = dalloc(d_perm, newcap)) return rc;
    if (int rc = dalloc(d_scratch, newcap)) return rc;
    if (has_bgk) {
        if (int rc = dalloc(d_sel_slot, newcap)) return rc;
    }
    // diffuse-wall queue: at most half of the simulators may reach a diffuse wall in one step
    wq_cap = std::max<uint32_t>(newcap / 2, 1024u);
    if (int rc = dalloc(d_wq_slot, wq_cap)) return rc;
    if (int rc = dalloc(d_wq_info, wq_cap)) return rc;
    if (int rc = dalloc(d_wq_tau, wq_cap)) return rc;
    if (has_weighting) {
        dfree(d_cstart);
        dfree(d_wclone);
        dfree(d_wcoff);
        dfree(d_wbsum);
        if (int rc = dalloc(d_cstart, newcap)) return rc;
        if (int rc = dalloc(d_wclone, (size_t)newcap + 1)) return rc;
        if (int rc = dalloc(d_wcoff, (size_t)newcap + 1)) return rc;
        if (int rc = dalloc(d_wbsum, scan_bsum_words((uint64_t)newcap + 1))) return rc;
    }
    if (has_subcells) {
        dfree(d_scid);
        dfree(d_sclist);
        if (int rc = dalloc(d_scid, newcap)) return rc;
        if (int rc = dalloc(d_sclist, newcap)) return rc;
    }
    cap = newcap;
    if (nwf > 0 && ev_cap < std::max<uint32_t>(newcap / 8, 65536u))
        if (int rc = alloc_events(std::max<uint32_t>(newcap / 8, 65536u))) return rc;
    return UGX_OK;
}

// Event buffers of the wall measurements for n hits per step (the event
// list, its keys, and the per-face bucket copy).
int World::alloc_events(uint32_t n)
{
    drop_graphs();
    dfree(d_ev);
    dfree(d_ev_key);
    dfree(d_bkey);
    dfree(d_bidx);
    ev_cap = 0;
    if (int rc = dalloc(d_ev, n)) return rc;
    if (int rc = dalloc(d_ev_key, n)) return rc;
    if (int rc = dalloc(d_bkey, n)) return rc;
    if (int rc = dalloc(d_bidx, n)) return rc;
    ev_cap = n;
    return UGX_OK;
}

int World::ensure_capacity(uint64_t need)
{
    if (need <= cap) return UGX_OK;
    if (need > 0xF0000000ull) return set_err(UGX_ERR_CAPACITY, "more than 4e9 simulators requested");
    uint64_t nc = std::max<uint64_t>(need + need / 4 + 1024, 4096);
    if (nc > 0xF0000000ull) nc = 0xF0000000ull;
    return alloc_particles((uint32_t)nc);
}

template <class R>
WorldView<R> World::view() const
{
    WorldView<R> w{};
    auto mk = [&](int k) {
        Parts<R> p;
        p.x = (R*)pbuf[k][0]; p.y = (R*)pbuf[k][1]; p.z = (R*)pbuf[k][2];
        p.vx = (R*)pbuf[k][3]; p.vy = (R*)pbuf[k][4]; p.vz = (R*)pbuf[k][5];
        p.erot = (R*)pbuf[k][6];
        p.cell = (uint32_t*)pbuf[k][7]; p.sp = (uint32_t*)pbuf[k][8];
        return p;
    };
    w.a = mk(cur);
    w.b = mk(1 - cur);
    w.geom.plane = (const V4<R>*)d_plane;
    w.geom.delta = (const V4<R>*)d_delta;
    w.geom.nbr = d_nbr;
    w.bm = d_bm;
    w.inflow = d_inflow;
    w.n_inflow = n_inflow;
    w.max_ins = max_ins;
    w.cases = d_cases;
    w.ncases = (int32_t)cases.size();
    w.cell_case = d_cell_case;
    w.vol = d_vol;
    w.smax = d_smax;
    w.cmodel = d_cmodel;
    w.samp = d_samp;
    w.uspq = d_uspq;
    w.usps = d_usps;
    w.uspmax = d_uspmax;
    w.knacc = d_knacc;
    w.kn = d_kn;
    w.cc = d_cc;
    w.sp = Species{d_mass, d_diam, d_rotdof, d_omega, d_muref, d_prterm, d_local};
    w.has_hybrid = has_hybrid;
    w.tile_variant = tile_variant;
    w.cons_variant = cons_variant;
    w.move_grid = move_grid;
    w.kn_due = 0;
    w.kn_passes = 0;
    w.knf = d_knf;
    w.knaw = d_knaw;
    w.fkind = d_fkind;
    w.kw = d_kw;
    w.kw2 = d_kw2;
    w.ids2 = d_ids2;
    w.nb_off = d_nb_off;
    w.nb_list = d_nb_list;
    w.pr = Pairs{d_pmr, d_pC, d_pe, d_pchi, d_palpha};
    w.cnt = d_cnt;
    w.ocnt = d_ocnt;
    w.out7 = d_out7;
    w.pre = d_pre;
    w.rev = d_rev;
    w.off = d_off;
    w.perm = d_perm;
    w.ins_off = d_ins_off;
    w.scratch = d_scratch;
    w.tile_first = d_tile_first;
    w.tile_p0 = d_tile_p0;
    w.tile_case = d_tile_case;
    w.nsp = nsp_total;
    w.mom = d_mom;
    w.bgkp = d_bgkp;
    w.raise = d_raise;
    w.sel_slot = d_sel_slot;
    w.ns_max = ns_max;
    w.sum_level = sum_level;
    w.has_collide = has_collide;
    w.has_bgk = has_bgk;
    w.blocksum = d_blocksum;
    w.counters = d_counters;
    w.st = d_st;
    w.nc = nc;
    w.cap = cap;
    w.wq_slot = d_wq_slot;
    w.wq_info = d_wq_info;
    w.wq_tau = d_wq_tau;
    w.wq_cap = wq_cap;
    w.nwf = nwf;
    w.wface = d_wface;
    w.wgeo = d_wgeo;
    w.wcell = d_wcell;
    w.ev = d_ev;
    w.ev_key = d_ev_key;
    w.ev_cap = ev_cap;
    w.wcnt = d_wcnt;
    w.woff = d_woff;
    w.bkey = d_bkey;
    w.bidx = d_bidx;
    w.wsamp = d_wsamp;
    w.cw = d_cw;
    w.has_weighting = has_weighting;
    w.cstart = d_cstart;
    w.wclone = d_wclone;
    w.wcoff = d_wcoff;
    w.wbsum = d_wbsum;
    w.has_subcells = has_subcells;
    w.sclev = d_sclev;
    w.scoff = d_scoff;
    w.sctab = d_sctab;
    w.scid = d_scid;
    w.sclist = d_sclist;
    w.bblo = d_bblo;
    w.bbhi = d_bbhi;
    w.adacc = d_adacc;
    w.has_macro = has_macro;
    w.npts = npts;
    w.mi_cent = d_mi_cent;
    w.mi_fp = d_mi_fp;
    w.mi_pts = d_mi_pts;
    w.mi_poff = d_mi_poff;
    w.mi_pcell = d_mi_pcell;
    w.mi_pw = d_mi_pw;
    w.mfield = d_mfield;
    w.pfield = d_pfield;
    return w;
}

template WorldView<float> World::view<float>() const;
template WorldView<double> World::view<double>() const;

// Builds the device arrays for the cases in `specs` (all cells concatenated).
int World::build(std::vector<CaseSpec>& specs, int prec, int device, uint64_t capacity)
{
    precision = prec;
    HIP_TRY(hipSetDevice(device));
    HIP_TRY(hipStreamCreate(&stream));  // blocking: ordered with null-stream copies
    nc = 0;
    int nsp = 0, npair = 0;
    for (auto& s : specs) { nc += s.geom.nc; nsp += s.ns; npair += s.ns * s.ns; }
    move_grid = (precision == UGX_FP32_STATE) ? move_blocks<float>(device) : move_blocks<double>(device);
    if (move_grid == 0) return set_err(UGX_ERR_DEVICE, "occupancy query of the move kernel failed");
    // cell tiles keep a 16-bit species id per simulator
    if (nsp > 65535) return UGX_ERR_CAPACITY;
    nsp_total = nsp;
    std::vector<CaseDev> hc;
    std::vector<int32_t> cell_case(nc), nbr((size_t)nc * kMaxFaces), cmodel(nc), local(nsp), rotdof(nsp);
    std::vector<double> vol(nc), cc((size_t)nc * 3), plane((size_t)nc * kMaxFaces * 4), delta((size_t)nc * kMaxFaces * 4, 0.0);
    std::vector<double> diam(nsp);
    std::vector<KnFace> knf;
    std::vector<int8_t> fkind;
    for (auto& sp_ : specs)
        if (sp_.model.collision_model == UGX_COLL_HYBRID) has_hybrid = 1;
    if (has_hybrid) {
        knf.assign((size_t)nc * kMaxFaces, KnFace{0.0, 0.0, 0.0, 0.0, 0.0, 0.0});
        fkind.assign((size_t)nc * kMaxFaces, kFaceSkip);
    }
    std::vector<double> mass(nsp), omega(nsp), muref(nsp), prterm(nsp), pmr(npair), pC(npair), pe(npair), pchi(npair), palpha(npair);
    std::vector<BModel> bm;
    std::vector<InflowFS> inflow;
    std::vector<int32_t> wface_all((size_t)nc * kMaxFaces, -1), wcell_all;
    std::vector<double> wgeo_all;
    wf_begin.assign(1, 0);
    int cbase = 0, sbase = 0, pbase = 0;
    max_ins = 0;
    for (size_t ci = 0; ci < specs.size(); ++ci) {
        CaseSpec& s = specs[ci];
        const MeshGeom& g = s.geom;
        CaseDev d{};
        d.dt = s.model.delta_t;
        d.fn = s.model.n_equivalent_particles;
        d.inv_zrot = (s.model.Zrot > 0.0) ? 1.0 / s.model.Zrot : 0.0;
        d.theta = s.model.theta;
        d.tref = s.model.Tref;
        for (int k = 0; k < 3; ++k) d.mask[k] = g.mask[k];
        d.seed_lo = (uint32_t)s.seed;
        d.seed_hi = (uint32_t)(s.seed >> 32);
        d.step = 0;
        d.sample_interval = s.model.sample_interval > 0 ? s.model.sample_interval : 1;
        d.coll_model = s.model.collision_model;
        d.dsmc_model = s.model.dsmc_model;
        d.bgk_model = s.model.bgk_model;
        d.ns = s.ns;
        d.sp_base = sbase;
        d.pair_base = pbase;
        d.bm_base = (int32_t)bm.size();
        d.cell_begin = cbase;
        d.cell_end = cbase + g.nc;
        d.inflow_begin = (int32_t)inflow.size();
        d.active = 1;
        d.decomposition_interval = s.model.decomposition_interval > 0 ? s.model.decomposition_interval : 100;
        d.smoothing_passes = s.model.smoothing_passes;
        d.breakdown_max = s.model.breakdown_max;
        d.kn_theta = s.model.kn_theta;
        d.sampler_full = s.model.sampler_full ? 1 : 0;
        d.reset_at_decomposition = s.model.reset_at_decomposition;
        d.weighted = s.model.cell_weighted ? 1 : 0;
        if (d.weighted) has_weighting = 1;
        d.reset_until_time = s.model.reset_until_time;
        d.time = 0.0;
        d.adaptive = s.model.adaptive ? 1 : 0;
        // macroInterpolation reaches the BGK-family draws only
        d.macro = (s.model.macro_interpolation && s.model.bgk_model != UGX_BGK_NONE &&
                   (s.model.collision_model == UGX_COLL_BGK || s.model.collision_model == UGX_COLL_HYBRID)) ? 1 : 0;
        if (d.macro) has_macro = 1;
        if (d.adaptive) has_adaptive = 1;
        {
            int lvl = d.sampler_full ? 2 : 0;
            const bool bgk_on = (s.model.collision_model == UGX_COLL_BGK || s.model.collision_model == UGX_COLL_HYBRID);
            if (bgk_on && s.model.bgk_model == UGX_BGK_ESBGK) lvl = std::max(lvl, 1);
            if (bgk_on && s.model.bgk_model == UGX_BGK_USP) lvl = 2;
            d.sum_level = lvl;
            sum_level = std::max(sum_level, lvl);
            ns_max = std::max(ns_max, s.ns);
            const bool dsmc_on = (s.model.collision_model == UGX_COLL_DSMC || s.model.collision_model == UGX_COLL_HYBRID)
                                 && s.model.dsmc_model != UGX_DSMC_NONE;
            if (bgk_on && s.model.bgk_model != UGX_BGK_NONE) { has_bgk = 1; has_collide = 1; }
            if (dsmc_on) has_collide = 1;
        }
        int cm = (s.model.collision_model == UGX_COLL_DSMC) ? 1 : 0;  // hybrid starts all-BGK (uniGasCloud.C:726-732)
        for (int c = 0; c < g.nc; ++c) {
            int gc = cbase + c;
            cell_case[gc] = (int32_t)ci;
            vol[gc] = g.vol[c];
            cmodel[gc] = cm;
            for (int d3 = 0; d3 < 3; ++d3) cc[3 * (size_t)gc + d3] = g.cc[3 * (size_t)c + d3];
            for (int f = 0; f < kMaxFaces; ++f) {
                size_t lk = (size_t)c * kMaxFaces + f, gk = (size_t)gc * kMaxFaces + f;
                for (int q = 0; q < 4; ++q) plane[gk * 4 + q] = g.plane[lk * 4 + q];
                for (int q = 0; q < 3; ++q) delta[gk * 4 + q] = g.delta[lk * 3 + q];
                int32_t code = g.nbr[lk];
                if (code >= 0) code += cbase;
                else if (code != kNbrEmpty) code = -1 - (d.bm_base + (-1 - code));
                nbr[gk] = code;
                if (has_hybrid) {
                    // face interpolation data, in the CPU oracle's operation order
                    const double* pl = &g.plane[lk * 4];
                    const double* dl = &g.delta[lk * 3];
                    const int8_t kind = g.fkind[lk];
                    KnFace kf{g.farea[lk], 0.0, 0.0, pl[0], pl[1], pl[2]};
                    if (kind == kFaceInternal || kind == kFaceCyclic) {
                        const double dP = std::fabs(pl[3]);
                        const double dN = std::fabs((pl[0] * dl[0] + pl[1] * dl[1] + pl[2] * dl[2]) + pl[3]);
                        kf.w = dN / (dP + dN);
                        kf.dist = std::sqrt(dl[0] * dl[0] + dl[1] * dl[1] + dl[2] * dl[2]);
                    }
                    knf[gk] = kf;
                    fkind[gk] = kind;
                }
            }
        }
        {
            const int32_t wb = wf_begin.back();
            for (size_t k = 0; k < g.wface.size(); ++k)
                if (g.wface[k] >= 0) wface_all[(size_t)cbase * kMaxFaces + k] = wb + g.wface[k];
            for (int k = 0; k < g.nwf; ++k) wcell_all.push_back(cbase + g.wcell[k]);
            wgeo_all.insert(wgeo_all.end(), g.wgeo.begin(), g.wgeo.end());
            wf_begin.push_back(wb + g.nwf);
        }
        for (auto b : g.bm) bm.push_back(b);
        for (auto e : g.inflow) {
            e.cell += cbase;
            e.species += sbase;
            e.ccase = (int32_t)ci;
            inflow.push_back(e);
            max_ins += (uint32_t)std::floor(e.nbar) + 1u;
        }
        d.inflow_end = (int32_t)inflow.size();
        for (int k = 0; k < s.ns; ++k) {
            const auto& sd = s.species[k];
            int gk = sbase + k;
            mass[gk] = sd.mass;
            omega[gk] = sd.omega;
            diam[gk] = sd.diameter;
            rotdof[gk] = sd.rotational_dof;
            local[gk] = k;
            double a = sd.alpha, om = sd.omega, dd = sd.diameter, m = sd.mass;
            muref[gk] = 1.25 * (1.0 + a) * (2.0 + a) * std::sqrt(m * kB * s.model.Tref)
                        / (a * (5.0 - 2.0 * om) * (7.0 - 2.0 * om) * std::sqrt(kPi) * (dd * dd));
            double z = (double)sd.rotational_dof;
            prterm[gk] = (5.0 + z) / (7.5 + z);
        }
        for (int p = 0; p < s.ns; ++p)
            for (int q = 0; q < s.ns; ++q) {
                int k = pbase + p * s.ns + q;
                double mP = s.species[p].mass, mQ = s.species[q].mass;
                double mR = mP * mQ / (mP + mQ);
                double dPQ = 0.5 * (s.species[p].diameter + s.species[q].diameter);
                double om = 0.5 * (s.species[p].omega + s.species[q].omega);
                pmr[k] = mR;
                pC[k] = kPi * dPQ * dPQ * std::pow(2.0 * kB * s.model.Tref / mR, om - 0.5) / std::exp(std::lgamma(2.5 - om));
                pe[k] = 1.0 - om;
                pchi[k] = 2.5 - om;
                palpha[k] = 0.5 * (s.species[p].alpha + s.species[q].alpha);
            }
        hc.push_back(d);
        s.cell_begin = cbase;
        cbase += g.nc;
        sbase += s.ns;
        pbase += s.ns * s.ns;
    }
    n_inflow = (int32_t)inflow.size();
    cases = hc;
    // precision-dependent geometry
    int rc;
    size_t rs = rsize();
    HIP_TRY(hipMalloc(&d_plane, (size_t)nc * kMaxFaces * 4 * rs));
    HIP_TRY(hipMalloc(&d_delta, (size_t)nc * kMaxFaces * 4 * rs));
    if (precision == UGX_FP32_STATE) {
        std::vector<float> pf(plane.size()), df(delta.size());
        for (size_t i = 0; i < plane.size(); ++i) pf[i] = (float)plane[i];
        for (size_t i = 0; i < delta.size(); ++i) df[i] = (float)delta[i];
        HIP_TRY(hipMemcpy(d_plane, pf.data(), pf.size() * 4, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d_delta, df.data(), df.size() * 4, hipMemcpyHostToDevice));
    } else {
        HIP_TRY(hipMemcpy(d_plane, plane.data(), plane.size() * 8, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d_delta, delta.data(), delta.size() * 8, hipMemcpyHostToDevice));
    }
    auto up = [&](auto*& dst, const auto& v) -> int {
        using T = std::remove_reference_t<decltype(*dst)>;
        if (int r = dalloc(dst, v.size())) return r;
        if (!v.empty()) {
            hipError_t e = hipMemcpy((void*)dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
            if (e != hipSuccess) return set_err(UGX_ERR_DEVICE, "upload failed: %s", hipGetErrorString(e));
        }
        return UGX_OK;
    };
    if ((rc = up(d_nbr, nbr))) return rc;
    {
        // for each face, the first face of the neighbour behind it that points back
        std::vector<int8_t> rev((size_t)nc * kMaxFaces, (int8_t)-1);
        for (size_t c = 0; c < (size_t)nc; ++c)
            for (int f = 0; f < kMaxFaces; ++f) {
                const int32_t nb = nbr[c * kMaxFaces + f];
                if (nb < 0) continue;
                for (int g = 0; g < kMaxFaces; ++g)
                    if (nbr[(size_t)nb * kMaxFaces + g] == (int32_t)c) { rev[c * kMaxFaces + f] = (int8_t)g; break; }
            }
        if ((rc = up(d_rev, rev))) return rc;
    }
    if ((rc = up(d_bm, bm))) return rc;
    if ((rc = up(d_inflow, inflow))) return rc;
    if ((rc = up(d_cases, hc))) return rc;
    if ((rc = up(d_cell_case, cell_case))) return rc;
    if ((rc = up(d_vol, vol))) return rc;
    if ((rc = up(d_cc, cc))) return rc;
    {
        std::vector<double> lo((size_t)nc * 3), hi((size_t)nc * 3);
        size_t o = 0;
        for (const auto& sp_ : specs) {
            std::copy(sp_.geom.bblo.begin(), sp_.geom.bblo.end(), lo.begin() + (ptrdiff_t)o);
            std::copy(sp_.geom.bbhi.begin(), sp_.geom.bbhi.end(), hi.begin() + (ptrdiff_t)o);
            o += sp_.geom.bblo.size();
        }
        if ((rc = up(d_bblo, lo))) return rc;
        if ((rc = up(d_bbhi, hi))) return rc;
    }
    if (has_macro) {
        // the cases' interpolation geometry, point and cell ids made world-wide
        std::vector<double> cent((size_t)nc * 3, 0.0), pts, pw;
        std::vector<int32_t> fp((size_t)nc * kMaxFaces * kMiFaceVerts, -1), pcell;
        std::vector<uint32_t> poff(1, 0u);
        int32_t pb = 0;
        for (size_t ci = 0; ci < specs.size(); ++ci) {
            const MeshGeom& gi = specs[ci].geom;
            const int cb = hc[ci].cell_begin;
            if (!hc[ci].macro) continue;
            std::copy(gi.mi_cent.begin(), gi.mi_cent.end(), cent.begin() + 3 * (size_t)cb);
            for (size_t k = 0; k < gi.mi_fp.size(); ++k)
                fp[(size_t)cb * kMaxFaces * kMiFaceVerts + k] = gi.mi_fp[k] < 0 ? -1 : gi.mi_fp[k] + pb;
            pts.insert(pts.end(), gi.mi_pts.begin(), gi.mi_pts.end());
            const uint32_t eb = poff.back();
            for (size_t p = 0; p + 1 < gi.mi_poff.size(); ++p) poff.push_back(eb + gi.mi_poff[p + 1]);
            for (int32_t c : gi.mi_pcell) pcell.push_back(c + cb);
            pw.insert(pw.end(), gi.mi_pw.begin(), gi.mi_pw.end());
            pb += (int32_t)(gi.mi_pts.size() / 3);
        }
        npts = (uint32_t)pb;
        if ((rc = up(d_mi_cent, cent))) return rc;
        if ((rc = up(d_mi_fp, fp))) return rc;
        if ((rc = up(d_mi_pts, pts))) return rc;
        if ((rc = up(d_mi_poff, poff))) return rc;
        if ((rc = up(d_mi_pcell, pcell))) return rc;
        if ((rc = up(d_mi_pw, pw))) return rc;
        std::vector<double> zf((size_t)nc * kMiRows, 0.0), zp((size_t)npts * kMiRows, 0.0);
        if ((rc = up(d_mfield, zf))) return rc;
        if ((rc = up(d_pfield, zp))) return rc;
    }
    if ((rc = up(d_cmodel, cmodel))) return rc;
    if ((rc = up(d_mass, mass))) return rc;
    if ((rc = up(d_rotdof, rotdof))) return rc;
    if ((rc = up(d_omega, omega))) return rc;
    if ((rc = up(d_diam, diam))) return rc;
    if (has_hybrid) {
        if ((rc = up(d_knf, knf))) return rc;
        std::vector<double2> knaw(knf.size());
        for (size_t k = 0; k < knf.size(); ++k) knaw[k] = make_double2(knf[k].area, knf[k].w);
        if ((rc = up(d_knaw, knaw))) return rc;
        if ((rc = up(d_fkind, fkind))) return rc;
        if ((rc = dalloc(d_kw, (size_t)nc * 8))) return rc;
        if ((rc = dalloc(d_kw2, (size_t)nc * 8))) return rc;
        if ((rc = dalloc(d_ids2, (size_t)nc))) return rc;
    }
    if ((rc = up(d_muref, muref))) return rc;
    if ((rc = up(d_prterm, prterm))) return rc;
    if ((rc = up(d_local, local))) return rc;
    if ((rc = up(d_pmr, pmr))) return rc;
    if ((rc = up(d_pC, pC))) return rc;
    if ((rc = up(d_pe, pe))) return rc;
    if ((rc = up(d_pchi, pchi))) return rc;
    if ((rc = up(d_palpha, palpha))) return rc;
    std::vector<double> zeros((size_t)nc, 0.0), ones((size_t)nc, 1.0);
    if ((rc = up(d_smax, zeros))) return rc;
    if ((rc = up(d_uspmax, ones))) return rc;
    if ((rc = up(d_kn, zeros))) return rc;
    std::vector<double> z3((size_t)nc * 3, 0.0), z6((size_t)nc * 6, 0.0), z8((size_t)nc * kKnRows, 0.0), zs((size_t)nc * kNSamp, 0.0);
    if ((rc = up(d_uspq, z3))) return rc;
    if ((rc = up(d_usps, z6))) return rc;
    if ((rc = up(d_knacc, z8))) return rc;
    if (has_adaptive && (rc = up(d_adacc, z8))) return rc;
    adapt_ratio.assign((size_t)nc * 3, 0.0);
    if ((rc = up(d_samp, zs))) return rc;
    std::vector<uint32_t> zc((size_t)nc + 1, 0u);
    if ((rc = up(d_cnt, zc))) return rc;
    if ((rc = up(d_ocnt, zc))) return rc;
    {
        std::vector<uint32_t> z7((size_t)nc * 7, 0u), z8((size_t)nc * 8, 0u);
        if ((rc = up(d_out7, z7))) return rc;
        if ((rc = up(d_pre, z8))) return rc;
    }
    if ((rc = up(d_off, zc))) return rc;
    std::vector<uint32_t> zi((size_t)n_inflow + 1, 0u);
    if ((rc = up(d_ins_off, zi))) return rc;
    if (has_bgk) {
        if ((rc = dalloc(d_mom, (size_t)nc * kNMom))) return rc;
        if ((rc = dalloc(d_bgkp, (size_t)nc))) return rc;
        if ((rc = dalloc(d_raise, (size_t)nc))) return rc;
    }
    cw_host.assign((size_t)nc, 1.0);
    if ((rc = up(d_cw, cw_host))) return rc;
    sclev_host.assign((size_t)nc, 0x010101u);
    nwf = wf_begin.back();
    if (nwf > 0) {
        if ((rc = up(d_wface, wface_all))) return rc;
        if ((rc = up(d_wgeo, wgeo_all))) return rc;
        if ((rc = up(d_wcell, wcell_all))) return rc;
        if ((rc = dalloc(d_wcnt, (size_t)nwf + 1))) return rc;
        if ((rc = dalloc(d_woff, (size_t)nwf + 1))) return rc;
        std::vector<double> zw((size_t)nwf * kNS * kWRows, 0.0);
        if ((rc = up(d_wsamp, zw))) return rc;
    }
    size_t nbs = scan_bsum_words(std::max<size_t>(std::max<size_t>((size_t)nc * kNS + 2, (size_t)nwf + 1),
                                                  (size_t)n_inflow + 1));
    if ((rc = dalloc(d_blocksum, nbs))) return rc;
    std::vector<unsigned long long> zcnt(specs.size() * kCntNum, 0ull);
    if ((rc = up(d_counters, zcnt))) return rc;
    DevState st{};
    std::vector<DevState> vst(1, st);
    if ((rc = up(d_st, vst))) return rc;
    specs_host = specs;
    samp_reset_steps.assign(specs.size(), 0);
    adapt_time.assign(specs.size(), 0.0);
    tsamp.assign(specs.size(), 0.0);
    cap = 0;
    uint64_t want = capacity > 0 ? capacity : 4096;
    if ((rc = alloc_particles((uint32_t)std::min<uint64_t>(want, 0xF0000000ull)))) return rc;
    if (has_hybrid)
        if ((rc = build_neighbourhoods())) return rc;
    return UGX_OK;
}

int World::build_neighbourhoods()
{
    uint32_t* cnt = nullptr;
    uint32_t* total = nullptr;
    int rc;
    if ((rc = dalloc(cnt, (size_t)nc + 1)) || (rc = dalloc(d_nb_off, (size_t)nc + 1)) || (rc = dalloc(total, 1))) {
        dfree(cnt); dfree(total);
        return rc;
    }
    if (precision == UGX_FP32_STATE) launch_nb_count<float>(view<float>(), cnt, d_nb_off, d_blocksum, total, stream);
    else launch_nb_count<double>(view<double>(), cnt, d_nb_off, d_blocksum, total, stream);
    uint32_t n = 0;
    hipError_t e = hipMemcpyAsync(&n, total, sizeof n, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    dfree(cnt);
    dfree(total);
    if (e != hipSuccess) return set_err(UGX_ERR_DEVICE, "neighbourhood count failed: %s", hipGetErrorString(e));
    if ((rc = dalloc(d_nb_list, (size_t)std::max<uint32_t>(n, 1u)))) return rc;
    if (precision == UGX_FP32_STATE) launch_nb_fill<float>(view<float>(), d_nb_off, d_nb_list, stream);
    else launch_nb_fill<double>(view<double>(), d_nb_off, d_nb_list, stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(stream));
    return UGX_OK;
}

int World::check_state()
{
    DevState st{};
    HIP_TRY(hipMemcpyAsync(&st, d_st, sizeof st, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    if (st.overflow) return set_err(UGX_ERR_CAPACITY, "particle capacity (%u) exceeded on the device", cap);
    if (st.evover)
        return set_err(UGX_ERR_CAPACITY, "a step had more wall hits than the %u-event buffer holds", ev_cap);
    // room for twice the busiest step's wall hits seen so far
    const uint32_t peak = std::max(st.nev_peak, st.nev);
    if (nwf > 0 && (uint64_t)peak * 2u > ev_cap) {
        const uint64_t want = std::min<uint64_t>((uint64_t)peak * 4u, 0xF0000000ull);
        if (int rc = alloc_events((uint32_t)want)) return rc;
    }
    return UGX_OK;
}

int World::count(uint32_t& n)
{
    HIP_TRY(hipMemcpyAsync(&n, &d_st->n, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    n_bound = n;
    n_bound_valid = true;
    return UGX_OK;
}

int World::sort_now()
{
    if (precision == UGX_FP32_STATE) launch_sort<float>(view<float>(), stream);
    else launch_sort<double>(view<double>(), stream);
    HIP_TRY(hipGetLastError());
    cur = 1 - cur;
    return UGX_OK;
}

static constexpr int kTimeBatch = 32;

int World::flush_timing()
{
    if (tev_used == 0) return UGX_OK;
    HIP_TRY(hipEventSynchronize(tev[(size_t)tev_used * (kPhases + 1) - 1]));
    for (int s = 0; s < tev_used; ++s) {
        hipEvent_t* m = &tev[(size_t)s * (kPhases + 1)];
        for (int p = 0; p < kPhases; ++p) {
            float ms = 0.0f;
            HIP_TRY(hipEventElapsedTime(&ms, m[p], m[p + 1]));
            phase_ms[p] += (double)ms;
        }
    }
    timed_steps += tev_used;
    tev_used = 0;
    return UGX_OK;
}

int World::step(int n)
{
    for (int i = 0; i < n; ++i) {
        // The capacity grows between steps, before the device could run out:
        // a step adds at most max_ins simulators (its inflow), so while the
        // bound plus one step's inflow fits, no synchronisation is needed.
        // When it may not, the count is read and the buffers are regrown
        // with room for 64 more steps of inflow (graphs are captured again).
        // With cell weighting a step may also add clones: the bound assumes
        // at most n/16 + 1024 per step (more is reported as an overflow).
        const uint64_t step_add = (uint64_t)max_ins + (has_weighting ? n_bound / 16 + 1024 : 0);
        if (!n_bound_valid || (step_add > 0 && n_bound + step_add > cap)) {
            if (int rc = check_state()) return rc;
            uint32_t nn = 0;
            if (int rc = count(nn)) return rc;
            const uint64_t add = (uint64_t)max_ins + (has_weighting ? (uint64_t)nn / 16 + 1024 : 0);
            const uint64_t need = (uint64_t)nn + 64ull * max_ins + (has_weighting ? 4 * add : 0);
            if (add > 0 && need > cap)
                if (int rc = ensure_capacity(need + need / 4)) return rc;
        }
        hipEvent_t* marks = nullptr;
        if (timing) {
            if (tev.empty()) {
                tev.resize((size_t)kTimeBatch * (kPhases + 1));
                for (auto& e : tev) HIP_TRY(hipEventCreate(&e));
            }
            if (tev_used == kTimeBatch)
                if (int rc = flush_timing()) return rc;
            marks = &tev[(size_t)tev_used * (kPhases + 1)];
            tev_used++;
        }
        int kn_due = 0, kn_passes = 0;
        for (const auto& c : cases)
            if (c.active && c.coll_model == UGX_COLL_HYBRID && ((c.step + 1u) % (uint32_t)c.decomposition_interval) == 0u) {
                kn_due = 1;
                kn_passes = std::max(kn_passes, c.smoothing_passes);
            }
        // the tuner times each cell-tile configuration on one of the first
        // kTileVariants steps, with or without step timing
        // and, as the flow develops, it times the conservation kernel's two
        // occupancies on the collide phase of the last kConsVariants steps of
        // every kConsRetune (not on decomposition steps)
        const bool tile_tuning = autotune && tune_step < kTileVariants;
        const uint64_t phase = steps_done % kConsRetune;
        const bool cons_tuning = autotune && !tile_tuning && !kn_due && precision == UGX_FP32_STATE && has_bgk
                                 && !has_macro && phase >= kConsRetune - kConsVariants;
        const bool tuning = tile_tuning || cons_tuning;
        int variant = tile_variant, cvariant = cons_variant;
        if (tuning) {
            if (!marks) {
                if (tune_ev.empty()) {
                    tune_ev.resize(kPhases + 1);
                    for (auto& e : tune_ev) HIP_TRY(hipEventCreate(&e));
                }
                marks = tune_ev.data();
            }
            if (tile_tuning) variant = tune_step;
            if (cons_tuning) cvariant = (int)(phase - (kConsRetune - kConsVariants));
        }
        const bool graph = use_graphs && !marks && !kn_due;
        if (graph && !gexec[cur]) {
            hipGraph_t g = nullptr;
            HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            if (precision == UGX_FP32_STATE) launch_step<float>(view<float>(), stream, nullptr);
            else launch_step<double>(view<double>(), stream, nullptr);
            HIP_TRY(hipStreamEndCapture(stream, &g));
            hipError_t e = hipGraphInstantiate(&gexec[cur], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) { gexec[cur] = nullptr; return set_err(UGX_ERR_DEVICE, "graph instantiate failed: %s", hipGetErrorString(e)); }
        }
        if (graph) {
            HIP_TRY(hipGraphLaunch(gexec[cur], stream));
        } else if (precision == UGX_FP32_STATE) {
            WorldView<float> w = view<float>();
            w.kn_due = kn_due;
            w.kn_passes = kn_passes;
            w.tile_variant = variant;
            w.cons_variant = cvariant;
            launch_step<float>(w, stream, marks);
        } else {
            WorldView<double> w = view<double>();
            w.kn_due = kn_due;
            w.kn_passes = kn_passes;
            w.tile_variant = variant;
            w.cons_variant = cvariant;
            launch_step<double>(w, stream, marks);
        }
        if (tuning) {
            HIP_TRY(hipEventSynchronize(marks[kPhases]));
            float ms = 0.0f;
            if (tile_tuning) {
                HIP_TRY(hipEventElapsedTime(&ms, marks[6], marks[7]));
                tune_ms[variant] += ms;
                if (++tune_step == kTileVariants) {
                    int best = 0;
                    for (int v = 1; v < kTileVariants; ++v) if (tune_ms[v] < tune_ms[best]) best = v;
                    tile_variant = best;
                    drop_graphs();
                }
            } else {
                HIP_TRY(hipEventElapsedTime(&ms, marks[7], marks[8]));
                cons_ms[cvariant] = ms;
                if (cvariant == kConsVariants - 1) {
                    const int best = (cons_ms[1] < cons_ms[0]) ? 1 : 0;
                    if (best != cons_variant) {
                        cons_variant = best;
                        drop_graphs();
                    }
                }
            }
        }
        steps_done += 1;
        cur = 1 - cur;
        for (size_t ci = 0; ci < cases.size(); ++ci) {
            CaseDev& c = cases[ci];
            if (c.active) {
                if (((c.step + 1u) % (uint32_t)c.sample_interval) == 0u) tsamp[ci] = tsamp[ci] + c.dt;
                if (c.adaptive) adapt_time[ci] = adapt_time[ci] + c.dt;
                c.time = c.time + c.dt;
            }
            c.step += 1;
        }
        n_bound += max_ins + (has_weighting ? n_bound / 16 + 1024 : 0);
        // the dynamic adapter at the end of every adaptationInterval-th step
        // (uniGasCloud.C:862, uniGasDynamicAdapter.C:549)
        for (size_t ci = 0; ci < cases.size(); ++ci) {
            const CaseDev& c = cases[ci];
            if (c.active && c.adaptive && (c.step % (uint32_t)specs_host[ci].model.adaptation_interval) == 0u)
                if (int rc = adapt_case((int)ci)) return rc;
        }
    }
    HIP_TRY(hipGetLastError());
    return UGX_OK;
}

// Appends host particles of case `ci` (global positions or local) and sorts.
int World::upload(int ci, const ugx_particle_soa* p, const MeshGeom& g)
{
    uint32_t n0 = 0;
    if (int rc = count(n0)) return rc;
    if (int rc = ensure_capacity((uint64_t)n0 + (uint64_t)p->n)) return rc;
    const size_t n = (size_t)p->n;
    const CaseDev& cd = cases[ci];
    std::vector<double> xyz[7];
    for (auto& v : xyz) v.resize(n);
    std::vector<uint32_t> cell(n), sp(n);
    for (size_t i = 0; i < n; ++i) {
        int c = p->cell[i];
        if (c < 0 || c >= g.nc) return set_err(UGX_ERR_ARG, "particle %zu: cell %d out of range", i, c);
        int s = p->species ? p->species[i] : 0;
        if (s < 0 || s >= cd.ns) return set_err(UGX_ERR_ARG, "particle %zu: species %d out of range", i, s);
        for (int d = 0; d < 3; ++d) {
            xyz[d][i] = p->local_position ? p->local_position[3 * i + d] : p->position[3 * i + d] - g.cc[3 * (size_t)c + d];
            xyz[3 + d][i] = p->velocity[3 * i + d];
        }
        xyz[6][i] = p->erot ? p->erot[i] : 0.0;
        cell[i] = (uint32_t)(cd.cell_begin + c);
        sp[i] = (uint32_t)(cd.sp_base + s);
    }
    size_t rs = rsize();
    for (int a = 0; a < 7; ++a) {
        if (precision == UGX_FP32_STATE) {
            std::vector<float> f(n);
            for (size_t i = 0; i < n; ++i) f[i] = (float)xyz[a][i];
            HIP_TRY(hipMemcpy((char*)pbuf[cur][a] + (size_t)n0 * rs, f.data(), n * rs, hipMemcpyHostToDevice));
        } else {
            HIP_TRY(hipMemcpy((char*)pbuf[cur][a] + (size_t)n0 * rs, xyz[a].data(), n * rs, hipMemcpyHostToDevice));
        }
    }
    HIP_TRY(hipMemcpy((char*)pbuf[cur][7] + (size_t)n0 * 4, cell.data(), n * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy((char*)pbuf[cur][8] + (size_t)n0 * 4, sp.data(), n * 4, hipMemcpyHostToDevice));
    uint32_t ntot = n0 + (uint32_t)n;
    HIP_TRY(hipMemcpy(&d_st->n, &ntot, sizeof(uint32_t), hipMemcpyHostToDevice));
    uint32_t zero = 0;
    HIP_TRY(hipMemcpy(&d_st->n_ins, &zero, sizeof(uint32_t), hipMemcpyHostToDevice));
    n_bound_valid = false;
    if (int rc = sort_now()) return rc;
    HIP_TRY(hipStreamSynchronize(stream));
    return check_state();
}

int World::fill(int ci, const ugx_fill_desc* f, const MeshGeom& g, bool sort, const FillScratch* fixed)
{
    const CaseDev& cd = cases[ci];
    uint32_t n0 = 0;
    if (int rc = count(n0)) return rc;
    double ntot = 0.0, expected = 0.0;
    for (int s = 0; s < cd.ns; ++s) ntot += f->number_density[s];
    // an adaptive case starts its time step and subcell levels from the
    // filled gas (uniGasMeshFill.C:100-109)
    if (cd.adaptive) {
        if (fixed) return set_err(UGX_ERR_ARG, "an adaptive case is filled with allocation (ugx_fill)");
        if (int rc = adapt_fill(ci, f)) return rc;
    }
    // a cell-weighted case takes its cell weights from this fill (uniGasMeshFill.C:111-168)
    if (cd.weighted) {
        if (f->exact_per_cell > 0) return set_err(UGX_ERR_ARG, "a cell-weighted case is filled by number density, not per cell");
        std::vector<double> wts;
        if (int rc = fill_cell_weights(specs_host[ci], ntot, sclev_host.data() + cd.cell_begin, wts)) return rc;
        if (int rc = set_cell_weights(ci, wts.data())) return rc;
    }
    // capacity from the expected count
    if (f->exact_per_cell > 0) expected = (double)f->exact_per_cell * g.nc + (double)g.nc * cd.ns;
    else {
        double VW = 0.0;
        for (int c = 0; c < g.nc; ++c) VW += g.vol[c] / cw_host[(size_t)cd.cell_begin + c];
        expected = ntot / cd.fn * VW * 1.05 + (double)g.nc * cd.ns + 64.0;
    }
    if (fixed) {
        // never grown here: the fill must fit (checked again exactly after the count)
        if ((double)n0 + expected > (double)cap)
            return set_err(UGX_ERR_CAPACITY, "the fill may need %.0f simulators beyond %u; the case holds %u "
                                             "(set ugx_options.particle_capacity)", expected, n0, cap);
    } else if (int rc = ensure_capacity((uint64_t)n0 + (uint64_t)expected)) {
        return rc;
    }
    const int ns = cd.ns, ncl = g.nc;
    const double* d_lo = d_bblo + (size_t)cd.cell_begin * 3;
    const double* d_hi = d_bbhi + (size_t)cd.cell_begin * 3;
    double* d_par = nullptr;
    uint32_t* d_fc = nullptr;
    uint32_t* d_fo = nullptr;
    auto cleanup = [&]() { if (!fixed) { dfree(d_par); dfree(d_fc); dfree(d_fo); } };
    FillParams fp{};
    for (int s = 0; s < kNS; ++s) fp.nd[s] = f->number_density[s];
    for (int d = 0; d < 3; ++d) { fp.T[d] = f->T[d]; fp.U[d] = f->velocity[d]; }
    if (fixed) {
        d_par = fixed->params;
        d_fc = fixed->fcnt;
        d_fo = fixed->foff;
    } else {
        int rc = UGX_OK;
        if ((rc = dalloc(d_par, kFillParams)) || (rc = dalloc(d_fc, (size_t)ncl * ns + 1)) || (rc = dalloc(d_fo, (size_t)ncl * ns + 1))) {
            cleanup();
            return rc;
        }
    }
    // the parameters go to the device through a kernel argument, not a copy
    launch_fill_params(fp, d_par, stream);
    const double* d_nd = d_par;
    const double* d_T = d_par + kNS;
    const double* d_U = d_par + kNS + 3;
    if (precision == UGX_FP32_STATE)
        launch_fill<float>(view<float>(), stream, ci, d_nd, d_T, f->rot_temperature, d_U, f->exact_per_cell, d_lo, d_hi, d_fc, d_fo, ns, ncl);
    else
        launch_fill<double>(view<double>(), stream, ci, d_nd, d_T, f->rot_temperature, d_U, f->exact_per_cell, d_lo, d_hi, d_fc, d_fo, ns, ncl);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    cleanup();
    if (e != hipSuccess) return set_err(UGX_ERR_DEVICE, "fill failed: %s", hipGetErrorString(e));
    if (int r2 = check_state()) return r2;
    if (sort) {
        if (int r2 = sort_now()) return r2;
        HIP_TRY(hipStreamSynchronize(stream));
    }
    // (sigma_T c_r)max = pi d^2 c_mp(T) of the most abundant species (uniGasMeshFill.C:284-296)
    int best = 0;
    for (int s = 0; s < ns; ++s) if (f->number_density[s] > f->number_density[best]) best = s;
    const ugx_species_desc& sb = specs_host[ci].species[best];
    double Tm = (f->T[0] + f->T[1] + f->T[2]) / 3.0;
    double s0 = kPi * sb.diameter * sb.diameter * std::sqrt(2.0 * kB * Tm / sb.mass);
    launch_set_f64(d_smax + cd.cell_begin, s0, (uint32_t)ncl, stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(stream));
    uint32_t n1 = 0;
    if (int r2 = count(n1)) return r2;
    if (fixed) return UGX_OK;
    // room for the filled simulators and 16 steps of inflow; ugx_step grows
    // the buffers further when the inflow needs it
    return ensure_capacity((uint64_t)n1 + (uint64_t)max_ins * 16);
}

int World::set_cell_weights(int ci, const double* wts)
{
    const CaseDev& cd = cases[ci];
    const int ncl = cd.cell_end - cd.cell_begin;
    for (int c = 0; c < ncl; ++c)
        if (!(wts[c] > 0.0) || !std::isfinite(wts[c])) return set_err(UGX_ERR_ARG, "cell weight %d is not > 0", c);
    HIP_TRY(hipStreamSynchronize(stream));
    std::copy(wts, wts + ncl, cw_host.begin() + cd.cell_begin);
    HIP_TRY(hipMemcpy(d_cw + cd.cell_begin, wts, sizeof(double) * ncl, hipMemcpyHostToDevice));
    update_max_ins();
    return UGX_OK;
}

int World::set_time_step(int ci, double dt)
{
    if (!(dt > 0.0) || !std::isfinite(dt)) return set_err(UGX_ERR_STATE, "adapted time step %g is not > 0", dt);
    HIP_TRY(hipStreamSynchronize(stream));
    CaseDev& cd = cases[ci];
    cd.dt = dt;
    HIP_TRY(hipMemcpy((char*)(d_cases + ci) + offsetof(CaseDev, dt), &dt, sizeof(double), hipMemcpyHostToDevice));
    // the freestream counts are per time step (uniGasGeneralBoundary.C:154-165)
    MeshGeom& g = specs_host[ci].geom;
    for (size_t k = 0; k < g.inflow.size(); ++k) {
        InflowFS& e = g.inflow[k];
        const double* pt = &g.inflow_parts[3 * k];
        e.nbar_num = pt[0] * dt * pt[1] * pt[2];
        e.nbar = e.nbar_num / e.nbar_den;
        InflowFS d = e;
        d.cell += cd.cell_begin;
        d.species += cd.sp_base;
        d.ccase = ci;
        HIP_TRY(hipMemcpy(d_inflow + cd.inflow_begin + (int)k, &d, sizeof(InflowFS), hipMemcpyHostToDevice));
    }
    update_max_ins();
    return UGX_OK;
}

int World::apply_adaptation(int ci, const AdaptOut& a)
{
    const CaseDev& cd = cases[ci];
    const int ncl = cd.cell_end - cd.cell_begin;
    std::copy(a.ratio.begin(), a.ratio.end(), adapt_ratio.begin() + 3 * (size_t)cd.cell_begin);
    if (a.dt != cd.dt)
        if (int rc = set_time_step(ci, a.dt)) return rc;
    if (!a.levels.empty()) {
        bool changed = false;
        for (int c = 0; c < ncl; ++c)
            if (sclev_host[(size_t)cd.cell_begin + c] != a.levels[c]) {
                sclev_host[(size_t)cd.cell_begin + c] = a.levels[c];
                changed = true;
            }
        if (changed)
            if (int rc = upload_subcells()) return rc;
    }
    return UGX_OK;
}

int World::reweight(int ci, const std::vector<double>& old_w)
{
    const CaseDev& cd = cases[ci];
    const int ncl = cd.cell_end - cd.cell_begin;
    // room for the clones: each simulator of cell c gets at most ceil(W_o / W_n) - 1
    std::vector<uint32_t> off((size_t)nc + 1);
    HIP_TRY(hipStreamSynchronize(stream));
    HIP_TRY(hipMemcpy(off.data(), d_off, sizeof(uint32_t) * off.size(), hipMemcpyDeviceToHost));
    uint32_t n = 0;
    if (int rc = count(n)) return rc;
    double extra = 0.0;
    for (int c = 0; c < ncl; ++c) {
        const double r = old_w[c] / cw_host[(size_t)cd.cell_begin + c];
        if (r > 1.0) extra += std::ceil(r - 1.0) * (double)(off[(size_t)cd.cell_begin + c + 1] - off[(size_t)cd.cell_begin + c]);
    }
    if (int rc = ensure_capacity((uint64_t)n + (uint64_t)extra + 1024u)) return rc;
    if (!d_cw_old)
        if (int rc = dalloc(d_cw_old, (size_t)nc)) return rc;
    std::vector<double> full(cw_host);
    std::copy(old_w.begin(), old_w.end(), full.begin() + cd.cell_begin);
    HIP_TRY(hipMemcpy(d_cw_old, full.data(), sizeof(double) * (size_t)nc, hipMemcpyHostToDevice));
    if (precision == UGX_FP32_STATE) launch_reweight<float>(view<float>(), d_cw_old, ci, stream);
    else launch_reweight<double>(view<double>(), d_cw_old, ci, stream);
    HIP_TRY(hipGetLastError());
    if (int rc = sort_now()) return rc;
    if (int rc = check_state()) return rc;
    n_bound_valid = false;
    return UGX_OK;
}

int World::adapt_fill(int ci, const ugx_fill_desc* f)
{
    const CaseDev& cd = cases[ci];
    const int ncl = cd.cell_end - cd.cell_begin;
    // the filled gas; the adapter takes one translational temperature
    const double T = (f->T[0] + f->T[1] + f->T[2]) / 3.0;
    std::vector<int32_t> cm((size_t)ncl);
    HIP_TRY(hipMemcpy(cm.data(), d_cmodel + cd.cell_begin, sizeof(int32_t) * ncl, hipMemcpyDeviceToHost));
    AdaptOut a;
    if (int rc = adapt_initial(specs_host[ci], f->number_density, T, f->velocity, cm.data(), a)) return rc;
    return apply_adaptation(ci, a);
}

int World::adapt_case(int ci)
{
    const CaseDev& cd = cases[ci];
    const int ncl = cd.cell_end - cd.cell_begin;
    HIP_TRY(hipStreamSynchronize(stream));
    std::vector<double> acc((size_t)ncl * kKnRows);
    std::vector<int32_t> cm((size_t)ncl);
    HIP_TRY(hipMemcpy(acc.data(), d_adacc + (size_t)cd.cell_begin * kKnRows, sizeof(double) * acc.size(),
                      hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(cm.data(), d_cmodel + cd.cell_begin, sizeof(int32_t) * ncl, hipMemcpyDeviceToHost));
    AdaptOut a;
    if (int rc = adapt_step(specs_host[ci], acc.data(), adapt_time[ci], cw_host.data() + cd.cell_begin, cm.data(),
                            cd.dt, adapt_ratio.data() + 3 * (size_t)cd.cell_begin, sclev_host.data() + cd.cell_begin, a))
        return rc;
    if (int rc = apply_adaptation(ci, a)) return rc;
    if (!a.cw.empty()) {
        // the new weights, then every simulator from its cell's old weight (uniGasCloud.C:228-232)
        std::vector<double> old_w(cw_host.begin() + cd.cell_begin, cw_host.begin() + cd.cell_end);
        if (int rc = set_cell_weights(ci, a.cw.data())) return rc;
        if (int rc = reweight(ci, old_w)) return rc;
    }
    // the sums start again (uniGasDynamicAdapter.C:676-694)
    std::fill(acc.begin(), acc.end(), 0.0);
    HIP_TRY(hipMemcpy(d_adacc + (size_t)cd.cell_begin * kKnRows, acc.data(), sizeof(double) * acc.size(),
                      hipMemcpyHostToDevice));
    adapt_time[ci] = 0.0;
    return UGX_OK;
}

int World::set_subcell_levels(int ci, const int32_t* lv)
{
    const CaseDev& cd = cases[ci];
    const int ncl = cd.cell_end - cd.cell_begin;
    const double* mask = specs_host[ci].geom.mask;
    for (int c = 0; c < ncl; ++c)
        for (int d = 0; d < 3; ++d) {
            const int32_t l = lv[3 * (size_t)c + d];
            if (l < 1 || l > kMaxSubcellLevels)
                return set_err(UGX_ERR_ARG, "subcell level %d of cell %d is outside 1..%d", l, c, kMaxSubcellLevels);
            if (mask[d] == 0.0 && l != 1) return set_err(UGX_ERR_ARG, "subcell level > 1 in the empty direction %d", d);
        }
    for (int c = 0; c < ncl; ++c)
        sclev_host[(size_t)cd.cell_begin + c] = (uint32_t)lv[3 * (size_t)c] | ((uint32_t)lv[3 * (size_t)c + 1] << 8) |
                                                ((uint32_t)lv[3 * (size_t)c + 2] << 16);
    return upload_subcells();
}

int World::upload_subcells()
{
    HIP_TRY(hipStreamSynchronize(stream));
    drop_graphs();
    std::vector<uint32_t> off((size_t)nc);
    uint64_t tot = 0;
    int any = 0;
    for (int c = 0; c < nc; ++c) {
        const uint32_t p = sclev_host[(size_t)c];
        const uint32_t nsub = (p & 0xffu) * ((p >> 8) & 0xffu) * ((p >> 16) & 0xffu);
        off[(size_t)c] = (uint32_t)tot;
        if (nsub > 1u) { any = 1; tot += 2ull * nsub + 1ull; }
    }
    if (tot > 0xF0000000ull) return set_err(UGX_ERR_CAPACITY, "subcell tables too large");
    dfree(d_sctab);
    if (!any) {
        // no cell is split: the tables are released and the NTC takes the whole cell
        dfree(d_sclev); dfree(d_scoff); dfree(d_scid); dfree(d_sclist);
        has_subcells = 0;
        return UGX_OK;
    }
    if (!d_sclev) {
        if (int rc = dalloc(d_sclev, (size_t)nc)) return rc;
        if (int rc = dalloc(d_scoff, (size_t)nc)) return rc;
    }
    if (int rc = dalloc(d_sctab, (size_t)tot)) return rc;
    HIP_TRY(hipMemcpy(d_sclev, sclev_host.data(), sizeof(uint32_t) * (size_t)nc, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d_scoff, off.data(), sizeof(uint32_t) * (size_t)nc, hipMemcpyHostToDevice));
    if (!has_subcells) {
        has_subcells = 1;
        if (cap > 0) {
            if (int rc = dalloc(d_scid, cap)) return rc;
            if (int rc = dalloc(d_sclist, cap)) return rc;
        }
    }
    return UGX_OK;
}

// Bound of one step's inflow: floor + 1 of every face-species' expected count
// at its cell's weight.
void World::update_max_ins()
{
    max_ins = 0;
    for (size_t ci = 0; ci < specs_host.size(); ++ci)
        for (const auto& e : specs_host[ci].geom.inflow) {
            const double nb = e.nbar_num / (e.nbar_den * cw_host[(size_t)cases[ci].cell_begin + e.cell]);
            max_ins += (uint32_t)std::floor(nb) + 1u;
        }
    n_bound_valid = false;
}

int World::pause(int ci)
{
    static const int32_t zero = 0;
    cases[ci].active = 0;
    HIP_TRY(hipMemcpyAsync((char*)(d_cases + ci) + offsetof(CaseDev, active), &zero, sizeof(int32_t),
                           hipMemcpyHostToDevice, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return UGX_OK;
}

CellState World::cell_state()
{
    return CellState{d_smax, d_cmodel, d_samp, d_uspq, d_usps, d_uspmax, d_knacc, d_kn, nc};
}

int World::adopt(World& old, const std::vector<int>& old_of_new)
{
    HIP_TRY(hipStreamSynchronize(old.stream));
    std::vector<uint32_t> ooff((size_t)old.nc + 1);
    HIP_TRY(hipMemcpy(ooff.data(), old.d_off, sizeof(uint32_t) * ooff.size(), hipMemcpyDeviceToHost));
    std::vector<CaseDev> ocd(old.cases.size());
    HIP_TRY(hipMemcpy(ocd.data(), old.d_cases, sizeof(CaseDev) * ocd.size(), hipMemcpyDeviceToHost));
    uint64_t total = 0;
    for (size_t i = 0; i < old_of_new.size(); ++i) {
        const int k = old_of_new[i];
        if (k >= 0) total += ooff[old.cases[k].cell_end] - ooff[old.cases[k].cell_begin];
    }
    if (int rc = ensure_capacity(total + total / 4 + 1024)) return rc;
    const CellState os = old.cell_state(), ns_ = cell_state();
    uint32_t dst = 0;
    for (size_t i = 0; i < old_of_new.size(); ++i) {
        const int k = old_of_new[i];
        if (k < 0) continue;
        const CaseDev& oc = old.cases[k];
        CaseDev& nc_ = cases[i];
        const uint32_t ncl = (uint32_t)(oc.cell_end - oc.cell_begin);
        launch_carry_cells(os, ns_, (uint32_t)oc.cell_begin, (uint32_t)nc_.cell_begin, ncl, stream);
        HIP_TRY(hipMemcpyAsync(d_counters + (size_t)i * kCntNum, old.d_counters + (size_t)k * kCntNum,
                               sizeof(unsigned long long) * kCntNum, hipMemcpyDeviceToDevice, stream));
        const uint32_t b = ooff[oc.cell_begin], n = ooff[oc.cell_end] - b;
        const int32_t cd = nc_.cell_begin - oc.cell_begin, sd = nc_.sp_base - oc.sp_base;
        if (precision == UGX_FP32_STATE)
            launch_carry_particles<float>(old.view<float>().a, view<float>().a, b, n, dst, cd, sd, stream);
        else
            launch_carry_particles<double>(old.view<double>().a, view<double>().a, b, n, dst, cd, sd, stream);
        dst += n;
        std::copy(old.cw_host.begin() + oc.cell_begin, old.cw_host.begin() + oc.cell_end,
                  cw_host.begin() + nc_.cell_begin);
        HIP_TRY(hipMemcpyAsync(d_cw + nc_.cell_begin, old.d_cw + oc.cell_begin, sizeof(double) * ncl,
                               hipMemcpyDeviceToDevice, stream));
        const int32_t nwk = old.wf_begin[k + 1] - old.wf_begin[k];
        if (nwk > 0)
            HIP_TRY(hipMemcpyAsync(d_wsamp + (size_t)wf_begin[i] * kNS * kWRows,
                                   old.d_wsamp + (size_t)old.wf_begin[k] * kNS * kWRows,
                                   sizeof(double) * (size_t)nwk * kNS * kWRows, hipMemcpyDeviceToDevice, stream));
        nc_.step = ocd[k].step;
        nc_.kn_time = ocd[k].kn_time;
        nc_.time = ocd[k].time;
        nc_.dt = ocd[k].dt;
        samp_reset_steps[i] = old.samp_reset_steps[k];
        tsamp[i] = old.tsamp[k];
        adapt_time[i] = old.adapt_time[k];
        std::copy(old.adapt_ratio.begin() + 3 * (size_t)oc.cell_begin, old.adapt_ratio.begin() + 3 * (size_t)oc.cell_end,
                  adapt_ratio.begin() + 3 * (size_t)nc_.cell_begin);
        if (d_adacc && old.d_adacc)
            HIP_TRY(hipMemcpyAsync(d_adacc + (size_t)nc_.cell_begin * kKnRows, old.d_adacc + (size_t)oc.cell_begin * kKnRows,
                                   sizeof(double) * (size_t)ncl * kKnRows, hipMemcpyDeviceToDevice, stream));
        if (ocd[k].dt != specs_host[i].model.delta_t) {
            // an adapted time step: the inflow counts follow it
            MeshGeom& gi = specs_host[i].geom;
            gi.inflow = old.specs_host[k].geom.inflow;
            for (size_t q = 0; q < gi.inflow.size(); ++q) {
                InflowFS d = gi.inflow[q];
                d.cell += nc_.cell_begin;
                d.species += nc_.sp_base;
                d.ccase = (int32_t)i;
                HIP_TRY(hipMemcpyAsync(d_inflow + nc_.inflow_begin + (int)q, &d, sizeof(InflowFS), hipMemcpyHostToDevice, stream));
                HIP_TRY(hipStreamSynchronize(stream));
            }
        }
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(d_cases, cases.data(), sizeof(CaseDev) * cases.size(), hipMemcpyHostToDevice, stream));
    HIP_TRY(hipMemcpyAsync(&d_st->n, &dst, sizeof(uint32_t), hipMemcpyHostToDevice, stream));
    update_max_ins();
    HIP_TRY(hipStreamSynchronize(stream));
    bool split = false;
    for (size_t i = 0; i < old_of_new.size(); ++i) {
        const int k = old_of_new[i];
        if (k < 0) continue;
        const CaseDev &oc = old.cases[k], &nc_ = cases[i];
        std::copy(old.sclev_host.begin() + oc.cell_begin, old.sclev_host.begin() + oc.cell_end,
                  sclev_host.begin() + nc_.cell_begin);
        split = split || old.has_subcells;
    }
    if (split)
        if (int rc = upload_subcells()) return rc;
    return UGX_OK;
}

int World::download(int ci, ugx_particle_soa* p, const MeshGeom& g)
{
    const CaseDev& cd = cases[ci];
    uint32_t n = 0;
    if (int rc = count(n)) return rc;
    std::vector<uint32_t> off((size_t)nc + 1);
    HIP_TRY(hipMemcpy(off.data(), d_off, sizeof(uint32_t) * off.size(), hipMemcpyDeviceToHost));
    uint32_t b = off[cd.cell_begin], e = off[cd.cell_end];
    size_t m = e - b;
    if ((size_t)p->n < m) { p->n = (int64_t)m; return set_err(UGX_ERR_CAPACITY, "download arrays hold %lld, need %zu", (long long)p->n, m); }
    size_t rs = rsize();
    std::vector<double> arr[7];
    for (int a = 0; a < 7; ++a) {
        arr[a].resize(m);
        if (precision == UGX_FP32_STATE) {
            std::vector<float> f(m);
            if (m) HIP_TRY(hipMemcpy(f.data(), (char*)pbuf[cur][a] + (size_t)b * rs, m * rs, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < m; ++i) arr[a][i] = (double)f[i];
        } else if (m) {
            HIP_TRY(hipMemcpy(arr[a].data(), (char*)pbuf[cur][a] + (size_t)b * rs, m * rs, hipMemcpyDeviceToHost));
        }
    }
    std::vector<uint32_t> cell(m), sp(m);
    if (m) {
        HIP_TRY(hipMemcpy(cell.data(), (char*)pbuf[cur][7] + (size_t)b * 4, m * 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(sp.data(), (char*)pbuf[cur][8] + (size_t)b * 4, m * 4, hipMemcpyDeviceToHost));
    }
    for (size_t i = 0; i < m; ++i) {
        int c = (int)cell[i] - cd.cell_begin;
        for (int d = 0; d < 3; ++d) {
            if (p->local_position) p->local_position[3 * i + d] = arr[d][i];
            if (p->position) p->position[3 * i + d] = arr[d][i] + g.cc[3 * (size_t)c + d];
            if (p->velocity) p->velocity[3 * i + d] = arr[3 + d][i];
        }
        if (p->erot) p->erot[i] = arr[6][i];
        if (p->cell) p->cell[i] = c;
        if (p->species) p->species[i] = (int)sp[i] - cd.sp_base;
    }
    p->n = (int64_t)m;
    return UGX_OK;
}

int World::case_counters(int ci, ugx_counters* out)
{
    std::vector<unsigned long long> cnt(kCntNum);
    HIP_TRY(hipMemcpyAsync(cnt.data(), d_counters + (size_t)ci * kCntNum, sizeof(unsigned long long) * kCntNum,
                           hipMemcpyDeviceToHost, stream));
    std::vector<uint32_t> off(2);
    HIP_TRY(hipMemcpyAsync(&off[0], d_off + cases[ci].cell_begin, 4, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipMemcpyAsync(&off[1], d_off + cases[ci].cell_end, 4, hipMemcpyDeviceToHost, stream));
    std::vector<int32_t> cm((size_t)(cases[ci].cell_end - cases[ci].cell_begin));
    HIP_TRY(hipMemcpyAsync(cm.data(), d_cmodel + cases[ci].cell_begin, cm.size() * 4, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    std::memset(out, 0, sizeof *out);
    out->n_particles = off[1] - off[0];
    out->steps = cases[ci].step;
    out->candidates = (int64_t)cnt[kCntCandidates];
    out->collisions = (int64_t)cnt[kCntCollisions];
    out->bgk_collisions = (int64_t)cnt[kCntBgk];
    out->inserted = (int64_t)cnt[kCntInserted];
    out->deleted = (int64_t)cnt[kCntDeleted];
    out->wall_hits = (int64_t)cnt[kCntWallHits];
    out->track_errors = (int64_t)cnt[kCntTrackErrors];
    out->n_sampled_steps = (int64_t)cnt[kCntSampled] - samp_reset_steps[ci];
    out->particle_steps = (int64_t)cnt[kCntParticleSteps];
    out->cloned = (int64_t)cnt[kCntCloned];
    out->weight_deleted = (int64_t)cnt[kCntWeightDeleted];
    out->time_sampled = tsamp[ci];
    int64_t k = 0;
    for (int32_t v : cm) k += v;
    out->n_dsmc_cells = k;
    return UGX_OK;
}

int World::case_samples(int ci, double* out, int64_t len)
{
    const int ncl = cases[ci].cell_end - cases[ci].cell_begin;
    if (len < (int64_t)kNSamp * ncl) return set_err(UGX_ERR_ARG, "sample buffer holds %lld, need %d", (long long)len, kNSamp * ncl);
    // device layout [cell][row] -> ABI layout [row][cell]
    std::vector<double> cm((size_t)ncl * kNSamp);
    HIP_TRY(hipMemcpyAsync(cm.data(), d_samp + (size_t)cases[ci].cell_begin * kNSamp, sizeof(double) * cm.size(),
                           hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    for (int c = 0; c < ncl; ++c)
        for (int k = 0; k < kNSamp; ++k) out[(size_t)k * ncl + c] = cm[(size_t)c * kNSamp + k];
    return UGX_OK;
}
int World::case_state(int ci, ugx_case_state* s, bool upload)
{
    static_assert(kCntNum <= UGX_N_COUNTERS, "counter slots");
    CaseDev& cd = cases[ci];
    const int ncl = cd.cell_end - cd.cell_begin;
    if (s->n_cells != ncl) return set_err(UGX_ERR_ARG, "state holds %lld cells, the case has %d", (long long)s->n_cells, ncl);
    if (!s->sigma_max || !s->cell_model || !s->samples || !s->usp_heat_flux || !s->usp_stress || !s->usp_max_prob ||
        !s->kn_sums || !s->kn)
        return set_err(UGX_ERR_ARG, "every state array is required");
    HIP_TRY(hipStreamSynchronize(stream));
    const size_t b = (size_t)cd.cell_begin, n = (size_t)ncl;
    struct Arr { void* host; void* dev; size_t bytes; };
    const Arr arrs[] = {
        {s->sigma_max, d_smax + b, sizeof(double) * n},
        {s->cell_model, d_cmodel + b, sizeof(int32_t) * n},
        {s->samples, d_samp + b * kNSamp, sizeof(double) * n * kNSamp},
        {s->usp_heat_flux, d_uspq + b * 3, sizeof(double) * n * 3},
        {s->usp_stress, d_usps + b * 6, sizeof(double) * n * 6},
        {s->usp_max_prob, d_uspmax + b, sizeof(double) * n},
        {s->kn_sums, d_knacc + b * kKnRows, sizeof(double) * n * kKnRows},
        {s->kn, d_kn + b, sizeof(double) * n},
    };
    const int64_t nwk = wf_begin[ci + 1] - wf_begin[ci];
    if (s->wall_samples && s->n_wall_faces != nwk)
        return set_err(UGX_ERR_ARG, "state holds %lld wall faces, the case has %lld", (long long)s->n_wall_faces,
                       (long long)nwk);
    if (s->wall_samples)
        if (int rc = case_wall_samples(ci, s->wall_samples, nwk * kNS * kWRows, upload)) return rc;
    if (s->cell_weight) {
        if (upload) {
            if (int rc = set_cell_weights(ci, s->cell_weight)) return rc;
        } else {
            std::copy(cw_host.begin() + cd.cell_begin, cw_host.begin() + cd.cell_end, s->cell_weight);
        }
    }
    unsigned long long cnt[kCntNum];
    if (!upload) {
        for (const Arr& a : arrs) HIP_TRY(hipMemcpy(a.host, a.dev, a.bytes, hipMemcpyDeviceToHost));
        // step and hybrid time advance on the device
        CaseDev dc;
        HIP_TRY(hipMemcpy(&dc, d_cases + ci, sizeof(CaseDev), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(cnt, d_counters + (size_t)ci * kCntNum, sizeof cnt, hipMemcpyDeviceToHost));
        s->steps = (int64_t)dc.step;
        s->kn_time = dc.kn_time;
        s->sampled_at_reset = samp_reset_steps[ci];
        for (int k = 0; k < UGX_N_COUNTERS; ++k) s->counters[k] = (k < kCntNum) ? (uint64_t)cnt[k] : 0u;
        s->delta_t = dc.dt;
        s->time = dc.time;
        s->time_sampled = tsamp[ci];
        s->adapt_time = adapt_time[ci];
        if (s->subcell_levels)
            for (size_t c = 0; c < n; ++c) {
                const uint32_t p = sclev_host[b + c];
                for (int d = 0; d < 3; ++d) s->subcell_levels[3 * c + d] = (int32_t)((p >> (8 * d)) & 0xffu);
            }
        if (s->adapt_sums) {
            if (d_adacc) HIP_TRY(hipMemcpy(s->adapt_sums, d_adacc + b * kKnRows, sizeof(double) * n * kKnRows,
                                           hipMemcpyDeviceToHost));
            else std::fill(s->adapt_sums, s->adapt_sums + n * kKnRows, 0.0);
        }
        if (s->adapt_ratio) std::copy(adapt_ratio.begin() + 3 * b, adapt_ratio.begin() + 3 * (b + n), s->adapt_ratio);
        return UGX_OK;
    }
    if (s->steps < 0 || s->steps > (int64_t)UINT32_MAX) return set_err(UGX_ERR_ARG, "step count out of range");
    if (s->delta_t > 0.0 && s->delta_t != cd.dt)
        if (int rc = set_time_step(ci, s->delta_t)) return rc;
    if (s->subcell_levels)
        if (int rc = set_subcell_levels(ci, s->subcell_levels)) return rc;
    if (s->adapt_sums && d_adacc)
        HIP_TRY(hipMemcpy(d_adacc + b * kKnRows, s->adapt_sums, sizeof(double) * n * kKnRows, hipMemcpyHostToDevice));
    if (s->adapt_ratio) std::copy(s->adapt_ratio, s->adapt_ratio + 3 * n, adapt_ratio.begin() + 3 * b);
    // a state without the sampled time (NaN) takes the sampled steps' count times the step
    tsamp[ci] = std::isfinite(s->time_sampled)
                    ? s->time_sampled
                    : (double)((int64_t)s->counters[kCntSampled] - s->sampled_at_reset) * cases[ci].dt;
    adapt_time[ci] = s->adapt_time;
    for (const Arr& a : arrs) HIP_TRY(hipMemcpy(a.dev, a.host, a.bytes, hipMemcpyHostToDevice));
    for (int k = 0; k < kCntNum; ++k) cnt[k] = (unsigned long long)s->counters[k];
    HIP_TRY(hipMemcpy(d_counters + (size_t)ci * kCntNum, cnt, sizeof cnt, hipMemcpyHostToDevice));
    // the host copy decides the decomposition steps; the device copy draws
    CaseDev dc;
    HIP_TRY(hipMemcpy(&dc, d_cases + ci, sizeof(CaseDev), hipMemcpyDeviceToHost));
    dc.step = (uint32_t)s->steps;
    dc.kn_time = s->kn_time;
    dc.time = s->time;
    cd.step = dc.step;
    cd.kn_time = dc.kn_time;
    cd.time = dc.time;
    HIP_TRY(hipMemcpy(d_cases + ci, &dc, sizeof(CaseDev), hipMemcpyHostToDevice));
    samp_reset_steps[ci] = s->sampled_at_reset;
    return UGX_OK;
}

int World::case_samples_reset(int ci)
{
    HIP_TRY(hipStreamSynchronize(stream));
    const int ncl = cases[ci].cell_end - cases[ci].cell_begin;
    std::vector<double> z((size_t)ncl * kNSamp, 0.0);
    HIP_TRY(hipMemcpy(d_samp + (size_t)cases[ci].cell_begin * kNSamp, z.data(), sizeof(double) * z.size(), hipMemcpyHostToDevice));
    const size_t nw = (size_t)(wf_begin[ci + 1] - wf_begin[ci]) * kNS * kWRows;
    if (nw > 0) {
        std::vector<double> zw(nw, 0.0);
        HIP_TRY(hipMemcpy(d_wsamp + (size_t)wf_begin[ci] * kNS * kWRows, zw.data(), sizeof(double) * nw,
                          hipMemcpyHostToDevice));
    }
    unsigned long long v = 0;
    HIP_TRY(hipMemcpy(&v, d_counters + (size_t)ci * kCntNum + kCntSampled, sizeof v, hipMemcpyDeviceToHost));
    samp_reset_steps[ci] = (int64_t)v;
    tsamp[ci] = 0.0;
    return UGX_OK;
}

int World::case_wall_samples(int ci, double* buf, int64_t len, bool upload)
{
    const size_t nw = (size_t)(wf_begin[ci + 1] - wf_begin[ci]) * kNS * kWRows;
    if (len < (int64_t)nw) return set_err(UGX_ERR_ARG, "wall sum buffer holds %lld, need %zu", (long long)len, nw);
    if (nw == 0) return UGX_OK;
    HIP_TRY(hipStreamSynchronize(stream));
    double* dev = d_wsamp + (size_t)wf_begin[ci] * kNS * kWRows;
    if (upload) HIP_TRY(hipMemcpy(dev, buf, sizeof(double) * nw, hipMemcpyHostToDevice));
    else HIP_TRY(hipMemcpy(buf, dev, sizeof(double) * nw, hipMemcpyDeviceToHost));
    return UGX_OK;
}

}  // namespace ugx
