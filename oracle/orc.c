/*
 * orc.c - CPU oracle of the uniGasFoam particle time step.
 * TEST INFRASTRUCTURE ONLY (checker for the HIP path; CPU baseline).
 *
 * Restates, in FP64 C with OpenMP, the functions of the reference on the
 * path uniGasCloud::evolve (src/lagrangian/uniGas/clouds/uniGasCloud.C:821-869).
 * Paths below are relative to src/lagrangian/uniGas/ unless stated.
 *
 * Conventions shared with the HIP path (they define bit-exactness):
 *  - positions are stored relative to a per-cell reference point (the mean
 *    of the cell's vertices); a face crossing adds the frame shift
 *    c_from - c_to (+ the cyclic translation);
 *  - cells are convex with <= 6 planar faces, stored as outward planes
 *    n.x = d in the cell frame (faces in increasing face id);
 *  - every random draw comes from a Philox stream keyed by (seed, kind,
 *    entity, step) - see orc_prim.h;
 *  - per-cell sums run sequentially over the cell's simulators in sorted
 *    order (stable by cell, then previous slot);
 *  - XnParticle sums are F_N times the raw sums (cell and radial weights are
 *    1: weighting is out of scope);
 *  - vibrational and electronic modes are not exchanged (out of scope; the
 *    reference's single-level electronic branch is a no-op besides RNG use).
 */
#include "orc.h"
#include "orc_prim.h"

#include <float.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define KB UGX_KB
#define PI_D 3.14159265358979323846
/* macroInterpolation: field rows (T, U, Pr, collision frequency, p, q, then
 * the pressure tensor (ES-BGK) or stress (USP) as xx xy xz yy yz zz), points
 * kept per face, and the tolerance of the tet search */
#define MIR 16
#define MI_FV 8
#define MI_TOL 1.0e-12
#define VSMALL 1.0e-300
#define SMALL 1.0e-15
#define MAXF UGX_MAX_CELL_FACES
#define NBR_EMPTY (-2147483647 - 1)
/* cell-face kinds for the hybrid decomposition's face interpolation */
#define KNR 16            /* hybrid time-sum rows per cell */
#define KN_NB_CAP 256     /* neighbourhood list capacity (3 levels of <= 6 faces: <= 187) */
#define FACE_SKIP 0       /* padded slot or empty patch: no face value */
#define FACE_INTERNAL 1
#define FACE_CYCLIC 2
#define FACE_BOUNDARY 3   /* zeroGradient: face value = cell value */
#define FACE_SYMMETRY 4   /* symmetryPlane: vectors lose their normal component */
#define CELL_DEAD (-1)
#define MAX_CROSSINGS 1000
#define NS UGX_MAX_SPECIES
#define NSAMP UGX_S_COUNT
#define ORC_SQRT_HALF 7.07106781186547524401e-01   /* 1/sqrt(2): the BGK-family draws' N(0, 1/2) */
#define ORC_THIRD (1.0 / 3.0)

static __thread char g_err[512];
const char* orc_last_error(void) { return g_err; }
static int fail(int code, const char* msg) { snprintf(g_err, sizeof g_err, "%s", msg); return code; }

typedef struct {
    int kind;          /* ugx_boundary_kind for the patch (specular/diffuse/deletion) */
    double T, U[3];
} bmodel;

typedef struct {
    int cell, species, desc;
    double nbar;       /* expected simulators per step (Bird 4.22) at cell weight 1 */
    double nbar_num, nbar_den;   /* its numerator and F_N part of the denominator */
    double flux_an, flux_mps, flux_br;   /* face area x number density, most probable speed, flux bracket */
    int nv;
    double vloc[8][3]; /* face vertices in the cell frame */
    double cum[6];     /* cumulative triangle area fractions (fan from vertex 0) */
    double n[3], t1[3], t2[3];
    double T, rotT, U[3];
} inflow_fs;

/* One interaction with a measured wall face: velocity and E_rot before and
 * after, the wall face and the sort key (species << 42 | slot << 10 |
 * crossing). */
typedef struct {
    uint64_t key;
    int32_t wf;
    double u0[3], u1[3], e0, e1;
} wall_ev;

#define WROWS UGX_W_COUNT

struct orc_world {
    /* mesh */
    int nc;
    double* cc;       /* [nc*3] */
    double* vol;      /* [nc] */
    double* bblo;     /* [nc*3] bounding box in the cell frame */
    double* bbhi;
    int32_t* ncf;     /* [nc] */
    double* plane;    /* [nc*6*4] */
    int32_t* nbr;     /* [nc*6] */
    double* delta;    /* [nc*6*3] */
    double* farea;    /* [nc*6] face area (0 for padded and empty faces) */
    int8_t* fkind;    /* [nc*6] FACE_* */
    double mask[3];   /* solution directions (0 for empty directions) */
    int npatch;
    bmodel* bm;       /* [npatch] */
    int nin;
    inflow_fs* inflow;
    /* species */
    int ns;
    double mass[NS], diam[NS], omega[NS], alpha[NS], muref[NS], prterm[NS];
    int rotdof[NS];
    double pmr[NS * NS], pC[NS * NS], pe[NS * NS], pchi[NS * NS], palpha[NS * NS];
    /* model */
    ugx_model_desc model;
    uint64_t seed;
    uint32_t step;
    /* particles */
    int64_t n, cap;
    double *x, *v, *erot;
    int32_t *cell, *sp;
    double *x2, *v2, *erot2;
    int32_t *cell2, *sp2;
    int64_t* off;     /* [nc+1] */
    /* per cell */
    double* smax;
    double* samp;     /* [NSAMP*nc] */
    int32_t* cmodel;  /* 1 DSMC, 0 BGK */
    double* uspq;     /* [nc*3] previous heat flux */
    double* usps;     /* [nc*6] previous stress */
    double* uspmax;   /* [nc] */
    double* knacc;    /* hybrid: time sums since the last decomposition [nc*KNR] */
    double* kn;       /* [nc] */
    double kn_time;
    int kn_steps;
    /* measured wall faces (uniGasPatchBoundary measurePropertiesAtWall_) */
    int nwf;
    int32_t* wface;   /* [nc*6] wall face id of a cell face, or -1 */
    double* wgeo;     /* [nwf*4] unit outward normal, area */
    double* wsamp;    /* [nwf*NS*WROWS] wall time sums */
    wall_ev* ev;      /* this step's interactions */
    int64_t nev, evcap;
    int32_t* wcell;   /* [nwf] cell of each measured wall face */
    /* cell weighting */
    double* cw;       /* [nc] cell weights (1 without weighting) */
    int32_t* cstart;  /* [cap] cell each simulator started the step in */
    /* subcells of the NTC partner selection: levels per cell and axis */
    int32_t* sclev;   /* [nc*3] (1 in every direction by default) */
    /* time and the dynamic adapter */
    double time;      /* simulated time at the start of the step */
    double* adacc;    /* [nc*KNR] adapter time sums since the last adaptation */
    double* adratio;  /* [nc*3] cell size to mean free path ratio of the last adaptation */
    double adapt_time;
    /* macroInterpolation */
    int macro;
    int npts;
    double* mi_cent;   /* [nc*3] OpenFOAM cell centroid in the cell frame */
    int32_t* mi_fp;    /* [nc*6*8] point ids of the cell's faces (ascending face order), -1 padded */
    double* mi_pts;    /* [npts*3] */
    uint32_t* mi_poff; /* [npts+1] */
    int32_t* mi_pcell;
    double* mi_pw;
    double* mfield;    /* [nc*16] */
    double* pfield;    /* [npts*16] */
    /* counters */
    ugx_counters cnt;
};

/* largest subcell level along an axis (uniGasDynamicAdapter.C:49) */
#define ORC_MAX_SUBCELL_LEVELS 10

static void smooth_field(const orc_world* w, const double* f, double* out, int ncomp);
static int fill_weights(orc_world* w, double ntot);
static int adapt(orc_world* w);
static void smooth_max_ratio(const orc_world* w, double* f, double max_ratio);
static int cell_weighting(orc_world* w, int64_t n0, int64_t nall, int64_t* nclone, const double* cw_old);
static int adapt_fill(orc_world* w, const ugx_fill_desc* fl);

/* ------------------------------------------------------------- geometry */

static void cross3(const double a[3], const double b[3], double r[3])
{
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

static double dot3(const double a[3], const double b[3])
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* Face centre (vertex mean) and area vector (fan about the centre). */
static void face_geom(const ugx_mesh_desc* m, int f, double pc[3], double A[3])
{
    int b = m->face_offsets[f], e = m->face_offsets[f + 1], k = e - b;
    double s[3] = {0.0, 0.0, 0.0};
    for (int i = b; i < e; ++i)
        for (int d = 0; d < 3; ++d) s[d] = s[d] + m->points[3 * m->face_points[i] + d];
    for (int d = 0; d < 3; ++d) pc[d] = s[d] / (double)k;
    double a[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < k; ++i) {
        const double* p0 = &m->points[3 * m->face_points[b + i]];
        const double* p1 = &m->points[3 * m->face_points[b + (i + 1) % k]];
        double u[3], w[3], c[3];
        for (int d = 0; d < 3; ++d) { u[d] = p0[d] - pc[d]; w[d] = p1[d] - pc[d]; }
        cross3(u, w, c);
        for (int d = 0; d < 3; ++d) a[d] = a[d] + c[d];
    }
    for (int d = 0; d < 3; ++d) A[d] = 0.5 * a[d];
}

/* The geometry of the cellPoint interpolation (OpenFOAM interpolationCellPoint
 * with volPointInterpolation): face centroids and areas of the fan about the
 * vertex mean, cell centroids of the face pyramids about the mean of the face
 * centroids (owner faces, then neighbour faces), per cell its faces' points in
 * ascending face order, and per point the inverse-distance weights of the cells
 * around it, or for a point on a wall, patch or symmetry face of those faces'
 * owner values; points joined by cyclic patches share one normalised set. */
static int find_root(int* up, int a)
{
    while (up[a] != a) { up[a] = up[up[a]]; a = up[a]; }
    return a;
}

static double pdist(const double* a, const double* b)
{
    double d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
    return sqrt(d0 * d0 + d1 * d1 + d2 * d2);
}

static int cmp_int(const void* a, const void* b)
{
    int x = *(const int*)a, y = *(const int*)b;
    return (x > y) - (x < y);
}

static int build_interp(orc_world* w, const ugx_mesh_desc* m)
{
    int nc = m->n_cells, nf = m->n_faces, np = m->n_points, ni = m->n_internal_faces;
    const double* P = m->points;
    double* fc = calloc((size_t)nf * 3, sizeof(double));
    double* fa = calloc((size_t)nf * 3, sizeof(double));
    for (int f = 0; f < nf; ++f) {
        int b = m->face_offsets[f], n = m->face_offsets[f + 1] - b;
        const int32_t* fp = &m->face_points[b];
        double* C = &fc[3 * f];
        double* A = &fa[3 * f];
        if (n == 3) {
            double u[3], v[3];
            for (int d = 0; d < 3; ++d) C[d] = (1.0 / 3.0) * (P[3 * fp[0] + d] + P[3 * fp[1] + d] + P[3 * fp[2] + d]);
            for (int d = 0; d < 3; ++d) { u[d] = P[3 * fp[1] + d] - P[3 * fp[0] + d]; v[d] = P[3 * fp[2] + d] - P[3 * fp[0] + d]; }
            cross3(u, v, A);
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
            for (int d = 0; d < 3; ++d) { c[d] = p0[d] + p1[d] + mean[d]; u[d] = p1[d] - p0[d]; v[d] = mean[d] - p0[d]; }
            cross3(u, v, nn);
            double a = sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
            for (int d = 0; d < 3; ++d) { sumN[d] = sumN[d] + nn[d]; sumAc[d] = sumAc[d] + a * c[d]; }
            sumA = sumA + a;
        }
        if (sumA < 1.0e-150) {
            for (int d = 0; d < 3; ++d) { C[d] = mean[d]; A[d] = 0.0; }
        } else {
            for (int d = 0; d < 3; ++d) { C[d] = (1.0 / 3.0) * sumAc[d] / sumA; A[d] = 0.5 * sumN[d]; }
        }
    }
    double* cest = calloc((size_t)nc * 3, sizeof(double));
    double* cctr = calloc((size_t)nc * 3, sizeof(double));
    double* cvol = calloc((size_t)nc, sizeof(double));
    int* nfc = calloc((size_t)nc, sizeof(int));
    for (int side = 0; side < 2; ++side)
        for (int f = 0; f < (side == 0 ? nf : ni); ++f) {
            int c = side == 0 ? m->owner[f] : m->neighbour[f];
            for (int d = 0; d < 3; ++d) cest[3 * c + d] = cest[3 * c + d] + fc[3 * f + d];
            nfc[c] += 1;
        }
    for (int c = 0; c < nc; ++c)
        for (int d = 0; d < 3; ++d) cest[3 * c + d] = cest[3 * c + d] / (double)nfc[c];
    for (int side = 0; side < 2; ++side)
        for (int f = 0; f < (side == 0 ? nf : ni); ++f) {
            int c = side == 0 ? m->owner[f] : m->neighbour[f];
            const double* C = &fc[3 * f];
            const double* A = &fa[3 * f];
            const double* E = &cest[3 * c];
            double pyr = side == 0 ? A[0] * (C[0] - E[0]) + A[1] * (C[1] - E[1]) + A[2] * (C[2] - E[2])
                                   : A[0] * (E[0] - C[0]) + A[1] * (E[1] - C[1]) + A[2] * (E[2] - C[2]);
            if (pyr < 1.0e-300) pyr = 1.0e-300;
            for (int d = 0; d < 3; ++d) cctr[3 * c + d] = cctr[3 * c + d] + pyr * (0.75 * C[d] + 0.25 * E[d]);
            cvol[c] = cvol[c] + pyr;
        }
    w->mi_cent = calloc((size_t)nc * 3, sizeof(double));
    for (int c = 0; c < nc; ++c)
        for (int d = 0; d < 3; ++d) {
            double x = fabs(cvol[c]) > 1.0e-300 ? cctr[3 * c + d] / cvol[c] : cest[3 * c + d];
            cctr[3 * c + d] = x;
            w->mi_cent[3 * c + d] = x - w->cc[3 * c + d];
        }
    /* faces of each cell, and per point its faces and cells */
    w->mi_fp = malloc(sizeof(int32_t) * (size_t)nc * MAXF * MI_FV);
    for (size_t k = 0; k < (size_t)nc * MAXF * MI_FV; ++k) w->mi_fp[k] = -1;
    int* nface = calloc((size_t)nc, sizeof(int));
    int* pfn = calloc((size_t)np + 1, sizeof(int));
    int* pcn = calloc((size_t)np + 1, sizeof(int));
    for (int f = 0; f < nf; ++f)
        for (int i = m->face_offsets[f]; i < m->face_offsets[f + 1]; ++i) {
            pfn[m->face_points[i] + 1] += 1;
            pcn[m->face_points[i] + 1] += (f < ni) ? 2 : 1;
        }
    for (int p = 0; p < np; ++p) { pfn[p + 1] += pfn[p]; pcn[p + 1] += pcn[p]; }
    int* pf = malloc(sizeof(int) * (size_t)(pfn[np] > 0 ? pfn[np] : 1));
    int* pcl = malloc(sizeof(int) * (size_t)(pcn[np] > 0 ? pcn[np] : 1));
    int* pfi = calloc((size_t)np, sizeof(int));
    int* pci = calloc((size_t)np, sizeof(int));
    for (int f = 0; f < nf; ++f) {
        int b = m->face_offsets[f], n = m->face_offsets[f + 1] - b;
        if (n > MI_FV) return fail(UGX_ERR_MESH, "macroInterpolation: face with more than 8 points");
        for (int side = 0; side < (f < ni ? 2 : 1); ++side) {
            int c = side == 0 ? m->owner[f] : m->neighbour[f];
            int j = nface[c]++;
            for (int i = 0; i < n; ++i) w->mi_fp[((size_t)c * MAXF + j) * MI_FV + i] = m->face_points[b + i];
        }
        for (int i = 0; i < n; ++i) {
            int p = m->face_points[b + i];
            pf[pfn[p] + pfi[p]++] = f;
            for (int side = 0; side < (f < ni ? 2 : 1); ++side)
                pcl[pcn[p] + pci[p]++] = side == 0 ? m->owner[f] : m->neighbour[f];
        }
    }
    /* each point's cells ascending, without repeats */
    int* pcu = calloc((size_t)np, sizeof(int));
    for (int p = 0; p < np; ++p) {
        int* v = &pcl[pcn[p]];
        int n = pcn[p + 1] - pcn[p], k = 0;
        qsort(v, (size_t)n, sizeof(int), cmp_int);
        for (int i = 0; i < n; ++i) if (i == 0 || v[i] != v[i - 1]) v[k++] = v[i];
        pcu[p] = k;
    }
    int8_t* vface = calloc((size_t)nf, 1);
    int8_t* ppoint = calloc((size_t)np, 1);
    int* up = malloc(sizeof(int) * (size_t)np);
    for (int p = 0; p < np; ++p) up[p] = p;
    for (int pa = 0; pa < m->n_patches; ++pa) {
        int t = m->patch_type[pa];
        if (t == UGX_PATCH_EMPTY) continue;
        if (t == UGX_PATCH_CYCLIC) {
            int pb = m->patch_neighbour ? m->patch_neighbour[pa] : -1;
            if (pb < pa) continue;
            for (int i = 0; i < m->patch_size[pa]; ++i) {
                int f0 = m->patch_start[pa] + i, f1 = m->patch_start[pb] + i;
                double T[3];
                for (int d = 0; d < 3; ++d) T[d] = fc[3 * f1 + d] - fc[3 * f0 + d];
                double sz = sqrt(fa[3 * f0] * fa[3 * f0] + fa[3 * f0 + 1] * fa[3 * f0 + 1] + fa[3 * f0 + 2] * fa[3 * f0 + 2]);
                for (int a = m->face_offsets[f0]; a < m->face_offsets[f0 + 1]; ++a) {
                    int p0 = m->face_points[a], best = -1;
                    double bd = 1.0e300;
                    for (int q = m->face_offsets[f1]; q < m->face_offsets[f1 + 1]; ++q) {
                        int p1 = m->face_points[q];
                        double x[3];
                        for (int d = 0; d < 3; ++d) x[d] = P[3 * p0 + d] + T[d];
                        double dd = pdist(x, &P[3 * p1]);
                        if (dd < bd) { bd = dd; best = p1; }
                    }
                    if (best < 0 || bd > 1.0e-6 * sqrt(sz)) return fail(UGX_ERR_MESH, "cyclic faces do not match point for point");
                    int r0 = find_root(up, p0), r1 = find_root(up, best);
                    if (r0 != r1) { if (r0 < r1) up[r1] = r0; else up[r0] = r1; }
                }
            }
            continue;
        }
        for (int f = m->patch_start[pa]; f < m->patch_start[pa] + m->patch_size[pa]; ++f) {
            vface[f] = 1;
            for (int a = m->face_offsets[f]; a < m->face_offsets[f + 1]; ++a) ppoint[m->face_points[a]] = 1;
        }
    }
    /* classes of joined points (members ascending), their weights */
    int* root = malloc(sizeof(int) * (size_t)np);
    int* mcount = calloc((size_t)np + 1, sizeof(int));
    for (int p = 0; p < np; ++p) { root[p] = find_root(up, p); mcount[root[p] + 1] += 1; }
    for (int p = 0; p < np; ++p) mcount[p + 1] += mcount[p];
    int* mem = malloc(sizeof(int) * (size_t)(np > 0 ? np : 1));
    int* mfill = calloc((size_t)np, sizeof(int));
    for (int p = 0; p < np; ++p) mem[mcount[root[p]] + mfill[root[p]]++] = p;
    /* entries per class */
    int* cls_n = calloc((size_t)np, sizeof(int));
    int total = 0;
    for (int r = 0; r < np; ++r) {
        int n0 = mcount[r], n1 = mcount[r + 1];
        if (n1 == n0) continue;
        int patch = 0;
        for (int k = n0; k < n1; ++k) patch = patch || ppoint[mem[k]];
        int cnt = 0;
        for (int k = n0; k < n1; ++k) {
            int p = mem[k];
            if (patch) { for (int i = pfn[p]; i < pfn[p + 1]; ++i) if (vface[pf[i]]) cnt++; }
            else cnt += pcu[p];
        }
        cls_n[r] = cnt;
    }
    w->mi_poff = calloc((size_t)np + 1, sizeof(uint32_t));
    for (int p = 0; p < np; ++p) w->mi_poff[p + 1] = w->mi_poff[p] + (uint32_t)cls_n[root[p]];
    total = (int)w->mi_poff[np];
    w->mi_pcell = malloc(sizeof(int32_t) * (size_t)(total > 0 ? total : 1));
    w->mi_pw = malloc(sizeof(double) * (size_t)(total > 0 ? total : 1));
    int32_t* ec = malloc(sizeof(int32_t) * (size_t)(total > 0 ? total : 1));
    double* ew = malloc(sizeof(double) * (size_t)(total > 0 ? total : 1));
    for (int r = 0; r < np; ++r) {
        int n0 = mcount[r], n1 = mcount[r + 1];
        if (n1 == n0) continue;
        int patch = 0, e = 0;
        for (int k = n0; k < n1; ++k) patch = patch || ppoint[mem[k]];
        for (int k = n0; k < n1; ++k) {
            int p = mem[k];
            if (patch) {
                for (int i = pfn[p]; i < pfn[p + 1]; ++i)
                    if (vface[pf[i]]) { ec[e] = m->owner[pf[i]]; ew[e] = 1.0 / pdist(&P[3 * p], &fc[3 * pf[i]]); e++; }
            } else {
                for (int i = 0; i < pcu[p]; ++i) { int c = pcl[pcn[p] + i]; ec[e] = c; ew[e] = 1.0 / pdist(&P[3 * p], &cctr[3 * c]); e++; }
            }
        }
        double sum = 0.0;
        for (int k = 0; k < e; ++k) sum = sum + ew[k];
        for (int k = 0; k < e; ++k) ew[k] = ew[k] / sum;
        for (int k = n0; k < n1; ++k) {
            int p = mem[k];
            for (int q = 0; q < e; ++q) { w->mi_pcell[w->mi_poff[p] + q] = ec[q]; w->mi_pw[w->mi_poff[p] + q] = ew[q]; }
        }
    }
    w->npts = np;
    w->mi_pts = malloc(sizeof(double) * (size_t)np * 3);
    memcpy(w->mi_pts, P, sizeof(double) * (size_t)np * 3);
    w->mfield = calloc((size_t)nc * MIR, sizeof(double));
    w->pfield = calloc((size_t)np * MIR, sizeof(double));
    void* tmp[] = {fc, fa, cest, cctr, cvol, nfc, nface, pfn, pcn, pf, pcl, pfi, pci, pcu, vface, ppoint, up, root,
                   mcount, mem, mfill, cls_n, ec, ew};
    for (size_t i = 0; i < sizeof tmp / sizeof tmp[0]; ++i) free(tmp[i]);
    return UGX_OK;
}

static int build_mesh(orc_world* w, const ugx_mesh_desc* m, const ugx_boundary_desc* bnd, int nb)
{
    int nc = m->n_cells, nf = m->n_faces;
    w->nc = nc;
    w->cc = calloc((size_t)nc * 3, sizeof(double));
    w->vol = calloc((size_t)nc, sizeof(double));
    w->bblo = calloc((size_t)nc * 3, sizeof(double));
    w->bbhi = calloc((size_t)nc * 3, sizeof(double));
    w->ncf = calloc((size_t)nc, sizeof(int32_t));
    w->plane = calloc((size_t)nc * MAXF * 4, sizeof(double));
    w->nbr = calloc((size_t)nc * MAXF, sizeof(int32_t));
    w->delta = calloc((size_t)nc * MAXF * 3, sizeof(double));
    w->farea = calloc((size_t)nc * MAXF, sizeof(double));
    w->fkind = calloc((size_t)nc * MAXF, sizeof(int8_t));
    int32_t* cf = malloc(sizeof(int32_t) * (size_t)nc * MAXF);
    double* pc = malloc(sizeof(double) * (size_t)nf * 3);
    double* A = malloc(sizeof(double) * (size_t)nf * 3);
    int* fpatch = malloc(sizeof(int) * (size_t)nf);
    for (int f = 0; f < nf; ++f) { face_geom(m, f, &pc[3 * f], &A[3 * f]); fpatch[f] = -1; }
    for (int p = 0; p < m->n_patches; ++p)
        for (int f = m->patch_start[p]; f < m->patch_start[p] + m->patch_size[p]; ++f) fpatch[f] = p;
    /* faces per cell in increasing face id */
    for (int f = 0; f < nf; ++f) {
        int o = m->owner[f];
        if (w->ncf[o] >= MAXF) return fail(UGX_ERR_MESH, "cell with more than 6 faces");
        cf[o * MAXF + w->ncf[o]++] = f;
        if (f < m->n_internal_faces) {
            int nbc = m->neighbour[f];
            if (w->ncf[nbc] >= MAXF) return fail(UGX_ERR_MESH, "cell with more than 6 faces");
            cf[nbc * MAXF + w->ncf[nbc]++] = f;
        }
    }
    /* reference point: mean of unique vertices (first appearance order) */
    for (int c = 0; c < nc; ++c) {
        int ids[64], nid = 0;
        for (int j = 0; j < w->ncf[c]; ++j) {
            int f = cf[c * MAXF + j];
            for (int i = m->face_offsets[f]; i < m->face_offsets[f + 1]; ++i) {
                int pid = m->face_points[i], seen = 0;
                for (int q = 0; q < nid; ++q) if (ids[q] == pid) { seen = 1; break; }
                if (!seen && nid < 64) ids[nid++] = pid;
            }
        }
        double s[3] = {0.0, 0.0, 0.0};
        for (int q = 0; q < nid; ++q)
            for (int d = 0; d < 3; ++d) s[d] = s[d] + m->points[3 * ids[q] + d];
        for (int d = 0; d < 3; ++d) w->cc[3 * c + d] = s[d] / (double)nid;
        for (int d = 0; d < 3; ++d) { w->bblo[3 * c + d] = DBL_MAX; w->bbhi[3 * c + d] = -DBL_MAX; }
        for (int q = 0; q < nid; ++q)
            for (int d = 0; d < 3; ++d) {
                double r = m->points[3 * ids[q] + d] - w->cc[3 * c + d];
                if (r < w->bblo[3 * c + d]) w->bblo[3 * c + d] = r;
                if (r > w->bbhi[3 * c + d]) w->bbhi[3 * c + d] = r;
            }
    }
    /* boundary models per patch */
    w->npatch = m->n_patches;
    w->bm = calloc((size_t)(m->n_patches > 0 ? m->n_patches : 1), sizeof(bmodel));
    for (int p = 0; p < m->n_patches; ++p) {
        int t = m->patch_type[p];
        w->bm[p].kind = (t == UGX_PATCH_SYMMETRY) ? UGX_BND_SPECULAR : UGX_BND_DELETION;
        int found = 0;
        for (int b = 0; b < nb; ++b) {
            if (bnd[b].patch != p) continue;
            int k = bnd[b].kind;
            if (k == UGX_BND_SPECULAR || k == UGX_BND_DIFFUSE || k == UGX_BND_DELETION) {
                w->bm[p].kind = k;
                w->bm[p].T = bnd[b].temperature;
                for (int d = 0; d < 3; ++d) w->bm[p].U[d] = bnd[b].velocity[d];
                found = 1;
            }
        }
        if (t == UGX_PATCH_WALL && !found) return fail(UGX_ERR_ARG, "wall patch without a boundary model");
    }
    /* measured wall faces, numbered in patch then face order */
    int* wid = malloc(sizeof(int) * (size_t)(nf > 0 ? nf : 1));
    for (int f = 0; f < nf; ++f) wid[f] = -1;
    w->nwf = 0;
    for (int p = 0; p < m->n_patches; ++p)
        if (m->patch_type[p] == UGX_PATCH_WALL && (w->bm[p].kind == UGX_BND_SPECULAR || w->bm[p].kind == UGX_BND_DIFFUSE))
            for (int f = m->patch_start[p]; f < m->patch_start[p] + m->patch_size[p]; ++f) wid[f] = w->nwf++;
    w->wgeo = calloc((size_t)(w->nwf > 0 ? w->nwf : 1) * 4, sizeof(double));
    w->wcell = calloc((size_t)(w->nwf > 0 ? w->nwf : 1), sizeof(int32_t));
    w->wsamp = calloc((size_t)(w->nwf > 0 ? w->nwf : 1) * NS * WROWS, sizeof(double));
    w->wface = malloc(sizeof(int32_t) * (size_t)nc * MAXF);
    for (int k = 0; k < nc * MAXF; ++k) w->wface[k] = -1;
    for (int f = 0; f < nf; ++f) {
        if (wid[f] < 0) continue;
        const double* a = &A[3 * f];
        double mag = sqrt(dot3(a, a));
        double* gq = &w->wgeo[4 * (size_t)wid[f]];
        for (int d = 0; d < 3; ++d) gq[d] = a[d] / mag;
        gq[3] = mag;
        w->wcell[wid[f]] = m->owner[f];
    }
    /* solution directions from empty patches */
    for (int d = 0; d < 3; ++d) w->mask[d] = 1.0;
    for (int p = 0; p < m->n_patches; ++p) {
        if (m->patch_type[p] != UGX_PATCH_EMPTY || m->patch_size[p] == 0) continue;
        const double* a = &A[3 * m->patch_start[p]];
        int ax = 0;
        for (int d = 1; d < 3; ++d) if (fabs(a[d]) > fabs(a[ax])) ax = d;
        w->mask[ax] = 0.0;
    }
    /* planes, neighbours, shifts, volumes */
    for (int c = 0; c < nc; ++c) {
        double vs = 0.0;
        for (int j = 0; j < MAXF; ++j) {
            double* pl = &w->plane[(c * MAXF + j) * 4];
            if (j >= w->ncf[c]) { pl[0] = pl[1] = pl[2] = 0.0; pl[3] = 1.0; w->nbr[c * MAXF + j] = NBR_EMPTY; continue; }
            int f = cf[c * MAXF + j];
            double sg = (m->owner[f] == c) ? 1.0 : -1.0;
            double a[3] = {A[3 * f], A[3 * f + 1], A[3 * f + 2]};
            double r[3];
            for (int d = 0; d < 3; ++d) r[d] = pc[3 * f + d] - w->cc[3 * c + d];
            double dt = dot3(a, r);
            vs = vs + sg * dt;
            double mag = sqrt(dot3(a, a));
            double nn[3];
            for (int d = 0; d < 3; ++d) nn[d] = sg * (a[d] / mag);
            pl[0] = nn[0]; pl[1] = nn[1]; pl[2] = nn[2];
            pl[3] = dot3(nn, r);
            w->farea[c * MAXF + j] = mag;
            double* dl = &w->delta[(c * MAXF + j) * 3];
            if (f < m->n_internal_faces) {
                int o = (m->owner[f] == c) ? m->neighbour[f] : m->owner[f];
                w->nbr[c * MAXF + j] = o;
                w->fkind[c * MAXF + j] = FACE_INTERNAL;
                for (int d = 0; d < 3; ++d) dl[d] = w->cc[3 * c + d] - w->cc[3 * o + d];
            } else {
                int p = fpatch[f];
                if (p < 0) return fail(UGX_ERR_MESH, "boundary face outside every patch");
                int t = m->patch_type[p];
                if (t == UGX_PATCH_EMPTY) {
                    w->nbr[c * MAXF + j] = NBR_EMPTY;
                    w->farea[c * MAXF + j] = 0.0;
                } else if (t == UGX_PATCH_CYCLIC) {
                    w->fkind[c * MAXF + j] = FACE_CYCLIC;
                    int q = m->patch_neighbour ? m->patch_neighbour[p] : -1;
                    if (q < 0 || m->patch_size[q] != m->patch_size[p]) return fail(UGX_ERR_MESH, "cyclic patch without a matching partner");
                    int g = m->patch_start[q] + (f - m->patch_start[p]);
                    int o = m->owner[g];
                    w->nbr[c * MAXF + j] = o;
                    for (int d = 0; d < 3; ++d)
                        dl[d] = (w->cc[3 * c + d] - w->cc[3 * o + d]) + (pc[3 * g + d] - pc[3 * f + d]);
                } else {
                    w->nbr[c * MAXF + j] = -1 - p;
                    w->fkind[c * MAXF + j] = (t == UGX_PATCH_SYMMETRY) ? FACE_SYMMETRY : FACE_BOUNDARY;
                    w->wface[c * MAXF + j] = wid[f];
                }
            }
        }
        w->vol[c] = vs / 3.0;
        if (!(w->vol[c] > 0.0)) return fail(UGX_ERR_MESH, "cell with non-positive volume");
    }
    /* freestream inflow face-species entries (uniGasGeneralBoundary.C:115-169) */
    int nin = 0;
    for (int b = 0; b < nb; ++b)
        if (bnd[b].kind == UGX_BND_FREESTREAM)
            for (int s = 0; s < w->ns; ++s)
                if (bnd[b].number_density[s] > 0.0) nin += m->patch_size[bnd[b].patch];
    w->nin = nin;
    w->inflow = calloc((size_t)(nin > 0 ? nin : 1), sizeof(inflow_fs));
    int k = 0;
    const double sqrtPi = sqrt(PI_D);
    for (int b = 0; b < nb; ++b) {
        if (bnd[b].kind != UGX_BND_FREESTREAM) continue;
        int p = bnd[b].patch;
        for (int f = m->patch_start[p]; f < m->patch_start[p] + m->patch_size[p]; ++f) {
            for (int s = 0; s < w->ns; ++s) {
                double nd = bnd[b].number_density[s];
                if (!(nd > 0.0)) continue;
                inflow_fs* e = &w->inflow[k++];
                int c = m->owner[f];
                e->cell = c; e->species = s; e->desc = b;
                e->T = bnd[b].temperature; e->rotT = bnd[b].rot_temperature;
                for (int d = 0; d < 3; ++d) e->U[d] = bnd[b].velocity[d];
                const double* sF = &A[3 * f];
                double fA = sqrt(dot3(sF, sF));
                double mps = sqrt(2.0 * KB * e->T / w->mass[s]);
                double nrm[3] = {-sF[0] / fA, -sF[1] / fA, -sF[2] / fA};
                double sCos = dot3(e->U, nrm) / mps;
                e->flux_an = fA * nd;
                e->flux_mps = mps;
                e->flux_br = exp(-(sCos * sCos)) + sqrtPi * sCos * (1.0 + erf(sCos));
                e->nbar_num = e->flux_an * w->model.delta_t * e->flux_mps * e->flux_br;
                e->nbar_den = 2.0 * sqrtPi * w->model.n_equivalent_particles;
                e->nbar = e->nbar_num / e->nbar_den;
                int fb = m->face_offsets[f], fe = m->face_offsets[f + 1];
                e->nv = fe - fb;
                if (e->nv > 8) return fail(UGX_ERR_MESH, "inflow face with more than 8 vertices");
                for (int i = 0; i < e->nv; ++i)
                    for (int d = 0; d < 3; ++d)
                        e->vloc[i][d] = m->points[3 * m->face_points[fb + i] + d] - w->cc[3 * c + d];
                double tot = 0.0, ar[6];
                for (int t = 0; t < e->nv - 2; ++t) {
                    double u[3], ww[3], cr[3];
                    for (int d = 0; d < 3; ++d) { u[d] = e->vloc[t + 1][d] - e->vloc[0][d]; ww[d] = e->vloc[t + 2][d] - e->vloc[0][d]; }
                    cross3(u, ww, cr);
                    ar[t] = 0.5 * sqrt(dot3(cr, cr));
                    tot = tot + ar[t];
                }
                double run = 0.0;
                for (int t = 0; t < e->nv - 2; ++t) { run = run + ar[t]; e->cum[t] = run / tot; }
                e->cum[e->nv - 3] = 1.0;
                for (int d = 0; d < 3; ++d) e->n[d] = nrm[d];
                double t1[3];
                for (int d = 0; d < 3; ++d) t1[d] = pc[3 * f + d] - m->points[3 * m->face_points[fb] + d];
                double tm = sqrt(dot3(t1, t1));
                for (int d = 0; d < 3; ++d) e->t1[d] = t1[d] / tm;
                double t2[3];
                cross3(e->n, e->t1, t2);
                double t2m = sqrt(dot3(t2, t2));
                for (int d = 0; d < 3; ++d) e->t2[d] = t2[d] / t2m;
            }
        }
    }
    free(cf); free(pc); free(A); free(fpatch); free(wid);
    return UGX_OK;
}

/* ------------------------------------------------------- kinetic helpers */

/* uniGasCloud::equipartitionRotationalEnergy, uniGasCloud.C:975-1017 */
static double equip_erot(orc_rng* r, double T, int dof)
{
    if (dof <= 0) return 0.0;
    if (dof == 2) return -orc_log(orc_uniform(r)) * KB * T;
    double a = 0.5 * (double)dof - 1.0;
    double eps = orc_uniform(r);
    double x = 0.0, P = -1.0;
    for (int it = 0; it < 100000; ++it) {
        x = 10.0 * orc_uniform(r);
        P = orc_pow(x / a, a) * orc_exp(a - x);
        if (!(P < eps)) break;
    }
    return x * KB * T;
}

/* uniGasCloud::postCollisionRotationalEnergy, uniGasCloud.C:1129-1189 */
static double post_rot_ratio(orc_rng* r, int dof, double chiB)
{
    if (dof == 2) return 1.0 - orc_pow(orc_uniform(r), 1.0 / chiB);
    double chiA = 0.5 * (double)dof;
    double am1 = chiA - 1.0, bm1 = chiB - 1.0;
    if (am1 < SMALL && bm1 < SMALL) return orc_uniform(r);
    double eps = orc_uniform(r);
    double ratio = 0.0;
    for (int it = 0; it < 100000; ++it) {
        ratio = orc_uniform(r);
        double P;
        if (am1 < SMALL) P = orc_pow(1.0 - ratio, bm1);
        else if (bm1 < SMALL) P = orc_pow(1.0 - ratio, am1);
        else P = orc_pow((am1 + bm1) * ratio / am1, am1) * orc_pow((am1 + bm1) * (1.0 - ratio) / bm1, bm1);
        if (!(P < eps)) break;
    }
    return ratio;
}

/* uniGasPatchBoundary::diffuseReflection, uniGasPatchBoundary.C:305-388 */
static void diffuse_reflect(orc_rng* r, const double nw[3], double U[3], double* erot,
                            double T, const double Uw[3], double mass, int rotdof)
{
    double udn = dot3(U, nw);
    double Ut[3];
    for (int d = 0; d < 3; ++d) Ut[d] = U[d] - udn * nw[d];
    double mUt = sqrt(dot3(Ut, Ut));
    for (int it = 0; it < 1000 && mUt < SMALL; ++it) {
        for (int d = 0; d < 3; ++d) U[d] = U[d] * (0.8 + 0.2 * orc_uniform(r));
        udn = dot3(U, nw);
        for (int d = 0; d < 3; ++d) Ut[d] = U[d] - udn * nw[d];
        mUt = sqrt(dot3(Ut, Ut));
    }
    double t1[3], t2[3];
    for (int d = 0; d < 3; ++d) t1[d] = Ut[d] / mUt;
    cross3(nw, t1, t2);
    double sig = sqrt(KB * T / mass);
    double g1, g2;
    orc_gauss2(r, &g1, &g2);
    double un = sqrt(-2.0 * orc_log(orc_uniform(r)));
    for (int d = 0; d < 3; ++d) U[d] = sig * ((g1 * t1[d] + g2 * t2[d]) - un * nw[d]) + Uw[d];
    *erot = equip_erot(r, T, rotdof);
}

/* ------------------------------------------------------------------ move */

static int sampling_step(const orc_world* w)
{
    return ((w->step + 1) % (uint32_t)w->model.sample_interval) == 0;
}

/* Records one wall interaction of a sampling step (measurePropertiesBefore/
 * AfterControl are called around the wall model, uniGasDiffuseWallPatch.C
 * and uniGasSpecularWallPatch.C controlParticle). */
static void wall_event(orc_world* w, int wf, int sp, int64_t slot, int it, const double u0[3], const double u1[3],
                       double e0, double e1)
{
    wall_ev e;
    e.key = ((uint64_t)sp << 42) | ((uint64_t)slot << 10) | (uint64_t)it;
    e.wf = wf;
    for (int d = 0; d < 3; ++d) { e.u0[d] = u0[d]; e.u1[d] = u1[d]; }
    e.e0 = e0;
    e.e1 = e1;
#pragma omp critical(orc_wall_events)
    {
        if (w->nev == w->evcap) {
            w->evcap = w->evcap ? 2 * w->evcap : 4096;
            w->ev = realloc(w->ev, sizeof(wall_ev) * (size_t)w->evcap);
        }
        w->ev[w->nev++] = e;
    }
}

static int cmp_ev(const void* a, const void* b)
{
    const wall_ev* p = a;
    const wall_ev* q = b;
    if (p->wf != q->wf) return p->wf < q->wf ? -1 : 1;
    return (p->key < q->key) ? -1 : (p->key > q->key);
}

/* before/after terms of one velocity and E_rot, uniGasPatchBoundary.C:147-184, 223-254 */
static void wall_terms(const double U[3], double e, const double n[3], double fA, double m, int rotdof,
                       double a[WROWS])
{
    double Un = U[0] * n[0] + U[1] * n[1] + U[2] * n[2];
    double inv = 1.0 / fmax(fabs(Un) * fA, VSMALL);
    double UU = U[0] * U[0] + U[1] * U[1] + U[2] * U[2];
    a[UGX_W_RHON] += inv;
    if (rotdof > 0) a[UGX_W_RHON_INT] += inv;
    a[UGX_W_RHOM] += m * inv;
    a[UGX_W_LINEAR_KE] += 0.5 * m * UU * inv;
    a[UGX_W_MCC] += m * UU * inv;
    a[UGX_W_MU] += m * U[0] * inv;
    a[UGX_W_MV] += m * U[1] * inv;
    a[UGX_W_MW] += m * U[2] * inv;
    a[UGX_W_EROT] += e * inv;
    a[UGX_W_ROTDOF] += (double)rotdof * inv;
}

/* The step's wall sums (boundaryMeasurements, uniGasPatchBoundary.C:130-302)
 * added into the time sums at a sampling step (uniGasVolFields.C:800-830);
 * each face's interactions are summed per species in key order. */
static void wall_reduce(orc_world* w)
{
    if (w->nev == 0) return;
    qsort(w->ev, (size_t)w->nev, sizeof(wall_ev), cmp_ev);
    const double dt = w->model.delta_t;
    int64_t i = 0;
    while (i < w->nev) {
        const int wf = w->ev[i].wf;
        const int sp = (int)(w->ev[i].key >> 42);
        const double* gq = &w->wgeo[4 * (size_t)wf];
        const double n[3] = {gq[0], gq[1], gq[2]};
        const double fA = gq[3];
        const double m = w->mass[sp];
        const int rotdof = w->rotdof[sp];
        /* nParticle F_N W_c of the face's cell (uniGasPatchBoundary.C:292-294) */
        const double FN = w->model.n_equivalent_particles * w->cw[w->wcell[wf]];
        double a[WROWS];
        for (int r = 0; r < WROWS; ++r) a[r] = 0.0;
        for (; i < w->nev && w->ev[i].wf == wf && (int)(w->ev[i].key >> 42) == sp; ++i) {
            const wall_ev* e = &w->ev[i];
            wall_terms(e->u0, e->e0, n, fA, m, rotdof, a);
            double preIE = 0.5 * m * (e->u0[0] * e->u0[0] + e->u0[1] * e->u0[1] + e->u0[2] * e->u0[2]) + e->e0;
            wall_terms(e->u1, e->e1, n, fA, m, rotdof, a);
            double postIE = 0.5 * m * (e->u1[0] * e->u1[0] + e->u1[1] * e->u1[1] + e->u1[2] * e->u1[2]) + e->e1;
            a[UGX_W_Q] += FN * (preIE - postIE) / (dt * fA);
            for (int d = 0; d < 3; ++d) a[UGX_W_FX + d] += FN * (m * e->u0[d] - m * e->u1[d]) / (dt * fA);
        }
        double* dst = &w->wsamp[((size_t)wf * NS + sp) * WROWS];
        for (int r = 0; r < WROWS; ++r) dst[r] += dt * a[r];
    }
    w->nev = 0;
}

/* uniGasParcel::move (parcels/uniGasParcel.C:35-108) over a convex-cell face
 * tracker, with the wall dispatch of uniGasParcel::hitWallPatch (:131-145).
 * Returns 1 if the simulator survives. */
static int track(orc_world* w, int32_t* cellp, double x[3], double U[3], double* erot, int sp, int64_t slot,
                 double tau, orc_rng* r, int64_t* hits, int64_t* errs)
{
    const int meas = w->nwf > 0 && sampling_step(w);
    int c = *cellp;
    double vt[3];
    for (int d = 0; d < 3; ++d) vt[d] = U[d] * w->mask[d];
    for (int it = 0; it < MAX_CROSSINGS; ++it) {
        const double* pl = &w->plane[(size_t)c * MAXF * 4];
        const int32_t* nb = &w->nbr[(size_t)c * MAXF];
        double tmin = tau;
        int fmin = -1;
        for (int f = 0; f < MAXF; ++f) {
            if (nb[f] == NBR_EMPTY) continue;
            const double* q = &pl[4 * f];
            double dn = q[0] * vt[0] + q[1] * vt[1] + q[2] * vt[2];
            if (dn > 0.0) {
                double t = (q[3] - (q[0] * x[0] + q[1] * x[1] + q[2] * x[2])) / dn;
                if (t < 0.0) t = 0.0;
                if (t < tmin) { tmin = t; fmin = f; }
            }
        }
        if (fmin < 0) {
            for (int d = 0; d < 3; ++d) x[d] = x[d] + vt[d] * tau;
            *cellp = c;
            return 1;
        }
        for (int d = 0; d < 3; ++d) x[d] = x[d] + vt[d] * tmin;
        tau = tau - tmin;
        int code = nb[fmin];
        if (code >= 0) {
            const double* dl = &w->delta[((size_t)c * MAXF + fmin) * 3];
            for (int d = 0; d < 3; ++d) x[d] = x[d] + dl[d];
            c = code;
            continue;
        }
        const bmodel* bm = &w->bm[-1 - code];
        const double* nw = &pl[4 * fmin];
        if (bm->kind == UGX_BND_DELETION) { *cellp = CELL_DEAD; return 0; }
        *hits += 1;
        const int wf = meas ? w->wface[(size_t)c * MAXF + fmin] : -1;
        const double pre[3] = {U[0], U[1], U[2]};
        const double e0 = *erot;
        if (bm->kind == UGX_BND_SPECULAR) {
            /* uniGasPatchBoundary::specularReflection, :390-403 */
            double udn = U[0] * nw[0] + U[1] * nw[1] + U[2] * nw[2];
            if (udn > 0.0)
                for (int d = 0; d < 3; ++d) U[d] = U[d] - 2.0 * udn * nw[d];
        } else {
            diffuse_reflect(r, nw, U, erot, bm->T, bm->U, w->mass[sp], w->rotdof[sp]);
        }
        if (wf >= 0) wall_event(w, wf, sp, slot, it, pre, U, e0, *erot);
        for (int d = 0; d < 3; ++d) vt[d] = U[d] * w->mask[d];
    }
    *errs += 1;
    *cellp = c;
    return 1;
}

/* ---------------------------------------------------------------- inflow */

/* uniGasGeneralBoundary::insertParcels, uniGasGeneralBoundary.C:537-761:
 * position on the face, Bird eq. 12.5 normal speed, tangential Maxwellian,
 * equipartition E_rot; then the first move over (1 - R) deltaT
 * (newParcel, uniGasParcel.C:47-53). */
static int make_inflow(orc_world* w, const inflow_fs* e, orc_rng* r, double x[3], double U[3],
                       double* erot, int32_t* cell, int64_t slot, int64_t* hits, int64_t* errs)
{
    double us = orc_uniform(r);
    int tri = 0;
    for (int t = 0; t < e->nv - 2; ++t) { tri = t; if (e->cum[t] >= us) break; }
    double s = sqrt(orc_uniform(r));
    double tt = orc_uniform(r);
    const double* a = e->vloc[0];
    const double* b = e->vloc[tri + 1];
    const double* c = e->vloc[tri + 2];
    for (int d = 0; d < 3; ++d) x[d] = a[d] * (1.0 - s) + b[d] * (s * (1.0 - tt)) + c[d] * (s * tt);
    double m = w->mass[e->species];
    double mps = sqrt(2.0 * KB * e->T / m);
    double udn = dot3(e->U, e->n);
    double sCos = udn / mps;
    double uNormal;
    if (fabs(udn) > VSMALL) {
        double cA = sCos + sqrt(sCos * sCos + 2.0);
        double cB = 0.5 * (1.0 + sCos * (sCos - sqrt(sCos * sCos + 2.0)));
        double scale = 3.0;
        if (sCos < -3.0) scale = fabs(sCos) + 1.0;
        uNormal = 0.0;
        for (int it = 0; it < 100000; ++it) {
            double th = scale * (2.0 * orc_uniform(r) - 1.0);
            uNormal = th + sCos;
            double P = (uNormal < 0.0) ? -1.0 : 2.0 * uNormal / cA * orc_exp(cB - th * th);
            if (!(P < orc_uniform(r))) break;
        }
    } else {
        uNormal = sqrt(-orc_log(orc_uniform(r)));
    }
    double g1, g2;
    orc_gauss2(r, &g1, &g2);
    double sig = sqrt(KB * e->T / m);
    double ut1 = dot3(e->t1, e->U), ut2 = dot3(e->t2, e->U);
    for (int d = 0; d < 3; ++d)
        U[d] = ((sig * (g1 * e->t1[d] + g2 * e->t2[d]) + ut1 * e->t1[d]) + ut2 * e->t2[d]) + mps * uNormal * e->n[d];
    *erot = equip_erot(r, e->rotT, w->rotdof[e->species]);
    *cell = e->cell;
    double frac = orc_uniform(r);
    return track(w, cell, x, U, erot, e->species, slot, (1.0 - frac) * w->model.delta_t, r, hits, errs);
}

/* ------------------------------------------------------------ collisions */

/* sigma_T c_r for VHS/VSS/LB (variableHardSphere.C:72-115):
 * pi d^2 (2 k Tref/(mR cR^2))^(w-1/2) / Gamma(5/2-w) cR = C_pq (cR^2)^(1-w). */
static double sigma_cr(const orc_world* w, int p, int q, const double* Up, const double* Uq)
{
    double d0 = Up[0] - Uq[0], d1 = Up[1] - Uq[1], d2 = Up[2] - Uq[2];
    double cR2 = d0 * d0 + d1 * d1 + d2 * d2;
    if (cR2 < VSMALL) return 0.0;
    int k = p * w->ns + q;
    return w->pC[k] * orc_exp(w->pe[k] * orc_log(cR2));
}

/* Isotropic (VHS) or VSS scattering of the relative velocity with speed cR,
 * then the centre-of-mass split (variableHardSphere.C:118-162,
 * variableSoftSphere.C:116-176, LB-VSS :389-424). */
static void scatter(orc_rng* r, int vss, double alpha, double cR, const double crc[3],
                    double mP, double mQ, double* UP, double* UQ)
{
    double Ucm[3];
    for (int d = 0; d < 3; ++d) Ucm[d] = (mP * UP[d] + mQ * UQ[d]) / (mP + mQ);
    double cosT = vss ? 2.0 * orc_pow(orc_uniform(r), 1.0 / alpha) - 1.0 : 2.0 * orc_uniform(r) - 1.0;
    double sinT = sqrt(1.0 - cosT * cosT);
    double sp, cp;
    orc_sincos2pi(orc_uniform(r), &sp, &cp);
    double c[3];
    double D = sqrt(crc[1] * crc[1] + crc[2] * crc[2]);
    if (vss && D > 0.0) {
        c[0] = cosT * crc[0] + sinT * sp * D;
        c[1] = cosT * crc[1] + sinT * (cR * crc[2] * cp - crc[0] * crc[1] * sp) / D;
        c[2] = cosT * crc[2] - sinT * (cR * crc[1] * cp + crc[0] * crc[2] * sp) / D;
    } else {
        c[0] = cR * cosT;
        c[1] = cR * (sinT * cp);
        c[2] = cR * (sinT * sp);
    }
    for (int d = 0; d < 3; ++d) {
        UP[d] = Ucm[d] + c[d] * mQ / (mP + mQ);
        UQ[d] = Ucm[d] - c[d] * mP / (mP + mQ);
    }
}

/* dsmcCollisionModel::collide for VHS, VSS and the serial Larsen-Borgnakke
 * variants (LarsenBorgnakkeVariableHardSphere.C:125-416). */
static void binary_collide(orc_world* w, orc_rng* r, int p, int q, double* UP, double* UQ,
                           double* eP, double* eQ)
{
    int model = w->model.dsmc_model;
    int k = p * w->ns + q;
    double mP = w->mass[p], mQ = w->mass[q];
    double crc[3] = {UP[0] - UQ[0], UP[1] - UQ[1], UP[2] - UQ[2]};
    double cR2 = crc[0] * crc[0] + crc[1] * crc[1] + crc[2] * crc[2];
    double cR = sqrt(cR2);
    int vss = (model == UGX_DSMC_VSS || model == UGX_DSMC_LB_VSS);
    if (model == UGX_DSMC_LB_VHS || model == UGX_DSMC_LB_VSS) {
        double mR = w->pmr[k];
        double Et = 0.5 * mR * cR2;
        double chiB = w->pchi[k];
        double invZ = 1.0 / w->model.Zrot;
        double preP = *eP, preQ = *eQ;
        if (w->rotdof[p] > 0 && invZ > orc_uniform(r)) {
            double Ec = Et + preP;
            double ratio = post_rot_ratio(r, w->rotdof[p], chiB);
            *eP = ratio * Ec;
            Et = Ec - *eP;
        }
        if (w->rotdof[q] > 0 && invZ > orc_uniform(r)) {
            double Ec = Et + preQ;
            double ratio = post_rot_ratio(r, w->rotdof[q], chiB);
            *eQ = ratio * Ec;
            Et = Ec - *eQ;
        }
        double cRn = sqrt(2.0 * Et / mR);
        if (vss && cR > 0.0)
            for (int d = 0; d < 3; ++d) crc[d] = crc[d] * (cRn / cR);
        cR = cRn;
    }
    scatter(r, vss, w->palpha[k], cR, crc, mP, mQ, UP, UQ);
}

/* Canonical per-cell sum of n values in sorted order: eight partial sums,
 * partial j adding v[j], v[j+8], v[j+16], ... in order to +0.0, combined by
 * the tree ((P0+P1)+(P2+P3)) + ((P4+P5)+(P6+P7)).  This is the order of an
 * eight-lane group reduction on the GPU.  The reference sums in list order
 * (cellMeasurements.C:417-507); any fixed order is an equally valid sum. */
/* Timing mode for the CPU baseline: plain sequential sums (the canonical
 * order only matters for bit parity with the GPU, not for the physics). */
static int g_fast_sums = 0;
void orc_set_fast_sums(int on) { g_fast_sums = on; }

static double canon_sum(const double* v, int64_t n)
{
    if (g_fast_sums) {
        double acc = 0.0;
        for (int64_t i = 0; i < n; ++i) acc += v[i];
        return acc;
    }
    double P[8];
    for (int j = 0; j < 8; ++j) {
        double s = 0.0;
        for (int64_t i = j; i < n; i += 8) s = s + v[i];
        P[j] = s;
    }
    return ((P[0] + P[1]) + (P[2] + P[3])) + ((P[4] + P[5]) + (P[6] + P[7]));
}

/* Per-cell sums of cellMeasurements::calculateFields (cellMeasurements.C:408-513):
 * the per-simulator contributions of each row, reduced with canon_sum. */
static void cell_sums(const orc_world* w, int64_t b, int64_t e, double S[NSAMP])
{
    int64_t n = e - b;
    double* v = malloc(sizeof(double) * (size_t)(n > 0 ? n : 1) * NSAMP);
    for (int64_t i = 0; i < n; ++i) {
        int s = w->sp[b + i];
        double m = w->mass[s];
        double vx = w->v[3 * (b + i)], vy = w->v[3 * (b + i) + 1], vz = w->v[3 * (b + i) + 2];
        double er = w->erot[b + i];
        double mcc = m * (vx * vx + vy * vy + vz * vz);
        double* r = &v[(size_t)i * NSAMP];
        r[UGX_S_N] = 1.0;
        r[UGX_S_M] = m;
        r[UGX_S_MU] = m * vx;
        r[UGX_S_MV] = m * vy;
        r[UGX_S_MW] = m * vz;
        r[UGX_S_MCC] = mcc;
        r[UGX_S_EROT] = er;
        r[UGX_S_ROTDOF] = (double)w->rotdof[s];
        r[UGX_S_MUU] = m * (vx * vx);
        r[UGX_S_MUV] = m * (vx * vy);
        r[UGX_S_MUW] = m * (vx * vz);
        r[UGX_S_MVV] = m * (vy * vy);
        r[UGX_S_MVW] = m * (vy * vz);
        r[UGX_S_MWW] = m * (vz * vz);
        r[UGX_S_MCCU] = mcc * vx;
        r[UGX_S_MCCV] = mcc * vy;
        r[UGX_S_MCCW] = mcc * vz;
        r[UGX_S_EROTU] = er * vx;
        r[UGX_S_EROTV] = er * vy;
        r[UGX_S_EROTW] = er * vz;
        for (int q = 0; q < NS; ++q) {
            r[UGX_S_SPECIES_N + q] = (s == q) ? 1.0 : 0.0;
            r[UGX_S_SPECIES_MCC + q] = (s == q) ? mcc : 0.0;
        }
    }
    double* col = malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    for (int k = 0; k < NSAMP; ++k) {
        for (int64_t i = 0; i < n; ++i) col[i] = v[(size_t)i * NSAMP + k];
        S[k] = canon_sum(col, n);
    }
    free(col);
    free(v);
}

/* Subcell of a simulator along one axis (noTimeCounter.C:141-155): the
 * virtual Cartesian subcells split the cell's bounding box into lv equal
 * slabs; x and lo are in the cell frame.  The index is truncated like the
 * reference's label() and kept inside [0, lv - 1] (a simulator on the box's
 * upper face would otherwise index one slab past the last). */
static int subcell_axis(double x, double lo, double len, int lv)
{
    double t = (double)lv * (x - lo) / len;
    int k = (int)t;
    if (k < 0) k = 0;
    if (k > lv - 1) k = lv - 1;
    return k;
}

/* Subcell lists of a cell (noTimeCounter.C:133-161): sid[i] the subcell of
 * the cell's i-th simulator, start[s]..start[s+1] the range of list[] that
 * holds subcell s's simulators in ascending order (the reference appends
 * them to its per-subcell lists in cell order).  The subcell index weights
 * the axes x, y, z by 1, lx, lx*ly; levels of empty directions are 1. */
static void subcell_lists(const orc_world* w, int c, int64_t b, int64_t nC, const int lv[3], int nsub,
                          int32_t* sid, int32_t* start, int32_t* list)
{
    double lo[3], len[3];
    for (int d = 0; d < 3; ++d) { lo[d] = w->bblo[3 * c + d]; len[d] = w->bbhi[3 * c + d] - lo[d]; }
    for (int s = 0; s <= nsub; ++s) start[s] = 0;
    for (int64_t i = 0; i < nC; ++i) {
        const double* x = &w->x[3 * (b + i)];
        int k0 = subcell_axis(x[0], lo[0], len[0], lv[0]);
        int k1 = subcell_axis(x[1], lo[1], len[1], lv[1]);
        int k2 = subcell_axis(x[2], lo[2], len[2], lv[2]);
        int s = k0 + lv[0] * (k1 + lv[1] * k2);
        sid[i] = s;
        start[s + 1] += 1;
    }
    for (int s = 1; s <= nsub; ++s) start[s] += start[s - 1];
    int32_t* cur = malloc(sizeof(int32_t) * (size_t)nsub);
    for (int s = 0; s < nsub; ++s) cur[s] = start[s];
    for (int64_t i = 0; i < nC; ++i) list[cur[sid[i]]++] = (int32_t)i;
    free(cur);
}

/* noTimeCounter::collide for one cell (noTimeCounter.C:83-312).  With more
 * than one subcell the second candidate comes from the first's subcell
 * when that holds two or more simulators, else from the whole cell
 * (:193-225). */
static void ntc_cell(orc_world* w, int c, int64_t b, int64_t nC, int64_t* cand, int64_t* coll)
{
    if (nC < 2) return;
    const int lv[3] = {w->sclev[3 * c], w->sclev[3 * c + 1], w->sclev[3 * c + 2]};
    const int nsub = lv[0] * lv[1] * lv[2];
    int32_t *sid = NULL, *start = NULL, *list = NULL;
    if (nsub > 1) {
        sid = malloc(sizeof(int32_t) * (size_t)nC);
        list = malloc(sizeof(int32_t) * (size_t)nC);
        start = malloc(sizeof(int32_t) * (size_t)(nsub + 1));
        subcell_lists(w, c, b, nC, lv, nsub, sid, start, list);
    }
    orc_rng r;
    orc_rng_init(&r, w->seed, ORC_RNG_NTC, 0u, (uint32_t)c, w->step);
    double smax0 = w->smax[c];
    double smax = smax0;
    double sel = 0.5 * (double)nC * (double)(nC - 1) * w->model.n_equivalent_particles * w->cw[c] * smax0
                 * w->model.delta_t / w->vol[c];
    double fl = floor(sel);
    int64_t ncand = (int64_t)fl;
    if (orc_uniform(&r) < sel - fl) ncand += 1;
    *cand += ncand;
    for (int64_t k = 0; k < ncand; ++k) {
        uint32_t P = orc_index(&r, (uint32_t)nC);
        uint32_t Q;
        uint32_t nSC = 0, s0 = 0;
        if (nsub > 1) { s0 = (uint32_t)start[sid[P]]; nSC = (uint32_t)start[sid[P] + 1] - s0; }
        if (nSC > 1) {
            do { Q = (uint32_t)list[s0 + orc_index(&r, nSC)]; } while (Q == P);
        } else {
            do { Q = orc_index(&r, (uint32_t)nC); } while (Q == P);
        }
        int64_t ip = b + P, iq = b + Q;
        int sp = w->sp[ip], sq = w->sp[iq];
        double* UP = &w->v[3 * ip];
        double* UQ = &w->v[3 * iq];
        double scr = sigma_cr(w, sp, sq, UP, UQ);
        if (scr > smax) smax = scr;
        if (scr / smax0 > orc_uniform(&r)) {
            binary_collide(w, &r, sp, sq, UP, UQ, &w->erot[ip], &w->erot[iq]);
            *coll += 1;
        }
    }
    w->smax[c] = smax;
    free(sid);
    free(start);
    free(list);
}

/* unifiedStochasticParticleSBGK::samplePostCollisionVelocity
 * (unifiedStochasticParticleSBGK.C:1047-1096): xi ~ N(0, I/2) accepted with
 * P(xi) = 1 + cQ/(p u0) (q.xi)(|xi|^2/2.5 - 1) + cS/p sigma:(xi xi - |xi|^2/3 I)
 * against the cell's maxProb m0; a draw with m0 < P < 10 is taken at once and
 * raises maxProb (largest such P in *raise).  sh: xx xy xz yy yz zz. */
static void usp_draw(orc_rng* rv, double cQ, double cS, double p, double u0, const double q[3], const double sh[6],
                     double m0, double* raise, double xi[3])
{
    /* GaussNormal()/sqrt(2), |xi|^2/3, /2.5 and prob/maxProb of the reference
     * as products with the reciprocals, and the constants of P divided out
     * once: a restatement that changes the last bit of some draws only */
    const double A = cQ / (p * u0), B = cS / p, inv_m0 = 1.0 / m0;
    for (int it = 0; it < 100000; ++it) {
        double g[3];
        orc_gauss3(rv, g);
        for (int d = 0; d < 3; ++d) xi[d] = g[d] * ORC_SQRT_HALF;
        double vSq = xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2];
        double vTr = vSq * ORC_THIRD;
        double prob = 1.0 + A * (q[0] * xi[0] + q[1] * xi[1] + q[2] * xi[2]) * (vSq * 0.4 - 1.0)
                      + B * (sh[0] * (xi[0] * xi[0] - vTr) + sh[3] * (xi[1] * xi[1] - vTr)
                             + sh[5] * (xi[2] * xi[2] - vTr) + 2.0 * sh[1] * xi[0] * xi[1]
                             + 2.0 * sh[2] * xi[0] * xi[2] + 2.0 * sh[4] * xi[1] * xi[2]);
        if (prob > m0 && prob < 10.0) { if (prob > *raise) *raise = prob; break; }
        if (orc_uniform(rv) < prob * inv_m0) break;
    }
}

/* Test hook: n USP draws (streams {i, 0, 0} of the BGK velocity kind). */
int orc_usp_sample(double cQ, double cS, double p, double u0, const double* q, const double* sh, double m0,
                   int64_t n, uint64_t seed, double* xi, double* raise)
{
    *raise = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        orc_rng rv;
        orc_rng_init(&rv, seed, ORC_RNG_BGK_V, (uint32_t)i, 0u, 0u);
        usp_draw(&rv, cQ, cS, p, u0, q, sh, m0, raise, &xi[3 * i]);
    }
    return UGX_OK;
}

/* BGK-family relaxation for one cell: calculateProperties
 * (stochasticParticleBGK.C:287-583, stochasticParticleESBGK.C:311-665,
 * unifiedStochasticParticleSBGK.C:379-800), collide (:628-729, :721-834,
 * :871-994), conserveMomentumAndEnergy (:731-780) and the post-collision
 * samplers (:782-796, :887-906, :1047-1096).
 * Restated for parallel sampling, with the same distributions:
 *  - the reference's five list shuffles followed by "take the first n" are
 *    one partial Fisher-Yates draw of n distinct simulators (cell stream);
 *  - each selected simulator draws its new velocity from its own stream
 *    (ORC_RNG_BGK_V, sub = selection index);
 *  - USP: every draw of the step is tested against the cell's maxProb at the
 *    start of the step; a draw with maxProb < P < 10 is accepted and the
 *    largest such P becomes the new maxProb (the value the reference's
 *    sequential raises end on); without a raise maxProb decays by 0.999. */
/* ---------------------------------------------------- macroInterpolation
 * interpolationCellPoint (OpenFOAM) of the BGK-family fields at a
 * simulator's position: linear on the tets (cell centroid, then per face in
 * ascending order the fan (p0, p_i, p_i+1)), the cell's value at the
 * centroid and point values at the points. */

static void mi_interp(const orc_world* w, int c, const double x[3], int nr, double* out)
{
    const double* ref = &w->cc[3 * c];
    double a[3], r[3];
    for (int d = 0; d < 3; ++d) { a[d] = w->mi_cent[3 * c + d]; r[d] = x[d] - a[d]; }
    double best = -1.0e300, bl[4] = {1.0, 0.0, 0.0, 0.0};
    int32_t bid[3] = {-1, -1, -1};
    int found = 0;
    for (int j = 0; j < MAXF && !found; ++j) {
        const int32_t* fp = &w->mi_fp[((size_t)c * MAXF + j) * MI_FV];
        if (fp[0] < 0) break;
        int nv = 3;
        while (nv < MI_FV && fp[nv] >= 0) ++nv;
        double e1[3];
        for (int d = 0; d < 3; ++d) e1[d] = (w->mi_pts[3 * (size_t)fp[0] + d] - ref[d]) - a[d];
        for (int t = 1; t + 1 < nv; ++t) {
            double e2[3], e3[3];
            for (int d = 0; d < 3; ++d) {
                e2[d] = (w->mi_pts[3 * (size_t)fp[t] + d] - ref[d]) - a[d];
                e3[d] = (w->mi_pts[3 * (size_t)fp[t + 1] + d] - ref[d]) - a[d];
            }
            double n23[3] = {e2[1] * e3[2] - e2[2] * e3[1], e2[2] * e3[0] - e2[0] * e3[2], e2[0] * e3[1] - e2[1] * e3[0]};
            double det = e1[0] * n23[0] + e1[1] * n23[1] + e1[2] * n23[2];
            if (det == 0.0) continue;
            double rx3[3] = {r[1] * e3[2] - r[2] * e3[1], r[2] * e3[0] - r[0] * e3[2], r[0] * e3[1] - r[1] * e3[0]};
            double e2r[3] = {e2[1] * r[2] - e2[2] * r[1], e2[2] * r[0] - e2[0] * r[2], e2[0] * r[1] - e2[1] * r[0]};
            double l1 = (r[0] * n23[0] + r[1] * n23[1] + r[2] * n23[2]) / det;
            double l2 = (e1[0] * rx3[0] + e1[1] * rx3[1] + e1[2] * rx3[2]) / det;
            double l3 = (e1[0] * e2r[0] + e1[1] * e2r[1] + e1[2] * e2r[2]) / det;
            double l0 = ((1.0 - l1) - l2) - l3;
            double mn = fmin(fmin(l0, l1), fmin(l2, l3));
            if (mn > best) {
                best = mn;
                bl[0] = l0; bl[1] = l1; bl[2] = l2; bl[3] = l3;
                bid[0] = fp[0]; bid[1] = fp[t]; bid[2] = fp[t + 1];
            }
            if (mn >= -MI_TOL) { found = 1; break; }
        }
    }
    const double* fc = &w->mfield[(size_t)MIR * c];
    for (int q = 0; q < nr; ++q) {
        double v = bl[0] * fc[q];
        for (int k = 0; k < 3; ++k)
            if (bid[k] >= 0) v = v + bl[k + 1] * w->pfield[(size_t)MIR * bid[k] + q];
        out[q] = v;
    }
}

/* volPointInterpolation of the step's cell fields to the points. */
static void mi_point_fields(orc_world* w)
{
    for (int p = 0; p < w->npts; ++p) {
        double v[MIR];
        for (int q = 0; q < MIR; ++q) v[q] = 0.0;
        for (uint32_t k = w->mi_poff[p]; k < w->mi_poff[p + 1]; ++k) {
            double wt = w->mi_pw[k];
            const double* f = &w->mfield[(size_t)MIR * w->mi_pcell[k]];
            for (int q = 0; q < MIR; ++q) v[q] = v[q] + wt * f[q];
        }
        for (int q = 0; q < MIR; ++q) w->pfield[(size_t)MIR * p + q] = v[q];
    }
}

static void bgk_cell(orc_world* w, int c, int64_t b, int64_t nC, const double S[NSAMP],
                     uint32_t* idx, int64_t* nrel_out, int update_state, int fields_only)
{
    int model = w->model.bgk_model;
    double FN = w->model.n_equivalent_particles * w->cw[c], dt = w->model.delta_t, V = w->vol[c];
    double N = S[UGX_S_N];
    int usp = (model == UGX_BGK_USP);
    int perform = 1;
    double T = 0.0, p = 0.0, U[3] = {0.0, 0.0, 0.0}, Pt[6] = {0, 0, 0, 0, 0, 0};
    double q[3], sh[6];
    for (int d = 0; d < 3; ++d) q[d] = w->uspq[3 * c + d];
    for (int d = 0; d < 6; ++d) sh[d] = w->usps[6 * c + d];
    double Pr = 0.0, nu = 0.0;
    if (N > VSMALL) {
        double rhoN = FN * N / V;
        for (int d = 0; d < 3; ++d) U[d] = S[UGX_S_MU + d] / S[UGX_S_M];
        double uu = U[0] * U[0] + U[1] * U[1] + U[2] * U[2];
        T = (S[UGX_S_MCC] - S[UGX_S_M] * uu) / (3.0 * KB * N);
        p = rhoN * KB * T;
        const int ij[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
        for (int k = 0; k < 6; ++k)
            Pt[k] = rhoN * (S[UGX_S_MUU + k] / N - (S[UGX_S_M] / N) * U[ij[k][0]] * U[ij[k][1]]);
        if (usp) {
            double tr = (Pt[0] + Pt[3] + Pt[5]) / 3.0;
            for (int k = 0; k < 6; ++k) sh[k] = Pt[k];
            sh[0] = sh[0] - tr; sh[3] = sh[3] - tr; sh[5] = sh[5] - tr;
            const int row[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
            for (int d = 0; d < 3; ++d)
                q[d] = rhoN * (0.5 * (S[UGX_S_MCCU + d] / N) - 0.5 * (S[UGX_S_MCC] / N) * U[d]
                               + S[UGX_S_EROTU + d] / N - (S[UGX_S_EROT] / N) * U[d])
                       - Pt[row[d][0]] * U[0] - Pt[row[d][1]] * U[1] - Pt[row[d][2]] * U[2];
        }
        if (N > (usp ? 2.0 : 1.0)) {
            double f1 = N / (N - 1.0);
            p = f1 * p;
            T = f1 * T;
            for (int k = 0; k < 6; ++k) Pt[k] = f1 * Pt[k];
            if (usp) {
                double f2 = N * N / (N - 1.0) / (N - 2.0);
                for (int d = 0; d < 3; ++d) q[d] = f2 * q[d];
                for (int k = 0; k < 6; ++k) sh[k] = f1 * sh[k];
            }
        } else {
            perform = 0;
        }
        if (T > VSMALL) {
            double visc = 0.0;
            for (int s = 0; s < w->ns; ++s) {
                double ns_ = S[UGX_S_SPECIES_N + s];
                visc = visc + ns_ * w->muref[s] * orc_pow(T / w->model.Tref, w->omega[s]);
                Pr = Pr + ns_ * w->prterm[s];
            }
            visc = visc / N;
            Pr = Pr / N;
            nu = (model == UGX_BGK_ESBGK) ? Pr * p / visc : p / visc;
        } else {
            perform = 0;
        }
    } else {
        perform = 0;
    }
    if (usp && (update_state || fields_only)) {
        double th = w->model.theta;
        for (int d = 0; d < 3; ++d)
            q[d] = th * q[d] / (1.0 + 0.5 * Pr * nu * dt) + (1.0 - th) * w->uspq[3 * c + d];
        for (int k = 0; k < 6; ++k)
            sh[k] = th * sh[k] / (1.0 + 0.5 * nu * dt) + (1.0 - th) * w->usps[6 * c + k];
        if (!fields_only) {
            for (int d = 0; d < 3; ++d) w->uspq[3 * c + d] = q[d];
            for (int k = 0; k < 6; ++k) w->usps[6 * c + k] = sh[k];
        }
    }
    if (fields_only) {
        /* the cell's fields for the cellPoint interpolation (every cell) */
        double* mf = &w->mfield[(size_t)MIR * c];
        mf[0] = T;
        for (int d = 0; d < 3; ++d) mf[1 + d] = U[d];
        mf[4] = Pr;
        mf[5] = nu;
        mf[6] = p;
        for (int d = 0; d < 3; ++d) mf[7 + d] = usp ? q[d] : 0.0;
        for (int k = 0; k < 6; ++k) mf[10 + k] = usp ? sh[k] : Pt[k];
        return;
    }
    const double m0 = w->uspmax[c];
    double raise = 0.0;
    if (perform && w->cmodel[c] == 0) {
        orc_rng r;
        orc_rng_init(&r, w->seed, ORC_RNG_BGK, 0u, (uint32_t)c, w->step);
        double pc = N * (1.0 - orc_exp(-nu * dt));
        double fl = floor(pc);
        int64_t nrel = (int64_t)fl;
        if (orc_uniform(&r) < pc - fl) nrel += 1;
        if (nrel > nC) nrel = nC;
        *nrel_out += nrel;
        for (int64_t i = 0; i < nC; ++i) idx[i] = (uint32_t)i;
        for (int64_t i = 0; i < nrel; ++i) {
            uint32_t j = (uint32_t)i + orc_index(&r, (uint32_t)(nC - i));
            uint32_t t = idx[i]; idx[i] = idx[j]; idx[j] = t;
        }
        double Smat[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (model == UGX_BGK_ESBGK) {
            double a = 0.5 * (1.0 - Pr) / Pr;
            const int full[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
            for (int k = 0; k < 9; ++k) {
                double dlt = (k == 0 || k == 4 || k == 8) ? 1.0 : 0.0;
                Smat[k] = dlt - a * (Pt[full[k]] / p - dlt);
            }
        }
        double tau = 0.5 * nu * dt;
        double ex = 1.0 + 2.0 / (orc_exp(2.0 * tau) - 1.0);
        double cQ = 2.0 * (1.0 - Pr * tau * ex);
        double cS = 1.0 - tau * ex;
        double s2kT = sqrt(2.0 * KB * T);
        for (int64_t i = 0; i < nrel; ++i) {
            int64_t pi = b + idx[i];
            orc_rng rv;
            orc_rng_init(&rv, w->seed, ORC_RNG_BGK_V, (uint32_t)i, (uint32_t)c, w->step);
            /* the draw's parameters: the cell's, or interpolated at the
             * simulator (stochasticParticleBGK.C:680-687,
             * stochasticParticleESBGK.C:779-789, unifiedStochasticParticleSBGK.C:933-946) */
            double dU[3] = {U[0], U[1], U[2]}, dS2kT = s2kT, dp = p, dcQ = cQ, dcS = cS, dq[3] = {q[0], q[1], q[2]};
            double dsh[6], dSm[9];
            for (int k = 0; k < 6; ++k) dsh[k] = sh[k];
            for (int k = 0; k < 9; ++k) dSm[k] = Smat[k];
            if (w->macro) {
                double f[MIR];
                mi_interp(w, c, &w->x[3 * pi], model == UGX_BGK_BGK ? 4 : MIR, f);
                for (int d = 0; d < 3; ++d) dU[d] = f[1 + d];
                dS2kT = sqrt(2.0 * KB * f[0]);
                if (model == UGX_BGK_ESBGK) {
                    double iPr = f[4];
                    dp = f[6];
                    double a = 0.5 * (1.0 - iPr) / iPr;
                    const int full[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
                    for (int k = 0; k < 9; ++k) {
                        double dlt = (k == 0 || k == 4 || k == 8) ? 1.0 : 0.0;
                        dSm[k] = dlt - a * (f[10 + full[k]] / dp - dlt);
                    }
                } else if (model == UGX_BGK_USP) {
                    double iPr = f[4], inu = f[5];
                    dp = f[6];
                    for (int d = 0; d < 3; ++d) dq[d] = f[7 + d];
                    for (int k = 0; k < 6; ++k) dsh[k] = f[10 + k];
                    double itau = 0.5 * inu * dt;
                    double iex = 1.0 + 2.0 / (orc_exp(2.0 * itau) - 1.0);
                    dcQ = 2.0 * (1.0 - iPr * itau * iex);
                    dcS = 1.0 - itau * iex;
                }
            }
            double u0 = dS2kT / sqrt(w->mass[w->sp[pi]]);
            double xi[3];
            if (model == UGX_BGK_USP) {
                usp_draw(&rv, dcQ, dcS, dp, u0, dq, dsh, m0, &raise, xi);
            } else {
                double g[3];
                orc_gauss3(&rv, g);
                for (int d = 0; d < 3; ++d) xi[d] = g[d] * ORC_SQRT_HALF;
                if (model == UGX_BGK_ESBGK) {
                    double y[3];
                    for (int d = 0; d < 3; ++d) y[d] = dSm[3 * d] * xi[0] + dSm[3 * d + 1] * xi[1] + dSm[3 * d + 2] * xi[2];
                    for (int d = 0; d < 3; ++d) xi[d] = y[d];
                }
            }
            for (int d = 0; d < 3; ++d) w->v[3 * pi + d] = dU[d] + u0 * xi[d];
        }
        /* conserveMomentumAndEnergy (sums in canonical order) */
        double ke, mom[3];
        {
            double* col = malloc(sizeof(double) * 4 * (size_t)(nC > 0 ? nC : 1));
            for (int64_t i = 0; i < nC; ++i) {
                int64_t j = b + i;
                double m = w->mass[w->sp[j]];
                double vx = w->v[3 * j], vy = w->v[3 * j + 1], vz = w->v[3 * j + 2];
                col[i] = m * (vx * vx + vy * vy + vz * vz);
                col[nC + i] = m * vx;
                col[2 * nC + i] = m * vy;
                col[3 * nC + i] = m * vz;
            }
            ke = canon_sum(col, nC);
            for (int d = 0; d < 3; ++d) mom[d] = canon_sum(col + (d + 1) * nC, nC);
            free(col);
        }
        double pU[3];
        for (int d = 0; d < 3; ++d) pU[d] = mom[d] / S[UGX_S_M];
        double pT = (ke - S[UGX_S_M] * (pU[0] * pU[0] + pU[1] * pU[1] + pU[2] * pU[2])) / (3.0 * KB * (N - 1.0));
        if (pT > VSMALL) {
            double sc = sqrt(T / pT);
            for (int64_t i = b; i < b + nC; ++i)
                for (int d = 0; d < 3; ++d) w->v[3 * i + d] = U[d] + (w->v[3 * i + d] - pU[d]) * sc;
        }
    }
    if (usp && update_state) w->uspmax[c] = (raise > 0.0) ? raise : m0 * 0.999;
}

/* ------------------------------------------------------------ lifecycle */

static void* xcalloc(size_t n, size_t s) { void* p = calloc(n ? n : 1, s); return p; }

int orc_create(const ugx_mesh_desc* mesh, const ugx_species_desc* species, int n_species,
               const ugx_model_desc* model, const ugx_boundary_desc* bnd, int n_bnd,
               uint64_t seed, int64_t capacity, orc_world** out)
{
    if (!mesh || !species || !model || !out || n_species < 1 || n_species > NS)
        return fail(UGX_ERR_ARG, "bad arguments");
    orc_world* w = calloc(1, sizeof(orc_world));
    w->model = *model;
    if (w->model.sample_interval < 1) w->model.sample_interval = 1;
    if (w->model.decomposition_interval < 1) w->model.decomposition_interval = 100;
    w->seed = seed;
    w->ns = n_species;
    for (int s = 0; s < n_species; ++s) {
        w->mass[s] = species[s].mass;
        w->diam[s] = species[s].diameter;
        w->omega[s] = species[s].omega;
        w->alpha[s] = species[s].alpha;
        w->rotdof[s] = species[s].rotational_dof;
        double a = w->alpha[s], om = w->omega[s], d = w->diam[s], m = w->mass[s];
        w->muref[s] = 1.25 * (1.0 + a) * (2.0 + a) * sqrt(m * KB * model->Tref)
                      / (a * (5.0 - 2.0 * om) * (7.0 - 2.0 * om) * sqrt(PI_D) * (d * d));
        double z = (double)w->rotdof[s];
        w->prterm[s] = (5.0 + z) / (7.5 + z);
    }
    for (int p = 0; p < n_species; ++p)
        for (int q = 0; q < n_species; ++q) {
            int k = p * n_species + q;
            double mP = w->mass[p], mQ = w->mass[q];
            double mR = mP * mQ / (mP + mQ);
            double dPQ = 0.5 * (w->diam[p] + w->diam[q]);
            double om = 0.5 * (w->omega[p] + w->omega[q]);
            w->pmr[k] = mR;
            w->pC[k] = PI_D * dPQ * dPQ * pow(2.0 * KB * model->Tref / mR, om - 0.5) / exp(lgamma(2.5 - om));
            w->pe[k] = 1.0 - om;
            w->pchi[k] = 2.5 - om;
            w->palpha[k] = 0.5 * (w->alpha[p] + w->alpha[q]);
        }
    int rc = build_mesh(w, mesh, bnd, n_bnd);
    if (rc) { orc_destroy(w); return rc; }
    /* macroInterpolation reaches the BGK-family draws only */
    w->macro = model->macro_interpolation && model->bgk_model != UGX_BGK_NONE
               && (model->collision_model == UGX_COLL_BGK || model->collision_model == UGX_COLL_HYBRID);
    if (w->macro) {
        rc = build_interp(w, mesh);
        if (rc) { orc_destroy(w); return rc; }
    }
    int nc = w->nc;
    w->cap = capacity > 0 ? capacity : 1024;
    w->x = xcalloc((size_t)w->cap * 3, sizeof(double));
    w->v = xcalloc((size_t)w->cap * 3, sizeof(double));
    w->erot = xcalloc((size_t)w->cap, sizeof(double));
    w->cell = xcalloc((size_t)w->cap, sizeof(int32_t));
    w->sp = xcalloc((size_t)w->cap, sizeof(int32_t));
    w->x2 = xcalloc((size_t)w->cap * 3, sizeof(double));
    w->v2 = xcalloc((size_t)w->cap * 3, sizeof(double));
    w->erot2 = xcalloc((size_t)w->cap, sizeof(double));
    w->cell2 = xcalloc((size_t)w->cap, sizeof(int32_t));
    w->sp2 = xcalloc((size_t)w->cap, sizeof(int32_t));
    w->off = xcalloc((size_t)nc + 1, sizeof(int64_t));
    w->smax = xcalloc((size_t)nc, sizeof(double));
    w->samp = xcalloc((size_t)nc * NSAMP, sizeof(double));
    w->cmodel = xcalloc((size_t)nc, sizeof(int32_t));
    w->uspq = xcalloc((size_t)nc * 3, sizeof(double));
    w->usps = xcalloc((size_t)nc * 6, sizeof(double));
    w->uspmax = xcalloc((size_t)nc, sizeof(double));
    w->knacc = xcalloc((size_t)nc * KNR, sizeof(double));
    w->kn = xcalloc((size_t)nc, sizeof(double));
    w->cw = xcalloc((size_t)nc, sizeof(double));
    for (int c = 0; c < nc; ++c) w->cw[c] = 1.0;
    w->cstart = xcalloc((size_t)w->cap, sizeof(int32_t));
    w->sclev = xcalloc((size_t)nc * 3, sizeof(int32_t));
    for (int c = 0; c < 3 * nc; ++c) w->sclev[c] = 1;
    w->adacc = xcalloc((size_t)nc * KNR, sizeof(double));
    w->adratio = xcalloc((size_t)nc * 3, sizeof(double));
    if (w->model.adaptive) {
        /* adaptiveProperties defaults (uniGasDynamicAdapter.C:198-216) */
        ugx_model_desc* a = &w->model;
        if (a->adaptation_interval == 0) a->adaptation_interval = 20;
        if (a->adaptation_smoothing_passes == 0) a->adaptation_smoothing_passes = 25;
        if (a->adaptation_smoothing_passes < 0) a->adaptation_smoothing_passes = 0;
        if (a->max_time_step_mct_ratio == 0.0) a->max_time_step_mct_ratio = 0.2;
        if (a->max_courant_number == 0.0) a->max_courant_number = 0.5;
        if (a->max_subcell_size_mfp_ratio == 0.0) a->max_subcell_size_mfp_ratio = 0.5;
    }
    int cm = (model->collision_model == UGX_COLL_DSMC) ? 1 : 0;  /* hybrid starts all-BGK (uniGasCloud.C:726-732) */
    for (int c = 0; c < nc; ++c) { w->cmodel[c] = cm; w->uspmax[c] = 1.0; }
    *out = w;
    return UGX_OK;
}

void orc_destroy(orc_world* w)
{
    if (!w) return;
    void* ptrs[] = {w->cc, w->vol, w->bblo, w->bbhi, w->ncf, w->plane, w->nbr, w->delta, w->farea, w->fkind, w->bm, w->inflow,
                    w->x, w->v, w->erot, w->cell, w->sp, w->x2, w->v2, w->erot2, w->cell2, w->sp2,
                    w->off, w->smax, w->samp, w->cmodel, w->uspq, w->usps, w->uspmax, w->knacc, w->kn,
                    w->wface, w->wgeo, w->wsamp, w->ev, w->wcell, w->cw, w->cstart, w->sclev,
                    w->adacc, w->adratio, w->mi_cent, w->mi_fp, w->mi_pts, w->mi_poff, w->mi_pcell, w->mi_pw,
                    w->mfield, w->pfield};
    for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; ++i) free(ptrs[i]);
    free(w);
}

int orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

int64_t orc_count(orc_world* w) { return w->n; }

/* Stable counting sort by cell over [0, n); dead simulators dropped
 * (the cell occupancy of CloudWithModels.C:109-138, as contiguous runs). */
static void sort_by_cell(orc_world* w, int64_t n)
{
    int nc = w->nc;
    int64_t* cnt = w->off;
    memset(cnt, 0, sizeof(int64_t) * ((size_t)nc + 1));
    for (int64_t i = 0; i < n; ++i) if (w->cell[i] >= 0) cnt[w->cell[i] + 1] += 1;
    for (int c = 0; c < nc; ++c) cnt[c + 1] += cnt[c];
    int64_t* cur = malloc(sizeof(int64_t) * (size_t)(nc > 0 ? nc : 1));
    for (int c = 0; c < nc; ++c) cur[c] = cnt[c];
    for (int64_t i = 0; i < n; ++i) {
        int c = w->cell[i];
        if (c < 0) continue;
        int64_t j = cur[c]++;
        for (int d = 0; d < 3; ++d) { w->x2[3 * j + d] = w->x[3 * i + d]; w->v2[3 * j + d] = w->v[3 * i + d]; }
        w->erot2[j] = w->erot[i];
        w->cell2[j] = c;
        w->sp2[j] = w->sp[i];
    }
    free(cur);
    double* t;
    int32_t* ti;
    t = w->x; w->x = w->x2; w->x2 = t;
    t = w->v; w->v = w->v2; w->v2 = t;
    t = w->erot; w->erot = w->erot2; w->erot2 = t;
    ti = w->cell; w->cell = w->cell2; w->cell2 = ti;
    ti = w->sp; w->sp = w->sp2; w->sp2 = ti;
    w->n = cnt[nc];
}

/* The time step from now on, and the freestream counts per step that
 * follow it (uniGasGeneralBoundary.C:154-165). */
static void set_delta_t(orc_world* w, double dt)
{
    w->model.delta_t = dt;
    for (int k = 0; k < w->nin; ++k) {
        inflow_fs* e = &w->inflow[k];
        e->nbar_num = e->flux_an * dt * e->flux_mps * e->flux_br;
        e->nbar = e->nbar_num / e->nbar_den;
    }
}

/* ---------------------------------------------------------- the adapter
 * uniGasDynamicAdapter (uniGasDynamicAdapter.C) of an adaptive case. */

/* calculateAdaptationQuantities (uniGasDynamicAdapter.C:224-314): per cell
 * from the number densities nsp[], temperature T and velocity U, the time
 * step over the mean collision time (Bird eq. 1.38 with 4.74), the Courant
 * numbers max(c_mp, U_d) dt / L_d of the cell's box, and the box size over
 * the mean free path (Bird eq. 4.76-4.77). */
static void adapt_quantities(const orc_world* w, int c, double rhoN, double T, const double U[3], const double* nsp,
                             double dt, const double* prev, double* mct, double co[3], double ratio[3])
{
    int ns = w->ns;
    double Tref = w->model.Tref;
    if (!(T > SMALL)) {
        *mct = 0.0;
        for (int d = 0; d < 3; ++d) { co[d] = 0.0; ratio[d] = prev[d]; }
        return;
    }
    double u0 = 0.0, mcr[NS] = {0.0}, mfp[NS] = {0.0};
    for (int i = 0; i < ns; ++i) {
        for (int q = 0; q < ns; ++q) {
            double dPQ = 0.5 * (w->diam[i] + w->diam[q]);
            double omPQ = 0.5 * (w->omega[i] + w->omega[q]);
            double massRatio = w->mass[i] / w->mass[q];
            if (nsp[q] > VSMALL && T > VSMALL) {
                double mr = w->mass[i] * w->mass[q] / (w->mass[i] + w->mass[q]);
                mcr[i] = mcr[i] + 2.0 * sqrt(PI_D) * dPQ * dPQ * nsp[q] * pow(T / Tref, 1.0 - omPQ) * sqrt(2.0 * KB * Tref / mr);
                mfp[i] = mfp[i] + PI_D * dPQ * dPQ * nsp[q] * pow(Tref / T, omPQ - 0.5) * sqrt(1.0 + massRatio);
            }
        }
        if (nsp[i] > VSMALL) {
            double c0 = sqrt(2.0 * KB / w->mass[i] * T);
            if (c0 > u0) u0 = c0;
        }
    }
    double MCR = 0.0, MFP = 0.0;
    for (int i = 0; i < ns; ++i) {
        if (!(nsp[i] > VSMALL)) continue;
        double lam = 1.0 / mfp[i];
        MCR = MCR + mcr[i] * nsp[i] / rhoN;
        MFP = MFP + lam * nsp[i] / rhoN;
    }
    *mct = dt * MCR;
    for (int d = 0; d < 3; ++d) {
        double L = w->bbhi[3 * c + d] - w->bblo[3 * c + d];
        double v = U[d] > u0 ? U[d] : u0;
        co[d] = v * dt / L;
        ratio[d] = L / MFP;
    }
}

/* calculateTimeStep (uniGasDynamicAdapter.C:316-372): the largest ratio
 * over the DSMC cells (none under the USP model) and the largest Courant
 * number of the solution directions each scale the step to their limit. */
static double adapt_time_step(const orc_world* w, const double* mct, const double* co, double dt)
{
    const ugx_model_desc* m = &w->model;
    int usp = m->collision_model != UGX_COLL_DSMC && m->bgk_model == UGX_BGK_USP;
    double a = 0.0, b = 0.0;
    for (int c = 0; c < w->nc; ++c)
        if (!usp && w->cmodel[c] != 0 && a < mct[c]) a = mct[c];
    for (int c = 0; c < w->nc; ++c)
        for (int d = 0; d < 3; ++d)
            if (w->mask[d] != 0.0 && b < co[3 * c + d]) b = co[3 * c + d];
    if (a > VSMALL && b < VSMALL) return dt * (m->max_time_step_mct_ratio / a);
    if (a < VSMALL && b > VSMALL) return dt * (m->max_courant_number / b);
    if (a > VSMALL && b > VSMALL) {
        double x = m->max_time_step_mct_ratio / a, y = m->max_courant_number / b;
        return dt * (y < x ? y : x);
    }
    return dt;
}

/* calculateSubCellLevels (uniGasDynamicAdapter.C:374-414): a DSMC cell
 * takes ceil(ratio / maxSubCellSizeMFPRatio) clamped to 1..10 per solution
 * direction, any other cell 1. */
static void adapt_levels(orc_world* w, const double* ratio)
{
    double lim = w->model.max_subcell_size_mfp_ratio;
    for (int c = 0; c < w->nc; ++c)
        for (int d = 0; d < 3; ++d) {
            int l = 1;
            if (w->mask[d] != 0.0 && w->cmodel[c] == 1) {
                double x = ceil(ratio[3 * c + d] / lim);
                if (x < 1.0) x = 1.0;
                if (x > (double)ORC_MAX_SUBCELL_LEVELS) x = (double)ORC_MAX_SUBCELL_LEVELS;
                l = (int)x;
            }
            w->sclev[3 * c + d] = l;
        }
}

/* One smoothing pass of an ncomp-component field in place. */
static void smooth_inplace(const orc_world* w, double* f, int ncomp)
{
    double* t = malloc(sizeof(double) * (size_t)w->nc * (size_t)ncomp);
    smooth_field(w, f, t, ncomp);
    memcpy(f, t, sizeof(double) * (size_t)w->nc * (size_t)ncomp);
    free(t);
}

/* setInitialConfiguration with the filled gas (uniGasDynamicAdapter.C:700-770):
 * all three fields smoothed, then the time step and the subcell levels. */
static int adapt_fill(orc_world* w, const ugx_fill_desc* fl)
{
    int nc = w->nc;
    double rhoN = 0.0;
    for (int s = 0; s < w->ns; ++s) rhoN = rhoN + fl->number_density[s];
    double T = (fl->T[0] + fl->T[1] + fl->T[2]) / 3.0;
    double* mct = calloc((size_t)nc, sizeof(double));
    double* co = calloc((size_t)nc * 3, sizeof(double));
    double zero[3] = {0.0, 0.0, 0.0};
    for (int c = 0; c < nc; ++c)
        adapt_quantities(w, c, rhoN, T, fl->velocity, fl->number_density, w->model.delta_t, zero, &mct[c], &co[3 * c],
                         &w->adratio[3 * c]);
    for (int pass = 0; pass < w->model.adaptation_smoothing_passes; ++pass) {
        smooth_inplace(w, mct, 1);
        smooth_inplace(w, co, 3);
        smooth_inplace(w, w->adratio, 3);
    }
    if (w->model.time_step_adaptation) set_delta_t(w, adapt_time_step(w, mct, co, w->model.delta_t));
    if (w->model.subcell_adaptation) adapt_levels(w, w->adratio);
    free(mct);
    free(co);
    return UGX_OK;
}

/* fvc::smooth(W, 1.3) (OpenFOAM fvcSmooth.C, called at
 * uniGasDynamicAdapter.C:456): the smallest field at least W in which no
 * cell is below a face neighbour's value over 2.3, i.e. each cell the
 * largest W_j / 2.3^d over cells j at d faces from it.  OpenFOAM reaches it
 * with a FaceCellWave whose 1% propagation tolerance may leave a cell up to
 * 1% below this fixed point; that tolerance is not restated. */
static void smooth_max_ratio(const orc_world* w, double* f, double max_ratio)
{
    int changed = 1;
    while (changed) {
        changed = 0;
        for (int c = 0; c < w->nc; ++c)
            for (int j = 0; j < MAXF; ++j) {
                int k = c * MAXF + j;
                if (w->fkind[k] != FACE_INTERNAL && w->fkind[k] != FACE_CYCLIC) continue;
                double v = f[w->nbr[k]] / max_ratio;
                if (v > f[c]) { f[c] = v; changed = 1; }
            }
    }
}

/* Cell weight adaptation (uniGasDynamicAdapter.C:416-503, 650-672): new
 * weights n V / (N_pps N_sub F_N) from the averaged densities, the max-ratio
 * smoothing, then area-average passes capped at n V / (N_min N_sub F_N)
 * until neighbour ratios are below 1 + maxCellWeightRatio, blended into the
 * old weights with theta = 0.2; every simulator then goes from its cell's
 * old weight to the new one (uniGasCloud.C:1353-1425) and the cloud is
 * sorted again. */
static int adapt_weights(orc_world* w, const double* rhoN)
{
    const ugx_model_desc* m = &w->model;
    int nc = w->nc;
    double FN = m->n_equivalent_particles;
    int npps = m->particles_per_subcell;
    int nmin = m->min_particles_per_subcell > 0 ? m->min_particles_per_subcell : npps;
    int max_passes = m->max_smoothing_passes > 0 ? m->max_smoothing_passes : 500;
    double max_ratio = m->max_cell_weight_ratio > 0.0 ? m->max_cell_weight_ratio : 0.05;
    const double theta = 0.2;
    double* wn = malloc(sizeof(double) * (size_t)nc);
    double* tmp = malloc(sizeof(double) * (size_t)nc);
    double* nsub = malloc(sizeof(double) * (size_t)nc);
    for (int c = 0; c < nc; ++c)
        nsub[c] = (double)w->sclev[3 * c] * (double)w->sclev[3 * c + 1] * (double)w->sclev[3 * c + 2];
    for (int c = 0; c < nc; ++c) wn[c] = (rhoN[c] * w->vol[c]) / ((double)npps * nsub[c] * FN * 1.0);
    smooth_max_ratio(w, wn, 1.0 + 1.3);
    int passes = 0;
    double max_r;
    do {
        passes++;
        smooth_field(w, wn, tmp, 1);
        memcpy(wn, tmp, sizeof(double) * (size_t)nc);
        for (int c = 0; c < nc; ++c) {
            double lim = (rhoN[c] * w->vol[c]) / ((double)nmin * nsub[c] * FN * 1.0);
            double v = wn[c] < lim ? wn[c] : lim;
            wn[c] = v > SMALL ? v : SMALL;
        }
        max_r = 1.0e-300;
        for (int c = 0; c < nc; ++c)
            for (int j = 0; j < MAXF; ++j) {
                int k = c * MAXF + j;
                if (w->fkind[k] != FACE_INTERNAL) continue;
                int o = w->nbr[k];
                double a = wn[c] / wn[o], b = wn[o] / wn[c];
                double r = a > b ? a : b;
                if (r > max_r) max_r = r;
            }
    } while (max_r > 1.0 + max_ratio && passes < max_passes);
    /* the blend; the simulators go from the old weights to it */
    for (int c = 0; c < nc; ++c) {
        tmp[c] = w->cw[c];
        w->cw[c] = theta * wn[c] + (1.0 - theta) * w->cw[c];
    }
    int64_t ncl = 0;
    int rc = cell_weighting(w, w->n, w->n, &ncl, tmp);
    if (rc == UGX_OK) sort_by_cell(w, w->n + ncl);
    free(wn);
    free(tmp);
    free(nsub);
    return rc;
}

/* adapt (uniGasDynamicAdapter.C:526-700) at the end of every
 * adaptationInterval-th step: the fields averaged over the time since the
 * last adaptation (an empty cell takes the smallest density, split evenly
 * over the species, a cold one the lowest temperature), the size ratio
 * smoothed, then the time step and subcell levels; the sums start again. */
static int adapt(orc_world* w)
{
    int nc = w->nc, ns = w->ns;
    double FN = w->model.n_equivalent_particles, tav = w->adapt_time, dt = w->model.delta_t;
    if (!(tav > 0.0)) return fail(UGX_ERR_STATE, "adaptation without sampled time");
    double* rhoN = calloc((size_t)nc, sizeof(double));
    double* T = calloc((size_t)nc, sizeof(double));
    double* U = calloc((size_t)nc * 3, sizeof(double));
    double* mct = calloc((size_t)nc, sizeof(double));
    double* co = calloc((size_t)nc * 3, sizeof(double));
    double* ratio = calloc((size_t)nc * 3, sizeof(double));
    double minRhoN = 1.0e300, minT = 1.0e300;
    for (int c = 0; c < nc; ++c) {
        const double* a = &w->adacc[KNR * c];
        if (!(a[0] > VSMALL)) continue;
        double FW = FN * w->cw[c], V = w->vol[c];
        rhoN[c] = FW * a[0] / (V * tav);
        double rhoM = FW * a[1] / (V * tav);
        for (int d = 0; d < 3; ++d) U[3 * c + d] = FW * a[2 + d] / (rhoM * V * tav);
        double ke = 0.5 * (FW * a[5]) / (V * tav);
        double rn = FW * a[0] / (V * tav);
        double uu = U[3 * c] * U[3 * c] + U[3 * c + 1] * U[3 * c + 1] + U[3 * c + 2] * U[3 * c + 2];
        T[c] = 2.0 / (3.0 * KB * rn) * (ke - 0.5 * rhoM * uu);
        if (T[c] > VSMALL) {
            if (minRhoN > rhoN[c]) minRhoN = rhoN[c];
            if (minT > T[c]) minT = T[c];
        }
    }
    for (int c = 0; c < nc; ++c) {
        const double* a = &w->adacc[KNR * c];
        double FW = FN * w->cw[c], V = w->vol[c];
        double nsp[NS] = {0.0};
        for (int s = 0; s < ns; ++s) nsp[s] = FW * a[8 + s] / (V * tav);
        if (rhoN[c] == 0.0) {
            rhoN[c] = minRhoN;
            for (int s = 0; s < ns; ++s) nsp[s] = rhoN[c] / (double)ns;
        }
        if (T[c] == 0.0) T[c] = minT;
        adapt_quantities(w, c, rhoN[c], T[c], &U[3 * c], nsp, dt, &w->adratio[3 * c], &mct[c], &co[3 * c], &ratio[3 * c]);
    }
    for (int pass = 0; pass < w->model.adaptation_smoothing_passes; ++pass) smooth_inplace(w, ratio, 3);
    if (w->model.time_step_adaptation) set_delta_t(w, adapt_time_step(w, mct, co, dt));
    if (w->model.subcell_adaptation) adapt_levels(w, ratio);
    int rc = UGX_OK;
    if (w->model.cell_weight_adaptation) rc = adapt_weights(w, rhoN);
    memcpy(w->adratio, ratio, sizeof(double) * (size_t)nc * 3);
    memset(w->adacc, 0, sizeof(double) * (size_t)nc * KNR);
    w->adapt_time = 0.0;
    free(rhoN); free(T); free(U); free(mct); free(co); free(ratio);
    return rc;
}

/* uniGasMeshFill::setInitialConfiguration (uniGasMeshFill.C:58-297): per
 * cell and species a stochastically rounded count (whole cell rather than
 * per tet), positions uniform in the cell (rejection in its bounding box),
 * Maxwellian velocities at per-axis temperatures, equipartition E_rot;
 * (sigma_T c_r)max = pi d^2 c_mp(T) of the most abundant species. */
int orc_fill(orc_world* w, const ugx_fill_desc* fl)
{
    int nc = w->nc, ns = w->ns;
    double FN = w->model.n_equivalent_particles;
    double ntot = 0.0;
    int best = 0;
    for (int s = 0; s < ns; ++s) { ntot += fl->number_density[s]; if (fl->number_density[s] > fl->number_density[best]) best = s; }
    if (w->model.adaptive) {
        /* the time step and subcell levels from the filled gas (uniGasMeshFill.C:100-109) */
        int rc = adapt_fill(w, fl);
        if (rc) return rc;
    }
    if (w->model.cell_weighted) {
        if (fl->exact_per_cell > 0) return fail(UGX_ERR_ARG, "a cell-weighted case is filled by number density, not per cell");
        int rc = fill_weights(w, ntot);
        if (rc) return rc;
    }
    int64_t* cnt = malloc(sizeof(int64_t) * (size_t)nc * ns);
    int64_t total = w->n;
    for (int c = 0; c < nc; ++c)
        for (int s = 0; s < ns; ++s) {
            int64_t k;
            if (fl->exact_per_cell > 0) {
                k = (ntot > 0.0) ? (int64_t)floor((double)fl->exact_per_cell * fl->number_density[s] / ntot + 0.5) : 0;
            } else {
                orc_rng r;
                orc_rng_init(&r, w->seed, ORC_RNG_FILL, (uint32_t)s, (uint32_t)c, 0u);
                double req = fl->number_density[s] / FN * w->vol[c] / w->cw[c];
                double f0 = floor(req);
                k = (int64_t)f0;
                if (orc_uniform(&r) < req - f0) k += 1;
            }
            cnt[c * ns + s] = k;
            total += k;
        }
    if (total > w->cap) { free(cnt); return fail(UGX_ERR_CAPACITY, "particle capacity exceeded by fill"); }
    int64_t j = w->n;
    for (int c = 0; c < nc; ++c)
        for (int s = 0; s < ns; ++s) {
            orc_rng r;
            orc_rng_init(&r, w->seed, ORC_RNG_FILL, (uint32_t)s, (uint32_t)c, 0u);
            if (fl->exact_per_cell <= 0) (void)orc_uniform(&r);
            double m = w->mass[s];
            double sig[3];
            for (int d = 0; d < 3; ++d) sig[d] = sqrt(KB * fl->T[d] / m);
            for (int64_t i = 0; i < cnt[c * ns + s]; ++i, ++j) {
                double x[3] = {0.0, 0.0, 0.0};
                for (int it = 0; it < 1000; ++it) {
                    double y[3];
                    for (int d = 0; d < 3; ++d) {
                        double u = orc_uniform(&r);
                        y[d] = w->bblo[3 * c + d] + u * (w->bbhi[3 * c + d] - w->bblo[3 * c + d]);
                    }
                    int inside = 1;
                    for (int f = 0; f < w->ncf[c]; ++f) {
                        const double* q = &w->plane[((size_t)c * MAXF + f) * 4];
                        if (q[0] * y[0] + q[1] * y[1] + q[2] * y[2] > q[3]) { inside = 0; break; }
                    }
                    if (inside) { x[0] = y[0]; x[1] = y[1]; x[2] = y[2]; break; }
                }
                double g[3];
                orc_gauss3(&r, g);
                for (int d = 0; d < 3; ++d) {
                    w->x[3 * j + d] = x[d] * w->mask[d] + 0.0;
                    w->v[3 * j + d] = sig[d] * g[d] + fl->velocity[d];
                }
                w->erot[j] = equip_erot(&r, fl->rot_temperature, w->rotdof[s]);
                w->cell[j] = c;
                w->sp[j] = s;
            }
        }
    free(cnt);
    sort_by_cell(w, j);
    double Tm = (fl->T[0] + fl->T[1] + fl->T[2]) / 3.0;
    double s0 = PI_D * w->diam[best] * w->diam[best] * sqrt(2.0 * KB * Tm / w->mass[best]);
    for (int c = 0; c < nc; ++c) w->smax[c] = s0;
    return UGX_OK;
}

int orc_upload(orc_world* w, const ugx_particle_soa* p)
{
    if (p->n > w->cap) return fail(UGX_ERR_CAPACITY, "particle capacity exceeded by upload");
    for (int64_t i = 0; i < p->n; ++i) {
        int c = p->cell[i];
        if (c < 0 || c >= w->nc) return fail(UGX_ERR_ARG, "particle cell out of range");
        for (int d = 0; d < 3; ++d) {
            w->x[3 * i + d] = p->local_position ? p->local_position[3 * i + d]
                                                : p->position[3 * i + d] - w->cc[3 * c + d];
            w->v[3 * i + d] = p->velocity[3 * i + d];
        }
        w->erot[i] = p->erot ? p->erot[i] : 0.0;
        w->cell[i] = c;
        w->sp[i] = p->species ? p->species[i] : 0;
    }
    sort_by_cell(w, p->n);
    return UGX_OK;
}

int orc_download(orc_world* w, ugx_particle_soa* p)
{
    if (p->n < w->n) { p->n = w->n; return fail(UGX_ERR_CAPACITY, "download arrays too small"); }
    for (int64_t i = 0; i < w->n; ++i) {
        int c = w->cell[i];
        for (int d = 0; d < 3; ++d) {
            if (p->local_position) p->local_position[3 * i + d] = w->x[3 * i + d];
            if (p->position) p->position[3 * i + d] = w->x[3 * i + d] + w->cc[3 * c + d];
            if (p->velocity) p->velocity[3 * i + d] = w->v[3 * i + d];
        }
        if (p->erot) p->erot[i] = w->erot[i];
        if (p->cell) p->cell[i] = c;
        if (p->species) p->species[i] = w->sp[i];
    }
    p->n = w->n;
    return UGX_OK;
}

/* ------------------------------------------------------------- hybrid */

/* localKnudsen::decompose (hybridDecomposition/derived/localKnudsen/localKnudsen.C:216-582),
 * restated on the cell-face graph: time means since the last decomposition ->
 * smoothing (face-average of neighbours) -> gradient-length Kn -> EMA ->
 * DSMC where Kn > breakdownMax. */
static void hybrid_switch(orc_world* w);

/* -------------------------------------------------------- cell weighting */

/* Cell weights of a weighted fill (uniGasMeshFill.C:111-168), one subcell
 * per cell and no radial weight: W_c = n V_c / (N_pps F_N), then passes of
 * fvc::average(fvc::interpolate(W)) (zeroGradient boundaries) each followed
 * by W_c = max(min(n V_c / (N_min F_N), W_c), SMALL), until no internal face
 * has max(W_o/W_n, W_n/W_o) > 1 + maxCellWeightRatio or maxSmoothingPasses
 * passes were made. */
static int fill_weights(orc_world* w, double ntot)
{
    const ugx_model_desc* m = &w->model;
    double FN = m->n_equivalent_particles;
    int npps = m->particles_per_subcell;
    int nmin = m->min_particles_per_subcell > 0 ? m->min_particles_per_subcell : npps;
    int max_passes = m->max_smoothing_passes > 0 ? m->max_smoothing_passes : 500;
    double max_ratio = m->max_cell_weight_ratio > 0.0 ? m->max_cell_weight_ratio : 0.05;
    if (npps <= 0) return fail(UGX_ERR_ARG, "a cell-weighted case needs particles_per_subcell > 0");
    if (!(ntot > 0.0)) return fail(UGX_ERR_ARG, "a cell-weighted fill needs a number density > 0");
    int nc = w->nc;
    double* tmp = malloc(sizeof(double) * (size_t)nc);
    /* nSubCells of each cell (uniGasMeshFill.C:118-120) */
    double* nsub = malloc(sizeof(double) * (size_t)nc);
    for (int c = 0; c < nc; ++c)
        nsub[c] = (double)w->sclev[3 * c] * (double)w->sclev[3 * c + 1] * (double)w->sclev[3 * c + 2];
    for (int c = 0; c < nc; ++c) w->cw[c] = (ntot * w->vol[c]) / ((double)npps * nsub[c] * FN * 1.0);
    int passes = 0;
    double max_r;
    do {
        passes++;
        smooth_field(w, w->cw, tmp, 1);
        memcpy(w->cw, tmp, sizeof(double) * (size_t)nc);
        for (int c = 0; c < nc; ++c) {
            double lim = (ntot * w->vol[c]) / ((double)nmin * nsub[c] * FN * 1.0);
            double v = w->cw[c] < lim ? w->cw[c] : lim;
            w->cw[c] = v > SMALL ? v : SMALL;
        }
        max_r = 1.0e-300;
        for (int c = 0; c < nc; ++c)
            for (int j = 0; j < MAXF; ++j) {
                int k = c * MAXF + j;
                if (w->fkind[k] != FACE_INTERNAL) continue;
                int o = w->nbr[k];
                double a = w->cw[c] / w->cw[o], b = w->cw[o] / w->cw[c];
                double r = a > b ? a : b;
                if (r > max_r) max_r = r;
            }
    } while (max_r > 1.0 + max_ratio && passes < max_passes);
    free(tmp);
    free(nsub);
    return UGX_OK;
}

/* uniGasCloud::cellWeighting (uniGasCloud.C:1353-1425) over the slots
 * [0, nall) after the move: a simulator whose weight fell from W_o (its
 * start cell's; a new one's insertion cell's) to W_n gets a clone per whole
 * unit of W_o/W_n - 1 and one more with the probability of the rest; one
 * whose weight rose is removed with probability 1 - W_o/W_n.  Draws from a
 * stream per simulator (slot n0 + i of the inflow keyed as 2^31 | i); the
 * clones are copies appended in the order of their parents. */
static int cell_weighting(orc_world* w, int64_t n0, int64_t nall, int64_t* nclone, const double* cw_old)
{
    int32_t* kc = malloc(sizeof(int32_t) * (size_t)(nall > 0 ? nall : 1));
    int64_t tot = 0, wdel = 0;
    for (int64_t i = 0; i < nall; ++i) {
        int32_t k = 0;
        int c = w->cell[i];
        if (c != CELL_DEAD) {
            /* after the move: from the start cell's weight; after an
             * adaptation: from the cell's previous weight (own stream) */
            double wo = cw_old ? cw_old[c] : w->cw[w->cstart[i]], wn = w->cw[c];
            if (wo != wn) {
                uint32_t id = (i < n0) ? (uint32_t)i : (0x80000000u | (uint32_t)(i - n0));
                orc_rng r;
                orc_rng_init(&r, w->seed, ORC_RNG_WEIGHT, cw_old ? 1u : 0u, id, w->step);
                if (wo > wn) {
                    double prob = wo / wn - 1.0;
                    while (prob > 1.0) { k += 1; prob = prob - 1.0; }
                    if (prob > orc_uniform(&r)) k += 1;
                } else if (wo / wn < orc_uniform(&r)) {
                    w->cell[i] = CELL_DEAD;
                    wdel += 1;
                }
            }
        }
        kc[i] = k;
        tot += k;
    }
    if (nall + tot > w->cap) { free(kc); return fail(UGX_ERR_CAPACITY, "particle capacity exceeded by clones"); }
    int64_t j = nall;
    for (int64_t i = 0; i < nall; ++i)
        for (int32_t q = 0; q < kc[i]; ++q, ++j) {
            for (int d = 0; d < 3; ++d) { w->x[3 * j + d] = w->x[3 * i + d]; w->v[3 * j + d] = w->v[3 * i + d]; }
            w->erot[j] = w->erot[i];
            w->cell[j] = w->cell[i];
            w->sp[j] = w->sp[i];
        }
    free(kc);
    w->cnt.cloned += tot;
    w->cnt.weight_deleted += wdel;
    *nclone = tot;
    return UGX_OK;
}

int orc_cell_weights(orc_world* w, double* out, int64_t len)
{
    if (len < w->nc) return fail(UGX_ERR_ARG, "buffer too small");
    memcpy(out, w->cw, sizeof(double) * (size_t)w->nc);
    return UGX_OK;
}

int orc_set_cell_weights(orc_world* w, const double* in, int64_t len)
{
    if (len < w->nc) return fail(UGX_ERR_ARG, "buffer too small");
    for (int c = 0; c < w->nc; ++c)
        if (!(in[c] > 0.0)) return fail(UGX_ERR_ARG, "cell weight not > 0");
    memcpy(w->cw, in, sizeof(double) * (size_t)w->nc);
    return UGX_OK;
}

/* Test hook: the adapter's time sums and summed time set directly, then
 * one adaptation. */
int orc_adapt_inject(orc_world* w, const double* acc, double tav)
{
    memcpy(w->adacc, acc, sizeof(double) * (size_t)w->nc * KNR);
    w->adapt_time = tav;
    return adapt(w);
}

/* Test hook: cell fields [nc*16] set, point values computed, then the
 * cellPoint interpolation at n positions (global coordinates) in the given
 * cells into out [n*16]. */
int orc_mi_probe(orc_world* w, const double* fields, int64_t n, const double* x, const int32_t* cells, double* out)
{
    if (!w->macro) return fail(UGX_ERR_STATE, "macroInterpolation is off");
    memcpy(w->mfield, fields, sizeof(double) * (size_t)w->nc * MIR);
    mi_point_fields(w);
    for (int64_t i = 0; i < n; ++i) {
        int c = cells[i];
        double xl[3];
        for (int d = 0; d < 3; ++d) xl[d] = x[3 * i + d] - w->cc[3 * c + d];
        mi_interp(w, c, xl, MIR, &out[MIR * i]);
    }
    return UGX_OK;
}

int orc_smooth_max_ratio(orc_world* w, double* f, double max_ratio)
{
    smooth_max_ratio(w, f, max_ratio);
    return UGX_OK;
}

int orc_set_cell_model(orc_world* w, const int32_t* in, int64_t len)
{
    if (len < w->nc) return fail(UGX_ERR_ARG, "buffer too small");
    memcpy(w->cmodel, in, sizeof(int32_t) * (size_t)w->nc);
    return UGX_OK;
}

double orc_delta_t(orc_world* w) { return w->model.delta_t; }

double orc_time(orc_world* w) { return w->time; }

int orc_subcell_levels(orc_world* w, int32_t* out, int64_t len)
{
    if (len < 3 * (int64_t)w->nc) return fail(UGX_ERR_ARG, "buffer too small");
    memcpy(out, w->sclev, sizeof(int32_t) * 3 * (size_t)w->nc);
    return UGX_OK;
}

int orc_set_subcell_levels(orc_world* w, const int32_t* in, int64_t len)
{
    if (len < 3 * (int64_t)w->nc) return fail(UGX_ERR_ARG, "buffer too small");
    for (int64_t i = 0; i < 3 * (int64_t)w->nc; ++i) {
        if (in[i] < 1 || in[i] > ORC_MAX_SUBCELL_LEVELS) return fail(UGX_ERR_ARG, "subcell level out of range");
        if (w->mask[i % 3] == 0.0 && in[i] != 1) return fail(UGX_ERR_ARG, "subcell level > 1 in an empty direction");
    }
    memcpy(w->sclev, in, sizeof(int32_t) * 3 * (size_t)w->nc);
    return UGX_OK;
}

/* ------------------------------------------------------------------ step */

int orc_step_phases(orc_world* w, int mask)
{
    int64_t n0 = w->n;
    int64_t nins = 0;
    if (mask & 1) {
        /* inflow: uniGasBoundaries::controlBeforeMove (uniGasBoundaries.C:531-538) */
        int64_t* cnt = malloc(sizeof(int64_t) * (size_t)(w->nin > 0 ? w->nin : 1));
        for (int k = 0; k < w->nin; ++k) {
            orc_rng r;
            orc_rng_init(&r, w->seed, ORC_RNG_INFLOW_N, 0u, (uint32_t)k, w->step);
            /* divided by F_N W_c of the face's cell (uniGasGeneralBoundary.C:154-165) */
            double nb = w->inflow[k].nbar_num / (w->inflow[k].nbar_den * w->cw[w->inflow[k].cell]);
            double fl = floor(nb);
            int64_t c = (int64_t)fl;
            if (c < 0) c = 0;
            if (orc_uniform(&r) < nb - fl) c += 1;
            cnt[k] = c;
            nins += c;
        }
        if (n0 + nins > w->cap) { free(cnt); return fail(UGX_ERR_CAPACITY, "particle capacity exceeded by inflow"); }
        int64_t hits = 0, errs = 0, dead = 0;
        int64_t base = n0;
        for (int k = 0; k < w->nin; ++k) {
            const inflow_fs* e = &w->inflow[k];
            for (int64_t i = 0; i < cnt[k]; ++i) {
                int64_t j = base + i;
                orc_rng r;
                orc_rng_init(&r, w->seed, ORC_RNG_INFLOW_P, (uint32_t)i, (uint32_t)k, w->step);
                w->sp[j] = e->species;
                w->cstart[j] = e->cell;   /* a new simulator carries its insertion cell's weight */
                if (!make_inflow(w, e, &r, &w->x[3 * j], &w->v[3 * j], &w->erot[j], &w->cell[j], j, &hits, &errs)) dead++;
            }
            base += cnt[k];
        }
        free(cnt);
        /* move: Cloud<uniGasParcel>::move -> uniGasParcel::move */
#pragma omp parallel for reduction(+:hits, errs, dead) schedule(static)
        for (int64_t i = 0; i < n0; ++i) {
            orc_rng r;
            orc_rng_init(&r, w->seed, ORC_RNG_MOVE, 0u, (uint32_t)i, w->step);
            w->cstart[i] = w->cell[i];
            if (!track(w, &w->cell[i], &w->x[3 * i], &w->v[3 * i], &w->erot[i], w->sp[i], i,
                       w->model.delta_t, &r, &hits, &errs)) dead++;
        }
        wall_reduce(w);
        w->cnt.inserted += nins;
        w->cnt.deleted += dead;
        w->cnt.wall_hits += hits;
        w->cnt.track_errors += errs;
        if (w->model.cell_weighted) {
            int64_t ncl = 0;
            int rc = cell_weighting(w, n0, n0 + nins, &ncl, NULL);
            if (rc) return rc;
            nins += ncl;
        }
        w->n = n0 + nins;
    }
    if (mask & 2) sort_by_cell(w, w->n);
    if (mask & 4) {
        /* cellMeas_.calculateFields + fields_.calculateFields + collisions()
         * (uniGasCloud.C:846-850, :236-248) */
        int sample = ((w->step + 1) % (uint32_t)w->model.sample_interval) == 0;
        double dt = w->model.delta_t;
        int cm = w->model.collision_model;
        int64_t cand = 0, coll = 0, nrel = 0;
        uint32_t* idx = malloc(sizeof(uint32_t) * (size_t)(w->n > 0 ? w->n : 1));
        int hybrid = (cm == UGX_COLL_HYBRID);
        if (w->macro) {
            /* every cell's fields, then the point values, before any draw
             * (stochasticParticleBGK.C:638-648) */
#pragma omp parallel for schedule(dynamic, 64)
            for (int c = 0; c < w->nc; ++c) {
                int64_t b = w->off[c], nC = w->off[c + 1] - w->off[c];
                double S[NSAMP];
                cell_sums(w, b, b + nC, S);
                int64_t nr0 = 0;
                bgk_cell(w, c, b, nC, S, idx + b, &nr0, 0, 1);
            }
            mi_point_fields(w);
        }
#pragma omp parallel for reduction(+:cand, coll, nrel) schedule(dynamic, 64)
        for (int c = 0; c < w->nc; ++c) {
            int64_t b = w->off[c], nC = w->off[c + 1] - w->off[c];
            double S[NSAMP];
            cell_sums(w, b, b + nC, S);
            if (sample)
                for (int k = 0; k < NSAMP; ++k)
                    if (w->model.sampler_full || k < UGX_S_MUU || k >= UGX_S_BASE)
                        w->samp[(size_t)k * w->nc + c] += dt * S[k];
            if (hybrid) {
                for (int k = 0; k < 8; ++k) w->knacc[KNR * c + k] += dt * S[k];
                for (int k = 0; k < NS; ++k) w->knacc[KNR * c + 8 + k] += dt * S[UGX_S_SPECIES_N + k];
            }
            if (w->model.adaptive) {
                /* the adapter's cell measurements (uniGasDynamicAdapter.C:526-547) */
                for (int k = 0; k < 8; ++k) w->adacc[KNR * c + k] += dt * S[k];
                for (int k = 0; k < NS; ++k) w->adacc[KNR * c + 8 + k] += dt * S[UGX_S_SPECIES_N + k];
            }
            if ((cm == UGX_COLL_DSMC || hybrid) && w->cmodel[c] == 1 && w->model.dsmc_model != UGX_DSMC_NONE)
                ntc_cell(w, c, b, nC, &cand, &coll);
            if ((cm == UGX_COLL_BGK || hybrid) && w->model.bgk_model != UGX_BGK_NONE)
                bgk_cell(w, c, b, nC, S, idx + b, &nrel, 1, 0);
        }
        free(idx);
        if (sample) { w->cnt.n_sampled_steps += 1; w->cnt.time_sampled += dt; }
        if (hybrid) { w->kn_time += dt; w->kn_steps += 1; }
        w->cnt.candidates += cand;
        w->cnt.collisions += coll;
        w->cnt.bgk_collisions += nrel;
    }
    if ((mask & 8) && w->model.collision_model == UGX_COLL_HYBRID
        && ((w->step + 1) % (uint32_t)w->model.decomposition_interval) == 0)
        hybrid_switch(w);
    if (mask & 16) {
        w->step += 1; w->cnt.steps += 1; w->cnt.particle_steps += w->n;
        if (w->model.adaptive) w->adapt_time = w->adapt_time + w->model.delta_t;
        w->time = w->time + w->model.delta_t;
        if (w->model.adaptive && (w->step % (uint32_t)w->model.adaptation_interval) == 0u) {
            int rc = adapt(w);
            if (rc) return rc;
        }
    }
    return UGX_OK;
}

int orc_step(orc_world* w, int n)
{
    for (int i = 0; i < n; ++i) {
        int rc = orc_step_phases(w, 31);
        if (rc) return rc;
    }
    return UGX_OK;
}

/* ------------------------------------------------ hybrid decomposition */

/* fvc::average(fvc::interpolate(f)) of a cell field on the cell-face
 * representation (localKnudsen.C:287-297, 407-417): face values by linear
 * interpolation with the weight w = dN / (dP + dN) of the cell's own value
 * (dP = distance of the face plane from the cell reference point, dN the
 * same for the neighbour, both along the face normal); zeroGradient boundary
 * faces carry the cell value, symmetry faces the cell value less its normal
 * component (vectors), empty faces none; the cell value is the
 * area-weighted mean of its face values.  ncomp = 1 (scalar) or 3. */
static void smooth_field(const orc_world* w, const double* f, double* out, int ncomp)
{
    for (int c = 0; c < w->nc; ++c) {
        double num[3] = {0.0, 0.0, 0.0}, den = 0.0;
        for (int j = 0; j < MAXF; ++j) {
            int k = c * MAXF + j;
            int kind = w->fkind[k];
            if (kind == FACE_SKIP) continue;
            double a = w->farea[k];
            const double* pl = &w->plane[4 * k];
            double fv[3];
            if (kind == FACE_INTERNAL || kind == FACE_CYCLIC) {
                int o = w->nbr[k];
                const double* dl = &w->delta[3 * k];
                double dP = fabs(pl[3]);
                double dN = fabs((pl[0] * dl[0] + pl[1] * dl[1] + pl[2] * dl[2]) + pl[3]);
                double wt = dN / (dP + dN);
                for (int q = 0; q < ncomp; ++q) fv[q] = wt * f[ncomp * c + q] + (1.0 - wt) * f[ncomp * o + q];
            } else if (kind == FACE_SYMMETRY && ncomp == 3) {
                const double* u = &f[3 * c];
                double un = pl[0] * u[0] + pl[1] * u[1] + pl[2] * u[2];
                for (int q = 0; q < 3; ++q) fv[q] = u[q] - un * pl[q];
            } else {
                for (int q = 0; q < ncomp; ++q) fv[q] = f[ncomp * c + q];
            }
            for (int q = 0; q < ncomp; ++q) num[q] = num[q] + a * fv[q];
            den = den + a;
        }
        for (int q = 0; q < ncomp; ++q) out[ncomp * c + q] = (den > 0.0) ? num[q] / den : f[ncomp * c + q];
    }
}

/* Cells within nLevels face-neighbour steps (internal faces only), the cell
 * included: uniGasHybridDecomposition::fetchCellNeighborhood
 * (uniGasHybridDecomposition.C:127-174). */
static int neighbourhood(const orc_world* w, int c, int levels, int* list, int cap)
{
    int n = 0, lo = 0;
    list[n++] = c;
    for (int l = 0; l < levels; ++l) {
        int hi = n;
        for (int i = lo; i < hi; ++i)
            for (int j = 0; j < MAXF; ++j) {
                int k = list[i] * MAXF + j;
                if (w->fkind[k] != FACE_INTERNAL) continue;
                int o = w->nbr[k], seen = 0;
                for (int q = 0; q < n; ++q) if (list[q] == o) { seen = 1; break; }
                if (!seen && n < cap) list[n++] = o;
            }
        lo = hi;
    }
    return n;
}

/* Refinement passes of the model ids in w->cmodel: 3 passes, neighbourhood
 * of 3 levels, fraction 0.4 (uniGasHybridDecomposition.C:73-75). */
static void refine_models(orc_world* w)
{
    int nc = w->nc;
    int32_t* ids = malloc(sizeof(int32_t) * (size_t)nc);
    int list[KN_NB_CAP];
    for (int pass = 0; pass < 3; ++pass) {
        for (int loop = 0; loop < 2; ++loop) {
            int me = (loop == 0) ? 1 : 0;          /* 1: DSMC cells, 0: BGK cells */
            memcpy(ids, w->cmodel, sizeof(int32_t) * nc);
            for (int c = 0; c < nc; ++c) {
                if (ids[c] != me) continue;
                int same = 0, other = 0;
                for (int j = 0; j < MAXF; ++j) {
                    int k = c * MAXF + j;
                    if (w->fkind[k] != FACE_INTERNAL) continue;
                    if (ids[w->nbr[k]] == me) same++; else other++;
                }
                if (same == 0 || (same == 1 && other > 1)) { w->cmodel[c] = 1 - me; continue; }
                int n = neighbourhood(w, c, 3, list, KN_NB_CAP);
                int cnt = 0;
                for (int q = 0; q < n; ++q) if (ids[list[q]] == me) cnt++;
                if ((double)cnt < 0.4 * (double)n) w->cmodel[c] = 1 - me;
            }
        }
    }
    free(ids);
}

/* localKnudsen::decompose at a decomposition step (localKnudsen.C:243-562).
 * Time sums (knacc, KNR rows per cell: N, M, M u, M v, M w, M |u|^2, -, -,
 * then the per-species counts) are Delta-t weighted raw cell sums; the
 * XnParticle sums of the reference are F_N times them.  The refinement
 * passes (:436-545) decide every cell of a loop from the model ids as they
 * were at the start of that loop (the reference updates in place in cell
 * order, which makes the result depend on the cell numbering). */
static void hybrid_switch(orc_world* w)
{
    int nc = w->nc;
    double FN0 = w->model.n_equivalent_particles, tav = w->kn_time;
    double bmax = w->model.breakdown_max, th = w->model.kn_theta;
    int passes = w->model.smoothing_passes;
    double* rhoN = calloc((size_t)nc, sizeof(double));
    double* sc = calloc((size_t)nc * 3, sizeof(double));      /* rhoM, p, T */
    double* U = calloc((size_t)nc * 3, sizeof(double));
    double* tmp = calloc((size_t)nc * 3, sizeof(double));
    double* rhoM = calloc((size_t)nc, sizeof(double));
    double* p = calloc((size_t)nc, sizeof(double));
    double* T = calloc((size_t)nc, sizeof(double));
    for (int c = 0; c < nc; ++c) {
        const double* a = &w->knacc[KNR * c];
        double V = w->vol[c];
        double FN = FN0 * w->cw[c];
        if (a[0] > VSMALL) {
            rhoN[c] = FN * a[0] / (tav * V);
            rhoM[c] = FN * a[1] / (tav * V);
            double rhoMMean = FN * a[1] / (V * tav);
            for (int d = 0; d < 3; ++d) U[3 * c + d] = FN * a[2 + d] / (rhoMMean * V * tav);
            double ke = 0.5 * (FN * a[5]) / (V * tav);
            double rhoNMean = FN * a[0] / (V * tav);
            double uu = U[3 * c] * U[3 * c] + U[3 * c + 1] * U[3 * c + 1] + U[3 * c + 2] * U[3 * c + 2];
            T[c] = 2.0 / (3.0 * KB * rhoNMean) * (ke - 0.5 * rhoMMean * uu);
            p[c] = rhoN[c] * KB * T[c];
        }
    }
    for (int pass = 0; pass < passes; ++pass) {
        smooth_field(w, rhoM, tmp, 1); memcpy(rhoM, tmp, sizeof(double) * nc);
        smooth_field(w, p, tmp, 1); memcpy(p, tmp, sizeof(double) * nc);
        smooth_field(w, T, tmp, 1); memcpy(T, tmp, sizeof(double) * nc);
        smooth_field(w, U, tmp, 3); memcpy(U, tmp, sizeof(double) * nc * 3);
    }
    for (int c = 0; c < nc; ++c) {
        double gr = 0.0, gt = 0.0, gu = 0.0;
        double mU = sqrt(U[3 * c] * U[3 * c] + U[3 * c + 1] * U[3 * c + 1] + U[3 * c + 2] * U[3 * c + 2]);
        for (int j = 0; j < MAXF; ++j) {
            int k = c * MAXF + j;
            if (w->fkind[k] != FACE_INTERNAL) continue;
            int o = w->nbr[k];
            const double* dl = &w->delta[3 * k];
            double dist = sqrt(dl[0] * dl[0] + dl[1] * dl[1] + dl[2] * dl[2]);
            double mUo = sqrt(U[3 * o] * U[3 * o] + U[3 * o + 1] * U[3 * o + 1] + U[3 * o + 2] * U[3 * o + 2]);
            double v;
            v = fabs(rhoM[o] - rhoM[c]) / dist; if (gr < v) gr = v;
            v = fabs(T[o] - T[c]) / dist; if (gt < v) gt = v;
            v = fabs(mUo - mU) / dist; if (gu < v) gu = v;
        }
        double inst;
        const double* a = &w->knacc[KNR * c];
        if (a[0] > VSMALL && T[c] > VSMALL) {
            double V = w->vol[c];
            double FN = FN0 * w->cw[c];
            double smfp[NS];
            for (int i = 0; i < w->ns; ++i) {
                smfp[i] = 0.0;
                for (int q = 0; q < w->ns; ++q) {
                    double dPQ = 0.5 * (w->diam[i] + w->diam[q]);
                    double omPQ = 0.5 * (w->omega[i] + w->omega[q]);
                    double mr = w->mass[i] / w->mass[q];
                    double nq = FN * a[8 + q];
                    if (nq > VSMALL) {
                        double nD = nq / V;
                        /* Bird eq. 4.76 */
                        smfp[i] += PI_D * dPQ * dPQ * nD * orc_pow(w->model.Tref / T[c], omPQ - 0.5) * sqrt(1.0 + mr);
                    }
                }
            }
            double mfp = 0.0;
            for (int i = 0; i < w->ns; ++i) {
                double ni = FN * a[8 + i];
                if (ni > VSMALL) {
                    smfp[i] = 1.0 / smfp[i];
                    mfp += smfp[i] * ni / (rhoN[c] * V);      /* Bird eq. 4.77 */
                }
            }
            double u0 = sqrt(2.0 * KB / (rhoM[c] / rhoN[c]) * T[c]);
            double kr = mfp * gr / rhoM[c];
            double kt = mfp * gt / T[c];
            double ku = mfp * gu / ((mU < u0) ? u0 : mU);
            inst = (kr < kt) ? kt : kr;
            inst = (inst < ku) ? ku : inst;
        } else {
            inst = 2.0 * bmax;
        }
        w->kn[c] = th * inst + (1.0 - th) * w->kn[c];
    }
    for (int pass = 0; pass < passes; ++pass) {
        smooth_field(w, w->kn, tmp, 1);
        memcpy(w->kn, tmp, sizeof(double) * nc);
    }
    for (int c = 0; c < nc; ++c) w->cmodel[c] = (w->kn[c] > bmax) ? 1 : 0;
    refine_models(w);
    /* reset (localKnudsen.C:520-542): time = (step + 1) deltaT */
    double time = w->time + w->model.delta_t;
    double until = (w->model.reset_until_time > 0.0) ? w->model.reset_until_time : 1.0e300;
    if (w->model.reset_at_decomposition && time < until + 0.5 * w->model.delta_t) {
        memset(w->knacc, 0, sizeof(double) * KNR * (size_t)nc);
        w->kn_time = 0.0;
    }
    w->kn_steps = 0;
    free(rhoN); free(sc); free(U); free(tmp); free(rhoM); free(p); free(T);
}

/* ------------------------------------------------------------ accessors */

int orc_samples(orc_world* w, double* out, int64_t len)
{
    if (len < (int64_t)NSAMP * w->nc) return fail(UGX_ERR_ARG, "sample buffer too small");
    memcpy(out, w->samp, sizeof(double) * NSAMP * (size_t)w->nc);
    return UGX_OK;
}

int orc_counters(orc_world* w, ugx_counters* out)
{
    *out = w->cnt;
    out->n_particles = w->n;
    int64_t k = 0;
    for (int c = 0; c < w->nc; ++c) k += w->cmodel[c];
    out->n_dsmc_cells = k;
    return UGX_OK;
}

int orc_sigma_max(orc_world* w, double* out, int64_t len)
{
    if (len < w->nc) return fail(UGX_ERR_ARG, "buffer too small");
    memcpy(out, w->smax, sizeof(double) * (size_t)w->nc);
    return UGX_OK;
}

int orc_set_sigma_max(orc_world* w, const double* in, int64_t len)
{
    if (len < w->nc) return fail(UGX_ERR_ARG, "buffer too small");
    memcpy(w->smax, in, sizeof(double) * (size_t)w->nc);
    return UGX_OK;
}

/* Test hooks of the hybrid decomposition (oracle only). */
int orc_hybrid_set(orc_world* w, const double* knacc, double tav)
{
    memcpy(w->knacc, knacc, sizeof(double) * KNR * (size_t)w->nc);
    w->kn_time = tav;
    return UGX_OK;
}

int orc_hybrid_decompose(orc_world* w)
{
    hybrid_switch(w);
    return UGX_OK;
}

int orc_smooth(orc_world* w, const double* in, double* out, int ncomp)
{
    if (ncomp != 1 && ncomp != 3) return fail(UGX_ERR_ARG, "ncomp must be 1 or 3");
    smooth_field(w, in, out, ncomp);
    return UGX_OK;
}

int orc_refine(orc_world* w, int32_t* ids)
{
    memcpy(w->cmodel, ids, sizeof(int32_t) * (size_t)w->nc);
    refine_models(w);
    memcpy(ids, w->cmodel, sizeof(int32_t) * (size_t)w->nc);
    return UGX_OK;
}

int orc_kn(orc_world* w, double* out, int64_t len)
{
    if (len < w->nc) return fail(UGX_ERR_ARG, "buffer too small");
    memcpy(out, w->kn, sizeof(double) * (size_t)w->nc);
    return UGX_OK;
}

int orc_cell_model(orc_world* w, int32_t* out, int64_t len)
{
    if (len < w->nc) return fail(UGX_ERR_ARG, "buffer too small");
    memcpy(out, w->cmodel, sizeof(int32_t) * (size_t)w->nc);
    return UGX_OK;
}

int orc_geometry(orc_world* w, double* cc, double* vol, int32_t* nfaces, double* planes,
                 int32_t* nbr, double* delta)
{
    size_t nc = (size_t)w->nc;
    if (cc) memcpy(cc, w->cc, sizeof(double) * 3 * nc);
    if (vol) memcpy(vol, w->vol, sizeof(double) * nc);
    if (nfaces) memcpy(nfaces, w->ncf, sizeof(int32_t) * nc);
    if (planes) memcpy(planes, w->plane, sizeof(double) * 4 * MAXF * nc);
    if (nbr) memcpy(nbr, w->nbr, sizeof(int32_t) * MAXF * nc);
    if (delta) memcpy(delta, w->delta, sizeof(double) * 3 * MAXF * nc);
    return UGX_OK;
}

int orc_face_data(orc_world* w, double* farea, int32_t* fkind)
{
    for (int k = 0; k < w->nc * MAXF; ++k) { farea[k] = w->farea[k]; fkind[k] = w->fkind[k]; }
    return UGX_OK;
}

int orc_inflow_nbar(orc_world* w, double* out, int64_t len, int64_t* n_out)
{
    *n_out = w->nin;
    for (int k = 0; k < w->nin && k < len; ++k) out[k] = w->inflow[k].nbar;
    return UGX_OK;
}

/* ------------------------------------------------------- primitive hooks */

void orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) { orc_philox(ctr, key, out); }

/* fn: 0 log, 1 exp, 2 sin(2 pi u), 3 cos(2 pi u), 4 pow(x, y=in2) */
void orc_mathfn(int fn, const double* in, const double* in2, double* out, int64_t n)
{
    for (int64_t i = 0; i < n; ++i) {
        double s, c;
        switch (fn) {
        case 0: out[i] = orc_log(in[i]); break;
        case 1: out[i] = orc_exp(in[i]); break;
        case 2: orc_sincos2pi(in[i], &s, &c); out[i] = s; break;
        case 3: orc_sincos2pi(in[i], &s, &c); out[i] = c; break;
        default: out[i] = orc_pow(in[i], in2[i]); break;
        }
    }
}

/* n draws of a stream: kind 0 uniform, 1 index in [0, m), 2 gauss pairs */
void orc_stream(uint64_t seed, uint32_t kind, uint32_t sub, uint32_t id, uint32_t step, int what,
                uint32_t m, double* out, int64_t n)
{
    orc_rng r;
    orc_rng_init(&r, seed, kind, sub, id, step);
    for (int64_t i = 0; i < n; ++i) {
        if (what == 0) out[i] = orc_uniform(&r);
        else if (what == 1) out[i] = (double)orc_index(&r, m);
        else { double a, b; orc_gauss2(&r, &a, &b); out[i] = a; if (i + 1 < n) out[++i] = b; }
    }
}

/* ------------------------------------------------------- physics hooks */

/* Test hook: n independent collisions of one pair (species p, q) with the
 * same pre-collision state, each drawn from its own stream (NTC kind,
 * entity i), through binary_collide: the dsmcCollisionModel::collide of the
 * case's model (variableHardSphere.C:118-162, variableSoftSphere.C:116-176,
 * LarsenBorgnakke*.C).  UP/UQ [3] and eP/eQ are the pre-collision state;
 * outP/outQ [3n] and erP/erQ [n] receive the results. */
int orc_pair_collide(orc_world* w, int p, int q, const double* UP, const double* UQ, double eP, double eQ,
                     int64_t n, uint64_t seed, double* outP, double* outQ, double* erP, double* erQ)
{
    if (p < 0 || q < 0 || p >= w->ns || q >= w->ns) return fail(UGX_ERR_ARG, "species out of range");
    for (int64_t i = 0; i < n; ++i) {
        orc_rng r;
        orc_rng_init(&r, seed, ORC_RNG_NTC, 0u, (uint32_t)i, 0u);
        double a[3] = {UP[0], UP[1], UP[2]}, b[3] = {UQ[0], UQ[1], UQ[2]};
        double ea = eP, eb = eQ;
        binary_collide(w, &r, p, q, a, b, &ea, &eb);
        for (int d = 0; d < 3; ++d) { outP[3 * i + d] = a[d]; outQ[3 * i + d] = b[d]; }
        if (erP) erP[i] = ea;
        if (erQ) erQ[i] = eb;
    }
    return UGX_OK;
}

/* Test hooks: the USP per-cell state carried between steps (relaxed heat
 * flux q [nc*3], relaxed stress [nc*6] xx xy xz yy yz zz, maxProb [nc];
 * unifiedStochasticParticleSBGK.C:774-783, :859-863).  NULL skips an array. */
int orc_usp_state(orc_world* w, double* q, double* sh, double* maxp)
{
    size_t nc = (size_t)w->nc;
    if (q) memcpy(q, w->uspq, sizeof(double) * 3 * nc);
    if (sh) memcpy(sh, w->usps, sizeof(double) * 6 * nc);
    if (maxp) memcpy(maxp, w->uspmax, sizeof(double) * nc);
    return UGX_OK;
}

int orc_set_usp_state(orc_world* w, const double* q, const double* sh, const double* maxp)
{
    size_t nc = (size_t)w->nc;
    if (q) memcpy(w->uspq, q, sizeof(double) * 3 * nc);
    if (sh) memcpy(w->usps, sh, sizeof(double) * 6 * nc);
    if (maxp) memcpy(w->uspmax, maxp, sizeof(double) * nc);
    return UGX_OK;
}

/* Test hook: zero the sampler sums (the sampler reset at output of
 * uniGasVolFields.C:1432-1502, as ugx_samples_reset). */
int orc_samples_reset(orc_world* w)
{
    memset(w->samp, 0, sizeof(double) * NSAMP * (size_t)w->nc);
    memset(w->wsamp, 0, sizeof(double) * NS * WROWS * (size_t)w->nwf);
    return UGX_OK;
}

int64_t orc_wall_face_count(orc_world* w) { return w->nwf; }

int orc_wall_samples(orc_world* w, double* out, int64_t len)
{
    if (len < (int64_t)w->nwf * NS * WROWS) return fail(UGX_ERR_ARG, "wall sum buffer too small");
    memcpy(out, w->wsamp, sizeof(double) * NS * WROWS * (size_t)w->nwf);
    return UGX_OK;
}
