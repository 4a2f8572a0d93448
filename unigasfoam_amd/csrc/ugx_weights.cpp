// ugx_weights.cpp - cell weights of a cell-weighted case's fill
// (src/lagrangian/uniGas/uniGasInitialisation/derived/uniGasMeshFill/
// uniGasMeshFill.C:111-168), without radial weighting (RWF = 1):
//   W_c = n V_c / (N_pps N_sub,c F_N)  (N_sub,c the cell's subcell count);
//   repeat: W = fvc::average(fvc::interpolate(W)) (zeroGradient boundaries),
//           W_c = max(min(n V_c / (N_min N_sub,c F_N), W_c), SMALL),
//   while some internal face has max(W_o/W_n, W_n/W_o) > 1 + maxCellWeightRatio
//   and fewer than maxSmoothingPasses passes were made.
// The smoothing pass is the face interpolation and area average of the
// hybrid decomposition (linear weight dN / (dP + dN) on internal and cyclic
// faces, the cell value on boundary faces, empty faces skipped).
#include <cmath>

#include "ugx_world.hpp"

namespace ugx {


static void smooth_pass(const MeshGeom& g, const std::vector<double>& f, std::vector<double>& out)
{
    for (int c = 0; c < g.nc; ++c) {
        double num = 0.0, den = 0.0;
        for (int j = 0; j < kMaxFaces; ++j) {
            const size_t k = (size_t)c * kMaxFaces + j;
            const int kind = g.fkind[k];
            if (kind == kFaceSkip) continue;
            const double a = g.farea[k];
            const double* pl = &g.plane[4 * k];
            double fv;
            if (kind == kFaceInternal || kind == kFaceCyclic) {
                const int o = g.nbr[k];
                const double* dl = &g.delta[3 * k];
                const double dP = std::fabs(pl[3]);
                const double dN = std::fabs((pl[0] * dl[0] + pl[1] * dl[1] + pl[2] * dl[2]) + pl[3]);
                const double wt = dN / (dP + dN);
                fv = wt * f[c] + (1.0 - wt) * f[o];
            } else {
                fv = f[c];
            }
            num = num + a * fv;
            den = den + a;
        }
        out[c] = (den > 0.0) ? num / den : f[c];
    }
}

int fill_cell_weights(const CaseSpec& spec, double ntot, const uint32_t* lv, std::vector<double>& w)
{
    const MeshGeom& g = spec.geom;
    const ugx_model_desc& m = spec.model;
    const double FN = m.n_equivalent_particles;
    const int npps = m.particles_per_subcell;
    const int nmin = m.min_particles_per_subcell > 0 ? m.min_particles_per_subcell : npps;
    const int max_passes = m.max_smoothing_passes > 0 ? m.max_smoothing_passes : 500;
    const double max_ratio = m.max_cell_weight_ratio > 0.0 ? m.max_cell_weight_ratio : 0.05;
    if (npps <= 0) return set_err(UGX_ERR_ARG, "a cell-weighted case needs particles_per_subcell > 0");
    if (!(ntot > 0.0)) return set_err(UGX_ERR_ARG, "a cell-weighted fill needs a number density > 0");
    w.assign((size_t)g.nc, 1.0);
    std::vector<double> nsub((size_t)g.nc, 1.0);
    if (lv)
        for (int c = 0; c < g.nc; ++c)
            nsub[c] = (double)(lv[c] & 0xffu) * (double)((lv[c] >> 8) & 0xffu) * (double)((lv[c] >> 16) & 0xffu);
    for (int c = 0; c < g.nc; ++c) w[c] = (ntot * g.vol[c]) / ((double)npps * nsub[c] * FN * 1.0);
    std::vector<double> tmp((size_t)g.nc);
    int passes = 0;
    double max_r;
    do {
        passes++;
        smooth_pass(g, w, tmp);
        w.swap(tmp);
        for (int c = 0; c < g.nc; ++c)
            w[c] = std::max(std::min((ntot * g.vol[c]) / ((double)nmin * nsub[c] * FN * 1.0), w[c]), kSmall);
        max_r = 1.0e-300;
        for (int c = 0; c < g.nc; ++c)
            for (int j = 0; j < kMaxFaces; ++j) {
                const size_t k = (size_t)c * kMaxFaces + j;
                if (g.fkind[k] != kFaceInternal) continue;
                const int o = g.nbr[k];
                const double r = std::max(w[c] / w[o], w[o] / w[c]);
                if (r > max_r) max_r = r;
            }
    } while (max_r > 1.0 + max_ratio && passes < max_passes);
    return UGX_OK;
}

}  // namespace ugx
