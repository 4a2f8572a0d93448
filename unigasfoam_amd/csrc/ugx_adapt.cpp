// ugx_adapt.cpp - the dynamic adapter of an adaptive case
// (src/lagrangian/uniGas/dynamicAdaptation/uniGasDynamicAdapter.C), host
// side, once at the fill and once every adaptationInterval steps:
//  * per cell, from a number density per species, a temperature and a
//    velocity: the time step over the mean collision time (Bird eq. 1.38,
//    4.74), the Courant number per axis from max(most probable speed, U_d)
//    and the cell's bounding box, and the box size over the mean free path
//    (Bird eq. 4.76-4.77) (calculateAdaptationQuantities, :224-314);
//  * the time step scaled so that the largest ratio of the DSMC cells and
//    the largest Courant number of the solution directions meet their
//    limits (calculateTimeStep, :316-372);
//  * the subcell levels of the DSMC cells: ceil(size ratio / limit) in
//    1..10 per solution direction (calculateSubCellLevels, :374-414);
//  * the cell weights of a weighted case (calculateCellWeightFactor and
//    smoothCellWeightFactor, :416-503, blended at :661-668).
// At the fill the gas is the filled one and all three fields are smoothed
// (setInitialConfiguration, :700-770); afterwards it is the time average of
// the cell measurements since the last adaptation and only the size ratio
// is smoothed (adapt, :526-700).
#include <algorithm>
#include <cmath>

#include "ugx_world.hpp"

namespace ugx {

static constexpr double kVGreat = 1.0e300;
static constexpr int kMinSubcellLevels = 1;

// calculateAdaptationQuantities (uniGasDynamicAdapter.C:224-314) for cell c
// of the case; nsp[ns] the species number densities.
static void cell_quantities(const CaseSpec& spec, int c, double rhoN, double T, const double U[3], const double* nsp,
                            double dt, const double* prev, double& mct, double co[3], double ratio[3])
{
    const MeshGeom& g = spec.geom;
    const int ns = spec.ns;
    const double Tref = spec.model.Tref;
    if (T > kSmall) {
        double u0 = 0.0;
        double mcr[kNS] = {}, mfp[kNS] = {};
        for (int i = 0; i < ns; ++i) {
            const ugx_species_desc& si = spec.species[i];
            for (int q = 0; q < ns; ++q) {
                const ugx_species_desc& sq = spec.species[q];
                const double dPQ = 0.5 * (si.diameter + sq.diameter);
                const double omegaPQ = 0.5 * (si.omega + sq.omega);
                const double massRatio = si.mass / sq.mass;
                if (nsp[q] > kVSmall && T > kVSmall) {
                    const double mr = si.mass * sq.mass / (si.mass + sq.mass);
                    mcr[i] = mcr[i] + 2.0 * std::sqrt(kPi) * dPQ * dPQ * nsp[q] * std::pow(T / Tref, 1.0 - omegaPQ)
                                          * std::sqrt(2.0 * kB * Tref / mr);
                    mfp[i] = mfp[i] + kPi * dPQ * dPQ * nsp[q] * std::pow(Tref / T, omegaPQ - 0.5)
                                          * std::sqrt(1.0 + massRatio);
                }
            }
            if (nsp[i] > kVSmall) u0 = std::max(u0, std::sqrt(2.0 * kB / si.mass * T));
        }
        double MCR = 0.0, MFP = 0.0;
        for (int i = 0; i < ns; ++i) {
            if (!(nsp[i] > kVSmall)) continue;
            const double lam = 1.0 / mfp[i];
            MCR = MCR + mcr[i] * nsp[i] / rhoN;
            MFP = MFP + lam * nsp[i] / rhoN;
        }
        mct = dt * MCR;
        for (int d = 0; d < 3; ++d) {
            const double len = g.bbhi[3 * (size_t)c + d] - g.bblo[3 * (size_t)c + d];
            co[d] = std::max(u0, U[d]) * dt / len;
            ratio[d] = len / MFP;
        }
    } else {
        mct = 0.0;
        for (int d = 0; d < 3; ++d) { co[d] = 0.0; ratio[d] = prev[d]; }
    }
}

// fvc::average(fvc::interpolate(f)) of an ncomp-component cell field: the
// face values (linear weight on internal and cyclic faces, the cell's own on
// boundary faces) averaged by face area, empty faces skipped.
static void smooth_field(const MeshGeom& g, std::vector<double>& f, int ncomp)
{
    std::vector<double> out(f.size());
    for (int c = 0; c < g.nc; ++c) {
        double num[3] = {0.0, 0.0, 0.0}, den = 0.0;
        for (int j = 0; j < kMaxFaces; ++j) {
            const size_t k = (size_t)c * kMaxFaces + j;
            const int kind = g.fkind[k];
            if (kind == kFaceSkip) continue;
            const double a = g.farea[k];
            const double* pl = &g.plane[4 * k];
            const bool inner = kind == kFaceInternal || kind == kFaceCyclic;
            double wt = 1.0;
            int o = c;
            if (inner) {
                o = g.nbr[k];
                const double* dl = &g.delta[3 * k];
                const double dP = std::fabs(pl[3]);
                const double dN = std::fabs((pl[0] * dl[0] + pl[1] * dl[1] + pl[2] * dl[2]) + pl[3]);
                wt = dN / (dP + dN);
            }
            // a symmetryPlane patch holds a vector field's reflection: U - n (n.U)
            const bool mirror = kind == kFaceSymmetry && ncomp == 3;
            const double* u = &f[(size_t)c * ncomp];
            const double un = mirror ? pl[0] * u[0] + pl[1] * u[1] + pl[2] * u[2] : 0.0;
            for (int d = 0; d < ncomp; ++d) {
                const double fv = inner    ? wt * f[(size_t)c * ncomp + d] + (1.0 - wt) * f[(size_t)o * ncomp + d]
                                  : mirror ? u[d] - un * pl[d]
                                           : f[(size_t)c * ncomp + d];
                num[d] = num[d] + a * fv;
            }
            den = den + a;
        }
        for (int d = 0; d < ncomp; ++d)
            out[(size_t)c * ncomp + d] = (den > 0.0) ? num[d] / den : f[(size_t)c * ncomp + d];
    }
    f.swap(out);
}

// calculateTimeStep (uniGasDynamicAdapter.C:316-372).
static double adapted_time_step(const CaseSpec& spec, const std::vector<double>& mct, const std::vector<double>& co,
                                const int32_t* cmodel, double dt)
{
    const ugx_model_desc& m = spec.model;
    const MeshGeom& g = spec.geom;
    // the reference names the BGK model only in BGK and hybrid cases
    const bool usp = m.collision_model != UGX_COLL_DSMC && m.bgk_model == UGX_BGK_USP;
    double imct = 0.0, ico = 0.0;
    for (int c = 0; c < g.nc; ++c)
        if (!usp && cmodel[c] != 0 && imct < mct[c]) imct = mct[c];
    for (int c = 0; c < g.nc; ++c)
        for (int d = 0; d < 3; ++d)
            if (g.mask[d] != 0.0 && ico < co[3 * (size_t)c + d]) ico = co[3 * (size_t)c + d];
    if (imct > kVSmall && ico < kVSmall) dt = dt * (m.max_time_step_mct_ratio / imct);
    else if (imct < kVSmall && ico > kVSmall) dt = dt * (m.max_courant_number / ico);
    else if (imct > kVSmall && ico > kVSmall)
        dt = dt * std::min(m.max_time_step_mct_ratio / imct, m.max_courant_number / ico);
    return dt;
}

// calculateSubCellLevels (uniGasDynamicAdapter.C:374-414), packed x | y << 8 | z << 16.
static void subcell_levels(const CaseSpec& spec, const std::vector<double>& ratio, const int32_t* cmodel,
                           uint32_t* lv)
{
    const MeshGeom& g = spec.geom;
    const double lim = spec.model.max_subcell_size_mfp_ratio;
    for (int c = 0; c < g.nc; ++c) {
        uint32_t p = 0u;
        for (int d = 0; d < 3; ++d) {
            int l = 1;
            if (g.mask[d] != 0.0) {
                if (cmodel[c] == 1) {
                    const double x = std::ceil(ratio[3 * (size_t)c + d] / lim);
                    l = (int)std::min((double)kMaxSubcellLevels, std::max((double)kMinSubcellLevels, x));
                } else {
                    l = kMinSubcellLevels;
                }
            }
            p |= (uint32_t)l << (8 * d);
        }
        lv[c] = p;
    }
}

int adapt_initial(const CaseSpec& spec, const double* nd, double T, const double U[3], const int32_t* cmodel,
                  AdaptOut& out)
{
    const MeshGeom& g = spec.geom;
    const ugx_model_desc& m = spec.model;
    double rhoN = 0.0;
    for (int s = 0; s < spec.ns; ++s) rhoN = rhoN + nd[s];
    std::vector<double> mct((size_t)g.nc), co((size_t)g.nc * 3);
    out.ratio.assign((size_t)g.nc * 3, 0.0);
    const double zero[3] = {0.0, 0.0, 0.0};
    for (int c = 0; c < g.nc; ++c)
        cell_quantities(spec, c, rhoN, T, U, nd, m.delta_t, zero, mct[c], &co[3 * (size_t)c], &out.ratio[3 * (size_t)c]);
    for (int pass = 0; pass < m.adaptation_smoothing_passes; ++pass) {
        smooth_field(g, mct, 1);
        smooth_field(g, co, 3);
        smooth_field(g, out.ratio, 3);
    }
    out.dt = m.time_step_adaptation ? adapted_time_step(spec, mct, co, cmodel, m.delta_t) : m.delta_t;
    out.levels.clear();
    if (m.subcell_adaptation) {
        out.levels.resize((size_t)g.nc);
        subcell_levels(spec, out.ratio, cmodel, out.levels.data());
    }
    return UGX_OK;
}

void smooth_max_ratio(const MeshGeom& g, std::vector<double>& f, double max_ratio)
{
    bool changed = true;
    while (changed) {
        changed = false;
        for (int c = 0; c < g.nc; ++c)
            for (int j = 0; j < kMaxFaces; ++j) {
                const size_t k = (size_t)c * kMaxFaces + j;
                if (g.fkind[k] != kFaceInternal && g.fkind[k] != kFaceCyclic) continue;
                const double v = f[g.nbr[k]] / max_ratio;
                if (v > f[c]) { f[c] = v; changed = true; }
            }
    }
}

// Cell weight adaptation (uniGasDynamicAdapter.C:416-503, 650-672): the
// weights of the averaged densities at particlesPerSubCell per subcell,
// fvc::smooth(W, 1.3), area-average passes capped at minParticlesPerSubCell
// per subcell until neighbour ratios are below 1 + maxCellWeightRatio,
// then blended into the old weights with theta = 0.2.
static void adapt_weights(const CaseSpec& spec, const std::vector<double>& rhoN, const uint32_t* lv,
                          const double* cw_old, std::vector<double>& cw)
{
    const MeshGeom& g = spec.geom;
    const ugx_model_desc& m = spec.model;
    const int nc = g.nc;
    const double FN = m.n_equivalent_particles;
    const int npps = m.particles_per_subcell;
    const int nmin = m.min_particles_per_subcell > 0 ? m.min_particles_per_subcell : npps;
    const int max_passes = m.max_smoothing_passes > 0 ? m.max_smoothing_passes : 500;
    const double max_ratio = m.max_cell_weight_ratio > 0.0 ? m.max_cell_weight_ratio : 0.05;
    const double theta = 0.2;
    std::vector<double> nsub((size_t)nc), wn((size_t)nc);
    for (int c = 0; c < nc; ++c)
        nsub[c] = (double)(lv[c] & 0xffu) * (double)((lv[c] >> 8) & 0xffu) * (double)((lv[c] >> 16) & 0xffu);
    for (int c = 0; c < nc; ++c) wn[c] = (rhoN[c] * g.vol[c]) / ((double)npps * nsub[c] * FN * 1.0);
    smooth_max_ratio(g, wn, 1.0 + 1.3);
    int passes = 0;
    double max_r;
    do {
        passes++;
        smooth_field(g, wn, 1);
        for (int c = 0; c < nc; ++c)
            wn[c] = std::max(std::min((rhoN[c] * g.vol[c]) / ((double)nmin * nsub[c] * FN * 1.0), wn[c]), kSmall);
        max_r = 1.0e-300;
        for (int c = 0; c < nc; ++c)
            for (int j = 0; j < kMaxFaces; ++j) {
                const size_t k = (size_t)c * kMaxFaces + j;
                if (g.fkind[k] != kFaceInternal) continue;
                const int o = g.nbr[k];
                const double r = std::max(wn[c] / wn[o], wn[o] / wn[c]);
                if (r > max_r) max_r = r;
            }
    } while (max_r > 1.0 + max_ratio && passes < max_passes);
    cw.resize((size_t)nc);
    for (int c = 0; c < nc; ++c) cw[c] = theta * wn[c] + (1.0 - theta) * cw_old[c];
}

int adapt_step(const CaseSpec& spec, const double* acc, double tav, const double* cw, const int32_t* cmodel,
               double dt, const double* prev, const uint32_t* levels, AdaptOut& out)
{
    const MeshGeom& g = spec.geom;
    const ugx_model_desc& m = spec.model;
    const int nc = g.nc, ns = spec.ns;
    const double FN = m.n_equivalent_particles;
    if (!(tav > 0.0)) return set_err(UGX_ERR_STATE, "adaptation without sampled time");
    std::vector<double> rhoN((size_t)nc), T((size_t)nc), U((size_t)nc * 3);
    double minRhoN = kVGreat, minT = kVGreat;
    // time-averaged fields of the cell measurements (uniGasDynamicAdapter.C:549-582)
    for (int c = 0; c < nc; ++c) {
        const double* a = &acc[(size_t)kKnRows * c];
        if (a[0] > kVSmall) {
            const double FW = FN * cw[c], V = g.vol[c];
            rhoN[c] = FW * a[0] / (V * tav);
            const double rhoM = FW * a[1] / (V * tav);
            for (int d = 0; d < 3; ++d) U[3 * (size_t)c + d] = FW * a[2 + d] / (rhoM * V * tav);
            const double ke = 0.5 * (FW * a[5]) / (V * tav);
            const double rhoNMean = FW * a[0] / (V * tav);
            const double* u = &U[3 * (size_t)c];
            T[c] = 2.0 / (3.0 * kB * rhoNMean) * (ke - 0.5 * rhoM * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
            if (T[c] > kVSmall) {
                if (minRhoN > rhoN[c]) minRhoN = rhoN[c];
                if (minT > T[c]) minT = T[c];
            }
        } else {
            rhoN[c] = 0.0;
            T[c] = 0.0;
            for (int d = 0; d < 3; ++d) U[3 * (size_t)c + d] = 0.0;
        }
    }
    // size ratios (:584-621): an empty cell takes the smallest density,
    // split evenly over the species, and a cold one the lowest temperature
    std::vector<double> mct((size_t)nc), co((size_t)nc * 3);
    out.ratio.assign((size_t)nc * 3, 0.0);
    for (int c = 0; c < nc; ++c) {
        const double* a = &acc[(size_t)kKnRows * c];
        const double FW = FN * cw[c], V = g.vol[c];
        double nsp[kNS] = {};
        for (int s = 0; s < ns; ++s) nsp[s] = FW * a[8 + s] / (V * tav);
        if (rhoN[c] == 0.0) {
            rhoN[c] = minRhoN;
            for (int s = 0; s < ns; ++s) nsp[s] = rhoN[c] / (double)ns;
        }
        if (T[c] == 0.0) T[c] = minT;
        cell_quantities(spec, c, rhoN[c], T[c], &U[3 * (size_t)c], nsp, dt, &prev[3 * (size_t)c], mct[c],
                        &co[3 * (size_t)c], &out.ratio[3 * (size_t)c]);
    }
    for (int pass = 0; pass < m.adaptation_smoothing_passes; ++pass) smooth_field(g, out.ratio, 3);
    out.dt = m.time_step_adaptation ? adapted_time_step(spec, mct, co, cmodel, dt) : dt;
    out.levels.clear();
    if (m.subcell_adaptation) {
        out.levels.resize((size_t)nc);
        subcell_levels(spec, out.ratio, cmodel, out.levels.data());
    }
    out.cw.clear();
    if (m.cell_weight_adaptation)
        adapt_weights(spec, rhoN, out.levels.empty() ? levels : out.levels.data(), cw, out.cw);
    return UGX_OK;
}

}  // namespace ugx
