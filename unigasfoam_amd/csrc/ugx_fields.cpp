// ugx_fields.cpp - write-time cell fields of uniGasVolFields from the raw
// time sums (src/lagrangian/uniGas/macroscopicProperties/derived/volumetric/
// uniGasVolFields/uniGasVolFields.C:839-1252), with cell and radial weights
// equal to 1 so that every "XnParticle" sum is F_N times the raw sum.
// Vibrational and electronic modes carry no energy here (out of scope), so
// their temperatures and degrees of freedom are zero.
//
// Extras for shear flows (not volFields outputs in the reference): the
// time-averaged pressure tensor and heat-flux vector, restated from the
// moment formulas of unifiedStochasticParticleSBGK.C:486-550.
#include <cmath>
#include <cstring>

#include "ugx_world.hpp"

namespace ugx {

static constexpr double kNA = 6.02214076e23;

int compute_fields(const CaseSpec& spec, const double* A, int ncl, const ugx_counters& cnt, const double* cw,
                   ugx_fields* f)
{
    const double tAv = cnt.time_sampled;
    const double nAv = (double)cnt.n_sampled_steps;
    const double FN0 = spec.model.n_equivalent_particles;
    const double Tref = spec.model.Tref;
    const int ns = spec.ns;
    auto S = [&](int k, int c) { return A[(size_t)k * ncl + c]; };
    for (int c = 0; c < ncl; ++c) {
        // every simulator of cell c stands for F_N W_c molecules when sampled
        const double FN = FN0 * (cw ? cw[c] : 1.0);
        const double V = spec.geom.vol[c];
        const double aN = S(UGX_S_N, c);
        double rhoN = 0.0, rhoM = 0.0, U[3] = {0.0, 0.0, 0.0}, T = 0.0, p = 0.0, uMean = 0.001;
        if (aN > kVSmall && tAv > 0.0) {
            uMean = aN / tAv;
            rhoN = (FN * aN) / (tAv * V);
            rhoM = (FN * S(UGX_S_M, c)) / (tAv * V);
            const double rhoMMean = FN * S(UGX_S_M, c) / (V * tAv);
            for (int d = 0; d < 3; ++d) U[d] = FN * S(UGX_S_MU + d, c) / (rhoMMean * V * tAv);
            const double ke = 0.5 * FN * S(UGX_S_MCC, c) / (V * tAv);
            const double rhoNMean = FN * aN / (V * tAv);
            T = 2.0 / (3.0 * kB * rhoNMean) * (ke - 0.5 * rhoMMean * (U[0] * U[0] + U[1] * U[1] + U[2] * U[2]));
            p = rhoN * kB * T;
        }
        double Trot = 0.0;
        if (S(UGX_S_ROTDOF, c) > kVSmall && tAv > kVSmall)
            Trot = (2.0 / kB) * ((S(UGX_S_EROT, c) / tAv) / (S(UGX_S_ROTDOF, c) / tAv));
        double nRotDof = (aN > kVSmall) ? S(UGX_S_ROTDOF, c) / aN : 0.0;
        double Tov = (3.0 * T + nRotDof * Trot) / (3.0 + nRotDof);
        // Mach number (uniGasVolFields.C:1087-1122)
        double Cp = 0.0, Cv = 0.0, mm = 0.0;
        if (aN > kVSmall) {
            for (int s = 0; s < ns; ++s) {
                double fr = S(UGX_S_SPECIES_N + s, c) / aN;
                mm += spec.species[s].mass * fr;
                Cp += (5.0 + spec.species[s].rotational_dof) * fr;
                Cv += (3.0 + spec.species[s].rotational_dof) * fr;
            }
        }
        double Cv_p = Cv / kNA;
        double gasR = (mm > kVSmall) ? kB / mm : 0.0;
        double gamma = (aN > kVSmall && Cv > 0.0) ? Cp / Cv : 0.0;
        double Ma = 0.0;
        if (aN > kVSmall && T > kVSmall) Ma = std::sqrt(U[0] * U[0] + U[1] * U[1] + U[2] * U[2]) / std::sqrt(gamma * gasR * T);
        // mean free path and collision rate (Bird 4.74, 4.76, 4.77)
        double MFP = 0.0, MCR = 0.0;
        if (rhoN > kVSmall && T > kVSmall) {
            for (int i = 0; i < ns; ++i) {
                double mfpi = 0.0, mcri = 0.0;
                for (int q = 0; q < ns; ++q) {
                    if (!(S(UGX_S_SPECIES_N + q, c) > kVSmall)) continue;
                    const auto& a = spec.species[i];
                    const auto& b = spec.species[q];
                    double dPQ = 0.5 * (a.diameter + b.diameter);
                    double om = 0.5 * (a.omega + b.omega);
                    double nq = FN * S(UGX_S_SPECIES_N + q, c) / (V * tAv);
                    double mr = a.mass * b.mass / (a.mass + b.mass);
                    mfpi += kPi * dPQ * dPQ * nq * std::pow(Tref / T, om - 0.5) * std::sqrt(1.0 + a.mass / b.mass);
                    mcri += 2.0 * std::sqrt(kPi) * dPQ * dPQ * nq * std::pow(T / Tref, 1.0 - om) * std::sqrt(2.0 * kB * Tref / mr);
                }
                if (mfpi > kVSmall) mfpi = 1.0 / mfpi;
                double np = FN * S(UGX_S_SPECIES_N + i, c) / (V * tAv);
                MFP += mfpi * np / rhoN;
                MCR += mcri * np / rhoN;
            }
        }
        if (MFP < kVSmall) MFP = 1.0e300;
        // statistical errors (Hadjiconstantinou et al.; uniGasVolFields.C:1242-1252)
        double dE = 0.0, vE = 0.0, tE = 0.0, pE = 0.0;
        if (uMean > kVSmall && Ma > kVSmall && gamma > kVSmall && Cv_p > kVSmall && nAv > 0.0) {
            dE = 1.0 / std::sqrt(uMean * nAv);
            vE = dE * (1.0 / (Ma * std::sqrt(gamma)));
            tE = dE * std::sqrt(kB / Cv_p);
            pE = std::sqrt(gamma) / std::sqrt(uMean * nAv);
        }
        if (f->rhoN) f->rhoN[c] = rhoN;
        if (f->rhoM) f->rhoM[c] = rhoM;
        if (f->U_mean) for (int d = 0; d < 3; ++d) f->U_mean[3 * (size_t)c + d] = U[d];
        if (f->translational_T) f->translational_T[c] = T;
        if (f->rotational_T) f->rotational_T[c] = Trot;
        if (f->overall_T) f->overall_T[c] = Tov;
        if (f->p) f->p[c] = p;
        if (f->Ma) f->Ma[c] = Ma;
        if (f->mfp) f->mfp[c] = MFP;
        if (f->mcr) f->mcr[c] = MCR;
        if (f->density_error) f->density_error[c] = dE;
        if (f->velocity_error) f->velocity_error[c] = vE;
        if (f->temperature_error) f->temperature_error[c] = tE;
        if (f->pressure_error) f->pressure_error[c] = pE;
        if (f->shear_stress || f->heat_flux) {
            double P[6] = {0, 0, 0, 0, 0, 0}, q[3] = {0, 0, 0};
            if (aN > kVSmall) {
                const int ia[6] = {0, 0, 0, 1, 1, 2}, ib[6] = {0, 1, 2, 1, 2, 2};
                const double aM = S(UGX_S_M, c);
                for (int k = 0; k < 6; ++k)
                    P[k] = rhoN * (S(UGX_S_MUU + k, c) / aN - (aM / aN) * U[ia[k]] * U[ib[k]]);
                const int r0[3] = {0, 1, 2}, r1[3] = {1, 3, 4}, r2[3] = {2, 4, 5};
                for (int d = 0; d < 3; ++d)
                    q[d] = rhoN * (0.5 * (S(UGX_S_MCCU + d, c) / aN) - 0.5 * (S(UGX_S_MCC, c) / aN) * U[d]
                                   + S(UGX_S_EROTU + d, c) / aN - (S(UGX_S_EROT, c) / aN) * U[d])
                           - P[r0[d]] * U[0] - P[r1[d]] * U[1] - P[r2[d]] * U[2];
            }
            if (f->shear_stress) for (int k = 0; k < 6; ++k) f->shear_stress[6 * (size_t)c + k] = P[k];
            if (f->heat_flux) for (int d = 0; d < 3; ++d) f->heat_flux[3 * (size_t)c + d] = q[d];
        }
    }
    return UGX_OK;
}

// Wall values (uniGasVolFields.C:1262-1350) per measured wall face, the
// species sums added together; faces never hit give zeros, as there.
int compute_wall_fields(const CaseSpec& spec, const double* W, const ugx_counters& cnt, const double* cw,
                        ugx_wall_fields* f)
{
    const MeshGeom& g = spec.geom;
    const double tAv = cnt.time_sampled;
    for (int k = 0; k < g.nwf; ++k) {
        // nParticle F_N W_c of the face's cell (uniGasVolFields.C:1276-1278)
        const double FN = spec.model.n_equivalent_particles * (cw ? cw[g.wcell[k]] : 1.0);
        double a[kWRows] = {};
        for (int s = 0; s < spec.ns; ++s)
            for (int r = 0; r < kWRows; ++r) a[r] += W[((size_t)k * kNS + s) * kWRows + r];
        double rhoN = 0.0, rhoM = 0.0, U[3] = {0.0, 0.0, 0.0}, T = 0.0;
        if (a[UGX_W_RHON] > kVSmall && tAv > 0.0) {
            rhoN = a[UGX_W_RHON] * FN / tAv;
            rhoM = a[UGX_W_RHOM] * FN / tAv;
            for (int d = 0; d < 3; ++d) U[d] = a[UGX_W_MU + d] * FN / (rhoM * tAv);
            const double rhoMMean = a[UGX_W_RHOM] * FN / tAv;
            const double keMean = a[UGX_W_LINEAR_KE] * FN / tAv;
            const double rhoNMean = a[UGX_W_RHON] * FN / tAv;
            T = 2.0 / (3.0 * kB * rhoNMean) * (keMean - 0.5 * rhoMMean * (U[0] * U[0] + U[1] * U[1] + U[2] * U[2]));
        }
        const double Trot = (a[UGX_W_ROTDOF] > kVSmall) ? (2.0 / kB) * (a[UGX_W_EROT] / a[UGX_W_ROTDOF]) : 0.0;
        const double nRotDof = (a[UGX_W_RHON] > kVSmall) ? a[UGX_W_ROTDOF] / a[UGX_W_RHON] : 0.0;
        const double Tov = (3.0 * T + nRotDof * Trot) / (3.0 + nRotDof);
        double q = 0.0, fD[3] = {0.0, 0.0, 0.0};
        if (tAv > 0.0) {
            q = a[UGX_W_Q] / tAv;
            for (int d = 0; d < 3; ++d) fD[d] = a[UGX_W_FX + d] / tAv;
        }
        const double* gq = &g.wgeo[(size_t)k * kWGeo];
        const double n[3] = {gq[0], gq[1], gq[2]}, t1[3] = {gq[4], gq[5], gq[6]};
        double t2[3] = {n[1] * t1[2] - n[2] * t1[1], n[2] * t1[0] - n[0] * t1[2], n[0] * t1[1] - n[1] * t1[0]};
        const double m2 = std::sqrt(t2[0] * t2[0] + t2[1] * t2[1] + t2[2] * t2[2]);
        if (m2 > kVSmall) for (int d = 0; d < 3; ++d) t2[d] /= m2;
        const double fn = fD[0] * n[0] + fD[1] * n[1] + fD[2] * n[2];
        const double f1 = fD[0] * t1[0] + fD[1] * t1[1] + fD[2] * t1[2];
        const double f2 = fD[0] * t2[0] + fD[1] * t2[1] + fD[2] * t2[2];
        if (f->rhoN) f->rhoN[k] = rhoN;
        if (f->rhoM) f->rhoM[k] = rhoM;
        if (f->U_mean) for (int d = 0; d < 3; ++d) f->U_mean[3 * (size_t)k + d] = U[d];
        if (f->translational_T) f->translational_T[k] = T;
        if (f->rotational_T) f->rotational_T[k] = Trot;
        if (f->overall_T) f->overall_T[k] = Tov;
        if (f->heat_flux) f->heat_flux[k] = q;
        if (f->force_density) for (int d = 0; d < 3; ++d) f->force_density[3 * (size_t)k + d] = fD[d];
        if (f->pressure) f->pressure[k] = fn;
        if (f->shear) f->shear[k] = std::sqrt(f1 * f1 + f2 * f2);
    }
    return UGX_OK;
}

}  // namespace ugx
