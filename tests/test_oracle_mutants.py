"""Planted mistakes: each pin of the oracle must fail when a plausible error
is written into a copy of oracle/orc.c.

For every entry the test edits one snippet of orc.c (the snippet must occur
exactly once, so an edit of the oracle that moves it fails here loudly),
compiles the copy with the oracle's flags into a temporary library, and runs
the named pins against that library through ``orc.using``: every one of them
must raise AssertionError.  The unedited oracle passes the same pins in
tests/test_oracle_pins.py and tests/test_oracle_physics.py.
"""
import concurrent.futures
import os
import shutil
import subprocess

import pytest

import test_oracle_physics as phys
import test_oracle_adapt as adapt
import test_oracle_macro as macro
import test_oracle_pins as pins
import test_oracle_subcells as subcells
import test_oracle_wall as wall
import test_oracle_weighting as weighting
from oracle import orc

ORC_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
CFLAGS = ["-O2", "-std=c11", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-shared"]

# name: (snippet, replacement, pins that must fail)
MUTANTS = {
    # a tet weight dropped from the centroid's barycentric coordinate (interpolationCellPoint)
    "macro_barycentric": (
        "            double l0 = ((1.0 - l1) - l2) - l3;\n            double mn",
        "            double l0 = (1.0 - l1) - l2;\n            double mn",
        [lambda: macro.test_constant_field_is_reproduced(False), macro.test_linear_field_is_exact_in_interior_cells]),
    # the draws with the cell's fields although macroInterpolation is on (stochasticParticleBGK.C:680-687)
    "macro_draw_cell_values": (
        "            if (w->macro) {\n                double f[MIR];",
        "            if (0) {\n                double f[MIR];",
        [lambda: macro.test_relaxed_draws_keep_the_local_shear(True)]),
    # weights blended with the weights swapped (uniGasDynamicAdapter.C:665)
    "adapt_weight_blend": (
        "w->cw[c] = theta * wn[c] + (1.0 - theta) * w->cw[c];",
        "w->cw[c] = (1.0 - theta) * wn[c] + theta * w->cw[c];",
        [adapt.test_weight_adaptation_blends_and_reweights]),
    # fvc::smooth's ratio taken as the coefficient itself, 1.3 for 1 + 1.3 (:456)
    "adapt_max_ratio_coefficient": (
        "    if (w->fkind[k] != FACE_INTERNAL && w->fkind[k] != FACE_CYCLIC) continue;\n                double v = f[w->nbr[k]] / max_ratio;",
        "    if (w->fkind[k] != FACE_INTERNAL && w->fkind[k] != FACE_CYCLIC) continue;\n                double v = f[w->nbr[k]] / (max_ratio - 1.0);",
        [adapt.test_max_ratio_smoothing_is_the_decayed_maximum]),
    # the re-weighting after an adaptation from the start cells instead of the old weights (uniGasCloud.C:228-232)
    "adapt_reweight_old_weights": (
        "double wo = cw_old ? cw_old[c] : w->cw[w->cstart[i]], wn = w->cw[c];",
        "double wo = w->cw[w->cstart[i]], wn = w->cw[c];",
        [adapt.test_weight_adaptation_blends_and_reweights, adapt.test_weight_adaptation_keeps_the_molecule_count]),
    # collision rate with the exponent omega for 1 - omega (uniGasDynamicAdapter.C:262-263, Bird eq. 4.74)
    "adapt_collision_rate_exponent": (
        "pow(T / Tref, 1.0 - omPQ)",
        "pow(T / Tref, omPQ)",
        [adapt.test_fill_mct_limited_step_and_levels, adapt.test_adaptation_from_time_sums]),
    # mean free path without the mass-ratio factor (:266, Bird eq. 4.76)
    "adapt_mfp_mass_ratio": (
        "sqrt(1.0 + massRatio)",
        "sqrt(massRatio)",
        [adapt.test_fill_mct_limited_step_and_levels, adapt.test_empty_direction_counts_for_nothing]),
    # Courant number from the most probable speed alone (:297-299)
    "adapt_courant_ignores_drift": (
        "double v = U[d] > u0 ? U[d] : u0;",
        "double v = u0;",
        [adapt.test_fill_courant_limited_step_with_drift]),
    # empty directions counted in the largest Courant number (:344-350)
    "adapt_courant_empty_direction": (
        "if (w->mask[d] != 0.0 && b < co[3 * c + d]) b = co[3 * c + d];",
        "if (b < co[3 * c + d]) b = co[3 * c + d];",
        [adapt.test_empty_direction_counts_for_nothing]),
    # the collision-time limit taken over BGK-family cells too (:326-334)
    "adapt_mct_all_cells": (
        "if (!usp && w->cmodel[c] != 0 && a < mct[c]) a = mct[c];",
        "if (a < mct[c]) a = mct[c];",
        [lambda: adapt.test_bgk_cells_count_only_for_courant("stochasticParticleBGK"),
         adapt.test_hybrid_dsmc_cells_set_step_and_levels_usp_cells_do_not]),
    # an empty cell keeps zero density instead of the smallest (:597-604)
    "adapt_empty_cell_density": (
        "            rhoN[c] = minRhoN;\n",
        "            rhoN[c] = rhoN[c];\n",
        [adapt.test_adaptation_from_time_sums]),
    # NTC second candidate always from the whole cell (noTimeCounter.C:193-225)
    "ntc_subcells_ignored": (
        "        if (nSC > 1) {\n            do { Q = (uint32_t)list[",
        "        if (0) {\n            do { Q = (uint32_t)list[",
        [lambda: subcells.test_subcell_momentum_is_conserved((2, 2, 1)),
         lambda: subcells.test_subcell_momentum_is_conserved((1, 1, 4))]),
    # subcell index with the y axis unweighted: subcells (1,0) and (0,1) share a list (:151-155)
    "subcell_index_weights": (
        "int s = k0 + lv[0] * (k1 + lv[1] * k2);",
        "int s = k0 + k1 + lv[1] * k2;",
        [lambda: subcells.test_subcell_momentum_is_conserved((2, 2, 1)),
         lambda: subcells.test_subcell_momentum_is_conserved((2, 3, 1))]),
    # VSS deflection exponent R^alpha for R^(1/alpha) (variableSoftSphere.C:147)
    "vss_deflection_exponent": (
        "2.0 * orc_pow(orc_uniform(r), 1.0 / alpha) - 1.0",
        "2.0 * orc_pow(orc_uniform(r), alpha) - 1.0",
        [lambda: pins.test_vss_deflection_law("variableSoftSphere", 1.4),
         lambda: pins.test_vss_deflection_law("LarsenBorgnakkeVariableSoftSphere", 1.4)]),
    # BGK reference viscosity doubled (stochasticParticleBGK.C:548-552)
    "bgk_mu_ref_doubled": (
        "w->muref[s] = 1.25 * (1.0 + a)",
        "w->muref[s] = 2.5 * (1.0 + a)",
        [lambda: pins.test_bgk_relaxed_count_chapman_enskog("stochasticParticleBGK"),
         lambda: pins.test_bgk_relaxed_count_chapman_enskog("stochasticParticleESBGK"),
         lambda: pins.test_bgk_anisotropy_decay("stochasticParticleBGK"),
         lambda: pins.test_bgk_anisotropy_decay("stochasticParticleESBGK")]),
    # ES-BGK frequency without Pr (stochasticParticleESBGK.C:641)
    "esbgk_nu_without_pr": (
        "nu = (model == UGX_BGK_ESBGK) ? Pr * p / visc : p / visc;",
        "nu = p / visc;",
        [lambda: pins.test_bgk_relaxed_count_chapman_enskog("stochasticParticleESBGK"),
         lambda: pins.test_bgk_anisotropy_decay("stochasticParticleESBGK")]),
    # USP heat flux without the E_rot c term (unifiedStochasticParticleSBGK.C:518-550)
    "usp_heat_flux_without_erot": (
        "\n                               + S[UGX_S_EROTU + d] / N - (S[UGX_S_EROT] / N) * U[d])",
        ")",
        [lambda: pins.test_usp_relaxed_moments(1.0)]),
    # USP heat flux scaled N/(N-1) instead of N^2/((N-1)(N-2)) (:553-559)
    "usp_heat_flux_factor": (
        "double f2 = N * N / (N - 1.0) / (N - 2.0);",
        "double f2 = N / (N - 1.0);",
        [lambda: pins.test_usp_relaxed_moments(1.0)]),
    # USP stress relaxed by 1/(1 + nu dt) (:780)
    "usp_stress_relaxation_full_step": (
        "sh[k] = th * sh[k] / (1.0 + 0.5 * nu * dt)",
        "sh[k] = th * sh[k] / (1.0 + nu * dt)",
        [lambda: pins.test_usp_relaxed_moments(1.0)]),
    # USP heat flux relaxed without Pr (:779)
    "usp_heat_flux_relaxation_without_pr": (
        "q[d] = th * q[d] / (1.0 + 0.5 * Pr * nu * dt)",
        "q[d] = th * q[d] / (1.0 + 0.5 * nu * dt)",
        [lambda: pins.test_usp_relaxed_moments(1.0)]),
    # USP theta blend with the wrong weight on the previous step (:779-780)
    "usp_theta_blend": (
        "+ (1.0 - th) * w->uspq[3 * c + d];",
        "+ th * w->uspq[3 * c + d];",
        [lambda: pins.test_usp_relaxed_moments(0.7)]),
    # USP sampling coefficient cQ without Pr (:1063)
    "usp_cq_without_pr": (
        "double cQ = 2.0 * (1.0 - Pr * tau * ex);",
        "double cQ = 2.0 * (1.0 - tau * ex);",
        [pins.test_usp_relaxation_is_trapezoidal]),
    # USP maxProb decay 0.99 instead of 0.999 (:58, :861)
    "usp_max_prob_decay": (
        "m0 * 0.999",
        "m0 * 0.99",
        [pins.test_usp_max_prob_decay_and_raise]),
    # NTC accepting against the running (sigma c_r)max (noTimeCounter.C:263-268)
    "ntc_running_max": (
        "if (scr / smax0 > orc_uniform(&r))",
        "if (scr / smax > orc_uniform(&r))",
        [pins.test_ntc_accepts_against_start_of_cell_max]),
    # NTC selection number without the factor 1/2 (noTimeCounter.C:184)
    "ntc_half_dropped": (
        "double sel = 0.5 * (double)nC * (double)(nC - 1)",
        "double sel = (double)nC * (double)(nC - 1)",
        [phys.test_ntc_candidates_and_vhs_rate]),
    # VHS cross-section exponent (variableHardSphere.C:109-114)
    "vhs_exponent": (
        "w->pe[k] = 1.0 - om;",
        "w->pe[k] = 1.5 - om;",
        [phys.test_ntc_candidates_and_vhs_rate]),
    # Larsen-Borgnakke energy without the rotational energy of P (LB-VHS.C:301-312)
    "lb_erot_left_out": (
        "double Ec = Et + preP;",
        "double Ec = Et;",
        [lambda: phys.test_collisions_conserve_momentum_and_energy(
            "LarsenBorgnakkeVariableHardSphere", (phys.N2, phys.O2))]),
    # LB 2-DoF ratio exponent (uniGasCloud.C:1137-1140)
    "lb_ratio_exponent": (
        "if (dof == 2) return 1.0 - orc_pow(orc_uniform(r), 1.0 / chiB);",
        "if (dof == 2) return 1.0 - orc_pow(orc_uniform(r), chiB);",
        [phys.test_lb_rotational_equilibrium]),
    # BGK conservation temperature with N instead of N - 1 (stochasticParticleBGK.C:761-777)
    "bgk_conservation_n_minus_1": (
        "/ (3.0 * KB * (N - 1.0));",
        "/ (3.0 * KB * N);",
        [lambda: phys.test_bgk_conserves_each_cell("stochasticParticleBGK")]),
    # ES-BGK S matrix sign (stochasticParticleESBGK.C:902-904)
    "esbgk_sign": (
        "Smat[k] = dlt - a * (Pt[full[k]] / p - dlt);",
        "Smat[k] = dlt + a * (Pt[full[k]] / p - dlt);",
        [lambda: phys.test_bgk_strong_relaxation_limit("stochasticParticleESBGK")]),
    # diffuse wall normal speed sqrt(-ln R) instead of sqrt(-2 ln R) (uniGasPatchBoundary.C:356-371)
    "diffuse_normal_speed": (
        "double un = sqrt(-2.0 * orc_log(orc_uniform(r)));",
        "double un = sqrt(-orc_log(orc_uniform(r)));",
        [phys.test_diffuse_walls_thermalise_gas]),
    # inflow count without the exp(-s^2) term (uniGasGeneralBoundary.C:140-157)
    "inflow_exp_term": (
        "e->flux_br = exp(-(sCos * sCos)) + sqrtPi",
        "e->flux_br = 0.0 + sqrtPi",
        [phys.test_inflow_count_bird_4_22]),
    # wall estimator without the face area (uniGasPatchBoundary.C:155)
    "wall_weight_without_area": (
        "double inv = 1.0 / fmax(fabs(Un) * fA, VSMALL);",
        "double inv = 1.0 / fmax(fabs(Un), VSMALL);",
        [lambda: wall.test_diffuse_wall_sums_match_kinetic_theory(wall.AR),
         wall.test_specular_wall_is_exact_and_gives_the_pressure]),
    # wall heat flux after minus before (uniGasPatchBoundary.C:295)
    "wall_heat_flux_sign": (
        "a[UGX_W_Q] += FN * (preIE - postIE) / (dt * fA);",
        "a[UGX_W_Q] += FN * (postIE - preIE) / (dt * fA);",
        [lambda: wall.test_diffuse_wall_sums_match_kinetic_theory(wall.AR)]),
    # wall energy without E_rot (uniGasPatchBoundary.C:187-190, 257-260)
    "wall_energy_without_erot": (
        "e->u0[2] * e->u0[2]) + e->e0;",
        "e->u0[2] * e->u0[2]);",
        [lambda: wall.test_diffuse_wall_sums_match_kinetic_theory(wall.N2)]),
    # wall force from the momentum after minus before (uniGasPatchBoundary.C:296)
    "wall_force_sign": (
        "a[UGX_W_FX + d] += FN * (m * e->u0[d] - m * e->u1[d]) / (dt * fA);",
        "a[UGX_W_FX + d] += FN * (m * e->u1[d] - m * e->u0[d]) / (dt * fA);",
        [wall.test_moving_wall_tangential_force, wall.test_specular_wall_is_exact_and_gives_the_pressure]),
    # only the before-interaction terms measured (uniGasPatchBoundary.C:217-254)
    "wall_after_terms_dropped": (
        "        wall_terms(e->u1, e->e1, n, fA, m, rotdof, a);\n",
        "",
        [lambda: wall.test_diffuse_wall_sums_match_kinetic_theory(wall.N2),
         wall.test_moving_wall_tangential_force]),
    # E_rot before and after swapped (uniGasPatchBoundary.C:177, 247)
    "wall_erot_swapped": (
        "wall_event(w, wf, sp, slot, it, pre, U, e0, *erot);",
        "wall_event(w, wf, sp, slot, it, pre, U, *erot, e0);",
        [lambda: wall.test_diffuse_wall_sums_match_kinetic_theory(wall.N2)]),
    # every step summed, not only the sampling steps (uniGasVolFields.C:728-830)
    "wall_every_step": (
        "const int meas = w->nwf > 0 && sampling_step(w);",
        "const int meas = w->nwf > 0;",
        [wall.test_only_sampling_steps_are_summed]),
    # clone probability W_o/W_n instead of W_o/W_n - 1 (uniGasCloud.C:1367)
    "weight_clone_probability": (
        "double prob = wo / wn - 1.0;",
        "double prob = wo / wn;",
        [weighting.test_clones_and_removals_keep_the_physical_density]),
    # removal test inverted (uniGasCloud.C:1409-1418)
    "weight_removal_inverted": (
        "} else if (wo / wn < orc_uniform(&r)) {",
        "} else if (wn / wo < orc_uniform(&r)) {",
        [weighting.test_clones_and_removals_keep_the_physical_density]),
    # NTC selection number without the cell weight (noTimeCounter.C:184)
    "weight_ntc": (
        "w->model.n_equivalent_particles * w->cw[c] * smax0",
        "w->model.n_equivalent_particles * smax0",
        [weighting.test_collision_rate_per_volume_independent_of_weight]),
    # inflow count without the cell weight (uniGasGeneralBoundary.C:165)
    "weight_inflow": (
        "(w->inflow[k].nbar_den * w->cw[w->inflow[k].cell])",
        "(w->inflow[k].nbar_den)",
        [weighting.test_inflow_count_divided_by_the_cell_weight]),
    # fill count without the cell weight (uniGasMeshFill.C:203-205)
    "weight_fill": (
        "w->vol[c] / w->cw[c];",
        "w->vol[c];",
        [weighting.test_fill_weights_uniform_mesh_exact, weighting.test_fill_weights_graded_mesh_smoothed_and_capped]),
    # smoothing capped at N_pps instead of minParticlesPerSubCell (uniGasMeshFill.C:141-142)
    "weight_cap": (
        "double lim = (ntot * w->vol[c]) / ((double)nmin * nsub[c] * FN * 1.0);",
        "double lim = (ntot * w->vol[c]) / ((double)npps * nsub[c] * FN * 1.0);",
        [weighting.test_fill_weights_graded_mesh_smoothed_and_capped]),
}


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    gcc = shutil.which(os.environ.get("CC", "gcc"))
    if gcc is None:
        pytest.skip("no C compiler")
    src = open(os.path.join(ORC_DIR, "orc.c")).read()
    base = tmp_path_factory.mktemp("mutants")
    root = base / "oracle"           # orc.h includes ../include/ugx.h
    os.makedirs(root)
    os.makedirs(base / "include")
    for name in ("orc.h", "orc_prim.h"):
        shutil.copy(os.path.join(ORC_DIR, name), root / name)
    shutil.copy(os.path.join(ORC_DIR, "..", "include", "ugx.h"), base / "include" / "ugx.h")

    def build(name):
        old, new, _ = MUTANTS[name]
        assert src.count(old) == 1, f"{name}: snippet occurs {src.count(old)} times in orc.c"
        c = root / f"{name}.c"
        c.write_text(src.replace(old, new))
        so = root / f"lib{name}.so"
        subprocess.run([gcc, *CFLAGS, "-o", str(so), str(c), "-lm"], check=True, capture_output=True)
        return name, str(so)

    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        return dict(ex.map(build, sorted(MUTANTS)))


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_planted_mistake_is_caught(name, mutant_libs):
    _, _, checks = MUTANTS[name]
    with orc.using(mutant_libs[name]):
        for i, check in enumerate(checks):
            with pytest.raises(AssertionError):
                check()
