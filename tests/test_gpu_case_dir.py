"""Cases read from OpenFOAM dictionaries, run through the C ABI on the GPU.

cylinderWake gathers the reference tutorials' switches on a small mesh.
platePlume is a three-block N2/O2 plate flow with freestream inflow,
deletion, a diffuse wall at 500 K, symmetry and empty patches and
Larsen-Borgnakke collisions (config C3's physics on 2400 cells); couetteN2
is config C2.  The FP64 GPU path equals the oracle bit for bit on both; the
FP32-state run of the plate case agrees with the FP64 one statistically.
"""
import os
import shutil

import numpy as np
import pytest

from oracle.orc import OracleCase
from unigasfoam_amd import case_dir, foam_io
from unigasfoam_amd.core import Case
from unigasfoam_amd.run import run_case

from test_gpu_parity import assert_identical

pytestmark = pytest.mark.gpu
CASES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cases")


def test_cylinder_wake_case_identical_to_oracle(require_gpu):
    """cylinderWake: the reference tutorials' feature set on 160 cells: an
    arc-edged half annulus around a diffuse cylinder, freestream inflow,
    symmetry and empty patches, hybrid USP/DSMC with macroInterpolation,
    cell weighting, and time-step, subcell and cell-weight adaptation every
    five steps.  Bit-identical to the oracle, including the adapted time
    step, the simulated time, the subcell levels and the weights."""
    spec, fills, run = case_dir.read_case(os.path.join(CASES, "cylinderWake"))
    assert spec.macro_interpolation and spec.adaptive and spec.cell_weight_adaptation and spec.collision_model == "hybrid"
    g = Case(spec, precision="fp64", capacity=1 << 18)
    o = OracleCase(spec, capacity=1 << 18)
    for f in fills:
        g.fill(**f)
        o.fill(**f)
    for k in range(3):
        g.step(5)
        o.step(5)
        what = f"cylinderWake after {5 * (k + 1)} steps"
        assert_identical(g, o, what)
        assert g.delta_t() == o.delta_t() and g.time() == o.time(), what
        assert np.array_equal(g.subcell_levels(), o.subcell_levels()), what
        assert np.array_equal(g.cell_weights(), o.cell_weights()), what
    c = g.counters()
    assert c["collisions"] > 100 and c["bgk_collisions"] > 1000 and c["inserted"] > 100 and c["cloned"] > 100
    assert 0 < g.cell_model().sum() < spec.mesh.n_cells and g.subcell_levels().max() > 1
    g.close()
    o.close()


def test_plate_case_identical_to_oracle(require_gpu):
    spec, fills, run = case_dir.read_case(os.path.join(CASES, "platePlume"))
    g = Case(spec, precision="fp64", capacity=1 << 18)
    o = OracleCase(spec, capacity=1 << 18)
    for f in fills:
        g.fill(**f)
        o.fill(**f)
    for k in range(3):
        g.step(10)
        o.step(10)
        assert_identical(g, o, f"platePlume after {10 * (k + 1)} steps")
    c = g.counters()
    assert c["inserted"] > 1000 and c["deleted"] > 1000 and c["wall_hits"] > 0 and c["collisions"] > 1000


def test_run_driver_writes_fields_and_matches_oracle(require_gpu, tmp_path):
    case = tmp_path / "couette"
    shutil.copytree(os.path.join(CASES, "couetteN2"), case)
    spec, fills, run = case_dir.read_case(str(case), seed=7)
    n_steps = 40
    with open(os.devnull, "w") as devnull:
        g, written = run_case(str(case), precision="fp64", end_time=n_steps * run.delta_t, seed=7,
                              out=devnull, write=False)
    o = OracleCase(spec, capacity=1 << 18)
    for f in fills:
        o.fill(**f)
    o.step(n_steps)
    assert_identical(g, o, "couette from dictionaries")
    # the written vol fields: the sampler at the end time, under uniGasVolFields' names
    fields = g.fields()
    patch, face = g.wall_faces()
    wf = g.wall_fields()
    foam_io.write_fields(str(case), "2e-07", spec.mesh, fields, run.field_name,
                         wall={"patch": patch, "face": face, "fields": wf})
    rho = foam_io.read_internal_field(str(case / "2e-07" / "rhoN_N2"))
    assert np.allclose(rho, fields["rhoN"], rtol=1e-15)
    assert abs(rho.mean() / 1.3e22 - 1.0) < 0.01
    # the walls' measured values: 273 K walls under a gas filled at 273 K
    for k, name in enumerate(spec.mesh.patches[i].name for i in patch):
        q = foam_io.read_patch_field(str(case / "2e-07" / "surfaceHeatTransfer_N2"), name)
        assert np.array_equal(q, wf["heat_flux"][k:k + 1])
        p = foam_io.read_patch_field(str(case / "2e-07" / "p_N2"), name)
        assert p[0] == wf["pressure"][k] and abs(p[0] / (1.3e22 * 1.380649e-23 * 273.0) - 1.0) < 0.1
    g.close()


def test_plate_case_fp32_state_agrees_with_fp64(require_gpu):
    """C3's precision on the plate case: FP32 particle state with FP64
    samplers against the all-FP64 run, same seed; the inflow counts are
    computed in FP64 in both and agree exactly, the sampled domain means of
    density and temperature within their sampling noise."""
    spec, fills, run = case_dir.read_case(os.path.join(CASES, "platePlume"))
    res = {}
    for prec in ("fp64", "fp32"):
        g = Case(spec, precision=prec, capacity=1 << 19)
        for f in fills:
            g.fill(**f)
        g.step(100)
        g.reset_samples()
        g.step(200)
        g.sync()
        f = g.fields()
        c = g.counters()
        res[prec] = (f, c)
        g.close()
    (f64, c64), (f32, c32) = res["fp64"], res["fp32"]
    assert c64["inserted"] == c32["inserted"]
    n = c64["n_particles"]
    assert abs(c32["n_particles"] - n) < 5 * np.sqrt(n)
    for k in ("rhoN", "translational_T", "rotational_T"):
        a, b = f64[k].mean(), f32[k].mean()
        assert abs(a - b) <= 0.01 * abs(a), (k, a, b)
