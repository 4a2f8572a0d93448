"""macroInterpolation on the GPU (interpolationCellPoint at
stochasticParticleBGK.C:645-688, stochasticParticleESBGK.C:741-790,
unifiedStochasticParticleSBGK.C:893-946): the BGK-family draws with the
fields interpolated at each relaxed simulator are bit-identical to the
oracle for the three models (a Couette channel with cyclic sides joining
points across the channel, a graded wall box, the hybrid nozzle with
inflow), and the FP32 state runs with finite velocities."""
import numpy as np
import pytest

from oracle.orc import OracleCase
from unigasfoam_amd import abi, workloads
from unigasfoam_amd.core import Case

from test_gpu_parity import assert_identical
from test_oracle_physics import box_case

pytestmark = pytest.mark.gpu

BGK_MODELS = ["stochasticParticleBGK", "stochasticParticleESBGK", "unifiedStochasticParticleSBGK"]


def couette_macro(bgk):
    spec, fill = workloads.couette(n_cells=24, per_cell=60, collision_model="bgk", bgk_model=bgk, seed=11, dt=5e-8)
    spec.macro_interpolation = True
    return spec, fill


def graded_box_macro(bgk):
    spec = box_case(n=(6, 4, 3), side=2e-4, fn=2e7, collision="bgk", bgk=bgk, dt=2e-8, kind=abi.BND_DIFFUSE,
                    grading=(5.0, 1.0, 1.0), seed=13, T_wall=400.0)
    spec.macro_interpolation = True
    return spec, {"number_density": {"Ar": 1.0e23}, "T": 300.0, "velocity": (200.0, 0.0, 0.0)}


def nozzle_macro():
    spec, fill = workloads.nozzle(nx=10, ny=6, nz=6, per_cell=20, n_stag=2.4e24)
    spec.decomposition_interval = 4
    spec.smoothing_passes = 2
    spec.macro_interpolation = True
    return spec, fill


CASES = {f"couette_{b}": (lambda b=b: couette_macro(b)) for b in BGK_MODELS}
CASES.update({f"graded_{b}": (lambda b=b: graded_box_macro(b)) for b in BGK_MODELS})
CASES["hybrid_nozzle_usp"] = nozzle_macro


@pytest.mark.parametrize("name", sorted(CASES))
def test_macro_interpolation_identical_to_oracle(require_gpu, name):
    spec, fill = CASES[name]()
    g = Case(spec, precision="fp64", capacity=1 << 17)
    o = OracleCase(spec, capacity=1 << 17)
    for c in (g, o):
        c.fill(**fill)
    for k in range(3):
        g.step(4)
        o.step(4)
        assert_identical(g, o, f"{name} after {4 * (k + 1)} steps")
    assert g.counters()["bgk_collisions"] > 500
    g.close()
    o.close()


def test_macro_interpolation_changes_the_draws(require_gpu):
    spec, fill = couette_macro("stochasticParticleBGK")
    out = []
    for macro in (False, True):
        spec.macro_interpolation = macro
        g = Case(spec, precision="fp64", capacity=1 << 17)
        g.fill(**fill)
        g.step(3)
        g.sync()
        out.append(g.particles()["velocity"])
        g.close()
    assert out[0].shape == out[1].shape and not np.array_equal(out[0], out[1])


@pytest.mark.parametrize("bgk", BGK_MODELS)
def test_macro_interpolation_fp32_runs(require_gpu, bgk):
    spec, fill = couette_macro(bgk)
    g = Case(spec, precision="fp32", capacity=1 << 17)
    g.fill(**fill)
    g.step(20)
    g.sync()
    v = g.particles()["velocity"]
    assert np.isfinite(v).all()
    assert g.counters()["bgk_collisions"] > 1000
    g.close()
