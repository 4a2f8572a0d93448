"""The dynamic adapter on the GPU (uniGasDynamicAdapter.C; uniGasMeshFill.C:
100-168): the time step and subcell levels set at the fill and every
adaptation interval are those of the oracle, and the run stays
bit-identical to the oracle across the changes (collisions, a freestream
inflow whose counts follow the time step, cell weights of a weighted fill
with subcells and their adaptation with the re-weighting of every
simulator, the hybrid switch), the same alone and in a batch, and
through a restart between two adaptations."""
import numpy as np
import pytest

from oracle.orc import OracleCase
from unigasfoam_amd import abi, workloads
from unigasfoam_amd.batch import Batch
from unigasfoam_amd.core import Case

from test_gpu_parity import assert_identical
from test_oracle_adapt import adaptive_box, mfp, symmetric_box, AR
from test_oracle_physics import SIDES, box_case
from unigasfoam_amd.spec import Boundary

pytestmark = pytest.mark.gpu


def inflow_case(weighted=False):
    nd = 5.0e20 if weighted else 1.0e21
    bnd = [Boundary("xmin", abi.BND_FREESTREAM, temperature=250.0, velocity=(400.0, 0.0, 0.0),
                    number_density={"Ar": nd}),
           Boundary("xmax", abi.BND_DELETION)] + [Boundary(p, abi.BND_DIFFUSE, temperature=300.0) for p in SIDES[2:]]
    spec = box_case(n=(6, 3, 3), side=0.01, fn=5.0e10, collision="dsmc", dt=2e-7, seed=3, bnd=bnd,
                    grading=(3.0, 1.0, 1.0) if weighted else None)
    for p in ("xmin", "xmax"):
        spec.mesh.patches[spec.mesh.patch_index(p)].type = abi.PATCH
    spec.adaptive = True
    spec.time_step_adaptation = True
    spec.subcell_adaptation = True
    spec.adaptation_interval = 6
    if weighted:
        spec.cell_weighted = True
        spec.particles_per_subcell = 4
        spec.max_cell_weight_ratio = 0.5
        spec.cell_weight_adaptation = True
    return spec, {"number_density": {"Ar": nd}, "T": 300.0}


def nozzle_case():
    spec, fill = workloads.nozzle(nx=10, ny=6, nz=6, per_cell=20, n_stag=2.4e24,
                                  bgk_model="stochasticParticleESBGK")
    spec.decomposition_interval = 4
    spec.smoothing_passes = 2
    spec.adaptive = True
    spec.time_step_adaptation = True
    spec.subcell_adaptation = True
    spec.adaptation_interval = 5
    return spec, fill


def closed_box():
    return adaptive_box(L=1.5e-4, interval=5, fn=2e8), {"number_density": {"Ar": 1.2e22}, "T": 300.0}


def symmetry_box():
    """A symmetryPlane at y = 0: the smoothed size ratios reflect on it."""
    n, T = 1.2e22, 300.0
    return (symmetric_box(2.25 * mfp(AR, n, T), passes=3, interval=5, fn=4e8),
            {"number_density": {"Ar": n}, "T": T})


CASES = {"closed_box": closed_box, "symmetry_box": symmetry_box, "inflow": inflow_case,
         "inflow_weighted": lambda: inflow_case(True),
         "hybrid_nozzle": nozzle_case}


def same_adaptation(g, o, what):
    assert g.delta_t() == o.delta_t(), f"{what}: time step {g.delta_t()} vs {o.delta_t()}"
    assert g.time() == o.time(), f"{what}: time"
    assert np.array_equal(g.subcell_levels(), o.subcell_levels()), f"{what}: subcell levels differ"


@pytest.mark.parametrize("name", sorted(CASES))
def test_adaptive_case_identical_to_oracle(require_gpu, name):
    spec, fill = CASES[name]()
    g = Case(spec, precision="fp64", capacity=1 << 18)
    o = OracleCase(spec, capacity=1 << 18)
    for c in (g, o):
        c.fill(**fill)
    same_adaptation(g, o, f"{name} fill")
    assert g.delta_t() != spec.delta_t
    assert np.array_equal(g.cell_weights(), o.cell_weights())
    seen = {g.delta_t()}
    for k in range(4):
        g.step(4)
        o.step(4)
        assert_identical(g, o, f"{name} after {4 * (k + 1)} steps")
        same_adaptation(g, o, f"{name} after {4 * (k + 1)} steps")
        seen.add(g.delta_t())
    assert len(seen) >= 3, "the time step should change at the fill and at each adaptation"
    if name != "hybrid_nozzle":
        assert g.subcell_levels().max() > 1
    g.close()
    o.close()


def test_adaptive_restart_between_adaptations(require_gpu):
    spec, fill = inflow_case()
    a = Case(spec, precision="fp64", capacity=1 << 18)
    a.fill(**fill)
    a.step(9)                       # one adaptation done, three steps into the next interval
    a.sync()
    ck = a.checkpoint()
    a.step(8)
    a.sync()
    b = Case.restart(spec, ck, capacity=1 << 18)
    b.step(8)
    b.sync()
    assert a.delta_t() == b.delta_t() and a.time() == b.time()
    assert np.array_equal(a.subcell_levels(), b.subcell_levels())
    pa, pb = a.particles(), b.particles()
    for k in ("local", "velocity", "cell"):
        assert np.array_equal(pa[k], pb[k]), k
    assert np.array_equal(a.samples(), b.samples())
    assert a.counters()["time_sampled"] == b.counters()["time_sampled"]
    a.close()
    b.close()


def test_adaptive_case_alone_equals_batch(require_gpu):
    specs = [inflow_case(), closed_box()]
    alone = []
    for s, fill in specs:
        c = Case(s, precision="fp64", capacity=1 << 18)
        c.fill(**fill)
        c.step(13)
        c.sync()
        alone.append((c.samples(), c.particles()["velocity"], c.delta_t(), c.subcell_levels()))
        c.close()
    b = Batch(precision="fp64", use_graphs=False)
    slots = [b.admit(s, fill, 1 << 16) for s, fill in specs]
    b.step(13)
    b.sync()
    for (sa, va, dta, lva), sl in zip(alone, slots):
        c = b.case(sl)
        assert np.array_equal(c.samples(), sa)
        assert np.array_equal(c.particles()["velocity"], va)
        assert c.delta_t() == dta
        assert np.array_equal(c.subcell_levels(), lva)
    b.close()
