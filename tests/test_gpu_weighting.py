"""Cell weighting on the GPU (uniGasCloud.C:1353-1425; uniGasMeshFill.C:
111-168): bit-identical to the oracle, the same alone and in a batch, carried
through a restart, and the sampled physical density flat where the weights
change tenfold (the write-time fields carry F_N W_c per simulator)."""
import numpy as np
import pytest

from oracle.orc import OracleCase
from unigasfoam_amd import abi
from unigasfoam_amd.batch import Batch
from unigasfoam_amd.core import Case

from test_gpu_parity import assert_identical
from test_oracle_physics import SIDES, box_case
from test_oracle_weighting import weighted
from unigasfoam_amd.spec import Boundary

pytestmark = pytest.mark.gpu


def graded_box(npps=120, dsmc="variableHardSphere", seed=7):
    spec = box_case(n=(8, 2, 2), side=0.01, fn=2.0e9, collision="dsmc", dsmc=dsmc, dt=4e-6,
                    grading=(25.0, 1.0, 1.0), seed=seed)
    return weighted(spec, npps=npps, ratio=0.5)


def inflow_box(seed=3):
    spec = box_case(n=(6, 3, 3), side=0.01, fn=5.0e10, collision="dsmc", dt=1e-6, grading=(4.0, 1.0, 1.0), seed=seed,
                    bnd=[Boundary("xmin", abi.BND_FREESTREAM, temperature=250.0, velocity=(300.0, 0.0, 0.0),
                                  number_density={"Ar": 1.0e20}),
                         Boundary("xmax", abi.BND_DELETION)] + [Boundary(p, abi.BND_DIFFUSE, temperature=300.0)
                                                                 for p in SIDES[2:]])
    spec.mesh.patches[spec.mesh.patch_index("xmin")].type = abi.PATCH
    spec.mesh.patches[spec.mesh.patch_index("xmax")].type = abi.PATCH
    return weighted(spec, npps=8)


@pytest.mark.parametrize("make", [graded_box, inflow_box], ids=["graded_box", "inflow_box"])
def test_weighted_case_identical_to_oracle(require_gpu, make):
    spec = make()
    g = Case(spec, precision="fp64", capacity=1 << 18)
    o = OracleCase(spec, capacity=1 << 18)
    for c in (g, o):
        c.fill({"Ar": 1.0e20}, 300.0)
    assert np.array_equal(g.cell_weights(), o.cell_weights())
    for k in range(3):
        g.step(10)
        o.step(10)
        assert_identical(g, o, f"{make.__name__} after {10 * (k + 1)} steps")
    c = g.counters()
    assert c["cloned"] > 10 and c["weight_deleted"] > 10
    g.close()
    o.close()


def test_weighted_inflow_case_alone_equals_batch(require_gpu):
    specs = [inflow_box(seed=3), graded_box(npps=60, seed=4)]
    fill = {"number_density": {"Ar": 1.0e20}, "T": 300.0}
    alone = []
    for s in specs:
        c = Case(s, precision="fp64", capacity=1 << 18)
        c.fill(**fill)
        c.step(25)
        c.sync()
        alone.append((c.samples(), c.counters(), c.particles()["velocity"]))
        c.close()
    b = Batch(precision="fp64", use_graphs=False)
    slots = [b.admit(s, fill, 1 << 16) for s in specs]
    b.step(25)
    b.sync()
    for (sa, ca, va), sl in zip(alone, slots):
        c = b.case(sl)
        assert np.array_equal(c.samples(), sa)
        assert np.array_equal(c.particles()["velocity"], va)
        cb = c.counters()
        for k in ("n_particles", "cloned", "weight_deleted", "inserted", "collisions"):
            assert cb[k] == ca[k], k
    b.close()


def test_weighted_restart_continues(require_gpu):
    spec = inflow_box()
    a = Case(spec, capacity=1 << 18)
    a.fill({"Ar": 1.0e20}, 300.0)
    a.step(8)
    ck = a.checkpoint()
    assert np.array_equal(ck["state"]["cell_weight"], a.cell_weights())
    b = Case.restart(spec, ck, capacity=1 << 18)
    assert np.array_equal(b.cell_weights(), a.cell_weights())
    a.step(9)
    b.step(9)
    assert_identical(a, OracleLike(b), "restart")
    a.close()
    b.close()


class OracleLike:
    """A GPU case seen through the oracle's accessors (for assert_identical)."""

    def __init__(self, c):
        self.c = c

    def particles(self):
        return self.c.particles()

    def counters(self):
        return self.c.counters()

    def sigma_max(self):
        return self.c.sigma_max()

    def samples(self):
        return self.c.samples()

    def wall_samples(self):
        return self.c.wall_samples()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_weighted_fields_give_the_physical_density(require_gpu, precision):
    spec = graded_box(npps=400)
    g = Case(spec, precision=precision, capacity=1 << 20)
    g.fill({"Ar": 1.0e20}, 300.0)
    w = g.cell_weights()
    assert w.max() / w.min() > 8.0
    g.reset_samples()
    g.step(300)
    f = g.fields()
    rho = f["rhoN"].reshape(2, 2, 8).mean((0, 1))
    # flat along the graded axis against its mean; the mean near the filled 1e20
    assert np.all(np.abs(rho / rho.mean() - 1.0) < 0.04), rho / rho.mean()
    assert abs(rho.mean() / 1.0e20 - 1.0) < 0.2
    assert np.allclose(f["translational_T"].mean(), 300.0, rtol=0.02)
    g.close()
