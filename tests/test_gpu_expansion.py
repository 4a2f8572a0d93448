"""The 1 bar nozzle expansion on the GPU: a case started from the isentropic
state (cell weights, (sigma_T c_r)max and simulators uploaded) runs
bit-identically to the oracle started from the same state, through the
weighting's clones and removals, the reservoir and background inflows and
the hybrid decomposition; the FP32 state keeps the per-cell load."""
import numpy as np
import pytest

from oracle.orc import OracleCase
from unigasfoam_amd import workloads
from unigasfoam_amd.core import Case

from test_gpu_parity import assert_identical

pytestmark = pytest.mark.gpu


def small_expansion(per_cell=20):
    spec, st = workloads.nozzle_expansion(nx=16, ny=6, nz=6, per_cell=per_cell)
    spec.decomposition_interval = 4
    return spec, st


def test_expansion_identical_to_oracle(require_gpu):
    spec, st = small_expansion()
    g = Case(spec, precision="fp64", capacity=1 << 17)
    o = OracleCase(spec, capacity=1 << 17)
    for c in (g, o):
        workloads.apply_state(c, st)
    for k in range(4):
        g.step(3)
        o.step(3)
        assert_identical(g, o, f"expansion after {3 * (k + 1)} steps")
    cg = g.counters()
    assert cg["bgk_collisions"] > 1000
    assert np.array_equal(g.cell_model(), o.cell_model())
    g.close()
    o.close()


def test_expansion_fp32_keeps_the_load(require_gpu):
    spec, st = small_expansion(per_cell=50)
    g = Case(spec, precision="fp32", capacity=1 << 18)
    workloads.apply_state(g, st)
    n0 = g.count()
    g.step(40)
    g.sync()
    n1 = g.count()
    assert 0.7 * n0 < n1 < 1.4 * n0, (n0, n1)
    assert np.isfinite(g.particles()["velocity"]).all()
    g.close()


def test_conservation_occupancy_variants_agree(require_gpu):
    """The autotuner runs the FP32 conservation kernel's second occupancy on
    step 255 (and keeps whichever is faster after it); a run continued from
    a checkpoint at step 250 retunes from scratch and keeps the first one
    through step 260, so both instances must give the same simulators."""
    spec, st = small_expansion(per_cell=30)
    a = Case(spec, precision="fp32", capacity=1 << 18)
    workloads.apply_state(a, st)
    a.step(250)
    a.sync()
    ck = a.checkpoint()
    a.step(10)
    a.sync()
    b = Case.restart(spec, ck, capacity=1 << 18)
    b.step(10)
    b.sync()
    pa, pb = a.particles(), b.particles()
    assert pa["cell"].size == pb["cell"].size
    for k in ("cell", "local", "velocity"):
        assert np.array_equal(pa[k], pb[k]), k
    assert a.counters()["bgk_collisions"] > 0
    a.close()
    b.close()


def test_expansion_many_inflow_faces_identical(require_gpu):
    """More than 4096 inflow face-species: the counts are drawn per face in
    parallel and scanned in place, the same as the oracle's."""
    spec, st = workloads.nozzle_expansion(nx=8, ny=48, nz=48, per_cell=2)
    spec.decomposition_interval = 4
    g = Case(spec, precision="fp64", capacity=1 << 19)
    o = OracleCase(spec, capacity=1 << 19)
    for c in (g, o):
        workloads.apply_state(c, st)
    for k in range(3):
        g.step(2)
        o.step(2)
        assert_identical(g, o, f"many inflow faces, after {2 * (k + 1)} steps")
    assert g.counters()["inserted"] > 0
    g.close()
    o.close()
