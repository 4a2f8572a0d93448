"""Capacity growth between steps: a box filled from empty through a
freestream patch (its other walls specular) grows far past its first
particle buffers; the GPU run
equals the oracle bit for bit (FP64), with and without captured graphs, and
no step's inflow is dropped.  A case restarted with automatic capacity keeps
going bit for bit (the restart reserves only what it uploads)."""
import numpy as np
import pytest

from oracle.orc import OracleCase
from unigasfoam_amd import abi, workloads
from unigasfoam_amd.core import Case
from unigasfoam_amd.mesh import box_mesh
from unigasfoam_amd.spec import Boundary, CaseSpec

from test_gpu_parity import assert_identical

pytestmark = pytest.mark.gpu
SIDES = ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")


def inflow_box(seed=11):
    pt = {k: abi.WALL for k in SIDES}
    pt["xmin"] = abi.PATCH
    mesh = box_mesh((0.0, 0.0, 0.0), (4e-3, 2e-3, 2e-3), (8, 4, 4), pt)
    bnd = [Boundary(p, abi.BND_SPECULAR) for p in SIDES if p != "xmin"]
    bnd.append(Boundary("xmin", abi.BND_FREESTREAM, temperature=300.0, velocity=(400.0, 0.0, 0.0),
                        rot_temperature=300.0, number_density={"Ar": 1e20}))
    return CaseSpec(mesh=mesh, species=[workloads.ARGON], delta_t=2e-7, n_equivalent_particles=2e7,
                    collision_model="dsmc", dsmc_model="variableHardSphere", boundaries=bnd, seed=seed)


@pytest.mark.parametrize("graphs", [False, True])
def test_empty_box_fills_past_four_times_its_capacity(require_gpu, graphs):
    spec = inflow_box()
    g = Case(spec, precision="fp64", capacity=0, use_graphs=graphs)
    o = OracleCase(spec, capacity=1 << 20)
    first_cap = 4096          # the buffers of a case created with automatic capacity
    done = 0
    for chunk in (10, 30, 60, 100):
        g.step(chunk - done)
        o.step(chunk - done)
        done = chunk
        assert_identical(g, o, f"inflow box after {done} steps")
    n = g.count()
    assert n > 4 * first_cap, n
    c = g.counters()
    # every step's inflow went in: the inserted count is the oracle's (which
    # never runs out of room), and the open inflow patch also takes some out
    assert c["inserted"] == o.counters()["inserted"] > n
    g.close()


def test_restart_with_automatic_capacity_continues_bit_for_bit(require_gpu):
    spec, fill = workloads.nozzle(nx=10, ny=6, nz=6, per_cell=20)
    a = Case(spec, precision="fp64")
    a.fill(**fill)
    a.step(20)
    ck = a.checkpoint()
    b = Case.restart(spec, ck)            # capacity 0: sized by the upload alone
    a.step(60)
    b.step(60)
    pa, pb = a.particles(), b.particles()
    for k in ("cell", "species", "local", "velocity", "erot"):
        assert np.array_equal(pa[k], pb[k]), k
    assert a.counters()["inserted"] == b.counters()["inserted"] > 0
