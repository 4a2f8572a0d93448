"""NTC subcells on the GPU (noTimeCounter.C:97-161, 193-225): with levels
uploaded per cell the collisions are bit-identical to the oracle (VHS and
Larsen-Borgnakke VSS, a graded mesh, a 2D mesh with an empty direction),
and subcells change the collisions while the candidate count stays that
of the whole cell."""
import numpy as np
import pytest

from oracle.orc import OracleCase
from unigasfoam_amd import abi
from unigasfoam_amd.core import Case
from unigasfoam_amd.spec import Species

from test_gpu_parity import assert_identical
from test_oracle_physics import box_case

pytestmark = pytest.mark.gpu

N2 = Species("N2", mass=4.65e-26, diameter=4.17e-10, omega=0.74, alpha=1.36, rotational_dof=2)


def random_levels(nc, seed, dims=(0, 1, 2)):
    rng = np.random.default_rng(seed)
    lv = np.ones((nc, 3), dtype=np.int32)
    for d in dims:
        lv[:, d] = rng.integers(1, 5, nc)
    lv[::5] = 1                       # some cells keep the whole cell
    return lv


def cases():
    a = box_case(n=(4, 4, 2), fn=1e10, collision="dsmc", dsmc="variableHardSphere", kind=abi.BND_DIFFUSE, seed=3)
    b = box_case(n=(6, 2, 2), fn=2e10, collision="dsmc", dsmc="LarsenBorgnakkeVariableSoftSphere", species=(N2,),
                 kind=abi.BND_DIFFUSE, grading=(6.0, 1.0, 1.0), seed=5)
    c = box_case(n=(5, 4, 1), fn=1e10, collision="dsmc", dsmc="variableHardSphere", kind=abi.BND_DIFFUSE, seed=9)
    for name in ("zmin", "zmax"):
        c.mesh.patches[c.mesh.patch_index(name)].type = abi.EMPTY
    c.boundaries = [bd for bd in c.boundaries if bd.patch not in ("zmin", "zmax")]
    return {"vhs_box": (a, (0, 1, 2)), "lb_vss_graded": (b, (0, 1, 2)), "vhs_2d": (c, (0, 1))}


@pytest.mark.parametrize("name", ["vhs_box", "lb_vss_graded", "vhs_2d"])
def test_subcells_identical_to_oracle(require_gpu, name):
    spec, dims = cases()[name]
    sp = spec.species[0].name
    lv = random_levels(spec.mesh.n_cells, seed=len(name), dims=dims)
    g = Case(spec, precision="fp64", capacity=1 << 17)
    o = OracleCase(spec, capacity=1 << 17)
    for c in (g, o):
        c.fill({sp: 1.0e20}, 300.0)
        c.set_subcell_levels(lv)
    assert np.array_equal(g.subcell_levels(), lv)
    for k in range(3):
        g.step(8)
        o.step(8)
        assert_identical(g, o, f"{name} after {8 * (k + 1)} steps")
    assert g.counters()["collisions"] > 1000
    g.close()
    o.close()


def test_subcells_change_collisions_not_candidates(require_gpu):
    spec, _ = cases()["vhs_box"]
    out = []
    for lv in ((1, 1, 1), (3, 3, 2)):
        g = Case(spec, precision="fp64", capacity=1 << 17)
        g.fill({"Ar": 1.0e20}, 300.0)
        g.set_subcell_levels(lv)
        g.step(1)
        g.sync()
        out.append((g.counters()["candidates"], g.particles()["velocity"]))
        g.close()
    assert out[0][0] == out[1][0] > 0
    assert not np.array_equal(out[0][1], out[1][1])


def test_subcell_levels_are_validated(require_gpu):
    spec, _ = cases()["vhs_2d"]
    g = Case(spec, precision="fp64", capacity=1 << 12)
    for bad in ((1, 1, 2), (11, 1, 1), (0, 1, 1)):
        with pytest.raises(Exception):
            g.set_subcell_levels(bad)
    g.set_subcell_levels((1, 1, 1))
    assert (g.subcell_levels() == 1).all()
    g.close()
