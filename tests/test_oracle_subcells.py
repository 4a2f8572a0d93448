"""Pins of the oracle's NTC subcells (noTimeCounter.C:97-161, 193-225).

With subcells, the second candidate of a pair is drawn from the first
candidate's virtual Cartesian subcell whenever that subcell holds two or
more simulators, and from the whole cell otherwise.  The selection number
does not depend on the subcells.  Observable consequences, checked with the
collision phase alone (no move):

  * a binary collision conserves the pair's momentum, so with every
    subcell holding two or more simulators the momentum of each subcell is
    conserved to rounding, while with the whole cell as one subcell it is
    not;
  * a simulator alone in its subcell still collides, with partners from
    the whole cell;
  * the candidate count of a step is the same with and without subcells;
  * a simulator on the cell's upper face belongs to the last subcell;
  * the subcells of x and y are distinct, i.e. the index weights the axes
    so that no two subcells share a list (per-subcell conservation on a
    2 x 3 split).
"""
import math

import numpy as np
import pytest

from oracle import orc
from unigasfoam_amd import abi
from unigasfoam_amd.mesh import box_mesh
from unigasfoam_amd.spec import CaseSpec, Species

KB = abi.KB
AR = Species("Ar", mass=6.63e-26, diameter=4.17e-10, omega=0.81)
SIDES = ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")
COLLIDE, ADVANCE = 4, 16
SIDE = 1.0e-3
DT = 1.0e-6


def one_cell(seed=7, fn=1.0, empty_z=False):
    pt = {k: abi.CYCLIC for k in SIDES}
    if empty_z:
        pt["zmin"] = pt["zmax"] = abi.EMPTY
    mesh = box_mesh((0.0, 0.0, 0.0), (SIDE, SIDE, SIDE), (1, 1, 1), pt)
    return CaseSpec(mesh=mesh, species=[AR], delta_t=DT, n_equivalent_particles=fn, collision_model="dsmc",
                    dsmc_model="variableHardSphere", Tref=273.0, boundaries=[], seed=seed)


def maxwellian(n, T, rng):
    return rng.normal(0.0, math.sqrt(KB * T / AR.mass), (n, 3))


def make_case(pos, vel, levels, candidates=3000.0, seed=7):
    """One periodic cell holding the given simulators; F_N chosen for about
    `candidates` NTC candidates per step at the start value of
    (sigma_T c_r)max."""
    n = len(pos)
    smax = math.pi * AR.diameter ** 2 * 2000.0
    fn = candidates / (0.5 * n * (n - 1) * smax * DT / SIDE ** 3)
    spec = one_cell(seed=seed, fn=fn)
    o = orc.OracleCase(spec, capacity=n + 16)
    o.upload(velocity=vel, cell=np.zeros(n, dtype=np.int32), position=pos)
    o.set_sigma_max(np.full(1, smax))
    o.set_subcell_levels(levels)
    return o


def subcell_of(pos, levels):
    k = [np.clip(np.floor(levels[d] * pos[:, d] / SIDE).astype(int), 0, levels[d] - 1) for d in range(3)]
    return k[0] + levels[0] * (k[1] + levels[1] * k[2])


def momenta(p, levels, nsub):
    s = subcell_of(p["position"], levels)
    mv = AR.mass * p["velocity"]
    return np.array([mv[s == q].sum(axis=0) for q in range(nsub)]), s


def collide_steps(o, k):
    c0 = o.counters()
    for _ in range(k):
        o.step_phases(COLLIDE | ADVANCE)
    c1 = o.counters()
    return c1["candidates"] - c0["candidates"], c1["collisions"] - c0["collisions"]


@pytest.mark.parametrize("levels", [(2, 2, 1), (2, 3, 1), (1, 1, 4)])
def test_subcell_momentum_is_conserved(levels):
    rng = np.random.default_rng(11)
    n = 480
    pos = rng.uniform(0.0, SIDE, (n, 3))
    vel = maxwellian(n, 400.0, rng)
    nsub = int(np.prod(levels))
    scale = AR.mass * np.abs(vel).sum() / nsub
    o = make_case(pos, vel, levels)
    m0, s = momenta(o.particles(), levels, nsub)
    assert np.bincount(s, minlength=nsub).min() >= 2
    cand, coll = collide_steps(o, 3)
    m1, _ = momenta(o.particles(), levels, nsub)
    o.close()
    assert coll > 1000
    assert np.abs(m1 - m0).max() < 1e-12 * scale
    # the same simulators with the whole cell as one subcell exchange momentum between the parts
    o = make_case(pos, vel, (1, 1, 1))
    collide_steps(o, 3)
    m2, _ = momenta(o.particles(), levels, nsub)
    o.close()
    assert np.abs(m2 - m0).max() > 1e-3 * scale


def test_lone_simulator_takes_partners_from_the_whole_cell():
    rng = np.random.default_rng(3)
    n = 200
    pos = rng.uniform(0.0, SIDE, (n, 3))
    pos[:, 0] *= 0.45                      # all in the lower x half ...
    pos[0, 0] = 0.8 * SIDE                 # ... but one
    vel = maxwellian(n, 400.0, rng)
    o = make_case(pos, vel, (2, 1, 1), candidates=4000.0)
    p0 = o.particles()
    collide_steps(o, 4)
    p1 = o.particles()
    o.close()
    lone0 = np.flatnonzero(p0["position"][:, 0] > 0.5 * SIDE)
    lone1 = np.flatnonzero(p1["position"][:, 0] > 0.5 * SIDE)
    assert len(lone0) == len(lone1) == 1
    assert not np.allclose(p1["velocity"][lone1[0]], p0["velocity"][lone0[0]])
    # total momentum is conserved; the lower half's alone is not
    tot0, tot1 = p0["velocity"].sum(axis=0), p1["velocity"].sum(axis=0)
    assert np.abs(tot1 - tot0).max() < 1e-9 * np.abs(p0["velocity"]).sum()
    low0 = p0["velocity"][p0["position"][:, 0] < 0.5 * SIDE].sum(axis=0)
    low1 = p1["velocity"][p1["position"][:, 0] < 0.5 * SIDE].sum(axis=0)
    assert np.abs(low1 - low0).max() > 1.0


def test_candidate_count_does_not_depend_on_subcells():
    rng = np.random.default_rng(5)
    n = 300
    pos = rng.uniform(0.0, SIDE, (n, 3))
    vel = maxwellian(n, 300.0, rng)
    got = []
    for lv in ((1, 1, 1), (3, 2, 2)):
        o = make_case(pos, vel, lv)
        got.append(collide_steps(o, 1)[0])
        o.close()
    assert got[0] == got[1] > 0


def test_upper_face_belongs_to_the_last_subcell():
    """Two simulators exactly on the cell's upper x face form the upper
    subcell's pair: they collide only with each other."""
    rng = np.random.default_rng(9)
    n = 120
    pos = rng.uniform(0.0, SIDE, (n, 3))
    pos[:, 0] *= 0.4
    pos[:2, 0] = SIDE
    vel = maxwellian(n, 400.0, rng)
    o = make_case(pos, vel, (2, 1, 1), candidates=6000.0)
    p0 = o.particles()
    collide_steps(o, 3)
    p1 = o.particles()
    o.close()
    up0 = p0["velocity"][p0["position"][:, 0] > 0.5 * SIDE]
    up1 = p1["velocity"][p1["position"][:, 0] > 0.5 * SIDE]
    assert len(up0) == len(up1) == 2
    assert not np.allclose(np.sort(up1, axis=0), np.sort(up0, axis=0))
    assert np.abs(up1.sum(axis=0) - up0.sum(axis=0)).max() < 1e-9 * np.abs(up0).sum()


def test_levels_are_validated():
    o = orc.OracleCase(one_cell(empty_z=True), capacity=16)
    with pytest.raises(orc.OracleError):
        o.set_subcell_levels((1, 1, 2))        # an empty direction
    with pytest.raises(orc.OracleError):
        o.set_subcell_levels((11, 1, 1))
    with pytest.raises(orc.OracleError):
        o.set_subcell_levels((0, 1, 1))
    o.set_subcell_levels((3, 10, 1))
    assert o.subcell_levels().tolist() == [[3, 10, 1]]
    o.close()
