"""The 1 bar nozzle expansion's starting state (workloads.nozzle_expansion).

What fixes it:
  * the area-Mach relation of isentropic quasi-one-dimensional flow, here
    solved independently on the general form (any gamma, scipy's brentq)
    at gamma = 5/3, and the known values A/A* = 1 at M = 1 and
    ((3 + M^2)/4)^2 / M at M = 2;
  * continuity: the mass flux n U A is the same in every column, and the
    reservoir state (p0 / k T0 at T0) is the stagnation state of every
    column (T (1 + M^2/3) = T0, n (T0/T)^(3/2) = n0);
  * every simulator lies inside its cell (the cell's face planes, from the
    oracle's geometry), and per_cell simulators of weight F_N W_c in a cell
    of the geometry's volume V_c represent the column's density.
"""
import math

import numpy as np
import pytest
from scipy.optimize import brentq

from oracle import orc
from unigasfoam_amd import abi, workloads

KB = abi.KB


def area_ratio(M, g):
    return (1.0 / M) * ((2.0 / (g + 1.0)) * (1.0 + 0.5 * (g - 1.0) * M * M)) ** ((g + 1.0) / (2.0 * (g - 1.0)))


def test_area_mach_relation():
    g = 5.0 / 3.0
    assert workloads.isentropic_mach(1.0, True) == pytest.approx(1.0, abs=1e-7)
    assert area_ratio(2.0, g) == pytest.approx(49.0 / 32.0, rel=1e-14)
    for r in (1.2, 4.0, 16.0, 36.0):
        sub = brentq(lambda M: area_ratio(M, g) - r, 1e-6, 1.0, xtol=1e-15, rtol=1e-15)
        sup = brentq(lambda M: area_ratio(M, g) - r, 1.0, 100.0, xtol=1e-15, rtol=1e-15)
        assert workloads.isentropic_mach(r, False) == pytest.approx(sub, rel=1e-10)
        assert workloads.isentropic_mach(r, True) == pytest.approx(sup, rel=1e-10)


def test_columns_are_one_isentropic_expansion():
    spec, st = workloads.nozzle_expansion(nx=24, ny=6, nz=6, per_cell=4)
    c = st["columns"]
    n0, T0 = 1.0e5 / (KB * 300.0), 300.0
    assert np.allclose(c["T"] * (1.0 + c["M"] ** 2 / 3.0), T0, rtol=1e-12)
    assert np.allclose(c["n"] * (T0 / c["T"]) ** 1.5, n0, rtol=1e-12)
    xs = 0.02 * np.arange(25) / 24
    hf = workloads.nozzle_half_width(xs)
    hc = 0.5 * (hf[:-1] + hf[1:])
    flux = c["n"] * c["U"] * hc ** 2
    assert np.allclose(flux, flux[0], rtol=1e-9)
    # subsonic before the throat, supersonic after, a hundredfold fall in density
    assert (c["M"][c["x"] < 0.005] < 1.0).all() and (c["M"][c["x"] > 0.005] > 1.0).all()
    assert c["n"][0] / c["n"][-1] > 50.0


def test_simulators_inside_their_cells_at_the_column_density():
    nx, ny, nz, pc = 12, 5, 4, 30
    spec, st = workloads.nozzle_expansion(nx=nx, ny=ny, nz=nz, per_cell=pc)
    o = orc.OracleCase(spec, capacity=16)
    g = o.geometry()
    cell, pos = st["cell"], st["position"]
    assert np.bincount(cell, minlength=spec.mesh.n_cells).tolist() == [pc] * spec.mesh.n_cells
    q = g["planes"][cell]                                   # [n, 6, 4]
    rel = pos - g["cc"][cell]
    side = np.einsum("nfk,nk->nf", q[..., :3], rel) - q[..., 3]
    used = g["fkind"][cell] != 0
    assert (side[used] <= 1e-15).all()
    col = np.arange(spec.mesh.n_cells) % nx
    n_rep = pc * spec.n_equivalent_particles * st["cell_weights"] / g["vol"]
    assert np.allclose(n_rep, st["columns"]["n"][col], rtol=1e-12)
    assert st["cell_weights"][0] == pytest.approx(1.0, rel=1e-14)
    o.close()


def test_expansion_reads_into_the_oracle_and_steps():
    spec, st = workloads.nozzle_expansion(nx=12, ny=4, nz=4, per_cell=10)
    spec.decomposition_interval = 3
    o = orc.OracleCase(spec, capacity=1 << 16)
    workloads.apply_state(o, st)
    n0 = len(o.particles()["cell"])
    o.step(6)
    n1 = len(o.particles()["cell"])
    assert 0.5 * n0 < n1 < 2.0 * n0
    assert o.counters()["bgk_collisions"] > 0
    o.close()
