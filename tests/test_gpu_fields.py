"""Write-time fields (uniGasVolFields.C:760-1252 as restated by
ugx_fields_compute) against the closed forms of a known state.

A periodic box holds a drifting Maxwellian of N2 (n, T, T_rot, U) and no
collisions, so the simulator set keeps its velocity distribution and every
cell samples it.  The expected values are those of the state, not the field
formulas: rhoN = n exactly (the samples count every simulator each step),
U and T of the simulator set, T_rot, p = n k T, the pressure tensor n k T
times the identity, zero heat flux, Ma = |U| / sqrt(7/5 k T / m), and the
VHS mean free path and collision rate of Bird (1994) eqs. 4.65 and 4.64.
"""
import math

import numpy as np
import pytest

from unigasfoam_amd import abi
from unigasfoam_amd.core import Case
from unigasfoam_amd.mesh import box_mesh
from unigasfoam_amd.spec import CaseSpec
from unigasfoam_amd.workloads import N2

pytestmark = pytest.mark.gpu

KB = abi.KB
SIDES = ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")


def test_fields_of_a_drifting_maxwellian(require_gpu):
    n_cells, side, n, T, T_rot = 8, 0.01, 1.0e20, 400.0, 250.0
    U = np.array([150.0, -50.0, 30.0])
    fn = 1.0e8
    mesh = box_mesh((0.0, 0.0, 0.0), (side, side, side), (n_cells,) * 3, {k: abi.CYCLIC for k in SIDES})
    spec = CaseSpec(mesh=mesh, species=[N2], delta_t=1.0e-7, n_equivalent_particles=fn,
                    collision_model="none", Tref=273.0, boundaries=[], seed=21, sampler_full=True)
    c = Case(spec, precision="fp64", capacity=1 << 21)
    c.fill(number_density={"N2": n}, T=T, rot_T=T_rot, velocity=tuple(U))
    c.reset_samples()
    c.step(30)
    c.sync()
    p = c.particles()
    N = p["cell"].size
    assert N > 500000
    v = p["velocity"]
    vbar = v.mean(0)
    T_set = N2.mass * ((v - vbar) ** 2).sum() / (3.0 * KB * N)
    Trot_set = p["erot"].sum() / (KB * N)
    f = c.fields()
    vol = side ** 3 / n_cells ** 3
    n_set = N * fn / (vol * n_cells ** 3)

    # the simulator set is the fill's: its moments match the requested state
    assert abs(n_set / n - 1.0) < 5.0 / math.sqrt(N)
    sig = math.sqrt(KB * T / N2.mass)
    assert np.all(np.abs(vbar - U) < 5.0 * sig / math.sqrt(N))
    assert abs(T_set / T - 1.0) < 5.0 * math.sqrt(2.0 / (3.0 * N))
    assert abs(Trot_set / T_rot - 1.0) < 5.0 / math.sqrt(N)

    # fields of that set
    assert abs(f["rhoN"].mean() / n_set - 1.0) < 1e-12
    assert abs(f["rhoM"].mean() / (n_set * N2.mass) - 1.0) < 1e-12
    w = f["rhoN"] / f["rhoN"].sum()
    Uf = (w[:, None] * f["U_mean"]).sum(0)
    assert np.all(np.abs(Uf - vbar) < 1e-9 * sig)
    Tf = f["translational_T"].mean()
    assert abs(Tf / T_set - 1.0) < 5e-3
    assert abs(f["rotational_T"].mean() / Trot_set - 1.0) < 5e-3
    assert abs(f["overall_T"].mean() / ((3.0 * T_set + 2.0 * Trot_set) / 5.0) - 1.0) < 5e-3
    nkT = n_set * KB * T_set
    assert abs(f["p"].mean() / nkT - 1.0) < 5e-3
    P = f["shear_stress"].mean(0)                      # xx, xy, xz, yy, yz, zz
    for k in (0, 3, 5):
        assert abs(P[k] / nkT - 1.0) < 1e-2
    for k in (1, 2, 4):
        assert abs(P[k] / nkT) < 1e-2
    q = f["heat_flux"].mean(0)
    assert np.all(np.abs(q) < 1e-2 * nkT * sig)
    Ma = np.linalg.norm(vbar) / math.sqrt(1.4 * KB * T_set / N2.mass)
    assert abs(f["Ma"].mean() / Ma - 1.0) < 1e-2
    d, om, Tref = N2.diameter, N2.omega, 273.0
    mfp = 1.0 / (math.sqrt(2.0) * math.pi * d * d * n_set * (Tref / T_set) ** (om - 0.5))
    assert abs(f["mfp"].mean() / mfp - 1.0) < 1e-2
    nu = 4.0 * d * d * n_set * math.sqrt(math.pi * KB * Tref / N2.mass) * (T_set / Tref) ** (1.0 - om)
    assert abs(f["mcr"].mean() / nu - 1.0) < 1e-2
    c.close()


def test_fields_through_openfoam_files(require_gpu, tmp_path):
    """A case on a mesh read back from constant/polyMesh, and its sampled
    fields written under the uniGasVolFields names and read back."""
    from unigasfoam_amd import foam_io
    mesh = box_mesh((0.0, 0.0, 0.0), (0.004, 0.004, 0.004), (4, 4, 4), {k: abi.CYCLIC for k in SIDES})
    foam_io.write_polymesh(mesh, tmp_path)
    spec = CaseSpec(mesh=foam_io.read_polymesh(tmp_path), species=[N2], delta_t=1.0e-7, n_equivalent_particles=1.0e9,
                    collision_model="dsmc", dsmc_model="LarsenBorgnakkeVariableHardSphere", Tref=273.0, boundaries=[],
                    seed=5)
    c = Case(spec, precision="fp64", capacity=1 << 17)
    c.fill(number_density={"N2": 1.0e20}, T=300.0, rot_T=300.0, velocity=(50.0, 0.0, 0.0))
    c.step(10)
    f = c.fields()
    written = foam_io.write_fields(tmp_path, "1e-06", spec.mesh, f)
    assert "rhoN_uniGasFields" in written and "UMean_uniGasFields" in written
    for key, (stem, _, _) in foam_io.FIELD_FILES.items():
        back = foam_io.read_internal_field(tmp_path / "1e-06" / f"{stem}_uniGasFields")
        assert np.array_equal(back, np.asarray(f[key]).reshape(back.shape), equal_nan=True), key
    c.close()
