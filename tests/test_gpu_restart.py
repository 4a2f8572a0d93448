"""Checkpoint and restart (uniGasCloud.C:577-703 start-up from a written
cloud, uniGasVolFields.C:550-669 resumable sampler sums): a case restarted
from a checkpoint continues bit for bit like the case that was not stopped.

The property is independent of size and of the physics: simulators (in
storage form and slot order), sampler sums, (sigma_T c_r)max, model ids,
Kn_GLL and every counter must be identical after the same steps.  The
cases cover the hybrid switch (decompositions after the restart), inflow,
deletion and diffuse walls, Larsen-Borgnakke DSMC, the BGK family, FP32
state, and the round trip through the .npz file.
"""
import numpy as np
import pytest

from unigasfoam_amd import workloads
from unigasfoam_amd.core import Case, checkpoint_load, checkpoint_save

pytestmark = pytest.mark.gpu


def results(c):
    c.sync()
    p = c.particles()
    out = {"p_" + k: p[k] for k in ("local", "velocity", "erot", "cell", "species")}
    out.update(samples=c.samples(), sigma_max=c.sigma_max(), cell_model=c.cell_model(), kn=c.kn())
    cnt = c.counters()
    for k in ("n_particles", "steps", "candidates", "collisions", "bgk_collisions", "inserted", "deleted",
              "wall_hits", "track_errors", "n_sampled_steps", "particle_steps", "n_dsmc_cells"):
        out["c_" + k] = cnt[k]
    st = c.state()
    out.update(kn_sums=st["kn_sums"], usp_heat_flux=st["usp_heat_flux"], usp_stress=st["usp_stress"],
               usp_max_prob=st["usp_max_prob"], wall_samples=c.wall_samples())
    return out


def assert_same(a, b, what):
    assert a.keys() == b.keys()
    for k in a:
        if not np.array_equal(np.asarray(a[k]), np.asarray(b[k])):
            raise AssertionError(f"{what}: {k} differs after the restart")


def run_restart(spec, fill, before, after, precision="fp64", capacity=1 << 17, path=None):
    a = Case(spec, precision=precision, capacity=capacity)
    a.fill(**fill)
    a.step(before)
    a.sync()
    ck = a.checkpoint()
    if path is not None:
        checkpoint_save(path, ck)
        ck = checkpoint_load(path)
    a.step(after)
    ra = results(a)
    a.close()
    b = Case.restart(spec, ck, capacity=capacity)
    b.step(after)
    rb = results(b)
    b.close()
    return ra, rb


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_restart_hybrid_nozzle(require_gpu, precision, tmp_path):
    """Hybrid USP/DSMC with inflow, deletion and diffuse walls; the restart
    falls between decompositions (every 3 steps) and two more follow."""
    spec, fill = workloads.nozzle(nx=10, ny=6, nz=6, per_cell=20, n_stag=2.4e24)
    spec.decomposition_interval = 3
    spec.smoothing_passes = 2
    ra, rb = run_restart(spec, fill, before=4, after=7, precision=precision, path=tmp_path / "nozzle.npz")
    assert_same(ra, rb, f"nozzle {precision}")
    assert 0 < ra["cell_model"].sum() < ra["cell_model"].size, "expected both collision models"
    assert ra["c_inserted"] > 0 and ra["c_deleted"] > 0


def test_restart_couette_lb(require_gpu):
    spec, fill = workloads.couette(n_cells=40, per_cell=30)
    ra, rb = run_restart(spec, fill, before=10, after=15)
    assert_same(ra, rb, "couette LB")
    assert ra["c_collisions"] > 0 and ra["c_wall_hits"] > 0 and ra["wall_samples"].any()


@pytest.mark.parametrize("bgk", ["stochasticParticleESBGK", "unifiedStochasticParticleSBGK"])
def test_restart_bgk(require_gpu, bgk):
    spec, fill = workloads.couette(n_cells=30, per_cell=40, collision_model="bgk", bgk_model=bgk)
    ra, rb = run_restart(spec, fill, before=6, after=8)
    assert_same(ra, rb, bgk)
    assert ra["c_bgk_collisions"] > 0


def test_restart_after_sampler_reset(require_gpu):
    """The sampled-step count is kept relative to the last sampler reset."""
    spec, fill = workloads.couette(n_cells=20, per_cell=30)
    a = Case(spec, capacity=1 << 16)
    a.fill(**fill)
    a.step(5)
    a.reset_samples()
    a.step(3)
    ck = a.checkpoint()
    b = Case.restart(spec, ck, capacity=1 << 16)
    assert b.counters()["n_sampled_steps"] == a.counters()["n_sampled_steps"] == 3
    a.step(4)
    b.step(4)
    assert_same(results(a), results(b), "sampler reset")


def test_start_from_the_reference_cloud_layout(require_gpu, tmp_path):
    """A case written in uniGasFoam's restart layout (lagrangian/uniGas,
    the restart cell fields, uniform/time) and started again from it holds
    the same simulators (positions to rounding of the cell frame), the same
    (sigma_T c_r)max and model ids, and steps on."""
    from unigasfoam_amd import foam_io
    spec, fill = workloads.nozzle(nx=10, ny=6, nz=6, per_cell=20, n_stag=2.4e24)
    spec.decomposition_interval = 3
    spec.smoothing_passes = 2
    a = Case(spec, precision="fp64", capacity=1 << 17)
    a.fill(**fill)
    a.step(7)
    foam_io.write_cloud(str(tmp_path), "7e-08", a)
    b = foam_io.case_from_cloud(spec, str(tmp_path), "7e-08", precision="fp64", capacity=1 << 17)
    pa, pb = a.particles(), b.particles()
    order_a = np.lexsort((pa["velocity"][:, 0], pa["cell"]))
    order_b = np.lexsort((pb["velocity"][:, 0], pb["cell"]))
    for k in ("cell", "species", "velocity", "erot"):
        assert np.array_equal(pa[k][order_a], pb[k][order_b]), k
    assert np.allclose(pa["position"][order_a], pb["position"][order_b], rtol=0.0, atol=1e-15)
    assert np.array_equal(a.sigma_max(), b.sigma_max())
    assert np.array_equal(a.cell_model(), b.cell_model())
    assert 0 < b.cell_model().sum() < b.cell_model().size
    b.step(5)
    b.sync()
    assert np.isfinite(b.particles()["velocity"]).all() and b.counters()["collisions"] > 0
    a.close()
    b.close()
