"""FP32 particle state against an all-FP64 GPU run of the same case.

Ensembles of independent replicas (compare_replicas: replica r of both
precisions starts from the same meshFill, the differences are paired, and
the standard errors come from the scatter over replicas) resolve the domain
means of 1-D Couette flows to about 0.03 %: temperature, pressure and the
rotational and overall temperatures must agree within 0.1 %, the velocity
components, the shear stress P_xy and the heat flux q_y within 0.1 % of
their scales or four standard errors, and the cell fields within three
standard errors in at least 95 % of the cells.  The DSMC case (N2,
Larsen-Borgnakke) exercises the FP32 move, walls and collisions; the USP
case the single-precision BGK-family draws.  The plate case adds two
species, freestream inflow, deletion and symmetry (C3's physics)."""
import os

import pytest

from unigasfoam_amd import case_dir, workloads
from unigasfoam_amd.precision import compare_replicas

pytestmark = pytest.mark.gpu
SCALARS = ("translational_T", "rotational_T", "overall_T", "p")


def check(r, what, scalar_bar=1e-3):
    for name, v in r.items():
        if name == "rhoN":
            assert v["mean_rel_diff"] < 1e-12, (what, name, v)        # closed box: exact count
            continue
        if name in SCALARS:
            assert v["mean_rel_diff"] < scalar_bar, (what, name, v)
        else:
            assert v["mean_rel_diff"] < max(1e-3, 4.0 * v["mean_rel_sigma"]), (what, name, v)
        assert v["z"] < 4.0, (what, name, v)
        assert v["cells_within_3sigma"] > 0.95, (what, name, v)


@pytest.mark.parametrize("model", ["dsmc", "usp"])
def test_fp32_state_within_a_tenth_of_a_percent(require_gpu, model):
    kw = {} if model == "dsmc" else dict(collision_model="bgk", bgk_model="unifiedStochasticParticleSBGK")
    spec, fill = workloads.couette(n_cells=100, per_cell=1000, U_wall=250.0, **kw)
    spec.sampler_full = True
    r = compare_replicas(spec, fill, warmup=1000, sample=2000, replicas=48)
    # the sample resolves 0.1 %: the standard error of T is below 0.05 %
    assert r["translational_T"]["mean_rel_sigma"] < 5e-4, r["translational_T"]
    check(r, model)


def test_plate_fp32_state_agrees_with_fp64(require_gpu):
    """Two species, inflow, deletion, diffuse wall, LB: the transient plate
    flow in 16 paired replicas."""
    spec, fills, run = case_dir.read_case(os.path.join(os.path.dirname(__file__), "cases", "platePlume"))
    assert len(fills) == 1
    r = compare_replicas(spec, fills[0], warmup=100, sample=200, replicas=16)
    for name, v in r.items():
        assert v["mean_rel_diff"] < max(1e-3, 4.0 * v["mean_rel_sigma"]), (name, v)
        assert v["cells_within_3sigma"] > 0.95, (name, v)
