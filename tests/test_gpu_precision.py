"""FP32 particle state against an all-FP64 GPU run of the same case: the
sampled fields agree within their statistical error, and domain means
within 0.1 % wherever the sample resolves that."""
import pytest

from unigasfoam_amd import workloads
from unigasfoam_amd.precision import compare

pytestmark = pytest.mark.gpu


def test_fp32_state_fields_match_fp64(require_gpu):
    spec, fill = workloads.couette(n_cells=100, per_cell=200, U_wall=250.0)
    r = compare(spec, fill, warmup=200, sample=2000, blocks=10)
    for name, v in r.items():
        # within 3 combined standard errors, and within 0.1 % when the sample resolves 0.1 %
        assert v["mean_rel_diff"] < max(1e-3, 3.0 * v["mean_rel_sigma"]), (name, v)
        # batch-means errors from 10 blocks: about 97 % of cells inside 3 sigma
        assert v["cells_within_3sigma"] > 0.85, (name, v)
    assert r["rhoN"]["mean_rel_diff"] < 1e-12        # closed box: the simulator count is exact
