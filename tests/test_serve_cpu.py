"""Host logic of the served form: the steps at which the loop looks at a
request (end of warmup, every error check, end of sampling) and the error
criterion on write-time fields."""
import numpy as np

from unigasfoam_amd import workloads
from unigasfoam_amd.serve import make_request, next_milestone


def test_fixed_count_request_milestones():
    spec, fill = workloads.couette(n_cells=10, per_cell=20)
    r = make_request(0, spec, fill, 20, warmup=7, samples_per_cell=200)
    assert r.error_target == 0.0 and r.sample_steps == 10
    assert next_milestone(r, 0) == 7
    assert next_milestone(r, 7) == 17 and next_milestone(r, 12) == 17


def test_error_target_request_milestones_and_criterion():
    spec, fill = workloads.couette(n_cells=10, per_cell=20)
    r = make_request(0, spec, fill, 20, warmup=5, error_target=0.05, check_every=4, max_sample_steps=30)
    assert r.sample_steps == 30 and r.total_steps == 35
    assert next_milestone(r, 0) == 5
    assert [next_milestone(r, d) for d in (5, 6, 9, 33)] == [9, 9, 13, 35]
    f = {"rhoN": np.array([1.0, 1.0, 0.0]), "density_error": np.array([0.04, 0.05, 0.9])}
    assert r.converged(f)                              # the empty cell does not count
    f["density_error"][1] = 0.051
    assert not r.converged(f)
    assert not r.converged({"rhoN": np.zeros(3), "density_error": np.zeros(3)})
