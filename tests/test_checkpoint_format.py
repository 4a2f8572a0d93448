"""Checkpoint file and restart-state layout, without a GPU: the .npz round
trip keeps every array and dtype, and the ctypes mirror of ugx_case_state
has the field offsets a C compiler gives the header's struct."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from unigasfoam_amd import abi
from unigasfoam_amd.core import CHECKPOINT_PARTICLE_KEYS, checkpoint_load, checkpoint_save

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def synthetic_checkpoint(n=37, nc=11, seed=3):
    rng = np.random.default_rng(seed)
    particles = {"local": rng.normal(size=(n, 3)), "velocity": rng.normal(size=(n, 3)), "erot": rng.random(n),
                 "cell": rng.integers(0, nc, n).astype(np.int32), "species": rng.integers(0, 2, n).astype(np.int32)}
    state = {}
    for name, rows, dt in abi.STATE_ARRAYS:
        a = rng.normal(size=(nc, rows)) if dt == "f8" else rng.integers(0, 2, (nc, rows))
        state[name] = a.astype(dt)
    state.update(steps=np.int64(123), kn_time=np.float64(4.5e-7), sampled_at_reset=np.int64(17),
                 counters=rng.integers(0, 2 ** 40, abi.N_COUNTERS).astype(np.uint64))
    return {"meta": {"precision": "fp32", "seed": 2 ** 40 + 5, "n_cells": nc}, "particles": particles,
            "state": state}


def test_checkpoint_file_round_trip(tmp_path):
    ck = synthetic_checkpoint()
    path = tmp_path / "ck.npz"
    checkpoint_save(path, ck)
    back = checkpoint_load(path)
    assert back["meta"] == ck["meta"]
    for k in CHECKPOINT_PARTICLE_KEYS:
        a, b = ck["particles"][k], back["particles"][k]
        assert a.dtype == b.dtype and np.array_equal(a, b), k
    assert back["state"].keys() == ck["state"].keys()
    for k, v in ck["state"].items():
        b = back["state"][k]
        assert np.asarray(v).dtype == b.dtype and np.array_equal(np.asarray(v), b), k


def test_case_state_layout_matches_header(tmp_path):
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    fields = [f for f, _ in abi.CaseState._fields_]
    src = tmp_path / "layout.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "ugx.h"\nint main(void) {\n'
                   + "".join(f'  printf("%zu\\n", offsetof(ugx_case_state, {f}));\n' for f in fields)
                   + '  printf("%zu\\n", sizeof(ugx_case_state));\n  return 0;\n}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    want = [getattr(abi.CaseState, f).offset for f in fields] + [ctypes.sizeof(abi.CaseState)]
    assert got == want
