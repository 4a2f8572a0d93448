"""Pins the oracle's primitives: Philox-4x32-10 against the published
Random123 known-answer vectors, the deterministic elementary functions
against libm (numpy) to a few ulp, and the stream helpers' distributions.
The product's __host__ __device__ copies of the same primitives are compiled
for the host and compared with the oracle bit for bit (test_abi_cpu.py)."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from oracle import orc

HERE = os.path.dirname(os.path.abspath(__file__))


def L():
    lib = orc.lib()
    lib.orc_philox4x32.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    lib.orc_mathfn.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int64]
    lib.orc_stream.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.c_uint32,
                               C.POINTER(C.c_double), C.c_int64]
    return lib


def mathfn(fn, x, y=None):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y if y is not None else x, dtype=np.float64)
    out = np.zeros_like(x)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    L().orc_mathfn(fn, dp(x), dp(y), dp(out), len(x))
    return out


def stream(what, n, m=0, seed=99, kind=1, sub=0, ident=7, step=3):
    out = np.zeros(n)
    L().orc_stream(seed, kind, sub, ident, step, what, m, out.ctypes.data_as(C.POINTER(C.c_double)), n)
    return out


def test_philox_known_answers():
    kat = json.load(open(os.path.join(HERE, "golden", "philox_kat.json")))
    for v in kat["vectors"]:
        ctr = (C.c_uint32 * 4)(*[int(h, 16) for h in v["ctr"]])
        key = (C.c_uint32 * 2)(*[int(h, 16) for h in v["key"]])
        out = (C.c_uint32 * 4)()
        L().orc_philox4x32(ctr, key, out)
        assert [f"{x:08x}" for x in out] == v["out"]


def ulps(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b) / np.spacing(np.maximum(np.abs(a), np.abs(b)))


def test_log_accuracy():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.random(200000), 10 ** rng.uniform(-300, 300, 200000), [1.0, 0.5, 2.0, 2 ** -53]])
    assert np.max(ulps(mathfn(0, x), np.log(x))) <= 2.0
    assert mathfn(0, np.array([1.0]))[0] == 0.0


def test_exp_accuracy():
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.uniform(-700, 700, 200000), rng.uniform(-1, 1, 200000), [0.0]])
    assert np.max(ulps(mathfn(1, x), np.exp(x))) <= 2.0
    assert mathfn(1, np.array([-800.0]))[0] == 0.0


def test_sincos_accuracy():
    rng = np.random.default_rng(3)
    u = np.concatenate([rng.random(200000), [0.25, 0.5, 0.75, 1.0, 2 ** -53]])
    s, c = mathfn(2, u), mathfn(3, u)
    # reference in extended precision: 2*pi*u rounded to double alone is off
    # by up to ~1e-15 near u = 1
    ul = u.astype(np.longdouble)
    two_pi = np.longdouble("6.283185307179586476925286766559")
    ref_s, ref_c = np.sin(two_pi * ul), np.cos(two_pi * ul)
    # absolute error is the meaningful bound near the zeros of sin/cos
    assert np.max(np.abs(s - ref_s)) < 3e-16
    assert np.max(np.abs(c - ref_c)) < 3e-16
    assert mathfn(2, np.array([0.5]))[0] == 0.0 and mathfn(3, np.array([0.25]))[0] == 0.0


def test_pow_accuracy():
    rng = np.random.default_rng(4)
    x = rng.uniform(1e-6, 10, 100000)
    y = rng.uniform(-3, 3, 100000)
    rel = np.abs(mathfn(4, x, y) / np.power(x, y) - 1)
    assert np.max(rel) < 1e-14


def test_uniform_stream_distribution():
    u = stream(0, 1_000_000)
    assert u.min() > 0.0 and u.max() <= 1.0
    assert abs(u.mean() - 0.5) < 5 * np.sqrt(1 / 12 / len(u))
    assert abs(u.var() - 1 / 12) < 5e-4
    # different counters give different streams; same counters repeat exactly
    assert not np.array_equal(stream(0, 100, ident=7), stream(0, 100, ident=8))
    assert np.array_equal(stream(0, 100), stream(0, 100))


def test_index_stream_is_uniform():
    k = stream(1, 600_000, m=6).astype(int)
    cnt = np.bincount(k, minlength=6)
    assert k.min() == 0 and k.max() == 5
    chi2 = np.sum((cnt - 1e5) ** 2 / 1e5)
    assert chi2 < 25.0  # 5 dof, p ~ 1e-4


def test_gauss_stream_moments():
    g = stream(2, 1_000_000)
    assert abs(g.mean()) < 5e-3
    assert abs(g.var() - 1.0) < 5e-3
    assert abs(np.mean(g ** 4) - 3.0) < 3e-2
