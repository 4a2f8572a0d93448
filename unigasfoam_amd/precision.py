"""Error of an FP32-state run against the same case run all-FP64 on the GPU.

Both runs start from the same meshFill (the FP32 run stores it rounded) and
advance the same number of steps; their trajectories then differ at the
rounding level and grow apart, so the comparison is statistical.  Each run
samples in `blocks` consecutive blocks; the block-to-block scatter of a
field gives its standard error (batch means, which accounts for the
correlation of consecutive steps).  For each sampled field (uniGasVolFields
names) this reports

  mean_rel_diff        |<f32> - <f64>| / |<f64>|, <.> the mean over cells
  mean_rel_sigma       combined standard error of that difference
  cells_within_3sigma  fraction of cells whose difference is below three
                       combined standard errors
"""
from __future__ import annotations

import numpy as np

from . import abi
from .core import Case

KB = abi.KB
FIELDS = ("rhoN", "translational_T", "rotational_T", "p")


def run_blocks(spec, fill, warmup, sample, blocks, precision):
    c = Case(spec, precision=precision, use_graphs=True)
    c.fill(**fill)
    c.step(warmup)
    per = max(sample // blocks, 1)
    out = []
    for _ in range(blocks):
        c.reset_samples()
        c.step(per)
        c.sync()
        f = c.fields()
        out.append({k: (f[k].copy() if k != "U_mean" else f[k].copy()) for k in FIELDS + ("U_mean",)})
    c.close()
    return out


def _stats(blocks, key, comp=None):
    a = np.array([b[key] if comp is None else b[key][:, comp] for b in blocks])   # [blocks, cells]
    mean = a.mean(0)
    err = a.std(0, ddof=1) / np.sqrt(a.shape[0])
    dom = a.mean(1)                       # domain mean per block
    return mean, err, dom.mean(), dom.std(ddof=1) / np.sqrt(a.shape[0])


def compare(spec, fill, warmup=100, sample=400, blocks=8):
    b64 = run_blocks(spec, fill, warmup, sample, blocks, "fp64")
    b32 = run_blocks(spec, fill, warmup, sample, blocks, "fp32")
    has_rot = any(s.rotational_dof > 0 for s in spec.species)
    out = {}
    keys = [(k, None) for k in FIELDS if k != "rotational_T" or has_rot] + [("U_mean", d) for d in range(3)]
    for key, comp in keys:
        m64, e64, d64, s64 = _stats(b64, key, comp)
        m32, e32, d32, s32 = _stats(b32, key, comp)
        ok = (e64 > 0) & (e32 > 0)
        z = np.abs(m32 - m64)[ok] / np.sqrt(e64[ok] ** 2 + e32[ok] ** 2)
        name = key if comp is None else "U" + "xyz"[comp]
        if comp is None:
            scale = abs(d64)
        else:
            # velocity components can average to ~0: scale by the thermal speed
            m = np.mean([s.mass for s in spec.species])
            scale = float(np.sqrt(KB * np.mean([b["translational_T"] for b in b64]) / m))
        out[name] = {
            "mean_rel_diff": float(abs(d32 - d64) / scale) if scale > 0 else 0.0,
            "mean_rel_sigma": float(np.hypot(s32, s64) / scale) if scale > 0 else 0.0,
            "cells_within_3sigma": float(np.mean(z < 3.0)) if z.size else 1.0,
        }
    return out
