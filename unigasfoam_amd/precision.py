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


def run_replicas(spec, fill, warmup, sample, replicas, precision, seed0=1000):
    """Fields of `replicas` independent copies of a case (seeds seed0, seed0+1,
    ...), run together as one batch: warmup steps, a sampler reset, then
    `sample` sampled steps."""
    from .batch import Batch
    b = Batch(precision=precision, use_graphs=True)
    n = spec.mesh.n_cells * fill.get("exact_per_cell", 0) or None
    slots = [b.admit(spec, fill, int(2 * n) if n else 0, seed=seed0 + r) for r in range(replicas)]
    if any(s is None for s in slots):
        b.close()
        raise RuntimeError("the replicas do not fit one batch")
    b.step(warmup)
    for s in slots:
        b.case(s).reset_samples()
    b.step(sample)
    b.sync()
    out = [b.case(s).fields() for s in slots]
    b.close()
    return out


REPLICA_FIELDS = (("rhoN", None), ("translational_T", None), ("rotational_T", None), ("overall_T", None),
                  ("p", None), ("U_mean", 0), ("U_mean", 1), ("U_mean", 2), ("shear_stress", 1),
                  ("heat_flux", 1))


def compare_replicas(spec, fill, warmup=1000, sample=2000, replicas=16, seed0=1000):
    """FP32-state against all-FP64 as two ensembles of independent replicas.

    Replica r of both ensembles starts from the same meshFill (seed
    seed0 + r; the FP32 one stores it rounded), so the replicas are compared
    in pairs: the paired difference cancels what a replica inherits from its
    own fill (its total energy, say), which otherwise dominates the scatter
    of a short run.  The standard errors come from the scatter of the paired
    differences over the independent replicas (unlike batch means of one
    run, whose blocks are correlated in a slowly varying flow).  Per field:
    the mean paired difference of
    the domain mean, relative to its scale (|domain mean| of the FP64
    ensemble; the thermal speed for velocity components, the mean pressure
    for the shear stress P_xy, the thermal energy flux p c_th for q_y), its
    combined standard error on the same scale, and the fraction of cells
    whose ensemble means differ by less than three combined standard errors.
    """
    f64 = run_replicas(spec, fill, warmup, sample, replicas, "fp64", seed0)
    f32 = run_replicas(spec, fill, warmup, sample, replicas, "fp32", seed0)
    has_rot = any(s.rotational_dof > 0 for s in spec.species)
    m = float(np.mean([s.mass for s in spec.species]))
    T0 = float(np.mean([f["translational_T"].mean() for f in f64]))
    p0 = float(np.mean([f["p"].mean() for f in f64]))
    c_th = float(np.sqrt(KB * T0 / m))
    out = {}
    for key, comp in REPLICA_FIELDS:
        if key == "rotational_T" and not has_rot:
            continue
        if key in ("shear_stress", "heat_flux") and not spec.sampler_full:
            continue
        a64 = np.array([f[key] if comp is None else f[key][:, comp] for f in f64])   # [replicas, cells]
        a32 = np.array([f[key] if comp is None else f[key][:, comp] for f in f32])
        dd = a32.mean(1) - a64.mean(1)                   # paired differences of the domain means
        diff = dd.mean()
        se = dd.std(ddof=1) / np.sqrt(replicas)
        d64 = a64.mean(1)
        if key == "U_mean":
            scale = c_th
        elif key == "shear_stress":
            scale = p0
        elif key == "heat_flux":
            scale = p0 * c_th
        else:
            scale = abs(d64.mean())
        dc = a32 - a64                                    # paired differences per cell
        e = dc.std(0, ddof=1) / np.sqrt(replicas)
        ok = e > 0
        z = np.abs(dc.mean(0))[ok] / e[ok]
        name = key if comp is None else key + "_" + "xyz"[comp]
        out[name] = {"mean_rel_diff": float(abs(diff) / scale), "mean_rel_sigma": float(se / scale),
                     "z": float(abs(diff) / se) if se > 0 else 0.0,
                     "cells_within_3sigma": float(np.mean(z < 3.0)) if z.size else 1.0}
    return out
