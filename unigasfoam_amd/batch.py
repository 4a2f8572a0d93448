"""Batched form: many independent cases in one device world (ugx_batch_*).

A Batch advances all of its cases with the same kernel launches.  Cases are
admitted with their meshFill and retired when done; both take effect at the
next step(), when the library rebuilds the world and carries every remaining
case's particles, samplers and per-cell state on the device.  A case gives
the same results alone (core.Case) and in any batch.
"""
from __future__ import annotations

import ctypes as C

from . import abi
from .core import Case, check, lib
from .spec import CaseSpec, fill_desc


class BatchCase(Case):
    """Non-owning per-case view of a batch slot: samples, fields, counters,
    particles, sigma_max, ... as for a plain Case; stepping goes through the
    batch."""

    def __init__(self, batch: "Batch", slot: int, spec: CaseSpec):
        self.L = batch.L
        self.spec = spec
        self.nc = spec.mesh.n_cells
        self.slot = slot
        h = C.c_void_p()
        check(self.L.ugx_batch_case(batch.h, slot, C.byref(h)), self.L)
        self.h = h

    def close(self):
        self.h = None

    def step(self, n=1):
        raise RuntimeError("a batch case advances with Batch.step()")


def footprint(spec: CaseSpec, expected_particles: int, precision="fp32") -> int:
    L = lib()
    md, _sp, _ns, m, _bd, _nb = spec.descriptors()
    out = C.c_int64()
    prec = abi.FP32_STATE if precision == "fp32" else abi.FP64
    check(L.ugx_case_footprint(C.byref(md), C.byref(m), int(expected_particles), prec, C.byref(out)), L)
    return out.value


class Batch:
    def __init__(self, precision="fp32", hbm_budget=0, max_cases=0, use_graphs=True, device=0):
        self.L = lib()
        o = abi.BatchOptions()
        o.precision = abi.FP32_STATE if precision == "fp32" else abi.FP64
        o.device = device
        o.hbm_budget = int(hbm_budget)
        o.max_cases = int(max_cases)
        o.use_graphs = 1 if use_graphs else 0
        h = C.c_void_p()
        check(self.L.ugx_batch_create(C.byref(o), C.byref(h)), self.L)
        self.h = h
        self.precision = precision
        self.cases = {}

    def close(self):
        if getattr(self, "h", None):
            for c in self.cases.values():
                c.close()
            self.cases.clear()
            self.L.ugx_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def admit(self, spec: CaseSpec, fill: dict | None, expected_particles: int, seed=None):
        """Queues a case; returns its slot, or None when it does not fit."""
        md, sp, ns, m, bd, nb = spec.descriptors()
        f = fill_desc(spec, **fill) if fill is not None else None
        slot = C.c_int(-1)
        rc = self.L.ugx_batch_admit(self.h, C.byref(md), sp, ns, C.byref(m), bd, nb,
                                    C.byref(f) if f is not None else None,
                                    spec.seed if seed is None else seed, int(expected_particles), C.byref(slot))
        if rc == abi.ERR_CAPACITY:
            return None
        check(rc, self.L)
        self.cases[slot.value] = BatchCase(self, slot.value, spec)
        return slot.value

    def case(self, slot) -> BatchCase:
        return self.cases[slot]

    def retire(self, slot):
        c = self.cases.pop(slot)
        c.close()
        check(self.L.ugx_batch_retire(self.h, int(slot)), self.L)

    def step(self, n=1):
        check(self.L.ugx_batch_step(self.h, int(n)), self.L)

    def sync(self):
        check(self.L.ugx_batch_sync(self.h), self.L)

    def info(self):
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        check(self.L.ugx_batch_info(self.h, C.byref(a), C.byref(b), C.byref(c)), self.L)
        return {"cases": a.value, "bytes": b.value, "particles": c.value}
