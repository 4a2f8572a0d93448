"""ctypes binding of the CPU oracle (oracle/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  Same interface as
unigasfoam_amd.core.Case so that parity tests drive both identically.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os

import numpy as np

from unigasfoam_amd import abi
from unigasfoam_amd.core import FIELD_NAMES, alloc_soa, make_upload_soa
from unigasfoam_amd.spec import CaseSpec, fill_desc

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liborc.so")
_lib = None


class OracleError(RuntimeError):
    pass


def _load(path):
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built; run `make -C oracle`")
    L = C.CDLL(path)
    _bind(L)
    return L


@contextlib.contextmanager
def using(path):
    """Run the enclosed code against another build of the oracle (a library
    compiled from an edited copy of orc.c); cases created inside bind it."""
    global _lib
    saved = _lib
    _lib = _load(path)
    try:
        yield _lib
    finally:
        _lib = saved


def lib():
    global _lib
    if _lib is None:
        _lib = _load(LIB_PATH)
    return _lib


def _bind(L):
    vp = C.c_void_p
    L.orc_last_error.restype = C.c_char_p
    L.orc_create.argtypes = [C.POINTER(abi.MeshDesc), C.POINTER(abi.SpeciesDesc), C.c_int,
                             C.POINTER(abi.ModelDesc), C.POINTER(abi.BoundaryDesc), C.c_int,
                             C.c_uint64, C.c_int64, C.POINTER(vp)]
    L.orc_destroy.argtypes = [vp]
    L.orc_fill.argtypes = [vp, C.POINTER(abi.FillDesc)]
    L.orc_upload.argtypes = [vp, C.POINTER(abi.ParticleSoA)]
    L.orc_download.argtypes = [vp, C.POINTER(abi.ParticleSoA)]
    L.orc_count.argtypes = [vp]
    L.orc_count.restype = C.c_int64
    L.orc_step.argtypes = [vp, C.c_int]
    L.orc_step_phases.argtypes = [vp, C.c_int]
    L.orc_samples.argtypes = [vp, abi.DBL_P, C.c_int64]
    L.orc_counters.argtypes = [vp, C.POINTER(abi.Counters)]
    L.orc_sigma_max.argtypes = [vp, abi.DBL_P, C.c_int64]
    L.orc_set_sigma_max.argtypes = [vp, abi.DBL_P, C.c_int64]
    L.orc_cell_model.argtypes = [vp, abi.I32_P, C.c_int64]
    L.orc_kn.argtypes = [vp, abi.DBL_P, C.c_int64]
    L.orc_hybrid_set.argtypes = [vp, abi.DBL_P, C.c_double]
    L.orc_hybrid_decompose.argtypes = [vp]
    L.orc_smooth.argtypes = [vp, abi.DBL_P, abi.DBL_P, C.c_int]
    L.orc_refine.argtypes = [vp, abi.I32_P]
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_set_fast_sums.argtypes = [C.c_int]
    L.orc_geometry.argtypes = [vp, abi.DBL_P, abi.DBL_P, abi.I32_P, abi.DBL_P, abi.I32_P, abi.DBL_P]
    L.orc_face_data.argtypes = [vp, abi.DBL_P, abi.I32_P]
    L.orc_inflow_nbar.argtypes = [vp, abi.DBL_P, C.c_int64, C.POINTER(C.c_int64)]
    L.orc_usp_sample.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, abi.DBL_P, abi.DBL_P,
                                 C.c_double, C.c_int64, C.c_uint64, abi.DBL_P, C.POINTER(C.c_double)]
    L.orc_pair_collide.argtypes = [vp, C.c_int, C.c_int, abi.DBL_P, abi.DBL_P, C.c_double, C.c_double,
                                   C.c_int64, C.c_uint64, abi.DBL_P, abi.DBL_P, abi.DBL_P, abi.DBL_P]
    L.orc_usp_state.argtypes = [vp, abi.DBL_P, abi.DBL_P, abi.DBL_P]
    L.orc_set_usp_state.argtypes = [vp, abi.DBL_P, abi.DBL_P, abi.DBL_P]
    L.orc_samples_reset.argtypes = [vp]
    L.orc_wall_face_count.argtypes = [vp]
    L.orc_wall_face_count.restype = C.c_int64
    L.orc_wall_samples.argtypes = [vp, abi.DBL_P, C.c_int64]
    L.orc_cell_weights.argtypes = [vp, abi.DBL_P, C.c_int64]
    L.orc_set_cell_weights.argtypes = [vp, abi.DBL_P, C.c_int64]
    L.orc_subcell_levels.argtypes = [vp, abi.I32_P, C.c_int64]
    L.orc_set_subcell_levels.argtypes = [vp, abi.I32_P, C.c_int64]
    L.orc_delta_t.argtypes = [vp]
    L.orc_delta_t.restype = C.c_double
    L.orc_time.argtypes = [vp]
    L.orc_time.restype = C.c_double
    L.orc_adapt_inject.argtypes = [vp, abi.DBL_P, C.c_double]
    L.orc_set_cell_model.argtypes = [vp, abi.I32_P, C.c_int64]
    L.orc_smooth_max_ratio.argtypes = [vp, abi.DBL_P, C.c_double]
    L.orc_mi_probe.argtypes = [vp, abi.DBL_P, C.c_int64, abi.DBL_P, abi.I32_P, abi.DBL_P]


def _check(rc):
    if rc != 0:
        raise OracleError(f"oracle error {rc}: {lib().orc_last_error().decode()}")


def usp_sample(cQ, cS, p, u0, q, sigma, m0, n, seed=1):
    """n USP velocity draws (units of u0) and the largest raise of maxProb."""
    q = np.ascontiguousarray(q, dtype=np.float64)
    sh = np.ascontiguousarray(sigma, dtype=np.float64)
    xi = np.zeros((int(n), 3))
    raise_ = C.c_double(0.0)
    _check(lib().orc_usp_sample(cQ, cS, p, u0, abi.dptr(q), abi.dptr(sh), m0, int(n), int(seed),
                                abi.dptr(xi), C.byref(raise_)))
    return xi, raise_.value


def set_threads(n):
    return lib().orc_set_threads(int(n))


def set_fast_sums(on):
    lib().orc_set_fast_sums(1 if on else 0)


class OracleCase:
    def __init__(self, spec: CaseSpec, capacity=1 << 20, seed=None):
        self.L = lib()
        self.spec = spec
        self.nc = spec.mesh.n_cells
        md, sp, ns, m, bd, nb = spec.descriptors()
        h = C.c_void_p()
        _check(self.L.orc_create(C.byref(md), sp, ns, C.byref(m), bd, nb,
                                 spec.seed if seed is None else seed, capacity, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def fill(self, number_density, T, rot_T=None, velocity=(0.0, 0.0, 0.0), exact_per_cell=0):
        f = fill_desc(self.spec, number_density, T, rot_T, velocity, exact_per_cell)
        _check(self.L.orc_fill(self.h, C.byref(f)))

    def upload(self, velocity, cell, species=None, erot=None, position=None, local=None):
        soa, keep = make_upload_soa(velocity, cell, species, erot, position, local)
        _check(self.L.orc_upload(self.h, C.byref(soa)))

    def step(self, n=1):
        _check(self.L.orc_step(self.h, int(n)))

    def step_phases(self, mask):
        _check(self.L.orc_step_phases(self.h, int(mask)))

    def sync(self):
        pass

    def count(self):
        return int(self.L.orc_count(self.h))

    def particles(self):
        n = self.count()
        soa, a = alloc_soa(max(n, 1))
        _check(self.L.orc_download(self.h, C.byref(soa)))
        return {k: v[: soa.n] for k, v in a.items()}

    def cell_weights(self):
        out = np.zeros(self.nc)
        _check(self.L.orc_cell_weights(self.h, abi.dptr(out), self.nc))
        return out

    def set_cell_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float64)
        _check(self.L.orc_set_cell_weights(self.h, abi.dptr(w), self.nc))

    def delta_t(self):
        return float(self.L.orc_delta_t(self.h))

    def time(self):
        return float(self.L.orc_time(self.h))

    def adapt_inject(self, acc, tav):
        """Sets the adapter's time sums [n_cells, 16] and their time, then adapts once."""
        acc = np.ascontiguousarray(acc, dtype=np.float64).reshape(self.nc, abi.KN_ROWS)
        _check(self.L.orc_adapt_inject(self.h, abi.dptr(acc), float(tav)))

    def mi_probe(self, fields, x, cells):
        """cellPoint interpolation of per-cell fields [n_cells, 16] at global
        positions x [n, 3] lying in cells [n]; returns [n, 16]."""
        fields = np.ascontiguousarray(fields, dtype=np.float64).reshape(self.nc, 16)
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        cells = np.ascontiguousarray(cells, dtype=np.int32)
        out = np.zeros((len(x), 16))
        _check(self.L.orc_mi_probe(self.h, abi.dptr(fields), len(x), abi.dptr(x), abi.iptr(cells), abi.dptr(out)))
        return out

    def smooth_max_ratio(self, f, max_ratio):
        f = np.array(f, dtype=np.float64)
        _check(self.L.orc_smooth_max_ratio(self.h, abi.dptr(f), float(max_ratio)))
        return f

    def set_cell_model(self, ids):
        ids = np.ascontiguousarray(np.broadcast_to(np.asarray(ids, dtype=np.int32), (self.nc,)))
        _check(self.L.orc_set_cell_model(self.h, abi.iptr(ids), self.nc))

    def subcell_levels(self):
        out = np.zeros((self.nc, 3), dtype=np.int32)
        _check(self.L.orc_subcell_levels(self.h, abi.iptr(out), out.size))
        return out

    def set_subcell_levels(self, lv):
        lv = np.ascontiguousarray(np.broadcast_to(np.asarray(lv, dtype=np.int32), (self.nc, 3)))
        _check(self.L.orc_set_subcell_levels(self.h, abi.iptr(lv), lv.size))

    def wall_face_count(self):
        return int(self.L.orc_wall_face_count(self.h))

    def wall_samples(self):
        """Wall time sums, [face, species, row] as Case.wall_samples."""
        out = np.zeros((self.wall_face_count(), abi.MAX_SPECIES, abi.W_COUNT))
        _check(self.L.orc_wall_samples(self.h, abi.dptr(out), out.size))
        return out

    def samples(self):
        out = np.zeros((abi.S_COUNT, self.nc))
        _check(self.L.orc_samples(self.h, abi.dptr(out), out.size))
        return out

    def reset_samples(self):
        _check(self.L.orc_samples_reset(self.h))

    def counters(self):
        k = abi.Counters()
        _check(self.L.orc_counters(self.h, C.byref(k)))
        return k.as_dict()

    def sigma_max(self):
        out = np.zeros(self.nc)
        _check(self.L.orc_sigma_max(self.h, abi.dptr(out), self.nc))
        return out

    def set_sigma_max(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        _check(self.L.orc_set_sigma_max(self.h, abi.dptr(v), self.nc))

    def pair_collide(self, p, q, UP, UQ, eP=0.0, eQ=0.0, n=1, seed=1):
        """n collisions of one pair from the same pre-collision state."""
        a = np.ascontiguousarray(UP, dtype=np.float64)
        b = np.ascontiguousarray(UQ, dtype=np.float64)
        oP, oQ = np.zeros((int(n), 3)), np.zeros((int(n), 3))
        eA, eB = np.zeros(int(n)), np.zeros(int(n))
        _check(self.L.orc_pair_collide(self.h, int(p), int(q), abi.dptr(a), abi.dptr(b), float(eP), float(eQ),
                                       int(n), int(seed), abi.dptr(oP), abi.dptr(oQ), abi.dptr(eA), abi.dptr(eB)))
        return oP, oQ, eA, eB

    def usp_state(self):
        q, sh, mp = np.zeros((self.nc, 3)), np.zeros((self.nc, 6)), np.zeros(self.nc)
        _check(self.L.orc_usp_state(self.h, abi.dptr(q), abi.dptr(sh), abi.dptr(mp)))
        return q, sh, mp

    def set_usp_state(self, q=None, sh=None, max_prob=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (q, sh, max_prob)]
        ptrs = [None if a is None else abi.dptr(a) for a in arrs]
        _check(self.L.orc_set_usp_state(self.h, *ptrs))

    def cell_model(self):
        out = np.zeros(self.nc, dtype=np.int32)
        _check(self.L.orc_cell_model(self.h, abi.iptr(out), self.nc))
        return out

    def kn(self):
        out = np.zeros(self.nc)
        _check(self.L.orc_kn(self.h, abi.dptr(out), self.nc))
        return out

    # hybrid decomposition hooks (tests only)
    KN_ROWS = 16

    def hybrid_set(self, knacc, tav):
        a = np.ascontiguousarray(knacc, dtype=np.float64).reshape(self.nc, self.KN_ROWS)
        _check(self.L.orc_hybrid_set(self.h, abi.dptr(a), float(tav)))

    def hybrid_decompose(self):
        _check(self.L.orc_hybrid_decompose(self.h))

    def smooth(self, field):
        f = np.ascontiguousarray(field, dtype=np.float64)
        ncomp = 1 if f.ndim == 1 else f.shape[1]
        out = np.zeros_like(f)
        _check(self.L.orc_smooth(self.h, abi.dptr(f), abi.dptr(out), ncomp))
        return out

    def refine(self, ids):
        a = np.ascontiguousarray(ids, dtype=np.int32).copy()
        _check(self.L.orc_refine(self.h, abi.iptr(a)))
        return a

    def geometry(self):
        n = self.nc
        g = {"cc": np.zeros((n, 3)), "vol": np.zeros(n), "nfaces": np.zeros(n, dtype=np.int32),
             "planes": np.zeros((n, 6, 4)), "nbr": np.zeros((n, 6), dtype=np.int32), "delta": np.zeros((n, 6, 3)),
             "farea": np.zeros((n, 6)), "fkind": np.zeros((n, 6), dtype=np.int32)}
        _check(self.L.orc_geometry(self.h, abi.dptr(g["cc"]), abi.dptr(g["vol"]), abi.iptr(g["nfaces"]),
                                   abi.dptr(g["planes"]), abi.iptr(g["nbr"]), abi.dptr(g["delta"])))
        _check(self.L.orc_face_data(self.h, abi.dptr(g["farea"]), abi.iptr(g["fkind"])))
        nb = np.zeros(1 << 16)
        k = C.c_int64()
        _check(self.L.orc_inflow_nbar(self.h, abi.dptr(nb), len(nb), C.byref(k)))
        g["nbar"] = nb[: k.value]
        return g


__all__ = ["OracleCase", "set_threads", "lib", "FIELD_NAMES"]
