"""ctypes binding of libugx.so (the HIP library) and the plain-solver Case.

The library is loaded from this package directory; importing this module on
a machine without the built library raises, and every call that reaches the
device fails loudly when no GPU is present - there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi
from .spec import CaseSpec, fill_desc

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libugx.so")
_lib = None

# Every symbol include/ugx.h declares.
EXPORTS = [
    "ugx_version", "ugx_last_error", "ugx_case_create", "ugx_case_destroy", "ugx_case_fill", "ugx_case_fill_noalloc",
    "ugx_particles_upload", "ugx_particles_download", "ugx_particle_count", "ugx_step", "ugx_sync",
    "ugx_samples_download", "ugx_samples_reset", "ugx_fields_compute", "ugx_counters_get",
    "ugx_sigma_max_download", "ugx_sigma_max_upload", "ugx_cell_model_download", "ugx_kn_download",
    "ugx_cell_volumes", "ugx_cell_centres", "ugx_geometry", "ugx_state_download", "ugx_state_upload",
    "ugx_wall_face_count", "ugx_wall_faces", "ugx_wall_samples_download", "ugx_wall_fields_compute",
    "ugx_cell_weights_download", "ugx_cell_weights_upload",
    "ugx_subcell_levels_download", "ugx_subcell_levels_upload", "ugx_time_get",
    "ugx_phase_name", "ugx_timing_enable", "ugx_timing_get", "ugx_tile_config", "ugx_probe_copy",
    "ugx_batch_create", "ugx_batch_destroy", "ugx_case_footprint", "ugx_batch_admit", "ugx_batch_retire",
    "ugx_batch_step", "ugx_batch_sync", "ugx_batch_case", "ugx_batch_info",
]


class UgxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ugx error {code}: {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.ugx_version.restype = C.c_char_p
        L.ugx_last_error.restype = C.c_char_p
        vp = C.c_void_p
        L.ugx_case_create.argtypes = [C.POINTER(abi.MeshDesc), C.POINTER(abi.SpeciesDesc), C.c_int,
                                      C.POINTER(abi.ModelDesc), C.POINTER(abi.BoundaryDesc), C.c_int,
                                      C.POINTER(abi.Options), C.POINTER(vp)]
        L.ugx_case_destroy.argtypes = [vp]
        L.ugx_case_fill.argtypes = [vp, C.POINTER(abi.FillDesc)]
        L.ugx_case_fill_noalloc.argtypes = [vp, C.POINTER(abi.FillDesc)]
        L.ugx_particles_upload.argtypes = [vp, C.POINTER(abi.ParticleSoA)]
        L.ugx_particles_download.argtypes = [vp, C.POINTER(abi.ParticleSoA)]
        L.ugx_particle_count.argtypes = [vp, C.POINTER(C.c_int64)]
        L.ugx_step.argtypes = [vp, C.c_int]
        L.ugx_sync.argtypes = [vp]
        L.ugx_samples_download.argtypes = [vp, abi.DBL_P, C.c_int64]
        L.ugx_samples_reset.argtypes = [vp]
        L.ugx_fields_compute.argtypes = [vp, C.POINTER(abi.Fields)]
        L.ugx_counters_get.argtypes = [vp, C.POINTER(abi.Counters)]
        L.ugx_sigma_max_download.argtypes = [vp, abi.DBL_P, C.c_int64]
        L.ugx_sigma_max_upload.argtypes = [vp, abi.DBL_P, C.c_int64]
        L.ugx_cell_model_download.argtypes = [vp, abi.I32_P, C.c_int64]
        L.ugx_state_download.argtypes = [vp, C.POINTER(abi.CaseState)]
        L.ugx_state_upload.argtypes = [vp, C.POINTER(abi.CaseState)]
        L.ugx_kn_download.argtypes = [vp, abi.DBL_P, C.c_int64]
        L.ugx_cell_weights_download.argtypes = [vp, abi.DBL_P, C.c_int64]
        L.ugx_cell_weights_upload.argtypes = [vp, abi.DBL_P, C.c_int64]
        L.ugx_subcell_levels_download.argtypes = [vp, abi.I32_P, C.c_int64]
        L.ugx_subcell_levels_upload.argtypes = [vp, abi.I32_P, C.c_int64]
        L.ugx_time_get.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.ugx_wall_face_count.argtypes = [vp, C.POINTER(C.c_int64)]
        L.ugx_wall_faces.argtypes = [vp, abi.I32_P, abi.I32_P, C.c_int64]
        L.ugx_wall_samples_download.argtypes = [vp, abi.DBL_P, C.c_int64]
        L.ugx_wall_fields_compute.argtypes = [vp, C.POINTER(abi.WallFields)]
        L.ugx_cell_volumes.argtypes = [vp, abi.DBL_P, C.c_int64]
        L.ugx_cell_centres.argtypes = [vp, abi.DBL_P, C.c_int64]
        L.ugx_geometry.argtypes = [C.POINTER(abi.MeshDesc), C.POINTER(abi.SpeciesDesc), C.c_int,
                                   C.POINTER(abi.ModelDesc), C.POINTER(abi.BoundaryDesc), C.c_int,
                                   abi.DBL_P, abi.DBL_P, abi.I32_P, abi.DBL_P, abi.I32_P, abi.DBL_P,
                                   abi.DBL_P, C.c_int64, C.POINTER(C.c_int64)]
        L.ugx_phase_name.restype = C.c_char_p
        L.ugx_phase_name.argtypes = [C.c_int]
        L.ugx_timing_enable.argtypes = [vp, C.c_int]
        L.ugx_timing_get.argtypes = [vp, abi.DBL_P, C.c_int, C.POINTER(C.c_int64)]
        L.ugx_probe_copy.argtypes = [C.c_int, C.c_int64, C.c_int, C.POINTER(C.c_double)]
        i32p = C.POINTER(C.c_int32)
        L.ugx_tile_config.argtypes = [vp, i32p, i32p, i32p, i32p, i32p]
        L.ugx_batch_create.argtypes = [C.POINTER(abi.BatchOptions), C.POINTER(vp)]
        L.ugx_batch_destroy.argtypes = [vp]
        L.ugx_case_footprint.argtypes = [C.POINTER(abi.MeshDesc), C.POINTER(abi.ModelDesc), C.c_int64, C.c_int,
                                         C.POINTER(C.c_int64)]
        L.ugx_batch_admit.argtypes = [vp, C.POINTER(abi.MeshDesc), C.POINTER(abi.SpeciesDesc), C.c_int,
                                      C.POINTER(abi.ModelDesc), C.POINTER(abi.BoundaryDesc), C.c_int,
                                      C.POINTER(abi.FillDesc), C.c_uint64, C.c_int64, C.POINTER(C.c_int)]
        L.ugx_batch_retire.argtypes = [vp, C.c_int]
        L.ugx_batch_step.argtypes = [vp, C.c_int]
        L.ugx_batch_sync.argtypes = [vp]
        L.ugx_batch_case.argtypes = [vp, C.c_int, C.POINTER(vp)]
        L.ugx_batch_info.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        _lib = L
    return _lib


N_PHASES = 9


def phase_names():
    return [lib().ugx_phase_name(i).decode() for i in range(N_PHASES)]


def probe_copy(nbytes, reps=20, device=0):
    """Seconds per device copy of nbytes (reads nbytes, writes nbytes)."""
    s = C.c_double()
    check(lib().ugx_probe_copy(device, int(nbytes), int(reps), C.byref(s)))
    return s.value


def check(rc, L=None):
    if rc != 0:
        L = L or lib()
        raise UgxError(rc, L.ugx_last_error().decode(errors="replace"))


def alloc_soa(n):
    a = {
        "local": np.zeros((n, 3)), "position": np.zeros((n, 3)), "velocity": np.zeros((n, 3)),
        "erot": np.zeros(n), "cell": np.zeros(n, dtype=np.int32), "species": np.zeros(n, dtype=np.int32),
    }
    soa = abi.ParticleSoA()
    soa.n = n
    soa.local_position = abi.dptr(a["local"])
    soa.position = abi.dptr(a["position"])
    soa.velocity = abi.dptr(a["velocity"])
    soa.erot = abi.dptr(a["erot"])
    soa.cell = abi.iptr(a["cell"])
    soa.species = abi.iptr(a["species"])
    return soa, a


def make_upload_soa(velocity, cell, species=None, erot=None, position=None, local=None):
    n = len(cell)
    keep = {
        "velocity": np.ascontiguousarray(velocity, dtype=np.float64),
        "cell": np.ascontiguousarray(cell, dtype=np.int32),
        "species": np.ascontiguousarray(species if species is not None else np.zeros(n), dtype=np.int32),
        "erot": np.ascontiguousarray(erot if erot is not None else np.zeros(n), dtype=np.float64),
    }
    soa = abi.ParticleSoA()
    soa.n = n
    if local is not None:
        keep["local"] = np.ascontiguousarray(local, dtype=np.float64)
        soa.local_position = abi.dptr(keep["local"])
    else:
        keep["position"] = np.ascontiguousarray(position, dtype=np.float64)
        soa.position = abi.dptr(keep["position"])
    soa.velocity = abi.dptr(keep["velocity"])
    soa.cell = abi.iptr(keep["cell"])
    soa.species = abi.iptr(keep["species"])
    soa.erot = abi.dptr(keep["erot"])
    return soa, keep


FIELD_NAMES = ["rhoN", "rhoM", "U_mean", "translational_T", "rotational_T", "overall_T", "p", "Ma",
               "mfp", "mcr", "density_error", "velocity_error", "temperature_error", "pressure_error",
               "shear_stress", "heat_flux"]


class Case:
    """Plain GPU solver for one case (uniGasCloud with evolve() on the GPU)."""

    def __init__(self, spec: CaseSpec, precision="fp64", capacity=0, device=0, seed=None, use_graphs=False):
        self.L = lib()
        self.spec = spec
        self.nc = spec.mesh.n_cells
        md, sp, ns, m, bd, nb = spec.descriptors()
        opt = abi.Options()
        opt.precision = abi.FP32_STATE if precision == "fp32" else abi.FP64
        opt.device = device
        opt.particle_capacity = capacity
        opt.seed = spec.seed if seed is None else seed
        opt.use_graphs = 1 if use_graphs else 0
        h = C.c_void_p()
        check(self.L.ugx_case_create(C.byref(md), sp, ns, C.byref(m), bd, nb, C.byref(opt), C.byref(h)), self.L)
        self.h = h
        self.precision = precision
        self.seed = int(opt.seed)

    def close(self):
        if getattr(self, "h", None):
            self.L.ugx_case_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def fill(self, number_density, T, rot_T=None, velocity=(0.0, 0.0, 0.0), exact_per_cell=0):
        f = fill_desc(self.spec, number_density, T, rot_T, velocity, exact_per_cell)
        check(self.L.ugx_case_fill(self.h, C.byref(f)), self.L)

    def fill_noalloc(self, number_density, T, rot_T=None, velocity=(0.0, 0.0, 0.0), exact_per_cell=0):
        """The same fill with no device allocation (fixed per-GPU scratch; the
        case's particle capacity must already hold it)."""
        f = fill_desc(self.spec, number_density, T, rot_T, velocity, exact_per_cell)
        check(self.L.ugx_case_fill_noalloc(self.h, C.byref(f)), self.L)

    def upload(self, velocity, cell, species=None, erot=None, position=None, local=None):
        soa, keep = make_upload_soa(velocity, cell, species, erot, position, local)
        check(self.L.ugx_particles_upload(self.h, C.byref(soa)), self.L)

    def step(self, n=1):
        check(self.L.ugx_step(self.h, int(n)), self.L)

    def sync(self):
        check(self.L.ugx_sync(self.h), self.L)

    def count(self):
        n = C.c_int64()
        check(self.L.ugx_particle_count(self.h, C.byref(n)), self.L)
        return n.value

    def particles(self):
        n = self.count()
        soa, a = alloc_soa(max(n, 1))
        check(self.L.ugx_particles_download(self.h, C.byref(soa)), self.L)
        return {k: v[: soa.n] for k, v in a.items()}

    def samples(self):
        out = np.zeros((abi.S_COUNT, self.nc))
        check(self.L.ugx_samples_download(self.h, abi.dptr(out), out.size), self.L)
        return out

    def reset_samples(self):
        check(self.L.ugx_samples_reset(self.h), self.L)

    def counters(self):
        k = abi.Counters()
        check(self.L.ugx_counters_get(self.h, C.byref(k)), self.L)
        return k.as_dict()

    def fields(self):
        n = self.nc
        arrs = {name: np.zeros(n * (3 if name in ("U_mean", "heat_flux") else 6 if name == "shear_stress" else 1))
                for name in FIELD_NAMES}
        f = abi.Fields()
        for name in FIELD_NAMES:
            setattr(f, name, abi.dptr(arrs[name]))
        check(self.L.ugx_fields_compute(self.h, C.byref(f)), self.L)
        arrs["U_mean"] = arrs["U_mean"].reshape(n, 3)
        arrs["heat_flux"] = arrs["heat_flux"].reshape(n, 3)
        arrs["shear_stress"] = arrs["shear_stress"].reshape(n, 6)
        return arrs

    def cell_weights(self):
        """Cell weights W_c (a simulator of cell c stands for F_N W_c molecules)."""
        out = np.zeros(self.nc)
        check(self.L.ugx_cell_weights_download(self.h, abi.dptr(out), self.nc), self.L)
        return out

    def set_cell_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float64)
        check(self.L.ugx_cell_weights_upload(self.h, abi.dptr(w), self.nc), self.L)

    def delta_t(self):
        """The current time step (an adaptive case changes it)."""
        dt, t = C.c_double(), C.c_double()
        check(self.L.ugx_time_get(self.h, C.byref(dt), C.byref(t)), self.L)
        return dt.value

    def time(self):
        dt, t = C.c_double(), C.c_double()
        check(self.L.ugx_time_get(self.h, C.byref(dt), C.byref(t)), self.L)
        return t.value

    def subcell_levels(self):
        """NTC subcell levels per cell and axis, [n_cells, 3]."""
        out = np.zeros((self.nc, 3), dtype=np.int32)
        check(self.L.ugx_subcell_levels_download(self.h, abi.iptr(out), out.size), self.L)
        return out

    def set_subcell_levels(self, lv):
        lv = np.ascontiguousarray(np.broadcast_to(np.asarray(lv, dtype=np.int32), (self.nc, 3)))
        check(self.L.ugx_subcell_levels_upload(self.h, abi.iptr(lv), lv.size), self.L)

    def wall_face_count(self):
        n = C.c_int64()
        check(self.L.ugx_wall_face_count(self.h, C.byref(n)), self.L)
        return n.value

    def wall_faces(self):
        """(patch, face within the patch) of every measured wall face."""
        n = self.wall_face_count()
        patch = np.zeros(n, dtype=np.int32)
        face = np.zeros(n, dtype=np.int32)
        check(self.L.ugx_wall_faces(self.h, abi.iptr(patch), abi.iptr(face), n), self.L)
        return patch, face

    def wall_samples(self):
        """Raw wall sums, [face, species, row] (rows: abi.W_*)."""
        n = self.wall_face_count()
        out = np.zeros((n, abi.MAX_SPECIES, abi.W_COUNT))
        check(self.L.ugx_wall_samples_download(self.h, abi.dptr(out), out.size), self.L)
        return out

    def wall_fields(self):
        """Write-time wall values per measured wall face."""
        n = self.wall_face_count()
        arrs = {k: np.zeros((n, 3) if k in ("U_mean", "force_density") else n) for k in abi.WALL_FIELD_NAMES}
        f = abi.WallFields()
        for k in abi.WALL_FIELD_NAMES:
            setattr(f, k, abi.dptr(arrs[k]))
        check(self.L.ugx_wall_fields_compute(self.h, C.byref(f)), self.L)
        return arrs

    def sigma_max(self):
        out = np.zeros(self.nc)
        check(self.L.ugx_sigma_max_download(self.h, abi.dptr(out), self.nc), self.L)
        return out

    def set_sigma_max(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        check(self.L.ugx_sigma_max_upload(self.h, abi.dptr(v), self.nc), self.L)

    def cell_model(self):
        out = np.zeros(self.nc, dtype=np.int32)
        check(self.L.ugx_cell_model_download(self.h, abi.iptr(out), self.nc), self.L)
        return out

    def kn(self):
        out = np.zeros(self.nc)
        check(self.L.ugx_kn_download(self.h, abi.dptr(out), self.nc), self.L)
        return out

    def _state_struct(self, arrs):
        s = abi.CaseState()
        s.n_cells = self.nc
        for name, _, dt in abi.STATE_ARRAYS:
            setattr(s, name, abi.iptr(arrs[name]) if dt == "i4" else abi.dptr(arrs[name]))
        w = arrs.get("wall_samples")
        if w is not None:
            s.n_wall_faces = w.shape[0]
            s.wall_samples = abi.dptr(w)
        if arrs.get("cell_weight") is not None:
            s.cell_weight = abi.dptr(arrs["cell_weight"])
        if arrs.get("subcell_levels") is not None:
            s.subcell_levels = abi.iptr(arrs["subcell_levels"])
        if arrs.get("adapt_sums") is not None:
            s.adapt_sums = abi.dptr(arrs["adapt_sums"])
        if arrs.get("adapt_ratio") is not None:
            s.adapt_ratio = abi.dptr(arrs["adapt_ratio"])
        return s

    def state(self):
        """Restart state besides the simulators (ugx_state_download): step
        count, hybrid time, counters, the per-cell arrays (cell-major) and
        the wall sums."""
        arrs = {name: np.zeros((self.nc, rows), dtype=dt) for name, rows, dt in abi.STATE_ARRAYS}
        arrs["wall_samples"] = np.zeros((self.wall_face_count(), abi.MAX_SPECIES, abi.W_COUNT))
        arrs["cell_weight"] = np.zeros(self.nc)
        arrs["subcell_levels"] = np.zeros((self.nc, 3), dtype=np.int32)
        arrs["adapt_sums"] = np.zeros((self.nc, abi.KN_ROWS))
        arrs["adapt_ratio"] = np.zeros((self.nc, 3))
        s = self._state_struct(arrs)
        check(self.L.ugx_state_download(self.h, C.byref(s)), self.L)
        arrs.update(steps=np.int64(s.steps), kn_time=np.float64(s.kn_time),
                    sampled_at_reset=np.int64(s.sampled_at_reset),
                    counters=np.array(s.counters[:], dtype=np.uint64),
                    delta_t=np.float64(s.delta_t), time=np.float64(s.time),
                    time_sampled=np.float64(s.time_sampled), adapt_time=np.float64(s.adapt_time))
        return arrs

    def restore_state(self, st):
        """Upload a state from state() (ugx_state_upload)."""
        arrs = {name: np.ascontiguousarray(np.asarray(st[name], dtype=dt).reshape(self.nc, rows))
                for name, rows, dt in abi.STATE_ARRAYS}
        if st.get("cell_weight") is not None:
            arrs["cell_weight"] = np.ascontiguousarray(np.asarray(st["cell_weight"], dtype=np.float64).reshape(self.nc))
        if st.get("wall_samples") is not None:
            arrs["wall_samples"] = np.ascontiguousarray(np.asarray(st["wall_samples"], dtype=np.float64).reshape(
                -1, abi.MAX_SPECIES, abi.W_COUNT))
        for name, rows, dt in (("subcell_levels", 3, np.int32), ("adapt_sums", abi.KN_ROWS, np.float64),
                               ("adapt_ratio", 3, np.float64)):
            if st.get(name) is not None:
                arrs[name] = np.ascontiguousarray(np.asarray(st[name], dtype=dt).reshape(self.nc, rows))
        s = self._state_struct(arrs)
        s.steps = int(st["steps"])
        s.kn_time = float(st["kn_time"])
        s.sampled_at_reset = int(st["sampled_at_reset"])
        # a state without the time fields (an older checkpoint) keeps the
        # case's time step, and its time follows from the step count
        s.delta_t = float(st.get("delta_t", 0.0))
        dt = s.delta_t if s.delta_t > 0.0 else self.spec.delta_t
        s.time = float(st.get("time", int(st["steps"]) * dt))
        s.time_sampled = float(st["time_sampled"]) if "time_sampled" in st else float("nan")
        s.adapt_time = float(st.get("adapt_time", 0.0))
        cnt = np.asarray(st["counters"], dtype=np.uint64)
        for k in range(abi.N_COUNTERS):
            s.counters[k] = int(cnt[k])
        check(self.L.ugx_state_upload(self.h, C.byref(s)), self.L)

    def checkpoint(self):
        """The simulators in storage form (local positions, slot order) and
        the restart state: Case.restart continues from it bit for bit."""
        p = self.particles()
        meta = {"precision": self.precision, "seed": self.seed, "n_cells": self.nc}
        return {"meta": meta, "particles": {k: p[k] for k in CHECKPOINT_PARTICLE_KEYS}, "state": self.state()}

    @classmethod
    def restart(cls, spec: CaseSpec, ck, capacity=0, device=0, use_graphs=False):
        """A new case of the same spec, precision and seed continuing from ck."""
        meta = ck["meta"]
        if int(meta["n_cells"]) != spec.mesh.n_cells:
            raise ValueError("checkpoint is for a mesh of %d cells" % int(meta["n_cells"]))
        c = cls(spec, precision=str(meta["precision"]), capacity=capacity, device=device,
                seed=int(meta["seed"]), use_graphs=use_graphs)
        p = ck["particles"]
        if len(p["cell"]):
            c.upload(velocity=p["velocity"], cell=p["cell"], species=p["species"], erot=p["erot"], local=p["local"])
        c.restore_state(ck["state"])
        return c

    def timing(self, on=True):
        check(self.L.ugx_timing_enable(self.h, 1 if on else 0), self.L)

    def phase_times(self):
        """{phase: total ms} over the timed steps, and the number of steps."""
        ms = np.zeros(N_PHASES)
        n = C.c_int64()
        check(self.L.ugx_timing_get(self.h, abi.dptr(ms), N_PHASES, C.byref(n)), self.L)
        return dict(zip(phase_names(), ms.tolist())), n.value

    def tile_config(self):
        """The cell-tile configuration in use and whether the tuner has finished."""
        v = [C.c_int32() for _ in range(5)]
        check(self.L.ugx_tile_config(self.h, *[C.byref(x) for x in v]), self.L)
        return {"variant": v[0].value, "threads": v[1].value, "window": v[2].value, "lds_simulators": v[3].value,
                "tuned": bool(v[4].value)}

    def cell_volumes(self):
        out = np.zeros(self.nc)
        check(self.L.ugx_cell_volumes(self.h, abi.dptr(out), self.nc), self.L)
        return out

    def cell_centres(self):
        out = np.zeros((self.nc, 3))
        check(self.L.ugx_cell_centres(self.h, abi.dptr(out), out.size), self.L)
        return out


CHECKPOINT_PARTICLE_KEYS = ("local", "velocity", "erot", "cell", "species")


def checkpoint_save(path, ck):
    """Write a checkpoint from Case.checkpoint() as one .npz (no pickled objects)."""
    flat = {"meta_precision": np.array(ck["meta"]["precision"]), "meta_seed": np.uint64(ck["meta"]["seed"]),
            "meta_n_cells": np.int64(ck["meta"]["n_cells"])}
    for k in CHECKPOINT_PARTICLE_KEYS:
        flat["p_" + k] = np.asarray(ck["particles"][k])
    for k, v in ck["state"].items():
        flat["s_" + k] = np.asarray(v)
    np.savez(path, **flat)


def checkpoint_load(path):
    """Read a checkpoint written by checkpoint_save."""
    with np.load(path, allow_pickle=False) as z:
        meta = {"precision": str(z["meta_precision"]), "seed": int(z["meta_seed"]), "n_cells": int(z["meta_n_cells"])}
        particles = {k: z["p_" + k] for k in CHECKPOINT_PARTICLE_KEYS}
        state = {k[2:]: z[k] for k in z.files if k.startswith("s_")}
    return {"meta": meta, "particles": particles, "state": state}


def geometry(spec: CaseSpec):
    """Kernel-form geometry computed by the library's host code (no GPU)."""
    L = lib()
    md, sp, ns, m, bd, nb = spec.descriptors()
    n = spec.mesh.n_cells
    g = {"cc": np.zeros((n, 3)), "vol": np.zeros(n), "nfaces": np.zeros(n, dtype=np.int32),
         "planes": np.zeros((n, 6, 4)), "nbr": np.zeros((n, 6), dtype=np.int32), "delta": np.zeros((n, 6, 3)),
         "nbar": np.zeros(1 << 16)}
    nin = C.c_int64()
    check(L.ugx_geometry(C.byref(md), sp, ns, C.byref(m), bd, nb, abi.dptr(g["cc"]), abi.dptr(g["vol"]),
                         abi.iptr(g["nfaces"]), abi.dptr(g["planes"]), abi.iptr(g["nbr"]), abi.dptr(g["delta"]),
                         abi.dptr(g["nbar"]), len(g["nbar"]), C.byref(nin)), L)
    g["nbar"] = g["nbar"][: nin.value]
    return g
