"""CPU-side checks of the product library: it loads, exports every symbol
include/ugx.h declares, converts meshes exactly as the oracle does (bit for
bit), and its __host__ __device__ primitives round exactly as the oracle's.
No compute call reaches a device here."""
import os
import re
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from unigasfoam_amd import core, workloads
from oracle.orc import OracleCase

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_header_symbols():
    L = core.lib()
    hdr = open(os.path.join(ROOT, "include", "ugx.h")).read()
    declared = sorted(set(re.findall(r"^\s*(?:const\s+char\s*\*|int)\s+(ugx_\w+)\s*\(", hdr, re.M)))
    assert declared, "no declarations parsed"
    for name in declared:
        assert hasattr(L, name), f"libugx.so does not export {name}"
    assert sorted(core.EXPORTS) == declared
    assert L.ugx_version().startswith(b"ugx")


def _same_bits(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint8), np.ascontiguousarray(b).view(np.uint8))


@pytest.mark.parametrize("name", ["heat_bath", "couette_cyclic", "couette_empty", "cylinder", "nozzle"])
def test_geometry_bit_identical_to_oracle(name):
    if name == "heat_bath":
        spec, _ = workloads.heat_bath(n_cells=5)
    elif name == "couette_cyclic":
        spec, _ = workloads.couette(n_cells=20)
    elif name == "couette_empty":
        spec, _ = workloads.couette(n_cells=20, periodic="empty")
    elif name == "cylinder":
        spec, _ = workloads.cylinder(n_theta=16, n_r=8, target=1e4)
    else:
        spec, _ = workloads.nozzle(nx=8, ny=4, nz=4, per_cell=5)
    g = core.geometry(spec)
    o = OracleCase(spec, capacity=1024).geometry()
    for k in ("cc", "vol", "nfaces", "planes", "nbr", "delta", "nbar"):
        assert _same_bits(g[k], o[k]), f"{name}: {k} differs"


def test_mesh_invariants():
    """Closed cells (sum of outward area vectors = 0 <=> planes valid),
    positive volumes summing to the domain volume."""
    spec, _ = workloads.cylinder(n_theta=40, n_r=20, target=1e4)
    g = core.geometry(spec)
    # half annulus of straight-sided cells: exact polygon area * depth
    th = np.linspace(np.pi, 0.0, 41)
    m = 20
    w = 10.0 ** (np.arange(m) / (m - 1))
    e = np.concatenate([[0.0], np.cumsum(w)]) / np.sum(w)
    r = 0.1 + 0.9 * e
    area = 0.0
    for i in range(40):
        for j in range(20):
            pts = [(r[j] * np.cos(th[i]), r[j] * np.sin(th[i])), (r[j] * np.cos(th[i + 1]), r[j] * np.sin(th[i + 1])),
                   (r[j + 1] * np.cos(th[i + 1]), r[j + 1] * np.sin(th[i + 1])), (r[j + 1] * np.cos(th[i]), r[j + 1] * np.sin(th[i]))]
            area += 0.5 * abs(sum(pts[k][0] * pts[(k + 1) % 4][1] - pts[(k + 1) % 4][0] * pts[k][1] for k in range(4)))
    assert np.all(g["vol"] > 0)
    assert abs(g["vol"].sum() - area * 0.01) < 1e-12 * area
    # planes are unit normals for real faces; padded slots are (0,0,0,1)
    n = np.linalg.norm(g["planes"][..., :3], axis=-1)
    real = np.arange(6)[None, :] < g["nfaces"][:, None]
    assert np.allclose(n[real], 1.0, atol=1e-14)
    assert np.all(g["planes"][~real][:, 3] == 1.0)
    # the reference point lies inside every cell (all d > 0)
    assert np.all(g["planes"][real][:, 3] > 0)


def test_device_primitives_match_oracle_on_host(tmp_path):
    """Compiles ugx_dev.hpp's __host__ __device__ primitives for the host and
    checks them bit for bit against the oracle's orc_prim.h."""
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    src = os.path.join(ROOT, "tests", "native", "prim_parity.cpp")
    exe = str(tmp_path / "prim_parity")
    subprocess.run([hipcc, "-O2", "-std=c++17", "-ffp-contract=off", "-x", "hip", "--offload-arch=gfx950",
                    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                    "-I", os.path.join(ROOT, "unigasfoam_amd", "csrc"), "-I", os.path.join(ROOT, "oracle"),
                    src, "-o", exe], check=True, capture_output=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    assert "mismatches 0" in out, out
