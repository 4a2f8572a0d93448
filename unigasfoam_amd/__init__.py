"""unigasfoam_amd - MI355X-native particle time step of uniGasFoam.

The product is libugx.so (HIP kernels for gfx950 behind the C ABI of
include/ugx.h); this package holds its Python binding, the polyMesh and
dictionary front ends, and the benchmark workloads.
"""
from . import abi  # noqa: F401
from .mesh import PolyMesh, box_mesh, channel_mesh, graded_box_mesh, mapped_box_mesh  # noqa: F401
from .spec import Boundary, CaseSpec, Species  # noqa: F401

__all__ = ["abi", "PolyMesh", "box_mesh", "channel_mesh", "graded_box_mesh", "mapped_box_mesh",
           "Boundary", "CaseSpec", "Species"]


def Case(*args, **kwargs):
    """Plain GPU solver (see unigasfoam_amd.core.Case)."""
    from .core import Case as _Case
    return _Case(*args, **kwargs)
