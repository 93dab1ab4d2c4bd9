"""plasmonx_amd: MI355X (gfx950) frequency-response solver for omegaFQ,
omegaFQFmu and BEM (DPCM / IEFPCM) nanoplasmonics.

The compute path is libpxh.so (HIP, C ABI in include/pxh.h); this package
holds the host side: ctypes binding, targets and units, material tables,
synthetic geometries, the solve driver and the request-serving loop.
"""
from . import _lib
from ._lib import Context, PxhError
from .model import AtomSystem, AtomType, Mesh, system_from_fixture, mesh_from_fixture
from .solver import Algorithm, Field, solve_dynamic_field

__all__ = ["Context", "PxhError", "AtomSystem", "AtomType", "Mesh", "Algorithm", "Field", "solve_dynamic_field",
           "system_from_fixture", "mesh_from_fixture"]
