"""galaxy_amd: MI355X (gfx950) implementation of GALAXY's polar-grid
particle-mesh step (reference: J. A. Sellwood's GALAXY v15, SRC/lib15).

The compute path is libgalhip.so (HIP kernels behind the C ABI declared in
include/galhip.h); this package holds its ctypes binding, the host-side
mirror of the lib15 interface (lib15.P2DGrid), and the seeded model set-up
(disc, configs).
"""
from ._lib import GalhipError, load  # noqa: F401

__all__ = ["GalhipError", "load", "P2DGrid"]


def __getattr__(name):
    if name == "P2DGrid":
        from .lib15 import P2DGrid
        return P2DGrid
    raise AttributeError(name)
