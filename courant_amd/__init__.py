"""MI355X FDTD engine behind CourantEM's courant.h C ABI.

courant_amd.lib     ctypes binding of libcourant.so (and any library with the same ABI)
courant_amd.scenes  scene builders (parity scene, GCPW workloads)
courant_amd.build   in-tree build of libcourant.so
"""
from . import lib, scenes  # noqa: F401
