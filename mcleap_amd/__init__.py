"""mcleap_amd: kinetic 2-D Delaunay maintenance of moving points on MI355X.

The engine is libmcleap_hip.so (hand-written gfx950 HIP kernels behind the C
ABI in include/mcleap_hip.h).  This package holds the build script and a
Python mirror of the reference's host interface (HostTriangulation,
DeviceTriangulation, GridCount2d) used by tests and benchmarks.
"""
from ._lib import (GRID_MORTON, GRID_OPT_BLOCK, GRID_PERIODIC, GRID_ROWMAJOR, MOTION_GAUSS, MOTION_SHEAR, MOVE_FLIPFLOP,
                   MOVE_FLIPFLOP_DELAUNAY, MOVE_UNTANGLING, MOVE_UNTANGLING_DELAUNAY, OPT_BIG_GRID_MIN, OPT_GRAPH,
                   OPT_RELOC_TAIL, OPT_REVERT_TAIL, OPT_STAR_ENABLE, OPT_STAR_GIVE_UP, OPT_STAR_HEAD_PASSES,
                   OPT_STAR_MIN_VERTICES, OPT_TAIL_THRESHOLD, MclError, lib)
from .triangulation import (DeviceTriangulation, GridCount2d, GridCount2dBatch, HostTriangulation,
                            HostTriangulationBatch)

__all__ = [
    "HostTriangulation", "HostTriangulationBatch", "DeviceTriangulation", "GridCount2d", "GridCount2dBatch",
    "MclError", "lib",
    "MOTION_GAUSS", "MOTION_SHEAR", "GRID_ROWMAJOR", "GRID_MORTON", "GRID_PERIODIC", "GRID_OPT_BLOCK",
    "OPT_STAR_MIN_VERTICES", "OPT_STAR_ENABLE", "OPT_STAR_GIVE_UP", "OPT_GRAPH", "OPT_TAIL_THRESHOLD",
    "OPT_REVERT_TAIL", "OPT_RELOC_TAIL", "OPT_BIG_GRID_MIN", "OPT_STAR_HEAD_PASSES",
    "MOVE_FLIPFLOP", "MOVE_FLIPFLOP_DELAUNAY", "MOVE_UNTANGLING", "MOVE_UNTANGLING_DELAUNAY",
]
