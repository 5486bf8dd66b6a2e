"""The autotuner's table (mcleap_amd/autotune.py writes it, libmcleap_hip.so
reads it at creation): per workload class (domain and log2 of the vertex
count), the launch and round-loop parameters that measured fastest.

The library applies the triangulation and grid options itself; this module
gives Python callers the same entry (for the FRNN grid geometry,
frnn_cells_per_radius, which the caller chooses when it creates a grid) and
the record of what is in use."""
import json
import math
import os

from ._lib import lib

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_PATH = os.path.join(HERE, "tuning.json")


def table_path():
    """the table the library loaded ("" when none): MCL_TUNING, else
    tuning.json next to the library"""
    p = lib().mcl_tuning_path()
    return p.decode() if p else ""


def load(path=None):
    path = path or table_path()
    if not path or not os.path.exists(path):
        return []
    with open(path) as f:
        return json.load(f).get("entries", [])


def lookup(domain, nv, path=None):
    """the entry the library applies to a mesh of this domain ("torus", "box"
    or "ensemble") and vertex count: same domain, log2 within 2, nearest"""
    l = round(math.log2(nv)) if nv > 1 else 0
    best = None
    for e in load(path):
        if e.get("domain") != domain or abs(e.get("log2_nv", -99) - l) > 2:
            continue
        if best is None or abs(e["log2_nv"] - l) < abs(best["log2_nv"] - l):
            best = e
    return best


def cells_per_radius(domain, nv, path=None):
    """FRNN grid cells per query radius for this workload class (1: cells of
    the radius, as the reference; 2: half-radius cells, a 5 x 5 search)"""
    e = lookup(domain, nv, path)
    return int(e.get("frnn_cells_per_radius", 1)) if e else 1
