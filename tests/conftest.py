"""Shared fixtures.  GPU tests are marked ``gpu``; everything else runs on
the CPU (oracle, host logic, ABI checks)."""
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libnbots_amd.so)")


def load_golden(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))


def load_json(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def problem_from_golden(g):
    """Rebuild an nbots_amd.synth.Problem from a golden fixture's inputs."""
    from nbots_amd.synth import Problem
    elem_E = g["elem_E"] if "elem_E" in g and g["elem_E"].size else None
    p = Problem(int(g.get("nx", 0)), int(g.get("ny", 0)), g["nod"], g["conn"], g["vtx"],
                g["sgm_ptr"], g["sgm_nodes"], E=float(g["E"]), nu=float(g["nu"]),
                thickness=float(g["thickness"]), elem_E=elem_E)
    for k in ("dv", "nv", "ds", "ns"):
        setattr(p, k, (g[k + "_id"].astype(np.uint32), g[k + "_mask"].astype(np.uint8),
                       g[k + "_val"].astype(np.float64)))
    return p


SMALL_CASES = ["cant_2x1", "cant_4x2_jit", "cant_8x4_jit", "traction_16x8", "mbb_16x8"]
