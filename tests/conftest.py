"""Shared fixtures.  GPU tests are marked ``gpu``; everything else runs on
the CPU (oracle, host logic, ABI checks)."""
import json
import math
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
    if "kirsch_sgm" in g:
        sx, sy = (int(s) for s in g["kirsch_sgm"])
        p.ns_fn = ((sx, (1, 1), kirsch_xface), (sy, (1, 1), kirsch_yface))
    return p


def kirsch_stress(x, y):
    """Kirsch's stress around a circular hole of radius 0.5 in an infinite
    plate under a far-field x traction of 1e4 (the plate-with-a-hole test,
    static_elasticity2D.c:189-203), in the same operation order."""
    a, tx = 0.5, 1e4
    r = math.sqrt(x * x + y * y)
    th = math.atan2(y, x)
    q2 = (a / r) * (a / r)
    q4 = 1.5 * (q2 * q2)
    c2, c4 = math.cos(2.0 * th), math.cos(4.0 * th)
    s2, s4 = math.sin(2.0 * th), math.sin(4.0 * th)
    return (tx * (1.0 - q2 * (1.5 * c2 + c4) + q4 * c4),
            tx * (-q2 * (0.5 * c2 - c4) - q4 * c4),
            tx * (-q2 * (0.5 * s2 + s4) + q4 * s4))


def kirsch_xface(x, y):
    s = kirsch_stress(x, y)
    return s[0], s[2]


def kirsch_yface(x, y):
    s = kirsch_stress(x, y)
    return s[2], s[1]


def vm_stress(sxx, syy, sxy):
    """von Mises stress (formulas.c:65-68)."""
    return np.sqrt(sxx * sxx + syy * syy - sxx * syy + 3.0 * (sxy * sxy))


def kirsch_vm_error(p, stress):
    """Mean |1 - vm/vm_Kirsch| over the T3 Gauss points (the centroids), the
    plate-with-a-hole test's measure (static_elasticity2D.c:141-187)."""
    c = p.conn.reshape(-1, 3)
    xy = p.nod.reshape(-1, 2)
    w = 1.0 / 3.0
    gx = w * xy[c[:, 0], 0] + w * xy[c[:, 1], 0] + w * xy[c[:, 2], 0]
    gy = w * xy[c[:, 0], 1] + w * xy[c[:, 1], 1] + w * xy[c[:, 2], 1]
    ana = np.array([kirsch_stress(x, y) for x, y in zip(gx, gy)])
    s = stress.reshape(-1, 3)
    vm = vm_stress(s[:, 0], s[:, 1], s[:, 2])
    return float(np.mean(np.abs(1.0 - vm / vm_stress(ana[:, 0], ana[:, 1], ana[:, 2]))))


SMALL_CASES = ["cant_2x1", "cant_4x2_jit", "cant_8x4_jit", "traction_16x8", "mbb_16x8"]
