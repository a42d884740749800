"""Captured-graph replay of the solvers' repeated steps against kernel-by-kernel
launches.

The FP64 PCG replays four iterations per HIP graph launch and the Chebyshev
flexible CG one outer iteration; between replays the host polls the device
state (a copy and a synchronisation), and the last batch's kernels return at
once after the stop.  Every replay runs the same kernels with the same
arguments as the eager launches, so every solve below must give the same
bits, the same iteration counts and the same residuals either way.  Each
solve replays its graph hundreds of times, with that eager work between the
replays, and the SIMP loop does so across several solves whose matrix and
preconditioner bound change in between.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode,kw", [
    ("fp64", dict(true_residual=True)),
    ("reference", dict()),
    ("mixed16", dict(true_residual=True)),
    ("mixed64", dict(true_residual=True, precond_bits=64)),
])
def test_replay_matches_eager_launches(mode, kw):
    from nbots_amd import synth
    from nbots_amd.fem import Fem, SOLVE_FP64, SOLVE_MIXED, SOLVE_REFERENCE, solve_opts
    p = synth.cantilever_traction(96, 48, seed=7)
    m = {"fp64": SOLVE_FP64, "reference": SOLVE_REFERENCE}.get(mode, SOLVE_MIXED)
    # the reference mode's tolerance is absolute on ||Ku - F||
    tol = 1e-9 if mode == "reference" else 1e-8
    res = {}
    for eager in (True, False):
        print(f"[replay] {mode} eager={eager} 96x48 cantilever", flush=True)
        f = Fem.from_problem(p)
        info = f.solve(m, tol=tol, max_iter=400000, opts=solve_opts(replay=not eager, **kw))
        res[eager] = (f.displacement(), info)
        f.destroy()
    (ue, ie), (ur, ir) = res[True], res[False]
    assert ie["status"] == 0 and ir["status"] == 0
    # long enough for several batches of replays (PCG: 4 iterations per
    # replay; flexible CG: one outer iteration)
    assert ie["iters"] > (16 if m == SOLVE_MIXED else 256), ie["iters"]
    for k in ("iters", "outer", "fp64_iters", "rel_res", "tol_reached", "fallback"):
        assert ie[k] == ir[k], (k, ie[k], ir[k])
    assert np.array_equal(ue, ur), np.abs(ue - ur).max()


def test_replay_simp_history_matches_eager():
    from nbots_amd import synth
    from nbots_amd.fem import Fem, Simp, SOLVE_FP64, SOLVE_MIXED, MIXED_AMG_FCG, solve_opts
    p = synth.mbb(96, 48, rho=0.5)
    hist = {}
    cases = ((SOLVE_MIXED, {}), (SOLVE_FP64, {}), ("amg", {"variant": MIXED_AMG_FCG}))
    for eager in (True, False):
        for mode, extra in cases:
            print(f"[replay] SIMP mode={mode} eager={eager} 96x48 MBB", flush=True)
            f = Fem.from_problem(p)
            s = Simp(f, rmin=1.5, volfrac=0.5, penal=3.0, E0=1.0, Emin=1e-9, move=0.2)
            o = solve_opts(simp=True, true_residual=True, replay=not eager, **extra)
            h = []
            for _ in range(4):
                C, change, info = s.iterate(SOLVE_MIXED if mode == "amg" else mode, tol=1e-8,
                                            opts=o)
                h.append((C, change, info["iters"], info["rel_res"]))
            x, xp, dc, _ = s.state()
            hist[(mode, eager)] = (h, x, xp)
            s.destroy()
            f.destroy()
    for mode, _ in cases:
        he, xe, xpe = hist[(mode, True)]
        hr, xr, xpr = hist[(mode, False)]
        assert he == hr
        assert np.array_equal(xe, xr) and np.array_equal(xpe, xpr)
